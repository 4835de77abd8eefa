// The Brusselator's host build (brusselator.hpp, host compiler only, no device
// code) as the C callbacks the CPU oracle takes for a system of the caller's
// (orc_set_user_system): built into tests/cpp/libva_brusselator_host.so.
#include "brusselator.hpp"

extern "C" void brusselator_rhs(const double *x, const double *p, double *f)
{
    Brusselator::rhs(x, p, f);
}

extern "C" void brusselator_vjp(const double *x, const double *p, const double *w, double *wjx,
                                double *wjp)
{
    Brusselator::vjp(x, p, w, wjx, wjp);
}
