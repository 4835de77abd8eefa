// va_inst_glv10.hip -- kernel instantiations for generalized Lotka-Volterra, N = 10.
#include "va_kernels_impl.hpp"

namespace va {

bool kernels_lotka_volterra_10(int tableau, int adaptive, KernelSet *out)
{
    return pick_tableau<LotkaVolterra<10>>(tableau, adaptive, out);
}

}  // namespace va
