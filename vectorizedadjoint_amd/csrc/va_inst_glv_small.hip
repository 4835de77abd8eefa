// va_inst_glv_small.hip -- kernel instantiations for generalized Lotka-Volterra, N = 2..5.
#include "va_kernels_impl.hpp"

namespace va {

bool kernels_lotka_volterra_small(int n, int tableau, int adaptive, KernelSet *out)
{
    switch (n) {
    case 2: return pick_tableau<LotkaVolterra<2>>(tableau, adaptive, out);
    case 3: return pick_tableau<LotkaVolterra<3>>(tableau, adaptive, out);
    case 4: return pick_tableau<LotkaVolterra<4>>(tableau, adaptive, out);
    case 5: return pick_tableau<LotkaVolterra<5>>(tableau, adaptive, out);
    default: return false;
    }
}

}  // namespace va
