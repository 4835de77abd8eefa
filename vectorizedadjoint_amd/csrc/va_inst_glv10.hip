// va_inst_glv10.hip -- kernel instantiations for generalized Lotka-Volterra, N = 10
// (one-lane-per-instance and row-per-lane layouts).
#include "va_glv_rows.hpp"

namespace va {

bool kernels_lotka_volterra_10(int tableau, int adaptive, int instances_per_wave,
                               KernelSet *out)
{
    if (instances_per_wave == 64) return pick_tableau<LotkaVolterra<10>>(tableau, adaptive, out);
    if (instances_per_wave != 0 && instances_per_wave != GlvRows<10>::GPW) return false;
    return pick_glv_rows<10>(tableau, adaptive, out);
}

}  // namespace va
