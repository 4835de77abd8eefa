// va_inst_vdp.hip -- kernel instantiations for the Van der Pol system.
#include "va_observed.hpp"

namespace va {

bool kernels_van_der_pol(int tableau, int adaptive, KernelSet *out)
{
    if (!pick_tableau<VanDerPol>(tableau, adaptive, out)) return false;
    pick_observed<VanDerPol>(tableau, adaptive, out);
    return true;
}

}  // namespace va
