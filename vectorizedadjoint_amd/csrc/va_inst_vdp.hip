// va_inst_vdp.hip -- kernel instantiations for the Van der Pol system.
#include "va_kernels_impl.hpp"

namespace va {

bool kernels_van_der_pol(int tableau, int adaptive, KernelSet *out)
{
    return pick_tableau<VanDerPol>(tableau, adaptive, out);
}

}  // namespace va
