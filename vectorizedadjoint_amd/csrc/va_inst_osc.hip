// va_inst_osc.hip -- kernel instantiations for the damped oscillators.
#include "va_kernels_impl.hpp"

namespace va {

bool kernels_oscillator(int tableau, int adaptive, KernelSet *out)
{
    return pick_tableau<Oscillator>(tableau, adaptive, out);
}

bool kernels_oscillator_fixed_k(int tableau, int adaptive, KernelSet *out)
{
    return pick_tableau<OscillatorFixedK>(tableau, adaptive, out);
}

}  // namespace va
