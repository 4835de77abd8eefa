// va_inst_osc.hip -- kernel instantiations for the damped oscillators.
#include "va_observed.hpp"

namespace va {

bool kernels_oscillator(int tableau, int adaptive, KernelSet *out)
{
    if (!pick_tableau<Oscillator>(tableau, adaptive, out)) return false;
    pick_observed<Oscillator>(tableau, adaptive, out);
    return true;
}

bool kernels_oscillator_fixed_k(int tableau, int adaptive, KernelSet *out)
{
    if (!pick_tableau<OscillatorFixedK>(tableau, adaptive, out)) return false;
    pick_observed<OscillatorFixedK>(tableau, adaptive, out);
    return true;
}

}  // namespace va
