// va_inst_glv_rows_a.hip -- row-per-lane kernel instantiations for generalized Lotka-Volterra,
// N = 2..9 (examples/GeneralizedLotkaVolterra/main.cpp:125-143 takes N
// from the command line).  One translation unit per range so they compile in
// parallel.
#include "va_glv_rows.hpp"

namespace va {

bool kernels_lotka_volterra_rows_a(int n, int tableau, int adaptive, KernelSet *out)
{
    switch (n) {
    case 2: return pick_glv_rows_core<2>(tableau, adaptive, out);
    case 3: return pick_glv_rows_core<3>(tableau, adaptive, out);
    case 4: return pick_glv_rows_core<4>(tableau, adaptive, out);
    case 5: return pick_glv_rows_core<5>(tableau, adaptive, out);
    case 6: return pick_glv_rows_core<6>(tableau, adaptive, out);
    case 7: return pick_glv_rows_core<7>(tableau, adaptive, out);
    case 8: return pick_glv_rows_core<8>(tableau, adaptive, out);
    case 9: return pick_glv_rows_core<9>(tableau, adaptive, out);
    default: return false;
    }
}

}  // namespace va
