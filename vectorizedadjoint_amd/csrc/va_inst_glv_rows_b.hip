// va_inst_glv_rows_b.hip -- row-per-lane kernel instantiations for generalized Lotka-Volterra,
// N = 11..20 (examples/GeneralizedLotkaVolterra/main.cpp:125-143 takes N
// from the command line).  One translation unit per range so they compile in
// parallel.
#include "va_glv_rows.hpp"

namespace va {

bool kernels_lotka_volterra_rows_b(int n, int tableau, int adaptive, KernelSet *out)
{
    switch (n) {
    case 11: return pick_glv_rows_core<11>(tableau, adaptive, out);
    case 12: return pick_glv_rows_core<12>(tableau, adaptive, out);
    case 13: return pick_glv_rows_core<13>(tableau, adaptive, out);
    case 14: return pick_glv_rows_core<14>(tableau, adaptive, out);
    case 15: return pick_glv_rows_core<15>(tableau, adaptive, out);
    case 16: return pick_glv_rows_core<16>(tableau, adaptive, out);
    case 17: return pick_glv_rows_core<17>(tableau, adaptive, out);
    case 18: return pick_glv_rows_core<18>(tableau, adaptive, out);
    case 19: return pick_glv_rows_core<19>(tableau, adaptive, out);
    case 20: return pick_glv_rows_core<20>(tableau, adaptive, out);
    default: return false;
    }
}

}  // namespace va
