// va_inst_glv_rows_c.hip -- row-per-lane kernel instantiations for generalized Lotka-Volterra,
// N = 21..32 (examples/GeneralizedLotkaVolterra/main.cpp:125-143 takes N
// from the command line).  One translation unit per range so they compile in
// parallel.
#include "va_glv_rows.hpp"

namespace va {

bool kernels_lotka_volterra_rows_c(int n, int tableau, int adaptive, KernelSet *out)
{
    switch (n) {
    case 21: return pick_glv_rows_core<21>(tableau, adaptive, out);
    case 22: return pick_glv_rows_core<22>(tableau, adaptive, out);
    case 23: return pick_glv_rows_core<23>(tableau, adaptive, out);
    case 24: return pick_glv_rows_core<24>(tableau, adaptive, out);
    case 25: return pick_glv_rows_core<25>(tableau, adaptive, out);
    case 26: return pick_glv_rows_core<26>(tableau, adaptive, out);
    case 27: return pick_glv_rows_core<27>(tableau, adaptive, out);
    case 28: return pick_glv_rows_core<28>(tableau, adaptive, out);
    case 29: return pick_glv_rows_core<29>(tableau, adaptive, out);
    case 30: return pick_glv_rows_core<30>(tableau, adaptive, out);
    case 31: return pick_glv_rows_core<31>(tableau, adaptive, out);
    case 32: return pick_glv_rows_core<32>(tableau, adaptive, out);
    default: return false;
    }
}

}  // namespace va
