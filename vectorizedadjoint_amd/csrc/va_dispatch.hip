// va_dispatch.hip -- non-template helper kernels of the plain form and the
// (system, N, tableau, adaptive) -> kernel-set dispatch.  The template
// instantiations live in va_inst_*.hip so they compile in parallel.
#include "va_kernels_impl.hpp"

namespace va {

bool kernels_van_der_pol(int tableau, int adaptive, KernelSet *out);
bool kernels_oscillator(int tableau, int adaptive, KernelSet *out);
bool kernels_oscillator_fixed_k(int tableau, int adaptive, KernelSet *out);
bool kernels_lotka_volterra_small(int n, int tableau, int adaptive, KernelSet *out);
bool kernels_lotka_volterra_10(int tableau, int adaptive, int instances_per_wave,
                               KernelSet *out);
bool kernels_lotka_volterra_shared(int n, int tableau, int adaptive, KernelSet *out);
bool kernels_lotka_volterra_rows_a(int n, int tableau, int adaptive, KernelSet *out);
bool kernels_lotka_volterra_rows_b(int n, int tableau, int adaptive, KernelSet *out);
bool kernels_lotka_volterra_rows_c(int n, int tableau, int adaptive, KernelSet *out);

__global__ void k_plain_finish(double *lambda_out, const double *lambda_in, double *mu,
                               const double *pbar, const int32_t *status, int64_t B, int n,
                               int np)
{
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const bool ok = status[b] == LANE_OK;
    const double qnan = __builtin_nan("");
    for (int i = 0; i < n; ++i)
        lambda_out[(int64_t)i * B + b] = ok ? lambda_in[(int64_t)i * B + b] : qnan;
    for (int k = 0; k < np; ++k) {
        const int64_t ix = (int64_t)k * B + b;
        mu[ix] = ok ? mu[ix] + pbar[ix] : qnan;
    }
}

__global__ void k_plain_cost(int cost, const double *w, const double *yT, double *J,
                             double *lambda, const int32_t *status, int64_t B, int n)
{
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    double Jv = 0.0;
    for (int i = 0; i < n; ++i) {
        const double yi = yT[(int64_t)i * B + b];
        double li;
        if (cost == COST_SUM) {
            Jv = Jv + yi;
            li = 1.0;
        } else if (cost == COST_SQUARED_NORM) {
            Jv = fma(yi, yi, Jv);
            li = 2.0 * yi;
        } else {
            Jv = fma(w[i], yi, Jv);
            li = w[i];
        }
        lambda[(int64_t)i * B + b] = li;
    }
    J[b] = status[b] == LANE_OK ? Jv : __builtin_nan("");
}

hipError_t launch_plain_finish(double *lambda_out, const double *lambda_in, double *mu,
                               const double *pbar, const int32_t *status, int64_t B, int n, int np,
                               int grid, int block, hipStream_t s)
{
    hipLaunchKernelGGL(k_plain_finish, dim3(grid), dim3(block), 0, s, lambda_out, lambda_in, mu,
                       pbar, status, B, n, np);
    return hipGetLastError();
}

hipError_t launch_plain_cost(int cost, const double *cost_w, const double *yT, double *J,
                             double *lambda, const int32_t *status, int64_t B, int n, int grid,
                             int block, hipStream_t s)
{
    hipLaunchKernelGGL(k_plain_cost, dim3(grid), dim3(block), 0, s, cost, cost_w, yT, J, lambda,
                       status, B, n);
    return hipGetLastError();
}

enum { SYS_VDP = 1, SYS_OSC = 2, SYS_OSC_K = 3, SYS_GLV = 4, SYS_GLV_SHARED = 5 };

bool get_kernels(int system, int n, int tableau, int adaptive, int instances_per_wave,
                 KernelSet *out)
{
    const bool one_lane = instances_per_wave == 0 || instances_per_wave == 64;
    switch (system) {
    case SYS_VDP: return n == 2 && one_lane && kernels_van_der_pol(tableau, adaptive, out);
    case SYS_OSC: return n == 2 && one_lane && kernels_oscillator(tableau, adaptive, out);
    case SYS_OSC_K:
        return n == 2 && one_lane && kernels_oscillator_fixed_k(tableau, adaptive, out);
    case SYS_GLV:
        if (n == 10) return kernels_lotka_volterra_10(tableau, adaptive, instances_per_wave, out);
        // one lane per instance (instances_per_wave 64): N = 2..5; one group of N
        // lanes per instance (0 or 64/N, the default): N = 2..32
        if (instances_per_wave == 64) return kernels_lotka_volterra_small(n, tableau, adaptive, out);
        if (n < 2 || n > 32 || (instances_per_wave != 0 && instances_per_wave != 64 / n))
            return false;
        if (n <= 9) return kernels_lotka_volterra_rows_a(n, tableau, adaptive, out);
        if (n <= 20) return kernels_lotka_volterra_rows_b(n, tableau, adaptive, out);
        return kernels_lotka_volterra_rows_c(n, tableau, adaptive, out);
    case SYS_GLV_SHARED: return kernels_lotka_volterra_shared(n, tableau, adaptive, out);
    default: return false;
    }
}

}  // namespace va
