// The Brusselator, a system that is not built into the library, written once as
// a caller writes a device functor (right-hand side and vector-Jacobian
// products).  va_user_system.hip compiles it for the device (VA_DEVICE_SYSTEM);
// brusselator_host.cpp compiles the same source with the host compiler alone as
// C callbacks for the CPU oracle.
//
//   x0' = a + x0^2 x1 - (b + 1) x0
//   x1' = b x0 - x0^2 x1,            p = (a, b)
#pragma once

#ifdef __HIPCC__
#define BRUSSELATOR_HD __host__ __device__
#else
#define BRUSSELATOR_HD
#endif

struct Brusselator {
    static constexpr int N = 2, P = 2;
    static constexpr int FLOP_RHS = 7, FLOP_VJP = 11;
    BRUSSELATOR_HD static void rhs(const double *x, const double *p, double *f)
    {
        const double x0 = x[0], x1 = x[1], a = p[0], b = p[1];
        const double u = (x0 * x0) * x1;
        f[0] = (a + u) - (b + 1.0) * x0;
        f[1] = b * x0 - u;
    }
    BRUSSELATOR_HD static void vjp(const double *x, const double *p, const double *w,
                                   double *wjx, double *wjp)
    {
        const double x0 = x[0], x1 = x[1], b = p[1];
        const double sq = x0 * x0;
        const double v = (2.0 * x0) * x1;
        wjx[0] = w[0] * (v - (b + 1.0)) + w[1] * (b - v);
        wjx[1] = (w[0] - w[1]) * sq;
        wjp[0] = w[0];
        wjp[1] = (w[1] - w[0]) * x0;
    }
};
