// device_system.hpp -- systems of the caller's own on the device (hipcc only).
//
// The reference accepts any templated system functor and tapes it with AADC
// (lib/include/AadData.hpp:124-171, recordDriverRHSFunction at
// lib/include/Driver.hpp:95-100).  Here a caller's system is a device functor
// with its right-hand side and its vector-Jacobian products, compiled in the
// caller's own hipcc translation unit together with the library's kernel
// templates:
//
//   struct Brusselator {
//       static constexpr int N = 2, P = 2;   // state size, parameters per instance
//       // f = f(x, p)
//       __host__ __device__ static void rhs(const double *x, const double *p, double *f);
//       // wJx = w^T df/dx, wJp = w^T df/dp (both overwritten)
//       __host__ __device__ static void vjp(const double *x, const double *p,
//                                           const double *w, double *wJx, double *wJp);
//       // optional: static constexpr int FLOP_RHS = ..., FLOP_VJP = ...;
//   };
//   VA_DEVICE_SYSTEM(Brusselator, my_brusselator_kernels)
//
// The macro defines `extern "C" const va_kernel_table *my_brusselator_kernels(
// int32_t tableau, int32_t adaptive)`, which returns the kernel table for one
// stepper (NULL where there is none) to pass to va_create_with_kernels with
// pb.system = VA_SYSTEM_USER.  The table holds the persistent forward and
// reverse, the plain form and, for the adaptive error steppers, the
// observation-cost kernels, all one lane per instance.  The right-hand side is
// autonomous (the reference's stage-time handling, detail/backpropagation.hpp:48,
// 127, is only exact for such systems); the reverse adds each wJp entry to the
// parameter adjoint as the reference adds wJalpha (detail/backpropagation.hpp:
// 220-223).
//
// Build: hipcc --offload-arch=gfx950 -O3 -std=c++17 -ffp-contract=off -fPIC
//        -I <repo>/include -I <repo>/vectorizedadjoint_amd/csrc -c system.hip
// and link the object into the caller's program or shared library; the
// program links libvectorizedadjoint_amd as well.
#pragma once

#include <type_traits>

#include "../vectorizedadjoint_amd.h"
#include "va_observed.hpp"

namespace va {

template <class F, class = void>
struct UserFlops {
    static constexpr int rhs = 0, vjp = 0;
};
template <class F>
struct UserFlops<F, std::void_t<decltype(F::FLOP_RHS), decltype(F::FLOP_VJP)>> {
    static constexpr int rhs = F::FLOP_RHS, vjp = F::FLOP_VJP;
};

// The kernels' system interface over a caller's device functor.
template <class F>
struct UserSystem {
    static constexpr int N = F::N, P = F::P;
    static constexpr int FLOP_RHS = UserFlops<F>::rhs, FLOP_VJP = UserFlops<F>::vjp;
    static_assert(N >= 1 && N <= 64 && P >= 1 && P <= 1024, "unsupported system size");
    struct Params {
        double p[P];
    };
    __device__ static void load(Params &q, const double *p, int64_t B, int64_t b, double)
    {
#pragma unroll
        for (int k = 0; k < P; ++k) q.p[k] = p[(int64_t)k * B + b];
    }
    VA_HD static void rhs(const Params &q, const double *x, double *f) { F::rhs(x, q.p, f); }
    VA_HD static void vjp(const Params &q, const double *x, const double *w, double *wjx,
                          double *pbar)
    {
        double wjp[P];
        F::vjp(x, q.p, w, wjx, wjp);
#pragma unroll
        for (int k = 0; k < P; ++k) pbar[k] = pbar[k] + wjp[k];
    }
};

// Every (tableau, adaptive) kernel table of one functor, built once.
template <class F>
struct UserTables {
    KernelSet set[2][6];
    bool ok[2][6];
    UserTables()
    {
        for (int ad = 0; ad < 2; ++ad)
            for (int tab = 1; tab <= 5; ++tab) {
                ok[ad][tab] = pick_tableau<UserSystem<F>>(tab, ad, &set[ad][tab]);
                if (ok[ad][tab]) pick_observed<UserSystem<F>>(tab, ad, &set[ad][tab]);
            }
    }
    static const va_kernel_table *get(int32_t tableau, int32_t adaptive)
    {
        static const UserTables t;  // thread-safe one-time construction
        if (tableau < 1 || tableau > 5 || adaptive < 0 || adaptive > 1) return nullptr;
        if (!t.ok[adaptive][tableau]) return nullptr;
        return reinterpret_cast<const va_kernel_table *>(&t.set[adaptive][tableau]);
    }
};

}  // namespace va

// Defines extern "C" const va_kernel_table *Name(int32_t tableau, int32_t adaptive)
// for the device functor Functor (at global scope).
#define VA_DEVICE_SYSTEM(Functor, Name)                                             \
    extern "C" const va_kernel_table *Name(int32_t tableau, int32_t adaptive)       \
    {                                                                               \
        return va::UserTables<Functor>::get(tableau, adaptive);                     \
    }
