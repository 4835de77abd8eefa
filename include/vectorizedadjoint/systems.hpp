// systems.hpp -- binding a user system functor to a device system.
//
// The reference records the user's templated functor with AADC
// (lib/include/AadData.hpp:124-171, recordDriverRHSFunction at
// lib/include/Driver.hpp:95-100); the system concept is
//   template <class T> void operator()(const std::vector<T> &x, std::vector<T> &dxdt,
//                                      const std::vector<T> &p, T t)
// (doc/source/harmonicOscillator.rst:61-85).  Device code cannot run a host
// functor, so a functor is bound to one of the built-in device systems, which
// carry hand-written vector-Jacobian products:
//   - explicitly:  VA_BIND_SYSTEM(MyFunctor, VA_SYSTEM_VAN_DER_POL, 0.0)
//   - to a device functor of the caller's own (any system): compile the device
//     functor with VA_DEVICE_SYSTEM(DeviceFunctor, my_kernels) in a hipcc
//     translation unit (vectorizedadjoint/device_system.hpp) and bind the host
//     functor to it with VA_BIND_DEVICE_SYSTEM(MyFunctor, my_kernels)
//   - or by identification: the functor is evaluated on the host at 24 fixed
//     probe points spread over a wide box of states and parameters, each at its
//     own time (zero and non-zero, positive and negative), and compared with
//     each built-in right-hand side (the oscillator's constant k is recovered
//     from the probes).  The built-in systems are autonomous, so a functor that
//     depends on t, or that changes form across the box (a piecewise
//     right-hand side), matches none and raises std::invalid_argument.
#pragma once

#include <cmath>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../vectorizedadjoint_amd.h"

namespace vectorizedadjoint {

struct device_system {
    int id = 0;
    double constant = 0.0;  // va_problem.system_constant
    // VA_SYSTEM_USER: the kernel table of the caller's device functor per stepper
    const va_kernel_table *(*kernels)(int32_t tableau, int32_t adaptive) = nullptr;
};

template <class System>
struct system_binding {
    static constexpr bool bound = false;
};

template <class System, class = void>
struct binding_kernels {
    static constexpr const va_kernel_table *(*get)(int32_t, int32_t) = nullptr;
};
template <class System>
struct binding_kernels<System, std::void_t<decltype(&system_binding<System>::kernels)>> {
    static constexpr const va_kernel_table *(*get)(int32_t, int32_t) =
        &system_binding<System>::kernels;
};

}  // namespace vectorizedadjoint

// Explicit binding of a functor type to a device system (at global scope).
#define VA_BIND_SYSTEM(Type, SystemId, Constant)                              \
    template <>                                                               \
    struct vectorizedadjoint::system_binding<Type> {                          \
        static constexpr bool bound = true;                                   \
        static constexpr int id = (SystemId);                                 \
        static constexpr double constant = (Constant);                        \
    };

// Binding of a functor type to the kernels of a device functor of the caller's
// (VA_DEVICE_SYSTEM(DeviceFunctor, TableFn) in a hipcc translation unit that is
// linked into the program), at global scope.
#define VA_BIND_DEVICE_SYSTEM(Type, TableFn)                                  \
    extern "C" const va_kernel_table *TableFn(int32_t, int32_t);              \
    template <>                                                               \
    struct vectorizedadjoint::system_binding<Type> {                          \
        static constexpr bool bound = true;                                   \
        static constexpr int id = VA_SYSTEM_USER;                             \
        static constexpr double constant = 0.0;                               \
        static const va_kernel_table *kernels(int32_t tab, int32_t ad)        \
        {                                                                     \
            return TableFn(tab, ad);                                          \
        }                                                                     \
    };

namespace vectorizedadjoint {
namespace detail_systems {

inline double probe_value(int k, int i, double lo, double hi)
{
    // fixed, irrational-looking probe coordinates (no RNG state)
    double u = std::fmod(0.6180339887498949 * (k * 131 + i * 17 + 7), 1.0);
    return lo + (hi - lo) * u;
}

// probe k's time: 0 for the first probe, then spread over [-7, 13]
inline double probe_time(int k)
{
    return k == 0 ? 0.0 : -7.0 + 20.0 * std::fmod(0.7548776662466927 * (k * 37 + 3), 1.0);
}

inline bool close(double a, double b, double scale)
{
    return std::fabs(a - b) <= 1e-12 * (scale + std::fabs(b));
}

template <class System>
bool eval(System &sys, const std::vector<double> &x, std::vector<double> &f,
          const std::vector<double> &p, double t)
{
    f.assign(x.size(), 0.0);
    sys(x, f, p, t);
    return true;
}

}  // namespace detail_systems

// Identify the device system that the functor implements for (Nin, Npar).
template <class System>
device_system identify_system(System &sys, int n, int np)
{
    using namespace detail_systems;
    if constexpr (system_binding<System>::bound) {
        return device_system{system_binding<System>::id, system_binding<System>::constant,
                             binding_kernels<System>::get};
    } else {
        const int kProbes = 24;
        std::vector<double> x(n), f(n), p(np);
        bool vdp = (n == 2 && np == 1), osc = (n == 2 && np == 2), osck = (n == 2 && np == 1);
        bool glv = (np == n + n * n);
        double kconst = 0.0;
        if (osck) {
            // at x = (1, 0) the functor returns f1 = -k * 1 - mu * 0 = -k exactly
            x = {1.0, 0.0};
            p = {0.75};
            eval(sys, x, f, p, 0.0);
            kconst = -f[1];
        }
        for (int k = 0; k < kProbes; ++k) {
            for (int i = 0; i < n; ++i) x[i] = probe_value(k, i, -4.3, 5.1);
            for (int j = 0; j < np; ++j) p[j] = probe_value(k, n + j, -2.9, 3.7);
            eval(sys, x, f, p, probe_time(k));
            if (n == 2) {
                const double x0 = x[0], x1 = x[1];
                if (vdp)
                    vdp = close(f[0], x1, 1.0) &&
                          close(f[1], p[0] * ((1.0 - x0 * x0) * x1 - x0), 1.0);
                if (osc) osc = close(f[0], x1, 1.0) && close(f[1], -p[0] * x0 - p[1] * x1, 1.0);
                if (osck)
                    osck = close(f[0], x1, 1.0) && close(f[1], -kconst * x0 - p[0] * x1, 1.0);
            }
            if (glv) {
                for (int i = 0; i < n && glv; ++i) {
                    double g = p[i];
                    for (int j = 0; j < n; ++j) g += p[n * (i + 1) + j] * x[j];
                    glv = close(f[i], x[i] * g, 1.0 + std::fabs(x[i]) * n);
                }
            }
        }
        if (vdp) return device_system{VA_SYSTEM_VAN_DER_POL, 0.0};
        if (osc) return device_system{VA_SYSTEM_OSCILLATOR, 0.0};
        if (osck) return device_system{VA_SYSTEM_OSCILLATOR_FIXED_K, kconst};
        if (glv) return device_system{VA_SYSTEM_LOTKA_VOLTERRA, 0.0};
        throw std::invalid_argument(
            "vectorizedadjoint: the system functor matches none of the device systems "
            "(Van der Pol, oscillator, oscillator with constant k, generalized Lotka-Volterra) "
            "for Nin=" + std::to_string(n) + ", Npar=" + std::to_string(np) +
            "; bind it with VA_BIND_SYSTEM if it is one of them");
    }
}

}  // namespace vectorizedadjoint
