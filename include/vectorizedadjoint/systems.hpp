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
//   - or by identification: the functor is evaluated on the host at a few
//     fixed probe points and compared with each built-in right-hand side
//     (the oscillator's constant k is recovered from the probes).  A functor
//     that matches none raises std::invalid_argument.
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
};

template <class System>
struct system_binding {
    static constexpr bool bound = false;
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

namespace vectorizedadjoint {
namespace detail_systems {

inline double probe_value(int k, int i, double lo, double hi)
{
    // fixed, irrational-looking probe coordinates (no RNG state)
    double u = std::fmod(0.6180339887498949 * (k * 131 + i * 17 + 7), 1.0);
    return lo + (hi - lo) * u;
}

inline bool close(double a, double b, double scale)
{
    return std::fabs(a - b) <= 1e-12 * (scale + std::fabs(b));
}

template <class System>
bool eval(System &sys, const std::vector<double> &x, std::vector<double> &f,
          const std::vector<double> &p)
{
    f.assign(x.size(), 0.0);
    sys(x, f, p, 0.0);
    return true;
}

}  // namespace detail_systems

// Identify the device system that the functor implements for (Nin, Npar).
template <class System>
device_system identify_system(System &sys, int n, int np)
{
    using namespace detail_systems;
    if constexpr (system_binding<System>::bound) {
        return device_system{system_binding<System>::id, system_binding<System>::constant};
    } else {
        const int kProbes = 4;
        std::vector<double> x(n), f(n), p(np);
        bool vdp = (n == 2 && np == 1), osc = (n == 2 && np == 2), osck = (n == 2 && np == 1);
        bool glv = (np == n + n * n);
        double kconst = 0.0;
        for (int k = 0; k < kProbes; ++k) {
            for (int i = 0; i < n; ++i) x[i] = probe_value(k, i, -1.3, 1.7);
            for (int j = 0; j < np; ++j) p[j] = probe_value(k, n + j, 0.4, 2.2);
            eval(sys, x, f, p);
            if (n == 2) {
                const double x0 = x[0], x1 = x[1];
                if (vdp)
                    vdp = close(f[0], x1, 1.0) &&
                          close(f[1], p[0] * ((1.0 - x0 * x0) * x1 - x0), 1.0);
                if (osc) osc = close(f[0], x1, 1.0) && close(f[1], -p[0] * x0 - p[1] * x1, 1.0);
                if (osck) {
                    const double kk = -(f[1] + p[0] * x1) / x0;
                    if (k == 0) kconst = kk;
                    osck = close(f[0], x1, 1.0) && close(f[1], -kconst * x0 - p[0] * x1, 1.0);
                }
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
