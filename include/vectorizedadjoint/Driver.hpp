// Driver.hpp -- the reference's Driver handle and its set-up functions, over the
// C ABI in vectorizedadjoint_amd.h.
//
// Reference: lib/include/Driver.hpp:15-114 and lib/include/StateStorage.hpp:4-43.
// What is the same: names, signatures, shapes (lambda [Nout][Nin], mu
// [Nout][Npar] borrowed by pointer, asserted at setCostGradients), the
// accessors GetNin/GetNout/GetNpar/GetT/GetTime/GetDt/GetState, Rhs.
// What differs (the device owns the work):
//   - the trajectory lives in device memory in the library's handle; GetT /
//     GetTime / GetState / GetDt read it back on first use after a forward;
//   - GetDt(n) returns the accepted step h_n stored by the forward, not
//     t[n+1]-t[n] (they agree to an ulp);
//   - recordDriverRHSFunction binds the functor to a device system
//     (systems.hpp) instead of taping it; Rhs evaluates the functor on the host;
//   - PushBackTime / PushBackState are not provided: the store is written by
//     the device forward only.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cassert>
#include <cstdint>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../vectorizedadjoint_amd.h"
#include "odeint.hpp"
#include "systems.hpp"

namespace vectorizedadjoint {

namespace detail_driver {

inline void check(va_status s, const char *what)
{
    if (s != VA_OK) throw std::runtime_error(std::string(what) + ": " + va_last_error());
}

inline void hip_check(hipError_t e, const char *what)
{
    if (e != hipSuccess)
        throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// One device array of T; grows on demand, never shrinks.
template <class T>
struct device_array {
    T *ptr = nullptr;
    size_t n = 0;
    device_array() = default;
    device_array(const device_array &) = delete;
    device_array &operator=(const device_array &) = delete;
    ~device_array()
    {
        if (ptr) (void)hipFree(ptr);
    }
    T *reserve(size_t count)
    {
        if (count > n) {
            if (ptr) hip_check(hipFree(ptr), "hipFree");
            ptr = nullptr;
            n = 0;
            hip_check(hipMalloc((void **)&ptr, sizeof(T) * (count ? count : 1)), "hipMalloc");
            n = count;
        }
        return ptr;
    }
    template <class V>
    T *upload(const V &v)
    {
        reserve(v.size());
        std::vector<T> tmp(v.begin(), v.end());
        if (!tmp.empty())
            hip_check(hipMemcpy(ptr, tmp.data(), sizeof(T) * tmp.size(), hipMemcpyHostToDevice),
                      "hipMemcpy");
        return ptr;
    }
    void download(T *dst, size_t count) const
    {
        if (count)
            hip_check(hipMemcpy(dst, ptr, sizeof(T) * count, hipMemcpyDeviceToHost), "hipMemcpy");
    }
};

}  // namespace detail_driver

class Driver
{
  public:
    // Borrowed cost gradients (Driver.hpp:22-23): lambda [Nout][Nin], mu [Nout][Npar].
    std::vector<std::vector<double>> *p_lambda = nullptr;
    std::vector<std::vector<double>> *p_mu = nullptr;

    Driver(int Nin, int Nout, int Npar) : m_Nin(Nin), m_Nout(Nout), m_Npar(Npar)
    {
        if (Nin < 1 || Nout < 1 || Npar < 0)
            throw std::invalid_argument("Driver: Nin, Nout >= 1 and Npar >= 0 required");
    }

    // Driver(stepper, system, Nin, Nout, Npar) (Driver.hpp:31-37): records the
    // tableau and the system up front.
    template <class Stepper, class System>
    Driver(Stepper, System system, int Nin, int Nout, int Npar) : Driver(Nin, Nout, Npar)
    {
        m_tableau = odeint::stepper_info<Stepper>::tableau;
        bind_rhs(system);
    }

    Driver(const Driver &) = delete;
    Driver &operator=(const Driver &) = delete;

    ~Driver()
    {
        if (m_handle) va_destroy(m_handle);
    }

    int GetNin() const { return m_Nin; }
    int GetNout() const { return m_Nout; }
    int GetNpar() const { return m_Npar; }

    // Number of stored points (accepted steps + 1), StateStorage.hpp:20.
    int GetT() const
    {
        fetch_rows();
        return (int)m_rows_count;
    }
    double GetTime(int n) const
    {
        fetch_rows();
        return m_rows.at((size_t)n * row_width());
    }
    // Accepted step taken from point n (StateStorage.hpp:22 recomputes it as t[n+1]-t[n]).
    double GetDt(int n) const
    {
        fetch_rows();
        return m_rows.at((size_t)n * row_width() + 1);
    }
    template <typename State>
    void GetState(State &u, int n) const
    {
        fetch_rows();
        const double *row = &m_rows.at((size_t)n * row_width());
        for (size_t i = 0; i < (size_t)u.size() && i < (size_t)m_Nin; ++i) u[i] = row[2 + i];
    }

    // Driver::Rhs (Driver.hpp:69-78): f(u, alphas, t) of the recorded system.
    template <typename State, typename Time>
    void Rhs(const State &u, State &dudt, const State &alphas, const Time &time)
    {
        if (!m_rhs) throw std::runtime_error("Must call recordDriverRHSFunction() first!");
        std::vector<double> x(u.begin(), u.end()), f(u.size()), p(alphas.begin(), alphas.end());
        m_rhs(x, f, p, (double)time);
        for (size_t i = 0; i < (size_t)dudt.size() && i < f.size(); ++i) dudt[i] = f[i];
    }

    // ---- state of the drop-in (not part of the reference's interface) ----
    int tableau() const { return m_tableau; }
    bool has_rhs() const { return (bool)m_rhs; }
    const device_system &recorded_system() const { return m_recorded; }
    va_handle *handle() const { return m_handle; }
    const device_system &forward_system() const { return m_fwd_system; }
    int forward_tableau() const { return m_fwd_tableau; }

    template <class System>
    void bind_rhs(System system)
    {
        m_recorded = identify_system(system, m_Nin, m_Npar);
        m_rhs = [system](const std::vector<double> &x, std::vector<double> &f,
                         const std::vector<double> &p, double t) mutable { system(x, f, p, t); };
    }

    void set_tableau(int tableau) { m_tableau = tableau; }

    // Runs one forward solve of a single instance on the device; grows the
    // trajectory capacity and solves again when the stored steps do not fit.
    // Returns the lane status; x0 receives y(T).
    template <class State>
    int forward(const device_system &sys, int tableau, bool adaptive, double abs_tol,
                double rel_tol, State &x0, const State &alphas, double t0, double t1, double dt,
                int32_t *n_accepted)
    {
        if ((int)x0.size() != m_Nin || (int)alphas.size() != m_Npar)
            throw std::invalid_argument("runge_kutta: state or parameter size differs from the "
                                        "Driver's Nin / Npar");
        int32_t cap = adaptive ? 4096 : fixed_step_count(t0, t1, dt) + 2;
        const int64_t max_rows_total = (int64_t)1 << 26;
        for (;;) {
            ensure_handle(sys, tableau, adaptive, abs_tol, rel_tol, t0, t1, dt, cap);
            double *y = m_y.upload(x0);
            const double *p = m_p.upload(alphas);
            int32_t *counts = m_counts.reserve(3);
            detail_driver::check(va_forward(m_handle, y, p, counts, counts + 1, counts + 2,
                                            nullptr),
                                 "va_forward");
            int32_t host[3] = {0, 0, 0};
            m_counts.download(host, 3);
            m_rows_valid = false;
            m_forward_done = true;
            if (host[2] == VA_LANE_MAX_STEPS && (int64_t)cap * 4 <= max_rows_total) {
                cap *= 4;
                continue;
            }
            std::vector<double> yT(m_Nin);
            m_y.download(yT.data(), (size_t)m_Nin);
            for (int i = 0; i < m_Nin; ++i) x0[i] = yT[i];
            *n_accepted = host[0];
            return host[2];
        }
    }

    bool forward_done() const { return m_forward_done; }

    // Device scratch for the reverse (lambda, mu, parameters, sensitivity).
    detail_driver::device_array<double> m_lam, m_mu, m_par, m_sens;

  private:
    static int32_t fixed_step_count(double t0, double t1, double dt)
    {
        if (!(dt != 0.0) || !((t1 - t0) / dt >= 0.0))
            throw std::invalid_argument("runge_kutta: dt must be non-zero and point from t0 to t1");
        const double n = (t1 - t0) / dt;
        if (n > 1e9) throw std::invalid_argument("runge_kutta: too many fixed steps");
        return (int32_t)n;
    }

    void ensure_handle(const device_system &sys, int tableau, bool adaptive, double abs_tol,
                       double rel_tol, double t0, double t1, double dt, int32_t cap)
    {
        va_problem pb;
        va_problem_init(&pb);
        pb.system = sys.id;
        pb.system_constant = sys.constant;
        pb.tableau = tableau;
        pb.adaptive = adaptive ? 1 : 0;
        pb.n_state = m_Nin;
        pb.n_param = m_Npar;
        pb.batch = 1;
        pb.t0 = t0;
        pb.t1 = t1;
        pb.dt0 = dt;
        if (adaptive) {
            pb.abs_tol = abs_tol;
            pb.rel_tol = rel_tol;
        }
        pb.max_steps = cap;
        if (m_handle && same_problem(pb, m_pb) && sys.kernels == m_fwd_system.kernels) return;
        if (m_handle) va_destroy(m_handle);
        m_handle = nullptr;
        if (sys.id == VA_SYSTEM_USER) {
            const va_kernel_table *k = sys.kernels ? sys.kernels(tableau, adaptive ? 1 : 0)
                                                   : nullptr;
            if (!k)
                throw std::invalid_argument("runge_kutta: the bound device system has no "
                                            "kernels for this stepper");
            detail_driver::check(va_create_with_kernels(&pb, k, &m_handle),
                                 "va_create_with_kernels");
        } else {
            detail_driver::check(va_create(&pb, &m_handle), "va_create");
        }
        m_pb = pb;
        m_fwd_system = sys;
        m_fwd_tableau = tableau;
    }

    static bool same_problem(const va_problem &a, const va_problem &b)
    {
        return a.system == b.system && a.system_constant == b.system_constant &&
               a.tableau == b.tableau && a.adaptive == b.adaptive && a.t0 == b.t0 &&
               a.t1 == b.t1 && a.dt0 == b.dt0 && a.abs_tol == b.abs_tol &&
               a.rel_tol == b.rel_tol && a.max_steps == b.max_steps;
    }

    size_t row_width() const { return (size_t)m_Nin + 2; }

    void fetch_rows() const
    {
        if (m_rows_valid) return;
        m_rows.clear();
        m_rows_count = 0;
        if (m_handle && m_forward_done) {
            int64_t n = 0;
            detail_driver::check(va_trajectory(m_handle, 0, nullptr, 0, &n, nullptr),
                                 "va_trajectory");
            m_rows.resize((size_t)n * row_width());
            detail_driver::check(va_trajectory(m_handle, 0, m_rows.data(), n, &n, nullptr),
                                 "va_trajectory");
            m_rows_count = n;
        }
        m_rows_valid = true;
    }

    int m_Nin, m_Nout, m_Npar;
    int m_tableau = 0;  // constructDriverButcherTableau
    device_system m_recorded;  // recordDriverRHSFunction
    std::function<void(const std::vector<double> &, std::vector<double> &,
                       const std::vector<double> &, double)>
        m_rhs;

    va_handle *m_handle = nullptr;
    va_problem m_pb{};
    device_system m_fwd_system;  // system and tableau the stored trajectory was made with
    int m_fwd_tableau = 0;
    bool m_forward_done = false;
    detail_driver::device_array<double> m_y, m_p;
    detail_driver::device_array<int32_t> m_counts;
    mutable std::vector<double> m_rows;
    mutable int64_t m_rows_count = 0;
    mutable bool m_rows_valid = false;
};

// delete_driver_handle (Driver.hpp:81-85).
inline void delete_driver_handle(void *ptr)
{
    delete static_cast<Driver *>(ptr);
}

// constructDriverButcherTableau (Driver.hpp:87-93): the tableau the reverse
// sweep uses.  It must be the forward stepper's tableau (the device reverse
// recomputes the forward's stages); a plain stepper or its controlled form
// are both accepted.
template <class Stepper>
void constructDriverButcherTableau(Driver &driver, Stepper)
{
    driver.set_tableau(odeint::stepper_info<Stepper>::tableau);
}

// recordDriverRHSFunction (Driver.hpp:95-100).
template <typename System>
void recordDriverRHSFunction(Driver &driver, System system)
{
    driver.bind_rhs(system);
}

// setCostGradients (Driver.hpp:103-114): lambda [Nout][Nin] (overwritten with
// dJ/dy0 by adjointSolve), mu [Nout][Npar] (dJ/dp added by adjointSolve).
inline void setCostGradients(Driver &driver, std::vector<std::vector<double>> &lambda,
                             std::vector<std::vector<double>> &mu)
{
    assert((int)lambda.size() == driver.GetNout());
    assert((int)lambda[0].size() == driver.GetNin());
    assert((int)mu.size() == driver.GetNout());
    assert((int)mu[0].size() == driver.GetNpar());
    driver.p_lambda = &lambda;
    driver.p_mu = &mu;
}

}  // namespace vectorizedadjoint
