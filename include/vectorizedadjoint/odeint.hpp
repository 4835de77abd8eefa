// odeint.hpp -- stepper types of the drop-in API.
//
// The reference selects its forward stepper with Boost.odeint types
// (examples/VanDerPol/main.cpp:49,76, examples/HarmonicOscillator/main.cpp:160,
// examples/GeneralizedLotkaVolterra/main.cpp:95,178).  Here those names are
// plain tags that select a device tableau; nothing of Boost is used.
// `namespace odeint` inside vectorizedadjoint lets the examples' spelling
// `odeint::runge_kutta_fehlberg78<std::vector<double>>` resolve after
// `using namespace vectorizedadjoint;`.
#pragma once

#include <stdexcept>
#include <string>

#include "../vectorizedadjoint_amd.h"

namespace vectorizedadjoint {
namespace odeint {

template <class State>
struct euler {
    using state_type = State;
    static constexpr int tableau = VA_TABLEAU_EULER;
    static constexpr bool has_error_estimate = false;
};

template <class State>
struct runge_kutta4 {
    using state_type = State;
    static constexpr int tableau = VA_TABLEAU_RK4;
    static constexpr bool has_error_estimate = false;
};

template <class State>
struct runge_kutta4_classic : runge_kutta4<State> {};

template <class State>
struct runge_kutta_cash_karp54 {
    using state_type = State;
    static constexpr int tableau = VA_TABLEAU_CASH_KARP54;
    static constexpr bool has_error_estimate = true;
};

// Not supported by the reference's adjoint (ButcherTable.hpp:247-250); supported here.
template <class State>
struct runge_kutta_dopri5 {
    using state_type = State;
    static constexpr int tableau = VA_TABLEAU_DOPRI5;
    static constexpr bool has_error_estimate = true;
};

template <class State>
struct runge_kutta_fehlberg78 {
    using state_type = State;
    static constexpr int tableau = VA_TABLEAU_FEHLBERG78;
    static constexpr bool has_error_estimate = true;
};

// make_controlled<Stepper>(abs, rel): the adaptive (controlled) stepper of
// Boost.odeint 1.74 with default_error_checker / default_step_adjuster.
template <class ErrorStepper>
struct controlled_runge_kutta {
    using stepper_type = ErrorStepper;
    using state_type = typename ErrorStepper::state_type;
    double abs_tol, rel_tol;
    static_assert(ErrorStepper::has_error_estimate,
                  "make_controlled needs an error stepper (Cash-Karp 5(4), DOPRI5, Fehlberg 7(8))");
};

template <class ErrorStepper>
controlled_runge_kutta<ErrorStepper> make_controlled(double abs_tol, double rel_tol,
                                                     ErrorStepper = ErrorStepper())
{
    return controlled_runge_kutta<ErrorStepper>{abs_tol, rel_tol};
}

// odeint's null_observer
struct null_observer {
    template <class State, class Time>
    void operator()(const State &, Time) const
    {
    }
};

// odeint's step_adjustment_error (failed_step_checker: more than 500 tries in
// one step); thrown by runge_kutta like the reference.
struct step_adjustment_error : std::runtime_error {
    explicit step_adjustment_error(const std::string &m) : std::runtime_error(m) {}
};

template <class S>
struct stepper_info {
    static constexpr int tableau = S::tableau;
    static constexpr bool adaptive = false;
    static double abs_tol(const S &) { return 0.0; }
    static double rel_tol(const S &) { return 0.0; }
};

template <class E>
struct stepper_info<controlled_runge_kutta<E>> {
    static constexpr int tableau = E::tableau;
    static constexpr bool adaptive = true;
    static double abs_tol(const controlled_runge_kutta<E> &c) { return c.abs_tol; }
    static double rel_tol(const controlled_runge_kutta<E> &c) { return c.rel_tol; }
};

}  // namespace odeint
}  // namespace vectorizedadjoint
