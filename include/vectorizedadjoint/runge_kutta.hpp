// runge_kutta.hpp -- the forward sweep of the drop-in API.
//
// Reference: lib/include/runge_kutta.hpp:17-59 (three overloads) and
// lib/include/detail/runge_kutta.hpp:36-118 (fixed-step loop for plain
// steppers, controlled loop for make_controlled steppers).  Same signature and
// semantics: x0 is overwritten with y(T); the return value is the number of
// steps (accepted steps for a controlled stepper); the observer sees
// (y_n, t_n) at every stored point, in order, after the device solve; a step
// that needs more than 500 tries throws odeint::step_adjustment_error.
#pragma once

#include <cstddef>

#include "Driver.hpp"

namespace vectorizedadjoint {

template <class Stepper, class System, class State, class Time, class Observer>
size_t runge_kutta(Stepper stepper, System system, State &start_state, const State &alphas,
                   Time start_time, const Time end_time, Time dt, Driver &driver,
                   Observer observer)
{
    using info = odeint::stepper_info<Stepper>;
    const device_system sys = identify_system(system, driver.GetNin(), driver.GetNpar());
    int32_t n_acc = 0;
    const int status = driver.forward(sys, info::tableau, info::adaptive,
                                      info::abs_tol(stepper), info::rel_tol(stepper), start_state,
                                      alphas, (double)start_time, (double)end_time, (double)dt,
                                      &n_acc);
    if (status == VA_LANE_MAX_FAIL)
        throw odeint::step_adjustment_error(
            "Max number of iterations exceeded (500). A new step size was not found.");
    if (status == VA_LANE_MAX_STEPS)
        throw std::runtime_error("runge_kutta: trajectory capacity exceeded");
    const int T = driver.GetT();
    State x(start_state);
    for (int n = 0; n < T; ++n) {
        driver.GetState(x, n);
        observer(x, (Time)driver.GetTime(n));
    }
    return (size_t)n_acc;
}

// Without observer (runge_kutta.hpp:47-59).
template <class Stepper, class System, class State, class Time>
size_t runge_kutta(Stepper stepper, System system, State &start_state, const State &alphas,
                   Time start_time, const Time end_time, Time dt, Driver &driver)
{
    return runge_kutta(stepper, system, start_state, alphas, start_time, end_time, dt, driver,
                       odeint::null_observer());
}

}  // namespace vectorizedadjoint
