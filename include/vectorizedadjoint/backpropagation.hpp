// backpropagation.hpp -- the reverse sweep of the drop-in API.
//
// Reference: lib/include/backpropagation.hpp:18-87 (public adjointSolve,
// computeSensitivityMatrix[NoSIMD]) over lib/include/detail/backpropagation.hpp:
// 280-434.  Same contract: precondition failures are printed to stdout and
// swallowed; lambda [Nout][Nin] is overwritten with dJ/dy0 and dJ/dp is added
// to mu [Nout][Npar]; the sensitivity matrix is [Nin][Npar] with entry (i, k) =
// d y_i(T) / d p_k.  All Nout seeds go through one device reverse sweep.
#pragma once

#include <iostream>
#include <stdexcept>
#include <vector>

#include "Driver.hpp"

namespace vectorizedadjoint {

namespace detail_backprop {

// The reference's three precondition checks (backpropagation.hpp:23-37) plus
// the ones the device path adds: a forward must exist and have been made with
// the recorded system and tableau.
inline void check_setup(const Driver &driver, bool need_gradients)
{
    if (need_gradients && !(driver.p_lambda && driver.p_mu))
        throw std::runtime_error("Must call setCostGradients() first!");
    if (!driver.tableau())
        throw std::runtime_error("Must call constructDriverButcherTableau() to set Butcher Tableau!");
    if (!driver.has_rhs())
        throw std::runtime_error("Must call recordDriverRHSFunction() to record the RHS with "
                                 "automatic differentiation!");
    if (!driver.forward_done() || !driver.handle())
        throw std::runtime_error("Must call runge_kutta() before the reverse pass!");
    if (driver.tableau() != driver.forward_tableau())
        throw std::runtime_error("The Butcher tableau differs from the forward stepper's: the "
                                 "reverse pass recomputes the forward stages");
    if (driver.recorded_system().id != driver.forward_system().id ||
        driver.recorded_system().constant != driver.forward_system().constant)
        throw std::runtime_error("The recorded RHS differs from the system of the forward pass");
}

template <class State>
void adjoint_solve(Driver &driver, const State &parameters)
{
    const int Nout = driver.GetNout(), N = driver.GetNin(), P = driver.GetNpar();
    if ((int)parameters.size() != P)
        throw std::runtime_error("adjointSolve: parameter vector size differs from Npar");
    std::vector<double> lam((size_t)Nout * N), mu((size_t)Nout * (P ? P : 1), 0.0);
    for (int o = 0; o < Nout; ++o) {
        for (int i = 0; i < N; ++i) lam[(size_t)o * N + i] = (*driver.p_lambda)[o][i];
        for (int k = 0; k < P; ++k) mu[(size_t)o * P + k] = (*driver.p_mu)[o][k];
    }
    double *dl = driver.m_lam.upload(lam);
    double *dm = driver.m_mu.upload(mu);
    const double *dp = driver.m_par.upload(parameters);
    detail_driver::check(va_adjoint(driver.handle(), dp, Nout, dl, dm, nullptr), "va_adjoint");
    driver.m_lam.download(lam.data(), lam.size());
    driver.m_mu.download(mu.data(), (size_t)Nout * P);
    for (int o = 0; o < Nout; ++o) {
        for (int i = 0; i < N; ++i) (*driver.p_lambda)[o][i] = lam[(size_t)o * N + i];
        for (int k = 0; k < P; ++k) (*driver.p_mu)[o][k] = mu[(size_t)o * P + k];
    }
}

template <class State>
std::vector<std::vector<double>> sensitivity_matrix(Driver &driver, const State &parameters)
{
    const int N = driver.GetNin(), P = driver.GetNpar();
    if ((int)parameters.size() != P)
        throw std::runtime_error("computeSensitivityMatrix: parameter vector size differs from Npar");
    const double *dp = driver.m_par.upload(parameters);
    double *S = driver.m_sens.reserve((size_t)N * P);
    detail_driver::check(va_sensitivity_matrix(driver.handle(), dp, S, nullptr),
                         "va_sensitivity_matrix");
    std::vector<double> host((size_t)N * P);
    driver.m_sens.download(host.data(), host.size());
    std::vector<std::vector<double>> J(N, std::vector<double>(P));
    for (int i = 0; i < N; ++i)
        for (int k = 0; k < P; ++k) J[i][k] = host[(size_t)i * P + k];
    return J;
}

}  // namespace detail_backprop

// adjointSolve (backpropagation.hpp:18-43).
template <class State>
void adjointSolve(Driver &driver, const State &parameters)
{
    try {
        detail_backprop::check_setup(driver, true);
        detail_backprop::adjoint_solve(driver, parameters);
    } catch (std::exception &e) {
        std::cout << e.what() << std::endl;
    }
}

// computeSensitivityMatrix (backpropagation.hpp:67-87).
template <class State>
auto computeSensitivityMatrix(Driver &driver, const State &parameters)
{
    try {
        detail_backprop::check_setup(driver, false);
        return detail_backprop::sensitivity_matrix(driver, parameters);
    } catch (std::exception &e) {
        std::cout << e.what() << std::endl;
        return std::vector<std::vector<double>>(0, std::vector<double>(0));
    }
}

// computeSensitivityMatrixNoSIMD (backpropagation.hpp:46-65): the same result;
// on the device every seed is a lane already.
template <class State>
auto computeSensitivityMatrixNoSIMD(Driver &driver, const State &parameters)
{
    return computeSensitivityMatrix(driver, parameters);
}

}  // namespace vectorizedadjoint
