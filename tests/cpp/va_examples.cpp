// va_examples.cpp -- the reference's three example drivers, written against the
// drop-in C++ API (include/vectorizedadjoint/lib.hpp) and run on the device.
//
// Scenarios (reference drivers they follow):
//   vdp  -- examples/VanDerPol/main.cpp:52-125: Fehlberg 7(8) controlled,
//           mu = 1e3, t in [0, 0.5], dt0 = 1e-3, Nout = N identity seeds.
//   ho   -- examples/HarmonicOscillator/main.cpp:35-85: RK4 fixed step dt = 0.01
//           on [0, 10], k = 1 held by the functor, mu = 0.151, lambda = y(T).
//   glv  -- examples/GeneralizedLotkaVolterra/main.cpp:121-215: Cash-Karp 5(4)
//           controlled, x0 = 0.1, parameters from alphasfile_cpp.csv, t in
//           [0, 10], dt0 = 1e-3, Nout = N identity seeds; plus the sensitivity
//           matrix.
//   user -- a system the library does not know (the Brusselator), its host functor
//           bound with VA_BIND_DEVICE_SYSTEM to the device functor compiled in
//           tests/cpp/va_user_system.hip (recordDriverRHSFunction of an arbitrary
//           functor, Driver.hpp:95-100): DOPRI5 controlled, a = 1, b = 3,
//           x0 = (1.2, 3.1), t in [0, 10], Nout = N identity seeds.
//   pre  -- the precondition messages of adjointSolve (backpropagation.hpp:23-42).
//   fn   -- the batched AdjointFunction (forward / backward over 64 instances).
//
// Usage: va_examples <scenario> [tol] [N] [alphas.csv]
// Prints one JSON object with the results (all doubles to 17 digits).
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "vectorizedadjoint/lib.hpp"

using namespace vectorizedadjoint;

namespace {

struct VanDerPol {
    template <typename T>
    void operator()(const std::vector<T> &x, std::vector<T> &dxdt, const std::vector<T> &mu,
                    const T) const
    {
        dxdt[0] = x[1];
        dxdt[1] = mu[0] * ((1.0 - x[0] * x[0]) * x[1] - x[0]);
    }
};

class HarmonicOscillator
{
    double k = 1.0;

  public:
    template <typename T>
    void operator()(const std::vector<T> &r, std::vector<T> &drdt, const std::vector<T> &mu,
                    const T) const
    {
        drdt[0] = r[1];
        drdt[1] = -k * r[0] - mu[0] * r[1];
    }
};

struct LotkaVolterra {
    template <typename T>
    void operator()(const std::vector<T> &x, std::vector<T> &dxdt, const std::vector<T> &p, T)
    {
        const int n = (int)x.size();
        for (int i = 0; i < n; ++i) {
            T acc = 0.0;
            for (int j = 0; j < n; ++j) acc += p[n * (i + 1) + j] * x[j];
            dxdt[i] = x[i] * (p[i] + acc);
        }
    }
};

// A system of the caller's: the host functor (used by Driver::Rhs) and, bound to
// it, the device functor's kernels from libva_user_system.so.
struct Brusselator {
    template <typename T>
    void operator()(const std::vector<T> &x, std::vector<T> &dxdt, const std::vector<T> &p, T) const
    {
        dxdt[0] = p[0] + x[0] * x[0] * x[1] - (p[1] + 1.0) * x[0];
        dxdt[1] = p[1] * x[0] - x[0] * x[0] * x[1];
    }
};

}  // namespace

VA_BIND_DEVICE_SYSTEM(Brusselator, va_brusselator_kernels)

namespace {

// Counts observer calls and keeps the first and last (t, x).
struct Recorder {
    int *calls;
    std::vector<double> *first, *last;
    template <class State, class Time>
    void operator()(const State &x, Time t)
    {
        if (*calls == 0) {
            first->assign(1, (double)t);
            first->insert(first->end(), x.begin(), x.end());
        }
        last->assign(1, (double)t);
        last->insert(last->end(), x.begin(), x.end());
        ++*calls;
    }
};

void put(const char *key, const std::vector<double> &v, bool comma = true)
{
    std::printf("\"%s\": [", key);
    for (size_t i = 0; i < v.size(); ++i) std::printf("%s%.17g", i ? ", " : "", v[i]);
    std::printf("]%s\n", comma ? "," : "");
}

void put(const char *key, const std::vector<std::vector<double>> &m, bool comma = true)
{
    std::printf("\"%s\": [", key);
    for (size_t r = 0; r < m.size(); ++r) {
        std::printf("%s[", r ? ", " : "");
        for (size_t i = 0; i < m[r].size(); ++i) std::printf("%s%.17g", i ? ", " : "", m[r][i]);
        std::printf("]");
    }
    std::printf("]%s\n", comma ? "," : "");
}

std::vector<double> read_alphas(const std::string &path)
{
    std::vector<double> out;
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open " + path);
    std::string line, cell;
    while (std::getline(f, line)) {
        std::stringstream ss(line);
        while (std::getline(ss, cell, ',')) {
            if (cell.find_first_not_of(" \t\r\n") == std::string::npos) continue;
            out.push_back(std::stod(cell));
        }
    }
    return out;
}

// Forward + identity-seed adjoint + trajectory summary, shared by the scenarios.
template <class Stepper, class Plain, class System>
void run(const char *name, Stepper stepper, Plain plain, System sys, std::vector<double> x0,
         std::vector<double> p, double t0, double t1, double dt, int nout_identity,
         bool sensitivity)
{
    const int N = (int)x0.size(), P = (int)p.size();
    const int Nout = nout_identity ? N : 1;
    Driver driver(N, Nout, P);
    const std::vector<double> y0 = x0;
    int calls = 0;
    std::vector<double> first, last;
    const size_t steps =
        runge_kutta(stepper, sys, x0, p, t0, t1, dt, driver, Recorder{&calls, &first, &last});

    std::vector<std::vector<double>> lambda(Nout, std::vector<double>(N, 0.0));
    std::vector<std::vector<double>> mu(Nout, std::vector<double>(P, 0.0));
    if (nout_identity)
        for (int i = 0; i < N; ++i) lambda[i][i] = 1.0;
    else
        lambda[0] = x0;  // d(|y|^2/2)/dy(T)
    setCostGradients(driver, lambda, mu);
    constructDriverButcherTableau(driver, plain);
    recordDriverRHSFunction(driver, sys);
    adjointSolve(driver, p);

    std::printf("{\"scenario\": \"%s\", \"steps\": %zu, \"stored\": %d, \"observer_calls\": %d,\n",
                name, steps, driver.GetT(), calls);
    put("y0", y0);
    put("yT", x0);
    put("p", p);
    put("observer_first", first);
    put("observer_last", last);
    std::vector<double> h(driver.GetT()), t(driver.GetT());
    for (int n = 0; n < driver.GetT(); ++n) {
        h[n] = driver.GetDt(n);
        t[n] = driver.GetTime(n);
    }
    put("h", h);
    put("t", t);
    std::vector<double> rhs(N);
    driver.Rhs(y0, rhs, p, t0);
    put("rhs_y0", rhs);
    if (sensitivity) put("sensitivity", computeSensitivityMatrix(driver, p));
    put("lambda", lambda);
    put("mu", mu, false);
    std::printf("}\n");
}

int preconditions()
{
    // adjointSolve before setCostGradients / constructDriverButcherTableau /
    // recordDriverRHSFunction prints the reference's messages and returns.
    VanDerPol vdp;
    std::vector<double> x0 = {2.0, 0.0}, mu = {1.0};
    Driver d(2, 1, 1);
    runge_kutta(odeint::make_controlled<odeint::runge_kutta_dopri5<std::vector<double>>>(1e-8, 1e-8),
                vdp, x0, mu, 0.0, 1.0, 1e-3, d);
    adjointSolve(d, mu);
    std::vector<std::vector<double>> lam(1, std::vector<double>(2, 1.0)), m(1, std::vector<double>(1));
    setCostGradients(d, lam, m);
    adjointSolve(d, mu);
    constructDriverButcherTableau(d, odeint::runge_kutta4<std::vector<double>>());
    adjointSolve(d, mu);
    recordDriverRHSFunction(d, vdp);
    adjointSolve(d, mu);  // tableau differs from the forward's
    constructDriverButcherTableau(d, odeint::runge_kutta_dopri5<std::vector<double>>());
    adjointSolve(d, mu);  // succeeds
    std::printf("{\"scenario\": \"pre\", \"lambda\": [%.17g, %.17g], \"mu\": %.17g}\n", lam[0][0],
                lam[0][1], m[0][0]);
    return 0;
}

// AdjointFunction over a batch: Van der Pol, DOPRI5 tol 1e-8 on [0, 20], B = 64,
// y0 = (2 + 0.01 s_b, 0), mu_b = 0.5 + 4.5 (b + 0.5) / B, gy = 2 y(T)
// (the gradient of |y(T)|^2).
int batched()
{
    const int B = 64;
    va_problem pb;
    va_problem_init(&pb);
    pb.system = VA_SYSTEM_VAN_DER_POL;
    pb.tableau = VA_TABLEAU_DOPRI5;
    pb.n_state = 2;
    pb.n_param = 1;
    pb.batch = B;
    pb.t1 = 20.0;
    pb.abs_tol = pb.rel_tol = 1e-8;
    std::vector<double> y(2 * B), p(B);
    for (int b = 0; b < B; ++b) {
        y[b] = 2.0 + 0.01 * ((b % 7) - 3);
        y[B + b] = 0.0;
        p[b] = 0.5 + 4.5 * (b + 0.5) / B;
    }
    const std::vector<double> y0 = y;
    detail_driver::device_array<double> dy, dp, dg, dgy0, dgp;
    dy.upload(y);
    dp.upload(p);
    AdjointFunction fn(pb);
    fn.forward(dy.ptr, dp.ptr);
    dy.download(y.data(), y.size());
    std::vector<double> gy(2 * B);
    for (int i = 0; i < 2 * B; ++i) gy[i] = 2.0 * y[i];
    dg.upload(gy);
    dgy0.reserve(2 * B);
    dgp.reserve(B);
    fn.backward(dp.ptr, dg.ptr, nullptr, dgy0.ptr, dgp.ptr);
    std::vector<double> gy0(2 * B), gp(B);
    dgy0.download(gy0.data(), gy0.size());
    dgp.download(gp.data(), gp.size());
    std::printf("{\"scenario\": \"fn\", \"batch\": %d,\n", B);
    put("y0", y0);
    put("p", p);
    put("yT", y);
    put("dJdy0", gy0);
    put("dJdp", gp, false);
    std::printf("}\n");
    return 0;
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc < 2) {
        std::cerr << "usage: " << argv[0] << " vdp|ho|glv|user|pre|fn [tol] [N] [alphas.csv]\n";
        return 1;
    }
    const std::string which = argv[1];
    const double tol = argc > 2 ? std::stod(argv[2]) : 1e-5;
    using state = std::vector<double>;
    try {
        if (which == "vdp") {
            typedef odeint::runge_kutta_fehlberg78<state> stepper_type;
            state mu = {1e3};
            state x0 = {2.0, -2.0 / 3.0 + 10.0 / (81.0 * mu[0]) - 292.0 / (2187.0 * mu[0] * mu[0])};
            run("vdp", odeint::make_controlled<stepper_type>(tol, tol), stepper_type(), VanDerPol(),
                x0, mu, 0.0, 0.5, 0.001, 1, true);
        } else if (which == "ho") {
            typedef odeint::runge_kutta4<state> stepper_type;
            run("ho", stepper_type(), stepper_type(), HarmonicOscillator(), state{0.0, 1.0},
                state{0.151}, 0.0, 10.0, 0.01, 0, true);
        } else if (which == "glv") {
            const int N = argc > 3 ? std::stoi(argv[3]) : 10;
            if (argc < 5) throw std::runtime_error("glv needs the alphas csv path");
            state alphas = read_alphas(argv[4]);
            if ((int)alphas.size() != N + N * N) throw std::runtime_error("csv size != N + N*N");
            typedef odeint::runge_kutta_cash_karp54<state> stepper_type;
            run("glv", odeint::make_controlled<stepper_type>(tol, tol), stepper_type(),
                LotkaVolterra(), state(N, 0.1), alphas, 0.0, 10.0, 1e-3, 1, true);
        } else if (which == "user") {
            typedef odeint::runge_kutta_dopri5<state> stepper_type;
            run("user", odeint::make_controlled<stepper_type>(tol, tol), stepper_type(),
                Brusselator(), state{1.2, 3.1}, state{1.0, 3.0}, 0.0, 10.0, 1e-3, 1, true);
        } else if (which == "pre") {
            return preconditions();
        } else if (which == "fn") {
            return batched();
        } else {
            std::cerr << "unknown scenario " << which << "\n";
            return 1;
        }
    } catch (std::exception &e) {
        std::cerr << "error: " << e.what() << "\n";
        return 2;
    }
    return 0;
}
