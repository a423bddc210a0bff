// The reference's validation problem (examples/validation/validation.hpp, main.cpp) on the
// GPU host surface: 2D heat u_t = Lap u on [0,1]^2, zero Dirichlet data on every side,
// u(0) = sin(pi x) sin(pi y), whose exact solution decays as exp(-2 pi^2 t). before()
// projects the initial state (projection + solve with the boundary rows zeroed, as the
// problem's own solve does), each step assembles the explicit rhs, zeroes its boundary rows
// and solves, and after() prints the relative L2 and H1 errors at T in percent.
//
//   validation_gpu <p> <n> <nsteps> [out.data]   (T = 1, dt = T / nsteps, nsteps / 10 steps)
//
// With out.data, the final solution is also sampled on the output manager's grid of 200
// intervals per side (validation.hpp:24, output_manager.hpp:52-75) and written as gnuplot
// rows.
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <string>

#include "ads_gpu/output.hpp"
#include "ads_gpu/simulation.hpp"

namespace {

struct value_type {
    double val, dx, dy;
};

class validation : public ads_gpu::gpu_simulation_2d {
    std::string out_;

public:
    validation(const ads_gpu::config_2d& c, std::string out) : gpu_simulation_2d{c}, out_{std::move(out)} {}

private:
    static constexpr double k = 2 * M_PI * M_PI;

    static double sc(double t) { return std::exp(-k * t); }
    static double fi(double x, double y) { return std::sin(x * M_PI) * std::sin(y * M_PI); }

    static value_type solution(double x, double y, double t) {
        return {sc(t) * fi(x, y), sc(t) * M_PI * std::cos(x * M_PI) * std::sin(y * M_PI),
                sc(t) * M_PI * std::sin(x * M_PI) * std::cos(y * M_PI)};
    }

    // validation.hpp:31-41: the problem's solve zeroes the boundary rows first
    void solve_zeroed() {
        zero_boundary();
        solve();
    }

    void before() override {
        for (int d = 0; d < 2; ++d) {
            fix_left(d);
            fix_right(d);
        }
        prepare_matrices();
        projection(ADSG_U, [](double x, double y) { return fi(x, y) * sc(0); });
        solve_zeroed();
    }

    void before_step(int /*iter*/, double /*t*/) override { swap_state(); }

    void step(int /*iter*/, double /*t*/) override {
        compute_rhs(ads_gpu::rhs_form::heat(steps.dt));
        solve_zeroed();
    }

    void after() override {
        const double T = steps.dt * steps.step_count;
        auto sol = [T](double x, double y) { return solution(x, y, T); };
        const double eL2 = errorL2(ADSG_U, sol) / normL2(sol) * 100;
        const double eH1 = errorH1(ADSG_U, sol) / normH1(sol) * 100;
        std::cout << std::setprecision(17) << eL2 << "  " << eH1 << std::endl;
        if (!out_.empty()) {
            const auto xs = ads_gpu::output::linspace(0, 1, 200), ys = xs;
            std::ofstream os{out_};
            ads_gpu::output::gnuplot_2d(os, xs, ys, sample(ADSG_U, xs, ys));
        }
    }
};

}  // namespace

int main(int argc, char* argv[]) {
    if (argc < 4) {
        std::cerr << "Usage: validation_gpu <p> <n> <nsteps>" << std::endl;
        return 2;
    }
    const int p = std::atoi(argv[1]);
    const int n = std::atoi(argv[2]);
    int nsteps = std::atoi(argv[3]);
    if (n <= 0 || nsteps < 10) {
        std::cerr << "validation_gpu: n must be positive and nsteps at least 10" << std::endl;
        return 2;
    }
    const double T = 1.0;
    const double dt = T / nsteps;
    nsteps /= 10;
    try {
        ads_gpu::dim_config dim{p, n};
        ads_gpu::timesteps_config steps{nsteps, dt};
        ads_gpu::config_2d c{dim, dim, steps, 1, ADSG_FP64};
        validation sim{c, argc > 4 ? argv[4] : ""};
        sim.run();
    } catch (const std::exception& e) {
        std::cerr << "validation_gpu: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
