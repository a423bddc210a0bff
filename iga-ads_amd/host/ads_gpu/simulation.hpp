// C++ host surface of the MI355X ADS step: the reference's simulation_base /
// simulation_2d / simulation_3d vocabulary (IGA-ADS 0.1.0) over the C ABI in adsgpu.h.
//
// A problem class derives from gpu_simulation_2d or gpu_simulation_3d exactly as it would
// from ads::simulation_2d / simulation_3d, overrides the same private hooks (before,
// before_step, step, after_step, after: simulation_base.hpp:16-24) and calls the same
// protected helpers (prepare_matrices, solve, ads_solve, compute_rhs, fix_dof, ...). The
// difference is where the element loop lives: instead of a lambda run by an executor over
// elements() (validation.hpp:100-118), compute_rhs takes a description of the integrand
// (rhs_form) that the device kernels evaluate. State vectors live on the GPU; they are
// copied to the host only when the problem asks for them (state(), set_state()).
//
// Header-only; link with libadsgpu.so.
#pragma once

#include <array>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "adsgpu.h"

namespace ads_gpu {

// ads::dim_config (simulation/config.hpp:11-29)
struct dim_config {
    int p;
    int elements;
    double a;
    double b;
    int quad_order;
    int repeated_nodes;

    dim_config(int p, int elements, double a, double b, int quad_order, int repeated_nodes)
    : p{p}, elements{elements}, a{a}, b{b}, quad_order{quad_order}, repeated_nodes{repeated_nodes} {}

    dim_config(int p, int elements, double a = 0, double b = 1, int repeated_nodes = 0)
    : dim_config{p, elements, a, b, p + 1, repeated_nodes} {}
};

// ads::timesteps_config (simulation/config.hpp:31-39)
struct timesteps_config {
    int step_count;
    double dt;
    timesteps_config(int step_count, double dt) : step_count{step_count}, dt{dt} {}
};

// ads::config_2d / config_3d (simulation/config.hpp:46-56), plus the step arithmetic
struct config_2d {
    dim_config x, y;
    timesteps_config steps;
    int derivatives;
    int precision = ADSG_FP64;
};

struct config_3d {
    dim_config x, y, z;
    timesteps_config steps;
    int derivatives;
    int precision = ADSG_FP64;
};

class error : public std::runtime_error {
public:
    int code;
    error(int code, const std::string& what) : std::runtime_error(what), code(code) {}
};

// the integrand of compute_rhs (see adsg_rhs_form in adsgpu.h)
struct rhs_form : adsg_rhs_form {
    explicit rhs_form(double dt) : adsg_rhs_form{} {
        mass = 1.0;
        this->dt = dt;
        grad[0] = grad[1] = grad[2] = 1.0;
    }
    // explicit heat: u*v - dt*grad u . grad v (validation.hpp:110-115)
    static rhs_form heat(double dt) { return rhs_form{dt}; }
    // one explicit direction with factor dt/2 (implicit.hpp:132-182)
    static rhs_form split(double dt, int explicit_dir, int dims) {
        rhs_form f{0.5 * dt};
        for (int d = 0; d < 3; ++d) f.grad[d] = (d == explicit_dir && d < dims) ? 1.0 : 0.0;
        return f;
    }
    rhs_form& with_reaction(double r) {
        reaction = r;
        return *this;
    }
    rhs_form& with_kfield() {
        use_kfield = 1;
        return *this;
    }
    rhs_form& with_source() {
        use_source = 1;
        return *this;
    }
};

// ads::simulation_base (simulation_base.hpp:11-36, simulation_base.cpp:11-20)
class simulation_base {
protected:
    timesteps_config steps;

private:
    virtual void before() {}
    virtual void after() {}
    virtual void before_step(int /*iter*/, double /*t*/) {}
    virtual void step(int /*iter*/, double /*t*/) {}
    virtual void after_step(int /*iter*/, double /*t*/) {}

public:
    explicit simulation_base(const timesteps_config& steps) : steps{steps} {}
    simulation_base(const simulation_base&) = delete;
    simulation_base& operator=(const simulation_base&) = delete;
    virtual ~simulation_base() = default;

    void run() {
        before();
        for (int i = 0; i < steps.step_count; ++i) {
            double t = i * steps.dt;
            before_step(i, t);
            step(i, t);
            after_step(i, t);
        }
        after();
    }
};

// the device-side part shared by 2D and 3D
class gpu_simulation : public simulation_base {
protected:
    adsg_sim* h_ = nullptr;
    int dim_ = 0;
    int quad_count_[3] = {0, 0, 0};  // Gauss points per direction

    void check(int rc, const char* what) const {
        if (rc != ADSG_OK) throw error(rc, std::string(what) + ": " + adsg_last_error(h_));
    }

    gpu_simulation(const adsg_config& cfg, const timesteps_config& steps) : simulation_base{steps}, dim_{cfg.dim} {
        int rc = adsg_create(&cfg, &h_);
        if (rc != ADSG_OK) throw error(rc, std::string("adsg_create: ") + adsg_last_error(nullptr));
        const adsg_dim_config* dc[3] = {&cfg.x, &cfg.y, &cfg.z};
        for (int d = 0; d < dim_; ++d) quad_count_[d] = dc[d]->elements * (dc[d]->p + 1);
    }

    static adsg_dim_config to_c(const dim_config& d) {
        return adsg_dim_config{d.p, d.elements, d.a, d.b, d.quad_order, d.repeated_nodes};
    }

    // dimension::fix_dof / fix_left / fix_right (dimension.cpp:23-29, dimension.hpp:45-52)
    void fix_dof(int dir, int k) { check(adsg_fix_dof(h_, dir, k), "fix_dof"); }
    void fix_left(int dir) { fix_dof(dir, 0); }
    void fix_right(int dir) { fix_dof(dir, dofs(dir) - 1); }

    // a second 1D matrix per direction, mass*M + stiff*K (implicit.hpp:46-64)
    void set_matrix(int dir, int slot, double mass, double stiff) {
        check(adsg_set_matrix(h_, dir, slot, mass, stiff), "set_matrix");
    }

    // simulation_2d.hpp:56-59: factor every matrix (FP64) and upload the factors
    void prepare_matrices() {
        int rc = adsg_prepare_matrices(h_);
        if (rc > 0) throw error(rc, "prepare_matrices: zero pivot in a band matrix");
        check(rc, "prepare_matrices");
    }

    // simulation_2d.hpp:41: solve(rhs) = ads_solve(rhs, buffer, x.data(), y.data()[, z])
    void solve() { check(adsg_ads_solve(h_, nullptr), "ads_solve"); }
    // ads_solve with a chosen matrix per direction (solver.hpp:222-226)
    void ads_solve(const std::array<int, 3>& slots) { check(adsg_ads_solve(h_, slots.data()), "ads_solve"); }

    // before_step's swap(u, u_prev) (validation.hpp:73-76)
    void swap_state() { check(adsg_swap(h_), "swap"); }
    // the element loop: u := rhs assembled from u_prev (validation.hpp:95-119)
    void compute_rhs(const rhs_form& f) { check(adsg_compute_rhs(h_, &f), "compute_rhs"); }
    // validation.hpp:31-41
    void zero_boundary() { check(adsg_zero_boundary(h_), "zero_boundary"); }
    // nsteps of swap / compute_rhs / [zero_boundary] / solve from one captured graph
    void advance(int nsteps, const rhs_form& f, bool dirichlet = false) {
        check(adsg_step(h_, nsteps, &f, dirichlet ? ADSG_STEP_DIRICHLET : 0), "step");
    }
    // nsteps of the implicit step u += A^-1 r(u), A_d = the matrix in slots[d]
    void advance_delta(int nsteps, const rhs_form& f, const std::array<int, 3>& slots,
                       bool refine = false, bool dirichlet = false) {
        int flags = (dirichlet ? ADSG_STEP_DIRICHLET : 0) | (refine ? ADSG_STEP_REFINE : 0);
        check(adsg_step_delta(h_, nsteps, &f, slots.data(), flags), "step_delta");
    }
    // dgbtrs 'T' per direction (adjoint steps)
    void ads_solve_transposed(const std::array<int, 3>& slots) {
        check(adsg_ads_solve_transposed(h_, slots.data()), "ads_solve_transposed");
    }
    // generalized ADS: one band matrix per line of direction dir (solver.hpp:43-195)
    void set_line_matrices(int dir, int slot, int kl, int ku, const std::vector<double>& ab) {
        check(adsg_set_line_matrices(h_, dir, slot, kl, ku, ab.data()), "set_line_matrices");
    }
    // normL2 / normH1 (basic_simulation_2d.hpp:291-316) of u, on the device
    double normL2() {
        double v = 0;
        check(adsg_norm(h_, ADSG_U, ADSG_NORM_L2, &v), "norm");
        return v;
    }
    double normH1() {
        double v = 0;
        check(adsg_norm(h_, ADSG_U, ADSG_NORM_H1, &v), "norm");
        return v;
    }

    void set_vector(int which, const std::vector<double>& v) {
        if (static_cast<long>(v.size()) != size()) throw error(ADSG_EINVAL, "set_vector: wrong size");
        check(adsg_set_vector(h_, which, v.data()), "set_vector");
    }
    std::vector<double> vector(int which) {
        std::vector<double> v(size());
        check(adsg_get_vector(h_, which, v.data()), "get_vector");
        return v;
    }
    void set_state(const std::vector<double>& v) { set_vector(ADSG_U, v); }
    std::vector<double> state() { return vector(ADSG_U); }

    // basis_data x[e][q]: the Gauss points of direction dir, element-major
    std::vector<double> quad_points(int dir) const {
        if (dir < 0 || dir >= dim_) throw error(ADSG_EINVAL, "quad_points: bad direction");
        std::vector<double> x(quad_count_[dir]);
        check(adsg_quad_points(h_, dir, x.data()), "quad_points");
        return x;
    }
    // projection(v, f) (simulation_2d.hpp:43-46 -> projection.hpp:11-153): f evaluated at the
    // Gauss points on the host, its load vector added into vector `which` on the device
    void project_values(int which, const std::vector<double>& fq) {
        check(adsg_project(h_, which, fq.data()), "project");
    }
    // bspline::eval / eval_ders of vector `which` on the tensor grid px x py [x pz] (x fastest;
    // with ders the derivative planes follow the values)
    std::vector<double> sample(int which, const std::vector<double>& px, const std::vector<double>& py,
                               const std::vector<double>& pz = {}, bool ders = false) {
        const int counts[3] = {static_cast<int>(px.size()), static_cast<int>(py.size()),
                               dim_ == 3 ? static_cast<int>(pz.size()) : 1};
        const size_t total = px.size() * py.size() * (dim_ == 3 ? pz.size() : 1);
        std::vector<double> out(total * (ders ? 1 + dim_ : 1));
        check(adsg_sample(h_, which, counts, px.data(), py.data(), dim_ == 3 ? pz.data() : nullptr,
                          ders ? 1 : 0, out.data()),
              "sample");
        return out;
    }
    // error / norm functionals of values at the Gauss points (planes: value, d/dx, d/dy[, d/dz])
    double error_values(int which, int kind, const std::vector<double>& sq) {
        double v = 0;
        check(adsg_error(h_, which, kind, sq.data(), &v), "error");
        return v;
    }

public:
    ~gpu_simulation() override { adsg_destroy(h_); }
    int dofs(int dir) const { return adsg_dofs(h_, dir); }
    long size() const { return adsg_size(h_); }
    adsg_sim* handle() { return h_; }
    void synchronize() { check(adsg_synchronize(h_), "synchronize"); }
};

// ads::simulation_2d (simulation_2d.hpp:22-147)
class gpu_simulation_2d : public gpu_simulation {
    static adsg_config make(const config_2d& c) {
        adsg_config k{};
        k.dim = 2;
        k.x = to_c(c.x);
        k.y = to_c(c.y);
        k.step_count = c.steps.step_count;
        k.dt = c.steps.dt;
        k.derivatives = c.derivatives;
        k.precision = c.precision;
        return k;
    }

protected:
    std::array<int, 2> shape() const { return {dofs(0), dofs(1)}; }

    // projection(v, f) with f(x, y) (simulation_2d.hpp:43-46)
    template <typename Function>
    void projection(int which, Function&& f) {
        const auto px = quad_points(0), py = quad_points(1);
        std::vector<double> fq(px.size() * py.size());
        for (size_t j = 0; j < py.size(); ++j)
            for (size_t i = 0; i < px.size(); ++i) fq[i + px.size() * j] = f(px[i], py[j]);
        project_values(which, fq);
    }

    // errorL2 / errorH1 (basic_simulation_2d.hpp:318-336) of vector `which` against sol(x, y),
    // which returns a value_type {val, dx, dy}; which = -1 gives normL2 / normH1 of sol
    // (:266-290)
    template <typename Sol>
    double errorL2(int which, Sol&& sol) { return error_values(which, ADSG_NORM_L2, planes(sol)); }
    template <typename Sol>
    double errorH1(int which, Sol&& sol) { return error_values(which, ADSG_NORM_H1, planes(sol)); }
    template <typename Sol>
    double normL2(Sol&& sol) { return errorL2(-1, sol); }
    template <typename Sol>
    double normH1(Sol&& sol) { return errorH1(-1, sol); }
    using gpu_simulation::normH1;  // of u
    using gpu_simulation::normL2;

private:
    template <typename Sol>
    std::vector<double> planes(Sol&& sol) {
        const auto px = quad_points(0), py = quad_points(1);
        const size_t Q = px.size() * py.size();
        std::vector<double> sq(3 * Q);
        for (size_t j = 0; j < py.size(); ++j)
            for (size_t i = 0; i < px.size(); ++i) {
                const auto v = sol(px[i], py[j]);
                const size_t k = i + px.size() * j;
                sq[k] = v.val;
                sq[Q + k] = v.dx;
                sq[2 * Q + k] = v.dy;
            }
        return sq;
    }

public:
    explicit gpu_simulation_2d(const config_2d& c) : gpu_simulation{make(c), c.steps} {}
};

// ads::simulation_3d (simulation_3d.hpp:22-152)
class gpu_simulation_3d : public gpu_simulation {
    static adsg_config make(const config_3d& c) {
        adsg_config k{};
        k.dim = 3;
        k.x = to_c(c.x);
        k.y = to_c(c.y);
        k.z = to_c(c.z);
        k.step_count = c.steps.step_count;
        k.dt = c.steps.dt;
        k.derivatives = c.derivatives;
        k.precision = c.precision;
        return k;
    }

protected:
    std::array<int, 3> shape() const { return {dofs(0), dofs(1), dofs(2)}; }

    // projection(v, f) with f(x, y, z) (simulation_3d.hpp:45-48)
    template <typename Function>
    void projection(int which, Function&& f) {
        const auto px = quad_points(0), py = quad_points(1), pz = quad_points(2);
        std::vector<double> fq(px.size() * py.size() * pz.size());
        for (size_t k = 0; k < pz.size(); ++k)
            for (size_t j = 0; j < py.size(); ++j)
                for (size_t i = 0; i < px.size(); ++i)
                    fq[i + px.size() * (j + py.size() * k)] = f(px[i], py[j], pz[k]);
        project_values(which, fq);
    }

    // errorL2 / errorH1 (basic_simulation_3d.hpp:336-400) of vector `which` against
    // sol(x, y, z) returning {val, dx, dy, dz}; which = -1 gives normL2 / normH1 of sol
    template <typename Sol>
    double errorL2(int which, Sol&& sol) { return error_values(which, ADSG_NORM_L2, planes(sol)); }
    template <typename Sol>
    double errorH1(int which, Sol&& sol) { return error_values(which, ADSG_NORM_H1, planes(sol)); }
    template <typename Sol>
    double normL2(Sol&& sol) { return errorL2(-1, sol); }
    template <typename Sol>
    double normH1(Sol&& sol) { return errorH1(-1, sol); }
    using gpu_simulation::normH1;  // of u
    using gpu_simulation::normL2;

private:
    template <typename Sol>
    std::vector<double> planes(Sol&& sol) {
        const auto px = quad_points(0), py = quad_points(1), pz = quad_points(2);
        const size_t Q = px.size() * py.size() * pz.size();
        std::vector<double> sq(4 * Q);
        for (size_t k = 0; k < pz.size(); ++k)
            for (size_t j = 0; j < py.size(); ++j)
                for (size_t i = 0; i < px.size(); ++i) {
                    const auto v = sol(px[i], py[j], pz[k]);
                    const size_t m = i + px.size() * (j + py.size() * k);
                    sq[m] = v.val;
                    sq[Q + m] = v.dx;
                    sq[2 * Q + m] = v.dy;
                    sq[3 * Q + m] = v.dz;
                }
        return sq;
    }

public:
    explicit gpu_simulation_3d(const config_3d& c) : gpu_simulation{make(c), c.steps} {}
};

}  // namespace ads_gpu
