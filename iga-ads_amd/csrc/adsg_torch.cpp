// libtorch autograd Functions over the ADS heat step with a B-spline diffusion field k
// (HeatK), and with k and a B-spline source field f (HeatKF): u(T) = step^N(u0; k[, f]), each
// step the explicit step of adsg_step with the form {mass 1, dt, k field[, source f]}
// (validation.hpp:95-119 with k as in flow.hpp:73-100). The backward pass is the
// hand-written discrete adjoint run over the same steps in reverse, with the forward states
// recomputed from checkpoints kept every `every` steps (adsg_forward_checkpointed /
// adsg_adjoint_checkpointed); it returns dL/dk, dL/df (dt M w per step) and dL/du0.
//
// Host code only (plain C++ over the C ABI); tensors are FP64 on the GPU and are exchanged
// with the handle by device-to-device copies.
#include <torch/extension.h>

#include <hip/hip_runtime_api.h>
#include <link.h>

#include <climits>
#include <cstdlib>
#include <cstring>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "adsgpu.h"

namespace {

using torch::Tensor;
using torch::autograd::AutogradContext;
using torch::autograd::tensor_list;

void check(adsg_sim* h, int rc, const char* what) {
    if (rc != ADSG_OK) throw std::runtime_error(std::string(what) + ": " + adsg_last_error(h));
}

// torch's kernels run on its own stream: finish them before the handle reads the tensors
void device_sync() {
    if (hipDeviceSynchronize() != hipSuccess) throw std::runtime_error("hipDeviceSynchronize failed");
}

adsg_rhs_form make_form(double dt, bool source = false) {
    adsg_rhs_form f{};
    f.mass = 1.0;
    f.dt = dt;
    f.grad[0] = f.grad[1] = f.grad[2] = 1.0;
    f.use_kfield = 1;
    f.use_source = source ? 1 : 0;
    return f;
}

// torch ships its own libamdhip64; when libadsgpu.so is loaded first, torch later maps its
// copy beside the one the library uses, and device pointers must not cross between two
// runtimes. Refuse to run in that case (loading torch first gives one runtime, torch's).
int collect_hip_runtime(dl_phdr_info* info, size_t, void* data) {
    const char* name = info->dlpi_name;
    if (name && *name) {
        const char* base = std::strrchr(name, '/');
        base = base ? base + 1 : name;
        if (std::strstr(base, "libamdhip64")) {
            char real[PATH_MAX];
            static_cast<std::set<std::string>*>(data)->insert(realpath(name, real) ? real : name);
        }
    }
    return 0;
}

void check_single_hip_runtime() {
    std::set<std::string> paths;
    dl_iterate_phdr(collect_hip_runtime, &paths);
    if (paths.size() > 1) {
        std::string list;
        for (const auto& p : paths) list += (list.empty() ? "" : ", ") + p;
        throw std::runtime_error("two HIP runtimes are loaded in this process (" + list +
                                 "): load torch before libadsgpu.so so that both use torch's runtime");
    }
}

void check_vector(const Tensor& t, long n, const char* name) {
    TORCH_CHECK(t.is_cuda(), name, " must be a GPU tensor");
    TORCH_CHECK(t.scalar_type() == torch::kFloat64, name, " must be float64");
    TORCH_CHECK(t.is_contiguous() && t.numel() == n, name, " must be contiguous with ", n, " entries");
}

struct HeatK : public torch::autograd::Function<HeatK> {
    static Tensor forward(AutogradContext* ctx, Tensor k, Tensor u0, int64_t handle, int64_t steps,
                          int64_t every, double dt, bool dirichlet) {
        auto* h = reinterpret_cast<adsg_sim*>(handle);
        const long n = adsg_size(h);
        TORCH_CHECK(h && n > 0, "bad handle");
        TORCH_CHECK(steps >= 1 && every >= 1, "steps and every must be positive");
        check_vector(k, n, "k");
        check_vector(u0, n, "u0");
        const int flags = dirichlet ? ADSG_STEP_DIRICHLET : 0;
        const adsg_rhs_form form = make_form(dt);
        device_sync();
        check(h, adsg_set_vector_device(h, ADSG_KFIELD, k.data_ptr<double>()), "set k");
        check(h, adsg_set_vector_device(h, ADSG_U, u0.data_ptr<double>()), "set u0");
        check(h, adsg_step(h, static_cast<int>(steps), &form, flags), "steps");
        Tensor uT = torch::empty_like(u0);
        check(h, adsg_get_vector_device(h, ADSG_U, uT.data_ptr<double>()), "get u(T)");
        ctx->save_for_backward({k, u0});
        ctx->saved_data["handle"] = handle;
        ctx->saved_data["steps"] = steps;
        ctx->saved_data["every"] = every;
        ctx->saved_data["dt"] = dt;
        ctx->saved_data["dirichlet"] = dirichlet;
        return uT;
    }

    // the forward is run again with checkpoints every `every` steps (the handle may have run
    // other work since), then the adjoint recomputes each segment from its checkpoint
    static tensor_list backward(AutogradContext* ctx, tensor_list grads) {
        auto saved = ctx->get_saved_variables();
        Tensor k = saved[0], u0 = saved[1];
        auto* h = reinterpret_cast<adsg_sim*>(ctx->saved_data["handle"].toInt());
        const int steps = static_cast<int>(ctx->saved_data["steps"].toInt());
        const int every = static_cast<int>(ctx->saved_data["every"].toInt());
        const double dt = ctx->saved_data["dt"].toDouble();
        const int flags = ctx->saved_data["dirichlet"].toBool() ? ADSG_STEP_DIRICHLET : 0;
        const adsg_rhs_form form = make_form(dt);
        Tensor lam = grads[0].to(torch::kFloat64).contiguous();
        Tensor dk = torch::zeros_like(k);
        Tensor du0 = torch::empty_like(u0);
        device_sync();
        check(h, adsg_set_vector_device(h, ADSG_KFIELD, k.data_ptr<double>()), "set k");
        check(h, adsg_set_vector_device(h, ADSG_U, u0.data_ptr<double>()), "set u0");
        check(h, adsg_forward_checkpointed(h, steps, every, &form, flags), "checkpointed forward");
        check(h, adsg_set_vector_device(h, ADSG_GRAD_K, dk.data_ptr<double>()), "clear dL/dk");
        check(h, adsg_set_vector_device(h, ADSG_U, lam.data_ptr<double>()), "dL/du(T)");
        check(h, adsg_adjoint_checkpointed(h, &form, flags), "adjoint");
        check(h, adsg_get_vector_device(h, ADSG_GRAD_K, dk.data_ptr<double>()), "dL/dk");
        check(h, adsg_get_vector_device(h, ADSG_U, du0.data_ptr<double>()), "dL/du0");
        return {dk, du0, Tensor(), Tensor(), Tensor(), Tensor(), Tensor()};
    }
};

struct HeatKF : public torch::autograd::Function<HeatKF> {
    static Tensor forward(AutogradContext* ctx, Tensor k, Tensor f, Tensor u0, int64_t handle,
                          int64_t steps, int64_t every, double dt, bool dirichlet) {
        auto* h = reinterpret_cast<adsg_sim*>(handle);
        const long n = adsg_size(h);
        TORCH_CHECK(h && n > 0, "bad handle");
        TORCH_CHECK(steps >= 1 && every >= 1, "steps and every must be positive");
        check_vector(k, n, "k");
        check_vector(f, n, "f");
        check_vector(u0, n, "u0");
        const int flags = dirichlet ? ADSG_STEP_DIRICHLET : 0;
        const adsg_rhs_form form = make_form(dt, true);
        device_sync();
        check(h, adsg_set_vector_device(h, ADSG_KFIELD, k.data_ptr<double>()), "set k");
        check(h, adsg_set_vector_device(h, ADSG_SOURCE, f.data_ptr<double>()), "set f");
        check(h, adsg_set_vector_device(h, ADSG_U, u0.data_ptr<double>()), "set u0");
        check(h, adsg_step(h, static_cast<int>(steps), &form, flags), "steps");
        Tensor uT = torch::empty_like(u0);
        check(h, adsg_get_vector_device(h, ADSG_U, uT.data_ptr<double>()), "get u(T)");
        ctx->save_for_backward({k, f, u0});
        ctx->saved_data["handle"] = handle;
        ctx->saved_data["steps"] = steps;
        ctx->saved_data["every"] = every;
        ctx->saved_data["dt"] = dt;
        ctx->saved_data["dirichlet"] = dirichlet;
        return uT;
    }

    static tensor_list backward(AutogradContext* ctx, tensor_list grads) {
        auto saved = ctx->get_saved_variables();
        Tensor k = saved[0], f = saved[1], u0 = saved[2];
        auto* h = reinterpret_cast<adsg_sim*>(ctx->saved_data["handle"].toInt());
        const int steps = static_cast<int>(ctx->saved_data["steps"].toInt());
        const int every = static_cast<int>(ctx->saved_data["every"].toInt());
        const double dt = ctx->saved_data["dt"].toDouble();
        const int flags = ctx->saved_data["dirichlet"].toBool() ? ADSG_STEP_DIRICHLET : 0;
        const adsg_rhs_form form = make_form(dt, true);
        Tensor lam = grads[0].to(torch::kFloat64).contiguous();
        Tensor dk = torch::zeros_like(k);
        Tensor df = torch::zeros_like(f);
        Tensor du0 = torch::empty_like(u0);
        device_sync();
        check(h, adsg_set_vector_device(h, ADSG_KFIELD, k.data_ptr<double>()), "set k");
        check(h, adsg_set_vector_device(h, ADSG_SOURCE, f.data_ptr<double>()), "set f");
        check(h, adsg_set_vector_device(h, ADSG_U, u0.data_ptr<double>()), "set u0");
        check(h, adsg_forward_checkpointed(h, steps, every, &form, flags), "checkpointed forward");
        check(h, adsg_set_vector_device(h, ADSG_GRAD_K, dk.data_ptr<double>()), "clear dL/dk");
        check(h, adsg_set_vector_device(h, ADSG_GRAD_F, df.data_ptr<double>()), "clear dL/df");
        check(h, adsg_set_vector_device(h, ADSG_U, lam.data_ptr<double>()), "dL/du(T)");
        check(h, adsg_adjoint_checkpointed(h, &form, flags), "adjoint");
        check(h, adsg_get_vector_device(h, ADSG_GRAD_K, dk.data_ptr<double>()), "dL/dk");
        check(h, adsg_get_vector_device(h, ADSG_GRAD_F, df.data_ptr<double>()), "dL/df");
        check(h, adsg_get_vector_device(h, ADSG_U, du0.data_ptr<double>()), "dL/du0");
        return {dk, df, du0, Tensor(), Tensor(), Tensor(), Tensor(), Tensor()};
    }
};

Tensor heat_k(Tensor k, Tensor u0, int64_t handle, int64_t steps, int64_t every, double dt,
              bool dirichlet) {
    check_single_hip_runtime();
    return HeatK::apply(k, u0, handle, steps, every, dt, dirichlet);
}

Tensor heat_kf(Tensor k, Tensor f, Tensor u0, int64_t handle, int64_t steps, int64_t every,
               double dt, bool dirichlet) {
    check_single_hip_runtime();
    return HeatKF::apply(k, f, u0, handle, steps, every, dt, dirichlet);
}

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
    m.def("heat_k", &heat_k,
          "u(T) after `steps` explicit ADS heat steps with diffusion field k (autograd: "
          "hand-written adjoint with checkpoints every `every` steps)",
          py::arg("k"), py::arg("u0"), py::arg("handle"), py::arg("steps"), py::arg("every"),
          py::arg("dt"), py::arg("dirichlet"));
    m.def("heat_kf", &heat_kf,
          "u(T) after `steps` explicit ADS heat steps with diffusion field k and source field f "
          "(autograd in k, f and u0: hand-written adjoint with checkpoints every `every` steps)",
          py::arg("k"), py::arg("f"), py::arg("u0"), py::arg("handle"), py::arg("steps"),
          py::arg("every"), py::arg("dt"), py::arg("dirichlet"));
}
