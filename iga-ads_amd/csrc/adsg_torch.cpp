// libtorch autograd Function over the ADS heat step with a B-spline diffusion field k:
// u(T) = step^N(u0; k), each step the explicit step of adsg_step with the form
// {mass 1, dt, k field} (validation.hpp:95-119 with k as in flow.hpp:73-100). The backward
// pass is the hand-written discrete adjoint (adsg_adjoint_step) run over the same steps in
// reverse, with the forward states recomputed from checkpoints kept every `every` steps.
//
// Host code only (plain C++ over the C ABI); tensors are FP64 on the GPU and are exchanged
// with the handle by device-to-device copies.
#include <torch/extension.h>

#include <hip/hip_runtime_api.h>

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

adsg_rhs_form make_form(double dt) {
    adsg_rhs_form f{};
    f.mass = 1.0;
    f.dt = dt;
    f.grad[0] = f.grad[1] = f.grad[2] = 1.0;
    f.use_kfield = 1;
    return f;
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
        const int64_t nchk = (steps + every - 1) / every;
        Tensor chk = torch::empty({nchk, n}, u0.options());
        device_sync();
        check(h, adsg_set_vector_device(h, ADSG_KFIELD, k.data_ptr<double>()), "set k");
        check(h, adsg_set_vector_device(h, ADSG_U, u0.data_ptr<double>()), "set u0");
        for (int64_t c = 0; c < nchk; ++c) {
            check(h, adsg_get_vector_device(h, ADSG_U, chk[c].data_ptr<double>()), "checkpoint");
            const int64_t todo = std::min(every, steps - c * every);
            check(h, adsg_step(h, static_cast<int>(todo), &form, flags), "step");
        }
        Tensor uT = torch::empty_like(u0);
        check(h, adsg_get_vector_device(h, ADSG_U, uT.data_ptr<double>()), "get u(T)");
        ctx->save_for_backward({k, chk});
        ctx->saved_data["handle"] = handle;
        ctx->saved_data["steps"] = steps;
        ctx->saved_data["every"] = every;
        ctx->saved_data["dt"] = dt;
        ctx->saved_data["dirichlet"] = dirichlet;
        return uT;
    }

    static tensor_list backward(AutogradContext* ctx, tensor_list grads) {
        auto saved = ctx->get_saved_variables();
        Tensor k = saved[0], chk = saved[1];
        auto* h = reinterpret_cast<adsg_sim*>(ctx->saved_data["handle"].toInt());
        const int64_t steps = ctx->saved_data["steps"].toInt();
        const int64_t every = ctx->saved_data["every"].toInt();
        const double dt = ctx->saved_data["dt"].toDouble();
        const int flags = ctx->saved_data["dirichlet"].toBool() ? ADSG_STEP_DIRICHLET : 0;
        const adsg_rhs_form form = make_form(dt);
        const long n = adsg_size(h);
        Tensor lam = grads[0].to(torch::kFloat64).contiguous().clone();
        Tensor dk = torch::zeros_like(k);
        Tensor seg = torch::empty({every, n}, k.options());
        device_sync();
        check(h, adsg_set_vector_device(h, ADSG_KFIELD, k.data_ptr<double>()), "set k");
        check(h, adsg_set_vector_device(h, ADSG_GRAD_K, dk.data_ptr<double>()), "clear dL/dk");
        const int64_t nchk = chk.size(0);
        for (int64_t c = nchk - 1; c >= 0; --c) {
            const int64_t start = c * every, stop = std::min(steps, start + every);
            // recompute the forward states of this segment from its checkpoint
            check(h, adsg_set_vector_device(h, ADSG_U, chk[c].data_ptr<double>()), "restore");
            for (int64_t s = start; s < stop; ++s) {
                check(h, adsg_get_vector_device(h, ADSG_U, seg[s - start].data_ptr<double>()), "keep state");
                if (s + 1 < stop) check(h, adsg_step(h, 1, &form, flags), "recompute");
            }
            for (int64_t s = stop - 1; s >= start; --s) {
                check(h, adsg_set_vector_device(h, ADSG_U_PREV, seg[s - start].data_ptr<double>()), "u(n)");
                check(h, adsg_set_vector_device(h, ADSG_U, lam.data_ptr<double>()), "lam(n+1)");
                check(h, adsg_adjoint_step(h, &form, flags), "adjoint step");
                check(h, adsg_get_vector_device(h, ADSG_U, lam.data_ptr<double>()), "lam(n)");
            }
        }
        check(h, adsg_get_vector_device(h, ADSG_GRAD_K, dk.data_ptr<double>()), "dL/dk");
        return {dk, lam, Tensor(), Tensor(), Tensor(), Tensor(), Tensor()};
    }
};

Tensor heat_k(Tensor k, Tensor u0, int64_t handle, int64_t steps, int64_t every, double dt,
              bool dirichlet) {
    return HeatK::apply(k, u0, handle, steps, every, dt, dirichlet);
}

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
    m.def("heat_k", &heat_k,
          "u(T) after `steps` explicit ADS heat steps with diffusion field k (autograd: "
          "hand-written adjoint with checkpoints every `every` steps)",
          py::arg("k"), py::arg("u0"), py::arg("handle"), py::arg("steps"), py::arg("every"),
          py::arg("dt"), py::arg("dirichlet"));
}
