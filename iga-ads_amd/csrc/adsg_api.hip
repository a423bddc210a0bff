// C ABI of the MI355X ADS step (declared in include/adsgpu.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "adsg_sim.hpp"

namespace {

thread_local std::string g_create_error;

}  // namespace

using adsg::dev_dir;
using adsg::dev_slot;

namespace {

// bytes per coefficient of a vector
size_t elt_of(const adsg_sim* h, int which) {
    if (which == ADSG_U || which == ADSG_U_PREV || which == ADSG_GRAD_K || which == ADSG_GRAD_F)
        return sizeof(double);
    return h->precision == ADSG_FP32 ? sizeof(float) : sizeof(double);
}

int fail(adsg_sim* h, int code, const std::string& msg) {
    if (h) h->err = msg;
    return code;
}

int hip_fail(adsg_sim* h, hipError_t e, const char* what) {
    return fail(h, ADSG_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(h, expr)                                      \
    do {                                                      \
        hipError_t e_ = (expr);                               \
        if (e_ != hipSuccess) return hip_fail((h), e_, #expr); \
    } while (0)

// phase boundary i of a timed step: 1 after the zero fill, 2 after the element kernels,
// 3 after the boundary handling, 4..6 after the x, y, z sweeps
void mark(adsg_sim* h, int i) {
    if (h->marks) (void)hipEventRecord(h->marks[i], h->stream);
}

// element-kernel choice passed to launch_rhs: the variant, and the column shape of variant 5
int rhs_code(const adsg_sim* h) {
    return h->rhs_variant | (h->rhs_col_shape << 4) | (h->rhs_tiling && h->spill ? adsg::kRhsOneLaunch : 0);
}

// the spill array of the one-launch tiling, allocated when the tiling is first selected
// (never inside a captured step); FP64-sized in both precisions, since FP32 handles also run
// FP64 element loops (the adjoint's and the norms' products)
int ensure_spill(adsg_sim* h) {
    if (h->spill || h->general) return ADSG_OK;
    const int el[3] = {h->dirs[0].tab.elements, h->dirs[1].tab.elements, h->dim == 3 ? h->dirs[2].tab.elements : 1};
    const size_t bytes = static_cast<size_t>(adsg::rhs_spill_slots(h->dim, h->dirs[0].tab.p, el)) * sizeof(double);
    if (bytes && hipMalloc(&h->spill, bytes) != hipSuccess) {
        h->spill = nullptr;
        return fail(h, ADSG_ENOMEM, "one-launch tiling: no memory for the spill array");
    }
    return ADSG_OK;
}

// a device copy of bytes at *dst (freed and replaced if set)
int upload_bytes(adsg_sim* h, void** dst, const void* data, size_t bytes) {
    if (*dst && !h->in_arena) (void)hipFree(*dst);
    *dst = nullptr;
    if (hipMalloc(dst, bytes) != hipSuccess) return fail(h, ADSG_ENOMEM, "hipMalloc failed for tables");
    HIP_TRY(h, hipMemcpy(*dst, data, bytes, hipMemcpyHostToDevice));
    return ADSG_OK;
}

template <typename T>
int upload(adsg_sim* h, void** dst, const std::vector<double>& src) {
    std::vector<T> tmp(std::max<size_t>(src.size(), 1), T(0));
    for (size_t i = 0; i < src.size(); ++i) tmp[i] = static_cast<T>(src[i]);
    return upload_bytes(h, dst, tmp.data(), tmp.size() * sizeof(T));
}

// band storage (ld = 3p+1, A(i,j) at j*ld + 2p + i - j) -> rows M[i][b] = A(i, i+b-p)
std::vector<double> band_rows(const std::vector<double>& ab, int n, int p) {
    const int ld = 3 * p + 1, w = 2 * p + 1;
    std::vector<double> r(static_cast<size_t>(n) * w, 0.0);
    for (int i = 0; i < n; ++i)
        for (int b = 0; b < w; ++b) {
            const int j = i + b - p;
            if (j >= 0 && j < n) r[static_cast<size_t>(i) * w + b] = ab[static_cast<size_t>(j) * ld + 2 * p + i - j];
        }
    return r;
}

int upload_tables(adsg_sim* h, dev_dir& d) {
    int rc;
    // the element tables zero-padded to whole kTablePad-element tiles, and their weighted copies
    const int per = d.tab.nq * (d.tab.p + 1);
    const int padded = (d.tab.elements + adsg::kTablePad - 1) / adsg::kTablePad * adsg::kTablePad;
    std::vector<double> val(static_cast<size_t>(padded) * per, 0.0), der(val.size(), 0.0);
    std::vector<double> bval(val.size(), 0.0), bder(val.size(), 0.0);
    for (int e = 0; e < d.tab.elements; ++e)
        for (int q = 0; q < d.tab.nq; ++q)
            for (int i = 0; i <= d.tab.p; ++i) {
                const size_t k = (static_cast<size_t>(e) * d.tab.nq + q) * (d.tab.p + 1) + i;
                const double w = d.tab.wj[static_cast<size_t>(e) * d.tab.nq + q];
                val[k] = d.tab.val[k];
                der[k] = d.tab.der[k];
                bval[k] = d.tab.val[k] * w;
                bder[k] = d.tab.der[k] * w;
            }
    if (h->precision == ADSG_FP32) {
        if ((rc = upload<float>(h, &d.val, val))) return rc;
        if ((rc = upload<float>(h, &d.der, der))) return rc;
        if ((rc = upload<float>(h, &d.bval, bval))) return rc;
        if ((rc = upload<float>(h, &d.bder, bder))) return rc;
        if ((rc = upload<float>(h, &d.wj, d.tab.wj))) return rc;
    } else {
        if ((rc = upload<double>(h, &d.val, val))) return rc;
        if ((rc = upload<double>(h, &d.der, der))) return rc;
        if ((rc = upload<double>(h, &d.bval, bval))) return rc;
        if ((rc = upload<double>(h, &d.bder, bder))) return rc;
        if ((rc = upload<double>(h, &d.wj, d.tab.wj))) return rc;
    }
    // form_matrix.cpp:8-42 on the same tables: the Gram and stiffness matrices, unfixed, in
    // FP64 in both precisions (the separable rhs contracts in FP64)
    const std::vector<double> M = band_rows(adsg::band_form(d.tab, 1.0, 0.0), d.tab.dofs, d.tab.p);
    const std::vector<double> K = band_rows(adsg::band_form(d.tab, 0.0, 1.0), d.tab.dofs, d.tab.p);
    if ((rc = upload<double>(h, &d.kM, M))) return rc;
    if ((rc = upload<double>(h, &d.kK, K))) return rc;
    // first DOF of every element
    return upload_bytes(h, reinterpret_cast<void**>(&d.first), d.tab.first.data(),
                        sizeof(int) * std::max<size_t>(d.tab.first.size(), 1));
}

// the device tables of every direction, in the handle's precision TT
template <typename TT>
void dir_tables(const adsg_sim* h, adsg::dev_tables<TT>* t) {
    for (int d = 0; d < h->dim; ++d) {
        t[d].val = static_cast<const TT*>(h->dirs[d].val);
        t[d].der = static_cast<const TT*>(h->dirs[d].der);
        t[d].wj = static_cast<const TT*>(h->dirs[d].wj);
        t[d].bval = static_cast<const TT*>(h->dirs[d].bval);
        t[d].bder = static_cast<const TT*>(h->dirs[d].bder);
        t[d].elements = h->dirs[d].tab.elements;
        t[d].dofs = h->dirs[d].tab.dofs;
        t[d].first = h->dirs[d].first;
        t[d].nq = h->dirs[d].tab.nq;
        t[d].p = h->dirs[d].tab.p;
    }
}

// the element loop of the handle: the general kernel for configurations outside the
// sum-factorised kernels', else the tuned variant
// (overwrite: out := the element sums, its start value unused; else out += them; defer: a
// one-launch tiling describes its fix-up there instead of running it, see sweep_all)
template <typename T>
hipError_t launch_elements(const adsg_sim* h, const T* up, T* out, const adsg::dev_tables<T>* t,
                           const adsg::rhs_params<T>& prm, bool overwrite = false,
                           adsg::rhs_fixup<T>* defer = nullptr) {
    if (h->general) {
        if (overwrite) {
            const hipError_t e = adsg::launch_fill<T>(out, h->N, T(0), h->stream);
            if (e != hipSuccess) return e;
        }
        return adsg::launch_rhs_general<T>(h->dim, up, out, t, prm, h->stream);
    }
    adsg::rhs_params<T> q = prm;
    q.spill = h->rhs_tiling ? static_cast<T*>(h->spill) : nullptr;
    q.defer = defer;  // the one-launch tiling leaves its fix-up to the caller
    return adsg::launch_rhs<T>(h->dim, h->dirs[0].tab.p, up, out, t, q,
                               rhs_code(h) | (overwrite ? adsg::kRhsOverwrite : 0), h->stream);
}

// entry points whose kernels assume one degree and p+1 Gauss points in every direction
int refuse_general(adsg_sim* h, const char* what) {
    return fail(h, ADSG_EINVAL, std::string(what) + ": not available for configurations with quad_order != p+1, "
                                "repeated interior knots or a degree per direction");
}

// the slot's device factors (pointers into a batch's arena are only forgotten)
void free_slot(dev_slot& s, bool in_arena = false) {
    for (void** p : {&s.lower, &s.upper, &s.rdiag, &s.pack, &s.packt}) {
        if (*p && !in_arena) (void)hipFree(*p);
        *p = nullptr;
    }
    if (s.poff && !in_arena) (void)hipFree(s.poff);
    s.poff = nullptr;
    if (s.arow && !in_arena) (void)hipFree(s.arow);
    s.arow = nullptr;
}

void drop_graphs(adsg_sim* h) {
    for (auto& kv : h->graphs) (void)hipGraphExecDestroy(kv.second);
    h->graphs.clear();
}

bool valid_dir(const adsg_sim* h, int dir) { return h && dir >= 0 && dir < h->dim; }

adsg::line_geometry geometry(const adsg_sim* h, int dir) {
    const long nx = h->dirs[0].tab.dofs, ny = h->dirs[1].tab.dofs;
    const long nz = h->dim == 3 ? h->dirs[2].tab.dofs : 1;
    adsg::line_geometry g;
    if (dir == 0) {  // line l starts at l * nx
        g.n_lines = ny * nz;
        g.A = 1;
        g.B = nx;
        g.S = 1;
    } else if (dir == 1) {
        g.n_lines = nx * nz;
        g.A = nx;
        g.B = nx * ny;
        g.S = nx;
    } else {
        g.n_lines = nx * ny;
        g.A = nx * ny;
        g.B = 0;
        g.S = nx * ny;
    }
    return g;
}

// mode: 0 = A^-1, 1 = A^-T, 2 = A^-1 with one FP64-residual refinement (FP32 data)
template <typename T>
int sweep_dir(adsg_sim* h, T* u, int d, int sl, int mode = 0, const adsg::rhs_fixup<T>* fix = nullptr,
              const adsg::seg_apply* apply = nullptr, bool* applied = nullptr, const T* src = nullptr) {
    if (applied) *applied = false;
    // src: the sweep reads its lines there and writes them to u (the segmented sweep); the
    // other sweeps work in place on a copy
    auto in_place = [&]() -> hipError_t {
        if (!src) return hipSuccess;
        const hipError_t e = hipMemcpyAsync(u, src, sizeof(T) * h->N, hipMemcpyDeviceToDevice, h->stream);
        src = nullptr;
        return e;
    };
    if (sl < 0 || sl >= static_cast<int>(h->dirs[d].slots.size()) || !h->dirs[d].slots[sl].factored)
        return fail(h, ADSG_ESTATE, "ads_solve: matrix slot not set or not factored");
    const dev_slot& s = h->dirs[d].slots[sl];
    if (s.lines > 0) {
        if (mode != 0) return fail(h, ADSG_EINVAL, "ads_solve: per-line matrices support the plain solve only");
        HIP_TRY(h, in_place());
        adsg::own_factors<T> of;
        of.lower = static_cast<const T*>(s.lower);
        of.upper = static_cast<const T*>(s.upper);
        of.rdiag = static_cast<const T*>(s.rdiag);
        of.poff = s.poff;
        of.n = s.lf.n;
        of.kl = s.lf.kl;
        of.ku = s.lf.ku_eff;
        of.pivoted = s.lf.pivoted;
        hipError_t e = adsg::launch_line_sweep_own<T>(u, geometry(h, d), of, h->stream);
        return e == hipSuccess ? ADSG_OK : hip_fail(h, e, "line sweep (per-line matrices)");
    }
    adsg::dev_factors<T> f;
    f.lower = static_cast<const T*>(s.lower);
    f.upper = static_cast<const T*>(s.upper);
    f.rdiag = static_cast<const T*>(s.rdiag);
    f.poff = s.poff;
    f.n = s.lf.n;
    f.kl = s.lf.kl;
    f.ku = s.lf.ku_eff;
    f.pivoted = s.lf.pivoted;
    const bool seg_ok = s.seg.nseg > 0;
    if (mode != 0 && !seg_ok)
        return fail(h, ADSG_EINVAL,
                    "ads_solve: transposed and refined solves need unpivoted factors and lines of "
                    "2..16 segments");
    hipError_t e = hipErrorInvalidValue;
    if (!(seg_ok && (mode != 0 || h->sweep_variant != 1))) HIP_TRY(h, in_place());
    if (seg_ok && (mode != 0 || h->sweep_variant != 1)) {
        adsg::seg_factors<T> sf;
        sf.f = f;
        sf.pack = static_cast<const T*>(s.pack);
        sf.packt = static_cast<const T*>(s.packt);
        sf.arow = s.arow;
        sf.m = s.seg.m;
        sf.nseg = s.seg.nseg;
        sf.mseg = s.seg.mseg;
        sf.transposed = mode == 1;
        sf.refine = mode == 2;
        sf.direct = h->sweep_variant == 2;  // x lines read lane by lane (small meshes)
        sf.src = src;
        e = adsg::launch_line_sweep_seg<T, T>(u, geometry(h, d), sf, h->stream, fix, apply, applied);
        if (e == hipErrorInvalidValue && mode != 0)
            return fail(h, ADSG_EINVAL, "ads_solve: line staging does not fit the LDS for this solve");
    }
    if (e == hipErrorInvalidValue) HIP_TRY(h, in_place());  // the fallback below works in place
    if (e == hipErrorInvalidValue && mode == 0 && fix && fix->pending()) {
        e = adsg::launch_rhs_fixup<T>(u, *fix, h->stream);  // the sweep below cannot take it
        if (e != hipSuccess) return hip_fail(h, e, "rhs fix-up");
        e = hipErrorInvalidValue;
    }
    if (e == hipErrorInvalidValue && mode == 0) e = adsg::launch_line_sweep<T, T>(u, geometry(h, d), f, h->stream);
    if (e == hipErrorInvalidValue) return fail(h, ADSG_EINVAL, "ads_solve: unsupported band widths (kl, ku)");
    if (e != hipSuccess) return hip_fail(h, e, "line sweep");
    return ADSG_OK;
}

// ads_solve: one sweep per direction in the order x, y, z (solver.hpp:35-41)
// The generalized ADS (solver.hpp:43-195) solves the direction with per-line matrices first
// and the others after it in cyclic order; the standard directions commute.
// fix: the element loop's deferred fix-up and boundary zeroing of u, in one kernel before
// the sweeps; a zeroing alone is done by the x sweep as it loads its lines when that sweep
// runs first as the segmented plain solve. apply: the step's increment update, done by the
// last sweep as it stores its lines where that is a segmented sweep of strided lines
// (*applied tells; the caller runs the update otherwise).
template <typename T>
int sweep_all(adsg_sim* h, T* u, const int* slots, int mode = 0, const adsg::rhs_fixup<T>* fix = nullptr,
              adsg::seg_apply* apply = nullptr, bool* applied = nullptr, const T* src = nullptr) {
    if (applied) *applied = false;
    int special = -1;
    for (int d = 0; d < h->dim; ++d) {
        const int sl = slots ? slots[d] : 0;
        if (sl >= 0 && sl < static_cast<int>(h->dirs[d].slots.size()) && h->dirs[d].slots[sl].lines > 0) {
            if (special >= 0) return fail(h, ADSG_EINVAL, "ads_solve: at most one direction with per-line matrices");
            special = d;
        }
    }
    const int start = special < 0 ? 0 : special;
    bool fused = false;
    if (fix && fix->pending()) {
        const int sl = slots ? slots[0] : 0;
        const bool seg_x = sl >= 0 && sl < static_cast<int>(h->dirs[0].slots.size()) &&
                           h->dirs[0].slots[sl].factored && h->dirs[0].slots[sl].lines == 0 &&
                           h->dirs[0].slots[sl].seg.nseg > 0 && h->sweep_variant != 1;
        fused = !fix->spill && start == 0 && mode == 0 && seg_x;  // zeroing only: in the x sweep
        if (!fused) {
            if (src) {  // the fix-up works on u: the sweeps then run in place
                HIP_TRY(h, hipMemcpyAsync(u, src, sizeof(T) * h->N, hipMemcpyDeviceToDevice, h->stream));
                src = nullptr;
            }
            const hipError_t e = adsg::launch_rhs_fixup<T>(u, *fix, h->stream);
            if (e != hipSuccess) return hip_fail(h, e, "rhs fix-up");
        }
    }
    for (int i = 0; i < 3; ++i) {
        if (i < h->dim) {
            const int d = (start + i) % h->dim;
            const bool last = i == h->dim - 1 && apply && d != 0;  // d = 0: x lines, stored through LDS
            if (last) apply->dir = d;
            int rc = sweep_dir<T>(h, u, d, slots ? slots[d] : 0, mode, i == 0 && fused ? fix : nullptr,
                                  last ? apply : nullptr, last ? applied : nullptr, i == 0 ? src : nullptr);
            if (rc) return rc;
        }
        mark(h, 4 + i);
    }
    return ADSG_OK;
}

// delta: the integrand exactly as given, from u_prev (its FP32 shadow in FP32 mode) into
// buffer; otherwise compute_rhs semantics (FP64: into u; FP32: rhs - M u_prev into buffer)
// b_i = sum_e sum_q w J f(x_q) B_i(x_q) for the source f: the element kernel with the
// form {mass 0, dt 1, no gradient, source f}, in the step precision
template <typename T>
int source_load_impl(adsg_sim* h) {
    if (!h->source_load && hipMalloc(&h->source_load, h->N * sizeof(T)) != hipSuccess)
        return fail(h, ADSG_ENOMEM, "source: no memory for its load vector");
    adsg::dev_tables<T> t[3];
    dir_tables(h, t);
    adsg::rhs_params<T> prm;
    prm.mass = T(0);
    prm.dt = T(1);
    prm.gx = prm.gy = prm.gz = T(0);
    prm.source = static_cast<const T*>(h->vec[ADSG_SOURCE]);
    T* b = static_cast<T*>(h->source_load);
    hipError_t e = adsg::launch_fill<T>(b, h->N, T(0), h->stream);
    if (e == hipSuccess)
        e = launch_elements<T>(h, prm.source, b, t, prm);
    if (e != hipSuccess) return hip_fail(h, e, "source load vector");
    h->source_load_ok = true;
    return ADSG_OK;
}

// called before any step that may be captured: the load vector is computed eagerly
int ensure_source_load(adsg_sim* h, const adsg_rhs_form* form) {
    if (!form->use_source || h->source_load_ok) return ADSG_OK;
    if (!h->vec[ADSG_SOURCE]) return fail(h, ADSG_ESTATE, "compute_rhs: source not set");
    return h->precision == ADSG_FP32 ? source_load_impl<float>(h) : source_load_impl<double>(h);
}

template <typename T>
int rhs_impl(adsg_sim* h, const adsg_rhs_form* form, bool delta = false, adsg::rhs_fixup<T>* defer = nullptr) {
    adsg::dev_tables<T> t[3];
    dir_tables(h, t);
    adsg::rhs_params<T> prm;
    prm.mass = static_cast<T>(form->mass);
    prm.dt = static_cast<T>(form->dt);
    prm.gx = static_cast<T>(form->grad[0]);
    prm.gy = static_cast<T>(form->grad[1]);
    prm.gz = static_cast<T>(form->grad[2]);
    prm.reaction = static_cast<T>(form->reaction);
    if (form->use_kfield) {
        if (!h->vec[ADSG_KFIELD]) return fail(h, ADSG_ESTATE, "compute_rhs: k field not set");
        prm.kfield = static_cast<const T*>(h->vec[ADSG_KFIELD]);
    }
    // the source term dt (f, v) is linear in f and independent of u: it enters as dt * b,
    // b the source's load vector (ensure_source_load), and the element kernel runs without it
    if (form->use_source && (!h->vec[ADSG_SOURCE] || !h->source_load_ok))
        return fail(h, ADSG_ESTATE, "compute_rhs: source not set");
    T* rhs;
    const T* up;
    if constexpr (sizeof(T) == sizeof(float)) {
        // increment form: rhs - M u_prev, i.e. the integrand without its mass term
        if (!delta) prm.mass = static_cast<T>(form->mass - 1.0);
        rhs = static_cast<T*>(h->vec[ADSG_BUFFER]);
        up = static_cast<const T*>(h->shadow[1]);
    } else {
        rhs = static_cast<T*>(h->vec[delta ? ADSG_BUFFER : ADSG_U]);
        up = static_cast<const T*>(h->vec[ADSG_U_PREV]);
    }
    // separable integrand (no k field, no reaction): the 1D-matrix form, which writes every
    // entry, so only a source needs a start value
    const bool kron = h->rhs_kron && h->dim == 3 && !h->general && !form->use_kfield && form->reaction == 0.0;
    hipError_t e = hipSuccess;
    // zero(rhs): done by the element launch (launch_elements with overwrite), which needs no
    // fill in the one-launch tiling
    if (form->use_source)
        e = adsg::launch_scale<T>(rhs, static_cast<const T*>(h->source_load), h->N, prm.dt, h->stream);
    if (e != hipSuccess) return hip_fail(h, e, "rhs start");
    mark(h, 1);
    if (kron) {
        adsg::kron_dir kd[3];
        for (int d = 0; d < 3; ++d) {
            kd[d].M = static_cast<const double*>(h->dirs[d].kM);
            kd[d].K = static_cast<const double*>(h->dirs[d].kK);
            kd[d].n = h->dirs[d].tab.dofs;
        }
        // the form's constants as given (FP64), the FP32 increment form's mass - 1 included
        const double mass = sizeof(T) == sizeof(float) && !delta ? form->mass - 1.0 : form->mass;
        e = adsg::launch_rhs_kron<T>(h->dim, h->dirs[0].tab.p, up, rhs, kd, mass, form->dt, form->grad,
                                     form->use_source, h->stream);
    } else {
        e = launch_elements<T>(h, up, rhs, t, prm, !form->use_source, defer);
    }
    if (e != hipSuccess) return hip_fail(h, e, "rhs assembly");
    mark(h, 2);
    if constexpr (sizeof(T) == sizeof(float)) {
        h->pending = !delta;
        h->pending_dirichlet = false;
    }
    return ADSG_OK;
}

void extents(const adsg_sim* h, int& nx, int& ny, int& nz) {
    nx = h->dirs[0].tab.dofs;
    ny = h->dirs[1].tab.dofs;
    nz = h->dim == 3 ? h->dirs[2].tab.dofs : 1;
}

// the increment update u = u_prev + d (boundary 0 when dirichlet; FP32 shadow of u if given)
adsg::seg_apply increment_update(const adsg_sim* h, float* shadow, bool dirichlet) {
    adsg::seg_apply a;
    a.x_prev = static_cast<const double*>(h->vec[ADSG_U_PREV]);
    a.x_new = static_cast<double*>(h->vec[ADSG_U]);
    a.shadow = shadow;
    extents(h, a.n[0], a.n[1], a.n[2]);
    a.dirichlet = dirichlet;
    return a;
}

int zero_boundary_impl(adsg_sim* h) {
    int nx, ny, nz;
    extents(h, nx, ny, nz);
    hipError_t e;
    if (h->precision == ADSG_FP32) {
        // fixed rows: the rhs entry is 0, so the increment entry is 0 - (M_f u_prev)_b
        if (!h->pending) return fail(h, ADSG_ESTATE, "zero_boundary (FP32): call compute_rhs first");
        if (h->mixed_p)
            return fail(h, ADSG_EINVAL, "zero_boundary (FP32): the boundary residual needs one degree in every direction");
        for (int d = 0; d < h->dim; ++d)
            if (!h->dirs[d].rows)
                return fail(h, ADSG_ESTATE, "zero_boundary (FP32): slot 0 must be a factored degree-p band matrix");
        e = adsg::launch_boundary_residual(
            static_cast<float*>(h->vec[ADSG_BUFFER]), static_cast<const float*>(h->shadow[1]), h->dim,
            nx, ny, nz, h->dirs[0].tab.p, static_cast<const float*>(h->dirs[0].rows),
            static_cast<const float*>(h->dirs[1].rows),
            h->dim == 3 ? static_cast<const float*>(h->dirs[2].rows) : nullptr, h->stream);
        h->pending_dirichlet = true;
    } else {
        e = adsg::launch_zero_boundary<double>(h->dim, static_cast<double*>(h->vec[ADSG_U]), nx, ny,
                                               nz, h->stream);
    }
    return e == hipSuccess ? ADSG_OK : hip_fail(h, e, "zero boundary");
}

int solve_impl(adsg_sim* h, const int* slots, const adsg::rhs_fixup<double>* fix64 = nullptr,
               const adsg::rhs_fixup<float>* fix32 = nullptr) {
    if (h->precision == ADSG_FP64)
        return sweep_all<double>(h, static_cast<double*>(h->vec[ADSG_U]), slots, 0, fix64);
    // FP32: u = u_prev + A^-1 (rhs - M u_prev); valid when A is the Gram matrix of every
    // direction (slot 0), which is what makes buffer = rhs - A u_prev
    if (!h->pending) return fail(h, ADSG_ESTATE, "ads_solve (FP32): call compute_rhs first");
    if (slots)
        for (int d = 0; d < h->dim; ++d)
            if (slots[d] != 0)
                return fail(h, ADSG_EINVAL,
                            "ads_solve (FP32): increments are defined against the Gram matrices (slot 0); "
                            "use adsg_step_delta for other matrices");
    adsg::seg_apply ap = increment_update(h, static_cast<float*>(h->shadow[0]), h->pending_dirichlet);
    bool applied = false;
    int rc = sweep_all<float>(h, static_cast<float*>(h->vec[ADSG_BUFFER]), nullptr, 0, fix32, &ap, &applied);
    if (rc) return rc;
    int nx, ny, nz;
    extents(h, nx, ny, nz);
    hipError_t e = applied ? hipSuccess
                           : adsg::launch_apply_increment(
                                 static_cast<const double*>(h->vec[ADSG_U_PREV]),
                                 static_cast<const float*>(h->vec[ADSG_BUFFER]), static_cast<double*>(h->vec[ADSG_U]),
                                 static_cast<float*>(h->shadow[0]), h->dim, nx, ny, nz, h->pending_dirichlet, h->stream);
    if (e != hipSuccess) return hip_fail(h, e, "apply increment");
    h->pending = false;
    h->pending_dirichlet = false;
    return ADSG_OK;
}

void swap_state(adsg_sim* h) {
    std::swap(h->vec[ADSG_U], h->vec[ADSG_U_PREV]);
    std::swap(h->shadow[0], h->shadow[1]);
}

int rhs_dispatch(adsg_sim* h, const adsg_rhs_form* form) {
    return h->precision == ADSG_FP32 ? rhs_impl<float>(h, form) : rhs_impl<double>(h, form);
}

// one explicit step as validation.hpp:73-81 (before_step swap, compute_rhs, solve)
// (the element loop's one-launch fix-up and, in FP64, the boundary zeroing are left to
// sweep_all: rhs_fixup)
template <typename T>
void fixup_extents(const adsg_sim* h, adsg::rhs_fixup<T>& fx) {
    extents(h, fx.n[0], fx.n[1], fx.n[2]);
}


int one_step(adsg_sim* h, const adsg_rhs_form* form, int flags) {
    swap_state(h);
    const bool dir = flags & ADSG_STEP_DIRICHLET;
    int rc;
    if (h->precision == ADSG_FP64) {
        adsg::rhs_fixup<double> fx;
        if ((rc = rhs_impl<double>(h, form, false, &fx))) return rc;
        fixup_extents(h, fx);
        fx.zero_boundary = dir;
        mark(h, 3);
        return solve_impl(h, nullptr, &fx, nullptr);
    }
    // FP32: the boundary residual reads the complete increment, so with Dirichlet rows the
    // fix-up runs in the element phase
    adsg::rhs_fixup<float> fx;
    if ((rc = rhs_impl<float>(h, form, false, dir ? nullptr : &fx))) return rc;
    if (dir && (rc = zero_boundary_impl(h))) return rc;
    mark(h, 3);
    return solve_impl(h, nullptr, nullptr, &fx);
}

// one delta step: u = u_prev + A^-1 r(u_prev)
int one_delta_step(adsg_sim* h, const adsg_rhs_form* form, const int* slots, int flags) {
    swap_state(h);
    int nx, ny, nz;
    extents(h, nx, ny, nz);
    const bool dir = flags & ADSG_STEP_DIRICHLET;
    int rc;
    hipError_t e;
    if (h->precision == ADSG_FP32) {
        adsg::rhs_fixup<float> fx;
        if ((rc = rhs_impl<float>(h, form, true, &fx))) return rc;
        float* r = static_cast<float*>(h->vec[ADSG_BUFFER]);
        fixup_extents(h, fx);
        fx.zero_boundary = dir;
        mark(h, 3);
        adsg::seg_apply ap = increment_update(h, static_cast<float*>(h->shadow[0]), dir);
        bool applied = false;
        if ((rc = sweep_all<float>(h, r, slots, (flags & ADSG_STEP_REFINE) ? 2 : 0, &fx, &ap, &applied))) return rc;
        e = applied ? hipSuccess
                    : adsg::launch_apply_increment(static_cast<const double*>(h->vec[ADSG_U_PREV]), r,
                                                   static_cast<double*>(h->vec[ADSG_U]),
                                                   static_cast<float*>(h->shadow[0]), h->dim, nx, ny, nz, dir,
                                                   h->stream);
    } else {
        adsg::rhs_fixup<double> fx;
        if ((rc = rhs_impl<double>(h, form, true, &fx))) return rc;
        double* r = static_cast<double*>(h->vec[ADSG_BUFFER]);
        fixup_extents(h, fx);
        fx.zero_boundary = dir;
        mark(h, 3);
        adsg::seg_apply ap = increment_update(h, nullptr, dir);
        bool applied = false;
        if ((rc = sweep_all<double>(h, r, slots, 0, &fx, &ap, &applied))) return rc;
        e = applied ? hipSuccess
                    : adsg::launch_add_increment(static_cast<const double*>(h->vec[ADSG_U_PREV]), r,
                                                 static_cast<double*>(h->vec[ADSG_U]), h->dim, nx, ny, nz, dir,
                                                 h->stream);
    }
    return e == hipSuccess ? ADSG_OK : hip_fail(h, e, "apply increment");
}

// R^T v = R v for the linear part of form (no source, no reaction), into out (zeroed first)
int apply_linear_form(adsg_sim* h, const adsg_rhs_form* form, const double* v, double* out) {
    adsg::dev_tables<double> t[3];
    dir_tables(h, t);
    adsg::rhs_params<double> prm;
    prm.mass = form->mass;
    prm.dt = form->dt;
    prm.gx = form->grad[0];
    prm.gy = form->grad[1];
    prm.gz = form->grad[2];
    prm.reaction = 0.0;
    if (form->use_kfield) {
        if (!h->vec[ADSG_KFIELD]) return fail(h, ADSG_ESTATE, "adjoint: k field not set");
        prm.kfield = static_cast<const double*>(h->vec[ADSG_KFIELD]);
    }
    hipError_t e;
    if (h->rhs_kron && h->dim == 3 && !h->general && !form->use_kfield) {  // separable: the 1D-matrix form
        mark(h, 1);
        adsg::kron_dir kd[3];
        for (int d = 0; d < 3; ++d) {
            kd[d].M = static_cast<const double*>(h->dirs[d].kM);
            kd[d].K = static_cast<const double*>(h->dirs[d].kK);
            kd[d].n = h->dirs[d].tab.dofs;
        }
        const double grad[3] = {prm.gx, prm.gy, prm.gz};
        e = adsg::launch_rhs_kron<double>(h->dim, h->dirs[0].tab.p, v, out, kd, prm.mass, prm.dt, grad, false,
                                          h->stream);
        if (e != hipSuccess) return hip_fail(h, e, "rhs assembly");
        mark(h, 2);
        return ADSG_OK;
    }
    if ((e = adsg::launch_fill<double>(out, h->N, 0.0, h->stream)) != hipSuccess) return hip_fail(h, e, "zero");
    mark(h, 1);
    e = launch_elements<double>(h, v, out, t, prm);
    if (e != hipSuccess) return hip_fail(h, e, "rhs assembly");
    mark(h, 2);
    return ADSG_OK;
}

// out = M v with M the Gram matrix of the whole basis (no fixed rows): the element loop of
// the form {mass 1, dt 0}, or its 1D-matrix form when that path is selected
int gram_apply(adsg_sim* h, const double* v, double* out) {
    hipError_t e;
    if (h->rhs_kron && h->dim == 3 && !h->general) {
        adsg::kron_dir kd[3];
        for (int d = 0; d < 3; ++d) {
            kd[d].M = static_cast<const double*>(h->dirs[d].kM);
            kd[d].K = static_cast<const double*>(h->dirs[d].kK);
            kd[d].n = h->dirs[d].tab.dofs;
        }
        const double none[3] = {0.0, 0.0, 0.0};
        e = adsg::launch_rhs_kron<double>(h->dim, h->dirs[0].tab.p, v, out, kd, 1.0, 0.0, none, false, h->stream);
    } else {
        adsg::dev_tables<double> t[3];
        dir_tables(h, t);
        adsg::rhs_params<double> prm;
        prm.mass = 1.0;
        prm.dt = 0.0;
        e = adsg::launch_fill<double>(out, h->N, 0.0, h->stream);
        if (e == hipSuccess)
            e = launch_elements<double>(h, v, out, t, prm);
    }
    return e == hipSuccess ? ADSG_OK : hip_fail(h, e, "Gram product");
}

// dL/df += dt M w for forms with a source, when the caller has set ADSG_GRAD_F; scratch is
// overwritten
int accumulate_source_gradient(adsg_sim* h, const adsg_rhs_form* form, const double* w, double* scratch) {
    if (!form->use_source || !h->vec[ADSG_GRAD_F]) return ADSG_OK;
    int rc = gram_apply(h, w, scratch);
    if (rc) return rc;
    hipError_t e = adsg::launch_axpy<double>(static_cast<double*>(h->vec[ADSG_GRAD_F]), scratch, h->N, form->dt,
                                             h->stream);
    return e == hipSuccess ? ADSG_OK : hip_fail(h, e, "source gradient");
}

// 2D with a k field in the one-launch tiling: the element loop for R(k) w (into lam) and
// the k-field gradient (into ADSG_GRAD_K) in one launch over all tiles and one fix-up launch
// for both (launch_adjoint2); false when that path does not apply (nothing launched)
bool adjoint_pair_2d(adsg_sim* h, const adsg_rhs_form* form, const double* w, const double* un, double* lam,
                     int& rc) {
    rc = ADSG_OK;
    if (h->dim != 2 || !form->use_kfield || h->general || !h->rhs_tiling || !h->spill || !h->vec[ADSG_KFIELD])
        return false;
    if (!h->spill2) {
        const int el[3] = {h->dirs[0].tab.elements, h->dirs[1].tab.elements, 1};
        const size_t bytes = static_cast<size_t>(adsg::rhs_spill_slots(2, h->dirs[0].tab.p, el)) * sizeof(double);
        if (hipMalloc(&h->spill2, bytes) != hipSuccess) {
            h->spill2 = nullptr;
            return false;  // the separate passes need no second array
        }
    }
    adsg::dev_tables<double> t[3];
    dir_tables(h, t);
    adsg::rhs_params<double> pr, pk;
    pr.mass = form->mass;
    pr.dt = pk.dt = form->dt;
    pr.gx = pk.gx = form->grad[0];
    pr.gy = pk.gy = form->grad[1];
    pr.gz = pk.gz = form->grad[2];
    pr.reaction = 0.0;
    pr.kfield = static_cast<const double*>(h->vec[ADSG_KFIELD]);
    mark(h, 1);
    const hipError_t e = adsg::launch_adjoint2<double>(
        h->dirs[0].tab.p, w, un, lam, static_cast<double*>(h->vec[ADSG_GRAD_K]), t, pr, pk,
        static_cast<double*>(h->spill), static_cast<double*>(h->spill2), h->stream);
    mark(h, 2);
    if (e != hipSuccess) rc = hip_fail(h, e, "adjoint element passes");
    return true;
}

// One step of the adjoint of one_step (FP64): see adsg_adjoint_step in adsgpu.h. u_prev:
// the forward state of the step (default: ADSG_U_PREV); ADSG_U_PREV is overwritten either way
// when the form has a source.
int one_adjoint_step(adsg_sim* h, const adsg_rhs_form* form, int flags, const double* u_prev = nullptr) {
    int nx, ny, nz;
    extents(h, nx, ny, nz);
    const bool dir = flags & ADSG_STEP_DIRICHLET;
    double* lam = static_cast<double*>(h->vec[ADSG_U]);
    double* w = static_cast<double*>(h->vec[ADSG_BUFFER]);
    const double* un = u_prev ? u_prev : static_cast<const double*>(h->vec[ADSG_U_PREV]);
    // w = Z M^-T lam: the first transposed sweep reads lam and writes w, the zeroing Z by the
    // last one as it stores its lines
    adsg::seg_apply zs;
    extents(h, zs.n[0], zs.n[1], zs.n[2]);
    zs.dirichlet = dir;
    bool zeroed = false;
    int rc = sweep_all<double>(h, w, nullptr, 1, nullptr, dir ? &zs : nullptr, &zeroed, lam);
    if (rc) return rc;
    hipError_t e;
    if (dir && !zeroed && (e = adsg::launch_zero_boundary<double>(h->dim, w, nx, ny, nz, h->stream)) != hipSuccess)
        return hip_fail(h, e, "zero boundary");
    mark(h, 3);
    int rp = ADSG_OK;
    const bool paired = adjoint_pair_2d(h, form, w, un, lam, rp);
    if (rp) return rp;
    if (form->use_kfield && !paired) {
        adsg::dev_tables<double> t[3];
        dir_tables(h, t);
        adsg::rhs_params<double> prm;
        prm.dt = form->dt;
        prm.gx = form->grad[0];
        prm.gy = form->grad[1];
        prm.gz = form->grad[2];
        prm.spill = h->rhs_tiling ? static_cast<double*>(h->spill) : nullptr;  // one launch and its fix-up
        e = adsg::launch_kgrad<double>(h->dim, h->dirs[0].tab.p, un, w,
                                       static_cast<double*>(h->vec[ADSG_GRAD_K]), t, prm, h->stream);
        if (e != hipSuccess) return hip_fail(h, e, "k-field gradient");
    }
    // the source enters as dt M f: dL/df += dt M w (u_prev, read above, is free now)
    int rs = accumulate_source_gradient(h, form, w, static_cast<double*>(h->vec[ADSG_U_PREV]));
    if (rs) return rs;
    return paired ? ADSG_OK : apply_linear_form(h, form, w, lam);
}

// One step of the adjoint of one_delta_step (FP64) for a linear form r(u) = R u + c:
//   forward  u' = Z (u + A^-1 Z r(u))           (Z: zero the boundary rows when dirichlet)
//   adjoint  lam := mu + R^T Z A^-T mu, mu = Z lam
// R is symmetric (mass*M - dt*K, with K the k-weighted stiffness), so R^T v is the integrand
// of form without its source term, assembled from v. u_prev is used as scratch.
int one_adjoint_delta_step(adsg_sim* h, const adsg_rhs_form* form, const int* slots, int flags) {
    int nx, ny, nz;
    extents(h, nx, ny, nz);
    const bool dir = flags & ADSG_STEP_DIRICHLET;
    double* lam = static_cast<double*>(h->vec[ADSG_U]);
    double* v = static_cast<double*>(h->vec[ADSG_BUFFER]);
    double* w = static_cast<double*>(h->vec[ADSG_U_PREV]);
    hipError_t e;
    if (dir && (e = adsg::launch_zero_boundary<double>(h->dim, lam, nx, ny, nz, h->stream)) != hipSuccess)
        return hip_fail(h, e, "zero boundary");
    HIP_TRY(h, hipMemcpyAsync(v, lam, sizeof(double) * h->N, hipMemcpyDeviceToDevice, h->stream));
    mark(h, 3);
    adsg::seg_apply zs;  // the zeroing of v by the last transposed sweep as it stores its lines
    extents(h, zs.n[0], zs.n[1], zs.n[2]);
    zs.dirichlet = dir;
    bool zeroed = false;
    int rc = sweep_all<double>(h, v, slots, 1, nullptr, dir ? &zs : nullptr, &zeroed);
    if (rc) return rc;
    if (dir && !zeroed && (e = adsg::launch_zero_boundary<double>(h->dim, v, nx, ny, nz, h->stream)) != hipSuccess)
        return hip_fail(h, e, "zero boundary");
    // r = R u + dt M f: dL/df += dt M v
    if ((rc = accumulate_source_gradient(h, form, v, w))) return rc;
    if ((rc = apply_linear_form(h, form, v, w))) return rc;
    e = adsg::launch_add_increment(lam, w, lam, h->dim, nx, ny, nz, false, h->stream);
    return e == hipSuccess ? ADSG_OK : hip_fail(h, e, "adjoint update");
}

std::string form_key(const adsg_rhs_form* f) {
    char buf[256];
    std::snprintf(buf, sizeof buf, "%a,%a,%a,%a,%a,%a,%d,%d", f->mass, f->dt, f->grad[0], f->grad[1],
                  f->grad[2], f->reaction, f->use_kfield, f->use_source);
    return buf;
}

// Factor every line's matrix of a per-line slot (FP64, dgbtrf) and upload the factors
// interleaved by line in the order of geometry(h, dir). The reference numbers the lines of
// direction d by the remaining indices in cyclic order (d+1 fastest, solver.hpp:100-150),
// which differs from geometry's order only for y lines in 3D (z fastest there, x here).
int prepare_line_matrices(adsg_sim* h, int dir, dev_slot& s) {
    const int n = h->dirs[dir].tab.dofs, kl = s.kl, ku = s.ku, ld = 2 * kl + ku + 1;
    const long L = s.lines;
    std::vector<adsg::line_factors> lf(L);
    int first_info = 0;
    bool pivoted = false;
    int kue = ku;
    for (long r = 0; r < L; ++r) {
        std::vector<double> ab(s.ab.begin() + r * ld * n, s.ab.begin() + (r + 1) * ld * n);
        int info = adsg::factor_for_lines(ab, n, kl, ku, lf[r]);
        if (info && !first_info) first_info = info;
        pivoted = pivoted || lf[r].pivoted;
    }
    if (first_info) return first_info;
    if (pivoted) kue = std::min(kl + ku, std::max(n - 1, 0));
    const long nx = h->dirs[0].tab.dofs, nz = h->dim == 3 ? h->dirs[2].tab.dofs : 1;
    auto ref_line = [&](long l) -> long {
        if (h->dim == 3 && dir == 1) return (l / nx) + nz * (l % nx);  // l = x + nx*z -> z + nz*x
        return l;
    };
    std::vector<double> lower(static_cast<size_t>(n) * std::max(kl, 1) * L, 0.0);
    std::vector<double> upper(static_cast<size_t>(n) * std::max(kue, 1) * L, 0.0);
    std::vector<double> rdiag(static_cast<size_t>(n) * L, 0.0);
    std::vector<int> poff(static_cast<size_t>(n) * L, 0);
    for (long l = 0; l < L; ++l) {
        const adsg::line_factors& f = lf[ref_line(l)];
        for (int j = 0; j < n; ++j) {
            for (int t = 0; t < kl; ++t) lower[(static_cast<size_t>(j) * kl + t) * L + l] = f.lower[static_cast<size_t>(j) * kl + t];
            for (int t = 0; t < f.ku_eff; ++t)
                upper[(static_cast<size_t>(j) * kue + t) * L + l] = f.upper[static_cast<size_t>(j) * f.ku_eff + t];
            rdiag[static_cast<size_t>(j) * L + l] = f.rdiag[j];
            poff[static_cast<size_t>(j) * L + l] = f.poff[j];
        }
    }
    free_slot(s, h->in_arena);
    int rc;
    if (h->precision == ADSG_FP32) {
        if ((rc = upload<float>(h, &s.lower, lower))) return rc;
        if ((rc = upload<float>(h, &s.upper, upper))) return rc;
        if ((rc = upload<float>(h, &s.rdiag, rdiag))) return rc;
    } else {
        if ((rc = upload<double>(h, &s.lower, lower))) return rc;
        if ((rc = upload<double>(h, &s.upper, upper))) return rc;
        if ((rc = upload<double>(h, &s.rdiag, rdiag))) return rc;
    }
    if ((rc = upload_bytes(h, reinterpret_cast<void**>(&s.poff), poff.data(), sizeof(int) * poff.size())))
        return rc;
    s.lf = adsg::line_factors{};
    s.lf.n = n;
    s.lf.kl = kl;
    s.lf.ku = ku;
    s.lf.ku_eff = kue;
    s.lf.pivoted = pivoted;
    s.seg = adsg::line_segments{};
    s.factored = true;
    return ADSG_OK;
}

enum step_kind { EXPLICIT_STEP, DELTA_STEP, ADJOINT_DELTA_STEP };

int step_of(adsg_sim* h, step_kind kind, const adsg_rhs_form* form, const int* slots, int flags) {
    switch (kind) {
        case DELTA_STEP:
            return one_delta_step(h, form, slots, flags);
        case ADJOINT_DELTA_STEP:
            return one_adjoint_delta_step(h, form, slots, flags);
        default:
            return one_step(h, form, flags);
    }
}

int get_graph(adsg_sim* h, step_kind kind, const adsg_rhs_form* form, int flags, int steps,
              const int* slots, hipGraphExec_t* out) {
    std::string k = form_key(form) + "|" + std::to_string(static_cast<int>(kind));
    if (slots) k += " " + std::to_string(slots[0]) + " " + std::to_string(slots[1]) + " " +
                    std::to_string(h->dim == 3 ? slots[2] : 0);
    auto key = std::make_tuple(k, flags, h->vec[ADSG_U], steps);
    auto it = h->graphs.find(key);
    if (it != h->graphs.end()) {
        *out = it->second;
        return ADSG_OK;
    }
    void* u0 = h->vec[ADSG_U];
    void* p0 = h->vec[ADSG_U_PREV];
    void* s0 = h->shadow[0];
    void* s1 = h->shadow[1];
    HIP_TRY(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    int rc = ADSG_OK;
    for (int s = 0; s < steps && rc == ADSG_OK; ++s)
        rc = step_of(h, kind, form, slots, flags);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(h->stream, &g);
    h->vec[ADSG_U] = u0;  // capture only recorded the swaps; restore the logical state
    h->vec[ADSG_U_PREV] = p0;
    h->shadow[0] = s0;
    h->shadow[1] = s1;
    h->pending = false;
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return hip_fail(h, e, "hipStreamEndCapture");
    hipGraphExec_t ex = nullptr;
    e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return hip_fail(h, e, "hipGraphInstantiate");
    h->graphs[key] = ex;
    *out = ex;
    return ADSG_OK;
}

// explicit and delta steps swap u and u_prev once per step, so a two-step graph returns
// the pointers to where they were; the adjoint step updates u in place
int run_graphs(adsg_sim* h, step_kind kind, int nsteps, const adsg_rhs_form* form, int flags,
               const int* slots) {
    hipGraphExec_t g2 = nullptr, g1 = nullptr;
    int rc;
    if (kind != ADJOINT_DELTA_STEP && (rc = ensure_source_load(h, form))) return rc;  // not captured
    if (nsteps >= 2) {
        if ((rc = get_graph(h, kind, form, flags, 2, slots, &g2))) return rc;
        for (int i = 0; i < nsteps / 2; ++i) HIP_TRY(h, hipGraphLaunch(g2, h->stream));
    }
    if (nsteps % 2) {
        if ((rc = get_graph(h, kind, form, flags, 1, slots, &g1))) return rc;
        HIP_TRY(h, hipGraphLaunch(g1, h->stream));
        if (kind != ADJOINT_DELTA_STEP) swap_state(h);
    }
    return ADSG_OK;
}

}  // namespace

// the handle of adsg_create; with a shared stream (not owned) and without state vectors for
// the problems of a batch, whose vectors live in the batch's arenas
int adsg::create_sim(const adsg_config* cfg, hipStream_t shared, bool vectors, adsg_sim** out) {
    if (!cfg || !out) {
        g_create_error = "adsg_create: null argument";
        return ADSG_EINVAL;
    }
    *out = nullptr;
    if (cfg->dim != 2 && cfg->dim != 3) {
        g_create_error = "adsg_create: dim must be 2 or 3";
        return ADSG_EINVAL;
    }
    if (cfg->precision != ADSG_FP64 && cfg->precision != ADSG_FP32) {
        g_create_error = "adsg_create: precision must be ADSG_FP64 or ADSG_FP32";
        return ADSG_EINVAL;
    }
    auto* h = new adsg_sim;
    h->dim = cfg->dim;
    h->precision = cfg->precision;
    h->step_count = cfg->step_count;
    h->dt = cfg->dt;
    const adsg_dim_config* dc[3] = {&cfg->x, &cfg->y, &cfg->z};
    h->N = 1;
    for (int d = 0; d < h->dim; ++d) {
        std::string err;
        if (!adsg::build_direction(h->dirs[d].tab, dc[d]->p, dc[d]->elements, dc[d]->a, dc[d]->b,
                                   dc[d]->quad_order, dc[d]->repeated_nodes, err)) {
            g_create_error = "adsg_create: direction " + std::to_string(d) + ": " + err;
            delete h;
            return ADSG_EINVAL;
        }
        // configurations outside the sum-factorised kernels' take the general element loop
        const int nq = dc[d]->quad_order > 0 ? dc[d]->quad_order : dc[d]->p + 1;
        if (nq != dc[d]->p + 1 || dc[d]->repeated_nodes != 0 || dc[d]->p != dc[0]->p) h->general = true;
        if (dc[d]->p != dc[0]->p) h->mixed_p = true;
        h->N *= h->dirs[d].tab.dofs;
        dev_slot s0;
        s0.set = true;
        s0.kl = s0.ku = dc[d]->p;
        s0.ab = adsg::band_form(h->dirs[d].tab, 1.0, 0.0);  // gram_matrix_1d (dimension.cpp:17)
        h->dirs[d].slots.push_back(std::move(s0));
    }
    // default element-kernel variant (adsg_autotune refines it): the one measured fastest for
    // the degree and precision. p = 3: FP64 2 (variants 0 and 1 need more registers than allow
    // two waves per SIMD), FP32 5 (element columns on the matrix cores); p <= 2: 4 from 2^21
    // elements (enough tiles of 512 elements to fill the device twice over), else 1. 3D
    // handles start with the one-launch tiling (ADSG_TUNE_RHS_TILING), which was faster for
    // every variant measured (p = 2 at 128^3, p = 3 at 192^3, both precisions); so do 2D
    // handles of at most 2048 tiles of 16 x 16 elements, whose colour launches are too small to
    // fill the device (128^2 p = 2 with a k field: 0.046 -> 0.029 ms per step). Larger 2D
    // meshes keep the colour launches, as the batches do (with thousands of tiles per launch
    // the fix-up pass costs more than the colour tails it removes).
    {
        long E = 1;
        for (int d = 0; d < h->dim; ++d) E *= h->dirs[d].tab.elements;
        const int p = dc[0]->p;
        h->rhs_variant = p == 3 ? (h->precision == ADSG_FP32 ? 5 : 2) : E >= (1L << 21) ? 4 : 1;
        h->rhs_col_shape = 1;  // variant 5: two columns per wave, one chunk per workgroup
        const long tiles16 = static_cast<long>((h->dirs[0].tab.elements + 15) / 16) * ((h->dirs[1].tab.elements + 15) / 16);
        h->rhs_tiling = h->general ? 0 : h->dim == 3 || tiles16 <= 2048 ? 1 : 0;
    }
    if (hipGetDevice(&h->device) != hipSuccess) {
        g_create_error = "adsg_create: hipGetDevice failed (no HIP device?)";
        delete h;
        return ADSG_EHIP;
    }
    if (shared) {
        h->stream = shared;
        h->owns_stream = false;
    } else if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        g_create_error = "adsg_create: hipStreamCreate failed (no HIP device?)";
        delete h;
        return ADSG_EHIP;
    }
    for (int d = 0; d < h->dim; ++d) {
        if (upload_tables(h, h->dirs[d])) {
            g_create_error = "adsg_create: " + h->err;
            adsg_destroy(h);
            return ADSG_EHIP;
        }
    }
    for (int v = ADSG_U; v <= ADSG_BUFFER && vectors; ++v) {
        if (hipMalloc(&h->vec[v], h->N * elt_of(h, v)) != hipSuccess) {
            g_create_error = "adsg_create: device allocation of the state vectors failed";
            adsg_destroy(h);
            return ADSG_ENOMEM;
        }
        if (hipMemsetAsync(h->vec[v], 0, h->N * elt_of(h, v), h->stream) != hipSuccess) {
            g_create_error = "adsg_create: clearing the state vectors failed";
            adsg_destroy(h);
            return ADSG_EHIP;
        }
    }
    if (h->precision == ADSG_FP32 && vectors) {
        for (auto& sp : h->shadow) {
            if (hipMalloc(&sp, h->N * sizeof(float)) != hipSuccess ||
                hipMemsetAsync(sp, 0, h->N * sizeof(float), h->stream) != hipSuccess) {
                g_create_error = "adsg_create: device allocation of the FP32 shadows failed";
                adsg_destroy(h);
                return ADSG_ENOMEM;
            }
        }
    }
    // the one-launch tiling's spill array (without it the handle falls back to colour launches)
    if (h->rhs_tiling && ensure_spill(h) != ADSG_OK) h->rhs_tiling = 0;
    (void)hipStreamSynchronize(h->stream);
    *out = h;
    return ADSG_OK;
}

namespace {

long quad_point_count(const adsg_sim* h) {
    long q = 1;
    for (int d = 0; d < h->dim; ++d) q *= static_cast<long>(h->dirs[d].tab.elements) * h->dirs[d].tab.nq;
    return q;
}

int project_impl(adsg_sim* h, int which, const double* fq_dev) {
    if (h->general) return refuse_general(h, "project");
    if (!h->vec[which]) {
        if (hipMalloc(&h->vec[which], h->N * sizeof(double)) != hipSuccess)
            return fail(h, ADSG_ENOMEM, "project: device allocation failed");
        HIP_TRY(h, hipMemsetAsync(h->vec[which], 0, h->N * sizeof(double), h->stream));
        drop_graphs(h);
    }
    if (which == ADSG_SOURCE) h->source_load_ok = false;
    double* v = static_cast<double*>(h->vec[which]);
    hipError_t e;
    if (h->precision == ADSG_FP32) {
        adsg::dev_tables<float> t[3];
        dir_tables(h, t);
        e = adsg::launch_project<float>(h->dim, h->dirs[0].tab.p, fq_dev, v, t, h->stream);
    } else {
        adsg::dev_tables<double> t[3];
        dir_tables(h, t);
        e = adsg::launch_project<double>(h->dim, h->dirs[0].tab.p, fq_dev, v, t, h->stream);
    }
    if (e != hipSuccess) return hip_fail(h, e, "project");
    if (h->precision == ADSG_FP32 && (which == ADSG_U || which == ADSG_U_PREV)) {
        e = adsg::launch_convert_d2f(v, static_cast<float*>(h->shadow[which]), h->N, h->stream);
        if (e != hipSuccess) return hip_fail(h, e, "project: shadow");
    }
    return ADSG_OK;
}

bool projectable(const adsg_sim* h, int which) {
    return which >= ADSG_U && which <= ADSG_GRAD_F && elt_of(h, which) == sizeof(double);
}

// NaN / Inf count of an FP64 vector of the handle (u and u_prev are FP64 in both
// precisions); ADSG_ENONFINITE when there is any, with the count in the message
int check_finite_impl(adsg_sim* h, int which, const char* where) {
    if (!h->vec[which]) return fail(h, ADSG_ESTATE, std::string(where) + ": vector not allocated");
    if (elt_of(h, which) != sizeof(double))
        return fail(h, ADSG_EINVAL, std::string(where) + ": FP64 vectors only (u, u_prev, gradients; any vector of FP64 handles)");
    if (!h->nonfinite && hipMalloc(reinterpret_cast<void**>(&h->nonfinite), sizeof(unsigned)) != hipSuccess)
        return fail(h, ADSG_ENOMEM, std::string(where) + ": no memory for the count");
    unsigned bad = 0;
    HIP_TRY(h, hipMemsetAsync(h->nonfinite, 0, sizeof(unsigned), h->stream));  // eager, never captured
    hipError_t e = adsg::launch_count_nonfinite(static_cast<const double*>(h->vec[which]), h->N, h->nonfinite, h->stream);
    if (e != hipSuccess) return hip_fail(h, e, where);
    HIP_TRY(h, hipMemcpyAsync(&bad, h->nonfinite, sizeof(unsigned), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    if (bad)
        return fail(h, ADSG_ENONFINITE, std::string(where) + ": " + std::to_string(bad) + " of " +
                                            std::to_string(h->N) + " coefficients are NaN or Inf");
    return ADSG_OK;
}

}  // namespace

extern "C" {

const char* adsg_last_error(const adsg_sim* h) { return h ? h->err.c_str() : g_create_error.c_str(); }

int adsg_create(const adsg_config* cfg, adsg_sim** out) { return adsg::create_sim(cfg, nullptr, true, out); }

void adsg_destroy(adsg_sim* h) {
    if (!h) return;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    drop_graphs(h);
    for (auto& d : h->dirs) {
        for (auto& s : d.slots) free_slot(s, h->in_arena);
        for (void** p : {&d.val, &d.der, &d.wj, &d.bval, &d.bder, &d.rows, &d.kM, &d.kK})
            if (*p && !h->in_arena) (void)hipFree(*p);
        if (d.first && !h->in_arena) (void)hipFree(d.first);
    }
    for (auto& v : h->vec)
        if (v) (void)hipFree(v);
    for (auto& v : h->shadow)
        if (v) (void)hipFree(v);
    if (h->norm_scratch) (void)hipFree(h->norm_scratch);
    if (h->nonfinite) (void)hipFree(h->nonfinite);
    if (h->source_load) (void)hipFree(h->source_load);
    if (h->spill) (void)hipFree(h->spill);
    if (h->spill2) (void)hipFree(h->spill2);
    if (h->ckpt) (void)hipFree(h->ckpt);
    if (h->ck_seg) (void)hipFree(h->ck_seg);
    if (h->stream && h->owns_stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

int adsg_dofs(const adsg_sim* h, int dir) { return valid_dir(h, dir) ? h->dirs[dir].tab.dofs : ADSG_EINVAL; }
long adsg_size(const adsg_sim* h) { return h ? h->N : ADSG_EINVAL; }
int adsg_precision(const adsg_sim* h) { return h ? h->precision : ADSG_EINVAL; }

int adsg_fix_dof(adsg_sim* h, int dir, int k) {
    if (!valid_dir(h, dir)) return fail(h, ADSG_EINVAL, "fix_dof: bad direction");
    dev_dir& d = h->dirs[dir];
    if (k < 0 || k >= d.tab.dofs) return fail(h, ADSG_EINVAL, "fix_dof: dof out of range");
    dev_slot& s = d.slots[0];
    if (s.kl != d.tab.p || s.ku != d.tab.p)
        return fail(h, ADSG_EINVAL, "fix_dof: slot 0 is a user band matrix");
    adsg::band_fix_dof(s.ab, d.tab.dofs, d.tab.p, k);
    s.factored = false;
    h->prepared = false;
    drop_graphs(h);
    return ADSG_OK;
}

int adsg_set_matrix(adsg_sim* h, int dir, int slot, double mass, double stiff) {
    if (!valid_dir(h, dir)) return fail(h, ADSG_EINVAL, "set_matrix: bad direction");
    if (slot < 1 || slot > 64) return fail(h, ADSG_EINVAL, "set_matrix: slot must be in 1..64");
    dev_dir& d = h->dirs[dir];
    if (static_cast<int>(d.slots.size()) <= slot) d.slots.resize(slot + 1);
    dev_slot& s = d.slots[slot];
    free_slot(s, h->in_arena);
    s = dev_slot{};
    s.set = true;
    s.kl = s.ku = d.tab.p;
    s.ab = adsg::band_form(d.tab, mass, stiff);
    h->prepared = false;
    drop_graphs(h);
    return ADSG_OK;
}

int adsg_set_band_matrix(adsg_sim* h, int dir, int slot, int kl, int ku, const double* ab) {
    if (!valid_dir(h, dir) || !ab) return fail(h, ADSG_EINVAL, "set_band_matrix: bad argument");
    if (slot < 0 || slot > 64) return fail(h, ADSG_EINVAL, "set_band_matrix: slot must be in 0..64");
    if (kl < 1 || kl > 3 || ku < 1 || ku > 3)
        return fail(h, ADSG_EINVAL, "set_band_matrix: kl and ku must be in 1..3");
    dev_dir& d = h->dirs[dir];
    if (static_cast<int>(d.slots.size()) <= slot) d.slots.resize(slot + 1);
    dev_slot& s = d.slots[slot];
    free_slot(s, h->in_arena);
    s = dev_slot{};
    s.set = true;
    s.kl = kl;
    s.ku = ku;
    const size_t n = static_cast<size_t>(2 * kl + ku + 1) * d.tab.dofs;
    s.ab.assign(ab, ab + n);
    h->prepared = false;
    drop_graphs(h);
    return ADSG_OK;
}

int adsg_set_line_matrices(adsg_sim* h, int dir, int slot, int kl, int ku, const double* ab) {
    if (!valid_dir(h, dir) || !ab) return fail(h, ADSG_EINVAL, "set_line_matrices: bad argument");
    if (slot < 1 || slot > 64) return fail(h, ADSG_EINVAL, "set_line_matrices: slot must be in 1..64");
    if (kl < 0 || ku < 0 || kl > 8 || ku > 8)
        return fail(h, ADSG_EINVAL, "set_line_matrices: kl and ku must be in 0..8");
    dev_dir& d = h->dirs[dir];
    if (static_cast<int>(d.slots.size()) <= slot) d.slots.resize(slot + 1);
    dev_slot& s = d.slots[slot];
    free_slot(s, h->in_arena);
    s = dev_slot{};
    s.set = true;
    s.kl = kl;
    s.ku = ku;
    s.lines = h->N / d.tab.dofs;
    const size_t total = static_cast<size_t>(2 * kl + ku + 1) * d.tab.dofs * s.lines;
    s.ab.assign(ab, ab + total);
    h->prepared = false;
    drop_graphs(h);
    return ADSG_OK;
}

int adsg_prepare_matrices(adsg_sim* h) {
    if (!h) return ADSG_EINVAL;
    int first_info = 0;
    for (int dd = 0; dd < h->dim; ++dd) {
        dev_dir& d = h->dirs[dd];
        for (auto& s : d.slots) {
            if (!s.set || s.factored) continue;
            if (s.lines > 0) {
                int rc = prepare_line_matrices(h, dd, s);
                if (rc > 0) {
                    if (!first_info) first_info = rc;
                    continue;
                }
                if (rc) return rc;
                continue;
            }
            int info = adsg::factor_for_lines(s.ab, d.tab.dofs, s.kl, s.ku, s.lf);
            if (info != 0) {
                if (!first_info) first_info = info;
                continue;
            }
            free_slot(s, h->in_arena);
            int rc;
            if (h->precision == ADSG_FP32) {
                if ((rc = upload<float>(h, &s.lower, s.lf.lower))) return rc;
                if ((rc = upload<float>(h, &s.upper, s.lf.upper))) return rc;
                if ((rc = upload<float>(h, &s.rdiag, s.lf.rdiag))) return rc;
            } else {
                if ((rc = upload<double>(h, &s.lower, s.lf.lower))) return rc;
                if ((rc = upload<double>(h, &s.upper, s.lf.upper))) return rc;
                if ((rc = upload<double>(h, &s.rdiag, s.lf.rdiag))) return rc;
            }
            s.seg = adsg::line_segments{};
            // up to 16 segment waves (1024 lanes) of 16 or 32 rows; x lines are staged through
            // LDS where that fits (150 KB: FP64 14 waves of 16 rows or 8 of 32, FP32 16 or 13)
            // and read lane by lane otherwise
            if (adsg::segment_lines(s.lf, s.seg, 16, 16)) {
                if (h->precision == ADSG_FP32) {
                    if ((rc = upload<float>(h, &s.pack, s.seg.pack))) return rc;
                    if ((rc = upload<float>(h, &s.packt, s.seg.packt))) return rc;
                } else {
                    if ((rc = upload<double>(h, &s.pack, s.seg.pack))) return rc;
                    if ((rc = upload<double>(h, &s.packt, s.seg.packt))) return rc;
                }
                // rows of the unfactored matrix: arow[j*w + t + kl] = A(j, j+t), t in [-kl, ku]
                const int n = s.lf.n, w = s.kl + s.ku + 1, ld = 2 * s.kl + s.ku + 1;
                std::vector<double> rows(static_cast<size_t>(n) * w, 0.0);
                for (int j = 0; j < n; ++j)
                    for (int t = -s.kl; t <= s.ku; ++t)
                        if (j + t >= 0 && j + t < n)
                            rows[static_cast<size_t>(j) * w + t + s.kl] =
                                s.ab[static_cast<size_t>(j + t) * ld + s.kl + s.ku - t];
                if ((rc = upload<double>(h, reinterpret_cast<void**>(&s.arow), rows))) return rc;
            } else {
                s.seg = adsg::line_segments{};
            }
            if ((rc = upload_bytes(h, reinterpret_cast<void**>(&s.poff), s.lf.poff.data(), sizeof(int) * s.lf.n)))
                return rc;
            s.factored = true;
        }
        // FP32 increments need the rows of the (fixed) slot-0 matrix for boundary residuals
        const dev_slot& s0 = d.slots[0];
        if (h->precision == ADSG_FP32 && s0.factored && s0.kl == d.tab.p && s0.ku == d.tab.p) {
            const int n = d.tab.dofs, p = d.tab.p, w = 2 * p + 1, ld = 3 * p + 1;
            std::vector<double> rows(static_cast<size_t>(n) * w, 0.0);
            for (int i = 0; i < n; ++i)
                for (int t = 0; t < w; ++t) {
                    const int j = i - p + t;
                    if (j >= 0 && j < n) rows[static_cast<size_t>(i) * w + t] = s0.ab[static_cast<size_t>(j) * ld + 2 * p + i - j];
                }
            int rc = upload<float>(h, &d.rows, rows);
            if (rc) return rc;
        }
    }
    h->prepared = first_info == 0;
    drop_graphs(h);
    if (first_info) return fail(h, first_info, "prepare_matrices: zero pivot (dgbtrf info > 0)");
    return ADSG_OK;
}

int adsg_set_vector(adsg_sim* h, int which, const double* host) {
    if (!h || !host || which < ADSG_U || which > ADSG_GRAD_F) return fail(h, ADSG_EINVAL, "set_vector: bad argument");
    if (which == ADSG_SOURCE) h->source_load_ok = false;
    const size_t es = elt_of(h, which);
    if (!h->vec[which]) {
        if (hipMalloc(&h->vec[which], h->N * es) != hipSuccess)
            return fail(h, ADSG_ENOMEM, "set_vector: device allocation failed");
        drop_graphs(h);
    }
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    if (es == sizeof(float)) {
        std::vector<float> tmp(h->N);
        for (long i = 0; i < h->N; ++i) tmp[i] = static_cast<float>(host[i]);
        HIP_TRY(h, hipMemcpy(h->vec[which], tmp.data(), h->N * sizeof(float), hipMemcpyHostToDevice));
    } else {
        HIP_TRY(h, hipMemcpy(h->vec[which], host, h->N * sizeof(double), hipMemcpyHostToDevice));
    }
    if (h->precision == ADSG_FP32 && (which == ADSG_U || which == ADSG_U_PREV)) {
        hipError_t e = adsg::launch_convert_d2f(static_cast<const double*>(h->vec[which]),
                                                static_cast<float*>(h->shadow[which]), h->N, h->stream);
        if (e != hipSuccess) return hip_fail(h, e, "set_vector: shadow");
        HIP_TRY(h, hipStreamSynchronize(h->stream));
    }
    return ADSG_OK;
}

int adsg_get_vector(adsg_sim* h, int which, double* host) {
    if (!h || !host || which < ADSG_U || which > ADSG_GRAD_F || !h->vec[which])
        return fail(h, ADSG_EINVAL, "get_vector: bad argument");
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    if (elt_of(h, which) == sizeof(float)) {
        std::vector<float> tmp(h->N);
        HIP_TRY(h, hipMemcpy(tmp.data(), h->vec[which], h->N * sizeof(float), hipMemcpyDeviceToHost));
        for (long i = 0; i < h->N; ++i) host[i] = tmp[i];
    } else {
        HIP_TRY(h, hipMemcpy(host, h->vec[which], h->N * sizeof(double), hipMemcpyDeviceToHost));
    }
    return ADSG_OK;
}

void* adsg_device_ptr(adsg_sim* h, int which) {
    if (!h || which < ADSG_U || which > ADSG_GRAD_F) return nullptr;
    if (which == ADSG_SOURCE) h->source_load_ok = false;  // the caller may write through it
    return h->vec[which];
}

int adsg_set_vector_device(adsg_sim* h, int which, const double* dev) {
    if (!h || !dev || which < ADSG_U || which > ADSG_GRAD_F)
        return fail(h, ADSG_EINVAL, "set_vector_device: bad argument");
    if (which == ADSG_SOURCE) h->source_load_ok = false;
    if (elt_of(h, which) != sizeof(double))
        return fail(h, ADSG_EINVAL, "set_vector_device: FP32 vectors are set from the host");
    if (!h->vec[which]) {
        if (hipMalloc(&h->vec[which], h->N * sizeof(double)) != hipSuccess)
            return fail(h, ADSG_ENOMEM, "set_vector_device: device allocation failed");
        drop_graphs(h);
    }
    HIP_TRY(h, hipMemcpyAsync(h->vec[which], dev, h->N * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    if (h->precision == ADSG_FP32 && (which == ADSG_U || which == ADSG_U_PREV)) {
        hipError_t e = adsg::launch_convert_d2f(static_cast<const double*>(h->vec[which]),
                                                static_cast<float*>(h->shadow[which]), h->N, h->stream);
        if (e != hipSuccess) return hip_fail(h, e, "set_vector_device: shadow");
    }
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    return ADSG_OK;
}

int adsg_get_vector_device(adsg_sim* h, int which, double* dev) {
    if (!h || !dev || which < ADSG_U || which > ADSG_GRAD_F || !h->vec[which])
        return fail(h, ADSG_EINVAL, "get_vector_device: bad argument");
    if (elt_of(h, which) != sizeof(double))
        return fail(h, ADSG_EINVAL, "get_vector_device: FP32 vectors are read to the host");
    HIP_TRY(h, hipMemcpyAsync(dev, h->vec[which], h->N * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    return ADSG_OK;
}

int adsg_swap(adsg_sim* h) {
    if (!h) return ADSG_EINVAL;
    swap_state(h);
    return ADSG_OK;
}

int adsg_compute_rhs(adsg_sim* h, const adsg_rhs_form* form) {
    if (!h || !form) return fail(h, ADSG_EINVAL, "compute_rhs: null argument");
    if (int rc = ensure_source_load(h, form)) return rc;
    return rhs_dispatch(h, form);
}

int adsg_zero_boundary(adsg_sim* h) {
    if (!h) return ADSG_EINVAL;
    return zero_boundary_impl(h);
}

int adsg_ads_solve(adsg_sim* h, const int* slots) {
    if (!h) return ADSG_EINVAL;
    return solve_impl(h, slots);
}

int adsg_step(adsg_sim* h, int nsteps, const adsg_rhs_form* form, int flags) {
    if (!h || !form || nsteps < 0) return fail(h, ADSG_EINVAL, "step: bad argument");
    if (!h->prepared) return fail(h, ADSG_ESTATE, "step: prepare_matrices first");
    if (nsteps == 0) return ADSG_OK;
    // the check runs after the replayed steps, so the captured graphs do not depend on it
    const int rc = run_graphs(h, EXPLICIT_STEP, nsteps, form, flags & ~ADSG_STEP_CHECK_FINITE, nullptr);
    if (rc || !(flags & ADSG_STEP_CHECK_FINITE)) return rc;
    return check_finite_impl(h, ADSG_U, "step");
}

int adsg_step_delta(adsg_sim* h, int nsteps, const adsg_rhs_form* form, const int* slots, int flags) {
    if (!h || !form || !slots || nsteps < 0) return fail(h, ADSG_EINVAL, "step_delta: bad argument");
    if (!h->prepared) return fail(h, ADSG_ESTATE, "step_delta: prepare_matrices first");
    if ((flags & ADSG_STEP_REFINE) && h->precision != ADSG_FP32)
        return fail(h, ADSG_EINVAL, "step_delta: refinement applies to FP32 handles");
    if (nsteps == 0) return ADSG_OK;
    const int rc = run_graphs(h, DELTA_STEP, nsteps, form, flags & ~ADSG_STEP_CHECK_FINITE, slots);
    if (rc || !(flags & ADSG_STEP_CHECK_FINITE)) return rc;
    return check_finite_impl(h, ADSG_U, "step_delta");
}

int adsg_check_finite(adsg_sim* h, int which) {
    if (!h || which < ADSG_U || which > ADSG_GRAD_F) return fail(h, ADSG_EINVAL, "check_finite: bad argument");
    return check_finite_impl(h, which, "check_finite");
}

int adsg_adjoint_step(adsg_sim* h, const adsg_rhs_form* form, int flags) {
    if (!h || !form) return fail(h, ADSG_EINVAL, "adjoint_step: bad argument");
    if (!h->prepared) return fail(h, ADSG_ESTATE, "adjoint_step: prepare_matrices first");
    if (h->precision != ADSG_FP64) return fail(h, ADSG_EINVAL, "adjoint_step: FP64 handles only");
    if (form->reaction != 0.0)
        return fail(h, ADSG_EINVAL, "adjoint_step: the form must be linear in u (reaction = 0)");
    if (form->use_kfield && !h->vec[ADSG_GRAD_K])
        return fail(h, ADSG_ESTATE, "adjoint_step: set ADSG_GRAD_K (the accumulator) first");
    if (form->use_kfield && h->general) return refuse_general(h, "adjoint_step (k-field gradient)");
    return one_adjoint_step(h, form, flags);
}

int adsg_forward_checkpointed(adsg_sim* h, int nsteps, int every, const adsg_rhs_form* form, int flags) {
    if (!h || !form || nsteps < 1 || every < 1) return fail(h, ADSG_EINVAL, "forward_checkpointed: bad argument");
    if (!h->prepared) return fail(h, ADSG_ESTATE, "forward_checkpointed: prepare_matrices first");
    if (h->precision != ADSG_FP64) return fail(h, ADSG_EINVAL, "forward_checkpointed: FP64 handles only");
    if (form->use_kfield && h->general) return refuse_general(h, "forward_checkpointed (k-field gradient)");
    const int nchk = (nsteps + every - 1) / every;
    const size_t need = static_cast<size_t>(nchk) * h->N;
    if (h->ckpt_cap < need) {
        if (h->ckpt) (void)hipFree(h->ckpt);
        h->ckpt = nullptr;
        h->ckpt_cap = 0;
        if (hipMalloc(reinterpret_cast<void**>(&h->ckpt), need * sizeof(double)) != hipSuccess)
            return fail(h, ADSG_ENOMEM, "forward_checkpointed: no memory for the checkpoints");
        h->ckpt_cap = need;
    }
    for (int c = 0; c < nchk; ++c) {
        HIP_TRY(h, hipMemcpyAsync(h->ckpt + static_cast<size_t>(c) * h->N, h->vec[ADSG_U],
                                  h->N * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
        const int todo = std::min(every, nsteps - c * every);
        int rc = run_graphs(h, EXPLICIT_STEP, todo, form, flags, nullptr);
        if (rc) return rc;
    }
    h->ck_steps = nsteps;
    h->ck_every = every;
    h->ck_flags = flags;
    h->ck_form = form_key(form);
    return ADSG_OK;
}

namespace {

// One segment of the checkpointed adjoint after its checkpoint is restored into u: the
// states u(start) .. u(stop-1) saved into ck_seg while the len - 1 explicit steps recompute
// them (one_step itself when captured, the one-step graph otherwise), lambda restored from
// ck_seg[every], then the len adjoint steps, each reading its forward state from ck_seg.
// Its copies are kernels (out = 1 * in, exact): the runtime's copies, eager or as graph
// nodes, measured from 2 to 4 µs each and made the whole adjoint vary from machine to machine.
int adjoint_segment(adsg_sim* h, const adsg_rhs_form* form, int flags, int len, bool captured) {
    auto copy = [&](double* dst, const void* src) {
        return adsg::launch_scale<double>(dst, static_cast<const double*>(src), h->N, 1.0, h->stream);
    };
    for (int s = 0; s < len; ++s) {
        HIP_TRY(h, copy(h->ck_seg + static_cast<size_t>(s) * h->N, h->vec[ADSG_U]));
        if (s + 1 < len) {
            const int rc = captured ? one_step(h, form, flags) : run_graphs(h, EXPLICIT_STEP, 1, form, flags, nullptr);
            if (rc) return rc;
        }
    }
    HIP_TRY(h, copy(static_cast<double*>(h->vec[ADSG_U]), h->ck_seg + static_cast<size_t>(h->ck_every) * h->N));
    for (int s = len - 1; s >= 0; --s) {
        // the step's forward state read where the recompute left it
        const int rc = one_adjoint_step(h, form, flags, h->ck_seg + static_cast<size_t>(s) * h->N);
        if (rc) return rc;
    }
    return ADSG_OK;
}

}  // namespace

int adsg_adjoint_checkpointed(adsg_sim* h, const adsg_rhs_form* form, int flags) {
    if (!h || !form) return fail(h, ADSG_EINVAL, "adjoint_checkpointed: bad argument");
    if (!h->ck_steps || h->ck_form != form_key(form) || h->ck_flags != flags)
        return fail(h, ADSG_ESTATE, "adjoint_checkpointed: no forward_checkpointed run with this form and flags");
    if (form->reaction != 0.0)
        return fail(h, ADSG_EINVAL, "adjoint_checkpointed: the form must be linear in u (reaction = 0)");
    if (form->use_kfield && !h->vec[ADSG_GRAD_K])
        return fail(h, ADSG_ESTATE, "adjoint_checkpointed: set ADSG_GRAD_K (the accumulator) first");
    const int nsteps = h->ck_steps, every = h->ck_every;
    const int nchk = (nsteps + every - 1) / every;
    const size_t need = static_cast<size_t>(every + 1) * h->N;  // segment states + lambda
    if (h->ck_seg_cap < need) {
        drop_graphs(h);  // the adjoint segment graphs hold the old buffer
        if (h->ck_seg) (void)hipFree(h->ck_seg);
        h->ck_seg = nullptr;
        h->ck_seg_cap = 0;
        if (hipMalloc(reinterpret_cast<void**>(&h->ck_seg), need * sizeof(double)) != hipSuccess)
            return fail(h, ADSG_ENOMEM, "adjoint_checkpointed: no memory for a segment");
        h->ck_seg_cap = need;
    }
    const size_t bytes = h->N * sizeof(double);
    double* lam_save = h->ck_seg + static_cast<size_t>(every) * h->N;
    int rc;
    if ((rc = ensure_source_load(h, form))) return rc;  // not captured
    for (int c = nchk - 1; c >= 0; --c) {
        const int start = c * every, stop = std::min(nsteps, start + every), len = stop - start;
        HIP_TRY(h, hipMemcpyAsync(lam_save, h->vec[ADSG_U], bytes, hipMemcpyDeviceToDevice, h->stream));
        // recompute the segment's states u(start) .. u(stop-1) from its checkpoint
        HIP_TRY(h, hipMemcpyAsync(h->vec[ADSG_U], h->ckpt + static_cast<size_t>(c) * h->N, bytes,
                                  hipMemcpyDeviceToDevice, h->stream));
        if (c == nchk - 1) {
            // the first segment eagerly (its adjoint steps allocate what they need on first use)
            if ((rc = adjoint_segment(h, form, flags, len, false))) return rc;
            continue;
        }
        // the others replay one graph of the segment (recompute, copies and adjoint steps on
        // buffers that stay fixed), captured once per starting state; its recompute swaps u
        // and u_prev once per step, as the eager steps do
        const std::string k = "adjoint segment|" + form_key(form) + "|" + std::to_string(every);
        const auto key = std::make_tuple(k, flags, h->vec[ADSG_U], len);
        auto it = h->graphs.find(key);
        hipGraphExec_t ex = nullptr;
        if (it != h->graphs.end()) {
            ex = it->second;
        } else {
            void* u0 = h->vec[ADSG_U];
            void* p0 = h->vec[ADSG_U_PREV];
            void* s0 = h->shadow[0];
            void* s1 = h->shadow[1];
            HIP_TRY(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
            rc = adjoint_segment(h, form, flags, len, true);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(h->stream, &g);
            h->vec[ADSG_U] = u0;  // capture only recorded the swaps
            h->vec[ADSG_U_PREV] = p0;
            h->shadow[0] = s0;
            h->shadow[1] = s1;
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess) return hip_fail(h, e, "hipStreamEndCapture");
            e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) return hip_fail(h, e, "hipGraphInstantiate");
            h->graphs[key] = ex;
        }
        HIP_TRY(h, hipGraphLaunch(ex, h->stream));
        if ((len - 1) % 2) swap_state(h);
    }
    return ADSG_OK;
}

int adsg_adjoint_step_delta(adsg_sim* h, int nsteps, const adsg_rhs_form* form, const int* slots,
                            int flags) {
    if (!h || !form || !slots || nsteps < 0) return fail(h, ADSG_EINVAL, "adjoint_step_delta: bad argument");
    if (!h->prepared) return fail(h, ADSG_ESTATE, "adjoint_step_delta: prepare_matrices first");
    if (h->precision != ADSG_FP64) return fail(h, ADSG_EINVAL, "adjoint_step_delta: FP64 handles only");
    if (form->reaction != 0.0)
        return fail(h, ADSG_EINVAL, "adjoint_step_delta: the form must be linear in u (reaction = 0)");
    if (!h->vec[ADSG_BUFFER] || !h->vec[ADSG_U_PREV]) return fail(h, ADSG_ESTATE, "adjoint_step_delta: vectors not allocated");
    if (nsteps == 0) return ADSG_OK;
    return run_graphs(h, ADJOINT_DELTA_STEP, nsteps, form, flags, slots);
}

namespace {

// block partial sums of launch_norm: one per block of 256 lanes, at most kNormBlocks blocks
// (the kernel strides over the elements, so every mesh fits), then the total
constexpr int kNormBlocks = 1024;

int norm_blocks(const adsg_sim* h) {
    long E = 1;
    for (int d = 0; d < h->dim; ++d) E *= h->dirs[d].tab.elements;
    return static_cast<int>(std::min<long>(kNormBlocks, (E + 255) / 256));
}

// sqrt of the functional of launch_norm: u = vec[which] (null for which = -1), sq on the
// device; partial holds `blocks` block sums and the total after them (device memory)
int norm_impl(adsg_sim* h, int which, const double* sq, int kind, double* out, double* partial, int blocks) {
    if (h->general) return refuse_general(h, "norm");
    const double* u = which >= 0 ? static_cast<const double*>(h->vec[which]) : nullptr;
    hipError_t e;
    if (h->precision == ADSG_FP32) {
        adsg::dev_tables<float> t[3];
        dir_tables(h, t);
        e = adsg::launch_norm<float>(h->dim, h->dirs[0].tab.p, u, sq, t, kind, partial, blocks,
                                     partial + blocks, h->stream);
    } else {
        adsg::dev_tables<double> t[3];
        dir_tables(h, t);
        e = adsg::launch_norm<double>(h->dim, h->dirs[0].tab.p, u, sq, t, kind, partial, blocks,
                                      partial + blocks, h->stream);
    }
    if (e != hipSuccess) return hip_fail(h, e, "norm");
    double total = 0;
    HIP_TRY(h, hipMemcpyAsync(&total, partial + blocks, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    *out = std::sqrt(total);
    return ADSG_OK;
}

// the same with the handle's own partial-sum buffer (allocated on first use)
int norm_impl_own(adsg_sim* h, int which, const double* sq, int kind, double* out) {
    if (!h->norm_scratch &&
        hipMalloc(reinterpret_cast<void**>(&h->norm_scratch), sizeof(double) * (kNormBlocks + 1)) != hipSuccess)
        return fail(h, ADSG_ENOMEM, "norm: no memory for partial sums");
    return norm_impl(h, which, sq, kind, out, h->norm_scratch, norm_blocks(h));
}

}  // namespace

// Partial sums of adsg_norm_noalloc: one device array per GPU, held until the process ends.
// Host code never uses this name as an address (on the host it is only a stand-in for the
// symbol); it takes the device address from scratch_ptr, which resolves it per device.
static __device__ __align__(256) double scratch[kNormBlocks + 1];

namespace {

std::mutex g_scratch_mu;
std::map<int, double*> g_scratch_dev;  // device ordinal -> address of that device's scratch

// the address of `scratch` on the handle's device, resolved once per device with that
// device current (each GPU holds its own copy of the array at its own address)
int scratch_ptr(adsg_sim* h, double** out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    auto it = g_scratch_dev.find(h->device);
    if (it != g_scratch_dev.end()) {
        *out = it->second;
        return ADSG_OK;
    }
    int cur = 0;
    HIP_TRY(h, hipGetDevice(&cur));
    if (cur != h->device) HIP_TRY(h, hipSetDevice(h->device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != h->device) (void)hipSetDevice(cur);
    if (e != hipSuccess || !p) return hip_fail(h, e != hipSuccess ? e : hipErrorInvalidSymbol, "norm_noalloc: scratch address");
    g_scratch_dev[h->device] = static_cast<double*>(p);
    *out = static_cast<double*>(p);
    return ADSG_OK;
}

}  // namespace

int adsg_norm(adsg_sim* h, int which, int kind, double* out) {
    if (!h || !out || (which != ADSG_U && which != ADSG_U_PREV) || kind < 0 || kind > 2)
        return fail(h, ADSG_EINVAL, "norm: bad argument (U or U_PREV; kind 0, 1 or 2)");
    return norm_impl_own(h, which, nullptr, kind, out);
}

int adsg_error(adsg_sim* h, int which, int kind, const double* sq, double* out) {
    if (!h || !out || !sq || (which != ADSG_U && which != ADSG_U_PREV && which != -1) || kind < 0 || kind > 2)
        return fail(h, ADSG_EINVAL, "error: bad argument (U, U_PREV or -1; kind 0, 1 or 2)");
    const size_t planes = kind == 0 ? 1 : 1 + h->dim;
    const size_t bytes = planes * quad_point_count(h) * sizeof(double);
    void* d = nullptr;
    if (hipMalloc(&d, bytes) != hipSuccess) return fail(h, ADSG_ENOMEM, "error: no memory for the function values");
    int rc = ADSG_OK;
    hipError_t e = hipMemcpyAsync(d, sq, bytes, hipMemcpyHostToDevice, h->stream);
    if (e != hipSuccess) rc = hip_fail(h, e, "error: upload");
    if (rc == ADSG_OK) rc = norm_impl_own(h, which, static_cast<const double*>(d), kind, out);
    (void)hipStreamSynchronize(h->stream);
    (void)hipFree(d);
    return rc;
}

int adsg_norm_noalloc(adsg_sim* h, int which, int kind, const double* sq_dev, double* out) {
    if (!h || !out || kind < 0 || kind > 2 || (which != ADSG_U && which != ADSG_U_PREV && which != -1) ||
        (which == -1 && !sq_dev))
        return fail(h, ADSG_EINVAL, "norm_noalloc: bad argument (U, U_PREV, or -1 with sq; kind 0, 1 or 2)");
    if (which >= 0 && !h->vec[which]) return fail(h, ADSG_ESTATE, "norm_noalloc: vector not allocated");
    const int blocks = norm_blocks(h);
    if (blocks < 1 || blocks > kNormBlocks)
        return fail(h, ADSG_EINVAL, "norm_noalloc: the reduction needs " + std::to_string(blocks) +
                                        " partial sums; the static scratch holds " + std::to_string(kNormBlocks));
    double* partial = nullptr;
    int rc = scratch_ptr(h, &partial);
    if (rc) return rc;
    return norm_impl(h, which, sq_dev, kind, out, partial, blocks);
}

int adsg_project(adsg_sim* h, int which, const double* fq) {
    if (!h || !fq || !projectable(h, which)) return fail(h, ADSG_EINVAL, "project: bad argument (an FP64 vector)");
    const size_t bytes = quad_point_count(h) * sizeof(double);
    void* d = nullptr;
    if (hipMalloc(&d, bytes) != hipSuccess) return fail(h, ADSG_ENOMEM, "project: no memory for f");
    int rc = ADSG_OK;
    hipError_t e = hipMemcpyAsync(d, fq, bytes, hipMemcpyHostToDevice, h->stream);
    if (e != hipSuccess) rc = hip_fail(h, e, "project: upload");
    if (rc == ADSG_OK) rc = project_impl(h, which, static_cast<const double*>(d));
    e = hipStreamSynchronize(h->stream);
    if (rc == ADSG_OK && e != hipSuccess) rc = hip_fail(h, e, "project");
    (void)hipFree(d);
    return rc;
}

int adsg_project_device(adsg_sim* h, int which, const double* fq) {
    if (!h || !fq || !projectable(h, which)) return fail(h, ADSG_EINVAL, "project_device: bad argument (an FP64 vector)");
    int rc = project_impl(h, which, fq);
    if (rc) return rc;
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    return ADSG_OK;
}

int adsg_sample(adsg_sim* h, int which, const int* counts, const double* px, const double* py,
                const double* pz, int ders, double* out) {
    const double* pts[3] = {px, py, pz};
    if (!h || !counts || !out || !projectable(h, which) || !h->vec[which])
        return fail(h, ADSG_EINVAL, "sample: bad argument (an FP64 vector that is set)");
    if (h->mixed_p) return fail(h, ADSG_EINVAL, "sample: needs one degree in every direction");
    long total = 1;
    for (int d = 0; d < h->dim; ++d) {
        if (counts[d] <= 0 || !pts[d]) return fail(h, ADSG_EINVAL, "sample: empty direction");
        total *= counts[d];
    }
    const int p = h->dirs[0].tab.p, m = p + 1;
    // per direction: [point][value 0..p | derivative 0..p] and the first nonzero function
    std::vector<double> tab[3];
    std::vector<int> off[3];
    for (int d = 0; d < h->dim; ++d) {
        const auto& knot = h->dirs[d].tab.knot;
        tab[d].resize(static_cast<size_t>(counts[d]) * 2 * m);
        off[d].resize(counts[d]);
        for (int i = 0; i < counts[d]; ++i) {
            const int span = adsg::span_of(pts[d][i], knot, p);
            adsg::basis_and_derivative(span, pts[d][i], knot, p, &tab[d][static_cast<size_t>(i) * 2 * m],
                                       &tab[d][static_cast<size_t>(i) * 2 * m + m]);
            off[d][i] = span - p;
        }
    }
    const size_t planes = ders ? 1 + h->dim : 1;
    size_t bytes = planes * total * sizeof(double);
    for (int d = 0; d < h->dim; ++d) bytes += tab[d].size() * sizeof(double) + off[d].size() * sizeof(int) + 256;
    char* buf = nullptr;
    if (hipMalloc(reinterpret_cast<void**>(&buf), bytes) != hipSuccess)
        return fail(h, ADSG_ENOMEM, "sample: no memory for the grid");
    int rc = ADSG_OK;
    double* dout = reinterpret_cast<double*>(buf);
    size_t at = planes * total * sizeof(double);
    const double* dtab[3] = {nullptr, nullptr, nullptr};
    const int* doff[3] = {nullptr, nullptr, nullptr};
    for (int d = 0; d < h->dim && rc == ADSG_OK; ++d) {
        at = (at + 255) / 256 * 256;
        dtab[d] = reinterpret_cast<const double*>(buf + at);
        hipError_t e = hipMemcpyAsync(buf + at, tab[d].data(), tab[d].size() * sizeof(double),
                                      hipMemcpyHostToDevice, h->stream);
        at += tab[d].size() * sizeof(double);
        doff[d] = reinterpret_cast<const int*>(buf + at);
        if (e == hipSuccess)
            e = hipMemcpyAsync(buf + at, off[d].data(), off[d].size() * sizeof(int), hipMemcpyHostToDevice, h->stream);
        at += off[d].size() * sizeof(int);
        if (e != hipSuccess) rc = hip_fail(h, e, "sample: upload");
    }
    if (rc == ADSG_OK) {
        const int c3[3] = {counts[0], counts[1], h->dim == 3 ? counts[2] : 1};
        hipError_t e = adsg::launch_sample(h->dim, p, static_cast<const double*>(h->vec[which]),
                                           h->dirs[0].tab.dofs, h->dirs[1].tab.dofs, dtab, doff, c3,
                                           ders ? 1 : 0, dout, h->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(out, dout, planes * total * sizeof(double), hipMemcpyDeviceToHost, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess) rc = hip_fail(h, e, "sample");
    }
    (void)hipStreamSynchronize(h->stream);
    (void)hipFree(buf);
    return rc;
}

int adsg_quad_points(const adsg_sim* h, int dir, double* x) {
    if (!h || !x || dir < 0 || dir >= h->dim) return ADSG_EINVAL;
    const auto& xq = h->dirs[dir].tab.xq;
    std::copy(xq.begin(), xq.end(), x);
    return ADSG_OK;
}

int adsg_ads_solve_transposed(adsg_sim* h, const int* slots) {
    if (!h) return ADSG_EINVAL;
    if (h->precision != ADSG_FP64) return fail(h, ADSG_EINVAL, "ads_solve_transposed: FP64 handles only");
    return sweep_all<double>(h, static_cast<double*>(h->vec[ADSG_U]), slots, 1);
}

int adsg_synchronize(adsg_sim* h) {
    if (!h) return ADSG_EINVAL;
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    return ADSG_OK;
}

int adsg_time_phases(adsg_sim* h, int nsteps, const adsg_rhs_form* form, const int* slots, int flags,
                     double* ms) {
    if (!h || !form || !ms || nsteps <= 0) return fail(h, ADSG_EINVAL, "time_phases: bad argument");
    if (!h->prepared) return fail(h, ADSG_ESTATE, "time_phases: prepare_matrices first");
    const step_kind kind = slots ? DELTA_STEP : EXPLICIT_STEP;
    if (int rc = ensure_source_load(h, form)) return rc;
    hipEvent_t ev[8];
    for (auto& e : ev) HIP_TRY(h, hipEventCreate(&e));
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    int rc = ADSG_OK;
    h->marks = ev;
    for (int s = 0; s < nsteps && rc == ADSG_OK; ++s) {
        (void)hipEventRecord(ev[0], h->stream);
        rc = step_of(h, kind, form, slots, flags);
        (void)hipEventRecord(ev[7], h->stream);
        if (rc) break;
        hipError_t e = hipEventSynchronize(ev[7]);
        if (e != hipSuccess) {
            rc = hip_fail(h, e, "hipEventSynchronize");
            break;
        }
        float t;
        const int from[7] = {1, 2, 3, 4, 5, 0, 0};
        const int to[7] = {2, 3, 4, 5, 6, 7, 1};
        for (int i = 0; i < 7; ++i) {
            (void)hipEventElapsedTime(&t, ev[from[i]], ev[to[i]]);
            acc[i] += t;
        }
    }
    h->marks = nullptr;
    for (auto& e : ev) (void)hipEventDestroy(e);
    for (int i = 0; i < 7; ++i) ms[i] = acc[i] / nsteps;
    return rc;
}

int adsg_set_tuning(adsg_sim* h, int key, int value) {
    if (!h) return ADSG_EINVAL;
    int* slot = key == ADSG_TUNE_RHS_VARIANT     ? &h->rhs_variant
                : key == ADSG_TUNE_SWEEP_VARIANT ? &h->sweep_variant
                : key == ADSG_TUNE_RHS_KRONECKER ? &h->rhs_kron
                : key == ADSG_TUNE_RHS_COLUMNS   ? &h->rhs_col_shape
                : key == ADSG_TUNE_RHS_TILING    ? &h->rhs_tiling
                : key == ADSG_TUNE_CACHE         ? &h->tune_cache
                                                 : nullptr;
    const int max_value = key == ADSG_TUNE_RHS_VARIANT     ? 5
                          : key == ADSG_TUNE_RHS_COLUMNS   ? 3
                          : key == ADSG_TUNE_SWEEP_VARIANT ? 2
                                                           : 1;
    if (!slot || value < 0 || value > max_value) return fail(h, ADSG_EINVAL, "set_tuning: unknown key or value");
    if (key == ADSG_TUNE_RHS_TILING && value == 1) {
        const int rc = ensure_spill(h);
        if (rc) return rc;
    }
    if (*slot != value && key != ADSG_TUNE_CACHE) drop_graphs(h);
    *slot = value;
    return ADSG_OK;
}

int adsg_get_tuning(const adsg_sim* h, int key) {
    if (!h) return ADSG_EINVAL;
    if (key == ADSG_TUNE_RHS_VARIANT) return h->rhs_variant;
    if (key == ADSG_TUNE_SWEEP_VARIANT) return h->sweep_variant;
    if (key == ADSG_TUNE_RHS_KRONECKER) return h->rhs_kron;
    if (key == ADSG_TUNE_RHS_COLUMNS) return h->rhs_col_shape;
    if (key == ADSG_TUNE_RHS_TILING) return h->rhs_tiling;
    if (key == ADSG_TUNE_CACHE) return h->tune_cache;
    return ADSG_EINVAL;
}

}  // extern "C"

namespace {

// The autotuner's results per configuration, shared by the handles of the process: the
// device, dimension, precision, degree, element count and Gauss points of every direction,
// the kind of form (mass term, k field, source, reaction), delta or explicit steps and the
// step flags. The element and sweep variants' relative speed depends on these only.
std::mutex g_tune_mutex;
std::map<std::string, std::array<int, 4>> g_tune_cache;  // {variant, column shape, tiling, sweep}

std::string tune_key(const adsg_sim* h, const adsg_rhs_form* form, const int* slots, int flags) {
    std::string k = std::to_string(h->device) + "/" + std::to_string(h->dim) + "/" + std::to_string(h->precision);
    for (int d = 0; d < h->dim; ++d)
        k += "/" + std::to_string(h->dirs[d].tab.p) + "," + std::to_string(h->dirs[d].tab.elements) + "," +
             std::to_string(h->dirs[d].tab.nq) + "," + std::to_string(h->dirs[d].tab.repeated_nodes);
    k += "/" + std::to_string(form->mass == 0.0) + std::to_string(form->use_kfield != 0) +
         std::to_string(form->use_source != 0) + std::to_string(form->reaction != 0.0);
    k += "/" + std::to_string(slots != nullptr) + "/" + std::to_string(flags);
    return k;
}

}  // namespace

extern "C" {

int adsg_autotune(adsg_sim* h, const adsg_rhs_form* form, const int* slots, int flags) {
    if (!h || !form) return fail(h, ADSG_EINVAL, "autotune: bad argument");
    if (!h->prepared) return fail(h, ADSG_ESTATE, "autotune: prepare_matrices first");
    const std::string key = tune_key(h, form, slots, flags);
    if (h->tune_cache) {
        std::array<int, 4> c{};
        bool hit = false;
        {
            std::lock_guard<std::mutex> lock(g_tune_mutex);
            auto it = g_tune_cache.find(key);
            if (it != g_tune_cache.end()) {
                c = it->second;
                hit = true;
            }
        }
        if (hit && (c[2] == 0 || ensure_spill(h) == ADSG_OK)) {
            h->rhs_variant = c[0];
            h->rhs_col_shape = c[1];
            h->rhs_tiling = c[2];
            h->sweep_variant = c[3];
            drop_graphs(h);
            return ADSG_OK;
        }
    }
    // keep the caller's state: copies of u, u_prev (and their FP32 shadows)
    std::vector<void*> saved;
    std::vector<std::pair<void*, size_t>> what = {{h->vec[ADSG_U], h->N * sizeof(double)},
                                                  {h->vec[ADSG_U_PREV], h->N * sizeof(double)}};
    if (h->precision == ADSG_FP32) {
        what.push_back({h->shadow[0], h->N * sizeof(float)});
        what.push_back({h->shadow[1], h->N * sizeof(float)});
    }
    void* ptrs[4] = {h->vec[ADSG_U], h->vec[ADSG_U_PREV], h->shadow[0], h->shadow[1]};
    for (auto& w : what) {
        void* p = nullptr;
        if (hipMalloc(&p, w.second) != hipSuccess) {
            for (void* q : saved) (void)hipFree(q);
            return fail(h, ADSG_ENOMEM, "autotune: no memory for the state copy");
        }
        (void)hipMemcpyAsync(p, w.first, w.second, hipMemcpyDeviceToDevice, h->stream);
        saved.push_back(p);
    }
    int best = h->rhs_variant, best_shape = h->rhs_col_shape, best_sweep = h->sweep_variant;
    int best_tiling = h->rhs_tiling;
    double best_ms = 1e30, best_sweep_ms = 1e30;
    int rc = ADSG_OK;
    // candidates {variant, shape of variant 5, tiling}: variants 0..4 in colour launches, the
    // column variant (5) in each of its shapes, and (3D) variants 0, 1, 2, 4 in one launch
    std::vector<std::array<int, 3>> cand;
    for (int v = 0; v <= 4; ++v) cand.push_back({v, 0, 0});
    for (int c = 0; c <= 3; ++c) cand.push_back({5, c, 0});
    if (!h->general && ensure_spill(h) == ADSG_OK && h->spill)
        for (int v : (h->dim == 3 ? std::vector<int>{0, 1, 2, 4} : std::vector<int>{0})) cand.push_back({v, 0, 1});
    for (size_t k = 0; k < cand.size() && rc == ADSG_OK; ++k) {
        const int v = cand[k][0];
        h->rhs_variant = v;
        h->rhs_col_shape = cand[k][1];
        h->rhs_tiling = cand[k][2];
        // the sweep variants 0, 1, 2 timed with the element variants of the same number in
        // colour launches
        const bool sweep_run = v <= 2 && cand[k][2] == 0;
        h->sweep_variant = sweep_run ? v : best_sweep;
        double ms[7];
        rc = adsg_time_phases(h, 1, form, slots, flags, ms);  // warm-up (code load, caches)
        // 8 timed steps: variants 1 and 3 are a few per cent apart at full-GPU sizes
        if (!rc) rc = adsg_time_phases(h, 8, form, slots, flags, ms);
        if (!rc && ms[0] < best_ms) {
            best_ms = ms[0];
            best = v;
            best_shape = h->rhs_col_shape;
            best_tiling = h->rhs_tiling;
        }
        const double sweeps = ms[2] + ms[3] + ms[4];
        if (!rc && sweep_run && sweeps < best_sweep_ms) {
            best_sweep_ms = sweeps;
            best_sweep = v;
        }
    }
    // restore the state and its pointer assignment
    h->vec[ADSG_U] = ptrs[0];
    h->vec[ADSG_U_PREV] = ptrs[1];
    h->shadow[0] = ptrs[2];
    h->shadow[1] = ptrs[3];
    for (size_t i = 0; i < what.size(); ++i)
        (void)hipMemcpyAsync(what[i].first, saved[i], what[i].second, hipMemcpyDeviceToDevice, h->stream);
    (void)hipStreamSynchronize(h->stream);
    for (void* q : saved) (void)hipFree(q);
    h->pending = h->pending_dirichlet = false;
    h->rhs_variant = best;
    h->rhs_col_shape = best_shape;
    h->rhs_tiling = best_tiling;
    h->sweep_variant = best_sweep;
    drop_graphs(h);
    if (rc == ADSG_OK && h->tune_cache) {
        std::lock_guard<std::mutex> lock(g_tune_mutex);
        g_tune_cache[key] = {best, best_shape, best_tiling, best_sweep};
    }
    return rc;
}

int adsg_band_factorize(int n, int kl, int ku, double* ab, int* ipiv) {
    if (n < 0 || kl < 0 || ku < 0 || !ab || !ipiv) return ADSG_EINVAL;
    std::vector<double> b(ab, ab + static_cast<size_t>(2 * kl + ku + 1) * n);
    std::vector<int> p;
    int info = adsg::band_lu(b, n, kl, ku, p);
    std::memcpy(ab, b.data(), sizeof(double) * b.size());
    std::memcpy(ipiv, p.data(), sizeof(int) * n);
    return info;
}

}  // extern "C"
