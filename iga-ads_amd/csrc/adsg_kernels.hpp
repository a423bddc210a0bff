// Kernel launch interface of the MI355X ADS step (device code in adsg_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace adsg {

// Factors of one band matrix for the line sweeps (layout documented in adsg_setup.hpp).
template <typename F>
struct dev_factors {
    const F* lower = nullptr;  // [n*kl]
    const F* upper = nullptr;  // [n*ku]
    const F* rdiag = nullptr;  // [n]
    const int* poff = nullptr; // [n], only read when pivoted
    int n = 0, kl = 0, ku = 0;
    bool pivoted = false;
};

// One direction of the line solve: n_lines lines of length f.n, line l starts at
// (l % A) + (l / A) * B and has stride S. S == 1 selects the LDS-staged contiguous sweep.
struct line_geometry {
    long n_lines = 0;
    long A = 1, B = 0, S = 1;
};

template <typename T, typename F>
hipError_t launch_line_sweep(T* u, const line_geometry& g, const dev_factors<F>& f,
                             hipStream_t s);

// Segmented sweep of unpivoted factors: every line is cut into nseg segments of length m
// (the last one shorter) solved concurrently. pack holds, per row j of the line padded to
// nseg*m rows, the factor entries row j uses and its influence coefficients: G (the
// influence of the previous segment's last kl values on y_j) and H (of the next segment's
// first ku values on x_j); packt the same for the transposed solve (layout in
// adsg_setup.hpp). refine: one refinement of every line solve against the FP64 residual
// w - A z, with arow the unfactored matrix as rows of kl+ku+1 entries (columns j-kl ..
// j+ku).
template <typename F>
struct seg_factors {
    dev_factors<F> f;
    const F* pack = nullptr;   // [nseg*mseg][2(kl+ku)+1]: mseg rows per segment
    const F* packt = nullptr;  // [nseg*mseg][2(kl+ku)+1]
    const double* arow = nullptr;  // [n*(kl+ku+1)]
    int m = 0, nseg = 0, mseg = 0;  // mseg: register capacity per lane (16 or 32)
    bool transposed = false, refine = false;
    bool direct = false;  // lines with S == 1 read by each lane directly, not staged through LDS
    const F* src = nullptr;  // when set, the lines are read from here and written to u (out of place)
};

// The one-launch element loop's fix-up (adding the spill entries of the earlier tiles that
// share a DOF) and the Dirichlet zeroing of every boundary DOF, left to the caller of the
// element loop: both in one pass over the DOFs (k_rhs_fixup: fix-up sum added, then the
// boundary set to 0, the order of the separate passes, so the same bits), or, without spill
// entries, the zeroing alone in the x-line sweep as it loads its lines. (The spill sums
// inside the sweep measured slower: each lane's gathers for its ~15 rows run one row after
// another.)
template <typename T>
struct rhs_fixup {
    const T* spill = nullptr;  // null: no spill entries to add
    long off[8] = {};          // the spill layout of rhs_params (spill_off, spill_ext)
    int ext[3][2] = {};
    int p = 0;
    int tile[3] = {1, 1, 1}, nt[3] = {1, 1, 1};  // tile size (elements) and tile count per direction
    int n[3] = {1, 1, 1};                        // DOFs per direction (1 beyond the dimension)
    bool zero_boundary = false;
    bool pending() const { return spill != nullptr || zero_boundary; }
};

// The increment update of the delta and FP32 steps, x_new = x_prev + d with the boundary
// entries 0 when dirichlet (k_add_increment / k_apply_increment; shadow = FP32 copy of x_new
// when set), done by the last sweep of the step as it stores its lines instead of storing d;
// without x_prev, the last sweep (also of a transposed solve) stores d with its boundary
// entries 0 when dirichlet (k_zero_boundary). dir: the direction of that sweep; n: DOFs per
// direction (1 beyond the dimension).
struct seg_apply {
    const double* x_prev = nullptr;
    double* x_new = nullptr;
    float* shadow = nullptr;
    int n[3] = {1, 1, 1};
    int dir = 0;
    bool dirichlet = false;
};

// fix: a deferred boundary zeroing (rhs_fixup without spill entries) to apply while the x
// lines (geometry with S == 1, the plain solve) are loaded, or null; the caller applies
// other fix-ups first (launch_rhs_fixup). apply: the increment update to do while lines with
// S != 1 are stored (not the transposed solve), or null: *applied tells whether it was done.
template <typename T, typename F>
hipError_t launch_line_sweep_seg(T* u, const line_geometry& g, const seg_factors<F>& f,
                                 hipStream_t s, const rhs_fixup<T>* fix = nullptr,
                                 const seg_apply* apply = nullptr, bool* applied = nullptr);

// Per-line factors (one band matrix per line, lines of geometry g in its order): entry k
// of line l at k * n_lines + l, widths kl and ku (ku = kl + ku of the matrix when pivoted).
template <typename T>
struct own_factors {
    const T* lower = nullptr;  // [n*kl][L]
    const T* upper = nullptr;  // [n*ku][L]
    const T* rdiag = nullptr;  // [n][L]
    const int* poff = nullptr; // [n][L]
    int n = 0, kl = 0, ku = 0;
    bool pivoted = false;
};

template <typename T>
hipError_t launch_line_sweep_own(T* u, const line_geometry& g, const own_factors<T>& f,
                                 hipStream_t s);

// Tables of one direction on the device: val/der [elements][p+1][p+1] (element, Gauss
// point, local function) and wj [elements][p+1]; bval/bder are val/der times the weight
// wj of their Gauss point. val, der, bval and bder are zero-padded to a whole number of
// kTablePad elements, so that a tile of elements can be copied without bounds checks.
constexpr int kTablePad = 16;
template <typename T>
struct dev_tables {
    const T* val = nullptr;
    const T* der = nullptr;
    const T* wj = nullptr;
    const T* bval = nullptr;
    const T* bder = nullptr;
    int elements = 0;
    int dofs = 0;
    const int* first = nullptr;  // first DOF of every element (general element loop)
    int nq = 0;                  // Gauss points per element (general element loop)
    int p = 0;                   // degree (general element loop)
};

// Integrand of compute_rhs at a Gauss point:
//   mass*u*v - dt*k*(gx*ux*vx + gy*uy*vy + gz*uz*vz) + dt*reaction*u*(1-u)*v + dt*f*v
template <typename T>
struct rhs_params {
    T mass = 1, dt = 0, gx = 1, gy = 1, gz = 1, reaction = 0;
    const T* kfield = nullptr;  // coefficients of k on the same basis, or null (k = 1)
    const T* source = nullptr;  // coefficients of f on the same basis, or null (f = 0)
    // one-launch tiling (element-kernel code bit 8): rhs_spill_slots(dim, p, elements) values
    // for the block entries a tile does not own; null selects the colour launches. The
    // entries go to 7 arrays, one per neighbour offset delta = dx + 2 dy + 4 dz (1..7) of the
    // owner, at spill_off[delta]; array delta has extent spill_ext[d][delta_d] in direction d
    // (the DOF count, or P entries per tile where delta_d = 1). Set by launch_rhs.
    T* spill = nullptr;
    long spill_off[8] = {};
    int spill_ext[3][2] = {};
    // host side only: when set, a one-launch element loop leaves its fix-up to the caller and
    // describes it here (defer->spill null when there is none to do)
    rhs_fixup<T>* defer = nullptr;
    // set by launch_rhs (and for batches) when the rhs start value is not used: the owners of
    // the one-launch tiling, and the first colour to reach a DOF in the 2D colour launches,
    // store their entries instead of adding them
    bool store = false;
};

// variant: kernel tuning choice (0 = coefficient block held in registers, 1 = re-read from
// LDS, ..., 5 = element columns); bits 4..7 carry the shape of variant 5, bit 8 selects one
// launch over all tiles with a spill fix-up instead of the 2^dim parity-colour launches (2D;
// 3D variants 0, 1, 2 and 4 with forms without a k field or a source in the kernel;
// prm.spill set).
template <typename T>
hipError_t launch_rhs(int dim, int p, const T* u_prev, T* rhs, const dev_tables<T>* dirs,
                      const rhs_params<T>& prm, int variant, hipStream_t s);
constexpr int kRhsOneLaunch = 1 << 8;
// code bit 9: the rhs start value is not used (rhs := the element sums): the one-launch
// tiling then stores instead of adding and needs no zero fill; other paths fill first
constexpr int kRhsOverwrite = 1 << 9;
// values the spill array of the one-launch tiling needs for degree p and these element
// counts (the largest over the variants that use it)
long rhs_spill_slots(int dim, int p, const int elements[3]);
// a deferred fix-up (rhs_params::defer, launch_line_sweep_seg) as a kernel of its own over
// the DOFs of rhs
template <typename T>
hipError_t launch_rhs_fixup(T* rhs, const rhs_fixup<T>& fx, hipStream_t s);

// The element loop for any quadrature order, repeated interior knots and a degree per
// direction (1..3), one lane per element (adsg_rhs_gen.hip); dirs[d].first, nq and p set.
template <typename T>
hipError_t launch_rhs_general(int dim, const T* u_prev, T* rhs, const dev_tables<T>* dirs, const rhs_params<T>& prm,
                              hipStream_t s);

// Variant 5 (3D, p = 2 or 3, forms without a k field or a source in the kernel): element
// columns with the x-y contractions on the matrix cores (adsg_rhs_col.hip); shape 0..3
// selects columns per wave (bit 0: two) and chunks per z-range (bit 1: three).
template <typename T>
bool rhs3_col_supported(int p, const rhs_params<T>& prm);
template <typename T>
hipError_t launch_rhs3_col(int p, const T* u_prev, T* rhs, const dev_tables<T>* dirs, const rhs_params<T>& prm,
                           int shape, hipStream_t s);

// Batched 2D element kernels: one item per problem (degree p, no k field, no source).
template <typename T>
struct rhs2_item {
    const T* up = nullptr;
    T* rhs = nullptr;  // the problem's rhs (its first DOF)
    dev_tables<T> X, Y;
    rhs_params<T> prm;
    int ntx2[4] = {0, 0, 0, 0};   // tiles of colour c per tile row
    int first[4] = {0, 0, 0, 0};  // first block of the item in colour c's launch
};
// tiles of each colour for tile shape 0 (16 x 16 elements) or 1 (8 x 8)
void rhs2_batch_tiles(int shape, int elements_x, int elements_y, int ntx2[4], int count[4]);
// the problems of degree p in one launch per colour
template <typename T>
hipError_t launch_rhs2_batch(int p, int shape, const rhs2_item<T>* items, int nitems, int colour, int blocks,
                             hipStream_t s);
// spill values of a 2D problem in the one-launch tiling with tiles of tx x ty elements, and
// (with off, ext) their layout as rhs_params holds it
long rhs2_spill_layout(int p, int tx, int ty, const int elements[2], long off[8], int ext[3][2]);

// Batched line sweeps of unpivoted factors: one item per (problem, direction), problems of
// every degree (kl = ku = p, from f) in one launch; 64 lines per block (x lines staged
// through LDS, y lines one per lane), first = first block.
template <typename T>
struct sweep_item {
    T* u = nullptr;
    line_geometry g;
    dev_factors<T> f;
    int first = 0;
};
template <typename T>
hipError_t launch_sweep_batch(const sweep_item<T>* items, int nitems, int blocks, hipStream_t s);

// Batched segmented sweeps (unpivoted factors): one item per (problem, direction) with its
// packed rows (line_segments::pack, segmented for at most kBatchSegments segments of at
// most 32 rows), problems of every degree p in one launch, 64 lines per block of
// kBatchSegments waves; first = first block of the item.
constexpr int kBatchSegments = 9;
template <typename T>
struct seg_item {
    T* u = nullptr;
    line_geometry g;
    const T* pack = nullptr;
    int n = 0, m = 0, nseg = 0, p = 0;
    int first = 0;
    // when x_new is set (the last sweep of an FP32 step): the solved values are not stored;
    // x_new = x_prev + value and shadow = its FP32 copy are (same offsets as u; no boundary rows)
    const double* x_prev = nullptr;
    double* x_new = nullptr;
    float* shadow = nullptr;
};
template <typename T>
hipError_t launch_sweep_seg_batch(const seg_item<T>* items, int nitems, int blocks, hipStream_t s);

// d/dk of w^T rhs(u; k) for the gradient term of prm, added into out:
//   out_b += -dt sum_e sum_q w_q J_e B_b(x_q) sum_d grad_d (d_d u)(d_d w)
// on the tiles of the element loop (FP64; dofs = elements + p), in one launch and the fix-up
// of its spill arrays when prm.spill is set (rhs_spill_slots(dim, p, ...) values), else one
// launch per parity colour
template <typename T>
hipError_t launch_kgrad(int dim, int p, const T* u, const T* w, T* out, const dev_tables<T>* dirs,
                        const rhs_params<T>& prm, hipStream_t s);
// The 2D explicit adjoint step's two element passes in one launch over all tiles and one
// fix-up launch for both: out_r := the element loop of prm_r (k field set) for w, out_k +=
// the k-field gradient (prm_k's dt and gradient weights) of u and w; spill_r, spill_k: two
// spill arrays of rhs_spill_slots(2, p, ...) values (FP64)
template <typename T>
hipError_t launch_adjoint2(int p, const T* w, const T* u, T* out_r, T* out_k, const dev_tables<T>* dirs,
                           const rhs_params<T>& prm_r, const rhs_params<T>& prm_k, T* spill_r, T* spill_k,
                           hipStream_t s);

// Squared norm / error functional: sum_e sum_q w J phi(d), d = u_h - s, u_h the spline of
// the FP64 coefficients u (0 when u is null), s given at the Gauss points as planes
// [value | d/dx | d/dy (| d/dz)] (0 when sq is null); phi = d^2 (kind 0), |grad d|^2 (1),
// d^2 + |grad d|^2 (2); `blocks` partial sums in partial[], total in out[0].
template <typename TT>
hipError_t launch_norm(int dim, int p, const double* u, const double* sq, const dev_tables<TT>* dirs,
                       int kind, double* partial, int blocks, double* out, hipStream_t s);

// One direction's 1D Gram (M) and stiffness (K) matrices as rows of 2p+1 entries,
// M[i*(2p+1) + b] = M(i, i+b-p) (0 outside the matrix), n rows, in FP64.
struct kron_dir {
    const double* M = nullptr;
    const double* K = nullptr;
    int n = 0;
};
// rhs (=, or += with accumulate) (mass Mz(x)My(x)Mx - dt sum_d grad_d K_d (x) M (x) M) u: the
// element-loop rhs of separable integrands (3D). u and rhs in T; the contractions in FP64
// in both precisions (the stiffness rows sum to zero, and FP32 sums of their products lose
// the digits that the element kernels keep by adding in FP64).
template <typename T>
hipError_t launch_rhs_kron(int dim, int p, const T* u, T* rhs, const kron_dir* dirs, double mass,
                           double dt, const double* grad, bool accumulate, hipStream_t s);

// Load vector of the L2 projection of f (given at the Gauss points, fq[ix + NX*(iy +
// NY*iz)], ix = e_x*(p+1) + q_x, NX = elements_x*(p+1)), added into out (FP64)
template <typename TT>
hipError_t launch_project(int dim, int p, const double* fq, double* out, const dev_tables<TT>* dirs,
                          hipStream_t s);

// The spline u (FP64, nx x ny [x nz] coefficients) on a tensor grid of counts[d] points per
// direction: tab[d] holds per point the p+1 basis values then the p+1 derivatives, off[d]
// the first nonzero function (span - p); out: values, then with ders the d/dx, d/dy[, d/dz]
// planes, x fastest
hipError_t launch_sample(int dim, int p, const double* u, long nx, long ny, const double* const* tab,
                         const int* const* off, const int* counts, int ders, double* out, hipStream_t s);

template <typename T>
hipError_t launch_fill(T* p, long n, T v, hipStream_t s);

// out = a * in
template <typename T>
hipError_t launch_scale(T* out, const T* in, long n, T a, hipStream_t s);
// out += a * x
template <typename T>
hipError_t launch_axpy(T* out, const T* x, long n, T a, hipStream_t s);

template <typename T>
hipError_t launch_zero_boundary(int dim, T* u, int nx, int ny, int nz, hipStream_t s);

// *count += number of NaN / Inf entries among v[0..n)
hipError_t launch_count_nonfinite(const double* v, long n, unsigned* count, hipStream_t s);

hipError_t launch_convert_d2f(const double* in, float* out, long n, hipStream_t s);
hipError_t launch_convert_f2d(const float* in, double* out, long n, hipStream_t s);

// FP32 increment step: x_new = x_prev + delta (FP64), shadow = (float) x_new; with
// dirichlet, boundary coefficients of x_new are set to 0 exactly.
hipError_t launch_apply_increment(const double* x_prev, const float* delta, double* x_new,
                                  float* shadow, int dim, int nx, int ny, int nz, bool dirichlet,
                                  hipStream_t s);
// FP64 form: x_new = x_prev + delta, boundary coefficients set to 0 with dirichlet.
hipError_t launch_add_increment(const double* x_prev, const double* delta, double* x_new, int dim,
                                int nx, int ny, int nz, bool dirichlet, hipStream_t s);
// Residual of the fixed boundary rows: delta_b = 0 - (M_f u)_b, where M_f is the Kronecker
// product of the 1D Gram matrices with their boundary rows fixed (dimension.cpp:23-29).
// rows_d holds the fixed 1D matrix of direction d as n_d rows of 2p+1 entries (columns
// i-p .. i+p).
hipError_t launch_boundary_residual(float* delta, const float* shadow, int dim, int nx, int ny,
                                    int nz, int p, const float* rows_x, const float* rows_y,
                                    const float* rows_z, hipStream_t s);

}  // namespace adsg
