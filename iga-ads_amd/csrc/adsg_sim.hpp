// Internal state of an adsg_sim handle, shared by the C ABI and the batch code.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/adsgpu.h"
#include "adsg_kernels.hpp"
#include "adsg_setup.hpp"

namespace adsg {

struct dev_slot {
    bool set = false;
    bool factored = false;
    int kl = 0, ku = 0;
    std::vector<double> ab;  // unfactored band, column size 2kl+ku+1
    adsg::line_factors lf;
    // device factors in the precision of the sweeps
    void* lower = nullptr;
    void* upper = nullptr;
    void* rdiag = nullptr;
    int* poff = nullptr;
    // segmented sweep (unpivoted factors only)
    adsg::line_segments seg;
    void* pack = nullptr;   // packed rows of the segmented sweep
    void* packt = nullptr;  // the same for the transposed solve
    double* arow = nullptr;  // FP64 rows of the unfactored matrix, for refinement
    // lines > 0: one matrix per line (generalized ADS), ab holds them all in the
    // reference's line order; the device factors are interleaved by line (own_factors)
    long lines = 0;
};

struct dev_dir {
    adsg::direction tab;
    std::vector<dev_slot> slots;
    void* val = nullptr;  // [elements][nq][p+1] in the step precision (padded, dev_tables)
    void* der = nullptr;
    void* wj = nullptr;
    void* bval = nullptr;  // val, der times the Gauss weight (padded like val)
    void* bder = nullptr;
    void* rows = nullptr;  // FP32 mode: slot-0 matrix as n rows of 2p+1 entries
    // the 1D Gram and stiffness matrices as rows of 2p+1 entries (step precision), for the
    // separable-integrand rhs (launch_rhs_kron)
    void* kM = nullptr;
    void* kK = nullptr;
    int* first = nullptr;  // first DOF of every element (bspline.cpp:162-172)
};

}  // namespace adsg

struct adsg_sim {
    int dim = 0;
    int precision = ADSG_FP64;
    int step_count = 0;
    double dt = 0;
    adsg::dev_dir dirs[3];
    long N = 0;
    hipStream_t stream = nullptr;
    bool owns_stream = true;
    int device = 0;  // the HIP device current at creation, which holds every buffer
    // a configuration outside the sum-factorised kernels' (quad_order != p+1, repeated
    // interior knots or a degree per direction): the general element loop assembles the rhs,
    // and the entry points that assume one degree and p+1 Gauss points refuse the handle
    bool general = false;
    bool mixed_p = false;
    // the table and factor pointers point into a batch's arenas and are not freed one by one
    bool in_arena = false;
    // u and u_prev are FP64 in both precisions; buffer, kfield and source are in the step
    // precision. In FP32 mode shadow[0]/shadow[1] are FP32 copies of u/u_prev (they swap
    // with them) and a step is computed as an increment: compute_rhs assembles
    // rhs - M u_prev in FP32 into buffer, ads_solve turns it into M^-1 (rhs - M u_prev) with
    // FP32 sweeps on FP64-computed factors and adds it to u_prev in FP64.
    void* vec[7] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    void* shadow[2] = {nullptr, nullptr};
    bool pending = false;            // FP32: buffer holds an increment based on u_prev
    bool pending_dirichlet = false;  // FP32: boundary rows of the increment were fixed
    bool prepared = false;
    int rhs_variant = 1;    // element-kernel tuning choice (set per degree at creation)
    int rhs_col_shape = 0;  // variant 5: columns per wave and chunks per z-range (ADSG_TUNE_RHS_COLUMNS)
    int sweep_variant = 0;  // 0: segmented line sweeps where possible, 1: one lane per line
    int rhs_kron = 0;       // 1: separable integrands through the 1D matrices (adsg_set_tuning)
    // 1: element kernels in one launch over all tiles, the entries a tile does not own
    // spilled to `spill` (rhs_spill_slots values, FP64-sized) and added after
    int rhs_tiling = 0;
    void* spill = nullptr;
    void* spill2 = nullptr;  // a second spill array of the same size (the 2D adjoint's paired passes)
    int tune_cache = 1;  // 1: adsg_autotune reuses / records the process's results per configuration
    std::string err;
    hipEvent_t* marks = nullptr;  // set only while adsg_time_phases records phase boundaries
    double* norm_scratch = nullptr;  // adsg_norm: block partial sums and the total
    unsigned* nonfinite = nullptr;   // adsg_check_finite: count of NaN / Inf entries
    // load vector of the source, b_i = sum_e sum_q w J f B_i (step precision): forms with a
    // source start their rhs from dt * b; recomputed after the source changes
    void* source_load = nullptr;
    bool source_load_ok = false;
    // adsg_forward_checkpointed / adsg_adjoint_checkpointed: u every `ck_every` steps, and the
    // recomputed states of one segment plus a slot for lambda
    double* ckpt = nullptr;
    double* ck_seg = nullptr;
    size_t ckpt_cap = 0, ck_seg_cap = 0;  // in coefficients
    int ck_steps = 0, ck_every = 0, ck_flags = 0;
    std::string ck_form;
    // captured step graphs keyed by (form, flags, u pointer, steps in graph)
    std::map<std::tuple<std::string, int, void*, int>, hipGraphExec_t> graphs;
};


namespace adsg {
// adsg_create with a shared stream (not owned by the handle) and, without `vectors`, no
// state vectors (the problems of a batch keep theirs in the batch's arenas)
int create_sim(const adsg_config* cfg, hipStream_t shared, bool vectors, adsg_sim** out);
}  // namespace adsg
