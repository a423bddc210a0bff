// Device set-up of the problems of a batch (adsg_batch_setup.hip): the 1D basis tables of
// every direction (basis_data.cpp:63-114 for p+1 Gauss points and simple knots), the 1D Gram
// matrices (form_matrix.cpp:8-42, gram_matrix_1d of dimension.cpp:8-18), their band LU
// (dgbtrf, band_solve.hpp:16-18) and the packed rows of the segmented line sweeps, for all
// directions of all problems in one launch per stage. The arithmetic follows the host
// set-up (adsg_setup.cpp: make_knots, span_of, basis_and_derivative, band_form, band_lu,
// factor_for_lines, segment_lines) operation for operation, without contractions, so both
// produce the same bits.
#pragma once

#include <hip/hip_runtime.h>

namespace adsg {

// one direction of one problem; T-typed arrays are void* (the batch's step precision)
struct bs_dir {
    int p = 0, ne = 0, n = 0;  // degree, elements, DOFs (ne + p)
    double a = 0, b = 1;       // the interval
    const double* gx = nullptr;  // Gauss-Legendre nodes and weights on [-1, 1], p + 1 points
    const double* gw = nullptr;
    void* val = nullptr;  // [ne][p+1][p+1] values, derivatives; [ne][p+1] weight x Jacobian
    void* der = nullptr;
    void* wj = nullptr;
    double* ab = nullptr;  // FP64 band of the Gram matrix, column size 3p+1 (work array)
    void* lower = nullptr;  // factors as the line sweeps read them (line_factors layout)
    void* upper = nullptr;
    void* rdiag = nullptr;
    int* poff = nullptr;
    int m = 0, nseg = 0, mseg = 0;  // segment shape (nseg 0: no packed rows)
    void* pack = nullptr;           // [nseg * mseg][4p + 1] packed rows (line_segments::pack)
    int* info = nullptr;  // 0, or the first zero pivot (1-based), or -1 when a pivot is needed
};

// tables of dirs[0 .. count) (val, der, wj)
template <typename T>
hipError_t launch_batch_tables(const bs_dir* dirs, int count, hipStream_t s);
// Gram bands (ab), their unpivoted LU with the factors (lower, upper, rdiag, poff, info), and
// the packed rows where nseg > 0
template <typename T>
hipError_t launch_batch_factors(const bs_dir* dirs, int count, hipStream_t s);

}  // namespace adsg
