// Host-side set-up of one direction of the tensor-product B-spline space: knot vector,
// Gauss points, tabulated basis values/derivatives per element, 1D band matrices and their
// FP64 LU factors, rearranged into the flat per-line arrays the HIP sweeps read.
//
// This mirrors what the reference does once per run in dimension's constructor
// (dimension.cpp:8-18 -> bspline::create_basis, basis_data, gram_matrix_1d) and in
// prepare_matrices (simulation_2d.hpp:56-59 -> lin::factorize / dgbtrf).
#pragma once

#include <string>
#include <vector>

namespace adsg {

// ---- B-spline basis (bspline/bspline.cpp) --------------------------------------------
std::vector<double> make_knots(double a, double b, int p, int elements, int repeated_nodes);
int span_of(double x, const std::vector<double>& knot, int p);
void basis_and_derivative(int span, double x, const std::vector<double>& knot, int p, double* val,
                          double* der);
bool gauss_legendre(int n, double* X, double* W);

// Tables of one direction (basis_data.cpp:63-114): for element e and Gauss point q,
// val[(e*nq + q)*(p+1) + i] is the value of the i-th local basis function (global dof
// first[e] + i), der[...] its derivative; wj[e*nq + q] = w_q * J_e; xq the point.
struct direction {
    int p = 0;
    int elements = 0;
    int dofs = 0;
    int nq = 0;
    int repeated_nodes = 0;
    double a = 0, b = 1;
    std::vector<double> knot, val, der, wj, xq, w, J;
    std::vector<int> first;
};

bool build_direction(direction& d, int p, int elements, double a, double b, int quad_order,
                     int repeated_nodes, std::string& err);

// ---- 1D band matrices in LAPACK layout (band_matrix.hpp:66-73), kl = ku = p ----------
// alpha * Gram + beta * stiffness (form_matrix.cpp:8-42; implicit.hpp:46-64 for alpha = 1)
std::vector<double> band_form(const direction& d, double alpha, double beta);
// dimension::fix_dof (dimension.cpp:23-29)
void band_fix_dof(std::vector<double>& ab, int n, int p, int k);
// dgbtrf for kl, ku (partial pivoting); returns LAPACK info (0 ok, j>0 zero pivot)
int band_lu(std::vector<double>& ab, int n, int kl, int ku, std::vector<int>& ipiv);

// Factors of one band matrix rearranged for the GPU line sweeps:
//   lower[j*kl + t-1]   multiplier L(j+t, j) of column j            (t = 1..kl)
//   upper[j*ku_eff + t-1] = U(j-t, j), the column above the diagonal  (t = 1..ku_eff)
//   rdiag[j] = 1 / U(j, j),  poff[j] = ipiv[j] - 1 - j  (row interchange offset, 0..kl)
// ku_eff = ku when no interchange happened (then U keeps the original bandwidth), kl+ku
// otherwise. The FP64 band matrix itself is kept for refinement residuals.
struct line_factors {
    int n = 0, kl = 0, ku = 0, ku_eff = 0;
    bool pivoted = false;
    std::vector<double> lower, upper, rdiag;
    std::vector<int> poff;
    std::vector<double> band;  // the unfactored matrix, band storage ldab = 2kl+ku+1
};

int factor_for_lines(const std::vector<double>& ab_in, int n, int kl, int ku, line_factors& out);

// Segmentation of the lines of unpivoted factors for the segmented sweep: nseg segments of
// length m (mseg = register capacity 16 or 32, nseg <= 16), and the influence
// coefficients gf [n*kl] (tail of the previous segment on y_j) and hb [n*ku] (head of the
// next segment on x_j), computed in FP64 from the factors. Returns false when the line
// does not fit the scheme (pivoted factors, too short or too long lines).
struct line_segments {
    int m = 0, nseg = 0, mseg = 0;
    std::vector<double> gf, hb;    // A = L U
    std::vector<double> gft, hbt;  // A^T = U^T L^T: gft [n*ku], hbt [n*kl]
    // Everything row j of a sweep reads, packed per row (w = 2(kl+ku)+1 values) in blocks
    // of mseg rows per segment (row j = s*m + k at s*mseg + k; the rows k >= m of a block
    // and the rows j >= n are identity rows): for A = L U
    //   [L(j,j-t), t=1..kl | U(j,j+t), t=1..ku | 1/U(j,j) | gf[j][0..kl) | hb[j][0..ku)]
    // and for A^T (packt)
    //   [U(j-t,j), t=1..ku | L(j+t,j), t=1..kl | 1/U(j,j) | gft[j][0..ku) | hbt[j][0..kl)]
    // with zeros where the band leaves the matrix.
    std::vector<double> pack, packt;
};
// The fewest segments of <= 16 rows if that count is <= max_nseg16, else the fewest of
// <= 32 rows if <= max_nseg32 (the callers' limits keep the kernels' LDS within budget).
bool segment_lines(const line_factors& f, line_segments& out, int max_nseg16, int max_nseg32);
// The segment shape segment_lines picks for lines of n rows with half-bandwidth km:
// m rows per segment, nseg segments in blocks of mseg rows; false when none fits.
bool segment_shape(int n, int km, int max_nseg16, int max_nseg32, int& m, int& nseg, int& mseg);

}  // namespace adsg
