// CPU oracle for the ADS time step -- TEST INFRASTRUCTURE ONLY (see ads_oracle.h).
//
// A plain restatement of the reference's CPU step, kept deliberately close to the
// reference's own loop orders so that its results agree with the reference to
// rounding. It is the checker of the HIP path and is never linked into it.
#include "ads_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

namespace {

// util.hpp:17-26 -- lerp(t, a, b) and lerp(i, n, a, b)
inline double lerp_t(double t, double a, double b) { return (1 - t) * a + t * b; }
inline double lerp_i(int i, int n, double a, double b) {
    double t = static_cast<double>(i) / static_cast<double>(n);
    return lerp_t(t, a, b);
}
inline int clampi(int v, int lo, int hi) { return v < lo ? lo : v > hi ? hi : v; }

}  // namespace

extern "C" {

// ---------------------------------------------------------------------------------------
// splitmix64 (Steele, Lea, Flood 2014) -> (x >> 11) * 2^-53
unsigned long long orc_splitmix64(unsigned long long* state) {
    unsigned long long z = (*state += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

double orc_uniform(unsigned long long* state) {
    return static_cast<double>(orc_splitmix64(state) >> 11) * 0x1.0p-53;
}

// ---------------------------------------------------------------------------------------
// Gauss-Legendre nodes/weights on [-1,1] for 2..64 points, ascending, as tabulated by
// quad/gauss.hpp and indexed through gauss_data.cpp:8-38. Computed here from the
// definition: Newton iteration on the three-term Legendre recurrence in extended
// precision, rounded once to double, with the symmetry x_i = -x_{n-1-i} imposed.
int orc_gauss(int n, double* X, double* W) {
    if (n < 2 || n > 64) return -1;
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int i = 0; i < (n + 1) / 2; ++i) {
        long double x = std::cos(pi * (i + 0.75L) / (n + 0.5L));
        long double dp = 0;
        for (int it = 0; it < 100; ++it) {
            long double p0 = 1, p1 = x;
            for (int k = 2; k <= n; ++k) {
                long double p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (x * p1 - p0) / (x * x - 1);
            long double dx = p1 / dp;
            x -= dx;
            if (std::fabs(dx) < 1e-21L) {
                // one more evaluation at the converged point for the weight
                p0 = 1;
                p1 = x;
                for (int k = 2; k <= n; ++k) {
                    long double p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
                    p0 = p1;
                    p1 = p2;
                }
                dp = n * (x * p1 - p0) / (x * x - 1);
                break;
            }
        }
        long double w = 2 / ((1 - x * x) * dp * dp);
        int hi = n - 1 - i;  // x is the i-th largest root
        X[hi] = static_cast<double>(x);
        X[i] = -static_cast<double>(x);
        W[hi] = W[i] = static_cast<double>(w);
    }
    if (n % 2 == 1) X[n / 2] = 0.0;
    return 0;
}

// ---------------------------------------------------------------------------------------
// bspline.cpp:26-43 -- clamped knot vector with optional repeated interior nodes
int orc_knot_size(int p, int elements, int repeated_nodes) {
    int points = elements + 1;
    int r = repeated_nodes + 1;
    return 2 * (p + 1) + (points - 2) * r;
}

void orc_create_basis(double a, double b, int p, int elements, int repeated_nodes, double* knot) {
    int points = elements + 1;
    int r = repeated_nodes + 1;
    int ks = orc_knot_size(p, elements, repeated_nodes);
    for (int i = 0; i <= p; ++i) {
        knot[i] = a;
        knot[ks - i - 1] = b;
    }
    for (int i = 1; i < points - 1; ++i)
        for (int j = 0; j < r; ++j) knot[p + 1 + (i - 1) * r + j] = lerp_i(i, elements, a, b);
}

// bspline.cpp:61-81 -- span by bisection between begin_idx = p and end_idx = size-p-1
int orc_find_span(double x, const double* knot, int knot_size, int p) {
    int low = p;
    int high = knot_size - p - 1;
    if (x >= knot[high]) return high - 1;
    if (x <= knot[low]) return low;
    int idx = (low + high) / 2;
    while (x < knot[idx] || x >= knot[idx + 1]) {
        if (x < knot[idx])
            high = idx;
        else
            low = idx;
        idx = (low + high) / 2;
    }
    return idx;
}

// bspline.cpp:83-100 -- Cox-de Boor triangle (Piegl & Tiller A2.2)
void orc_eval_basis(int i, double x, const double* U, int p, double* out) {
    std::vector<double> left(p + 1), right(p + 1);
    out[0] = 1;
    for (int j = 1; j <= p; ++j) {
        left[j] = x - U[i + 1 - j];
        right[j] = U[i + j] - x;
        double saved = 0;
        for (int r = 0; r < j; ++r) {
            double tmp = out[r] / (right[r + 1] + left[j - r]);
            out[r] = saved + right[r + 1] * tmp;
            saved = left[j - r] * tmp;
        }
        out[j] = saved;
    }
}

// bspline.cpp:102-160 -- values and derivatives (Piegl & Tiller A2.3)
void orc_eval_basis_with_derivatives(int i, double x, const double* U, int p, int der,
                                     double* out) {
    const int m = p + 1;
    std::vector<double> ndu(m * m), a(2 * (der + 1 > m ? der + 1 : m)), left(m), right(m);
    const int ac = (der + 1 > m ? der + 1 : m);
    auto NDU = [&](int r, int c) -> double& { return ndu[r * m + c]; };
    auto A = [&](int r, int c) -> double& { return a[r * ac + c]; };
    auto OUT = [&](int k, int j) -> double& { return out[k * m + j]; };

    NDU(0, 0) = 1;
    for (int j = 1; j <= p; ++j) {
        left[j] = x - U[i + 1 - j];
        right[j] = U[i + j] - x;
        double saved = 0;
        for (int r = 0; r < j; ++r) {
            NDU(j, r) = right[r + 1] + left[j - r];
            double tmp = NDU(r, j - 1) / NDU(j, r);
            NDU(r, j) = saved + right[r + 1] * tmp;
            saved = left[j - r] * tmp;
        }
        NDU(j, j) = saved;
    }
    for (int j = 0; j <= p; ++j) OUT(0, j) = NDU(j, p);
    for (int r = 0; r <= p; ++r) {
        int s1 = 0, s2 = 1;
        A(0, 0) = 1;
        for (int k = 1; k <= der; ++k) {
            double d = 0;
            int rk = r - k, pk = p - k;
            if (r >= k) {
                A(s2, 0) = A(s1, 0) / NDU(pk + 1, rk);
                d = A(s2, 0) * NDU(rk, pk);
            }
            int j1 = (rk >= -1) ? 1 : -rk;
            int j2 = (r - 1 <= pk) ? k - 1 : p - r;
            for (int j = j1; j <= j2; ++j) {
                A(s2, j) = (A(s1, j) - A(s1, j - 1)) / NDU(pk + 1, rk + j);
                d += A(s2, j) * NDU(rk + j, pk);
            }
            if (r <= pk) {
                A(s2, k) = -A(s1, k - 1) / NDU(pk + 1, r);
                d += A(s2, k) * NDU(r, pk);
            }
            OUT(k, r) = d;
            std::swap(s1, s2);
        }
    }
    int r = p;
    for (int k = 1; k <= der; ++k) {
        for (int j = 0; j <= p; ++j) OUT(k, j) *= r;
        r *= (p - k);
    }
}

// bspline.cpp:162-172
int orc_first_nonzero_dofs(const double* knot, int knot_size, int p, int* out) {
    int e = 0;
    for (int i = p; i + 1 < knot_size - p; ++i)
        if (knot[i] != knot[i + 1]) out[e++] = i - p;
    return e;
}

// bspline.cpp:174-194
int orc_elements_supporting_dofs(const double* knot, int knot_size, int p, int* first, int* last) {
    int dofs = knot_size - p - 1;
    int e = 0;
    for (int i = 0; i < dofs; ++i) {
        int ee = e - 1;
        for (int j = 0; j < p + 1; ++j)
            if (knot[i + j] != knot[i + j + 1]) ++ee;
        first[i] = e;
        last[i] = ee;
        if (knot[i] != knot[i + 1]) ++e;
    }
    return dofs;
}

// ---------------------------------------------------------------------------------------
// basis_data.cpp:63-114 with elem_division = 1
int orc_basis_data(const double* knot, int knot_size, int p, int der, int quad_order, double* b,
                   double* x, double* J, double* w, int* first_dofs) {
    std::vector<double> points;
    points.push_back(knot[0]);  // bspline.hpp:20-27: the distinct knots
    for (int i = 1; i < knot_size; ++i)
        if (knot[i] != knot[i - 1]) points.push_back(knot[i]);
    int elements = static_cast<int>(points.size()) - 1;
    orc_first_nonzero_dofs(knot, knot_size, p, first_dofs);

    std::vector<double> X(quad_order), W(quad_order);
    if (orc_gauss(quad_order, X.data(), W.data()) != 0) return -1;
    for (int k = 0; k < quad_order; ++k) w[k] = W[k];

    const int m = p + 1;
    std::vector<double> tmp((der + 1) * m);
    for (int e = 0; e < elements; ++e) {
        double x1 = points[e], x2 = points[e + 1];
        J[e] = 0.5 * (x2 - x1);
        for (int k = 0; k < quad_order; ++k) {
            double t = 0.5 * (X[k] + 1);
            double xx = lerp_t(t, x1, x2);
            x[e * quad_order + k] = xx;
            int span = orc_find_span(xx, knot, knot_size, p);
            orc_eval_basis_with_derivatives(span, xx, knot, p, der, tmp.data());
            std::memcpy(b + (static_cast<long>(e) * quad_order + k) * (der + 1) * m, tmp.data(),
                        sizeof(double) * (der + 1) * m);
        }
    }
    return elements;
}

// ---------------------------------------------------------------------------------------
// band storage helper: lin::band_matrix(kl, ku, n) has row_offset = kl (band_matrix.hpp:31-32,
// 66-73) and column size 2kl+ku+1
static inline double& band_at(double* ab, int kl, int ku, int i, int j) {
    int ldab = 2 * kl + ku + 1;
    return ab[j * ldab + kl + ku + i - j];
}

// form_matrix.cpp:8-24 (and :26-42 for derivatives); loop order e, q, a, b
static void form_1d(double* ab, int p, int elements, int quad_order, int der, const double* b,
                    const double* J, const double* w, const int* first_dofs, int kind, double h) {
    const int m = p + 1;
    for (int e = 0; e < elements; ++e) {
        for (int q = 0; q < quad_order; ++q) {
            int first = first_dofs[e];
            int last = first + p;
            const double* bq = b + (static_cast<long>(e) * quad_order + q) * (der + 1) * m;
            for (int a = 0; a + first <= last; ++a) {
                for (int c = 0; c + first <= last; ++c) {
                    int ia = a + first, ib = c + first;
                    double va = bq[a], vb = bq[c];
                    double da = bq[m + a], db = bq[m + c];
                    double v;
                    if (kind == 0)
                        v = va * vb * w[q] * J[e];
                    else if (kind == 1)
                        v = da * db * w[q] * J[e];
                    else  // implicit.hpp:46-64: (va*vb + h*da*db) * w * J
                        v = (va * vb + h * da * db) * w[q] * J[e];
                    band_at(ab, p, p, ia, ib) += v;
                }
            }
        }
    }
}

void orc_gram_matrix_1d(double* ab, int n, int p, int elements, int quad_order, int der,
                        const double* b, const double* J, const double* w, const int* first_dofs) {
    (void)n;
    form_1d(ab, p, elements, quad_order, der, b, J, w, first_dofs, 0, 0);
}

void orc_stiffness_matrix_1d(double* ab, int n, int p, int elements, int quad_order, int der,
                             const double* b, const double* J, const double* w,
                             const int* first_dofs) {
    (void)n;
    form_1d(ab, p, elements, quad_order, der, b, J, w, first_dofs, 1, 0);
}

void orc_mass_plus_stiffness_1d(double* ab, int n, int p, int elements, int quad_order, int der,
                                const double* b, const double* J, const double* w,
                                const int* first_dofs, double h) {
    (void)n;
    form_1d(ab, p, elements, quad_order, der, b, J, w, first_dofs, 2, h);
}

// dimension.cpp:23-29
void orc_fix_dof(double* ab, int n, int p, int k) {
    int last = n - 1;
    for (int i = clampi(k - p, 0, last); i <= clampi(k + p, 0, last); ++i)
        band_at(ab, p, p, k, i) = 0;
    band_at(ab, p, p, k, k) = 1;
}

// dgbmv: y = op(A) x for count right-hand sides (band_matrix.hpp:91-106). a points at the
// ku+1+kl band rows; element (i,j) at a[(ku+i-j) + j*lda].
void orc_band_multiply(const double* a, int rows, int cols, int kl, int ku, int lda,
                       const double* x, double* y, int count, char trans) {
    bool nt = (trans == 'N' || trans == 'n');
    int leny = nt ? rows : cols;
    int lenx = nt ? cols : rows;
    for (int c = 0; c < count; ++c) {
        const double* xx = x + static_cast<long>(lenx) * c;
        double* yy = y + static_cast<long>(leny) * c;
        for (int i = 0; i < leny; ++i) yy[i] = 0;
        for (int j = 0; j < cols; ++j) {
            int i0 = std::max(0, j - ku), i1 = std::min(rows - 1, j + kl);
            if (nt) {
                double t = xx[j];
                for (int i = i0; i <= i1; ++i) yy[i] += t * a[(ku + i - j) + j * lda];
            } else {
                double t = 0;
                for (int i = i0; i <= i1; ++i) t += a[(ku + i - j) + j * lda] * xx[i];
                yy[j] += t;
            }
        }
    }
}

// ---------------------------------------------------------------------------------------
// LAPACK dgbtrf restated. For kl < NB (always here: kl = p <= 3) LAPACK's dgbtrf runs the
// unblocked dgbtf2: partial pivoting by columns, multipliers scaled by the reciprocal of the
// pivot, rank-1 update of the trailing band. Written with 1-based helpers to follow it.
int orc_dgbtrf(int m, int n, int kl, int ku, double* ab, int ldab, int* ipiv) {
    auto AB = [&](int i, int j) -> double& { return ab[(i - 1) + static_cast<long>(j - 1) * ldab]; };
    int info = 0;
    int kv = ku + kl;
    for (int j = ku + 2; j <= std::min(kv, n); ++j)
        for (int i = kv - j + 2; i <= kl; ++i) AB(i, j) = 0;
    int ju = 1;
    for (int j = 1; j <= std::min(m, n); ++j) {
        if (j + kv <= n)
            for (int i = 1; i <= kl; ++i) AB(i, j + kv) = 0;
        int km = std::min(kl, m - j);
        // idamax over AB(kv+1 .. kv+1+km, j)
        int jp = 1;
        double best = std::fabs(AB(kv + 1, j));
        for (int t = 2; t <= km + 1; ++t) {
            double v = std::fabs(AB(kv + t, j));
            if (v > best) {
                best = v;
                jp = t;
            }
        }
        ipiv[j - 1] = jp + j - 1;
        if (AB(kv + jp, j) != 0) {
            ju = std::max(ju, std::min(j + ku + jp - 1, n));
            if (jp != 1) {
                // dswap(ju-j+1, AB(kv+jp,j), ldab-1, AB(kv+1,j), ldab-1)
                for (int t = 0; t < ju - j + 1; ++t) {
                    double& x1 = ab[(kv + jp - 1) + static_cast<long>(j - 1) * ldab + t * (ldab - 1)];
                    double& x2 = ab[(kv) + static_cast<long>(j - 1) * ldab + t * (ldab - 1)];
                    std::swap(x1, x2);
                }
            }
            if (km > 0) {
                double r = 1.0 / AB(kv + 1, j);
                for (int t = 0; t < km; ++t) AB(kv + 2 + t, j) *= r;
                if (ju > j) {
                    // dger(km, ju-j, -1, AB(kv+2,j),1, AB(kv,j+1),ldab-1, AB(kv+1,j+1),ldab-1)
                    for (int c = 0; c < ju - j; ++c) {
                        double yv = ab[(kv - 1) + static_cast<long>(j) * ldab + c * (ldab - 1)];
                        if (yv != 0) {
                            double temp = -yv;
                            for (int t = 0; t < km; ++t) {
                                ab[(kv) + static_cast<long>(j) * ldab + c * (ldab - 1) + t] +=
                                    AB(kv + 2 + t, j) * temp;
                            }
                        }
                    }
                }
            }
        } else if (info == 0) {
            info = j;
        }
    }
    return info;
}

// LAPACK dgbtrs restated: 'N' = row interchanges + unit-L updates, then dtbsv (upper,
// no-transpose) per right-hand side; 'T' = dtbsv (upper, transpose) then L^T updates and
// interchanges in reverse order.
void orc_dgbtrs(char trans, int n, int kl, int ku, int nrhs, const double* ab, int ldab,
                const int* ipiv, double* b, int ldb) {
    auto AB = [&](int i, int j) -> double { return ab[(i - 1) + static_cast<long>(j - 1) * ldab]; };
    auto B = [&](int i, int c) -> double& { return b[(i - 1) + static_cast<long>(c - 1) * ldb]; };
    if (n == 0 || nrhs == 0) return;
    int kd = ku + kl + 1;
    int k = kl + ku;  // superdiagonals of U
    bool notran = (trans == 'N' || trans == 'n');
    if (notran) {
        if (kl > 0) {
            for (int j = 1; j <= n - 1; ++j) {
                int lm = std::min(kl, n - j);
                int l = ipiv[j - 1];
                if (l != j)
                    for (int c = 1; c <= nrhs; ++c) std::swap(B(l, c), B(j, c));
                for (int c = 1; c <= nrhs; ++c) {
                    double bj = B(j, c);
                    if (bj != 0) {
                        double temp = -bj;
                        for (int t = 1; t <= lm; ++t) B(j + t, c) += AB(kd + t, j) * temp;
                    }
                }
            }
        }
        for (int c = 1; c <= nrhs; ++c) {
            // dtbsv('U','N','N', n, k, ab, ldab, B(1,c), 1)
            for (int j = n; j >= 1; --j) {
                if (B(j, c) != 0) {
                    B(j, c) = B(j, c) / AB(k + 1, j);
                    double temp = B(j, c);
                    int L = k + 1 - j;
                    for (int i = j - 1; i >= std::max(1, j - k); --i) B(i, c) -= temp * AB(L + i, j);
                }
            }
        }
    } else {
        for (int c = 1; c <= nrhs; ++c) {
            // dtbsv('U','T','N', n, k, ab, ldab, B(1,c), 1)
            for (int j = 1; j <= n; ++j) {
                double temp = B(j, c);
                int L = k + 1 - j;
                for (int i = std::max(1, j - k); i <= j - 1; ++i) temp -= AB(L + i, j) * B(i, c);
                temp = temp / AB(k + 1, j);
                B(j, c) = temp;
            }
        }
        if (kl > 0) {
            for (int j = n - 1; j >= 1; --j) {
                int lm = std::min(kl, n - j);
                for (int c = 1; c <= nrhs; ++c) {
                    double temp = 0;
                    for (int t = 1; t <= lm; ++t) temp += B(j + t, c) * AB(kd + t, j);
                    B(j, c) += -1.0 * temp;
                }
                int l = ipiv[j - 1];
                if (l != j)
                    for (int c = 1; c <= nrhs; ++c) std::swap(B(l, c), B(j, c));
            }
        }
    }
}

// ---------------------------------------------------------------------------------------
// cyclic_transpose.hpp:17-63: out(i1,...,i_{R-1}, i0) = in(i0, i1, ..., i_{R-1})
void orc_cyclic_transpose(int rank, const int* sizes, const double* in, double* out) {
    long s0 = sizes[0];
    long rest = 1;
    for (int d = 1; d < rank; ++d) rest *= sizes[d];
    for (long r = 0; r < rest; ++r)
        for (long i0 = 0; i0 < s0; ++i0) out[r + rest * i0] = in[i0 + s0 * r];
}

// solver.hpp:30-41 (solve_with_dim_data), :148-160 (odd-rank swap), :163-166 (1D)
static void ads_solve_impl(char trans, int rank, const int* sizes, double* rhs, double* buf,
                           const double* const* ab, const int* const* ipiv, const int* kl,
                           const int* ku) {
    long total = 1;
    for (int d = 0; d < rank; ++d) total *= sizes[d];
    std::vector<int> cur_sizes(sizes, sizes + rank);
    double* cur = rhs;
    double* other = buf;
    for (int d = 0; d < rank; ++d) {
        int n = cur_sizes[0];
        int nrhs = static_cast<int>(total / n);
        int ldab = 2 * kl[d] + ku[d] + 1;  // solver_ctx.hpp:23-24
        orc_dgbtrs(trans, n, kl[d], ku[d], nrhs, ab[d], ldab, ipiv[d], cur, n);
        if (rank == 1) break;
        orc_cyclic_transpose(rank, cur_sizes.data(), cur, other);
        std::rotate(cur_sizes.begin(), cur_sizes.begin() + 1, cur_sizes.end());
        std::swap(cur, other);
    }
    if (cur != rhs) std::memcpy(rhs, cur, sizeof(double) * total);
}

void orc_ads_solve(int rank, const int* sizes, double* rhs, double* buf, const double* const* ab,
                   const int* const* ipiv, const int* kl, const int* ku) {
    ads_solve_impl('N', rank, sizes, rhs, buf, ab, ipiv, kl, ku);
}

void orc_ads_solve_transposed(int rank, const int* sizes, double* rhs, double* buf,
                              const double* const* ab, const int* const* ipiv, const int* kl,
                              const int* ku) {
    ads_solve_impl('T', rank, sizes, rhs, buf, ab, ipiv, kl, ku);
}

}  // extern "C"

// =======================================================================================
// simulation object: dimension (dimension.hpp:20-59) + simulation_2d/3d helpers
// (simulation_2d.hpp:38-140, simulation_3d.hpp:38-145)
struct orc_dim {
    int p = 0, elements = 0, n = 0, q = 0, der = 1;
    std::vector<double> knot, b, x, J, w;
    std::vector<int> first;
    std::vector<std::vector<double>> ab;  // per slot, LAPACK band storage kl=ku=p
    std::vector<std::vector<int>> ipiv;
    std::vector<int> factored;
};

struct orc_sim {
    int dim = 0;
    orc_dim d[3];
    long size = 0;
    std::vector<double> buffer;  // simulation_2d.cpp:19 buffer{shape()}
};

namespace {

struct fval {
    double val = 0, dx = 0, dy = 0, dz = 0;
};

inline const double* tab(const orc_dim& D, int e, int q) {
    return D.b.data() + (static_cast<long>(e) * D.q + q) * (D.der + 1) * (D.p + 1);
}

// eval_basis (simulation_2d.hpp:99-115, simulation_3d.hpp:98-118) for local indices
inline fval eval_basis2(const double* tx, const double* ty, int mx, int my, int ax, int ay) {
    double B1 = tx[ax], B2 = ty[ay];
    double dB1 = tx[mx + ax], dB2 = ty[my + ay];
    return {B1 * B2, dB1 * B2, B1 * dB2, 0};
}
inline fval eval_basis3(const double* tx, const double* ty, const double* tz, int mx, int my,
                        int mz, int ax, int ay, int az) {
    double B1 = tx[ax], B2 = ty[ay], B3 = tz[az];
    double dB1 = tx[mx + ax], dB2 = ty[my + ay], dB3 = tz[mz + az];
    return {B1 * B2 * B3, dB1 * B2 * B3, B1 * dB2 * B3, B1 * B2 * dB3};
}

// the integrand of compute_rhs (validation.hpp:110-115 for heat; other forms documented in
// ads_oracle.h)
inline double integrand(const orc_rhs_form* f, int dim, const fval& u, const fval& v, double k,
                        double src) {
    double gp = f->grad[0] * u.dx * v.dx + f->grad[1] * u.dy * v.dy;
    if (dim == 3) gp += f->grad[2] * u.dz * v.dz;
    double val;
    if (f->kfield)
        val = f->mass * u.val * v.val - f->dt * (k * gp);
    else
        val = f->mass * u.val * v.val - f->dt * gp;
    if (f->reaction != 0) val += f->dt * f->reaction * u.val * (1 - u.val) * v.val;
    if (f->source) val += f->dt * src * v.val;
    return val;
}

// one element of the hoisted loop (validation.hpp:100-118); the element contribution is
// returned in U (local first-index-fastest order)
void element_rhs(const orc_sim* s, const orc_rhs_form* f, const double* up, const int* e,
                 double* U, double* rhs_direct) {
    const int dim = s->dim;
    const orc_dim& X = s->d[0];
    const orc_dim& Y = s->d[1];
    const orc_dim& Z = s->d[2];
    const int mx = X.p + 1, my = Y.p + 1, mz = dim == 3 ? Z.p + 1 : 1;
    const long nx = X.n, ny = Y.n;
    const int fx = X.first[e[0]], fy = Y.first[e[1]], fz = dim == 3 ? Z.first[e[2]] : 0;
    double J = X.J[e[0]] * Y.J[e[1]];
    if (dim == 3) J = J * Z.J[e[2]];
    const int nloc = mx * my * mz;
    for (int i = 0; i < nloc; ++i) U[i] = 0;
    const int qz_n = dim == 3 ? Z.q : 1;
    for (int qx = 0; qx < X.q; ++qx)
        for (int qy = 0; qy < Y.q; ++qy)
            for (int qz = 0; qz < qz_n; ++qz) {
                double w = X.w[qx] * Y.w[qy];
                if (dim == 3) w = w * Z.w[qz];
                const double* tx = tab(X, e[0], qx);
                const double* ty = tab(Y, e[1], qy);
                const double* tz = dim == 3 ? tab(Z, e[2], qz) : nullptr;
                // eval_fun (simulation_2d.hpp:117-125): u += c * B over dofs_on_element
                fval u, kv, sv;
                for (int bx = 0; bx < mx; ++bx)
                    for (int by = 0; by < my; ++by)
                        for (int bz = 0; bz < mz; ++bz) {
                            long gi = (fx + bx) + nx * ((fy + by) + ny * (fz + bz));
                            fval B = dim == 3 ? eval_basis3(tx, ty, tz, mx, my, mz, bx, by, bz)
                                              : eval_basis2(tx, ty, mx, my, bx, by);
                            double c = up[gi];
                            u.val += c * B.val;
                            u.dx += c * B.dx;
                            u.dy += c * B.dy;
                            u.dz += c * B.dz;
                            if (f->kfield) kv.val += f->kfield[gi] * B.val;
                            if (f->source) sv.val += f->source[gi] * B.val;
                        }
                for (int ax = 0; ax < mx; ++ax)
                    for (int ay = 0; ay < my; ++ay)
                        for (int az = 0; az < mz; ++az) {
                            fval v = dim == 3 ? eval_basis3(tx, ty, tz, mx, my, mz, ax, ay, az)
                                              : eval_basis2(tx, ty, mx, my, ax, ay);
                            double val = integrand(f, dim, u, v, kv.val, sv.val);
                            if (rhs_direct) {
                                long gi = (fx + ax) + nx * ((fy + ay) + ny * (fz + az));
                                rhs_direct[gi] += val * w * J;
                            } else {
                                U[ax + mx * (ay + my * az)] += val * w * J;
                            }
                        }
            }
}

// update_global_rhs (simulation_2d.hpp:135-140)
void scatter(const orc_sim* s, const int* e, const double* U, double* rhs) {
    const orc_dim& X = s->d[0];
    const orc_dim& Y = s->d[1];
    const orc_dim& Z = s->d[2];
    const int mx = X.p + 1, my = Y.p + 1, mz = s->dim == 3 ? Z.p + 1 : 1;
    const long nx = X.n, ny = Y.n;
    const int fx = X.first[e[0]], fy = Y.first[e[1]], fz = s->dim == 3 ? Z.first[e[2]] : 0;
    for (int ax = 0; ax < mx; ++ax)
        for (int ay = 0; ay < my; ++ay)
            for (int az = 0; az < mz; ++az)
                rhs[(fx + ax) + nx * ((fy + ay) + ny * (fz + az))] += U[ax + mx * (ay + my * az)];
}

long element_count(const orc_sim* s) {
    long E = static_cast<long>(s->d[0].elements) * s->d[1].elements;
    if (s->dim == 3) E *= s->d[2].elements;
    return E;
}

// elements() iterates with the last index fastest (util/iter/product.hpp:45-54)
void element_index(const orc_sim* s, long lin, int* e) {
    if (s->dim == 2) {
        e[1] = static_cast<int>(lin % s->d[1].elements);
        e[0] = static_cast<int>(lin / s->d[1].elements);
        e[2] = 0;
    } else {
        e[2] = static_cast<int>(lin % s->d[2].elements);
        long r = lin / s->d[2].elements;
        e[1] = static_cast<int>(r % s->d[1].elements);
        e[0] = static_cast<int>(r / s->d[1].elements);
    }
}

}  // namespace

extern "C" {

orc_sim* orc_sim_create(int dim, const int* p, const int* elements, const double* a,
                        const double* b, const int* quad_order, const int* repeated_nodes) {
    if (dim != 2 && dim != 3 && dim != 1) return nullptr;
    auto* s = new orc_sim;
    s->dim = dim;
    s->size = 1;
    for (int i = 0; i < dim; ++i) {
        orc_dim& D = s->d[i];
        int rn = repeated_nodes ? repeated_nodes[i] : 0;
        D.p = p[i];
        D.q = quad_order && quad_order[i] > 0 ? quad_order[i] : p[i] + 1;  // config.hpp:27-28
        D.der = 1;
        int ks = orc_knot_size(D.p, elements[i], rn);
        D.knot.resize(ks);
        orc_create_basis(a ? a[i] : 0.0, b ? b[i] : 1.0, D.p, elements[i], rn, D.knot.data());
        D.n = ks - D.p - 1;
        // distinct knots give the element count
        int E = 0;
        for (int k = 1; k < ks; ++k)
            if (D.knot[k] != D.knot[k - 1]) ++E;
        D.elements = E;
        D.b.resize(static_cast<long>(E) * D.q * (D.der + 1) * (D.p + 1));
        D.x.resize(static_cast<long>(E) * D.q);
        D.J.resize(E);
        D.w.resize(D.q);
        D.first.resize(E);
        if (orc_basis_data(D.knot.data(), ks, D.p, D.der, D.q, D.b.data(), D.x.data(), D.J.data(),
                           D.w.data(), D.first.data()) < 0) {
            delete s;
            return nullptr;
        }
        // dimension.cpp:8-18: M{p,p,n} + gram_matrix_1d
        D.ab.assign(1, std::vector<double>(static_cast<long>(3 * D.p + 1) * D.n, 0.0));
        D.ipiv.assign(1, std::vector<int>(D.n, 0));
        D.factored.assign(1, 0);
        orc_gram_matrix_1d(D.ab[0].data(), D.n, D.p, E, D.q, D.der, D.b.data(), D.J.data(),
                           D.w.data(), D.first.data());
        s->size *= D.n;
    }
    s->buffer.assign(s->size, 0.0);
    return s;
}

void orc_sim_destroy(orc_sim* s) { delete s; }

int orc_sim_dofs(const orc_sim* s, int dir) { return s->d[dir].n; }
int orc_sim_size(const orc_sim* s) { return static_cast<int>(s->size); }

int orc_sim_fix_dof(orc_sim* s, int dir, int k) {
    orc_dim& D = s->d[dir];
    if (k < 0 || k >= D.n) return -1;
    orc_fix_dof(D.ab[0].data(), D.n, D.p, k);
    return 0;
}

int orc_sim_set_matrix(orc_sim* s, int dir, int slot, double mass, double stiff) {
    orc_dim& D = s->d[dir];
    if (slot < 1) return -1;
    if (static_cast<int>(D.ab.size()) <= slot) {
        D.ab.resize(slot + 1);
        D.ipiv.resize(slot + 1);
        D.factored.resize(slot + 1, 0);
    }
    D.ab[slot].assign(static_cast<long>(3 * D.p + 1) * D.n, 0.0);
    D.ipiv[slot].assign(D.n, 0);
    D.factored[slot] = 0;
    if (mass == 1.0) {
        orc_mass_plus_stiffness_1d(D.ab[slot].data(), D.n, D.p, D.elements, D.q, D.der, D.b.data(),
                                   D.J.data(), D.w.data(), D.first.data(), stiff);
    } else {
        std::vector<double> M(D.ab[slot].size(), 0.0), K(D.ab[slot].size(), 0.0);
        orc_gram_matrix_1d(M.data(), D.n, D.p, D.elements, D.q, D.der, D.b.data(), D.J.data(),
                           D.w.data(), D.first.data());
        orc_stiffness_matrix_1d(K.data(), D.n, D.p, D.elements, D.q, D.der, D.b.data(), D.J.data(),
                                D.w.data(), D.first.data());
        for (size_t i = 0; i < M.size(); ++i) D.ab[slot][i] = mass * M[i] + stiff * K[i];
    }
    return 0;
}

int orc_sim_get_matrix(const orc_sim* s, int dir, int slot, double* ab_out, int* ipiv_out) {
    const orc_dim& D = s->d[dir];
    if (slot < 0 || slot >= static_cast<int>(D.ab.size())) return -1;
    if (ab_out) std::memcpy(ab_out, D.ab[slot].data(), sizeof(double) * D.ab[slot].size());
    if (ipiv_out) std::memcpy(ipiv_out, D.ipiv[slot].data(), sizeof(int) * D.n);
    return 0;
}

// simulation_2d.hpp:56-59 prepare_matrices -> lin::factorize (band_solve.hpp:16-18)
int orc_sim_prepare_matrices(orc_sim* s) {
    int first_info = 0;
    for (int i = 0; i < s->dim; ++i) {
        orc_dim& D = s->d[i];
        for (size_t k = 0; k < D.ab.size(); ++k) {
            if (D.factored[k] || D.ab[k].empty()) continue;
            int info = orc_dgbtrf(D.n, D.n, D.p, D.p, D.ab[k].data(), 3 * D.p + 1, D.ipiv[k].data());
            D.factored[k] = 1;
            if (info != 0 && first_info == 0) first_info = info;
        }
    }
    return first_info;
}

void orc_sim_compute_rhs(const orc_sim* s, const orc_rhs_form* f, const double* u_prev,
                         double* rhs) {
    for (long i = 0; i < s->size; ++i) rhs[i] = 0;  // zero(rhs)
    long E = element_count(s);
    std::vector<double> U(64);
    int e[3];
    for (long l = 0; l < E; ++l) {
        element_index(s, l, e);
        if (f->as_shipped) {
            element_rhs(s, f, u_prev, e, U.data(), rhs);
        } else {
            element_rhs(s, f, u_prev, e, U.data(), nullptr);
            scatter(s, e, U.data(), rhs);
        }
    }
}

// the sequential element loop of orc_sim_compute_rhs over elements [first, first + count)
// of the reference's element order, added into rhs (not zeroed): a bounded sample of a full
// assembly for timing on one core
void orc_sim_compute_rhs_sample(const orc_sim* s, const orc_rhs_form* f, const double* u_prev,
                                double* rhs, long first, long count) {
    const long E = element_count(s);
    std::vector<double> U(64);
    int e[3];
    for (long l = std::max(0L, first); l < std::min(E, first + count); ++l) {
        element_index(s, l, e);
        element_rhs(s, f, u_prev, e, U.data(), nullptr);
        scatter(s, e, U.data(), rhs);
    }
}

void orc_sim_compute_rhs_threaded(const orc_sim* s, const orc_rhs_form* f, const double* u_prev,
                                  double* rhs, int threads) {
    for (long i = 0; i < s->size; ++i) rhs[i] = 0;
    long E = element_count(s);
    if (threads < 1) threads = 1;
    std::mutex lock;  // executor.synchronized (galois.hpp:27-32)
    auto worker = [&](int t) {
        std::vector<double> U(64);
        int e[3];
        long chunk = (E + threads - 1) / threads;
        long lo = t * chunk, hi = std::min(E, lo + chunk);
        for (long l = lo; l < hi; ++l) {
            element_index(s, l, e);
            element_rhs(s, f, u_prev, e, U.data(), nullptr);
            std::lock_guard<std::mutex> g(lock);
            scatter(s, e, U.data(), rhs);
        }
    };
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t) pool.emplace_back(worker, t);
    for (auto& th : pool) th.join();
}

// validation.hpp:31-41 (2D: zero the four sides); 3D: the six faces
void orc_sim_zero_boundary(const orc_sim* s, double* v) {
    long nx = s->d[0].n, ny = s->d[1].n, nz = s->dim == 3 ? s->d[2].n : 1;
    for (long k = 0; k < nz; ++k)
        for (long j = 0; j < ny; ++j)
            for (long i = 0; i < nx; ++i) {
                bool bnd = i == 0 || i == nx - 1 || j == 0 || j == ny - 1;
                if (s->dim == 3) bnd = bnd || k == 0 || k == nz - 1;
                if (bnd) v[i + nx * (j + ny * k)] = 0;
            }
}

static void sim_solve(orc_sim* s, double* rhs, const int* slots, char trans) {
    int sizes[3];
    const double* ab[3];
    const int* ip[3];
    int kl[3], ku[3];
    for (int i = 0; i < s->dim; ++i) {
        const orc_dim& D = s->d[i];
        int slot = slots ? slots[i] : 0;
        sizes[i] = D.n;
        ab[i] = D.ab[slot].data();
        ip[i] = D.ipiv[slot].data();
        kl[i] = ku[i] = D.p;
    }
    ads_solve_impl(trans, s->dim, sizes, rhs, s->buffer.data(), ab, ip, kl, ku);
}

// simulation_2d.hpp:41 solve -> ads_solve(rhs, buffer, x.data(), y.data())
void orc_sim_solve(orc_sim* s, double* rhs, const int* slots) { sim_solve(s, rhs, slots, 'N'); }
void orc_sim_solve_transposed(orc_sim* s, double* rhs, const int* slots) {
    sim_solve(s, rhs, slots, 'T');
}

int orc_sim_tables(const orc_sim* s, int dir, double* b, double* x, double* J, double* w,
                   int* first_dofs) {
    const orc_dim& D = s->d[dir];
    if (b) std::memcpy(b, D.b.data(), sizeof(double) * D.b.size());
    if (x) std::memcpy(x, D.x.data(), sizeof(double) * D.x.size());
    if (J) std::memcpy(J, D.J.data(), sizeof(double) * D.J.size());
    if (w) std::memcpy(w, D.w.data(), sizeof(double) * D.w.size());
    if (first_dofs) std::memcpy(first_dofs, D.first.data(), sizeof(int) * D.first.size());
    return D.elements;
}

// implicit.hpp:194-205 (L2norm): sum_e sum_q u(x_q)^2 w J
double orc_sim_l2norm_sq(const orc_sim* s, const double* u) {
    orc_rhs_form f{};
    double E = 0;
    long n = element_count(s);
    int e[3];
    const orc_dim& X = s->d[0];
    const orc_dim& Y = s->d[1];
    const orc_dim& Z = s->d[2];
    const int dim = s->dim;
    const int mx = X.p + 1, my = Y.p + 1, mz = dim == 3 ? Z.p + 1 : 1;
    (void)f;
    for (long l = 0; l < n; ++l) {
        element_index(s, l, e);
        double J = X.J[e[0]] * Y.J[e[1]];
        if (dim == 3) J = J * Z.J[e[2]];
        const int qz_n = dim == 3 ? Z.q : 1;
        for (int qx = 0; qx < X.q; ++qx)
            for (int qy = 0; qy < Y.q; ++qy)
                for (int qz = 0; qz < qz_n; ++qz) {
                    double w = X.w[qx] * Y.w[qy];
                    if (dim == 3) w = w * Z.w[qz];
                    const double* tx = tab(X, e[0], qx);
                    const double* ty = tab(Y, e[1], qy);
                    const double* tz = dim == 3 ? tab(Z, e[2], qz) : nullptr;
                    double val = 0;
                    for (int bx = 0; bx < mx; ++bx)
                        for (int by = 0; by < my; ++by)
                            for (int bz = 0; bz < mz; ++bz) {
                                long gi = (X.first[e[0]] + bx) +
                                          X.n * ((Y.first[e[1]] + by) +
                                                 static_cast<long>(Y.n) * ((dim == 3 ? Z.first[e[2]] : 0) + bz));
                                fval B = dim == 3 ? eval_basis3(tx, ty, tz, mx, my, mz, bx, by, bz)
                                                  : eval_basis2(tx, ty, mx, my, bx, by);
                                val += u[gi] * B.val;
                            }
                    E += val * val * w * J;
                }
    }
    return E;
}

// Derivative of w^T rhs(u; k) with respect to the coefficients k_b of a B-spline k field,
// for the form's gradient term -dt * k * sum_d grad_d u_d v_d (validation.hpp:110-115 with
// k as in flow.hpp:73-100): out_b += -dt sum_e sum_q w_q J_e B_b(x_q) sum_d grad_d
// (d_d u)(x_q) (d_d w)(x_q). No counterpart in the reference (it has no adjoint code).
// projection.hpp:55-110: element loop (x outermost), Gauss-point loop, local-DOF loop;
// J, w and B are products over the directions taken in order from 1.
void orc_sim_project(const orc_sim* s, const double* fq, double* v) {
    const int dim = s->dim;
    const orc_dim* D[3] = {&s->d[0], &s->d[1], &s->d[2]};
    const int ez_n = dim == 3 ? D[2]->elements : 1;
    const int qz_n = dim == 3 ? D[2]->q : 1;
    const int az_n = dim == 3 ? D[2]->p + 1 : 1;
    const long NX = static_cast<long>(D[0]->elements) * D[0]->q;
    const long NY = static_cast<long>(D[1]->elements) * D[1]->q;
    for (int ex = 0; ex < D[0]->elements; ++ex)
        for (int ey = 0; ey < D[1]->elements; ++ey)
            for (int ez = 0; ez < ez_n; ++ez) {
                const int e[3] = {ex, ey, ez};
                double J = 1;
                for (int i = 0; i < dim; ++i) J *= D[i]->J[e[i]];
                for (int qx = 0; qx < D[0]->q; ++qx)
                    for (int qy = 0; qy < D[1]->q; ++qy)
                        for (int qz = 0; qz < qz_n; ++qz) {
                            const int q[3] = {qx, qy, qz};
                            double w = 1;
                            for (int i = 0; i < dim; ++i) w *= D[i]->w[q[i]];
                            const long ix = static_cast<long>(ex) * D[0]->q + qx;
                            const long iy = static_cast<long>(ey) * D[1]->q + qy;
                            const long iz = dim == 3 ? static_cast<long>(ez) * D[2]->q + qz : 0;
                            const double fx = fq[ix + NX * (iy + NY * iz)];
                            for (int ax = 0; ax <= D[0]->p; ++ax)
                                for (int ay = 0; ay <= D[1]->p; ++ay)
                                    for (int az = 0; az < az_n; ++az) {
                                        const int a[3] = {ax, ay, az};
                                        double B = 1;
                                        long idx[3] = {0, 0, 0};
                                        for (int i = 0; i < dim; ++i) {
                                            B *= tab(*D[i], e[i], q[i])[a[i]];
                                            idx[i] = a[i] + D[i]->first[e[i]];
                                        }
                                        const long gi = idx[0] + D[0]->n * (idx[1] + static_cast<long>(D[1]->n) * idx[2]);
                                        v[gi] += fx * B * w * J;
                                    }
                        }
            }
}

// error(): element loop, Gauss-point loop; uu = eval(u, e, q), d = uu - fun(x),
// error += norm(d) * w * J
double orc_sim_error_sq(const orc_sim* s, const double* u, const double* sq, int kind) {
    const int dim = s->dim;
    const orc_dim& X = s->d[0];
    const orc_dim& Y = s->d[1];
    const orc_dim& Z = s->d[2];
    const int mx = X.p + 1, my = Y.p + 1, mz = dim == 3 ? Z.p + 1 : 1;
    const long NX = static_cast<long>(X.elements) * X.q, NY = static_cast<long>(Y.elements) * Y.q;
    const long Q = NX * NY * (dim == 3 ? static_cast<long>(Z.elements) * Z.q : 1);
    const long n = element_count(s);
    int e[3];
    double err = 0;
    for (long l = 0; l < n; ++l) {
        element_index(s, l, e);
        double J = X.J[e[0]] * Y.J[e[1]];
        if (dim == 3) J = J * Z.J[e[2]];
        const int qz_n = dim == 3 ? Z.q : 1;
        for (int qx = 0; qx < X.q; ++qx)
            for (int qy = 0; qy < Y.q; ++qy)
                for (int qz = 0; qz < qz_n; ++qz) {
                    double w = X.w[qx] * Y.w[qy];
                    if (dim == 3) w = w * Z.w[qz];
                    const double* tx = tab(X, e[0], qx);
                    const double* ty = tab(Y, e[1], qy);
                    const double* tz = dim == 3 ? tab(Z, e[2], qz) : nullptr;
                    fval v;
                    if (u)
                        for (int bx = 0; bx < mx; ++bx)
                            for (int by = 0; by < my; ++by)
                                for (int bz = 0; bz < mz; ++bz) {
                                    const long gi = (X.first[e[0]] + bx) +
                                                    X.n * ((Y.first[e[1]] + by) +
                                                           static_cast<long>(Y.n) * ((dim == 3 ? Z.first[e[2]] : 0) + bz));
                                    fval B = dim == 3 ? eval_basis3(tx, ty, tz, mx, my, mz, bx, by, bz)
                                                      : eval_basis2(tx, ty, mx, my, bx, by);
                                    v.val += u[gi] * B.val;
                                    v.dx += u[gi] * B.dx;
                                    v.dy += u[gi] * B.dy;
                                    v.dz += u[gi] * B.dz;
                                }
                    const long pt = (static_cast<long>(e[0]) * X.q + qx) +
                                    NX * ((static_cast<long>(e[1]) * Y.q + qy) +
                                          NY * (dim == 3 ? static_cast<long>(e[2]) * Z.q + qz : 0));
                    fval d = v;
                    d.val -= sq[pt];
                    if (kind != 0) {
                        d.dx -= sq[Q + pt];
                        d.dy -= sq[2 * Q + pt];
                        if (dim == 3) d.dz -= sq[3 * Q + pt];
                    }
                    const double g2 = d.dx * d.dx + d.dy * d.dy + (dim == 3 ? d.dz * d.dz : 0.0);
                    const double phi = kind == 0 ? d.val * d.val : kind == 1 ? g2 : d.val * d.val + g2;
                    err += phi * w * J;
                }
    }
    return err;
}

// eval: span by find_span, the p+1 nonzero functions (and derivatives) by A2.3, offset
// span - p, value = sum over the tensor product of local functions (x outermost)
void orc_sim_sample(const orc_sim* s, const double* u, const int* counts, const double* px,
                    const double* py, const double* pz, int ders, double* out) {
    const int dim = s->dim;
    const double* pts[3] = {px, py, pz};
    std::vector<std::vector<double>> tabv(dim), tabd(dim);
    std::vector<std::vector<int>> off(dim);
    for (int d = 0; d < dim; ++d) {
        const orc_dim& D = s->d[d];
        const int m = D.p + 1;
        tabv[d].resize(static_cast<size_t>(counts[d]) * m);
        tabd[d].resize(static_cast<size_t>(counts[d]) * m);
        off[d].resize(counts[d]);
        std::vector<double> buf(2 * m);
        for (int i = 0; i < counts[d]; ++i) {
            const int span = orc_find_span(pts[d][i], D.knot.data(), static_cast<int>(D.knot.size()), D.p);
            orc_eval_basis_with_derivatives(span, pts[d][i], D.knot.data(), D.p, 1, buf.data());
            for (int a = 0; a < m; ++a) {
                tabv[d][static_cast<size_t>(i) * m + a] = buf[a];
                tabd[d][static_cast<size_t>(i) * m + a] = buf[m + a];
            }
            off[d][i] = span - D.p;
        }
    }
    const int cz = dim == 3 ? counts[2] : 1;
    const long P = static_cast<long>(counts[0]) * counts[1] * cz;
    const int mx = s->d[0].p + 1, my = s->d[1].p + 1, mz = dim == 3 ? s->d[2].p + 1 : 1;
    const long nx = s->d[0].n, ny = s->d[1].n;
    for (int k = 0; k < cz; ++k)
        for (int j = 0; j < counts[1]; ++j)
            for (int i = 0; i < counts[0]; ++i) {
                double v = 0, gx = 0, gy = 0, gz = 0;
                for (int a = 0; a < mx; ++a)
                    for (int b = 0; b < my; ++b)
                        for (int c = 0; c < mz; ++c) {
                            const long gi = (off[0][i] + a) + nx * ((off[1][j] + b) + ny * (dim == 3 ? off[2][k] + c : 0));
                            const double bx = tabv[0][static_cast<size_t>(i) * mx + a];
                            const double by = tabv[1][static_cast<size_t>(j) * my + b];
                            const double bz = dim == 3 ? tabv[2][static_cast<size_t>(k) * mz + c] : 1.0;
                            const double dx = tabd[0][static_cast<size_t>(i) * mx + a];
                            const double dy = tabd[1][static_cast<size_t>(j) * my + b];
                            const double dz = dim == 3 ? tabd[2][static_cast<size_t>(k) * mz + c] : 0.0;
                            v += u[gi] * bx * by * bz;
                            gx += u[gi] * dx * by * bz;
                            gy += u[gi] * bx * dy * bz;
                            gz += u[gi] * bx * by * dz;
                        }
                const long pt = i + static_cast<long>(counts[0]) * (j + static_cast<long>(counts[1]) * k);
                out[pt] = v;
                if (ders) {
                    out[P + pt] = gx;
                    out[2 * P + pt] = gy;
                    if (dim == 3) out[3 * P + pt] = gz;
                }
            }
}

int orc_sim_quad_points(const orc_sim* s, int dir, double* out) {
    if (dir < 0 || dir >= s->dim) return -1;
    const orc_dim& D = s->d[dir];
    for (int e = 0; e < D.elements; ++e)
        for (int q = 0; q < D.q; ++q) out[e * D.q + q] = D.x[static_cast<long>(e) * D.q + q];
    return 0;
}

void orc_sim_kgrad(const orc_sim* s, const orc_rhs_form* f, const double* u, const double* w,
                   double* out) {
    const long n = element_count(s);
    int e[3];
    const orc_dim& X = s->d[0];
    const orc_dim& Y = s->d[1];
    const orc_dim& Z = s->d[2];
    const int dim = s->dim;
    const int mx = X.p + 1, my = Y.p + 1, mz = dim == 3 ? Z.p + 1 : 1;
    for (long l = 0; l < n; ++l) {
        element_index(s, l, e);
        double J = X.J[e[0]] * Y.J[e[1]];
        if (dim == 3) J = J * Z.J[e[2]];
        const int qz_n = dim == 3 ? Z.q : 1;
        auto gidx = [&](int bx, int by, int bz) {
            return (X.first[e[0]] + bx) +
                   X.n * ((Y.first[e[1]] + by) +
                          static_cast<long>(Y.n) * ((dim == 3 ? Z.first[e[2]] : 0) + bz));
        };
        for (int qx = 0; qx < X.q; ++qx)
            for (int qy = 0; qy < Y.q; ++qy)
                for (int qz = 0; qz < qz_n; ++qz) {
                    double wq = X.w[qx] * Y.w[qy];
                    if (dim == 3) wq = wq * Z.w[qz];
                    const double* tx = tab(X, e[0], qx);
                    const double* ty = tab(Y, e[1], qy);
                    const double* tz = dim == 3 ? tab(Z, e[2], qz) : nullptr;
                    fval gu, gw;
                    for (int bx = 0; bx < mx; ++bx)
                        for (int by = 0; by < my; ++by)
                            for (int bz = 0; bz < mz; ++bz) {
                                const long gi = gidx(bx, by, bz);
                                fval B = dim == 3 ? eval_basis3(tx, ty, tz, mx, my, mz, bx, by, bz)
                                                  : eval_basis2(tx, ty, mx, my, bx, by);
                                gu.dx += u[gi] * B.dx;
                                gu.dy += u[gi] * B.dy;
                                gu.dz += u[gi] * B.dz;
                                gw.dx += w[gi] * B.dx;
                                gw.dy += w[gi] * B.dy;
                                gw.dz += w[gi] * B.dz;
                            }
                    double g = f->grad[0] * gu.dx * gw.dx + f->grad[1] * gu.dy * gw.dy;
                    if (dim == 3) g += f->grad[2] * gu.dz * gw.dz;
                    g *= -f->dt * wq * J;
                    for (int bx = 0; bx < mx; ++bx)
                        for (int by = 0; by < my; ++by)
                            for (int bz = 0; bz < mz; ++bz) {
                                fval B = dim == 3 ? eval_basis3(tx, ty, tz, mx, my, mz, bx, by, bz)
                                                  : eval_basis2(tx, ty, mx, my, bx, by);
                                out[gidx(bx, by, bz)] += g * B.val;
                            }
                }
    }
}

// basic_simulation_2d.hpp:42-52, :291-316 (3D: basic_simulation_3d.hpp): the squared norm
// sum_e sum_q w J phi(u) with phi = u^2 (kind 0, L2), |grad u|^2 (kind 1, H1_0) or
// u^2 + |grad u|^2 (kind 2, H1); the reference's norm() returns its square root.
double orc_sim_norm_sq(const orc_sim* s, const double* u, int kind) {
    const long n = element_count(s);
    int e[3];
    const orc_dim& X = s->d[0];
    const orc_dim& Y = s->d[1];
    const orc_dim& Z = s->d[2];
    const int dim = s->dim;
    const int mx = X.p + 1, my = Y.p + 1, mz = dim == 3 ? Z.p + 1 : 1;
    double val = 0;
    for (long l = 0; l < n; ++l) {
        element_index(s, l, e);
        double J = X.J[e[0]] * Y.J[e[1]];
        if (dim == 3) J = J * Z.J[e[2]];
        const int qz_n = dim == 3 ? Z.q : 1;
        for (int qx = 0; qx < X.q; ++qx)
            for (int qy = 0; qy < Y.q; ++qy)
                for (int qz = 0; qz < qz_n; ++qz) {
                    double w = X.w[qx] * Y.w[qy];
                    if (dim == 3) w = w * Z.w[qz];
                    const double* tx = tab(X, e[0], qx);
                    const double* ty = tab(Y, e[1], qy);
                    const double* tz = dim == 3 ? tab(Z, e[2], qz) : nullptr;
                    fval v;
                    for (int bx = 0; bx < mx; ++bx)
                        for (int by = 0; by < my; ++by)
                            for (int bz = 0; bz < mz; ++bz) {
                                const long gi = (X.first[e[0]] + bx) +
                                                X.n * ((Y.first[e[1]] + by) +
                                                       static_cast<long>(Y.n) * ((dim == 3 ? Z.first[e[2]] : 0) + bz));
                                fval B = dim == 3 ? eval_basis3(tx, ty, tz, mx, my, mz, bx, by, bz)
                                                  : eval_basis2(tx, ty, mx, my, bx, by);
                                v.val += u[gi] * B.val;
                                v.dx += u[gi] * B.dx;
                                v.dy += u[gi] * B.dy;
                                v.dz += u[gi] * B.dz;
                            }
                    const double g2 = v.dx * v.dx + v.dy * v.dy + (dim == 3 ? v.dz * v.dz : 0.0);
                    const double phi = kind == 0 ? v.val * v.val : kind == 1 ? g2 : v.val * v.val + g2;
                    val += phi * w * J;
                }
    }
    return val;
}

// Greville abscissae xi_i = (t_{i+1} + ... + t_{i+p}) / p
void orc_greville(const double* knot, int knot_size, int p, double* out) {
    int n = knot_size - p - 1;
    for (int i = 0; i < n; ++i) {
        double s = 0;
        for (int k = 1; k <= p; ++k) s += knot[i + k];
        out[i] = p > 0 ? s / p : knot[i];
    }
}

// sum of `count` Gaussians exp(-|x-c|^2 / (2 sigma^2)) sampled at Greville points, plus
// uniform noise in [-noise, noise]. Draw order: per Gaussian its centre (x, y[, z]) then
// sigma; then one noise draw per coefficient in memory order.
void orc_init_gaussians(const orc_sim* s, unsigned long long seed, int count, double sigma_lo,
                        double sigma_hi, double centre_lo, double centre_hi, double noise,
                        double* u) {
    unsigned long long st = seed;
    std::vector<double> cx(count), cy(count), cz(count), sg(count);
    for (int g = 0; g < count; ++g) {
        cx[g] = centre_lo + (centre_hi - centre_lo) * orc_uniform(&st);
        cy[g] = centre_lo + (centre_hi - centre_lo) * orc_uniform(&st);
        cz[g] = s->dim == 3 ? centre_lo + (centre_hi - centre_lo) * orc_uniform(&st) : 0.0;
        sg[g] = sigma_lo + (sigma_hi - sigma_lo) * orc_uniform(&st);
    }
    std::vector<double> gx(s->d[0].n), gy(s->d[1].n), gz(s->dim == 3 ? s->d[2].n : 1, 0.0);
    orc_greville(s->d[0].knot.data(), static_cast<int>(s->d[0].knot.size()), s->d[0].p, gx.data());
    orc_greville(s->d[1].knot.data(), static_cast<int>(s->d[1].knot.size()), s->d[1].p, gy.data());
    if (s->dim == 3)
        orc_greville(s->d[2].knot.data(), static_cast<int>(s->d[2].knot.size()), s->d[2].p, gz.data());
    long nx = s->d[0].n, ny = s->d[1].n, nz = s->dim == 3 ? s->d[2].n : 1;
    for (long k = 0; k < nz; ++k)
        for (long j = 0; j < ny; ++j)
            for (long i = 0; i < nx; ++i) {
                double v = 0;
                for (int g = 0; g < count; ++g) {
                    double dx = gx[i] - cx[g], dy = gy[j] - cy[g], dz = gz[k] - cz[g];
                    double r2 = dx * dx + dy * dy + (s->dim == 3 ? dz * dz : 0.0);
                    v += std::exp(-r2 / (2 * sg[g] * sg[g]));
                }
                u[i + nx * (j + ny * k)] = v;
            }
    for (long l = 0; l < nx * ny * nz; ++l) u[l] += noise * (2 * orc_uniform(&st) - 1);
}

}  // extern "C"
