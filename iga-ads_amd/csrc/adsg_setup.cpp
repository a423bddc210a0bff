// Host set-up of the ADS step: B-spline tables, 1D band matrices, FP64 band LU.
#include "adsg_setup.hpp"

#include <algorithm>
#include <cmath>

namespace adsg {

namespace {
inline double interp(double t, double lo, double hi) { return (1 - t) * lo + t * hi; }
}  // namespace

// bspline.cpp:26-43: clamped ends (p+1 copies of a and b), interior breakpoints
// a + i/ne*(b-a) written as the reference's lerp, each repeated repeated_nodes+1 times.
std::vector<double> make_knots(double a, double b, int p, int elements, int repeated_nodes) {
    const int r = repeated_nodes + 1;
    const int size = 2 * (p + 1) + (elements - 1) * r;
    std::vector<double> k(size);
    for (int i = 0; i <= p; ++i) {
        k[i] = a;
        k[size - 1 - i] = b;
    }
    for (int i = 1; i < elements; ++i) {
        double t = static_cast<double>(i) / static_cast<double>(elements);
        double v = interp(t, a, b);
        for (int j = 0; j < r; ++j) k[p + 1 + (i - 1) * r + j] = v;
    }
    return k;
}

// bspline.cpp:61-81: bisection over [p, size-p-1], clamped at both ends
int span_of(double x, const std::vector<double>& k, int p) {
    int lo = p;
    int hi = static_cast<int>(k.size()) - p - 1;
    if (x >= k[hi]) return hi - 1;
    if (x <= k[lo]) return lo;
    int mid = (lo + hi) / 2;
    while (x < k[mid] || x >= k[mid + 1]) {
        (x < k[mid] ? hi : lo) = mid;
        mid = (lo + hi) / 2;
    }
    return mid;
}

// Values and first derivatives of the p+1 nonzero B-splines on knot span `span`
// (Piegl & Tiller, The NURBS Book, A2.3, restricted to one derivative; the reference's
// bspline.cpp:102-160 uses the same algorithm).
void basis_and_derivative(int span, double x, const std::vector<double>& k, int p, double* val,
                          double* der) {
    // tri[r][c]: upper triangle = basis functions of increasing degree, lower = knot gaps
    double tri[8][8];
    double left[8], right[8];
    tri[0][0] = 1;
    for (int j = 1; j <= p; ++j) {
        left[j] = x - k[span + 1 - j];
        right[j] = k[span + j] - x;
        double carry = 0;
        for (int r = 0; r < j; ++r) {
            tri[j][r] = right[r + 1] + left[j - r];
            double t = tri[r][j - 1] / tri[j][r];
            tri[r][j] = carry + right[r + 1] * t;
            carry = left[j - r] * t;
        }
        tri[j][j] = carry;
    }
    for (int j = 0; j <= p; ++j) val[j] = tri[j][p];
    if (p == 0) {
        der[0] = 0;
        return;
    }
    // first derivative: p * (N_{r-1,p-1}/gap - N_{r,p-1}/gap')
    for (int r = 0; r <= p; ++r) {
        double d = 0;
        double a0 = 0;
        const int rk = r - 1, pk = p - 1;
        if (r >= 1) {
            a0 = 1.0 / tri[pk + 1][rk];
            d = a0 * tri[rk][pk];
        }
        if (r <= pk) {
            double a1 = -1.0 / tri[pk + 1][r];
            d += a1 * tri[r][pk];
        }
        der[r] = d * p;
    }
}

// Gauss-Legendre rule on [-1,1] (quad/gauss.hpp tables hold the same nodes): Newton on the
// Legendre recurrence in long double, rounded once; nodes ascending, symmetric.
bool gauss_legendre(int n, double* X, double* W) {
    if (n < 2 || n > 64) return false;
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int i = 0; i < (n + 1) / 2; ++i) {
        long double x = std::cos(pi * (i + 0.75L) / (n + 0.5L));
        long double p1 = 0, p0 = 0, dp = 0;
        for (int it = 0; it < 100; ++it) {
            p0 = 1;
            p1 = x;
            for (int m = 2; m <= n; ++m) {
                long double p2 = ((2 * m - 1) * x * p1 - (m - 1) * p0) / m;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (x * p1 - p0) / (x * x - 1);
            long double step = p1 / dp;
            x -= step;
            if (std::fabs(step) < 1e-21L) break;
        }
        p0 = 1;
        p1 = x;
        for (int m = 2; m <= n; ++m) {
            long double p2 = ((2 * m - 1) * x * p1 - (m - 1) * p0) / m;
            p0 = p1;
            p1 = p2;
        }
        dp = n * (x * p1 - p0) / (x * x - 1);
        X[n - 1 - i] = static_cast<double>(x);
        X[i] = -static_cast<double>(x);
        W[i] = W[n - 1 - i] = static_cast<double>(2 / ((1 - x * x) * dp * dp));
    }
    if (n % 2) X[n / 2] = 0;
    return true;
}

bool build_direction(direction& d, int p, int elements, double a, double b, int quad_order,
                     int repeated_nodes, std::string& err) {
    if (p < 1 || p > 3) {
        err = "spline degree must be 1, 2 or 3";
        return false;
    }
    if (elements < 1) {
        err = "need at least one element per direction";
        return false;
    }
    if (repeated_nodes < 0 || repeated_nodes > p) {
        err = "repeated_nodes must be in 0..p (each interior knot repeated up to p+1 times)";
        return false;
    }
    int nq = quad_order > 0 ? quad_order : p + 1;  // config.hpp:27-28
    if (nq < 2 || nq > 64) {
        err = "quad_order must be in 2..64 (the Gauss rules of quad/gauss_data.cpp:8-38)";
        return false;
    }
    d.p = p;
    d.elements = elements;
    d.a = a;
    d.b = b;
    d.nq = nq;
    d.repeated_nodes = repeated_nodes;
    d.knot = make_knots(a, b, p, elements, repeated_nodes);
    d.dofs = static_cast<int>(d.knot.size()) - p - 1;
    std::vector<double> X(nq), W(nq);
    gauss_legendre(nq, X.data(), W.data());
    d.w = W;
    // breakpoints = distinct knots (bspline.hpp:20-27)
    std::vector<double> pts{d.knot[0]};
    for (size_t i = 1; i < d.knot.size(); ++i)
        if (d.knot[i] != d.knot[i - 1]) pts.push_back(d.knot[i]);
    const int m = p + 1;
    d.val.assign(static_cast<size_t>(elements) * nq * m, 0);
    d.der.assign(static_cast<size_t>(elements) * nq * m, 0);
    d.wj.assign(static_cast<size_t>(elements) * nq, 0);
    d.xq.assign(static_cast<size_t>(elements) * nq, 0);
    d.J.assign(elements, 0);
    d.first.assign(elements, 0);
    // first_nonzero_dofs (bspline.cpp:162-172)
    {
        int e = 0;
        for (int i = p; i + 1 < static_cast<int>(d.knot.size()) - p; ++i)
            if (d.knot[i] != d.knot[i + 1]) d.first[e++] = i - p;
    }
    for (int e = 0; e < elements; ++e) {
        double x1 = pts[e], x2 = pts[e + 1];
        double J = 0.5 * (x2 - x1);
        d.J[e] = J;
        for (int q = 0; q < nq; ++q) {
            double x = interp(0.5 * (X[q] + 1), x1, x2);
            d.xq[e * nq + q] = x;
            int s = span_of(x, d.knot, p);
            basis_and_derivative(s, x, d.knot, p, &d.val[(e * nq + q) * m], &d.der[(e * nq + q) * m]);
            d.wj[e * nq + q] = W[q] * J;
        }
    }
    return true;
}

std::vector<double> band_form(const direction& d, double alpha, double beta) {
    const int p = d.p, m = p + 1, ld = 3 * p + 1;
    std::vector<double> ab(static_cast<size_t>(ld) * d.dofs, 0.0);
    for (int e = 0; e < d.elements; ++e)
        for (int q = 0; q < d.nq; ++q) {
            const double* v = &d.val[(e * d.nq + q) * m];
            const double* g = &d.der[(e * d.nq + q) * m];
            double w = d.w[q], J = d.J[e];
            for (int i = 0; i < m; ++i)
                for (int j = 0; j < m; ++j) {
                    int gi = d.first[e] + i, gj = d.first[e] + j;
                    double s;
                    if (beta == 0)
                        s = alpha * v[i] * v[j] * w * J;
                    else if (alpha == 0)
                        s = beta * g[i] * g[j] * w * J;
                    else
                        s = (alpha * v[i] * v[j] + beta * g[i] * g[j]) * w * J;
                    ab[static_cast<size_t>(gj) * ld + 2 * p + gi - gj] += s;
                }
        }
    return ab;
}

void band_fix_dof(std::vector<double>& ab, int n, int p, int k) {
    const int ld = 3 * p + 1;
    for (int j = std::max(0, k - p); j <= std::min(n - 1, k + p); ++j)
        ab[static_cast<size_t>(j) * ld + 2 * p + k - j] = 0;
    ab[static_cast<size_t>(k) * ld + 2 * p] = 1;
}

// Band LU with partial pivoting by columns, the algorithm of LAPACK's dgbtf2 (which dgbtrf
// runs for these bandwidths): element (i,j) of the working matrix is at
// ab[j*ldab + kl+ku + i-j]; rows may fill in up to kl+ku above the diagonal.
int band_lu(std::vector<double>& ab, int n, int kl, int ku, std::vector<int>& ipiv) {
    const int ld = 2 * kl + ku + 1, kv = kl + ku;
    auto at = [&](int i, int j) -> double& { return ab[static_cast<size_t>(j) * ld + kv + i - j]; };
    ipiv.assign(n, 0);
    // fill-in rows above the original band start at zero
    for (int j = 0; j < n; ++j)
        for (int i = std::max(0, j - kv); i < j - ku; ++i) at(i, j) = 0;
    int info = 0;
    int ju = 0;  // last column touched by the U rows so far
    for (int j = 0; j < n; ++j) {
        int km = std::min(kl, n - 1 - j);
        int piv = j;
        double best = std::fabs(at(j, j));
        for (int i = j + 1; i <= j + km; ++i)
            if (std::fabs(at(i, j)) > best) {
                best = std::fabs(at(i, j));
                piv = i;
            }
        ipiv[j] = piv + 1;
        if (at(piv, j) == 0) {
            if (info == 0) info = j + 1;
            continue;
        }
        ju = std::max(ju, std::min(j + ku + (piv - j), n - 1));
        if (piv != j)
            for (int c = j; c <= ju; ++c) std::swap(at(piv, c), at(j, c));
        if (km > 0) {
            double r = 1.0 / at(j, j);
            for (int i = j + 1; i <= j + km; ++i) at(i, j) *= r;
            for (int c = j + 1; c <= ju; ++c) {
                double u = at(j, c);
                if (u != 0)
                    for (int i = j + 1; i <= j + km; ++i) at(i, c) -= at(i, j) * u;
            }
        }
    }
    return info;
}

int factor_for_lines(const std::vector<double>& ab_in, int n, int kl, int ku, line_factors& f) {
    const int ld = 2 * kl + ku + 1, kv = kl + ku;
    std::vector<double> ab = ab_in;
    std::vector<int> ipiv;
    int info = band_lu(ab, n, kl, ku, ipiv);
    f.n = n;
    f.kl = kl;
    f.ku = ku;
    f.band = ab_in;
    f.pivoted = false;
    for (int j = 0; j < n; ++j)
        if (ipiv[j] != j + 1) f.pivoted = true;
    f.ku_eff = f.pivoted ? std::min(kv, std::max(n - 1, 0)) : ku;
    if (f.ku_eff < 0) f.ku_eff = 0;
    f.lower.assign(static_cast<size_t>(n) * std::max(kl, 1), 0.0);
    f.upper.assign(static_cast<size_t>(n) * std::max(f.ku_eff, 1), 0.0);
    f.rdiag.assign(n, 0.0);
    f.poff.assign(n, 0);
    for (int j = 0; j < n; ++j) {
        for (int t = 1; t <= kl; ++t)
            if (j + t < n) f.lower[static_cast<size_t>(j) * kl + t - 1] = ab[static_cast<size_t>(j) * ld + kv + t];
        for (int t = 1; t <= f.ku_eff; ++t)
            if (j - t >= 0) f.upper[static_cast<size_t>(j) * f.ku_eff + t - 1] = ab[static_cast<size_t>(j) * ld + kv - t];
        double piv = ab[static_cast<size_t>(j) * ld + kv];
        f.rdiag[j] = piv != 0 ? 1.0 / piv : 0.0;
        f.poff[j] = ipiv[j] - 1 - j;
    }
    return info;
}

bool segment_shape(int n, int km, int max_nseg16, int max_nseg32, int& m, int& nseg, int& mseg) {
    mseg = nseg = m = 0;
    // the fewest segments of at most `rows` rows (balanced lengths, the last one at least km
    // rows long) within the caller's count limit: the serial chain between segments costs
    // more than the extra waves of shorter segments gain (measured on 130-row lines: 9
    // segments of 15 rows beat 13 of 10)
    auto pick = [&](int max_s, int rows) {
        if (mseg) return;
        const int s = (n + rows - 1) / rows;
        if (s < 2 || s > max_s) return;
        const int mm = (n + s - 1) / s;
        if (mm > rows || mm < km || n - (s - 1) * mm < km) return;
        mseg = rows;
        nseg = s;
        m = mm;
    };
    pick(max_nseg16, 16);
    pick(max_nseg32, 32);  // lines too long for segments of 16 rows
    return mseg != 0;
}

bool segment_lines(const line_factors& f, line_segments& out, int max_nseg16, int max_nseg32) {
    const int n = f.n, kl = f.kl, ku = f.ku_eff;
    const int km = std::max(kl, ku);
    if (f.pivoted || kl < 1 || ku < 1 || kl != ku) return false;
    int mseg = 0, nseg = 0, m = 0;
    if (!segment_shape(n, km, max_nseg16, max_nseg32, m, nseg, mseg)) return false;
    out.m = m;
    out.nseg = nseg;
    out.mseg = mseg;
    out.gf.assign(static_cast<size_t>(n) * kl, 0.0);
    out.hb.assign(static_cast<size_t>(n) * ku, 0.0);
    out.gft.assign(static_cast<size_t>(n) * ku, 0.0);
    out.hbt.assign(static_cast<size_t>(n) * kl, 0.0);
    auto L = [&](int j, int t) { return f.lower[static_cast<size_t>(j - t) * kl + t - 1]; };  // L(j, j-t)
    auto U = [&](int j, int t) { return f.upper[static_cast<size_t>(j + t) * ku + t - 1]; };  // U(j, j+t)
    std::vector<double> y(n + km + 1);
    for (int s = 0; s < nseg; ++s) {
        const int j0 = s * m, j1 = std::min(n, j0 + m);
        if (s > 0) {
            // forward: unit value at tail position j0-kl+q, zero right-hand side
            for (int q = 0; q < kl; ++q) {
                std::fill(y.begin(), y.end(), 0.0);
                y[j0 - kl + q] = 1.0;
                for (int j = j0; j < j1; ++j) {
                    double acc = 0;
                    for (int t = 1; t <= kl; ++t)
                        if (j - t >= j0 - kl) acc -= L(j, t) * y[j - t];
                    y[j] = acc;
                    out.gf[static_cast<size_t>(j) * kl + q] = acc;
                }
            }
        }
        if (s > 0) {
            // transposed forward (U^T): unit value at tail position j0-ku+q
            for (int q = 0; q < ku; ++q) {
                std::fill(y.begin(), y.end(), 0.0);
                y[j0 - ku + q] = 1.0;
                for (int j = j0; j < j1; ++j) {
                    double acc = 0;
                    for (int t = 1; t <= ku; ++t)
                        if (j - t >= j0 - ku) acc -= f.upper[static_cast<size_t>(j) * ku + t - 1] * y[j - t];
                    acc *= f.rdiag[j];
                    y[j] = acc;
                    out.gft[static_cast<size_t>(j) * ku + q] = acc;
                }
            }
        }
        if (s < nseg - 1) {
            // transposed backward (L^T): unit value at head position j1+q
            for (int q = 0; q < kl; ++q) {
                std::fill(y.begin(), y.end(), 0.0);
                y[j1 + q] = 1.0;
                for (int j = j1 - 1; j >= j0; --j) {
                    double acc = 0;
                    for (int t = 1; t <= kl; ++t)
                        if (j + t < std::min(n, j1 + kl)) acc -= f.lower[static_cast<size_t>(j) * kl + t - 1] * y[j + t];
                    y[j] = acc;
                    out.hbt[static_cast<size_t>(j) * kl + q] = acc;
                }
            }
        }
        if (s < nseg - 1) {
            // backward: unit value at head position j1+q
            for (int q = 0; q < ku; ++q) {
                std::fill(y.begin(), y.end(), 0.0);
                y[j1 + q] = 1.0;
                for (int j = j1 - 1; j >= j0; --j) {
                    double acc = 0;
                    for (int t = 1; t <= ku; ++t)
                        if (j + t < std::min(n, j1 + ku)) acc -= U(j, t) * y[j + t];
                    acc *= f.rdiag[j];
                    y[j] = acc;
                    out.hb[static_cast<size_t>(j) * ku + q] = acc;
                }
            }
        }
    }
    // packed rows for the sweep kernels (see adsg_setup.hpp): segment s owns the block of
    // mseg rows s*mseg .. s*mseg+mseg-1, its m rows first, then identity rows (no coupling,
    // unit diagonal, no influence), so that every segment runs mseg uniform steps
    const int np = nseg * mseg, w = 2 * (kl + ku) + 1;
    out.pack.assign(static_cast<size_t>(np) * w, 0.0);
    out.packt.assign(static_cast<size_t>(np) * w, 0.0);
    for (int row = 0; row < np; ++row) {
        const int sg = row / mseg, k = row % mseg, j = sg * m + k;
        double* r = &out.pack[static_cast<size_t>(row) * w];
        double* rt = &out.packt[static_cast<size_t>(row) * w];
        if (k >= m || j >= n) {
            r[kl + ku] = rt[ku + kl] = 1.0;
            continue;
        }
        for (int t = 1; t <= kl; ++t) {
            if (j - t >= 0) r[t - 1] = L(j, t);                                            // L(j, j-t)
            if (j + t < n) rt[ku + t - 1] = f.lower[static_cast<size_t>(j) * kl + t - 1];  // L(j+t, j)
        }
        for (int t = 1; t <= ku; ++t) {
            if (j + t < n) r[kl + t - 1] = U(j, t);                                        // U(j, j+t)
            if (j - t >= 0) rt[t - 1] = f.upper[static_cast<size_t>(j) * ku + t - 1];      // U(j-t, j)
        }
        r[kl + ku] = rt[ku + kl] = f.rdiag[j];
        for (int q = 0; q < kl; ++q) r[kl + ku + 1 + q] = out.gf[static_cast<size_t>(j) * kl + q];
        for (int q = 0; q < ku; ++q) r[2 * kl + ku + 1 + q] = out.hb[static_cast<size_t>(j) * ku + q];
        for (int q = 0; q < ku; ++q) rt[ku + kl + 1 + q] = out.gft[static_cast<size_t>(j) * ku + q];
        for (int q = 0; q < kl; ++q) rt[2 * ku + kl + 1 + q] = out.hbt[static_cast<size_t>(j) * kl + q];
    }
    return true;
}

}  // namespace adsg
