// Device set-up of the problems of a batch (declared in adsg_batch_setup.hpp): basis tables,
// Gram bands, band LU and packed sweep rows for every direction of every problem, each stage
// one launch over all of them. Every function below restates one host routine of
// adsg_setup.cpp for the batch's configurations (p + 1 Gauss points, simple knots) with the
// same operations in the same order; contraction is off so that no product and sum fuse
// into one rounding, and the results equal the host's bit for bit.
#include "adsg_batch_setup.hpp"

#include <hip/hip_runtime.h>

#pragma clang fp contract(off)

namespace adsg {

namespace {

__device__ inline double bs_interp(double t, double lo, double hi) { return (1 - t) * lo + t * hi; }

// knot i of the clamped knot vector with simple interior knots (make_knots; bspline.cpp:26-43)
__device__ inline double bs_knot(const bs_dir& d, int i) {
    const int size = 2 * (d.p + 1) + d.ne - 1;
    if (i <= d.p) return d.a;
    if (i >= size - 1 - d.p) return d.b;
    const double t = static_cast<double>(i - d.p) / static_cast<double>(d.ne);
    return bs_interp(t, d.a, d.b);
}

// span_of (bspline.cpp:61-81): bisection over [p, size - p - 1]; the loop halves the bracket
// every pass, and the pass cap only guards against malformed input
__device__ inline int bs_span(const bs_dir& d, double x) {
    const int size = 2 * (d.p + 1) + d.ne - 1;
    int lo = d.p, hi = size - d.p - 1;
    if (x >= bs_knot(d, hi)) return hi - 1;
    if (x <= bs_knot(d, lo)) return lo;
    int mid = (lo + hi) / 2;
    for (int pass = 0; pass < 64 && (x < bs_knot(d, mid) || x >= bs_knot(d, mid + 1)); ++pass) {
        if (x < bs_knot(d, mid))
            hi = mid;
        else
            lo = mid;
        mid = (lo + hi) / 2;
    }
    return mid;
}

// basis_and_derivative: values and first derivatives of the P+1 B-splines on the span
// (Piegl & Tiller A2.3, one derivative)
template <int P>
__device__ inline void bs_basis(const bs_dir& d, int span, double x, double (&val)[P + 1], double (&der)[P + 1]) {
    double tri[P + 1][P + 1];
    double left[P + 1], right[P + 1];
    tri[0][0] = 1;
#pragma unroll
    for (int j = 1; j <= P; ++j) {
        left[j] = x - bs_knot(d, span + 1 - j);
        right[j] = bs_knot(d, span + j) - x;
        double carry = 0;
#pragma unroll
        for (int r = 0; r < j; ++r) {
            tri[j][r] = right[r + 1] + left[j - r];
            const double t = tri[r][j - 1] / tri[j][r];
            tri[r][j] = carry + right[r + 1] * t;
            carry = left[j - r] * t;
        }
        tri[j][j] = carry;
    }
#pragma unroll
    for (int j = 0; j <= P; ++j) val[j] = tri[j][P];
#pragma unroll
    for (int r = 0; r <= P; ++r) {
        double dd = 0;
        double a0 = 0;
        constexpr int pk = P - 1;
        if (r >= 1) {
            a0 = 1.0 / tri[pk + 1][r - 1];
            dd = a0 * tri[r - 1][pk];
        }
        if (r <= pk) {
            const double a1 = -1.0 / tri[pk + 1][r];
            dd += a1 * tri[r][pk];
        }
        der[r] = dd * P;
    }
}

// Gauss point q of element e: the point, and the element's half length (build_direction)
__device__ inline void bs_point(const bs_dir& d, int e, int q, double& x, double& J) {
    const double x1 = bs_knot(d, d.p + e), x2 = bs_knot(d, d.p + e + 1);
    J = 0.5 * (x2 - x1);
    x = bs_interp(0.5 * (d.gx[q] + 1), x1, x2);
}

// build_direction's tables of one direction (threads tid, tid + nt, ... over (e, q))
template <typename T, int P>
__device__ void bs_tables_dir(const bs_dir& d, int tid, int nt) {
    constexpr int M = P + 1;
    T* val = static_cast<T*>(d.val);
    T* der = static_cast<T*>(d.der);
    T* wj = static_cast<T*>(d.wj);
    for (int k = tid; k < d.ne * M; k += nt) {
        const int e = k / M, q = k % M;
        double x, J;
        bs_point(d, e, q, x, J);
        double v[M], g[M];
        bs_basis<P>(d, bs_span(d, x), x, v, g);
#pragma unroll
        for (int i = 0; i < M; ++i) {
            val[static_cast<long>(k) * M + i] = static_cast<T>(v[i]);
            der[static_cast<long>(k) * M + i] = static_cast<T>(g[i]);
        }
        wj[k] = static_cast<T>(d.gw[q] * J);
    }
}

// band_form(alpha 1, beta 0) of one direction, column by column: entry (gi, j) at
// ab[j * (3P+1) + 2P + gi - j] sums alpha v_i v_j w J over the elements e = j-P .. j (in
// increasing order) and their Gauss points (in increasing order), the host's order
template <int P>
__device__ void bs_gram_dir(const bs_dir& d, int tid, int nt) {
    constexpr int M = P + 1, LD = 3 * P + 1;
    const double alpha = 1.0;
    for (int j = tid; j < d.n; j += nt) {
        double* col = d.ab + static_cast<long>(j) * LD;
        for (int r = 0; r < LD; ++r) col[r] = 0.0;
        const int e_lo = j - P > 0 ? j - P : 0, e_hi = j < d.ne - 1 ? j : d.ne - 1;
        for (int e = e_lo; e <= e_hi; ++e)
            for (int q = 0; q < M; ++q) {
                double x, J;
                bs_point(d, e, q, x, J);
                double v[M], g[M];
                bs_basis<P>(d, bs_span(d, x), x, v, g);
                double vj = v[0];
#pragma unroll
                for (int c = 1; c < M; ++c)
                    if (c == j - e) vj = v[c];
                const double w = d.gw[q];
#pragma unroll
                for (int a = 0; a < M; ++a) {
                    const double s = alpha * v[a] * vj * w * J;
                    col[2 * P + e + a - j] += s;
                }
            }
    }
}

// band_lu without row interchanges (the batch's sweeps take unpivoted factors; a column whose
// largest entry lies below the diagonal stops with info -1, a zero pivot with its index):
// one lane, the column updates depend on each other
template <int P>
__device__ int bs_lu_dir(const bs_dir& d) {
    constexpr int KL = P, KU = P, LD = 3 * P + 1, KV = 2 * P;
    const int n = d.n;
    double* ab = d.ab;
    auto at = [&](int i, int j) -> double& { return ab[static_cast<long>(j) * LD + KV + i - j]; };
    int ju = 0;
    for (int j = 0; j < n; ++j) {
        const int km = KL < n - 1 - j ? KL : n - 1 - j;
        int piv = j;
        double best = fabs(at(j, j));
        for (int i = j + 1; i <= j + km; ++i)
            if (fabs(at(i, j)) > best) {
                best = fabs(at(i, j));
                piv = i;
            }
        if (piv != j) return -1;
        if (at(j, j) == 0) return j + 1;
        const int jr = j + KU < n - 1 ? j + KU : n - 1;
        ju = ju > jr ? ju : jr;
        if (km > 0) {
            const double r = 1.0 / at(j, j);
            for (int i = j + 1; i <= j + km; ++i) at(i, j) *= r;
            for (int c = j + 1; c <= ju; ++c) {
                const double u = at(j, c);
                if (u != 0)
                    for (int i = j + 1; i <= j + km; ++i) at(i, c) -= at(i, j) * u;
            }
        }
    }
    return 0;
}

// factor_for_lines' arrays and segment_lines' packed rows from the factored band, by the NT
// lanes of the workgroup: rows of the arrays and of the packs per lane, then one lane per
// unit response (segment, tail or head position), which are independent of each other
template <typename T, int P>
__device__ void bs_pack_dir(const bs_dir& d, int tid, int nt) {
    constexpr int KL = P, KU = P, LD = 3 * P + 1, KV = 2 * P, W = 2 * (KL + KU) + 1;
    const int n = d.n;
    const double* ab = d.ab;
    T* lower = static_cast<T*>(d.lower);
    T* upper = static_cast<T*>(d.upper);
    T* rdiag = static_cast<T*>(d.rdiag);
    for (int j = tid; j < n; j += nt) {
        const double* cj = ab + static_cast<long>(j) * LD;
#pragma unroll
        for (int t = 1; t <= KL; ++t) lower[static_cast<long>(j) * KL + t - 1] = static_cast<T>(j + t < n ? cj[KV + t] : 0.0);
#pragma unroll
        for (int t = 1; t <= KU; ++t) upper[static_cast<long>(j) * KU + t - 1] = static_cast<T>(j - t >= 0 ? cj[KV - t] : 0.0);
        const double pv = cj[KV];
        rdiag[j] = static_cast<T>(pv != 0 ? 1.0 / pv : 0.0);
        d.poff[j] = 0;
    }
    if (d.nseg <= 0) return;
    // L(j, j-t), U(j, j+t) and 1 / U(j, j) of the factored band
    auto L = [&](int j, int t) { return ab[static_cast<long>(j - t) * LD + KV + t]; };
    auto U = [&](int j, int t) { return ab[static_cast<long>(j + t) * LD + KV - t]; };
    auto R = [&](int j) { return 1.0 / ab[static_cast<long>(j) * LD + KV]; };
    const int m = d.m, nseg = d.nseg, mseg = d.mseg;
    T* pack = static_cast<T*>(d.pack);
    // every row: [L | U | 1/U(j,j) | gf (0) | hb (0)], identity rows past the segments
    for (int row = tid; row < nseg * mseg; row += nt) {
        const int sg = row / mseg, k = row % mseg, j = sg * m + k;
        T* r = pack + static_cast<long>(row) * W;
        for (int c = 0; c < W; ++c) r[c] = T(0);
        if (k >= m || j >= n) {
            r[KL + KU] = T(1);
            continue;
        }
#pragma unroll
        for (int t = 1; t <= KL; ++t)
            if (j - t >= 0) r[t - 1] = static_cast<T>(L(j, t));
#pragma unroll
        for (int t = 1; t <= KU; ++t)
            if (j + t < n) r[KL + t - 1] = static_cast<T>(U(j, t));
        r[KL + KU] = static_cast<T>(R(j));
    }
    __syncthreads();  // the rows are written; the unit responses fill in their gf, hb slots
    const int nf = (nseg - 1) * KL;  // forward tasks: segments 1 .. nseg-1, tail positions
    for (int task = tid; task < nf + (nseg - 1) * KU; task += nt) {
        if (task < nf) {
            // forward unit response to tail position j0 - KL + q (gf); y[0 .. KL) holds
            // y_{j-1} .. y_{j-KL}
            const int s = 1 + task / KL, q = task % KL;
            const int j0 = s * m, j1 = n < j0 + m ? n : j0 + m;
            double y[KL];
#pragma unroll
            for (int t = 1; t <= KL; ++t) y[t - 1] = t == KL - q ? 1.0 : 0.0;
            for (int j = j0; j < j1; ++j) {
                double acc = 0;
#pragma unroll
                for (int t = 1; t <= KL; ++t) acc -= L(j, t) * y[t - 1];
#pragma unroll
                for (int t = KL - 1; t >= 1; --t) y[t] = y[t - 1];
                y[0] = acc;
                pack[static_cast<long>(s * mseg + j - j0) * W + KL + KU + 1 + q] = static_cast<T>(acc);
            }
        } else {
            // backward unit response to head position j1 + q (hb); z[0 .. KU) holds
            // y_{j+1} .. y_{j+KU}
            const int s = (task - nf) / KU, q = (task - nf) % KU;
            const int j0 = s * m, j1 = n < j0 + m ? n : j0 + m;
            const int lim = n < j1 + KU ? n : j1 + KU;
            double z[KU];
#pragma unroll
            for (int t = 1; t <= KU; ++t) z[t - 1] = t == q + 1 ? 1.0 : 0.0;
            for (int j = j1 - 1; j >= j0; --j) {
                double acc = 0;
#pragma unroll
                for (int t = 1; t <= KU; ++t)
                    if (j + t < lim) acc -= U(j, t) * z[t - 1];
                acc *= R(j);
#pragma unroll
                for (int t = KU - 1; t >= 1; --t) z[t] = z[t - 1];
                z[0] = acc;
                pack[static_cast<long>(s * mseg + j - j0) * W + 2 * KL + KU + 1 + q] = static_cast<T>(acc);
            }
        }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_bs_tables(const bs_dir* __restrict__ dirs) {
    const bs_dir d = dirs[blockIdx.x];
    if (d.p == 1) bs_tables_dir<T, 1>(d, threadIdx.x, 256);
    if (d.p == 2) bs_tables_dir<T, 2>(d, threadIdx.x, 256);
    if (d.p == 3) bs_tables_dir<T, 3>(d, threadIdx.x, 256);
}

__global__ __launch_bounds__(256) void k_bs_gram(const bs_dir* __restrict__ dirs) {
    const bs_dir d = dirs[blockIdx.x];
    if (d.p == 1) bs_gram_dir<1>(d, threadIdx.x, 256);
    if (d.p == 2) bs_gram_dir<2>(d, threadIdx.x, 256);
    if (d.p == 3) bs_gram_dir<3>(d, threadIdx.x, 256);
}

// one workgroup of 64 lanes per direction: lane 0 factors the band, then all lanes write the
// factor arrays and the packed rows
template <typename T>
__global__ __launch_bounds__(64) void k_bs_factor(const bs_dir* __restrict__ dirs) {
    const bs_dir d = dirs[blockIdx.x];
    __shared__ int info;
    if (threadIdx.x == 0) {
        info = d.p == 1 ? bs_lu_dir<1>(d) : d.p == 2 ? bs_lu_dir<2>(d) : bs_lu_dir<3>(d);
        *d.info = info;
    }
    __syncthreads();  // the factored band (global, written by lane 0) and info
    if (info != 0) return;
    if (d.p == 1) bs_pack_dir<T, 1>(d, threadIdx.x, 64);
    if (d.p == 2) bs_pack_dir<T, 2>(d, threadIdx.x, 64);
    if (d.p == 3) bs_pack_dir<T, 3>(d, threadIdx.x, 64);
}

}  // namespace

template <typename T>
hipError_t launch_batch_tables(const bs_dir* dirs, int count, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_bs_tables<T>, dim3(count), dim3(256), 0, s, dirs);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_batch_factors(const bs_dir* dirs, int count, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_bs_gram, dim3(count), dim3(256), 0, s, dirs);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_bs_factor<T>, dim3(count), dim3(64), 0, s, dirs);
    return hipGetLastError();
}

template hipError_t launch_batch_tables<double>(const bs_dir*, int, hipStream_t);
template hipError_t launch_batch_tables<float>(const bs_dir*, int, hipStream_t);
template hipError_t launch_batch_factors<double>(const bs_dir*, int, hipStream_t);
template hipError_t launch_batch_factors<float>(const bs_dir*, int, hipStream_t);

}  // namespace adsg
