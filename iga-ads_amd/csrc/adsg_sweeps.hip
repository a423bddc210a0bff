// HIP line sweeps of the MI355X ADS solve (gfx950), one lane per line: the dgbtrs 'N'
// substitution (band_solve.hpp:20-31) of every line of one direction, in place, in the
// natural first-index-fastest layout, single or batched over many problems. y/z lines are read directly
// (consecutive lanes own consecutive lines, so each step is one coalesced access); x lines
// are contiguous and are staged through LDS, which is the transpose the reference performs
// with cyclic_transpose (cyclic_transpose.hpp:17-63) between its dgbtrs calls.
#include <hip/hip_runtime.h>

#include "adsg_kernels.hpp"

namespace adsg {

namespace {

constexpr int CH = 16;  // line-sweep chunk (values in flight per lane and per pass)

// ------------------------------------------------------------------------------------
// line sweeps
// ------------------------------------------------------------------------------------

// Forward elimination over one chunk [j0, j0+CH): w[0..KL] is the window b[j..j+KL]
// (partially eliminated); cur holds the raw chunk, nxt the raw next chunk. Column order of
// LAPACK's dgbtrs: interchange rows j and ipiv(j), then b[j+t] -= L(j+t,j) * b[j].
template <typename T, typename F, int KL, bool PIV>
__device__ __forceinline__ void forward_chunk(T (&cur)[CH], const T (&nxt)[CH], T (&w)[KL + 1],
                                              int j0, int n, const dev_factors<F>& f) {
#pragma unroll
    for (int k = 0; k < CH; ++k) {
        const int j = j0 + k;
        if (j < n) {
            if constexpr (PIV) {
                const int off = f.poff[j];
#pragma unroll
                for (int t = 1; t <= KL; ++t) {
                    if (off == t) {
                        T tmp = w[0];
                        w[0] = w[t];
                        w[t] = tmp;
                    }
                }
            }
            const T y = w[0];
#pragma unroll
            for (int t = 1; t <= KL; ++t)
                w[t] -= static_cast<T>(f.lower[static_cast<long>(j) * KL + t - 1]) * y;
            cur[k] = y;
#pragma unroll
            for (int t = 0; t < KL; ++t) w[t] = w[t + 1];
            const int src = k + KL + 1;
            w[KL] = (src < CH) ? cur[src < CH ? src : 0] : nxt[src >= CH ? src - CH : 0];
        }
    }
}

// Back substitution over one chunk [j0, j0+CH), highest index first: v[KU] = b[j],
// v[KU-t] = b[j-t]; x_j = b_j / U(j,j), then b[j-t] -= U(j-t, j) * x_j (dtbsv, upper,
// column order). prv holds the raw chunk below.
template <typename T, typename F, int KU>
__device__ __forceinline__ void backward_chunk(T (&cur)[CH], const T (&prv)[CH], T (&v)[KU + 1],
                                               int j0, const dev_factors<F>& f) {
#pragma unroll
    for (int k = CH - 1; k >= 0; --k) {
        const int j = j0 + k;
        if (j >= 0) {
            const T x = v[KU] * static_cast<T>(f.rdiag[j]);
#pragma unroll
            for (int t = 1; t <= KU; ++t)
                v[KU - t] -= static_cast<T>(f.upper[static_cast<long>(j) * KU + t - 1]) * x;
            cur[k] = x;
#pragma unroll
            for (int t = KU; t > 0; --t) v[t] = v[t - 1];
            const int src = k - KU - 1;
            v[0] = (src >= 0) ? cur[src >= 0 ? src : 0] : prv[src < 0 ? src + CH : 0];
        }
    }
}

template <typename T>
__device__ __forceinline__ void load_chunk(T (&r)[CH], const T* p, long S, int j0, int n) {
#pragma unroll
    for (int k = 0; k < CH; ++k) {
        const int j = j0 + k;
        r[k] = (j >= 0 && j < n) ? p[static_cast<long>(j) * S] : T(0);
    }
}

template <typename T>
__device__ __forceinline__ void store_chunk(const T (&r)[CH], T* p, long S, int j0, int n) {
#pragma unroll
    for (int k = 0; k < CH; ++k) {
        const int j = j0 + k;
        if (j >= 0 && j < n) p[static_cast<long>(j) * S] = r[k];
    }
}

// Lines with stride S > 1 (y and z directions): lane = line.
template <typename T, typename F, int KL, int KU, bool PIV>
__device__ __forceinline__ void strided_line(T* __restrict__ u, long l, long A, long B, long S,
                                             const dev_factors<F>& f) {
    T* p = u + (l % A) + (l / A) * B;
    const int n = f.n;
    T cur[CH], other[CH];

    load_chunk(cur, p, S, 0, n);
    load_chunk(other, p, S, CH, n);
    T w[KL + 1];
#pragma unroll
    for (int t = 0; t <= KL; ++t) w[t] = cur[t];
    for (int j0 = 0; j0 < n; j0 += CH) {
        forward_chunk<T, F, KL, PIV>(cur, other, w, j0, n, f);
        store_chunk(cur, p, S, j0, n);
#pragma unroll
        for (int k = 0; k < CH; ++k) cur[k] = other[k];
        load_chunk(other, p, S, j0 + 2 * CH, n);
    }

    const int top = n - CH;
    load_chunk(cur, p, S, top, n);
    load_chunk(other, p, S, top - CH, n);
    T v[KU + 1];
#pragma unroll
    for (int t = 0; t <= KU; ++t) v[KU - t] = cur[CH - 1 - t];
    for (int j0 = top; j0 > -CH; j0 -= CH) {
        backward_chunk<T, F, KU>(cur, other, v, j0, f);
        store_chunk(cur, p, S, j0, n);
#pragma unroll
        for (int k = 0; k < CH; ++k) cur[k] = other[k];
        load_chunk(other, p, S, j0 - 2 * CH, n);
    }
}

template <typename T, typename F, int KL, int KU, bool PIV>
__global__ __launch_bounds__(256) void k_sweep_strided(T* __restrict__ u, long n_lines, long A,
                                                       long B, long S, dev_factors<F> f) {
    const long l = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (l >= n_lines) return;
    strided_line<T, F, KL, KU, PIV>(u, l, A, B, S, f);
}

// Contiguous lines (x direction): one wave owns 64 lines; each chunk of CH values of its
// lines is loaded coalesced (16 lanes per 128-byte line segment), transposed through LDS
// into per-lane registers, and written back the same way.
template <typename T>
__device__ __forceinline__ void stage_in(T (&r)[CH], T (*tile)[CH + 1], const T* u, long line0,
                                         long n_lines, int n, int j0) {
    const int lane = threadIdx.x;
#pragma unroll
    for (int rr = 0; rr < CH; ++rr) {
        const int idx = rr * 64 + lane;
        const int ln = idx / CH, k = idx % CH;
        const long line = line0 + ln;
        const int j = j0 + k;
        tile[ln][k] = (line < n_lines && j >= 0 && j < n) ? u[line * n + j] : T(0);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < CH; ++k) r[k] = tile[lane][k];
    __syncthreads();
}

template <typename T>
__device__ __forceinline__ void stage_out(const T (&r)[CH], T (*tile)[CH + 1], T* u, long line0,
                                          long n_lines, int n, int j0) {
    const int lane = threadIdx.x;
#pragma unroll
    for (int k = 0; k < CH; ++k) tile[lane][k] = r[k];
    __syncthreads();
#pragma unroll
    for (int rr = 0; rr < CH; ++rr) {
        const int idx = rr * 64 + lane;
        const int ln = idx / CH, k = idx % CH;
        const long line = line0 + ln;
        const int j = j0 + k;
        if (line < n_lines && j >= 0 && j < n) u[line * n + j] = tile[ln][k];
    }
    __syncthreads();
}

template <typename T, typename F, int KL, int KU, bool PIV>
__device__ __forceinline__ void contig_lines(T* __restrict__ u, long line0, long n_lines,
                                             const dev_factors<F>& f, T (*tile)[CH + 1]) {
    const int n = f.n;
    T cur[CH], other[CH];

    stage_in(cur, tile, u, line0, n_lines, n, 0);
    stage_in(other, tile, u, line0, n_lines, n, CH);
    T w[KL + 1];
#pragma unroll
    for (int t = 0; t <= KL; ++t) w[t] = cur[t];
    for (int j0 = 0; j0 < n; j0 += CH) {
        forward_chunk<T, F, KL, PIV>(cur, other, w, j0, n, f);
        stage_out(cur, tile, u, line0, n_lines, n, j0);
#pragma unroll
        for (int k = 0; k < CH; ++k) cur[k] = other[k];
        stage_in(other, tile, u, line0, n_lines, n, j0 + 2 * CH);
    }

    const int top = n - CH;
    stage_in(cur, tile, u, line0, n_lines, n, top);
    stage_in(other, tile, u, line0, n_lines, n, top - CH);
    T v[KU + 1];
#pragma unroll
    for (int t = 0; t <= KU; ++t) v[KU - t] = cur[CH - 1 - t];
    for (int j0 = top; j0 > -CH; j0 -= CH) {
        backward_chunk<T, F, KU>(cur, other, v, j0, f);
        stage_out(cur, tile, u, line0, n_lines, n, j0);
#pragma unroll
        for (int k = 0; k < CH; ++k) cur[k] = other[k];
        stage_in(other, tile, u, line0, n_lines, n, j0 - 2 * CH);
    }
}

template <typename T, typename F, int KL, int KU, bool PIV>
__global__ __launch_bounds__(64) void k_sweep_contig(T* __restrict__ u, long n_lines,
                                                     dev_factors<F> f) {
    __shared__ T tile[64][CH + 1];
    contig_lines<T, F, KL, KU, PIV>(u, static_cast<long>(blockIdx.x) * 64, n_lines, f, tile);
}

// Lines with their own factors (the generalized ADS of solver.hpp:43-195, one matrix per
// line of the special direction): lane = line, factors interleaved by line (entry k of line
// l at k*L + l) so that the lanes of a wave read consecutive addresses. dgbtrs 'N' order:
// interchange, then column updates of L; back substitution by columns of U.
template <typename T>
__global__ __launch_bounds__(256) void k_sweep_own(T* __restrict__ u, long n_lines, long A, long B,
                                                   long S, int n, int kl, int ku,
                                                   const T* __restrict__ lower,
                                                   const T* __restrict__ upper,
                                                   const T* __restrict__ rdiag,
                                                   const int* __restrict__ poff, int pivoted) {
    const long l = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (l >= n_lines) return;
    T* p = S == 1 ? u + l * n : u + (l % A) + (l / A) * B;
    const long L = n_lines;
    for (int j = 0; j < n; ++j) {
        if (pivoted) {
            const int o = poff[j * L + l];
            if (o != 0) {
                const T t = p[j * S];
                p[j * S] = p[(j + o) * S];
                p[(j + o) * S] = t;
            }
        }
        const T bj = p[j * S];
        for (int t = 1; t <= kl && j + t < n; ++t)
            p[(j + t) * S] -= lower[(static_cast<long>(j) * kl + t - 1) * L + l] * bj;
    }
    for (int j = n - 1; j >= 0; --j) {
        const T xj = p[j * S] * rdiag[static_cast<long>(j) * L + l];
        p[j * S] = xj;
        for (int t = 1; t <= ku && j - t >= 0; ++t)
            p[(j - t) * S] -= upper[(static_cast<long>(j) * ku + t - 1) * L + l] * xj;
    }
}

// Many problems in one launch (unpivoted factors): block b belongs to the last item whose
// first <= b; x lines (S == 1) are staged through LDS, other lines are one per lane.
template <typename T, int P>
__device__ __forceinline__ void sweep_batch_item(const sweep_item<T>& it, int b, T (&tile)[64][CH + 1]) {
    const long line0 = static_cast<long>(b - it.first) * 64;
    if (it.g.S == 1) {
        contig_lines<T, T, P, P, false>(it.u, line0, it.g.n_lines, it.f, tile);
    } else {
        const long l = line0 + threadIdx.x;
        if (l < it.g.n_lines) strided_line<T, T, P, P, false>(it.u, l, it.g.A, it.g.B, it.g.S, it.f);
    }
}

// every degree in one launch (kl = ku = p of the item's factors)
template <typename T>
__global__ __launch_bounds__(64) void k_sweep_batch(const sweep_item<T>* __restrict__ items, int nitems) {
    __shared__ T tile[64][CH + 1];
    const int b = blockIdx.x;
    int lo = 0, hi = nitems - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (items[mid].first <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    const sweep_item<T>& it = items[lo];
    const int p = it.f.kl;  // uniform over the workgroup
    if (p == 1) sweep_batch_item<T, 1>(it, b, tile);
    if (p == 2) sweep_batch_item<T, 2>(it, b, tile);
    if (p == 3) sweep_batch_item<T, 3>(it, b, tile);
}

template <typename T, typename F, int KL, int KU, bool PIV>
hipError_t sweep_dispatch_piv(T* u, const line_geometry& g, const dev_factors<F>& f,
                              hipStream_t s) {
    if (g.n_lines <= 0 || f.n <= 0) return hipSuccess;
    if (g.S == 1) {
        const unsigned blocks = static_cast<unsigned>((g.n_lines + 63) / 64);
        hipLaunchKernelGGL((k_sweep_contig<T, F, KL, KU, PIV>), dim3(blocks), dim3(64), 0, s, u,
                           g.n_lines, f);
    } else {
        const unsigned blocks = static_cast<unsigned>((g.n_lines + 255) / 256);
        hipLaunchKernelGGL((k_sweep_strided<T, F, KL, KU, PIV>), dim3(blocks), dim3(256), 0, s, u,
                           g.n_lines, g.A, g.B, g.S, f);
    }
    return hipGetLastError();
}

template <typename T, typename F, int KL, int KU>
hipError_t sweep_dispatch(T* u, const line_geometry& g, const dev_factors<F>& f, hipStream_t s) {
    return f.pivoted ? sweep_dispatch_piv<T, F, KL, KU, true>(u, g, f, s)
                     : sweep_dispatch_piv<T, F, KL, KU, false>(u, g, f, s);
}

}  // namespace

template <typename T, typename F>
hipError_t launch_line_sweep(T* u, const line_geometry& g, const dev_factors<F>& f,
                             hipStream_t s) {
    const int kl = f.kl, ku = f.ku;
#define ADSG_SWEEP(KL_, KU_) \
    if (kl == KL_ && ku == KU_) return sweep_dispatch<T, F, KL_, KU_>(u, g, f, s);
    ADSG_SWEEP(1, 1)
    ADSG_SWEEP(1, 2)
    ADSG_SWEEP(1, 3)
    ADSG_SWEEP(2, 2)
    ADSG_SWEEP(2, 3)
    ADSG_SWEEP(2, 4)
    ADSG_SWEEP(3, 3)
    ADSG_SWEEP(3, 4)
    ADSG_SWEEP(3, 5)
    ADSG_SWEEP(3, 6)
#undef ADSG_SWEEP
    return hipErrorInvalidValue;
}


template <typename T>
hipError_t launch_sweep_batch(const sweep_item<T>* items, int nitems, int blocks, hipStream_t s) {
    if (blocks <= 0 || nitems <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sweep_batch<T>, dim3(blocks), dim3(64), 0, s, items, nitems);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_line_sweep_own(T* u, const line_geometry& g, const own_factors<T>& f,
                                 hipStream_t s) {
    if (g.n_lines <= 0 || f.n <= 0) return hipSuccess;
    const unsigned blocks = static_cast<unsigned>((g.n_lines + 255) / 256);
    hipLaunchKernelGGL(k_sweep_own<T>, dim3(blocks), dim3(256), 0, s, u, g.n_lines, g.A, g.B, g.S,
                       f.n, f.kl, f.ku, f.lower, f.upper, f.rdiag, f.poff, f.pivoted ? 1 : 0);
    return hipGetLastError();
}

template hipError_t launch_line_sweep_own<double>(double*, const line_geometry&,
                                                  const own_factors<double>&, hipStream_t);
template hipError_t launch_line_sweep_own<float>(float*, const line_geometry&,
                                                 const own_factors<float>&, hipStream_t);
template hipError_t launch_sweep_batch<double>(const sweep_item<double>*, int, int, hipStream_t);
template hipError_t launch_sweep_batch<float>(const sweep_item<float>*, int, int, hipStream_t);
template hipError_t launch_line_sweep<double, double>(double*, const line_geometry&,
                                                      const dev_factors<double>&, hipStream_t);
template hipError_t launch_line_sweep<float, float>(float*, const line_geometry&,
                                                    const dev_factors<float>&, hipStream_t);

}  // namespace adsg
