// HIP line sweeps of the MI355X ADS solve (gfx950): the dgbtrs 'N' substitution
// (band_solve.hpp:20-31) of every line of one direction, in place, in the natural
// first-index-fastest layout, and its transpose (dgbtrs 'T'). y/z lines are read directly
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
__global__ __launch_bounds__(256) void k_sweep_strided(T* __restrict__ u, long n_lines, long A,
                                                       long B, long S, dev_factors<F> f) {
    const long l = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (l >= n_lines) return;
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
__global__ __launch_bounds__(64) void k_sweep_contig(T* __restrict__ u, long n_lines,
                                                     dev_factors<F> f) {
    __shared__ T tile[64][CH + 1];
    const long line0 = static_cast<long>(blockIdx.x) * 64;
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

// ------------------------------------------------------------------------------------
// segmented line sweeps (unpivoted factors)
//
// A workgroup owns 64 lines; wave w owns segment w of all 64 (lane = line), so every
// line is solved by nseg waves at once and each coefficient is read and written once.
// Forward: y = y0 + G t, where y0 solves the segment with a zero tail and t is the true
// tail of the previous segment; the tails are chained through LDS by one wave. Backward
// likewise with the heads of the next segment. Recurrences are in row form:
//   A = L U:      y_j = b_j - sum_t L(j, j-t) y_{j-t},     x_j = (y_j - sum_t U(j, j+t) x_{j+t}) / U(j,j)
//   A^T = U^T L^T: y_j = (b_j - sum_t U(j-t, j) y_{j-t}) / U(j,j),  x_j = y_j - sum_t L(j+t, j) x_{j+t}
// With REF (FP32 data), the solve is followed by one refinement: the residual w - A z is
// formed in FP64 from the unfactored FP64 matrix and solved again.
// ------------------------------------------------------------------------------------
template <typename T, typename F, int KL, int KU, int MSEG, bool TRANS>
__device__ __forceinline__ void seg_solve(T (&vals)[MSEG], int len, int j0, int seg, int lane,
                                          T* bound, const seg_factors<F>& sf) {
    constexpr int KM = KL > KU ? KL : KU;
    constexpr int HF = TRANS ? KU : KL;  // forward history length
    constexpr int HB = TRANS ? KL : KU;  // backward history length
    const dev_factors<F>& f = sf.f;
    const int n = f.n, m = sf.m, nseg = sf.nseg;
    const F* gfw = TRANS ? sf.gft : sf.gf;  // [n*HF]
    const F* gbw = TRANS ? sf.hbt : sf.hb;  // [n*HB]

    // forward, zero tail
    {
        T hist[HF];
#pragma unroll
        for (int t = 0; t < HF; ++t) hist[t] = T(0);
#pragma unroll
        for (int k = 0; k < MSEG; ++k) {
            if (k < len) {
                const int j = j0 + k;
                T acc = vals[k];
#pragma unroll
                for (int t = 1; t <= HF; ++t) {
                    if (j - t >= 0) {
                        const F c = TRANS ? f.upper[static_cast<long>(j) * KU + t - 1]
                                          : f.lower[static_cast<long>(j - t) * KL + t - 1];
                        acc -= static_cast<T>(c) * hist[t - 1];
                    }
                }
                if constexpr (TRANS) acc *= static_cast<T>(f.rdiag[j]);
#pragma unroll
                for (int t = HF - 1; t > 0; --t) hist[t] = hist[t - 1];
                hist[0] = acc;
                vals[k] = acc;
            }
        }
#pragma unroll
        for (int t = 0; t < HF; ++t) {  // tail position len-HF+t
            T tv = T(0);
#pragma unroll
            for (int k = 0; k < MSEG; ++k)
                if (k == len - HF + t) tv = vals[k];
            bound[(seg * KM + t) * 64 + lane] = tv;
        }
    }
    __syncthreads();
    if (seg == 0) {
        for (int s = 1; s < nseg; ++s) {
            const int jt = s * m + (n - s * m < m ? n - s * m : m) - HF;  // first tail position of s
            T prev[HF];
#pragma unroll
            for (int t = 0; t < HF; ++t) prev[t] = bound[((s - 1) * KM + t) * 64 + lane];
#pragma unroll
            for (int t = 0; t < HF; ++t) {
                T acc = bound[(s * KM + t) * 64 + lane];
#pragma unroll
                for (int q = 0; q < HF; ++q)
                    acc += static_cast<T>(gfw[static_cast<long>(jt + t) * HF + q]) * prev[q];
                bound[(s * KM + t) * 64 + lane] = acc;
            }
        }
    }
    __syncthreads();
    if (seg > 0) {
        T tail[HF];
#pragma unroll
        for (int t = 0; t < HF; ++t) tail[t] = bound[((seg - 1) * KM + t) * 64 + lane];
#pragma unroll
        for (int k = 0; k < MSEG; ++k) {
            if (k < len) {
                T acc = vals[k];
#pragma unroll
                for (int q = 0; q < HF; ++q)
                    acc += static_cast<T>(gfw[static_cast<long>(j0 + k) * HF + q]) * tail[q];
                vals[k] = acc;
            }
        }
    }
    __syncthreads();  // every wave has read its tail before the heads overwrite it

    // backward, zero head
    {
        T hist[HB];
#pragma unroll
        for (int t = 0; t < HB; ++t) hist[t] = T(0);
#pragma unroll
        for (int k = MSEG - 1; k >= 0; --k) {
            if (k < len) {
                const int j = j0 + k;
                T acc = vals[k];
#pragma unroll
                for (int t = 1; t <= HB; ++t) {
                    if (j + t < n) {
                        const F c = TRANS ? f.lower[static_cast<long>(j) * KL + t - 1]
                                          : f.upper[static_cast<long>(j + t) * KU + t - 1];
                        acc -= static_cast<T>(c) * hist[t - 1];
                    }
                }
                if constexpr (!TRANS) acc *= static_cast<T>(f.rdiag[j]);
#pragma unroll
                for (int t = HB - 1; t > 0; --t) hist[t] = hist[t - 1];
                hist[0] = acc;
                vals[k] = acc;
            }
        }
#pragma unroll
        for (int t = 0; t < HB; ++t) bound[(seg * KM + t) * 64 + lane] = vals[t];
    }
    __syncthreads();
    if (seg == 0) {
        for (int s = nseg - 2; s >= 0; --s) {
            const int jh = s * m;  // first head position of segment s
            T next[HB];
#pragma unroll
            for (int t = 0; t < HB; ++t) next[t] = bound[((s + 1) * KM + t) * 64 + lane];
#pragma unroll
            for (int t = 0; t < HB; ++t) {
                T acc = bound[(s * KM + t) * 64 + lane];
#pragma unroll
                for (int q = 0; q < HB; ++q)
                    acc += static_cast<T>(gbw[static_cast<long>(jh + t) * HB + q]) * next[q];
                bound[(s * KM + t) * 64 + lane] = acc;
            }
        }
    }
    __syncthreads();
    if (seg < nseg - 1) {
        T head[HB];
#pragma unroll
        for (int t = 0; t < HB; ++t) head[t] = bound[((seg + 1) * KM + t) * 64 + lane];
#pragma unroll
        for (int k = 0; k < MSEG; ++k) {
            if (k < len) {
                T acc = vals[k];
#pragma unroll
                for (int q = 0; q < HB; ++q)
                    acc += static_cast<T>(gbw[static_cast<long>(j0 + k) * HB + q]) * head[q];
                vals[k] = acc;
            }
        }
    }
    __syncthreads();  // bound is free again
}

template <typename T, typename F, int KL, int KU, int MSEG, bool CONTIG, bool TRANS, bool REF>
__global__ __launch_bounds__(1024) void k_sweep_seg(T* __restrict__ u, long n_lines, long A,
                                                    long B, long S, seg_factors<F> sf) {
    extern __shared__ double smem_raw[];
    constexpr int KM = KL > KU ? KL : KU;
    constexpr int KW = KL + KU + 1;
    const int n = sf.f.n, m = sf.m, nseg = sf.nseg;
    const int lane = threadIdx.x & 63;
    const int seg = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const long line0 = static_cast<long>(blockIdx.x) * 64;
    const long line = line0 + lane;
    const int j0 = seg * m;
    const int len = n - j0 < m ? n - j0 : m;  // >= KM by construction

    T* bound = reinterpret_cast<T*>(smem_raw);  // [nseg][KM][64]
    T* halo = bound + nseg * KM * 64;           // REF: [nseg][KL+KU][64]
    T(*stage)[MSEG + 1] =
        reinterpret_cast<T(*)[MSEG + 1]>(halo + (REF ? nseg * (KL + KU) * 64 : 0)) + seg * 64;

    T vals[MSEG];
    T* p = nullptr;
    if constexpr (CONTIG) {
#pragma unroll
        for (int r = 0; r < MSEG; ++r) {
            const int idx = r * 64 + lane;
            const int ln = idx / MSEG, k = idx % MSEG;
            const long g = line0 + ln;
            stage[ln][k] = (g < n_lines && k < len) ? u[g * n + j0 + k] : T(0);
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < MSEG; ++k) vals[k] = stage[lane][k];
    } else {
        p = u + (line % A) + (line / A) * B;
        const bool act = line < n_lines;
#pragma unroll
        for (int k = 0; k < MSEG; ++k)
            vals[k] = (act && k < len) ? p[static_cast<long>(j0 + k) * S] : T(0);
    }

    T win[REF ? MSEG : 1];
    if constexpr (REF) {
#pragma unroll
        for (int k = 0; k < MSEG; ++k) win[k] = vals[k];
    }

    seg_solve<T, F, KL, KU, MSEG, TRANS>(vals, len, j0, seg, lane, bound, sf);

    if constexpr (REF) {
        // neighbours' values for the band product: last KL of the previous segment,
        // first KU of the next one
#pragma unroll
        for (int t = 0; t < KL; ++t) {
            T tv = T(0);
#pragma unroll
            for (int k = 0; k < MSEG; ++k)
                if (k == len - KL + t) tv = vals[k];
            halo[(seg * (KL + KU) + t) * 64 + lane] = tv;
        }
#pragma unroll
        for (int t = 0; t < KU; ++t) halo[(seg * (KL + KU) + KL + t) * 64 + lane] = vals[t];
        __syncthreads();
        T prevt[KL], nexth[KU];
#pragma unroll
        for (int t = 0; t < KL; ++t)
            prevt[t] = seg > 0 ? halo[((seg - 1) * (KL + KU) + t) * 64 + lane] : T(0);
#pragma unroll
        for (int t = 0; t < KU; ++t)
            nexth[t] = seg < nseg - 1 ? halo[((seg + 1) * (KL + KU) + KL + t) * 64 + lane] : T(0);
        T res[MSEG];
#pragma unroll
        for (int k = 0; k < MSEG; ++k) {
            double r = static_cast<double>(win[k]);
            if (k < len) {
                const int j = j0 + k;
#pragma unroll
                for (int t = -KL; t <= KU; ++t) {
                    const int kk = k + t;
                    if (j + t < 0 || j + t >= n) continue;
                    T z;
                    if (kk < 0) {
                        z = prevt[kk + KL < 0 ? 0 : kk + KL];
                    } else if (kk >= len) {
                        z = T(0);
#pragma unroll
                        for (int q = 0; q < KU; ++q)
                            if (kk - len == q) z = nexth[q];
                    } else {
                        z = vals[kk >= MSEG ? MSEG - 1 : kk];
                    }
                    r -= sf.arow[static_cast<long>(j) * KW + t + KL] * static_cast<double>(z);
                }
            }
            res[k] = k < len ? static_cast<T>(r) : T(0);
        }
        seg_solve<T, F, KL, KU, MSEG, TRANS>(res, len, j0, seg, lane, bound, sf);
#pragma unroll
        for (int k = 0; k < MSEG; ++k) vals[k] += res[k];
    }

    if constexpr (CONTIG) {
#pragma unroll
        for (int k = 0; k < MSEG; ++k) stage[lane][k] = vals[k];
        __syncthreads();
#pragma unroll
        for (int r = 0; r < MSEG; ++r) {
            const int idx = r * 64 + lane;
            const int ln = idx / MSEG, k = idx % MSEG;
            const long g = line0 + ln;
            if (g < n_lines && k < len) u[g * n + j0 + k] = stage[ln][k];
        }
    } else {
        if (line < n_lines) {
#pragma unroll
            for (int k = 0; k < MSEG; ++k)
                if (k < len) p[static_cast<long>(j0 + k) * S] = vals[k];
        }
    }
}

template <typename T, typename F, int KL, int KU, int MSEG, bool TRANS, bool REF>
hipError_t seg_dispatch_m(T* u, const line_geometry& g, const seg_factors<F>& sf, hipStream_t s) {
    constexpr int KM = KL > KU ? KL : KU;
    const unsigned blocks = static_cast<unsigned>((g.n_lines + 63) / 64);
    const unsigned threads = static_cast<unsigned>(sf.nseg * 64);
    size_t shm = static_cast<size_t>(sf.nseg) * KM * 64 * sizeof(T);
    if (REF) shm += static_cast<size_t>(sf.nseg) * (KL + KU) * 64 * sizeof(T);
    if (g.S == 1) {
        shm += static_cast<size_t>(sf.nseg) * 64 * (MSEG + 1) * sizeof(T);
        if (shm > 150 * 1024) return hipErrorInvalidValue;  // staging does not fit the LDS
        auto kern = k_sweep_seg<T, F, KL, KU, MSEG, true, TRANS, REF>;
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(shm));
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(threads), shm, s, u, g.n_lines, g.A, g.B, g.S, sf);
    } else {
        auto kern = k_sweep_seg<T, F, KL, KU, MSEG, false, TRANS, REF>;
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(shm));
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(threads), shm, s, u, g.n_lines, g.A, g.B, g.S, sf);
    }
    return hipGetLastError();
}

template <typename T, typename F, int KL, int KU, bool TRANS, bool REF>
hipError_t seg_dispatch(T* u, const line_geometry& g, const seg_factors<F>& sf, hipStream_t s) {
    if (g.n_lines <= 0 || sf.f.n <= 0) return hipSuccess;
    if (sf.mseg == 16) return seg_dispatch_m<T, F, KL, KU, 16, TRANS, REF>(u, g, sf, s);
    if (sf.mseg == 32) return seg_dispatch_m<T, F, KL, KU, 32, TRANS, REF>(u, g, sf, s);
    return hipErrorInvalidValue;
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

template <typename T, typename F>
hipError_t launch_line_sweep_seg(T* u, const line_geometry& g, const seg_factors<F>& f,
                                 hipStream_t s) {
    const int kl = f.f.kl, ku = f.f.ku;
    if (f.f.pivoted) return hipErrorInvalidValue;
    // refinement is an FP32 form (FP64 sweeps are already at the FP64 residual level)
    if (f.refine && (f.transposed || !f.arow)) return hipErrorInvalidValue;
#define ADSG_SEG(KL_, KU_)                                                         \
    if (kl == KL_ && ku == KU_) {                                                  \
        if (f.transposed) return seg_dispatch<T, F, KL_, KU_, true, false>(u, g, f, s); \
        if constexpr (sizeof(T) == sizeof(float))                                   \
            if (f.refine) return seg_dispatch<T, F, KL_, KU_, false, true>(u, g, f, s); \
        if (f.refine) return hipErrorInvalidValue;                                  \
        return seg_dispatch<T, F, KL_, KU_, false, false>(u, g, f, s);                  \
    }
    ADSG_SEG(1, 1)
    ADSG_SEG(2, 2)
    ADSG_SEG(3, 3)
#undef ADSG_SEG
    return hipErrorInvalidValue;
}

template hipError_t launch_line_sweep<double, double>(double*, const line_geometry&,
                                                      const dev_factors<double>&, hipStream_t);
template hipError_t launch_line_sweep<float, float>(float*, const line_geometry&,
                                                    const dev_factors<float>&, hipStream_t);
template hipError_t launch_line_sweep_seg<double, double>(double*, const line_geometry&,
                                                          const seg_factors<double>&, hipStream_t);
template hipError_t launch_line_sweep_seg<float, float>(float*, const line_geometry&,
                                                        const seg_factors<float>&, hipStream_t);

}  // namespace adsg
