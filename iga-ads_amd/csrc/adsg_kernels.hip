// HIP kernels of the MI355X ADS time step (gfx950).
//
//  * line sweeps: the dgbtrs 'N' substitution (band_solve.hpp:20-31) of every line of one
//    direction, in place, in the natural first-index-fastest layout. y/z lines are read
//    directly (consecutive lanes own consecutive lines, so each step is one coalesced
//    access); x lines are contiguous and are staged through LDS in chunks, which is the
//    transpose the reference performs with cyclic_transpose (cyclic_transpose.hpp:17-63)
//    between its dgbtrs calls. Factors are uniform across lanes (scalar loads).
//  * element-loop RHS (compute_rhs, validation.hpp:95-119): one workgroup per tile of
//    elements, one lane per element, sum factorisation through the 1D tables, local
//    contributions combined in LDS, and tiles of one parity colour (2^d colours) adding
//    their DOF blocks into the global rhs without atomics.
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
// tail (last kl values of y) of the previous segment; the tails are chained through LDS
// by one wave. Backward likewise with the heads (first ku values of x) of the next
// segment. Recurrences are in row form: y_j = b_j - sum_t L(j, j-t) y_{j-t},
// x_j = (y_j - sum_t U(j, j+t) x_{j+t}) / U(j, j).
// ------------------------------------------------------------------------------------
template <typename T, typename F, int KL, int KU, int MSEG, bool CONTIG>
__global__ __launch_bounds__(1024) void k_sweep_seg(T* __restrict__ u, long n_lines, long A,
                                                    long B, long S, seg_factors<F> sf) {
    extern __shared__ double smem_raw[];
    constexpr int KM = KL > KU ? KL : KU;
    const dev_factors<F>& f = sf.f;
    const int n = f.n, m = sf.m, nseg = sf.nseg;
    const int lane = threadIdx.x & 63;
    const int seg = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const long line0 = static_cast<long>(blockIdx.x) * 64;
    const long line = line0 + lane;
    const int j0 = seg * m;
    const int len = n - j0 < m ? n - j0 : m;  // >= KM by construction

    T* bound = reinterpret_cast<T*>(smem_raw);  // [nseg][KM][64] tails, then heads
    T(*stage)[MSEG + 1] = reinterpret_cast<T(*)[MSEG + 1]>(bound + nseg * KM * 64) + seg * 64;

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

    // forward, zero tail
    {
        T hist[KL];
#pragma unroll
        for (int t = 0; t < KL; ++t) hist[t] = T(0);
#pragma unroll
        for (int k = 0; k < MSEG; ++k) {
            if (k < len) {
                const int j = j0 + k;
                T acc = vals[k];
#pragma unroll
                for (int t = 1; t <= KL; ++t)
                    if (j - t >= 0) acc -= static_cast<T>(f.lower[static_cast<long>(j - t) * KL + t - 1]) * hist[t - 1];
#pragma unroll
                for (int t = KL - 1; t > 0; --t) hist[t] = hist[t - 1];
                hist[0] = acc;
                vals[k] = acc;
            }
        }
#pragma unroll
        for (int t = 0; t < KL; ++t) {  // tail position len-KL+t
            T tv = T(0);
#pragma unroll
            for (int k = 0; k < MSEG; ++k)
                if (k == len - KL + t) tv = vals[k];
            bound[(seg * KM + t) * 64 + lane] = tv;
        }
    }
    __syncthreads();
    if (seg == 0) {
        for (int s = 1; s < nseg; ++s) {
            const int jt = s * m + (n - s * m < m ? n - s * m : m) - KL;  // first tail position of s
            T prev[KL];
#pragma unroll
            for (int t = 0; t < KL; ++t) prev[t] = bound[((s - 1) * KM + t) * 64 + lane];
#pragma unroll
            for (int t = 0; t < KL; ++t) {
                T acc = bound[(s * KM + t) * 64 + lane];
#pragma unroll
                for (int q = 0; q < KL; ++q)
                    acc += static_cast<T>(sf.gf[static_cast<long>(jt + t) * KL + q]) * prev[q];
                bound[(s * KM + t) * 64 + lane] = acc;
            }
        }
    }
    __syncthreads();
    if (seg > 0) {
        T tail[KL];
#pragma unroll
        for (int t = 0; t < KL; ++t) tail[t] = bound[((seg - 1) * KM + t) * 64 + lane];
#pragma unroll
        for (int k = 0; k < MSEG; ++k) {
            if (k < len) {
                T acc = vals[k];
#pragma unroll
                for (int q = 0; q < KL; ++q)
                    acc += static_cast<T>(sf.gf[static_cast<long>(j0 + k) * KL + q]) * tail[q];
                vals[k] = acc;
            }
        }
    }
    __syncthreads();  // every wave has read its tail before the heads overwrite it

    // backward, zero head
    {
        T hist[KU];
#pragma unroll
        for (int t = 0; t < KU; ++t) hist[t] = T(0);
#pragma unroll
        for (int k = MSEG - 1; k >= 0; --k) {
            if (k < len) {
                const int j = j0 + k;
                T acc = vals[k];
#pragma unroll
                for (int t = 1; t <= KU; ++t)
                    if (j + t < n) acc -= static_cast<T>(f.upper[static_cast<long>(j + t) * KU + t - 1]) * hist[t - 1];
                acc *= static_cast<T>(f.rdiag[j]);
#pragma unroll
                for (int t = KU - 1; t > 0; --t) hist[t] = hist[t - 1];
                hist[0] = acc;
                vals[k] = acc;
            }
        }
#pragma unroll
        for (int t = 0; t < KU; ++t) bound[(seg * KM + t) * 64 + lane] = vals[t];
    }
    __syncthreads();
    if (seg == 0) {
        for (int s = nseg - 2; s >= 0; --s) {
            const int jh = s * m;  // first head position of segment s
            T next[KU];
#pragma unroll
            for (int t = 0; t < KU; ++t) next[t] = bound[((s + 1) * KM + t) * 64 + lane];
#pragma unroll
            for (int t = 0; t < KU; ++t) {
                T acc = bound[(s * KM + t) * 64 + lane];
#pragma unroll
                for (int q = 0; q < KU; ++q)
                    acc += static_cast<T>(sf.hb[static_cast<long>(jh + t) * KU + q]) * next[q];
                bound[(s * KM + t) * 64 + lane] = acc;
            }
        }
    }
    __syncthreads();
    if (seg < nseg - 1) {
        T head[KU];
#pragma unroll
        for (int t = 0; t < KU; ++t) head[t] = bound[((seg + 1) * KM + t) * 64 + lane];
#pragma unroll
        for (int k = 0; k < MSEG; ++k) {
            if (k < len) {
                T acc = vals[k];
#pragma unroll
                for (int q = 0; q < KU; ++q)
                    acc += static_cast<T>(sf.hb[static_cast<long>(j0 + k) * KU + q]) * head[q];
                vals[k] = acc;
            }
        }
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

template <typename T, typename F, int KL, int KU, int MSEG>
hipError_t seg_dispatch_m(T* u, const line_geometry& g, const seg_factors<F>& sf, hipStream_t s) {
    constexpr int KM = KL > KU ? KL : KU;
    const unsigned blocks = static_cast<unsigned>((g.n_lines + 63) / 64);
    const unsigned threads = static_cast<unsigned>(sf.nseg * 64);
    size_t shm = static_cast<size_t>(sf.nseg) * KM * 64 * sizeof(T);
    if (g.S == 1) {
        shm += static_cast<size_t>(sf.nseg) * 64 * (MSEG + 1) * sizeof(T);
        // staging would not fit the 160 KiB of LDS: one lane per whole line instead
        if (shm > 150 * 1024) return sweep_dispatch<T, F, KL, KU>(u, g, sf.f, s);
        auto kern = k_sweep_seg<T, F, KL, KU, MSEG, true>;
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(shm));
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(threads), shm, s, u, g.n_lines, g.A, g.B, g.S, sf);
    } else {
        auto kern = k_sweep_seg<T, F, KL, KU, MSEG, false>;
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(shm));
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(threads), shm, s, u, g.n_lines, g.A, g.B, g.S, sf);
    }
    return hipGetLastError();
}

template <typename T, typename F, int KL, int KU>
hipError_t seg_dispatch(T* u, const line_geometry& g, const seg_factors<F>& sf, hipStream_t s) {
    if (g.n_lines <= 0 || sf.f.n <= 0) return hipSuccess;
    if (sf.mseg == 16) return seg_dispatch_m<T, F, KL, KU, 16>(u, g, sf, s);
    if (sf.mseg == 32) return seg_dispatch_m<T, F, KL, KU, 32>(u, g, sf, s);
    return hipErrorInvalidValue;
}

// ------------------------------------------------------------------------------------
// element-loop RHS, 3D
// ------------------------------------------------------------------------------------
// CREG: hold the element's coefficient block in registers (fewer LDS reads) or re-read it
// from LDS for every x Gauss point (fewer registers: two waves per SIMD).
//
// One lane per element; the x/y tables of the tile's elements are staged in LDS, the
// forward ones as tabulated and the backward ones with the quadrature weight (and, for the
// derivative tables, the constant -dt*grad_d) folded in, so that the only per-point work
// is the integrand itself. The element contributions are summed per wave into private DOF
// planes with LDS atomic adds: a wave's adds are executed in instruction order, so the sums
// are deterministic, and waves never share a plane. One barrier later the planes of the
// (P+1) waves that touch a DOF plane are added and the tile's block is added into the rhs.
template <typename T, int P, int TX, int TY, int TZ, bool KF, bool SRC, bool CREG>
__global__ __launch_bounds__(TX* TY* TZ, CREG ? 1 : 2) void k_rhs3(
    const T* __restrict__ up, T* __restrict__ rhs, dev_tables<T> X, dev_tables<T> Y,
    dev_tables<T> Z, rhs_params<T> prm, int colour, int ntx2, int nty2) {
    constexpr int M = P + 1, NT = TX * TY * TZ;
    constexpr int BX = TX + P, BY = TY + P, BZ = TZ + P, NB = BX * BY * BZ, PL = BX * BY;
    static_assert(TX * TY == 64, "a wave must cover exactly one z-layer of the tile");
    __shared__ T blk[NB];
    __shared__ T kb[KF ? NB : 1];
    __shared__ T sb[SRC ? NB : 1];
    __shared__ T fvx[TX * M * M], fdx[TX * M * M], bvx[TX * M * M], bdx[TX * M * M];
    __shared__ T fvy[TY * M * M], fdy[TY * M * M], bvy[TY * M * M], bdy[TY * M * M];
    // wave lz owns planes lz*M .. lz*M+P (DOF z = lz + c); FP64 in both precisions (the FP64
    // LDS add is the fast one on this part)
    __shared__ double acc[TZ * M * PL];

    const int tid = threadIdx.x;
    int t = blockIdx.x;
    const int tx = 2 * (t % ntx2) + (colour & 1);
    t /= ntx2;
    const int ty = 2 * (t % nty2) + ((colour >> 1) & 1);
    t /= nty2;
    const int tz = 2 * t + ((colour >> 2) & 1);
    const int ex0 = tx * TX, ey0 = ty * TY, ez0 = tz * TZ;
    const int nx = X.dofs, ny = Y.dofs, nz = Z.dofs;

    // the tile's DOF block (first(e) = e for these knot vectors, so dof origin = element origin)
    for (int i = tid; i < NB; i += NT) {
        const int bx = i % BX, r = i / BX, by = r % BY, bz = r / BY;
        const int gx = ex0 + bx, gy = ey0 + by, gz = ez0 + bz;
        const bool in = gx < nx && gy < ny && gz < nz;
        const long g = gx + static_cast<long>(nx) * (gy + static_cast<long>(ny) * gz);
        blk[i] = in ? up[g] : T(0);
        if constexpr (KF) kb[i] = in ? prm.kfield[g] : T(0);
        if constexpr (SRC) sb[i] = in ? prm.source[g] : T(0);
    }
    const T cx = -prm.dt * prm.gx, cy = -prm.dt * prm.gy;
    for (int i = tid; i < TX * M * M; i += NT) {
        const int e = ex0 + i / (M * M), q = (i / M) % M;
        const bool in = e < X.elements;
        const T v = in ? X.val[static_cast<long>(ex0) * M * M + i] : T(0);
        const T d = in ? X.der[static_cast<long>(ex0) * M * M + i] : T(0);
        const T w = in ? X.wj[static_cast<long>(e) * M + q] : T(0);
        fvx[i] = v;
        fdx[i] = d;
        bvx[i] = v * w;
        bdx[i] = cx * (d * w);
    }
    for (int i = tid; i < TY * M * M; i += NT) {
        const int e = ey0 + i / (M * M), q = (i / M) % M;
        const bool in = e < Y.elements;
        const T v = in ? Y.val[static_cast<long>(ey0) * M * M + i] : T(0);
        const T d = in ? Y.der[static_cast<long>(ey0) * M * M + i] : T(0);
        const T w = in ? Y.wj[static_cast<long>(e) * M + q] : T(0);
        fvy[i] = v;
        fdy[i] = d;
        bvy[i] = v * w;
        bdy[i] = cy * (d * w);
    }
    for (int i = tid; i < TZ * M * PL; i += NT) acc[i] = 0.0;
    __syncthreads();

    const int lx = tid % TX, ly = (tid / TX) % TY;
    const int lz = __builtin_amdgcn_readfirstlane(tid / (TX * TY));
    const int ez = ez0 + lz;

    T R[M][M][M];  // the element's contribution, (a, b, c) = local (x, y, z) function
#pragma unroll
    for (int a = 0; a < M; ++a)
#pragma unroll
        for (int b = 0; b < M; ++b)
#pragma unroll
            for (int c = 0; c < M; ++c) R[a][b][c] = T(0);

    if (ez < Z.elements) {  // wave-uniform
        // z tables of the wave's element layer: uniform addresses in uniform control flow,
        // read with scalar loads
        T Bz[M][M], Dz[M][M], wz[M];
        const T* zv = Z.val + static_cast<long>(ez) * M * M;
        const T* zd = Z.der + static_cast<long>(ez) * M * M;
#pragma unroll
        for (int q = 0; q < M; ++q) {
            wz[q] = Z.wj[static_cast<long>(ez) * M + q];
#pragma unroll
            for (int c = 0; c < M; ++c) {
                Bz[q][c] = zv[q * M + c];
                Dz[q][c] = zd[q * M + c];
            }
        }
        const T mass = prm.mass, cz = -prm.dt * prm.gz;
        if (ex0 + lx < X.elements && ey0 + ly < Y.elements) {
            T cf[CREG ? M : 1][CREG ? M : 1][CREG ? M : 1];
            T kc[KF && CREG ? M : 1][KF && CREG ? M : 1][KF && CREG ? M : 1];
            T sc[SRC && CREG ? M : 1][SRC && CREG ? M : 1][SRC && CREG ? M : 1];
            if constexpr (CREG) {
#pragma unroll
                for (int a = 0; a < M; ++a)
#pragma unroll
                    for (int b = 0; b < M; ++b)
#pragma unroll
                        for (int c = 0; c < M; ++c) {
                            const int li = (lx + a) + BX * ((ly + b) + BY * (lz + c));
                            cf[a][b][c] = blk[li];
                            if constexpr (KF) kc[a][b][c] = kb[li];
                            if constexpr (SRC) sc[a][b][c] = sb[li];
                        }
            }
#pragma unroll 1
            for (int qx = 0; qx < M; ++qx) {
                T bx[M], dx[M];
#pragma unroll
                for (int a = 0; a < M; ++a) {
                    bx[a] = fvx[(lx * M + qx) * M + a];
                    dx[a] = fdx[(lx * M + qx) * M + a];
                }
                // stage x: contract the element's first index
                T A0[M][M], A1[M][M], KA[KF ? M : 1][KF ? M : 1], SA[SRC ? M : 1][SRC ? M : 1];
#pragma unroll
                for (int b = 0; b < M; ++b)
#pragma unroll
                    for (int c = 0; c < M; ++c) {
                        T s0 = 0, s1 = 0, sk = 0, ss = 0;
#pragma unroll
                        for (int a = 0; a < M; ++a) {
                            const int li = (lx + a) + BX * ((ly + b) + BY * (lz + c));
                            const T cv = CREG ? cf[CREG ? a : 0][CREG ? b : 0][CREG ? c : 0] : blk[li];
                            s0 += bx[a] * cv;
                            s1 += dx[a] * cv;
                            if constexpr (KF)
                                sk += bx[a] * (CREG ? kc[KF && CREG ? a : 0][KF && CREG ? b : 0][KF && CREG ? c : 0] : kb[li]);
                            if constexpr (SRC)
                                ss += bx[a] * (CREG ? sc[SRC && CREG ? a : 0][SRC && CREG ? b : 0][SRC && CREG ? c : 0] : sb[li]);
                        }
                        A0[b][c] = s0;
                        A1[b][c] = s1;
                        if constexpr (KF) KA[b][c] = sk;
                        if constexpr (SRC) SA[b][c] = ss;
                    }
                T K0[M][M], Kx[M][M];
#pragma unroll
                for (int b = 0; b < M; ++b)
#pragma unroll
                    for (int c = 0; c < M; ++c) K0[b][c] = Kx[b][c] = T(0);
#pragma unroll 1
                for (int qy = 0; qy < M; ++qy) {
                    T by[M], dy[M];
#pragma unroll
                    for (int b = 0; b < M; ++b) {
                        by[b] = fvy[(ly * M + qy) * M + b];
                        dy[b] = fdy[(ly * M + qy) * M + b];
                    }
                    // stage y for this Gauss point
                    T a00[M], a01[M], a10[M], ka[KF ? M : 1], sa[SRC ? M : 1];
#pragma unroll
                    for (int c = 0; c < M; ++c) {
                        T s00 = 0, s01 = 0, s10 = 0, sk = 0, ss = 0;
#pragma unroll
                        for (int b = 0; b < M; ++b) {
                            s00 += by[b] * A0[b][c];
                            s01 += dy[b] * A0[b][c];
                            s10 += by[b] * A1[b][c];
                            if constexpr (KF) sk += by[b] * KA[b][c];
                            if constexpr (SRC) ss += by[b] * SA[b][c];
                        }
                        a00[c] = s00;
                        a01[c] = s01;
                        a10[c] = s10;
                        if constexpr (KF) ka[c] = sk;
                        if constexpr (SRC) sa[c] = ss;
                    }
                    T H0[M], Hx[M], Hy[M];
#pragma unroll
                    for (int c = 0; c < M; ++c) H0[c] = Hx[c] = Hy[c] = T(0);
#pragma unroll
                    for (int qz = 0; qz < M; ++qz) {
                        // stage z, the integrand, and stage z transposed
                        T u = 0, uz = 0, uy = 0, ux = 0, kv = 0, fv = 0;
#pragma unroll
                        for (int c = 0; c < M; ++c) {
                            u += Bz[qz][c] * a00[c];
                            uz += Dz[qz][c] * a00[c];
                            uy += Bz[qz][c] * a01[c];
                            ux += Bz[qz][c] * a10[c];
                            if constexpr (KF) kv += Bz[qz][c] * ka[c];
                            if constexpr (SRC) fv += Bz[qz][c] * sa[c];
                        }
                        T g0 = mass * u;
                        if (prm.reaction != T(0)) g0 += prm.dt * prm.reaction * u * (T(1) - u);
                        if constexpr (SRC) g0 += prm.dt * fv;
                        T gw = wz[qz];
                        if constexpr (KF) gw *= kv;
                        g0 *= wz[qz];
                        const T gx = gw * ux, gy = gw * uy, gz = (cz * gw) * uz;
#pragma unroll
                        for (int c = 0; c < M; ++c) {
                            H0[c] += Bz[qz][c] * g0 + Dz[qz][c] * gz;
                            Hx[c] += Bz[qz][c] * gx;
                            Hy[c] += Bz[qz][c] * gy;
                        }
                    }
                    // stage y transposed (weights and -dt*grad_y folded into the tables)
#pragma unroll
                    for (int b = 0; b < M; ++b) {
                        const T wb = bvy[(ly * M + qy) * M + b], wd = bdy[(ly * M + qy) * M + b];
#pragma unroll
                        for (int c = 0; c < M; ++c) {
                            K0[b][c] += wb * H0[c] + wd * Hy[c];
                            Kx[b][c] += wb * Hx[c];
                        }
                    }
                }
                // stage x transposed
#pragma unroll
                for (int a = 0; a < M; ++a) {
                    const T wb = bvx[(lx * M + qx) * M + a], wd = bdx[(lx * M + qx) * M + a];
#pragma unroll
                    for (int b = 0; b < M; ++b)
#pragma unroll
                        for (int c = 0; c < M; ++c) R[a][b][c] += wb * K0[b][c] + wd * Kx[b][c];
                }
            }
        }
    }
    // sum into the wave's planes: in one instruction the 64 lanes hit 64 distinct entries
    double* mine = acc + lz * M * PL;
#pragma unroll
    for (int c = 0; c < M; ++c)
#pragma unroll
        for (int b = 0; b < M; ++b)
#pragma unroll
            for (int a = 0; a < M; ++a)
                __hip_atomic_fetch_add(mine + c * PL + (lx + a) + BX * (ly + b),
                                       static_cast<double>(R[a][b][c]), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WAVEFRONT);
    __syncthreads();

    // DOF plane k of the block collects plane k - lz of the waves lz = k-P .. k
    for (int i = tid; i < NB; i += NT) {
        const int bxy = i % PL, bz = i / PL;
        const int gx = ex0 + bxy % BX, gy = ey0 + bxy / BX, gz = ez0 + bz;
        if (gx >= nx || gy >= ny || gz >= nz) continue;
        double sum = 0;
#pragma unroll
        for (int d = 0; d < M; ++d) {
            const int w = bz - d;
            if (w >= 0 && w < TZ) sum += acc[(w * M + d) * PL + bxy];
        }
        rhs[gx + static_cast<long>(nx) * (gy + static_cast<long>(ny) * gz)] += static_cast<T>(sum);
    }
}

// ------------------------------------------------------------------------------------
// element-loop RHS, 2D
// ------------------------------------------------------------------------------------
template <typename T, int P, int TX, int TY, bool KF, bool SRC>
__global__ __launch_bounds__(TX* TY) void k_rhs2(const T* __restrict__ up, T* __restrict__ rhs,
                                                 dev_tables<T> X, dev_tables<T> Y,
                                                 rhs_params<T> prm, int colour, int ntx2) {
    constexpr int M = P + 1, NT = TX * TY;
    constexpr int BX = TX + P, BY = TY + P, NB = BX * BY;
    constexpr int WY = 64 / TX;  // element rows per wave
    constexpr int NW = NT / 64;  // waves
    constexpr int WR = WY + P;   // DOF rows touched by one wave
    static_assert(TX <= 64 && 64 % TX == 0 && TY % WY == 0, "waves must cover whole element rows");
    __shared__ T blk[NB];
    __shared__ T kb[KF ? NB : 1];
    __shared__ T sb[SRC ? NB : 1];
    __shared__ T fvx[TX * M * M], fdx[TX * M * M], bvx[TX * M * M], bdx[TX * M * M];
    __shared__ T fvy[TY * M * M], fdy[TY * M * M], bvy[TY * M * M], bdy[TY * M * M];
    __shared__ double acc[NW * WR * BX];  // wave w owns DOF rows w*WY .. w*WY+WR-1

    const int tid = threadIdx.x;
    const int t = blockIdx.x;
    const int tx = 2 * (t % ntx2) + (colour & 1);
    const int ty = 2 * (t / ntx2) + ((colour >> 1) & 1);
    const int ex0 = tx * TX, ey0 = ty * TY;
    const int nx = X.dofs, ny = Y.dofs;

    for (int i = tid; i < NB; i += NT) {
        const int bx = i % BX, by = i / BX;
        const int gx = ex0 + bx, gy = ey0 + by;
        const bool in = gx < nx && gy < ny;
        const long g = gx + static_cast<long>(nx) * gy;
        blk[i] = in ? up[g] : T(0);
        if constexpr (KF) kb[i] = in ? prm.kfield[g] : T(0);
        if constexpr (SRC) sb[i] = in ? prm.source[g] : T(0);
    }
    const T cx = -prm.dt * prm.gx, cy = -prm.dt * prm.gy;
    for (int i = tid; i < TX * M * M; i += NT) {
        const int e = ex0 + i / (M * M), q = (i / M) % M;
        const bool in = e < X.elements;
        const T v = in ? X.val[static_cast<long>(ex0) * M * M + i] : T(0);
        const T d = in ? X.der[static_cast<long>(ex0) * M * M + i] : T(0);
        const T w = in ? X.wj[static_cast<long>(e) * M + q] : T(0);
        fvx[i] = v;
        fdx[i] = d;
        bvx[i] = v * w;
        bdx[i] = cx * (d * w);
    }
    for (int i = tid; i < TY * M * M; i += NT) {
        const int e = ey0 + i / (M * M), q = (i / M) % M;
        const bool in = e < Y.elements;
        const T v = in ? Y.val[static_cast<long>(ey0) * M * M + i] : T(0);
        const T d = in ? Y.der[static_cast<long>(ey0) * M * M + i] : T(0);
        const T w = in ? Y.wj[static_cast<long>(e) * M + q] : T(0);
        fvy[i] = v;
        fdy[i] = d;
        bvy[i] = v * w;
        bdy[i] = d * w;
    }
    for (int i = tid; i < NW * WR * BX; i += NT) acc[i] = 0.0;
    __syncthreads();

    const int lx = tid % TX, ly = tid / TX;
    const int wave = __builtin_amdgcn_readfirstlane(tid / 64);
    const int wy = ly - wave * WY;  // element row inside the wave
    T R[M][M];
#pragma unroll
    for (int a = 0; a < M; ++a)
#pragma unroll
        for (int b = 0; b < M; ++b) R[a][b] = T(0);
    if (ex0 + lx < X.elements && ey0 + ly < Y.elements) {
        T By[M][M], Dy[M][M];
#pragma unroll
        for (int q = 0; q < M; ++q)
#pragma unroll
            for (int b = 0; b < M; ++b) {
                By[q][b] = fvy[(ly * M + q) * M + b];
                Dy[q][b] = fdy[(ly * M + q) * M + b];
            }
        const T mass = prm.mass;
#pragma unroll
        for (int qx = 0; qx < M; ++qx) {
            T bx[M], dx[M];
#pragma unroll
            for (int a = 0; a < M; ++a) {
                bx[a] = fvx[(lx * M + qx) * M + a];
                dx[a] = fdx[(lx * M + qx) * M + a];
            }
            T A0[M], A1[M], KA[KF ? M : 1], SA[SRC ? M : 1];
#pragma unroll
            for (int b = 0; b < M; ++b) {
                T s0 = 0, s1 = 0, sk = 0, ss = 0;
#pragma unroll
                for (int a = 0; a < M; ++a) {
                    const int li = (lx + a) + BX * (ly + b);
                    const T cv = blk[li];
                    s0 += bx[a] * cv;
                    s1 += dx[a] * cv;
                    if constexpr (KF) sk += bx[a] * kb[li];
                    if constexpr (SRC) ss += bx[a] * sb[li];
                }
                A0[b] = s0;
                A1[b] = s1;
                if constexpr (KF) KA[b] = sk;
                if constexpr (SRC) SA[b] = ss;
            }
            T K0[M], Kx[M];
#pragma unroll
            for (int b = 0; b < M; ++b) K0[b] = Kx[b] = T(0);
#pragma unroll
            for (int qy = 0; qy < M; ++qy) {
                T u = 0, uy = 0, ux = 0, kv = 0, fv = 0;
#pragma unroll
                for (int b = 0; b < M; ++b) {
                    u += By[qy][b] * A0[b];
                    uy += Dy[qy][b] * A0[b];
                    ux += By[qy][b] * A1[b];
                    if constexpr (KF) kv += By[qy][b] * KA[b];
                    if constexpr (SRC) fv += By[qy][b] * SA[b];
                }
                T g0 = mass * u;
                if (prm.reaction != T(0)) g0 += prm.dt * prm.reaction * u * (T(1) - u);
                if constexpr (SRC) g0 += prm.dt * fv;
                T gx = ux, gy = cy * uy;
                if constexpr (KF) {
                    gx *= kv;
                    gy *= kv;
                }
                // stage y transposed (weights folded into the backward y tables)
#pragma unroll
                for (int b = 0; b < M; ++b) {
                    const T wb = bvy[(ly * M + qy) * M + b], wd = bdy[(ly * M + qy) * M + b];
                    K0[b] += wb * g0 + wd * gy;
                    Kx[b] += wb * gx;
                }
            }
#pragma unroll
            for (int a = 0; a < M; ++a) {
                const T wb = bvx[(lx * M + qx) * M + a], wd = bdx[(lx * M + qx) * M + a];
#pragma unroll
                for (int b = 0; b < M; ++b) R[a][b] += wb * K0[b] + wd * Kx[b];
            }
        }
    }
    double* mine = acc + wave * WR * BX;
#pragma unroll
    for (int b = 0; b < M; ++b)
#pragma unroll
        for (int a = 0; a < M; ++a)
            __hip_atomic_fetch_add(mine + (wy + b) * BX + lx + a, static_cast<double>(R[a][b]),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    __syncthreads();

    for (int i = tid; i < NB; i += NT) {
        const int bx = i % BX, by = i / BX;
        const int gx = ex0 + bx, gy = ey0 + by;
        if (gx >= nx || gy >= ny) continue;
        double sum = 0;
        for (int w = 0; w < NW; ++w) {
            const int r = by - w * WY;
            if (r >= 0 && r < WR) sum += acc[(w * WR + r) * BX + bx];
        }
        rhs[gx + static_cast<long>(nx) * gy] += static_cast<T>(sum);
    }
}

// tile shapes (elements) per dimension and degree: one lane per element
template <int D, int P>
struct tile;
template <int P>
struct tile<2, P> {
    static constexpr int X = 16, Y = 16, Z = 1;
};
template <>
struct tile<3, 1> {
    static constexpr int X = 8, Y = 8, Z = 4;
};
template <>
struct tile<3, 2> {
    static constexpr int X = 8, Y = 8, Z = 4;
};
template <>
struct tile<3, 3> {
    static constexpr int X = 8, Y = 8, Z = 3;
};

template <typename T, int P, bool KF, bool SRC, bool CREG>
hipError_t rhs3_colours(const T* up, T* rhs, const dev_tables<T>* d, const rhs_params<T>& prm,
                        hipStream_t s) {
    using TL = tile<3, P>;
    const int ntx = (d[0].elements + TL::X - 1) / TL::X;
    const int nty = (d[1].elements + TL::Y - 1) / TL::Y;
    const int ntz = (d[2].elements + TL::Z - 1) / TL::Z;
    for (int c = 0; c < 8; ++c) {
        const int cx = c & 1, cy = (c >> 1) & 1, cz = (c >> 2) & 1;
        const int nx2 = (ntx - cx + 1) / 2, ny2 = (nty - cy + 1) / 2, nz2 = (ntz - cz + 1) / 2;
        const long blocks = static_cast<long>(nx2) * ny2 * nz2;
        if (blocks == 0) continue;
        hipLaunchKernelGGL((k_rhs3<T, P, TL::X, TL::Y, TL::Z, KF, SRC, CREG>), dim3(blocks),
                           dim3(TL::X * TL::Y * TL::Z), 0, s, up, rhs, d[0], d[1], d[2], prm, c,
                           nx2, ny2);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

template <typename T, int P, bool KF, bool SRC>
hipError_t rhs2_colours(const T* up, T* rhs, const dev_tables<T>* d, const rhs_params<T>& prm,
                        hipStream_t s) {
    using TL = tile<2, P>;
    const int ntx = (d[0].elements + TL::X - 1) / TL::X;
    const int nty = (d[1].elements + TL::Y - 1) / TL::Y;
    for (int c = 0; c < 4; ++c) {
        const int cx = c & 1, cy = (c >> 1) & 1;
        const int nx2 = (ntx - cx + 1) / 2, ny2 = (nty - cy + 1) / 2;
        const long blocks = static_cast<long>(nx2) * ny2;
        if (blocks == 0) continue;
        hipLaunchKernelGGL((k_rhs2<T, P, TL::X, TL::Y, KF, SRC>), dim3(blocks),
                           dim3(TL::X * TL::Y), 0, s, up, rhs, d[0], d[1], prm, c, nx2);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

template <typename T, int P, bool CREG>
hipError_t rhs3_fields(const T* up, T* rhs, const dev_tables<T>* d, const rhs_params<T>& prm,
                       hipStream_t s) {
    const bool kf = prm.kfield != nullptr, src = prm.source != nullptr;
    if (kf && src) return rhs3_colours<T, P, true, true, CREG>(up, rhs, d, prm, s);
    if (kf) return rhs3_colours<T, P, true, false, CREG>(up, rhs, d, prm, s);
    if (src) return rhs3_colours<T, P, false, true, CREG>(up, rhs, d, prm, s);
    return rhs3_colours<T, P, false, false, CREG>(up, rhs, d, prm, s);
}

template <typename T, int P>
hipError_t rhs_fields(int dim, const T* up, T* rhs, const dev_tables<T>* d,
                      const rhs_params<T>& prm, int variant, hipStream_t s) {
    const bool kf = prm.kfield != nullptr, src = prm.source != nullptr;
    if (dim == 3) {
        if (variant == 1) return rhs3_fields<T, P, false>(up, rhs, d, prm, s);
        return rhs3_fields<T, P, true>(up, rhs, d, prm, s);
    }
    if (kf && src) return rhs2_colours<T, P, true, true>(up, rhs, d, prm, s);
    if (kf) return rhs2_colours<T, P, true, false>(up, rhs, d, prm, s);
    if (src) return rhs2_colours<T, P, false, true>(up, rhs, d, prm, s);
    return rhs2_colours<T, P, false, false>(up, rhs, d, prm, s);
}

// ------------------------------------------------------------------------------------
// small utilities
// ------------------------------------------------------------------------------------
template <typename T>
__global__ void k_fill(T* __restrict__ p, long n, T v) {
    for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<long>(gridDim.x) * blockDim.x)
        p[i] = v;
}

template <typename T>
__global__ void k_zero_boundary(T* __restrict__ u, int dim, int nx, int ny, int nz) {
    const long total = static_cast<long>(nx) * ny * nz;
    for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
         i += static_cast<long>(gridDim.x) * blockDim.x) {
        const int x = static_cast<int>(i % nx);
        const long r = i / nx;
        const int y = static_cast<int>(r % ny), z = static_cast<int>(r / ny);
        bool b = x == 0 || x == nx - 1 || y == 0 || y == ny - 1;
        if (dim == 3) b = b || z == 0 || z == nz - 1;
        if (b) u[i] = T(0);
    }
}

__global__ void k_d2f(const double* __restrict__ in, float* __restrict__ out, long n) {
    for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<long>(gridDim.x) * blockDim.x)
        out[i] = static_cast<float>(in[i]);
}

__global__ void k_f2d(const float* __restrict__ in, double* __restrict__ out, long n) {
    for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<long>(gridDim.x) * blockDim.x)
        out[i] = static_cast<double>(in[i]);
}

__device__ __forceinline__ bool on_boundary(long i, int dim, int nx, int ny, int nz) {
    const int x = static_cast<int>(i % nx);
    const long r = i / nx;
    const int y = static_cast<int>(r % ny), z = static_cast<int>(r / ny);
    bool b = x == 0 || x == nx - 1 || y == 0 || y == ny - 1;
    if (dim == 3) b = b || z == 0 || z == nz - 1;
    return b;
}

__global__ void k_apply_increment(const double* __restrict__ xp, const float* __restrict__ d,
                                  double* __restrict__ xn, float* __restrict__ sh, long n, int dim,
                                  int nx, int ny, int nz, int dirichlet) {
    for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<long>(gridDim.x) * blockDim.x) {
        double v = xp[i] + static_cast<double>(d[i]);
        if (dirichlet && on_boundary(i, dim, nx, ny, nz)) v = 0.0;
        xn[i] = v;
        sh[i] = static_cast<float>(v);
    }
}

__global__ void k_boundary_residual(float* __restrict__ d, const float* __restrict__ sh, long n,
                                    int dim, int nx, int ny, int nz, int p,
                                    const float* __restrict__ rx, const float* __restrict__ ry,
                                    const float* __restrict__ rz) {
    const int w = 2 * p + 1;
    for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<long>(gridDim.x) * blockDim.x) {
        if (!on_boundary(i, dim, nx, ny, nz)) continue;
        const int x = static_cast<int>(i % nx);
        const long r = i / nx;
        const int y = static_cast<int>(r % ny), z = static_cast<int>(r / ny);
        float acc = 0.0f;
        const int z0 = dim == 3 ? -p : 0, z1 = dim == 3 ? p : 0;
        for (int c = z0; c <= z1; ++c) {
            const int zz = z + c;
            if (zz < 0 || zz >= nz) continue;
            const float wz = dim == 3 ? rz[static_cast<long>(z) * w + c + p] : 1.0f;
            if (wz == 0.0f) continue;
            for (int b = -p; b <= p; ++b) {
                const int yy = y + b;
                if (yy < 0 || yy >= ny) continue;
                const float wy = ry[static_cast<long>(y) * w + b + p];
                if (wy == 0.0f) continue;
                for (int a = -p; a <= p; ++a) {
                    const int xx = x + a;
                    if (xx < 0 || xx >= nx) continue;
                    acc += rx[static_cast<long>(x) * w + a + p] * wy * wz *
                           sh[xx + static_cast<long>(nx) * (yy + static_cast<long>(ny) * zz)];
                }
            }
        }
        d[i] = -acc;
    }
}

inline unsigned grid_for(long n) {
    long b = (n + 255) / 256;
    return static_cast<unsigned>(b < 4096 ? (b > 0 ? b : 1) : 4096);
}

}  // namespace

// ------------------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------------------
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
#define ADSG_SEG(KL_, KU_) \
    if (kl == KL_ && ku == KU_) return seg_dispatch<T, F, KL_, KU_>(u, g, f, s);
    ADSG_SEG(1, 1)
    ADSG_SEG(2, 2)
    ADSG_SEG(3, 3)
#undef ADSG_SEG
    return hipErrorInvalidValue;
}

template <typename T>
hipError_t launch_rhs(int dim, int p, const T* u_prev, T* rhs, const dev_tables<T>* dirs,
                      const rhs_params<T>& prm, int variant, hipStream_t s) {
    switch (p) {
        case 1:
            return rhs_fields<T, 1>(dim, u_prev, rhs, dirs, prm, variant, s);
        case 2:
            return rhs_fields<T, 2>(dim, u_prev, rhs, dirs, prm, variant, s);
        case 3:
            return rhs_fields<T, 3>(dim, u_prev, rhs, dirs, prm, variant, s);
        default:
            return hipErrorInvalidValue;
    }
}

template <typename T>
hipError_t launch_fill(T* p, long n, T v, hipStream_t s) {
    hipLaunchKernelGGL((k_fill<T>), dim3(grid_for(n)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_zero_boundary(int dim, T* u, int nx, int ny, int nz, hipStream_t s) {
    const long n = static_cast<long>(nx) * ny * nz;
    hipLaunchKernelGGL((k_zero_boundary<T>), dim3(grid_for(n)), dim3(256), 0, s, u, dim, nx, ny, nz);
    return hipGetLastError();
}

hipError_t launch_convert_d2f(const double* in, float* out, long n, hipStream_t s) {
    hipLaunchKernelGGL(k_d2f, dim3(grid_for(n)), dim3(256), 0, s, in, out, n);
    return hipGetLastError();
}

hipError_t launch_convert_f2d(const float* in, double* out, long n, hipStream_t s) {
    hipLaunchKernelGGL(k_f2d, dim3(grid_for(n)), dim3(256), 0, s, in, out, n);
    return hipGetLastError();
}

hipError_t launch_apply_increment(const double* x_prev, const float* delta, double* x_new,
                                  float* shadow, int dim, int nx, int ny, int nz, bool dirichlet,
                                  hipStream_t s) {
    const long n = static_cast<long>(nx) * ny * nz;
    hipLaunchKernelGGL(k_apply_increment, dim3(grid_for(n)), dim3(256), 0, s, x_prev, delta, x_new,
                       shadow, n, dim, nx, ny, nz, dirichlet ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_boundary_residual(float* delta, const float* shadow, int dim, int nx, int ny,
                                    int nz, int p, const float* rows_x, const float* rows_y,
                                    const float* rows_z, hipStream_t s) {
    const long n = static_cast<long>(nx) * ny * nz;
    hipLaunchKernelGGL(k_boundary_residual, dim3(grid_for(n)), dim3(256), 0, s, delta, shadow, n,
                       dim, nx, ny, nz, p, rows_x, rows_y, rows_z);
    return hipGetLastError();
}

template hipError_t launch_line_sweep<double, double>(double*, const line_geometry&,
                                                      const dev_factors<double>&, hipStream_t);
template hipError_t launch_line_sweep<float, float>(float*, const line_geometry&,
                                                    const dev_factors<float>&, hipStream_t);
template hipError_t launch_line_sweep_seg<double, double>(double*, const line_geometry&,
                                                          const seg_factors<double>&, hipStream_t);
template hipError_t launch_line_sweep_seg<float, float>(float*, const line_geometry&,
                                                        const seg_factors<float>&, hipStream_t);
template hipError_t launch_rhs<double>(int, int, const double*, double*, const dev_tables<double>*,
                                       const rhs_params<double>&, int, hipStream_t);
template hipError_t launch_rhs<float>(int, int, const float*, float*, const dev_tables<float>*,
                                      const rhs_params<float>&, int, hipStream_t);
template hipError_t launch_fill<double>(double*, long, double, hipStream_t);
template hipError_t launch_fill<float>(float*, long, float, hipStream_t);
template hipError_t launch_zero_boundary<double>(int, double*, int, int, int, hipStream_t);
template hipError_t launch_zero_boundary<float>(int, float*, int, int, int, hipStream_t);

}  // namespace adsg
