// Element-loop RHS, 3D, in columns of elements (variant 5): compute_rhs
// (validation.hpp:95-119, heat_3d.hpp:49-67) for forms without a k field or a source in the
// kernel, with the x-y contractions on the matrix cores (MFMA 16x16x4, FP64 or FP32).
//
// One wave integrates one column of elements (ex, ey, ez = z0 .. z1-1), walking it in chunks
// of 16 DOF planes. Per chunk (planes kz0 .. kz0+15, elements kz0 .. kz0+15-p):
//  1. x-y stage, forward: for every plane kz, the values of u, u_x and u_y at the (p+1)^2
//     Gauss points (qx, qy) of the column from the (p+1)^2 coefficients u(ex+a, ey+b, kz):
//     one dense GEMM T = A U with A[(point, field)][(a, b)] = X_f(qx, a) Y_f(qy, b) (rows
//     padded to 16-row blocks of 4 points x 4 fields, columns to K-steps of 4) and U[(a, b)]
//     [kz] gathered from the state, the 16 planes being the 16 columns of the MFMA. The
//     accumulator layout leaves lane (g, j) with the 3 fields of point 4m+g at plane j in
//     the registers of block m.
//  2. z stage, the integrand and the z stage transposed on the vector ALU: element kz0+j
//     needs planes j .. j+p, read from the lanes j+c of the same 16-lane row (DPP row shift
//     left); its contributions to planes j .. j+p go back the same way (row shift right).
//  3. x-y stage, backward: R[(a, b)][kz] = sum A'[(a, b)][(point, field)] S[(point, field)]
//     [kz], the accumulators of step 2 being the B operand directly (K-step (m, field)).
//  4. The waves of the workgroup (one tile of TX x TY columns) write R to LDS; one barrier
//     later every DOF of the tile's block sums its (p+1)^2 column contributions in a fixed
//     order and adds the sum to the rhs (return-less global add). Planes shared with the
//     next chunk are carried in LDS, so every DOF gets one global add per launch. Tiles are
//     coloured by the parity of their x, y and z-range indices (8 launches): tiles of one
//     colour share no DOF, and the result does not depend on timing.
//
// The sum-factorised x-y stage of the per-element kernels is computed here once per DOF
// plane instead of once per element and plane (p+1 times less), and on the matrix cores;
// the vector ALU keeps the per-element z stage. Summation orders differ from the other
// variants (equal to rounding).
#include <hip/hip_runtime.h>

#include "adsg_kernels.hpp"

namespace adsg {

namespace {

typedef double d4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

// 16x16x4 MFMA of T: A[i][k] from lane i + 16k, B[k][j] from lane j + 16k; accumulator
// register r of lane j + 16g holds row row(g, r) of column j (the two precisions differ).
template <typename T>
struct mf;
template <>
struct mf<double> {
    using acc = d4v;
    static __device__ __forceinline__ acc run(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static constexpr __device__ int row(int g, int r) { return g + 4 * r; }
    static constexpr __device__ int grp(int i) { return i % 4; }  // g of row i
    static constexpr __device__ int reg(int i) { return i / 4; }  // r of row i
};
template <>
struct mf<float> {
    using acc = f4v;
    static __device__ __forceinline__ acc run(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static constexpr __device__ int row(int g, int r) { return 4 * g + r; }
    static constexpr __device__ int grp(int i) { return i / 4; }
    static constexpr __device__ int reg(int i) { return i % 4; }
};

// DPP within rows of 16 lanes: CTRL 0x100 + c takes the value of lane l + c (row shift
// left), 0x110 + c the value of lane l - c (row shift right); 0 where that lane is outside
// the row
template <int CTRL>
__device__ __forceinline__ float row_move(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xf, 0xf, false));
}
template <int CTRL>
__device__ __forceinline__ double row_move(double v) {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(b), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(b >> 32), CTRL, 0xf, 0xf, false);
    return __builtin_bit_cast(double, (static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}
template <int C, typename T>
__device__ __forceinline__ T from_above(T v) {  // lane j gets lane j + C (same row)
    if constexpr (C == 0) return v;
    else return row_move<0x100 + C>(v);
}
template <int C, typename T>
__device__ __forceinline__ T from_below(T v) {  // lane j gets lane j - C (same row)
    if constexpr (C == 0) return v;
    else return row_move<0x110 + C>(v);
}

constexpr int kXcds = 8;  // MI355X: 8 accelerator complex dies, 32 CUs each
__device__ __forceinline__ int xcd_contiguous(int b, int n) {
    const int per = n / kXcds, rem = n % kXcds;
    const int x = b % kXcds, i = b / kXcds;
    return x < rem ? x * (per + 1) + i : rem * (per + 1) + (x - rem) * per + i;
}

template <int P>
struct col_shape {
    static constexpr int M = P + 1;
    static constexpr int NPT = M * M;         // Gauss points (qx, qy) of a column
    static constexpr int MB = (NPT + 3) / 4;  // 16-row blocks of 4 points x 4 fields
    static constexpr int KD = M * M;          // coefficients (a, b) of a column's plane
    static constexpr int KS = (KD + 3) / 4;   // K-steps of the forward GEMM
    static constexpr int CH = 16 - P;         // elements per chunk of 16 planes
};

template <typename T, int P, int TX, int TY, int NCH>
struct col_lds {
    using SH = col_shape<P>;
    static constexpr int M = SH::M, BX = TX + P, BY = TY + P, ZR = NCH * SH::CH;
    // padded strides (LDS banks): rows of a plane, planes of a column, a z element's table
    static constexpr int RS = 17, RR = 16 * RS + 4, ZS = M * M + 1, WS = M + 1;
    T R[TX * TY][RR];              // per column: [plane][row (a, b)] of the backward GEMM
    double carry[2][BX * BY * P];  // planes shared with the next chunk
    T xv[TX][M][M], xd[TX][M][M], xw[TX][M];
    T yv[TY][M][M], yd[TY][M][M], yw[TY][M];
    T zv[ZR][ZS], zd[ZR][ZS], zw[ZR][WS];
};

template <typename T, int P, int FM, int TX, int TY, int CPW, int NCH>
__global__ __launch_bounds__(64 * TX * TY / CPW) void k_rhs3_col(const T* __restrict__ up, T* __restrict__ rhs,
                                                           dev_tables<T> X, dev_tables<T> Y,
                                                           dev_tables<T> Z, rhs_params<T> prm,
                                                           int colour, int ntx2, int nty2) {
    using SH = col_shape<P>;
    using MF = mf<T>;
    using acc = typename MF::acc;
    using LDS = col_lds<T, P, TX, TY, NCH>;
    constexpr int M = SH::M, NPT = SH::NPT, MB = SH::MB, KD = SH::KD, KS = SH::KS, CH = SH::CH;
    constexpr int NT = 64 * TX * TY / CPW, BX = LDS::BX, BY = LDS::BY, ZR = LDS::ZR;
    static_assert((TX * TY) % CPW == 0, "whole columns per wave");
    // FM: 0 the heat form (mass term), 1 with the reaction term, 2 no mass term
    constexpr bool MS = FM != 2, RX = FM == 1;
    // row blocks per pass: FP64 at p = 3 takes two passes to stay within the register file
    constexpr int MH = sizeof(T) == sizeof(double) && P == 3 ? 2 : SH::MB;
    __shared__ LDS S;

    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int g = lane >> 4, j = lane & 15;
    int t = xcd_contiguous(blockIdx.x, gridDim.x);
    const int tx = 2 * (t % ntx2) + (colour & 1);
    t /= ntx2;
    const int ty = 2 * (t % nty2) + ((colour >> 1) & 1);
    t /= nty2;
    const int tz = 2 * t + ((colour >> 2) & 1);
    const int ex0 = tx * TX, ey0 = ty * TY, z0 = tz * ZR;
    const int z1 = min(z0 + ZR, Z.elements);
    const int nx = X.dofs, ny = Y.dofs, nz = Z.dofs;

    // the tile's 1D tables (zero outside the mesh)
    for (int i = tid; i < TX * M * M; i += NT) {
        const bool in = ex0 + i / (M * M) < X.elements;
        (&S.xv[0][0][0])[i] = in ? X.val[static_cast<long>(ex0) * M * M + i] : T(0);
        (&S.xd[0][0][0])[i] = in ? X.der[static_cast<long>(ex0) * M * M + i] : T(0);
    }
    for (int i = tid; i < TX * M; i += NT)
        (&S.xw[0][0])[i] = ex0 + i / M < X.elements ? X.wj[static_cast<long>(ex0) * M + i] : T(0);
    for (int i = tid; i < TY * M * M; i += NT) {
        const bool in = ey0 + i / (M * M) < Y.elements;
        (&S.yv[0][0][0])[i] = in ? Y.val[static_cast<long>(ey0) * M * M + i] : T(0);
        (&S.yd[0][0][0])[i] = in ? Y.der[static_cast<long>(ey0) * M * M + i] : T(0);
    }
    for (int i = tid; i < TY * M; i += NT)
        (&S.yw[0][0])[i] = ey0 + i / M < Y.elements ? Y.wj[static_cast<long>(ey0) * M + i] : T(0);
    for (int i = tid; i < ZR * M * M; i += NT) {
        const int e = i / (M * M), k = i % (M * M);
        const bool in = z0 + e < z1;
        S.zv[e][k] = in ? Z.val[static_cast<long>(z0) * M * M + i] : T(0);
        S.zd[e][k] = in ? Z.der[static_cast<long>(z0) * M * M + i] : T(0);
    }
    for (int i = tid; i < ZR * M; i += NT)
        S.zw[i / M][i % M] = z0 + i / M < z1 ? Z.wj[static_cast<long>(z0) * M + i] : T(0);
    for (int i = tid; i < BX * BY * P; i += NT) S.carry[0][i] = 0.0;
    __syncthreads();

    const T mass = prm.mass, rdt = prm.dt * prm.reaction;
    const T gxc = -prm.dt * prm.gx, gyc = -prm.dt * prm.gy, gzc = -prm.dt * prm.gz;

    // table factor of field f (0: u, 1: u_x, 2: u_y) in x and in y
    auto xf = [&](int cx, int f, int q, int a) -> T { return f == 1 ? S.xd[cx][q][a] : S.xv[cx][q][a]; };
    auto yf = [&](int cy, int f, int q, int b) -> T { return f == 2 ? S.yd[cy][q][b] : S.yv[cy][q][b]; };

    // the B operand of the forward GEMM for column col at chunk kzs: U[(a, b)][kzs + j] for
    // this lane's K indices 4s + g (zero outside the mesh and for padded K)
    auto load_u = [&](int col, int kzs, T(&o)[KS]) {
        const int ex = ex0 + col % TX, ey = ey0 + col / TX, plane = kzs + j;
        const bool in = ex < X.elements && ey < Y.elements && kzs < z1 && plane < nz;
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const int kk = 4 * s + g;
            const int a = kk % M, b = kk / M;
            o[s] = in && kk < KD ? up[(ex + a) + static_cast<long>(nx) * ((ey + b) + static_cast<long>(ny) * plane)]
                                 : T(0);
        }
    };
    T uv[KS];
    load_u(w * CPW, z0, uv);

    int par = 0;
    int kz0 = z0;
    for (; kz0 < z1; kz0 += CH, par ^= 1) {
#pragma unroll
      for (int ci = 0; ci < CPW; ++ci) {
        const int col = w * CPW + ci;
        const int cx = col % TX, cy = col / TX;
        const int ex = ex0 + cx, ey = ey0 + cy;
        // start loading the next column's (or the next chunk's) coefficients now: their
        // latency then passes during this column's work
        T un[KS];
        load_u(ci + 1 < CPW ? col + 1 : w * CPW, ci + 1 < CPW ? kz0 : kz0 + CH, un);
        if (ex < X.elements && ey < Y.elements) {  // wave-uniform
            const int le = kz0 - z0 + j;  // element kz0 + j of the z-range
            const bool ev = j < CH && kz0 + j < z1;
            const int lz = ev ? le : 0;
            // two backward accumulators (even and odd row blocks): two independent MFMA chains
            acc Rr[2] = {acc{0, 0, 0, 0}, acc{0, 0, 0, 0}};
            // the row blocks in passes of MH (fewer live registers), each through steps 1-3
#pragma unroll
            for (int h = 0; h < MB; h += MH) {
                // 1. forward x-y GEMM: T[m] = sum_s A[m][s] U[s]
                acc Tm[MH];
#pragma unroll
                for (int mm = 0; mm < MH; ++mm) Tm[mm] = acc{0, 0, 0, 0};
#pragma unroll
                for (int s = 0; s < KS; ++s) {
                    const int kk = 4 * s + g;  // K index of this lane (A and B)
                    const int a = kk % M, b = kk / M;
#pragma unroll
                    for (int mm = 0; mm < MH; ++mm) {
                        const int m = h + mm;
                        if (m < MB) {
                            const int i = j;  // A row of this lane
                            const int pt = 4 * m + MF::grp(i), f = MF::reg(i);
                            const int qx = pt % M, qy = pt / M;
                            const T av = pt < NPT && f < 3 && kk < KD ? xf(cx, f, qx, a) * yf(cy, f, qy, b) : T(0);
                            Tm[mm] = MF::run(av, uv[s], Tm[mm]);
                        }
                    }
                }
                // 2. z stage, integrand, z stage transposed for element kz0 + j
#pragma unroll
                for (int mm = 0; mm < MH; ++mm) {
                    const int m = h + mm;
                    if (m >= MB) break;
                    const int pt = 4 * m + g;
                    const int qx = pt % M, qy = pt / M;
                    const T wxy = pt < NPT ? S.xw[cx][qx] * S.yw[cy][qy] : T(0);
                    T tv[M], tdx[M], tdy[M];
                    tv[0] = Tm[mm][0];
                    tdx[0] = Tm[mm][1];
                    tdy[0] = Tm[mm][2];
                    tv[1] = from_above<1>(Tm[mm][0]);
                    tdx[1] = from_above<1>(Tm[mm][1]);
                    tdy[1] = from_above<1>(Tm[mm][2]);
                    if constexpr (P >= 2) {
                        tv[2] = from_above<2>(Tm[mm][0]);
                        tdx[2] = from_above<2>(Tm[mm][1]);
                        tdy[2] = from_above<2>(Tm[mm][2]);
                    }
                    if constexpr (P >= 3) {
                        tv[3] = from_above<3>(Tm[mm][0]);
                        tdx[3] = from_above<3>(Tm[mm][1]);
                        tdy[3] = from_above<3>(Tm[mm][2]);
                    }
                    T Wv[M], Wx[M], Wy[M];
#pragma unroll
                    for (int c = 0; c < M; ++c) Wv[c] = Wx[c] = Wy[c] = T(0);
#pragma unroll
                    for (int qz = 0; qz < M; ++qz) {
                        T u = 0, uz = 0, ux = 0, uy = 0;
#pragma unroll
                        for (int c = 0; c < M; ++c) {
                            const T bz = S.zv[lz][qz * M + c], dz = S.zd[lz][qz * M + c];
                            if constexpr (MS) u += bz * tv[c];
                            uz += dz * tv[c];
                            ux += bz * tdx[c];
                            uy += bz * tdy[c];
                        }
                        const T ww = ev ? wxy * S.zw[lz][qz] : T(0);
                        T rv = T(0);
                        if constexpr (MS) {
                            rv = mass * u;
                            if constexpr (RX) rv += rdt * u * (T(1) - u);
                            rv *= ww;
                        }
                        const T rx = (ww * gxc) * ux, ry = (ww * gyc) * uy, rz = (ww * gzc) * uz;
#pragma unroll
                        for (int c = 0; c < M; ++c) {
                            const T bz = S.zv[lz][qz * M + c], dz = S.zd[lz][qz * M + c];
                            if constexpr (MS) Wv[c] += bz * rv;
                            Wv[c] += dz * rz;
                            Wx[c] += bz * rx;
                            Wy[c] += bz * ry;
                        }
                    }
                    // plane j collects c = 0 .. P of elements j - c
                    T sv = Wv[0], sx = Wx[0], sy = Wy[0];
                    sv += from_below<1>(Wv[1]);
                    sx += from_below<1>(Wx[1]);
                    sy += from_below<1>(Wy[1]);
                    if constexpr (P >= 2) {
                        sv += from_below<2>(Wv[2]);
                        sx += from_below<2>(Wx[2]);
                        sy += from_below<2>(Wy[2]);
                    }
                    if constexpr (P >= 3) {
                        sv += from_below<3>(Wv[3]);
                        sx += from_below<3>(Wx[3]);
                        sy += from_below<3>(Wy[3]);
                    }
                    // 3. backward x-y GEMM: R += sum_f A'[m][f] S[m][f], K index of this lane
                    // = (point pt, field f), output row (a, b) = lane row j
                    const int i = j;
                    const int a = i % M, b = i / M;
                    const T sf[3] = {sv, sx, sy};
#pragma unroll
                    for (int f = 0; f < 3; ++f) {
                        const T av = pt < NPT && i < KD ? xf(cx, f, qx, a) * yf(cy, f, qy, b) : T(0);
                        Rr[m & 1] = MF::run(av, sf[f], Rr[m & 1]);
                    }
                }
            }
            // 4. the column's contribution to its planes, [plane][row]
            const acc Rs = Rr[0] + Rr[1];
#pragma unroll
            for (int r = 0; r < 4; ++r) S.R[col][j * LDS::RS + MF::row(g, r)] = Rs[r];
        }
#pragma unroll
        for (int s = 0; s < KS; ++s) uv[s] = un[s];
      }
        __syncthreads();
        // every DOF of the tile's block: the column contributions in a fixed order
        for (int o = tid; o < BX * BY * 16; o += NT) {
            const int bx = o % BX, by = (o / BX) % BY, jj = o / (BX * BY);
            double sum = 0.0;
#pragma unroll
            for (int b = 0; b < M; ++b)
#pragma unroll
                for (int a = 0; a < M; ++a) {
                    const int ccx = bx - a, ccy = by - b;
                    if (ccx >= 0 && ccx < TX && ccy >= 0 && ccy < TY && ex0 + ccx < X.elements &&
                        ey0 + ccy < Y.elements)
                        sum += static_cast<double>(S.R[ccx + TX * ccy][jj * LDS::RS + a + M * b]);
                }
            const int ci = (by * BX + bx) * P;
            if (jj < P) sum += S.carry[par][ci + jj];
            if (jj >= CH) {
                S.carry[par ^ 1][ci + jj - CH] = sum;
            } else {
                const int gx = ex0 + bx, gy = ey0 + by, gz = kz0 + jj;
                if (gx < nx && gy < ny && gz < nz)
                    unsafeAtomicAdd(rhs + (gx + static_cast<long>(nx) * (gy + static_cast<long>(ny) * gz)),
                                    static_cast<T>(sum));
            }
        }
        __syncthreads();
    }
    // the planes carried out of the last chunk
    for (int o = tid; o < BX * BY * P; o += NT) {
        const int bx = o % BX, by = (o / BX) % BY, jj = o / (BX * BY);
        const int gx = ex0 + bx, gy = ey0 + by, gz = kz0 + jj;
        if (gx < nx && gy < ny && gz < nz)
            unsafeAtomicAdd(rhs + (gx + static_cast<long>(nx) * (gy + static_cast<long>(ny) * gz)),
                            static_cast<T>(S.carry[par][(by * BX + bx) * P + jj]));
    }
}

template <typename T, int P, int FM, int CPW, int NCH>
hipError_t col_colours(const T* up, T* rhs, const dev_tables<T>* d, const rhs_params<T>& prm, hipStream_t s) {
    constexpr int TX = 4, TY = 4, ZR = NCH * col_shape<P>::CH;
    const int ntx = (d[0].elements + TX - 1) / TX;
    const int nty = (d[1].elements + TY - 1) / TY;
    const int ntz = (d[2].elements + ZR - 1) / ZR;
    for (int c = 0; c < 8; ++c) {
        const int cx = c & 1, cy = (c >> 1) & 1, cz = (c >> 2) & 1;
        const int nx2 = (ntx - cx + 1) / 2, ny2 = (nty - cy + 1) / 2, nz2 = (ntz - cz + 1) / 2;
        const long tiles = static_cast<long>(nx2) * ny2 * nz2;
        if (tiles == 0) continue;
        hipLaunchKernelGGL((k_rhs3_col<T, P, FM, TX, TY, CPW, NCH>), dim3(tiles), dim3(64 * TX * TY / CPW), 0, s,
                           up, rhs, d[0], d[1], d[2], prm, c, nx2, ny2);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// shape: bit 0 = two columns per wave (else one; FP64 at p = 3 always two, its registers
// do not fit four waves per SIMD), bit 1 = three chunks per z-range (else one)
template <typename T, int P, int FM>
hipError_t col_shapes(const T* up, T* rhs, const dev_tables<T>* d, const rhs_params<T>& prm, int shape,
                      hipStream_t s) {
    constexpr bool only2 = sizeof(T) == sizeof(double) && P == 3;
    const bool two = only2 || (shape & 1);
    if (shape & 2)
        return two ? col_colours<T, P, FM, 2, 3>(up, rhs, d, prm, s) : col_colours<T, P, FM, only2 ? 2 : 1, 3>(up, rhs, d, prm, s);
    return two ? col_colours<T, P, FM, 2, 1>(up, rhs, d, prm, s) : col_colours<T, P, FM, only2 ? 2 : 1, 1>(up, rhs, d, prm, s);
}

template <typename T, int P>
hipError_t col_forms(const T* up, T* rhs, const dev_tables<T>* d, const rhs_params<T>& prm, int shape, hipStream_t s) {
    if (prm.reaction != T(0)) return col_shapes<T, P, 1>(up, rhs, d, prm, shape, s);
    if (prm.mass == T(0)) return col_shapes<T, P, 2>(up, rhs, d, prm, shape, s);
    return col_shapes<T, P, 0>(up, rhs, d, prm, shape, s);
}

}  // namespace

template <typename T>
bool rhs3_col_supported(int p, const rhs_params<T>& prm) {
    return (p == 2 || p == 3) && !prm.kfield && !prm.source;
}

template <typename T>
hipError_t launch_rhs3_col(int p, const T* u_prev, T* rhs, const dev_tables<T>* dirs, const rhs_params<T>& prm,
                           int shape, hipStream_t s) {
    if (!rhs3_col_supported(p, prm) || shape < 0 || shape > 3) return hipErrorInvalidValue;
    return p == 2 ? col_forms<T, 2>(u_prev, rhs, dirs, prm, shape, s) : col_forms<T, 3>(u_prev, rhs, dirs, prm, shape, s);
}

template bool rhs3_col_supported<double>(int, const rhs_params<double>&);
template bool rhs3_col_supported<float>(int, const rhs_params<float>&);
template hipError_t launch_rhs3_col<double>(int, const double*, double*, const dev_tables<double>*,
                                            const rhs_params<double>&, int, hipStream_t);
template hipError_t launch_rhs3_col<float>(int, const float*, float*, const dev_tables<float>*,
                                           const rhs_params<float>&, int, hipStream_t);

}  // namespace adsg
