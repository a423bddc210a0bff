// HIP kernel of the element-loop RHS for integrands that separate by direction (gfx950):
// mass u v - dt sum_d g_d (d_d u)(d_d v) with constant mass and g (plus a source, which enters
// as dt times its load vector). The p+1-point Gauss rule integrates these products of
// B-splines exactly element by element, so the element loop of compute_rhs
// (validation.hpp:95-119, heat_3d.hpp:49-67) sums to
//   rhs = (mass Mz(x)My(x)Mx - dt [gx Mz(x)My(x)Kx + gy Mz(x)Ky(x)Mx + gz Kz(x)My(x)Mx]) u
// with the 1D Gram and stiffness matrices of form_matrix.cpp:8-42 (the same quadrature):
// the same sums in another order. Applied as three banded contractions: a workgroup owns a
// 32 x 8 column of DOFs and streams through a range of z-planes, staging each u plane with
// its halo in LDS (the next plane's loads in flight meanwhile), contracting x, then y, and
// keeping the last 2p+1 planes' y results in registers for the z contraction. u and rhs are
// stored in T; the matrices, the staged contractions and the sums are FP64.
#include <hip/hip_runtime.h>

#include "adsg_kernels.hpp"

namespace adsg {

namespace {

template <typename T, int P, bool ACC>
__global__ __launch_bounds__(256) void k_rhs_kron3(const T* __restrict__ u, T* __restrict__ rhs,
                                                   kron_dir X, kron_dir Y, kron_dir Z, double cm,
                                                   double cx, double cy, double cz, int zchunk) {
    using C = double;  // contraction precision
    constexpr int B = 2 * P + 1, TX = 32, TY = 8, UX = TX + 2 * P, UY = TY + 2 * P;
    constexpr int NU = (UX * UY + 255) / 256;  // staged plane entries per thread
    __shared__ T U[UY * UX];
    __shared__ C XM[UY * TX], XK[UY * TX];
    const int tid = threadIdx.x, i = tid % TX, j = tid / TX;
    const int x0 = blockIdx.x * TX, y0 = blockIdx.y * TY, z0 = blockIdx.z * zchunk;
    const int nx = X.n, ny = Y.n, nz = Z.n;
    const int z1 = z0 + zchunk < nz ? z0 + zchunk : nz;
    const int gx = x0 + i, gy = y0 + j;
    // this thread's rows of the 1D matrices: A(g, g + b - P), b = 0..2P
    C mx[B], kx[B], my[B], ky[B];
#pragma unroll
    for (int b = 0; b < B; ++b) {
        mx[b] = gx < nx ? X.M[static_cast<long>(gx) * B + b] : C(0);
        kx[b] = gx < nx ? X.K[static_cast<long>(gx) * B + b] : C(0);
        my[b] = gy < ny ? Y.M[static_cast<long>(gy) * B + b] : C(0);
        ky[b] = gy < ny ? Y.K[static_cast<long>(gy) * B + b] : C(0);
    }
    T pre[NU];
    auto fetch = [&](int kz) {
#pragma unroll
        for (int s = 0; s < NU; ++s) {
            const int idx = tid + s * 256;
            const int ux = idx % UX, uy = idx / UX;
            const int px = x0 - P + ux, py = y0 - P + uy;
            const bool in = idx < UX * UY && px >= 0 && px < nx && py >= 0 && py < ny && kz >= 0 && kz < nz;
            pre[s] = in ? u[px + static_cast<long>(nx) * (py + static_cast<long>(ny) * kz)] : T(0);
        }
    };
    // y-stage results of the last 2P+1 planes (q[2P] the newest)
    C qmm[B], qkm[B], qmk[B];
#pragma unroll
    for (int c = 0; c < B; ++c) qmm[c] = qkm[c] = qmk[c] = C(0);
    fetch(z0 - P);
    for (int kz = z0 - P; kz < z1 + P; ++kz) {
#pragma unroll
        for (int s = 0; s < NU; ++s) {
            const int idx = tid + s * 256;
            if (idx < UX * UY) U[idx] = pre[s];
        }
        __syncthreads();
        if (kz + 1 < z1 + P) fetch(kz + 1);  // in flight during this plane
        // x-stage on rows jj = j and j + TY of the staged tile
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int jj = j + h * TY;
            if (jj < UY) {
                C sm = C(0), sk = C(0);
#pragma unroll
                for (int b = 0; b < B; ++b) {
                    const C v = U[jj * UX + i + b];
                    sm += mx[b] * v;
                    sk += kx[b] * v;
                }
                XM[jj * TX + i] = sm;
                XK[jj * TX + i] = sk;
            }
        }
        __syncthreads();
        // y-stage
        C ymm = C(0), ykm = C(0), ymk = C(0);
#pragma unroll
        for (int b = 0; b < B; ++b) {
            const C am = XM[(j + b) * TX + i], ak = XK[(j + b) * TX + i];
            ymm += my[b] * am;
            ykm += ky[b] * am;
            ymk += my[b] * ak;
        }
#pragma unroll
        for (int c = 0; c < B - 1; ++c) {
            qmm[c] = qmm[c + 1];
            qkm[c] = qkm[c + 1];
            qmk[c] = qmk[c + 1];
        }
        qmm[B - 1] = ymm;
        qkm[B - 1] = ykm;
        qmk[B - 1] = ymk;
        // z-stage for plane ko = kz - P (its planes ko-P .. ko+P are in the queue)
        const int ko = kz - P;
        if (ko >= z0 && ko < z1 && gx < nx && gy < ny) {
            const C* mz = Z.M + static_cast<long>(ko) * B;
            const C* kzr = Z.K + static_cast<long>(ko) * B;
            C smm = C(0), smk = C(0), skm = C(0), szz = C(0);
#pragma unroll
            for (int c = 0; c < B; ++c) {
                smm += mz[c] * qmm[c];
                smk += mz[c] * qmk[c];
                skm += mz[c] * qkm[c];
                szz += kzr[c] * qmm[c];
            }
            C val = cm * smm;
            val += cx * smk;
            val += cy * skm;
            val += cz * szz;
            const long g = gx + static_cast<long>(nx) * (gy + static_cast<long>(ny) * ko);
            rhs[g] = static_cast<T>(ACC ? static_cast<C>(rhs[g]) + val : val);
        }
    }
}

template <typename T, int P, bool ACC>
hipError_t kron3(const T* u, T* rhs, const kron_dir* d, double cm, double cx, double cy, double cz,
                 hipStream_t s) {
    constexpr int zchunk = 16;
    const dim3 grid((d[0].n + 31) / 32, (d[1].n + 7) / 8, (d[2].n + zchunk - 1) / zchunk);
    hipLaunchKernelGGL((k_rhs_kron3<T, P, ACC>), grid, dim3(256), 0, s, u, rhs, d[0], d[1], d[2], cm,
                       cx, cy, cz, zchunk);
    return hipGetLastError();
}

}  // namespace

template <typename T>
hipError_t launch_rhs_kron(int dim, int p, const T* u, T* rhs, const kron_dir* dirs, double mass,
                           double dt, const double* grad, bool accumulate, hipStream_t s) {
    if (dim != 3) return hipErrorInvalidValue;
    const double cx = -dt * grad[0], cy = -dt * grad[1], cz = -dt * grad[2];
#define ADSG_KRON(P_)                                                                        \
    if (p == P_) return accumulate ? kron3<T, P_, true>(u, rhs, dirs, mass, cx, cy, cz, s)    \
                                   : kron3<T, P_, false>(u, rhs, dirs, mass, cx, cy, cz, s);
    ADSG_KRON(1)
    ADSG_KRON(2)
    ADSG_KRON(3)
#undef ADSG_KRON
    return hipErrorInvalidValue;
}

template hipError_t launch_rhs_kron<double>(int, int, const double*, double*, const kron_dir*, double,
                                            double, const double*, bool, hipStream_t);
template hipError_t launch_rhs_kron<float>(int, int, const float*, float*, const kron_dir*, double,
                                           double, const double*, bool, hipStream_t);

}  // namespace adsg
