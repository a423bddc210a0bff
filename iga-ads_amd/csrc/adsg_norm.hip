// HIP kernels of the norm and error functionals (gfx950). (The adjoint's k-field gradient
// runs on the element tiles of adsg_rhs.hip: kgrad2_tile, kgrad3_element.)
#include <hip/hip_runtime.h>

#include "adsg_kernels.hpp"

namespace adsg {

namespace {

// Squared norm and error functionals (basic_simulation_2d.hpp:42-52, :266-362): per element
// and Gauss point w J phi(d), d = u_h - s, with u_h the spline of the coefficients u (or 0
// when u is null) and s a function given at the Gauss points as planes of Q values
// [value | d/dx | d/dy (| d/dz)] (or 0 when sq is null); phi = d^2 (kind 0), |grad d|^2 (1)
// or both (2). One lane per element (grid-stride), FP64 sums reduced in a fixed order:
// lanes -> block partial -> one total.
template <typename TT, int DIM, int P>
__global__ __launch_bounds__(256) void k_norm(const double* __restrict__ u, const double* __restrict__ sq,
                                              long Q, dev_tables<TT> tx, dev_tables<TT> ty,
                                              dev_tables<TT> tz, int kind, double* __restrict__ partial) {
    constexpr int n = P + 1;
    constexpr int NZ = DIM == 3 ? n : 1;
    __shared__ double red[256];
    const long E = static_cast<long>(tx.elements) * ty.elements * (DIM == 3 ? tz.elements : 1);
    const long nx = tx.dofs, ny = ty.dofs;
    const long NX = static_cast<long>(tx.elements) * n, NY = static_cast<long>(ty.elements) * n;
    double acc = 0;
    for (long l = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; l < E;
         l += static_cast<long>(gridDim.x) * blockDim.x) {
        const int ex = static_cast<int>(l % tx.elements);
        const long r = l / tx.elements;
        const int ey = static_cast<int>(r % ty.elements), ez = static_cast<int>(r / ty.elements);
        double U[NZ][n][n];
        for (int c = 0; c < NZ; ++c)
            for (int b = 0; b < n; ++b)
                for (int a = 0; a < n; ++a)
                    U[c][b][a] = u ? u[(ex + a) + nx * ((ey + b) + ny * static_cast<long>(ez + c))] : 0.0;
        const TT* vx = tx.val + static_cast<long>(ex) * n * n;
        const TT* dx = tx.der + static_cast<long>(ex) * n * n;
        const TT* vy = ty.val + static_cast<long>(ey) * n * n;
        const TT* dy = ty.der + static_cast<long>(ey) * n * n;
        const TT* vz = DIM == 3 ? tz.val + static_cast<long>(ez) * n * n : nullptr;
        const TT* dz = DIM == 3 ? tz.der + static_cast<long>(ez) * n * n : nullptr;
        // sum factorisation: contract x (values and derivatives), then y, then z, per Gauss
        // point of the outer directions (9 n^4 products per element instead of 4 n^6)
#pragma unroll 1
        for (int qx = 0; qx < n; ++qx) {
            double A0[NZ][n], A1[NZ][n];
#pragma unroll
            for (int c = 0; c < NZ; ++c)
#pragma unroll
                for (int b = 0; b < n; ++b) {
                    double s0 = 0, s1 = 0;
#pragma unroll
                    for (int a = 0; a < n; ++a) {
                        s0 += double(vx[qx * n + a]) * U[c][b][a];
                        s1 += double(dx[qx * n + a]) * U[c][b][a];
                    }
                    A0[c][b] = s0;
                    A1[c][b] = s1;
                }
#pragma unroll 1
            for (int qy = 0; qy < n; ++qy) {
                double a00[NZ], a01[NZ], a10[NZ];
#pragma unroll
                for (int c = 0; c < NZ; ++c) {
                    double s00 = 0, s01 = 0, s10 = 0;
#pragma unroll
                    for (int b = 0; b < n; ++b) {
                        const double by = vy[qy * n + b], bdy = dy[qy * n + b];
                        s00 += by * A0[c][b];
                        s01 += bdy * A0[c][b];
                        s10 += by * A1[c][b];
                    }
                    a00[c] = s00;
                    a01[c] = s01;
                    a10[c] = s10;
                }
#pragma unroll
                for (int qz = 0; qz < NZ; ++qz) {
                    double v = 0, gx = 0, gy = 0, gz = 0;
                    if constexpr (DIM == 3) {
#pragma unroll
                        for (int c = 0; c < NZ; ++c) {
                            const double bz = vz[qz * n + c], bdz = dz[qz * n + c];
                            v += bz * a00[c];
                            gz += bdz * a00[c];
                            gy += bz * a01[c];
                            gx += bz * a10[c];
                        }
                    } else {
                        v = a00[0];
                        gy = a01[0];
                        gx = a10[0];
                    }
                    if (sq) {
                        const long pt = (static_cast<long>(ex) * n + qx) +
                                        NX * ((static_cast<long>(ey) * n + qy) + NY * (static_cast<long>(ez) * n + qz));
                        v -= sq[pt];
                        if (kind != 0) {
                            gx -= sq[Q + pt];
                            gy -= sq[2 * Q + pt];
                            if (DIM == 3) gz -= sq[3 * Q + pt];
                        }
                    }
                    double w = double(tx.wj[ex * n + qx]) * double(ty.wj[ey * n + qy]);
                    if (DIM == 3) w *= double(tz.wj[ez * n + qz]);
                    const double g2 = gx * gx + gy * gy + gz * gz;
                    acc += w * (kind == 0 ? v * v : kind == 1 ? g2 : v * v + g2);
                }
            }
        }
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (static_cast<int>(threadIdx.x) < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_sum_partials(const double* __restrict__ partial, int count,
                                                      double* __restrict__ out) {
    __shared__ double red[256];
    double acc = 0;
    for (int i = threadIdx.x; i < count; i += 256) acc += partial[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (static_cast<int>(threadIdx.x) < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

template <typename TT, int P>
hipError_t norm_dim(int dim, const double* u, const double* sq, const dev_tables<TT>* d, int kind,
                    double* partial, int blocks, double* out, hipStream_t s) {
    const dev_tables<TT> none{};
    long Q = 1;
    for (int i = 0; i < dim; ++i) Q *= static_cast<long>(d[i].elements) * (P + 1);
    if (dim == 2)
        hipLaunchKernelGGL((k_norm<TT, 2, P>), dim3(blocks), dim3(256), 0, s, u, sq, Q, d[0], d[1],
                           none, kind, partial);
    else
        hipLaunchKernelGGL((k_norm<TT, 3, P>), dim3(blocks), dim3(256), 0, s, u, sq, Q, d[0], d[1],
                           d[2], kind, partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, partial, blocks, out);
    return hipGetLastError();
}

}  // namespace

template <typename TT>
hipError_t launch_norm(int dim, int p, const double* u, const double* sq, const dev_tables<TT>* dirs,
                       int kind, double* partial, int blocks, double* out, hipStream_t s) {
    if (blocks <= 0 || (dim != 2 && dim != 3)) return hipErrorInvalidValue;
    switch (p) {
        case 1:
            return norm_dim<TT, 1>(dim, u, sq, dirs, kind, partial, blocks, out, s);
        case 2:
            return norm_dim<TT, 2>(dim, u, sq, dirs, kind, partial, blocks, out, s);
        case 3:
            return norm_dim<TT, 3>(dim, u, sq, dirs, kind, partial, blocks, out, s);
        default:
            return hipErrorInvalidValue;
    }
}

template hipError_t launch_norm<double>(int, int, const double*, const double*, const dev_tables<double>*,
                                        int, double*, int, double*, hipStream_t);
template hipError_t launch_norm<float>(int, int, const double*, const double*, const dev_tables<float>*,
                                       int, double*, int, double*, hipStream_t);

}  // namespace adsg
