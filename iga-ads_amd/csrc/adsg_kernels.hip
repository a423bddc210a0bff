// Small HIP kernels of the MI355X ADS step (gfx950): fills, boundary handling, precision
// conversions and the increment updates.
#include <hip/hip_runtime.h>

#include "adsg_kernels.hpp"

namespace adsg {

namespace {

template <typename T>
__global__ void k_fill(T* __restrict__ p, long n, T v) {
    for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<long>(gridDim.x) * blockDim.x)
        p[i] = v;
}

template <typename T>
__global__ void k_scale(T* __restrict__ out, const T* __restrict__ in, long n, T a) {
    for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<long>(gridDim.x) * blockDim.x)
        out[i] = a * in[i];
}

template <typename T>
__global__ void k_axpy(T* __restrict__ out, const T* __restrict__ x, long n, T a) {
    for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<long>(gridDim.x) * blockDim.x)
        out[i] += a * x[i];
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

__global__ void k_add_increment(const double* __restrict__ xp, const double* __restrict__ d,
                                double* __restrict__ xn, long n, int dim, int nx, int ny, int nz,
                                int dirichlet) {
    for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<long>(gridDim.x) * blockDim.x) {
        double v = xp[i] + d[i];
        if (dirichlet && on_boundary(i, dim, nx, ny, nz)) v = 0.0;
        xn[i] = v;
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

// number of NaN / Inf entries of v, added into *count (one atomic add per wave with any)
__global__ __launch_bounds__(256) void k_count_nonfinite(const double* __restrict__ v, long n,
                                                         unsigned* __restrict__ count) {
    unsigned bad = 0;
    for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<long>(gridDim.x) * blockDim.x)
        bad += !isfinite(v[i]);
    for (int off = 32; off > 0; off >>= 1) bad += __shfl_xor(bad, off);
    if ((threadIdx.x & 63) == 0 && bad) atomicAdd(count, bad);
}

inline unsigned grid_for(long n) {
    long b = (n + 255) / 256;
    return static_cast<unsigned>(b < 4096 ? (b > 0 ? b : 1) : 4096);
}

}  // namespace

// ------------------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------------------
template <typename T>
hipError_t launch_fill(T* p, long n, T v, hipStream_t s) {
    hipLaunchKernelGGL((k_fill<T>), dim3(grid_for(n)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_scale(T* out, const T* in, long n, T a, hipStream_t s) {
    hipLaunchKernelGGL((k_scale<T>), dim3(grid_for(n)), dim3(256), 0, s, out, in, n, a);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_axpy(T* out, const T* x, long n, T a, hipStream_t s) {
    hipLaunchKernelGGL((k_axpy<T>), dim3(grid_for(n)), dim3(256), 0, s, out, x, n, a);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_zero_boundary(int dim, T* u, int nx, int ny, int nz, hipStream_t s) {
    const long n = static_cast<long>(nx) * ny * nz;
    hipLaunchKernelGGL((k_zero_boundary<T>), dim3(grid_for(n)), dim3(256), 0, s, u, dim, nx, ny, nz);
    return hipGetLastError();
}

hipError_t launch_count_nonfinite(const double* v, long n, unsigned* count, hipStream_t s) {
    hipLaunchKernelGGL(k_count_nonfinite, dim3(grid_for(n)), dim3(256), 0, s, v, n, count);
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

hipError_t launch_add_increment(const double* x_prev, const double* delta, double* x_new, int dim,
                                int nx, int ny, int nz, bool dirichlet, hipStream_t s) {
    const long n = static_cast<long>(nx) * ny * nz;
    hipLaunchKernelGGL(k_add_increment, dim3(grid_for(n)), dim3(256), 0, s, x_prev, delta, x_new, n,
                       dim, nx, ny, nz, dirichlet ? 1 : 0);
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

template hipError_t launch_fill<double>(double*, long, double, hipStream_t);
template hipError_t launch_fill<float>(float*, long, float, hipStream_t);
template hipError_t launch_scale<double>(double*, const double*, long, double, hipStream_t);
template hipError_t launch_scale<float>(float*, const float*, long, float, hipStream_t);
template hipError_t launch_axpy<double>(double*, const double*, long, double, hipStream_t);
template hipError_t launch_zero_boundary<double>(int, double*, int, int, int, hipStream_t);
template hipError_t launch_zero_boundary<float>(int, float*, int, int, int, hipStream_t);

}  // namespace adsg
