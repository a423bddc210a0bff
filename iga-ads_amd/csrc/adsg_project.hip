// HIP kernels of the L2 projection's load vector and of grid sampling (gfx950). projection(v, f)
// (simulation_2d.hpp:43-46, simulation_3d.hpp:45-48 -> projection.hpp:11-153) adds
//   v_i += sum_e sum_q f(x_q) B_i(x_q) w_q J_e
// for f given at the Gauss points. One lane per element, sum-factorised over the Gauss
// points direction by direction, FP64 accumulation. Elements whose indices agree modulo
// p+1 in every direction share no DOF, so the (p+1)^dim colour launches add their local
// blocks with plain read-modify-writes, deterministically.
#include <hip/hip_runtime.h>

#include "adsg_kernels.hpp"

namespace adsg {

namespace {

template <typename TT, int DIM, int P>
__global__ __launch_bounds__(256) void k_project(const double* __restrict__ fq, double* __restrict__ out,
                                                 dev_tables<TT> tx, dev_tables<TT> ty,
                                                 dev_tables<TT> tz, int cx, int cy, int cz, int mx,
                                                 int my, int mz) {
    constexpr int n = P + 1;
    constexpr int NZ = DIM == 3 ? n : 1;
    const long lane = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (lane >= static_cast<long>(mx) * my * mz) return;
    const int ix = static_cast<int>(lane % mx);
    const long r = lane / mx;
    const int iy = static_cast<int>(r % my), iz = static_cast<int>(r / my);
    const int ex = cx + n * ix, ey = cy + n * iy, ez = DIM == 3 ? cz + n * iz : 0;
    const long NX = static_cast<long>(tx.elements) * n, NY = static_cast<long>(ty.elements) * n;
    const TT* vx = tx.val + static_cast<long>(ex) * n * n;
    const TT* vy = ty.val + static_cast<long>(ey) * n * n;
    const TT* vz = DIM == 3 ? tz.val + static_cast<long>(ez) * n * n : nullptr;
    // x stage: X[qz][qy][a] = sum_qx f w B_a(x_qx)
    double X[NZ][n][n];
    for (int qz = 0; qz < NZ; ++qz)
        for (int qy = 0; qy < n; ++qy) {
            double w = double(ty.wj[ey * n + qy]);
            if (DIM == 3) w *= double(tz.wj[ez * n + qz]);
            const long row = static_cast<long>(ex) * n +
                             NX * ((static_cast<long>(ey) * n + qy) + NY * (static_cast<long>(ez) * n + qz));
            for (int a = 0; a < n; ++a) X[qz][qy][a] = 0.0;
            for (int qx = 0; qx < n; ++qx) {
                const double fw = fq[row + qx] * (double(tx.wj[ex * n + qx]) * w);
                for (int a = 0; a < n; ++a) X[qz][qy][a] += fw * double(vx[qx * n + a]);
            }
        }
    // y stage: Y[qz][b][a] = sum_qy B_b(y_qy) X[qz][qy][a]
    double Y[NZ][n][n];
    for (int qz = 0; qz < NZ; ++qz)
        for (int b = 0; b < n; ++b)
            for (int a = 0; a < n; ++a) {
                double s = 0.0;
                for (int qy = 0; qy < n; ++qy) s += double(vy[qy * n + b]) * X[qz][qy][a];
                Y[qz][b][a] = s;
            }
    // z stage and the add into the element's DOF block
    const long nx = tx.dofs, ny = ty.dofs;
    for (int c = 0; c < NZ; ++c)
        for (int b = 0; b < n; ++b)
            for (int a = 0; a < n; ++a) {
                double s;
                if constexpr (DIM == 3) {
                    s = 0.0;
                    for (int qz = 0; qz < n; ++qz) s += double(vz[qz * n + c]) * Y[qz][b][a];
                } else {
                    s = Y[0][b][a];
                }
                out[(ex + a) + nx * ((ey + b) + ny * static_cast<long>(ez + c))] += s;
            }
}

template <typename TT, int DIM, int P>
hipError_t project_colours(const double* fq, double* out, const dev_tables<TT>* d, hipStream_t s) {
    constexpr int n = P + 1;
    const dev_tables<TT> none{};
    const int cz_n = DIM == 3 ? n : 1;
    for (int cz = 0; cz < cz_n; ++cz)
        for (int cy = 0; cy < n; ++cy)
            for (int cx = 0; cx < n; ++cx) {
                const int mx = (d[0].elements - cx + n - 1) / n;
                const int my = (d[1].elements - cy + n - 1) / n;
                const int mz = DIM == 3 ? (d[2].elements - cz + n - 1) / n : 1;
                if (mx <= 0 || my <= 0 || mz <= 0) continue;
                const long lanes = static_cast<long>(mx) * my * mz;
                hipLaunchKernelGGL((k_project<TT, DIM, P>), dim3(static_cast<unsigned>((lanes + 255) / 256)),
                                   dim3(256), 0, s, fq, out, d[0], d[1], DIM == 3 ? d[2] : none, cx, cy,
                                   cz, mx, my, mz);
                hipError_t e = hipGetLastError();
                if (e != hipSuccess) return e;
            }
    return hipSuccess;
}

template <typename TT, int P>
hipError_t project_dim(int dim, const double* fq, double* out, const dev_tables<TT>* d, hipStream_t s) {
    if (dim == 2) return project_colours<TT, 2, P>(fq, out, d, s);
    if (dim == 3) return project_colours<TT, 3, P>(fq, out, d, s);
    return hipErrorInvalidValue;
}

// The spline u on a tensor grid of points (bspline/eval.hpp:73-234): per direction and
// point, off[i] = span - p and tab[i] = the p+1 nonzero basis values then their derivatives;
// one lane per point (x fastest), FP64. out: values, and with ders the derivative planes.
template <int DIM, int P>
__global__ __launch_bounds__(256) void k_sample(const double* __restrict__ u, long nx, long ny,
                                                const double* __restrict__ tx, const int* __restrict__ ox,
                                                const double* __restrict__ ty, const int* __restrict__ oy,
                                                const double* __restrict__ tz, const int* __restrict__ oz,
                                                int cx, int cy, int cz, int ders, double* __restrict__ out) {
    constexpr int m = P + 1;
    constexpr int MZ = DIM == 3 ? m : 1;
    const long total = static_cast<long>(cx) * cy * cz;
    const long pt = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (pt >= total) return;
    const int i = static_cast<int>(pt % cx);
    const long r = pt / cx;
    const int j = static_cast<int>(r % cy), k = static_cast<int>(r / cy);
    const double* bx = tx + static_cast<long>(i) * 2 * m;
    const double* by = ty + static_cast<long>(j) * 2 * m;
    const double* bz = DIM == 3 ? tz + static_cast<long>(k) * 2 * m : nullptr;
    const int x0 = ox[i], y0 = oy[j], z0 = DIM == 3 ? oz[k] : 0;
    double v = 0, gx = 0, gy = 0, gz = 0;
    for (int a = 0; a < m; ++a)
        for (int b = 0; b < m; ++b)
            for (int c = 0; c < MZ; ++c) {
                const double cu = u[(x0 + a) + nx * ((y0 + b) + ny * static_cast<long>(z0 + c))];
                const double vz = DIM == 3 ? bz[c] : 1.0, dz = DIM == 3 ? bz[m + c] : 0.0;
                v += cu * bx[a] * by[b] * vz;
                gx += cu * bx[m + a] * by[b] * vz;
                gy += cu * bx[a] * by[m + b] * vz;
                gz += cu * bx[a] * by[b] * dz;
            }
    out[pt] = v;
    if (ders) {
        out[total + pt] = gx;
        out[2 * total + pt] = gy;
        if (DIM == 3) out[3 * total + pt] = gz;
    }
}

template <int DIM>
hipError_t sample_dim(int p, const double* u, long nx, long ny, const double* const* tab,
                      const int* const* off, const int* counts, int ders, double* out, hipStream_t s) {
    const long total = static_cast<long>(counts[0]) * counts[1] * (DIM == 3 ? counts[2] : 1);
    const unsigned blocks = static_cast<unsigned>((total + 255) / 256);
    const int cz = DIM == 3 ? counts[2] : 1;
#define ADSG_SAMPLE(P_)                                                                             \
    if (p == P_) {                                                                                  \
        hipLaunchKernelGGL((k_sample<DIM, P_>), dim3(blocks), dim3(256), 0, s, u, nx, ny, tab[0], off[0], \
                           tab[1], off[1], tab[2], off[2], counts[0], counts[1], cz, ders, out);  \
        return hipGetLastError();                                                                   \
    }
    ADSG_SAMPLE(1)
    ADSG_SAMPLE(2)
    ADSG_SAMPLE(3)
#undef ADSG_SAMPLE
    return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_sample(int dim, int p, const double* u, long nx, long ny, const double* const* tab,
                         const int* const* off, const int* counts, int ders, double* out, hipStream_t s) {
    if (counts[0] <= 0 || counts[1] <= 0 || (dim == 3 && counts[2] <= 0)) return hipSuccess;
    if (dim == 2) return sample_dim<2>(p, u, nx, ny, tab, off, counts, ders, out, s);
    if (dim == 3) return sample_dim<3>(p, u, nx, ny, tab, off, counts, ders, out, s);
    return hipErrorInvalidValue;
}

template <typename TT>
hipError_t launch_project(int dim, int p, const double* fq, double* out, const dev_tables<TT>* dirs,
                          hipStream_t s) {
    switch (p) {
        case 1:
            return project_dim<TT, 1>(dim, fq, out, dirs, s);
        case 2:
            return project_dim<TT, 2>(dim, fq, out, dirs, s);
        case 3:
            return project_dim<TT, 3>(dim, fq, out, dirs, s);
        default:
            return hipErrorInvalidValue;
    }
}

template hipError_t launch_project<double>(int, int, const double*, double*, const dev_tables<double>*,
                                           hipStream_t);
template hipError_t launch_project<float>(int, int, const double*, double*, const dev_tables<float>*,
                                          hipStream_t);

}  // namespace adsg
