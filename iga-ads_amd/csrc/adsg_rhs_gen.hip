// Element-loop RHS for any quadrature order, repeated interior knots and a degree per
// direction (dim_config of simulation/config.hpp:11-56: quad_order, repeated_nodes, p per
// direction): the loop of validation.hpp:100-118 as written there, one lane per element.
// At every Gauss point the state (and k, f) and its gradient are evaluated from the
// element's (p_x+1)(p_y+1)(p_z+1) coefficients, the integrand is formed and its products
// with the element's test functions are accumulated; the element's block is then added to
// the rhs at its first DOFs (bspline.cpp:162-172 first_nonzero_dofs, which repeated knots
// move). Elements whose indices agree modulo p_d + 1 in every direction share no DOF, so
// each of the prod_d (p_d + 1) colour launches adds with plain read-modify-writes and the
// sum does not depend on timing. This path trades speed for generality: the sum-factorised
// kernels cover the default configurations (p+1 Gauss points, simple knots, one degree).
#include <hip/hip_runtime.h>

#include "adsg_kernels.hpp"

namespace adsg {

namespace {

template <typename T, int DIM, int PX, int PY, int PZ>
__global__ __launch_bounds__(256) void k_rhs_gen(const T* __restrict__ up, T* __restrict__ rhs, dev_tables<T> X,
                                                 dev_tables<T> Y, dev_tables<T> Z, rhs_params<T> prm, int cx,
                                                 int cy, int cz, int mx, int my, int mz) {
    constexpr int MX = PX + 1, MY = PY + 1, MZ = DIM == 3 ? PZ + 1 : 1;
    const long l = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (l >= static_cast<long>(mx) * my * mz) return;
    const int ex = cx + MX * static_cast<int>(l % mx);
    const long r = l / mx;
    const int ey = cy + MY * static_cast<int>(r % my);
    const int ez = DIM == 3 ? cz + MZ * static_cast<int>(r / my) : 0;
    if (ex >= X.elements || ey >= Y.elements || (DIM == 3 && ez >= Z.elements)) return;
    const long nx = X.dofs, ny = Y.dofs;
    const int fx = X.first[ex], fy = Y.first[ey], fz = DIM == 3 ? Z.first[ez] : 0;
    const bool KF = prm.kfield != nullptr, SRC = prm.source != nullptr;

    T U[MZ][MY][MX], Kc[MZ][MY][MX], Fc[MZ][MY][MX];
    double R[MZ][MY][MX];
#pragma unroll
    for (int c = 0; c < MZ; ++c)
#pragma unroll
        for (int b = 0; b < MY; ++b)
#pragma unroll
            for (int a = 0; a < MX; ++a) {
                const long g = (fx + a) + nx * ((fy + b) + ny * static_cast<long>(fz + c));
                U[c][b][a] = up[g];
                Kc[c][b][a] = KF ? prm.kfield[g] : T(0);
                Fc[c][b][a] = SRC ? prm.source[g] : T(0);
                R[c][b][a] = 0.0;
            }
    const int qz_end = DIM == 3 ? Z.nq : 1;
    for (int qz = 0; qz < qz_end; ++qz) {
        T bz[MZ], dz[MZ];
        T wz = T(1);
        if constexpr (DIM == 3) {
#pragma unroll
            for (int c = 0; c < MZ; ++c) {
                bz[c] = Z.val[(static_cast<long>(ez) * Z.nq + qz) * MZ + c];
                dz[c] = Z.der[(static_cast<long>(ez) * Z.nq + qz) * MZ + c];
            }
            wz = Z.wj[static_cast<long>(ez) * Z.nq + qz];
        } else {
            bz[0] = T(1);
            dz[0] = T(0);
        }
        for (int qy = 0; qy < Y.nq; ++qy) {
            T by[MY], dy[MY];
#pragma unroll
            for (int b = 0; b < MY; ++b) {
                by[b] = Y.val[(static_cast<long>(ey) * Y.nq + qy) * MY + b];
                dy[b] = Y.der[(static_cast<long>(ey) * Y.nq + qy) * MY + b];
            }
            const T wyz = Y.wj[static_cast<long>(ey) * Y.nq + qy] * wz;
            for (int qx = 0; qx < X.nq; ++qx) {
                T bx[MX], dx[MX];
#pragma unroll
                for (int a = 0; a < MX; ++a) {
                    bx[a] = X.val[(static_cast<long>(ex) * X.nq + qx) * MX + a];
                    dx[a] = X.der[(static_cast<long>(ex) * X.nq + qx) * MX + a];
                }
                const T w = X.wj[static_cast<long>(ex) * X.nq + qx] * wyz;
                // eval_fun (simulation_3d.hpp:120-128): value and gradient at the point
                T u = 0, ux = 0, uy = 0, uz = 0, k = 0, f = 0;
#pragma unroll
                for (int c = 0; c < MZ; ++c)
#pragma unroll
                    for (int b = 0; b < MY; ++b)
#pragma unroll
                        for (int a = 0; a < MX; ++a) {
                            const T v = bz[c] * by[b] * bx[a];
                            u += v * U[c][b][a];
                            ux += bz[c] * by[b] * dx[a] * U[c][b][a];
                            uy += bz[c] * dy[b] * bx[a] * U[c][b][a];
                            if constexpr (DIM == 3) uz += dz[c] * by[b] * bx[a] * U[c][b][a];
                            k += v * Kc[c][b][a];
                            f += v * Fc[c][b][a];
                        }
                if (!KF) k = T(1);
                // the integrand of compute_rhs against v = B_a B_b B_c and its gradient
                const T g0 = w * (prm.mass * u + prm.dt * prm.reaction * u * (T(1) - u) + (SRC ? prm.dt * f : T(0)));
                const T gk = -w * prm.dt * k;
                const T gx = gk * prm.gx * ux, gy = gk * prm.gy * uy, gz = gk * prm.gz * uz;
#pragma unroll
                for (int c = 0; c < MZ; ++c)
#pragma unroll
                    for (int b = 0; b < MY; ++b)
#pragma unroll
                        for (int a = 0; a < MX; ++a) {
                            T s = g0 * (bz[c] * by[b] * bx[a]) + gx * (bz[c] * by[b] * dx[a]) +
                                  gy * (bz[c] * dy[b] * bx[a]);
                            if constexpr (DIM == 3) s += gz * (dz[c] * by[b] * bx[a]);
                            R[c][b][a] += static_cast<double>(s);
                        }
            }
        }
    }
    // update_global_rhs (simulation_3d.hpp:138-145): the colour's elements share no DOF
#pragma unroll
    for (int c = 0; c < MZ; ++c)
#pragma unroll
        for (int b = 0; b < MY; ++b)
#pragma unroll
            for (int a = 0; a < MX; ++a) {
                const long g = (fx + a) + nx * ((fy + b) + ny * static_cast<long>(fz + c));
                rhs[g] += static_cast<T>(R[c][b][a]);
            }
}

template <typename T, int DIM, int PX, int PY, int PZ>
hipError_t gen_colours(const T* up, T* rhs, const dev_tables<T>* d, const rhs_params<T>& prm, hipStream_t s) {
    constexpr int MX = PX + 1, MY = PY + 1, MZ = DIM == 3 ? PZ + 1 : 1;
    const dev_tables<T> none{};
    for (int cz = 0; cz < MZ; ++cz)
        for (int cy = 0; cy < MY; ++cy)
            for (int cx = 0; cx < MX; ++cx) {
                const int mx = (d[0].elements - cx + MX - 1) / MX;
                const int my = (d[1].elements - cy + MY - 1) / MY;
                const int mz = DIM == 3 ? (d[2].elements - cz + MZ - 1) / MZ : 1;
                if (mx <= 0 || my <= 0 || mz <= 0) continue;
                const long lanes = static_cast<long>(mx) * my * mz;
                hipLaunchKernelGGL((k_rhs_gen<T, DIM, PX, PY, PZ>), dim3(static_cast<unsigned>((lanes + 255) / 256)),
                                   dim3(256), 0, s, up, rhs, d[0], d[1], DIM == 3 ? d[2] : none, prm, cx, cy, cz, mx,
                                   my, mz);
                hipError_t e = hipGetLastError();
                if (e != hipSuccess) return e;
            }
    return hipSuccess;
}

template <typename T, int DIM, int PX, int PY>
hipError_t gen_pz(int pz, const T* up, T* rhs, const dev_tables<T>* d, const rhs_params<T>& prm, hipStream_t s) {
    if constexpr (DIM == 2) return gen_colours<T, 2, PX, PY, 1>(up, rhs, d, prm, s);
    switch (pz) {
        case 1:
            return gen_colours<T, DIM, PX, PY, 1>(up, rhs, d, prm, s);
        case 2:
            return gen_colours<T, DIM, PX, PY, 2>(up, rhs, d, prm, s);
        case 3:
            return gen_colours<T, DIM, PX, PY, 3>(up, rhs, d, prm, s);
        default:
            return hipErrorInvalidValue;
    }
}

template <typename T, int DIM, int PX>
hipError_t gen_py(int py, int pz, const T* up, T* rhs, const dev_tables<T>* d, const rhs_params<T>& prm,
                  hipStream_t s) {
    switch (py) {
        case 1:
            return gen_pz<T, DIM, PX, 1>(pz, up, rhs, d, prm, s);
        case 2:
            return gen_pz<T, DIM, PX, 2>(pz, up, rhs, d, prm, s);
        case 3:
            return gen_pz<T, DIM, PX, 3>(pz, up, rhs, d, prm, s);
        default:
            return hipErrorInvalidValue;
    }
}

template <typename T, int DIM>
hipError_t gen_px(const T* up, T* rhs, const dev_tables<T>* d, const rhs_params<T>& prm, hipStream_t s) {
    const int py = d[1].p, pz = DIM == 3 ? d[2].p : 1;
    switch (d[0].p) {
        case 1:
            return gen_py<T, DIM, 1>(py, pz, up, rhs, d, prm, s);
        case 2:
            return gen_py<T, DIM, 2>(py, pz, up, rhs, d, prm, s);
        case 3:
            return gen_py<T, DIM, 3>(py, pz, up, rhs, d, prm, s);
        default:
            return hipErrorInvalidValue;
    }
}

}  // namespace

template <typename T>
hipError_t launch_rhs_general(int dim, const T* u_prev, T* rhs, const dev_tables<T>* dirs, const rhs_params<T>& prm,
                              hipStream_t s) {
    for (int i = 0; i < dim; ++i)
        if (!dirs[i].first || dirs[i].nq < 1 || dirs[i].p < 1 || dirs[i].p > 3) return hipErrorInvalidValue;
    if (dim == 2) return gen_px<T, 2>(u_prev, rhs, dirs, prm, s);
    if (dim == 3) return gen_px<T, 3>(u_prev, rhs, dirs, prm, s);
    return hipErrorInvalidValue;
}

template hipError_t launch_rhs_general<double>(int, const double*, double*, const dev_tables<double>*,
                                               const rhs_params<double>&, hipStream_t);
template hipError_t launch_rhs_general<float>(int, const float*, float*, const dev_tables<float>*,
                                              const rhs_params<float>&, hipStream_t);

}  // namespace adsg
