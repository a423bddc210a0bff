// Thin extern "C" driver over the reference's own, unmodified sources -- TEST
// INFRASTRUCTURE ONLY. Compiled by oracle/Makefile from the reference tree where it lies
// (bspline/bspline.cpp, quad/gauss_data.cpp and the header-only solver.hpp + lin/*) and
// linked to the image's LAPACK (MKL) into oracle/_ref/libadsref.so. It validates the
// restatement in ads_oracle.cpp; it is never shipped or used by the product path.
//
// The reference's basis_data/form_matrix/simulation need Boost, which the image lacks,
// so they are not built here.
#include <algorithm>
#include <array>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

#include "ads/bspline/bspline.hpp"
#include "ads/bspline/eval.hpp"
#include "ads/lin/band_matrix.hpp"
#include "ads/lin/band_solve.hpp"
#include "ads/lin/solver_ctx.hpp"
#include "ads/lin/tensor.hpp"
#include "ads/output/gnuplot.hpp"
#include "ads/output/grid.hpp"
#include "ads/output/range.hpp"
#include "ads/quad/gauss.hpp"
#include "ads/solver.hpp"

namespace bsp = ads::bspline;
namespace lin = ads::lin;

static bsp::basis make_basis(const double* knot, int knot_size, int p) {
    return bsp::basis{bsp::knot_vector(knot, knot + knot_size), p};
}

// coefficient tensors (first index fastest) as the callables bspline::eval reads
struct coeffs2 {
    const double* u;
    int nx;
    double operator()(int i, int j) const { return u[i + static_cast<long>(nx) * j]; }
};
struct coeffs3 {
    const double* u;
    int nx, ny;
    double operator()(int i, int j, int k) const { return u[i + static_cast<long>(nx) * (j + static_cast<long>(ny) * k)]; }
};

extern "C" {

int ref_create_basis(double a, double b, int p, int elements, int repeated_nodes, double* knot) {
    auto B = bsp::create_basis(a, b, p, elements, repeated_nodes);
    if (knot) std::copy(B.knot.begin(), B.knot.end(), knot);
    return B.knot_size();
}

int ref_create_basis_plain(double a, double b, int p, int elements, double* knot) {
    auto B = bsp::create_basis(a, b, p, elements);
    if (knot) std::copy(B.knot.begin(), B.knot.end(), knot);
    return B.knot_size();
}

int ref_find_span(double x, const double* knot, int knot_size, int p) {
    auto B = make_basis(knot, knot_size, p);
    return bsp::find_span(x, B);
}

void ref_eval_basis(int span, double x, const double* knot, int knot_size, int p, double* out) {
    auto B = make_basis(knot, knot_size, p);
    bsp::eval_ctx ctx(p);
    bsp::eval_basis(span, x, B, out, ctx);
}

// out: (der+1) x (p+1)
void ref_eval_basis_with_derivatives(int span, double x, const double* knot, int knot_size, int p,
                                     int der, double* out) {
    auto B = make_basis(knot, knot_size, p);
    bsp::eval_ctx ctx(p);  // the context basis_data uses (basis_data.cpp:87)
    std::vector<double*> rows(der + 1);
    for (int k = 0; k <= der; ++k) rows[k] = out + k * (p + 1);
    bsp::eval_basis_with_derivatives(span, x, B, rows.data(), der, ctx);
}

int ref_first_nonzero_dofs(const double* knot, int knot_size, int p, int* out) {
    auto v = bsp::first_nonzero_dofs(make_basis(knot, knot_size, p));
    std::copy(v.begin(), v.end(), out);
    return static_cast<int>(v.size());
}

int ref_elements_supporting_dofs(const double* knot, int knot_size, int p, int* first, int* last) {
    auto v = bsp::elements_supporting_dofs(make_basis(knot, knot_size, p));
    for (size_t i = 0; i < v.size(); ++i) {
        first[i] = v[i].first;
        last[i] = v[i].second;
    }
    return static_cast<int>(v.size());
}

int ref_gauss(int n, double* X, double* W) {
    if (n < 2 || n > 64) return -1;
    std::memcpy(X, ads::quad::gauss::Xs[n], sizeof(double) * n);
    std::memcpy(W, ads::quad::gauss::Ws[n], sizeof(double) * n);
    return 0;
}

// lin::factorize on an n x n band_matrix(kl, ku, n) whose full buffer is ab (in/out)
int ref_factorize(int n, int kl, int ku, double* ab, int* ipiv) {
    lin::band_matrix M{kl, ku, n};
    std::memcpy(M.full_buffer(), ab, sizeof(double) * M.column_size() * n);
    lin::solver_ctx ctx{M};
    lin::factorize(M, ctx);
    std::memcpy(ab, M.full_buffer(), sizeof(double) * M.column_size() * n);
    std::copy(ctx.pivot_vector.begin(), ctx.pivot_vector.end(), ipiv);
    return ctx.info;
}

// lin::solve_with_factorized with nrhs right-hand sides of length n
void ref_solve_with_factorized(int n, int kl, int ku, const double* ab, const int* ipiv, double* b,
                               int nrhs) {
    lin::band_matrix M{kl, ku, n};
    std::memcpy(M.full_buffer(), ab, sizeof(double) * M.column_size() * n);
    lin::solver_ctx ctx{M};
    std::copy(ipiv, ipiv + n, ctx.pivot_vector.begin());
    lin::solve_with_factorized(M, b, ctx, nrhs);
}

void ref_cyclic_transpose(int rank, const int* sizes, const double* in, double* out) {
    if (rank == 2) {
        auto a = lin::as_tensor(const_cast<double*>(in), std::array<int, 2>{sizes[0], sizes[1]});
        lin::cyclic_transpose(a, out);
    } else if (rank == 3) {
        auto a = lin::as_tensor(const_cast<double*>(in),
                                std::array<int, 3>{sizes[0], sizes[1], sizes[2]});
        lin::cyclic_transpose(a, out);
    }
}

// ads_solve(rhs, buf, dim_data...) for rank 2/3 with already factored matrices
void ref_ads_solve(int rank, const int* sizes, double* rhs, const double* const* ab,
                   const int* const* ipiv, const int* kl, const int* ku) {
    std::vector<lin::band_matrix> M;
    std::vector<lin::solver_ctx> C;
    M.reserve(3);
    C.reserve(3);
    for (int d = 0; d < rank; ++d) {
        M.emplace_back(kl[d], ku[d], sizes[d]);
        std::memcpy(M.back().full_buffer(), ab[d], sizeof(double) * M.back().column_size() * sizes[d]);
        C.emplace_back(M.back());
        std::copy(ipiv[d], ipiv[d] + sizes[d], C.back().pivot_vector.begin());
    }
    long total = 1;
    for (int d = 0; d < rank; ++d) total *= sizes[d];
    if (rank == 2) {
        lin::tensor<double, 2> r{{sizes[0], sizes[1]}}, buf{{sizes[0], sizes[1]}};
        std::memcpy(r.data(), rhs, sizeof(double) * total);
        ads::ads_solve(r, buf, ads::dim_data{M[0], C[0]}, ads::dim_data{M[1], C[1]});
        std::memcpy(rhs, r.data(), sizeof(double) * total);
    } else if (rank == 3) {
        lin::tensor<double, 3> r{{sizes[0], sizes[1], sizes[2]}},
            buf{{sizes[0], sizes[1], sizes[2]}};
        std::memcpy(r.data(), rhs, sizeof(double) * total);
        ads::ads_solve(r, buf, ads::dim_data{M[0], C[0]}, ads::dim_data{M[1], C[1]},
                       ads::dim_data{M[2], C[2]});
        std::memcpy(rhs, r.data(), sizeof(double) * total);
    }
}

// bspline::eval / eval_ders (bspline/eval.hpp:104-234): the spline of u (first index
// fastest) at (x, y[, z]); out = value (ders = 0) or value, d/dx, d/dy[, d/dz] (ders = 1)
void ref_eval(int dim, const double* pt, const double* u, const double* const* knot,
              const int* knot_size, const int* p, int ders, double* out) {
    std::vector<bsp::basis> b;
    for (int d = 0; d < dim; ++d) b.push_back(make_basis(knot[d], knot_size[d], p[d]));
    if (dim == 2) {
        coeffs2 U{u, b[0].dofs()};
        if (!ders) {
            bsp::eval_ctx cx(p[0]), cy(p[1]);
            out[0] = bsp::eval(pt[0], pt[1], U, b[0], b[1], cx, cy);
        } else {
            bsp::eval_ders_ctx cx(p[0], 1), cy(p[1], 1);
            auto v = bsp::eval_ders(pt[0], pt[1], U, b[0], b[1], cx, cy);
            out[0] = v.val;
            out[1] = v.dx;
            out[2] = v.dy;
        }
    } else {
        coeffs3 U{u, b[0].dofs(), b[1].dofs()};
        if (!ders) {
            bsp::eval_ctx cx(p[0]), cy(p[1]), cz(p[2]);
            out[0] = bsp::eval(pt[0], pt[1], pt[2], U, b[0], b[1], b[2], cx, cy, cz);
        } else {
            bsp::eval_ders_ctx cx(p[0], 1), cy(p[1], 1), cz(p[2], 1);
            auto v = bsp::eval_ders(pt[0], pt[1], pt[2], U, b[0], b[1], b[2], cx, cy, cz);
            out[0] = v.val;
            out[1] = v.dx;
            out[2] = v.dy;
            out[3] = v.dz;
        }
    }
}

// output::gnuplot_printer<2> with the output manager's format (fixed_format(10, 18),
// output_manager.hpp:23, gnuplot.hpp:39-55): vals[i + nx*j] at (xs[i], ys[j]); the text is
// copied into out (at most cap bytes) and its length returned
long ref_gnuplot2(const double* xs, int nx, const double* ys, int ny, const double* vals, char* out,
                  long cap) {
    std::vector<double> X(xs, xs + nx), Y(ys, ys + ny);
    auto grid = ads::output::make_grid(ads::output::from_container(X), ads::output::from_container(Y));
    ads::output::gnuplot_printer<2> printer{ads::output::fixed_format(10, 18)};
    auto value = [&](int i, int j) { return vals[i + static_cast<long>(nx) * j]; };
    std::ostringstream os;
    printer.print(os, grid, value);
    const std::string t = os.str();
    if (out) std::memcpy(out, t.data(), std::min<long>(cap, static_cast<long>(t.size())));
    return static_cast<long>(t.size());
}

}  // extern "C"
