/*
 * CPU oracle for the alternating-direction (ADS) time step of IGA-ADS.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain, reference-faithful restatement of the
 * reference's CPU step, used by tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py as the checker. The product library (libadsgpu) never links,
 * loads or calls it.
 *
 * Every function cites the reference file:line whose algorithm it restates.
 * Memory layout follows the reference's lin::tensor: first index fastest
 * (lin/tensor/base.hpp:14, util/multi_array/ordering/reverse.hpp:28-31), and the
 * band storage of lin::band_matrix (lin/band_matrix.hpp:66-73): element (i,j) of an
 * n x n matrix with bandwidths kl, ku and row_offset = kl lives at
 * ab[j*ldab + kl + ku + i - j], ldab = 2*kl + ku + 1.
 *
 * Pins (what each function is checked against) are listed in tests/test_oracle_*.py.
 */
#ifndef ADS_ORACLE_H
#define ADS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- random inputs: splitmix64 -> uniform [0,1) with 53 random bits ---- */
unsigned long long orc_splitmix64(unsigned long long* state);
double orc_uniform(unsigned long long* state);

/* ---- Gauss-Legendre quadrature on [-1,1], ascending nodes (quad/gauss.hpp) ---- */
int orc_gauss(int n, double* X, double* W);

/* ---- B-splines (bspline/bspline.cpp) ---- */
int orc_knot_size(int p, int elements, int repeated_nodes);
void orc_create_basis(double a, double b, int p, int elements, int repeated_nodes, double* knot);
int orc_find_span(double x, const double* knot, int knot_size, int p);
void orc_eval_basis(int span, double x, const double* knot, int p, double* out);
void orc_eval_basis_with_derivatives(int span, double x, const double* knot, int p, int der,
                                     double* out /* (der+1) x (p+1), row k = k-th derivative */);
int orc_first_nonzero_dofs(const double* knot, int knot_size, int p, int* out);
int orc_elements_supporting_dofs(const double* knot, int knot_size, int p, int* first, int* last);

/* ---- basis_data tables (basis_data.cpp:63-114), elem_division = 1 ----
 * b: [elements][quad_order][der+1][p+1], x: [elements][quad_order], J: [elements], w: [quad_order]
 * first_dofs: [elements]. Returns the number of elements. */
int orc_basis_data(const double* knot, int knot_size, int p, int der, int quad_order, double* b,
                   double* x, double* J, double* w, int* first_dofs);

/* ---- 1D band matrices (form_matrix.cpp:8-42), kl = ku = p, added into ab ---- */
void orc_gram_matrix_1d(double* ab, int n, int p, int elements, int quad_order, int der,
                        const double* b, const double* J, const double* w, const int* first_dofs);
void orc_stiffness_matrix_1d(double* ab, int n, int p, int elements, int quad_order, int der,
                             const double* b, const double* J, const double* w,
                             const int* first_dofs);
/* (va*vb + h*da*db) form of examples/implicit/implicit.hpp:46-64 (h multiplies the stiffness) */
void orc_mass_plus_stiffness_1d(double* ab, int n, int p, int elements, int quad_order, int der,
                                const double* b, const double* J, const double* w,
                                const int* first_dofs, double h);
/* dimension::fix_dof (dimension.cpp:23-29) */
void orc_fix_dof(double* ab, int n, int p, int k);
/* banded y = op(A) x (band_matrix.hpp:91-106 -> dgbmv), count right-hand sides */
void orc_band_multiply(const double* ab, int rows, int cols, int kl, int ku, int ldab,
                       const double* x, double* y, int count, char trans);

/* ---- LAPACK restatements: dgbtrf (via dgbtf2) and dgbtrs 'N'/'T' ---- */
int orc_dgbtrf(int m, int n, int kl, int ku, double* ab, int ldab, int* ipiv);
void orc_dgbtrs(char trans, int n, int kl, int ku, int nrhs, const double* ab, int ldab,
                const int* ipiv, double* b, int ldb);

/* ---- tensors (lin/tensor/cyclic_transpose.hpp:17-63) ---- */
void orc_cyclic_transpose(int rank, const int* sizes, const double* in, double* out);

/* ---- standard ADS (solver.hpp:30-41, 148-166); per-dimension factored matrices ----
 * ab[d], ipiv[d] factored by orc_dgbtrf with kl[d], ku[d]; result returned in rhs
 * (the odd-rank swap of solver.hpp:156-159 is done as a copy into rhs). */
void orc_ads_solve(int rank, const int* sizes, double* rhs, double* buf, const double* const* ab,
                   const int* const* ipiv, const int* kl, const int* ku);
/* transposed ADS: solves (A_1 (x) ... (x) A_d)^T, i.e. dgbtrs 'T' per direction */
void orc_ads_solve_transposed(int rank, const int* sizes, double* rhs, double* buf,
                              const double* const* ab, const int* const* ipiv, const int* kl,
                              const int* ku);

/* ---- simulation object: the simulation_2d/3d surface restated ---- */
typedef struct orc_sim orc_sim;

typedef struct orc_rhs_form {
    double mass;        /* multiplies u*v (1 for every heat form) */
    double dt;          /* scalar in front of the gradient term */
    double grad[3];     /* per-direction selector inside the gradient product */
    double reaction;    /* coefficient of dt*u*(1-u)*v (reaction-diffusion) */
    const double* kfield; /* optional coefficients of k(x) on the same basis (variable diffusion) */
    const double* source; /* optional coefficients of f(x) on the same basis: + dt*f*v */
    int as_shipped;     /* 1: heat_3d.hpp:49-67 loop (direct scatter per (e,q,a)) */
} orc_rhs_form;

orc_sim* orc_sim_create(int dim, const int* p, const int* elements, const double* a,
                        const double* b, const int* quad_order, const int* repeated_nodes);
void orc_sim_destroy(orc_sim* s);
int orc_sim_dofs(const orc_sim* s, int dir);
int orc_sim_size(const orc_sim* s);
/* matrix slots per direction: slot 0 = Gram M (dimension::M); others set by the caller */
int orc_sim_fix_dof(orc_sim* s, int dir, int k);
int orc_sim_set_matrix(orc_sim* s, int dir, int slot, double mass, double stiff);
int orc_sim_get_matrix(const orc_sim* s, int dir, int slot, double* ab_out, int* ipiv_out);
int orc_sim_prepare_matrices(orc_sim* s); /* dgbtrf every slot; returns first nonzero info */
void orc_sim_compute_rhs(const orc_sim* s, const orc_rhs_form* f, const double* u_prev,
                         double* rhs);
void orc_sim_zero_boundary(const orc_sim* s, double* v);
void orc_sim_solve(orc_sim* s, double* rhs, const int* slots);
void orc_sim_solve_transposed(orc_sim* s, double* rhs, const int* slots);
int orc_sim_tables(const orc_sim* s, int dir, double* b, double* x, double* J, double* w,
                   int* first_dofs);
/* L2-projection right-hand side of a coefficient field (projection.hpp:11-153 pattern) and
 * the weighted norm sum_e sum_q w J u(x_q)^2 (implicit.hpp L2norm) */
double orc_sim_l2norm_sq(const orc_sim* s, const double* u);
/* squared L2 (kind 0), H1_0 seminorm (1) or H1 norm (2) of u (basic_simulation_2d.hpp:42-52) */
double orc_sim_norm_sq(const orc_sim* s, const double* u, int kind);

/* ---- threaded element loop: executor contract for_each + synchronized
 * (executor/sequential.hpp:12-26, executor/galois.hpp:27-44) with std::thread workers ---- */
void orc_sim_compute_rhs_threaded(const orc_sim* s, const orc_rhs_form* f, const double* u_prev,
                                  double* rhs, int threads);
/* the sequential loop over elements [first, first + count) only, added into rhs: a bounded
 * sample of one assembly for single-core timing */
void orc_sim_compute_rhs_sample(const orc_sim* s, const orc_rhs_form* f, const double* u_prev,
                                double* rhs, long first, long count);

/* ---- initial states of the benchmark problems ---- */
/* d/dk_b of w^T rhs(u; k) for the gradient term of form (-dt k grad u . grad v), added
 * into out (no counterpart in the reference; the k-field adjoint). */
void orc_sim_kgrad(const orc_sim* s, const orc_rhs_form* f, const double* u, const double* w,
                   double* out);
/* projector<N> (projection.hpp:11-153, via simulation_2d.hpp:43-46): v_i += sum_e sum_q
 * f(x_q) B_i(x_q) w_q J_e, f given at the Gauss points: fq[ix + NX*(iy + NY*iz)] with
 * ix = e_x*q + q_x, NX = elements_x*q (likewise y, z). Adds into v. */
void orc_sim_project(const orc_sim* s, const double* fq, double* v);
/* error(u, Ux, Uy, norm, fun) (basic_simulation_2d.hpp:318-336, basic_simulation_3d.hpp:
 * 336-357): sum_e sum_q phi(u_h(x_q) - s(x_q)) w J (squared, no sqrt), u_h the spline of u
 * (0 when u is NULL), s at the Gauss points as planes [value | d/dx | d/dy (| d/dz)] of
 * Q values (only the value plane for kind 0); phi: 0 L2, 1 H10, 2 H1. */
double orc_sim_error_sq(const orc_sim* s, const double* u, const double* sq, int kind);
/* bspline::eval / eval_ders (bspline/eval.hpp:73-234) of the spline u on the tensor grid
 * px[0..counts[0]) x py x pz (x fastest): out[pt] the value; with ders, planes
 * [value | d/dx | d/dy (| d/dz)] of prod(counts) values each. */
void orc_sim_sample(const orc_sim* s, const double* u, const int* counts, const double* px,
                    const double* py, const double* pz, int ders, double* out);
/* the Gauss points of direction dir (basis_data x[e][q]), elements*q values */
int orc_sim_quad_points(const orc_sim* s, int dir, double* out);
void orc_greville(const double* knot, int knot_size, int p, double* out);
void orc_init_gaussians(const orc_sim* s, unsigned long long seed, int count, double sigma_lo,
                        double sigma_hi, double centre_lo, double centre_hi, double noise,
                        double* u);

#ifdef __cplusplus
}
#endif

#endif
