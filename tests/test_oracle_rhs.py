"""Pins of the oracle's element-loop RHS (compute_rhs, validation.hpp:95-119 form) and of
one full explicit step, by formulations that share no code with the element loop:

* Kronecker form: with p+1 Gauss points the heat integrand is integrated exactly, so the
  assembled rhs equals (M - dt K) c with M = Mz (x) My (x) Mx and K the sum of the three
  stiffness-in-one-direction products, all built from the pinned 1D matrices.
* Quadrature-matrix form: global basis matrices at all quadrature points (numpy), which
  covers the forms that are not integrated exactly (variable k, reaction, source).
"""
import numpy as np
import pytest

import pyoracle as po


def _dense(ab, n, p):
    A = np.zeros((n, n))
    for i in range(n):
        for j in range(max(0, i - p), min(n, i + p + 1)):
            A[i, j] = ab[j, 2 * p + i - j]
    return A


def _mats(sim):
    """1D Gram and stiffness matrices of every direction (before any fix_dof)."""
    Ms, Ks = [], []
    for d in range(sim.dim):
        ab, _ = sim.matrix(d, 0)
        Ms.append(_dense(ab, sim.shape[d], sim.p[d]))
        sim.set_matrix(d, 7, 0.0, 1.0)
        abk, _ = sim.matrix(d, 7)
        Ks.append(_dense(abk, sim.shape[d], sim.p[d]))
    return Ms, Ks


def _apply(mats, c, shape):
    """(A_{d-1} (x) ... (x) A_0) c for first-index-fastest c."""
    t = c.reshape(shape[::-1])
    for d, A in enumerate(mats):
        axis = len(shape) - 1 - d
        t = np.moveaxis(np.tensordot(A, t, axes=([1], [axis])), 0, axis)
    return t.reshape(-1)


@pytest.mark.parametrize("dim,p,ne", [(2, 1, (5, 4)), (2, 2, (6, 5)), (2, 3, (7, 6)),
                                      (3, 1, (3, 4, 2)), (3, 2, (4, 3, 5)), (3, 3, (3, 4, 3))])
def test_heat_rhs_equals_kronecker_form(dim, p, ne):
    sim = po.Sim(dim, [p] * dim, list(ne))
    rng = np.random.default_rng(dim * 10 + p)
    c = rng.standard_normal(sim.size)
    dt = 1e-3
    Ms, Ks = _mats(sim)
    expect = _apply(Ms, c, sim.shape)
    for d in range(dim):
        mats = list(Ms)
        mats[d] = Ks[d]
        expect = expect - dt * _apply(mats, c, sim.shape)
    got = sim.compute_rhs(c, po.heat_form(dt, dim))
    assert po.rel_l2(got, expect) < 2e-14
    shipped = sim.compute_rhs(c, po.heat_form(dt, dim, as_shipped=True))
    assert po.rel_l2(shipped, got) < 4e-15
    threaded = sim.compute_rhs(c, po.heat_form(dt, dim), threads=3)
    assert po.rel_l2(threaded, got) < 4e-15


def test_split_rhs_equals_kronecker_form():
    # implicit.hpp:132-156 (compute_rhs_1): u*v - 0.5*dt*u_y*v_y
    sim = po.Sim(2, [2, 2], [6, 7])
    rng = np.random.default_rng(1)
    c = rng.standard_normal(sim.size)
    dt = 1e-2
    Ms, Ks = _mats(sim)
    expect = _apply(Ms, c, sim.shape) - 0.5 * dt * _apply([Ms[0], Ks[1]], c, sim.shape)
    got = sim.compute_rhs(c, po.heat_form(0.5 * dt, 2, grad=(0.0, 1.0, 0.0)))
    assert po.rel_l2(got, expect) < 2e-14


def _quadrature_matrices(sim, d):
    """Global basis values / derivatives at all Gauss points of direction d, plus the
    weights w_q * J_e, from point evaluation of the spline basis (no tables)."""
    p = sim.p[d]
    ne = sim.elements[d]
    k = po.knots(0.0, 1.0, p, ne)
    X, W = po.gauss(p + 1)
    rows_b, rows_d, wts = [], [], []
    for e in range(ne):
        a, b = k[p + e], k[p + e + 1]
        for q in range(p + 1):
            x = a + (b - a) * (X[q] + 1) / 2
            s = po.find_span(x, k, p)
            v = po.eval_basis_with_derivatives(s, x, k, p, 1)
            rb = np.zeros(sim.shape[d])
            rd = np.zeros(sim.shape[d])
            rb[s - p: s + 1] = v[0]
            rd[s - p: s + 1] = v[1]
            rows_b.append(rb)
            rows_d.append(rd)
            wts.append(W[q] * (b - a) / 2)
    return np.array(rows_b), np.array(rows_d), np.array(wts)


def _qform_rhs(sim, c, dt, kfield=None, reaction=0.0, source=None):
    """Quadrature-matrix restatement (2D): rhs = B^T W [u + dt r u(1-u) + dt f]
    - dt sum_i D_i^T W [k du/dx_i]."""
    Bx, Dx, wx = _quadrature_matrices(sim, 0)
    By, Dy, wy = _quadrature_matrices(sim, 1)
    B = np.kron(By, Bx)
    DX = np.kron(By, Dx)
    DY = np.kron(Dy, Bx)
    w = np.kron(wy, wx)
    u = B @ c
    ux = DX @ c
    uy = DY @ c
    k = B @ kfield if kfield is not None else np.ones_like(u)
    f = B @ source if source is not None else np.zeros_like(u)
    val = u + dt * reaction * u * (1 - u) + dt * f
    return B.T @ (w * val) - dt * (DX.T @ (w * k * ux) + DY.T @ (w * k * uy))


@pytest.mark.parametrize("p", [1, 2, 3])
def test_rhs_forms_match_quadrature_matrix_form(p):
    sim = po.Sim(2, [p, p], [6, 5])
    rng = np.random.default_rng(100 + p)
    c = rng.uniform(0, 1, sim.size)
    kf = np.exp(rng.uniform(-0.5, 0.5, sim.size))
    src = rng.standard_normal(sim.size)
    dt = 3e-3
    cases = [
        (dict(), dict()),
        (dict(kfield=kf), dict(kfield=kf)),
        (dict(reaction=1.0), dict(reaction=1.0)),
        (dict(source=src), dict(source=src)),
        (dict(kfield=kf, reaction=0.7, source=src), dict(kfield=kf, reaction=0.7, source=src)),
    ]
    for okw, qkw in cases:
        got = sim.compute_rhs(c, po.heat_form(dt, 2, **okw))
        expect = _qform_rhs(sim, c, dt, **qkw)
        assert po.rel_l2(got, expect) < 1e-13, okw.keys()


def test_constant_state_is_steady_for_heat():
    # u = 1 without boundary conditions: rhs = M 1, and the ADS step returns 1
    for dim, p, ne in [(2, 2, (8, 8)), (3, 2, (4, 5, 3)), (3, 3, (3, 3, 4)), (2, 1, (9, 4))]:
        sim = po.Sim(dim, [p] * dim, list(ne))
        assert sim.prepare_matrices() == 0
        u = np.ones(sim.size)
        rhs = sim.compute_rhs(u, po.heat_form(1e-4, dim))
        sim.solve(rhs)
        assert np.max(np.abs(rhs - 1.0)) < 5e-12  # rounding x cond(M) (up to ~2e4 at p=3)


def test_mass_only_rhs_inverts_to_state():
    # with dt = 0 the rhs is M c and one ADS solve recovers c (simulation_2d.hpp:41)
    sim = po.Sim(3, [2, 2, 2], [5, 4, 6])
    sim.prepare_matrices()
    c = np.random.default_rng(9).standard_normal(sim.size)
    rhs = sim.compute_rhs(c, po.heat_form(0.0, 3))
    sim.solve(rhs)
    assert po.rel_l2(rhs, c) < 1e-13


def test_dirichlet_step_keeps_boundary_zero():
    # validation.hpp:31-63 flow: fixed rows + zeroed rhs sides -> boundary coefficients 0
    sim = po.Sim(2, [2, 2], [10, 10])
    sim.fix_all_boundaries()
    sim.prepare_matrices()
    u = sim.init_gaussians(5, 1)
    for _ in range(3):
        rhs = sim.compute_rhs(u, po.heat_form(1e-5, 2))
        sim.zero_boundary(rhs)
        sim.solve(rhs)
        u = rhs
    g = u.reshape(sim.shape[::-1])
    assert np.all(g[0, :] == 0) and np.all(g[-1, :] == 0)
    assert np.all(g[:, 0] == 0) and np.all(g[:, -1] == 0)


def test_implicit_matrix_is_mass_plus_half_dt_stiffness():
    # implicit.hpp:46-64 builds K = M + 0.5*dt*K_stiff with the (va*vb + h*da*db) form
    sim = po.Sim(2, [3, 3], [6, 6])
    Ms, Ks = _mats(sim)
    sim.set_matrix(0, 1, 1.0, 0.5 * 1e-2)
    ab, _ = sim.matrix(0, 1)
    assert np.allclose(_dense(ab, sim.shape[0], 3), Ms[0] + 0.5e-2 * Ks[0], rtol=1e-14, atol=1e-17)


def test_heat_amplification_of_a_mode():
    # for u = M^{-1}-eigen-mode of the generalized problem K v = lam M v (1D), one explicit
    # step multiplies by (1 - dt*lam); checked on a 2D separable mode
    sim = po.Sim(2, [2, 2], [12, 12])
    Ms, Ks = _mats(sim)
    import scipy.linalg

    lam, V = scipy.linalg.eigh(Ks[0], Ms[0])
    v = V[:, 5]
    c = np.kron(v, v)  # symmetric mesh: same mode in x and y
    sim.prepare_matrices()
    dt = 1e-5
    rhs = sim.compute_rhs(c, po.heat_form(dt, 2))
    sim.solve(rhs)
    assert po.rel_l2(rhs, (1 - 2 * dt * lam[5]) * c) < 1e-12


@pytest.mark.parametrize("dim,p,ne", [(2, 2, (3, 4)), (3, 2, (2, 3, 2)), (2, 3, (3, 2))])
def test_kgrad_equals_rhs_with_unit_k_fields(dim, p, ne):
    # w^T rhs(u; k) is linear in k, so d/dk_b = w^T rhs(u; k = e_b) with the mass term off
    o = po.Sim(dim, [p] * dim, list(ne))
    rng = np.random.default_rng(40 + dim)
    u = rng.standard_normal(o.size)
    w = rng.standard_normal(o.size)
    grad = (1.0, 0.7, 1.3)
    g = o.kgrad(u, w, po.heat_form(0.25, dim, grad=grad))
    ref = np.zeros(o.size)
    for b in range(o.size):
        e = np.zeros(o.size)
        e[b] = 1.0
        f = po.heat_form(0.25, dim, grad=grad, kfield=e)
        f.mass = 0.0
        ref[b] = w @ o.compute_rhs(u, f)
    assert np.allclose(g, ref, rtol=1e-12, atol=1e-14)
    assert np.abs(g).max() > 1e-3


@pytest.mark.parametrize("dim,p,ne", [(2, 2, (5, 4)), (3, 1, (3, 2, 4)), (2, 3, (4, 3))])
def test_norms_equal_quadratic_forms_of_the_1d_matrices(dim, p, ne):
    # exact quadrature: ||u||_L2^2 = u^T (M_z (x) M_y (x) M_x) u and
    # |u|_H1^2 = u^T (sum_d K_d (x) M_others) u, from the pinned 1D matrices
    o = po.Sim(dim, [p] * dim, list(ne))
    M, K = [], []
    for d in range(dim):
        n = o.shape[d]
        M.append(_dense(o.matrix(d, 0)[0], n, p))
        o.set_matrix(d, 1, 0.0, 1.0)
        K.append(_dense(o.matrix(d, 1)[0], n, p))

    def kron(mats):  # first index fastest: kron(A_z, kron(A_y, A_x))
        out = mats[0]
        for m in mats[1:]:
            out = np.kron(m, out)
        return out

    Mt = kron(M)
    Kt = sum(kron([K[d] if d == e else M[d] for d in range(dim)]) for e in range(dim))
    u = np.random.default_rng(70 + dim).standard_normal(o.size)
    l2, h10, h1 = (o.norm_sq(u, k) for k in range(3))
    assert abs(l2 - u @ Mt @ u) <= 1e-12 * abs(l2)
    assert abs(h10 - u @ Kt @ u) <= 1e-12 * abs(h10)
    assert abs(h1 - (l2 + h10)) <= 1e-12 * h1
    assert abs(l2 - o.l2norm_sq(u)) <= 1e-13 * l2


def test_sample_loop_partitions_the_sequential_loop():
    """compute_rhs_sample over consecutive element ranges replays the sequential loop's
    additions in the same order, so the pieces give the full assembly bit for bit (the
    single-core CPU baseline times such a sample)."""
    sim = po.Sim(3, [2, 2, 2], [6, 5, 4])
    sim.prepare_matrices()
    u = np.random.default_rng(11).standard_normal(sim.size)
    form = po.heat_form(1e-3, 3)
    full = sim.compute_rhs(u, form)
    E = 6 * 5 * 4
    rhs = sim.zeros()
    for first in range(0, E, 37):
        sim.compute_rhs_sample(u, form, first, 37, rhs)
    assert np.array_equal(rhs, full)
    assert np.array_equal(sim.compute_rhs_sample(u, form, 0, E), full)


# ------------------------------------------------------------------ general configurations
# dim_config (simulation/config.hpp:11-29) beyond the defaults: quad_order != p+1, repeated
# interior knots (bspline.cpp:26-43), a degree per direction.

GENERAL = [
    # dim, p, elements, quad_order, repeated_nodes
    (2, (2, 2), (6, 5), (5, 4), (0, 0)),
    (2, (3, 1), (5, 7), (0, 0), (0, 0)),
    (3, (1, 2, 3), (4, 3, 5), (0, 0, 0), (0, 0, 0)),
    (2, (2, 3), (5, 4), (0, 0), (1, 1)),
    (3, (2, 2, 2), (4, 3, 4), (4, 3, 5), (1, 0, 2)),
]


def _gsim(dim, p, ne, q, r):
    return po.Sim(dim, list(p), list(ne), quad_order=list(q), repeated_nodes=list(r))


@pytest.mark.parametrize("dim,p,ne,q,r", GENERAL)
def test_general_config_1d_matrices_closed_forms(dim, p, ne, q, r):
    """Gram and stiffness matrices of each direction: with quad_order >= p+1 the Gauss rule is
    exact, so they equal the default rule's matrices; column sums of the Gram matrix are the
    integrals of the B-splines, (t_{j+p+1} - t_j) / (p+1), also with repeated knots; the
    stiffness rows sum to zero (the derivative of the partition of unity)."""
    sim = _gsim(dim, p, ne, q, r)
    ref = po.Sim(dim, list(p), list(ne), repeated_nodes=list(r))
    Ms, Ks = _mats(sim)
    Mr, Kr = _mats(ref)
    for d in range(dim):
        assert np.abs(Ms[d] - Mr[d]).max() < 1e-15 * max(1.0, np.abs(Mr[d]).max()) * 16
        assert np.abs(Ks[d] - Kr[d]).max() < 1e-12 * np.abs(Kr[d]).max()
        k = po.knots(0.0, 1.0, p[d], ne[d], r[d])
        n = sim.shape[d]
        integrals = np.array([(k[j + p[d] + 1] - k[j]) / (p[d] + 1) for j in range(n)])
        assert np.allclose(Ms[d].sum(axis=0), integrals, rtol=0, atol=1e-15)
        assert np.abs(Ks[d].sum(axis=1)).max() < 1e-11 * np.abs(Ks[d]).max()


@pytest.mark.parametrize("dim,p,ne,q,r", GENERAL)
def test_general_config_heat_rhs_equals_kronecker_form(dim, p, ne, q, r):
    sim = _gsim(dim, p, ne, q, r)
    rng = np.random.default_rng(77 + dim)
    c = rng.standard_normal(sim.size)
    dt = 1e-3
    Ms, Ks = _mats(sim)
    expect = _apply(Ms, c, sim.shape)
    for d in range(dim):
        mats = list(Ms)
        mats[d] = Ks[d]
        expect = expect - dt * _apply(mats, c, sim.shape)
    got = sim.compute_rhs(c, po.heat_form(dt, dim))
    assert po.rel_l2(got, expect) < 2e-14


def _general_quadrature_matrices(p, ne, nq, r):
    """Global basis values / derivatives at the nq Gauss points of every element of one
    direction (elements between distinct knots), from point evaluation of the basis."""
    k = po.knots(0.0, 1.0, p, ne, r)
    n = len(k) - p - 1
    X, W = po.gauss(nq)
    pts = np.unique(k)
    rows_b, rows_d, wts = [], [], []
    for e in range(ne):
        a, b = pts[e], pts[e + 1]
        for qq in range(nq):
            x = a + (b - a) * (X[qq] + 1) / 2
            s = po.find_span(x, k, p)
            v = po.eval_basis_with_derivatives(s, x, k, p, 1)
            rb, rd = np.zeros(n), np.zeros(n)
            rb[s - p: s + 1] = v[0]
            rd[s - p: s + 1] = v[1]
            rows_b.append(rb)
            rows_d.append(rd)
            wts.append(W[qq] * (b - a) / 2)
    return np.array(rows_b), np.array(rows_d), np.array(wts)


@pytest.mark.parametrize("p,ne,nq,r", [((3, 2), (5, 4), (2, 3), (0, 0)), ((2, 3), (4, 5), (2, 6), (1, 0)),
                                       ((1, 3), (6, 3), (4, 3), (0, 2))])
def test_general_config_forms_match_quadrature_matrix_form(p, ne, nq, r):
    """2D, every form, against global quadrature matrices built by point evaluation: covers
    quad_order < p+1 (under-integration, where no exact form exists), repeated knots and a
    degree per direction."""
    sim = po.Sim(2, list(p), list(ne), quad_order=list(nq), repeated_nodes=list(r))
    Bx, Dx, wx = _general_quadrature_matrices(p[0], ne[0], nq[0], r[0])
    By, Dy, wy = _general_quadrature_matrices(p[1], ne[1], nq[1], r[1])
    B, DX, DY, w = np.kron(By, Bx), np.kron(By, Dx), np.kron(Dy, Bx), np.kron(wy, wx)
    rng = np.random.default_rng(5)
    c = rng.uniform(0, 1, sim.size)
    kf = np.exp(rng.uniform(-0.5, 0.5, sim.size))
    src = rng.standard_normal(sim.size)
    dt = 3e-3
    for kw in (dict(), dict(kfield=kf, reaction=0.7, source=src)):
        u, ux, uy = B @ c, DX @ c, DY @ c
        kv = B @ kw["kfield"] if "kfield" in kw else np.ones_like(u)
        fv = B @ kw["source"] if "source" in kw else np.zeros_like(u)
        val = u + dt * kw.get("reaction", 0.0) * u * (1 - u) + dt * fv
        expect = B.T @ (w * val) - dt * (DX.T @ (w * kv * ux) + DY.T @ (w * kv * uy))
        got = sim.compute_rhs(c, po.heat_form(dt, 2, **kw))
        assert po.rel_l2(got, expect) < 1e-13, kw.keys()
