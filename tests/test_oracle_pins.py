"""Pins of the CPU oracle (oracle/ads_oracle.cpp) before it is trusted as the checker.

Each oracle function is checked against (i) the reference's own known-answer tests,
(ii) the reference's own sources compiled unmodified (oracle/_ref, where the reference
tree is present), (iii) an independent LAPACK (scipy) and (iv) what mathematics fixes
(closed forms, exactness, invariants, brute-force Kronecker solves).
"""
import numpy as np
import pytest
import scipy.linalg
import scipy.linalg.lapack as lapack

import pyoracle as po

REF = po.ref_lib()
needs_ref = pytest.mark.skipif(REF is None, reason="reference build oracle/_ref not present")


# ---------------------------------------------------------------- quadrature
@pytest.mark.parametrize("n", list(range(2, 65)))
def test_gauss_matches_numpy_and_integrates_monomials(n):
    X, W = po.gauss(n)
    xn, wn = np.polynomial.legendre.leggauss(n)
    assert np.max(np.abs(X - xn)) <= 4.5e-15
    assert np.max(np.abs(W - wn)) <= 4.5e-15
    assert np.all(np.diff(X) > 0)
    # exact for degree <= 2n-1 (checked up to degree 12 where cancellation stays benign)
    for k in range(0, min(2 * n - 1, 12) + 1):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs(np.dot(W, X ** k) - exact) < 1e-14


@needs_ref
def test_gauss_matches_reference_tables():
    worst = 0.0
    for n in range(2, 65):
        X, W = po.gauss(n)
        Xr, Wr = np.zeros(n), np.zeros(n)
        assert REF.ref_gauss(n, po._d(Xr), po._d(Wr)) == 0
        worst = max(worst, np.max(np.abs(X - Xr)), np.max(np.abs(W - Wr)))
        if n <= 7:
            assert np.array_equal(X, Xr) and np.array_equal(W, Wr)
    assert worst <= 2.3e-16


# ---------------------------------------------------------------- B-splines
def test_knot_vector_known_answer():
    # bspline_test.cpp:14-16
    assert list(po.knots(0.0, 1.0, 2, 4)) == [0, 0, 0, 0.25, 0.5, 0.75, 1, 1, 1]


def test_find_span_known_answers():
    # bspline_test.cpp:18-37
    k = po.knots(0.0, 1.0, 2, 4)
    cases = {-1: 2, 2.0: 5, 0.0: 2, 1.0: 5, 0.25: 3, 0.75: 5, 0.1: 2, 0.3: 3, 0.7: 4, 0.9: 5}
    for x, s in cases.items():
        assert po.find_span(x, k, 2) == s
    # bspline_test.cpp:39-63 -- repeated nodes, p = 3
    kr = np.array([0, 0, 0, 0, 1, 1, 2, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5], dtype=float)
    cases = {-4: 3, 8: 13, 0: 3, 5: 13, 1: 5, 2: 6, 3: 9, 4: 13, 0.3: 3, 1.2: 5, 2.8: 6,
             3.1: 9, 4.9: 13}
    for x, s in cases.items():
        assert po.find_span(x, kr, 3) == s


def test_first_nonzero_dofs_known_answer():
    # bspline_test.cpp:65-67
    k = po.knots(0.0, 1.0, 2, 4)
    out = np.zeros(4, dtype=np.int32)
    assert po.lib().orc_first_nonzero_dofs(po._d(k), len(k), 2, po._i(out)) == 4
    assert list(out) == [0, 1, 2, 3]


@pytest.mark.parametrize("p", [1, 2, 3])
def test_partition_of_unity_and_derivative_sums(p):
    # eval_test.cpp:14-79, p=2 with 5 elements there; all p here
    k = po.knots(0.0, 1.0, p, 5)
    for x in np.linspace(0, 1, 101):
        s = po.find_span(x, k, p)
        v = po.eval_basis_with_derivatives(s, x, k, p, min(2, p))
        assert abs(v[0].sum() - 1.0) < 1e-14
        for d in range(1, v.shape[0]):
            assert abs(v[d].sum()) < 1e-10
        assert np.allclose(po.eval_basis(s, x, k, p), v[0], rtol=0, atol=1e-15)


@pytest.mark.parametrize("p", [1, 2, 3])
def test_derivative_is_derivative_of_value(p):
    k = po.knots(0.0, 1.0, p, 7)
    h = 1e-6
    for x in np.linspace(0.013, 0.987, 37):
        s = po.find_span(x, k, p)
        d = po.eval_basis_with_derivatives(s, x, k, p, 1)[1]
        fd = (po.eval_basis(s, x + h, k, p) - po.eval_basis(s, x - h, k, p)) / (2 * h)
        if po.find_span(x + h, k, p) == s == po.find_span(x - h, k, p):
            assert np.allclose(d, fd, atol=1e-6 * (1 + np.abs(d).max()))


@needs_ref
@pytest.mark.parametrize("p,ne,rn", [(1, 4, 0), (2, 4, 0), (3, 5, 0), (2, 64, 0), (3, 6, 1), (2, 7, 2)])
def test_bspline_bitwise_vs_reference(p, ne, rn):
    k = po.knots(0.0, 1.0, p, ne, rn)
    kr = np.zeros(len(k))
    assert REF.ref_create_basis(0.0, 1.0, p, ne, rn, po._d(kr)) == len(k)
    assert np.array_equal(k, kr)
    for x in np.linspace(-0.1, 1.1, 301):
        s = po.find_span(x, k, p)
        assert s == REF.ref_find_span(x, po._d(kr), len(kr), p)
        o = po.eval_basis_with_derivatives(s, x, k, p, 1)
        r = np.zeros((2, p + 1))
        REF.ref_eval_basis_with_derivatives(s, x, po._d(kr), len(kr), p, 1, po._d(r))
        assert np.array_equal(o, r)
    f1 = np.zeros(len(k), dtype=np.int32)
    f2 = np.zeros(len(k), dtype=np.int32)
    n1 = po.lib().orc_first_nonzero_dofs(po._d(k), len(k), p, po._i(f1))
    n2 = REF.ref_first_nonzero_dofs(po._d(kr), len(kr), p, po._i(f2))
    assert n1 == n2 and np.array_equal(f1[:n1], f2[:n2])
    a1, b1, a2, b2 = (np.zeros(len(k), dtype=np.int32) for _ in range(4))
    m1 = po.lib().orc_elements_supporting_dofs(po._d(k), len(k), p, po._i(a1), po._i(b1))
    m2 = REF.ref_elements_supporting_dofs(po._d(kr), len(kr), p, po._i(a2), po._i(b2))
    assert m1 == m2 and np.array_equal(a1, a2) and np.array_equal(b1, b2)


# ---------------------------------------------------------------- 1D matrices
CLOSED = {
    # interior rows of the 1D Gram / stiffness matrices on a uniform mesh, h = 1/ne
    1: ([1, 4, 1], 6.0, [-1, 2, -1], 1.0),
    2: ([1, 26, 66, 26, 1], 120.0, [-1, -2, 6, -2, -1], 6.0),
    3: ([1, 120, 1191, 2416, 1191, 120, 1], 5040.0, [-1, -24, -15, 80, -15, -24, -1], 120.0),
}


def _dense_from_band(ab, n, kl, ku):
    A = np.zeros((n, n))
    for i in range(n):
        for j in range(max(0, i - kl), min(n, i + ku + 1)):
            A[i, j] = ab[j, kl + ku + i - j]
    return A


@pytest.mark.parametrize("p", [1, 2, 3])
def test_gram_and_stiffness_closed_forms(p):
    ne = 12
    s = po.Sim(1, [p], [ne])
    h = 1.0 / ne
    ab, _ = s.matrix(0)
    M = _dense_from_band(ab, s.shape[0], p, p)
    s.set_matrix(0, 1, 0.0, 1.0)
    abk, _ = s.matrix(0, 1)
    K = _dense_from_band(abk, s.shape[0], p, p)
    mrow, mden, krow, kden = CLOSED[p]
    mid = s.shape[0] // 2
    assert np.allclose(M[mid, mid - p: mid + p + 1], np.array(mrow) * h / mden, rtol=1e-13)
    assert np.allclose(K[mid, mid - p: mid + p + 1], np.array(krow) / (kden * h), rtol=1e-12)
    # integrals of the basis sum to 1, constants are in the kernel of K, both symmetric
    assert abs(M.sum() - 1.0) < 1e-14
    assert np.max(np.abs(K.sum(axis=1))) < 1e-10
    assert np.allclose(M, M.T, atol=1e-17) and np.allclose(K, K.T, atol=1e-12)


def test_fix_dof_row_identity():
    s = po.Sim(1, [2], [6])
    ab0, _ = s.matrix(0)
    s.fix_dof(0, 0)
    s.fix_dof(0, s.shape[0] - 1)
    ab, _ = s.matrix(0)
    M0 = _dense_from_band(ab0, s.shape[0], 2, 2)
    M = _dense_from_band(ab, s.shape[0], 2, 2)
    e0 = np.zeros(s.shape[0])
    e0[0] = 1
    assert np.array_equal(M[0], e0) and np.array_equal(M[-1], e0[::-1])
    # dimension.cpp:23-29 clears the row only; the column keeps its entries
    assert np.array_equal(M[1:-1, 0], M0[1:-1, 0])


# ---------------------------------------------------------------- band LU
def test_band_solve_known_answer():
    # band_solve_test.cpp:15-46 (kl=1, ku=2, n=6, 4 right-hand sides, 1e-5)
    kl, ku, n, d = 1, 2, 6, 4
    ab = np.zeros((n, 2 * kl + ku + 1))
    for i in range(n):
        for j in range(max(0, i - kl), min(n, i + ku + 1)):
            ab[j, kl + ku + i - j] = (i + 1) * 10 + j + 1
    b = np.array([[(j + 1) * (i + 1) for i in range(n)] for j in range(d)], dtype=float)
    ip, info = po.dgbtrf(ab, n, kl, ku)
    assert info == 0
    po.dgbtrs(ab, ip, b, n, kl, ku)
    sol = np.array([0.230377, -0.126052, -0.0016554, -0.00111222, 0.203603, -0.109609])
    assert np.allclose(b, np.outer(np.arange(1, d + 1), sol), atol=1e-5)


def test_band_multiply_known_answer():
    # band_solve_test.cpp:48-58: band_matrix{1,1,4,3} with row_offset 0, two vectors
    kl, ku, rows, cols = 1, 1, 4, 3
    lda = kl + ku + 1
    a = np.zeros((cols, lda))
    for (i, j) in [(0, 0), (1, 1), (2, 2), (3, 2)]:
        a[j, ku + i - j] = 1
    x = np.array([1, 2, 3, 4, 5, 6], dtype=float)
    y = np.zeros(8)
    po.lib().orc_band_multiply(po._d(a), rows, cols, kl, ku, lda, po._d(x), po._d(y), 2, b"N")
    assert list(y) == [1, 2, 3, 3, 4, 5, 6, 6]


def _random_band(rng, n, kl, ku):
    ab = np.zeros((n, 2 * kl + ku + 1))
    for i in range(n):
        for j in range(max(0, i - kl), min(n, i + ku + 1)):
            ab[j, kl + ku + i - j] = rng.standard_normal()
    return ab


def test_dgbtrf_dgbtrs_match_scipy_lapack():
    rng = np.random.default_rng(7)
    for _ in range(150):
        n = int(rng.integers(1, 30))
        kl, ku = int(rng.integers(0, 4)), int(rng.integers(0, 4))
        ab = _random_band(rng, n, kl, ku)
        A = _dense_from_band(ab, n, kl, ku)
        if np.linalg.cond(A) > 1e6:
            continue
        mine = ab.copy()
        ip, info = po.dgbtrf(mine, n, kl, ku)
        # scipy's dgbtrf takes the (2kl+ku+1, n) Fortran band array
        lu, piv, sinfo = lapack.dgbtrf(np.asfortranarray(ab.T), kl, ku)
        assert info == sinfo
        if info:
            continue
        assert np.array_equal(ip, piv + 1)
        assert np.allclose(mine.T, lu, rtol=1e-12, atol=1e-12)
        b = rng.standard_normal((3, n))
        for tr, op in (("N", A), ("T", A.T)):
            x = b.copy()
            po.dgbtrs(mine, ip, x, n, kl, ku, tr)
            xs, _ = lapack.dgbtrs(lu, kl, ku, np.asfortranarray(b.T), piv, trans=0 if tr == "N" else 1)
            assert np.allclose(x.T, xs, rtol=1e-10, atol=1e-10)
            assert np.allclose(op @ x.T, b.T, atol=1e-8 * (1 + np.abs(x).max()))


def test_dgbtrf_reports_zero_pivot():
    n, kl, ku = 4, 1, 1
    ab = np.zeros((n, 2 * kl + ku + 1))
    for i in range(n):
        ab[i, kl + ku] = 1.0
    ab[2, kl + ku] = 0.0  # A(2,2) = 0, nothing below it in column 2 either
    ip, info = po.dgbtrf(ab, n, kl, ku)
    assert info == 3


@needs_ref
def test_dgbtrf_dgbtrs_match_reference_build():
    rng = np.random.default_rng(11)
    for _ in range(100):
        n = int(rng.integers(1, 40))
        kl, ku = int(rng.integers(0, 4)), int(rng.integers(0, 4))
        ab = _random_band(rng, n, kl, ku)
        mine = ab.copy()
        ip, info = po.dgbtrf(mine, n, kl, ku)
        ref = ab.copy()
        ipr = np.zeros(n, dtype=np.int32)
        assert REF.ref_factorize(n, kl, ku, po._d(ref), po._i(ipr)) == info
        assert np.array_equal(ip, ipr)
        assert np.max(np.abs(mine - ref)) <= 1e-13 * max(1.0, np.max(np.abs(ref)))
        if info or np.linalg.cond(_dense_from_band(ab, n, kl, ku)) > 1e6:
            continue
        b = rng.standard_normal((2, n))
        x1, x2 = b.copy(), b.copy()
        po.dgbtrs(mine, ip, x1, n, kl, ku)
        REF.ref_solve_with_factorized(n, kl, ku, po._d(ref), po._i(ipr), po._d(x2), 2)
        assert np.max(np.abs(x1 - x2)) <= 1e-11 * max(1.0, np.max(np.abs(x2)))


# ---------------------------------------------------------------- tensors + ADS
def test_cyclic_transpose_known_answer_and_round_trip():
    # tensor_test.cpp:49-80
    k, n, m = 2, 3, 2
    a = np.zeros(k * n * m)
    e = np.zeros(k * n * m)

    def A(i, j, l):
        return i + k * (j + n * l)

    def E(i, j, l):
        return i + n * (j + m * l)

    for i in range(k):
        for j in range(n):
            for l in range(m):
                a[A(i, j, l)] = e[E(j, l, i)] = 100 * (i + 1) + 10 * (j + 1) + (l + 1)
    out = po.cyclic_transpose([k, n, m], a)
    assert np.array_equal(out, e)
    a2 = po.cyclic_transpose([n, m, k], out)
    a3 = po.cyclic_transpose([m, k, n], a2)
    assert np.array_equal(a3, a)


def _band_from_rows(kl, ku, n, entries):
    ab = np.zeros((n, 2 * kl + ku + 1))
    for (i, j), v in entries.items():
        ab[j, kl + ku + i - j] = v
    return ab


MX = {(0, 0): 1, (0, 1): 2, (1, 0): 2, (1, 1): 3, (1, 2): 1, (2, 1): -1, (2, 2): 4, (3, 2): 1, (3, 3): 3}
MY = {(0, 0): 2, (0, 1): 3, (1, 0): 1, (1, 1): 2, (1, 2): -3, (2, 1): 2, (2, 2): 2}
MZ = {(0, 0): 1, (0, 1): 2, (1, 0): 3, (1, 1): 4}


def _ads(rank, sizes, rhs, mats):
    abs_, ips = [], []
    for (n, ent) in zip(sizes, mats):
        ab = _band_from_rows(1, 1, n, ent)
        ip, info = po.dgbtrf(ab, n, 1, 1)
        assert info == 0
        abs_.append(ab)
        ips.append(ip)
    import ctypes as C

    abp = (po._dp * rank)(*[po._d(x) for x in abs_])
    ipp = (po._ip * rank)(*[po._i(x) for x in ips])
    kl = np.ones(rank, dtype=np.int32)
    s = np.array(sizes, dtype=np.int32)
    buf = np.zeros_like(rhs)
    po.lib().orc_ads_solve.argtypes = [C.c_int, po._ip, po._dp, po._dp, C.POINTER(po._dp),
                                       C.POINTER(po._ip), po._ip, po._ip]
    po.lib().orc_ads_solve(rank, po._i(s), po._d(rhs), po._d(buf), abp, ipp, po._i(kl), po._i(kl))
    return rhs, abs_, ips


def test_ads_known_answers():
    # solver_test.cpp:101-162 (standard 2D) and :164-255 (standard 3D)
    rhs2 = np.array([61, 127, 86, 123, -48, -96, -48, -64, 92, 188, 112, 156], dtype=float)
    out, _, _ = _ads(2, [4, 3], rhs2, [MX, MY])
    assert np.allclose(out, np.arange(1, 13), rtol=1e-12)
    rhs3 = np.array([543, 1101, 618, 849, -144, -288, -144, -192, 564, 1140, 624, 852,
                     1147, 2329, 1322, 1821, -336, -672, -336, -448, 1220, 2468, 1360, 1860],
                    dtype=float)
    out, _, _ = _ads(3, [4, 3, 2], rhs3, [MX, MY, MZ])
    assert np.allclose(out, np.arange(1, 25), rtol=1e-12)


def test_ads_one_dimension_known_answer():
    # solver_test.cpp:63-99
    ab = _band_from_rows(1, 1, 4, MX)
    ip, _ = po.dgbtrf(ab, 4, 1, 1)
    b = np.array([[5, 11, 10, 15]], dtype=float)
    po.dgbtrs(ab, ip, b, 4, 1, 1)
    assert np.allclose(b[0], [1, 2, 3, 4], rtol=1e-12)


def test_ads_equals_kronecker_dense_solve():
    rng = np.random.default_rng(3)
    for sizes in ([7, 9], [5, 6, 7], [4, 1, 3]):
        mats, dense = [], []
        for n in sizes:
            ent = {}
            for i in range(n):
                for j in range(max(0, i - 2), min(n, i + 3)):
                    ent[(i, j)] = rng.standard_normal() + (6.0 if i == j else 0.0)
            mats.append(ent)
            dense.append(_dense_from_band(_band_from_rows(1, 1, n, {k: v for k, v in ent.items() if abs(k[0] - k[1]) <= 1}), n, 1, 1))
        mats = [{k: v for k, v in ent.items() if abs(k[0] - k[1]) <= 1} for ent in mats]
        K = dense[0]
        for D in dense[1:]:
            K = np.kron(D, K)  # first index fastest: u = (A_z (x) A_y (x) A_x) vec
        b = rng.standard_normal(int(np.prod(sizes)))
        x, _, _ = _ads(len(sizes), sizes, b.copy(), mats)
        assert np.allclose(x, np.linalg.solve(K, b), rtol=1e-11, atol=1e-12)


@needs_ref
def test_ads_matches_reference_build():
    import ctypes as C

    rng = np.random.default_rng(5)
    for sizes in ([7, 9], [5, 6, 7], [12, 3, 4]):
        rank = len(sizes)
        mats = []
        for n in sizes:
            ent = {}
            for i in range(n):
                for j in range(max(0, i - 1), min(n, i + 2)):
                    ent[(i, j)] = rng.standard_normal()
            mats.append(ent)
        b = rng.standard_normal(int(np.prod(sizes)))
        x, abs_, ips = _ads(rank, sizes, b.copy(), mats)
        abp = (po._dp * rank)(*[po._d(a) for a in abs_])
        ipp = (po._ip * rank)(*[po._i(i) for i in ips])
        kl = np.ones(rank, dtype=np.int32)
        s = np.array(sizes, dtype=np.int32)
        y = b.copy()
        REF.ref_ads_solve(rank, po._i(s), po._d(y), abp, ipp, po._i(kl), po._i(kl))
        assert np.allclose(x, y, rtol=1e-10, atol=1e-12)
        for _ in range(1):
            t = rng.standard_normal(int(np.prod(sizes)))
            o1 = po.cyclic_transpose(sizes, t)
            o2 = np.zeros_like(t)
            REF.ref_cyclic_transpose(rank, po._i(s), po._d(t), po._d(o2))
            assert np.array_equal(o1, o2)


@pytest.mark.parametrize("dim", [2, 3])
def test_sim_transposed_solve_equals_dense_transposed_solve(dim):
    # The adjoint needs (A_x (x) A_y (x) A_z)^-T: pinned against a dense solve with the
    # transposed Kronecker matrix, on non-symmetric matrices (a fixed dof breaks symmetry)
    # and on the implicit M + c K slot.
    p, ne = [2, 3, 2][:dim], [4, 3, 5][:dim]
    o = po.Sim(dim, p, ne)
    for d in range(dim):
        o.fix_dof(d, 0 if d % 2 == 0 else o.shape[d] - 1)
        o.set_matrix(d, 1, 1.0, 0.37)
    mats = {(d, sl): _dense_from_band(o.matrix(d, sl)[0], o.shape[d], o.p[d], o.p[d])
            for d in range(dim) for sl in (0, 1)}
    assert o.prepare_matrices() == 0
    for slots in ([0] * dim, [1] * dim, [1, 0, 1][:dim]):
        K = mats[(0, slots[0])]
        for d in range(1, dim):
            K = np.kron(mats[(d, slots[d])], K)
        if 0 in slots:  # slot 1 (unfixed M + c K) is symmetric, slot 0 is not
            assert not np.allclose(K, K.T)
        rng = np.random.default_rng(11 + dim)
        b = rng.standard_normal(o.size)
        x = b.copy()
        o.solve(x, slots=slots + [0] * (3 - dim), transposed=True)
        # cond(K) reaches 1e5 here: compare in relative L2
        assert po.rel_l2(x, np.linalg.solve(K.T, b)) < 1e-11
        y = b.copy()
        o.solve(y, slots=slots + [0] * (3 - dim))
        assert po.rel_l2(y, np.linalg.solve(K, b)) < 1e-11
        if 0 in slots:
            assert po.rel_l2(x, y) > 1e-3  # the transpose is not the plain solve here
