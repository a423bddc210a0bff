"""GPU parity of the generalized ADS (one band matrix per line of a special direction,
solver.hpp:43-195): the reference's own known answers (solver_test.cpp:257-534, integer
solutions 1..N) and random per-line matrices against a dense numpy restatement (per-line
solves along the special direction first, then the Kronecker solve of the others)."""
import numpy as np
import pytest

import adsgpu as ag

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)


def _band_rows(n, rows, kl=1, ku=1):
    """LAPACK full buffer (n, 2kl+ku+1) from rows of band entries starting at column i-kl
    (make_matrix in solver_test.cpp:31-40)."""
    ab = np.zeros((n, 2 * kl + ku + 1))
    for i, r in enumerate(rows):
        j = max(i - kl, 0)
        for v in r:
            ab[j, kl + ku + i - j] = v
            j += 1
    return ab


A2 = [[1, 2], [3, 4]]                      # 2x2: rows as full rows
B1 = [[2, 3], [1, 2, -3], [2, 2]]
B2 = [[1, 2], [3, 1, 2], [2, -1]]
A3 = [[2, 3], [1, 2, -3], [2, 2]]
B3 = [[1, 2], [3, 4]]
C = {  # C{a}{b}
    (1, 1): [[1, 2], [2, 3, 1], [1, -4, 0], [1, 3]],
    (2, 1): [[1, 3], [2, -2, 1], [3, -2, 1], [1, 1]],
    (3, 1): [[3, 0], [1, 2, 1], [2, -1, 3], [3, 1]],
    (1, 2): [[2, 3], [-1, 3, 2], [2, 4, 1], [-2, 1]],
    (2, 2): [[-1, -2], [3, 2, 1], [2, -1, 1], [1, 1]],
    (3, 2): [[2, 1], [-1, 3, 2], [4, 2, 1], [-3, 2]],
}
C_LINES = [C[(1, 1)], C[(2, 1)], C[(3, 1)], C[(1, 2)], C[(2, 2)], C[(3, 2)]]


def _sim(sizes):
    """A handle whose directions have the given dof counts (p = 1, ne = n - 1)."""
    return ag.Simulation(len(sizes), [1] * len(sizes), [n - 1 for n in sizes])


def _run(sizes, std, special, lines, rhs):
    g = _sim(sizes)
    slots = [0] * len(sizes)
    for d, rows in std.items():
        g.set_band_matrix(d, 0, 1, 1, _band_rows(sizes[d], rows))
    g.set_line_matrices(special, 1, 1, 1, np.stack([_band_rows(sizes[special], r) for r in lines]))
    slots[special] = 1
    assert g.prepare_matrices() == 0
    g.set(ag.U, np.array(rhs, dtype=float))
    g.ads_solve(slots)
    out = g.get(ag.U)
    g.close()
    return out


def test_generalized_2d_known_answers():
    _log("case: generalized ADS 2D known answers (solver_test.cpp:257-346)")
    expect = np.arange(1, 7, dtype=float)
    # special dim second: rhs (na=2, nb=3), rhs(i, j) at i + 2 j
    rhs = [43, 61, -24, 136, 56, 11]
    assert np.allclose(_run([2, 3], {0: A2}, 1, [B1, B2], rhs), expect, rtol=1e-12)
    # special dim first: rhs (nb=3, na=2)
    rhs = [54, -12, 54, 71, 149, 19]
    assert np.allclose(_run([3, 2], {1: A2}, 0, [B1, B2], rhs), expect, rtol=1e-12)


def _t3(shape, entries):
    out = np.zeros(int(np.prod(shape)))
    for (i, j, k), v in entries.items():
        out[i + shape[0] * (j + shape[1] * k)] = v
    return out


def test_generalized_3d_known_answers():
    _log("case: generalized ADS 3D known answers (solver_test.cpp:348-534)")
    expect = np.arange(1, 25, dtype=float)
    first = {(0, 0, 0): 543, (1, 0, 0): 1101, (2, 0, 0): -618, (3, 0, 0): 849,
             (0, 1, 0): -192, (1, 1, 0): -48, (2, 1, 0): -96, (3, 1, 0): -96,
             (0, 2, 0): 540, (1, 2, 0): 768, (2, 2, 0): 828, (3, 2, 0): 828,
             (0, 0, 1): 1900, (1, 0, 1): 1681, (2, 0, 1): 2968, (3, 0, 1): -394,
             (0, 1, 1): 336, (1, 1, 1): -672, (2, 1, 1): -224, (3, 1, 1): -224,
             (0, 2, 1): 1192, (1, 2, 1): 1748, (2, 2, 1): 3024, (3, 2, 1): -388}
    out = _run([4, 3, 2], {1: A3, 2: B3}, 0, C_LINES, _t3((4, 3, 2), first))
    assert np.allclose(out, expect, rtol=1e-12)
    second = {(0, 0, 0): 351, (1, 0, 0): 1325, (0, 1, 0): 732, (1, 1, 0): 1382,
              (0, 2, 0): -501, (1, 2, 0): 2471, (0, 3, 0): 718, (1, 3, 0): -293,
              (0, 0, 1): -384, (1, 0, 1): 672, (0, 1, 1): -96, (1, 1, 1): -1344,
              (0, 2, 1): -192, (1, 2, 1): -448, (0, 3, 1): -192, (1, 3, 1): -448,
              (0, 0, 2): 492, (1, 0, 2): 1196, (0, 1, 2): 752, (1, 1, 2): 1912,
              (0, 2, 2): 872, (1, 2, 2): 3276, (0, 3, 2): 872, (1, 3, 2): -380}
    out = _run([2, 4, 3], {0: B3, 2: A3}, 1, C_LINES, _t3((2, 4, 3), second))
    assert np.allclose(out, expect, rtol=1e-12)
    third = {(0, 0, 0): 342, (1, 0, 0): -48, (2, 0, 0): 162, (0, 1, 0): 1210, (1, 1, 0): 84,
             (2, 1, 0): 522, (0, 0, 1): 774, (1, 0, 1): -12, (2, 0, 1): 504, (0, 1, 1): 1934,
             (1, 1, 1): -168, (2, 1, 1): 1648, (0, 0, 2): -792, (1, 0, 2): -24, (2, 0, 2): 864,
             (0, 1, 2): 3542, (1, 1, 2): -56, (2, 1, 2): 2674, (0, 0, 3): 1206, (1, 0, 3): -24,
             (2, 0, 3): 864, (0, 1, 3): -326, (1, 1, 3): -56, (2, 1, 3): -118}
    out = _run([3, 2, 4], {0: A3, 1: B3}, 2, C_LINES, _t3((3, 2, 4), third))
    assert np.allclose(out, expect, rtol=1e-12)


def _dense(ab, n, kl, ku):
    A = np.zeros((n, n))
    for i in range(n):
        for j in range(max(0, i - kl), min(n, i + ku + 1)):
            A[i, j] = ab[j, kl + ku + i - j]
    return A


@pytest.mark.parametrize("special", [0, 1, 2])
def test_generalized_3d_random_against_dense(special):
    _log(f"case: generalized ADS random 3D, special direction {special}")
    p, ne = 2, (9, 7, 8)
    g = ag.Simulation(3, [p] * 3, list(ne))
    shape = g.shape
    rng = np.random.default_rng(300 + special)
    n = shape[special]
    others = [d for d in range(3) if d != special]
    lines = int(np.prod([shape[d] for d in others]))
    kl, ku = 2, 1
    abs_ = np.zeros((lines, n, 2 * kl + ku + 1))
    for l in range(lines):
        for i in range(n):
            for j in range(max(0, i - kl), min(n, i + ku + 1)):
                abs_[l, j, kl + ku + i - j] = rng.standard_normal() + (5.0 if i == j else 0.0)
    g.set_line_matrices(special, 1, kl, ku, abs_)
    slots = [0, 0, 0]
    slots[special] = 1
    assert g.prepare_matrices() == 0
    b = rng.standard_normal(g.size)
    g.set(ag.U, b)
    g.ads_solve(slots)
    x = g.get(ag.U)
    # dense restatement: per-line solves along `special` (lines in cyclic order of the other
    # two indices, special+1 fastest), then the Gram-matrix solves of the other directions
    t = b.reshape(shape[2], shape[1], shape[0]).transpose(2, 1, 0).copy()  # t[i0, i1, i2]
    d1, d2 = (special + 1) % 3, (special + 2) % 3
    for i2 in range(shape[d2]):
        for i1 in range(shape[d1]):
            l = i1 + shape[d1] * i2
            idx = [slice(None)] * 3
            idx[d1], idx[d2] = i1, i2
            A = _dense(abs_[l], n, kl, ku)
            t[tuple(idx)] = np.linalg.solve(A, t[tuple(idx)])
    import pyoracle as po

    o = po.Sim(3, [p] * 3, list(ne))
    for d in others:
        M = _dense(o.matrix(d, 0)[0], shape[d], p, p)
        t = np.moveaxis(np.tensordot(np.linalg.inv(M), np.moveaxis(t, d, 0), axes=1), 0, d)
    ref = t.transpose(2, 1, 0).reshape(-1)
    assert po.rel_l2(x, ref) <= 1e-11
    g.close()
