"""Grid sampling and output formats (no GPU): the oracle's sampler (orc_sim_sample) against
the reference's own bspline::eval / eval_ders compiled in oracle/_ref and against linear
precision of B-splines; the Python writers against the reference's gnuplot_printer<2>
(byte for byte), the VTK layout of output/vtk.hpp, and the coefficient-dump format of
implicit.hpp:121-127 / tools/error.cpp:16-30."""
import ctypes as C
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import pyoracle as po
from adsgpu import output as out
from adsgpu import linspace

REF = po.ref_lib()
needs_ref = pytest.mark.skipif(REF is None, reason="reference build oracle/_ref not present")


def _knots(p, ne, a=0.0, b=1.0):
    k = np.zeros(po.lib().orc_knot_size(p, ne, 0))
    po.lib().orc_create_basis(a, b, p, ne, 0, po._d(k))
    return k


def _points(rng, a, b, n):
    """random points, the end points and a few knots (span edges)"""
    pts = np.concatenate([rng.uniform(a, b, n), [a, b], linspace(a, b, 4)])
    return np.sort(pts)


@needs_ref
@pytest.mark.parametrize("dim,p,ne", [(2, 1, (5, 4)), (2, 2, (6, 7)), (2, 3, (5, 5)),
                                      (3, 1, (3, 4, 2)), (3, 2, (4, 3, 5)), (3, 3, (3, 3, 4))])
def test_sampler_matches_reference_eval(dim, p, ne):
    s = po.Sim(dim, [p] * dim, list(ne))
    rng = np.random.default_rng(9)
    u = rng.standard_normal(s.size)
    pts = [_points(rng, 0.0, 1.0, 3) for _ in range(dim)]
    got = s.sample(u, pts, ders=True)
    knots = [_knots(p, ne[d]) for d in range(dim)]
    kp = (po._dp * dim)(*[po._d(k) for k in knots])
    ks = np.array([k.size for k in knots], dtype=np.int32)
    ps = np.array([p] * dim, dtype=np.int32)
    counts = [len(q) for q in pts]
    P = int(np.prod(counts))
    worst = 0.0
    for idx in np.ndindex(*counts[::-1]):
        i = idx[::-1]
        pt = np.array([pts[d][i[d]] for d in range(dim)])
        r = np.zeros(1 + dim)
        REF.ref_eval(dim, po._d(pt), po._d(u), kp, po._i(ks), po._i(ps), 1, po._d(r))
        flat = int(np.ravel_multi_index(idx, counts[::-1]))
        for c in range(1 + dim):
            worst = max(worst, abs(got[c * P + flat] - r[c]))
    assert worst <= 1e-13


@pytest.mark.parametrize("dim,p", [(2, 1), (2, 2), (3, 3)])
def test_sampler_reproduces_linear_functions(dim, p):
    """Greville abscissae as coefficients reproduce x (linear precision), so sum over
    directions of the x_d-Greville spline equals x + y (+ z) with gradient (1, 1, 1)."""
    ne = [5, 4, 3][:dim]
    s = po.Sim(dim, [p] * dim, ne)
    g = []
    for d in range(dim):
        k = _knots(p, ne[d])
        xi = np.zeros(s.shape[d])
        po.lib().orc_greville(po._d(k), k.size, p, po._d(xi))
        g.append(xi)
    grids = np.meshgrid(*g, indexing="ij")
    u = np.sum(grids, axis=0).transpose().ravel()  # c_ij = xi_i + eta_j (x fastest)
    rng = np.random.default_rng(4)
    pts = [_points(rng, 0.0, 1.0, 5) for _ in range(dim)]
    got = s.sample(u, pts, ders=True).reshape(1 + dim, -1)
    X = [q.transpose().ravel() for q in np.meshgrid(*pts, indexing="ij")]
    assert np.max(np.abs(got[0] - np.sum(X, axis=0))) <= 1e-14
    for d in range(dim):
        assert np.max(np.abs(got[1 + d] - 1.0)) <= 1e-12


@needs_ref
def test_gnuplot_writer_matches_reference_printer():
    rng = np.random.default_rng(6)
    xs, ys = linspace(0.0, 1.0, 7), linspace(-1.0, 2.0, 5)
    vals = rng.standard_normal((ys.size, xs.size)) * np.array([1e-9, 1.0, 1e3, 1e7, -3.0, 0.5])[:, None]
    vals[0, 0] = -0.0
    text = out.gnuplot_2d(xs, ys, vals)
    flat = np.ascontiguousarray(vals).ravel()  # x fastest: vals[j, i] at index i + nx*j
    n = REF.ref_gnuplot2(po._d(xs), xs.size, po._d(ys), ys.size, po._d(flat), None, 0)
    buf = C.create_string_buffer(n)
    REF.ref_gnuplot2(po._d(xs), xs.size, po._d(ys), ys.size, po._d(flat), buf, n)
    assert text == buf.raw[:n].decode()


def test_vtk_layout():
    rng = np.random.default_rng(7)
    a, b = rng.standard_normal((3, 4, 5)), rng.standard_normal((3, 4, 5))
    text = out.vtk_3d(a, b)
    root = ET.fromstring(text)
    img = root.find("ImageData")
    assert img.get("WholeExtent") == "0 4 0 3 0 2" and img.get("origin") == "0 0 0"
    arr = img.find("Piece/PointData/DataArray")
    assert arr.get("NumberOfComponents") == "2" and arr.get("type") == "Float32"
    rows = np.array([[float(v) for v in r.split()] for r in arr.text.strip().splitlines()])
    assert rows.shape == (60, 2)
    assert np.max(np.abs(rows[:, 0] - a.ravel())) <= 5e-11  # x fastest
    assert np.max(np.abs(rows[:, 1] - b.ravel())) <= 5e-11
    assert all(len(r) == 36 for r in arr.text.split("\n") if r.strip())


def test_coefficient_dump_round_trip(tmp_path):
    shape = (4, 3)
    u = np.array([1.5, -2e-07, 0.0, 123456.7, 1.0 / 3, -0.25, 7.0, 8.5, 9.25, 1e10, -1e-300, 2.0])
    path = tmp_path / "out_1.data"
    out.write_coefficients(path, u, shape)
    lines = path.read_text().splitlines()
    assert lines[:3] == ["0 0 1.5", "0 1 0.333333", "0 2 9.25"]  # i outermost, u(i, j) = u[i + 4j]
    assert lines[3] == "1 0 -2e-07"
    back = out.read_coefficients(path, shape)
    assert np.allclose(back, u, rtol=1e-5, atol=0)
    # a dump written by the reference's loop (text fixture) reads back the same way
    ref = tmp_path / "ref.data"
    ref.write_text("0 0 1.5\n1 0 -2e-07\n0 1 3\n1 1 4\n")
    assert np.array_equal(out.read_coefficients(ref, (2, 2)), [1.5, -2e-07, 3.0, 4.0])
