"""numpy/ctypes face of the CPU oracle (liboracle.so) and of the reference build
(_ref/libadsref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker. The product path never imports it.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def build():
    """Compile liboracle.so (and _ref/libadsref.so where the reference tree is present)."""
    import subprocess

    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _lib = C.CDLL(path)
        L = _lib
        L.orc_uniform.restype = C.c_double
        L.orc_sim_create.restype = C.c_void_p
        L.orc_sim_create.argtypes = [C.c_int, _ip, _ip, _dp, _dp, _ip, _ip]
        for f in ("orc_sim_destroy",):
            getattr(L, f).argtypes = [C.c_void_p]
        L.orc_sim_dofs.argtypes = [C.c_void_p, C.c_int]
        L.orc_sim_size.argtypes = [C.c_void_p]
        L.orc_sim_fix_dof.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.orc_sim_set_matrix.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double]
        L.orc_sim_get_matrix.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _ip]
        L.orc_sim_prepare_matrices.argtypes = [C.c_void_p]
        L.orc_sim_compute_rhs.argtypes = [C.c_void_p, C.c_void_p, _dp, _dp]
        L.orc_sim_compute_rhs_threaded.argtypes = [C.c_void_p, C.c_void_p, _dp, _dp, C.c_int]
        L.orc_sim_compute_rhs_sample.argtypes = [C.c_void_p, C.c_void_p, _dp, _dp, C.c_long, C.c_long]
        L.orc_sim_zero_boundary.argtypes = [C.c_void_p, _dp]
        L.orc_sim_solve.argtypes = [C.c_void_p, _dp, _ip]
        L.orc_sim_solve_transposed.argtypes = [C.c_void_p, _dp, _ip]
        L.orc_sim_tables.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp, _ip]
        L.orc_sim_kgrad.argtypes = [C.c_void_p, C.c_void_p, _dp, _dp, _dp]
        L.orc_sim_project.argtypes = [C.c_void_p, _dp, _dp]
        L.orc_sim_quad_points.argtypes = [C.c_void_p, C.c_int, _dp]
        L.orc_sim_error_sq.argtypes = [C.c_void_p, _dp, _dp, C.c_int]
        L.orc_sim_error_sq.restype = C.c_double
        L.orc_sim_sample.argtypes = [C.c_void_p, _dp, _ip, _dp, _dp, _dp, C.c_int, _dp]
        L.orc_sim_l2norm_sq.argtypes = [C.c_void_p, _dp]
        L.orc_sim_l2norm_sq.restype = C.c_double
        L.orc_sim_norm_sq.argtypes = [C.c_void_p, _dp, C.c_int]
        L.orc_sim_norm_sq.restype = C.c_double
        L.orc_init_gaussians.argtypes = [C.c_void_p, C.c_ulonglong, C.c_int, C.c_double,
                                         C.c_double, C.c_double, C.c_double, C.c_double, _dp]
        L.orc_find_span.argtypes = [C.c_double, _dp, C.c_int, C.c_int]
        L.orc_eval_basis.argtypes = [C.c_int, C.c_double, _dp, C.c_int, _dp]
        L.orc_eval_basis_with_derivatives.argtypes = [C.c_int, C.c_double, _dp, C.c_int, C.c_int, _dp]
        L.orc_create_basis.argtypes = [C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, _dp]
        L.orc_dgbtrf.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _dp, C.c_int, _ip]
        L.orc_dgbtrs.argtypes = [C.c_char, C.c_int, C.c_int, C.c_int, C.c_int, _dp, C.c_int, _ip,
                                 _dp, C.c_int]
        L.orc_cyclic_transpose.argtypes = [C.c_int, _ip, _dp, _dp]
        L.orc_band_multiply.argtypes = [_dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp,
                                        C.c_int, C.c_char]
        L.orc_greville.argtypes = [_dp, C.c_int, C.c_int, _dp]
    return _lib


def ref_lib():
    """The reference build, or None where it was not built (e.g. on a machine without the
    reference tree)."""
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libadsref.so")
        if not os.path.exists(path):
            return None
        _ref = C.CDLL(path)
        R = _ref
        R.ref_create_basis.argtypes = [C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, _dp]
        R.ref_create_basis_plain.argtypes = [C.c_double, C.c_double, C.c_int, C.c_int, _dp]
        R.ref_find_span.argtypes = [C.c_double, _dp, C.c_int, C.c_int]
        R.ref_eval_basis.argtypes = [C.c_int, C.c_double, _dp, C.c_int, C.c_int, _dp]
        R.ref_eval_basis_with_derivatives.argtypes = [C.c_int, C.c_double, _dp, C.c_int, C.c_int,
                                                      C.c_int, _dp]
        R.ref_first_nonzero_dofs.argtypes = [_dp, C.c_int, C.c_int, _ip]
        R.ref_elements_supporting_dofs.argtypes = [_dp, C.c_int, C.c_int, _ip, _ip]
        R.ref_gauss.argtypes = [C.c_int, _dp, _dp]
        R.ref_factorize.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _ip]
        R.ref_solve_with_factorized.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _ip, _dp, C.c_int]
        R.ref_cyclic_transpose.argtypes = [C.c_int, _ip, _dp, _dp]
        R.ref_ads_solve.argtypes = [C.c_int, _ip, _dp, C.POINTER(_dp), C.POINTER(_ip), _ip, _ip]
        R.ref_eval.argtypes = [C.c_int, _dp, _dp, C.POINTER(_dp), _ip, _ip, C.c_int, _dp]
        R.ref_gnuplot2.argtypes = [_dp, C.c_int, _dp, C.c_int, _dp, C.c_char_p, C.c_long]
        R.ref_gnuplot2.restype = C.c_long
    return _ref


# ------------------------------------------------------------------------------------
class RhsForm(C.Structure):
    """orc_rhs_form: integrand  mass*u*v - dt*(k*)sum_i grad_i u_i v_i
    + dt*reaction*u(1-u)*v + dt*f*v  (validation.hpp:110-115 for the heat form)."""

    _fields_ = [
        ("mass", C.c_double),
        ("dt", C.c_double),
        ("grad", C.c_double * 3),
        ("reaction", C.c_double),
        ("kfield", _dp),
        ("source", _dp),
        ("as_shipped", C.c_int),
    ]


def heat_form(dt, dim=3, grad=(1.0, 1.0, 1.0), reaction=0.0, kfield=None, source=None,
              as_shipped=False, mass=1.0):
    f = RhsForm()
    f.mass = mass
    f.dt = dt
    for i in range(3):
        f.grad[i] = grad[i]
    f.reaction = reaction
    f._keep = []
    if kfield is not None:
        kf = np.ascontiguousarray(kfield, dtype=np.float64)
        f._keep.append(kf)
        f.kfield = _d(kf)
    if source is not None:
        sf = np.ascontiguousarray(source, dtype=np.float64)
        f._keep.append(sf)
        f.source = _d(sf)
    f.as_shipped = 1 if as_shipped else 0
    return f


class Sim:
    """The oracle's simulation_2d/3d: dimension setup, matrices, compute_rhs, ads_solve."""

    def __init__(self, dim, p, elements, a=None, b=None, quad_order=None, repeated_nodes=None):
        L = lib()
        self.dim = dim
        p = list(p) + [0] * (3 - len(p))
        elements = list(elements) + [1] * (3 - len(elements))
        self._p = np.array(p, dtype=np.int32)
        self._e = np.array(elements, dtype=np.int32)
        self._a = np.array(a if a is not None else [0.0] * 3, dtype=np.float64)
        self._b = np.array(b if b is not None else [1.0] * 3, dtype=np.float64)
        self._q = np.array(quad_order if quad_order is not None else [0] * 3, dtype=np.int32)
        self._r = np.array(repeated_nodes if repeated_nodes is not None else [0] * 3, dtype=np.int32)
        self.h = L.orc_sim_create(dim, _i(self._p), _i(self._e), _d(self._a), _d(self._b),
                                  _i(self._q), _i(self._r))
        if not self.h:
            raise ValueError("orc_sim_create failed")
        self.shape = tuple(L.orc_sim_dofs(self.h, i) for i in range(dim))
        self.size = int(np.prod(self.shape))
        self.p = p[:dim]
        self.elements = elements[:dim]

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_sim_destroy(self.h)
            self.h = None

    def zeros(self):
        return np.zeros(self.size, dtype=np.float64)

    def fix_dof(self, d, k):
        assert lib().orc_sim_fix_dof(self.h, d, k) == 0

    def fix_all_boundaries(self):
        for d in range(self.dim):
            self.fix_dof(d, 0)
            self.fix_dof(d, self.shape[d] - 1)

    def set_matrix(self, d, slot, mass, stiff):
        assert lib().orc_sim_set_matrix(self.h, d, slot, mass, stiff) == 0

    def matrix(self, d, slot=0):
        n = self.shape[d]
        p = self.p[d]
        ab = np.zeros((n, 3 * p + 1), dtype=np.float64)
        ip = np.zeros(n, dtype=np.int32)
        assert lib().orc_sim_get_matrix(self.h, d, slot, _d(ab), _i(ip)) == 0
        return ab, ip

    def prepare_matrices(self):
        return lib().orc_sim_prepare_matrices(self.h)

    def compute_rhs(self, u_prev, form, threads=0):
        u_prev = np.ascontiguousarray(u_prev, dtype=np.float64)
        rhs = self.zeros()
        if threads:
            lib().orc_sim_compute_rhs_threaded(self.h, C.byref(form), _d(u_prev), _d(rhs), threads)
        else:
            lib().orc_sim_compute_rhs(self.h, C.byref(form), _d(u_prev), _d(rhs))
        return rhs

    def compute_rhs_sample(self, u_prev, form, first, count, rhs=None):
        """The sequential element loop over elements [first, first + count) only, added into
        rhs (a bounded sample of one assembly, for single-core timing)."""
        u_prev = np.ascontiguousarray(u_prev, dtype=np.float64)
        if rhs is None:
            rhs = self.zeros()
        lib().orc_sim_compute_rhs_sample(self.h, C.byref(form), _d(u_prev), _d(rhs), first, count)
        return rhs

    def kgrad(self, u, w, form):
        """d/dk of w^T rhs(u; k) for the gradient term of form (k-field adjoint)."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        w = np.ascontiguousarray(w, dtype=np.float64)
        out = self.zeros()
        lib().orc_sim_kgrad(self.h, C.byref(form), _d(u), _d(w), _d(out))
        return out

    def quad_points(self, d):
        q = int(self._q[d]) if self._q[d] > 0 else self.p[d] + 1
        out = np.zeros(self.elements[d] * q)
        lib().orc_sim_quad_points(self.h, d, _d(out))
        return out

    def project(self, fq, v=None):
        """projection.hpp:55-110: v += load vector of f given at the Gauss points (x fastest)."""
        fq = np.ascontiguousarray(fq, dtype=np.float64).ravel()
        v = self.zeros() if v is None else v
        lib().orc_sim_project(self.h, _d(fq), _d(v))
        return v

    def error_sq(self, u, sq, kind=0):
        """error(u, ..., norm, fun)^2: u may be None (u_h = 0); sq planes at the Gauss points."""
        sq = np.ascontiguousarray(sq, dtype=np.float64).ravel()
        up = _d(np.ascontiguousarray(u, dtype=np.float64)) if u is not None else None
        return lib().orc_sim_error_sq(self.h, up, _d(sq), kind)

    def sample(self, u, pts, ders=False):
        """The spline u on the tensor grid pts[0] x pts[1] (x pts[2]), x fastest; with ders
        the planes [value, d/dx, d/dy(, d/dz)]."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        P = [np.ascontiguousarray(p, dtype=np.float64) for p in pts] + [np.zeros(1)] * (3 - len(pts))
        counts = np.array([len(p) for p in pts] + [1] * (3 - len(pts)), dtype=np.int32)
        n = int(np.prod(counts))
        out = np.zeros(n * ((1 + self.dim) if ders else 1))
        lib().orc_sim_sample(self.h, _d(u), _i(counts), _d(P[0]), _d(P[1]), _d(P[2]), int(ders), _d(out))
        return out

    def zero_boundary(self, v):
        lib().orc_sim_zero_boundary(self.h, _d(v))

    def solve(self, rhs, slots=None, transposed=False):
        assert rhs.dtype == np.float64 and rhs.flags.c_contiguous
        sl = np.array(slots if slots is not None else [0] * 3, dtype=np.int32)
        if transposed:
            lib().orc_sim_solve_transposed(self.h, _d(rhs), _i(sl))
        else:
            lib().orc_sim_solve(self.h, _d(rhs), _i(sl))
        return rhs

    def tables(self, d):
        n_e = self.elements[d]
        p = self.p[d]
        q = int(self._q[d]) if self._q[d] > 0 else p + 1
        b = np.zeros((n_e, q, 2, p + 1))
        x = np.zeros((n_e, q))
        J = np.zeros(n_e)
        w = np.zeros(q)
        first = np.zeros(n_e, dtype=np.int32)
        lib().orc_sim_tables(self.h, d, _d(b), _d(x), _d(J), _d(w), _i(first))
        return dict(b=b, x=x, J=J, w=w, first=first)

    def l2norm_sq(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        return lib().orc_sim_l2norm_sq(self.h, _d(u))

    def norm_sq(self, u, kind):
        """Squared L2 (0), H1_0 (1) or H1 (2) norm of the coefficient vector u."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        return lib().orc_sim_norm_sq(self.h, _d(u), kind)

    def init_gaussians(self, seed, count, sigma=(0.1, 0.1), centre=(0.3, 0.7), noise=1e-3):
        u = self.zeros()
        lib().orc_init_gaussians(self.h, seed, count, sigma[0], sigma[1], centre[0], centre[1],
                                 noise, _d(u))
        return u


# ------------------------------------------------------------------------------------
def knots(a, b, p, elements, repeated_nodes=0):
    L = lib()
    ks = L.orc_knot_size(p, elements, repeated_nodes)
    k = np.zeros(ks)
    L.orc_create_basis(a, b, p, elements, repeated_nodes, _d(k))
    return k


def find_span(x, knot, p):
    knot = np.ascontiguousarray(knot, dtype=np.float64)
    return lib().orc_find_span(x, _d(knot), len(knot), p)


def eval_basis_with_derivatives(span, x, knot, p, der):
    knot = np.ascontiguousarray(knot, dtype=np.float64)
    out = np.zeros((der + 1, p + 1))
    lib().orc_eval_basis_with_derivatives(span, x, _d(knot), p, der, _d(out))
    return out


def eval_basis(span, x, knot, p):
    knot = np.ascontiguousarray(knot, dtype=np.float64)
    out = np.zeros(p + 1)
    lib().orc_eval_basis(span, x, _d(knot), p, _d(out))
    return out


def gauss(n):
    X = np.zeros(n)
    W = np.zeros(n)
    assert lib().orc_gauss(n, _d(X), _d(W)) == 0
    return X, W


def dgbtrf(ab, n, kl, ku):
    """ab: (n, 2kl+ku+1) array = LAPACK column-major band storage; factored in place."""
    ip = np.zeros(n, dtype=np.int32)
    info = lib().orc_dgbtrf(n, n, kl, ku, _d(ab), 2 * kl + ku + 1, _i(ip))
    return ip, info


def dgbtrs(ab, ip, b, n, kl, ku, trans="N"):
    """b: (nrhs, n) C-contiguous = column-major n x nrhs; solved in place."""
    nrhs = b.size // n
    lib().orc_dgbtrs(trans.encode(), n, kl, ku, nrhs, _d(ab), 2 * kl + ku + 1, _i(ip), _d(b), n)
    return b


def cyclic_transpose(sizes, a):
    s = np.array(sizes, dtype=np.int32)
    out = np.zeros_like(a)
    lib().orc_cyclic_transpose(len(sizes), _i(s), _d(np.ascontiguousarray(a)), _d(out))
    return out


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))
