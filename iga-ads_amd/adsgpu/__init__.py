"""ctypes binding of libadsgpu.so (the C ABI in include/adsgpu.h).

The Python side is a thin driver for tests and the benchmark; the product is the C ABI
and the C++ host surface in iga-ads_amd/host/ads_gpu/simulation.hpp. Importing this
module loads the HIP library and raises if it has not been built: there is no CPU path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libadsgpu.so")

OK, EINVAL, EHIP, ENOMEM, ESTATE, ENONFINITE = 0, -1, -2, -3, -4, -5
FP64, FP32 = 0, 1
U, U_PREV, BUFFER, KFIELD, SOURCE, GRAD_K, GRAD_F = 0, 1, 2, 3, 4, 5, 6
STEP_DIRICHLET = 1
STEP_REFINE = 2
STEP_CHECK_FINITE = 4

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class DimConfig(C.Structure):
    _fields_ = [("p", C.c_int), ("elements", C.c_int), ("a", C.c_double), ("b", C.c_double),
                ("quad_order", C.c_int), ("repeated_nodes", C.c_int)]


class Config(C.Structure):
    _fields_ = [("dim", C.c_int), ("x", DimConfig), ("y", DimConfig), ("z", DimConfig),
                ("step_count", C.c_int), ("dt", C.c_double), ("derivatives", C.c_int),
                ("precision", C.c_int)]


class RhsForm(C.Structure):
    _fields_ = [("mass", C.c_double), ("dt", C.c_double), ("grad", C.c_double * 3),
                ("reaction", C.c_double), ("use_kfield", C.c_int), ("use_source", C.c_int)]


def rhs_form(dt, grad=(1.0, 1.0, 1.0), mass=1.0, reaction=0.0, kfield=False, source=False):
    f = RhsForm()
    f.mass = mass
    f.dt = dt
    for i in range(3):
        f.grad[i] = grad[i]
    f.reaction = reaction
    f.use_kfield = int(bool(kfield))
    f.use_source = int(bool(source))
    return f


class AdsgError(RuntimeError):
    pass


_lib = None

EXPORTS = [
    "adsg_create", "adsg_destroy", "adsg_last_error", "adsg_dofs", "adsg_size", "adsg_precision",
    "adsg_fix_dof", "adsg_set_matrix", "adsg_set_band_matrix", "adsg_prepare_matrices",
    "adsg_set_vector", "adsg_get_vector", "adsg_device_ptr", "adsg_swap", "adsg_compute_rhs",
    "adsg_zero_boundary", "adsg_ads_solve", "adsg_ads_solve_transposed", "adsg_step",
    "adsg_step_delta", "adsg_adjoint_step", "adsg_adjoint_step_delta", "adsg_synchronize",
    "adsg_set_vector_device", "adsg_get_vector_device", "adsg_norm", "adsg_set_line_matrices",
    "adsg_forward_checkpointed", "adsg_adjoint_checkpointed",
    "adsg_batch_create", "adsg_batch_destroy", "adsg_batch_last_error", "adsg_batch_count",
    "adsg_batch_size", "adsg_batch_dofs", "adsg_batch_prepare", "adsg_batch_set_forms",
    "adsg_batch_set_vector", "adsg_batch_get_vector", "adsg_batch_step", "adsg_batch_synchronize",
    "adsg_batch_set_tuning", "adsg_batch_get_tuning", "adsg_batch_autotune",
    "adsg_band_factorize", "adsg_time_phases", "adsg_set_tuning", "adsg_get_tuning",
    "adsg_autotune", "adsg_project", "adsg_project_device", "adsg_quad_points", "adsg_error",
    "adsg_sample", "adsg_norm_noalloc", "adsg_check_finite", "adsg_device_rates",
    "adsg_save_state", "adsg_load_state",
]
TUNE_RHS_VARIANT = 0
TUNE_SWEEP_VARIANT = 1
TUNE_RHS_KRONECKER = 3
TUNE_BATCH_TILE = 2
TUNE_RHS_COLUMNS = 4
TUNE_RHS_TILING = 5
TUNE_CACHE = 6


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise AdsgError(f"{LIB_PATH} is not built; run __graft_entry__.build() first")
        L = C.CDLL(LIB_PATH)
        L.adsg_create.argtypes = [C.POINTER(Config), C.POINTER(C.c_void_p)]
        L.adsg_destroy.argtypes = [C.c_void_p]
        L.adsg_destroy.restype = None
        L.adsg_last_error.argtypes = [C.c_void_p]
        L.adsg_last_error.restype = C.c_char_p
        L.adsg_dofs.argtypes = [C.c_void_p, C.c_int]
        L.adsg_size.argtypes = [C.c_void_p]
        L.adsg_size.restype = C.c_long
        L.adsg_precision.argtypes = [C.c_void_p]
        L.adsg_fix_dof.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.adsg_set_matrix.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double]
        L.adsg_set_band_matrix.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, _dp]
        L.adsg_prepare_matrices.argtypes = [C.c_void_p]
        L.adsg_set_vector.argtypes = [C.c_void_p, C.c_int, _dp]
        L.adsg_get_vector.argtypes = [C.c_void_p, C.c_int, _dp]
        L.adsg_device_ptr.argtypes = [C.c_void_p, C.c_int]
        L.adsg_device_ptr.restype = C.c_void_p
        L.adsg_swap.argtypes = [C.c_void_p]
        L.adsg_compute_rhs.argtypes = [C.c_void_p, C.POINTER(RhsForm)]
        L.adsg_zero_boundary.argtypes = [C.c_void_p]
        L.adsg_ads_solve.argtypes = [C.c_void_p, _ip]
        L.adsg_step.argtypes = [C.c_void_p, C.c_int, C.POINTER(RhsForm), C.c_int]
        L.adsg_ads_solve_transposed.argtypes = [C.c_void_p, _ip]
        L.adsg_step_delta.argtypes = [C.c_void_p, C.c_int, C.POINTER(RhsForm), _ip, C.c_int]
        L.adsg_adjoint_step_delta.argtypes = [C.c_void_p, C.c_int, C.POINTER(RhsForm), _ip, C.c_int]
        L.adsg_adjoint_step.argtypes = [C.c_void_p, C.POINTER(RhsForm), C.c_int]
        L.adsg_batch_create.argtypes = [C.POINTER(Config), C.c_int, C.POINTER(C.c_void_p)]
        L.adsg_batch_destroy.argtypes = [C.c_void_p]
        L.adsg_batch_destroy.restype = None
        L.adsg_batch_last_error.argtypes = [C.c_void_p]
        L.adsg_batch_last_error.restype = C.c_char_p
        L.adsg_batch_count.argtypes = [C.c_void_p]
        L.adsg_batch_size.argtypes = [C.c_void_p, C.c_int]
        L.adsg_batch_size.restype = C.c_long
        L.adsg_batch_dofs.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.adsg_batch_prepare.argtypes = [C.c_void_p]
        L.adsg_batch_set_forms.argtypes = [C.c_void_p, C.POINTER(RhsForm)]
        L.adsg_batch_set_vector.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp]
        L.adsg_batch_get_vector.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp]
        L.adsg_batch_step.argtypes = [C.c_void_p, C.c_int]
        L.adsg_batch_synchronize.argtypes = [C.c_void_p]
        L.adsg_batch_set_tuning.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.adsg_batch_get_tuning.argtypes = [C.c_void_p, C.c_int]
        L.adsg_batch_autotune.argtypes = [C.c_void_p, _dp]
        L.adsg_set_vector_device.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.adsg_norm.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp]
        L.adsg_forward_checkpointed.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(RhsForm), C.c_int]
        L.adsg_adjoint_checkpointed.argtypes = [C.c_void_p, C.POINTER(RhsForm), C.c_int]
        L.adsg_set_line_matrices.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, _dp]
        L.adsg_get_vector_device.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.adsg_synchronize.argtypes = [C.c_void_p]
        L.adsg_band_factorize.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _ip]
        L.adsg_time_phases.argtypes = [C.c_void_p, C.c_int, C.POINTER(RhsForm), _ip, C.c_int, _dp]
        L.adsg_set_tuning.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.adsg_get_tuning.argtypes = [C.c_void_p, C.c_int]
        L.adsg_autotune.argtypes = [C.c_void_p, C.POINTER(RhsForm), _ip, C.c_int]
        L.adsg_project.argtypes = [C.c_void_p, C.c_int, _dp]
        L.adsg_project_device.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.adsg_quad_points.argtypes = [C.c_void_p, C.c_int, _dp]
        L.adsg_error.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _dp]
        L.adsg_check_finite.argtypes = [C.c_void_p, C.c_int]
        L.adsg_device_rates.argtypes = [_dp]
        L.adsg_save_state.argtypes = [C.c_void_p, C.c_char_p, C.c_long, C.c_double]
        L.adsg_load_state.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_long), _dp]
        L.adsg_norm_noalloc.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, _dp]
        L.adsg_sample.argtypes = [C.c_void_p, C.c_int, _ip, _dp, _dp, _dp, C.c_int, _dp]
        _lib = L
    return _lib


def device_rates():
    """adsg_device_rates: measured FMA ceilings of the current device in TFLOP/s."""
    out = np.zeros(4)
    rc = lib().adsg_device_rates(out.ctypes.data_as(_dp))
    if rc != OK:
        raise AdsgError(f"device_rates: {rc}")
    return dict(zip(("valu_fp64", "valu_fp32", "mfma_fp64", "mfma_fp32"), out.tolist()))


def linspace(a, b, n):
    """util.hpp:28-34: n + 1 points (1 - t) a + t b, t = i / n."""
    t = np.arange(n + 1, dtype=np.float64) / float(n)
    return (1.0 - t) * a + t * b


class Simulation:
    """One adsg_sim handle: the simulation_2d / simulation_3d surface on the GPU."""

    def __init__(self, dim, p, elements, dt=0.0, steps=0, precision=FP64, a=None, b=None,
                 quad_order=None, repeated_nodes=None):
        """dim_config per direction (simulation/config.hpp:11-29): degree p, elements,
        interval [a, b], quad_order (Gauss points per element; None or 0 = p + 1) and
        repeated_nodes (extra copies of every interior knot)."""
        L = lib()
        cfg = Config()
        cfg.dim = dim
        ps = list(p) if hasattr(p, "__len__") else [p] * dim
        es = list(elements) if hasattr(elements, "__len__") else [elements] * dim
        for i, dc in enumerate((cfg.x, cfg.y, cfg.z)):
            if i < dim:
                dc.p = ps[i]
                dc.elements = es[i]
                dc.a = a[i] if a else 0.0
                dc.b = b[i] if b else 1.0
                dc.quad_order = quad_order[i] if quad_order else 0
                dc.repeated_nodes = repeated_nodes[i] if repeated_nodes else 0
        cfg.step_count = steps
        cfg.dt = dt
        cfg.derivatives = 1
        cfg.precision = precision
        h = C.c_void_p()
        rc = L.adsg_create(C.byref(cfg), C.byref(h))
        if rc != OK:
            raise AdsgError(f"adsg_create: {rc}: {L.adsg_last_error(None).decode()}")
        self.h = h
        self.dim = dim
        self.p = ps[:dim]
        self.elements = es[:dim]
        self.a = [a[i] if a else 0.0 for i in range(dim)]
        self.b = [b[i] if b else 1.0 for i in range(dim)]
        self.precision = precision
        self.shape = tuple(L.adsg_dofs(h, d) for d in range(dim))
        self.size = int(L.adsg_size(h))

    def close(self):
        if getattr(self, "h", None):
            lib().adsg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != OK:
            msg = lib().adsg_last_error(self.h).decode()
            raise AdsgError(f"{what}: {rc}: {msg}")
        return rc

    def fix_dof(self, d, k):
        self._check(lib().adsg_fix_dof(self.h, d, k), "fix_dof")

    def fix_all_boundaries(self):
        for d in range(self.dim):
            self.fix_dof(d, 0)
            self.fix_dof(d, self.shape[d] - 1)

    def set_matrix(self, d, slot, mass, stiff):
        self._check(lib().adsg_set_matrix(self.h, d, slot, mass, stiff), "set_matrix")

    def set_band_matrix(self, d, slot, kl, ku, ab):
        ab = np.ascontiguousarray(ab, dtype=np.float64)
        self._check(lib().adsg_set_band_matrix(self.h, d, slot, kl, ku, ab.ctypes.data_as(_dp)),
                    "set_band_matrix")

    def prepare_matrices(self):
        return self._check(lib().adsg_prepare_matrices(self.h), "prepare_matrices")

    def set(self, which, v):
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
        assert v.size == self.size
        self._check(lib().adsg_set_vector(self.h, which, v.ctypes.data_as(_dp)), "set_vector")

    def get(self, which=U):
        out = np.zeros(self.size)
        self._check(lib().adsg_get_vector(self.h, which, out.ctypes.data_as(_dp)), "get_vector")
        return out

    def swap(self):
        self._check(lib().adsg_swap(self.h), "swap")

    def compute_rhs(self, form):
        self._check(lib().adsg_compute_rhs(self.h, C.byref(form)), "compute_rhs")

    def zero_boundary(self):
        self._check(lib().adsg_zero_boundary(self.h), "zero_boundary")

    @staticmethod
    def _slots(slots):
        if slots is None:
            return None
        return (C.c_int * 3)(*(list(slots) + [0] * (3 - len(slots))))

    def ads_solve(self, slots=None):
        self._check(lib().adsg_ads_solve(self.h, self._slots(slots)), "ads_solve")

    def ads_solve_transposed(self, slots=None):
        """u := (A_x (x) A_y [(x) A_z])^-T u (dgbtrs 'T' per direction)."""
        self._check(lib().adsg_ads_solve_transposed(self.h, self._slots(slots)),
                    "ads_solve_transposed")

    def step(self, n, form, flags=0):
        self._check(lib().adsg_step(self.h, n, C.byref(form), flags), "step")

    def save_state(self, path, step=0, t=0.0):
        """adsg_save_state: u and u_prev with a checked header (configuration, step, time)."""
        self._check(lib().adsg_save_state(self.h, os.fsencode(path), step, t), "save_state")

    def load_state(self, path):
        """adsg_load_state: restore u and u_prev; returns (step, t) as saved."""
        st, t = C.c_long(0), np.zeros(1)
        self._check(lib().adsg_load_state(self.h, os.fsencode(path), C.byref(st), t.ctypes.data_as(_dp)),
                    "load_state")
        return int(st.value), float(t[0])

    def check_finite(self, which=U):
        """adsg_check_finite: the return code (OK or ENONFINITE) and the library's message."""
        rc = lib().adsg_check_finite(self.h, which)
        return rc, (lib().adsg_last_error(self.h).decode() if rc != OK else "")

    def step_delta(self, n, form, slots, flags=0):
        """n steps of u := u_prev + A_slots^-1 r(u_prev), r the integrand of form."""
        self._check(lib().adsg_step_delta(self.h, n, C.byref(form), self._slots(slots), flags),
                    "step_delta")

    def adjoint_step(self, form, flags=0):
        """One adjoint step of step(): u holds lam(n+1), u_prev the forward state u(n);
        u := lam(n) and dL/dk is added into GRAD_K when the form uses the k field."""
        self._check(lib().adsg_adjoint_step(self.h, C.byref(form), flags), "adjoint_step")

    def set_line_matrices(self, d, slot, kl, ku, abs_):
        """One LAPACK band matrix per line of direction d (generalized ADS); abs_ has shape
        (lines, n, 2kl+ku+1) in the reference's cyclic line order."""
        a = np.ascontiguousarray(abs_, dtype=np.float64)
        self._check(lib().adsg_set_line_matrices(self.h, d, slot, kl, ku, a.ctypes.data_as(_dp)),
                    "set_line_matrices")

    def norm(self, kind=0, which=U):
        """normL2 (kind 0), H1_0 seminorm (1) or normH1 (2) of a state vector, on the device."""
        out = np.zeros(1)
        self._check(lib().adsg_norm(self.h, which, kind, out.ctypes.data_as(_dp)), "norm")
        return float(out[0])

    def norm_noalloc(self, kind=0, which=U, sq_dev=None):
        """adsg_norm_noalloc: norm (sq_dev None) or error against function values held in
        device memory at address sq_dev (layout of error()); which=None with sq_dev gives
        the norm of those values. Allocates no device memory."""
        out = np.zeros(1)
        w = -1 if which is None else which
        self._check(lib().adsg_norm_noalloc(self.h, w, kind, sq_dev, out.ctypes.data_as(_dp)),
                    "norm_noalloc")
        return float(out[0])

    def quad_points(self, d):
        """Gauss points of direction d, element-major (elements*(p+1) abscissae)."""
        out = np.zeros(self.elements[d] * (self.p[d] + 1))
        self._check(lib().adsg_quad_points(self.h, d, out.ctypes.data_as(_dp)), "quad_points")
        return out

    def quad_grid(self):
        """Coordinate arrays of the Gauss points, indexed [z][y][x] (x fastest when raveled):
        f(*sim.quad_grid()) gives f at the Gauss points in the layout adsg_project reads."""
        pts = [self.quad_points(d) for d in range(self.dim)]
        return [g.transpose() for g in np.meshgrid(*pts, indexing="ij")]

    def error(self, s, kind=0, which=U):
        """error(u, ..., norm, fun) (basic_simulation_2d.hpp:318-336): s at the Gauss points,
        planes [value, d/dx, d/dy(, d/dz)] (value only for kind 0); which=None gives the norm
        of s itself (normL2/normH1(Ux, Uy, fun))."""
        sq = np.ascontiguousarray(s, dtype=np.float64)
        count = int(np.prod([self.elements[d] * (self.p[d] + 1) for d in range(self.dim)]))
        planes = 1 if kind == 0 else 1 + self.dim
        if sq.size < planes * count:
            raise ValueError(f"error: {sq.size} values for {planes} planes of {count} points")
        out = np.zeros(1)
        w = -1 if which is None else which
        self._check(lib().adsg_error(self.h, w, kind, sq.ravel().ctypes.data_as(_dp),
                                     out.ctypes.data_as(_dp)), "error")
        return float(out[0])

    def sample(self, pts=None, intervals=None, ders=False, which=U):
        """The spline of vector `which` on a tensor grid (bspline::eval / eval_ders): pts a
        list of abscissae per direction, or `intervals` per direction for the output
        manager's uniform grid linspace(a, b, intervals) (output/axis.hpp:15-25). Returns an
        array indexed [z][y][x] (with ders, a leading axis over value, d/dx, d/dy(, d/dz))."""
        if pts is None:
            iv = list(intervals) if hasattr(intervals, "__len__") else [intervals] * self.dim
            pts = [linspace(self.a[d], self.b[d], iv[d]) for d in range(self.dim)]
        P = [np.ascontiguousarray(p, dtype=np.float64) for p in pts] + [np.zeros(1)] * (3 - len(pts))
        counts = np.array([len(p) for p in pts] + [1] * (3 - len(pts)), dtype=np.int32)
        shape = tuple(len(p) for p in pts[::-1])
        planes = (1 + self.dim) if ders else 1
        out = np.zeros(planes * int(np.prod(shape)))
        self._check(lib().adsg_sample(self.h, which, counts.ctypes.data_as(_ip), P[0].ctypes.data_as(_dp),
                                      P[1].ctypes.data_as(_dp), P[2].ctypes.data_as(_dp), int(ders),
                                      out.ctypes.data_as(_dp)), "sample")
        return out.reshape((planes,) + shape) if ders else out.reshape(shape)

    def project(self, f, which=U):
        """projection(v, f) (projection.hpp:11-153): add the L2 projection's load vector of f
        into vector `which`; f is a callable of the coordinates (vectorised over numpy arrays)
        or its values at the Gauss points, x fastest. Follow with ads_solve() to get the
        coefficients, as the reference's before() does."""
        if callable(f):
            fq = np.ascontiguousarray(f(*self.quad_grid()), dtype=np.float64)
        else:
            fq = np.ascontiguousarray(f, dtype=np.float64)
        count = int(np.prod([self.elements[d] * (self.p[d] + 1) for d in range(self.dim)]))
        if fq.size != count:
            raise ValueError(f"project: {fq.size} values for {count} Gauss points")
        self._check(lib().adsg_project(self.h, which, fq.ravel().ctypes.data_as(_dp)), "project")

    def forward_checkpointed(self, n, every, form, flags=0):
        """n explicit steps from u keeping a checkpoint every `every` steps (adjoint runs)."""
        self._check(lib().adsg_forward_checkpointed(self.h, n, every, C.byref(form), flags),
                    "forward_checkpointed")

    def adjoint_checkpointed(self, form, flags=0):
        """u = dL/du(T) in, dL/du(0) out; dL/dk added into GRAD_K (k-field forms)."""
        self._check(lib().adsg_adjoint_checkpointed(self.h, C.byref(form), flags),
                    "adjoint_checkpointed")

    def set_device(self, which, ptr):
        """Copy from FP64 device memory at ptr (e.g. a tensor's data_ptr())."""
        self._check(lib().adsg_set_vector_device(self.h, which, C.c_void_p(ptr)), "set_vector_device")

    def get_device(self, which, ptr):
        self._check(lib().adsg_get_vector_device(self.h, which, C.c_void_p(ptr)), "get_vector_device")

    def adjoint_step_delta(self, n, form, slots, flags=0):
        """n adjoint steps of step_delta applied to u in place (u_prev, buffer clobbered)."""
        self._check(lib().adsg_adjoint_step_delta(self.h, n, C.byref(form), self._slots(slots),
                                                  flags), "adjoint_step_delta")

    def synchronize(self):
        self._check(lib().adsg_synchronize(self.h), "synchronize")

    PHASES = ("rhs", "boundary", "sweep_x", "sweep_y", "sweep_z", "step", "fill")

    def time_phases(self, n, form, slots=None, flags=0):
        """Average device ms per phase of n eager steps (explicit steps, or delta steps with
        slots): element kernels (with the rhs zero fill where one is needed), boundary
        handling, x/y/z sweeps, whole step, rhs start (a source's load vector)."""
        ms = np.zeros(7)
        self._check(lib().adsg_time_phases(self.h, n, C.byref(form), self._slots(slots), flags,
                                           ms.ctypes.data_as(_dp)), "time_phases")
        return dict(zip(self.PHASES, ms.tolist()))

    def set_tuning(self, key, value):
        self._check(lib().adsg_set_tuning(self.h, key, value), "set_tuning")

    def get_tuning(self, key):
        return lib().adsg_get_tuning(self.h, key)

    def autotune(self, form, slots=None, flags=0):
        self._check(lib().adsg_autotune(self.h, C.byref(form), self._slots(slots), flags),
                    "autotune")
        return self.get_tuning(TUNE_RHS_VARIANT)

    def device_ptr(self, which):
        return lib().adsg_device_ptr(self.h, which)


def hip_runtimes(maps_text=None):
    """Paths of the distinct HIP runtime libraries (libamdhip64) mapped into this process,
    read from /proc/self/maps (or from `maps_text`, the same format)."""
    if maps_text is None:
        with open("/proc/self/maps") as fh:
            maps_text = fh.read()
    paths = set()
    for line in maps_text.splitlines():
        parts = line.split(None, 5)
        if len(parts) == 6 and "libamdhip64" in os.path.basename(parts[5].strip()):
            paths.add(os.path.realpath(parts[5].strip()))
    return paths


def check_single_hip_runtime(maps_text=None):
    """Raise AdsgError when more than one HIP runtime is mapped. torch ships its own
    libamdhip64; when libadsgpu.so is loaded before torch, torch then loads its copy beside
    the one libadsgpu.so uses, and device pointers of one runtime must not reach the other.
    Importing torch before the first use of this module gives one runtime, torch's."""
    paths = hip_runtimes(maps_text)
    if len(paths) > 1:
        raise AdsgError("two HIP runtimes are loaded in this process (" + ", ".join(sorted(paths)) +
                        "): import torch before adsgpu loads libadsgpu.so, so that both use "
                        "torch's runtime")


def heat_k(sim, k, u0, steps, every, dt, dirichlet=True):
    """u(T) of `steps` explicit heat steps with diffusion field k, as a differentiable torch
    operation (libtorch autograd Function; backward = hand-written adjoint, checkpoints
    every `every` steps). k and u0: float64 GPU tensors of sim.size entries."""
    from . import adsg_torch  # needs torch; built beside libadsgpu.so

    check_single_hip_runtime()
    return adsg_torch.heat_k(k, u0, sim.h.value, steps, every, dt, dirichlet)


def heat_kf(sim, k, f, u0, steps, every, dt, dirichlet=True):
    """As heat_k with a source field f as well (the form {mass 1, dt, k field, source f}),
    differentiable in k, f and u0."""
    from . import adsg_torch

    check_single_hip_runtime()
    return adsg_torch.heat_kf(k, f, u0, sim.h.value, steps, every, dt, dirichlet)


class Batch:
    """Many independent 2D problems (degree 1..3, natural boundaries) advanced together with
    grouped launches and one captured graph per step count (adsg_batch_*)."""

    def __init__(self, problems, precision=FP32, interval=(0.0, 1.0)):
        """problems: list of (p, elements) pairs, square meshes on interval^2 ([0,1]^2 by
        default)."""
        L = lib()
        cfgs = (Config * len(problems))()
        for c, (p, ne) in zip(cfgs, problems):
            c.dim = 2
            for dc in (c.x, c.y):
                dc.p = p
                dc.elements = ne
                dc.a, dc.b = float(interval[0]), float(interval[1])
                dc.quad_order = 0
                dc.repeated_nodes = 0
            c.derivatives = 1
            c.precision = precision
        h = C.c_void_p()
        rc = L.adsg_batch_create(cfgs, len(problems), C.byref(h))
        if rc != OK:
            raise AdsgError(f"adsg_batch_create: {rc}: {L.adsg_batch_last_error(None).decode()}")
        self.h = h
        self.count = len(problems)
        self.sizes = [int(L.adsg_batch_size(h, i)) for i in range(self.count)]

    def _check(self, rc, what):
        if rc != OK:
            raise AdsgError(f"{what}: {rc}: {lib().adsg_batch_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None):
            lib().adsg_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def prepare(self):
        self._check(lib().adsg_batch_prepare(self.h), "batch_prepare")

    def set_forms(self, forms):
        arr = (RhsForm * self.count)(*forms)
        self._check(lib().adsg_batch_set_forms(self.h, arr), "batch_set_forms")

    def set(self, i, v, which=U):
        v = np.ascontiguousarray(v, dtype=np.float64)
        assert v.size == self.sizes[i]
        self._check(lib().adsg_batch_set_vector(self.h, i, which, v.ctypes.data_as(_dp)), "batch_set_vector")

    def get(self, i, which=U):
        v = np.empty(self.sizes[i])
        self._check(lib().adsg_batch_get_vector(self.h, i, which, v.ctypes.data_as(_dp)), "batch_get_vector")
        return v

    def step(self, n):
        self._check(lib().adsg_batch_step(self.h, n), "batch_step")

    def synchronize(self):
        self._check(lib().adsg_batch_synchronize(self.h), "batch_synchronize")

    def set_tuning(self, value, key=TUNE_BATCH_TILE):
        """TUNE_BATCH_TILE: element tile (0 = 16x16, 1 = 8x8); TUNE_SWEEP_VARIANT: 0 =
        segmented line sweeps, 1 = one lane per line."""
        self._check(lib().adsg_batch_set_tuning(self.h, key, value), "batch_set_tuning")

    def get_tuning(self, key=TUNE_BATCH_TILE):
        return lib().adsg_batch_get_tuning(self.h, key)

    def autotune(self):
        ms = np.zeros(2)
        self._check(lib().adsg_batch_autotune(self.h, ms.ctypes.data_as(_dp)), "batch_autotune")
        return self.get_tuning(), ms
