"""Seeded synthetic inputs of the benchmark workloads, as B-spline coefficient vectors.

Coefficients are the field sampled at the Greville abscissae of each direction's clamped
uniform knot vector (bspline.cpp:26-43 builds the knots; lerp as util.hpp:17-26), first
index fastest. Random draws come from splitmix64 (Steele, Lea, Flood 2014), mapped to
[0, 1) as (z >> 11) * 2^-53; the k-th draw of a stream seeded with s is mix(s + k*gamma),
so the noise draws are vectorised.
"""
import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix(z):
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


class Stream:
    """splitmix64 stream: uniform() draws in [0, 1)."""

    def __init__(self, seed):
        self.k = 0
        self.seed = np.uint64(seed)

    def uniform(self, n=None):
        count = 1 if n is None else int(n)
        with np.errstate(over="ignore"):
            ks = np.arange(self.k + 1, self.k + 1 + count, dtype=np.uint64)
            z = _mix(self.seed + ks * _GAMMA)
        self.k += count
        u = (z >> np.uint64(11)).astype(np.float64) * 2.0**-53
        return float(u[0]) if n is None else u


def knots(p, elements, a=0.0, b=1.0):
    """Clamped uniform knot vector: p+1 copies of a, the interior points, p+1 copies of b."""
    t = np.arange(1, elements, dtype=np.float64) / float(elements)
    interior = (1.0 - t) * a + t * b
    return np.concatenate([np.full(p + 1, a), interior, np.full(p + 1, b)])


def greville(p, elements, a=0.0, b=1.0):
    """Greville abscissae (averages of p consecutive interior knots), one per dof."""
    kn = knots(p, elements, a, b)
    n = len(kn) - p - 1
    if p == 0:
        return kn[:n].copy()
    out = np.zeros(n)
    for i in range(n):
        s = 0.0
        for k in range(1, p + 1):
            s += kn[i + k]
        out[i] = s / p
    return out


def gaussians(p, elements, seed, count, sigma=(0.1, 0.1), centre=(0.3, 0.7), noise=1e-3):
    """Sum of `count` Gaussians exp(-|x-c|^2 / (2 sigma^2)) plus uniform noise in
    [-noise, noise]. Draw order: per Gaussian its centre (x, y[, z]) then sigma; then one
    noise draw per coefficient in memory order."""
    dim = len(p)
    st = Stream(seed)
    cs, sg = [], []
    for _ in range(count):
        c = [centre[0] + (centre[1] - centre[0]) * st.uniform() for _ in range(dim)]
        cs.append(c)
        sg.append(sigma[0] + (sigma[1] - sigma[0]) * st.uniform())
    g = [greville(p[d], elements[d]) for d in range(dim)]
    shape = tuple(len(x) for x in g)
    u = np.zeros(shape[::-1])  # (z,) y, x: first index fastest in memory
    for c, s in zip(cs, sg):
        r2 = 0.0
        for d in range(dim):
            axis = [1] * dim
            axis[dim - 1 - d] = shape[d]
            dx = (g[d] - c[d]).reshape(axis)
            r2 = r2 + dx * dx
        u += np.exp(-r2 / (2 * s * s))
    u = u.reshape(-1)
    u += noise * (2 * st.uniform(u.size) - 1)
    return u


def sine_source(p, elements, seed, terms=4):
    """Smooth separable source: sum over `terms` of a * prod_d sin(k_d pi x_d), with
    amplitude a uniform in [-1, 1] and wave numbers k_d uniform in {1, 2, 3}. Draw order:
    per term a, then k_x, k_y[, k_z]."""
    dim = len(p)
    st = Stream(seed)
    g = [greville(p[d], elements[d]) for d in range(dim)]
    shape = tuple(len(x) for x in g)
    f = np.zeros(shape[::-1])
    for _ in range(terms):
        a = 2 * st.uniform() - 1
        prod = a
        for d in range(dim):
            k = 1 + min(int(3 * st.uniform()), 2)
            axis = [1] * dim
            axis[dim - 1 - d] = shape[d]
            prod = prod * np.sin(k * np.pi * g[d]).reshape(axis)
        f += prod
    return f.reshape(-1)


def ensemble(seed, count, ne_range=(32, 256), degrees=(1, 2, 3), eps_range=(1e-4, 1e-2)):
    """Problems of the batched reaction-diffusion ensemble u_t = eps Lap u + u (1 - u):
    per problem (draw order) elements per side uniform in ne_range, degree uniform in
    degrees, eps log-uniform in eps_range, and a seed for its initial state. Returns a list
    of dicts {p, ne, eps, seed}."""
    st = Stream(seed)
    out = []
    for _ in range(count):
        ne = ne_range[0] + min(int((ne_range[1] - ne_range[0] + 1) * st.uniform()), ne_range[1] - ne_range[0])
        p = degrees[min(int(len(degrees) * st.uniform()), len(degrees) - 1)]
        lo, hi = np.log(eps_range[0]), np.log(eps_range[1])
        eps = float(np.exp(lo + (hi - lo) * st.uniform()))
        s = int(st.uniform() * 2**53)
        out.append({"p": p, "ne": ne, "eps": eps, "seed": s})
    return out


def smooth_unit(p, ne, seed):
    """Smooth initial data in [0, 1]: three seeded Gaussians (sigma in [0.1, 0.3]) rescaled
    to span [0, 1]."""
    u = gaussians([p, p], [ne, ne], seed, 3, sigma=(0.1, 0.3), centre=(0.2, 0.8), noise=0.0)
    lo, hi = u.min(), u.max()
    return (u - lo) / (hi - lo) if hi > lo else np.zeros_like(u)
