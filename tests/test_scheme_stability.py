"""Why the 3D implicit heat step is the all-direction (approximately factored) theta scheme
rather than the three-substep directional splitting of elasticity/implicit.hpp:464-511.

Per separable Fourier mode with 1D generalised eigenvalues mu_d (K_d v = mu_d M_d v, from
the oracle's own 1D matrices), one step multiplies the mode by
  split (3 substeps, substep d implicit in d, explicit in the others, tau = dt/2 as in the
  2D scheme of implicit.hpp:102-111):  prod_d (1 - tau sum_{e != d} mu_e) / (1 + tau mu_d)
  delta form (u += A^-1 dt (f - K u), A = (x)_d (M_d + theta dt K_d), theta = 1/2):
      1 - dt sum_d mu_d / prod_d (1 + theta dt mu_d)
At the 3D p=3, 192^3, dt=1e-4 workload the split amplifies the stiffest modes (|g| >> 1),
the delta form keeps |g| <= 1 for every mode.
"""
import numpy as np
import scipy.linalg

import pyoracle as po


def _dense(ab, n, p):
    A = np.zeros((n, n))
    for i in range(n):
        for j in range(max(0, i - p), min(n, i + p + 1)):
            A[i, j] = ab[j, 2 * p + i - j]
    return A


def _mu(p, ne):
    o = po.Sim(1, [p], [ne])
    M = _dense(o.matrix(0, 0)[0], o.shape[0], p)
    o.set_matrix(0, 1, 0.0, 1.0)
    K = _dense(o.matrix(0, 1)[0], o.shape[0], p)
    return scipy.linalg.eigh(K, M, eigvals_only=True)


def test_generalised_eigenvalues_scale_like_closed_form():
    # mu_max h^2 = 12.0 for p = 1 (the [1, 4, 1] / [-1, 2, -1] pair); with natural ends the
    # boundary modes of p = 3 reach 47.2 (14.56 with Dirichlet ends)
    assert abs(_mu(1, 64).max() / 64 ** 2 - 12.0) < 1e-9
    mu = _mu(3, 192)
    assert abs(mu.min()) < 1e-6  # constants are in the kernel of K (natural BC)
    assert 40.0 < mu.max() / 192 ** 2 < 50.0


def test_three_substep_split_is_unstable_and_delta_form_is_not():
    dt, tau, theta = 1e-4, 0.5e-4, 0.5
    mu = _mu(3, 192)
    sample = np.concatenate([mu[:5], mu[len(mu) // 2:len(mu) // 2 + 3], mu[-5:]])
    gx, gy, gz = np.meshgrid(sample, sample, sample, indexing="ij")
    split = ((1 - tau * (gy + gz)) / (1 + tau * gx)) * ((1 - tau * (gx + gz)) / (1 + tau * gy)) \
        * ((1 - tau * (gx + gy)) / (1 + tau * gz))
    delta = 1 - dt * (gx + gy + gz) / ((1 + theta * dt * gx) * (1 + theta * dt * gy)
                                       * (1 + theta * dt * gz))
    assert np.abs(split).max() > 100.0          # about 6.8 per substep triple at the top
    assert np.abs(delta).max() <= 1.0 + 1e-12
    # both agree with exp(-dt sum mu) to first order on the smooth modes
    smooth = (gx + gy + gz) * dt < 1e-3
    ref = np.exp(-dt * (gx + gy + gz))
    assert np.abs(delta - ref)[smooth].max() < 1e-6
