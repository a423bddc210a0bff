"""NaN / Inf detection of the step (ADSG_STEP_CHECK_FINITE, adsg_check_finite). The
reference ignores such states (LAPACK's info is dropped, band_solve.hpp:16-31); here a step
asked to check returns ADSG_ENONFINITE instead of carrying them on silently."""
import numpy as np
import pytest

import adsgpu as ag

pytestmark = pytest.mark.gpu


def rc_of(fn):
    try:
        fn()
        return ag.OK, ""
    except ag.AdsgError as e:
        return int(str(e).split(":")[1]), str(e)


@pytest.mark.parametrize("precision", [ag.FP64, ag.FP32])
@pytest.mark.parametrize("bad", [np.nan, np.inf, -np.inf])
def test_step_reports_nonfinite_state(precision, bad):
    print(f"nonfinite: precision {precision} value {bad}", flush=True)
    sim = ag.Simulation(3, [2, 2, 2], [12, 10, 9], precision=precision)
    sim.prepare_matrices()
    rng = np.random.default_rng(3)
    u = rng.uniform(0.0, 1.0, sim.size)
    form = ag.rhs_form(1e-4)
    sim.set(ag.U, u)
    # a clean state passes the check, and the checked steps equal the unchecked ones
    assert rc_of(lambda: sim.step(3, form, ag.STEP_CHECK_FINITE)) == (ag.OK, "")
    a = sim.get(ag.U)
    sim.set(ag.U, u)
    sim.step(3, form)
    assert np.array_equal(a, sim.get(ag.U))
    assert sim.check_finite(ag.U) == (ag.OK, "")
    # one coefficient in the interior set to NaN / Inf
    u[sim.size // 2 + 7] = bad
    sim.set(ag.U, u)
    assert sim.check_finite(ag.U)[0] == ag.ENONFINITE
    rc, msg = rc_of(lambda: sim.step(2, form, ag.STEP_CHECK_FINITE))
    assert rc == ag.ENONFINITE, msg
    assert "NaN or Inf" in msg
    # without the flag the same steps run and report nothing (the reference's behaviour)
    sim.set(ag.U, u)
    assert rc_of(lambda: sim.step(2, form)) == (ag.OK, "")
    assert not np.all(np.isfinite(sim.get(ag.U)))
    sim.close()


def test_delta_step_reports_nonfinite_state():
    sim = ag.Simulation(2, [3, 3], [20, 18])
    for d in range(2):
        sim.set_matrix(d, 1, 1.0, 0.5e-3)
    sim.prepare_matrices()
    u = np.random.default_rng(4).uniform(0.0, 1.0, sim.size)
    form = ag.rhs_form(1e-3, mass=0.0)
    sim.set(ag.U, u)
    assert rc_of(lambda: sim.step_delta(2, form, [1, 1], ag.STEP_CHECK_FINITE)) == (ag.OK, "")
    u[0] = np.nan  # a corner coefficient: its support reaches the interior within a step
    sim.set(ag.U, u)
    rc, msg = rc_of(lambda: sim.step_delta(1, form, [1, 1], ag.STEP_CHECK_FINITE))
    assert rc == ag.ENONFINITE, msg
    sim.close()


def test_check_finite_rejects_bad_arguments():
    sim = ag.Simulation(2, [2, 2], [8, 8], precision=ag.FP32)
    assert ag.lib().adsg_check_finite(sim.h, 9) == ag.EINVAL
    assert ag.lib().adsg_check_finite(sim.h, ag.BUFFER) == ag.EINVAL  # FP32 buffer
    assert ag.lib().adsg_check_finite(sim.h, ag.KFIELD) == ag.ESTATE  # never set
    sim.close()

