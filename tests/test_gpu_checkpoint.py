"""Checkpoint and resume (adsg_save_state / adsg_load_state): a run interrupted after k steps,
saved, loaded into a fresh handle and continued equals the uninterrupted run bit for bit;
foreign, truncated and corrupted files are refused before the state changes. The reference
only writes text dumps of coefficients (implicit.hpp:121-127); it has no resume."""
import numpy as np
import pytest

import adsgpu as ag
import pyoracle as po

pytestmark = pytest.mark.gpu


def _sim(precision, ne=(20, 17, 19)):
    g = ag.Simulation(3, [2, 2, 2], list(ne), precision=precision)
    g.prepare_matrices()
    return g


@pytest.mark.parametrize("precision", [ag.FP64, ag.FP32])
def test_resume_equals_uninterrupted_run(tmp_path, precision):
    print(f"checkpoint: precision {precision}", flush=True)
    dt = 1e-5
    form = ag.rhs_form(dt)
    o = po.Sim(3, [2, 2, 2], [20, 17, 19])
    u0 = o.init_gaussians(seed=8, count=3, sigma=(0.05, 0.15), centre=(0.2, 0.8), noise=1e-3)
    a = _sim(precision)
    a.set(ag.U, u0)
    a.step(4, form)
    path = tmp_path / "state.adsg"
    a.save_state(path, step=4, t=4 * dt)
    a.step(3, form)
    ref = a.get(ag.U)
    b = _sim(precision)
    step, t = b.load_state(path)
    assert step == 4 and t == 4 * dt
    b.step(3, form)
    assert np.array_equal(b.get(ag.U), ref)
    a.close()
    b.close()


def test_bad_state_files_are_refused(tmp_path):
    g = _sim(ag.FP64)
    u = np.random.default_rng(2).standard_normal(g.size)
    g.set(ag.U, u)
    good = tmp_path / "good.adsg"
    g.save_state(good, step=7, t=0.5)
    raw = good.read_bytes()
    cases = {
        "corrupt.adsg": (raw[:200] + bytes([raw[200] ^ 1]) + raw[201:], "checksum"),
        "short.adsg": (raw[:-8], "truncated"),
        "long.adsg": (raw + b"\0", "longer"),
        "junk.adsg": (b"x" * 300, "not a state file"),
    }
    g.set(ag.U, np.zeros(g.size))
    for name, (data, why) in cases.items():
        p = tmp_path / name
        p.write_bytes(data)
        with pytest.raises(ag.AdsgError, match=why):
            g.load_state(p)
        assert not np.any(g.get(ag.U)), name  # the state was left alone
    other = _sim(ag.FP64, ne=(20, 17, 18))
    with pytest.raises(ag.AdsgError, match="another configuration"):
        other.load_state(good)
    assert g.load_state(good) == (7, 0.5)
    assert np.array_equal(g.get(ag.U), u)
    g.close()
    other.close()
