"""The libtorch autograd Function over the ADS heat step with a k field (c4 shape: 2D,
p=2, 128^2, 200 steps of dt = 1e-6, checkpoints every 20 steps), run in its own process by
torch_adjoint_check.py: torch must bring up the GPU runtime before libadsgpu.so is loaded,
which other tests in this process have already done the other way round. Bars: gradients
against the oracle's FP64 discrete adjoint <= 1e-10, 32 central differences (h = 1e-6)
<= 1e-5.
"""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu


def test_torch_function_gradient_matches_oracle_adjoint_and_fd():
    print("case: torch heat_k (separate process)", flush=True)
    here = os.path.dirname(os.path.abspath(__file__))
    r = subprocess.run([sys.executable, "-u", os.path.join(here, "torch_adjoint_check.py")],
                       capture_output=True, text=True, timeout=600)
    print(r.stdout[-3000:], r.stderr[-3000:], flush=True)
    assert r.returncode == 0
    assert "torch adjoint check: ok" in r.stdout
