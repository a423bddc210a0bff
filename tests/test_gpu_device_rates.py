"""adsg_device_rates: the vector and matrix FMA ceilings the benchmark reports beside the
element kernels' rates (no counterpart in the reference)."""
import numpy as np
import pytest

import adsgpu as ag

pytestmark = pytest.mark.gpu


def test_device_rates_are_measured():
    r = ag.device_rates()
    print("device rates (TFLOP/s):", r, flush=True)
    assert set(r) == {"valu_fp64", "valu_fp32", "mfma_fp64", "mfma_fp32"}
    assert all(np.isfinite(v) and v > 0 for v in r.values())
