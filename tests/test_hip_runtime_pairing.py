"""One HIP runtime per process for the libtorch Function. torch ships its own libamdhip64;
when libadsgpu.so is loaded before torch, torch maps its copy beside the one the library
uses. heat_k / heat_kf (Python and the C++ extension) must refuse that process before any
tensor pointer reaches the library. Runs on the CPU: loading the libraries needs no GPU."""
import os
import subprocess
import sys

import pytest

import adsgpu as ag

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "iga-ads_amd")

MAPS_ONE = """\
7f0000000000-7f0000100000 r--p 00000000 08:01 1234 /opt/rocm-7.2.0/lib/libamdhip64.so.7.2.70200
7f0000100000-7f0000800000 r-xp 00100000 08:01 1234 /opt/rocm-7.2.0/lib/libamdhip64.so.7.2.70200
7f0000900000-7f0000a00000 r--p 00000000 08:01 99 /usr/lib/x86_64-linux-gnu/libc.so.6
7ffd00000000-7ffd00021000 rw-p 00000000 00:00 0 [stack]
"""
MAPS_TWO = MAPS_ONE + (
    "7f1000000000-7f1000100000 r--p 00000000 08:01 777 "
    "/usr/local/lib/python3.10/dist-packages/torch/lib/libamdhip64.so\n")


def test_parse_maps():
    assert ag.hip_runtimes(MAPS_ONE) == {os.path.realpath("/opt/rocm-7.2.0/lib/libamdhip64.so.7.2.70200")}
    assert len(ag.hip_runtimes(MAPS_TWO)) == 2
    ag.check_single_hip_runtime(MAPS_ONE)
    with pytest.raises(ag.AdsgError, match="two HIP runtimes"):
        ag.check_single_hip_runtime(MAPS_TWO)
    assert ag.hip_runtimes("") == set()


def _have_torch_ext():
    return any(f.startswith("adsg_torch") and f.endswith(".so") for f in os.listdir(os.path.join(PKG, "adsgpu")))


SCRIPT = r"""
import sys
sys.path.insert(0, {pkg!r})
order = {order!r}
if order == "torch_first":
    import torch
import adsgpu
adsgpu.lib()
import torch
from adsgpu import adsg_torch
print("runtimes", len(adsgpu.hip_runtimes()))
try:
    adsgpu.check_single_hip_runtime()
    print("python: ok")
except adsgpu.AdsgError as e:
    print("python: refused:", e)
k = torch.zeros(4, dtype=torch.float64)
try:
    adsg_torch.heat_k(k, k, 0, 1, 1, 1e-6, True)
except RuntimeError as e:
    print("cpp:", str(e).splitlines()[0])
"""


@pytest.mark.skipif(not _have_torch_ext(), reason="libtorch extension not built")
@pytest.mark.parametrize("order", ["torch_first", "library_first"])
def test_import_order(order):
    r = subprocess.run([sys.executable, "-c", SCRIPT.format(pkg=PKG, order=order)],
                       capture_output=True, text=True, timeout=600)
    out = r.stdout
    assert r.returncode == 0, r.stderr[-2000:]
    if order == "torch_first":
        assert "runtimes 1" in out and "python: ok" in out, out
        # past the runtime check, the Function rejects the null handle itself
        assert "cpp: bad handle" in out, out
    else:
        assert "runtimes 2" in out and "python: refused" in out, out
        assert "cpp: two HIP runtimes" in out, out
