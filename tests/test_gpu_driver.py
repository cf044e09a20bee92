"""The lulesh-amd driver on the GPU: the reference's report for -s 30 run to
completion (932 cycles, final origin energy 2.025075e+05), the same
17-digit origin energy from the host set-up, the set-up in HBM and the
host-driven loop, and the quiet CSV line."""
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "lulesh-task-based-hpx_amd", "lulesh-amd")


def run(*args):
    print(f"\n[driver] {' '.join(args)}", flush=True)
    r = subprocess.run([EXE, *args], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    return r.stdout


def energy17(out):
    return re.search(r"Origin Energy \(17 digits\) = (\S+)", out).group(1)


def test_driver_s30_report_and_setups():
    if not os.path.exists(EXE):
        pytest.fail("lulesh-amd is not built")
    base = run("--s", "30")
    assert "Iteration count     =  932" in base
    assert re.search(r"Final Origin Energy =\s+2\.025075e\+05", base)
    dev = run("--s", "30", "--device-setup")
    host_loop = run("--s", "30", "--host-loop")
    assert energy17(dev) == energy17(base) == energy17(host_loop)
    assert "Iteration count     =  932" in dev and "Iteration count     =  932" in host_loop
    for key in ("MaxAbsDiff", "TotalAbsDiff", "MaxRelDiff"):
        a = re.search(key + r"\s+=\s+(\S+)", base).group(1)
        b = re.search(key + r"\s+=\s+(\S+)", dev).group(1)
        assert a == b, key


def test_driver_quiet_csv_and_mixed():
    q = run("--s", "20", "--i", "50", "--q", "--device-setup")
    fields = q.strip().split(",")
    assert fields[:4] == ["20", "11", "50", "1"] and float(fields[5]) > 0
    m = run("--s", "20", "--i", "50", "--precision", "mixed")
    assert "Iteration count     =  50" in m


def test_driver_fma_form():
    """--fma (fused multiply-adds, the reference build's arithmetic): the same
    report for -s 30, and the set-up in HBM gives the same 17 digits as the
    host set-up in that form."""
    f = run("--s", "30", "--fma")
    assert "Iteration count     =  932" in f
    assert re.search(r"Final Origin Energy =\s+2\.025075e\+05", f)
    assert "fused multiply-add" in f
    fd = run("--s", "30", "--fma", "--device-setup")
    assert energy17(fd) == energy17(f)
