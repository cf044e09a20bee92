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


def test_driver_viz_dump(tmp_path):
    """--v writes the final state for VisIt (lulesh.cc:2397-2400) in the
    working directory: the host set-up and the set-up in HBM give the same
    file (region numbers included), whole or split by --f; its energies are
    the run's (the 17-digit origin energy of the report, as float)."""
    from test_host_viz import read_vtk

    def run_in(d, *args):
        print(f"\n[driver] (in {d.name}) {' '.join(args)}", flush=True)
        r = subprocess.run([EXE, *args], capture_output=True, text=True, timeout=240, cwd=d)
        assert r.returncode == 0, r.stderr
        return r.stdout

    a, b, c = (tmp_path / n for n in ("host", "device", "split"))
    for d in (a, b, c):
        d.mkdir()
    out = run_in(a, "--s", "6", "--i", "40", "--v")
    run_in(b, "--s", "6", "--i", "40", "--v", "--device-setup")
    run_in(c, "--s", "6", "--i", "40", "--v", "--f", "3")
    assert os.listdir(a) == ["lulesh_plot_c40.vtk"]
    va = (a / "lulesh_plot_c40.vtk").read_bytes()
    assert (b / "lulesh_plot_c40.vtk").read_bytes() == va
    whole = read_vtk(a / "lulesh_plot_c40.vtk")
    assert whole["fields"]["CYCLE"][0] == 40 and whole["cells"].shape == (216, 9)
    e0 = float(energy17(out))
    assert whole["cell"]["e"][0] == pytest.approx(e0, rel=1e-7)
    assert len(set(whole["cell"]["regions"].tolist())) > 1
    assert (c / "lulesh_plot_c40.visit").read_text().startswith("!NBLOCKS 3\n")
    parts = [read_vtk(c / f"lulesh_plot_c40.{k:03d}.vtk") for k in range(3)]
    np_e = [p["cell"]["e"] for p in parts]
    import numpy as np
    np.testing.assert_array_equal(np.concatenate(np_e), whole["cell"]["e"])
