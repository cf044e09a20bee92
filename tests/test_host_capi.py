"""CPU checks of the product's host side and C-ABI surface (no GPU needed):
the library loads and exports every declared entry point, the host Domain
is bit-identical to the oracle's set-up, the host TimeIncrement and report
match the oracle, and argument errors are rejected before any device call.
"""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from oracle_lib import OracleDomain

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INT_FIELDS = ["nodelist", "lxim", "lxip", "letam", "letap", "lzetam", "lzetap", "elemBC",
              "regNumList", "symmX", "symmY", "symmZ", "nodeElemStart", "nodeElemCornerList"]


def declared_functions():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(re.findall(r"\b(lulesh_\w+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol(pyl):
    L = pyl.lib()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_headers_are_plain_c(tmp_path):
    """The boundary headers compile as strict C99 (plain pointers, sizes and
    structs: a C caller or an FFI generator can consume them) and as C++."""
    src = tmp_path / "abi.c"
    src.write_text('#include "lulesh_hip.h"\n#include "lulesh_domain.h"\n'
                   "int main(void) {\n"
                   "  return sizeof(lulesh_tuning) + sizeof(lulesh_time_state) +\n"
                   "         sizeof(lulesh_domain_view) == 0;\n}\n")
    inc = os.path.join(ROOT, "include")
    for cmd in (["gcc", "-std=c99", "-pedantic", "-Wall", "-Wextra", "-Werror"],
                ["g++", "-std=c++17", "-x", "c++", "-pedantic", "-Wall", "-Wextra", "-Werror"]):
        r = subprocess.run(cmd + ["-I", inc, "-c", str(src), "-o", str(tmp_path / "abi.o")],
                           capture_output=True, text=True, timeout=60)
        assert r.returncode == 0, (cmd[0], r.stderr)


@pytest.mark.parametrize("s,nr,b,c", [(1, 11, 1, 1), (2, 11, 1, 1), (5, 11, 1, 1), (9, 16, 1, 2),
                                      (12, 21, 1, 1), (7, 5, 2, 0), (6, 1, 1, 1), (30, 11, 1, 1)])
def test_host_domain_matches_oracle(pyl, s, nr, b, c):
    dom = pyl.Domain(s, nr, b, c)
    ora = OracleDomain(s, nr, b, c)
    for f in pyl.FIELDS:
        np.testing.assert_array_equal(dom.field(f), ora.f64(f), err_msg=f)
    for f in INT_FIELDS:
        np.testing.assert_array_equal(dom.field(f), ora.i32(f), err_msg=f)
    for r in range(nr):
        np.testing.assert_array_equal(dom.region(r), ora.region(r))
        assert pyl.lib().lulesh_region_rep(r, nr, c) == ora.rep(r)
    v = dom.view
    for k in ("deltatime", "time", "stoptime", "dtcourant", "dthydro", "dtmax", "dtfixed",
              "hgcoef", "qqc", "eosvmin", "eosvmax", "emin", "qlc_monoq", "qqc_monoq"):
        assert getattr(v, k) == ora.scalar(k), k


def test_host_time_increment_matches_oracle(pyl):
    rng = np.random.default_rng(0)
    dom = pyl.Domain(2)
    for trial in range(300):
        ora = OracleDomain(2)
        st = pyl.TimeState()
        st.time = float(rng.uniform(0, 1.2e-2))
        st.deltatime = float(10 ** rng.uniform(-6, -2))
        st.dtcourant = float(10 ** rng.uniform(-7, -1)) if rng.random() < 0.8 else 1e20
        st.dthydro = float(10 ** rng.uniform(-7, -1)) if rng.random() < 0.8 else 1e20
        st.cycle = int(rng.integers(0, 3))
        dom.set_time_state(st)
        for k in ("time", "deltatime", "dtcourant", "dthydro"):
            ora.set_scalar(k, getattr(st, k))
        ora.cycle = st.cycle
        dom.time_increment()
        ora.time_increment()
        got = dom.time_state()
        assert got.deltatime == ora.scalar("deltatime"), trial
        assert got.time == ora.scalar("time"), trial
        assert got.cycle == ora.cycle


def test_host_report_matches_oracle(pyl):
    rng = np.random.default_rng(1)
    s = 9
    dom = pyl.Domain(s)
    ora = OracleDomain(s)
    e = rng.standard_normal(s ** 3) * 1e3
    e[rng.integers(0, s ** 3, 50)] = 0.0
    dom.field("e")[:] = e
    ora.f64("e")[:] = e
    np.testing.assert_array_equal(dom.report(), ora.report())


def test_create_rejects_bad_views_without_device(pyl):
    L = pyl.lib()
    h = ctypes.c_void_p()
    assert L.lulesh_hip_create(None, 0, ctypes.byref(h)) == pyl.INVALID_ARGUMENT
    dom = pyl.Domain(3)
    bad = pyl.DomainView.from_buffer_copy(dom.view)
    bad.numElem = 26
    assert L.lulesh_hip_create(ctypes.byref(bad), 0, ctypes.byref(h)) == pyl.INVALID_ARGUMENT
    assert L.lulesh_hip_create(ctypes.byref(dom.view), 7, ctypes.byref(h)) == pyl.INVALID_ARGUMENT
    assert L.lulesh_domain_create(0, 11, 1, 1) is None
    assert L.lulesh_hip_leapfrog(None, 1e-6, None, None) == pyl.INVALID_ARGUMENT
    assert L.lulesh_hip_run(None, 10, None) == pyl.INVALID_ARGUMENT
    # device-side set-up: argument checks come before any device call
    for args in [(0, 11, 1, 1, 0), (10, 0, 1, 1, 0), (10, 11, -1, 1, 0), (10, 11, 1, -1, 0),
                 (10, 11, 1, 1, pyl.OPT_GENERIC), (10, 11, 1, 1, 7), (700, 11, 1, 1, 0),
                 (646, 11, 1, 1, 0),
                 (10, 21, 1, 100000, 0)]:
        assert L.lulesh_hip_create_sedov(*args, ctypes.byref(h)) == pyl.INVALID_ARGUMENT, args
    assert L.lulesh_hip_create_sedov(10, 11, 1, 1, 0, None) == pyl.INVALID_ARGUMENT
    assert L.lulesh_hip_get_time_state(None, None) == pyl.INVALID_ARGUMENT
    assert L.lulesh_hip_report(None, None) == pyl.INVALID_ARGUMENT


def test_create_rejects_region_lists_that_are_not_a_partition(pyl):
    """Region lists with a duplicated, missing or out-of-range element are
    refused before any device call (the material kernels need every element
    in exactly one region, as lulesh-init.cc:341-445 builds them)."""
    L = pyl.lib()
    h = ctypes.c_void_p()
    dom = pyl.Domain(4, 11)
    sizes = [int(dom.view.regElemSize[r]) for r in range(11)]
    r = max(range(11), key=lambda k: sizes[k])  # a region with at least two elements
    lst = dom.view.regElemlist[r]
    keep0, keep1 = lst[0], lst[1]
    try:
        lst[1] = keep0                       # duplicate: element keep1 has no region
        assert L.lulesh_hip_create(ctypes.byref(dom.view), 0, ctypes.byref(h)) == \
            pyl.INVALID_ARGUMENT
        lst[1] = 64                          # out of range (E = 64)
        assert L.lulesh_hip_create(ctypes.byref(dom.view), 0, ctypes.byref(h)) == \
            pyl.INVALID_ARGUMENT
        lst[1] = keep1
        dom.view.regElemSize[r] = sizes[r] - 1  # element keep-last missing
        assert L.lulesh_hip_create(ctypes.byref(dom.view), 0, ctypes.byref(h)) == \
            pyl.INVALID_ARGUMENT
    finally:
        lst[1] = keep1
        dom.view.regElemSize[r] = sizes[r]
    # keeping and not keeping the node forces at once
    both = pyl.OPT_NODE_FORCES | pyl.OPT_NO_NODE_FORCES
    assert L.lulesh_hip_create(ctypes.byref(dom.view), both, ctypes.byref(h)) == \
        pyl.INVALID_ARGUMENT
    # not keeping them is an option of lulesh_hip_create only (off already elsewhere)
    arr = (ctypes.POINTER(pyl.DomainView) * 1)(ctypes.pointer(dom.view))
    assert L.lulesh_hip_batch_create(arr, 1, pyl.OPT_NO_NODE_FORCES, ctypes.byref(h)) == \
        pyl.INVALID_ARGUMENT


def test_driver_flag_errors():
    exe = os.path.join(ROOT, "lulesh-task-based-hpx_amd", "lulesh-amd")
    if not os.path.exists(exe):
        pytest.skip("driver not built")
    r = subprocess.run([exe, "--bogus"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 255 and "usage" in r.stderr
    r = subprocess.run([exe, "--precision", "fp8"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 255
    r = subprocess.run([exe, "--device-setup", "--host-loop"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 255 and "host set-up" in r.stderr


def test_null_context_rejected_by_every_entry_point(pyl):
    """Every C-ABI entry point taking a context rejects NULL before any
    device call (the forwarding layer hands NULL to the default form)."""
    L = pyl.lib()
    inv = pyl.INVALID_ARGUMENT
    ts = pyl.TimeState()
    t = pyl.Tuning()
    buf = (ctypes.c_double * 4)()
    assert L.lulesh_hip_leapfrog(None, 1e-6, buf, buf) == inv
    assert L.lulesh_hip_run(None, 10, ctypes.byref(ts)) == inv
    assert L.lulesh_hip_phase(None, 0, 1e-6) == inv
    assert L.lulesh_hip_constraints_result(None, buf, buf) == inv
    assert L.lulesh_hip_download(None, 0, buf, 4) == inv
    assert L.lulesh_hip_upload(None, 0, buf, 4) == inv
    assert L.lulesh_hip_download_view(None, None) == inv
    assert L.lulesh_hip_set_timing(None, 1) == inv
    assert L.lulesh_hip_kernel_times(None, None, None, None, 0) == inv
    assert L.lulesh_hip_set_tuning(None, ctypes.byref(t)) == inv
    assert L.lulesh_hip_get_tuning(None, ctypes.byref(t)) == inv
    assert L.lulesh_hip_autotune(None, 4, ctypes.byref(t), buf) == inv
    assert L.lulesh_hip_batch_run(None, 10, ctypes.byref(ts)) == inv
    assert L.lulesh_hip_batch_download(None, 0, 0, buf, 4) == inv
    assert L.lulesh_hip_get_time_state(None, ctypes.byref(ts)) == inv
    assert L.lulesh_hip_report(None, buf) == inv
    assert L.lulesh_hip_device_bytes(None) == 0
    assert L.lulesh_hip_num_problems(None) == 0
    assert L.lulesh_hip_is_structured(None) == 0
    assert L.lulesh_hip_last_run_ms(None) == 0.0
    assert L.lulesh_hip_last_error(None) == b"null context"
    L.lulesh_hip_destroy(None)


def test_fma_form_validates_like_the_default_form(pyl):
    """LULESH_OPT_FMA selects the fused form; its argument checks are the
    default form's (no device call is made for invalid arguments)."""
    L = pyl.lib()
    h = ctypes.c_void_p()
    F = pyl.OPT_FMA
    assert L.lulesh_hip_create(None, F, ctypes.byref(h)) == pyl.INVALID_ARGUMENT
    dom = pyl.Domain(3)
    assert L.lulesh_hip_create(ctypes.byref(dom.view), 7 | F, ctypes.byref(h)) == \
        pyl.INVALID_ARGUMENT
    assert L.lulesh_hip_create(ctypes.byref(dom.view), F, None) == pyl.INVALID_ARGUMENT
    for args in [(0, 11, 1, 1, F), (10, 11, 1, 1, pyl.OPT_GENERIC | F), (700, 11, 1, 1, F)]:
        assert L.lulesh_hip_create_sedov(*args, ctypes.byref(h)) == pyl.INVALID_ARGUMENT, args
    assert h.value is None
