"""lulesh_hip_constraints_noalloc (no context, no allocation: the partial
minima sit in a fixed device array of the library) against the constraints
phase of a context (lulesh_hip_phase(CONSTRAINTS) + constraints_result, whose
slots the context allocates), on the same fields: two calls in a row, the
second over a smaller mesh, each bit for bit (minima are exact).  Device
memory for the inputs and the output comes from hipMalloc of the HIP runtime
the library itself uses."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def context_minima(pyl, s, cycles):
    """A separately rounded context advanced `cycles` cycles; its fields and
    the minima of its constraints phase."""
    dom = pyl.Domain(s, 11)
    dev = pyl.Device(dom)
    ts = dom.time_state()
    assert dev.run(cycles, ts) == 0
    f = {k: dev.download(k) for k in ("ss", "vdov", "arealg")}
    assert dev.phase(pyl.PHASE_CONSTRAINTS, ts.deltatime) == 0
    dtc, dth = dev.constraints_result()
    qqc, dvovmax = dom.view.qqc, dom.view.dvovmax
    dev.close()
    return f, dtc, dth, qqc, dvovmax


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime the library loaded."""

    def __init__(self):
        self.h = None
        for name in ("libamdhip64.so.7", "libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                self.h = ctypes.CDLL(name)
                break
            except OSError:
                continue
        assert self.h is not None, "HIP runtime not found"
        self.h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                     ctypes.c_int]
        self.h.hipFree.argtypes = [ctypes.c_void_p]
        self.bufs = []

    def upload(self, a):
        a = np.ascontiguousarray(a, np.float64)
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), max(a.nbytes, 8)) == 0
        self.bufs.append(p)
        assert self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def download(self, p, n):
        out = np.empty(n, np.float64)
        assert self.h.hipDeviceSynchronize() == 0
        assert self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


def test_constraints_noalloc_equals_context_phase(pyl):
    L = pyl.lib()
    fn = L.lulesh_hip_constraints_noalloc
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_void_p, ctypes.c_void_p]
    hip = Hip()
    try:
        out = hip.upload(np.full(2, -1.0))
        for s, cycles in ((24, 60), (9, 45)):  # the second call has fewer elements
            print(f"[noalloc] s={s} after {cycles} cycles", flush=True)
            f, dtc, dth, qqc, dvovmax = context_minima(pyl, s, cycles)
            assert dtc < 1e20 and dth < 1e20  # the blast constrains both
            d = {k: hip.upload(v) for k, v in f.items()}
            rc = fn(d["ss"], d["vdov"], d["arealg"], s ** 3, qqc, dvovmax, out, None)
            assert rc == 0
            got = hip.download(out, 2)
            print(f"[noalloc] context {dtc!r} {dth!r}, noalloc {got[0]!r} {got[1]!r}", flush=True)
            assert got[0] == dtc and got[1] == dth
        # no element: both minima stay 1e20; bad arguments are refused
        assert fn(None, None, None, 0, 2.0, 0.1, out, None) == 0
        assert hip.download(out, 2).tolist() == [1e20, 1e20]
        assert fn(None, None, None, 5, 2.0, 0.1, out, None) == pyl.INVALID_ARGUMENT
        assert fn(None, None, None, -1, 2.0, 0.1, out, None) == pyl.INVALID_ARGUMENT
        assert fn(None, None, None, 0, 2.0, 0.1, None, None) == pyl.INVALID_ARGUMENT
    finally:
        hip.free()
