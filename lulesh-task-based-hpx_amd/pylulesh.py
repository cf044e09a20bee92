"""ctypes binding of liblulesh_amd.so (include/lulesh_hip.h, include/lulesh_domain.h).

Used by the test suite and bench.py.  The library is the product: nothing
here computes physics; every call goes through the C-ABI, and a missing or
unloadable library is an error (there is no fallback path).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblulesh_amd.so")

OK, VOLUME_ERROR, QSTOP_ERROR, INVALID_ARGUMENT = 0, -1, -2, -3
FP64, MIXED = 0, 1
OPT_NODE_FORCES, OPT_GENERIC, OPT_FMA, OPT_NO_NODE_FORCES = 0x100, 0x200, 0x400, 0x800


def OPT_STREAMS(k):
    """Option splitting a batch over k streams (LULESH_OPT_STREAMS)."""
    return (int(k) & 0xF) << 12


PHASE_NODAL, PHASE_ELEMENTS, PHASE_MATERIAL, PHASE_CONSTRAINTS = 0, 1, 2, 3

FIELDS = ("x y z xd yd zd xdd ydd zdd fx fy fz nodalMass e p q ql qq v volo delv vdov "
          "arealg ss elemMass vnew delv_xi delv_eta delv_zeta delx_xi delx_eta "
          "delx_zeta").split()
FIELD_ID = {n: i for i, n in enumerate(FIELDS)}
NODE_FIELDS = set(FIELDS[:13])
# Q gradients: cycle temporaries, not stored by the fused elements + material
# pass (material_mode 2)
GRAD_FIELDS = set(FIELDS[-6:])

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int32)


class DomainView(ctypes.Structure):
    _fields_ = (
        [(n, ctypes.c_int32) for n in ("sizeX", "numElem", "numNode", "numReg", "cost", "pad0")]
        + [(n, _D) for n in "x y z xd yd zd xdd ydd zdd fx fy fz nodalMass".split()]
        + [(n, _I) for n in ("symmX", "symmY", "symmZ", "nodelist", "lxim", "lxip", "letam",
                             "letap", "lzetam", "lzetap", "elemBC")]
        + [(n, _D) for n in ("e p q ql qq v volo delv vdov arealg ss elemMass vnew delv_xi "
                             "delv_eta delv_zeta delx_xi delx_eta delx_zeta").split()]
        + [("regNumList", _I), ("regElemSize", _I), ("regElemlist", ctypes.POINTER(_I)),
           ("nodeElemStart", _I), ("nodeElemCornerList", _I)]
        + [(n, ctypes.c_double) for n in (
            "e_cut p_cut q_cut v_cut u_cut hgcoef ss4o3 qstop monoq_max_slope "
            "monoq_limiter_mult qlc_monoq qqc_monoq qqc eosvmax eosvmin pmin emin dvovmax "
            "refdens dtcourant dthydro dtfixed time deltatime deltatimemultlb "
            "deltatimemultub dtmax stoptime").split()]
        + [("cycle", ctypes.c_int32), ("pad1", ctypes.c_int32)]
    )


class TimeState(ctypes.Structure):
    _fields_ = [("time", ctypes.c_double), ("deltatime", ctypes.c_double),
                ("dtcourant", ctypes.c_double), ("dthydro", ctypes.c_double),
                ("cycle", ctypes.c_int32), ("status", ctypes.c_int32)]


class Tuning(ctypes.Structure):
    _fields_ = [("zplanes", ctypes.c_int32), ("block_elem", ctypes.c_int32),
                ("block_material", ctypes.c_int32), ("cycles_per_graph", ctypes.c_int32),
                ("material_mode", ctypes.c_int32), ("em_planes", ctypes.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib():
    """Load the library (raises if it is missing: there is no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    vp, i32, dbl = ctypes.c_void_p, ctypes.c_int32, ctypes.c_double
    sig = {
        "lulesh_domain_create": (vp, [i32, i32, i32, i32]),
        "lulesh_domain_destroy": (None, [vp]),
        "lulesh_domain_get_view": (ctypes.c_int, [vp, ctypes.POINTER(DomainView)]),
        "lulesh_domain_set_time_state": (ctypes.c_int, [vp, ctypes.POINTER(TimeState)]),
        "lulesh_domain_get_time_state": (ctypes.c_int, [vp, ctypes.POINTER(TimeState)]),
        "lulesh_time_increment": (None, [vp]),
        "lulesh_lagrange_leapfrog": (ctypes.c_int, [vp, vp]),
        "lulesh_verify_report": (None, [vp, _D]),
        "lulesh_region_rep": (i32, [i32, i32, i32]),
        "lulesh_hip_create": (ctypes.c_int, [ctypes.POINTER(DomainView), ctypes.c_int,
                                             ctypes.POINTER(vp)]),
        "lulesh_hip_leapfrog": (ctypes.c_int, [vp, dbl, _D, _D]),
        "lulesh_hip_run": (ctypes.c_int, [vp, i32, ctypes.POINTER(TimeState)]),
        "lulesh_hip_phase": (ctypes.c_int, [vp, ctypes.c_int, dbl]),
        "lulesh_hip_constraints_result": (ctypes.c_int, [vp, _D, _D]),
        "lulesh_hip_download": (ctypes.c_int, [vp, ctypes.c_int, vp, ctypes.c_size_t]),
        "lulesh_hip_upload": (ctypes.c_int, [vp, ctypes.c_int, vp, ctypes.c_size_t]),
        "lulesh_hip_download_view": (ctypes.c_int, [vp, ctypes.POINTER(DomainView)]),
        "lulesh_hip_set_timing": (ctypes.c_int, [vp, ctypes.c_int]),
        "lulesh_hip_kernel_times": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_char_p), _D,
                                                   ctypes.POINTER(ctypes.c_int64), ctypes.c_int]),
        "lulesh_hip_device_bytes": (ctypes.c_int64, [vp]),
        "lulesh_hip_last_error": (ctypes.c_char_p, [vp]),
        "lulesh_hip_destroy": (None, [vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    more = {
        "lulesh_hip_batch_create": (ctypes.c_int, [ctypes.POINTER(ctypes.POINTER(DomainView)),
                                                   i32, ctypes.c_int, ctypes.POINTER(vp)]),
        "lulesh_hip_batch_run": (ctypes.c_int, [vp, i32, ctypes.POINTER(TimeState)]),
        "lulesh_hip_batch_download": (ctypes.c_int, [vp, i32, ctypes.c_int, vp, ctypes.c_size_t]),
        "lulesh_hip_num_problems": (i32, [vp]),
        "lulesh_hip_last_run_ms": (dbl, [vp]),
        "lulesh_hip_is_structured": (i32, [vp]),
        "lulesh_hip_set_tuning": (ctypes.c_int, [vp, ctypes.POINTER(Tuning)]),
        "lulesh_hip_get_tuning": (ctypes.c_int, [vp, ctypes.POINTER(Tuning)]),
        "lulesh_hip_autotune": (ctypes.c_int, [vp, i32, ctypes.POINTER(Tuning), _D]),
        "lulesh_domain_scale_energy": (ctypes.c_int, [vp, dbl]),
        "lulesh_batch_params": (ctypes.c_int, [ctypes.c_uint64, i32, i32, i32, _I, _I, _D]),
        "lulesh_hip_create_sedov": (ctypes.c_int, [i32, i32, i32, i32, ctypes.c_int,
                                                   ctypes.POINTER(vp)]),
        "lulesh_hip_get_time_state": (ctypes.c_int, [vp, ctypes.POINTER(TimeState)]),
        "lulesh_hip_batch_create_sedov": (ctypes.c_int, [_I, _I, _D, i32, i32, i32, ctypes.c_int,
                                                         ctypes.POINTER(vp)]),
        "lulesh_hip_report": (ctypes.c_int, [vp, _D]),
        "lulesh_hip_stream_bandwidth": (ctypes.c_int, [ctypes.c_int64, i32, _D]),
        "lulesh_hip_fp64_fma_rate": (ctypes.c_int, [i32, _D]),
        "lulesh_domain_write_viz": (ctypes.c_int, [vp, ctypes.c_char_p, i32]),
    }
    for name, (res, args) in more.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


class LuleshError(RuntimeError):
    def __init__(self, rc, msg=""):
        super().__init__(f"lulesh error {rc}: {msg}")
        self.rc = rc


class Domain:
    """The host Domain (C++ class behind lulesh_domain.h)."""

    def __init__(self, s, nr=11, balance=1, cost=1):
        self.L = lib()
        self.h = self.L.lulesh_domain_create(s, nr, balance, cost)
        if not self.h:
            raise ValueError("lulesh_domain_create failed")
        self.s, self.nr, self.balance, self.cost = s, nr, balance, cost
        self.view = DomainView()
        self.L.lulesh_domain_get_view(self.h, ctypes.byref(self.view))

    def close(self):
        if getattr(self, "h", None):
            self.L.lulesh_domain_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @property
    def E(self):
        return self.view.numElem

    @property
    def N(self):
        return self.view.numNode

    def field(self, name):
        """Live numpy view of a Domain array."""
        v = self.view
        if name in FIELD_ID:
            n = self.N if name in NODE_FIELDS else self.E
            return np.ctypeslib.as_array(getattr(v, name), shape=(n,))
        lengths = {"nodelist": 8 * self.E, "nodeElemStart": self.N + 1,
                   "nodeElemCornerList": 8 * self.E, "symmX": (self.s + 1) ** 2,
                   "symmY": (self.s + 1) ** 2, "symmZ": (self.s + 1) ** 2}
        n = lengths.get(name, self.E)
        return np.ctypeslib.as_array(getattr(v, name), shape=(n,))

    def region(self, r):
        n = self.view.regElemSize[r]
        if n == 0:
            return np.zeros(0, np.int32)
        return np.ctypeslib.as_array(self.view.regElemlist[r], shape=(n,)).copy()

    def time_state(self):
        ts = TimeState()
        self.L.lulesh_domain_get_time_state(self.h, ctypes.byref(ts))
        return ts

    def set_time_state(self, ts):
        self.L.lulesh_domain_set_time_state(self.h, ctypes.byref(ts))

    def time_increment(self):
        self.L.lulesh_time_increment(self.h)

    def report(self):
        out = np.zeros(4)
        self.L.lulesh_verify_report(self.h, out.ctypes.data_as(_D))
        return out

    def refresh_view(self):
        self.L.lulesh_domain_get_view(self.h, ctypes.byref(self.view))

    def write_viz(self, directory, pieces=1):
        """The -v dump of the current host fields (lulesh_domain_write_viz)."""
        rc = self.L.lulesh_domain_write_viz(self.h, os.fsencode(directory), pieces)
        if rc != 0:
            raise LuleshError(rc, "lulesh_domain_write_viz")


class Device:
    """A device context (lulesh_hip_create) holding one Domain in HBM."""

    def __init__(self, dom, precision=FP64, opts=0):
        self.L = lib()
        self.dom = dom
        self.E, self.N = dom.E, dom.N
        h = ctypes.c_void_p()
        rc = self.L.lulesh_hip_create(ctypes.byref(dom.view), precision | opts, ctypes.byref(h))
        if rc != 0:
            raise LuleshError(rc, "lulesh_hip_create")
        self.h = h
        self.precision = precision

    @classmethod
    def sedov(cls, s, nr=11, balance=1, cost=1, precision=FP64, opts=0):
        """The Sedov problem built directly in HBM (lulesh_hip_create_sedov);
        no host Domain."""
        self = cls.__new__(cls)
        self.L = lib()
        self.dom = None
        self.E, self.N = s ** 3, (s + 1) ** 3
        h = ctypes.c_void_p()
        rc = self.L.lulesh_hip_create_sedov(s, nr, balance, cost, precision | opts,
                                             ctypes.byref(h))
        if rc != 0:
            raise LuleshError(rc, "lulesh_hip_create_sedov")
        self.h = h
        self.precision = precision
        return self

    def time_state(self):
        ts = TimeState()
        self._check(self.L.lulesh_hip_get_time_state(self.h, ctypes.byref(ts)), "time_state")
        return ts

    def report(self):
        out = np.zeros(4)
        self._check(self.L.lulesh_hip_report(self.h, out.ctypes.data_as(_D)), "report")
        return out

    def close(self):
        if getattr(self, "h", None):
            self.L.lulesh_hip_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, rc, what):
        if rc != 0:
            raise LuleshError(rc, f"{what}: {self.L.lulesh_hip_last_error(self.h).decode()}")

    def leapfrog(self, dt):
        c, h = ctypes.c_double(), ctypes.c_double()
        rc = self.L.lulesh_hip_leapfrog(self.h, dt, ctypes.byref(c), ctypes.byref(h))
        return rc, c.value, h.value

    def run(self, max_cycles, ts):
        return self.L.lulesh_hip_run(self.h, max_cycles, ctypes.byref(ts))

    def phase(self, ph, dt):
        return self.L.lulesh_hip_phase(self.h, ph, dt)

    def constraints_result(self):
        c, h = ctypes.c_double(), ctypes.c_double()
        self._check(self.L.lulesh_hip_constraints_result(self.h, ctypes.byref(c),
                                                         ctypes.byref(h)), "constraints")
        return c.value, h.value

    def download(self, name):
        n = self.N if name in NODE_FIELDS else self.E
        out = np.empty(n, np.float64)
        self._check(self.L.lulesh_hip_download(self.h, FIELD_ID[name], out.ctypes.data, n),
                    "download " + name)
        return out

    def upload(self, name, arr):
        arr = np.ascontiguousarray(arr, dtype=np.float64)
        self._check(self.L.lulesh_hip_upload(self.h, FIELD_ID[name], arr.ctypes.data, arr.size),
                    "upload " + name)

    def download_view(self):
        self._check(self.L.lulesh_hip_download_view(self.h, ctypes.byref(self.dom.view)),
                    "download_view")

    def set_timing(self, on):
        self.L.lulesh_hip_set_timing(self.h, 1 if on else 0)

    def kernel_times(self):
        cap = 16
        names = (ctypes.c_char_p * cap)()
        avg = (ctypes.c_double * cap)()
        cnt = (ctypes.c_int64 * cap)()
        n = self.L.lulesh_hip_kernel_times(self.h, names, avg, cnt, cap)
        return {names[i].decode(): (avg[i], cnt[i]) for i in range(n)}

    def device_bytes(self):
        return self.L.lulesh_hip_device_bytes(self.h)

    def structured(self):
        return bool(self.L.lulesh_hip_is_structured(self.h))

    def last_run_ms(self):
        return self.L.lulesh_hip_last_run_ms(self.h)

    def tuning(self):
        t = Tuning()
        self.L.lulesh_hip_get_tuning(self.h, ctypes.byref(t))
        return t

    def set_tuning(self, t):
        self._check(self.L.lulesh_hip_set_tuning(self.h, ctypes.byref(t)), "set_tuning")

    def autotune(self, cycles):
        t = Tuning()
        ms = ctypes.c_double()
        self._check(self.L.lulesh_hip_autotune(self.h, cycles, ctypes.byref(t), ctypes.byref(ms)),
                    "autotune")
        return t, ms.value


def batch_params(seed, nprob, smin=20, smax=80):
    """Seeded ensemble parameters (sizes, region counts, energy factors)."""
    L = lib()
    s = np.zeros(nprob, np.int32)
    r = np.zeros(nprob, np.int32)
    f = np.zeros(nprob, np.float64)
    rc = L.lulesh_batch_params(seed, nprob, smin, smax, s.ctypes.data_as(_I),
                               r.ctypes.data_as(_I), f.ctypes.data_as(_D))
    if rc != 0:
        raise LuleshError(rc, "lulesh_batch_params")
    return s, r, f


def make_domain(s, nr=11, factor=1.0, balance=1, cost=1):
    dom = Domain(s, nr, balance, cost)
    if factor != 1.0:
        rc = dom.L.lulesh_domain_scale_energy(dom.h, factor)
        if rc != 0:
            raise LuleshError(rc, "scale_energy")
        dom.refresh_view()
    return dom


class Batch:
    """Many Domains in one device context (lulesh_hip_batch_create)."""

    def __init__(self, doms, precision=FP64, opts=0):
        self.L = lib()
        self.doms = doms
        arr = (ctypes.POINTER(DomainView) * len(doms))(*[ctypes.pointer(d.view) for d in doms])
        h = ctypes.c_void_p()
        rc = self.L.lulesh_hip_batch_create(arr, len(doms), precision | opts, ctypes.byref(h))
        if rc != 0:
            raise LuleshError(rc, "lulesh_hip_batch_create")
        self.h = h
        self.sizes = [int(d.view.sizeX) for d in doms]

    @classmethod
    def sedov(cls, sizes, regions, factors, balance=1, cost=1, precision=FP64, opts=0):
        """The ensemble built directly in HBM (lulesh_hip_batch_create_sedov)."""
        self = cls.__new__(cls)
        self.L = lib()
        self.doms = None
        self.sizes = [int(x) for x in sizes]
        n = len(sizes)
        s = np.ascontiguousarray(sizes, np.int32)
        r = np.ascontiguousarray(regions, np.int32)
        f = np.ascontiguousarray(factors, np.float64)
        h = ctypes.c_void_p()
        rc = self.L.lulesh_hip_batch_create_sedov(s.ctypes.data_as(_I), r.ctypes.data_as(_I),
                                                   f.ctypes.data_as(_D), n, balance, cost,
                                                   precision | opts, ctypes.byref(h))
        if rc != 0:
            raise LuleshError(rc, "lulesh_hip_batch_create_sedov")
        self.h = h
        return self

    def close(self):
        if getattr(self, "h", None):
            self.L.lulesh_hip_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def run(self, max_cycles):
        states = (TimeState * len(self.sizes))()
        rc = self.L.lulesh_hip_batch_run(self.h, max_cycles, states)
        return rc, list(states)

    def download(self, prob, name):
        s = self.sizes[prob]
        n = (s + 1) ** 3 if name in NODE_FIELDS else s ** 3
        out = np.empty(n, np.float64)
        rc = self.L.lulesh_hip_batch_download(self.h, prob, FIELD_ID[name], out.ctypes.data, n)
        if rc != 0:
            raise LuleshError(rc, "batch_download")
        return out

    def last_run_ms(self):
        return self.L.lulesh_hip_last_run_ms(self.h)

    def tuning(self):
        t = Tuning()
        self.L.lulesh_hip_get_tuning(self.h, ctypes.byref(t))
        return t

    def set_tuning(self, t):
        rc = self.L.lulesh_hip_set_tuning(self.h, ctypes.byref(t))
        if rc != 0:
            raise LuleshError(rc, "set_tuning")

    def autotune(self, cycles):
        t = Tuning()
        ms = ctypes.c_double()
        rc = self.L.lulesh_hip_autotune(self.h, cycles, ctypes.byref(t), ctypes.byref(ms))
        if rc != 0:
            raise LuleshError(rc, "autotune")
        return t, ms.value


def run_device(s, its, nr=11, balance=1, cost=1, precision=FP64, opts=0):
    """Build a Domain, run `its` cycles (or to stoptime) on the device with
    the device-side time loop; returns (domain, device, time_state)."""
    dom = Domain(s, nr, balance, cost)
    dev = Device(dom, precision, opts)
    ts = dom.time_state()
    rc = dev.run(its, ts)
    if rc != 0:
        raise LuleshError(rc, "lulesh_hip_run")
    dev.download_view()
    dom.set_time_state(ts)
    return dom, dev, ts
