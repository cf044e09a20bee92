"""ctypes binding of the CPU oracle (oracle/liblulesh_oracle.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it, as the checker.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liblulesh_oracle.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle library missing: run `make -C oracle`")
        L = ctypes.CDLL(ORACLE_SO)
        vp, c_int, c_dbl = ctypes.c_void_p, ctypes.c_int32, ctypes.c_double
        dp = ctypes.POINTER(ctypes.c_double)
        L.ora_create.restype = vp
        L.ora_create.argtypes = [c_int, c_int, c_int, c_int]
        L.ora_destroy.argtypes = [vp]
        L.ora_field_f64.restype = dp
        L.ora_field_f64.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64)]
        L.ora_field_i32.restype = ctypes.POINTER(ctypes.c_int32)
        L.ora_field_i32.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64)]
        L.ora_get_scalar.restype = c_dbl
        L.ora_get_scalar.argtypes = [vp, ctypes.c_char_p]
        L.ora_set_scalar.argtypes = [vp, ctypes.c_char_p, c_dbl]
        L.ora_cycle.restype = c_int
        L.ora_cycle.argtypes = [vp]
        L.ora_set_cycle.argtypes = [vp, c_int]
        L.ora_num_reg.restype = c_int
        L.ora_num_reg.argtypes = [vp]
        L.ora_reg_elem_size.restype = c_int
        L.ora_reg_elem_size.argtypes = [vp, c_int]
        L.ora_reg_elem_list.restype = ctypes.POINTER(ctypes.c_int32)
        L.ora_reg_elem_list.argtypes = [vp, c_int]
        L.ora_reg_rep.restype = c_int
        L.ora_reg_rep.argtypes = [vp, c_int]
        L.ora_set_threads.argtypes = [c_int]
        L.ora_scale_energy.argtypes = [vp, c_dbl]
        L.ora_max_threads.restype = c_int
        L.ora_time_increment.argtypes = [vp]
        for fn in ("ora_lagrange_nodal", "ora_lagrange_elements",
                   "ora_apply_material_properties", "ora_lagrange_leapfrog"):
            getattr(L, fn).restype = ctypes.c_int
            getattr(L, fn).argtypes = [vp]
        L.ora_calc_corner_forces.restype = ctypes.c_int
        L.ora_calc_corner_forces.argtypes = [vp] + [dp] * 6
        L.ora_calc_time_constraints.argtypes = [vp]
        L.ora_run.restype = ctypes.c_int
        L.ora_run.argtypes = [vp, c_int]
        L.ora_report.argtypes = [vp, dp]
        L.ora_calc_elem_volume.restype = c_dbl
        L.ora_calc_elem_volume.argtypes = [dp, dp, dp]
        L.ora_calc_elem_shape_function_derivatives.argtypes = [dp, dp, dp, dp, dp]
        L.ora_calc_elem_node_normals.argtypes = [dp, dp, dp, dp]
        L.ora_calc_elem_volume_derivative.argtypes = [dp, dp, dp, dp]
        L.ora_calc_elem_characteristic_length.restype = c_dbl
        L.ora_calc_elem_characteristic_length.argtypes = [dp, dp, dp, c_dbl]
        L.ora_calc_elem_velocity_gradient.argtypes = [dp, dp, dp, dp, c_dbl, dp]
        L.ora_calc_elem_fb_hourglass.argtypes = [dp] * 6 + [c_dbl] * 4 + [dp]
        L.ora_eval_eos_elem.argtypes = [dp, dp]
        _lib = L
    return _lib


def dptr(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


F64_FIELDS = ("x y z xd yd zd xdd ydd zdd fx fy fz nodalMass e p q ql qq v volo delv "
              "vdov arealg ss elemMass vnew delv_xi delv_eta delv_zeta delx_xi delx_eta "
              "delx_zeta").split()
I32_FIELDS = ("nodelist lxim lxip letam letap lzetam lzetap elemBC regNumList symmX symmY "
              "symmZ nodeElemStart nodeElemCornerList").split()


class OracleDomain:
    """A Domain of the oracle; field arrays are live numpy views."""

    def __init__(self, s, nr=11, balance=1, cost=1):
        self.L = lib()
        self.h = self.L.ora_create(s, nr, balance, cost)
        if not self.h:
            raise ValueError("ora_create failed")
        self.s = s

    def scale_energy(self, factor):
        self.L.ora_scale_energy(self.h, float(factor))

    def close(self):
        if self.h:
            self.L.ora_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def f64(self, name):
        n = ctypes.c_int64()
        p = self.L.ora_field_f64(self.h, name.encode(), ctypes.byref(n))
        if not p:
            raise KeyError(name)
        return np.ctypeslib.as_array(p, shape=(n.value,))

    def i32(self, name):
        n = ctypes.c_int64()
        p = self.L.ora_field_i32(self.h, name.encode(), ctypes.byref(n))
        if not p:
            raise KeyError(name)
        return np.ctypeslib.as_array(p, shape=(n.value,))

    def field(self, name):
        return self.f64(name) if name in F64_FIELDS else self.i32(name)

    def scalar(self, name):
        return self.L.ora_get_scalar(self.h, name.encode())

    def set_scalar(self, name, v):
        self.L.ora_set_scalar(self.h, name.encode(), float(v))

    @property
    def cycle(self):
        return self.L.ora_cycle(self.h)

    @cycle.setter
    def cycle(self, c):
        self.L.ora_set_cycle(self.h, c)

    @property
    def num_reg(self):
        return self.L.ora_num_reg(self.h)

    def region(self, r):
        n = self.L.ora_reg_elem_size(self.h, r)
        if n == 0:
            return np.zeros(0, np.int32)
        return np.ctypeslib.as_array(self.L.ora_reg_elem_list(self.h, r), shape=(n,)).copy()

    def rep(self, r):
        return self.L.ora_reg_rep(self.h, r)

    def time_increment(self):
        self.L.ora_time_increment(self.h)

    def leapfrog(self):
        return self.L.ora_lagrange_leapfrog(self.h)

    def nodal(self):
        return self.L.ora_lagrange_nodal(self.h)

    def elements(self):
        return self.L.ora_lagrange_elements(self.h)

    def material(self):
        return self.L.ora_apply_material_properties(self.h)

    def constraints(self):
        self.L.ora_calc_time_constraints(self.h)

    def corner_forces(self):
        E = self.s ** 3
        arrs = [np.zeros(8 * E) for _ in range(6)]
        rc = self.L.ora_calc_corner_forces(self.h, *[dptr(a) for a in arrs])
        return rc, arrs

    def run(self, its):
        return self.L.ora_run(self.h, its)

    def report(self):
        out = np.zeros(4)
        self.L.ora_report(self.h, dptr(out))
        return out


def set_threads(n):
    lib().ora_set_threads(n)
