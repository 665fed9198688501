"""ctypes binding of the FP64 CPU oracle (librefcpu.so).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, as the checker or the timed CPU baseline.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from crashcat_amd import _abi

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "librefcpu.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int32)
        L.oc_create.restype = C.c_void_p
        L.oc_create.argtypes = [C.POINTER(_abi.Settings), C.POINTER(_abi.BodyDesc), C.c_int32,
                                C.POINTER(_abi.ShapeDesc), C.c_int32, C.POINTER(_abi.HullData),
                                C.POINTER(_abi.PointConstraintDesc), C.c_int32]
        L.oc_destroy.argtypes = [C.c_void_p]
        L.oc_step.argtypes = [C.c_void_p, C.c_double]
        L.oc_get_bodies.argtypes = [C.c_void_p, C.POINTER(_abi.BodyState)]
        L.oc_set_bodies.argtypes = [C.c_void_p, C.POINTER(_abi.BodyState)]
        L.oc_get_pairs.restype = C.c_int32
        L.oc_get_pairs.argtypes = [C.c_void_p, ip, C.c_int32]
        L.oc_get_contacts.restype = C.c_int32
        L.oc_get_contacts.argtypes = [C.c_void_p, C.POINTER(_abi.ContactOut), C.c_int32]
        L.oc_get_counters.argtypes = [C.c_void_p, C.POINTER(_abi.StepCounters)]
        L.oc_broadphase_aabbs.restype = C.c_int32
        L.oc_broadphase_aabbs.argtypes = [C.POINTER(C.c_float), C.POINTER(C.c_float), ip, ip, ip, C.c_int32,
                                          C.c_double, C.POINTER(_abi.Settings), ip, C.c_int32]
        _LIB = L
    return _LIB


BODY_STATE_DTYPE = np.dtype(_abi.BodyState)
CONTACT_DTYPE = np.dtype(_abi.ContactOut)


class OracleWorld:
    """One FP64 oracle world built from a Scene (one world of it)."""

    def __init__(self, scene, world=0):
        L = lib()
        offs = scene.offsets()
        b0, b1 = offs[world], offs[world + 1]
        self.b0 = b0
        bodies = scene.bodies()[b0:b1].copy()
        self._bodies = bodies
        shapes, hulls, self._keep = scene.c_shapes()
        pcs = [(a - b0, b - b0, la, lb) for (a, b, la, lb) in scene.point_constraints if b0 <= a < b1]
        pa = (_abi.PointConstraintDesc * max(len(pcs), 1))()
        for i, (a, b, la, lb) in enumerate(pcs):
            pa[i].body_a, pa[i].body_b, pa[i].enabled = a, b, 1
            pa[i].local_point_a[:] = la
            pa[i].local_point_b[:] = lb
        self._pcs = pa
        self._settings = scene.settings
        self.n = b1 - b0
        self.h = L.oc_create(C.byref(scene.settings), bodies.ctypes.data_as(C.POINTER(_abi.BodyDesc)), self.n,
                             shapes, len(scene.shapes), C.byref(hulls), pa, len(pcs))

    def __del__(self):
        try:
            lib().oc_destroy(self.h)
        except Exception:
            pass

    def step(self, dt=1.0 / 60.0, n=1):
        for _ in range(n):
            lib().oc_step(self.h, dt)

    def bodies(self):
        out = np.zeros(self.n, dtype=BODY_STATE_DTYPE)
        lib().oc_get_bodies(self.h, out.ctypes.data_as(C.POINTER(_abi.BodyState)))
        return out

    def set_bodies(self, states):
        st = np.ascontiguousarray(states, dtype=BODY_STATE_DTYPE)
        lib().oc_set_bodies(self.h, st.ctypes.data_as(C.POINTER(_abi.BodyState)))

    def pairs(self):
        n = lib().oc_get_pairs(self.h, None, 0)
        out = np.zeros((max(n, 1), 2), dtype=np.int32)
        lib().oc_get_pairs(self.h, out.ctypes.data_as(C.POINTER(C.c_int32)), n)
        return out[:n]

    def contacts(self):
        n = lib().oc_get_contacts(self.h, None, 0)
        out = np.zeros(max(n, 1), dtype=CONTACT_DTYPE)
        lib().oc_get_contacts(self.h, out.ctypes.data_as(C.POINTER(_abi.ContactOut)), n)
        return out[:n]

    def counters(self):
        c = _abi.StepCounters()
        lib().oc_get_counters(self.h, C.byref(c))
        return {k: getattr(c, k) for k, _ in _abi.StepCounters._fields_}


def broadphase_aabbs(mn, mx, motion, sleeping, layer, margin, settings):
    L = lib()
    mn = np.ascontiguousarray(mn, dtype=np.float32)
    mx = np.ascontiguousarray(mx, dtype=np.float32)
    ip = C.POINTER(C.c_int32)
    mo = np.ascontiguousarray(motion, dtype=np.int32)
    sl = np.ascontiguousarray(sleeping, dtype=np.int32)
    la = np.ascontiguousarray(layer, dtype=np.int32)
    n = len(mo)
    fp = C.POINTER(C.c_float)
    cnt = L.oc_broadphase_aabbs(mn.ctypes.data_as(fp), mx.ctypes.data_as(fp), mo.ctypes.data_as(ip),
                                sl.ctypes.data_as(ip), la.ctypes.data_as(ip), n, margin, C.byref(settings), None, 0)
    out = np.zeros((max(cnt, 1), 2), dtype=np.int32)
    L.oc_broadphase_aabbs(mn.ctypes.data_as(fp), mx.ctypes.data_as(fp), mo.ctypes.data_as(ip),
                          sl.ctypes.data_as(ip), la.ctypes.data_as(ip), n, margin, C.byref(settings),
                          out.ctypes.data_as(ip), cnt)
    return out[:cnt]
