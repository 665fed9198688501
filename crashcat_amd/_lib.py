"""ctypes binding of libcrashcat_amd.so (the HIP path).  There is no CPU fallback:
if the library or a GPU is missing, every entry point raises."""
import ctypes as C
import os

import numpy as np

from . import _abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcrashcat_amd.so")
_LIB = None

SYMBOLS = [
    "cc_create", "cc_destroy", "cc_last_error", "cc_upload_scene", "cc_step", "cc_sync", "cc_download_bodies",
    "cc_set_bodies", "cc_download_pairs", "cc_download_contacts", "cc_get_counters", "cc_broadphase_aabbs",
    "cc_get_phase_times", "cc_get_totals", "cc_get_solver_residual", "cc_build_triangle_mesh", "cc_download_removed_contacts", "cc_download_ccd_bodies",
    "cc_snapshot", "cc_restore", "cc_add_forces", "cc_download_removed_contacts_noalloc", "cc_fwd", "cc_bwd",
    "cc_set_contact_cache", "cc_download_contact_cache", "cc_set_sleep_state", "cc_set_point_impulses",
]


class CrashcatError(RuntimeError):
    pass


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise CrashcatError(f"{LIB_PATH} is not built; run crashcat_amd/build.py (hipcc, gfx950)")
    L = C.CDLL(LIB_PATH)
    ip = C.POINTER(C.c_int32)
    vp = C.c_void_p
    L.cc_create.argtypes = [C.POINTER(vp), C.POINTER(_abi.Config)]
    L.cc_destroy.argtypes = [vp]
    L.cc_last_error.argtypes = [vp]
    L.cc_last_error.restype = C.c_char_p
    L.cc_upload_scene.argtypes = [vp, C.POINTER(_abi.Settings), C.c_int32, ip, C.POINTER(_abi.BodyDesc), C.c_int32,
                                  C.POINTER(_abi.ShapeDesc), C.c_int32, C.POINTER(_abi.HullData),
                                  C.POINTER(_abi.PointConstraintDesc), C.c_int32, C.POINTER(_abi.MeshData)]
    dp = C.POINTER(C.c_double)
    L.cc_build_triangle_mesh.argtypes = [dp, C.c_int32, ip, C.c_int32, C.c_double, C.c_double, C.c_int32, dp, ip, dp, ip]
    L.cc_step.argtypes = [vp, C.c_double, C.c_int32]
    L.cc_sync.argtypes = [vp]
    L.cc_download_bodies.argtypes = [vp, C.POINTER(_abi.BodyState), C.c_int32]
    L.cc_set_bodies.argtypes = [vp, C.POINTER(_abi.BodyState), C.c_int32]
    L.cc_download_pairs.argtypes = [vp, ip, C.c_int32, ip]
    L.cc_download_contacts.argtypes = [vp, C.POINTER(_abi.ContactOut), C.c_int32, ip]
    L.cc_download_removed_contacts.argtypes = [vp, C.POINTER(_abi.ContactRemoved), C.c_int32, ip]
    L.cc_download_ccd_bodies.argtypes = [vp, C.POINTER(_abi.CcdBody), C.c_int32, ip]
    L.cc_download_removed_contacts_noalloc.argtypes = [vp, C.POINTER(_abi.ContactRemoved), C.c_int32, ip]
    L.cc_snapshot.argtypes = [vp, vp, C.c_int64, C.POINTER(C.c_int64)]
    L.cc_restore.argtypes = [vp, vp, C.c_int64]
    L.cc_get_counters.argtypes = [vp, C.POINTER(_abi.StepCounters)]
    fp = C.POINTER(C.c_float)
    L.cc_broadphase_aabbs.argtypes = [vp, fp, fp, ip, ip, ip, C.c_int32, C.c_double, ip, C.c_int32, ip]
    L.cc_get_phase_times.argtypes = [vp, C.POINTER(C.c_double), C.c_int32, ip]
    L.cc_get_totals.argtypes = [vp, C.POINTER(_abi.Totals)]
    L.cc_add_forces.argtypes = [vp, dp, dp, C.c_int32, C.c_int32]
    L.cc_get_solver_residual.argtypes = [vp, C.POINTER(C.c_double)]
    L.cc_set_contact_cache.argtypes = [vp, C.POINTER(_abi.CachedContact), C.c_int32, C.c_double]
    L.cc_download_contact_cache.argtypes = [vp, C.POINTER(_abi.CachedContact), C.c_int32, ip, dp]
    L.cc_set_sleep_state.argtypes = [vp, C.POINTER(_abi.BodySleep), C.c_int32]
    L.cc_set_point_impulses.argtypes = [vp, dp, C.c_int32]
    L.cc_fwd.argtypes = [vp, C.c_double, C.c_int32, C.c_int32]
    L.cc_bwd.argtypes = [vp, dp, dp, dp, dp, dp, dp]
    for s in SYMBOLS:
        getattr(L, s).restype = getattr(L, s).restype if s == "cc_last_error" else C.c_int
    _LIB = L
    return L


BODY_STATE_DTYPE = np.dtype(_abi.BodyState)
CONTACT_DTYPE = np.dtype(_abi.ContactOut)
REMOVED_DTYPE = np.dtype(_abi.ContactRemoved)
CCD_DTYPE = np.dtype(_abi.CcdBody)
# the step's phases, in the order cc_get_phase_times reports them
PHASES = ["accel", "grid", "cell_pairs", "pair_layout", "filter", "gjk", "epa", "manifold", "manifold_large", "islands",
          "solver", "cache"]
# phases that time one kernel on its own (with the launches of tiny bookkeeping kernels)
PHASE_KERNEL = {"accel": "k_accel", "cell_pairs": "k_bp_cells", "gjk": "k_np_gjk", "epa": "k_epa",
                "manifold": "k_np_manifold", "solver": "k_world"}


class Context:
    """One device context (a batch of worlds) behind the C ABI."""

    def __init__(self, precision="fp32", device=0, max_pairs_per_body=16, block_solver=0, max_contacts_per_body=8,
                 island_groups=0, block_bodies=0, graphs=True, solver_order="reference", measure_residual=False,
                 narrowphase="fp64"):
        """narrowphase: arithmetic of GJK and EPA in an FP32 context, "fp64" (from the FP32 poses,
        with FP64 shape data) or "fp32"; FP64 contexts always use FP64."""
        L = lib()
        cfg = _abi.Config()
        cfg.precision = _abi.CC_PRECISION_FP64 if precision == "fp64" else _abi.CC_PRECISION_FP32
        cfg.device = device
        cfg.max_pairs_per_body = max_pairs_per_body
        cfg.max_contacts_per_body = max_contacts_per_body
        cfg.block_solver = block_solver
        cfg.island_groups = island_groups
        cfg.block_bodies = block_bodies
        cfg.disable_graphs = 0 if graphs else 1
        if solver_order not in ("reference", "colored"):
            raise ValueError("solver_order is 'reference' or 'colored'")
        cfg.solver_order = _abi.CC_ORDER_COLORED if solver_order == "colored" else _abi.CC_ORDER_REFERENCE
        cfg.measure_residual = 1 if measure_residual else 0
        if narrowphase not in ("fp64", "fp32"):
            raise ValueError("narrowphase is 'fp64' or 'fp32'")
        cfg.narrowphase_fp32 = 1 if narrowphase == "fp32" else 0
        h = C.c_void_p()
        rc = L.cc_create(C.byref(h), C.byref(cfg))
        if rc != 0:
            raise CrashcatError(f"cc_create failed ({rc}): no usable HIP device")
        self.h = h
        self.n = 0
        self.precision = precision

    def _check(self, rc):
        if rc != 0:
            raise CrashcatError(lib().cc_last_error(self.h).decode() + f" (code {rc})")

    def close(self):
        if self.h:
            lib().cc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, scene):
        L = lib()
        bodies = scene.bodies()
        shapes, hulls, keep = scene.c_shapes()
        meshes = keep["meshes"]
        pcs, npc = scene.c_point_constraints()
        offs = scene.c_offsets()
        self._keep = (bodies, keep, pcs, offs)
        self.n = len(bodies)
        self.num_worlds = len(offs) - 1
        self._check(L.cc_upload_scene(self.h, C.byref(scene.settings), len(offs) - 1,
                                      offs.ctypes.data_as(C.POINTER(C.c_int32)),
                                      bodies.ctypes.data_as(C.POINTER(_abi.BodyDesc)), self.n, shapes,
                                      keep["num_shapes"], C.byref(hulls), pcs, npc, C.byref(meshes)))

    def step(self, dt=1.0 / 60.0, n=1):
        self._check(lib().cc_step(self.h, dt, n))

    def sync(self):
        self._check(lib().cc_sync(self.h))

    def bodies(self):
        out = np.zeros(self.n, dtype=BODY_STATE_DTYPE)
        self._check(lib().cc_download_bodies(self.h, out.ctypes.data_as(C.POINTER(_abi.BodyState)), self.n))
        return out

    def set_bodies(self, states):
        st = np.ascontiguousarray(states, dtype=BODY_STATE_DTYPE)
        self._check(lib().cc_set_bodies(self.h, st.ctypes.data_as(C.POINTER(_abi.BodyState)), self.n))

    def add_forces(self, force=None, torque=None, wake=True):
        """addForce / addTorque on every body: (num_bodies, 3) arrays, either may be None."""
        dp = C.POINTER(C.c_double)
        f = None if force is None else np.ascontiguousarray(force, dtype=np.float64).reshape(self.n, 3)
        t = None if torque is None else np.ascontiguousarray(torque, dtype=np.float64).reshape(self.n, 3)
        self._check(lib().cc_add_forces(self.h, None if f is None else f.ctypes.data_as(dp),
                                        None if t is None else t.ctypes.data_as(dp), self.n, int(bool(wake))))

    def fwd(self, dt, nsteps, checkpoint_every=10):
        """Differentiable forward: nsteps steps of dt with a checkpoint every checkpoint_every."""
        self._check(lib().cc_fwd(self.h, dt, nsteps, checkpoint_every))

    def bwd(self, dl_dcom, dl_dvel=None):
        """Adjoint of the last fwd: returns dL/d(initial v, w, com) as (n, 3) arrays and dL/d(friction
        of each body) as (n,), given dL/d(final com) (and optionally dL/d(final v))."""
        dp = C.POINTER(C.c_double)
        g = np.ascontiguousarray(dl_dcom, dtype=np.float64).reshape(self.n, 3)
        gv = None if dl_dvel is None else np.ascontiguousarray(dl_dvel, dtype=np.float64).reshape(self.n, 3)
        out = {k: np.zeros((self.n, 3)) for k in ("v0", "w0", "com0")}
        out["friction"] = np.zeros(self.n)
        self._check(lib().cc_bwd(self.h, g.ctypes.data_as(dp), None if gv is None else gv.ctypes.data_as(dp),
                                 out["v0"].ctypes.data_as(dp), out["w0"].ctypes.data_as(dp),
                                 out["com0"].ctypes.data_as(dp), out["friction"].ctypes.data_as(dp)))
        return out

    def pairs(self):
        cnt = C.c_int32()
        self._check(lib().cc_download_pairs(self.h, None, 0, C.byref(cnt)))
        out = np.zeros((max(cnt.value, 1), 2), dtype=np.int32)
        self._check(lib().cc_download_pairs(self.h, out.ctypes.data_as(C.POINTER(C.c_int32)), cnt.value, C.byref(cnt)))
        return out[:cnt.value]

    def contacts(self):
        cnt = C.c_int32()
        self._check(lib().cc_download_contacts(self.h, None, 0, C.byref(cnt)))
        out = np.zeros(max(cnt.value, 1), dtype=CONTACT_DTYPE)
        self._check(lib().cc_download_contacts(self.h, out.ctypes.data_as(C.POINTER(_abi.ContactOut)), cnt.value,
                                               C.byref(cnt)))
        return out[:cnt.value]

    def snapshot(self):
        """The state one step hands to the next, as bytes (cc_snapshot)."""
        n = C.c_int64()
        self._check(lib().cc_snapshot(self.h, None, 0, C.byref(n)))
        buf = C.create_string_buffer(n.value)
        self._check(lib().cc_snapshot(self.h, buf, n.value, C.byref(n)))
        return buf.raw[:n.value]

    def restore(self, blob):
        """Load a snapshot taken from a context holding the same upload (cc_restore)."""
        buf = C.create_string_buffer(bytes(blob), len(blob))
        self._check(lib().cc_restore(self.h, buf, len(blob)))

    def removed_contacts(self):
        """Contacts of the step before the last that the last step did not make again
        (onContactRemoved, src/contacts.ts:297-307), sorted by (body_a, body_b, sub ids)."""
        cnt = C.c_int32()
        self._check(lib().cc_download_removed_contacts(self.h, None, 0, C.byref(cnt)))
        out = np.zeros(max(cnt.value, 1), dtype=REMOVED_DTYPE)
        self._check(lib().cc_download_removed_contacts(self.h, out.ctypes.data_as(C.POINTER(_abi.ContactRemoved)),
                                                       cnt.value, C.byref(cnt)))
        return out[:cnt.value]

    def ccd_bodies(self):
        """The last step's CCD records (world.ccd.ccdBodies after updateWorld, src/update.ts:131-140),
        each world's in time-of-impact order."""
        cnt = C.c_int32()
        self._check(lib().cc_download_ccd_bodies(self.h, None, 0, C.byref(cnt)))
        out = np.zeros(max(cnt.value, 1), dtype=CCD_DTYPE)
        self._check(lib().cc_download_ccd_bodies(self.h, out.ctypes.data_as(C.POINTER(_abi.CcdBody)), cnt.value,
                                                 C.byref(cnt)))
        return out[:cnt.value]

    def counters(self):
        c = _abi.StepCounters()
        self._check(lib().cc_get_counters(self.h, C.byref(c)))
        return {k: getattr(c, k) for k, _ in _abi.StepCounters._fields_}

    def phase_times(self):
        buf = (C.c_double * 16)()
        cnt = C.c_int32()
        self._check(lib().cc_get_phase_times(self.h, buf, 16, C.byref(cnt)))
        return {PHASES[i] if i < len(PHASES) else str(i): buf[i] for i in range(cnt.value)}

    def set_contact_cache(self, records, previous_time_step):
        """Replace the cache the next step warm-starts from (cc_set_contact_cache)."""
        r = np.ascontiguousarray(records, dtype=np.dtype(_abi.CachedContact))
        self._check(lib().cc_set_contact_cache(self.h, r.ctypes.data_as(C.POINTER(_abi.CachedContact)), len(r),
                                               C.c_double(previous_time_step)))

    def contact_cache(self):
        """(records sorted by key, previous time step) of the last step (cc_download_contact_cache)."""
        n, pdt = C.c_int32(), C.c_double()
        self._check(lib().cc_download_contact_cache(self.h, None, 0, C.byref(n), C.byref(pdt)))
        out = np.zeros(max(n.value, 1), dtype=np.dtype(_abi.CachedContact))
        self._check(lib().cc_download_contact_cache(self.h, out.ctypes.data_as(C.POINTER(_abi.CachedContact)),
                                                    n.value, C.byref(n), None))
        return out[:n.value], pdt.value

    def set_sleep_state(self, sleep):
        s = np.ascontiguousarray(sleep, dtype=np.dtype(_abi.BodySleep))
        self._check(lib().cc_set_sleep_state(self.h, s.ctypes.data_as(C.POINTER(_abi.BodySleep)), len(s)))

    def set_point_impulses(self, lam):
        a = np.ascontiguousarray(lam, dtype=np.float64).reshape(-1, 3)
        self._check(lib().cc_set_point_impulses(self.h, a.ctypes.data_as(C.POINTER(C.c_double)), len(a)))

    def solver_residual(self):
        """The last step's velocity residual (cc_get_solver_residual): sum of |dlambda|, sum of
        normal impulses, largest |dlambda|, contact points."""
        out = (C.c_double * 4)()
        self._check(lib().cc_get_solver_residual(self.h, out))
        return {"sum": out[0], "lambda_n_sum": out[1], "max": out[2], "points": int(out[3])}

    def totals(self):
        t = _abi.Totals()
        self._check(lib().cc_get_totals(self.h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in _abi.Totals._fields_ if k != "reserved"}

    def broadphase_aabbs(self, mn, mx, motion, sleeping, layer, margin):
        mn = np.ascontiguousarray(mn, dtype=np.float32)
        mx = np.ascontiguousarray(mx, dtype=np.float32)
        mo = np.ascontiguousarray(motion, dtype=np.int32)
        sl = np.ascontiguousarray(sleeping, dtype=np.int32)
        la = np.ascontiguousarray(layer, dtype=np.int32)
        n = len(mo)
        ip = C.POINTER(C.c_int32)
        fp = C.POINTER(C.c_float)
        cnt = C.c_int32()
        cap = max(64 * n, 1)
        out = np.zeros((cap, 2), dtype=np.int32)
        self._check(lib().cc_broadphase_aabbs(self.h, mn.ctypes.data_as(fp), mx.ctypes.data_as(fp), mo.ctypes.data_as(ip),
                                              sl.ctypes.data_as(ip), la.ctypes.data_as(ip), n, margin,
                                              out.ctypes.data_as(ip), cap, C.byref(cnt)))
        self.n = n
        return out[:min(cnt.value, cap)]


def build_triangle_mesh(positions, indices, degenerate_tolerance=1e-6, active_edge_cos_threshold=0.996195,
                        max_leaf_triangles=8):
    """buildTriangleMesh (src/shapes/utils/triangle-mesh-builder.ts:62-220) on the host, through the
    library: returns (positions [V,3] float64, triangles [T,4] int32 = a, b, c, active-edge bits,
    nodes [N,8] float64 = the BVH node buffer)."""
    pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
    idx = np.ascontiguousarray(indices, dtype=np.int32).reshape(-1, 3)
    nv, nt = len(pos), len(idx)
    op = np.zeros((max(nv, 1), 3), dtype=np.float64)
    ot = np.zeros((max(nt, 1), 4), dtype=np.int32)
    on = np.zeros((max(2 * nt, 1), 8), dtype=np.float64)
    cnt = np.zeros(3, dtype=np.int32)
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int32)
    rc = lib().cc_build_triangle_mesh(pos.ctypes.data_as(dp), nv, idx.ctypes.data_as(ip), nt, degenerate_tolerance,
                                      active_edge_cos_threshold, max_leaf_triangles, op.ctypes.data_as(dp),
                                      ot.ctypes.data_as(ip), on.ctypes.data_as(dp), cnt.ctypes.data_as(ip))
    if rc != 0:
        raise CrashcatError(f"cc_build_triangle_mesh failed (code {rc})")
    return op[:cnt[0]].copy(), ot[:cnt[1]].copy(), on[:cnt[2]].copy()
