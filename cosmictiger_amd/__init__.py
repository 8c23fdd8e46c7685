"""MI355X-native CosmicTiger gravity kick: Python host mirror of the C-ABI.

This module binds ``libcosmictiger_amd.so`` (include/cosmictiger_amd.h) with
ctypes and mirrors the reference's host interface for the kick path:

* ``KickParams`` / ``KickReturn``: ``kick_params`` / ``kick_return``
  (cosmictiger/kick.hpp:92-186), same fields, same order.
* ``Box``: a resident box of particles (cosmictiger/particles.hpp:159-306)
  with ``tree_build`` (tree_create, src/tree.cpp:273), ``kick``
  (kick(params, 0, 0, root, {root}, {root}), src/kick.cpp:150) and ``drift``
  (cuda_drift, src/drift.cu:156).

The library must be built in-tree (``cosmictiger_amd.build``); importing
``lib()`` without it raises, there is no fallback path.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcosmictiger_amd.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "cosmictiger_amd.h")

CT_OK = 0
CT_E_ARG = -1
CT_E_OOM = -2
CT_E_HIP = -3
CT_E_ARENA = -4
CT_E_STATE = -5
CT_E_TREE = -6
CT_E_WALK = -7

# where ct_tree_build_ex builds the tree
TREE_AUTO = 0    # device, host fallback for trees deeper than 63 levels
TREE_HOST = 1
TREE_DEVICE = 2

# arithmetic modes of ct_kick_prec
PREC_F32C = 0  # FP32, FP64 sink accumulation per source tile (the default kick)
PREC_F32U = 1  # FP32, FP32 running sums
PREC_F64 = 2   # FP64 everywhere but the walk's acceptance tests

SELF_PHI = 693.0 / 256.0  # cosmictiger/gravity.hpp:98-102
EXPANSION_SIZE = 50
MULTIPOLE_SIZE = 37

_lib = None
_inited_device = None


class CTError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (code %d)" % (msg, code))
        self.code = code


class KickParams(ctypes.Structure):
    """kick_params (cosmictiger/kick.hpp:138-157)."""
    _fields_ = [("sign", ctypes.c_float), ("min_level", ctypes.c_int), ("min_rung", ctypes.c_int), ("a", ctypes.c_float),
                ("t0", ctypes.c_float), ("theta", ctypes.c_float), ("h", ctypes.c_float), ("eta", ctypes.c_float),
                ("GM", ctypes.c_float), ("vel_init", ctypes.c_bool), ("save_force", ctypes.c_bool),
                ("first_call", ctypes.c_bool), ("gpu", ctypes.c_bool), ("node_load", ctypes.c_float),
                ("max_dt", ctypes.c_float), ("ascending", ctypes.c_bool), ("descending", ctypes.c_bool),
                ("top", ctypes.c_bool), ("do_phi", ctypes.c_bool)]

    def __init__(self, **kw):
        super().__init__()
        # defaults of the reference constructor (kick.hpp:158-163)
        self.max_dt = 1e30
        self.do_phi = True
        self.vel_init = False
        self.sign = 1.0
        for k, v in kw.items():
            setattr(self, k, v)


class KickReturn(ctypes.Structure):
    """kick_return (cosmictiger/kick.hpp:92-100)."""
    _fields_ = [("max_rung", ctypes.c_int), ("pot", ctypes.c_double), ("kin", ctypes.c_double), ("dkin", ctypes.c_double),
                ("xmom", ctypes.c_double), ("ymom", ctypes.c_double), ("zmom", ctypes.c_double), ("nmom", ctypes.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class Node(ctypes.Structure):
    _fields_ = [("box_begin", ctypes.c_int32 * 3), ("box_end", ctypes.c_int32 * 3), ("child", ctypes.c_int32 * 2),
                ("parent", ctypes.c_int32), ("pbegin", ctypes.c_int32), ("pend", ctypes.c_int32), ("pos", ctypes.c_uint32 * 3),
                ("radius", ctypes.c_float), ("leaf", ctypes.c_int32), ("depth", ctypes.c_int32), ("multi", ctypes.c_float * 37)]


NODE_DTYPE = np.dtype([("box_begin", "<i4", 3), ("box_end", "<i4", 3), ("child", "<i4", 2), ("parent", "<i4"), ("pbegin", "<i4"),
                       ("pend", "<i4"), ("pos", "<u4", 3), ("radius", "<f4"), ("leaf", "<i4"), ("depth", "<i4"),
                       ("multi", "<f4", 37)])
assert NODE_DTYPE.itemsize == ctypes.sizeof(Node)

_P = ctypes.c_void_p
_I64P = ctypes.POINTER(ctypes.c_int64)

# ct_lc_entry (lc_entry of cosmictiger/lightcone.hpp:54-58)
LC_ENTRY = np.dtype([("x", np.int32), ("y", np.int32), ("z", np.int32), ("vx", np.float32), ("vy", np.float32), ("vz", np.float32),
                     ("pix", np.int32)])


class StepParams(ctypes.Structure):
    """Arguments of one call of the hierarchical step (kick_step_hierarchical,
    src/driver.cpp:382-593), ct_step_params of the C-ABI."""
    _fields_ = [("minrung", ctypes.c_int), ("max_rung", ctypes.c_int), ("minrung0", ctypes.c_int), ("scale", ctypes.c_double),
                ("tau", ctypes.c_double), ("t0", ctypes.c_double), ("theta", ctypes.c_double), ("bucket", ctypes.c_int),
                ("hsoft", ctypes.c_double), ("GM", ctypes.c_double), ("eta", ctypes.c_double), ("max_dt", ctypes.c_double),
                ("sign", ctypes.c_double), ("do_phi", ctypes.c_int), ("nocrop", ctypes.c_int), ("drag", ctypes.c_double),
                ("save_force", ctypes.c_int), ("precision", ctypes.c_int), ("tree_where", ctypes.c_int)]

    def __init__(self, **kw):
        super().__init__()
        self.scale = 1.0
        self.t0 = 1.0
        self.theta = 0.7
        self.bucket = 128
        self.eta = 0.25 / np.sqrt(SELF_PHI)
        self.max_dt = 1e30
        self.sign = 1.0
        self.do_phi = 1
        self.save_force = 1
        for k, v in kw.items():
            setattr(self, k, v)


class StepReturn(ctypes.Structure):
    _fields_ = [("kr", KickReturn), ("minrung", ctypes.c_int), ("nkicks", ctypes.c_int), ("clip_top", ctypes.c_int),
                ("tckin", ctypes.c_double), ("parts_processed", ctypes.c_double)]


def _declare(L):
    L.ct_init.argtypes = [ctypes.c_int]
    L.ct_ewald_init.argtypes = []
    L.ct_box_create.argtypes = [ctypes.c_int64, _P, _P, _P, _P, _P, ctypes.POINTER(_P)]
    L.ct_box_destroy.argtypes = [_P]
    L.ct_tree_build.argtypes = [_P, ctypes.c_double, ctypes.c_int, ctypes.c_double]
    L.ct_tree_build_ex.argtypes = [_P, ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_int]
    L.ct_tree_build_host.argtypes = [ctypes.c_int64, _P, _P, _P, _P, ctypes.c_double, ctypes.c_int, ctypes.c_double, _P,
                                     ctypes.c_int64, _I64P]
    L.ct_kick.argtypes = [_P, ctypes.POINTER(KickParams), ctypes.POINTER(KickReturn), _P]
    L.ct_kick_prec.argtypes = [_P, ctypes.POINTER(KickParams), ctypes.c_int, ctypes.POINTER(KickReturn), _P]
    L.ct_drift.argtypes = [_P, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_double, _P]
    L.ct_drift_lightcone.argtypes = [_P, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                     ctypes.c_int, ctypes.c_double, _I64P, _P]
    L.ct_lc_read.argtypes = [_P, _P, ctypes.c_int64, _I64P]
    L.ct_box_size.argtypes = [_P, _I64P, _I64P]
    L.ct_box_read.argtypes = [_P] + [_P] * 10
    L.ct_tree_read.argtypes = [_P, _P, ctypes.c_int64]
    L.ct_lists_read.argtypes = [_P, ctypes.c_int, _P, _P, ctypes.c_int64]
    L.ct_kick_stats.argtypes = [_P, _P, ctypes.c_int]
    L.ct_direct_ewald.argtypes = [_P, ctypes.c_int64, _P, _P, _P, ctypes.c_double, _P, _P, _P, _P, _P]
    L.ct_batch_create.argtypes = [_P, ctypes.c_int, ctypes.POINTER(_P)]
    L.ct_batch_destroy.argtypes = [_P]
    L.ct_batch_reset.argtypes = [_P, _P, ctypes.c_int]
    L.ct_batch_step.argtypes = [_P, _P, ctypes.c_int, _P, ctypes.c_int, _P, _P, _P, ctypes.c_int, _P]
    L.ct_batch_kick.argtypes = [_P, _P, ctypes.c_int, _P, ctypes.c_int, _P]
    L.ct_batch_drift.argtypes = [_P, ctypes.c_int, _P, _P, _P, _P]
    L.ct_batch_stats.argtypes = [_P, _P, ctypes.c_int]
    L.ct_box_sort_by_rung.argtypes = [_P, ctypes.c_int, _I64P]
    L.ct_box_current_range.argtypes = [_P, _I64P, _I64P]
    L.ct_box_push_rungs.argtypes = [_P]
    L.ct_box_pop_rungs.argtypes = [_P]
    L.ct_box_set_minrung.argtypes = [_P, ctypes.c_int]
    L.ct_box_rung_counts.argtypes = [_P, _P, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    L.ct_box_save.argtypes = [_P, ctypes.c_char_p, ctypes.c_int]
    L.ct_box_load.argtypes = [ctypes.c_char_p, ctypes.c_int64, ctypes.POINTER(_P), _I64P]
    L.ct_step_hierarchical.argtypes = [_P, ctypes.POINTER(StepParams), ctypes.POINTER(StepReturn), _P]
    L.ct_time_inc.argtypes = [ctypes.c_uint64, ctypes.c_int]
    L.ct_time_inc.restype = ctypes.c_uint64
    L.ct_time_min_rung.argtypes = [ctypes.c_uint64]
    L.ct_time_min_rung.restype = ctypes.c_int
    L.ct_fmm_host.argtypes = [ctypes.c_int, _P, _P]
    L.ct_fmm_device.argtypes = [ctypes.c_int, ctypes.c_int64, _P, _P]
    L.ct_last_error.argtypes = []
    L.ct_last_error.restype = ctypes.c_char_p
    for name in ("ct_init", "ct_ewald_init", "ct_box_create", "ct_box_destroy", "ct_tree_build", "ct_tree_build_ex",
                 "ct_tree_build_host", "ct_kick",
                 "ct_kick_prec",
                 "ct_drift", "ct_drift_lightcone", "ct_lc_read", "ct_box_size", "ct_box_read", "ct_tree_read", "ct_lists_read", "ct_kick_stats", "ct_fmm_host",
                 "ct_fmm_device", "ct_direct_ewald", "ct_batch_create", "ct_batch_destroy", "ct_batch_reset", "ct_batch_step", "ct_batch_kick", "ct_batch_drift",
                 "ct_batch_stats", "ct_box_sort_by_rung", "ct_box_current_range", "ct_box_push_rungs", "ct_box_pop_rungs",
                 "ct_box_set_minrung", "ct_box_rung_counts", "ct_step_hierarchical", "ct_box_save", "ct_box_load"):
        getattr(L, name).restype = ctypes.c_int


def lib():
    """The loaded C-ABI library.  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libcosmictiger_amd.so is not built; run cosmictiger_amd.build.build() "
                               "(or __graft_entry__.build())")
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def _check(rc):
    if rc != CT_OK:
        raise CTError(rc, lib().ct_last_error().decode())


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_P)


def init(device=0):
    """Select the device and upload the Ewald tables (ewald_const::init)."""
    global _inited_device
    if _inited_device != device:
        _check(lib().ct_init(device))
        _inited_device = device


# ---------------------------------------------------------------------------
# expansion kernels (for layout checks)
# ---------------------------------------------------------------------------

FMM_OPS = {"greens": 0, "ewald_greens": 1, "M2L": 2, "L2L": 3, "L2P": 4, "P2M": 5, "M2M": 6, "M2L2": 7}
_IN_STRIDE = {0: 3, 1: 3, 2: 87, 3: 53, 4: 53, 5: 3, 6: 40, 7: 87}
_OUT_STRIDE = {0: 50, 1: 50, 2: 50, 3: 50, 4: 4, 5: 37, 6: 37, 7: 4}


def fmm_host(op, inp):
    """Evaluate one expansion kernel on the host in double."""
    op = FMM_OPS.get(op, op)
    inp = np.ascontiguousarray(inp, dtype=np.float64)
    assert inp.size == _IN_STRIDE[op]
    out = np.zeros(_OUT_STRIDE[op])
    _check(lib().ct_fmm_host(op, _ptr(inp), _ptr(out)))
    return out


def fmm_device(op, inp):
    """Evaluate one expansion kernel on the device in float for a batch (n, in_stride)."""
    op = FMM_OPS.get(op, op)
    inp = np.ascontiguousarray(inp, dtype=np.float64).reshape(-1, _IN_STRIDE[op])
    out = np.zeros((inp.shape[0], _OUT_STRIDE[op]))
    _check(lib().ct_fmm_device(op, inp.shape[0], _ptr(inp), _ptr(out)))
    return out


# ---------------------------------------------------------------------------
# tree (host)
# ---------------------------------------------------------------------------

def tree_build_host(x, y, z, theta, bucket, hsoft):
    """Host tree build over copies of the positions.  Returns (x, y, z, perm, nodes)."""
    x = np.array(x, dtype=np.uint32)
    y = np.array(y, dtype=np.uint32)
    z = np.array(z, dtype=np.uint32)
    n = len(x)
    perm = np.zeros(n, dtype=np.int32)
    nn = ctypes.c_int64(0)
    cap = max(1024, 8 * n // max(bucket, 1) + 65536)
    for _ in range(4):
        nodes = np.zeros(cap, dtype=NODE_DTYPE)
        xs, ys, zs = x.copy(), y.copy(), z.copy()
        rc = lib().ct_tree_build_host(n, _ptr(xs), _ptr(ys), _ptr(zs), _ptr(perm), float(theta), int(bucket), float(hsoft),
                                      _ptr(nodes), cap, ctypes.byref(nn))
        if rc == CT_E_ARG and nn.value > cap:
            cap = nn.value
            continue
        _check(rc)
        return xs, ys, zs, perm, nodes[:nn.value].copy()
    raise CTError(CT_E_ARG, "tree_build_host: capacity")


# ---------------------------------------------------------------------------
# resident box
# ---------------------------------------------------------------------------

class Box:
    """A resident box: particles and tree in HBM, kicked and drifted there."""

    def __init__(self, x, y, z, vel=None, rung=None, device=0, _handle=None):
        init(device)
        if _handle is not None:
            self._h = _handle
            self.n = self.size()[0]
            return
        self.n = len(x)
        x = np.ascontiguousarray(x, dtype=np.uint32)
        y = np.ascontiguousarray(y, dtype=np.uint32)
        z = np.ascontiguousarray(z, dtype=np.uint32)
        vel = None if vel is None else np.ascontiguousarray(vel, dtype=np.float32).reshape(self.n, 3)
        rung = None if rung is None else np.ascontiguousarray(rung, dtype=np.int8)
        h = _P()
        _check(lib().ct_box_create(self.n, _ptr(x), _ptr(y), _ptr(z), _ptr(vel), _ptr(rung), ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().ct_box_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def tree_build(self, theta, bucket, hsoft):
        _check(lib().ct_tree_build(self._h, float(theta), int(bucket), float(hsoft)))

    def tree_build_ex(self, theta, bucket, hsoft, where=TREE_AUTO):
        """tree_build on the device (TREE_DEVICE), the host (TREE_HOST) or
        the device with the host as fallback (TREE_AUTO, what tree_build does)."""
        _check(lib().ct_tree_build_ex(self._h, float(theta), int(bucket), float(hsoft), int(where)))

    def kick(self, params, stream=None, precision=None):
        """One kick of the box; precision: None (the default FP32 kick) or one
        of PREC_F32C, PREC_F32U, PREC_F64."""
        ret = KickReturn()
        if precision is None:
            _check(lib().ct_kick(self._h, ctypes.byref(params), ctypes.byref(ret), stream))
        else:
            _check(lib().ct_kick_prec(self._h, ctypes.byref(params), int(precision), ctypes.byref(ret), stream))
        return ret

    def drift(self, rung, a, dt, drag=0.0, stream=None):
        """cuda_drift over the active range (src/drift.cu:156-195)."""
        _check(lib().ct_drift(self._h, int(rung), float(a), float(dt), float(drag), stream))

    def drift_lightcone(self, rung, a, dt, tau0, tau1, tau_max, nside, drag=0.0, stream=None, read=True):
        """The drift with the lightcone of cuda_drift_kernel (src/drift.cu:36-154):
        returns the crossings as an array of LC_ENTRY (in no particular order),
        or only their number with read=False (they stay with the box)."""
        n = ctypes.c_int64(0)
        _check(lib().ct_drift_lightcone(self._h, int(rung), float(a), float(dt), float(tau0), float(tau1), float(tau_max), int(nside),
                                        float(drag), ctypes.byref(n), stream))
        if not read:
            return n.value
        out = np.zeros(max(n.value, 1), dtype=LC_ENTRY)
        _check(lib().ct_lc_read(self._h, out.ctypes.data_as(ctypes.c_void_p), n.value, None))
        return out[:n.value]

    # ---- active range and rungs (src/particles.cpp:218-235,583-649,1020-1068)

    def sort_by_rung(self, min_rung):
        """particles_sort_by_rung: particles of rung >= min_rung to the end of
        the active range, which then starts at the returned index."""
        rb = ctypes.c_int64()
        _check(lib().ct_box_sort_by_rung(self._h, int(min_rung), ctypes.byref(rb)))
        return rb.value

    def current_range(self):
        """particles_current_range: (first, second) of the active range."""
        a, b = ctypes.c_int64(), ctypes.c_int64()
        _check(lib().ct_box_current_range(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def push_rungs(self):
        _check(lib().ct_box_push_rungs(self._h))

    def pop_rungs(self):
        _check(lib().ct_box_pop_rungs(self._h))

    def set_minrung(self, min_rung):
        _check(lib().ct_box_set_minrung(self._h, int(min_rung)))

    def rung_counts(self):
        """particles_rung_counts: counts[r] for r up to the highest rung present."""
        c = np.zeros(128, np.int64)
        nr = ctypes.c_int()
        _check(lib().ct_box_rung_counts(self._h, _ptr(c), 128, ctypes.byref(nr)))
        return c[:nr.value].copy()

    def step_hierarchical(self, params, stream=None):
        """One call of the kick ladder (kick_step_hierarchical,
        src/driver.cpp:382-593).  Returns a StepReturn."""
        out = StepReturn()
        _check(lib().ct_step_hierarchical(self._h, ctypes.byref(params), ctypes.byref(out), stream))
        return out

    def save(self, path, append=False):
        """particles_save layout (src/particles.cpp:1206-1227)."""
        _check(lib().ct_box_save(self._h, os.fsencode(path), int(bool(append))))

    @classmethod
    def load(cls, path, offset=0, device=0):
        """particles_load (src/particles.cpp:1182-1204): (box, offset past the record)."""
        init(device)
        h = _P()
        end = ctypes.c_int64()
        _check(lib().ct_box_load(os.fsencode(path), int(offset), ctypes.byref(h), ctypes.byref(end)))
        return cls(None, None, None, device=device, _handle=h), end.value

    def size(self):
        n = ctypes.c_int64()
        nn = ctypes.c_int64()
        _check(lib().ct_box_size(self._h, ctypes.byref(n), ctypes.byref(nn)))
        return n.value, nn.value

    def read(self):
        n = self.n
        out = {"x": np.zeros(n, np.uint32), "y": np.zeros(n, np.uint32), "z": np.zeros(n, np.uint32),
               "vel": np.zeros((n, 3), np.float32), "rung": np.zeros(n, np.int8), "gx": np.zeros(n, np.float32),
               "gy": np.zeros(n, np.float32), "gz": np.zeros(n, np.float32), "phi": np.zeros(n, np.float32),
               "id": np.zeros(n, np.int32)}
        _check(lib().ct_box_read(self._h, *[_ptr(out[k]) for k in ("x", "y", "z", "vel", "rung", "gx", "gy", "gz", "phi", "id")]))
        return out

    def nodes(self):
        _, nn = self.size()
        arr = np.zeros(nn, dtype=NODE_DTYPE)
        _check(lib().ct_tree_read(self._h, _ptr(arr), nn))
        return arr

    def lists(self, kind):
        _, nn = self.size()
        off = np.zeros(nn + 1, np.int64)
        _check(lib().ct_lists_read(self._h, int(kind), _ptr(off), None, 0))
        idx = np.zeros(max(int(off[-1]), 1), np.int32)
        _check(lib().ct_lists_read(self._h, int(kind), _ptr(off), _ptr(idx), len(idx)))
        return off, idx[:int(off[-1])]

    def direct_ewald(self, sx, sy, sz, hsoft, stream=None):
        """FP64 direct Ewald sum at the given fixed32 sink positions over every
        particle of the box, unit masses (gravity_analytic_call_kernel,
        src/analytic.cu:30).  Returns dict of float64 arrays phi, gx, gy, gz."""
        sx = np.ascontiguousarray(sx, dtype=np.uint32)
        sy = np.ascontiguousarray(sy, dtype=np.uint32)
        sz = np.ascontiguousarray(sz, dtype=np.uint32)
        m = len(sx)
        out = {k: np.zeros(m) for k in ("phi", "gx", "gy", "gz")}
        _check(lib().ct_direct_ewald(self._h, m, _ptr(sx), _ptr(sy), _ptr(sz), float(hsoft), _ptr(out["phi"]), _ptr(out["gx"]),
                                     _ptr(out["gy"]), _ptr(out["gz"]), stream))
        return out

    def stats(self):
        out = np.zeros(15)
        _check(lib().ct_kick_stats(self._h, _ptr(out), 15))
        return dict(zip(STAT_KEYS, out.tolist()))


STAT_KEYS = ["walk_ms", "cells_ms", "l2l_ms", "leaf_ms", "kick_ms", "pp_pairs", "cc", "ewald_cc", "cp_parts", "pc_pairs", "nodes",
             "sinks", "pp_ms", "leafkick_ms", "checks"]


class Batch:
    """Resident boxes kicked together: each phase of the kick is one launch over
    all of them (the reference's kick_workitem list, cosmictiger/kick.hpp:188,
    extended to many boxes).  Optionally replayed from a captured HIP graph."""

    def __init__(self, boxes):
        self.boxes = list(boxes)
        arr = (_P * len(self.boxes))(*[b._h for b in self.boxes])
        h = _P()
        _check(lib().ct_batch_create(ctypes.cast(arr, _P), len(self.boxes), ctypes.byref(h)))
        self._h = h

    def reset(self, boxes):
        """Point the batch at another set of boxes, keeping its arenas."""
        self.boxes = list(boxes)
        arr = (_P * len(self.boxes))(*[b._h for b in self.boxes])
        _check(lib().ct_batch_reset(self._h, ctypes.cast(arr, _P), len(self.boxes)))

    def close(self):
        if getattr(self, "_h", None):
            lib().ct_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def kick(self, params, precision=PREC_F32C, graph=False, stream=None):
        """params: one KickParams per box.  Returns the list of KickReturn."""
        n = len(self.boxes)
        parr = (KickParams * n)(*params)
        rets = (KickReturn * n)()
        _check(lib().ct_batch_kick(self._h, ctypes.cast(parr, _P), int(precision), ctypes.cast(rets, _P), int(bool(graph)), stream))
        return list(rets)

    def step(self, params, rung, a, dt, drag=None, precision=PREC_F32C, graph=False, stream=None):
        """Kick then drift every box (one captured graph with graph=True).
        Returns the list of KickReturn."""
        n = len(self.boxes)
        parr = (KickParams * n)(*params)
        rets = (KickReturn * n)()
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(a, np.float32), (n,)))
        dt = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, np.float32), (n,)))
        dg = None if drag is None else np.ascontiguousarray(np.broadcast_to(np.asarray(drag, np.float64), (n,)))
        _check(lib().ct_batch_step(self._h, ctypes.cast(parr, _P), int(precision), ctypes.cast(rets, _P), int(rung), _ptr(a), _ptr(dt),
                                   _ptr(dg), int(bool(graph)), stream))
        return list(rets)

    def drift(self, rung, a, dt, drag=None, stream=None):
        n = len(self.boxes)
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(a, np.float32), (n,)))
        dt = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, np.float32), (n,)))
        dg = None if drag is None else np.ascontiguousarray(np.broadcast_to(np.asarray(drag, np.float64), (n,)))
        _check(lib().ct_batch_drift(self._h, int(rung), _ptr(a), _ptr(dt), _ptr(dg), stream))

    def stats(self):
        out = np.zeros(15)
        _check(lib().ct_batch_stats(self._h, _ptr(out), 15))
        return dict(zip(STAT_KEYS, out.tolist()))


def default_kick_params(theta, h, GM, eta=None, do_phi=True, save_force=True):
    """kick_params of the top-level first kick (src/driver.cpp:481-508 with tau = 0)."""
    p = KickParams()
    p.sign = 1.0
    p.min_level = 0
    p.min_rung = 0
    p.a = 1.0
    p.t0 = 1.0
    p.theta = theta
    p.h = h
    p.eta = eta if eta is not None else 0.25 / np.sqrt(SELF_PHI)
    p.GM = GM
    p.save_force = save_force
    p.first_call = True
    p.gpu = True
    p.node_load = 0.0
    p.max_dt = 1e30
    p.ascending = True
    p.descending = True
    p.top = True
    p.do_phi = do_phi
    return p


def time_inc(t, max_rung):
    """inc() of cosmictiger/time.hpp:28-33 (the C-ABI's ct_time_inc)."""
    if max_rung > 0:
        t = (t + (1 << (64 - max_rung))) & ((1 << 64) - 1)
    return t


def time_min_rung(t):
    """min_rung() of cosmictiger/time.hpp:35-43 (the C-ABI's ct_time_min_rung)."""
    r = 64
    while (t & 1) == 0 and r != 0:
        r -= 1
        t >>= 1
    return r


def softening(n):
    """Derived softening of the reference options: (1/80) * self_phi * N^(-1/3)
    (src/options.cpp:124,165,217)."""
    return (1.0 / 80.0) * SELF_PHI * n ** (-1.0 / 3.0)
