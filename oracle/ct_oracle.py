"""ctypes binding of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Loaded by tests/, by ``__graft_entry__.smoke()`` and by bench.py's CPU-baseline
leg, as the checker of the MI355X kick.  The product never imports this file.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libctoracle.so")
_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


class KickParams(ctypes.Structure):
    # field order of kick_params, cosmictiger/kick.hpp:138-157
    _fields_ = [("sign", ctypes.c_float), ("min_level", ctypes.c_int), ("min_rung", ctypes.c_int), ("a", ctypes.c_float),
                ("t0", ctypes.c_float), ("theta", ctypes.c_float), ("h", ctypes.c_float), ("eta", ctypes.c_float),
                ("GM", ctypes.c_float), ("vel_init", ctypes.c_bool), ("save_force", ctypes.c_bool),
                ("first_call", ctypes.c_bool), ("gpu", ctypes.c_bool), ("node_load", ctypes.c_float),
                ("max_dt", ctypes.c_float), ("ascending", ctypes.c_bool), ("descending", ctypes.c_bool),
                ("top", ctypes.c_bool), ("do_phi", ctypes.c_bool)]


class KickReturn(ctypes.Structure):
    # field order of kick_return, cosmictiger/kick.hpp:92-100
    _fields_ = [("max_rung", ctypes.c_int), ("pot", ctypes.c_double), ("kin", ctypes.c_double), ("dkin", ctypes.c_double),
                ("xmom", ctypes.c_double), ("ymom", ctypes.c_double), ("zmom", ctypes.c_double), ("nmom", ctypes.c_double)]


class Node(ctypes.Structure):
    _fields_ = [("box_begin", ctypes.c_int32 * 3), ("box_end", ctypes.c_int32 * 3), ("child", ctypes.c_int32 * 2),
                ("pbegin", ctypes.c_int32), ("pend", ctypes.c_int32), ("pos", ctypes.c_uint32 * 3), ("radius", ctypes.c_float),
                ("leaf", ctypes.c_int32), ("depth", ctypes.c_int32), ("multi", ctypes.c_double * 37)]


class Tree(ctypes.Structure):
    _fields_ = [("nodes", ctypes.POINTER(Node)), ("nnodes", ctypes.c_int64), ("cap", ctypes.c_int64), ("root", ctypes.c_int),
                ("error", ctypes.c_int)]


class Lists(ctypes.Structure):
    _fields_ = [("nnodes", ctypes.c_int64), ("off", ctypes.POINTER(ctypes.c_int64) * 5), ("idx", ctypes.POINTER(ctypes.c_int32) * 5)]


_D = ctypes.POINTER(ctypes.c_double)
_F = ctypes.POINTER(ctypes.c_float)
_U32 = ctypes.POINTER(ctypes.c_uint32)
_I32 = ctypes.POINTER(ctypes.c_int32)
_I64 = ctypes.POINTER(ctypes.c_int64)
_I8 = ctypes.POINTER(ctypes.c_int8)


def _declare(L):
    L.oc_init.restype = None
    L.oc_trless_index.restype = ctypes.c_int
    L.oc_trless_index.argtypes = [ctypes.c_int] * 4
    L.oc_full.restype = ctypes.c_double
    L.oc_full.argtypes = [_D, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    for name in ("oc_greens", "oc_ewald_greens"):
        getattr(L, name).argtypes = [_D, _D]
        getattr(L, name).restype = None
    L.oc_m2l.argtypes = [_D, ctypes.c_int, _D, _D, ctypes.c_int]
    L.oc_m2l.restype = None
    L.oc_l2l.argtypes = [_D, _D, _D, ctypes.c_int]
    L.oc_l2l.restype = None
    L.oc_l2p.argtypes = [_D, _D, _D, ctypes.c_int]
    L.oc_l2p.restype = None
    L.oc_p2m.argtypes = [_D, _D]
    L.oc_p2m.restype = None
    L.oc_m2m.argtypes = [_D, _D, _D]
    L.oc_m2m.restype = None
    L.oc_gsoft.argtypes = [_D, _D, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double]
    L.oc_gsoft.restype = None
    L.oc_self_phi.restype = ctypes.c_double
    L.oc_ewald_nreal.restype = ctypes.c_int
    L.oc_ewald_nfour.restype = ctypes.c_int
    L.oc_ewald_real_index.argtypes = [ctypes.c_int, _D]
    L.oc_ewald_four_index.argtypes = [ctypes.c_int, _D]
    L.oc_distance.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
    L.oc_distance.restype = ctypes.c_float
    L.oc_fixed32_from_double.argtypes = [ctypes.c_double]
    L.oc_fixed32_from_double.restype = ctypes.c_uint32
    L.oc_tree_build.argtypes = [ctypes.POINTER(Tree), ctypes.c_int64, _U32, _U32, _U32, _I32, ctypes.c_int, ctypes.c_double,
                                ctypes.c_double, ctypes.c_int, ctypes.c_int64]
    L.oc_tree_build.restype = ctypes.c_int
    L.oc_tree_free.argtypes = [ctypes.POINTER(Tree)]
    L.oc_kick.argtypes = [ctypes.POINTER(Tree), ctypes.POINTER(KickParams), _U32, _U32, _U32, _F, _I8, _D, _D, _D, _D,
                          ctypes.POINTER(KickReturn), ctypes.POINTER(Lists), ctypes.c_int]
    L.oc_kick.restype = ctypes.c_int
    L.oc_kick_sample.argtypes = L.oc_kick.argtypes + [ctypes.c_int64, ctypes.c_int64]
    L.oc_kick_sample.restype = ctypes.c_int
    L.oc_lists_free.argtypes = [ctypes.POINTER(Lists)]
    L.oc_drift.argtypes = [ctypes.c_int64, _U32, _U32, _U32, _F, _I8, ctypes.c_int, ctypes.c_float, ctypes.c_float]
    L.oc_drift.restype = None
    L.oc_direct_ewald.argtypes = [ctypes.c_int64, _U32, _U32, _U32, ctypes.c_int64, _I64, ctypes.c_double, ctypes.c_double,
                                  _D, _D, _D, _D, ctypes.c_int]
    L.oc_direct_ewald.restype = None
    L.oc_ewald_psi.argtypes = [_D, _D, _D]
    L.oc_ewald_psi.restype = None
    L.oc_ewald_psi0.restype = ctypes.c_double


def _p(a, t):
    return a.ctypes.data_as(t)


def _dv(x):
    return np.ascontiguousarray(x, dtype=np.float64)


def greens(X):
    D = np.zeros(50)
    lib().oc_greens(_p(D, _D), _p(_dv(X), _D))
    return D


def ewald_greens(X):
    D = np.zeros(50)
    lib().oc_ewald_greens(_p(D, _D), _p(_dv(X), _D))
    return D


def m2l(M, D, pout=7, do_phi=True, L=None):
    L = np.zeros(50 if pout == 7 else 4) if L is None else _dv(L).copy()
    lib().oc_m2l(_p(L, _D), pout, _p(_dv(M), _D), _p(_dv(D), _D), int(do_phi))
    return L


def l2l(La, X, do_phi=True):
    Lb = np.zeros(50)
    lib().oc_l2l(_p(Lb, _D), _p(_dv(La), _D), _p(_dv(X), _D), int(do_phi))
    return Lb


def l2p(La, X, do_phi=True):
    L4 = np.zeros(4)
    lib().oc_l2p(_p(L4, _D), _p(_dv(La), _D), _p(_dv(X), _D), int(do_phi))
    return L4


def p2m(X):
    M = np.zeros(37)
    lib().oc_p2m(_p(M, _D), _p(_dv(X), _D))
    return M


def m2m(Ma, X):
    Mb = np.zeros(37)
    lib().oc_m2m(_p(Mb, _D), _p(_dv(Ma), _D), _p(_dv(X), _D))
    return Mb


def gsoft(q2, h):
    f = ctypes.c_double()
    p = ctypes.c_double()
    lib().oc_gsoft(ctypes.byref(f), ctypes.byref(p), q2, 1.0 / h, 1.0 / h ** 2, 1.0 / h ** 3)
    return f.value, p.value


def trless_index(a, b, c, P):
    return lib().oc_trless_index(a, b, c, P)


def ewald_tables():
    L = lib()
    L.oc_init()
    nr = L.oc_ewald_nreal()
    nf = L.oc_ewald_nfour()
    real = np.zeros((nr, 3))
    four = np.zeros((nf, 3))
    tmp = np.zeros(3)
    for i in range(nr):
        L.oc_ewald_real_index(i, _p(tmp, _D))
        real[i] = tmp
    for i in range(nf):
        L.oc_ewald_four_index(i, _p(tmp, _D))
        four[i] = tmp
    return real, four


def ewald_psi(X):
    psi = ctypes.c_double()
    g = np.zeros(3)
    lib().oc_ewald_psi(_p(_dv(X), _D), ctypes.byref(psi), _p(g, _D))
    return psi.value, g


def ewald_psi0():
    return lib().oc_ewald_psi0()


def distance(a, b):
    return lib().oc_distance(int(a) & 0xFFFFFFFF, int(b) & 0xFFFFFFFF)


def fixed32_from_double(x):
    return lib().oc_fixed32_from_double(float(x))


def self_phi():
    return lib().oc_self_phi()


class OracleTree:
    """Tree built by the oracle over copies of the given positions."""

    def __init__(self, x, y, z, bucket, theta, h, nthreads=8):
        x0 = np.array(x, dtype=np.uint32)
        y0 = np.array(y, dtype=np.uint32)
        z0 = np.array(z, dtype=np.uint32)
        self.n = len(x0)
        cap = 0
        for _ in range(6):
            self.x, self.y, self.z = x0.copy(), y0.copy(), z0.copy()
            self.id = np.arange(self.n, dtype=np.int32)
            self.t = Tree()
            rc = lib().oc_tree_build(ctypes.byref(self.t), self.n, _p(self.x, _U32), _p(self.y, _U32), _p(self.z, _U32),
                                     _p(self.id, _I32), int(bucket), float(theta), float(h), int(nthreads), int(cap))
            if rc == 2:
                cap = 2 * self.t.cap
                lib().oc_tree_free(ctypes.byref(self.t))
                continue
            break
        if rc:
            raise RuntimeError("oracle tree build failed: %d" % rc)
        self.nnodes = self.t.nnodes
        self.root = self.t.root
        arr = np.ctypeslib.as_array(self.t.nodes, shape=(self.nnodes,))
        self.nodes = arr

    def node_arrays(self):
        """Structured numpy copy of the node table."""
        dt = np.dtype([("box_begin", "<i4", 3), ("box_end", "<i4", 3), ("child", "<i4", 2), ("pbegin", "<i4"),
                       ("pend", "<i4"), ("pos", "<u4", 3), ("radius", "<f4"), ("leaf", "<i4"), ("depth", "<i4"),
                       ("multi", "<f8", 37)])
        buf = (ctypes.c_char * (self.nnodes * ctypes.sizeof(Node))).from_address(ctypes.addressof(self.t.nodes.contents))
        return np.frombuffer(bytes(buf), dtype=dt).copy()

    def __del__(self):
        try:
            lib().oc_tree_free(ctypes.byref(self.t))
        except Exception:
            pass


def kick(tree, params, vel, rung, want_forces=True, want_lists=False, nthreads=8, sample=None):
    """Oracle kick over a tree.  vel: (n,3) float32 in tree order (updated in
    place), rung: int8 (updated in place).  Returns (forces dict, kick_return,
    lists)."""
    n = tree.n
    vel = np.ascontiguousarray(vel, dtype=np.float32)
    assert vel.shape == (n, 3) and rung.dtype == np.int8
    g = [np.zeros(n) for _ in range(4)] if want_forces else None
    ret = KickReturn()
    lists = Lists() if want_lists else None
    sb, se = (0, 2 ** 62) if sample is None else sample
    rc = lib().oc_kick_sample(ctypes.byref(tree.t), ctypes.byref(params), _p(tree.x, _U32), _p(tree.y, _U32), _p(tree.z, _U32),
                              _p(vel, _F), _p(rung, _I8), *([_p(a, _D) for a in g] if g else [None] * 4), ctypes.byref(ret),
                              ctypes.byref(lists) if lists is not None else None, int(nthreads), int(sb), int(se))
    if rc:
        raise RuntimeError("oracle kick failed: %d" % rc)
    out_lists = None
    if lists is not None:
        out_lists = []
        for q in range(5):
            off = np.ctypeslib.as_array(lists.off[q], shape=(tree.nnodes + 1,)).copy()
            tot = int(off[-1])
            idx = np.ctypeslib.as_array(lists.idx[q], shape=(max(tot, 1),))[:tot].copy()
            out_lists.append((off, idx))
        lib().oc_lists_free(ctypes.byref(lists))
    forces = None
    if g:
        forces = {"gx": g[0], "gy": g[1], "gz": g[2], "phi": g[3]}
    return forces, ret, out_lists


def drift(x, y, z, vel, rung, rung_level, a, dt):
    x = np.array(x, dtype=np.uint32)
    y = np.array(y, dtype=np.uint32)
    z = np.array(z, dtype=np.uint32)
    vel = np.ascontiguousarray(vel, dtype=np.float32)
    rung = np.ascontiguousarray(rung, dtype=np.int8)
    lib().oc_drift(len(x), _p(x, _U32), _p(y, _U32), _p(z, _U32), _p(vel, _F), _p(rung, _I8), int(rung_level), float(a),
                   float(dt))
    return x, y, z


def direct_ewald(x, y, z, sinks, h, GM, nthreads=8):
    x = np.ascontiguousarray(x, dtype=np.uint32)
    y = np.ascontiguousarray(y, dtype=np.uint32)
    z = np.ascontiguousarray(z, dtype=np.uint32)
    sinks = np.ascontiguousarray(sinks, dtype=np.int64)
    out = [np.zeros(len(sinks)) for _ in range(4)]
    lib().oc_direct_ewald(len(x), _p(x, _U32), _p(y, _U32), _p(z, _U32), len(sinks), _p(sinks, _I64), float(h), float(GM),
                          *[_p(a, _D) for a in out], int(nthreads))
    return {"gx": out[0], "gy": out[1], "gz": out[2], "phi": out[3]}


def default_params(theta, h, GM, eta=None, do_phi=True):
    """kick_params of the reference's top-level first kick (src/driver.cpp:481-508)."""
    p = KickParams()
    p.sign = 1.0
    p.min_level = 0
    p.min_rung = 0
    p.a = 1.0
    p.t0 = 1.0
    p.theta = theta
    p.h = h
    p.eta = eta if eta is not None else 0.25 / np.sqrt(693.0 / 256.0)
    p.GM = GM
    p.vel_init = False
    p.save_force = True
    p.first_call = True
    p.gpu = True
    p.node_load = 0.0
    p.max_dt = 1e30
    p.ascending = True
    p.descending = True
    p.top = True
    p.do_phi = do_phi
    return p
