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
    L.oc_vec2pix_nest_many.argtypes = [ctypes.c_int64, ctypes.c_int64, _D, _I64]
    L.oc_vec2pix_nest_many.restype = None
    L.oc_pix2vec_nest_many.argtypes = [ctypes.c_int64, ctypes.c_int64, _I64, _D]
    L.oc_pix2vec_nest_many.restype = None
    L.oc_lightcone.argtypes = [ctypes.c_int64, _U32, _U32, _U32, _F, _I8, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                               ctypes.c_float, ctypes.c_float, ctypes.c_int, ctypes.c_void_p, ctypes.c_int64]
    L.oc_lightcone.restype = ctypes.c_int64
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


def vec2pix_nest(nside, vecs):
    """HEALPix nested pixel of each row of vecs (src/healpix.cu:75-172)."""
    v = np.ascontiguousarray(np.atleast_2d(vecs), dtype=np.float64)
    out = np.zeros(len(v), dtype=np.int64)
    lib().oc_vec2pix_nest_many(int(nside), len(v), _p(v, _D), _p(out, _I64))
    return out


def pix2vec_nest(nside, pix):
    """Unit vector of the centre of each nested pixel (HEALPix pix2ang_nest)."""
    p = np.ascontiguousarray(np.atleast_1d(pix), dtype=np.int64)
    out = np.zeros((len(p), 3))
    lib().oc_pix2vec_nest_many(int(nside), len(p), _p(p, _I64), _p(out, _D))
    return out


# lc_entry (cosmictiger/lightcone.hpp:54-58) plus the crossing time
LC_ENTRY = np.dtype([("x", np.int32), ("y", np.int32), ("z", np.int32), ("vx", np.float32), ("vy", np.float32), ("vz", np.float32),
                     ("pix", np.int32), ("t", np.float64)], align=True)


def lightcone(x, y, z, vel, rung, rung_level, a, dt, tau0, tau1, tau_max, nside):
    """Lightcone crossings of the drift (src/drift.cu:75-131), as an array of
    LC_ENTRY in particle order, then image order (xi, yi, zi in {-1, 0})."""
    x = np.ascontiguousarray(x, dtype=np.uint32)
    y = np.ascontiguousarray(y, dtype=np.uint32)
    z = np.ascontiguousarray(z, dtype=np.uint32)
    vel = np.ascontiguousarray(vel, dtype=np.float32)
    rung = np.ascontiguousarray(rung, dtype=np.int8)
    args = (len(x), _p(x, _U32), _p(y, _U32), _p(z, _U32), _p(vel, _F), _p(rung, _I8), int(rung_level), float(a), float(dt), float(tau0),
            float(tau1), float(tau_max), int(nside))
    n = lib().oc_lightcone(*args, None, 0)
    out = np.zeros(max(n, 1), dtype=LC_ENTRY)
    lib().oc_lightcone(*args, out.ctypes.data_as(ctypes.c_void_p), n)
    return out[:n]


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


# ---------------------------------------------------------------------------
# active range and the hierarchical step (src/particles.cpp:218-235,583-649,
# 1020-1068; src/driver.cpp:382-593), restated over the oracle's tree build,
# kick and drift.  Particles are kept in the order the steps leave them in;
# tdrift[i] accumulates the drift time of each particle (moves with it).
# ---------------------------------------------------------------------------

RUNG_DT = [2.0 ** -r for r in range(64)]  # rung_dt (cosmictiger/particles.hpp:37)


class OracleBox:
    def __init__(self, x, y, z, vel, rung):
        self.x = np.array(x, dtype=np.uint32)
        self.y = np.array(y, dtype=np.uint32)
        self.z = np.array(z, dtype=np.uint32)
        self.n = len(self.x)
        self.vel = np.array(vel, dtype=np.float32).reshape(self.n, 3)
        self.rung = np.array(rung, dtype=np.int8)
        self.id = np.arange(self.n, dtype=np.int32)
        self.tdrift = np.zeros(self.n)
        self.rung_begin = 0
        self.stack = []
        self.forces = None  # forces of the last kick, by particle id

    def _permute_range(self, b, perm):
        for name in ("x", "y", "z", "vel", "rung", "id", "tdrift"):
            a = getattr(self, name)
            a[b:] = a[b:][perm]

    def sort_by_rung(self, min_rung):
        """particles_sort_by_rung (src/particles.cpp:1029-1068): the active
        range keeps the particles of rung >= min_rung, moved to its end."""
        if min_rung == 0:
            self.rung_begin = 0
            return 0
        b = self.rung_begin
        low = self.rung[b:] < min_rung
        perm = np.concatenate([np.nonzero(low)[0], np.nonzero(~low)[0]])
        self._permute_range(b, perm)
        self.rung_begin = b + int(np.count_nonzero(low))
        return self.rung_begin

    def push_rungs(self):
        self.stack.append(self.rung_begin)

    def pop_rungs(self):
        self.rung_begin = self.stack.pop()

    def set_minrung(self, m):
        np.maximum(self.rung, np.int8(m), out=self.rung)

    def rung_counts(self):
        return np.bincount(self.rung.astype(np.int64), minlength=0) if self.n else np.zeros(0, np.int64)

    def tree_kick(self, params, theta, bucket, h, nthreads=8):
        """tree_create over the active range, then kick() of that tree
        (src/driver.cpp:470-526)."""
        b = self.rung_begin
        T = OracleTree(self.x[b:], self.y[b:], self.z[b:], bucket, theta, h, nthreads=nthreads)
        perm = T.id.astype(np.int64)
        self._permute_range(b, perm)
        vel = np.ascontiguousarray(self.vel[b:])
        rung = np.ascontiguousarray(self.rung[b:])
        forces, ret, _ = kick(T, params, vel, rung, want_forces=True, nthreads=nthreads)
        self.vel[b:] = vel
        self.rung[b:] = rung
        self.forces = {"id": self.id[b:].copy(), **{k: v.copy() for k, v in forces.items()}}
        return ret

    def drift(self, level, a, dt):
        """cuda_drift over the active range (src/drift.cu:156-195)."""
        b = self.rung_begin
        sel = self.rung[b:] == level
        self.x[b:], self.y[b:], self.z[b:] = drift(self.x[b:], self.y[b:], self.z[b:], self.vel[b:], self.rung[b:], level, a, dt)
        self.tdrift[b:][sel] += np.float32(dt)

    def by_id(self):
        o = np.argsort(self.id)
        return {"x": self.x[o], "y": self.y[o], "z": self.z[o], "vel": self.vel[o], "rung": self.rung[o], "tdrift": self.tdrift[o]}


def step_hierarchical(box, minrung, max_rung, minrung0, scale, tau, t0, theta, bucket, h, GM, eta, max_dt=1e30, sign=1.0,
                      do_phi=True, nocrop=False, drag=0.0, nthreads=8, trace=None):
    """kick_step_hierarchical (src/driver.cpp:382-593) over an OracleBox, with
    the expansion factor held at `scale` for the drifts.  Returns
    (kick_return of the top kick with max_rung set, minrung after the call,
    number of kicks).  trace (a list) receives one tuple per kick:
    (level, ascending, top, descending, rung_begin, drifted)."""
    levels = list(range(max_rung, minrung, -1)) + [minrung]
    ascending = True
    top = False
    clip_top = False
    kr = None
    nk = 0
    li = 0
    while li < len(levels):
        level = levels[li]
        if level == minrung:
            ascending = False
            top = True
        else:
            top = False
        if not ascending and not top:
            box.push_rungs()
        if not ascending:
            box.sort_by_rung(level)
        if top and minrung == minrung0:
            counts = box.rung_counts()
            if len(counts) > minrung0 + 1:
                fast = int(np.sum(counts[minrung0 + 1:]))
                slow = int(counts[minrung0])
                if 3 * fast > slow:
                    clip_top = not nocrop
        p = KickParams()
        p.top = True if (clip_top and level == minrung0 + 1) else top
        p.sign = sign
        p.ascending = ascending or top
        p.descending = False if (clip_top and top) else ((not ascending) or top)
        p.vel_init = False
        p.node_load = 0.0
        p.gpu = True
        p.min_level = 9
        p.save_force = True
        p.GM = GM
        p.eta = eta
        p.h = h
        p.a = scale
        p.first_call = tau == 0.0
        p.min_rung = level
        p.do_phi = do_phi and top
        p.t0 = t0
        p.max_dt = max_dt
        p.theta = theta
        rb = box.rung_begin
        tkr = box.tree_kick(p, theta, bucket, h, nthreads=nthreads)
        nk += 1
        if clip_top and top:
            box.set_minrung(minrung0 + 1)
        if top:
            kr = tkr
        if clip_top and top:
            levels.append(minrung0 + 1)
        elif (not ascending) or top:
            kr.max_rung = tkr.max_rung
            if tkr.max_rung > level:
                levels.append(level + 1)
        drifted = ((not ascending) or top) and not (clip_top and top)
        if drifted:
            box.drift(level, scale, RUNG_DT[level] * t0)
        if trace is not None:
            trace.append((level, bool(ascending), bool(top), bool(p.descending), rb, bool(drifted)))
        if ascending and not top:
            box.pop_rungs()
        li += 1
    if clip_top:
        minrung += 1
    return kr, minrung, nk


def run_steps(box, ncalls, t0, theta, bucket, h, GM, eta, max_dt=1e30, scale=1.0, nocrop=True, step=None, trace=None, cycles=1):
    """The driver's sub-step loop (src/driver.cpp:907-1087): minrung from the
    integer time, one hierarchical step, max_rung from its kick_return, tau
    advanced by t0 / 2^max_rung, the time by inc().  step (default
    step_hierarchical over an OracleBox) is called as step(minrung,
    max_rung, minrung0, tau); it returns (kick_return, minrung after).
    Stops after ncalls calls or when the time has returned to 0 `cycles`
    times."""
    itime = 0
    tau = 0.0
    max_rung = 0
    minrung0 = 0
    calls = []
    for _ in range(ncalls):
        m = 64
        t = itime
        while (t & 1) == 0 and m != 0:
            m -= 1
            t >>= 1
        this_minrung = max(m, minrung0)
        if step is None:
            kr, om, _ = step_hierarchical(box, this_minrung, max_rung, minrung0, scale, tau, t0, theta, bucket, h, GM, eta,
                                          max_dt=max_dt, nocrop=nocrop, trace=trace)
        else:
            kr, om = step(this_minrung, max_rung, minrung0, tau)
        if om != this_minrung:
            minrung0 += 1
        max_rung = kr.max_rung
        calls.append((this_minrung, max_rung))
        tau += t0 / (1 << max_rung)
        if max_rung > 0:
            itime = (itime + (1 << (64 - max_rung))) & ((1 << 64) - 1)
        if itime == 0:
            cycles -= 1
            if cycles == 0:
                break
    return calls, tau


# ---------------------------------------------------------------------------
# the CPU SIMD kick (oracle/ct_cpukick.cpp): the reference's CPU kick path
# in AVX2 FP32, the timed CPU baseline
# ---------------------------------------------------------------------------

_CPK_PATH = os.path.join(_HERE, "_build", "libctcpukick.so")
_cpk = None


def cpk_lib():
    global _cpk
    if _cpk is None:
        if not os.path.exists(_CPK_PATH):
            build()
        lib()  # oc_tree_build lives there
        _cpk = ctypes.CDLL(_CPK_PATH)
        _cpk.cpk_kick.argtypes = [ctypes.POINTER(Tree), ctypes.POINTER(KickParams), _U32, _U32, _U32, _F, _I8, _F, _F, _F, _F,
                                  ctypes.POINTER(KickReturn), ctypes.c_int, ctypes.c_int64, ctypes.c_int64, ctypes.c_int]
        _cpk.cpk_kick.restype = ctypes.c_int
    return _cpk


def cpu_simd_kick(tree, params, vel, rung, nthreads=8, sample=None, cpu_rules=False):
    """The CPU SIMD kick over an OracleTree (particles in tree order): vel
    (n,3) float32 and rung int8 updated in place.  Returns (forces as float32
    arrays, kick_return).  sample=(b, e) kicks only the particles [b, e).
    cpu_rules: the CPU path's own PC / CP test instead of the GPU path's."""
    n = tree.n
    vel = np.ascontiguousarray(vel, dtype=np.float32)
    assert vel.shape == (n, 3) and rung.dtype == np.int8
    g = [np.zeros(n, np.float32) for _ in range(4)]
    ret = KickReturn()
    sb, se = (0, n) if sample is None else sample
    rc = cpk_lib().cpk_kick(ctypes.byref(tree.t), ctypes.byref(params), _p(tree.x, _U32), _p(tree.y, _U32), _p(tree.z, _U32), _p(vel, _F),
                            _p(rung, _I8), *[_p(a, _F) for a in g], ctypes.byref(ret), int(nthreads), int(sb), int(se),
                            int(bool(cpu_rules)))
    if rc:
        raise RuntimeError("CPU SIMD kick failed: %d" % rc)
    return {"gx": g[0], "gy": g[1], "gz": g[2], "phi": g[3]}, ret
