"""Shared driver of the GPU-vs-oracle kick comparisons."""
import numpy as np


def node_map(pnodes, onodes, oroot):
    """Map product node index -> oracle node index by walking both trees."""
    m = -np.ones(len(pnodes), dtype=np.int64)
    stack = [(0, oroot)]
    while stack:
        a, b = stack.pop()
        m[a] = b
        if not pnodes[a]["leaf"]:
            stack.append((int(pnodes[a]["child"][0]), int(onodes[b]["child"][0])))
            stack.append((int(pnodes[a]["child"][1]), int(onodes[b]["child"][1])))
    return m


def run_pair(ct, oc, x, y, z, vel, rung, theta, bucket, h, GM, params_kw=None, nthreads=8, precision=None):
    """Kick the same box on the GPU and in the oracle.  Returns a dict of both
    sides' results, in original particle order where per-particle."""
    n = len(x)
    params_kw = params_kw or {}
    box = ct.Box(x, y, z, vel, rung)
    box.tree_build(theta, bucket, h)
    p = ct.default_kick_params(theta, h, GM)
    for k, v in params_kw.items():
        setattr(p, k, v)
    ret = box.kick(p, precision=precision)
    out = box.read()
    pnodes = box.nodes()
    lists = [box.lists(k) for k in range(5)]
    stats = box.stats()
    inv = np.empty(n, dtype=np.int64)
    inv[out["id"]] = np.arange(n)
    gpu = {k: (out[k][inv] if out[k].shape[0] == n else out[k]) for k in ("gx", "gy", "gz", "phi", "rung", "vel")}

    tr = oc.OracleTree(x, y, z, bucket, theta, h)
    onodes = tr.node_arrays()
    ovel = np.ascontiguousarray(vel[tr.id], dtype=np.float32) if vel is not None else np.zeros((n, 3), np.float32)
    orung = np.ascontiguousarray(rung[tr.id], dtype=np.int8) if rung is not None else np.zeros(n, np.int8)
    op = oc.default_params(theta, h, GM)
    for k, v in params_kw.items():
        setattr(op, k, v)
    F, oret, olists = oc.kick(tr, op, ovel, orung, want_lists=True, nthreads=nthreads)
    oinv = np.empty(n, dtype=np.int64)
    oinv[tr.id] = np.arange(n)
    ora = {"gx": F["gx"][oinv], "gy": F["gy"][oinv], "gz": F["gz"][oinv], "phi": F["phi"][oinv], "rung": orung[oinv],
           "vel": ovel[oinv]}
    box.close()
    return {"gpu": gpu, "oracle": ora, "ret": ret, "oret": oret, "pnodes": pnodes, "onodes": onodes, "lists": lists,
            "olists": olists, "map": node_map(pnodes, onodes, tr.root), "stats": stats, "tree": tr}


def compare_lists(res):
    """Per node and list kind, the GPU list equals the oracle list (same
    entries in the same order after mapping node ids)."""
    m = res["map"]
    bad = 0
    for kind in range(5):
        poff, pidx = res["lists"][kind]
        ooff, oidx = res["olists"][kind]
        for a in range(len(res["pnodes"])):
            b = m[a]
            pl = m[pidx[poff[a]:poff[a + 1]]]
            ol = oidx[ooff[b]:ooff[b + 1]]
            if len(pl) != len(ol) or not np.array_equal(pl, ol):
                bad += 1
    return bad


def force_errors(res):
    g = np.stack([res["gpu"]["gx"], res["gpu"]["gy"], res["gpu"]["gz"]], 1).astype(np.float64)
    o = np.stack([res["oracle"]["gx"], res["oracle"]["gy"], res["oracle"]["gz"]], 1)
    err = np.linalg.norm(g - o, axis=1) / np.maximum(np.linalg.norm(o, axis=1), 1e-300)
    perr = np.abs(res["gpu"]["phi"] - res["oracle"]["phi"]) / np.sqrt(np.mean(res["oracle"]["phi"] ** 2))
    return err, perr
