"""CPU checks of the product library's host side: expansion layout against the
oracle, tree build against the oracle's restatement (bit-exact on integer and
fixed32 fields), and the exported C-ABI."""
import ctypes
import re
import subprocess

import numpy as np
import pytest


def _rand_multipole(oc, rng, k=5):
    return sum(oc.p2m(rng.uniform(-2e-3, 2e-3, 3)) for _ in range(k))


@pytest.mark.parametrize("seed", [0, 1])
def test_generated_kernels_match_oracle(ct, oc, seed):
    """The straight-line kernels (host double instantiation) equal the oracle's
    generic multi-index loops."""
    rng = np.random.default_rng(seed)
    for _ in range(5):
        X = rng.uniform(-0.3, 0.3, 3)
        np.testing.assert_allclose(ct.fmm_host("greens", X), oc.greens(X), rtol=1e-12, atol=1e-14 * np.max(np.abs(oc.greens(X))))
        E = oc.ewald_greens(X)
        # product tables are FP32; the oracle's FP64
        assert np.max(np.abs(ct.fmm_host("ewald_greens", X) - E)) <= 1e-7 * np.max(np.abs(E))
        M = _rand_multipole(oc, rng)
        D = oc.greens(rng.uniform(0.03, 0.1, 3))
        ref = oc.m2l(M, D)
        assert np.max(np.abs(ct.fmm_host("M2L", np.concatenate([M, D])) - ref)) <= 1e-12 * np.max(np.abs(ref))
        ref2 = oc.m2l(M, D, pout=2)
        assert np.max(np.abs(ct.fmm_host("M2L2", np.concatenate([M, D])) - ref2)) <= 1e-12 * np.max(np.abs(ref2))
        Y = rng.uniform(-3e-3, 3e-3, 3)
        assert np.max(np.abs(ct.fmm_host("L2L", np.concatenate([ref, Y])) - oc.l2l(ref, Y))) <= 1e-12 * np.max(np.abs(ref))
        l2p = oc.l2p(ref, Y)
        assert np.max(np.abs(ct.fmm_host("L2P", np.concatenate([ref, Y])) - l2p)) <= 1e-12 * np.max(np.abs(l2p))
        P = rng.uniform(-2e-3, 2e-3, 3)
        assert np.max(np.abs(ct.fmm_host("P2M", P) - oc.p2m(P))) <= 1e-12 * np.max(np.abs(oc.p2m(P)))
        m2m = oc.m2m(M, Y)
        assert np.max(np.abs(ct.fmm_host("M2M", np.concatenate([M, Y])) - m2m)) <= 1e-12 * np.max(np.abs(m2m))


def _node_map(pn, on, oroot):
    m = -np.ones(len(pn), dtype=np.int64)
    stack = [(0, oroot)]
    while stack:
        a, b = stack.pop()
        m[a] = b
        if not pn[a]["leaf"]:
            stack.append((int(pn[a]["child"][0]), int(on[b]["child"][0])))
            stack.append((int(pn[a]["child"][1]), int(on[b]["child"][1])))
    return m


@pytest.mark.parametrize("kind,n,theta,bucket", [("uniform", 60000, 0.7, 128), ("clustered", 60000, 0.55, 64),
                                                  ("clustered", 20000, 0.4, 32), ("uniform", 0, 0.7, 128),
                                                  ("uniform", 1, 0.7, 128), ("uniform", 3, 0.55, 2)])
def test_tree_matches_oracle(ct, oc, kind, n, theta, bucket):
    from cosmictiger_amd import workloads
    if kind == "uniform":
        x, y, z = workloads.uniform_box(n, 7)
    else:
        x, y, z = workloads.nfw_box(n, 7, clustered_fraction=0.7, nhalos=25)
    h = ct.softening(max(n, 1))
    xs, ys, zs, perm, pn = ct.tree_build_host(x, y, z, theta, bucket, h)
    tr = oc.OracleTree(x, y, z, bucket, theta, h)
    on = tr.node_arrays()
    assert len(pn) == len(on)
    m = _node_map(pn, on, tr.root)
    assert np.all(m >= 0) and len(np.unique(m)) == len(m)
    b = on[m]
    for f in ("pbegin", "pend", "leaf", "depth", "radius"):
        assert np.array_equal(pn[f], b[f]), f
    for f in ("pos", "box_begin", "box_end"):
        assert np.array_equal(pn[f], b[f]), f
    # parents and children consistent, breadth-first depth order
    assert np.all(np.diff(pn["depth"]) >= 0)
    for i in range(1, len(pn)):
        p = pn["parent"][i]
        assert i in (pn["child"][p][0], pn["child"][p][1])
    # leaf membership (by original index) identical
    for i in np.nonzero(pn["leaf"])[0]:
        a = np.sort(perm[pn["pbegin"][i]:pn["pend"][i]])
        c = np.sort(tr.id[b["pbegin"][i]:b["pend"][i]])
        assert np.array_equal(a, c)
    # positions permuted consistently
    assert np.array_equal(xs, x[perm]) and np.array_equal(ys, y[perm]) and np.array_equal(zs, z[perm])
    # multipoles: FP tolerance, normalised by mass (1000 r)^l
    deg = np.zeros(37, int)
    for l in range(6):
        for a in range(l, -1, -1):
            for bb in range(l - a, -1, -1):
                c = l - a - bb
                if c <= 1 or (a, bb, c) == (0, 0, 2):
                    deg[oc.trless_index(a, bb, c, 6)] = l
    nz = b["multi"][:, 0] > 0
    scale = b["multi"][nz, 0:1] * (1000.0 * np.maximum(b["radius"][nz], 1e-9))[:, None] ** deg[None, :]
    if nz.any():
        assert np.max(np.abs(pn["multi"][nz] - b["multi"][nz]) / scale) <= 1e-6


def test_tree_invariants(ct):
    """Leaves hold <= bucket particles and satisfy the Ewald size rule; ranges
    partition [0, n); every member lies within radius of the centre."""
    from cosmictiger_amd import workloads
    n, theta, bucket = 40000, 0.55, 64
    x, y, z = workloads.nfw_box(n, 3, clustered_fraction=0.8, nhalos=15)
    h = ct.softening(n)
    xs, ys, zs, perm, nodes = ct.tree_build_host(x, y, z, theta, bucket, h)
    assert np.array_equal(np.sort(perm), np.arange(n))
    leaves = nodes[nodes["leaf"] == 1]
    assert np.all(leaves["pend"] - leaves["pbegin"] <= bucket)
    leaves = leaves[leaves["pend"] > leaves["pbegin"]]
    order = np.argsort(leaves["pbegin"])
    assert leaves["pbegin"][order][0] == 0 and leaves["pend"][order][-1] == n
    assert np.all(leaves["pbegin"][order][1:] == leaves["pend"][order][:-1])
    for nd in nodes[::7]:
        k = slice(nd["pbegin"], nd["pend"])
        if nd["pend"] == nd["pbegin"]:
            continue
        d = [((q[k].astype(np.int64) - int(c)) + 2 ** 31) % 2 ** 32 - 2 ** 31 for q, c in ((xs, nd["pos"][0]), (ys, nd["pos"][1]), (zs, nd["pos"][2]))]
        r = np.sqrt(sum(v.astype(np.float64) ** 2 for v in d)) / 2 ** 32
        assert r.max() <= nd["radius"] * (1 + 1e-6) + 2e-9


def test_abi_exports_every_declared_symbol(ct):
    text = open(ct.HEADER_PATH).read()
    names = sorted(set(re.findall(r"^\s*(?:int|uint64_t|const char\*)\s+(ct_\w+)\s*\(", text, flags=re.M)))
    assert len(names) >= 15
    L = ct.lib()
    for nm in names:
        assert hasattr(L, nm), nm


def test_abi_struct_layouts(ct):
    # kick_params: 9 scalars, 4 bools, 2 floats, 4 bools (cosmictiger/kick.hpp:138-157)
    assert ctypes.sizeof(ct.KickParams) == 52
    assert ct.KickParams.ascending.offset == 48
    assert ctypes.sizeof(ct.KickReturn) == 64
    assert ctypes.sizeof(ct.Node) == 216


def test_abi_struct_layouts_match_the_header(ct, tmp_path):
    """The ctypes mirrors have the C layout of the header's structs: a C
    program compiled against the header prints sizeof and every offsetof."""
    structs = {"ct_kick_params": ct.KickParams, "ct_kick_return": ct.KickReturn, "ct_step_params": ct.StepParams,
               "ct_step_return": ct.StepReturn}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "%s"' % ct.HEADER_PATH, "int main(void) {"]
    for cname, py in structs.items():
        lines.append('printf("%s %%zu\\n", sizeof(%s));' % (cname, cname))
        for f, _ in py._fields_:
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (cname, f, cname, f))
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.check_call(["gcc", "-std=c11", "-o", str(exe), str(src)])
    got = dict(l.split() for l in subprocess.check_output([str(exe)]).decode().split("\n") if l)
    for cname, py in structs.items():
        assert int(got[cname]) == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(got["%s.%s" % (cname, f)]) == getattr(py, f).offset, (cname, f)


def test_lc_entry_dtype_matches_the_header(ct, tmp_path):
    """ct.LC_ENTRY (the numpy record the lightcone entries are read into) has
    the layout of the header's ct_lc_entry."""
    dt = ct.LC_ENTRY
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "%s"' % ct.HEADER_PATH, "int main(void) {",
             'printf("size %zu\\n", sizeof(ct_lc_entry));']
    for f in dt.names:
        lines.append('printf("%s %%zu\\n", offsetof(ct_lc_entry, %s));' % (f, f))
    lines.append("return 0; }")
    src = tmp_path / "lc.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "lc"
    subprocess.check_call(["gcc", "-std=c11", "-o", str(exe), str(src)])
    got = dict(l.split() for l in subprocess.check_output([str(exe)]).decode().split("\n") if l)
    assert int(got["size"]) == dt.itemsize
    for f in dt.names:
        assert int(got[f]) == dt.fields[f][1], f


def test_time_counter_closed_form(ct):
    """inc / min_rung of the driver's integer time (cosmictiger/time.hpp:28-43):
    over one top-level step at max_rung R the counter visits 2^R sub-steps,
    sub-step k having min_rung = R - (trailing zeros of k), 0 at k = 0, and
    the library's C versions agree with the Python ones."""
    L = ct.lib()
    for R in range(0, 7):
        t = 0
        seen = []
        while True:
            seen.append(ct.time_min_rung(t))
            assert L.ct_time_min_rung(t) == seen[-1]
            assert L.ct_time_inc(t, R) == ct.time_inc(t, R)
            t = ct.time_inc(t, R)
            if t == 0 or R == 0:
                break
        assert len(seen) == 2 ** R
        for k, m in enumerate(seen):
            tz = (k & -k).bit_length() - 1 if k else None
            assert m == (0 if k == 0 else R - tz)


def test_abi_errors_without_device(ct):
    """Calls that need a device fail with an error code and a message when
    none has been selected; they never abort."""
    L = ct.lib()
    if ct._inited_device is not None:
        pytest.skip("device already initialised in this process")
    h = ctypes.c_void_p()
    x = np.zeros(4, np.uint32)
    rc = L.ct_box_create(4, ct._ptr(x), ct._ptr(x), ct._ptr(x), None, None, ctypes.byref(h))
    assert rc == ct.CT_E_STATE
    assert b"ct_init" in L.ct_last_error()
    assert L.ct_kick(None, None, None, None) != 0
