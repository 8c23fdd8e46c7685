#!/usr/bin/env python3
"""Generator of the straight-line Cartesian FMM kernels used by the kick.

Writes ``cosmictiger_amd/csrc/ct_expansions.hpp``: C++ templates that compile
for the host (double, used by the tree build) and for gfx950 device code
(float, used by the kick kernels).  Each function is written out as plain
scalar code over named locals so that the compiler keeps everything in VGPRs.

Storage layout (the drop-in contract with the reference's tensors):

* ``tensor_trless_sym<T,P>`` (cosmictiger/tensor.hpp:63-136) stores the
  components (a,b,c) with a+b+c < P and c <= 1, plus (0,0,2); the index is
  ``trless_index`` (cosmictiger/tensor.hpp:317-319).  Components with c >= 2
  (other than (0,0,2)) are recovered from the trace rule
  T(a,b,c) = -(T(a+2,b,c-2) + T(a,b+2,c-2)) (tensor.hpp:104-110), i.e. the
  stored tensor is traceless from degree 3 up and keeps its trace at degree 2.
* Local expansion L: P=7 (50 floats, degrees 0..6); multipole M: P=6
  (37 floats, degrees 0..5); M2P output: P=2 (4 floats: phi, gx, gy, gz).
* All lengths are multiplied by SCALE_FACTOR=1000 inside the kernels
  (cosmictiger/fmm_kernels.hpp:33-49), so L and D hold derivatives in the
  scaled frame and M holds moments in the scaled frame.

Mathematics (written from the Cartesian Taylor expansion, not from the
reference's generator):

* Green's function f(X) = -1/|X|.  D_n = d^n f (n a multi-index).  With
  u = X/r:  d^n (1/r) = (-1)^l/r^(l+1) sum_m (-1)^|m| (2l-2|m|-1)!!
  n!/(2^|m| m! (n-2m)!) u^(n-2m), l = |n|.
* Radial functions g(r): d^n g = sum_m n!/(2^|m| m! (n-2m)!) x^(n-2m)
  g_(l-|m|)(r) with g_k = (1/r d/dr)^k g.  Used for the Ewald correction.
* P2M: M_n = moments of y = -(x_particle - centre) (cosmictiger/
  fmm_kernel_codegen.cpp:1473-1475 negate X); degree <= 2 raw, degree >= 3
  the traceless projection
  TF(B)_n = sum_m (-1)^|m| (2l-2|m|-1)!!/(2l-1)!! n!/(2^|m| m!(n-2m)!)
            Tr^|m|(B)_(n-2m),  Tr^j(B)_p = sum_{|k|=j} j!/k! B_(p+2k).
* M2L: L_n += sum_m M_m D_(n+m) / m!  with |m| <= min(6-|n|, 5)
  (fmm_kernel_codegen.cpp:1310-1345 ranges).
* L2L / L2P: Lb_n = sum_k La_(n+k) x^k / k!  (Taylor shift), phi row only
  when do_phi (fmm_kernel_codegen.cpp:750-775).
* M2M: binomial shift of raw moments, then the projection above.

Usage: python gen_expansions.py [output_path]
"""
import math
import os
import sys
from fractions import Fraction

ORDER = 7
MORDER = ORDER - 1
SCALE = 1000.0
EWALD_ALPHA = 2.0


def dfact(n):
    r = 1
    while n > 1:
        r *= n
        n -= 2
    return r


def fact(n):
    return math.factorial(n)


def mfact(k):
    return fact(k[0]) * fact(k[1]) * fact(k[2])


def deg(n):
    return n[0] + n[1] + n[2]


def add(a, b):
    return (a[0] + b[0], a[1] + b[1], a[2] + b[2])


def sub(a, b):
    return (a[0] - b[0], a[1] - b[1], a[2] - b[2])


def scal(a, s):
    return (a[0] * s, a[1] * s, a[2] * s)


def trless_set(P):
    out = []
    for l in range(P):
        for a in range(l, -1, -1):
            for b in range(l - a, -1, -1):
                c = l - a - b
                if c <= 1 or (a, b, c) == (0, 0, 2):
                    out.append((a, b, c))
    return out


def trless_index(n, P):
    a, b, c = n
    return (a + b) * (a + b + 1) // 2 + b + (P * (P + 1) // 2) * (c == 1) + P * P * (c == 2)


def sym_set(P):
    out = []
    for l in range(P):
        for a in range(l, -1, -1):
            for b in range(l - a, -1, -1):
                out.append((a, b, l - a - b))
    return out


def half_multi(n):
    """All multi-indices m with 2m <= n componentwise."""
    out = []
    for i in range(n[0] // 2 + 1):
        for j in range(n[1] // 2 + 1):
            for k in range(n[2] // 2 + 1):
                out.append((i, j, k))
    return out


def multis_of_degree(j):
    out = []
    for a in range(j, -1, -1):
        for b in range(j - a, -1, -1):
            out.append((a, b, j - a - b))
    return out


def full_expr(n, P):
    """Full component n of a stored trless tensor, as {stored index: int}."""
    a, b, c = n
    if c >= 2 and (a, b, c) != (0, 0, 2):
        e1 = full_expr((a + 2, b, c - 2), P)
        e2 = full_expr((a, b + 2, c - 2), P)
        out = {}
        for k, v in list(e1.items()) + list(e2.items()):
            out[k] = out.get(k, 0) - v
        return {k: v for k, v in out.items() if v != 0}
    return {trless_index(n, P): 1}


def radial_coef(n, m):
    """n!/(2^|m| m! (n-2m)!)."""
    return Fraction(mfact(n), (2 ** deg(m)) * mfact(m) * mfact(sub(n, scal(m, 2))))


def greens_coef(n, m):
    """Integer coefficient K(n,m) of u^(n-2m) in r^(l+1) (-1)^l d^n(1/r)."""
    l = deg(n)
    k = (-1) ** deg(m) * dfact(2 * l - 2 * deg(m) - 1) * radial_coef(n, m)
    assert k.denominator == 1
    return int(k)


def tf_coef(n, m):
    """Traceless projection coefficient of Tr^|m|(B)_(n-2m)."""
    l = deg(n)
    return Fraction((-1) ** deg(m) * dfact(2 * l - 2 * deg(m) - 1), dfact(2 * l - 1)) * radial_coef(n, m)


def lit(x):
    x = float(x)
    if x == int(x) and abs(x) < 1e15:
        return "T(%d)" % int(x)
    return "T(%s)" % repr(x)


def vname(prefix, n):
    return "%s_%d_%d_%d" % (prefix, n[0], n[1], n[2])


class Emitter:
    def __init__(self):
        self.lines = []
        self.ind = 0
        self.flops = 0

    def __call__(self, s=""):
        self.lines.append("\t" * self.ind + s if s else "")

    def text(self):
        return "\n".join(self.lines) + "\n"


def emit_monomials(E, prefix, coords, needed):
    """Monomials coords^n for all n in `needed` (closed under lowering)."""
    have = {(0, 0, 0): None}
    order = sorted(set(needed), key=lambda n: (deg(n), n))
    for n in order:
        if deg(n) == 0:
            continue
        if deg(n) == 1:
            d = [i for i in range(3) if n[i] == 1][0]
            have[n] = vname(prefix, n)
            if coords[d] != vname(prefix, n):
                E("const T %s = %s;" % (vname(prefix, n), coords[d]))
            continue
        # split into two halves of nearly equal degree for accuracy
        best = None
        for k in sorted(have.keys(), key=lambda q: -deg(q)):
            if deg(k) == 0:
                continue
            j = sub(n, k)
            if min(j) < 0 or deg(j) == 0:
                continue
            if j in have:
                score = abs(deg(k) - deg(j))
                if best is None or score < best[0]:
                    best = (score, k, j)
        assert best is not None, n
        _, k, j = best
        E("const T %s = %s * %s;" % (vname(prefix, n), vname(prefix, k), vname(prefix, j)))
        E.flops += 1
        have[n] = vname(prefix, n)
    return have


def closure(ns):
    """Close a set of multi-indices under lowering (for monomial chains)."""
    out = set()
    stack = list(ns)
    while stack:
        n = stack.pop()
        if n in out:
            continue
        out.add(n)
        for d in range(3):
            if n[d] > 0:
                stack.append(tuple(n[i] - (1 if i == d else 0) for i in range(3)))
    return out


def term_sum(E, target, terms, init=None):
    """target = init + sum coef*var, accumulated with FMAs (terms in order)."""
    first = init is None
    expr = init
    for coef, var in terms:
        if coef == 0:
            continue
        if first:
            if coef == 1:
                expr = var
            elif coef == -1:
                expr = "-" + var
                E.flops += 1
            else:
                expr = "%s * %s" % (lit(coef), var)
                E.flops += 1
            first = False
            E("%s = %s;" % (target, expr))
        else:
            if coef == 1:
                E("%s += %s;" % (target, var))
                E.flops += 1
            elif coef == -1:
                E("%s -= %s;" % (target, var))
                E.flops += 1
            else:
                E("%s = ct_fma(%s, %s, %s);" % (target, lit(coef), var, target))
                E.flops += 2
    if first:
        E("%s = T(0);" % target)


def gen_greens():
    E = Emitter()
    P = ORDER
    E("// D_n = d^n(-1/r) at X (box units, scaled by 1000 inside), stored trless order 7.")
    E("// Returns the flop count of this kernel.")
    E("template<class T>")
    E("CT_HD inline int greens(T* __restrict__ D, T X0, T X1, T X2) {")
    E.ind += 1
    E("X0 *= T(%r); X1 *= T(%r); X2 *= T(%r);" % (SCALE, SCALE, SCALE))
    E("const T r2 = ct_fma(X0, X0, ct_fma(X1, X1, X2 * X2));")
    E("const T rinv = ct_select(r2 > T(0), ct_rsqrt(r2), T(0));")
    E.flops += 3 + 5 + 2
    E("const T u_1_0_0 = X0 * rinv;")
    E("const T u_0_1_0 = X1 * rinv;")
    E("const T u_0_0_1 = X2 * rinv;")
    E.flops += 3
    outs = trless_set(P)
    need = set()
    for n in outs:
        for m in half_multi(n):
            need.add(sub(n, scal(m, 2)))
    need = closure(need)
    need.discard((0, 0, 0))
    emit_monomials(E, "u", ["u_1_0_0", "u_0_1_0", "u_0_0_1"], sorted(need))
    # -(-1)^l / r^(l+1)
    E("const T nr0 = -rinv;")
    E.flops += 1
    for l in range(1, P):
        E("const T nr%d = -nr%d * rinv;" % (l, l - 1))
        E.flops += 2
    for n in outs:
        l = deg(n)
        terms = []
        for m in sorted(half_multi(n), key=lambda q: -deg(q)):
            p = sub(n, scal(m, 2))
            var = "T(1)" if deg(p) == 0 else vname("u", p)
            terms.append((greens_coef(n, m), var))
        E("{")
        E.ind += 1
        E("T acc;")
        if len(terms) == 1 and terms[0][1] == "T(1)":
            E("acc = %s;" % lit(terms[0][0]))
        else:
            term_sum(E, "acc", [(c, v) for c, v in terms])
        E("D[%d] = acc * nr%d;" % (trless_index(n, P), l))
        E.flops += 1
        E.ind -= 1
        E("}")
    E("return CT_FLOPS_GREENS;")
    E.ind -= 1
    E("}")
    return E


def gen_m2l(Pout, name):
    E = Emitter()
    P = ORDER
    PM = MORDER
    outs = trless_set(Pout)
    E("// L_n += sum_m M_m D_(n+m)/m!  (L: trless order %d, M: order 6, D: order 7)." % Pout)
    E("template<class T>")
    E("CT_HD inline int %s(T* __restrict__ L, const T* __restrict__ M, const T* __restrict__ D, bool do_phi) {" % name)
    E.ind += 1
    used_m = set()
    used_d = set()
    rows = {}
    for n in outs:
        l = deg(n)
        mmax = min(P - 1 - l, PM - 1)
        row = []
        for j in range(mmax, -1, -1):
            for m in multis_of_degree(j):
                row.append(m)
                used_m.add(m)
                used_d.add(add(n, m))
        rows[n] = row
    for m in sorted(used_m, key=lambda q: (deg(q), q)):
        ex = full_expr(m, PM)
        w = Fraction(1, mfact(m))
        parts = []
        for k, v in sorted(ex.items()):
            parts.append((v, "M[%d]" % k))
        E("T %s;" % vname("Mw", m))
        term_sum(E, vname("Mw", m), parts)
        if w != 1:
            E("%s *= %s;" % (vname("Mw", m), lit(float(w))))
            E.flops += 1
    for p in sorted(used_d, key=lambda q: (deg(q), q)):
        ex = full_expr(p, P)
        if len(ex) == 1 and list(ex.values())[0] == 1:
            E("const T %s = D[%d];" % (vname("Df", p), list(ex.keys())[0]))
            continue
        E("T %s;" % vname("Df", p))
        term_sum(E, vname("Df", p), [(v, "D[%d]" % k) for k, v in sorted(ex.items())])
    for n in outs:
        idx = trless_index(n, Pout)
        if deg(n) == 0:
            E("if (do_phi) {")
        else:
            E("{")
        E.ind += 1
        E("T acc = L[%d];" % idx)
        for m in rows[n]:
            E("acc = ct_fma(%s, %s, acc);" % (vname("Mw", m), vname("Df", add(n, m))))
            E.flops += 2
        E("L[%d] = acc;" % idx)
        E.ind -= 1
        E("}")
    E("return %s;" % ("CT_FLOPS_" + name.upper()))
    E.ind -= 1
    E("}")
    return E


def gen_l2l(two):
    E = Emitter()
    P = ORDER
    if two:
        outs = trless_set(2)
        name = "L2P"
        E("// Evaluate phi and grad phi (4 floats) of a trless order-7 local expansion at offset X.")
    else:
        outs = trless_set(P)
        name = "L2L"
        E("// Shift a trless order-7 local expansion by X (Lb must not alias La).")
    E("template<class T>")
    E("CT_HD inline int %s(T* __restrict__ Lb, const T* __restrict__ La, T X0, T X1, T X2, bool do_phi) {" % name)
    E.ind += 1
    E("X0 *= T(%r); X1 *= T(%r); X2 *= T(%r);" % (SCALE, SCALE, SCALE))
    E.flops += 3
    Pout = 2 if two else P
    used_k = set()
    used_p = set()
    rows = {}
    for n in outs:
        row = []
        for j in range(P - 1 - deg(n), 0, -1):
            for k in multis_of_degree(j):
                row.append(k)
                used_k.add(k)
                used_p.add(add(n, k))
        used_p.add(n)
        rows[n] = row
    emit_monomials(E, "x", ["X0", "X1", "X2"], sorted(closure(used_k) - {(0, 0, 0)}))
    for k in sorted(used_k, key=lambda q: (deg(q), q)):
        w = Fraction(1, mfact(k))
        if w != 1:
            E("const T %s = %s * %s;" % (vname("xw", k), lit(float(w)), vname("x", k)))
            E.flops += 1
        else:
            E("const T %s = %s;" % (vname("xw", k), vname("x", k)))
    for p in sorted(used_p, key=lambda q: (deg(q), q)):
        ex = full_expr(p, P)
        if len(ex) == 1 and list(ex.values())[0] == 1:
            E("const T %s = La[%d];" % (vname("Lf", p), list(ex.keys())[0]))
            continue
        E("T %s;" % vname("Lf", p))
        term_sum(E, vname("Lf", p), [(v, "La[%d]" % k) for k, v in sorted(ex.items())])
    for n in outs:
        idx = trless_index(n, Pout)
        E("{")
        E.ind += 1
        E("T acc = %s;" % vname("Lf", n))
        if deg(n) == 0:
            E("if (do_phi) {")
            E.ind += 1
        for k in rows[n]:
            E("acc = ct_fma(%s, %s, acc);" % (vname("Lf", add(n, k)), vname("xw", k)))
            E.flops += 2
        if deg(n) == 0:
            E.ind -= 1
            E("}")
        E("Lb[%d] = acc;" % idx)
        E.ind -= 1
        E("}")
    E("return %s;" % ("CT_FLOPS_" + name.upper()))
    E.ind -= 1
    E("}")
    return E


def emit_projection(E, outs, P, raw_name, dst, have_raw):
    """dst[trless_index(n)] = raw for degree <= 2, traceless projection above."""
    for n in outs:
        l = deg(n)
        idx = trless_index(n, P)
        if l <= 2:
            E("%s[%d] = %s;" % (dst, idx, raw_name(n)))
            continue
        # sum_m tf_coef(n,m) * sum_{|k|=|m|} |m|!/k! raw(n-2m+2k)
        coefs = {}
        for m in half_multi(n):
            c = tf_coef(n, m)
            j = deg(m)
            base = sub(n, scal(m, 2))
            for k in multis_of_degree(j):
                q = add(base, scal(k, 2))
                coefs[q] = coefs.get(q, Fraction(0)) + c * Fraction(fact(j), mfact(k))
        terms = [(float(v), raw_name(q)) for q, v in sorted(coefs.items(), key=lambda t: (-abs(t[1]), t[0])) if v != 0]
        E("{")
        E.ind += 1
        E("T acc;")
        term_sum(E, "acc", terms)
        E("%s[%d] = acc;" % (dst, idx))
        E.ind -= 1
        E("}")


def gen_p2m():
    E = Emitter()
    P = MORDER
    outs = trless_set(P)
    E("// Multipole (trless order 6) of a unit mass at offset X = particle - centre.")
    E("template<class T>")
    E("CT_HD inline int P2M(T* __restrict__ M, T X0, T X1, T X2) {")
    E.ind += 1
    E("const T y0 = X0 * T(%r), y1 = X1 * T(%r), y2 = X2 * T(%r);" % (-SCALE, -SCALE, -SCALE))
    E.flops += 3
    # raw monomials needed: y^q for q appearing in projection of every output
    needed = set()
    for n in outs:
        if deg(n) <= 2:
            needed.add(n)
            continue
        for m in half_multi(n):
            base = sub(n, scal(m, 2))
            for k in multis_of_degree(deg(m)):
                needed.add(add(base, scal(k, 2)))
    emit_monomials(E, "y", ["y0", "y1", "y2"], sorted(closure(needed) - {(0, 0, 0)}))

    def raw(q):
        return "T(1)" if deg(q) == 0 else vname("y", q)

    emit_projection(E, outs, P, raw, "M", None)
    E("return CT_FLOPS_P2M;")
    E.ind -= 1
    E("}")
    return E


def gen_m2m():
    E = Emitter()
    P = MORDER
    outs = trless_set(P)
    E("// Translate a trless order-6 multipole by X = child centre - parent centre.")
    E("template<class T>")
    E("CT_HD inline int M2M(T* __restrict__ Mb, const T* __restrict__ Ma, T X0, T X1, T X2) {")
    E.ind += 1
    E("const T y0 = X0 * T(%r), y1 = X1 * T(%r), y2 = X2 * T(%r);" % (-SCALE, -SCALE, -SCALE))
    E.flops += 3
    needed_raw = set()
    for n in outs:
        if deg(n) <= 2:
            needed_raw.add(n)
            continue
        for m in half_multi(n):
            base = sub(n, scal(m, 2))
            for k in multis_of_degree(deg(m)):
                needed_raw.add(add(base, scal(k, 2)))
    mono = set()
    ma_used = set()
    for n in needed_raw:
        for i in range(n[0] + 1):
            for j in range(n[1] + 1):
                for k in range(n[2] + 1):
                    kk = (i, j, k)
                    ma_used.add(kk)
                    mono.add(sub(n, kk))
    emit_monomials(E, "y", ["y0", "y1", "y2"], sorted(closure(mono) - {(0, 0, 0)}))
    for k in sorted(ma_used, key=lambda q: (deg(q), q)):
        ex = full_expr(k, P)
        if len(ex) == 1 and list(ex.values())[0] == 1:
            E("const T %s = Ma[%d];" % (vname("Ma", k), list(ex.keys())[0]))
            continue
        E("T %s;" % vname("Ma", k))
        term_sum(E, vname("Ma", k), [(v, "Ma[%d]" % kk) for kk, v in sorted(ex.items())])
    for n in sorted(needed_raw, key=lambda q: (deg(q), q)):
        terms = []
        for i in range(n[0] + 1):
            for j in range(n[1] + 1):
                for k in range(n[2] + 1):
                    kk = (i, j, k)
                    c = mfact(n) // (mfact(kk) * mfact(sub(n, kk)))
                    if kk == n:
                        continue
                    terms.append((c, kk))
        terms.sort(key=lambda t: (deg(t[1]), t[1]))
        E("T %s = %s;" % (vname("Mr", n), vname("Ma", n)))
        for c, kk in terms:
            mon = vname("y", sub(n, kk))
            if c == 1:
                E("%s = ct_fma(%s, %s, %s);" % (vname("Mr", n), mon, vname("Ma", kk), vname("Mr", n)))
                E.flops += 2
            else:
                E("%s = ct_fma(%s * %s, %s, %s);" % (vname("Mr", n), lit(c), mon, vname("Ma", kk), vname("Mr", n)))
                E.flops += 3
    emit_projection(E, outs, P, lambda q: vname("Mr", q), "Mb", None)
    E("return CT_FLOPS_M2M;")
    E.ind -= 1
    E("}")
    return E


def gen_ewald(masked=False):
    """Ewald correction D = d^n[f_periodic - f] at X, f = -1/r, f_periodic its
    zero-mean periodic sum (alpha = 2).  Inputs in box units; outputs in the
    scaled frame like greens().  masked: the variant for short-vector types
    (CPU SIMD), where the image cutoff is a lane mask: the block runs when any
    lane is inside it and the radial functions of the other lanes are zero."""
    E = Emitter()
    P = ORDER
    outs = trless_set(P)
    a = EWALD_ALPHA
    E("// Ewald correction of the Green's function (periodic sum minus nearest image),")
    E("// stored trless order 7 in the scaled frame.  X in box units.")
    E("// rimg: 3*nreal image offsets, fh: 3*nfour wave vectors (one of each +-h pair),")
    E("// fc: nfour*50 box-unit coefficients -2 c(h) (2 pi)^l h^n s_l, c(h) = exp(-pi^2 h^2/a^2)/(pi h^2),")
    E("// s_l = (+1, -1, -1, +1)[l %% 4] multiplying (cos, sin, cos, sin) of 2 pi h.X.")
    E("// TC: storage type of the coefficient table (float for FP32 kicks, double for FP64).")
    E("template<class T, class TC = float>")
    E("CT_HD inline int ewald_greens%s(T* __restrict__ D, T X0, T X1, T X2, const float* __restrict__ rimg, int nreal," % ("_masked" if masked else ""))
    E("\t\tconst float* __restrict__ fh, const TC* __restrict__ fc, int nfour) {")
    E.ind += 1
    for n in outs:
        E("T %s = T(0);" % vname("A", n))
    need = set()
    for n in outs:
        for m in half_multi(n):
            need.add(sub(n, scal(m, 2)))
    need = sorted(closure(need) - {(0, 0, 0)})

    def accumulate(prefix, gname, sign):
        for n in outs:
            l = deg(n)
            terms = []
            for m in sorted(half_multi(n), key=lambda q: -deg(q)):
                p = sub(n, scal(m, 2))
                c = float(radial_coef(n, m)) * sign
                var = "%s%d" % (gname, l - deg(m))
                if deg(p) > 0:
                    var = "%s * %s" % (vname(prefix, p), var)
                terms.append((c, var))
            for c, var in terms:
                if c == 1:
                    E("%s += %s;" % (vname("A", n), var))
                elif c == -1:
                    E("%s -= %s;" % (vname("A", n), var))
                else:
                    E("%s = ct_fma(%s, %s, %s);" % (vname("A", n), lit(c), var, vname("A", n)))

    # primary image: +erf(a r)/r, derivative functions as power series in r^2
    E("{")
    E.ind += 1
    E("const T s = ct_fma(X0, X0, ct_fma(X1, X1, X2 * X2));")
    # a_j: erf(a r)/r = sum_j a_j r^(2j)
    J = 30
    aj = [2.0 * a / math.sqrt(math.pi) * (-1) ** j * a ** (2 * j) / (math.factorial(j) * (2 * j + 1)) for j in range(J + 1)]
    for k in range(P):
        # h_k(r) = sum_{j>=k} a_j 2^k j!/(j-k)! s^(j-k)
        cs = [aj[j] * 2 ** k * math.factorial(j) / math.factorial(j - k) for j in range(k, J + 1)]
        # drop tiny tail (s <= 0.75, a^2 s <= 3)
        while len(cs) > 1 and abs(cs[-1]) * 0.75 ** (len(cs) - 1) < 1e-18 * abs(cs[0]):
            cs.pop()
        E("T h%d = %s;" % (k, lit(cs[-1])))
        for c in reversed(cs[:-1]):
            E("h%d = ct_fma(h%d, s, %s);" % (k, k, lit(c)))
    emit_monomials(E, "p", ["X0", "X1", "X2"], need)
    accumulate("p", "h", 1.0)
    E.ind -= 1
    E("}")
    # real-space images: -erfc(a r)/r
    E("for (int i = 0; i < nreal; i++) {")
    E.ind += 1
    E("const T d0 = X0 - T(rimg[3 * i + 0]);")
    E("const T d1 = X1 - T(rimg[3 * i + 1]);")
    E("const T d2 = X2 - T(rimg[3 * i + 2]);")
    E("const T r2 = ct_fma(d0, d0, ct_fma(d1, d1, d2 * d2));")
    if masked:
        E("const auto inside = r2 < T(%r);" % (2.6 * 2.6))
        E("if (ct_any(inside)) {")
        E.ind += 1
        E("const T rr2 = ct_select(inside, r2, T(1));")
        E("const T r = ct_sqrt(rr2);")
        E("const T rinv2 = T(1) / rr2;")
        E("const T ex = ct_select(inside, ct_exp(T(%r) * rr2), T(0));" % (-a * a))
        E("const T b0 = ct_select(inside, ct_erfc(T(%r) * r) / r, T(0));" % a)
    else:
        E("if (r2 < T(%r)) {" % (2.6 * 2.6))
        E.ind += 1
        E("const T r = ct_sqrt(r2);")
        E("const T rinv2 = T(1) / r2;")
        E("const T ex = ct_exp(T(%r) * r2);" % (-a * a))
        E("const T b0 = ct_erfc(T(%r) * r) / r;" % a)
    # B_k = ((2k-1) B_(k-1) + (2a^2)^k/(a sqrt(pi)) e) / r^2 ; g_k = (-1)^k B_k
    E("T e = T(%r) * ex;" % (1.0 / (a * math.sqrt(math.pi))))
    E("T b = b0;")
    E("const T g0 = b0;")
    for k in range(1, P):
        E("e *= T(%r);" % (2 * a * a))
        E("b = ct_fma(T(%d), b, e) * rinv2;" % (2 * k - 1))
        E("const T g%d = %sb;" % (k, "-" if k % 2 else ""))
    emit_monomials(E, "q", ["d0", "d1", "d2"], need)
    accumulate("q", "g", -1.0)
    E.ind -= 1
    E("}")
    E.ind -= 1
    E("}")
    # Fourier
    E("for (int i = 0; i < nfour; i++) {")
    E.ind += 1
    E("const T hx = ct_fma(T(fh[3 * i + 0]), X0, ct_fma(T(fh[3 * i + 1]), X1, T(fh[3 * i + 2]) * X2));")
    E("T sn, cn;")
    E("ct_sincos(T(%r) * hx, &sn, &cn);" % (2 * math.pi))
    E("const TC* __restrict__ c = fc + %d * i;" % (P * P + 1))
    for n in outs:
        l = deg(n)
        E("%s = ct_fma(T(c[%d]), %s, %s);" % (vname("A", n), trless_index(n, P), "cn" if l % 2 == 0 else "sn", vname("A", n)))
    E.ind -= 1
    E("}")
    E("%s += T(%r);" % (vname("A", (0, 0, 0)), math.pi / (a * a)))
    for n in outs:
        l = deg(n)
        E("D[%d] = %s * T(%r);" % (trless_index(n, P), vname("A", n), SCALE ** (-(l + 1))))
    E("return 0;")
    E.ind -= 1
    E("}")
    return E


def ewald_flops_estimate():
    P = ORDER
    outs = trless_set(P)
    terms = sum(len(half_multi(n)) for n in outs)
    return terms


HEADER = """// Generated by cosmictiger_amd/codegen/gen_expansions.py -- do not edit.
// Order-7 Cartesian FMM kernels in the traceless storage of the reference's
// tensor_trless_sym (cosmictiger/tensor.hpp:63-136, trless_index :317-319),
// lengths scaled by SCALE_FACTOR (cosmictiger/fmm_kernels.hpp:33-49).
#pragma once
#include "ct_math.hpp"

#define CT_EXPANSION_SIZE 50
#define CT_MULTIPOLE_SIZE 37
#define CT_SCALE_FACTOR 1000.0f
"""


def main():
    out = sys.argv[1] if len(sys.argv) > 1 else os.path.join(os.path.dirname(__file__), "..", "csrc", "ct_expansions.hpp")
    parts = []
    gens = [("GREENS", gen_greens()), ("M2L", gen_m2l(ORDER, "M2L")), ("M2L2", gen_m2l(2, "M2L2")),
            ("L2L", gen_l2l(False)), ("L2P", gen_l2l(True)), ("P2M", gen_p2m()), ("M2M", gen_m2m()),
            ("EWALD", gen_ewald()), ("EWALD_MASKED", gen_ewald(masked=True))]
    text = HEADER
    for name, E in gens:
        text += "#define CT_FLOPS_%s %d\n" % (name, E.flops)
    text += "#define CT_EWALD_TERMS_PER_IMAGE %d\n" % ewald_flops_estimate()
    text += "\nnamespace ctfmm {\n\n"
    for name, E in gens:
        text += E.text() + "\n"
    text += "}  // namespace ctfmm\n"
    with open(out, "w") as f:
        f.write(text)
    print("wrote %s (%d lines)" % (out, text.count("\n")))
    for name, E in gens:
        print("  %-6s flops %d" % (name, E.flops))


if __name__ == "__main__":
    main()
