"""Generate hq_local.h: the far-field local expansion of the mixed-precision
bunched walk (cell -> bunch M2L, and L2P), as straight-line FP32 code.

    python hierarquicos_amd/csrc/gen_local.py        # rewrites hq_local.h

Kernel: g(R) = (|R|^2 + eps^2)^(-1/2).  With rho = |R|^2 + eps^2, the
Cartesian derivatives are D_k = sum_l c(k,l) R^(k-2l) F_(|k|-|l|)(rho),
F_j = 2^j d^j g / d rho^j = (-1)^j (2j-1)!! rho^(-j-1/2), obtained by
differentiating term by term with  d_i [R^a F_j] = a_i R^(a-e_i) F_j +
R^(a+e_i) F_(j+1).  In units of 1/sqrt(rho) (R^ = R / sqrt(rho)) the F_j are
the constants (-1)^j (2j-1)!!, so D^_k is a polynomial in R^ with integer
coefficients and D_k = rho^(-(|k|+1)/2) D^_k.

A source cell with mass M and raw moments M_m about its centre of mass
(quadrupole Q, |m| = 2; octupole O, |m| = 3; the dipole vanishes) at
R = (bunch centre) - (centre of mass) gives the potential's local
coefficients  L_n = -sum_m (-1)^|m| / m! M_m D_(n+m)(R),  so that near the
bunch centre  phi(c + d) = sum_n L_n d^n / n!  and the acceleration is
a_i = -sum_n L_(n+e_i) d^n / n!.  Two sets of orders are generated: p5 takes
the monopole term to |n| <= 5, the quadrupole to |n| <= 3 and the octupole to
|n| <= 2 (derivative order 5 in all three; 55 coefficients); p6 takes them to
6, 4 and 3 (derivative order 6; 83 coefficients).  The coefficients are
stored scaled by h^|n| (h the bunch radius), so every stored value is bounded
and the evaluation uses d / h.
"""
import os
import re
from fractions import Fraction

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "hq_local.h")

ORDERS = {5: (5, 3, 2), 6: (6, 4, 3)}  # name -> (monopole, quadrupole, octupole) local orders


def midx(n):
    return [(i, j, n - i - j) for i in range(n, -1, -1) for j in range(n - i, -1, -1)]


def add(a, b):
    return (a[0] + b[0], a[1] + b[1], a[2] + b[2])


def fact(k):
    f = 1
    for v in k:
        for i in range(2, v + 1):
            f *= i
    return f


def dfact(j):  # (2j-1)!!
    r = 1
    for i in range(1, 2 * j, 2):
        r *= i
    return r


def derivative_poly(k):
    """D^_k as {monomial exponent: integer coefficient} (F_j substituted)."""
    terms = {((0, 0, 0), 0): 1}
    path = [0] * k[0] + [1] * k[1] + [2] * k[2]
    for i in path:
        new = {}
        for (a, j), c in terms.items():
            if a[i] > 0:
                aa = list(a)
                aa[i] -= 1
                key = (tuple(aa), j)
                new[key] = new.get(key, 0) + c * a[i]
            aa = list(a)
            aa[i] += 1
            key = (tuple(aa), j + 1)
            new[key] = new.get(key, 0) + c
        terms = new
    poly = {}
    for (a, j), c in terms.items():
        val = c * (-1) ** j * dfact(j)
        if val:
            poly[a] = poly.get(a, 0) + val
    return {a: c for a, c in poly.items() if c}


# source moments: multi-index -> (record slot, factor) with the record's
# scaling (r[1..6] = -3 Q, r[11..20] = 7.5 O in xxx, xyy, xzz, xxy, yyy, yzz,
# xxz, yyz, zzz, xyz order) undone; P_m = -(-1)^|m|/m! M_m ri^|m|
QUAD = {(2, 0, 0): 1, (0, 2, 0): 2, (0, 0, 2): 3, (1, 1, 0): 4, (1, 0, 1): 5, (0, 1, 1): 6}
OCT = {(3, 0, 0): 11, (1, 2, 0): 12, (1, 0, 2): 13, (2, 1, 0): 14, (0, 3, 0): 15,
       (0, 1, 2): 16, (2, 0, 1): 17, (0, 2, 1): 18, (0, 0, 3): 19, (1, 1, 1): 20}


def moment_factor(m):
    """P_m = factor * r[slot] * ri^|m|."""
    s = sum(m)
    base = Fraction(-1 if s == 0 else (-1) ** (s + 1), fact(m))  # -(-1)^|m| / m!
    if s == 0:
        return base
    if s == 2:
        return base * Fraction(-1, 3)
    return base * Fraction(2, 15)


def flt(x):
    """A float literal, rendered per language by render()."""
    x = float(x)
    r = repr(x)
    if "e" not in r and "." not in r:
        r += ".0"
    return "@" + r + "@"


def render(lines, lang):
    """C++ (FP32 literals, MP ternaries) or Python (for the CPU tests)."""
    out = []
    for ln in lines:
        if lang == "c":
            ln = re.sub(r"@([^@]+)@", r"\1f", ln)
            ln = re.sub(r"SEL(\d)\((.*)\);$",
                        lambda m: "MP %s ? %s : 0.0f;" % (">= 4" if m.group(1) == "4" else "== 8",
                                                          m.group(2)), ln)
            ln = ln.replace("RSQRT(", "__frsqrt_rn(")
        out.append(ln)
    return out


def orders(p):
    """(PMONO, PQUAD, POCT, KMAX, LOC, LIDX) of the order set p."""
    pmono, pquad, poct = ORDERS[p]
    kmax = max(pmono, pquad + 2, poct + 3)
    loc = [n for o in range(1, kmax + 1) for n in midx(o)]  # stored coefficients
    return pmono, pquad, poct, kmax, loc, {n: i for i, n in enumerate(loc)}


LOC = orders(5)[4]


def gen_m2l(p):
    PMONO, PQUAD, POCT, KMAX, LOC, LIDX = orders(p)
    out = []
    w = out.append
    w("// Local coefficients of one source cell (this lane's) about the bunch centre.")
    w("// R = bunch centre - cell centre of mass, h = bunch radius (> 0), r = the")
    w("// cell's FP32 record; adds h^|n| L_n to L[i] for the multi-index n = LOC_n[i]")
    w("// (the addition folded into each coefficient's last multiply-add).")
    w("template <int MP>")
    w("__device__ __forceinline__ void local_m2l_p%d(float Rx, float Ry, float Rz, float eps2, float h," % p)
    w("                                             const float* __restrict__ r, float* __restrict__ L) {")
    w("#pragma clang fp contract(off)")
    w("    const float rho = fmaf(Rx, Rx, fmaf(Ry, Ry, fmaf(Rz, Rz, eps2)));")
    w("    const float ri = RSQRT(rho);")
    w("    const float X = Rx * ri;")
    w("    const float Y = Ry * ri;")
    w("    const float Z = Rz * ri;")
    w("    const float sg = h * ri;")
    # monomials of X, Y, Z up to degree KMAX
    need = set()
    polys = {}
    for o in range(1, KMAX + 1):
        for k in midx(o):
            polys[k] = derivative_poly(k)
            need.update(polys[k].keys())
    names = {(1, 0, 0): "X", (0, 1, 0): "Y", (0, 0, 1): "Z"}
    for o in range(2, KMAX + 1):
        for a in midx(o):
            if a not in need:
                continue
            # a = b + e_i with b already defined (need closure)
            for i, v in ((0, "X"), (1, "Y"), (2, "Z")):
                if a[i] > 0:
                    b = list(a)
                    b[i] -= 1
                    b = tuple(b)
                    if b in names or sum(b) <= 1:
                        break
            nm = "m%d%d%d" % a
            bn = names[b] if b in names else "m%d%d%d" % b
            w("    const float %s = %s * %s;" % (nm, bn, v))
            names[a] = nm
            need.add(b)
    for k, p in polys.items():
        terms = sorted(p.items(), key=lambda t: -sum(t[0]))
        const = sum(c for a, c in terms if sum(a) == 0)
        terms = [(a, c) for a, c in terms if sum(a) > 0]
        expr = None
        for a, c in terms:
            m = names[a]
            if expr is None:
                if const:
                    expr = "fmaf(%s, %s, %s)" % (flt(c), m, flt(const))
                else:
                    expr = m if c == 1 else ("-" + m if c == -1 else "%s * %s" % (flt(c), m))
            else:
                expr = "fmaf(%s, %s, %s)" % (flt(c), m, expr)
        w("    const float D%d%d%d = %s;" % (k + (expr,)))
    # moments
    w("    const float ri2 = ri * ri;")
    w("    const float ri3 = ri2 * ri;")
    w("    const float P000 = -r[0];")
    cq = {}
    for m, slot in QUAD.items():
        cq.setdefault(moment_factor(m), []).append((m, slot))
    co = {}
    for m, slot in OCT.items():
        co.setdefault(moment_factor(m), []).append((m, slot))
    for i, f in enumerate(cq):
        w("    const float cq%d = SEL4(ri2 * %s);" % (i, flt(f)))
    for i, f in enumerate(co):
        w("    const float co%d = SEL8(ri3 * %s);" % (i, flt(f)))
    for i, (f, lst) in enumerate(cq.items()):
        for m, slot in lst:
            w("    const float P%d%d%d = SEL4(r[%d] * cq%d);" % (m + (slot, i)))
    for i, (f, lst) in enumerate(co.items()):
        for m, slot in lst:
            w("    const float P%d%d%d = SEL8(r[%d] * co%d);" % (m + (slot, i)))
    # scale factors s_j = ri * sg^j
    w("    const float s1 = ri * sg;")
    for j in range(2, KMAX + 1):
        w("    const float s%d = s%d * sg;" % (j, j - 1))
    w("    const float q0 = P000 * ri;")
    w("    const float t1 = q0 * sg;")
    for j in range(2, KMAX + 1):
        w("    const float t%d = t%d * sg;" % (j, j - 1))
    for idx, n in enumerate(LOC):
        o = sum(n)
        # multipole part (quadrupole |n| <= PQUAD, octupole |n| <= POCT)
        parts = []
        if o <= PQUAD:
            for m in QUAD:
                parts.append(("P%d%d%d" % m, "D%d%d%d" % add(n, m), "q"))
        if o <= POCT:
            for m in OCT:
                parts.append(("P%d%d%d" % m, "D%d%d%d" % add(n, m), "o"))
        mono_term = "fmaf(t%d, D%d%d%d, L[%d])" % ((o,) + n + (idx,))
        if not parts:
            w("    L[%d] = %s;" % (idx, mono_term))
            continue
        expr = None
        for p, d, kind in parts:
            expr = "%s * %s" % (p, d) if expr is None else "fmaf(%s, %s, %s)" % (p, d, expr)
        w("    L[%d] = fmaf(s%d, %s, %s);" % (idx, o, expr, mono_term))
    w("}")
    return out


def gen_l2p(p):
    PMONO, PQUAD, POCT, KMAX, LOC, LIDX = orders(p)
    out = []
    w = out.append
    w("")
    w("// Acceleration sum_n L_(n+e_i) u^n / n! at u = (target - bunch centre) / h from the")
    w("// scaled coefficients L of local_m2l_p%d; the caller multiplies by -1/h." % p)
    w("__device__ __forceinline__ void local_l2p_p%d(const float* __restrict__ L, float ux, float uy," % p)
    w("                                             float uz, float acc[3]) {")
    w("#pragma clang fp contract(off)")
    names = {(0, 0, 0): None, (1, 0, 0): "ux", (0, 1, 0): "uy", (0, 0, 1): "uz"}
    # u^n / n! = u^(n-e_i)/(n-e_i)! * u_i / n_i
    for o in range(2, KMAX):
        for a in midx(o):
            for i, v in ((0, "ux"), (1, "uy"), (2, "uz")):
                if a[i] > 0:
                    b = list(a)
                    b[i] -= 1
                    b = tuple(b)
                    break
            nm = "u%d%d%d" % a
            c = Fraction(1, a[i])
            if c == 1:
                w("    const float %s = %s * %s;" % (nm, names[b], v))
            else:
                w("    const float %s = (%s * %s) * %s;" % (nm, names[b], v, flt(c)))
            names[a] = nm
    for i in range(3):
        ei = tuple(1 if j == i else 0 for j in range(3))
        terms = []
        for o in range(0, KMAX):
            for n in midx(o):
                terms.append((LIDX[add(n, ei)], names[n]))
        # highest orders first, the constant term last
        terms.sort(key=lambda t: -(0 if t[1] is None else int(t[1][1]) + int(t[1][2]) + int(t[1][3])
                                   if t[1] and t[1].startswith("u") and len(t[1]) == 4 else 0))
        expr = None
        const = None
        for li, nm in terms:
            if nm is None:
                const = "L[%d]" % li
                continue
            expr = "L[%d] * %s" % (li, nm) if expr is None else "fmaf(L[%d], %s, %s)" % (li, nm, expr)
        w("    acc[%d] = %s + %s;" % (i, expr, const))
    w("}")
    return out


def python_source(p=5):
    """The two functions of order set p as Python (for CPU tests): m2l(MP, Rx, Ry, Rz,
    eps2, h, r, L, fmaf, RSQRT) adds into L; l2p(L, ux, uy, uz, acc, fmaf) fills acc."""
    def body(lines):
        out = []
        for ln in lines:
            t = ln.strip()
            if not t or t.startswith("//") or t.startswith("#") or t.startswith("template") \
                    or t.startswith("__device__") or t.startswith("const float* ") \
                    or t.startswith("float ") or t == "}":
                continue
            t = t.replace("const float ", "").rstrip(";")
            t = re.sub(r"@([^@]+)@", r"\1", t)
            t = re.sub(r"SEL(\d)\((.*)\)$",
                       lambda m: "(%s) if MP %s else 0.0" % (m.group(2), ">= 4" if m.group(1) == "4"
                                                             else "== 8"), t)
            out.append("    " + t)
        return out
    src = ["def m2l(MP, Rx, Ry, Rz, eps2, h, r, L, fmaf, RSQRT):"] + body(gen_m2l(p))
    src += ["", "def l2p(L, ux, uy, uz, acc, fmaf):"] + body(gen_l2p(p))
    return "\n".join(src) + "\n"


def main():
    lines = ["// Generated by gen_local.py from the formulas in its docstring; do not edit.",
             "#pragma once",
             "",
             "namespace hq {",
             ""]
    for p in sorted(ORDERS):
        pmono, pquad, poct, kmax, loc, lidx = orders(p)
        lines += ["constexpr int NLOC_P%d = %d;  // local coefficients, 1 <= |n| <= %d"
                  % (p, len(loc), kmax),
                  "// order set p%d: monopole %d, quadrupole %d, octupole %d; coefficient i is"
                  % (p, pmono, pquad, poct),
                  "// the multi-index (nx, ny, nz): " + " ".join("%d%d%d" % n for n in loc),
                  ""]
        lines += render(gen_m2l(p), "c")
        lines += render(gen_l2p(p), "c")
        lines += [""]
    lines += ["}  // namespace hq", ""]
    with open(OUT, "w") as f:
        f.write("\n".join(lines))
    print(OUT, len(lines), "lines")


if __name__ == "__main__":
    main()
