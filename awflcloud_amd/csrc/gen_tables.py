#!/usr/bin/env python3
"""Derive the polynomial / GLL / stencil transform tables of the ADER-WENO step
and emit them as a C/C++ header of hex-float literals (``awfl_tables.h``).

Every table is re-derived here from its mathematical definition, in exact
rational arithmetic where the definition is rational and in 200-bit mpmath
where it involves Gauss-Lobatto-Legendre (GLL) nodes.  The definitions follow
the reference's SageMath generator:

* GLL nodes/weights on [-1/2, 1/2] ........ sage/quadrature_utils.sage:15-50
* gll_to_coefs / coefs_to_gll (N x N) ...... sage/transformation_matrices.sage:8-16
* sten_to_coefs (cell averages -> coefs) ... transformation_matrices.sage:43-51
* coefs_to_deriv ............................ transformation_matrices.sage:181-188
* coefs_to_gll_lower (N coefs -> M pts) ..... transformation_matrices.sage:217-235
* weno_sten_to_coefs (hs+2 polynomials) ..... transformation_matrices.sage:238-260
* coefs_to_TV (truncated integrand) ......... transformation_matrices.sage:263-269,
                                               poly_utils.sage:16-26

Storage convention (same as the reference's SArray(in, out) "column-row"
matrices, sage/c_utils.sage:114-126): ``M[in][out]``.  The composite operators
built in ``Spatial::init`` (Spatial_euler3d_cons_expl_cart_fv_Agrid.h:274-309)
are formed here in IEEE double, summing the inner index in ascending order from
0.0, exactly as a plain triple loop does.

Run:  python gen_tables.py [output_header]
"""
from fractions import Fraction as Fr
import sys
import os

import mpmath as mp

mp.mp.prec = 200

ORDS = (3, 5, 7, 9)          # WENO orders (odd)
NMAX = 9                     # largest polynomial size / GLL count


# ----------------------------------------------------------------------------
# small exact / high-precision linear algebra helpers
# ----------------------------------------------------------------------------
def inv_exact(a):
    """Gauss-Jordan inverse of a square matrix of Fractions."""
    n = len(a)
    m = [list(row) + [Fr(int(i == j)) for j in range(n)] for i, row in enumerate(a)]
    for c in range(n):
        p = next(r for r in range(c, n) if m[r][c] != 0)
        m[c], m[p] = m[p], m[c]
        piv = m[c][c]
        m[c] = [v / piv for v in m[c]]
        for r in range(n):
            if r != c and m[r][c] != 0:
                f = m[r][c]
                m[r] = [vr - f * vc for vr, vc in zip(m[r], m[c])]
    return [row[n:] for row in m]


def inv_mp(a):
    return mp.inverse(mp.matrix(a))


# ----------------------------------------------------------------------------
# Gauss-Lobatto-Legendre nodes and weights on [-1/2, 1/2]
# ----------------------------------------------------------------------------
def legendre_and_deriv(n, x):
    """P_n(x) and P_n'(x) by the three-term recurrence."""
    p0, p1 = mp.mpf(1), x
    if n == 0:
        return p0, mp.mpf(0)
    for k in range(2, n + 1):
        p0, p1 = p1, ((2 * k - 1) * x * p1 - (k - 1) * p0) / k
    dp = n * (x * p1 - p0) / (x * x - 1) if abs(x) != 1 else mp.mpf(n * (n + 1)) / 2 * (x ** (n + 1))
    return p1, dp


def gll(npts):
    """GLL nodes (ascending) and weights on [-1/2, 1/2] (mpf)."""
    if npts == 1:
        return [mp.mpf(0)], [mp.mpf(1)]
    n = npts - 1
    xs = []
    for i in range(npts):
        x = -mp.cos(mp.pi * i / n)           # Chebyshev-Gauss-Lobatto start
        if 0 < i < n:
            for _ in range(200):             # Newton on P_n'(x) = 0
                p, dp = legendre_and_deriv(n, x)
                # second derivative from Legendre ODE: (1-x^2)P'' = 2xP' - n(n+1)P
                d2p = (2 * x * dp - n * (n + 1) * p) / (1 - x * x)
                dx = dp / d2p
                x -= dx
                if abs(dx) < mp.mpf(10) ** -70:
                    break
        xs.append(x)
    ws = []
    for x in xs:
        p, _ = legendre_and_deriv(n, x)
        ws.append(mp.mpf(2) / (n * npts * p * p))
    return [x / 2 for x in xs], [w / 2 for w in ws]


def exact_gll(npts):
    """Nodes as Fractions where they are rational (npts <= 3), else None."""
    if npts == 1:
        return [Fr(0)]
    if npts == 2:
        return [Fr(-1, 2), Fr(1, 2)]
    if npts == 3:
        return [Fr(-1, 2), Fr(0), Fr(1, 2)]
    return None


# ----------------------------------------------------------------------------
# table definitions (all returned as [in][out] of Fraction or mpf)
# ----------------------------------------------------------------------------
def coefs_to_gll_lower(n, m):
    """n monomial coefficients -> m GLL point values; m == 1 uses x = 0."""
    ex = exact_gll(m)
    pts = ex if ex is not None else gll(m)[0]
    return [[pts[i] ** j for i in range(m)] for j in range(n)]


def gll_to_coefs(n):
    c2g = coefs_to_gll_lower(n, n)                      # [coef][pt]
    vand = [[c2g[j][i] for j in range(n)] for i in range(n)]  # V[pt][coef]
    if exact_gll(n) is not None:
        inv = inv_exact(vand)
        return [[inv[j][i] for j in range(n)] for i in range(n)]  # [pt][coef]
    inv = inv_mp(vand)
    return [[inv[j, i] for j in range(n)] for i in range(n)]


def coefs_to_deriv(n):
    out = [[Fr(0)] * n for _ in range(n)]
    for j in range(1, n):
        out[j][j - 1] = Fr(j)
    return out


def cell_avg_row(lo, hi, ncoef):
    """Average of x^j over [lo, hi], j < ncoef (exact)."""
    return [(hi ** (j + 1) - lo ** (j + 1)) / (j + 1) / (hi - lo) for j in range(ncoef)]


def sten_to_coefs(n):
    hs = (n - 1) // 2
    a = [cell_avg_row(Fr(2 * i - 1, 2), Fr(2 * i + 1, 2), n) for i in range(-hs, hs + 1)]
    inv = inv_exact(a)                                   # inv[coef][cell]
    return [[inv[c][s] for c in range(n)] for s in range(n)]   # [cell][coef]


def weno_sten_to_coefs(n):
    """(hs+2) polynomials: hs+1 degree-hs candidates on sub-stencils, then the
    full order-n polynomial.  Result[poly][cell][coef], zero-padded to n x n."""
    hs = (n - 1) // 2
    locs = [Fr(2 * i - 1, 2) for i in range(-hs, hs + 2)]
    polys = []
    for j in range(hs + 1):
        a = [cell_avg_row(locs[i], locs[i + 1], hs + 1) for i in range(j, j + hs + 1)]
        inv = inv_exact(a)
        t = [[Fr(0)] * n for _ in range(n)]
        for s in range(hs + 1):
            for c in range(hs + 1):
                t[s][c] = inv[c][s]
        polys.append(t)
    a = [cell_avg_row(locs[i], locs[i + 1], n) for i in range(n)]
    inv = inv_exact(a)
    polys.append([[inv[c][s] for c in range(n)] for s in range(n)])
    return polys


def tv_terms(n):
    """Quadratic form of the truncated total-variation indicator.

    TV(a) = sum_{d=1}^{n-1} int_{-1/2}^{1/2} T_n[(p^(d))^2] dx with T_n keeping
    only monomials of degree < n (poly_utils.sage integrate_poly).  Returned as
    (j, k, coef) with j <= k, ordered by (k, j) as the reference's generated
    expression orders them; zero coefficients are dropped."""
    q = {}
    for d in range(1, n):
        # p^(d) = sum_j a_j * j!/(j-d)! x^(j-d)
        terms = [(j, Fr(mp_fact(j), mp_fact(j - d)), j - d) for j in range(d, n)]
        for (j1, c1, e1) in terms:
            for (j2, c2, e2) in terms:
                e = e1 + e2
                if e >= n:
                    continue
                integ = (Fr(1, 2) ** (e + 1) - Fr(-1, 2) ** (e + 1)) / (e + 1)
                key = (min(j1, j2), max(j1, j2))
                q[key] = q.get(key, Fr(0)) + c1 * c2 * integ
    out = [(j, k, c) for (j, k), c in q.items() if c != 0]
    out.sort(key=lambda t: (t[1], t[0]))
    return out


def mp_fact(k):
    f = 1
    for i in range(2, k + 1):
        f *= i
    return f


# WENO ideal weights and sigma: the tuned decimal constants of
# WenoLimiter.h:28-90, then convexified (divided by their sum + 1e-20) in double.
IDEAL_RAW = {
    3: (0.0343557947899881, [1.0, 1.0, 1224.61619926508]),
    5: (0.73564225445964, [1.0, 73.564225445964, 1.0, 1584.89319246111]),
    7: (0.125594321575479, [1.0, 7.35642254459641, 7.35642254459641, 1.0, 794.328234724281]),
    9: (0.0288539981181442, [1.0, 2.15766927997459, 2.40224886796286, 2.15766927997459, 1.0,
                             1136.12697719888]),
}


def ideal_weights(n):
    sigma, raw = IDEAL_RAW[n]
    s = 0.0
    for v in raw:
        s += v
    return sigma, [v / (s + 1.0e-20) for v in raw]


# ----------------------------------------------------------------------------
# double rounding and emission
# ----------------------------------------------------------------------------
def d(v):
    if isinstance(v, Fr):
        return float(v)
    return float(mp.mpf(v))


def matmul_cr(left, right, n_in, n_mid, n_out):
    """ret[i][j] = sum_k left[k][j] * right[i][k] in double, k ascending."""
    ret = [[0.0] * n_out for _ in range(n_in)]
    for i in range(n_in):
        for j in range(n_out):
            t = 0.0
            for k in range(n_mid):
                t += left[k][j] * right[i][k]
            ret[i][j] = t
    return ret


def fmt(x):
    if x == 0.0:
        return "0.0"
    return x.hex()


def emit_array(name, arr, dims):
    lines = []
    def rec(a, depth):
        if depth == len(dims) - 1:
            return "{" + ", ".join(fmt(v) for v in a) + "}"
        return "{" + ",\n    ".join(rec(x, depth + 1) for x in a) + "}"
    dstr = "".join("[%d]" % n for n in dims)
    lines.append("AWFL_TBL double %s%s = %s;" % (name, dstr, rec(arr, 0)))
    return "\n".join(lines)


def build():
    """Return a dict of every table in double precision (used by tests too)."""
    T = {}
    for n in range(1, NMAX + 1):
        pts, wts = gll(n)
        ex = exact_gll(n)
        T["gll_pts_%d" % n] = [float(v) for v in ex] if ex else [d(v) for v in pts]
        if n == 3:
            T["gll_wts_%d" % n] = [float(Fr(1, 6)), float(Fr(2, 3)), float(Fr(1, 6))]
        elif n == 2:
            T["gll_wts_%d" % n] = [0.5, 0.5]
        elif n == 1:
            T["gll_wts_%d" % n] = [1.0]
        else:
            T["gll_wts_%d" % n] = [d(v) for v in wts]
    for n in range(2, NMAX + 1):
        T["gll_to_coefs_%d" % n] = [[d(v) for v in row] for row in gll_to_coefs(n)]
        T["coefs_to_gll_%d" % n] = [[d(v) for v in row] for row in coefs_to_gll_lower(n, n)]
        T["coefs_to_deriv_%d" % n] = [[d(v) for v in row] for row in coefs_to_deriv(n)]
        T["tv_%d" % n] = [(j, k, float(c)) for (j, k, c) in tv_terms(n)]
    for o in ORDS:
        hs = (o - 1) // 2
        T["sten_to_coefs_%d" % o] = [[d(v) for v in row] for row in sten_to_coefs(o)]
        T["weno_recon_%d" % o] = [[[d(v) for v in row] for row in p] for p in weno_sten_to_coefs(o)]
        sig, idl = ideal_weights(o)
        T["sigma_%d" % o] = sig
        T["idl_%d" % o] = idl
        for m in range(1, o + 1):
            T["coefs_to_gll_lower_%d_%d" % (o, m)] = [[d(v) for v in row]
                                                   for row in coefs_to_gll_lower(o, m)]
        # composites of Spatial::init (S:291-309)
        for m in range(2, o + 1):
            c2gl = T["coefs_to_gll_lower_%d_%d" % (o, m)]
            s2c = T["sten_to_coefs_%d" % o]
            T["sten_to_gll_%d_%d" % (o, m)] = matmul_cr(c2gl, s2c, o, o, m)
    for m in range(2, NMAX + 1):
        g2c = T["gll_to_coefs_%d" % m]
        c2d = T["coefs_to_deriv_%d" % m]
        c2g = T["coefs_to_gll_%d" % m]
        inner = matmul_cr(c2d, g2c, m, m, m)
        T["deriv_%d" % m] = matmul_cr(c2g, inner, m, m, m)
    return T


def render(T):
    out = []
    out.append("/* awfl_tables.h -- GENERATED by gen_tables.py; do not edit.")
    out.append(" * Transform tables of the ADER-WENO step in IEEE double (hex literals).")
    out.append(" * Matrices are stored [in][out]; weno_recon is [poly][cell][coef]. */")
    out.append("#ifndef AWFL_TABLES_H")
    out.append("#define AWFL_TABLES_H")
    out.append("#ifdef __cplusplus")
    out.append("#define AWFL_TBL static constexpr")
    out.append("#else")
    out.append("#define AWFL_TBL static const")
    out.append("#endif")
    out.append("#define AWFL_TABLES_NMAX %d" % NMAX)
    for n in range(1, NMAX + 1):
        out.append(emit_array("awfl_gll_pts_%d" % n, T["gll_pts_%d" % n], [n]))
        out.append(emit_array("awfl_gll_wts_%d" % n, T["gll_wts_%d" % n], [n]))
    for n in range(2, NMAX + 1):
        for key in ("gll_to_coefs", "coefs_to_gll", "coefs_to_deriv", "deriv"):
            out.append(emit_array("awfl_%s_%d" % (key, n), T["%s_%d" % (key, n)], [n, n]))
        terms = T["tv_%d" % n]
        out.append("#define AWFL_TV_%d_NTERMS %d" % (n, len(terms)))
        out.append("AWFL_TBL int awfl_tv_%d_j[%d] = {%s};" % (n, len(terms), ", ".join(str(t[0]) for t in terms)))
        out.append("AWFL_TBL int awfl_tv_%d_k[%d] = {%s};" % (n, len(terms), ", ".join(str(t[1]) for t in terms)))
        out.append(emit_array("awfl_tv_%d_c" % n, [t[2] for t in terms], [len(terms)]))
    for o in ORDS:
        hs = (o - 1) // 2
        out.append(emit_array("awfl_sten_to_coefs_%d" % o, T["sten_to_coefs_%d" % o], [o, o]))
        out.append(emit_array("awfl_weno_recon_%d" % o, T["weno_recon_%d" % o], [hs + 2, o, o]))
        out.append(emit_array("awfl_idl_%d" % o, T["idl_%d" % o], [hs + 2]))
        out.append("AWFL_TBL double awfl_sigma_%d = %s;" % (o, fmt(T["sigma_%d" % o])))
        for m in range(1, o + 1):
            out.append(emit_array("awfl_coefs_to_gll_lower_%d_%d" % (o, m),
                                  T["coefs_to_gll_lower_%d_%d" % (o, m)], [o, m]))
        for m in range(2, o + 1):
            out.append(emit_array("awfl_sten_to_gll_%d_%d" % (o, m), T["sten_to_gll_%d_%d" % (o, m)], [o, m]))
    out.append("#endif /* AWFL_TABLES_H */")
    return "\n".join(out) + "\n"


if __name__ == "__main__":
    dst = sys.argv[1] if len(sys.argv) > 1 else os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                              "awfl_tables.h")
    text = render(build())
    old = open(dst).read() if os.path.exists(dst) else None
    if old != text:
        with open(dst, "w") as f:
            f.write(text)
