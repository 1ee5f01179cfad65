"""Pin the transform tables.

1. Every table the step uses, re-derived from its mathematical definition by
   awflcloud_amd/csrc/gen_tables.py, equals the reference's generated literals
   (tests/golden/tm_literals.json, numbers extracted from TransformMatrices.h).
2. The composite operators formed by the oracle's restatement of
   Spatial::init (S:274-315) equal the composites in the generated header.
3. Known answer of sage/derivatives.ipynb cell 1, and quadrature/derivative
   exactness that a wrong node, weight or transposition would break.
"""
import json
import os
import sys

import numpy as np
import pytest
from mpmath import mp, mpf

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "awflcloud_amd", "csrc"))
import gen_tables  # noqa: E402

FIX = json.load(open(os.path.join(ROOT, "tests", "golden", "tm_literals.json")))["tables"]
T = gen_tables.build()


def _lookup(tab, idx):
    v = tab
    for i in idx:
        v = v[i]
    return v


def _cmp(ours, key, abs_tol=0.0):
    worst = 0.0
    for e in FIX[key]["entries"]:
        idx, s = e[:-1], e[-1]
        ref = float(mpf(s))
        got = _lookup(ours, idx)
        d = abs(got - ref)
        if d > abs_tol:
            worst = max(worst, d / max(abs(ref), 1e-300))
    return worst


@pytest.mark.parametrize("n", range(2, 10))
def test_gll_and_polynomial_tables_match_reference(n):
    assert _cmp(T["gll_pts_%d" % n], "get_gll_points_%d" % n) == 0.0
    assert _cmp(T["gll_wts_%d" % n], "get_gll_weights_%d" % n) == 0.0
    assert _cmp(T["coefs_to_gll_%d" % n], "coefs_to_gll_%d_%d" % (n, n)) == 0.0
    assert _cmp(T["coefs_to_deriv_%d" % n], "coefs_to_deriv_%d_%d" % (n, n)) == 0.0
    # the reference's inversion leaves |x| < 1e-32 where the exact entry is 0
    assert _cmp(T["gll_to_coefs_%d" % n], "gll_to_coefs_%d_%d" % (n, n), abs_tol=1e-31) == 0.0


@pytest.mark.parametrize("o", [3, 5, 7, 9])
def test_stencil_and_weno_tables_match_reference(o):
    assert _cmp(T["sten_to_coefs_%d" % o], "sten_to_coefs_%d_%d" % (o, o)) == 0.0
    # weno_sten_to_coefs sets polynomials 0..hs+1 only
    assert _cmp(T["weno_recon_%d" % o], "weno_sten_to_coefs_%d_%d_%d" % (o, o, o)) == 0.0
    for m in range(1, o + 1):
        assert _cmp(T["coefs_to_gll_lower_%d_%d" % (o, m)], "coefs_to_gll_lower_%d_%d" % (o, m)) == 0.0


@pytest.mark.parametrize("n", range(2, 10))
def test_tv_quadratic_forms_match_reference(n):
    ref = FIX["coefs_to_tv_%d" % n]["terms"]
    ours = T["tv_%d" % n]
    assert [(j, k) for j, k, _ in ours] == [(t[0], t[1]) for t in ref]
    for (j, k, c), t in zip(ours, ref):
        # the reference literal carries the _fp suffix: long double, then double
        mp.prec = 64
        ld = +mpf(t[2])
        mp.prec = 53
        assert c == float(+ld), (n, j, k)
    mp.prec = 200


def test_tv_is_truncated_integral():
    """Q3: the ord-5 a4^2 coefficient keeps only integrand degrees < 5 (625.8,
    not the untruncated 625.8357...)."""
    t5 = {(j, k): c for j, k, c in T["tv_5"]}
    assert t5[(4, 4)] == 625.8
    t9 = {(j, k): c for j, k, c in T["tv_9"]}
    assert abs(t9[(4, 4)] - 625.8357142857143) < 1e-12


def test_derivative_known_answer_notebook():
    """sage/derivatives.ipynb cell 1: (g2d2g)^3 * s_to_g * [1,0,2,0.5,3] at dx=1
    = (-6.5, -4.08257569495584, 0.5, 5.08257569495584, 7.5)."""
    D = np.array(T["deriv_5"])            # [in][out]
    s2g = np.array(T["sten_to_gll_5_5"])  # [cell][pt]
    u = np.array([1.0, 0.0, 2.0, 0.5, 3.0])
    g = s2g.T @ u
    for _ in range(3):
        g = D.T @ g
    expect = np.array([-6.5, -4.08257569495584, 0.5, 5.08257569495584, 7.5])
    assert np.allclose(g, expect, rtol=0, atol=5e-12)


@pytest.mark.parametrize("n", range(2, 10))
def test_gll_quadrature_exactness(n):
    x = np.array(T["gll_pts_%d" % n])
    w = np.array(T["gll_wts_%d" % n])
    for deg in range(0, 2 * n - 2):
        exact = ((0.5) ** (deg + 1) - (-0.5) ** (deg + 1)) / (deg + 1)
        assert abs(np.dot(w, x ** deg) - exact) < 1e-14
    assert np.all(np.diff(x) > 0) and x[0] == -0.5 and x[-1] == 0.5


@pytest.mark.parametrize("n", range(2, 10))
def test_deriv_matrix_exact_on_polynomials(n):
    D = np.array(T["deriv_%d" % n])
    x = np.array(T["gll_pts_%d" % n])
    for deg in range(n):
        f = x ** deg
        df = deg * x ** (deg - 1) if deg else np.zeros(n)
        assert np.allclose(D.T @ f, df, atol=1e-11 * max(1, n ** 3))


@pytest.mark.parametrize("o", [3, 5, 7, 9])
def test_weno_candidates_reproduce_cell_averages(o):
    hs = (o - 1) // 2
    R = np.array(T["weno_recon_%d" % o])
    rng = np.random.default_rng(o)
    u = rng.standard_normal(o)

    def avg(coefs, lo, hi):
        return sum(c * (hi ** (j + 1) - lo ** (j + 1)) / (j + 1) for j, c in enumerate(coefs)) / (hi - lo)

    for i in range(hs + 1):
        a = R[i][:hs + 1, :hs + 1].T @ u[i:i + hs + 1]
        for s in range(hs + 1):
            c = i + s - hs
            assert abs(avg(a, c - 0.5, c + 0.5) - u[i + s]) < 1e-12
    a = R[hs + 1].T @ u
    for s in range(o):
        c = s - hs
        assert abs(avg(a, c - 0.5, c + 0.5) - u[s]) < 1e-11


def test_oracle_spatial_init_matches_generated_composites():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    for o, n in [(3, 2), (5, 3), (7, 4), (9, 5), (5, 5)]:
        O = pyoracle.Oracle(8, 8, 8, ord=o, ngll=n)
        s2g = O.table(2).reshape(o, n)
        D = O.table(3).reshape(n, n)
        assert np.array_equal(s2g, np.array(T["sten_to_gll_%d_%d" % (o, n)]))
        assert np.array_equal(D, np.array(T["deriv_%d" % n]))
        idl = O.table(4)
        assert np.array_equal(idl[:-1], np.array(T["idl_%d" % o]))
        assert idl[-1] == T["sigma_%d" % o]
        assert abs(idl[:-1].sum() - 1) < 1e-15
