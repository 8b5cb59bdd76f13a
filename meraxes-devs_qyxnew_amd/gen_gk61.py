"""Generate csrc/mxr_gk61.hpp: the 61-point Gauss-Kronrod rule on [-1, 1].

The spin-temperature tables integrate with an adaptive Gauss-Kronrod (30, 61) rule (the rule
GSL's gsl_integration_qag uses for GSL_INTEG_GAUSS61, called at
src/core/XRayHeatingFunctions.c:414 and :848).  The nodes and weights are computed here from
their definition, in 80-digit arithmetic, and rounded to double:

  * Gauss nodes: the 30 roots of the Legendre polynomial P_30, weights 2 / ((1 - x^2) P_30'(x)^2);
  * Kronrod nodes: the 31 roots of the Stieltjes polynomial E_31, the degree-31 polynomial
    orthogonal to every polynomial of degree < 31 under the weight P_30 on [-1, 1]
    (coefficients solved exactly in rationals);
  * Kronrod weights: the interpolatory weights of all 61 nodes (moment equations in the
    Legendre basis).

Layout (QUADPACK's): xgk[0..30] are the non-negative nodes in decreasing order, the odd
entries being the Gauss nodes and xgk[30] = 0; wgk[j] is the Kronrod weight of xgk[j]; wg[j]
is the Gauss weight of xgk[2 j + 1].

    python gen_gk61.py            # rewrites csrc/mxr_gk61.hpp
"""
from __future__ import annotations

import os

import mpmath as mp
import sympy as sp

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "csrc", "mxr_gk61.hpp")
NG = 30  # Gauss points; Kronrod adds NG + 1


def _moment(m: int) -> sp.Rational:
    return sp.Rational(2, m + 1) if m % 2 == 0 else sp.Integer(0)


def _integral(poly: sp.Poly) -> sp.Rational:
    return sum((c * _moment(m[0]) for m, c in poly.terms()), sp.Integer(0))


def stieltjes(x: sp.Symbol) -> sp.Poly:
    """E_{NG+1} = P_{NG+1} + sum of lower P_k of the same parity, orthogonal to x^j P_NG."""
    P = [sp.Poly(sp.legendre_poly(k, x), x) for k in range(NG + 2)]
    ks = list(range(NG + 1 - 2, -1, -2))  # NG+1 is odd: odd k below it
    js = [j for j in range(NG + 1) if (j + NG + NG + 1) % 2 == 0]  # even integrands only
    A = sp.zeros(len(js), len(ks))
    b = sp.zeros(len(js), 1)
    for r, j in enumerate(js):
        w = sp.Poly(x ** j, x) * P[NG]
        b[r] = -_integral(P[NG + 1] * w)
        for cidx, k in enumerate(ks):
            A[r, cidx] = _integral(P[k] * w)
    a = A.LUsolve(b)
    E = P[NG + 1]
    for cidx, k in enumerate(ks):
        E = E + P[k] * a[cidx]
    return E


def roots_in(poly_coeffs, brackets):
    f = lambda t: mp.polyval(poly_coeffs, t)
    out = []
    for lo, hi in brackets:
        lo, hi = mp.mpf(lo), mp.mpf(hi)
        flo = f(lo)
        for _ in range(300):
            mid = (lo + hi) / 2
            fm = f(mid)
            if (fm > 0) == (flo > 0):
                lo, flo = mid, fm
            else:
                hi = mid
        out.append((lo + hi) / 2)
    return out


def compute():
    mp.mp.dps = 80
    x = sp.Symbol("x")
    PG = sp.Poly(sp.legendre_poly(NG, x), x)
    pg = [mp.mpf(sp.Rational(c).p) / mp.mpf(sp.Rational(c).q) for c in PG.all_coeffs()]
    # Gauss nodes: Newton from the classical cosine guesses, polished in 80 digits
    gauss = []
    for i in range(1, NG + 1):
        t = mp.cos(mp.pi * (i - mp.mpf(1) / 4) / (NG + mp.mpf(1) / 2))
        for _ in range(100):
            p, dp = mp.legendre(NG, t), NG * (t * mp.legendre(NG, t) - mp.legendre(NG - 1, t)) / (t * t - 1)
            t = t - p / dp
        gauss.append(t)
    gauss = sorted(gauss)
    # Kronrod nodes interlace with the Gauss nodes: one per gap and one beyond each end
    E = stieltjes(x)
    ec = [mp.mpf(sp.Rational(c).p) / mp.mpf(sp.Rational(c).q) for c in E.all_coeffs()]
    edges = [mp.mpf(-1)] + gauss + [mp.mpf(1)]
    kron = roots_in(ec, [(edges[i], edges[i + 1]) for i in range(len(edges) - 1)])
    nodes = sorted(gauss + kron)
    n = len(nodes)
    # interpolatory weights: sum_i w_i P_k(x_i) = int P_k, k < n
    A = mp.matrix(n, n)
    rhs = mp.matrix(n, 1)
    for k in range(n):
        for i, t in enumerate(nodes):
            A[k, i] = mp.legendre(k, t)
        rhs[k] = 2 if k == 0 else 0
    w = mp.lu_solve(A, rhs)
    wk = {nodes[i]: w[i] for i in range(n)}
    wgauss = {t: 2 / ((1 - t * t) * (NG * (t * mp.legendre(NG, t) - mp.legendre(NG - 1, t)) / (t * t - 1)) ** 2)
              for t in gauss}
    # QUADPACK layout: non-negative nodes, decreasing
    pos = sorted([t for t in nodes if t >= -mp.mpf(10) ** -60], reverse=True)
    pos[-1] = mp.mpf(0)
    xgk = pos
    wgk = [wk[t] if t != 0 else wk[min(nodes, key=lambda s: abs(s))] for t in pos]
    gpos = sorted([t for t in gauss if t > 0], reverse=True)
    wg = [wgauss[t] for t in gpos]
    for j, t in enumerate(gpos):
        assert abs(xgk[2 * j + 1] - t) < mp.mpf(10) ** -70, "Gauss nodes must sit at the odd entries"
    # exactness: the Kronrod rule integrates x^m exactly for m <= 3 NG + 1
    for m in range(0, 3 * NG + 2, 2):
        s = sum(wk[t] * t ** m for t in nodes)
        assert abs(s - mp.mpf(2) / (m + 1)) < mp.mpf(10) ** -60, m
    return xgk, wg, wgk


def emit(xgk, wg, wgk) -> str:
    def fmt(v):
        return mp.nstr(v, 21, min_fixed=-1, max_fixed=-1, strip_zeros=False)

    lines = [
        "// mxr_gk61.hpp -- 61-point Gauss-Kronrod rule on [-1, 1] (generated by gen_gk61.py).",
        "// xgk: non-negative nodes, decreasing; odd entries are the 30-point Gauss nodes, xgk[30] = 0.",
        "// wgk: Kronrod weight of each xgk; wg: Gauss weight of xgk[2 j + 1].",
        "#pragma once",
        "",
        "namespace mxr {",
        "",
        "static constexpr double GK61_XGK[31] = {",
    ]
    lines += [f"  {fmt(v)}," for v in xgk]
    lines += ["};", "", "static constexpr double GK61_WGK[31] = {"]
    lines += [f"  {fmt(v)}," for v in wgk]
    lines += ["};", "", "static constexpr double GK61_WG[15] = {"]
    lines += [f"  {fmt(v)}," for v in wg]
    lines += ["};", "", "}  // namespace mxr", ""]
    return "\n".join(lines)


if __name__ == "__main__":
    xgk, wg, wgk = compute()
    with open(OUT, "w") as f:
        f.write(emit(xgk, wg, wgk))
    print(OUT)
