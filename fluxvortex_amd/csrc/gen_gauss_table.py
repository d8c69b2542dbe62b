"""Writes fv_gauss_table.h: piecewise polynomials of the Gaussian-erf smoothing
factors for rho = r / sigma in [0, 10].

With g(rho) = erf(rho / sqrt2) - sqrt(2/pi) rho exp(-rho^2 / 2) and
g'(rho) = sqrt(2/pi) rho^2 exp(-rho^2 / 2) (the Gaussian-erf kernel of
src/fluxvortex/kernel.py:26-31), the pair factors are

    G1 = g / rho^3                      (velocity:   CG = -G1 / (4 pi sigma^3))
    G2 = g' / rho^4 - 3 g / rho^5       (gradient:   CH = -G2 / (4 pi sigma^5))
    G3 = g' / rho^2 - g / rho^3         (vorticity:  CH r^2 + 2 CG = -G3 / (4 pi sigma^3))

all even, entire functions of rho (the 1/rho^n terms cancel).  Each is fitted
on 20 pieces [k/2, (k+1)/2) by Chebyshev interpolation of degree 12 at 50
significant digits (mpmath; the series of g below rho = 1, the closed forms
above), converted to monomials in the piece's local variable t = 4 rho - 2k - 1
in [-1, 1] and rounded to double.  Horner's rule in double then reproduces
each factor to within about one unit in the last place of its terms' scale
(tests/test_gauss_table.py).

    python fluxvortex_amd/csrc/gen_gauss_table.py      (rewrites the header)
"""
from pathlib import Path

import mpmath as mp

PIECES, WIDTH, DEG, ROW = 20, mp.mpf(1) / 2, 12, 14
mp.mp.dps = 50
S2P = mp.sqrt(2 / mp.pi)


def g(r):
    return mp.erf(r / mp.sqrt(2)) - S2P * r * mp.exp(-r * r / 2)


def gp(r):
    return S2P * r * r * mp.exp(-r * r / 2)


def factor(which, r):
    """G1, G2 or G3 (which = 1, 2, 3) at rho = r, mpmath."""
    r = mp.mpf(r)
    if r < 1:
        # g = S2P sum_n c_n rho^(2n+3) / (2n+3), g' = S2P sum_n c_n rho^(2n+2),
        # c_n = (-1)^n / (2^n n!)
        s = mp.mpf(0)
        for n in range(60):
            c = (-1) ** n / (mp.mpf(2) ** n * mp.factorial(n))
            if which == 1:
                s += c * r ** (2 * n) / (2 * n + 3)
            elif which == 2:
                if n:
                    s += c * r ** (2 * n - 2) * (1 - mp.mpf(3) / (2 * n + 3))
            else:
                s += c * r ** (2 * n) * (1 - mp.mpf(1) / (2 * n + 3))
        return S2P * s
    if which == 1:
        return g(r) / r ** 3
    if which == 2:
        return gp(r) / r ** 4 - 3 * g(r) / r ** 5
    return gp(r) / r ** 2 - g(r) / r ** 3


def term_scale(which, r):
    """Magnitude of the terms a factor is formed from (its error scale)."""
    r = mp.mpf(r)
    if r < 1:
        return abs(factor(which, r))
    a, b = abs(g(r)), abs(gp(r))
    return {1: a / r ** 3, 2: b / r ** 4 + 3 * a / r ** 5, 3: b / r ** 2 + a / r ** 3}[which]


def piece_coefficients(which, k):
    """Monomial coefficients (in t) of the degree-DEG Chebyshev interpolant
    of factor `which` on piece k."""
    a, b = k * WIDTH, (k + 1) * WIDTH
    n = DEG + 1
    nodes = [mp.cos(mp.pi * (j + mp.mpf(1) / 2) / n) for j in range(n)]
    vals = [factor(which, (b - a) / 2 * t + (a + b) / 2) for t in nodes]
    cheb = []
    for j in range(n):
        s = sum(vals[i] * mp.cos(mp.pi * j * (i + mp.mpf(1) / 2) / n) for i in range(n))
        cheb.append(2 * s / n if j else s / n)
    T = [[mp.mpf(1)], [mp.mpf(0), mp.mpf(1)]]
    for j in range(2, n):
        p = [mp.mpf(0)] + [2 * x for x in T[j - 1]]
        q = T[j - 2] + [mp.mpf(0)] * (len(p) - len(T[j - 2]))
        T.append([p[i] - q[i] for i in range(len(p))])
    mono = [mp.mpf(0)] * n
    for j in range(n):
        for i, x in enumerate(T[j]):
            mono[i] += cheb[j] * x
    return [float(x) for x in mono]


def table():
    """[piece][factor][coefficient], factors G1, G2, G3."""
    return [[piece_coefficients(w, k) for w in (1, 2, 3)] for k in range(PIECES)]


def header(tab):
    lines = [
        "// fv_gauss_table.h -- written by gen_gauss_table.py; do not edit.",
        "// Piecewise polynomials of the Gaussian-erf pair factors G1 = g / rho^3,",
        "// G2 = g' / rho^4 - 3 g / rho^5, G3 = g' / rho^2 - g / rho^3 on rho in [0, 10]:",
        f"// piece k covers [k/2, (k+1)/2), t = 4 rho - 2k - 1, coefficients of t^0 .. t^{DEG}",
        f"// (each row padded to {ROW} doubles).",
        "#pragma once",
        "",
        "namespace fv {",
        "",
        f"constexpr int kGaussPieces = {PIECES};",
        f"constexpr int kGaussDeg = {DEG};",
        f"constexpr int kGaussRow = {ROW};",
        "",
        "__device__ const double kGaussTab[kGaussPieces][3][kGaussRow] = {",
    ]
    for k, piece in enumerate(tab):
        lines.append(f"    {{  // [{k / 2:g}, {(k + 1) / 2:g})")
        for cs in piece:
            vals = ", ".join(repr(c) for c in cs) + ", 0.0" * (ROW - len(cs))
            lines.append(f"        {{{vals}}},")
        lines.append("    },")
    lines += ["};", "", "}  // namespace fv", ""]
    return "\n".join(lines)


if __name__ == "__main__":
    out = Path(__file__).with_name("fv_gauss_table.h")
    out.write_text(header(table()))
    print("wrote", out)
