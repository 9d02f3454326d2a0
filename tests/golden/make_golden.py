"""Writes weno_faces.json: WENO face values in exact rational arithmetic.

An independent restatement from the mathematical definitions, not from the
formulas the oracle or the kernels use: for each candidate stencil the
polynomial whose cell averages (unit cells centred on integers) match the
data is solved for exactly; the candidate value is that polynomial at the
face x = 1/2; the Jiang-Shu smoothness indicator is
    beta = sum_{l>=1} int_{-1/2}^{1/2} (d^l P / dx^l)^2 dx.
The nonlinear weights then follow Borges et al. (2008) for WENO-Z5
(tau = |b0 - b2|), Castro, Costa & Don (2011) for WENO-Z7
(tau = |b0 - b1 - b2 + b3|), Jiang & Shu (1996) for WENO-JS5
(alpha = d / (beta + eps)^2); ideal weights (1/10, 6/10, 3/10) and
(1/35, 12/35, 18/35, 4/35); eps = (2^-52)^2 for the Z schemes, the double
nearest 1e-6 for JS.

Run from the repository root: python tests/golden/make_golden.py
"""
import json
import os
import random
from fractions import Fraction as Fr


def poly_from_averages(offsets, vals):
    """Coefficients c_0..c_{r-1} of P(x) = sum c_m x^m with cell averages vals on [o-1/2, o+1/2]."""
    r = len(offsets)
    # average of x^m over [o-1/2, o+1/2] = ((o+1/2)^{m+1} - (o-1/2)^{m+1}) / (m+1)
    A = [[((Fr(o) + Fr(1, 2)) ** (m + 1) - (Fr(o) - Fr(1, 2)) ** (m + 1)) / (m + 1) for m in range(r)]
         for o in offsets]
    b = [Fr(v) for v in vals]
    # Gauss-Jordan in exact arithmetic
    n = r
    M = [row[:] + [b[i]] for i, row in enumerate(A)]
    for c in range(n):
        p = next(i for i in range(c, n) if M[i][c] != 0)
        M[c], M[p] = M[p], M[c]
        piv = M[c][c]
        M[c] = [x / piv for x in M[c]]
        for i in range(n):
            if i != c and M[i][c] != 0:
                f = M[i][c]
                M[i] = [a - f * bb for a, bb in zip(M[i], M[c])]
    return [M[i][n] for i in range(n)]


def deriv(c):
    return [m * c[m] for m in range(1, len(c))]


def integ_sq(c):
    """int_{-1/2}^{1/2} P(x)^2 dx."""
    s = Fr(0)
    for a in range(len(c)):
        for b in range(len(c)):
            m = a + b
            s += c[a] * c[b] * ((Fr(1, 2)) ** (m + 1) - (Fr(-1, 2)) ** (m + 1)) / (m + 1)
    return s


def candidate(offsets, vals):
    c = poly_from_averages(offsets, vals)
    q = sum(cm * Fr(1, 2) ** m for m, cm in enumerate(c))
    beta = Fr(0)
    d = c
    for _ in range(len(offsets) - 1):
        d = deriv(d)
        beta += integ_sq(d)
    return q, beta


EPS_Z = Fr(1, 2 ** 104)
EPS_JS = Fr(1e-6)


def weno_z5(v, p):
    cands = [candidate([-2, -1, 0], v[0:3]), candidate([-1, 0, 1], v[1:4]), candidate([0, 1, 2], v[2:5])]
    d = [Fr(1, 10), Fr(6, 10), Fr(3, 10)]
    tau = abs(cands[0][1] - cands[2][1])
    a = [d[k] * (1 + (tau / (cands[k][1] + EPS_Z)) ** p) for k in range(3)]
    return sum(a[k] * cands[k][0] for k in range(3)) / sum(a)


def weno_z7(v, p):
    cands = [candidate([-3, -2, -1, 0], v[0:4]), candidate([-2, -1, 0, 1], v[1:5]),
             candidate([-1, 0, 1, 2], v[2:6]), candidate([0, 1, 2, 3], v[3:7])]
    d = [Fr(1, 35), Fr(12, 35), Fr(18, 35), Fr(4, 35)]
    b = [c[1] for c in cands]
    tau = abs(b[0] - b[1] - b[2] + b[3])
    a = [d[k] * (1 + (tau / (b[k] + EPS_Z)) ** p) for k in range(4)]
    return sum(a[k] * cands[k][0] for k in range(4)) / sum(a)


def weno_js5(v):
    cands = [candidate([-2, -1, 0], v[0:3]), candidate([-1, 0, 1], v[1:4]), candidate([0, 1, 2], v[2:5])]
    d = [Fr(1, 10), Fr(6, 10), Fr(3, 10)]
    a = [d[k] / (cands[k][1] + EPS_JS) ** 2 for k in range(3)]
    return sum(a[k] * cands[k][0] for k in range(3)) / sum(a)


def stencils(n, rng):
    out = []
    for _ in range(n):
        kind = rng.choice(["random", "smooth", "jump", "scaled"])
        if kind == "random":
            v = [rng.uniform(-1, 1) for _ in range(7)]
        elif kind == "smooth":
            a, b, c = rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(-0.2, 0.2)
            v = [a + b * 0.1 * x + c * (0.1 * x) ** 3 for x in range(-3, 4)]
        elif kind == "jump":
            s = rng.randint(1, 6)
            v = [0.0 if x < s else 1.0 for x in range(7)]
        else:
            v = [300.0 + rng.uniform(-0.1, 0.1) for _ in range(7)]
        out.append(v)
    return out


def main():
    rng = random.Random(20261202)
    cases = []
    for v in stencils(48, rng):
        v5 = v[1:6]
        cases.append({
            "v7": [x.hex() for x in v],
            "weno_z5_p1": float(weno_z5(v5, 1)).hex(),
            "weno_z5_p2": float(weno_z5(v5, 2)).hex(),
            "weno_z7_p2": float(weno_z7(v, 2)).hex(),
            "weno_js5": float(weno_js5(v5)).hex(),
        })
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weno_faces.json")
    with open(path, "w") as f:
        json.dump({"doc": "exact-rational WENO face values at i+1/2; v7 = f[i-3..i+3] (weno_z5/js5 use "
                          "f[i-2..i+2] = v7[1:6]); floats as hex", "cases": cases}, f, indent=0)
    print(f"wrote {len(cases)} cases to {path}")


if __name__ == "__main__":
    main()
