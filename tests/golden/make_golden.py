"""Generate exact-arithmetic golden fixtures for the LGL basis and ln_mean.

Run:  python tests/golden/make_golden.py
Writes tests/golden/basis_N{2,3,4,5,6}.json and tests/golden/ln_mean.json.

The fixtures come from first principles, not from any code under test:
  * LGL nodes: the roots of (1 - x^2) P'_{N-1}(x), solved exactly by sympy
    (closed forms for N <= 5, 60-digit numerics above);
  * weights: 2 / (n (n+1) P_n(x_i)^2) with n = N - 1;
  * D: derivative of the Lagrange basis at the nodes, D[i, j] = l_j'(x_i);
  * D_split = 2 D - M^{-1} B with B = diag(-1, 0, ..., 0, 1);
  * D_hat = -M^{-1} D^T M;
  * lhat(+-1) = l_j(+-1) / w_j;
  * inverse Vandermonde of the normalized Legendre polynomials
    sqrt(m + 1/2) P_m(x_i) (the normalization of Trixi's vandermonde_legendre);
  * ln_mean(x, y) = (y - x) / log(y / x) in 50-digit mpmath, ln_mean(x, x) = x.
"""
import json
import os

import mpmath
import sympy as sp

HERE = os.path.dirname(os.path.abspath(__file__))
DIGITS = 40


def lgl_nodes(N):
    x = sp.Symbol("x")
    n = N - 1
    poly = sp.expand((1 - x**2) * sp.diff(sp.legendre(n, x), x))
    roots = sp.Poly(poly, x).all_roots()
    roots = sorted(roots, key=lambda r: float(sp.N(r, 50)))
    if N > 5:
        roots = [sp.Float(sp.N(r, 60), 60) for r in roots]
    return [sp.nsimplify(r) if N <= 5 else r for r in roots]


def basis(N):
    x = sp.Symbol("x")
    n = N - 1
    nodes = lgl_nodes(N)
    weights = [sp.simplify(sp.Rational(2, n * (n + 1)) / sp.legendre(n, xi) ** 2) for xi in nodes]
    lag = []
    for j in range(N):
        lj = sp.Integer(1)
        for k in range(N):
            if k != j:
                lj = lj * (x - nodes[k]) / (nodes[j] - nodes[k])
        lag.append(sp.expand(lj))
    D = sp.Matrix(N, N, lambda i, j: sp.simplify(sp.diff(lag[j], x).subs(x, nodes[i])))
    M = sp.diag(*weights)
    Bm = sp.zeros(N, N)
    Bm[0, 0], Bm[N - 1, N - 1] = -1, 1
    Dsplit = (2 * D - M.inv() * Bm).applyfunc(sp.simplify)
    Dhat = (-M.inv() * D.T * M).applyfunc(sp.simplify)
    lm = [sp.simplify(lag[j].subs(x, -1) / weights[j]) for j in range(N)]
    lp = [sp.simplify(lag[j].subs(x, 1) / weights[j]) for j in range(N)]
    V = sp.Matrix(N, N, lambda i, m: sp.sqrt(sp.Rational(2 * m + 1, 2)) * sp.legendre(m, nodes[i]))
    Vinv = V.inv().applyfunc(sp.simplify)

    def f(v):
        return sp.N(v, DIGITS).__str__()

    def mat(Mx):
        return [[f(Mx[i, j]) for j in range(N)] for i in range(N)]

    return {
        "N": N,
        "nodes": [f(v) for v in nodes],
        "weights": [f(v) for v in weights],
        "D": mat(D),
        "Dsplit": mat(Dsplit),
        "Dhat": mat(Dhat),
        "lhat_minus": [f(v) for v in lm],
        "lhat_plus": [f(v) for v in lp],
        "Vinv": mat(Vinv),
    }


def ln_mean_cases():
    mpmath.mp.dps = 50
    pts = []
    xs = [1.0, 0.5, 2.0, 1.3, 0.01, 100.0]
    rel = [0.0, 1e-12, 1e-8, 1e-5, 1e-3, 4.9e-3, 5.0e-3, 5.1e-3, 1e-2, 0.1, 0.5, 1.0, 3.0]
    for x in xs:
        for r in rel:
            y = x * (1.0 + r)
            X, Y = mpmath.mpf(x), mpmath.mpf(y)
            if X == Y:
                lm = X
            else:
                lm = (Y - X) / mpmath.log(Y / X)
            pts.append({"x": x, "y": y, "ln_mean": mpmath.nstr(lm, 30), "inv_ln_mean": mpmath.nstr(1 / lm, 30)})
    return pts


def main():
    for N in (2, 3, 4, 5, 6):
        with open(os.path.join(HERE, f"basis_N{N}.json"), "w") as fh:
            json.dump(basis(N), fh, indent=1)
    with open(os.path.join(HERE, "ln_mean.json"), "w") as fh:
        json.dump(ln_mean_cases(), fh, indent=1)


if __name__ == "__main__":
    main()
