"""Derive the Taylor coefficients of the Slater overlap integral from its closed form.

pi l^3 exp(x) S(x l; l, l (1 + t)) = sum_n t^n sum_k V[n][k] x^k, with S the closed form
S = F/(8 pi r), F = 4 a^2 b^2/D^3 (e_a - e_b) + r/D^2 (a e_a + b e_b), D = b^2 - a^2.
Prints the exact rationals V[n][k] for n <= 6 (k <= n + 2); --json prints [[ [num, den], ... ], ...].
"""
import json
import sys

import sympy as sp


def derive(order=6):
    t, x = sp.symbols("t x", positive=True)
    E = sp.exp(x*t/(1+t))     # exp(x - x/(1+t)) = e_b/e_a with a = l, b = l(1+t), r = x l
    expr = (4*(1+t)**2*(1-E)/(t**3*(2+t)**3) + x*(1+(1+t)*E)/(t**2*(2+t)**2))/(8*x)
    series = sp.expand(sp.series(expr, t, 0, order+1).removeO())
    table = []
    for n in range(order+1):
        poly = sp.Poly(sp.expand(series.coeff(t, n)), x)
        table.append([sp.Rational(poly.coeff_monomial(x**k)) for k in range(n+3)])
    return table


if __name__ == "__main__":
    table = derive()
    if "--json" in sys.argv:
        print(json.dumps([[[int(c.p), int(c.q)] for c in row] for row in table]))
    else:
        for n, row in enumerate(table):
            print(n, " ".join(str(c) for c in row))
