"""Coefficients of the FP32 arccos used by the table lookups (mcrt_coop.hip).

    python montecarloradiativetransfer.jl_amd/csrc/fit_acos.py [degree]

acos(a) = √(1-a)·p(a) on [0, 1], acos(-a) = π - acos(a).  p is the weighted
minimax polynomial of f(a) = acos(a)/√(1-a) (f(1) = √2) with weight √(1-a), so
that the error it minimises is the absolute error of acos itself: the Remez
exchange algorithm on a fine grid, evaluated in 40-digit arithmetic (mpmath).
Degree 6 gives |acos error| ≤ 8.8e-8 rad, below the FP32 rounding of the table
coordinate it feeds (μ rounded to FP32 moves θ by ≥ 6e-8 rad, θ itself has an
ulp of 2.4e-7 near π); degree 7 reaches 1.2e-8.  The FP64 form uses degree 17
(8.5e-17 rad, below an ulp of θ; degree 16 gives 5.4e-16).  Prints the
coefficients c0…cd of p(a) = Σ cₖ aᵏ and the maximum error.
"""
from __future__ import annotations

import sys

import mpmath as mp
import numpy as np

mp.mp.dps = 40


def f(a):
    a = mp.mpf(a)
    return mp.acos(a) / mp.sqrt(1 - a) if a < 1 else mp.sqrt(2)


def remez(deg: int, iters: int = 40, grid: int = 40001):
    n = deg + 2
    x = [(1 - mp.cos(mp.pi * k / (n - 1))) / 2 for k in range(n)]
    g = np.linspace(0.0, 1.0, grid)
    fg = [f(t) for t in g]
    best = None
    for _ in range(iters):
        A = mp.matrix(n, n)
        b = mp.matrix(n, 1)
        for i, xi in enumerate(x):
            w = mp.sqrt(1 - xi) if xi < 1 else mp.mpf("1e-30")
            for j in range(deg + 1):
                A[i, j] = xi ** j
            A[i, deg + 1] = (-1) ** i / w
            b[i] = f(xi)
        sol = mp.lu_solve(A, b)
        c = [sol[j] for j in range(deg + 1)]
        err = np.array([float(mp.sqrt(1 - t) * (mp.polyval(c[::-1], t) - fv)) if t < 1 else 0.0
                        for t, fv in zip(g, fg)])
        emax = float(np.max(np.abs(err)))
        if best is None or emax < best[1]:
            best = (c, emax)
        # the new reference: one extremum of largest |error| per sign run
        pts = []
        for k in range(len(g)):
            if k not in (0, len(g) - 1) and (err[k] - err[k - 1]) * (err[k + 1] - err[k]) > 0:
                continue
            if pts and np.sign(err[k]) == np.sign(err[pts[-1]]):
                if abs(err[k]) > abs(err[pts[-1]]):
                    pts[-1] = k
            else:
                pts.append(k)
        while len(pts) > n:
            pts.pop(0 if abs(err[pts[0]]) < abs(err[pts[-1]]) else -1)
        if len(pts) < n:
            break
        x = [mp.mpf(g[k]) for k in pts]
    return best


if __name__ == "__main__":
    d = int(sys.argv[1]) if len(sys.argv) > 1 else 6
    c, e = remez(d)
    print(f"degree {d}: max |acos error| = {e:.3e} rad")
    print(", ".join(mp.nstr(v, 17) for v in c))
