"""Monte Carlo error statistics for bin arrays (the reference has none).

With thousands of bins a correct estimator still puts about 0.27% of them
outside 3σ, so agreement between two estimates is judged by (1) the count of
|z| > 3 against a binomial upper bound and (2) χ²/dof, not by "every bin".
The per-bin σ comes from batch means (simulate(..., sigma=...) or a batch's
read(..., sigma=...)).
"""
from __future__ import annotations

import math

import numpy as np


def compare(a, b, sigma, *, floor=0.0, dead_rel=1e-12, dof=None, label="", verbose=True):
    """z-statistics of a - b with per-element standard error sigma.

    Elements with sigma <= floor (deterministic values) must agree to within
    floor + dead_rel·scale.  When sigma is itself estimated from dof degrees of
    freedom (batch means), z follows Student's t: the exceedance probability
    and E[z²] = dof/(dof-2) replace the Gaussian ones.
    Returns a dict of summary statistics and `ok`."""
    a = np.asarray(a, float).ravel()
    b = np.asarray(b, float).ravel()
    s = np.asarray(sigma, float).ravel()
    scale = max(np.max(np.abs(b)) if b.size else 0.0, 1e-300)
    live = s > floor
    z = np.zeros_like(a)
    z[live] = (a[live] - b[live]) / s[live]
    dead_err = np.max(np.abs(a[~live] - b[~live])) if (~live).any() else 0.0
    n = int(live.sum())
    k = int((np.abs(z[live]) > 3).sum())
    if dof:
        from scipy.stats import t as student_t
        p = float(2 * student_t.sf(3.0, dof))
        # E[z²] is finite only for dof > 2, and χ²/dof is a usable test only
        # when its spread is moderate: below 5 dof the exceedance count decides
        ez2 = dof / (dof - 2.0) if dof >= 5 else math.inf
    else:
        p, ez2 = 0.0027, 1.0
    kmax = n * p + 4 * math.sqrt(max(n * p * (1 - p), 1e-12)) + 3
    chi2 = float(np.mean(z[live] ** 2)) if n else 0.0
    chi2max = ez2 * (1 + 5 * math.sqrt(2.0 / max(n, 1))) + 0.05
    ok = bool(k <= kmax and chi2 <= chi2max and dead_err <= floor + dead_rel * scale)
    out = dict(n=n, exceed3=k, exceed3_max=kmax, chi2dof=chi2, chi2dof_max=chi2max,
               max_abs_z=float(np.max(np.abs(z))) if n else 0.0, dead_err=float(dead_err), ok=ok)
    if verbose:
        print(f"[{label}] n={n} |z|>3: {k} (≤{kmax:.1f}) χ²/dof={chi2:.3f} (≤{chi2max:.3f}) "
              f"max|z|={out['max_abs_z']:.2f} dead={dead_err:.2e} -> {'ok' if ok else 'FAIL'}",
              flush=True)
    return out


def flat(res_tuple):
    """simulate 7-tuple → flat vector in the oracle's raw layout (A, Irt, Icb)."""
    A = np.array(res_tuple[:5], float)
    Irt = np.asarray(res_tuple[5]).transpose(2, 1, 0).ravel()
    Icb = np.asarray(res_tuple[6]).transpose(2, 1, 0).ravel()
    return np.concatenate([A, Irt, Icb])


def compare_results(a, b, sigma_b, *, batches, same_stream=True, label="", verbose=False):
    """Agreement of two simulate 7-tuples: `a` against `b` with b's standard
    errors (a dict with "A", "Irt", "Icb") estimated from `batches` batch means
    (the "batches" entry of the solve's stats).
    With same_stream the two used the same Philox stream and only b's σ counts
    (√2σ otherwise)."""
    s = flat((*sigma_b["A"], sigma_b["Irt"], sigma_b["Icb"]))
    if not same_stream:
        s = s * math.sqrt(2.0)
    r = compare(flat(a), flat(b), s, floor=1e-15, dead_rel=1e-6, dof=max(1, batches - 1),
                label=label, verbose=verbose)
    r["exceed3_frac"] = r["exceed3"] / max(r["n"], 1)
    return r
