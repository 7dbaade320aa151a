"""The comparator has power: a mistake planted in the oracle (on the CPU) is
detected, while two independent correct estimates pass."""
import numpy as np

import mcrt
import oracle
from statcheck import compare, oracle_flat


def _run(omega, seed, n=3000):
    cfg, g, _ = mcrt.workloads.plane_rayleigh_tau1(n, omega=omega, max_orders=50, Ntheta=40, Nphi=8)
    return oracle.Solve(cfg, g).simulate([1.0, -1.0, 0.0, 0.0], seed, nbatch=32)


def test_independent_correct_estimates_pass():
    a, b = _run(1.0, 1), _run(1.0, 2)
    sig = np.sqrt(oracle.sigma_of_mean(a) ** 2 + oracle.sigma_of_mean(b) ** 2)
    assert compare(oracle_flat(a), oracle_flat(b), sig, floor=1e-15, label="same physics")["ok"]


def test_planted_albedo_error_is_detected():
    a, b = _run(1.0, 1), _run(0.97, 2)  # 3% absorption planted in one run
    sig = np.sqrt(oracle.sigma_of_mean(a) ** 2 + oracle.sigma_of_mean(b) ** 2)
    assert not compare(oracle_flat(a), oracle_flat(b), sig, floor=1e-15, label="planted")["ok"]
