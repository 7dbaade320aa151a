"""Write tests/golden/bhmie_golden.json: 50-digit mpmath evaluations of the
Bohren–Huffman series exactly as src/scatterers/mie.jl:129-216 truncates it
(nstop = ⌊max(x + 4∛x + 2, |m|x)⌋, nmax = nstop + 15, D_n downward from
D_nmax = 0, angles θ_i = i·π/(Nθ-1) with the mirrored half by parity).

The fixture pins the FP64 builders (oracle and product) at 1e-12 relative to
the table's largest magnitude.  Run:  python tests/golden/make_bhmie_golden.py
"""
import json
import math
from pathlib import Path

import mpmath as mp

mp.mp.dps = 50

CASES = [
    # (x, m, Nθ, angle stride of the stored subset)
    (2 * math.pi * 0.5 / 0.55, complex(1.5, 0.001), 1801, 60),  # the sphere Mie workload
    (0.5, complex(1.3, 1e-5), 181, 10),
    (20.0, complex(1.7, 0.1), 181, 10),
    (10.0, complex(1.45, 0.01), 361, 20),
    (3.0, complex(1.33, 0.0), 181, 10),
]


def theta_f64(i, N):
    # range(0.0, Float64(π), N)[i+1], correctly rounded
    if i == N - 1:
        return math.pi
    return float(mp.mpf(math.pi) * i / (N - 1))


def bhmie_mp(x, m, N, idx):
    x = mp.mpf(x)
    m = mp.mpc(m.real, m.imag)
    y = m * x
    nstop = int(math.floor(float(max(x + 4 * mp.cbrt(x) + 2, x * abs(m)))))
    nmax = nstop + 15
    d = [mp.mpc(0)] * (nmax + 1)
    for n in range(nmax - 1, 0, -1):
        d[n] = (n + 1) / y - 1 / (d[n + 1] + (n + 1) / y)
    half = (N + 1) // 2
    # angles needed: each stored index maps to a first-half index (direct or mirrored)
    need = sorted({i if i < half else N - 1 - i for i in idx})
    mu = {i: mp.cos(mp.mpf(theta_f64(i, N))) for i in need}
    pi0 = {i: mp.mpf(0) for i in need}
    pi1 = {i: mp.mpf(1) for i in need}
    S1 = {i: mp.mpc(0) for i in idx}
    S2 = {i: mp.mpc(0) for i in idx}
    psi0, psi1 = mp.cos(x), mp.sin(x)
    chi0, chi1 = -mp.sin(x), mp.cos(x)
    xi1 = mp.mpc(psi1, -chi1)
    a_prev = b_prev = None
    qsca = mp.mpf(0)
    asymm = mp.mpf(0)
    for n in range(1, nstop + 1):
        fn = mp.mpf(2 * n + 1) / (n * (n + 1))
        psi = (2 * n - 1) * psi1 / x - psi0
        chi = (2 * n - 1) * chi1 / x - chi0
        xi = mp.mpc(psi, -chi)
        an = ((d[n] / m + n / x) * psi - psi1) / ((d[n] / m + n / x) * xi - xi1)
        bn = ((d[n] * m + n / x) * psi - psi1) / ((d[n] * m + n / x) * xi - xi1)
        qsca += (abs(an) ** 2 + abs(bn) ** 2) * (2 * n + 1)
        asymm += (an.real * bn.real + an.imag * bn.imag) * fn
        if n > 1:
            asymm += (a_prev.real * an.real + a_prev.imag * an.imag + b_prev.real * bn.real
                      + b_prev.imag * bn.imag) * (n - 1) * (n + 1) / n
        p = -1 if n % 2 == 0 else 1
        pin = {}
        taun = {}
        for i in need:
            pin[i] = pi1[i]
            taun[i] = n * mu[i] * pin[i] - (n + 1) * pi0[i]
        for i in idx:
            if i < half:
                S1[i] += fn * (an * pin[i] + bn * taun[i])
                S2[i] += fn * (an * taun[i] + bn * pin[i])
            else:
                j = N - 1 - i
                if j < i:
                    S1[i] += fn * (an * pin[j] - bn * taun[j]) * p
                    S2[i] += fn * (-an * taun[j] + bn * pin[j]) * p
                else:  # the middle angle of odd N
                    S1[i] += fn * (an * pin[j] + bn * taun[j])
                    S2[i] += fn * (an * taun[j] + bn * pin[j])
        psi0, psi1 = psi1, psi
        chi0, chi1 = chi1, chi
        xi1 = mp.mpc(psi1, -chi1)
        a_prev, b_prev = an, bn
        for i in need:
            pi1[i] = ((2 * n + 1) * mu[i] * pin[i] - (n + 1) * pi0[i]) / n
            pi0[i] = pin[i]
    asymm *= 2 / qsca
    qsca *= 2 / x ** 2
    return nstop, qsca, asymm, S1, S2


def main():
    out = []
    for x, m, N, stride in CASES:
        idx = sorted(set(list(range(0, N, stride)) + [N - 1, 1, N - 2]))
        half = (N + 1) // 2
        idx_all = sorted(set(idx) | {0, N - 1})
        nstop, qsca, asymm, S1, S2 = bhmie_mp(x, m, N, idx_all)
        qext = 4 / mp.mpf(x) ** 2 * S1[0].real
        qback = 4 / mp.mpf(x) ** 2 * abs(S1[N - 1]) ** 2
        coeff = 4 / (qsca * mp.mpf(x) ** 2)
        rows = []
        for i in idx_all:
            s1, s2 = S1[i], S2[i]
            a1, a2 = abs(s1) ** 2, abs(s2) ** 2
            s21 = s2 * mp.conj(s1)
            rows.append({
                "i": i,
                "S1": [float(s1.real), float(s1.imag)],
                "S2": [float(s2.real), float(s2.imag)],
                "F11": float((a1 + a2) / 2 * coeff), "F12": float(-(a1 - a2) / 2 * coeff),
                "F33": float(s21.real * coeff), "F34": float(s21.imag * coeff),
            })
        out.append({"x": x, "m": [m.real, m.imag], "Ntheta": N, "nstop": nstop,
                    "q_ext": float(qext), "q_sca": float(qsca), "q_abs": float(qext - qsca),
                    "q_back": float(qback), "asymm": float(asymm), "angles": rows})
        print(f"x={x:.5f} m={m} N={N}: q_ext={float(qext):.15g} q_sca={float(qsca):.15g}")
    Path(__file__).with_name("bhmie_golden.json").write_text(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
