"""ctypes wrapper of liboracle.so — TEST INFRASTRUCTURE ONLY.

The oracle is the checker: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg load it; the product never does.  It takes the same plain
structs (include/mcrt.h) as the product, built from the same Python objects.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

_dp = C.POINTER(C.c_double)
_lib = None


def build(verbose: bool = False) -> Path:
    """Compile liboracle.so with the Makefile beside this file."""
    r = subprocess.run(["make", "-B", "-C", str(HERE), "liboracle.so"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)
    if verbose:
        print(r.stdout)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        L.or_uniform.restype = C.c_double
        L.or_uniform.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint32]
        L.or_uniform_f32.restype = C.c_float
        L.or_uniform_f32.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint32]
        L.or_philox4x32_10.argtypes = [C.POINTER(C.c_uint32)] * 3
        L.or_kepler_solver.restype = C.c_double
        L.or_kepler_solver.argtypes = [C.c_double, C.c_double]
        L.or_phase_sample.restype = C.c_double
        L.or_phase_sample.argtypes = [C.c_void_p, C.c_double]
        L.or_phase_eval.argtypes = [C.c_void_p, C.c_double, _dp]
        L.or_amplitude.argtypes = [C.c_void_p, C.c_double, _dp]
        for n in ("or_spherical_from_cartesian",):
            getattr(L, n).argtypes = [_dp, _dp]
        for n in ("or_ray_to_norm", "or_norm_to_ray", "or_get_basis"):
            getattr(L, n).argtypes = [_dp, _dp, _dp]
        L.or_rotate.argtypes = [_dp, C.c_double, C.c_double, _dp]
        L.or_update_I.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp]
        L.or_update_E.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp]
        L.or_E_from_I.argtypes = [_dp, _dp]
        L.or_I_from_E.argtypes = [_dp, _dp]
        L.or_forward_E.restype = C.c_double
        L.or_forward_E.argtypes = [C.c_void_p, _dp, _dp, C.c_int64, _dp, _dp]
        L.or_backward_E.restype = C.c_double
        L.or_backward_E.argtypes = [C.c_void_p, _dp, _dp, C.c_int64, _dp, _dp]
        L.or_bhmie.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int] + [_dp] * 7
        L.or_calculate_csrn.argtypes = [_dp, C.c_int, C.c_int, _dp]
        L.or_gausslegendre.argtypes = [C.c_int, _dp, _dp]
        L.or_initial_weights.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp]
        L.or_distance_to_boundary.restype = C.c_double
        L.or_distance_to_boundary.argtypes = [C.c_void_p, _dp, _dp, C.c_int]
        L.or_trace_ray.argtypes = [C.c_void_p, C.c_void_p, _dp, C.c_uint64, C.c_uint32, C.c_uint64,
                                   _dp, _dp, _dp, C.POINTER(C.c_uint64)]
        L.or_simulate.argtypes = [C.c_void_p, C.c_void_p, _dp, C.c_uint64, C.c_uint32, C.c_int64,
                                  C.c_int64, C.c_int, C.c_int, _dp, C.POINTER(C.c_uint64)]
        L.or_collect_cb_event.argtypes = [C.c_void_p, C.c_void_p, _dp, _dp, _dp, _dp, C.c_int,
                                          C.c_int64, _dp, _dp, C.POINTER(C.c_int), _dp]
        _lib = L
    return _lib


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def v(*xs):
    return np.array(xs, dtype=np.float64)


# ------------------------------------------------------------------ scalars
def uniform(seed, instance, ray, draw):
    return lib().or_uniform(seed, instance, ray, draw)


def uniform_f32(seed, instance, ray, draw):
    return lib().or_uniform_f32(seed, instance, ray, draw)


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().or_philox4x32_10(c, k, o)
    return list(o)


def kepler(M, e):
    return lib().or_kepler_solver(M, e)


def _vec_fn(name, x, f=None, nout=3):
    x, px = _d(x)
    out = np.zeros(nout)
    if f is None:
        getattr(lib(), name)(px, out.ctypes.data_as(_dp))
    else:
        f, pf = _d(f)
        getattr(lib(), name)(px, pf, out.ctypes.data_as(_dp))
    return out


def spherical_from_cartesian(x):
    return _vec_fn("or_spherical_from_cartesian", x, nout=4)


def ray_to_norm(x, f):
    return _vec_fn("or_ray_to_norm", x, f)


def norm_to_ray(x, f):
    return _vec_fn("or_norm_to_ray", x, f)


def get_basis(k):
    k, pk = _d(k)
    Eh = np.zeros(3)
    Ev = np.zeros(3)
    lib().or_get_basis(pk, Eh.ctypes.data_as(_dp), Ev.ctypes.data_as(_dp))
    return Eh, Ev


def rotate(I, c2, s2):
    I, pI = _d(I)
    o = np.zeros(4)
    lib().or_rotate(pI, c2, s2, o.ctypes.data_as(_dp))
    return o


class Sca:
    """A scatterer struct for the oracle built from an mcrt scatterer object."""

    def __init__(self, s):
        from mcrt.scatterers import to_struct
        self.keep = []
        self.st = to_struct(s, self.keep)

    @property
    def ptr(self):
        return C.cast(C.pointer(self.st), C.c_void_p)


def phase_eval(sca: Sca, mu):
    """(P11, P12, P22, P33, P34, P44) of phase_matrix(s, μ)."""
    P = np.zeros(6)
    lib().or_phase_eval(sca.ptr, mu, P.ctypes.data_as(_dp))
    return P


def phase_sample(sca: Sca, u):
    return lib().or_phase_sample(sca.ptr, u)


def amplitude(sca: Sca, mu):
    S = np.zeros(4)
    lib().or_amplitude(sca.ptr, mu, S.ctypes.data_as(_dp))
    return S[0] + 1j * S[1], S[2] + 1j * S[3]


def update_I(sca: Sca, I, k, f):
    I, pI = _d(I)
    k, pk = _d(k)
    f, pf = _d(f)
    o = np.zeros(4)
    Eh = np.zeros(3)
    Ev = np.zeros(3)
    lib().or_update_I(sca.ptr, pI, pk, pf, o.ctypes.data_as(_dp), Eh.ctypes.data_as(_dp),
                      Ev.ctypes.data_as(_dp))
    return o, Eh, Ev


def _cpack(E):
    E = np.asarray(E, dtype=np.complex128)
    return np.ascontiguousarray(np.stack([E.real, E.imag], 1).reshape(-1))


def update_E(sca: Sca, E, k, f):
    e, pe = _d(_cpack(E))
    k, pk = _d(k)
    f, pf = _d(f)
    o = np.zeros(4)
    Eh = np.zeros(3)
    Ev = np.zeros(3)
    lib().or_update_E(sca.ptr, pe, pk, pf, o.ctypes.data_as(_dp), Eh.ctypes.data_as(_dp),
                      Ev.ctypes.data_as(_dp))
    return np.array([o[0] + 1j * o[1], o[2] + 1j * o[3]]), Eh, Ev


def E_from_I(I):
    I, pI = _d(I)
    o = np.zeros(4)
    lib().or_E_from_I(pI, o.ctypes.data_as(_dp))
    return np.array([o[0] + 1j * o[1], o[2] + 1j * o[3]])


def I_from_E(E):
    e, pe = _d(_cpack(E))
    o = np.zeros(4)
    lib().or_I_from_E(pe, o.ctypes.data_as(_dp))
    return o


def forward_E(sca: Sca, E, kpath, f0, backward=False):
    e, pe = _d(_cpack(E))
    kp, pk = _d(np.asarray(kpath, dtype=np.float64).reshape(-1))
    f0, pf = _d(f0)
    o = np.zeros(4)
    fn = lib().or_backward_E if backward else lib().or_forward_E
    en = fn(sca.ptr, pe, pk, len(kp) // 3, pf, o.ctypes.data_as(_dp))
    return np.array([o[0] + 1j * o[1], o[2] + 1j * o[3]]), en


def bhmie(x, m, Ntheta=181, nextra=15, custom_nstop=0):
    m = complex(m)
    q = np.zeros(5)
    S1 = np.zeros(2 * Ntheta)
    S2 = np.zeros(2 * Ntheta)
    F = [np.zeros(Ntheta) for _ in range(4)]
    rc = lib().or_bhmie(x, m.real, m.imag, Ntheta, nextra, custom_nstop, q.ctypes.data_as(_dp),
                        S1.ctypes.data_as(_dp), S2.ctypes.data_as(_dp),
                        *[f.ctypes.data_as(_dp) for f in F])
    assert rc == 0
    return dict(q_ext=q[0], q_sca=q[1], q_abs=q[2], q_back=q[3], asymm=q[4],
                S1=S1[0::2] + 1j * S1[1::2], S2=S2[0::2] + 1j * S2[1::2], F11=F[0], F12=F[1],
                F33=F[2], F34=F[3])


def calculate_csrn(F11, n):
    F11, pF = _d(F11)
    out = np.zeros(n + 1)
    assert lib().or_calculate_csrn(pF, len(F11), n, out.ctypes.data_as(_dp)) == 0
    return out


def gausslegendre(n):
    x = np.zeros(n)
    w = np.zeros(n)
    assert lib().or_gausslegendre(n, x.ctypes.data_as(_dp), w.ctypes.data_as(_dp)) == 0
    return x, w


def initial_weights(kind, Nt, Np):
    a, b, c = np.zeros(Nt), np.zeros(Nt), np.zeros(Nt)
    assert lib().or_initial_weights(kind, Nt, Np, a.ctypes.data_as(_dp), b.ctypes.data_as(_dp),
                                    c.ctypes.data_as(_dp)) == 0
    return a, b, c


class Solve:
    """(cfg, geometry) marshalled once for the oracle."""

    def __init__(self, cfg, g):
        from mcrt.api import marshal
        self.cfg = cfg
        self.cs, self.gs, _, self.keep = marshal(cfg, g, [1.0, 0.0, 0.0, 0.0])

    @property
    def cptr(self):
        return C.cast(C.pointer(self.cs), C.c_void_p)

    @property
    def gptr(self):
        return C.cast(C.pointer(self.gs), C.c_void_p)

    def distance(self, x, k, idx=0):
        x, px = _d(x)
        k, pk = _d(k)
        return lib().or_distance_to_boundary(self.gptr, px, pk, idx)

    def trace_ray(self, I0, seed, ray, instance=0):
        cfg = self.cfg
        A = np.zeros(5)
        Irt = np.zeros(4 * cfg.Ntheta * cfg.Nphi)
        Icb = np.zeros(max(1, 4 * cfg.Nthetab * cfg.Nphib))
        cnt = (C.c_uint64 * 6)()
        i0, p0 = _d(I0)
        rc = lib().or_trace_ray(self.cptr, self.gptr, p0, seed, instance, ray, A.ctypes.data_as(_dp),
                                Irt.ctypes.data_as(_dp), Icb.ctypes.data_as(_dp), cnt)
        assert rc == 0, rc
        return A, Irt, Icb, list(cnt)

    def collect_cb_event(self, I0, I, k, x, idx, kpath, xpath, idxpath=None):
        """collect_cb_cpu! (cpu.jl:38-160) for one event: the event number is
        len(kpath); returns Icb[4, Nθb, Nϕb] of this event alone."""
        cfg = self.cfg
        kp = np.ascontiguousarray(np.asarray(kpath, dtype=np.float64).reshape(-1))
        xp = np.ascontiguousarray(np.asarray(xpath, dtype=np.float64).reshape(-1))
        n = len(kp) // 3
        ip = (C.c_int * n)(*(idxpath if idxpath is not None else [0] * n))
        Icb = np.zeros(4 * cfg.Nthetab * cfg.Nphib)
        a = [_d(v) for v in (I0, I, k, x)]
        rc = lib().or_collect_cb_event(self.cptr, self.gptr, a[0][1], a[1][1], a[2][1], a[3][1], idx,
                                       n, kp.ctypes.data_as(_dp), xp.ctypes.data_as(_dp), ip,
                                       Icb.ctypes.data_as(_dp))
        assert rc == 0, rc
        return Icb.reshape(cfg.Nphib, cfg.Nthetab, 4).transpose(2, 1, 0)

    def simulate(self, I0, seed, n_rays=None, ray_begin=0, nthreads=None, nbatch=32, instance=0):
        """Raw batch sums: returns dict(A, Irt, Icb) totals normalised by N, per-batch
        means (for σ), and counters."""
        cfg = self.cfg
        N = cfg.number_of_rays if n_rays is None else n_rays
        nbatch = max(1, min(nbatch, N))
        nrt = 4 * cfg.Ntheta * cfg.Nphi
        ncb = 4 * cfg.Nthetab * cfg.Nphib
        stride = 5 + nrt + ncb
        out = np.zeros(nbatch * stride)
        cnt = (C.c_uint64 * 6)()
        i0, p0 = _d(I0)
        nth = nthreads or min(os.cpu_count() or 1, 16)
        rc = lib().or_simulate(self.cptr, self.gptr, p0, seed, instance, ray_begin, ray_begin + N,
                               nth, nbatch, out.ctypes.data_as(_dp), cnt)
        assert rc == 0, rc
        B = out.reshape(nbatch, stride)
        tot = B.sum(axis=0) / N
        # per-batch estimates of the normalised quantities (batch sizes nearly equal)
        edges = [(N * b) // nbatch for b in range(nbatch + 1)]
        sizes = np.diff(edges).astype(float)
        means = B / sizes[:, None]
        return dict(
            A=tot[:5], Irt=tot[5:5 + nrt].reshape(cfg.Nphi, cfg.Ntheta, 4).transpose(2, 1, 0),
            Icb=tot[5 + nrt:].reshape(cfg.Nphib, cfg.Nthetab, 4).transpose(2, 1, 0),
            batch_means=means, batch_sizes=sizes, counters=list(cnt), N=N,
            nrt=nrt, ncb=ncb)


def sigma_of_mean(res):
    """Per-quantity standard error of the N-ray mean from batch means (flat layout
    A[5], Irt, Icb as in the raw buffers)."""
    m = res["batch_means"]
    w = res["batch_sizes"]
    B = m.shape[0]
    mu = (m * w[:, None]).sum(0) / w.sum()
    var = ((m - mu) ** 2 * w[:, None]).sum(0) / w.sum() * B / (B - 1)
    return np.sqrt(var / B)
