"""The C ABI (include/mcrt.h) and libmcrt_hip.so on the CPU: the library loads,
exports every declared entry point, the ctypes mirror matches the C layout,
and the host-side FP64 table builders agree with the oracle and the golden
series.  No GPU compute is called here."""
import ctypes as C
import json
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import mcrt
import oracle
from mcrt import _abi

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "mcrt.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(mcrt_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = _abi.load_library()
    names = declared_functions()
    assert len(names) >= 16
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(_abi.SIGNATURES), set(names) ^ set(_abi.SIGNATURES)
    out = subprocess.run(["nm", "-D", "--defined-only", str(_abi.LIB_PATH)], capture_output=True,
                         text=True).stdout
    for n in names:
        assert re.search(rf"\bT {n}\b", out), n


def test_struct_layout_matches_c(tmp_path):
    src = tmp_path / "layout.c"
    structs = {"mcrt_scatterer": _abi.mcrt_scatterer, "mcrt_layer": _abi.mcrt_layer,
               "mcrt_geometry": _abi.mcrt_geometry, "mcrt_config": _abi.mcrt_config,
               "mcrt_stats": _abi.mcrt_stats, "mcrt_request": _abi.mcrt_request,
               "mcrt_server_stats": _abi.mcrt_server_stats,
               "mcrt_launch_shape": _abi.mcrt_launch_shape}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){"]
    for sname, cls in structs.items():
        lines.append(f'printf("{sname} %zu\\n", sizeof({sname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{sname}.{fname} %zu\\n", offsetof({sname}, {fname}));')
    lines.append("return 0;}")
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-o", str(exe), str(src)], check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                  check=True).stdout.splitlines())
    for sname, cls in structs.items():
        assert int(got[sname]) == C.sizeof(cls), sname
        for fname, _ in cls._fields_:
            assert int(got[f"{sname}.{fname}"]) == getattr(cls, fname).offset, (sname, fname)


def test_version_and_no_device_error():
    lib = _abi.load_library()
    assert b"gfx950" in lib.mcrt_version()
    n = C.c_int(-1)
    assert lib.mcrt_device_count(C.byref(n)) == 0
    if n.value == 0:
        h = C.c_void_p()
        rc = lib.mcrt_context_create(0, C.byref(h))
        assert rc == _abi.MCRT_ERR_NO_DEVICE
        assert b"no HIP device" in lib.mcrt_last_error(None)


def test_product_bhmie_matches_oracle_and_golden():
    G = json.loads((Path(__file__).parent / "golden" / "bhmie_golden.json").read_text())
    for c in G:
        p = mcrt.bhmie(c["x"], complex(*c["m"]), Ntheta=c["Ntheta"])
        o = oracle.bhmie(c["x"], complex(*c["m"]), c["Ntheta"])
        for k in ("F11", "F12", "F33", "F34"):
            a = getattr(p, k)
            assert np.max(np.abs(a - o[k])) <= 1e-13 * np.max(np.abs(o[k])), k
        idx = [a["i"] for a in c["angles"]]
        for k in ("F11", "F12", "F33", "F34"):
            g = np.array([a[k] for a in c["angles"]])
            assert np.max(np.abs(g - getattr(p, k)[idx])) <= 1e-12 * np.max(np.abs(g))
        for k in ("q_ext", "q_sca", "q_back"):
            assert abs(getattr(p, k) - c[k]) <= 1e-12 * abs(c[k])


def test_product_csrn_and_quadrature_match_oracle():
    s = mcrt.workloads.c2_mie(1801)
    assert np.max(np.abs(s.csrn - oracle.calculate_csrn(s.F11, 1000))) < 1e-14
    for n in (3, 48, 90, 180):
        x, w = mcrt.gausslegendre(n)
        xo, wo = oracle.gausslegendre(n)
        assert np.max(np.abs(x - xo)) < 1e-15 and np.max(np.abs(w - wo)) < 1e-15
        xn, wn = np.polynomial.legendre.leggauss(n)
        assert np.max(np.abs(x - xn)) < 2e-15
    for kind, g in ((0, mcrt.PlaneGeometry(3.0, 0.0, [mcrt.Layer(d=1.0, scatterer=mcrt.RayleighScatterer(1, 1))])),
                    (1, mcrt.SphereGeometry(R=1.0, scatterer=mcrt.RayleighScatterer(1, 1)))):
        a = mcrt.initial_weights(g, 180, 48)
        b = oracle.initial_weights(kind, 180, 48)
        for u, v in zip(a, b):
            assert np.max(np.abs(u - v)) < 1e-15


def test_invalid_arguments_are_rejected():
    lib = _abi.load_library()
    with pytest.raises(_abi.ArgumentError):
        mcrt.bhmie(-1.0, 1.5, Ntheta=181)
    with pytest.raises(_abi.ArgumentError):
        mcrt.calculate_csrn(np.ones(1), 10)
    assert lib.mcrt_gauss_legendre(0, None, None) == _abi.MCRT_ERR_INVALID


def test_reference_validate_rules():
    r = mcrt.RayleighScatterer(1.0, 1.0)
    with pytest.raises(ValueError, match="θ must be"):
        mcrt.PlaneGeometry(1.0, 0.0, [mcrt.Layer(d=1.0, scatterer=r)]).validate()
    with pytest.raises(ValueError, match="ϕ must be"):
        mcrt.PlaneGeometry(3.0, 7.0, [mcrt.Layer(d=1.0, scatterer=r)]).validate()
    with pytest.raises(ValueError, match="at least one layer"):
        mcrt.PlaneGeometry(3.0, 0.0, []).validate()
    with pytest.raises(ValueError, match="Only the last layer"):
        mcrt.PlaneGeometry(3.0, 0.0, [mcrt.Layer(d=-1.0, scatterer=r), mcrt.Layer(d=1.0, scatterer=r)]).validate()
    with pytest.raises(ValueError, match="Radius"):
        mcrt.SphereGeometry(R=-1.0, scatterer=r, tau=1.0).validate()


def test_config_defaults_match_reference():
    g = mcrt.SphereGeometry(R=2.0, scatterer=mcrt.RayleighScatterer(0.9, 0.5))
    assert g.tau == 4.0 and g.omega == 0.9
    c = mcrt.Config(g)
    assert (c.number_of_rays, c.minimum_intensity, c.maximum_scattering_times, c.maximum_threads,
            c.compute_cb, c.tau0, c.Nphi, c.Ntheta, c.Nthetab, c.Nphib) == (
        1000, 1e-6, 10000, 8192, True, 50.0, 48, 64, 51, 8)
    assert np.allclose(c.phi, 2 * np.pi * np.arange(48) / 48)
    assert c.thetab[-1] == np.pi
    c2 = mcrt.Config(g, **{"Nθ": 32, "τ₀": 10.0})
    assert c2.Ntheta == 32 and c2.tau0 == 10.0 and len(c2.costheta) == 32


def test_no_cpu_fallback():
    g = mcrt.SphereGeometry(R=2.0, scatterer=mcrt.RayleighScatterer(0.9, 0.5))
    with pytest.raises(NotImplementedError):
        mcrt.simulate(mcrt.Config(g), g, [1, -1, 0, 0], use_gpu=False)


JULIA_TYPES = {"Int32": (4, 4), "UInt32": (4, 4), "Int64": (8, 8), "UInt64": (8, 8),
               "Float64": (8, 8), "NTuple{4, Float64}": (32, 8)}
JULIA_STRUCTS = {"McrtScatterer": "mcrt_scatterer", "McrtLayer": "mcrt_layer",
                 "McrtGeometry": "mcrt_geometry", "McrtConfig": "mcrt_config",
                 "McrtStats": "mcrt_stats", "McrtRequest": "mcrt_request"}


def test_julia_binding_mirrors_c_layout(tmp_path):
    """julia/src/hip.jl declares the C structs field for field: same names in
    the same order, and C alignment of its field types gives the offsets and
    sizes of include/mcrt.h (Julia is not installed here, so the binding is
    checked by its text)."""
    src = (ROOT / "julia" / "src" / "hip.jl").read_text()
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){"]
    parsed = {}
    for jname, cname in JULIA_STRUCTS.items():
        m = re.search(rf"^struct {jname}\b.*?\n(.*?)^end", src, flags=re.S | re.M)
        assert m, jname
        fields = []
        for line in m.group(1).splitlines():
            line = line.split("#")[0].strip()
            if not line:
                continue
            name, typ = [x.strip() for x in line.split("::", 1)]
            fields.append((name, typ))
        off, align = 0, 1
        layout = []
        for name, typ in fields:
            size, al = (8, 8) if typ.startswith("Ptr{") else JULIA_TYPES[typ]
            off = (off + al - 1) // al * al
            layout.append((name, off))
            off += size
            align = max(align, al)
        parsed[cname] = (layout, (off + align - 1) // align * align)
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for name, _ in layout:
            lines.append(f'printf("{cname}.{name} %zu\\n", offsetof({cname}, {name}));')
    lines.append("return 0;}")
    c = tmp_path / "jl_layout.c"
    c.write_text("\n".join(lines))
    exe = tmp_path / "jl_layout"
    subprocess.run(["gcc", "-o", str(exe), str(c)], check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                  check=True).stdout.splitlines())
    for cname, (layout, size) in parsed.items():
        assert int(got[cname]) == size, cname
        for name, off in layout:
            assert int(got[f"{cname}.{name}"]) == off, (cname, name)
        # every C field is mirrored
        assert len(layout) == len(getattr(_abi, cname)._fields_), cname


def test_julia_binding_calls_only_declared_functions():
    src = (ROOT / "julia" / "src" / "hip.jl").read_text()
    called = set(re.findall(r"ccall\(\(:(mcrt_[a-z0-9_]+),", src))
    assert {"mcrt_simulate", "mcrt_context_create", "mcrt_last_error", "mcrt_table_upload",
            "mcrt_table_release", "mcrt_batch_create", "mcrt_batch_read"} <= called
    assert called <= set(declared_functions()), called - set(declared_functions())
