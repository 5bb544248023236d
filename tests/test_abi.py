"""The C ABI as a C compiler sees it: programs compiled against include/dgh.h (tests/abi/) check the layout of
every descriptor and statistics struct against the Python binding's ctypes mirrors, call the host-only entry
point from C, and (GPU) run one right-hand side from a C caller, bit-identical to the binding's result."""
import ctypes as C
import json
import os
import subprocess

import numpy as np
import pytest

import hopefoam_amd.dgh as dgh

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
INC = os.path.join(ROOT, "include")
LIBDIR = os.path.dirname(dgh.lib_path())

MIRRORS = {"dgh_rhs_desc": dgh.RhsDesc, "dgh_lim_desc": dgh.LimDesc, "dgh_mlp_train_desc": dgh.MlpTrainDesc,
           "dgh_sipg_desc": dgh.SipgDesc, "dgh_sipg_stats": dgh.SipgStats, "dgh_ens_case": dgh.EnsCase,
           "dgh_ens_desc": dgh.EnsDesc, "dgh_bcfn_desc": dgh.BcFnDesc}


def _compile(src, out, link):
    cmd = ["gcc", "-std=c99", "-Wall", "-Werror", "-I" + INC, os.path.join(HERE, "abi", src), "-o", out]
    if link:
        cmd += ["-DWITH_LIBRARY", "-L" + LIBDIR, "-ldghip", "-Wl,-rpath," + LIBDIR, "-lm"]
    subprocess.run(cmd, check=True)
    return out


def test_struct_layouts_match_binding(tmp_path):
    exe = _compile("abi_layout.c", str(tmp_path / "abi_layout"), link=False)
    info = json.loads(subprocess.run([exe], check=True, capture_output=True, text=True).stdout)
    assert set(info["structs"]) == set(MIRRORS)
    for name, cls in MIRRORS.items():
        s = info["structs"][name]
        assert s["size"] == C.sizeof(cls), (name, s["size"], C.sizeof(cls))
        fields = {f: (getattr(cls, f).offset, getattr(cls, f).size) for f, _ in cls._fields_}
        assert fields == {k: tuple(v) for k, v in s["fields"].items()}, name
    for k, v in info["enums"].items():
        assert getattr(dgh, k) == v, k


def test_host_entry_point_from_c(tmp_path):
    """dgh_refelem_get called from C (no GPU needed): the order-3 mass matrix; its trace equals the binding's."""
    exe = _compile("abi_layout.c", str(tmp_path / "abi_layout_lib"), link=True)
    info = json.loads(subprocess.run([exe], check=True, capture_output=True, text=True).stdout)
    M = dgh.refelem_get(3, "M").reshape(10, 10)
    assert info["refelem_M3_count"] == 100
    assert info["refelem_M3_trace"] == float(np.trace(M))


@pytest.mark.gpu
def test_rhs_from_c_caller(tmp_path):
    """One Euler/Roe right-hand side from a C program linked against libdghip.so equals the same call
    through the Python binding, to round-off (the C program builds its inputs with libm, the test with numpy)."""
    src = os.path.join(HERE, "abi", "abi_rhs.c")
    exe = str(tmp_path / "abi_rhs")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I" + INC, src, "-o", exe, "-L" + LIBDIR, "-ldghip",
                    "-Wl,-rpath," + LIBDIR, "-lm"], check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True, timeout=120).stdout.split()
    K, Np = int(out[0]), int(out[1])
    got = np.array([float(v) for v in out[2:]]).reshape(4, K, Np)
    with dgh.Context(0) as ctx:
        m = ctx.mesh_structured(3, 5, 4, 0.0, 10.0, 0.0, 10.0, periodic=(True, True))
        x, y = m.nodes()
        rho = 1.0 + 0.2 * np.sin(0.6 * x) * np.cos(0.4 * y)
        u, v = 0.5 + 0.1 * np.cos(0.3 * y), -0.2 + 0.1 * np.sin(0.5 * x)
        p = 1.0 + 0.1 * rho
        q = np.stack([rho, rho * u, rho * v, p / 0.4 + 0.5 * rho * (u * u + v * v)])
        fq, fr = m.field(4, q), m.field(4)
        m.rhs(dgh.RhsDesc(dgh.DGH_EQ_EULER, dgh.DGH_FLUX_ROE, dgh.DGH_PREC_FP64, 0, 0.0, 0.0, 1.4), fq, fr)
        want = fr.download()
    # the C program's inputs pass through sin/cos of libm, the binding's through numpy: compare to round-off
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
