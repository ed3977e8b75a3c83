"""The C-ABI library loads without a GPU and exports every entry point the
public headers declare (no compute calls here)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "mlau_amd", "_lib", "libmlau_hip.so")
DROPIN = os.path.join(ROOT, "mlau_amd", "_lib", "libmlau_dropin.so")


def declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b((?:mlau|mhd)_\w+)\s*\(", txt)))


def exported(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    return {l.split()[-1] for l in out.splitlines() if l.strip()}


def test_library_loads():
    assert os.path.exists(LIB), "build first (__graft_entry__.build())"
    ctypes.CDLL(LIB)


def test_exports_every_declared_symbol():
    names = declared("mlau.h")
    assert "mlau_create" in names and "mlau_fd_ct_2d" in names
    syms = exported(LIB)
    missing = [n for n in names if n not in syms]
    assert not missing, missing


@pytest.mark.parametrize("header,must", [("mlau_ensemble.h", ("mlau_ens_add", "mlau_ens_run")),
                                         ("mlau_server.h", ("mlau_srv_submit", "mlau_srv_wait"))])
def test_engine_headers_exported(header, must):
    names = declared(header)
    assert all(n in names for n in must)
    missing = [n for n in names if n not in exported(LIB)]
    assert not missing, missing


def test_struct_layouts_match_ctypes(tmp_path):
    """The ctypes mirrors have the C structs' sizes and field offsets."""
    import mlau_amd as M
    from mlau_amd import ensemble as E
    src = tmp_path / "sz.c"
    src.write_text('#include <stddef.h>\n#include <stdio.h>\n#include "mlau_ensemble.h"\n'
                   'int main(void){printf("%zu %zu %zu %zu %zu %zu\\n", sizeof(mlau_grid), sizeof(mlau_scheme),'
                   ' sizeof(mlau_member), offsetof(mlau_member, tend), offsetof(mlau_member, max_steps),'
                   ' offsetof(mlau_scheme, precision)); return 0;}\n')
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = [ctypes.sizeof(M.Grid), ctypes.sizeof(M.Scheme), ctypes.sizeof(E.Member), E.Member.tend.offset,
            E.Member.max_steps.offset, M.Scheme.precision.offset]
    assert got == want


def test_dropin_exports_reference_names():
    """The shims carry the reference's C++ names and signatures
    (1D/SERIAL/shock/mhd_fd4c_1d.c:60-63, 2D/SERIAL/OTvortex/mhd_fd_ct_2d.c:91-94,
    3D/MPI/OTvortex/mhd_fd_ct_3d.c:99-103)."""
    assert os.path.exists(DROPIN)
    syms = exported(DROPIN)
    for mangled in ("_Z11mhd_fd4c_1dPdS_S_S_S_S_S_dddiid", "_Z12mhd_fd_ct_2dPdS_S_S_S_S_S_S_dddiiiid",
                    "_Z12mhd_fd_ct_3dPdS_S_S_S_S_S_S_ddddiiiiiidiiii"):
        assert mangled in syms, mangled


def test_python_layer_fails_loudly_without_library(tmp_path, monkeypatch):
    import mlau_amd
    monkeypatch.setattr(mlau_amd, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(mlau_amd, "_lib", None)
    with pytest.raises(mlau_amd.MLAUError):
        mlau_amd.lib()


def _grid_scheme(dim):
    import mlau_amd as M
    g = M.Grid()
    g.dim, g.nx, g.ny, g.nz = dim, 24, (24 if dim >= 2 else 1), (24 if dim == 3 else 1)
    g.xoff, g.yoff, g.zoff = 4, (4 if dim >= 2 else 0), (4 if dim == 3 else 0)
    g.dx = g.dy = g.dz = 0.1
    g.gamma = 5.0 / 3.0
    return g, M.scheme(4, 3, "lhlld", 1)


@pytest.mark.parametrize("dim,field,value,words", [
    (4, "dim", 4, "dim"), (2, "odr", 5, "odr"), (2, "rk", 0, "rk"), (2, "flux", 4, "flux"),
    (2, "precision", 3, "precision"), (2, "gamma", 1.0, "gamma"), (2, "xoff", 3, "xoff"),
    (2, "nx", 11, "nx"), (1, "dx", 0.0, "spacing"), (2, "dy", -1.0, "spacing"),
    (3, "dz", float("nan"), "spacing"), (1, "bc1d", 2, "bc1d"), (2, "ctw", 2, "ctw"),
    (2, "dnxs", (0, 1, 0, 0, 0, 0, 0, 0), "periodic"), (3, "dnzs", (2,) * 8, "flags")])
def test_create_refuses_bad_arguments(dim, field, value, words):
    """mlau_create checks the grid and scheme before it touches the device, so
    every refusal is an argument error with a message naming the field."""
    import mlau_amd as M
    g, s = _grid_scheme(dim if dim <= 3 else 2)
    if field == "dim":
        g.dim = value
    elif field in ("dnxs", "dnys", "dnzs"):
        getattr(s, field)[:] = value
    elif hasattr(g, field) and field not in ("odr", "rk", "flux", "ctw", "bc1d", "precision"):
        setattr(g, field, value)
    else:
        setattr(s, field, value)
    h = ctypes.c_void_p()
    rc = M.lib().mlau_create(ctypes.byref(g), ctypes.byref(s), ctypes.byref(h))
    assert rc == 1 and not h.value  # MLAU_ERR_ARG, no handle
    assert words in M.lib().mlau_last_error().decode()
