"""Reference-compatible output and backup files (host-only, no GPU).

Files are checked byte for byte against the layouts the reference drivers
write and its readers expect (2D/SERIAL/python/batch.py:41 reshapes an outdat
record to (8, ny, nx) float64; 2D/MPI/python/batch.py reads merge_outdat as
float32), and the FP32 per-rank records are run through the reference's own
merge tools (2D/MPI/common/merge.c, 3D/MPI/common/merge.c, built from source
under oracle/_ref when the reference tree is present)."""
import os
import pathlib
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from mlau_amd import io
from mlau_amd import problems as P
from oracle import oracle as O

MERGE2D = os.path.join(O.REF_DIR, "merge2d")
MERGE3D = os.path.join(O.REF_DIR, "merge3d")


@pytest.fixture
def short_dir():
    """The merge tools keep paths in 100-byte buffers: work under a short path."""
    d = tempfile.mkdtemp(prefix="mio", dir="/tmp")
    yield pathlib.Path(d)
    shutil.rmtree(d)


def text(d, name):
    with open(os.path.join(d, name)) as f:
        return f.read()


def test_2d_serial_fp64_records(tmp_path):
    f, p = P.vortex_2d(16, 4)
    g = io.grid_of(f, gamma=p["gamma"])
    x = (np.arange(g.nx) - 4) * p["dx"]  # 2D/SERIAL/OTvortex/init.hpp:13-14
    io.write_grid(str(tmp_path), g, x, x)
    io.write_output(str(tmp_path), g, True, 0, 0.0, f)
    f2 = f * 1.5
    io.write_output(str(tmp_path), g, False, 1, 0.0125, f2)
    assert text(tmp_path, "params.dat") == "%f\n" % p["gamma"]
    assert text(tmp_path, "offsets.dat") == "4 4\n"
    assert text(tmp_path, "t.dat") == "%f\n%f\n" % (0.0, 0.0125)
    assert text(tmp_path, "x.dat") == "".join("%.12f\n" % v for v in x)
    raw = (tmp_path / "outdat_00001.dat").read_bytes()
    assert raw == np.ascontiguousarray(f2).tobytes()
    # the reference reader's view of the record
    data = np.fromfile(str(tmp_path / "outdat_00000.dat"), dtype=np.float64).reshape((8, g.ny, g.nx))
    assert np.array_equal(data, f)
    assert np.array_equal(io.read_output(str(tmp_path), g, 1), f2)


def test_1d_records(tmp_path):
    f, p = P.brio_wu(64)
    g = io.grid_of(f, gamma=p["gamma"], bx=p["bx"])
    x = (np.arange(g.nx) - 4 + 0.5) * p["dx"] - 0.5
    io.write_grid(str(tmp_path), g, x)
    io.write_output(str(tmp_path), g, True, 0, 0.0, f)
    io.write_output(str(tmp_path), g, False, 1, 0.01, f)
    assert text(tmp_path, "params.dat") == "%.12f %.12f\n" % (p["bx"], p["gamma"])
    assert text(tmp_path, "xoff.dat") == "4\n"
    assert text(tmp_path, "t.dat") == "%.12f\n%.12f\n" % (0.0, 0.01)
    data = np.fromfile(str(tmp_path / "outdat_00001.dat"), dtype=np.float64).reshape((7, g.nx))
    assert np.array_equal(data, f)


def _mpi_fp32(tmp_path, f, gamma):
    g = io.grid_of(f, gamma=gamma)
    coords = [(np.arange(n) - 4) * 0.1 for n in (g.nx, g.ny, g.nz)[:g.dim]]
    io.write_grid(str(tmp_path), g, *coords, rank=0)
    io.write_output(str(tmp_path), g, True, 0, 0.0, f, fmt=io.F32, rank=0)
    io.write_output(str(tmp_path), g, False, 1, 0.5, 2 * f, fmt=io.F32, rank=0)
    rec = np.fromfile(str(tmp_path / "outdat_00001_00000.dat"), dtype=np.float32).reshape(f.shape)
    assert np.array_equal(rec, (2 * f).astype(np.float32))
    back = io.read_output(str(tmp_path), g, 0, fmt=io.F32, rank=0)
    assert np.array_equal(back, f.astype(np.float32).astype(np.float64))
    return g, coords


def test_2d_mpi_fp32_records_merge_in_reference_tool(short_dir):
    tmp_path = short_dir
    f, p = P.vortex_2d(12, 4)
    src = tmp_path / "run"
    src.mkdir()
    g, coords = _mpi_fp32(src, f, p["gamma"])
    assert text(src, "offsets.dat") == "4 4\n"
    if not os.path.exists(MERGE2D):
        pytest.skip("reference merge tool not built (no reference tree)")
    out = tmp_path / "merged"
    out.mkdir()
    subprocess.run([MERGE2D, str(src), str(out)], check=True, capture_output=True)
    inner = (slice(None), slice(4, -4), slice(4, -4))
    for cnt, mult in ((0, 1.0), (1, 2.0)):
        m = np.fromfile(str(out / ("merge_outdat_%05d.dat" % cnt)), dtype=np.float32).reshape((8, g.ny - 8, g.nx - 8))
        assert np.array_equal(m, (mult * f)[inner].astype(np.float32))
    assert np.allclose(np.loadtxt(str(out / "merge_x.dat")), coords[0][4:-4], atol=1e-6)
    assert text(out, "t.dat") == "%f\n%f\n" % (0.0, 0.5)


def test_3d_mpi_fp32_records_merge_in_reference_tool(short_dir):
    tmp_path = short_dir
    f, p = P.ot_vortex_3d(8)
    src = tmp_path / "run"
    src.mkdir()
    g, _ = _mpi_fp32(src, f, p["gamma"])
    assert text(src, "offsets.dat") == "4 4 4\n"
    if not os.path.exists(MERGE3D):
        pytest.skip("reference merge tool not built (no reference tree)")
    out = tmp_path / "merged"
    out.mkdir()
    subprocess.run([MERGE3D, str(src), str(out)], check=True, capture_output=True)
    m = np.fromfile(str(out / "merge_outdat_00001.dat"), dtype=np.float32).reshape((8, 8, 8, 8))
    assert np.array_equal(m, (2 * f)[:, 4:-4, 4:-4, 4:-4].astype(np.float32))


def test_backup_order_and_round_trip(tmp_path):
    f, p = P.vortex_2d(16, 4)
    io.bkup_save(str(tmp_path), f, 120, 3, 0.25, 1.5e-3, 0.3)
    assert text(tmp_path, "bkup_stamp.dat") == "120 3\n"
    assert np.array_equal(np.fromfile(str(tmp_path / "bkup_time.dat")), [0.25, 1.5e-3, 0.3])
    raw = np.fromfile(str(tmp_path / "bkup_data_00000.dat")).reshape(8, -1)
    order = [0, 1, 2, 3, 5, 6, 7, 4]  # ro,mx,my,mz,bx,by,bz,en (2D/MPI/OTvortex/bkup.hpp:68-75)
    for k in range(8):
        assert np.array_equal(raw[k], f[order[k]].ravel())
    g, st = io.bkup_load(str(tmp_path), f.shape)
    assert np.array_equal(g, f)
    assert st == dict(n=120, cnt=3, tim=0.25, dt=1.5e-3, trec=0.3)
    assert io.bkup_load(str(tmp_path / "nothing"), f.shape) is None
