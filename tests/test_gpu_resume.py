"""Checkpoint and resume on the GPU: a run stopped after k steps, written to
the reference's backup files and resumed in a fresh solver, ends bit-identical
(state, time, step count) to the run that never stopped (the reference's
bkup_save/bkup_load cycle, 2D/MPI/OTvortex/main.cpp:39,71)."""
import numpy as np
import pytest

import mlau_amd as M
from mlau_amd import problems as P

pytestmark = pytest.mark.gpu


def run(f0, p, sch, steps, tmp=None, stop=None):
    s = M.Solver(f0, sch, p["dx"], p.get("dy"), p.get("dz"), gamma=p["gamma"])
    s.cfl(0.4, p["vfloor"])
    if stop is None:
        s.step_cfl(steps, 0.4, p["vfloor"])
    else:
        s.step_cfl(stop, 0.4, p["vfloor"])
        s.checkpoint(str(tmp), cnt=1, trec=0.0)
        s.close()
        s = M.Solver(np.zeros_like(f0), sch, p["dx"], p.get("dy"), p.get("dz"), gamma=p["gamma"])
        st = s.resume(str(tmp))
        assert st["n"] == stop
        s.step_cfl(steps - stop, 0.4, p["vfloor"])
    out, t = s.download(), s.time()
    s.close()
    return out, t


@pytest.mark.parametrize("case", ["2d", "3d"])
def test_resume_is_bit_identical(gpu, tmp_path, case):
    if case == "2d":
        f0, p = P.vortex_2d(64, 4)
        sch = M.scheme(4, 3, "lhlld", 1)
    else:
        f0, p = P.ot_vortex_3d(16)
        sch = M.scheme(2, 3, "hlld", 1)
    a, ta = run(f0, p, sch, 24)
    b, tb = run(f0, p, sch, 24, tmp_path, stop=9)
    assert ta == tb
    assert np.array_equal(a, b)
