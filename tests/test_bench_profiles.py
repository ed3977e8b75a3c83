"""bench.py's per-kernel HBM traffic and FP64 instruction counts for the
2048^2 headline are copies of the committed rocprofv3 summaries it names;
check that they still agree with those files' tables."""
import importlib.util
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def table(path, col):
    """{bench name: value of column col} from a markdown table whose second
    column is the bench name ('update2d (first stage)' kept as its own key)."""
    out = {}
    for line in open(os.path.join(ROOT, path)):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) > col and re.match(r"^[a-z0-9_]+( \(.*\))?$", cells[1]) and cells[1] != "bench name":
            try:
                out[cells[1]] = float(cells[col])
            except ValueError:
                pass
    return out


def test_traffic_matches_profile():
    P = bench().PROFILE_2048
    t = table(P["traffic_file"], 4)
    assert t, "no table in " + P["traffic_file"]
    for k, v in P["traffic"].items():
        if k == "update2d":
            want = (t["update2d (first stage)"] + 2 * t["update2d (other stages)"]) / 3 * 1e6
        else:
            want = t[k] * 1e6
        assert abs(v - want) <= 1e-9 * want, (k, v, want)


def test_fp64_counts_match_profile():
    P = bench().PROFILE_2048
    t = table(P["fp64_file"], 3)
    assert t, "no table in " + P["fp64_file"]
    for k, v in P["fp64"].items():
        if k == "update2d":
            want = (t["update2d (first stage)"] + 2 * t["update2d"]) / 3 * 1e6
        else:
            want = t[k] * 1e6
        assert abs(v - want) <= 1e-9 * want, (k, v, want)


def test_more_than_one_gpu_is_refused():
    """bench.py runs on one GPU: --gpus above 1 exits with a message saying
    so, before it touches a device."""
    import subprocess
    import sys
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2", "--steps", "1"],
                         capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert out.returncode == 2
    assert "one GPU only" in out.stderr and not out.stdout.strip()
