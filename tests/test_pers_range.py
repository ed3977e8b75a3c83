"""Block-to-work maps of the kernels, compiled for the host from their own
source text with g++.  The persistent ensemble kernels' tile split (pers_range in
mlau_amd/csrc/mlau_ens.h) hands every tile of a list to exactly one block of the
grid, for grids that are and are not multiples of the XCD count and for grids
smaller than it; the cell kernels' XCD-banded block order (xcd_band in
mlau_amd/csrc/mlau_common.h) is a permutation for every grid size."""
import os
import re
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "mlau_amd", "csrc", "mlau_ens.h")
COMMON = os.path.join(ROOT, "mlau_amd", "csrc", "mlau_common.h")

DRIVER = r"""
#include <cstdio>
#include <vector>
struct Dim { unsigned x; };
static Dim blockIdx, gridDim;
#define __device__
#define __forceinline__ inline
constexpr unsigned kXcds = %(xcds)s;
%(src)s
int main()
{
  const unsigned grids[] = {1, 2, 3, 7, 8, 9, 15, 16, 17, 31, 64, 120, 180, 236, 240, 248, 255, 256, 257, 1000};
  const int totals[] = {0, 1, 2, 5, 7, 8, 9, 63, 100, 1023, 4096};
  for (unsigned g : grids)
    for (int total : totals) {
      std::vector<int> hits(total, 0);
      gridDim.x = g;
      for (unsigned b = 0; b < g; b++) {
        blockIdx.x = b;
        long guard = 0;
        for (PRange r = pers_range(total); r.t < r.end; r.t += r.stride) {
          if (r.t < 0 || r.t >= total || ++guard > total) { printf("bad g=%%u total=%%d block=%%u t=%%d\n", g, total, b, r.t); return 1; }
          hits[r.t]++;
        }
      }
      for (int t = 0; t < total; t++)
        if (hits[t] != 1) { printf("tile %%d of %%d hit %%d times with %%u blocks\n", t, total, hits[t], g); return 1; }
    }
  printf("ok\n");
  return 0;
}
"""


def test_every_tile_to_exactly_one_block(tmp_path):
    text = open(HDR).read()
    a = text.index("struct PRange {")
    m = re.search(r"__device__ __forceinline__ PRange pers_range\(int total\)\n\{.*?\n\}", text, flags=re.S)
    assert m and m.start() > a
    src = text[a:m.end()]
    xcds = re.search(r"constexpr unsigned kXcds = (\d+);", open(COMMON).read()).group(1)
    c = tmp_path / "pr.cpp"
    c.write_text(DRIVER % {"xcds": xcds, "src": src})
    exe = tmp_path / "pr"
    subprocess.run(["g++", "-O1", "-std=c++17", str(c), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stdout


BAND = r"""
#include <cstdio>
#include <vector>
#define __device__
#define __forceinline__ inline
constexpr unsigned kXcds = %(xcds)s;
%(src)s
int main()
{
  for (unsigned nb = 1; nb <= 3000; nb++) {
    std::vector<int> hits(nb, 0);
    for (unsigned b = 0; b < nb; b++) {
      const unsigned t = xcd_band(b, nb);
      if (t >= nb) { printf("nb=%%u block %%u -> %%u\n", nb, b, t); return 1; }
      hits[t]++;
    }
    for (unsigned t = 0; t < nb; t++)
      if (hits[t] != 1) { printf("nb=%%u: %%u hit %%d times\n", nb, t, hits[t]); return 1; }
  }
  printf("ok\n");
  return 0;
}
"""


def test_xcd_band_is_a_bijection(tmp_path):
    """The cell kernels' XCD-banded block order (xcd_band in mlau_common.h)
    permutes the blocks of any grid size."""
    text = open(COMMON).read()
    m = re.search(r"__device__ __forceinline__ unsigned xcd_band\(unsigned b, unsigned nb\)\n\{.*?\n\}", text,
                  flags=re.S)
    assert m
    xcds = re.search(r"constexpr unsigned kXcds = (\d+);", text).group(1)
    c = tmp_path / "band.cpp"
    c.write_text(BAND % {"xcds": xcds, "src": m.group(0)})
    exe = tmp_path / "band"
    subprocess.run(["g++", "-O1", "-std=c++17", str(c), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stdout
