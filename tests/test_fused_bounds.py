"""Index check of the fused 2D ensemble stage kernel (k_ens_stage2d in
mlau_amd/csrc/mlau_ens2d.hip), on the CPU.

The kernel forms primitives, x-face and y-face fluxes and corner E-fields of
a kFT x kFT tile and its halo in LDS, then updates the tile and writes the
periodic images into the ghost frame.  This test replays the kernel's index
arithmetic for every thread of a 512-thread block, phase by phase, for every
partial tile size (tw, th in 1..kFT) and for whole grids with ghost widths
4..6, and checks that

  * every LDS address a thread reads or writes lies inside its array;
  * every LDS value a live lane uses was written by an earlier phase (the
    values that the shuffle brings from lane - 1 included);
  * every global address read or written lies inside the nx * ny member;
  * the grid's cells are written exactly once per stage: each interior cell
    by its tile, each ghost cell as the image of exactly one interior cell,
    and every ghost cell the next stage's halo reads is among them.

The tile constants are read from the kernel source, so the check follows any
change of the tile shape."""
import os
import re

import numpy as np
import pytest

SRC = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "mlau_amd", "csrc", "mlau_ens2d.hip")


def tile_constants():
    text = open(SRC).read()
    env = {}
    for name, expr in re.findall(r"constexpr int (kF\w+) = ([^;]+);", text):
        env[name] = eval(expr.replace("/", "//"), {}, dict(env))  # positive integers: C division is floor
    return env


C = tile_constants()
FT, NT, GL = C["kFT"], C["kFThreads"], C["kFGL"]
GW, NW, NCELLS = C["kFGW"], C["kFWaves"], C["kFCells"]
PS, PW, XS, XR, YS = C["kFPS"], C["kFPW"], C["kFXS"], C["kFXR"], C["kFYS"]
NP = 8  # primitive arrays (FP_N)


class Lds:
    """Written-flags of the kernel's LDS arrays (FusedLds)."""

    def __init__(self):
        self.p = np.zeros(PW * PS, bool)   # p[*] and sup share the layout
        self.f6x = np.zeros(XS * XS, bool)  # f6 from the x pass
        self.f6y = np.zeros(XS * XS, bool)  # f6 from the y pass
        self.x2 = np.zeros(XR * XS, bool)
        self.y2 = np.zeros(XS * YS, bool)
        self.ez = np.zeros(XS * XS, bool)


def in_range(idx, n, what):
    idx = np.asarray(idx)
    assert idx.size == 0 or (idx.min() >= 0 and idx.max() < n), "%s: index %d..%d outside [0, %d)" % (
        what, idx.min(), idx.max(), n)


def read(flags, idx, what):
    in_range(idx, flags.size, what)
    idx = np.asarray(idx)
    assert np.all(flags[idx]), "%s: %d reads of unwritten LDS" % (what, int((~flags[idx]).sum()))


def write(flags, idx, what):
    in_range(idx, flags.size, what)
    flags[np.asarray(idx)] = True


def lanes():
    tid = np.arange(NT)
    return tid >> 6, tid & 63


def phase_prim(L, tw, th):
    t = np.arange(PW * PW)
    li, lj = t % PW, t // PW
    ok = (li < tw + 6) & (lj < th + 6)
    write(L.p, lj[ok] * PS + li[ok], "prim")
    return li[ok], lj[ok]


def sweep_lanes(lines, cells):
    """(line, position, live, source-lane live) of every lane of every group
    pass of a sweep over `lines` lines of `cells` cells."""
    wave, lane = lanes()
    groups = (lines + GW - 1) // GW
    out = []
    for gi0 in range(0, groups, NW):
        gi = gi0 + wave
        act = gi < groups  # waves past the last group skip the loop
        rr = lane // GL
        cc = lane - GL * rr
        line = GW * gi + rr
        live = act & (line < lines) & (cc < cells)
        line = np.minimum(line, lines - 1)
        cc = np.minimum(cc, cells - 1)
        src = np.where(lane % 64 > 0, np.roll(live, 1), False)  # shuffle from lane - 1 of the same wave
        out.append((act, line, cc, live, src))
    return out


def phase_x(L, tw, th):
    for act, row, cc, live, src in sweep_lanes(th + 4, tw + 2):
        q = (row + 1) * PS + (cc + 2)
        for m in range(5):
            in_range(q[act] - 2 + m, PW * PS, "x stencil")
            read(L.p, (q - 2 + m)[live], "x stencil")
        face = live & (cc >= 1)
        assert np.all(src[face]), "x sweep: a face takes its left state from a lane that is not live"
        for d in (-1, 0, 1):
            read(L.p, (q + d)[live], "x sup")
        for d in (-1, 0):
            for e in (-PS, 0, PS):
                read(L.p, (q + d + e)[face], "x dvy")
        fi = cc - 1
        write(L.x2, (row * XS + fi)[face], "x2")
        jr = row - 2
        inner = face & (jr >= 0) & (jr < th)
        write(L.f6x, (jr * XS + fi)[inner], "f6 x")


def tile_cells(tw, th):
    tid = np.arange(NT)
    for u in range(NCELLS):
        k = tid + NT * u
        ci, cj = k % FT, k // FT
        own = (k < FT * FT) & (ci < tw) & (cj < th)
        yield ci[own], cj[own]


def phase_xupdate(L, tw, th):
    for ci, cj in tile_cells(tw, th):
        read(L.f6x, cj * XS + ci, "x update")
        read(L.f6x, cj * XS + ci + 1, "x update")


def phase_y(L, tw, th):
    for act, col, rr, live, src in sweep_lanes(tw + 4, th + 2):
        q = (rr + 2) * PS + (col + 1)
        for m in range(5):
            in_range(q[act] + (m - 2) * PS, PW * PS, "y stencil")
            read(L.p, (q + (m - 2) * PS)[live], "y stencil")
        face = live & (rr >= 1)
        assert np.all(src[face]), "y sweep: a face takes its lower state from a lane that is not live"
        for d in (-PS, 0, PS):
            read(L.p, (q + d)[live], "y sup")
        for d in (-PS, 0):
            for e in (-1, 0, 1):
                read(L.p, (q + d + e)[face], "y dvx")
        fr = rr - 1
        write(L.y2, (fr * YS + col)[face], "y2")
        ic = col - 2
        inner = face & (ic >= 0) & (ic < tw)
        write(L.f6y, (fr * XS + ic)[inner], "f6 y")


def phase_ez(L, tw, th):
    t = np.arange(XS * XS)
    ei, ej = t % XS, t // XS
    ok = (ei <= tw) & (ej <= th)
    ei, ej = ei[ok], ej[ok]
    q = (ej + 3) * PS + (ei + 3)
    qa = q - PS
    for idx in (qa - 1, qa, q - 1, q):
        read(L.p, idx, "ez weight")
    for r in range(4):
        read(L.x2, (ej + r) * XS + ei, "ez x2")
        read(L.y2, ej * YS + ei + r, "ez y2")
    write(L.ez, ej * XS + ei, "ez")


def phase_update(L, tw, th):
    for ci, cj in tile_cells(tw, th):
        read(L.f6y, cj * XS + ci, "update fy")
        read(L.f6y, (cj + 1) * XS + ci, "update fy")
        read(L.ez, cj * XS + ci, "update ez")
        read(L.ez, (cj + 1) * XS + ci, "update ez")
        read(L.ez, cj * XS + ci + 1, "update ez")


def run_tile_lds(tw, th):
    L = Lds()
    phase_prim(L, tw, th)
    phase_x(L, tw, th)
    phase_xupdate(L, tw, th)
    phase_y(L, tw, th)
    phase_ez(L, tw, th)
    phase_update(L, tw, th)


def test_tile_constants_fit_the_lanes():
    assert FT + 2 <= GL and GW * GL == 64 and NW * 64 == NT and NCELLS * NT >= FT * FT
    # LDS of the block (FusedLds) within the CU's 160 KB
    lds = 8 * (NP * PW * PS + 6 * XS * XS + 2 * XR * XS + 2 * XS * YS + XS * XS) + PW * PS
    assert lds <= 160 * 1024, lds


@pytest.mark.parametrize("th", list(range(1, FT + 1)))
def test_lds_indices_every_partial_tile(th):
    for tw in range(1, FT + 1):
        run_tile_lds(tw, th)


def global_tile(nx, ny, xo, yo, a, c, tw, th, reads, writes):
    n = nx * ny
    # primitives: cells a-3..a+tw+2 x c-3..c+th+2, with s+1 (bx) and s+nx (by)
    li, lj = np.meshgrid(np.arange(tw + 6), np.arange(th + 6))
    s = (c - 3 + lj) * nx + (a - 3 + li)
    for idx in (s, s + 1, s + nx):
        in_range(idx, n, "prim global")
        reads[idx.ravel()] = True
    # x sweep: bx at cells a-1..a+tw (+1) on rows c-2..c+th+1, for every lane (clamped)
    for act, row, cc, live, src in sweep_lanes(th + 4, tw + 2):
        s = (c - 2 + row[act]) * nx + (a - 1 + cc[act])
        in_range(s, n, "x sweep global")
        in_range(s + 1, n, "x sweep global")
    for act, col, rr, live, src in sweep_lanes(tw + 4, th + 2):
        s = (c - 1 + rr[act]) * nx + (a - 2 + col[act])
        in_range(s, n, "y sweep global")
        in_range(s + nx, n, "y sweep global")
    # ez weight: bx at s-nx, s; by at s-1, s for corners a..a+tw x c..c+th
    ei, ej = np.meshgrid(np.arange(tw + 1), np.arange(th + 1))
    s = (c + ej) * nx + (a + ei)
    for idx in (s - nx, s, s - 1):
        in_range(idx, n, "ez global")
    # update and periodic images
    NX, NY = nx - 2 * xo, ny - 2 * yo
    for ci, cj in tile_cells(tw, th):
        i, j = a + ci, c + cj
        in_range(j * nx + i, n, "update read")
        gi = np.where(i < 2 * xo, i + NX, np.where(i >= nx - 2 * xo, i - NX, -1))
        gj = np.where(j < 2 * yo, j + NY, np.where(j >= ny - 2 * yo, j - NY, -1))
        targets = [j * nx + i, (j * nx + gi)[gi >= 0], (gj * nx + i)[gj >= 0],
                   (gj * nx + gi)[(gi >= 0) & (gj >= 0)]]
        for tg in targets:
            in_range(tg, n, "update write")
            np.add.at(writes, tg, 1)


@pytest.mark.parametrize("xo,yo", [(4, 4), (5, 4), (4, 6), (6, 5)])
@pytest.mark.parametrize("NX,NY", [(8, 8), (12, 10), (2 * FT, FT + 1), (FT + 13, 3 * FT - 5), (3 * FT + 1, 2 * FT + 27)])
def test_global_indices_and_ghost_images(xo, yo, NX, NY):
    nx, ny = NX + 2 * xo, NY + 2 * yo
    if NX < 2 * xo or NY < 2 * yo:
        pytest.skip("the fused kernel takes interiors of at least two ghost widths")
    reads = np.zeros(nx * ny, bool)
    writes = np.zeros(nx * ny, np.int64)
    ntx, nty = (NX + FT - 1) // FT, (NY + FT - 1) // FT
    for l in range(ntx * nty):
        a, c = xo + FT * (l % ntx), yo + FT * (l // ntx)
        tw, th = min(FT, nx - xo - a), min(FT, ny - yo - c)
        assert 1 <= tw <= FT and 1 <= th <= FT
        global_tile(nx, ny, xo, yo, a, c, tw, th, reads, writes)
    w = writes.reshape(ny, nx)
    # each cell written at most once; every cell the next stage reads is written
    assert w.max() == 1, "a cell is written by two sources"
    assert np.all(w.ravel()[reads] == 1), "the halo reads a ghost cell that no tile writes"
    # the interior and the whole ghost frame, corners included, are written
    assert np.all(w == 1)
