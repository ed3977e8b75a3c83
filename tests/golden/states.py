"""Seeded test inputs that are rebuilt from a seed instead of being stored.

A fixture that needs a large input (the 3D CFL check runs on the 58^3 grid
fixed by the reference driver's global.hpp) stores the seed, a SHA-256 of the
input bytes and the reference's outputs; the input itself is rebuilt here
from splitmix64 draws (no libm call, so every machine rebuilds the same
bytes).  Used by gen_golden.py and by the tests.
"""
import hashlib

import numpy as np

from mlau_amd.problems import splitmix64_vec


def sha256(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def cfl3d_state(seed, n=58, gamma=5.0 / 3.0, nan_cell=False):
    """Random 3D state (8, n, n, n) in argument order ro,mx,my,mz,en,bx,by,bz.

    Velocities have a large z component, so a check that left vz out of the
    speed (the 2D rule sqrt(vx^2+vy^2)) gives a different maximum.  With
    nan_cell, one interior cell gets gamma*p + B^2 < 0, whose fast speed is
    NaN: the reference's "if (vtmp > vmax)" skips that cell."""
    m = n * n * n
    u = splitmix64_vec(seed, 9 * m).reshape(9, n, n, n)
    ro = 0.5 + u[0]
    vx, vy = 2.0 * u[1] - 1.0, 2.0 * u[2] - 1.0
    vz = 6.0 * u[3] - 3.0
    bx, by, bz = 2.0 * u[4] - 1.0, 2.0 * u[5] - 1.0, 2.0 * u[6] - 1.0
    pr = 0.1 + u[7]
    en = pr / (gamma - 1.0) + 0.5 * ro * (vx * vx + vy * vy + vz * vz) + 0.5 * (bx * bx + by * by + bz * bz)
    f = np.stack([ro, ro * vx, ro * vy, ro * vz, en, bx, by, bz])
    if nan_cell:
        c = (n // 2, n // 3, n // 4)
        f[4][c] = -50.0
    return np.ascontiguousarray(f)
