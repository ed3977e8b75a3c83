"""Initial states of the benchmark problems, in the reference layout.

Every builder returns a float64 array with ghost cells (``xoff`` per side,
x fastest) in the reference's argument order -- ``(7, nx)`` =
ro,mx,my,mz,en,by,bz for 1D and ``(8, [nz,] ny, nx)`` = ro,mx,my,mz,en,bx,by,bz
for 2D/3D -- plus a dict of run parameters.  Face-centred B sits on the faces
i-1/2 (``bx[i]``), j-1/2 and k-1/2 like the reference
(2D/SERIAL/OTvortex/mhd_fd_ct_2d.c:179-181).  Ghosts are filled by periodic
copies (2D/SERIAL/OTvortex/boundary.c:28-57) so every ghost equals its image
bit for bit.  Random draws use splitmix64 so results never depend on numpy's
generator version.
"""
import math

import numpy as np

XOFF = 4
RO, MX, MY, MZ, EN, BX, BY, BZ = range(8)


# ------------------------------------------------------------------ utils
def splitmix64(seed, n):
    """n uniform doubles in [0,1) from splitmix64 (x>>11) * 2^-53."""
    out = np.empty(n)
    s = np.uint64(seed)
    with np.errstate(over="ignore"):
        for k in range(n):
            s = np.uint64(s + np.uint64(0x9E3779B97F4A7C15))
            z = s
            z = np.uint64((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9))
            z = np.uint64((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB))
            z = z ^ (z >> np.uint64(31))
            out[k] = float(z >> np.uint64(11)) * (1.0 / 9007199254740992.0)
    return out


def splitmix64_vec(seed, n):
    """Vectorised splitmix64 stream (same values as :func:`splitmix64`)."""
    with np.errstate(over="ignore"):
        k = np.arange(1, n + 1, dtype=np.uint64)
        s = np.uint64(seed) + k * np.uint64(0x9E3779B97F4A7C15)
        z = (s ^ (s >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def periodic_fill(f, xoff=XOFF):
    """Periodic ghost copy along every spatial axis, x first (boundary.c:28-57)."""
    nd = f.ndim - 1
    for ax in range(nd - 1, -1, -1):  # last array axis is x
        a = ax + 1
        n = f.shape[a]
        idx = np.arange(n)
        src = idx.copy()
        src[:xoff] = n - 2 * xoff + idx[:xoff]
        src[n - xoff:] = idx[n - xoff:] - (n - 2 * xoff)
        f[...] = np.take(f, src, axis=a)
    return f


def _da(a0, a1, a2, a3, odr):
    """Difference of a vector potential across a cell: the solver's calc_df
    (2D/SERIAL/OTvortex/mhd_fd_ct_2d.c:74-82) so that div B vanishes under the
    scheme's own operator (2D/SERIAL/loop/init.hpp:71-81)."""
    if odr <= 2:
        return a2 - a1
    return (27.0 * (a2 - a1) - (a3 - a0)) / 24.0


def _bc_centre(f0, f1, f2, f3, odr):
    if odr <= 2:
        return 0.5 * (f1 + f2)
    return 0.0625 * (9.0 * (f1 + f2) - (f0 + f3))


# --------------------------------------------------------------- 1D problems
def riemann_1d(nx, left, right, bx, gamma, xoff=XOFF):
    """1D Riemann problem on [0,1] with the jump at x=0.5 (cells with centre
    x <= 0.5 take the left state), as 1D/SERIAL/shock/main.cpp:55-111.
    left/right = (ro, p, vx, vy, vz, by, bz)."""
    n = nx + 2 * xoff
    dx = 1.0 / nx
    x = (np.arange(n) - xoff + 0.5) * dx - 0.5
    fl = x <= 0
    f = np.zeros((7, n))
    st = [np.where(fl, left[k], right[k]) for k in range(7)]
    ro, pr, vx, vy, vz, by, bz = st
    f[0] = ro
    f[1] = ro * vx
    f[2] = ro * vy
    f[3] = ro * vz
    f[5] = by
    f[6] = bz
    f[4] = pr / (gamma - 1) + 0.5 * (ro * (vx * vx + vy * vy + vz * vz) + (bx * bx + by * by + bz * bz))
    return f, dict(dim=1, nx=n, xoff=xoff, dx=dx, bx=bx, gamma=gamma)


def brio_wu(nx=1024, xoff=XOFF):
    """Brio & Wu shock tube (1D/SERIAL/shock/main.cpp:36-39,60-68)."""
    return riemann_1d(nx, (1.0, 1.0, 0, 0, 0, 1.0, 0), (0.125, 0.1, 0, 0, 0, -1.0, 0), 0.75, 2.0, xoff)


def brio_wu_steps(dt, t_end=0.1, dtrec=0.01):
    """Step count the reference driver takes to its output at t_end:
    nrec=(int)(dtrec/dt+0.5) steps per output (1D/SERIAL/shock/main.cpp:129)."""
    nrec = int(dtrec / dt + 0.5)
    return nrec * int(round(t_end / dtrec))


# --------------------------------------------------------------- 2D problems
def _fill_vector_potential_2d(f, az, dx, dy, odr):
    """Face B from corner A_z: bx = +D_y A, by = -D_x A (loop/init.hpp:63-81).
    az has shape (ny+3, nx+3) holding corners j-1/2-1 .. ny+1/2+1."""
    ny, nx = f.shape[1:]
    # corner (i-1/2, j-1/2) of cell (i,j) is az[j+1, i+1]
    A = lambda dj, di: az[1 + dj:1 + dj + ny, 1 + di:1 + di + nx]
    f[BX] = _da(A(-1, 0), A(0, 0), A(1, 0), A(2, 0), odr) / dy
    f[BY] = -_da(A(0, -1), A(0, 0), A(0, 1), A(0, 2), odr) / dx


def vortex_2d(n, odr=4, xoff=XOFF):
    """Periodic Orszag-Tang-type vortex on [0,1]^2: rho=25/(36 pi), p=5/(12 pi),
    v=(-sin 2 pi y, sin 2 pi x, 0), B=(-sin 2 pi y, sin 4 pi x, 0)/sqrt(4 pi)
    from A_z=[cos(2 pi y)/(2 pi)+cos(4 pi x)/(4 pi)]/sqrt(4 pi) on cell corners."""
    gamma = 5.0 / 3.0
    N = n + 2 * xoff
    dx = dy = 1.0 / n
    xc = (np.arange(N) - xoff + 0.5) * dx
    yc = (np.arange(N) - xoff + 0.5) * dy
    X, Y = np.meshgrid(xc, yc)
    f = np.zeros((8, N, N))
    ro0 = 25.0 / (36.0 * math.pi)
    p0 = 5.0 / (12.0 * math.pi)
    s4 = math.sqrt(4.0 * math.pi)
    vx = -np.sin(2 * math.pi * Y)
    vy = np.sin(2 * math.pi * X)
    xf = (np.arange(-1, N + 2) - xoff) * dx
    yf = (np.arange(-1, N + 2) - xoff) * dy
    XF, YF = np.meshgrid(xf, yf)
    az = (np.cos(2 * math.pi * YF) / (2 * math.pi) + np.cos(4 * math.pi * XF) / (4 * math.pi)) / s4
    _fill_vector_potential_2d(f, az, dx, dy, odr)
    f[RO] = ro0
    f[MX] = ro0 * vx
    f[MY] = ro0 * vy
    periodic_fill(f, xoff)
    bxc = _bc_centre(np.roll(f[BX], 1, 1), f[BX], np.roll(f[BX], -1, 1), np.roll(f[BX], -2, 1), odr)
    byc = _bc_centre(np.roll(f[BY], 1, 0), f[BY], np.roll(f[BY], -1, 0), np.roll(f[BY], -2, 0), odr)
    f[EN] = p0 / (gamma - 1) + 0.5 * (ro0 * (vx * vx + vy * vy) + (bxc * bxc + byc * byc))
    periodic_fill(f, xoff)
    return f, dict(dim=2, nx=N, ny=N, xoff=xoff, dx=dx, dy=dy, gamma=gamma, vfloor=1.0)


def blast_2d(n, pratio, bmag, bang, xoff=XOFF, gamma=5.0 / 3.0, rho=1.0, p_amb=0.1, r0=0.1):
    """Periodic MHD blast on [0,1]^2: disk r<r0 at p_amb*pratio, uniform B of
    magnitude bmag at angle bang (radians) (2D/SERIAL/blast/init.hpp)."""
    N = n + 2 * xoff
    dx = dy = 1.0 / n
    xc = (np.arange(N) - xoff + 0.5) * dx - 0.5
    X, Y = np.meshgrid(xc, xc)
    r = np.sqrt(X * X + Y * Y)
    pr = np.where(r <= r0, p_amb * pratio, p_amb)
    f = np.zeros((8, N, N))
    f[RO] = rho
    f[BX] = bmag * math.cos(bang)
    f[BY] = bmag * math.sin(bang)
    f[EN] = pr / (gamma - 1) + 0.5 * (f[BX] * f[BX] + f[BY] * f[BY])
    periodic_fill(f, xoff)
    return f, dict(dim=2, nx=N, ny=N, xoff=xoff, dx=dx, dy=dy, gamma=gamma, vfloor=1.0)


# --------------------------------------------------------------- 3D problems
def shear_layer_3d(n, mach=0.01, beta=100.0, seed=2, xoff=XOFF, gamma=5.0 / 3.0):
    """Low-Mach double shear layer on [0,1]^3: rho=1, p=1/(gamma M^2),
    vx=M(tanh((y-.25)/.05)-tanh((y-.75)/.05)-1), vy,vz=1e-3 M U[-1,1]
    (splitmix64 stream, seed), uniform Bx=sqrt(2p/beta)."""
    N = n + 2 * xoff
    dx = 1.0 / n
    c = (np.arange(N) - xoff + 0.5) * dx
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    p0 = 1.0 / (gamma * mach * mach)
    f = np.zeros((8, N, N, N))
    vx = mach * (np.tanh((Y - 0.25) / 0.05) - np.tanh((Y - 0.75) / 0.05) - 1.0)
    u = splitmix64_vec(seed, 2 * N ** 3).reshape(2, N, N, N)
    vy = 1e-3 * mach * (2.0 * u[0] - 1.0)
    vz = 1e-3 * mach * (2.0 * u[1] - 1.0)
    b0 = math.sqrt(2.0 * p0 / beta)
    f[RO] = 1.0
    f[MX] = vx
    f[MY] = vy
    f[MZ] = vz
    f[BX] = b0
    f[EN] = p0 / (gamma - 1) + 0.5 * (vx * vx + vy * vy + vz * vz + b0 * b0)
    periodic_fill(f, xoff)
    return f, dict(dim=3, nx=N, ny=N, nz=N, xoff=xoff, dx=dx, dy=dx, dz=dx, gamma=gamma, vfloor=0.0)


def ot_vortex_3d(n, xoff=XOFF):
    """3D Orszag-Tang vortex on [0,1]^3 (2D vortex extended along z with a
    vz = 0.1 sin 2 pi z perturbation), face B from the same A_z."""
    f2, p2 = vortex_2d(n, 4, xoff)
    N = n + 2 * xoff
    f = np.repeat(f2[:, None], N, axis=1)
    c = (np.arange(N) - xoff + 0.5) / n
    vz = 0.1 * np.sin(2 * math.pi * c)[:, None, None]
    ro = f[RO]
    f[MZ] = ro * vz
    f[EN] = f[EN] + 0.5 * ro * vz * vz
    periodic_fill(f, xoff)
    return f, dict(dim=3, nx=N, ny=N, nz=N, xoff=xoff, dx=1.0 / n, dy=1.0 / n, dz=1.0 / n, gamma=5.0 / 3.0,
                   vfloor=1.0)


# --------------------------------------------------------- batched ensemble
def blast_ensemble_draws(count=512, seed=3, nmin=64, nmax=512):
    """Parameters of a batched 2D blast ensemble, five splitmix64 draws per
    member (u0..u4):
      N      = 16 * floor(nmin * (nmax/nmin)**u0 / 16 + 1/2)  (log-uniform, multiple of 16)
      pratio = 10**(1 + 2*u1)                                  (log-uniform in [10, 1000))
      bmag   = 0.1 + 0.9*u2,  bang = 2*pi*u3
      flux   = "mlau" if u4 < 1/2 else "hlld"
    Returns a list of dicts (n, pratio, bmag, bang, flux)."""
    u = splitmix64_vec(seed, 5 * count).reshape(count, 5)
    out = []
    for k in range(count):
        n = 16 * int(math.floor(nmin * (nmax / nmin) ** u[k, 0] / 16.0 + 0.5))
        out.append(dict(n=n, pratio=10.0 ** (1.0 + 2.0 * u[k, 1]), bmag=0.1 + 0.9 * u[k, 2],
                        bang=2.0 * math.pi * u[k, 3], flux="mlau" if u[k, 4] < 0.5 else "hlld"))
    return out


# ------------------------------------------------------------ serving stream
SERVE_TEND = {1: 0.1, 2: 0.02, 3: 2e-4}


def serve_request_draws(count, seed=4):
    """Request parameters of a mixed serving stream, six splitmix64 draws per
    request (u0..u5); kind by u0 in thirds:
      1D Riemann: nx = floor(256 * 32**u1 + 1/2) in [256, 8192]; left/right rho
                  and p are the Brio-Wu states scaled by 0.8 + 0.4*u2..u5
                  (bx = 0.75, by = +-1, gamma = 2); MLAU; end time 0.1
      2D blast:   N = 16*floor(64 * 16**u1 / 16 + 1/2) in [64, 1024]; pressure
                  ratio 10**(1 + 2*u2), |B| = 0.1 + 0.9*u3, angle 2*pi*u4,
                  MLAU if u5 < 1/2 else HLLD; end time 0.02
      3D shear:   N = 16*floor(32 * 6**u1 / 16 + 1/2) in [32, 192]; Mach
                  10**(-3 + 2*u2); MLAU; end time 2e-4
    Returns a list of dicts with 'kind' (1, 2, 3) and the parameters."""
    u = splitmix64_vec(seed, 6 * count).reshape(count, 6)
    out = []
    for k in range(count):
        kind = 1 + min(int(u[k, 0] * 3.0), 2)
        if kind == 1:
            s = 0.8 + 0.4 * u[k, 2:6]
            out.append(dict(kind=1, nx=int(math.floor(256 * 32.0 ** u[k, 1] + 0.5)), rho_l=s[0], p_l=s[1],
                            rho_r=0.125 * s[2], p_r=0.1 * s[3], flux="mlau"))
        elif kind == 2:
            out.append(dict(kind=2, n=16 * int(math.floor(64 * 16.0 ** u[k, 1] / 16.0 + 0.5)),
                            pratio=10.0 ** (1.0 + 2.0 * u[k, 2]), bmag=0.1 + 0.9 * u[k, 3],
                            bang=2.0 * math.pi * u[k, 4], flux="mlau" if u[k, 5] < 0.5 else "hlld"))
        else:
            out.append(dict(kind=3, n=16 * int(math.floor(32 * 6.0 ** u[k, 1] / 16.0 + 0.5)),
                            mach=10.0 ** (-3.0 + 2.0 * u[k, 2]), flux="mlau"))
    return out


def serve_request(d):
    """Initial fields and grid parameters of one drawn request (see
    serve_request_draws).  Returns (fields, params, scheme dict)."""
    if d["kind"] == 1:
        f, p = riemann_1d(d["nx"], (d["rho_l"], d["p_l"], 0, 0, 0, 1.0, 0), (d["rho_r"], d["p_r"], 0, 0, 0, -1.0, 0),
                          0.75, 2.0)
        p["vfloor"] = 1.0
        return f, p, dict(odr=2, rk=3, flux=d["flux"], ctw=1, bc1d="open")
    if d["kind"] == 2:
        f, p = blast_2d(d["n"], d["pratio"], d["bmag"], d["bang"])
        return f, p, dict(odr=2, rk=3, flux=d["flux"], ctw=1)
    f, p = shear_layer_3d(d["n"], mach=d["mach"])
    return f, p, dict(odr=4, rk=3, flux=d["flux"], ctw=1)
