"""Initial conditions and boundary-hook parameters of the problems the
force-free step is run on (the Python counterpart of the user/*.F90 files'
userReadInput + userInitFields; the Fortran host has its own user files).

Every builder returns a ``Problem``: the field arrays in the reference's
layout (C-ordered ``(nz+4, ny+4, nx+4)`` views of Fortran arrays with two
ghost layers, FP64 on the host), the boundary flags, CC, the absorbing-layer
settings and the boundary-kernel kind and parameters.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

NG = 2
BC_NONE, BC_LOOP, BC_PSR, BC_UNIFORM = 0, 1, 2, 3
NAMES = ("ex", "ey", "ez", "bx", "by", "bz", "b0x", "b0y", "b0z")


@dataclasses.dataclass
class Problem:
    name: str
    n: tuple
    boundary: tuple
    cc: float
    fields: dict
    bc_kind: int = BC_NONE
    bc_params: tuple = ()
    absorb_buff: int = 20
    absorb_fixed: bool = False
    steps: int = 0
    g: int = 0  # 1 when B0 has a non-zero curl (algorithmic bytes count it)

    @property
    def cells(self) -> int:
        return int(np.prod(self.n))


def _zeros(n):
    nx, ny, nz = n
    return {k: np.zeros((nz + 2 * NG, ny + 2 * NG, nx + 2 * NG)) for k in NAMES}


def _interior(a):
    return a[NG:-NG, NG:-NG, NG:-NG]


def _coords(n, full=False):
    """Global index arrays (k, j, i) broadcastable over the box; with
    full=True they span the ghosts too."""
    nx, ny, nz = n
    lo = -NG if full else 0
    i = np.arange(lo, nx - lo, dtype=np.float64)[None, None, :]
    j = np.arange(lo, ny - lo, dtype=np.float64)[None, :, None]
    k = np.arange(lo, nz - lo, dtype=np.float64)[:, None, None]
    return i, j, k


# --- config: vacuum plane wave (user_emwave.F90 form) -----------------------
def plane_wave_eigen(kx: float, ky: float):
    """Discrete Yee eigenvector of the vacuum mode travelling along +k in the
    x-y plane with dB along z: (alpha, beta) multiply cos of the E phases."""
    sx, sy = math.sin(kx / 2), math.sin(ky / 2)
    s = math.hypot(sx, sy)
    return -sy / s, sx / s, 2.0 * s


def emwave(n=64, amp=1e-2, mode=(1, 1), b0=1.0, cc=0.5, axis="z", steps=200) -> Problem:
    """Plane wave over a uniform background: in the frame (a, b, c) that is
    (x, y, z) for axis='z', (y, z, x) for 'x' and (z, x, y) for 'y',
    B = b0 c^ + amp cos(k.x_Bc) c^, E = amp (alpha cos(k.x_Ea), beta cos(k.x_Eb)),
    k = 2 pi mode / n; the staggering of user_emwave.F90:58-63 (cell-edge E,
    face B) with the discrete eigenvector, so the mode is exact for the
    semi-discrete Yee operator."""
    assert np.isscalar(n)
    nn = (n, n, n)
    f = _zeros(nn)
    ka, kb = 2 * math.pi * mode[0] / n, 2 * math.pi * mode[1] / n
    al, be, _ = plane_wave_eigen(ka, kb)
    i, j, k = _coords(nn)
    # local-frame arrays indexed [c][b][a]
    a, b = i, j
    ea = amp * al * np.cos(ka * (a + 0.5) + kb * b) + 0 * k
    eb = amp * be * np.cos(ka * a + kb * (b + 0.5)) + 0 * k
    bc = b0 + amp * np.cos(ka * (a + 0.5) + kb * (b + 0.5)) + 0 * k
    perm = {"z": ((0, 1, 2), ("ex", "ey", "bz", "b0z")),
            "x": ((1, 2, 0), ("ey", "ez", "bx", "b0x")),
            "y": ((2, 0, 1), ("ez", "ex", "by", "b0y"))}[axis]
    axes, (na, nb, nc, n0) = perm
    _interior(f[na])[...] = np.transpose(ea, axes)
    _interior(f[nb])[...] = np.transpose(eb, axes)
    _interior(f[nc])[...] = np.transpose(bc, axes)
    f[n0][...] = b0
    return Problem(f"emwave{axis}", nn, (1, 1, 1), cc, f, steps=steps)


def emwave_closed_form(p: Problem, amp, mode, nsteps, b0=1.0, axis="z"):
    """Perturbation of B_c after nsteps: amp |G|^n cos(k.x - n psi) with
    G(z) = 1 + z + z^2/2 + z^3/6 at z = -i theta (SSP-RK3 on a linear
    operator), theta = 2 CC sqrt(sin^2(ka/2) + sin^2(kb/2))."""
    n = p.n[0]
    ka, kb = 2 * math.pi * mode[0] / n, 2 * math.pi * mode[1] / n
    _, _, s = plane_wave_eigen(ka, kb)
    th = p.cc * s
    z = -1j * th
    G = 1 + z + z * z / 2 + z ** 3 / 6
    i, j, k = _coords(p.n)
    ph = ka * (i + 0.5) + kb * (j + 0.5) + 0 * k
    val = amp * (G ** nsteps * np.exp(1j * ph)).real
    axes = {"z": (0, 1, 2), "x": (1, 2, 0), "y": (2, 0, 1)}[axis]
    return np.transpose(val, axes)


# --- config: uniform B0 plus seeded Alfvenic perturbation -----------------
def _band_limited(n, rng, kmin=1, kmax=8):
    k = np.fft.fftfreq(n, 1.0 / n)
    kz, ky, kx = np.meshgrid(k, k, k, indexing="ij")
    kk = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
    spec = (rng.standard_normal((n, n, n)) + 1j * rng.standard_normal((n, n, n)))
    spec[(kk < kmin) | (kk > kmax)] = 0.0
    return np.fft.ifftn(spec).real


def alfven(n=256, seed=1, amp=0.05, kmax=8, cc=0.45, steps=500) -> Problem:
    """B = z^ + curl_h A with A band-limited (1 <= |k| <= kmax), so div_h B = 0
    at t=0; E = grad_perp psi = -(v x z^) for the solenoidal drift
    v = (d_y psi, -d_x psi, 0).  rms|dB| = rms|E| = amp."""
    nn = (n, n, n)
    rng = np.random.default_rng(seed)
    f = _zeros(nn)
    ax, ay, az = (_band_limited(n, rng, 1, kmax) for _ in range(3))
    # arrays are [k][j][i]; roll(-1, axis) gives the +1 neighbour
    sh = lambda a, ax_: np.roll(a, -1, axis=ax_)
    dbx = (sh(az, 1) - az) - (sh(ay, 0) - ay)
    dby = (sh(ax, 0) - ax) - (sh(az, 2) - az)
    dbz = (sh(ay, 2) - ay) - (sh(ax, 1) - ax)
    s = amp / math.sqrt(np.mean(dbx ** 2 + dby ** 2 + dbz ** 2))
    psi = _band_limited(n, rng, 1, kmax)
    ex = sh(psi, 2) - psi
    ey = sh(psi, 1) - psi
    se = amp / math.sqrt(np.mean(ex ** 2 + ey ** 2))
    _interior(f["bx"])[...] = s * dbx
    _interior(f["by"])[...] = s * dby
    _interior(f["bz"])[...] = 1.0 + s * dbz
    _interior(f["ex"])[...] = se * ex
    _interior(f["ey"])[...] = se * ey
    f["b0z"][...] = 1.0
    return Problem("alfven", nn, (1, 1, 1), cc, f, steps=steps)


# --- config: twisted loop (user_loop.F90 form) -----------------------------
LOOP_COPY, LOOP_GHOST_DRIVER, LOOP_DRIVEN = 0, 1, 2


def loop(n=384, L0=None, a_L0=3.0, h_L0=2.0, r_L0=0.5, cc=0.45, twist_angle=None, seed=3,
         t0=None, steps=1000, twist=True, driver=LOOP_GHOST_DRIVER, dtype=np.float64) -> Problem:
    """user_loop.F90:27-93: two opposite sources at depth h*L0 below the
    floor, +-a*L0 either side of the box centre; B0 = B everywhere including
    ghosts; x, y periodic, z open with the floor/top hook.  L0 scales as
    20*n/256 (inputs/input.loop:34 on 256 cells).  With twist=True the
    floor hook also rotates footpoint 1 (the driver of the commented block
    user_loop.F90:122-150, r0 = r_L0*L0 as user_uniform.F90:38 scales it);
    beta_twist is set so that the centre of the footpoint turns by
    twist_angle (seeded in [pi/2, 3pi/2]) over `steps` steps.

    driver = LOOP_DRIVEN instead applies that driver on z <= 0 (the floor
    plane included) and closes the top as a conductor (ex = ey = 0, bz = the
    potential field), the boundary form of user_uniform.F90:113-142: the
    ghost-only driver and the floor/top copy never reach the interior
    update, so only this form twists the loop."""
    nn = (n, n, n)
    d = np.dtype(dtype).type
    L0 = 20.0 * n / 256 if L0 is None else L0
    if twist_angle is None:
        twist_angle = float(np.random.default_rng(seed).uniform(0.5 * math.pi, 1.5 * math.pi))
    t0 = steps / 10.0 if t0 is None else t0
    L0, a0, h0 = d(L0), d(a_L0), d(h_L0)
    xc = d(0.5) * d(n)
    yc1 = d(0.5) * d(n) - a0 * L0
    yc2 = d(0.5) * d(n) + a0 * L0
    zc = -h0 * L0
    bnorm = d(1000.0) * h0 ** 2 * L0 ** 2
    r0 = d(r_L0) * L0
    f = _zeros(nn)
    i, j, k = _coords(nn, full=True)

    def comp(rx, ry1, ry2, rz, num1, num2):
        rr1 = np.sqrt(rx * rx + ry1 * ry1 + rz * rz)
        rr2 = np.sqrt(rx * rx + ry2 * ry2 + rz * rz)
        return float(bnorm) * (num1 / rr1 ** 3 - num2 / rr2 ** 3)

    xcf, yc1f, yc2f, zcf = float(xc), float(yc1), float(yc2), float(zc)
    rx, ry1, ry2, rz = i - xcf, j + 0.5 - yc1f, j + 0.5 - yc2f, k + 0.5 - zcf
    f["bx"][...] = comp(rx, ry1, ry2, rz, rx, rx)
    rx, ry1, ry2, rz = i + 0.5 - xcf, j - yc1f, j - yc2f, k + 0.5 - zcf
    f["by"][...] = comp(rx, ry1, ry2, rz, ry1, ry2)
    rx, ry1, ry2, rz = i + 0.5 - xcf, j + 0.5 - yc1f, j + 0.5 - yc2f, k - zcf
    f["bz"][...] = comp(rx, ry1, ry2, rz, rz, rz)
    for c in "xyz":
        f["b0" + c][...] = f["b" + c]
    # footpoint angular speed is CC * omega (drift v = c E x B / B^2, E = -omega r x B)
    dist = float(h0 * L0) - (0.0 if driver == LOOP_DRIVEN else 1.0)
    ramp = sum(math.tanh(t / t0) for t in range(steps))
    beta = twist_angle * float(r0) * math.exp(dist / float(r0)) / (cc * cc * max(ramp, 1e-30))
    mode = float(driver) if twist else float(LOOP_COPY)
    params = (xcf, yc1f, yc2f, zcf, float(bnorm), float(r0), beta, float(t0), mode)
    return Problem("loop", nn, (1, 1, 0), cc, f, BC_LOOP, params, steps=steps, g=1)


# --- config: aligned rotator (user_psr.F90 form) ---------------------------
PSR_PI = 3.14159265  # user_psr.F90:11


def dipole_field(params, x, y, z, step=0):
    """user_psr.F90:345-374 getDipole, in FP64 on the host."""
    R, ang, om, xc, yc, zc = params[:6]
    nx, ny, nz = x - xc, y - yc, z - zc
    rr = 1.0 / np.sqrt(nx ** 2 + ny ** 2 + nz ** 2)
    nx, ny, nz = nx * rr, ny * rr, nz * rr
    rr = rr ** 3
    ph = om * step
    mux = R ** 3 * math.sin(ang) * math.cos(ph)
    muy = R ** 3 * math.sin(ang) * math.sin(ph)
    muz = R ** 3 * math.cos(ang)
    mdn = mux * nx + muy * ny + muz * nz
    return (3 * nx * mdn - mux) * rr, (3 * ny * mdn - muy) * rr, (3 * nz * mdn - muz) * rr


def psr(n=512, radius=20.0, r_lc=80.0, angle_deg=0.0, absorb_buff=40, cc=0.5, periods=2.0,
        absorb_fixed=True, dtype=np.float64) -> Problem:
    """user_psr.F90:26-69: dipole of radius R at the box centre (interior
    cells only, ghosts zero, B0 = 0), all boundaries absorbing.  The period
    puts the light cylinder at r_lc cells: P = 2 pi r_lc / c."""
    nn = (n, n, n)
    d = np.dtype(dtype).type
    period = 2 * math.pi * r_lc / cc
    pi = d(PSR_PI)
    ang = d(angle_deg) * pi / d(180.0)
    om = d(2.0) * pi / d(period)
    c = d(0.5) * d(n)
    params = (float(d(radius)), float(ang), float(om), float(c), float(c), float(c))
    f = _zeros(nn)
    i, j, k = _coords(nn)
    bx, _, _ = dipole_field(params, i, j + 0.5, k + 0.5)
    _, by, _ = dipole_field(params, i + 0.5, j, k + 0.5)
    _, _, bz = dipole_field(params, i + 0.5, j + 0.5, k)
    _interior(f["bx"])[...] = bx
    _interior(f["by"])[...] = by
    _interior(f["bz"])[...] = bz
    steps = int(round(periods * period))
    return Problem("psr", nn, (2, 2, 2), cc, f, BC_PSR, params, absorb_buff=absorb_buff,
                   absorb_fixed=absorb_fixed, steps=steps)


# --- user_uniform.F90 form ---------------------------------------------------
def uniform(n=64, L0=10.0, r_L0=1.0, beta_twist=0.1, t0=50.0, cc=0.45, steps=100) -> Problem:
    """user_uniform.F90:27-69: bz = 1/(1 + r/r0) about the vertical axis
    through the box centre, B0 = B, x, y periodic and z open; the hook
    rotates the floor (z <= 0)."""
    nn = (n, n, n) if np.isscalar(n) else tuple(n)
    xc, yc = 0.5 * nn[0], 0.5 * nn[1]
    r0 = r_L0 * L0
    f = _zeros(nn)
    i, j, k = _coords(nn, full=True)
    rr = np.sqrt((i + 0.5 - xc) ** 2 + (j + 0.5 - yc) ** 2 + 1e-7)
    f["bz"][...] = 1.0 / (1.0 + rr / r0) + 0 * k
    f["b0z"][...] = f["bz"]
    params = (xc, yc, r0, beta_twist, t0, 1.0)
    return Problem("uniform", nn, (1, 1, 0), cc, f, BC_UNIFORM, params, steps=steps, g=1)


def load(target, p: Problem):
    """Upload a problem into a Meshblock or an oracle (both expose set_fields
    and set_bc)."""
    target.set_fields(**p.fields)
    target.set_bc(p.bc_kind, p.bc_params)


# --- config: batched parameter scan -------------------------------------------
def scan(count=48, nmin=64, nmax=192, seed=11, steps=300, cc=0.45, sizes=None):
    """A parameter scan of independent instances with seeded, uneven grid
    sizes in [nmin, nmax]: a mix of driven twisted loops (user_loop field,
    footpoint twist angle seeded in [pi/2, 3pi/2]) and vacuum plane waves
    (user_emwave form, amplitude seeded in [1e-3, 1e-2], seeded in-plane
    mode)."""
    rng = np.random.default_rng(seed)
    out = []
    for q in range(count):
        n = int(sizes[q]) if sizes is not None else int(rng.integers(nmin, nmax + 1))
        if rng.random() < 0.5:
            p = loop(n, twist_angle=float(rng.uniform(0.5 * math.pi, 1.5 * math.pi)), steps=steps, cc=cc,
                     driver=LOOP_DRIVEN)
        else:
            mode = (int(rng.integers(1, 4)), int(rng.integers(0, 4)))
            p = emwave(n, amp=float(10 ** rng.uniform(-3, -2)), mode=mode, cc=cc, steps=steps)
        p.name = f"{p.name}{q}"
        out.append(p)
    return out


def batch_specs(probs):
    """(n, boundary, cc, bc_kind, bc_params) of each problem, for pytff.Batch."""
    return [(p.n, p.boundary, p.cc, p.bc_kind, p.bc_params) for p in probs]
