"""Initial conditions and boundary-hook parameters of the problems the
force-free step is run on (the Python counterpart of the user/*.F90 files'
userReadInput + userInitFields; the Fortran host has its own user files).

Every builder returns a ``Problem``: the field arrays in the reference's
layout (C-ordered ``(nz+4, ny+4, nx+4)`` views of Fortran arrays with two
ghost layers, FP64 on the host), the boundary flags, CC, the absorbing-layer
settings and the boundary-kernel kind and parameters.  Every problem except
the plane wave also carries the recipe of its initial condition for the
device (``init``): ``load(..., device=True)`` has libtff generate the fields
in HBM (tff_init_fields, tff_init_spectrum), and a builder called with
``host=False`` builds no host arrays at all, for boxes whose FP64 host copy
would take gigabytes.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

NG = 2
BC_NONE, BC_LOOP, BC_PSR, BC_UNIFORM = 0, 1, 2, 3
NAMES = ("ex", "ey", "ez", "bx", "by", "bz", "b0x", "b0y", "b0z")
# the device-side initial conditions (tff_init_fields' kinds)
IC_LOOP, IC_PSR, IC_UNIFORM = 1, 2, 3


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
    # ("fields", kind, params) or ("spectrum", modes, background), or None
    init: tuple = None

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
class ParkMiller:
    """The reference's generator (src/tools/auxiliary.F90:246-257,
    randomNum): the Lehmer / Park-Miller minimal standard, seed <- 16807
    seed mod (2^31 - 1), returning seed / 2^31, in FP64 as the reference
    does (DMOD on doubles: every product is below 2^46, so exact).  The
    seed starts at 123457 + rank (initializeRandomSeed, :296-301)."""

    def __init__(self, rank: int = 0):
        self.seed = 123457.0 + float(rank)

    def uniform(self) -> float:
        self.seed = math.fmod(16807.0 * self.seed, 2147483647.0)
        return self.seed / 2147483648.0

    def normal(self) -> float:
        """One standard normal draw by the Box-Muller transform of two
        uniforms (the cosine branch; u1 = 0 cannot occur)."""
        u1, u2 = self.uniform(), self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


def alfven_modes(n, seed=1, amp=0.05, kmax=8):
    """The seeded spectrum of alfven(): one row of 15 per wave vector k of
    the half space (kz > 0, or kz = 0 and ky > 0, or kz = ky = 0 and kx > 0)
    with 1 <= |k| <= kmax, holding kx ky kz and the complex amplitudes of
    ex ey ez bx by bz.  Each k draws complex normal amplitudes (the
    reference's Park-Miller generator started at rank `seed`, Box-Muller)
    for a vector potential A and a stream function psi; B's are those of the
    forward-difference curl of A (so div_h B = 0 to rounding) and E's those
    of the forward-difference gradient of psi in the x-y plane.  With 2 kmax < n no
    two modes alias and none is its own conjugate, so the box mean of a
    mode's square is |c|^2 / 2 (Parseval) and the scales below make
    rms|dB| = rms|E| = amp exactly."""
    nx, ny, nz = n
    if 2 * kmax >= min(n):
        raise ValueError(f"kmax {kmax} needs a box wider than {2 * kmax} cells")
    r = np.arange(-kmax, kmax + 1)
    kz, ky, kx = (a.ravel() for a in np.meshgrid(r, r, r, indexing="ij"))
    half = (kz > 0) | ((kz == 0) & ((ky > 0) | ((ky == 0) & (kx > 0))))
    k2 = kx * kx + ky * ky + kz * kz
    sel = half & (k2 >= 1) & (k2 <= kmax * kmax)
    kx, ky, kz = kx[sel], ky[sel], kz[sel]
    rng = ParkMiller(seed)
    z = np.array([rng.normal() for _ in range(kx.size * 8)]).reshape(kx.size, 4, 2)
    z = z[..., 0] + 1j * z[..., 1]
    a, psi = z[:, :3], z[:, 3]
    # forward difference of exp(i theta) along each axis
    dx = np.exp(2j * np.pi * kx / nx) - 1.0
    dy = np.exp(2j * np.pi * ky / ny) - 1.0
    dz = np.exp(2j * np.pi * kz / nz) - 1.0
    cb = np.stack([a[:, 2] * dy - a[:, 1] * dz, a[:, 0] * dz - a[:, 2] * dx, a[:, 1] * dx - a[:, 0] * dy], 1)
    ce = np.stack([psi * dx, psi * dy, 0.0 * psi], 1)
    cb *= amp / math.sqrt(0.5 * np.sum(np.abs(cb) ** 2))
    ce *= amp / math.sqrt(0.5 * np.sum(np.abs(ce) ** 2))
    modes = np.zeros((kx.size, 15))
    modes[:, 0], modes[:, 1], modes[:, 2] = kx, ky, kz
    c = np.concatenate([ce, cb], 1)
    modes[:, 3::2], modes[:, 4::2] = c.real, c.imag
    return modes


def spectrum_fields(n, modes, background):
    """The host form of tff_init_spectrum: component q of interior cell
    (i, j, k) = background[q] + sum over modes of Re[c_q exp(2 pi i (kx i /
    nx + ky j / ny + kz k / nz))], ghosts zero, B0 = the background B
    everywhere; the mode sum is one inverse FFT per component."""
    nx, ny, nz = n
    f = _zeros(n)
    idx = tuple(modes[:, a].astype(np.int64) % m for a, m in ((2, nz), (1, ny), (0, nx)))
    for q, name in enumerate(NAMES[:6]):
        spec = np.zeros((nz, ny, nx), dtype=np.complex128)
        spec[idx] = modes[:, 3 + 2 * q] + 1j * modes[:, 4 + 2 * q]
        _interior(f[name])[...] = background[q] + np.fft.ifftn(spec).real * (nx * ny * nz)
    for c in range(3):
        f[NAMES[6 + c]][...] = background[3 + c]
    return f


def alfven(n=256, seed=1, amp=0.05, kmax=8, cc=0.45, steps=500, host=True) -> Problem:
    """B = z^ + dB, E = the in-plane gradient of a stream function: the
    seeded band-limited spectrum of alfven_modes (1 <= |k| <= kmax), with
    div_h B = 0 at t = 0 and rms|dB| = rms|E| = amp; periodic box."""
    nn = (n, n, n) if np.isscalar(n) else tuple(n)
    modes = alfven_modes(nn, seed, amp, kmax)
    bg = (0.0, 0.0, 0.0, 0.0, 0.0, 1.0)
    f = spectrum_fields(nn, modes, bg) if host else None
    return Problem("alfven", nn, (1, 1, 1), cc, f, steps=steps, init=("spectrum", modes, bg))


# --- config: twisted loop (user_loop.F90 form) -----------------------------
LOOP_COPY, LOOP_GHOST_DRIVER, LOOP_DRIVEN = 0, 1, 2


def loop(n=384, L0=None, a_L0=3.0, h_L0=2.0, r_L0=0.5, cc=0.45, twist_angle=None, seed=3,
         t0=None, steps=1000, twist=True, driver=LOOP_GHOST_DRIVER, dtype=np.float64, host=True) -> Problem:
    """user_loop.F90:27-93: two opposite sources at depth h*L0 below the
    floor, +-a*L0 either side of the box centre; B0 = B everywhere including
    ghosts; x, y periodic, z open with the floor/top hook.  L0 scales as
    20*n/256 (inputs/input.loop:34 on 256 cells).  With twist=True the
    floor hook also rotates footpoint 1 (the driver of the commented block
    user_loop.F90:122-150, r0 = r_L0*L0 as user_uniform.F90:38 scales it);
    beta_twist is set so that the centre of the footpoint turns by
    twist_angle (seeded in [pi/2, 3pi/2]) over `steps` steps.

    driver = LOOP_DRIVEN instead applies that driver's ex, ey on z <= 0 (the
    floor plane included; bz only on the ghosts below it, so the floor's bz
    follows Faraday's law and div_h B is kept) and closes the top as a
    conductor (ex = ey = 0, bz = the potential field), the boundary form of
    user_uniform.F90:113-142: the ghost-only driver and the floor/top copy
    never reach the interior update, so only this form twists the loop."""
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
    xcf, yc1f, yc2f, zcf = float(xc), float(yc1), float(yc2), float(zc)
    f = None
    if host:
        f = _zeros(nn)
        i, j, k = _coords(nn, full=True)

        def comp(rx, ry1, ry2, rz, num1, num2):
            rr1 = np.sqrt(rx * rx + ry1 * ry1 + rz * rz)
            rr2 = np.sqrt(rx * rx + ry2 * ry2 + rz * rz)
            return float(bnorm) * (num1 / rr1 ** 3 - num2 / rr2 ** 3)

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
    return Problem("loop", nn, (1, 1, 0), cc, f, BC_LOOP, params, steps=steps, g=1,
                   init=("fields", IC_LOOP, params[:5]))


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
        absorb_fixed=True, dtype=np.float64, host=True) -> Problem:
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
    f = None
    if host:
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
                   absorb_fixed=absorb_fixed, steps=steps, init=("fields", IC_PSR, params))


# --- user_uniform.F90 form ---------------------------------------------------
def uniform(n=64, L0=10.0, r_L0=1.0, beta_twist=0.1, t0=50.0, cc=0.45, steps=100, host=True) -> Problem:
    """user_uniform.F90:27-69: bz = 1/(1 + r/r0) about the vertical axis
    through the box centre, B0 = B, x, y periodic and z open; the hook
    rotates the floor (z <= 0)."""
    nn = (n, n, n) if np.isscalar(n) else tuple(n)
    xc, yc = 0.5 * nn[0], 0.5 * nn[1]
    r0 = r_L0 * L0
    f = None
    if host:
        f = _zeros(nn)
        i, j, k = _coords(nn, full=True)
        rr = np.sqrt((i + 0.5 - xc) ** 2 + (j + 0.5 - yc) ** 2 + 1e-7)
        f["bz"][...] = 1.0 / (1.0 + rr / r0) + 0 * k
        f["b0z"][...] = f["bz"]
    params = (xc, yc, r0, beta_twist, t0, 1.0)
    return Problem("uniform", nn, (1, 1, 0), cc, f, BC_UNIFORM, params, steps=steps, g=1,
                   init=("fields", IC_UNIFORM, (xc, yc, r0, 1.0)))


# --- diagnostics: div_h B -------------------------------------------------------
def div_h_b(bx, by, bz) -> np.ndarray:
    """div_h B at the centre of every interior cell, from ghost-inclusive
    (k, j, i) arrays: (bx(i+1) - bx(i)) + (by(j+1) - by(j)) + (bz(k+1) - bz(k)),
    the cell-centred divergence of the Yee face fields, in FP64 (the stored
    ghosts supply the upper faces of the last interior cells)."""
    s, s1 = slice(NG, -NG), slice(NG + 1, -NG + 1 if NG > 1 else None)
    d = np.asarray(bx[s, s, s1], dtype=np.float64) - bx[s, s, s]
    d += np.asarray(by[s, s1, s], dtype=np.float64) - by[s, s, s]
    d += np.asarray(bz[s1, s, s], dtype=np.float64) - bz[s, s, s]
    return d


def divb_free_mask(p: Problem):
    """Interior cells whose div_h B only the curl update changes (None: all
    of them).  Excluded are the cells whose faces a boundary hook writes or
    the absorbing layer damps component by component: the rotator's
    conducting-sphere region (centre within 2R + 3 of the star) and every
    cell whose centre lies within 2 cells of the absorbing layer's inner
    radius (field_solver.F90:198-269: min(n)/2 - absorb_buff); the floor
    plane of the user_uniform hook, which pins bz there."""
    nx, ny, nz = p.n
    mask = None
    if p.bc_kind == BC_PSR or 2 in p.boundary:
        i, j, k = _coords(p.n)
        mask = np.ones((nz, ny, nx), dtype=bool)
        if 2 in p.boundary:
            xc, yc, zc = 0.5 * nx, 0.5 * ny, 0.5 * nz
            rabs = min(xc, yc, zc) - p.absorb_buff
            r2 = (i + 0.5 - xc) ** 2 + (j + 0.5 - yc) ** 2 + (k + 0.5 - zc) ** 2
            mask &= r2 < (rabs - 2.0) ** 2
            del r2
        if p.bc_kind == BC_PSR:
            R, xc, yc, zc = p.bc_params[0], p.bc_params[3], p.bc_params[4], p.bc_params[5]
            r2 = (i + 0.5 - xc) ** 2 + (j + 0.5 - yc) ** 2 + (k + 0.5 - zc) ** 2
            mask &= r2 >= (2.0 * R + 3.0) ** 2
    elif p.bc_kind == BC_UNIFORM:
        mask = np.ones((nz, ny, nx), dtype=bool)
        mask[0] = False
    return mask


def divb_report(p: Problem, d0: np.ndarray, d1: np.ndarray, bmax: float) -> dict:
    """div_h B at t = 0 (d0) and after the run (d1), and its change, over
    every interior cell and over the cells no hook or absorber writes
    (divb_free_mask); `bmax` = max |B| at t = 0, the scale the change is
    quoted against."""
    m = divb_free_mask(p)
    dd = np.abs(d1 - d0)
    out = {"max_t0": float(np.abs(d0).max()), "max_end": float(np.abs(d1).max()),
           "l2_end": float(np.sqrt(np.mean(d1 * d1))), "change_max": float(dd.max()),
           "b_max_t0": float(bmax)}
    if m is not None:
        out["change_max_free"] = float(dd[m].max()) if m.any() else 0.0
        out["change_l2_free"] = float(np.sqrt(np.mean(dd[m] ** 2))) if m.any() else 0.0
        out["free_cells"] = int(m.sum())
    else:
        out["change_max_free"] = out["change_max"]
        out["change_l2_free"] = float(np.sqrt(np.mean(dd * dd)))
        out["free_cells"] = int(dd.size)
    out["change_max_free_rel_bmax"] = out["change_max_free"] / bmax if bmax > 0 else 0.0
    return out


def load(target, p: Problem, device=False):
    """Upload a problem into a Meshblock or an oracle (both expose set_fields
    and set_bc).  device=True (or a problem built with host=False) has the
    Meshblock generate the initial fields in HBM from p.init instead."""
    if device or p.fields is None:
        if p.init is None:
            raise ValueError(f"problem {p.name} has no device initial condition")
        if not hasattr(target, "init_fields"):
            raise TypeError(f"{type(target).__name__} cannot generate fields; build the problem with host=True")
        if p.init[0] == "spectrum":
            target.init_spectrum(p.init[1], p.init[2])
        else:
            target.init_fields(p.init[1], p.init[2])
    else:
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
