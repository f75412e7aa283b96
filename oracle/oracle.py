"""Python handle on the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module, and only as the checker.  The product
path (``tristan-ff_amd``: libtff.so and its Fortran/Python hosts) never
imports it.

Two CPU implementations are wrapped:

* ``Oracle``     -- oracle/tff_oracle.c, a C restatement of the reference's
  force-free step (src/algorithms/field_solver.F90, the single-rank ghost
  fill of src/logistics/exchange_fields.F90:11-197, the stage order of
  src/tristanmainloop.F90:40-84 and the user/*.F90 boundary hooks), built in
  FP64 and in FP32 (the reference's shipped precision).
* ``RefSolver``  -- the reference's own m_fldsolver compiled from its sources
  by oracle/Makefile into oracle/_ref (present only where the reference tree
  was available at build time).

How each part of the oracle is pinned (tests/test_oracle_*.py):

=====================  ======================================================
initRKstep, rk3Step,   bit-for-bit against the reference's own compiled
cleanEpar (incl. the   field_solver.F90 (oracle/_ref), FP64 and FP32, on
absorbing rk3Update)   random fields with every boundary mode
exchangeFields (1 rk)  periodic-wrap invariant; untouched open ghosts;
                       brute-force comparison with numpy.pad(mode="wrap")
stage sequence + RK    closed form: discrete Yee plane-wave eigenmode grows by
coefficients (H1)      G(z)=1+z+z^2/2+z^3/6 per step, three orientations
div B, static states   div_h B conserved to roundoff; uniform B static
force-free current     closed-form current for linear E in uniform B;
                       limiter and E.B=0 projection against hand algebra
user_loop BC           floor copy by construction (ghost == k=0 plane);
                       driven form: footpoint field at sample points
user_uniform BC        properties its definition fixes: a 90-degree turn
                       about the box axis maps Ey onto Ex bit for bit
                       (radial field, Yee offsets), the tanh(t/t0) ramp by
                       step ratios, the e^(-r/r0) profile by radius ratios,
                       linearity in beta, Bz = the initial potential field
user_psr BC            closed form on uniform fields (tests/closed_forms.py:
                       radial and tangential relaxation toward the rotated
                       dipole and the corotation field, each with its own
                       width and shell), every cell of r < 2R; the widths
                       or shells swapped fail the same comparison
absorbing layer,       fixed mode (jg from j, delta capped at 1) and the
both modes             layer as coded (jg from i, uncapped): closed-form
                       damping factors of a static uniform field, every
                       cell and three cells by hand; as coded also
                       bit-for-bit against oracle/_ref
cleanEpar guard        fixed mode leaves E unchanged at |B|^2 = 0
output interpolation   bit-for-bit against the reference's own compiled
(helpers.F90:93-257)   interpFromEdges / interpFromFaces (oracle/_ref), FP64
                       and FP32, random fields and offsets; linear fields
                       interpolate exactly at nodes
output sampling        node count ceil(s / istep) per axis and the nodes
(write_output.F90)     0, istep, ... against brute force
=====================  ======================================================
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
NG = 2

FIELDS = ["ex", "ey", "ez", "bx", "by", "bz", "b0x", "b0y", "b0z",
          "enx", "eny", "enz", "bnx", "bny", "bnz",
          "dex", "dey", "dez", "dbx", "dby", "dbz", "rho"]
FID = {n: i for i, n in enumerate(FIELDS)}

BC_NONE, BC_LOOP, BC_PSR, BC_UNIFORM = 0, 1, 2, 3

# SSP-RK3 stage coefficients of src/tristanmainloop.F90:51,62,73
def rk_coefficients(dtype) -> list[float]:
    d = np.dtype(dtype).type
    return [float(x) for x in (d(1.0), d(0.0), d(1.0),
                               d(3.0) / d(4.0), d(1.0) / d(4.0), d(1.0) / d(4.0),
                               d(1.0) / d(3.0), d(2.0) / d(3.0), d(2.0) / d(3.0))]


def _lib(path: str) -> ctypes.CDLL:
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is not built (run `make -f oracle/Makefile`)")
    return ctypes.CDLL(path)


def build(ref: bool = True) -> None:
    """Compile the oracle (and oracle/_ref when the reference tree exists)."""
    import subprocess
    target = "all" if ref else "oracle"
    subprocess.run(["make", "-s", "-f", os.path.join(HERE, "Makefile"), target], check=True)


class Oracle:
    """A single meshblock of the reference's field state on the CPU."""

    _libs: dict = {}

    def __init__(self, n, boundary=(1, 1, 1), cc=0.45, absorb_buff=20,
                 absorb_fixed=False, dtype=np.float64):
        nx, ny, nz = (n, n, n) if np.isscalar(n) else tuple(n)
        self.dtype = np.dtype(dtype)
        name = "liboracle_f64.so" if self.dtype == np.float64 else "liboracle_f32.so"
        if name not in Oracle._libs:
            lib = _lib(os.path.join(HERE, "_build", name))
            lib.orc_create.restype = ctypes.c_void_p
            lib.orc_create.argtypes = [ctypes.c_int] * 14 + [ctypes.c_double]
            lib.orc_field.restype = ctypes.c_void_p
            lib.orc_field.argtypes = [ctypes.c_void_p, ctypes.c_int]
            for fn in ("orc_destroy", "orc_init_rk_step", "orc_clean_epar", "orc_preloop"):
                getattr(lib, fn).argtypes = [ctypes.c_void_p]
            lib.orc_rk3_step.argtypes = [ctypes.c_void_p] + [ctypes.c_double] * 3
            lib.orc_exchange_fields.argtypes = [ctypes.c_void_p, ctypes.c_int]
            lib.orc_user_bc.argtypes = [ctypes.c_void_p, ctypes.c_int]
            lib.orc_set_bc.argtypes = [ctypes.c_void_p, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_double), ctypes.c_int]
            lib.orc_step.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
            lib.orc_run.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_double), ctypes.c_void_p]
            lib.orc_reduce.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
            lib.orc_interp.argtypes = [ctypes.c_void_p] + [ctypes.c_double] * 3 + [ctypes.c_int] * 3 + [ctypes.c_void_p]
            lib.orc_output_dims.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
            lib.orc_output_fields.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
            Oracle._libs[name] = lib
        self.lib = Oracle._libs[name]
        self.n = (nx, ny, nz)
        self.boundary = tuple(boundary)
        self.h = self.lib.orc_create(nx, ny, nz, 0, 0, 0, nx, ny, nz, *boundary,
                                     absorb_buff, int(absorb_fixed), float(cc))
        if not self.h:
            raise MemoryError("orc_create failed")
        self.shape = (nz + 2 * NG, ny + 2 * NG, nx + 2 * NG)  # C order of a Fortran array
        self.coef = rk_coefficients(self.dtype)

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            self.lib.orc_destroy(h)
            self.h = None

    def field(self, name: str) -> np.ndarray:
        """Live view (k, j, i) of one field array including ghosts."""
        ptr = self.lib.orc_field(self.h, FID[name])
        ctype = ctypes.c_double if self.dtype == np.float64 else ctypes.c_float
        arr = np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctype)),
                                    shape=self.shape)
        return arr

    def set_fields(self, **arrays):
        for k, v in arrays.items():
            self.field(k)[...] = v

    def get(self, *names):
        return {n: self.field(n).copy() for n in names}

    def set_bc(self, kind: int, params=()):
        p = (ctypes.c_double * 16)(*[float(x) for x in params])
        self.lib.orc_set_bc(self.h, kind, p, len(params))

    def init_rk_step(self):
        self.lib.orc_init_rk_step(self.h)

    def rk3_step(self, c1, c2, c3):
        self.lib.orc_rk3_step(self.h, c1, c2, c3)

    def clean_epar(self):
        self.lib.orc_clean_epar(self.h)

    def exchange_fields(self, exchange_b0=False):
        self.lib.orc_exchange_fields(self.h, int(exchange_b0))

    def user_bc(self, step: int):
        self.lib.orc_user_bc(self.h, step)

    def preloop(self):
        self.lib.orc_preloop(self.h)

    def step(self, timestep: int):
        c = (ctypes.c_double * 9)(*self.coef)
        self.lib.orc_step(self.h, timestep, c)

    def run(self, first_step: int, nsteps: int, energy: bool = False):
        c = (ctypes.c_double * 9)(*self.coef)
        hist = np.zeros((max(nsteps, 1), 2)) if energy else None
        self.lib.orc_run(self.h, first_step, nsteps, c,
                         hist.ctypes.data if energy else None)
        return hist

    def reduce(self):
        out = (ctypes.c_double * 4)()
        self.lib.orc_reduce(self.h, out)
        return list(out)

    def interp(self, dx, dy, dz, i, j, k) -> np.ndarray:
        """interpFromEdges (E) and interpFromFaces (B) at offset (dx, dy, dz)
        from cell (i, j, k): [ex, ey, ez, bx, by, bz]."""
        out = np.empty(6, dtype=self.dtype)
        self.lib.orc_interp(self.h, dx, dy, dz, i, j, k, out.ctypes.data)
        return out

    def output_fields(self, istep: int) -> dict:
        """The six node-interpolated field datasets of writeFields_hdf5 for
        output_istep >= 2, each shaped (k, j, i)."""
        d = (ctypes.c_int * 3)()
        self.lib.orc_output_dims(self.h, istep, d)
        arrs = [np.empty((d[2], d[1], d[0]), dtype=self.dtype) for _ in range(6)]
        ptrs = (ctypes.c_void_p * 6)(*[a.ctypes.data for a in arrs])
        self.lib.orc_output_fields(self.h, istep, ptrs)
        return dict(zip(("ex", "ey", "ez", "bx", "by", "bz"), arrs))


class RefSolver:
    """The reference's own m_fldsolver (oracle/_ref), one global instance per
    precision because the reference keeps its state in module variables."""

    _libs: dict = {}

    def __init__(self, n, boundary=(1, 1, 1), cc=0.45, absorb_buff=20, dtype=np.float64):
        self.dtype = np.dtype(dtype)
        name = "libtffref_f64.so" if self.dtype == np.float64 else "libtffref_f32.so"
        if name not in RefSolver._libs:
            lib = _lib(os.path.join(HERE, "_ref", name))
            lib.ref_setup.argtypes = [ctypes.c_int] * 7 + [ctypes.c_double]
            lib.ref_xfer.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int]
            lib.ref_rk3_step.argtypes = [ctypes.c_double] * 3
            lib.ref_real_size.restype = ctypes.c_int
            lib.ref_run.argtypes = [ctypes.c_int]
            lib.ref_interp.argtypes = [ctypes.c_double] * 3 + [ctypes.c_int] * 3 + [ctypes.c_void_p]
            RefSolver._libs[name] = lib
        self.lib = RefSolver._libs[name]
        assert self.lib.ref_real_size() == self.dtype.itemsize
        nx, ny, nz = (n, n, n) if np.isscalar(n) else tuple(n)
        self.shape = (nz + 2 * NG, ny + 2 * NG, nx + 2 * NG)
        self.lib.ref_setup(nx, ny, nz, *boundary, absorb_buff, float(cc))

    @staticmethod
    def available(dtype=np.float64) -> bool:
        name = "libtffref_f64.so" if np.dtype(dtype) == np.float64 else "libtffref_f32.so"
        return os.path.exists(os.path.join(HERE, "_ref", name))

    def put(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=self.dtype).reshape(self.shape)
        self.lib.ref_xfer(FID[name], a.ctypes.data, 1)

    def get(self, name):
        a = np.empty(self.shape, dtype=self.dtype)
        self.lib.ref_xfer(FID[name], a.ctypes.data, 0)
        return a

    def init_rk_step(self):
        self.lib.ref_init_rk_step()

    def rk3_step(self, c1, c2, c3):
        self.lib.ref_rk3_step(c1, c2, c3)

    def clean_epar(self):
        self.lib.ref_clean_epar()

    def exchange(self):
        self.lib.ref_exchange()

    def run(self, nsteps: int):
        """nsteps timesteps of tristanmainloop.F90:50-83 (no user hook)."""
        self.lib.ref_run(nsteps)

    def interp(self, dx, dy, dz, i, j, k) -> np.ndarray:
        """The reference's own interpFromEdges / interpFromFaces
        (helpers.F90:93-257): [ex, ey, ez, bx, by, bz] in double."""
        out = np.empty(6, dtype=np.float64)
        self.lib.ref_interp(dx, dy, dz, i, j, k, out.ctypes.data)
        return out
