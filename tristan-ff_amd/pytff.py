"""Python host for libtff.so, mirroring the reference's Fortran interface.

The reference drives its solver through module procedures on module-global
arrays (src/algorithms/field_solver.F90: initRKstep, rk3Step, cleanEpar;
src/logistics/exchange_fields.F90: exchangeFields; user/*.F90:
userFieldBoundaryConditions).  ``Meshblock`` keeps those names and their
argument meaning on top of the C ABI of include/tff.h; arrays go in and out
in the reference's layout (Fortran order with two ghost layers), seen from
numpy as C-ordered ``(nz+4, ny+4, nx+4)`` arrays.

There is no CPU fallback: if libtff.so is missing or the GPU is absent the
constructor raises.
"""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
# TFF_LIBRARY selects another build of the library (e.g. a variant under test)
LIB_PATH = os.environ.get("TFF_LIBRARY") or os.path.join(PKG, "_build", "libtff.so")
HEADER = os.path.join(ROOT, "include", "tff.h")
NG = 2

FP64, FP32, FP32_BULK, FP32_COMP = 0, 1, 2, 3
OPEN, PERIODIC, ABSORB = 0, 1, 2
BC_NONE, BC_LOOP, BC_PSR, BC_UNIFORM = 0, 1, 2, 3
FIELDS = {"ex": 0, "ey": 1, "ez": 2, "bx": 3, "by": 4, "bz": 5,
          "b0x": 6, "b0y": 7, "b0z": 8, "enx": 9, "eny": 10, "enz": 11,
          "bnx": 12, "bny": 13, "bnz": 14}
_PREC_NAMES = {"fp64": FP64, "fp32": FP32, "fp32_bulk": FP32_BULK, "fp32_comp": FP32_COMP}


class TffError(RuntimeError):
    pass


class Config(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("boundary", ctypes.c_int * 3), ("absorb_buff", ctypes.c_int),
                ("absorb_fixed", ctypes.c_int), ("precision", ctypes.c_int),
                ("device", ctypes.c_int), ("cc", ctypes.c_double)]


class Instance(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("boundary", ctypes.c_int * 3), ("cc", ctypes.c_double),
                ("bc_kind", ctypes.c_int), ("nparams", ctypes.c_int),
                ("params", ctypes.c_double * 16)]


_lib = None


def kernel_source_hash() -> str:
    """A short SHA-256 over the HIP sources libtff.so is built from, so that
    a stored counter profile can be matched to the kernels it measured."""
    import hashlib
    d = os.path.join(PKG, "csrc")
    h = hashlib.sha256()
    for name in sorted(os.listdir(d)):
        if name.endswith((".h", ".hip")):
            h.update(name.encode())
            with open(os.path.join(d, name), "rb") as fh:
                h.update(fh.read())
    return h.hexdigest()[:16]


def header_functions() -> list[str]:
    """Names of the functions include/tff.h declares."""
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(tff_\w+)\s*\(", txt, flags=re.M)


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise TffError(f"{path} is not built; run `python tristan-ff_amd/build.py`")
    lib = ctypes.CDLL(path)
    P = ctypes.c_void_p
    lib.tff_create.argtypes = [ctypes.POINTER(Config), ctypes.POINTER(P)]
    lib.tff_destroy.argtypes = [P]
    lib.tff_destroy.restype = None
    lib.tff_last_error.argtypes = [P]
    lib.tff_last_error.restype = ctypes.c_char_p
    lib.tff_upload.argtypes = [P, ctypes.c_int, P, ctypes.c_int]
    lib.tff_download.argtypes = [P, ctypes.c_int, P, ctypes.c_int]
    lib.tff_set_bc.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_int]
    lib.tff_init_rk_step.argtypes = [P]
    lib.tff_rk3_step.argtypes = [P, ctypes.c_double, ctypes.c_double, ctypes.c_double]
    lib.tff_clean_epar.argtypes = [P]
    lib.tff_exchange_fields.argtypes = [P, ctypes.c_int]
    lib.tff_user_bc.argtypes = [P, ctypes.c_int]
    lib.tff_step.argtypes = [P, ctypes.c_int]
    lib.tff_run.argtypes = [P, ctypes.c_int, ctypes.c_int]
    lib.tff_reduce.argtypes = [P, ctypes.POINTER(ctypes.c_double)]
    lib.tff_output_dims.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    lib.tff_output_fields.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p), ctypes.c_int]
    lib.tff_output_fields_noalloc.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p), ctypes.c_int]
    lib.tff_output_samples.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p), ctypes.c_int]
    lib.tff_init_fields.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_int]
    lib.tff_init_spectrum.argtypes = [P, P, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    lib.tff_set_phase_timing.argtypes = [P, ctypes.c_int]
    lib.tff_phase_times.argtypes = [P, ctypes.POINTER(ctypes.c_double)]
    lib.tff_get_size.argtypes = [P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    lib.tff_sync.argtypes = [P]
    lib.tff_event_record.argtypes = [P, ctypes.c_int]
    lib.tff_event_elapsed.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_float)]
    lib.tff_set_kernel_timing.argtypes = [P, ctypes.c_int]
    lib.tff_stage_kernel_time.argtypes = [P, ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(ctypes.c_long)]
    lib.tff_batch_create.argtypes = [ctypes.POINTER(Instance), ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, ctypes.POINTER(P)]
    lib.tff_batch_destroy.argtypes = [P]
    lib.tff_batch_destroy.restype = None
    lib.tff_batch_last_error.argtypes = [P]
    lib.tff_batch_last_error.restype = ctypes.c_char_p
    lib.tff_batch_instance.argtypes = [P, ctypes.c_int]
    lib.tff_batch_instance.restype = P
    lib.tff_batch_run.argtypes = [P, ctypes.c_int, ctypes.c_int]
    lib.tff_batch_get_tuning.argtypes = [P] + [ctypes.POINTER(ctypes.c_int)] * 4
    lib.tff_set_path.argtypes = [P, ctypes.c_int]
    lib.tff_set_tuning.argtypes = [P, ctypes.c_int, ctypes.c_int]
    lib.tff_get_tuning.argtypes = [P] + [ctypes.POINTER(ctypes.c_int)] * 4
    lib.tff_get_layout.argtypes = [P, ctypes.POINTER(ctypes.c_longlong)]
    for name in header_functions():
        if name not in ("tff_destroy", "tff_last_error", "tff_batch_destroy",
                        "tff_batch_last_error", "tff_batch_instance"):
            getattr(lib, name).restype = ctypes.c_int
    _lib = lib
    return lib


def rk_coefficients(precision: int) -> list[tuple[float, float, float]]:
    """Stage coefficients of src/tristanmainloop.F90:51,62,73 as the
    working precision rounds them (default REAL literals for FP32)."""
    d = np.float64 if precision in (FP64, FP32_BULK) else np.float32
    one, two, three, four = d(1), d(2), d(3), d(4)
    return [(float(one), 0.0, float(one)),
            (float(three / four), float(one / four), float(one / four)),
            (float(one / three), float(two / three), float(two / three))]


class Meshblock:
    """One meshblock of E, B (and B0) resident in HBM."""

    def __init__(self, n, boundary=(1, 1, 1), cc=0.45, precision="fp64", absorb_buff=20,
                 absorb_fixed=False, device=0):
        self.lib = load_library()
        nx, ny, nz = (n, n, n) if np.isscalar(n) else tuple(int(v) for v in n)
        self.precision = _PREC_NAMES[precision] if isinstance(precision, str) else int(precision)
        cfg = Config(nx, ny, nz, (ctypes.c_int * 3)(*boundary), int(absorb_buff),
                     int(bool(absorb_fixed)), self.precision, int(device), float(cc))
        h = ctypes.c_void_p()
        rc = self.lib.tff_create(ctypes.byref(cfg), ctypes.byref(h))
        if rc != 0:
            raise TffError(f"tff_create failed with code {rc}")
        self.h = h
        self.n = (nx, ny, nz)
        self.cc = cc
        self.boundary = tuple(boundary)
        self.shape = (nz + 2 * NG, ny + 2 * NG, nx + 2 * NG)
        self.storage = np.float64 if self.precision == FP64 else np.float32

    @classmethod
    def borrow(cls, lib, h, n, boundary, cc, precision):
        """A Meshblock view of a handle owned elsewhere (a batch instance)."""
        self = cls.__new__(cls)
        self.lib = lib
        self.h = ctypes.c_void_p(h)
        self.n = tuple(n)
        self.cc = cc
        self.boundary = tuple(boundary)
        self.precision = precision
        nx, ny, nz = self.n
        self.shape = (nz + 2 * NG, ny + 2 * NG, nx + 2 * NG)
        self.storage = np.float64 if precision == FP64 else np.float32
        self._owner = False
        return self

    def close(self):
        if getattr(self, "h", None) and getattr(self, "_owner", True):
            self.lib.tff_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int, what: str):
        if rc != 0:
            msg = self.lib.tff_last_error(self.h)
            raise TffError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    # ---- data movement (userInitFields / output access to m_fields) ------
    def upload(self, name: str, arr: np.ndarray):
        a = np.asarray(arr)
        dt = 1 if a.dtype == np.float32 else 0
        a = np.ascontiguousarray(a, dtype=np.float32 if dt else np.float64)
        if a.shape != self.shape:
            raise ValueError(f"{name}: expected shape {self.shape}, got {a.shape}")
        self._check(self.lib.tff_upload(self.h, FIELDS[name], a.ctypes.data, dt), f"upload {name}")

    def download(self, name: str, dtype=np.float64) -> np.ndarray:
        out = np.empty(self.shape, dtype=dtype)
        dt = 1 if out.dtype == np.float32 else 0
        self._check(self.lib.tff_download(self.h, FIELDS[name], out.ctypes.data, dt),
                    f"download {name}")
        return out

    def set_fields(self, **arrays):
        for k, v in arrays.items():
            self.upload(k, v)

    def get(self, *names, dtype=np.float64):
        return {n: self.download(n, dtype) for n in names}

    # ---- the reference's procedures ---------------------------------------
    def set_bc(self, kind: int, params=()):
        p = (ctypes.c_double * 16)(*[float(x) for x in params])
        self._check(self.lib.tff_set_bc(self.h, kind, p, len(params)), "tff_set_bc")

    def initRKstep(self):
        self._check(self.lib.tff_init_rk_step(self.h), "initRKstep")

    def rk3Step(self, c1, c2, c3):
        self._check(self.lib.tff_rk3_step(self.h, c1, c2, c3), "rk3Step")

    def cleanEpar(self):
        self._check(self.lib.tff_clean_epar(self.h), "cleanEpar")

    def exchangeFields(self, exchangeB0: bool = False):
        self._check(self.lib.tff_exchange_fields(self.h, int(exchangeB0)), "exchangeFields")

    def userFieldBoundaryConditions(self, step: int):
        self._check(self.lib.tff_user_bc(self.h, int(step)), "userFieldBoundaryConditions")

    def preloop(self):
        """tristanmainloop.F90:40-42."""
        self.exchangeFields(exchangeB0=True)
        self.userFieldBoundaryConditions(0)

    def step(self, timestep: int):
        self._check(self.lib.tff_step(self.h, int(timestep)), "tff_step")

    def run(self, first_step: int, nsteps: int):
        self._check(self.lib.tff_run(self.h, int(first_step), int(nsteps)), "tff_run")

    def reduce(self) -> list[float]:
        out = (ctypes.c_double * 4)()
        self._check(self.lib.tff_reduce(self.h, out), "tff_reduce")
        return list(out)

    def output_fields(self, istep: int, dtype=np.float32, noalloc: bool = False) -> dict:
        """writeFields_hdf5's field datasets (output_istep >= 2): ex ey ez
        bx by bz node interpolants at every istep-th node, each shaped
        (k, j, i).  noalloc: through tff_output_fields_noalloc (no device
        allocation; outputs up to TFF_NOALLOC_SCRATCH_BYTES)."""
        d = (ctypes.c_int * 3)()
        self._check(self.lib.tff_output_dims(self.h, int(istep), d), "tff_output_dims")
        arrs = [np.empty((d[2], d[1], d[0]), dtype=dtype) for _ in range(6)]
        ptrs = (ctypes.c_void_p * 6)(*[a.ctypes.data for a in arrs])
        dt = 1 if np.dtype(dtype) == np.float32 else 0
        fn = "tff_output_fields_noalloc" if noalloc else "tff_output_fields"
        self._check(getattr(self.lib, fn)(self.h, int(istep), ptrs, dt), fn)
        return dict(zip(("ex", "ey", "ez", "bx", "by", "bz"), arrs))

    def output_samples(self, istep: int, dtype=np.float32) -> dict:
        """Raw ex..bz and rho at the output nodes (tff_output_samples)."""
        d = (ctypes.c_int * 3)()
        self._check(self.lib.tff_output_dims(self.h, int(istep), d), "tff_output_dims")
        arrs = [np.empty((d[2], d[1], d[0]), dtype=dtype) for _ in range(7)]
        ptrs = (ctypes.c_void_p * 7)(*[a.ctypes.data for a in arrs])
        dt = 1 if np.dtype(dtype) == np.float32 else 0
        self._check(self.lib.tff_output_samples(self.h, int(istep), ptrs, dt), "tff_output_samples")
        return dict(zip(("ex", "ey", "ez", "bx", "by", "bz", "rho"), arrs))

    def init_fields(self, kind: int, params):
        """tff_init_fields: a user file's initial condition set in HBM."""
        p = (ctypes.c_double * 16)(*[float(x) for x in params])
        self._check(self.lib.tff_init_fields(self.h, int(kind), p, len(params)), "tff_init_fields")

    def init_spectrum(self, modes: np.ndarray, background):
        """tff_init_spectrum: a periodic field from Fourier modes (rows of 15)."""
        m = np.ascontiguousarray(modes, dtype=np.float64).reshape(-1, 15)
        bg = (ctypes.c_double * 6)(*[float(x) for x in background])
        self._check(self.lib.tff_init_spectrum(self.h, m.ctypes.data, m.shape[0], bg), "tff_init_spectrum")

    def phase_timing(self, on: bool):
        self._check(self.lib.tff_set_phase_timing(self.h, int(on)), "tff_set_phase_timing")

    def phase_times(self) -> list[float]:
        """GPU ms of the solver, hook and exchange phases since the last call."""
        ms = (ctypes.c_double * 3)()
        self._check(self.lib.tff_phase_times(self.h, ms), "tff_phase_times")
        return list(ms)

    def sync(self):
        self._check(self.lib.tff_sync(self.h), "tff_sync")

    def set_path(self, fused: bool):
        """fused=True: one fused kernel per stage (default); False: the
        per-procedure kernels in the reference's order."""
        self._check(self.lib.tff_set_path(self.h, int(bool(fused))), "tff_set_path")

    def set_tuning(self, tile: int, kchunk: int = 0):
        self._check(self.lib.tff_set_tuning(self.h, int(tile), int(kchunk)), "tff_set_tuning")

    def get_tuning(self) -> dict:
        v = [ctypes.c_int() for _ in range(4)]
        self._check(self.lib.tff_get_tuning(self.h, *[ctypes.byref(x) for x in v]), "tff_get_tuning")
        return {"tile": v[0].value, "tx": v[1].value, "ty": v[2].value, "kchunk": v[3].value}

    def layout(self) -> dict:
        v = (ctypes.c_longlong * 5)()
        self._check(self.lib.tff_get_layout(self.h, v), "tff_get_layout")
        return {"px": v[0], "plane": v[1], "org": v[2], "nalloc": v[3], "esz": v[4]}

    @property
    def nalloc(self) -> int:
        return self.layout()["nalloc"]

    # ---- timing ---------------------------------------------------------
    def event(self, slot: int):
        self._check(self.lib.tff_event_record(self.h, slot), "tff_event_record")

    def elapsed_ms(self, a: int, b: int) -> float:
        ms = ctypes.c_float()
        self._check(self.lib.tff_event_elapsed(self.h, a, b, ctypes.byref(ms)), "tff_event_elapsed")
        return ms.value

    def kernel_timing(self, on: bool):
        self._check(self.lib.tff_set_kernel_timing(self.h, int(on)), "tff_set_kernel_timing")

    def stage_kernel_time(self):
        ms = ctypes.c_double()
        cnt = ctypes.c_long()
        self._check(self.lib.tff_stage_kernel_time(self.h, ctypes.byref(ms), ctypes.byref(cnt)),
                    "tff_stage_kernel_time")
        return ms.value, cnt.value


class Batch:
    """Independent meshblocks advanced together (tff_batch_*): one persistent
    launch of the fused stage kernel per RK stage for all of them."""

    def __init__(self, specs, precision="fp32_bulk", device=0):
        """specs: iterable of (n, boundary, cc, bc_kind, bc_params)."""
        self.lib = load_library()
        self.precision = _PREC_NAMES[precision] if isinstance(precision, str) else int(precision)
        specs = list(specs)
        arr = (Instance * len(specs))()
        for q, (n, boundary, cc, kind, params) in enumerate(specs):
            nx, ny, nz = (n, n, n) if np.isscalar(n) else tuple(n)
            arr[q].nx, arr[q].ny, arr[q].nz = nx, ny, nz
            arr[q].boundary = (ctypes.c_int * 3)(*boundary)
            arr[q].cc = float(cc)
            arr[q].bc_kind = int(kind)
            arr[q].nparams = len(params)
            arr[q].params = (ctypes.c_double * 16)(*[float(x) for x in params])
        h = ctypes.c_void_p()
        rc = self.lib.tff_batch_create(arr, len(specs), self.precision, int(device), ctypes.byref(h))
        if rc != 0:
            raise TffError(f"tff_batch_create failed with code {rc}")
        self.h = h
        self.members = []
        for q in range(len(specs)):
            ih = self.lib.tff_batch_instance(self.h, q)
            self.members.append(Meshblock.borrow(self.lib, ih, (arr[q].nx, arr[q].ny, arr[q].nz),
                                                 tuple(specs[q][1]), specs[q][2], self.precision))

    def __len__(self):
        return len(self.members)

    def __getitem__(self, q) -> Meshblock:
        return self.members[q]

    def run(self, first_step: int, nsteps: int):
        rc = self.lib.tff_batch_run(self.h, int(first_step), int(nsteps))
        if rc != 0:
            raise TffError(f"tff_batch_run failed ({rc}): {self.lib.tff_batch_last_error(self.h).decode()}")

    def get_tuning(self) -> dict:
        v = [ctypes.c_int() for _ in range(4)]
        self.lib.tff_batch_get_tuning(self.h, *[ctypes.byref(x) for x in v])
        return {"tile": v[0].value, "tx": v[1].value, "ty": v[2].value, "kchunk": v[3].value}

    def close(self):
        if getattr(self, "h", None):
            for m in self.members:
                m.h = None
            self.lib.tff_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
