"""Initial conditions generated in HBM (tff_init_fields, tff_init_spectrum)
against the host builders that the oracle runs from: every field of the
ghost-inclusive box, B0 included, for each user file's initial condition
and the seeded Alfvenic spectrum, in FP64 and in FP32 storage; then a few
steps from each start agree; malformed requests are refused."""
import numpy as np
import pytest

from helpers import FIELDS6, field_errors, normwise_max_rel

pytestmark = pytest.mark.gpu
pytff = pytest.importorskip("pytff")
problems = pytest.importorskip("problems")

ALL9 = FIELDS6 + ("b0x", "b0y", "b0z")
CASES = {
    "loop": lambda: problems.loop(30, twist_angle=1.0, steps=12),
    "loop_driven": lambda: problems.loop(28, twist_angle=1.0, steps=12, driver=problems.LOOP_DRIVEN),
    "psr": lambda: problems.psr(34, radius=5.0, r_lc=12.0, absorb_buff=6, angle_deg=30.0, periods=0.2),
    "uniform": lambda: problems.uniform((26, 22, 20), L0=4.0, steps=12),
    "alfven": lambda: problems.alfven((26, 24, 22), seed=3, kmax=5, amp=0.2),
}


def _mb(p, precision):
    return pytff.Meshblock(p.n, boundary=p.boundary, cc=p.cc, precision=precision, absorb_buff=p.absorb_buff,
                           absorb_fixed=p.absorb_fixed)


@pytest.mark.parametrize("precision", ["fp64", "fp32_bulk"])
@pytest.mark.parametrize("name", list(CASES))
def test_device_initial_condition_equals_host(name, precision):
    p = CASES[name]()
    print(f"case {name} {precision} n={p.n}", flush=True)
    host, dev = _mb(p, precision), _mb(p, precision)
    problems.load(host, p)
    problems.load(dev, p, device=True)
    dt = np.float64 if precision == "fp64" else np.float32
    for f in ALL9:
        a, b = host.download(f, dt), dev.download(f, dt)
        if precision == "fp64":
            # FP64 kernels (sqrt, sincos, division order) against numpy
            assert normwise_max_rel(b, a) <= 1e-14, (f, normwise_max_rel(b, a))
        else:
            # each value rounded once from FP64 on both sides: equal up to
            # a rounding boundary that the FP64 difference straddles
            ulp = np.spacing(np.abs(a).astype(np.float32))
            assert (np.abs(b - a) <= ulp).all(), f
    # a few steps from each start
    for g in (host, dev):
        g.preloop()
        g.run(0, 4)
    # E and B as vector fields (a component that is zero in exact
    # arithmetic holds only roundoff)
    err = field_errors(dev.download, host.download)
    print(name, precision, "after 4 steps", err)
    tol = 1e-12 if precision == "fp64" else 1e-5
    assert max(err.values()) <= tol


def test_spectrum_rejects_non_integer_wave_numbers():
    p = problems.alfven(16, seed=1, kmax=3)
    g = _mb(p, "fp64")
    bad = p.init[1].copy()
    bad[3, 1] += 0.5
    with pytest.raises(pytff.TffError, match="not an integer"):
        g.init_spectrum(bad, p.init[2])
    with pytest.raises(pytff.TffError):
        g.init_fields(problems.IC_PSR, (1.0, 2.0))  # six parameters needed
    with pytest.raises(pytff.TffError):
        g.init_fields(7, ())
