"""Mixed-precision form (FP32 stencil, limiter, flux, sources and pi/Pi
storage; FP64 T^{tau mu} accumulation, root-finds and EoS) against the FP64
form and the oracle.  Tolerance: 1e-4 relative per cell with the parity
floor (1e-4 of the field's scale), the north-star bound for e and u^mu."""
import numpy as np
import pytest

import gpuvh
from parity import (FLOOR_FRAC, Case, PI_FIELDS, U_FIELDS, compare, floor_scales, hotspot_state,
                    rel_err)

pytestmark = pytest.mark.gpu

MIXED_TOL = 1e-4


def say(*a):
    print(*a, flush=True)


def run_ctx(case, state, nsteps, precision, variant=None, monkeypatch=None):
    if variant is not None:
        monkeypatch.setenv("GVH_VARIANT", str(variant))
    ctx = case.context(precision)
    ctx.set_initial_state(case.t0, *state)
    ctx.steps(case.t0, 0, nsteps)
    ctx.synchronize()
    return ctx


CASES = [
    ("3d-table-bulk", Case(31, 29, 15, 0.1, 0.1, 0.01, 0.5, 0.16, bulk=True, eos=gpuvh.EOS_TABLE,
                           zeta_tpeak=0.16 / gpuvh.HBARC), 40),
    ("2d-conformal", Case(61, 61, 1, 0.1, 0.1, 0.01, 0.5, 0.2, bulk=False,
                          eos=gpuvh.EOS_CONFORMAL), 100),
    ("3d-wb-bulk", Case(21, 21, 11, 0.1, 0.1, 0.01, 0.5, 0.2, bulk=True, eos=gpuvh.EOS_WB), 30),
]


@pytest.mark.parametrize("variant", [0, 3])
@pytest.mark.parametrize("name,case,nsteps", CASES, ids=[c[0] for c in CASES])
def test_mixed_against_fp64(name, case, nsteps, variant, monkeypatch):
    """After nsteps the mixed form's e and u^mu agree with the FP64 form (itself
    the oracle's to 1e-10, and checked against it here) within 1e-4 per cell;
    pi^{mu nu} within 1e-3 of its scale."""
    say("mixed", name, "variant", variant, "steps", nsteps)
    state = hotspot_state(case, seed=5, n_hot=10, amp=6.0, u_noise=0.02)
    monkeypatch.setenv("GVH_VARIANT", str(variant))
    ref = run_ctx(case, state, nsteps, gpuvh.FP64)
    mix = run_ctx(case, state, nsteps, gpuvh.MIXED)
    tau = case.t0 + nsteps * case.dt
    errs = {}
    fields = [gpuvh.E] + U_FIELDS + PI_FIELDS
    refs = {f: ref.get(f) for f in fields}
    scales = floor_scales(refs, tau)
    for f in fields:
        errs[gpuvh.FIELD_NAMES[f]] = rel_err(mix.get(f), refs[f], scale=scales[f])
    say("  errors", {k: "%.2e" % v for k, v in errs.items()})
    for f in [gpuvh.E] + U_FIELDS:
        assert errs[gpuvh.FIELD_NAMES[f]] <= MIXED_TOL, gpuvh.FIELD_NAMES[f]
    for f in PI_FIELDS:
        assert errs[gpuvh.FIELD_NAMES[f]] <= 10 * MIXED_TOL, gpuvh.FIELD_NAMES[f]
    # pi fields come back from FP32 storage
    pixx = mix.get(gpuvh.PIXX)
    assert np.array_equal(pixx, pixx.astype(np.float32).astype(np.float64))
    d = mix.diagnostics()
    assert d["e_nan"] == 0 and d["newton_not_converged"] == 0
    ref.close()
    mix.close()


def test_mixed_freezeout_normals():
    """Freeze-out at T_fo = 0.15 GeV over three windows of a 3+1D table-EoS
    event: cells found by the mixed form match the FP64 form's, and the
    normals of common elements agree within 1e-4 of the largest normal."""
    case = Case(25, 25, 11, 0.1, 0.1, 0.01, 0.5, 0.16, bulk=True, eos=gpuvh.EOS_TABLE,
                zeta_tpeak=0.16 / gpuvh.HBARC)
    state = hotspot_state(case, seed=9, n_hot=12, amp=8.0)
    out = {}
    for prec in (gpuvh.FP64, gpuvh.MIXED):
        ctx = case.context(prec)
        ctx.set_initial_state(case.t0, *state)
        e_fo = ctx.energy_density_of_T(0.15 / gpuvh.HBARC)
        ctx.freezeout_init(10, e_fo, 0)
        t = case.t0
        rows = []
        for n in range(1, 30):
            ctx.freezeout_record(n)
            if n % 10 == 9:
                rows.append(ctx.freezeout_find(n, t))
            ctx.step(t)
            t = case.t0 + n * case.dt
        out[prec] = np.concatenate(rows)
        ctx.close()
    a, b = out[gpuvh.FP64], out[gpuvh.MIXED]
    say("freeze-out rows fp64", len(a), "mixed", len(b))
    assert len(a) > 50
    # match elements by position (tau, x, y, eta rounded to 1e-6)
    key = lambda r: tuple(np.round(r[:4], 6))
    ka = {key(r): r for r in a}
    common = [(ka[key(r)], r) for r in b if key(r) in ka]
    assert len(common) >= 0.99 * len(a)
    na = np.array([c[0][4:8] for c in common])
    nb = np.array([c[1][4:8] for c in common])
    err = np.abs(na - nb).max() / np.abs(na).max()
    say("  max normal error / max normal %.2e" % err)
    assert err <= MIXED_TOL
