"""Mixed-precision forms against the oracle (the FP64 CPU restatement of the
reference KT step):

- GVH_MIXED: FP32 stencil, limiter, flux, sources, pi/Pi storage and the 12
  half-site root-finds (tabulated and conformal EoS); FP64 T^{tau mu}
  accumulation and the cell's own root-find;
- GVH_MIXED_F64_ROOTS: the same with every root-find and its EoS lookups in
  FP64 (the split of FP32 flux, limiter, source and stencil arithmetic and
  FP32 pi, Pi storage against FP64 accumulation and root-finds).

Metric: per-cell relative error with a floor of MIXED_FLOOR (1%) of the
field's scale -- components that cross zero (u^x, u^y, u^eta, pi) are
resolved to FP32 precision in absolute terms, so below 1% of their maximum
the bar becomes absolute (1e-6 of the maximum).  Bound: 1e-4 for e and u^mu
(the tolerance stated for the mixed form), 1e-3 for pi^{mu nu}.  The
Lorentz-norm error max |du^mu| / u^tau is reported beside it."""
import numpy as np
import pytest

import gpuvh
from parity import Case, PI_FIELDS, U_FIELDS, floor_scales, hotspot_state, rel_err

pytestmark = pytest.mark.gpu

MIXED_TOL = 1e-4
MIXED_FLOOR = 1e-2


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


_ORACLE = {}


def oracle_fields(name, case, state, nsteps, fields):
    """The oracle's fields after nsteps, computed once per case."""
    if name not in _ORACLE:
        orc = case.oracle()
        orc.set_primitives(case.t0, *state)
        orc.steps(case.t0, 0, nsteps)
        _ORACLE[name] = {f: orc.get(f) for f in fields}
        orc.close()
    return _ORACLE[name]


@pytest.mark.parametrize("precision", [gpuvh.MIXED, gpuvh.MIXED_F64_ROOTS],
                         ids=["mixed", "mixed_f64_roots"])
@pytest.mark.parametrize("variant", [0, 3, 5])
@pytest.mark.parametrize("name,case,nsteps", CASES, ids=[c[0] for c in CASES])
def test_mixed_against_oracle(name, case, nsteps, variant, precision, monkeypatch):
    """After nsteps the mixed forms' e and u^mu agree with the oracle within
    1e-4 per cell; pi^{mu nu} within 1e-3 of its scale."""
    say("mixed", name, "variant", variant, "precision", precision, "steps", nsteps)
    state = hotspot_state(case, seed=5, n_hot=10, amp=6.0, u_noise=0.02)
    fields = [gpuvh.E] + U_FIELDS + PI_FIELDS
    refs = oracle_fields(name, case, state, nsteps, fields)
    monkeypatch.setenv("GVH_VARIANT", str(variant))
    mix = run_ctx(case, state, nsteps, precision)
    tau = case.t0 + nsteps * case.dt
    errs = {}
    scales = floor_scales(refs, tau)
    for f in fields:
        errs[gpuvh.FIELD_NAMES[f]] = rel_err(mix.get(f), refs[f], floor_frac=MIXED_FLOOR,
                                             scale=scales[f])
    du = max(np.abs(mix.get(f) - refs[f]).max() for f in U_FIELDS[1:])
    lorentz = du / np.abs(refs[gpuvh.UT]).min()
    say("  errors", {k: "%.2e" % v for k, v in errs.items()}, "|du|/u^tau %.2e" % lorentz)
    assert lorentz <= MIXED_TOL
    for f in [gpuvh.E] + U_FIELDS:
        assert errs[gpuvh.FIELD_NAMES[f]] <= MIXED_TOL, gpuvh.FIELD_NAMES[f]
    for f in PI_FIELDS:
        assert errs[gpuvh.FIELD_NAMES[f]] <= 10 * MIXED_TOL, gpuvh.FIELD_NAMES[f]
    # pi fields come back from FP32 storage
    pixx = mix.get(gpuvh.PIXX)
    assert np.array_equal(pixx, pixx.astype(np.float32).astype(np.float64))
    d = mix.diagnostics()
    assert d["e_nan"] == 0 and d["newton_not_converged"] == 0
    mix.close()


def test_mixed_freezeout_normals():
    """Freeze-out over three windows of a 3+1D table-EoS event, iso-energy
    density halfway between the event's minimum and maximum: the mixed form
    finds the FP64 form's cells in the same order, the normals agree within
    1e-4 of the largest normal, and e and u^mu interpolated on the surface
    (the rows' columns 8-11) within 1e-4 of their largest value."""
    case = Case(25, 25, 11, 0.1, 0.1, 0.01, 0.5, 0.16, bulk=True, eos=gpuvh.EOS_TABLE,
                zeta_tpeak=0.16 / gpuvh.HBARC)
    state = hotspot_state(case, seed=9, n_hot=12, amp=8.0)
    out = {}
    for prec in (gpuvh.FP64, gpuvh.MIXED):
        ctx = case.context(prec)
        ctx.set_initial_state(case.t0, *state)
        e_fo = 0.5 * (state[0].min() + state[0].max())
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
    # the same cells in the same order: positions within 1e-4 of a cell,
    # normals within 1e-4 of the largest normal
    assert len(a) == len(b)
    pos = np.abs(a[:, :4] - b[:, :4]).max()
    err = np.abs(a[:, 4:8] - b[:, 4:8]).max() / np.abs(a[:, 4:8]).max()
    say("  max position difference %.2e, normal error / max normal %.2e" % (pos, err))
    assert pos <= 1e-4 * 0.1
    assert err <= MIXED_TOL
    # u^x, u^y, u^eta, e on the surface
    for col, name in ((8, "ux"), (9, "uy"), (10, "un"), (11, "e")):
        scale = np.abs(a[:, col]).max()
        d = np.abs(a[:, col] - b[:, col]).max() / scale
        say("  surface %s: max |mixed - fp64| / max %.2e" % (name, d))
        assert d <= MIXED_TOL, name
