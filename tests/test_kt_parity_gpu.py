"""Per-cell parity of the HIP KT step against the CPU FP64 oracle.

Bar (FP64 form): after the stated number of steps every cell of e, u^mu and
pi^{mu nu} (and Pi when evolved) agrees with the oracle to 1e-10 relative.
The denominator is floored at 1e-4 of the tensor's largest component in the
orthonormal frame (tests/parity.py FLOOR_FRAC, floor_scales): a component
much smaller than its tensor is a sum of O(scale) terms that cancel, so its
rounding noise scales with the tensor, not with itself; there the bar is an
absolute 1e-14 x scale.
"""
import numpy as np
import pytest

import gpuvh
import oracle_lib as O
from gpuvh import ic
from parity import PI_FIELDS, U_FIELDS, Case, compare, hotspot_state, rel_err, run_both

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _check(errs, tol=TOL, report=None):
    bad = {k: v for k, v in errs.items() if not v <= tol}
    if report:
        for row in report:
            print("  %-5s rel %.2e at %s ref %.3e diff %.2e scale %.3e" % row)
    assert not bad, "relative errors above %g: %s (all: %s)" % (tol, bad, errs)


def test_c1_conformal_shear_200_steps():
    """61x61x1, dx=0.1, tau0=0.5, dtau=0.01, 200 steps, p=e/3, eta/s=0.2,
    Gaussian + 10 seeded hotspots, u at rest, pi=0 (seed 1)."""
    case = Case(61, 61, 1, 0.1, 0.1, 0.01, 0.5, 0.2, bulk=False, eos=gpuvh.EOS_CONFORMAL)
    e = ic.c1_event(seed=1)
    ut, ux, uy, un, pi10, Pi = ic.at_rest(e.shape)
    state = (e, e / 3.0, ut, ux, uy, un, pi10, Pi)
    orc, ctx = run_both(case, state, 200)
    rep = []
    errs = compare(orc, ctx, [gpuvh.E] + U_FIELDS + PI_FIELDS, tau=0.5 + 200 * 0.01, report=rep)
    print("C1 errors", errs)
    _check(errs, report=rep)
    assert ctx.diagnostics()["e_nan"] == 0


@pytest.mark.parametrize("eos", [gpuvh.EOS_WB, gpuvh.EOS_TABLE])
def test_3d_shear_bulk_small(eos):
    """21x21x11 3+1D event with shear and bulk, 20 steps, both non-conformal
    equations of state (Newton root-find path)."""
    zeta = 0.16 / gpuvh.HBARC if eos == gpuvh.EOS_TABLE else 1.01355
    case = Case(21, 21, 11, 0.1, 0.1, 0.01, 0.5, 0.16, bulk=True, eos=eos, zeta_tpeak=zeta)
    state = hotspot_state(case, seed=7, n_hot=12, amp=6.0)
    orc, ctx = run_both(case, state, 20)
    rep = []
    errs = compare(orc, ctx, [gpuvh.E] + U_FIELDS + PI_FIELDS + [gpuvh.PI], tau=0.5 + 20 * 0.01,
                   report=rep)
    print("3D errors", eos, errs)
    _check(errs, report=rep)


def test_regulation_active_with_flow():
    """Regulation below T_reg = 0.25 GeV with random pi and a flowing initial
    state (exercises regulateDissipativeCurrents' rescaling)."""
    case = Case(17, 15, 5, 0.1, 0.1, 0.01, 0.6, 0.2, bulk=True, eos=gpuvh.EOS_WB,
                T_reg_GeV=0.25)
    state = hotspot_state(case, seed=11, n_hot=8, amp=3.0, pi_noise=0.05, u_noise=0.05)
    orc, ctx = run_both(case, state, 10)
    rep = []
    errs = compare(orc, ctx, [gpuvh.E] + U_FIELDS + PI_FIELDS + [gpuvh.PI], tau=0.6 + 10 * 0.01,
                   report=rep)
    print("regulation errors", errs)
    _check(errs, report=rep)
    # the rescaling must actually have happened somewhere
    e = orc.get(O.E)
    assert np.any(e < orc_e_reg(case)), "no cell below the regulation energy density"


def orc_e_reg(case):
    return O.lib().kto_e_of_T(case.oracle().P, case.T_reg_GeV / gpuvh.HBARC)


@pytest.mark.parametrize("variant", [3, 5])
def test_2d_table_eos_with_flow_and_pi(variant, monkeypatch):
    """2+1D, tabulated EoS, bulk, flowing state with initial pi: 31x27x1,
    30 steps (ragged nx != ny); the default arrangement and the four-row
    blocks with the y stencil in LDS."""
    monkeypatch.setenv("GVH_VARIANT", str(variant))
    case = Case(31, 27, 1, 0.1, 0.1, 0.01, 0.5, 0.16, bulk=True, eos=gpuvh.EOS_TABLE,
                zeta_tpeak=0.16 / gpuvh.HBARC)
    state = hotspot_state(case, seed=3, n_hot=10, amp=6.0, pi_noise=0.02, u_noise=0.1)
    orc, ctx = run_both(case, state, 30)
    rep = []
    errs = compare(orc, ctx, [gpuvh.E] + U_FIELDS + PI_FIELDS + [gpuvh.PI], tau=0.5 + 30 * 0.01,
                   report=rep)
    print("2D table errors", errs)
    _check(errs, report=rep)


def test_ghost_cells_match_oracle():
    """Full ghosted grid (faces) after 3 steps matches the oracle's ghost fill."""
    case = Case(9, 7, 5, 0.1, 0.1, 0.01, 0.5, 0.2, bulk=True, eos=gpuvh.EOS_WB)
    state = hotspot_state(case, seed=5, n_hot=4, amp=4.0, u_noise=0.05)
    orc, ctx = run_both(case, state, 3)
    for f in (gpuvh.E, gpuvh.TTT, gpuvh.PIXX, gpuvh.UX):
        a = ctx.get(f, ghosted=True)
        b = orc.get(f, ghosted=True)
        # edges/corners are never written by either side; compare face cells
        nx, ny, nz = case.nx, case.ny, case.nz
        m = np.zeros(a.shape, bool)
        m[:, 2:ny + 2, 2:nz + 2] = True
        m[2:nx + 2, :, 2:nz + 2] = True
        m[2:nx + 2, 2:ny + 2, :] = True
        assert rel_err(a[m], b[m]) <= TOL, gpuvh.FIELD_NAMES[f]


def test_count_above_matches_numpy():
    case = Case(21, 19, 7, 0.1, 0.1, 0.01, 0.5, 0.2, bulk=False, eos=gpuvh.EOS_CONFORMAL)
    state = hotspot_state(case, seed=9)
    ctx = case.context()
    ctx.set_initial_state(case.t0, *state)
    e = ctx.get(gpuvh.E)
    for thr in (0.0, float(np.median(e)), float(e.max()), float(e.max()) + 1.0):
        assert ctx.count_above(thr) == int(np.sum(e > thr))


def test_field_roundtrip_and_errors():
    case = Case(5, 6, 3, 0.1, 0.1, 0.01, 0.5, 0.2, bulk=True, eos=gpuvh.EOS_WB)
    ctx = case.context()
    with pytest.raises(gpuvh.GvhError):
        ctx.step(0.5)  # no initial state yet
    state = hotspot_state(case, seed=2, n_hot=3)
    ctx.set_initial_state(case.t0, *state)
    g = np.arange((5 + 4) * (6 + 4) * (3 + 4), dtype=np.float64).reshape((9, 10, 7), order="F")
    ctx.set_field(gpuvh.PIXY, g)
    assert np.array_equal(ctx.get(gpuvh.PIXY, ghosted=True), g)
    assert np.array_equal(ctx.get(gpuvh.PIXY), g[2:7, 2:8, 2:5])
    with pytest.raises(gpuvh.GvhError):
        ctx.get(99)


@pytest.mark.parametrize("variant,strip", [(0, 16), (1, 4), (2, 8), (0, 1), (2, 100), (3, 16), (3, 5),
                                           (4, 16), (5, 16), (5, 8), (5, 4), (5, 5)])
def test_kernel_arrangements_agree(monkeypatch, variant, strip):
    """Every stage-kernel arrangement and traversal strip gives the oracle's
    result (23x21x9, tabulated EoS, shear+bulk, 8 steps)."""
    monkeypatch.setenv("GVH_VARIANT", str(variant))
    monkeypatch.setenv("GVH_STRIP", str(strip))
    case = Case(23, 21, 9, 0.1, 0.1, 0.01, 0.5, 0.16, bulk=True, eos=gpuvh.EOS_TABLE,
                zeta_tpeak=0.16 / gpuvh.HBARC)
    state = hotspot_state(case, seed=21, n_hot=8, amp=6.0, u_noise=0.03)
    orc, ctx = run_both(case, state, 8)
    errs = compare(orc, ctx, [gpuvh.E] + U_FIELDS + PI_FIELDS + [gpuvh.PI], tau=0.5 + 8 * 0.01)
    _check(errs)


def limiter_edge_state(case):
    """Plateaus (exactly equal neighbours), a jump, values alternating in
    sign or size from cell to cell, exact zeros and negative zeros in u,
    pi and Pi: every sign class of the limiter's differences, including
    a = -0 (-0 - +0), where the kernels' sign test differs from the
    reference's x < 0 only in the sign of a zero slope."""
    nx, ny, nz = case.nx, case.ny, case.nz
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    nz0 = np.copysign(0.0, -1.0)
    e = np.full(i.shape, 2.0)
    e[i < 6] = 6.0
    band = (j >= 8) & (j < 12) & (i >= 6)
    e[band] = np.where(i[band] % 2 == 0, 3.0, 3.5)
    ux = np.where(i < 9, 0.1, np.where(i >= 12, -0.1, 0.0))
    ux[(i >= 9) & (i < 12) & (j < 5)] = nz0
    uy = np.where(k % 2 == 0, 0.05 * (-1.0) ** j, 0.0)
    un = np.zeros(i.shape)
    un[:, :, 3] = nz0
    ut = np.sqrt(1 + ux ** 2 + uy ** 2 + (case.t0 * un) ** 2)
    pi10 = [np.zeros(i.shape) for _ in range(10)]
    pi10[4] = 0.01 * e * (-1.0) ** (i + j)                       # xx, alternating
    pi10[5] = np.where(j % 3 == 0, 0.005 * e, 0.0)               # xy, zeros ...
    pi10[5][i < 4] = nz0                                         # ... and -0
    pi10[7] = 0.01 * e                                           # yy, plateaus
    Pi = -0.002 * e * (-1.0) ** k
    return e, case.pressure(e), ut, ux, uy, un, pi10, Pi


def test_limiter_edge_states():
    """FP64: the edge-case state (plateaus, a jump, sign alternation, +-0)
    after 5 steps agrees with the oracle to 1e-10 per cell.  Mixed: e and
    u^mu within 1e-4 per cell (1% floor), pi within 1e-3."""
    case = Case(19, 17, 7, 0.1, 0.1, 0.01, 0.5, 0.16, bulk=True, eos=gpuvh.EOS_TABLE,
                zeta_tpeak=0.16 / gpuvh.HBARC)
    state = limiter_edge_state(case)
    fields = [gpuvh.E] + U_FIELDS + PI_FIELDS + [gpuvh.PI]
    orc, ctx = run_both(case, state, 5)
    rep = []
    errs = compare(orc, ctx, fields, tau=0.5 + 5 * 0.01, report=rep)
    print("limiter edge state, fp64", errs)
    _check(errs, report=rep)
    assert ctx.diagnostics()["e_nan"] == 0
    ctx.close()
    mix = case.context(gpuvh.MIXED)
    mix.set_initial_state(case.t0, *state)
    mix.steps(case.t0, 0, 5)
    mix.synchronize()
    from parity import floor_scales
    refs = {f: orc.get(f) for f in fields}
    scales = floor_scales(refs, 0.5 + 5 * 0.01)
    merr = {gpuvh.FIELD_NAMES[f]: rel_err(mix.get(f), refs[f], floor_frac=1e-2, scale=scales[f])
            for f in fields}
    print("limiter edge state, mixed", merr)
    for f in [gpuvh.E] + U_FIELDS:
        assert merr[gpuvh.FIELD_NAMES[f]] <= 1e-4, gpuvh.FIELD_NAMES[f]
    for f in PI_FIELDS + [gpuvh.PI]:
        assert merr[gpuvh.FIELD_NAMES[f]] <= 1e-3, gpuvh.FIELD_NAMES[f]
    mix.close()
    orc.close()


def test_full_size_benchmark_event():
    """The benchmark's own event at its full size: 201x201x101, dx = deta =
    0.1, tau0 = 0.5, dtau = 0.01, tabulated EoS, eta/s = 0.16, zeta/s peak
    0.16 GeV, 200 seeded hotspots (seed 2) times the eta plateau, u at rest,
    pi = Pi = 0.  After 4 KT steps every one of the 4.08 M cells agrees with
    the oracle to 1e-10, and the device counts no NaN e."""
    case = Case(201, 201, 101, 0.1, 0.1, 0.01, 0.5, 0.16, bulk=True, eos=gpuvh.EOS_TABLE,
                zeta_tpeak=0.16 / gpuvh.HBARC)
    e = ic.c2_event(201, 201, 101, 0.1, 0.1, 0.1, seed=2)
    ut, ux, uy, un, pi10, Pi = ic.at_rest(e.shape)
    state = (e, case.pressure(e), ut, ux, uy, un, pi10, Pi)
    print("full-size event: 4 steps on both sides", flush=True)
    orc, ctx = run_both(case, state, 4)
    print("full-size event: comparing", flush=True)
    rep = []
    errs = compare(orc, ctx, [gpuvh.E] + U_FIELDS + PI_FIELDS + [gpuvh.PI], tau=0.5 + 4 * 0.01,
                   report=rep)
    print("full-size errors", errs)
    _check(errs, report=rep)
    assert ctx.diagnostics()["e_nan"] == 0
    ctx.close()
    orc.close()
