"""Validity diagnostics on the device: gvh_check_validity (the validity
kernel) and the regulation factor that the stage kernel records, against the
oracle's restatement of HydrodynamicValidity.cu (itself pinned in
test_validity_oracle.py).

Parity bar: per cell |x - x_ref| <= tol max(|x_ref|, floor x median|x_ref|)
per field, NaN where the oracle has NaN.  On identical input states (the
oracle's state copied onto the device) tol = 1e-11, floor 1e-6: the kernel
itself.  After steps on both sides tol = 1e-10, floor 1e-4, the bar of the
state itself: the states agree to ~1e-13, and the central differences
(a difference of neighbours over 2h) take that to ~1e-11.  KnTauPi =
theta zeta/s / (15 a^2 T) with a = 0.333333f - c_s^2 is conditioned as 1/a:
its bar is multiplied per cell by max(1, 1/(3|a|)) (a is ~1e-7 where the
table's c_s^2 crosses 0.333333f)."""
import os

import numpy as np
import pytest

import fo_ref as F
import gpuvh
import oracle_lib as O
from parity import Case, hotspot_state

pytestmark = pytest.mark.gpu

NV = gpuvh.VALIDITY_FIELDS


def say(*a):
    print(*a, flush=True)


def device_validity(ctx, fields=range(NV)):
    return np.stack([ctx.get_validity(f) for f in fields])


def bulk_conditioning(orc):
    """max(1, 1/(3|a|)) per cell, a = 0.333333f - c_s^2(e) of the oracle."""
    e = orc.get(O.E)
    cs2 = np.array([O.eos(orc.P, x)[1] for x in e.ravel()]).reshape(e.shape)
    a = float(np.float32(0.333333)) - cs2
    return np.maximum(1.0, 1.0 / (3.0 * np.maximum(np.abs(a), 1e-300)))


def field_errors(got, ref, floor, cond=None):
    """{name: max per-cell error / conditioning}; NaN cells must coincide."""
    out = {}
    for f in range(len(ref)):
        a, b = got[f], ref[f]
        nan_a, nan_b = np.isnan(a), np.isnan(b)
        assert np.array_equal(nan_a, nan_b), gpuvh.VALIDITY_NAMES[f]
        fin = ~nan_b
        nz = np.abs(b[fin])
        nz = nz[nz > 0]
        if nz.size == 0:
            out[gpuvh.VALIDITY_NAMES[f]] = float(np.max(np.abs(a[fin]))) if fin.any() else 0.0
            continue
        den = np.maximum(np.abs(b[fin]), floor * float(np.median(nz)))
        r = np.abs(a[fin] - b[fin]) / den
        if cond is not None and f == O.V_KN_TAUBULK:
            r = r / cond[fin]
        out[gpuvh.VALIDITY_NAMES[f]] = float(np.max(r))
    return out


CASES = [
    ("3d-table-bulk", Case(25, 23, 11, 0.1, 0.1, 0.01, 0.5, 0.16, bulk=True, eos=gpuvh.EOS_TABLE,
                           zeta_tpeak=0.16 / gpuvh.HBARC)),
    ("3d-wb-bulk-reg", Case(19, 17, 9, 0.1, 0.1, 0.01, 0.5, 0.2, bulk=True, eos=gpuvh.EOS_WB,
                            T_reg_GeV=0.45)),
    ("2d-conformal", Case(41, 37, 1, 0.1, 0.1, 0.01, 0.5, 0.2, bulk=False,
                          eos=gpuvh.EOS_CONFORMAL)),
]
IDS = [c[0] for c in CASES]


def event(case, seed):
    return hotspot_state(case, seed=seed, n_hot=10, amp=6.0, u_noise=0.03, pi_noise=0.01)


@pytest.mark.parametrize("name,case", CASES, ids=IDS)
def test_validity_kernel_on_the_oracle_state(name, case):
    """The oracle's state after 6 steps, copied onto the device (q, e, p, u,
    u_prev on the ghosted grid): the ten validity numbers equal the oracle's
    per cell to 1e-11."""
    say("validity kernel", name)
    state = event(case, 4)
    orc = case.oracle()
    orc.set_primitives(case.t0, *state)
    orc.steps(case.t0, 0, 6)
    t = case.t0 + 6 * case.dt
    ctx = case.context()
    ctx.set_initial_state(case.t0, *state)
    nq = 15 if case.bulk else 14
    for f in list(range(nq)) + [O.E, O.P, O.UT, O.UX, O.UY, O.UN, O.UTP, O.UXP, O.UYP, O.UNP]:
        ctx.set_field(f, orc.get(f, ghosted=True))
    ref = orc.validity(t, prev_q=0)[:O.V_REG]
    ctx.check_validity(t, gpuvh.VALIDITY_CURRENT_PI)
    got = device_validity(ctx, range(O.V_REG))
    errs = field_errors(got, ref, floor=1e-6, cond=bulk_conditioning(orc))
    say("  errors", {k: "%.1e" % v for k, v in errs.items()})
    assert max(errs.values()) <= 1e-11
    ctx.close()
    orc.close()


@pytest.mark.parametrize("name,case", CASES, ids=IDS)
def test_validity_after_steps(name, case):
    """Before the first step and after 12 steps of both paths: the eleven
    validity arrays (pi and Pi of q^n of the last step, the regulation factor
    of the last step's end) agree with the oracle's to 1e-10; with
    GVH_VALIDITY_CURRENT_PI they follow the current pi."""
    say("validity after steps", name)
    state = event(case, 5)
    orc = case.oracle()
    orc.set_primitives(case.t0, *state)
    ctx = case.context()
    ctx.set_initial_state(case.t0, *state)
    ctx.validity_enable()
    ctx.check_validity(case.t0)
    errs0 = field_errors(device_validity(ctx), orc.validity(case.t0, prev_q=0), floor=1e-4,
                         cond=bulk_conditioning(orc))
    say("  step 0", {k: "%.1e" % v for k, v in errs0.items()})
    assert max(errs0.values()) <= 1e-10
    n = 12
    orc.steps(case.t0, 0, n)
    ctx.steps(case.t0, 0, n)
    t = case.t0 + n * case.dt
    ctx.check_validity(t)
    ref = orc.validity(t, prev_q=1)
    got = device_validity(ctx)
    cond = bulk_conditioning(orc)
    errs = field_errors(got, ref, floor=1e-4, cond=cond)
    say("  step %d" % n, {k: "%.1e" % v for k, v in errs.items()})
    assert max(errs.values()) <= 1e-10
    if name == "3d-wb-bulk-reg":
        # the regulation is active: factors below 1 on part of the lattice
        assert (ref[O.V_REG] < 0.999).sum() > 10
    ctx.check_validity(t, gpuvh.VALIDITY_CURRENT_PI)
    errs_c = field_errors(device_validity(ctx, range(O.V_REG)), orc.validity(t, prev_q=0)[:O.V_REG],
                          floor=1e-4, cond=cond)
    assert max(errs_c.values()) <= 1e-10
    # the two readings differ where pi changed over the step
    assert not np.allclose(ref[O.V_RE_PI], orc.validity(t, prev_q=0)[O.V_RE_PI], rtol=1e-6)
    ctx.close()
    orc.close()


def test_validity_mixed_precision():
    """A mixed-precision context (pi, Pi stored in FP32) gives the FP64
    context's validity numbers to 1e-3 (floor 1e-2) after 10 steps."""
    case = CASES[0][1]
    state = event(case, 6)
    out = {}
    for prec in (gpuvh.FP64, gpuvh.MIXED):
        ctx = case.context(prec)
        ctx.set_initial_state(case.t0, *state)
        ctx.validity_enable()
        ctx.steps(case.t0, 0, 10)
        ctx.check_validity(case.t0 + 10 * case.dt)
        out[prec] = device_validity(ctx)
        ctx.close()
    errs = field_errors(out[gpuvh.MIXED], out[gpuvh.FP64], floor=1e-2)
    say("  mixed vs fp64", {k: "%.1e" % v for k, v in errs.items()})
    assert max(errs.values()) <= 1e-3


def test_validity_argument_errors():
    case = CASES[2][1]
    ctx = case.context()
    with pytest.raises(gpuvh.GvhError):
        ctx.check_validity(case.t0)  # no state yet
    state = event(case, 7)
    ctx.set_initial_state(case.t0, *state)
    with pytest.raises(gpuvh.GvhError):
        ctx.get_validity(gpuvh.V_THETA)  # not enabled
    with pytest.raises(gpuvh.GvhError):
        ctx.check_validity(case.t0, 8)
    with pytest.raises(gpuvh.GvhError):
        ctx.check_validity(0.0)
    ctx.check_validity(case.t0)
    with pytest.raises(gpuvh.GvhError):
        ctx.get_validity(NV)
    reg = ctx.get_validity(gpuvh.V_REGULATIONS, ghosted=True)
    assert np.all(reg == 1.0)
    ctx.close()


def test_run_loop_validity_files(tmp_path):
    """gvh_run with the validity option writes the eleven fields at every
    FREQ snapshot under the reference's output names, equal to the restated
    run loop's validity arrays (printed with %.8f)."""
    from test_freezeout_gpu import EVENTS, event_case, oracle_eos_T_p, smooth_state, write_ic_files
    ev = EVENTS[2]
    case, geom, Tfo = event_case(ev)
    name, nx, ny, nz, dx, dz, dt = ev[:7]
    say("run loop validity", name)
    state = smooth_state(case, 8)
    root = str(tmp_path / "root")
    out = str(tmp_path / "out")
    os.makedirs(out)
    write_ic_files(root, state, nx, ny, nz, dx, dz, case.bulk)
    nt = 40
    L = gpuvh.lattice(nx, ny, nz, dx, dx, dz, dt, nt=nt)
    IC = gpuvh.InitialConditionParameters()
    IC.initialConditionType = 10
    H = gpuvh.hydro(0.5, 0.2, T_fo_GeV=Tfo)
    o = gpuvh.options(case.eos, case.bulk, gpuvh.FP64, case.zeta_tpeak, table=case.table)
    gpuvh.run(L, IC, H, root, out, o, gpuvh.run_options(validity=True, surface=False))
    orc = case.oracle()
    e_ic = np.float32(state[0] - 1e-3).astype(np.float64) + 1e-3
    orc.set_primitives(case.t0, e_ic, O.eos_p_batch(orc.P, e_ic), *state[2:])
    e_fo = O.e_of_T(orc.P, Tfo / gpuvh.HBARC)
    _, steps, _, snaps = F.run_loop(orc, geom, nt, e_fo, oracle_eos_T_p(orc.P), bool(case.bulk),
                                    surface=False, snapshot_fields=[O.E], validity=True)
    assert len(snaps) >= 4
    for tsnap, fields in snaps.items():
        ref_all = fields["validity"]
        for f, fname in enumerate(gpuvh.VALIDITY_NAMES):
            path = os.path.join(out, "%s_%.3f.dat" % (fname, tsnap))
            d = np.loadtxt(path)
            assert d.shape == (nx * ny * nz, 4), path
            v = d[:, 3].reshape((nx, ny, nz), order="F")
            ref = ref_all[f]
            scale = np.median(np.abs(ref)) if f != gpuvh.V_KN_TAUBULK else 0.0
            rtol = 1e-7 if f != gpuvh.V_KN_TAUBULK else 1e-4
            np.testing.assert_allclose(v, ref, rtol=rtol, atol=2e-8 + 1e-7 * scale, err_msg=path)
    orc.close()
