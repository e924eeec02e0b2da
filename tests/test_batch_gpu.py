"""Batches of events (gvh_batch_*): each event's final state, step count,
stop reason and freeze-out rows are the same in the mix as in a batch of its
own -- exactly, since an event's cells get the same arithmetic whatever runs
beside them -- in FP64 and mixed precision, with continuous admission under
a memory budget; and they match the run loop's rules on a context of its own
(gvh_step, gvh_freezeout_*, gvh_count_above) to the parity tolerance, the
two paths being separately compiled kernels."""
import numpy as np
import pytest

import gpuvh
from gpuvh import ic

pytestmark = pytest.mark.gpu

FOFREQ = 10


def say(*a):
    print(*a, flush=True)


def event_e(nx, ny, nz, dx, dz, seed, peak=40.0):
    ext = 0.5 * (nx - 1) * dx
    e2 = ic.hotspot_event_2d(nx, ny, dx, dx, seed, n_hot=8, hot_sigma=0.5, hot_amp_max=10.0,
                             base_peak=peak, base_sigma=0.5 * ext, centre_radius=0.4 * ext)
    eL = ic.longitudinal_profile(nz, dz, eta_flat=0.3 * nz * dz, eta_variance=0.1) \
        if nz > 1 else np.ones(1)
    return ic.quantise_like_reference(e2[:, :, None] * eL[None, None, :])


def T_at(table, e):
    """Temperature [GeV] of energy density e on the tabulated EoS."""
    nodes = np.array([table.node_e(k) for k in range(table.n)])
    return float(np.interp(e, nodes, table.T)) * gpuvh.HBARC


def make_events(specs, table):
    out = []
    for (nx, ny, nz, dx, dz, dt, nt, seed) in specs:
        e = event_e(nx, ny, nz, dx, dz, seed)
        L = gpuvh.lattice(nx, ny, nz, dx, dx, dz, dt, nt=nt)
        Tfo = T_at(table, float(np.median(e)))
        H = gpuvh.hydro(0.5, 0.16, T_fo_GeV=Tfo)
        out.append((L, H, e))
    return out


def run_alone(opts, L, H, e, surface, fo_flags=0, fofreq=FOFREQ):
    """The run loop's stepping, freeze-out windows and stops on one context."""
    ctx = gpuvh.Context(L, H, opts)
    shape = e.shape
    ut, ux, uy, un, pi10, Pi = ic.at_rest(shape)
    ctx.set_initial_state(H.initialProperTimePoint, e, ctx.equilibrium_pressure(e), ut, ux, uy, un)
    e_fo = ctx.energy_density_of_T(H.freezeoutTemperatureGeV / gpuvh.HBARC)
    if surface:
        ctx.freezeout_init(fofreq, e_fo, fo_flags)
    t0, dt, nt = H.initialProperTimePoint, L.latticeSpacingProperTime, L.numProperTimePoints
    rows, t, steps, frozen_count, frozen = [], t0, 0, 0, False
    for n in range(1, nt + 2):
        if surface:
            ctx.freezeout_record(n)
            if n % fofreq == fofreq - 1:
                rows.append(ctx.freezeout_find(n, t))
        if ctx.count_above(e_fo) == 0:
            frozen_count += 1
        if frozen_count >= fofreq + 1:
            frozen = True
            break
        ctx.step(t)
        steps += 1
        t = t0 + n * dt
    rows = np.concatenate(rows) if rows else np.zeros((0, 20))
    return ctx, rows, steps, frozen


FIELDS = [gpuvh.E, gpuvh.UT, gpuvh.UX, gpuvh.UY, gpuvh.UN, gpuvh.PIXX, gpuvh.PIXY, gpuvh.PINN,
          gpuvh.PI]


def run_batch(opts, events, surface, hbm_bytes=0, fofreq=FOFREQ):
    """(result, rows, {field: array}) per event, in submission order."""
    batch = gpuvh.Batch(opts, hbm_bytes=hbm_bytes, max_events=8, autotune=False, fofreq=fofreq)
    ids = [batch.submit(L, H, e, surface=surface, keep_state=True) for (L, H, e) in events]
    batch.run()
    res = {r["id"]: r for r in batch.poll()}
    assert set(res) == set(ids)
    out = []
    for i in ids:
        r = res[i]
        out.append((r, batch.rows(i, r["surface_rows"]), {f: batch.field(i, f) for f in FIELDS}))
    batch.close()
    return out


def assert_same(a, b, label):
    (ra, rows_a, fa), (rb, rows_b, fb) = a, b
    assert ra["steps"] == rb["steps"] and ra["frozen"] == rb["frozen"], label
    np.testing.assert_array_equal(rows_a, rows_b, err_msg=label)
    for f in FIELDS:
        np.testing.assert_array_equal(fa[f], fb[f], err_msg="%s %s" % (label, gpuvh.FIELD_NAMES[f]))


def assert_close_to_context(res, opts, ev, label, fofreq=FOFREQ):
    """Against the run loop on a context of its own: same steps and stop,
    fields to 1e-10 (floor 1e-4 of the field's maximum), rows to 1e-8."""
    r, rows, fields = res
    L, H, e = ev
    ctx, rows_c, steps, frozen = run_alone(opts, L, H, e, True, fofreq=fofreq)
    say("  ", label, "steps", r["steps"], "context", steps, "rows", len(rows), len(rows_c))
    assert r["steps"] == steps and bool(r["frozen"]) == frozen, label
    for f in FIELDS:
        ref = ctx.get(f)
        scale = np.abs(ref).max()
        if scale == 0:
            assert np.all(fields[f] == 0)
            continue
        err = np.max(np.abs(fields[f] - ref) / np.maximum(np.abs(ref), 1e-4 * scale))
        assert err <= 1e-10, "%s %s %.2e" % (label, gpuvh.FIELD_NAMES[f], err)
    assert len(rows) == len(rows_c), label
    if len(rows):
        np.testing.assert_allclose(rows, rows_c, rtol=1e-8, atol=1e-12 * np.abs(rows_c).max())
    ctx.close()


def test_windows_every_three_steps(monkeypatch):
    """fofreq 3, the smallest: a window's search overlaps the next steps and
    must be finished before the step that records into its ring half again
    (four steps later).  Four events of different shapes and lengths in one
    batch: steps, stop and rows as the run loop on contexts of their own
    (rows to 1e-8, fields to 1e-10), and bitwise as each event alone."""
    monkeypatch.setenv("GVH_VARIANT", "0")
    table = gpuvh.EosTable()
    opts = gpuvh.options(gpuvh.EOS_TABLE, True, gpuvh.FP64, 0.16 / gpuvh.HBARC, table=table)
    specs = [(31, 29, 1, 0.2, 0.1, 0.02, 23, 11), (21, 25, 1, 0.25, 0.1, 0.025, 17, 12),
             (13, 15, 9, 0.3, 0.3, 0.03, 20, 13), (11, 11, 7, 0.3, 0.3, 0.03, 14, 14)]
    events = make_events(specs, table)
    mix = run_batch(opts, events, True, fofreq=3)
    for k, ev in enumerate(events):
        assert len(mix[k][1]) > 0, "event %d has no surface rows" % k
        assert_close_to_context(mix[k], opts, ev, "event %d" % k, fofreq=3)
        assert_same(mix[k], run_batch(opts, [ev], True, fofreq=3)[0], "event %d" % k)


@pytest.mark.parametrize("precision", [gpuvh.FP64, gpuvh.MIXED], ids=["fp64", "mixed"])
def test_batch_matches_alone(precision, monkeypatch):
    """Five events of three 2+1D and two 3+1D shapes in one batch (two launch
    groups), with surfaces: bitwise equal to each event alone."""
    monkeypatch.setenv("GVH_VARIANT", "0")
    table = gpuvh.EosTable()
    opts = gpuvh.options(gpuvh.EOS_TABLE, True, precision, 0.16 / gpuvh.HBARC, table=table)
    specs = [(41, 41, 1, 0.2, 0.1, 0.02, 40, 1), (33, 37, 1, 0.2, 0.1, 0.02, 25, 2),
             (25, 29, 1, 0.25, 0.1, 0.025, 31, 3), (15, 15, 9, 0.3, 0.3, 0.03, 22, 4),
             (13, 11, 7, 0.3, 0.3, 0.03, 19, 5)]
    events = make_events(specs, table)
    mix = run_batch(opts, events, True)
    for k, ev in enumerate(events):
        alone = run_batch(opts, [ev], True)[0]
        say("  event", k, "steps", mix[k][0]["steps"], "rows", len(mix[k][1]))
        assert_same(mix[k], alone, "event %d" % k)
        if precision == gpuvh.FP64:
            assert_close_to_context(mix[k], opts, ev, "event %d" % k)


def test_continuous_batching_under_budget(monkeypatch):
    """A budget that holds about two events: the others wait, are admitted
    as running ones finish, and still equal their runs alone."""
    monkeypatch.setenv("GVH_VARIANT", "0")
    table = gpuvh.EosTable()
    opts = gpuvh.options(gpuvh.EOS_TABLE, True, gpuvh.FP64, 0.16 / gpuvh.HBARC, table=table)
    specs = [(41, 41, 1, 0.2, 0.1, 0.02, 12 + 7 * k, 10 + k) for k in range(5)]
    events = make_events(specs, table)
    probe = gpuvh.Batch(opts, max_events=8, autotune=False)
    probe.submit(*events[0], surface=True)
    probe.step()  # the footprint is booked at admission
    one = probe.stats()["hbm_used"]
    probe.close()
    batch = gpuvh.Batch(opts, hbm_bytes=int(2.5 * one), max_events=8, autotune=False)
    ids = [batch.submit(L, H, e, surface=True, keep_state=False) for (L, H, e) in events]
    max_running = 0
    while True:
        n = batch.step()
        max_running = max(max_running, batch.stats()["running"])
        if n == 0:
            break
    say("  max running", max_running)
    assert max_running == 2
    res = {r["id"]: r for r in batch.poll()}
    admits = sorted(res[i]["t_admit"] for i in ids)
    dones = sorted(res[i]["t_done"] for i in ids)
    assert admits[2] >= dones[0]  # the third waited for a slot of memory
    # each equal to its batch of one (state not kept: steps and surface rows)
    for i, ev in zip(ids, events):
        alone = run_batch(opts, [ev], True)[0]
        assert res[i]["steps"] == alone[0]["steps"]
        assert res[i]["surface_rows"] == len(alone[1])
        np.testing.assert_array_equal(batch.rows(i, len(alone[1])), alone[1])
    batch.close()


def test_stops_frozen_and_nt(monkeypatch):
    """An event that cools below e_fo everywhere stops by the frozen rule;
    one with a short nt stops after nt + 1 steps; both as alone."""
    monkeypatch.setenv("GVH_VARIANT", "0")
    table = gpuvh.EosTable()
    opts = gpuvh.options(gpuvh.EOS_TABLE, True, gpuvh.FP64, 0.16 / gpuvh.HBARC, table=table)
    e_hot = event_e(21, 21, 1, 0.2, 0.1, 7)
    L1 = gpuvh.lattice(21, 21, 1, 0.2, 0.2, 0.1, 0.05, nt=400)
    H1 = gpuvh.hydro(0.5, 0.16, T_fo_GeV=T_at(table, float(e_hot.max()) * 0.9))
    L2 = gpuvh.lattice(21, 21, 1, 0.2, 0.2, 0.1, 0.02, nt=7)
    H2 = gpuvh.hydro(0.5, 0.16, T_fo_GeV=0.05)
    events = [(L1, H1, e_hot), (L2, H2, e_hot)]
    out = run_batch(opts, events, True)
    assert out[0][0]["frozen"] == 1 and out[0][0]["steps"] < 400
    assert out[1][0]["frozen"] == 0 and out[1][0]["steps"] == 8
    for k, ev in enumerate(events):
        assert_close_to_context(out[k], opts, ev, "event %d" % k)


def test_autotune_keeps_state_and_results():
    """gvh_autotune returns a valid arrangement, leaves the state as it was,
    and the tuned context still matches an untuned one to 1e-10 after steps."""
    table = gpuvh.EosTable()
    opts = gpuvh.options(gpuvh.EOS_TABLE, True, gpuvh.FP64, 0.16 / gpuvh.HBARC, table=table)
    nx, ny, nz = 23, 21, 9
    e = event_e(nx, ny, nz, 0.1, 0.1, 3)
    L = gpuvh.lattice(nx, ny, nz, 0.1, 0.1, 0.1, 0.01)
    H = gpuvh.hydro(0.5, 0.16)
    ctxs = []
    for tune in (False, True):
        ctx = gpuvh.Context(L, H, opts)
        ut, ux, uy, un, pi10, Pi = ic.at_rest(e.shape)
        ctx.set_initial_state(0.5, e, ctx.equilibrium_pressure(e), ut, ux, uy, un)
        if tune:
            before = ctx.get(gpuvh.E, ghosted=True)
            v, st, ms = ctx.autotune(3)
            say("  tuned variant", v, "strip", st, "%.4f ms/step" % ms)
            assert v in (0, 3) and st in (4, 8, 16, 21) and ms > 0
            np.testing.assert_array_equal(ctx.get(gpuvh.E, ghosted=True), before)
        ctx.steps(0.5, 0, 10)
        ctxs.append(ctx)
    a, b = ctxs[0].get(gpuvh.E), ctxs[1].get(gpuvh.E)
    assert np.max(np.abs(a - b) / np.abs(a)) < 1e-10
    for c in ctxs:
        c.close()


def test_batch_argument_errors():
    opts = gpuvh.options(gpuvh.EOS_CONFORMAL, False, gpuvh.FP64)
    batch = gpuvh.Batch(opts, hbm_bytes=1 << 20, max_events=2, autotune=False)
    e = np.ones((41, 41, 1))
    L = gpuvh.lattice(41, 41, 1, 0.1, 0.1, 0.1, 0.01, nt=5)
    H = gpuvh.hydro(0.5, 0.2)
    with pytest.raises(gpuvh.GvhError):
        batch.submit(L, H, e, surface=True)   # larger than the 1 MiB budget
    with pytest.raises(gpuvh.GvhError):
        gpuvh._check(gpuvh.lib().gvh_batch_field(batch.h, 0, gpuvh.E, None))  # no such event
    batch.close()


def test_invalid_submission_between_valid_events(monkeypatch):
    """Submissions are checked as gvh_create checks a context (positive
    spacings and eta/s): an invalid event between two valid ones is refused
    at submit and the valid ones run to their ends, equal to their batches
    of one."""
    monkeypatch.setenv("GVH_VARIANT", "0")
    table = gpuvh.EosTable()
    opts = gpuvh.options(gpuvh.EOS_TABLE, True, gpuvh.FP64, 0.16 / gpuvh.HBARC, table=table)
    events = make_events([(21, 21, 1, 0.2, 0.1, 0.02, 9, 31), (17, 19, 1, 0.2, 0.1, 0.02, 7, 32)],
                         table)
    batch = gpuvh.Batch(opts, max_events=4, autotune=False)
    a = batch.submit(*events[0], keep_state=True)
    L, H, e = events[1]
    bad_h = gpuvh.hydro(0.5, 0.0)  # eta/s = 0
    with pytest.raises(gpuvh.GvhError):
        batch.submit(L, bad_h, e)
    bad_l = gpuvh.lattice(17, 19, 1, 0.0, 0.2, 0.1, 0.02, nt=7)  # dx = 0
    with pytest.raises(gpuvh.GvhError):
        batch.submit(bad_l, H, e)
    with pytest.raises(ValueError):
        batch.submit(L, H, e[:-1])  # fewer values than the lattice
    c = batch.submit(L, H, e, keep_state=True)
    batch.run()
    res = {r["id"]: r for r in batch.poll()}
    assert set(res) == {a, c}
    for i, ev in ((a, events[0]), (c, events[1])):
        assert res[i]["status"] == gpuvh.OK
        alone = run_batch(opts, [ev], False)[0]
        assert res[i]["steps"] == alone[0]["steps"] == ev[0].numProperTimePoints + 1
        for f in FIELDS:
            np.testing.assert_array_equal(batch.field(i, f), alone[2][f])
    batch.close()


def test_newton_budget_retires_only_its_event(monkeypatch):
    """An event submitted with a Newton iteration budget of 2 has root-finds
    that need more (half-site states a few percent from the cell's own e),
    so it is retired with status ENOCONV after its first step; the events
    beside it run to their ends, bitwise equal to their batches of one.  The
    same event with a budget it never reaches (1000) ends normally, equal to
    its run without a budget."""
    monkeypatch.setenv("GVH_VARIANT", "0")
    table = gpuvh.EosTable()
    opts = gpuvh.options(gpuvh.EOS_TABLE, True, gpuvh.FP64, 0.16 / gpuvh.HBARC, table=table)
    specs = [(25, 25, 1, 0.2, 0.1, 0.02, 12, 41), (21, 23, 1, 0.2, 0.1, 0.02, 15, 42),
             (13, 11, 7, 0.3, 0.3, 0.03, 9, 43)]
    events = make_events(specs, table)
    batch = gpuvh.Batch(opts, max_events=4, autotune=False)
    ids = [batch.submit(L, H, e, keep_state=True, newton_budget=(2 if k == 1 else 0))
           for k, (L, H, e) in enumerate(events)]
    batch.run()
    res = {r["id"]: r for r in batch.poll()}
    say("  results", [(res[i]["steps"], res[i]["status"], res[i]["newton_not_converged"])
                      for i in ids])
    bad = res[ids[1]]
    assert bad["status"] == gpuvh.ENOCONV
    assert bad["newton_not_converged"] > 0 and bad["steps"] <= 2
    for k in (0, 2):
        r = res[ids[k]]
        assert r["status"] == gpuvh.OK and r["newton_not_converged"] == 0
        alone = run_batch(opts, [events[k]], False)[0]
        assert r["steps"] == alone[0]["steps"]
        for f in FIELDS:
            np.testing.assert_array_equal(batch.field(ids[k], f), alone[2][f])
    batch.close()
    # a budget that is never reached changes nothing
    batch = gpuvh.Batch(opts, max_events=2, autotune=False)
    i = batch.submit(*events[1], keep_state=True, newton_budget=1000)
    batch.run()
    r = batch.poll()[0]
    assert r["status"] == gpuvh.OK and r["newton_not_converged"] == 0
    alone = run_batch(opts, [events[1]], False)[0]
    assert r["steps"] == alone[0]["steps"]
    for f in FIELDS:
        np.testing.assert_array_equal(batch.field(i, f), alone[2][f])
    batch.close()
