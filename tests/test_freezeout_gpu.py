"""Freeze-out surface on the device against the oracle: cornelius cells,
the history ring, window rows and the run loop (gvh_run, gvh_run_hydro)."""
import ctypes as C
import gzip
import os

import numpy as np
import pytest

import cornelius_lib as CL
import fo_ref as F
import gpuvh
import oracle_lib as O
from gpuvh import ic
from parity import Case

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cornelius")


def say(*a):
    print(*a, flush=True)


# ------------------------------------------------------------ cells
def oracle_cells(dim, cells, dx, v0):
    counts = np.zeros(len(cells), dtype=np.int32)
    cen = np.zeros((len(cells), 10, dim))
    nrm = np.zeros((len(cells), 10, dim))
    for i, c in enumerate(cells):
        a, b = CL.find(dim, c, dx, v0)
        counts[i] = len(a)
        cen[i, :len(a)] = a
        nrm[i, :len(a)] = b
    return counts, cen, nrm


@pytest.mark.parametrize("dim", [3, 4])
def test_device_cornelius_equals_oracle(dim):
    """The 10000 random cells of the reference's example, cells with corners
    on the iso-value, unit and anisotropic spacings: element counts, order,
    centroids and normals equal the oracle's bit for bit."""
    _, cu, hc = CL.example_inputs()
    cells = np.concatenate([cu if dim == 3 else hc, CL.degenerate_cells(dim, 3000, 21)])
    for dx in (np.ones(dim), np.array([0.01, 0.1, 0.2, 0.15][:dim])):
        say("cornelius dim", dim, "cells", len(cells), "dx", dx)
        got = gpuvh.cornelius_cells(dim, cells, dx, 0.5)
        ref = oracle_cells(dim, cells, dx, 0.5)
        np.testing.assert_array_equal(got[0], ref[0])
        for k in (1, 2):
            mask = np.arange(10)[None, :] < ref[0][:, None]
            a, b = got[k][mask], ref[k][mask]
            diff = np.abs(a - b).max() if a.size else 0.0
            assert diff == 0.0, "max difference %g" % diff


@pytest.mark.parametrize("dim", [3, 4])
def test_device_cornelius_golden_text(dim):
    """Device elements of the example's cells printed like the reference
    equal output_{3d,4d}.dat line for line (6 significant digits)."""
    _, cu, hc = CL.example_inputs()
    cells = cu if dim == 3 else hc
    counts, cen, nrm = gpuvh.cornelius_cells(dim, cells, np.ones(dim), 0.5)
    lines = CL.format_elements(dim, counts, cen, nrm)
    with gzip.open(os.path.join(GOLD, "output_%dd.dat.gz" % dim), "rt") as f:
        gold = f.read().splitlines()
    assert len(lines) == len(gold)
    bad = [(a, b) for a, b in zip(lines, gold) if a != b]
    for a, b in bad:
        assert np.allclose([float(x) for x in a.split()], [float(x) for x in b.split()],
                           rtol=2e-5, atol=2e-6), (a, b)
    assert len(bad) < 0.001 * len(gold)


# ------------------------------------------------------------ events
def smooth_state(case, seed):
    """A dense fireball (30 GeV/fm^3 Gaussian over half the lattice plus six
    hotspots), eta plateau profile in 3+1D, at rest, quantised like the
    reference's file reader."""
    ext = 0.5 * (case.nx - 1) * case.dx
    e2 = ic.hotspot_event_2d(case.nx, case.ny, case.dx, case.dx, seed, n_hot=6, hot_sigma=0.5,
                             hot_amp_max=8.0, base_peak=30.0, base_sigma=0.5 * ext,
                             centre_radius=0.4 * ext)
    eL = ic.longitudinal_profile(case.nz, case.dz, eta_flat=0.3 * case.nz * case.dz,
                                 eta_variance=0.1) if case.nz > 1 else np.ones(1)
    e = ic.quantise_like_reference(e2[:, :, None] * eL[None, None, :])
    ut, ux, uy, un, pi10, Pi = ic.at_rest(e.shape)
    return e, case.pressure(e), ut, ux, uy, un, pi10, Pi


# (name, nx, ny, nz, dx, dz, dt, T_fo GeV, eos, bulk)
EVENTS = [
    ("2d-conformal", 41, 41, 1, 0.2, 0.1, 0.02, 0.25, gpuvh.EOS_CONFORMAL, 0),
    ("3d-conformal", 15, 15, 9, 0.3, 0.3, 0.03, 0.3, gpuvh.EOS_CONFORMAL, 0),
    ("3d-wb-bulk", 15, 15, 9, 0.3, 0.3, 0.03, 0.2, gpuvh.EOS_WB, 1),
    ("2d-table-bulk", 21, 21, 1, 0.3, 0.1, 0.03, 0.2, gpuvh.EOS_TABLE, 1),
]


def event_case(ev):
    name, nx, ny, nz, dx, dz, dt, Tfo, eos, bulk = ev
    case = Case(nx, ny, nz, dx, dz, dt, 0.5, 0.2, bulk, eos)
    geom = dict(nx=nx, ny=ny, nz=nz, dt=dt, dx=dx, dy=dx, dz=dz, t0=0.5,
                dim=4 if nz > 1 else 3)
    return case, geom, Tfo


def oracle_eos_T_p(P):
    def f(e):
        p, _, T = O.eos(P, e)
        return T, p
    return f


@pytest.mark.parametrize("flags", [0, gpuvh.FO_EXACT_TAU | gpuvh.FO_NORMAL_TAU | gpuvh.FO_CURRENT_FIELDS])
@pytest.mark.parametrize("ev", EVENTS[:3], ids=[e[0] for e in EVENTS[:3]])
def test_window_rows_and_ring(ev, flags):
    """Over 35 steps of an event: every recorded slice equals the state the
    reference's host copy would hold (current e; u and pi current or one step
    behind by step parity unless FO_CURRENT_FIELDS), and every window's rows
    equal the restatement's rows from the same slices (T and P to 1e-14,
    everything else exactly)."""
    case, geom, Tfo = event_case(ev)
    say("window rows", ev[0], "flags", flags)
    state = smooth_state(case, 7)
    ctx = case.context()
    ctx.set_initial_state(case.t0, *state)
    P = case.oracle().P
    e_fo = O.e_of_T(P, Tfo / gpuvh.HBARC)
    assert ctx.energy_density_of_T(Tfo / gpuvh.HBARC) == pytest.approx(e_fo, rel=1e-12)
    fofreq = 10
    ctx.freezeout_init(fofreq, e_fo, flags)
    fields = [gpuvh.UX, gpuvh.UY, gpuvh.UN, gpuvh.E, gpuvh.PIXX, gpuvh.PIXY, gpuvh.PIXN,
              gpuvh.PIYY, gpuvh.PIYN, gpuvh.PI]
    t = case.t0
    prev = None
    nrows = 0
    for n in range(1, 36):
        cur = [ctx.get(f) for f in fields]
        if not case.bulk:
            cur[9] = np.zeros_like(cur[9])
        ctx.freezeout_record(n)
        logical = 1 if n % fofreq == 0 else n % fofreq + 1
        want = F.host_view(cur, prev, n - 1, flags)
        for v in range(10):
            np.testing.assert_array_equal(ctx.freezeout_slot(logical, v), want[v],
                                          err_msg="n=%d var=%s" % (n, F.VARS[v]))
        if n % fofreq == fofreq - 1:
            rows = ctx.freezeout_find(n, t)
            ring = F.Ring(fofreq)
            start = 2 if n <= fofreq else 0
            for k in range(start, fofreq + 1):
                ring.slots[k] = [ctx.freezeout_slot(k, v) for v in range(10)]
            ref = F.window_rows(ring, n, t, geom, e_fo, oracle_eos_T_p(P), flags)
            say("  n", n, "rows", len(rows))
            assert rows.shape == ref.shape
            cols = [c for c in range(20) if c not in (12, 13)]
            np.testing.assert_array_equal(rows[:, cols], ref[:, cols])
            np.testing.assert_allclose(rows[:, 12:14], ref[:, 12:14], rtol=1e-14, atol=0)
            nrows += len(rows)
        prev = cur
        ctx.step(t)
        t = case.t0 + n * case.dt
    assert nrows > 100
    ctx.close()


@pytest.mark.parametrize("nz", [1, 4])
def test_uniform_fluid_surface_closed_form(nz):
    """A uniform fluid at rest stays uniform in space (Bjorken flow): every
    cube of the slab in which e crosses e_fo is one element at the linearly
    interpolated crossing time with normal (t dx dy deta, 0, 0, 0),
    interpolated e = e_fo and T, P of e_fo; windows after the first place cell
    tau one step late."""
    nx, ny, dx, dz, dt, t0 = 5, 4, 0.2, 0.3, 0.05, 0.5
    case = Case(nx, ny, nz, dx, dz, dt, t0, 0.2, 0, gpuvh.EOS_CONFORMAL)
    shape = (nx, ny, nz)

    def fresh():
        ut, ux, uy, un, pi10, Pi = ic.at_rest(shape)
        e = np.full(shape, 12.0)
        ctx = case.context()
        ctx.set_initial_state(t0, e, e / 3, ut, ux, uy, un, pi10, Pi)
        return ctx

    # first pass: e after k steps (uniform in space)
    ctx = fresh()
    e_hist = []
    t = t0
    for n in range(1, 30):
        eg = ctx.get(gpuvh.E)
        assert np.all(eg == eg[0, 0, 0])
        e_hist.append(eg[0, 0, 0])
        ctx.step(t)
        t = t0 + n * dt
    ctx.close()
    e_fo = 0.5 * (e_hist[23] + e_hist[24])   # crossed between states 23 and 24
    fofreq = 10
    ctx = fresh()
    ctx.freezeout_init(fofreq, e_fo, 0)
    t = t0
    found = []
    for n in range(1, 30):
        ctx.freezeout_record(n)
        if n % fofreq == fofreq - 1:
            found.append((n, t, ctx.freezeout_find(n, t)))
        ctx.step(t)
        t = t0 + n * dt
    assert [len(f[2]) > 0 for f in found] == [False, False, True]
    n, tw, rows = found[2]
    ncube = (nx - 1) * (ny - 1) * max(nz - 1, 1)
    assert rows.shape == (ncube, 20)
    # window n = 29: slice k holds the state after 18 + k steps, so the
    # crossing lies in slice it = 5; the reference's cell tau for it is
    # t0 + (n - fofreq + it) dt, one step after that state's time
    ea, eb = e_hist[23], e_hist[24]
    frac = (ea - e_fo) / (ea - eb)
    cell_tau = t0 + (n - fofreq + 5) * dt
    vol = dx * dx * (dz if nz > 1 else 1.0)
    np.testing.assert_allclose(rows[:, 0], cell_tau + frac * dt, rtol=1e-12)
    np.testing.assert_allclose(rows[:, 4], tw * vol, rtol=1e-11)
    assert np.abs(rows[:, 5:8]).max() < 1e-13
    np.testing.assert_allclose(rows[:, 11], e_fo, rtol=1e-12)
    np.testing.assert_allclose(rows[:, 12], (e_fo / 13.8997) ** 0.25, rtol=1e-12)
    np.testing.assert_allclose(rows[:, 13], e_fo / 3, rtol=1e-12)
    ctx.close()


# ------------------------------------------------------------ run loop
def write_ic_files(root, state, nx, ny, nz, dx, dz, bulk):
    """Type-10 input files: x y z value per cell, x outermost, eta innermost.
    e.dat holds e - 1e-3 as a float (the reader adds 1e-3 back)."""
    e, p, ut, ux, uy, un, pi10, Pi = state
    os.makedirs(os.path.join(root, "input"), exist_ok=True)
    os.makedirs(os.path.join(root, "output"), exist_ok=True)
    xs = (np.arange(nx) - (nx - 1) / 2.0) * dx
    ys = (np.arange(ny) - (ny - 1) / 2.0) * dx
    zs = (np.arange(nz) - (nz - 1) / 2.0) * dz
    e_file = np.float32(e - 1e-3)
    assert np.array_equal(e_file.astype(np.float64) + 1e-3, e)
    arrays = {"e": e_file, "ut": ut, "ux": ux, "uy": uy, "un": un}
    names = ["pitt", "pitx", "pity", "pitn", "pixx", "pixy", "pixn", "piyy", "piyn", "pinn"]
    for k, nm in enumerate(names):
        arrays[nm] = pi10[k]
    if bulk:
        arrays["bulk"] = Pi
    for nm, a in arrays.items():
        a32 = np.asarray(a, dtype=np.float32)
        with open(os.path.join(root, "input", nm + ".dat"), "w") as f:
            for i in range(nx):
                for j in range(ny):
                    for k in range(nz):
                        f.write("%.6f %.6f %.6f %.9g\n" % (xs[i], ys[j], zs[k], a32[i, j, k]))


def parse_surface(path):
    with open(path) as f:
        lines = f.read().splitlines()
    return lines, np.array([[float(x) for x in ln.split()] for ln in lines]).reshape(-1, 20)


def read_snapshot(path):
    return np.loadtxt(path)


def check_run_against_restatement(case, geom, Tfo, oracle_state, nt, stats, root, out):
    """The run's steps, all-frozen stop, surface.dat and FREQ snapshots
    against the restated run loop driven by the oracle from oracle_state
    (surface text line for line, values that print differently within 1e-5
    relative)."""
    orc = case.oracle()
    orc.set_primitives(case.t0, *(oracle_state(orc.P) if callable(oracle_state) else oracle_state))
    e_fo = O.e_of_T(orc.P, Tfo / gpuvh.HBARC)
    snap_fields = [O.E, O.UX, O.PIXX]
    rows, steps, frozen, snaps = F.run_loop(orc, geom, nt, e_fo, oracle_eos_T_p(orc.P),
                                            bool(case.bulk), snapshot_fields=snap_fields)
    assert stats["steps"] == steps and bool(stats["frozen"]) == frozen
    assert stats["surface_rows"] == len(rows)
    lines, got = parse_surface(os.path.join(root, "output", "surface.dat"))
    ref_lines = [" ".join("%g" % x for x in r) for r in rows]
    assert len(lines) == len(ref_lines)
    differ = [i for i, (a, b) in enumerate(zip(lines, ref_lines)) if a != b]
    say("  rows", len(lines), "lines printed differently", len(differ))
    for i in differ:
        np.testing.assert_allclose(got[i], rows[i], rtol=1e-5, atol=1e-7 * np.abs(rows[i]).max())
    assert len(differ) <= 0.01 * len(lines) + 1
    # snapshots: file per field and FREQ time, %.8f values
    nx, ny, nz = geom["nx"], geom["ny"], geom["nz"]
    names = {O.E: "e", O.UX: "ux", O.PIXX: "pixx"}
    for tsnap, fields in snaps.items():
        for f, ref in fields.items():
            path = os.path.join(out, "%s_%.3f.dat" % (names[f], tsnap))
            d = read_snapshot(path)
            assert d.shape == (nx * ny * nz, 4)
            v = d[:, 3].reshape((nx, ny, nz), order="F")
            np.testing.assert_allclose(v, ref, rtol=0, atol=2e-8 + 1e-10 * np.abs(ref).max(),
                                       err_msg=path)
    return len(rows)


@pytest.mark.parametrize("ev", EVENTS, ids=[e[0] for e in EVENTS])
def test_run_loop_matches_restatement(ev, tmp_path):
    """gvh_run on type-10 input files: steps taken, the all-frozen stop,
    surface.dat and the FREQ snapshots equal the restated run loop driven by
    the oracle on the same files."""
    case, geom, Tfo = event_case(ev)
    name, nx, ny, nz, dx, dz, dt = ev[:7]
    say("run loop", name)
    state = smooth_state(case, 7)
    root = str(tmp_path / "root")
    out = str(tmp_path / "out")
    os.makedirs(out)
    write_ic_files(root, state, nx, ny, nz, dx, dz, case.bulk)
    nt = 300
    L = gpuvh.lattice(nx, ny, nz, dx, dx, dz, dt, nt=nt)
    IC = gpuvh.InitialConditionParameters()
    IC.initialConditionType = 10
    H = gpuvh.hydro(0.5, 0.2, T_fo_GeV=Tfo)
    o = gpuvh.options(case.eos, case.bulk, gpuvh.FP64, case.zeta_tpeak, table=case.table)
    stats = gpuvh.run(L, IC, H, root, out, o, gpuvh.run_options())
    say("  stats", stats)
    e_ic = np.float32(state[0] - 1e-3).astype(np.float64) + 1e-3
    orc_state = lambda P: (e_ic, O.eos_p_batch(P, e_ic)) + tuple(state[2:])  # noqa: E731
    check_run_against_restatement(case, geom, Tfo, orc_state, nt, stats, root, out)


def write_tmunu_file(root, cols, nx, ny, nz, dx, dz):
    """Type-11 input (InitialConditions.cpp:29-95): one row of 20 floats per
    cell, x y z e p ut ux uy un pitt pitx pity pitn pixx pixy pixn piyy piyn
    pinn Pi, x outermost and eta innermost; e holds e - 1e-3 (the reader adds
    1e-3 back).  cols: the 17 value columns as float32 arrays."""
    os.makedirs(os.path.join(root, "input"), exist_ok=True)
    os.makedirs(os.path.join(root, "output"), exist_ok=True)
    xs = (np.arange(nx) - (nx - 1) / 2.0) * dx
    ys = (np.arange(ny) - (ny - 1) / 2.0) * dx
    zs = (np.arange(nz) - (nz - 1) / 2.0) * dz
    with open(os.path.join(root, "input", "Tmunu.dat"), "w") as f:
        for i in range(nx):
            for j in range(ny):
                for k in range(nz):
                    vals = " ".join("%.9g" % c[i, j, k] for c in cols)
                    f.write("%.6f %.6f %.6f %s\n" % (xs[i], ys[j], zs[k], vals))


@pytest.mark.parametrize("ev", [EVENTS[2], EVENTS[3]], ids=[EVENTS[2][0], EVENTS[3][0]])
def test_run_loop_tmunu_file(ev, tmp_path):
    """gvh_run on a type-11 Tmunu.dat (InitialConditions.cpp:29-95) with flow,
    shear and bulk in every column: the run equals the restated run loop
    driven by the oracle from the file's float values.  The file's p is 0.2%
    above the EoS value, so a reader that recomputed p from e, dropped a
    column or shifted the columns would fail the comparison."""
    case, geom, Tfo = event_case(ev)
    name, nx, ny, nz, dx, dz, dt = ev[:7]
    say("run loop, Tmunu.dat", name)
    e = smooth_state(case, 11)[0]
    t0 = case.t0
    X, Y, Z = np.meshgrid((np.arange(nx) - (nx - 1) / 2.0) * dx,
                          (np.arange(ny) - (ny - 1) / 2.0) * dx,
                          (np.arange(nz) - (nz - 1) / 2.0) * dz, indexing="ij")
    ext = 0.5 * (nx - 1) * dx
    ux = 0.08 * np.tanh(X / ext) + 0.02 * np.sin(Y)
    uy = -0.06 * np.tanh(Y / ext)
    un = (0.01 * np.tanh(Z) / t0) if nz > 1 else np.zeros_like(X)
    ut = np.sqrt(1.0 + ux * ux + uy * uy + t0 * t0 * un * un)
    p_eos = case.pressure(e)
    p_file = 1.002 * p_eos
    w = 0.01 * p_eos
    pi10 = np.stack([0.3 * w, 0.1 * w * np.cos(X), -0.1 * w * np.sin(Y), 0.05 * w / t0,
                     w * (1 + 0.2 * np.tanh(X)), 0.2 * w, 0.05 * w / t0, -0.5 * w,
                     -0.04 * w / t0, -0.5 * w / (t0 * t0)])
    Pi = -0.02 * p_eos if case.bulk else np.zeros_like(e)
    cols = [np.float32(e - 1e-3), np.float32(p_file), np.float32(ut), np.float32(ux),
            np.float32(uy), np.float32(un)] + [np.float32(c) for c in pi10] + [np.float32(Pi)]
    root = str(tmp_path / "root")
    out = str(tmp_path / "out")
    os.makedirs(out)
    write_tmunu_file(root, cols, nx, ny, nz, dx, dz)
    nt = 120
    L = gpuvh.lattice(nx, ny, nz, dx, dx, dz, dt, nt=nt)
    IC = gpuvh.InitialConditionParameters()
    IC.initialConditionType = 11
    H = gpuvh.hydro(t0, 0.2, T_fo_GeV=Tfo)
    o = gpuvh.options(case.eos, case.bulk, gpuvh.FP64, case.zeta_tpeak, table=case.table)
    stats = gpuvh.run(L, IC, H, root, out, o, gpuvh.run_options())
    say("  stats", stats)
    f64 = [c.astype(np.float64) for c in cols]
    orc_state = (f64[0] + 1e-3, f64[1], f64[2], f64[3], f64[4], f64[5],
                 np.stack(f64[6:16]), f64[16] if case.bulk else np.zeros_like(e))
    nrows = check_run_against_restatement(case, geom, Tfo, orc_state, nt, stats, root, out)
    assert nrows > 0


def test_run_hydro_command_line(tmp_path, monkeypatch):
    """gvh_run_hydro with the reference's options reads rhic-conf style
    positional files from --config, runs in the working directory and writes
    the same surface as gvh_run; without -h it does nothing."""
    ev = EVENTS[0]
    case, geom, Tfo = event_case(ev)
    name, nx, ny, nz, dx, dz, dt = ev[:7]
    state = smooth_state(case, 7)
    root = tmp_path / "work"
    write_ic_files(str(root), state, nx, ny, nz, dx, dz, case.bulk)
    cfg = root / "rhic-conf"
    cfg.mkdir()
    (cfg / "lattice.properties").write_text(
        "numLatticePointsX=%d\nnumLatticePointsY=%d\nnumLatticePointsRapidity=%d\n"
        "numProperTimePoints=%d\nlatticeSpacingX=%g\nlatticeSpacingY=%g\n"
        "latticeSpacingRapidity=%g\nlatticeSpacingProperTime=%g" % (nx, ny, nz, 300, dx, dx, dz, dt))
    (cfg / "hydro.properties").write_text(
        "initialProperTimePoint=0.5\nshearViscosityToEntropyDensity=0.2\n"
        "freezeoutTemperatureGeV=%g\ninitializePimunuNavierStokes=0\nregulationTemperatureGeV=0" % Tfo)
    (cfg / "ic.properties").write_text(
        "initialConditionType=10\nnumberOfNucleonsPerNuclei=208\nnumberOfSourceFiles=0\n"
        "initialEnergyDensity=1\nscatteringCrossSectionNN=62\nimpactParameter=0\n"
        "fractionOfBinaryCollisions=0\nrapidityVariance=0.5\nrapidityMean=0\nsourceType=0")
    (root / "output").mkdir(exist_ok=True)
    o = gpuvh.options(case.eos, case.bulk, gpuvh.FP64, case.zeta_tpeak)
    ro = gpuvh.run_options(snapshots=False)
    monkeypatch.chdir(root)
    argv = [b"gpu-vh", b"--config", b"rhic-conf", b"-o", b"output", b"-h"]
    arr = (C.c_char_p * len(argv))(*argv)
    surf = root / "output" / "surface.dat"
    rc = gpuvh.lib().gvh_run_hydro(len(argv) - 1, arr, C.byref(o), C.byref(ro))  # no -h
    assert rc == 0 and not surf.exists()
    rc = gpuvh.lib().gvh_run_hydro(len(argv), arr, C.byref(o), C.byref(ro))
    assert rc == 0, gpuvh.lib().gvh_last_error()
    text_cli = surf.read_text()
    L = gpuvh.read_lattice_parameters(str(cfg))
    IC = gpuvh.read_ic_parameters(str(cfg))
    H = gpuvh.read_hydro_parameters(str(cfg))
    gpuvh.run(L, IC, H, str(root), str(root / "output"), o, ro)
    assert surf.read_text() == text_cli and len(text_cli) > 0


def test_freezeout_argument_errors():
    case = Case(6, 6, 1, 0.1, 0.1, 0.01, 0.5, 0.2, 0, gpuvh.EOS_CONFORMAL)
    ctx = case.context()
    with pytest.raises(gpuvh.GvhError):
        ctx.freezeout_record(1)          # not initialised
    with pytest.raises(gpuvh.GvhError):
        ctx.freezeout_init(2, 1.0)       # window too short
    ctx.freezeout_init(10, 1.0)
    with pytest.raises(gpuvh.GvhError):
        ctx.freezeout_record(1)          # no state yet
    ctx.close()
    line = Case(6, 1, 1, 0.1, 0.1, 0.01, 0.5, 0.2, 0, gpuvh.EOS_CONFORMAL).context()
    with pytest.raises(gpuvh.GvhError):
        line.freezeout_init(10, 1.0)     # 1+1D has no surface finder
    line.close()
