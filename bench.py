"""Benchmark of the KT viscous-hydro step on one MI355X.

Workload (default): the 3+1D event on a 201x201x101 lattice (dx=dy=deta=0.1,
tau0=0.5 fm/c, dtau=0.01), synthetic QCD-like tabulated EoS (2000 entries in
e), shear eta/s=0.16 plus bulk zeta/s peaked at T=0.16 GeV, initial e from
200 seeded transverse hotspots times the plateau-plus-Gaussian-tail profile
in eta, plain FP64.  A step is one full KT step (both RK stages) over the
whole lattice with the state resident in HBM.

Prints one JSON line: cell-updates/s (interior cells x steps / wall time),
the roofline of the stage kernel (HBM-bound stencil), the CPU baseline (the
FP64 oracle on the host cores, a bounded sample of the same event) and the
mixed-precision form on the same event (speed and error against FP64).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gpu-vh_amd"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0


def log(*a):
    """Progress on stderr (stdout carries only the JSON line)."""
    print("bench:", *a, file=sys.stderr, flush=True)


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")


def algorithmic_bytes_per_cell_update(bulk, precision):
    """Bytes moved per cell per KT step with perfect on-chip neighbour reuse:
    stage 1 reads q^n, e, p, u^n, u^{n-1} and writes q*, e*, p*, u*; stage 2
    reads q*, e*, p*, u*, u^n, q^n and writes q^{n+1}, e, p, u^{n+1}: five
    conserved-buffer accesses, four (e, p) and six velocity accesses.  A
    conserved buffer is 4 FP64 T^{tau mu} plus 10 (11 with bulk) pi/Pi values
    in FP64, or in FP32 in the mixed form."""
    npi = 11 if bulk else 10
    q_bytes = 4 * 8 + npi * (4 if precision == "mixed" else 8)
    return 5 * q_bytes + 4 * 2 * 8 + 6 * 4 * 8


def build_event(args):
    from gpuvh import ic

    return ic.c2_event(args.nx, args.ny, args.nz, 0.1, 0.1, 0.1, seed=args.seed)


def make_context(args, precision, eos, stream):
    import gpuvh

    L = gpuvh.lattice(args.nx, args.ny, args.nz, 0.1, 0.1, 0.1, args.dt)
    H = gpuvh.hydro(args.t0, 0.16, T_fo_GeV=0.15)
    opts = gpuvh.options(eos, True, precision, 0.16 / gpuvh.HBARC, stream=stream)
    return gpuvh.Context(L, H, opts)


def init_state(ctx, args, e):
    from gpuvh import ic

    ut, ux, uy, un, pi10, Pi = ic.at_rest(e.shape)
    p = ctx.equilibrium_pressure(e)
    ctx.set_initial_state(args.t0, e, p, ut, ux, uy, un, pi10, Pi)


def newton_fraction(ctx):
    """Fraction of interior cells whose root-finds take the Newton branch:
    the inversions of a cell use its own e as the guess, and a guess
    <= 0.1 fm^-4 selects the closed form e = M0 - M/M0
    (EnergyMomentumTensor.cu:87-88)."""
    import gpuvh

    return float(np.mean(ctx.get(gpuvh.E) > 0.1))


VALIDITY_BYTES_PER_CELL = 31 * 8  # reads e, p, u, u_prev, 10 pi, Pi; writes 10 values


def time_validity(ctx, args, stream, ncells, reps=10):
    """The validity kernel (gvh_check_validity) on the stepped state: HIP
    events on the context's stream around reps launches."""
    import torch

    t = args.t0 + (args.warmup + args.steps) * args.dt
    ctx.validity_enable()
    ctx.check_validity(t)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(reps):
        ctx.check_validity(t)
    b.record(stream)
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / reps
    gbs = VALIDITY_BYTES_PER_CELL * ncells / (ms * 1e-3) / 1e9
    log("validity kernel: %.4f ms, %.0f GB/s algorithmic" % (ms, gbs))
    return {"kernel": "validity_kernel (gvh_check_validity)", "ms": round(ms, 4),
            "bytes_per_cell": VALIDITY_BYTES_PER_CELL, "bound": "hbm",
            "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4)}


def measured_traffic(workload_key):
    """HBM bytes per step of the stage kernels from rocprofv3 FETCH_SIZE and
    WRITE_SIZE runs, calibrated against a copy kernel (profiles/traffic.json),
    or None when no measurement of this workload is on record."""
    try:
        with open(TRAFFIC_FILE) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    return rec.get(workload_key)


def mixed_error(args, e, stream, nsteps):
    """Max relative error of the mixed form against FP64 after nsteps on the
    same event: per cell with a floor of 1% of the field's maximum (u^x, u^y,
    u^eta and pi cross zero), and the Lorentz-norm error max|du|/u^tau.
    Cells where either form is not finite are counted, not compared."""
    import gpuvh

    ctxs = {}
    for prec in (gpuvh.FP64, gpuvh.MIXED):
        c = make_context(args, prec, gpuvh.EOS_TABLE, stream)
        init_state(c, args, e)
        c.steps(args.t0, 0, nsteps)
        c.synchronize()
        ctxs[prec] = c
    out = {}
    groups = {"e": [gpuvh.E], "u": [gpuvh.UT, gpuvh.UX, gpuvh.UY, gpuvh.UN],
              "pi": [gpuvh.PITT, gpuvh.PITX, gpuvh.PITY, gpuvh.PITN, gpuvh.PIXX, gpuvh.PIXY,
                     gpuvh.PIXN, gpuvh.PIYY, gpuvh.PIYN, gpuvh.PINN], "Pi": [gpuvh.PI]}
    du = 0.0
    nonfinite = 0
    for name, fields in groups.items():
        worst = 0.0
        for f in fields:
            a, b = ctxs[gpuvh.FP64].get(f), ctxs[gpuvh.MIXED].get(f)
            ok = np.isfinite(a) & np.isfinite(b)
            nonfinite = max(nonfinite, int((~ok).sum()))
            a, b = a[ok], b[ok]
            scale = float(np.abs(a).max()) if a.size else 0.0
            if scale == 0:
                continue
            worst = max(worst, float((np.abs(b - a) / np.maximum(np.abs(a), 1e-2 * scale)).max()))
            if name == "u" and f != gpuvh.UT:
                du = max(du, float(np.abs(b - a).max()))
        out[name] = worst
    ut = ctxs[gpuvh.FP64].get(gpuvh.UT)
    out["u_lorentz"] = du / float(np.nanmin(np.abs(ut)))
    # the mixed form's accuracy criterion (1e-4 in e and u^mu at freeze-out):
    # the fluid at or above the freeze-out energy
    # density (T >= 0.15 GeV); and where the global errors come from, by the
    # FP64 energy density of the cell
    e_fo = ctxs[gpuvh.FP64].energy_density_of_T(0.15 / gpuvh.HBARC)
    ref = {f: ctxs[gpuvh.FP64].get(f) for f in groups["u"] + groups["pi"] + [gpuvh.E]}
    got = {f: ctxs[gpuvh.MIXED].get(f) for f in ref}
    tau = args.t0 + nsteps * args.dt

    def region(m):
        r = {"cells": int(m.sum())}
        if not m.any():
            return r
        r["e"] = float("%.3e" % np.max(np.abs(got[gpuvh.E][m] - ref[gpuvh.E][m]) / ref[gpuvh.E][m]))
        r["u"] = float("%.3e" % max(
            float(np.max(np.abs(got[f][m] - ref[f][m]) * (tau if f == gpuvh.UN else 1.0)
                         / ref[gpuvh.UT][m])) for f in groups["u"]))
        pscale = max(float(np.max(np.abs(ref[f][m]))) for f in groups["pi"])
        if pscale > 0:
            r["pi"] = float("%.3e" % (max(float(np.max(np.abs(got[f][m] - ref[f][m])))
                                          for f in groups["pi"]) / pscale))
        return r

    eref = ref[gpuvh.E]
    out["at_or_above_e_fo"] = region(eref >= e_fo)
    out["by_energy_density"] = {"e >= e_fo": out["at_or_above_e_fo"],
                                "0.1 <= e < e_fo": region((eref >= 0.1) & (eref < e_fo)),
                                "e < 0.1 (vacuum branch)": region(eref < 0.1)}
    out["nonfinite_cells"] = nonfinite
    out["diagnostics_fp64"] = ctxs[gpuvh.FP64].diagnostics()
    out["diagnostics_mixed"] = ctxs[gpuvh.MIXED].diagnostics()
    for c in ctxs.values():
        c.close()
    return out


def batch_workload(n_events, seed, kind):
    """Seeded events of the batched (kind "batch": 2+1D, transverse grids odd
    121..261, 300..800 steps, 50..300 hotspots) or serving (kind "serve": half
    2+1D 121^2..201^2, half 3+1D 121^2..201^2 x 21..51, through freeze-out at
    T_fo = 0.15 GeV) workloads; event k uses seed + k."""
    import gpuvh
    from gpuvh import ic

    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_events):
        if kind == "batch":
            nx = 2 * int(rng.integers(60, 131)) + 1
            nz, nt = 1, int(rng.integers(300, 801))
            Tfo = 0.0  # fixed step count: never frozen
        else:
            nx = 2 * int(rng.integers(60, 101)) + 1
            nz = 1 if k % 2 == 0 else 2 * int(rng.integers(10, 26)) + 1
            nt, Tfo = 4000, 0.15
        nh = int(rng.integers(50, 301))
        e = ic.serving_event(nx, nx, nz, 0.1, 0.1, seed + k, nh)
        L = gpuvh.lattice(nx, nx, nz, 0.1, 0.1, 0.1, 0.01, nt=nt)
        H = gpuvh.hydro(0.5, 0.16, T_fo_GeV=Tfo)
        out.append((L, H, e))
    return out


def run_batch_mode(args, n_events):
    """2+1D events in one batch (mixed precision, autotuned) against the
    same events one by one in plain FP64: throughput and error.  Returns the
    summary line as a dict."""
    import torch
    import gpuvh

    table = gpuvh.EosTable()
    events = batch_workload(n_events, args.seed, "batch")
    cells = sum(L.numLatticePointsX * L.numLatticePointsY * (L.numProperTimePoints + 1)
                for L, H, e in events)
    log("batch: %d events, %.3g cell-updates" % (len(events), cells))
    res = {}
    keep = min(4, len(events))
    kept = {}
    for prec, name in ((gpuvh.MIXED, "mixed"), (gpuvh.FP64, "fp64")):
        opts = gpuvh.options(gpuvh.EOS_TABLE, True, prec, 0.16 / gpuvh.HBARC, table=table)
        b = gpuvh.Batch(opts, max_events=len(events), autotune=True)
        ids = [b.submit(L, H, e, keep_state=(k < keep)) for k, (L, H, e) in enumerate(events)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b.run()
        el = time.perf_counter() - t0
        res["batch_" + name] = el
        kept[name] = [{f: b.field(ids[k], f) for f in (gpuvh.E, gpuvh.UX, gpuvh.UY)}
                      for k in range(keep)]
        b.close()
        log("batch %s: %.3f s, %.3g cell-updates/s" % (name, el, cells / el))
    # one by one in plain FP64 on contexts of their own
    opts = gpuvh.options(gpuvh.EOS_TABLE, True, gpuvh.FP64, 0.16 / gpuvh.HBARC, table=table)
    from gpuvh import ic
    el = 0.0
    alone = []
    for k, (L, H, e) in enumerate(events):
        ctx = gpuvh.Context(L, H, opts)
        ut, ux, uy, un, pi10, Pi = ic.at_rest(e.shape)
        ctx.set_initial_state(0.5, e, ctx.equilibrium_pressure(e), ut, ux, uy, un)
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.steps(0.5, 0, L.numProperTimePoints + 1)
        ctx.synchronize()
        el += time.perf_counter() - t0
        if k < keep:
            alone.append({f: ctx.get(f) for f in (gpuvh.E, gpuvh.UX, gpuvh.UY)})
        ctx.close()
    res["alone_fp64"] = el
    log("one by one fp64: %.3f s, %.3g cell-updates/s" % (el, cells / el))
    errs = {"e": 0.0, "u": 0.0}
    for k in range(keep):
        for f, g in ((gpuvh.E, "e"), (gpuvh.UX, "u"), (gpuvh.UY, "u")):
            a, m = alone[k][f], kept["mixed"][k][f]
            den = np.maximum(np.abs(a), 1e-2 * np.abs(a).max())
            errs[g] = max(errs[g], float(np.max(np.abs(m - a) / den)))
    out = {"metric": "cell-updates/s, 64 batched 2+1D events", "value": round(cells / res["batch_mixed"], 1),
           "unit": "cell-updates/s", "n_gpus": 1, "higher_is_better": True, "dtype": "f32/f64 mixed",
           "data": "synthetic",
           "config": {"workload": "%d 2+1D events, transverse grids 121..261 (odd), dx 0.1, 300..800 "
                                  "steps, 50..300 hotspots on a Gaussian base, tabulated EoS, "
                                  "shear+bulk, autotuned" % len(events), "seed": args.seed},
           "events_per_s": round(len(events) / res["batch_mixed"], 2),
           "batch_mixed_s": round(res["batch_mixed"], 3), "batch_fp64_s": round(res["batch_fp64"], 3),
           "one_by_one_fp64_s": round(res["alone_fp64"], 3),
           "speedup_vs_one_by_one_fp64": round(res["alone_fp64"] / res["batch_mixed"], 3),
           "batch_fp64_cell_updates_per_s": round(cells / res["batch_fp64"], 1),
           "mixed_error_vs_fp64_alone": {k: float("%.3e" % v) for k, v in errs.items()},
           "error_metric": "max per-cell |mixed - fp64| / max(|fp64|, 1e-2 max|fp64|) over the "
                           "first %d events at their last step" % keep}
    return out


def run_serve_mode(args, n_events):
    """Serving loop: a Poisson stream of events at 80% of the measured
    capacity into a continuously batched engine; events/s and latency.
    Returns the summary line as a dict."""
    import gpuvh

    table = gpuvh.EosTable()
    events = batch_workload(n_events, args.seed, "serve")
    report = {}
    for prec, name in ((gpuvh.MIXED, "mixed"), (gpuvh.FP64, "fp64")):
        opts = gpuvh.options(gpuvh.EOS_TABLE, True, prec, 0.16 / gpuvh.HBARC, table=table)
        # capacity: everything submitted at once
        b = gpuvh.Batch(opts, max_events=len(events), autotune=True)
        for (L, H, e) in events:
            b.submit(L, H, e, surface=True)
        t0 = time.perf_counter()
        b.run()
        cap_s = time.perf_counter() - t0
        res = b.poll(len(events))
        b.close()
        capacity = len(events) / cap_s
        log("serve %s: capacity %.2f events/s (%d events in %.2f s)" % (name, capacity, len(res), cap_s))
        # Poisson arrivals at 80% of capacity
        rate = 0.8 * capacity
        rng = np.random.default_rng(args.seed + 1)
        arrivals = np.cumsum(rng.exponential(1.0 / rate, len(events)))
        b = gpuvh.Batch(opts, max_events=len(events), autotune=True)
        start = time.perf_counter()
        nsub, done = 0, []
        while len(done) < len(events):
            now = time.perf_counter() - start
            while nsub < len(events) and arrivals[nsub] <= now:
                L, H, e = events[nsub]
                b.submit(L, H, e, surface=True)
                nsub += 1
            running = b.step()
            done += b.poll(len(events))
            if running == 0 and nsub < len(events):
                time.sleep(max(0.0, arrivals[nsub] - (time.perf_counter() - start)))
        makespan = time.perf_counter() - start
        b.close()
        lat = np.array([r["t_done"] - r["t_submit"] for r in done])
        rows = int(sum(r["surface_rows"] for r in done))
        report[name] = {"capacity_events_per_s": round(capacity, 3), "offered_rate": round(rate, 3),
                        "events_per_s": round(len(done) / makespan, 3),
                        "latency_p50_s": round(float(np.percentile(lat, 50)), 3),
                        "latency_p99_s": round(float(np.percentile(lat, 99)), 3),
                        "latency_mean_s": round(float(lat.mean()), 3),
                        "steps_mean": round(float(np.mean([r["steps"] for r in done])), 1),
                        "frozen": int(sum(r["frozen"] for r in done)), "surface_rows": rows}
        log("serve %s: %s" % (name, report[name]))
    out = {"metric": "events/s at 80% of capacity, p50/p99 latency", "value": report["mixed"]["events_per_s"],
           "unit": "events/s", "n_gpus": 1, "higher_is_better": True, "dtype": "f32/f64 mixed",
           "data": "synthetic",
           "config": {"workload": "%d events, half 2+1D 121^2..201^2, half 3+1D 121^2..201^2 x 21..51, "
                                  "hotspots on a Gaussian base, tabulated EoS, shear+bulk, through "
                                  "freeze-out at T_fo 0.15 GeV with surfaces, Poisson arrivals"
                                  % len(events), "seed": args.seed},
           "mixed": report["mixed"], "fp64": report["fp64"]}
    return out


def cpu_info():
    """CPU model and core counts of this host from /proc/cpuinfo."""
    model, phys, logical = None, set(), 0
    try:
        with open("/proc/cpuinfo") as f:
            pid = core = None
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and model is None:
                    model = v
                elif k == "processor":
                    logical += 1
                elif k == "physical id":
                    pid = v
                elif k == "core id":
                    core = v
                    phys.add((pid, core))
    except OSError:
        pass
    return {"model": model, "logical_cpus": logical or os.cpu_count(),
            "physical_cores": len(phys) or None,
            "affinity_cpus": len(os.sched_getaffinity(0))}


def cpu_baseline(args, e_full, threads):
    """Time the oracle (FP64 CPU restatement of the reference KT step, test
    infrastructure) on the host cores, threads pinned close: KT steps of the
    same event on the full lattice, after one untimed step.  Cell-updates/s."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    os.environ["OMP_NUM_THREADS"] = str(threads)
    os.environ["OMP_PROC_BIND"] = "close"
    os.environ.setdefault("OMP_PLACES", "cores")
    import gpuvh
    import oracle_lib as O
    from gpuvh import ic

    tab = gpuvh.EosTable()
    P = O.make_params(args.nx, args.ny, args.nz, 0.1, 0.1, 0.1, args.dt, 0.16, 1, O.EOS_TABLE,
                      zeta_tpeak=0.16 / gpuvh.HBARC, table=tab)
    orc = O.Oracle(P)
    ut, ux, uy, un, pi10, Pi = ic.at_rest(e_full.shape)
    p = O.eos_p_batch(P, e_full)
    orc.set_primitives(args.t0, e_full, p, ut, ux, uy, un, pi10, Pi)
    orc.steps(args.t0, 0, 1)
    nsteps = args.cpu_steps
    t0 = time.perf_counter()
    orc.steps(args.t0, 1, nsteps)
    el = time.perf_counter() - t0
    orc.close()
    cells = args.nx * args.ny * args.nz
    return cells * nsteps / el, el, nsteps


def dense_event(args):
    """The benchmark lattice with no cell near vacuum (ic.c2_dense_event):
    the 200 hotspots on a 15 GeV/fm^3 transverse Gaussian base (sigma 6 fm)
    and an eta plateau wider than the lattice; it runs 500 steps."""
    from gpuvh import ic

    return ic.c2_dense_event(args.nx, args.ny, args.nz, 0.1, 0.1, 0.1, seed=args.seed)


def run_form(args, form, eos, stream, e, ncells, want_validity=False):
    """One form on one event: timed steps, kernel times from the same steps,
    the Newton-branch fraction at the start and end of the timed steps."""
    import gpuvh

    prec = {"fp64": gpuvh.FP64, "mixed": gpuvh.MIXED,
            "mixed_f64_roots": gpuvh.MIXED_F64_ROOTS}[form]
    ctx = make_context(args, prec, eos, stream.cuda_stream)
    init_state(ctx, args, e)
    ctx.steps(args.t0, 0, args.warmup)
    nf0 = newton_fraction(ctx)
    import torch
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    s1, s2, gpu_ms = ctx.steps_timed(args.t0, args.warmup, args.steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    s1, s2 = s1 / args.steps, s2 / args.steps
    nf1 = newton_fraction(ctx)
    diag = ctx.diagnostics()
    validity = time_validity(ctx, args, stream, ncells) if want_validity else None
    ctx.close()
    bpc = algorithmic_bytes_per_cell_update(True, "fp64" if form == "fp64" else "mixed")
    kt_ms = s1 + s2
    log("%s: %.4f ms/step, stage kernels %.4f + %.4f ms, Newton branch %.3f -> %.3f"
        % (form, 1e3 * elapsed / args.steps, s1, s2, nf0, nf1))
    return {"ms_per_step": 1e3 * elapsed / args.steps, "value": ncells * args.steps / elapsed,
            "stage_ms": [s1, s2], "kernel_ms": kt_ms, "gpu_ms_per_step": gpu_ms / args.steps,
            "bytes_per_cell_update": bpc,
            "achieved_gbs": bpc * ncells / (kt_ms * 1e-3) / 1e9,
            "newton_fraction": [round(nf0, 4), round(nf1, 4)], "diagnostics": diag,
            "validity": validity}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nx", type=int, default=201)
    ap.add_argument("--ny", type=int, default=201)
    ap.add_argument("--nz", type=int, default=101)
    ap.add_argument("--dt", type=float, default=0.01)
    ap.add_argument("--t0", type=float, default=0.5)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--precision", choices=["fp64", "mixed", "mixed_f64_roots"], default="fp64",
                    help="form of the headline line (the other is reported beside it)")
    ap.add_argument("--eos", choices=["table", "wb", "conformal"], default="table",
                    help="equation of state (the workload's is the 2000-node table)")
    ap.add_argument("--variant", type=int, default=None,
                    help="stage-kernel arrangement (library default if unset)")
    ap.add_argument("--error-steps", type=int, default=50,
                    help="steps after which mixed is compared with FP64 (0: skip)")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision report")
    ap.add_argument("--no-dense", action="store_true", help="skip the dense-event line")
    ap.add_argument("--cpu-steps", type=int, default=12)
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="oracle threads (default: OMP_NUM_THREADS, else the CPUs of this process)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch", action="store_true",
                    help="only the batched 2+1D events against one-by-one FP64 (own line)")
    ap.add_argument("--serve", action="store_true",
                    help="only the serving loop with Poisson arrivals (own line)")
    ap.add_argument("--events", type=int, default=64, help="events of --batch / --serve")
    ap.add_argument("--batch-events", type=int, default=64,
                    help="events of the batch summary in the default line (0: skip)")
    ap.add_argument("--serve-events", type=int, default=24,
                    help="events of the serving summary in the default line (0: skip)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a 1 GiB device copy (reference kernel for HBM counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (the KT step of one event on "
              "one MI355X); --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    if args.variant is not None:
        os.environ["GVH_VARIANT"] = str(args.variant)
    log("start (%dx%dx%d, %s, %s)" % (args.nx, args.ny, args.nz, args.eos, args.precision))
    import torch
    import gpuvh

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    stream = torch.cuda.current_stream()
    if args.batch:
        print(json.dumps(run_batch_mode(args, args.events)))
        return
    if args.serve:
        print(json.dumps(run_serve_mode(args, args.events)))
        return

    if args.copy_calibration:
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    e = build_event(args)
    log("event built")
    eos = {"table": gpuvh.EOS_TABLE, "wb": gpuvh.EOS_WB, "conformal": gpuvh.EOS_CONFORMAL}[args.eos]
    ncells = args.nx * args.ny * args.nz
    forms = [args.precision] + ([] if args.no_mixed else
                                [p for p in ("fp64", "mixed", "mixed_f64_roots")
                                 if p != args.precision])
    res = {form: run_form(args, form, eos, stream, e, ncells, want_validity=(form == "fp64"))
           for form in forms}

    head = res[args.precision]
    wkey = "%dx%dx%d/%s/%s/v%s" % (args.nx, args.ny, args.nz, args.eos, args.precision,
                                   os.environ.get("GVH_VARIANT", "default"))
    rec = measured_traffic(wkey)
    roofline = {"bound": "hbm", "achieved": round(head["achieved_gbs"], 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(head["achieved_gbs"] / HBM_PEAK_GBS, 4),
                "traffic": rec["bytes_per_step"] if rec else None,
                "traffic_unit": "HBM bytes per KT step (the two stage launches), rocprofv3 "
                                "FETCH_SIZE x fetch_factor + WRITE_SIZE x write_factor",
                "traffic_stage_launch_bytes": rec["stage_bytes"] if rec else None,
                "traffic_source": rec["source"] if rec else None,
                "kernel": "kt_flux_kernel (one launch per RK stage)",
                "stage_ms": [round(v, 4) for v in head["stage_ms"]],
                "timing": "HIP events around every stage kernel of the %d timed steps" % args.steps,
                "bytes_per_cell_update": head["bytes_per_cell_update"],
                "cells_per_launch": ncells}

    mixed = None
    if not args.no_mixed:
        other = "mixed" if args.precision == "fp64" else "fp64"
        m = res["mixed"]
        mixed = {"ms_per_step": round(m["ms_per_step"], 4), "value": round(m["value"], 1),
                 "speedup_vs_fp64": round(res["fp64"]["ms_per_step"] / m["ms_per_step"], 3),
                 "roofline_frac": round(m["achieved_gbs"] / HBM_PEAK_GBS, 4),
                 "stage_ms": [round(v, 4) for v in m["stage_ms"]],
                 "bytes_per_cell_update": m["bytes_per_cell_update"], "reported_beside": other,
                 "split": "FP32 stencil, limiter, flux, sources, pi/Pi storage and the 12 "
                          "half-site root-finds; FP64 T^{tau mu} accumulation, the cell's own "
                          "root-find and its EoS"}
        if "mixed_f64_roots" in res:
            m2 = res["mixed_f64_roots"]
            mixed["f64_roots"] = {
                "split": "as mixed, with every root-find and its EoS lookups in FP64",
                "ms_per_step": round(m2["ms_per_step"], 4),
                "speedup_vs_fp64": round(res["fp64"]["ms_per_step"] / m2["ms_per_step"], 3),
                "stage_ms": [round(v, 4) for v in m2["stage_ms"]]}
        if args.error_steps > 0 and args.eos == "table":
            log("mixed vs fp64 after %d steps" % args.error_steps)
            errs = mixed_error(args, e, stream.cuda_stream, args.error_steps)
            mixed["error_vs_fp64"] = {k: (float("%.3e" % v) if isinstance(v, float) else v)
                                      for k, v in errs.items()}
            mixed["error_steps"] = args.error_steps
            mixed["error_metric"] = ("max per-cell |mixed - fp64| / max(|fp64|, 1e-2 max|fp64|) "
                                     "per field group; u_lorentz = max|du^i| / min u^tau; "
                                     "at_or_above_e_fo: on cells with e >= e(T = 0.15 GeV), "
                                     "e relative, u max|du^mu|/u^tau, pi relative to max|pi|")

    dense = None
    if not args.no_dense and args.eos == "table":
        ed = dense_event(args)
        log("dense event built")
        dres = {form: run_form(args, form, eos, stream, ed, ncells)
                for form in (["fp64"] + ([] if args.no_mixed else ["mixed"]))}
        d = dres["fp64"]
        dense = {"workload": "the same lattice, 200 hotspots on a 15 GeV/fm^3 Gaussian base "
                             "(sigma 6 fm), eta plateau wider than the lattice (runs 500 steps)",
                 "value": round(d["value"], 1), "ms_per_step": round(d["ms_per_step"], 4),
                 "stage_ms": [round(v, 4) for v in d["stage_ms"]],
                 "roofline_frac": round(d["achieved_gbs"] / HBM_PEAK_GBS, 4),
                 "newton_fraction": d["newton_fraction"], "diagnostics": d["diagnostics"]}
        if "mixed" in dres:
            dense["mixed_ms_per_step"] = round(dres["mixed"]["ms_per_step"], 4)
            dense["mixed_speedup_vs_fp64"] = round(d["ms_per_step"] / dres["mixed"]["ms_per_step"], 3)

    batch = serve = None
    if args.batch_events > 0 and args.eos == "table":
        log("batch summary: %d events" % args.batch_events)
        batch = run_batch_mode(args, args.batch_events)
    if args.serve_events > 0 and args.eos == "table":
        log("serve summary: %d events" % args.serve_events)
        serve = run_serve_mode(args, args.serve_events)

    cpu = None
    if not args.no_cpu:
        env_threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
        threads = args.cpu_threads or env_threads or len(os.sched_getaffinity(0))
        info = cpu_info()
        log("cpu baseline: %d oracle steps on %d threads (%s)" % (args.cpu_steps, threads, info))
        v, el, ns = cpu_baseline(args, e, threads)
        log("cpu baseline: %.3g cell-updates/s" % v)
        cpu = {"value": round(v, 1), "unit": "cell-updates/s", "cores": threads,
               "kind": "port",
               "sample": "oracle (FP64 CPU restatement of the reference KT step, OpenMP, "
                         "OMP_PROC_BIND=close) on the full %dx%dx%d event, %d KT steps after one "
                         "untimed step, %.1f s" % (args.nx, args.ny, args.nz, ns, el),
               "cpu": info,
               "threads_note": "threads = OMP_NUM_THREADS of the environment (the host's CPU "
                               "share for one GPU) unless --cpu-threads is given"}

    out = {
        "metric": "cell-updates/s per KT step",
        "value": round(head["value"], 1),
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32/f64 mixed",
        "data": "synthetic",
        "config": {"workload": "3+1D KT step %dx%dx%d, %s EoS, shear+bulk, %s"
                               % (args.nx, args.ny, args.nz,
                                  "tabulated" if args.eos == "table" else args.eos, args.precision),
                   "lattice": [args.nx, args.ny, args.nz], "dtau": args.dt, "tau0": args.t0,
                   "eos": "tabulated 2000 nodes" if args.eos == "table" else args.eos,
                   "eta_over_s": 0.16, "zeta_peak_GeV": 0.16, "seed": args.seed,
                   "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "newton_fraction": head["newton_fraction"],
        "mixed": mixed,
        "dense": dense,
        "batch": batch,
        "serve": serve,
        "validity": res["fp64"]["validity"] if "fp64" in res else None,
        "diagnostics": head["diagnostics"],
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
