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


def time_steps(ctx, args):
    """Warm-up, then args.steps timed steps (wall clock, synchronized on both
    sides), then args.timed_steps steps with HIP events around each stage
    kernel.  Returns (seconds, [stage-1 ms], [stage-2 ms])."""
    import torch

    n = 0
    ctx.steps(args.t0, n, args.warmup)
    n += args.warmup
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    ctx.steps(args.t0, n, args.steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    n += args.steps
    s1, s2 = [], []
    for _ in range(args.timed_steps):
        a, b = ctx.step_timed(args.t0 + n * args.dt)
        n += 1
        s1.append(a)
        s2.append(b)
    return elapsed, s1, s2


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
    u^eta and pi cross zero), and the Lorentz-norm error max|du|/u^tau."""
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
    for name, fields in groups.items():
        worst = 0.0
        for f in fields:
            a, b = ctxs[gpuvh.FP64].get(f), ctxs[gpuvh.MIXED].get(f)
            scale = np.abs(a).max()
            if scale == 0:
                continue
            worst = max(worst, float((np.abs(b - a) / np.maximum(np.abs(a), 1e-2 * scale)).max()))
            if name == "u" and f != gpuvh.UT:
                du = max(du, float(np.abs(b - a).max()))
        out[name] = worst
    ut = ctxs[gpuvh.FP64].get(gpuvh.UT)
    out["u_lorentz"] = du / float(np.abs(ut).min())
    out["diagnostics_fp64"] = ctxs[gpuvh.FP64].diagnostics()
    out["diagnostics_mixed"] = ctxs[gpuvh.MIXED].diagnostics()
    for c in ctxs.values():
        c.close()
    return out


def cpu_baseline(args, e_full, threads):
    """Time the oracle (FP64 CPU restatement of the reference KT step, test
    infrastructure) on the host cores: KT steps of the same event on the full
    lattice.  Cell-updates/s."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    os.environ["OMP_NUM_THREADS"] = str(threads)
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
    nsteps = args.cpu_steps
    t0 = time.perf_counter()
    orc.steps(args.t0, 0, nsteps)
    el = time.perf_counter() - t0
    orc.close()
    cells = args.nx * args.ny * args.nz
    return cells * nsteps / el, el, nsteps


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
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64",
                    help="form of the headline line (the other is reported beside it)")
    ap.add_argument("--eos", choices=["table", "wb", "conformal"], default="table",
                    help="equation of state (the workload's is the 2000-node table)")
    ap.add_argument("--variant", type=int, default=None,
                    help="stage-kernel arrangement (library default if unset)")
    ap.add_argument("--timed-steps", type=int, default=5,
                    help="extra steps with HIP events around each stage kernel")
    ap.add_argument("--error-steps", type=int, default=100,
                    help="steps after which mixed is compared with FP64 (0: skip)")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision report")
    ap.add_argument("--cpu-steps", type=int, default=8)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
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
                                [p for p in ("fp64", "mixed") if p != args.precision])
    res = {}
    diag = None
    for form in forms:
        prec = gpuvh.FP64 if form == "fp64" else gpuvh.MIXED
        ctx = make_context(args, prec, eos, stream.cuda_stream)
        init_state(ctx, args, e)
        elapsed, s1, s2 = time_steps(ctx, args)
        if form == args.precision:
            diag = ctx.diagnostics()
        ctx.close()
        bpc = algorithmic_bytes_per_cell_update(True, form)
        kt_ms = float(np.mean(s1) + np.mean(s2))
        log("%s: %.4f ms/step, stage kernels %.4f + %.4f ms"
            % (form, 1e3 * elapsed / args.steps, np.mean(s1), np.mean(s2)))
        res[form] = {"ms_per_step": 1e3 * elapsed / args.steps,
                     "value": ncells * args.steps / elapsed,
                     "stage_ms": [float(np.mean(s1)), float(np.mean(s2))],
                     "kernel_ms": kt_ms, "bytes_per_cell_update": bpc,
                     "achieved_gbs": bpc * ncells / (kt_ms * 1e-3) / 1e9}

    head = res[args.precision]
    wkey = "%dx%dx%d/%s/%s/v%s" % (args.nx, args.ny, args.nz, args.eos, args.precision,
                                   os.environ.get("GVH_VARIANT", "default"))
    traffic = measured_traffic(wkey)
    roofline = {"bound": "hbm", "achieved": round(head["achieved_gbs"], 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(head["achieved_gbs"] / HBM_PEAK_GBS, 4),
                "traffic": traffic,
                "kernel": "kt_flux_kernel (one launch per RK stage)",
                "stage_ms": [round(v, 4) for v in head["stage_ms"]],
                "bytes_per_cell_update": head["bytes_per_cell_update"],
                "cells_per_launch": ncells}

    mixed = None
    if not args.no_mixed:
        other = "mixed" if args.precision == "fp64" else "fp64"
        m = res["mixed"]
        mixed = {"ms_per_step": round(m["ms_per_step"], 4), "value": round(m["value"], 1),
                 "speedup_vs_fp64": round(res["fp64"]["ms_per_step"] / m["ms_per_step"], 3),
                 "roofline_frac": round(m["achieved_gbs"] / HBM_PEAK_GBS, 4),
                 "bytes_per_cell_update": m["bytes_per_cell_update"], "reported_beside": other}
        if args.error_steps > 0 and args.eos == "table":
            log("mixed vs fp64 after %d steps" % args.error_steps)
            errs = mixed_error(args, e, stream.cuda_stream, args.error_steps)
            mixed["error_vs_fp64"] = {k: (float("%.3e" % v) if isinstance(v, float) else v)
                                      for k, v in errs.items()}
            mixed["error_steps"] = args.error_steps
            mixed["error_metric"] = ("max per-cell |mixed - fp64| / max(|fp64|, 1e-2 max|fp64|) "
                                     "per field group; u_lorentz = max|du^i| / min u^tau")

    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        log("cpu baseline: %d oracle steps on %d threads" % (args.cpu_steps, threads))
        v, el, ns = cpu_baseline(args, e, threads)
        log("cpu baseline: %.3g cell-updates/s" % v)
        cpu = {"value": round(v, 1), "unit": "cell-updates/s", "cores": threads,
               "kind": "port",
               "sample": "oracle (FP64 CPU restatement of the reference KT step, OpenMP) on "
                         "the full %dx%dx%d event, %d KT step(s), %.1f s"
                         % (args.nx, args.ny, args.nz, ns, el)}

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
        "mixed": mixed,
        "diagnostics": diag,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
