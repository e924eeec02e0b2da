"""Benchmark of the KT viscous-hydro step on one MI355X.

Workload (default): the 3+1D event on a 201x201x101 lattice (dx=dy=deta=0.1,
tau0=0.5 fm/c, dtau=0.01), synthetic QCD-like tabulated EoS (2000 entries in
e), shear eta/s=0.16 plus bulk zeta/s peaked at T=0.16 GeV, initial e from
200 seeded transverse hotspots times the plateau-plus-Gaussian-tail profile
in eta, plain FP64.  A step is one full KT step (both RK stages) over the
whole lattice with the state resident in HBM.

Prints one JSON line: cell-updates/s (interior cells x steps / wall time),
the roofline of the stage kernel (HBM-bound stencil) and the CPU baseline
(the FP64 oracle on the host cores, a bounded sample of the same event).
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
VALUES_PER_CELL_UPDATE = {True: 107, False: 100}  # 15 (bulk) or 14 conserved variables


def algorithmic_bytes_per_cell_update(bulk, precision):
    """Values moved per cell per KT step with perfect on-chip neighbour reuse:
    stage 1 reads q^n, e, p, u^n, u^{n-1} and writes q*, e*, p*, u*; stage 2
    reads q*, e*, p*, u*, u^n, q^n and writes q^{n+1}, e, p, u^{n+1}."""
    nq = 15 if bulk else 14
    vals = (nq + 2 + 4 + 4) + (nq + 2 + 4) + (nq + 2 + 4 + 4 + nq) + (nq + 2 + 4)
    assert vals == VALUES_PER_CELL_UPDATE[bulk]
    return 8 * vals


def build_event(args):
    from gpuvh import ic

    e = ic.c2_event(args.nx, args.ny, args.nz, 0.1, 0.1, 0.1, seed=args.seed)
    return e


def cpu_baseline(args, e_full, threads):
    """Time the oracle (FP64 CPU restatement of the reference KT step, test
    infrastructure) on the host cores: one KT step of the same event on the
    full lattice.  Cell-updates/s."""
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
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--eos", choices=["table", "wb", "conformal"], default="table",
                    help="equation of state (the workload's is the 2000-node table)")
    ap.add_argument("--variant", type=int, default=None,
                    help="stage-kernel arrangement (library default if unset)")
    ap.add_argument("--timed-steps", type=int, default=5,
                    help="extra steps with HIP events around each stage kernel")
    ap.add_argument("--cpu-steps", type=int, default=6)
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
    prec = gpuvh.FP64 if args.precision == "fp64" else gpuvh.MIXED
    L = gpuvh.lattice(args.nx, args.ny, args.nz, 0.1, 0.1, 0.1, args.dt)
    H = gpuvh.hydro(args.t0, 0.16, T_fo_GeV=0.15)
    eos = {"table": gpuvh.EOS_TABLE, "wb": gpuvh.EOS_WB, "conformal": gpuvh.EOS_CONFORMAL}[args.eos]
    opts = gpuvh.options(eos, True, prec, 0.16 / gpuvh.HBARC, stream=stream.cuda_stream)
    ctx = gpuvh.Context(L, H, opts)
    ut, ux, uy, un, pi10, Pi = gpuvh.ic.at_rest(e.shape)
    p = ctx.equilibrium_pressure(e)
    ctx.set_initial_state(args.t0, e, p, ut, ux, uy, un, pi10, Pi)

    n = 0
    ctx.steps(args.t0, n, args.warmup)
    n += args.warmup
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    ctx.steps(args.t0, n, args.steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    n += args.steps

    # per-launch durations of the two stage kernels, HIP events on the stream
    # the kernels run on
    s1, s2 = [], []
    for _ in range(args.timed_steps):
        a, b = ctx.step_timed(args.t0 + n * args.dt)
        n += 1
        s1.append(a)
        s2.append(b)
    diag = ctx.diagnostics()
    ncells = args.nx * args.ny * args.nz
    ms_per_step = 1e3 * elapsed / args.steps
    value = ncells * args.steps / elapsed

    bpc = algorithmic_bytes_per_cell_update(True, prec)
    kt_ms = float(np.mean(s1) + np.mean(s2))
    achieved = bpc * ncells / (kt_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": "kt_stage_kernel", "stage_ms": [round(float(np.mean(s1)), 4),
                                                          round(float(np.mean(s2)), 4)],
                "bytes_per_cell_update": bpc}

    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        v, el, ns = cpu_baseline(args, e, threads)
        cpu = {"value": round(v, 1), "unit": "cell-updates/s", "cores": threads,
               "kind": "port",
               "sample": "oracle (FP64 CPU restatement of the reference KT step, OpenMP) on "
                         "the full %dx%dx%d event, %d KT step(s), %.1f s"
                         % (args.nx, args.ny, args.nz, ns, el)}

    out = {
        "metric": "cell-updates/s per KT step",
        "value": round(value, 1),
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == gpuvh.FP64 else "f32/f64 mixed",
        "data": "synthetic",
        "config": {"workload": "3+1D KT step %dx%dx%d, %s EoS, shear+bulk, %s"
                               % (args.nx, args.ny, args.nz,
                                  "tabulated" if args.eos == "table" else args.eos, args.precision),
                   "lattice": [args.nx, args.ny, args.nz], "dtau": args.dt, "tau0": args.t0,
                   "eos": "tabulated 2000 nodes" if args.eos == "table" else args.eos,
                   "eta_over_s": 0.16,
                   "zeta_peak_GeV": 0.16, "seed": args.seed, "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "diagnostics": diag,
    }
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
