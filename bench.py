"""Benchmark of the explicit residual + RK path on one MI355X.

Workload (default): 3-D Euler isentropic vortex, 128^3 vertices on a periodic
[0, 10]^3 box, Yee vortex of strength 5 in a (1, 1, 0) free stream, MUSCL
with the van Albada limiter + Roe flux, classical RK4, CFL 0.8, FP64.
One bench step = one full RK4 time step (4 fused stage launches + the CFL
dt update) over the whole grid; the state is resident in HBM throughout.

Metric: FP64 cell-updates per second per RK stage
        = 128^3 vertices x 4 stages x K steps / wall time of the K steps.
Roofline: the dominant kernel (the fused 3-D stage) against HBM, with
algorithmic bytes of 160 B per cell-update (RK4 reads/writes 3, 5, 5, 3
five-component FP64 arrays over its four stages) divided by the mean stage
launch duration measured with HIP events on the library's stream.
CPU baseline: the oracle (a scalar C restatement of the reference
algorithm) on one host core, one RK4 step of the same 128^3 workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fvm-3d-f90_amd"))

HBM_PEAK_GBS = 8000.0
B_ALG = {"rk4_3d": 160.0, "ssp3_3d": 8.0 / 3.0 * 40.0, "ssp3_2d": 8.0 / 3.0 * 32.0}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=128, help="vertices per direction")
    ap.add_argument("--cfl", type=float, default=0.8)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--numerics", choices=("exact", "fast"), default="fast",
                    help="exact: bit-identical to the reference arithmetic; fast: FMA + "
                         "reciprocals, <= 1e-12 rel. after 1 step, <= 1e-9 after 1000 (tested)")
    ap.add_argument("--no-tune", action="store_true", help="skip the launch-shape autotuner")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a plain 1 GiB device copy (scale reference for the "
                         "FETCH_SIZE / WRITE_SIZE counters)")
    return ap.parse_args()


def calibrate_copy():
    """A plain copy kernel of known size: 2^27 doubles (1 GiB) read and written
    once per launch, 5 launches."""
    import torch
    x = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_per_launch": x.numel() * 8, "launches": 5}))


def cpu_baseline(p, h, args):
    """The oracle on one host core, a bounded sample of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O
    n = args.n
    t0 = time.perf_counter()
    O.oracle_solve(p, 3, (h, h, h), recon=O.MUSCL, limiter=O.LIM_VANALBADA, flux=O.ROE,
                   rk=O.RK4, kappa=0.33333, entfix=0.2, dt=0.0, cfl=args.cfl,
                   nsteps=args.cpu_steps, p_from_q=1, bc=(O.BC_PERIODIC,) * 6)
    sec = time.perf_counter() - t0
    cells = float(n) ** 3 * 4 * args.cpu_steps
    return {"value": cells / sec, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": f"{args.cpu_steps} RK4 step(s) of the {n}^3 vortex workload "
                      f"({cells:.3g} cell-updates, {sec:.1f} s, oracle/fvm_oracle.c)"}


def main():
    args = parse()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is described, not built)", file=sys.stderr)
        sys.exit(2)
    import fvmgpu
    from fvmgpu import cases

    n = args.n
    p, h, _ = cases.isentropic_vortex(n, seed=0)
    q = cases.to_conserved(p)
    ctx = fvmgpu.Context(0)
    ctx.scheme(recon=fvmgpu.MUSCL, limiter=fvmgpu.LIM_VANALBADA, flux=fvmgpu.ROE,
               rk=fvmgpu.RK4, kappa=0.33333, entropy_fix=0.2, gamma=1.4)
    blk = ctx.block((n, n, n), (h, h, h))
    blk.set_periodic()
    blk.upload(q)
    ctx.numerics(fvmgpu.NUMERICS_FAST if args.numerics == "fast" else fvmgpu.NUMERICS_EXACT)
    if not args.no_tune:
        ctx.autotune(reps=3)
    tile = blk.launch_shape()
    if args.warmup:
        ctx.step(args.warmup, cfl=args.cfl)
    ctx.sync()

    ctx.timing(True)
    ctx.sync()
    t0 = time.perf_counter()
    dt_last = ctx.step(args.steps, cfl=args.cfl)
    ctx.sync()
    t1 = time.perf_counter()
    stage_ms, launches, cell_updates = ctx.timing_read()
    ctx.timing(False)
    norms, nan = ctx.norms()
    if nan:
        raise SystemExit(f"bench.py: residual NaN or diverged, norms {norms}")

    sec = t1 - t0
    ncell = float(n) ** 3
    value = ncell * 4 * args.steps / sec
    avg_launch_s = stage_ms * 1e-3 / max(launches, 1)
    achieved = B_ALG["rk4_3d"] * ncell / avg_launch_s / 1e9

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "stage3d_muscl_roe_vortex128_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

    out = {
        "metric": "FP64 cell-updates/s per RK stage",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"isentropic-vortex-3d {n}^3 periodic, MUSCL-vanAlbada + Roe, "
                               f"RK4, CFL {args.cfl}, FP64",
                   "grid": [n, n, n], "stages_per_step": 4, "dt_last": dt_last,
                   "numerics": args.numerics, "tile_xy_chunk": list(tile),
                   "pct_hbm_roofline": 100.0 * value * B_ALG["rk4_3d"] / (HBM_PEAK_GBS * 1e9)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": f"fvm::{args.numerics}::stage3d_kernel<ReconMuscl, 300, "
                               f"{tile[0]}, {tile[1]}>",
                     "avg_launch_us": avg_launch_s * 1e6, "launches": launches,
                     "bytes_per_cell_update": B_ALG["rk4_3d"]},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(p, h, args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
