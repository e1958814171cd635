"""Benchmark of the explicit residual + RK path on one MI355X.

Default workload: 3-D Euler isentropic vortex, 128^3 vertices on a periodic
[0, 10]^3 box, Yee vortex of strength 5 in a (1, 1, 0) free stream, MUSCL
with the van Albada limiter + Roe flux, classical RK4, CFL 0.8, FP64.
One bench step = one full RK4 time step (4 fused stage launches + the CFL
dt update) over the whole grid; the state is resident in HBM throughout.
Setup before the W warm-up steps: 0.5 s of untimed steps while the GPU's
clocks settle, then the launch-shape autotuner (on settled clocks, so the
shapes it ranks are timed alike).  Numerics: the exact family
(bit for bit the oracle), the one that meets the 1000-step tolerance here.

Metric: FP64 cell-updates per second per RK stage
        = vertices x stages x steps / wall time of the steps.
Roofline: the dominant kernel (the fused stage) against HBM, with
algorithmic bytes per cell-update of 160 B for RK4 (it reads/writes 3, 5,
5, 3 five-component FP64 arrays over its four stages), 106.7 B for 3-D
SSP-RK3 (2, 3, 3 arrays) and 85.3 B for 2-D SSP-RK3 (four components),
divided by the mean stage launch duration measured with HIP events on the
library's stream.
CPU baseline: the oracle (a C restatement of the reference algorithm,
OpenMP over independent pencils as the reference has no parallel path) on
the host cores, and on one core, over a bounded sample of the same workload.

Other workloads (--workload):
  riemann400  2-D four-quadrant Riemann problem, 400 x 400 vertices, the
              reference's quadrant states, MUSCL-minmod + Roe, SSP-RK3,
              CFL 0.5, transmissive boundaries, run to t = 0.25.
  blob27      27 blocks in a 3 x 3 x 3 conforming layout, block sizes drawn
              from U{32..96} per direction (seed 3: 5.45 M vertices),
              density-blob state, WENO5-JS + AUSM+, SSP-RK3, CFL 0.5,
              block-to-block ghost exchange on the device; one batched
              (persistent when --persistent) launch per stage against one
              launch per block.
  tgv192      3-D Navier-Stokes Taylor-Green vortex, 192^3 vertices on the
              periodic [0, 2 pi)^3 box, Re 1600, Ma 0.1, Pr 0.71, Sutherland
              viscosity, WENO5-JS + Van Leer FVS, central viscous fluxes,
              RK4, CFL 0.5; reports the kinetic-energy decay rate over the
              timed steps (--t-end 10 runs the whole case).
  serving     Poisson arrivals (50 / s, seeded) of independent solves:
              2-D Riemann 64^2..512^2 (the reference's quadrant states)
              and 3-D vortex / blob 32^3..128^3
              (log-uniform sizes), schemes MUSCL-minmod + Roe or WENO5 +
              AUSM+, each to a fixed end time, through the continuous-
              batching server; against the same solves run one after the
              other through plain contexts.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fvm-3d-f90_amd"))

HBM_PEAK_GBS = 8000.0
FP64_VECTOR_PEAK_TFLOPS = 78.6   # MI355X FP64 vector (published)
B_ALG = {"rk4_3d": 160.0, "ssp3_3d": 8.0 / 3.0 * 40.0, "ssp3_2d": 8.0 / 3.0 * 32.0}
METRIC = "FP64 cell-updates/s per RK stage"
# the stage launches of every TIMING_EVERY-th step are bracketed by one HIP
# event pair (an event costs the stream a few microseconds: timing every
# launch would slow short stages)
TIMING_EVERY = 4


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (vortex128: 1000, blob27: 200; riemann400 runs to t_end)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="vortex128",
                    choices=("vortex128", "riemann400", "blob27", "tgv192", "serving"))
    ap.add_argument("--t-end", type=float, default=None,
                    help="tgv192: run to this time instead of --steps steps")
    ap.add_argument("--n", type=int, default=None, help="vertices per direction")
    ap.add_argument("--cfl", type=float, default=None)
    ap.add_argument("--seed", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-steps", type=int, default=None,
                    help="steps of the CPU baseline sample (default per workload, about "
                         "10-30 s of CPU work: vortex128 5, riemann400 30, blob27 40)")
    ap.add_argument("--numerics", choices=("exact", "fast"), default="exact",
                    help="exact (default): bit-identical to the oracle's arithmetic, so within "
                         "every stated tolerance; fast: FMA + refined reciprocals, <= 1e-12 rel. "
                         "after 1 step, but on the 128^3 vortex 1.9e-9 from exact after 1000 "
                         "steps (the flow amplifies one-ulp differences)")
    ap.add_argument("--no-tune", action="store_true", help="skip the launch-shape autotuner")
    ap.add_argument("--output-every", type=int, default=0,
                    help="vortex: write the output variables to a file every K steps inside "
                         "the timed region (0: no output)")
    ap.add_argument("--output-mode", choices=("snapshot", "download"), default="snapshot",
                    help="snapshot: asynchronous device snapshot saved on a background thread; "
                         "download: blocking download, q -> p and write on the host")
    ap.add_argument("--persistent", action="store_true",
                    help="blob27: persistent launch (grid of 2 workgroups per CU) instead of "
                         "one workgroup per tile")
    ap.add_argument("--requests", type=int, default=96, help="serving: number of requests")
    ap.add_argument("--rate", type=float, default=50.0, help="serving: arrivals per second")
    ap.add_argument("--max-active", type=int, default=256,
                    help="serving: solves resident on the GPU at once")
    ap.add_argument("--steps-per-sync", type=int, default=4,
                    help="serving: time steps per host synchronisation")
    ap.add_argument("--no-extras", action="store_true",
                    help="vortex128 only: skip the fast-family, workloads and serving objects "
                         "added after the timed region")
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


def oracle():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O
    return O


def cpu_measure(run, cells, sample):
    """The oracle on the host cores (OpenMP over pencils, OMP_NUM_THREADS or
    every core) and on one core, the same bounded sample both times."""
    O = oracle()
    n = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    used = O.set_threads(n)
    t0 = time.perf_counter()
    run(O)
    sec_n = time.perf_counter() - t0
    O.set_threads(1)
    t0 = time.perf_counter()
    run(O)
    sec_1 = time.perf_counter() - t0
    O.set_threads(used)
    return {"value": cells / sec_n, "unit": "cell-updates/s", "cores": used, "kind": "port",
            "sample": f"{sample} ({cells:.3g} cell-updates; {sec_n:.2f} s on {used} threads, "
                      f"{sec_1:.2f} s on 1; oracle/fvm_oracle.c)",
            "value_1core": cells / sec_1, "cpu_model": cpu_model(),
            "note": "kind port: the reference is a serial Fortran solver with SSP-RK3, fixed dt "
                    "and extrapolated faces only (no RK4, periodic faces or CFL step to run "
                    "this workload), so the C restatement of its algorithm is timed"}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def numerics_code(fvmgpu, args):
    return fvmgpu.NUMERICS_FAST if args.numerics == "fast" else fvmgpu.NUMERICS_EXACT


def roofline(stage_ms, launches, cells_per_launch, b_alg, traffic=None, kernel=""):
    avg_launch_s = stage_ms * 1e-3 / max(launches, 1)
    achieved = b_alg * cells_per_launch / avg_launch_s / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": kernel,
            "avg_launch_us": avg_launch_s * 1e6, "launches": launches,
            "bytes_per_cell_update": b_alg}


def line(args, value, steps, sec, config, roof, cpu):
    return {
        "metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1,
        "steps": steps, "warmup": args.warmup, "ms_per_step": sec * 1e3 / max(steps, 1),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic", "config": config, "roofline": roof, "cpu_baseline": cpu,
    }


# ---------------------------------------------------------------------------
# vortex128: the headline
# ---------------------------------------------------------------------------
def bench_vortex(args):
    import fvmgpu
    from fvmgpu import cases
    n = args.n or 128
    steps = args.steps or 1000
    cfl = args.cfl or 0.8
    p, h, _ = cases.isentropic_vortex(n, seed=args.seed or 0)
    q = cases.to_conserved(p)
    ctx = fvmgpu.Context(0)
    ctx.scheme(recon=fvmgpu.MUSCL, limiter=fvmgpu.LIM_VANALBADA, flux=fvmgpu.ROE,
               rk=fvmgpu.RK4, kappa=0.33333, entropy_fix=0.2, gamma=1.4)
    blk = ctx.block((n, n, n), (h, h, h))
    blk.set_periodic()
    blk.upload(q)
    ctx.numerics(numerics_code(fvmgpu, args))
    settle(ctx, cfl)
    if not args.no_tune:
        # shapes within a few per cent: long enough windows to rank them
        ctx.autotune(reps=getattr(args, "tune_reps", 10))
    tile = blk.launch_shape()
    reg = blk.launch_tile() >= 4          # the register-reconstruction kernel
    if args.warmup:
        ctx.step(args.warmup, cfl=cfl)
    ctx.sync()

    ctx.timing(TIMING_EVERY)
    ctx.sync()
    if args.output_every:
        t0, t1, dt_last, output = timed_with_output(fvmgpu, cases, ctx, blk, steps, cfl,
                                                    args.output_every, args.output_mode)
    else:
        output = None
        t0 = time.perf_counter()
        dt_last = ctx.step(steps, cfl=cfl)
        ctx.sync()
        t1 = time.perf_counter()
    stage_ms, launches, _ = ctx.timing_read()
    ctx.timing(False)
    norms, nan = ctx.norms()
    if nan:
        raise SystemExit(f"bench.py: residual NaN or diverged, norms {norms}")

    sec = t1 - t0
    ncell = float(n) ** 3
    value = ncell * 4 * steps / sec
    traffic, flops, source = None, None, None
    tpath = os.path.join(ROOT, "profiles", "stage3d_muscl_roe_vortex128_traffic.json")
    if os.path.exists(tpath) and n == 128:
        # counters of the kernel family and tile shape this run launched (the
        # autotuner's pick), measured by rocprofv3 --pmc and committed: they
        # describe the build named in the file, not this run
        with open(tpath) as f:
            tj = json.load(f)
        fam = tj.get("families", {}).get(args.numerics, {})
        prof = fam.get("shapes", {}).get(f"{tile[0]}x{tile[1]}" + ("-reg" if reg else ""), {})
        traffic = prof.get("hbm_bytes_per_launch")
        flops = prof.get("fp64_flops_per_cell_update")
        if prof:
            source = f"profiles/stage3d_muscl_roe_vortex128_traffic.json ({fam.get('measured', '?')})"
    roof = roofline(stage_ms, launches, ncell, B_ALG["rk4_3d"], traffic,
                    f"fvm::{args.numerics}::{'stage3d_reg_kernel' if reg else 'stage3d_kernel'}"
                    f"<ReconMuscl, 300, {tile[0]}, {tile[1]}>")
    if source:
        roof["traffic_source"] = source
    if flops:
        # the co-limiting roof: counted FP64 VALU work against the vector FP64 peak
        tflops = flops * ncell / (roof["avg_launch_us"] * 1e-6) / 1e12
        roof["fp64_valu"] = {"flops_per_cell_update": flops, "achieved": tflops,
                             "peak": FP64_VECTOR_PEAK_TFLOPS, "unit": "TFLOP/s",
                             "frac": tflops / FP64_VECTOR_PEAK_TFLOPS}
    config = {"workload": f"isentropic-vortex-3d {n}^3 periodic, MUSCL-vanAlbada + Roe, "
                          f"RK4, CFL {cfl}, FP64",
              "grid": [n, n, n], "stages_per_step": 4, "dt_last": dt_last,
              "numerics": args.numerics, "tile_xy_chunk": list(tile),
              "pct_hbm_roofline": 100.0 * value * B_ALG["rk4_3d"] / (HBM_PEAK_GBS * 1e9)}
    if output:
        config["output"] = output
    cpu = None
    if not args.no_cpu:
        cs = args.cpu_steps or 5
        cpu = cpu_measure(
            lambda O: O.oracle_solve(p, 3, (h, h, h), recon=O.MUSCL, limiter=O.LIM_VANALBADA,
                                     flux=O.ROE, rk=O.RK4, kappa=0.33333, entfix=0.2, dt=0.0,
                                     cfl=cfl, nsteps=cs, p_from_q=1,
                                     bc=(O.BC_PERIODIC,) * 6),
            ncell * 4 * cs, f"{cs} RK4 step(s) of the {n}^3 vortex workload")
    return line(args, value, steps, sec, config, roof, cpu)


def settle(ctx, cfl, seconds=0.5):
    """Untimed steps until the GPU has run the workload for `seconds`: a
    fresh box ramps its clocks over the first fraction of a second under
    load, and a 20-step timed window would otherwise measure the ramp.  Part
    of the setup, like the autotuner; the state advances, the timed region
    is unchanged."""
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        ctx.step(8, cfl=cfl)
        ctx.sync()


def timed_with_output(fvmgpu, cases, ctx, blk, steps, cfl, every, mode):
    """The timed steps in calls of `every`, each followed by writing the
    output variables (density, velocity, pressure) of the state to a file in
    a temporary directory (two files used in turn, removed afterwards)."""
    import tempfile
    if steps % every:
        raise SystemExit("bench.py: --steps must be a multiple of --output-every")
    tmp = tempfile.mkdtemp(prefix="fvm_out_")
    paths = [os.path.join(tmp, f"out{i}.fvmsnap") for i in range(2)]
    pending = []
    nout = 0
    t0 = time.perf_counter()
    for c in range(steps // every):
        dt_last = ctx.step(every, cfl=cfl)
        path = paths[c % 2]
        if mode == "snapshot":
            if len(pending) == 2:            # the file this one overwrites is written
                ctx.snapshot_release(pending.pop(0))
            sid = ctx.snapshot(fvmgpu.PRIMITIVE_VARS)
            ctx.snapshot_save(sid, path)
            pending.append(sid)
        else:
            pv = np.moveaxis(cases.to_primitive(blk.download()), -1, 0)
            with open(path, "wb") as f:
                f.write(np.ascontiguousarray(pv).tobytes())
        nout += 1
    for sid in pending:
        ctx.snapshot_release(sid)
    ctx.sync()
    t1 = time.perf_counter()
    nbytes = os.path.getsize(paths[0])
    for q in paths:
        if os.path.exists(q):
            os.remove(q)
    os.rmdir(tmp)
    return t0, t1, dt_last, {"every": every, "mode": mode, "outputs": nout,
                             "bytes_per_output": nbytes}


# ---------------------------------------------------------------------------
# riemann400
# ---------------------------------------------------------------------------
def bench_riemann(args):
    import fvmgpu
    from fvmgpu import cases
    n = args.n or 400
    cfl = args.cfl or 0.5
    t_end = 0.25
    p, dx = cases.riemann2d(n, seed=args.seed or 0)
    q = cases.to_conserved(p)
    ctx = fvmgpu.Context(0)
    ctx.scheme(recon=fvmgpu.MUSCL, limiter=fvmgpu.LIM_MINMOD, flux=fvmgpu.ROE,
               rk=fvmgpu.RK_SSP3, kappa=0.33333, entropy_fix=0.2, gamma=1.4)
    blk = ctx.block((n, n, 1), (dx, dx, 1.0))
    blk.upload(q)
    ctx.numerics(numerics_code(fvmgpu, args))
    settle(ctx, cfl)
    if not args.no_tune:
        ctx.autotune(reps=3)
    tile = blk.launch_shape()
    if args.warmup:
        ctx.step(args.warmup, cfl=cfl)
    blk.upload(q)                      # restart from the initial state
    ctx.set_time(0.0, t_end)
    ctx.sync()
    _, steps0 = ctx.time()             # steps of the setup (settling, tuning, warm-up)

    ctx.timing(TIMING_EVERY)
    t0 = time.perf_counter()
    chunk = 64
    while True:
        dt_last = ctx.step(chunk, cfl=cfl)
        t, steps = ctx.time()
        if t >= t_end * (1 - 1e-14):
            break
        # size the next batch from the last dt so few skipped steps are queued
        chunk = int(min(64, max(1, np.ceil((t_end - t) / dt_last) + 1)))
    ctx.sync()
    t1 = time.perf_counter()
    stage_ms, launches, cu = ctx.timing_read()
    ctx.timing(False)
    norms, nan = ctx.norms()
    if nan:
        raise SystemExit(f"bench.py: residual NaN or diverged, norms {norms}")
    _, steps = ctx.time()
    steps -= steps0
    sec = t1 - t0
    ncell = float(n) * n
    value = ncell * 3 * steps / sec
    kname = (f"fvm::{args.numerics}::stage2d_kernel<ReconMuscl, 300>" if tile[1] == 1 else
             f"fvm::{args.numerics}::stage2d_tile_kernel<ReconMuscl, 300, {tile[0]}, {tile[1]}>")
    roof = roofline(stage_ms, launches, ncell, B_ALG["ssp3_2d"], None, kname)
    config = {"workload": f"riemann-2d {n}x{n} four-quadrant (reference states), MUSCL-minmod "
                          f"+ Roe, SSP-RK3, CFL {cfl}, transmissive, t_end {t_end}, FP64",
              "grid": [n, n, 1], "stages_per_step": 3, "t_final": t, "numerics": args.numerics,
              "tile_xy_chunk": list(tile),
              "pct_hbm_roofline": 100.0 * value * B_ALG["ssp3_2d"] / (HBM_PEAK_GBS * 1e9)}
    cpu = None
    if not args.no_cpu:
        cs = args.cpu_steps or 30
        cpu = cpu_measure(
            lambda O: O.oracle_solve(p, 2, (dx, dx, 1.0), recon=O.MUSCL, limiter=O.LIM_MINMOD,
                                     flux=O.ROE, rk=O.RK_SSP3, kappa=0.33333, entfix=0.2, dt=0.0,
                                     cfl=cfl, nsteps=cs, p_from_q=1),
            ncell * 3 * cs, f"{cs} SSP-RK3 steps of the {n}x{n} Riemann workload")
        ref = cpu_reference(p, q, dx, cfl, cs, ncell * 3 * cs)
        if ref is not None:
            # the reference's own scheme modules on one core (the reference
            # has no threading of its own): the baseline; the port's figures
            # stay beside it
            port = cpu
            cpu = dict(ref)
            cpu["port_value"] = port["value"]
            cpu["port_cores"] = port["cores"]
            cpu["port_value_1core"] = port["value_1core"]
            cpu["port_sample"] = port["sample"]
    return line(args, value, steps, sec, config, roof, cpu)


def cpu_reference(p, q, dx, cfl, steps, cells):
    """The reference's Fortran scheme modules (oracle/_ref, compiled from its
    sources) driven over the same Riemann sample on one core.  The reference
    steps with a fixed dt: the CFL dt of the initial state.  None where the
    harness is not built."""
    O = oracle()
    if O.ref() is None:
        return None
    dt = O.cfl_dt(q, 2, (dx, dx, 1.0), 1.4, cfl)
    t0 = time.perf_counter()
    O.ref_solve(p, 2, (dx, dx, 1.0), recon=O.MUSCL, limiter=O.LIM_MINMOD, flux=O.ROE,
                kappa=0.33333, entfix=0.2, dt=dt, nsteps=steps, p_from_q=1)
    sec = time.perf_counter() - t0
    return {"value": cells / sec, "unit": "cell-updates/s", "cores": 1, "kind": "reference",
            "sample": f"{steps} SSP-RK3 steps (fixed dt {dt:.4g}) of the Riemann workload, "
                      f"{cells:.3g} cell-updates in {sec:.2f} s; oracle/_ref/libfvmref.so "
                      "(the reference's scheme modules)"}


# ---------------------------------------------------------------------------
# blob27
# ---------------------------------------------------------------------------
def build_blob27(fvmgpu, cases, seed, mode, args):
    sizes, offs, sp, states = cases.blob_blocks(seed)
    ctx = fvmgpu.Context(0)
    ctx.scheme(recon=fvmgpu.WENO, limiter=fvmgpu.LIM_NONE, flux=fvmgpu.AUSMPLUS,
               rk=fvmgpu.RK_SSP3, kappa=0.33333, entropy_fix=0.2, gamma=1.4, ref_compat=0)
    ctx.numerics(numerics_code(fvmgpu, args))
    ctx.launch_mode(mode)
    blocks = {}
    for key, pb in states.items():
        nk, nj, ni, _ = pb.shape
        b = ctx.block((ni, nj, nk), sp)
        b.upload(cases.to_conserved(pb))
        blocks[key] = b
    for (a, b_, c), blk in blocks.items():
        idx = [a, b_, c]
        for d in range(3):
            if idx[d] + 1 < 3:
                nb = list(idx)
                nb[d] += 1
                blk.connect(2 * d + 1, blocks[tuple(nb)], shared=True)
    return ctx, blocks, sizes, sp, states


def time_steps(ctx, steps, cfl, warmup):
    if warmup:
        ctx.step(warmup, cfl=cfl)
    ctx.sync()
    ctx.timing(TIMING_EVERY)
    t0 = time.perf_counter()
    ctx.step(steps, cfl=cfl)
    ctx.sync()
    sec = time.perf_counter() - t0
    stage_ms, launches, _ = ctx.timing_read()
    ctx.timing(False)
    norms, nan = ctx.norms()
    if nan:
        raise SystemExit(f"bench.py: residual NaN or diverged, norms {norms}")
    return sec, stage_ms, launches


def bench_blob27(args):
    import fvmgpu
    from fvmgpu import cases
    seed = 3 if args.seed is None else args.seed
    steps = args.steps or 200
    cfl = args.cfl or 0.5
    batched = fvmgpu.LAUNCH_PERSISTENT if args.persistent else fvmgpu.LAUNCH_BATCHED
    results = {}
    for name, mode in (("batched", batched), ("per_block", fvmgpu.LAUNCH_PER_BLOCK)):
        ctx, blocks, sizes, sp, states = build_blob27(fvmgpu, cases, seed, mode, args)
        if not args.no_tune:
            ctx.autotune(reps=2)
        shapes = sorted(set(b.launch_shape() for b in blocks.values()))
        sec, stage_ms, launches = time_steps(ctx, steps, cfl, args.warmup)
        results[name] = (sec, stage_ms, launches, shapes)
        ctx.close()
    ncell = float(sum(v.shape[0] * v.shape[1] * v.shape[2] for v in states.values()))
    sec, stage_ms, launches, shapes = results["batched"]
    value = ncell * 3 * steps / sec
    pb_sec = results["per_block"][0]
    roof = roofline(stage_ms, launches, ncell, B_ALG["ssp3_3d"], None,
                    f"fvm::{args.numerics}::stage3d_kernel<ReconWeno, 1303, *, *> over 27 blocks")
    config = {"workload": f"blob-3d 27 blocks 3x3x3 (sizes {sizes}, seed {seed}), WENO5-JS + "
                          f"AUSM+, SSP-RK3, CFL {cfl}, device ghost exchange, FP64",
              "vertices": ncell, "stages_per_step": 3, "numerics": args.numerics,
              "launch": "persistent" if args.persistent else "batched",
              "tile_xy_chunk": [list(s) for s in shapes],
              "per_block_launches_value": ncell * 3 * steps / pb_sec,
              "speedup_vs_per_block": pb_sec / sec,
              "pct_hbm_roofline": 100.0 * value * B_ALG["ssp3_3d"] / (HBM_PEAK_GBS * 1e9)}
    cpu = None
    if not args.no_cpu:
        key = max(states, key=lambda k: states[k].size)
        pb = states[key]
        cs = args.cpu_steps or 40
        cells = float(pb.shape[0] * pb.shape[1] * pb.shape[2]) * 3 * cs
        cpu = cpu_measure(
            lambda O: O.oracle_solve(pb, 3, sp, recon=O.WENO, limiter=O.LIM_NONE, flux=O.AUSMPLUS,
                                     rk=O.RK_SSP3, kappa=0.33333, entfix=0.2, dt=0.0, cfl=cfl,
                                     nsteps=cs, p_from_q=1, weno_textbook=1),
            cells, f"{cs} SSP-RK3 steps of the largest block {pb.shape[:3]} alone")
    return line(args, value, steps, sec, config, roof, cpu)


# ---------------------------------------------------------------------------
# tgv192
# ---------------------------------------------------------------------------
def bench_tgv(args):
    import fvmgpu
    from fvmgpu import cases
    n = args.n or 192
    cfl = args.cfl or 0.5
    steps = args.steps or 100
    p, h = cases.tgv(n)
    visc = cases.tgv_params(re=1600.0, ma=0.1, pr=0.71)
    q = cases.to_conserved(p)
    ctx = fvmgpu.Context(0)
    ctx.scheme(recon=fvmgpu.WENO, limiter=fvmgpu.LIM_NONE, flux=fvmgpu.VANLEER, rk=fvmgpu.RK4,
               kappa=0.33333, entropy_fix=0.2, gamma=1.4, ref_compat=0)
    ctx.numerics(numerics_code(fvmgpu, args))
    ctx.viscous(*visc)
    blk = ctx.block((n, n, n), (h, h, h))
    blk.set_periodic()
    blk.upload(q)
    if not args.no_tune:
        ctx.autotune(reps=2)
    tile = blk.launch_shape()
    if args.warmup:
        ctx.step(args.warmup, cfl=cfl)
    ctx.sync()
    t_a, steps_a = ctx.time()
    ek_a = cases.kinetic_energy(blk.download())
    if args.t_end:
        ctx.set_time(t_a, args.t_end)
    ctx.timing(TIMING_EVERY)
    t0 = time.perf_counter()
    if args.t_end:
        while True:
            dt = ctx.step(64, cfl=cfl)
            t, _ = ctx.time()
            if t >= args.t_end * (1 - 1e-14):
                break
    else:
        ctx.step(steps, cfl=cfl)
    ctx.sync()
    sec = time.perf_counter() - t0
    stage_ms, launches, _ = ctx.timing_read()
    ctx.timing(False)
    norms, nan = ctx.norms()
    if nan:
        raise SystemExit(f"bench.py: residual NaN or diverged, norms {norms}")
    t_b, steps_b = ctx.time()
    steps = steps_b - steps_a
    ek_b = cases.kinetic_energy(blk.download())
    ncell = float(n) ** 3
    value = ncell * 4 * steps / sec
    roof = roofline(stage_ms, launches, ncell, B_ALG["rk4_3d"], None,
                    f"fvm::{args.numerics}::viscous3d_kernel<32, 8> + stage3d_kernel<ReconWeno, "
                    f"301, {tile[0]}, {tile[1]}> per stage")
    config = {"workload": f"taylor-green-3d {n}^3 periodic [0,2pi)^3, Re 1600, Ma 0.1, Pr 0.71, "
                          "Sutherland, WENO5-JS + Van Leer FVS + central viscous, RK4, "
                          f"CFL {cfl}, FP64",
              "grid": [n, n, n], "stages_per_step": 4, "numerics": args.numerics,
              "tile_xy_chunk": list(tile), "t_start": t_a, "t_final": t_b,
              "kinetic_energy_start": ek_a, "kinetic_energy_final": ek_b,
              "kinetic_energy_decay_rate": (ek_a - ek_b) / (t_b - t_a),
              "pct_hbm_roofline": 100.0 * value * B_ALG["rk4_3d"] / (HBM_PEAK_GBS * 1e9)}
    cpu = None
    if not args.no_cpu:
        cpu = cpu_measure(
            lambda O: O.oracle_solve(p, 3, (h, h, h), recon=O.WENO, limiter=O.LIM_NONE,
                                     flux=O.VANLEER, rk=O.RK4, dt=0.0, cfl=cfl, nsteps=1,
                                     p_from_q=1, bc=(O.BC_PERIODIC,) * 6, weno_textbook=1,
                                     visc=visc),
            ncell * 4, f"1 RK4 step of the {n}^3 TGV workload")
    return line(args, value, steps, sec, config, roof, cpu)


# ---------------------------------------------------------------------------
# serving
# ---------------------------------------------------------------------------
# The request stream is fvmgpu.cases.serving_requests (end times
# cases.SERVE_T_END).  At 50 arrivals / s the served run ends soon after the
# last arrival (the rate is then set by the arrivals); the capacity run queues
# every request at t = 0 and measures what the GPU sustains.


def bench_serving(args):
    import fvmgpu
    from fvmgpu import cases
    seed = 0 if args.seed is None else args.seed
    arrivals, reqs = cases.serving_requests(seed, args.requests, args.rate)
    numerics = numerics_code(fvmgpu, args)

    def submit_args(r):
        a = {k: v for k, v in r.items() if k not in ("label", "p0")}
        a["numerics"] = numerics
        return a

    # warm-up: one tiny solve per (dimension, scheme) so no first-launch cost
    # lands in the timed run
    srv = fvmgpu.Server(max_active=256)
    for sch in ({k: reqs[0][k] for k in ("recon", "limiter", "flux", "rk", "ref_compat")},
                dict(recon=fvmgpu.WENO, limiter=fvmgpu.LIM_NONE, flux=fvmgpu.AUSMPLUS,
                     rk=fvmgpu.RK_SSP3, ref_compat=0),
                dict(recon=fvmgpu.MUSCL, limiter=fvmgpu.LIM_MINMOD, flux=fvmgpu.ROE,
                     rk=fvmgpu.RK_SSP3, ref_compat=1)):
        p2, dx = cases.riemann2d(24, seed=1)
        srv.submit(q0=cases.to_conserved(p2), spacing=(dx, dx, 1.0), t_end=1e-3, cfl=0.5,
                   numerics=numerics, **sch)
        x = np.linspace(0.0, 1.0, 12)
        srv.submit(q0=cases.to_conserved(cases.blob(x, x, x)), spacing=(1 / 11,) * 3, t_end=1e-3,
                   cfl=0.5, numerics=numerics, **sch)
    while srv.poll(100, args.steps_per_sync)[1]:
        pass
    srv.close()
    srv = fvmgpu.Server(max_active=args.max_active)
    t0 = time.perf_counter()
    ids = []
    nxt = 0
    done = 0
    while done < len(reqs):
        now = time.perf_counter() - t0
        while nxt < len(reqs) and arrivals[nxt] <= now:
            ids.append(srv.submit(**submit_args(reqs[nxt])))
            nxt += 1
        pend, act, done = srv.poll(1, args.steps_per_sync)
        if act == 0 and pend == 0 and nxt < len(reqs):
            time.sleep(max(0.0, min(arrivals[nxt] - (time.perf_counter() - t0), 0.001)))
    makespan = time.perf_counter() - t0
    lat, cell_updates, steps, infos = [], 0.0, 0, []
    for i, rid in zip(range(len(reqs)), ids):
        try:
            _, info = srv.result(rid)
        except fvmgpu.FvmError as e:
            raise SystemExit(f"bench.py: request {i} ({reqs[i]['label']}) failed: {e}")
        infos.append((rid, info))
        lat.append(info["t_done"] - arrivals[i])
        cell_updates += info["cell_updates"]
        steps += int(info["steps"])
    srv.close()
    served_value = cell_updates / makespan

    # capacity: the same requests all queued at t = 0
    srv = fvmgpu.Server(max_active=args.max_active)
    t0 = time.perf_counter()
    cap_ids = [srv.submit(**submit_args(r)) for r in reqs]
    while srv.poll(1, args.steps_per_sync)[2] < len(reqs):
        pass
    cap_makespan = time.perf_counter() - t0
    cap_lat = []
    for i, rid in enumerate(cap_ids):
        _, info = srv.result(rid)
        if info["status"] != 2:
            raise SystemExit(f"bench.py: request {i} ({reqs[i]['label']}) failed in the capacity run")
        cap_lat.append(info["t_done"])
    srv.close()
    cap_value = cell_updates / cap_makespan

    # sequential plain solves: one context per request, run to t_end
    ts = time.perf_counter()
    per_kind = {}
    service = []
    for r, (_, info) in zip(reqs, infos):
        t_req = time.perf_counter()
        _, nk, nj, ni = r["q0"].shape
        with fvmgpu.Context(0) as ctx:
            ctx.scheme(recon=r["recon"], limiter=r["limiter"], flux=r["flux"], rk=r["rk"],
                       ref_compat=r["ref_compat"])
            ctx.numerics(numerics)
            b = ctx.block((ni, nj, nk), r["spacing"])
            if any(r.get("periodic", ())):
                b.set_periodic()
            b.upload(r["q0"])
            ctx.set_time(0.0, r["t_end"])
            chunk = 16
            while True:
                dt = ctx.step(chunk, cfl=r["cfl"])
                t, _ = ctx.time()
                if t >= r["t_end"] * (1 - 1e-14):
                    break
                chunk = int(min(64, max(1, np.ceil((r["t_end"] - t) / dt) + 1)))
            b.download()
        service.append(time.perf_counter() - t_req)
        kind = r["label"].split()[0] + ("-weno" if r["recon"] == fvmgpu.WENO else "-muscl")
        k = per_kind.setdefault(kind, [0, 0.0, 0.0])
        k[0] += 1
        k[1] += info["cell_updates"]
        k[2] += service[-1]
    seq_sec = time.perf_counter() - ts
    seq_value = cell_updates / seq_sec

    lat = np.array(lat)
    # the same arrivals served first-come first-served by one plain solve at a
    # time, with the service times just measured
    fifo, free_at = [], 0.0
    for a, sv in zip(arrivals, service):
        free_at = max(free_at, a) + sv
        fifo.append(free_at - a)
    fifo = np.array(fifo)
    config = {"workload": f"serving: {len(reqs)} Poisson arrivals at {args.rate}/s (seed {seed}); "
                          "riemann2d 64^2..512^2, vortex3d / blob3d 32^3..128^3, log-uniform "
                          f"sizes, end times {cases.SERVE_T_END}, MUSCL-minmod + Roe or WENO5 + "
                          "AUSM+, SSP-RK3, CFL 0.5, FP64",
              "requests": len(reqs), "total_steps": steps, "cell_updates": cell_updates,
              "makespan_s": makespan, "p50_time_to_solution_s": float(np.percentile(lat, 50)),
              "p99_time_to_solution_s": float(np.percentile(lat, 99)),
              "sequential_plain_value": seq_value, "sequential_plain_s": seq_sec,
              "sequential_fifo_p50_s": float(np.percentile(fifo, 50)),
              "sequential_fifo_p99_s": float(np.percentile(fifo, 99)),
              "speedup_vs_sequential": served_value / seq_value,
              "capacity_value": cap_value, "capacity_makespan_s": cap_makespan,
              "capacity_speedup_vs_sequential": cap_value / seq_value,
              "capacity_p50_time_to_solution_s": float(np.percentile(cap_lat, 50)),
              "capacity_p99_time_to_solution_s": float(np.percentile(cap_lat, 99)),
              "steps_per_sync": args.steps_per_sync, "max_active": args.max_active,
              "numerics": args.numerics,
              "sequential_by_kind": {k: {"requests": v[0], "cell_updates_per_s": v[1] / v[2]}
                                     for k, v in sorted(per_kind.items())}}
    return line(args, served_value, steps, makespan, config, None, None)


# ---------------------------------------------------------------------------
# extras of the default run, after the headline's timed region
# ---------------------------------------------------------------------------
def rel_linf(a, b):
    """Relative L-inf per conserved component (max |b| with a floor of 1)."""
    return max(float(np.abs(a[c] - b[c]).max() / max(np.abs(b[c]).max(), 1.0))
               for c in range(a.shape[0]))


def gate_1000_steps():
    """The 1000-step tolerance on the headline workload, measured: the fast
    family against the exact family (which is the oracle bit for bit), beside
    the exact family's own spread when one initial value per vertex (rho E)
    moves by one ulp."""
    import fvmgpu
    from fvmgpu import cases
    p, h, _ = cases.isentropic_vortex(128, seed=0)
    q0 = cases.to_conserved(p)
    q1 = q0.copy()
    q1[4] = np.nextafter(q1[4], np.inf)
    out = {}
    for name, mode, q in (("exact", fvmgpu.NUMERICS_EXACT, q0), ("exact_ulp", fvmgpu.NUMERICS_EXACT, q1),
                          ("fast", fvmgpu.NUMERICS_FAST, q0)):
        with fvmgpu.Context(0) as ctx:
            ctx.scheme(recon=fvmgpu.MUSCL, limiter=fvmgpu.LIM_VANALBADA, flux=fvmgpu.ROE,
                       rk=fvmgpu.RK4, kappa=0.33333, entropy_fix=0.2, gamma=1.4)
            ctx.numerics(mode)
            b = ctx.block((128, 128, 128), (h, h, h))
            b.set_periodic()
            b.upload(q)
            ctx.step(1000, cfl=0.8)
            out[name] = b.download()
    return {"steps": 1000, "tolerance": 1e-9,
            "fast_vs_exact": rel_linf(out["fast"], out["exact"]),
            "exact_vs_exact_one_ulp_initial_state": rel_linf(out["exact_ulp"], out["exact"]),
            "exact_vs_oracle": "bit-identical (tests/test_gpu_fast.py, test_gpu_parity.py)"}


def summary(ln):
    r = ln.get("roofline") or {}
    out = {"workload": ln["config"]["workload"], "value": ln["value"], "unit": ln["unit"],
           "ms_per_step": ln["ms_per_step"], "steps": ln["steps"],
           "numerics": ln["config"].get("numerics")}
    if r:
        out.update(roofline_frac=r["frac"], avg_launch_us=r["avg_launch_us"], kernel=r["kernel"])
    if "tile_xy_chunk" in ln["config"]:
        out["tile_xy_chunk"] = ln["config"]["tile_xy_chunk"]
    return out


def extras(args):
    """The fast family on the headline workload, the other workloads at their
    stated sizes and the serving workload: separate runs after the headline's
    timed region (which they do not touch)."""
    ns = argparse.Namespace(**vars(args))
    ns.no_cpu, ns.output_every, ns.n, ns.cfl, ns.seed, ns.t_end = True, 0, None, None, None, None
    out = {}
    fast = argparse.Namespace(**vars(ns))
    fast.numerics, fast.tune_reps = "fast", 3
    out["fast_family"] = summary(bench_vortex(fast))
    out["fast_family"]["gate_1000_steps"] = gate_1000_steps()
    wl = {}
    r = argparse.Namespace(**vars(ns))
    r.steps, r.warmup = None, 5
    wl["riemann400"] = summary(bench_riemann(r))
    if not args.no_cpu:
        # the reference's own Fortran scheme modules on this workload, which
        # is the one the reference itself can run (its case/Riemann-2D setup)
        from fvmgpu import cases
        p, dx = cases.riemann2d(400, seed=0)
        ref = cpu_reference(p, cases.to_conserved(p), dx, 0.5, 10, 400.0 * 400 * 3 * 10)
        if ref is not None:
            ref["cpu_model"] = cpu_model()
            wl["riemann400"]["cpu_baseline"] = ref
    b = argparse.Namespace(**vars(ns))
    b.steps, b.warmup = 60, 5
    ln = bench_blob27(b)
    wl["blob27"] = summary(ln)
    wl["blob27"]["speedup_vs_per_block"] = ln["config"]["speedup_vs_per_block"]
    t = argparse.Namespace(**vars(ns))
    t.steps, t.warmup = 30, 3
    wl["tgv192"] = summary(bench_tgv(t))
    out["workloads"] = wl
    sv = argparse.Namespace(**vars(ns))
    ln = bench_serving(sv)
    c = ln["config"]
    out["serving"] = {k: c[k] for k in (
        "workload", "requests", "p50_time_to_solution_s", "p99_time_to_solution_s",
        "sequential_fifo_p50_s", "sequential_fifo_p99_s", "sequential_plain_value",
        "speedup_vs_sequential", "capacity_value", "capacity_speedup_vs_sequential",
        "capacity_p50_time_to_solution_s", "capacity_p99_time_to_solution_s", "numerics")}
    out["serving"]["value"] = ln["value"]
    out["serving"]["unit"] = "cell-updates/s (aggregate, served at 50 arrivals/s)"
    return out


def main():
    args = parse()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1:
        print(f"bench.py: this library runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is described, not built)", file=sys.stderr)
        sys.exit(2)
    fn = {"vortex128": bench_vortex, "riemann400": bench_riemann, "blob27": bench_blob27,
          "tgv192": bench_tgv, "serving": bench_serving}[args.workload]
    out = fn(args)
    if args.workload == "vortex128" and not args.no_extras and not args.output_every:
        out.update(extras(args))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
