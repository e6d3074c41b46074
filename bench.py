#!/usr/bin/env python3
"""Benchmark of the D2Q9 time step on one MI355X.

Workload (the single-large-domain configuration): D2Q9 TRT (Lambda = 3/16) on an
8192 x 4096 grid, uniform inflow u = 0.04 with radiating side boundaries and an absorbing
pressure outlet, a square cylinder of side 256 with 32 x 32 notches cut from each corner
centred at (2048, 2048), Re = 1000.  A "step" is one full time step of that domain: bulk
stream-collide, boundary links, outlet controller and the FP64 force reduction.

The headline value is GLUPS (fluid-node updates per second / 1e9) of the FP32-deviation
storage path with inputs resident in HBM; the same steps are run with FP64 storage for the
speed ratio and the Cd / Cl' / St error of FP32 against FP64.  The roofline object prices the
bulk kernel (measured with HIP events on the library's own stream) against MI355X HBM peak.
The CPU baseline is the oracle (C++ restatement of the reference solver, same algorithm, all
host cores) timed on a bounded number of steps of the same domain.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "D2Q9 GLUPS & %HBM peak, FP32 vs FP64; batched solves/s at |dCd|<0.5%"
BYTES_PER_NODE = {0: 144, 1: 72, 2: 36}  # 9 reads + 9 writes per fluid-node update
PREC_NAME = {0: "f64", 1: "f32", 2: "f16"}


def workload(api, nx=8192, ny=4096, u=0.04, d=256.0, re=1000.0, ramp_units=0.0):
    """ramp_units = 0: the run starts from the uniform stream (Channel::new with no ramp), so a
    20000-step (3.1 D/U) run has a drag transient to compare between precisions."""
    body = api.section("recessed", 32.0 / 256.0, 2048.0, 2048.0, d, 0.0)
    return api.ChannelSpec(nx=nx, ny=ny, inflow=api.INFLOW_UNIFORM_OPEN, u=u, nu=u * d / re, body=body,
                           collision=api.TRT, magic=3.0 / 16.0, ramp_steps=int(ramp_units * d / u))


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def run_gpu(vl, api, precision, steps, warmup, scale):
    spec = workload(api, nx=8192 // scale, ny=4096 // scale, d=256.0 / scale)
    t0 = time.time()
    ch = vl.make_channel(spec, precision)
    info = ch.sim.info()
    log(f"[bench] {PREC_NAME[precision]}: built {info.nx}x{info.ny}, {info.fluid_nodes} fluid nodes, "
        f"{info.link_count} links in {time.time() - t0:.1f} s")
    coef_w = ch.run(warmup, spec.body.args[4]) if warmup else np.zeros((0, 2))
    # timed region: the production path (replayed step-window graphs)
    t0 = time.perf_counter()
    coef = ch.run(steps, spec.body.args[4])
    secs = time.perf_counter() - t0
    # kernel pass: the same step launched directly with HIP events around every bulk kernel on
    # the stream it runs on (events cannot sit inside a replayed graph)
    ch.sim.set_timing(True)
    ch.run(min(300, steps), spec.body.args[4])
    bulk_ms, bulk_n, total_ms, tsteps = ch.sim.timing()
    ch.sim.set_timing(False)
    res = dict(precision=precision, fluid=int(info.fluid_nodes), bnodes=int(info.bnode_count), secs=secs,
               glups=info.fluid_nodes * steps / secs / 1e9, bulk_ms=bulk_ms / max(bulk_n, 1),
               step_ms=total_ms / max(tsteps, 1), coef=np.concatenate([coef_w, coef]), d=spec.body.args[4],
               dt=spec.u / spec.body.args[4])
    log(f"[bench] {PREC_NAME[precision]}: {steps} steps in {secs:.2f} s = {res['glups']:.2f} GLUPS, "
        f"bulk kernel {res['bulk_ms'] * 1e3:.1f} us, whole step {res['step_ms'] * 1e3:.1f} us")
    ch.close()
    return res


def coef_stats(vl, r):
    cd, cl = r["coef"][:, 0], r["coef"][:, 1]
    return vl.series_stats(cd, cl, r["dt"], 0.5 * len(cd) * r["dt"])


def cpu_baseline(api, seconds_target, scale):
    """The oracle on the same domain, all host cores (capped at 16), bounded step count."""
    import oracle
    be = oracle.backend()
    threads = max(1, min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()))
    spec = workload(api, nx=8192 // scale, ny=4096 // scale, d=256.0 / scale)
    t0 = time.time()
    ch = api.Channel(be, spec, threads)
    fluid = ch.sim.fluid_nodes
    log(f"[cpu] oracle built in {time.time() - t0:.1f} s, {threads} threads")
    t0 = time.perf_counter()
    ch.run(2, spec.body.args[4])
    per = (time.perf_counter() - t0) / 2
    steps = max(2, int(seconds_target / max(per, 1e-6)))
    t0 = time.perf_counter()
    ch.run(steps, spec.body.args[4])
    secs = time.perf_counter() - t0
    ch.close()
    g = fluid * steps / secs / 1e9
    log(f"[cpu] {steps} steps in {secs:.1f} s = {g * 1e3:.1f} MLUPS")
    return {"value": round(g, 5), "unit": "GLUPS", "cores": threads, "kind": "port",
            "sample": f"oracle (C++ restatement of the reference solver), {steps} steps of the "
                      f"{8192 // scale}x{4096 // scale} domain from t=0, {threads} threads, {secs:.1f} s"}


def batch_bench(vl, api, ncases, seed, sample):
    """The batched sweep: `ncases` seeded towing-tank cases (random star polygons, grid
    200-1400 x 80-600, D 10-60, Re 40-250, BGK/TRT, u = 0.05, 20 D/U each) in FP32 storage, one
    batch on one GPU; then `sample` of the same cases run one at a time (single-case path, FP32,
    and FP64 for the drag check)."""
    from vortexlab_amd import workloads
    cases = workloads.sweep_cases(seed, ncases)
    be = vl.backend()
    t0 = time.perf_counter()
    b = api.Batch(be, api.F32)
    for c in cases:
        b.add(c.spec, c.steps, c.d)
    b.build()
    t_build = time.perf_counter() - t0
    info = b.info()
    log(f"[batch] {ncases} cases, {info.fluid_nodes / 1e6:.1f} M fluid nodes, {info.lattice_updates / 1e12:.3f} T "
        f"updates; set-up {t_build:.1f} s; tiles {info.tile_w}x{info.tile_h}, grid {info.grid_bulk}")
    t0 = time.perf_counter()
    b.run()
    t_run = time.perf_counter() - t0
    results = [b.result(k) for k in range(ncases)]
    diverged = sum(1 for r in results if r[2] != 0)
    log(f"[batch] run {t_run:.2f} s = {ncases / t_run:.2f} cases/s, {info.lattice_updates / t_run / 1e9:.1f} GLUPS; "
        f"{diverged} diverged")
    rng = np.random.default_rng(seed + 1)
    pick = sorted(rng.choice(ncases, size=min(sample, ncases), replace=False).tolist())
    seq_secs, dcd = 0.0, []
    for k in pick:
        c = cases[k]
        ch = vl.make_channel(c.spec, api.F32)
        t0 = time.perf_counter()
        try:
            ch.run(c.steps, c.d)
        except api.Diverged:
            pass
        seq_secs += time.perf_counter() - t0
        ch.close()
        coef, done, st = results[k]
        if st == 0:
            ref = vl.make_channel(c.spec, api.F64)
            try:
                rc = ref.run(c.steps, c.d)
                h = c.steps // 2
                dcd.append(abs(coef[h:, 0].mean() / rc[h:, 0].mean() - 1.0))
            except api.Diverged:
                pass
            ref.close()
    seq_rate = len(pick) / seq_secs
    log(f"[batch] sequential single-case runs: {len(pick)} cases in {seq_secs:.2f} s = {seq_rate:.2f} cases/s; "
        f"max |dCd| vs FP64 {max(dcd) if dcd else float('nan'):.2e} over {len(dcd)} cases")
    b.close()
    return {"batched_cases_per_s": round(ncases / t_run, 3), "batched_glups": round(info.lattice_updates / t_run / 1e9, 2),
            "batched_cases": ncases, "batched_diverged": diverged, "batched_seconds": round(t_run, 2),
            "batched_setup_seconds": round(t_build, 2), "batched_tile": f"{info.tile_w}x{info.tile_h}",
            "sequential_cases_per_s": round(seq_rate, 3), "sequential_sample": len(pick),
            "batched_vs_sequential": round((ncases / t_run) / seq_rate, 2),
            "batched_max_dcd_vs_fp64": (max(dcd) if dcd else None), "batched_dcd_checked": len(dcd)}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--precision", choices=["f32", "f16", "f64"], default="f32")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison run")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="CPU-baseline sample length (0 = skip)")
    ap.add_argument("--scale", type=int, default=1, help="divide the domain by this (testing only)")
    ap.add_argument("--batch-cases", type=int, default=512, help="cases of the batched sweep (0 = skip)")
    ap.add_argument("--batch-sample", type=int, default=8, help="cases re-run one at a time for the comparison")
    ap.add_argument("--batch-seed", type=int, default=2024)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); multi-GPU runs are not part of "
              f"this build", file=sys.stderr)
        sys.exit(2)

    import vortexlab_amd as vl
    from vortexlab_amd import api
    if vl.device_count() < 1:
        raise SystemExit("no HIP device visible")
    prec = {"f64": api.F64, "f32": api.F32, "f16": api.F16}[args.precision]
    main_run = run_gpu(vl, api, prec, args.steps, args.warmup, args.scale)
    extra = {}
    if not args.no_fp64 and prec != api.F64:
        ref = run_gpu(vl, api, api.F64, args.steps, args.warmup, args.scale)
        s_main, s_ref = coef_stats(vl, main_run), coef_stats(vl, ref)
        cl_ref = s_ref.cl_rms if s_ref.cl_rms != 0 else 1.0
        extra = {
            "fp64_glups": round(ref["glups"], 3),
            "speedup_vs_fp64": round(main_run["glups"] / ref["glups"], 3),
            "cd_mean": s_main.cd_mean, "cd_mean_fp64": s_ref.cd_mean,
            "dcd_rel": abs(s_main.cd_mean / s_ref.cd_mean - 1.0),
            "dcl_rms_rel": abs(s_main.cl_rms / cl_ref - 1.0),
            "st": s_main.st, "st_fp64": s_ref.st,
            "fp64_bulk_frac_hbm": round(BYTES_PER_NODE[0] * (ref["fluid"] - ref["bnodes"]) / (ref["bulk_ms"] * 1e-3)
                                        / 1e9 / HBM_PEAK_GBS, 4),
        }
    # the dominant kernel (k_bulk) updates the interior fluid nodes; boundary nodes run concurrently
    traffic_alg = BYTES_PER_NODE[prec] * (main_run["fluid"] - main_run["bnodes"])
    achieved = traffic_alg / (main_run["bulk_ms"] * 1e-3) / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", f"bulk_traffic_{PREC_NAME[prec]}.json")
    if os.path.exists(tf) and args.scale == 1:
        with open(tf) as fh:
            traffic = json.load(fh).get("hbm_bytes_per_launch")
    import ctypes
    copy = ctypes.c_double()
    vl.backend().check(vl.backend().device_copy_bandwidth(1 << 30, 20, ctypes.byref(copy)))
    log(f"[bench] plain device copy (1 GiB): {copy.value:.0f} GB/s")
    if args.batch_cases > 0:
        extra.update(batch_bench(vl, api, args.batch_cases, args.batch_seed, args.batch_sample))
    cpu = cpu_baseline(api, args.cpu_seconds, args.scale) if args.cpu_seconds > 0 else None
    out = {
        "metric": METRIC, "value": round(main_run["glups"], 3), "unit": "GLUPS", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(main_run["secs"] / args.steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": PREC_NAME[prec],
        "data": "synthetic (uniform stream at t=0 past the notched square)",
        "config": {"workload": f"D2Q9 TRT 3/16 {8192 // args.scale}x{4096 // args.scale}, notched square "
                               f"D={256 // args.scale}, Re=1000, u=0.04, {PREC_NAME[prec]}-deviation storage",
                   "nx": 8192 // args.scale, "ny": 4096 // args.scale, "fluid_nodes": main_run["fluid"],
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": f"k_bulk<{PREC_NAME[prec]}>", "kernel_us": round(main_run["bulk_ms"] * 1e3, 2),
                     "algorithmic_bytes_per_launch": traffic_alg,
                     "copy_gbs": round(copy.value, 1), "frac_of_copy": round(achieved / copy.value, 4)},
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
