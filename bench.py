#!/usr/bin/env python3
"""Benchmark of the D2Q9 time step on one MI355X.

Workload (the single-large-domain configuration): D2Q9 TRT (Lambda = 3/16) on an
8192 x 4096 grid, uniform inflow u = 0.04 with radiating side boundaries and an absorbing
pressure outlet, a square cylinder of side 256 with 32 x 32 notches cut from each corner
centred at (2048, 2048), Re = 1000.  A "step" is one full time step of that domain: bulk
stream-collide, boundary links, outlet controller and the FP64 force reduction.

The headline value is GLUPS (fluid-node updates per second / 1e9) of the FP32-deviation
storage path over the --steps timed steps, with inputs resident in HBM.  Whatever --steps is,
the precision gate then runs the workload's full length (20000 steps) in FP32 and in FP64 from
the same start and reports the whole-run rates, their ratio and the Cd / Cl' / St error of FP32
against FP64 over the second half of the run (Series::stats), with pass / fail against the
|dCd| < 0.5 % bound.  The roofline object prices the dominant kernel (measured with HIP events
on the stream it runs on) against MI355X HBM peak.  The batched sweep (512 seeded cases) runs in
FP32 and again in FP64, so the drag of every finished case is checked against FP64.  The CPU
baseline is the oracle (C++ restatement of the reference solver, same algorithm, every core the
process may use) timed on a bounded number of steps of the same domain.
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
# vector instruction issue: one wave64 VALU instruction per CU per clock, 256 CUs at 2.4 GHz (the
# rate behind the published 78.6 TFLOP/s FP64 vector peak: 64 lanes x 2 flops per FMA)
VALU_PEAK_G = 256 * 2.4
METRIC = "D2Q9 GLUPS & %HBM peak, FP32 vs FP64; batched solves/s at |dCd|<0.5%"
BYTES_PER_NODE = {0: 144, 1: 72, 2: 36}  # 9 reads + 9 writes per fluid-node update
PREC_NAME = {0: "f64", 1: "f32", 2: "f16"}


def workload_notched(api, nx=8192, ny=4096, u=0.04, d=256.0, re=1000.0, ramp_units=0.0):
    """8192 x 4096, TRT 3/16, notched square D=256 at (2048, 2048), Re 1000.  ramp_units = 0: the
    run starts from the uniform stream (Channel::new with no ramp), so a 20000-step (3.1 D/U) run
    has a drag transient to compare between precisions."""
    body = api.section("recessed", 32.0 / 256.0, 2048.0, 2048.0, d, 0.0)
    return api.ChannelSpec(nx=nx, ny=ny, inflow=api.INFLOW_UNIFORM_OPEN, u=u, nu=u * d / re, body=body,
                           collision=api.TRT, magic=3.0 / 16.0, ramp_steps=int(ramp_units * d / u))


def workload_wake(api):
    """The wake case on 662 x 302: BGK, uniform inflow u = 0.05 between radiating ("free-slip")
    sides, absorbing pressure outlet ("zero-gradient" outflow), circle D = 20 at (160, 151),
    Re = 150; a 5-unit ramp and a 2-unit body kick as the reference's towing tank does."""
    u, d, re = 0.05, 20.0, 150.0
    unit = d / u
    return api.ChannelSpec(nx=662, ny=302, inflow=api.INFLOW_UNIFORM_OPEN, u=u, nu=u * d / re,
                           body=api.circle(160.0, 151.0, 10.0), collision=api.BGK, magic=0.0,
                           ramp_steps=int(5 * unit), kick=api.Kick(160.0, 151.0, 0.4 * u / d, int(5 * unit), int(2 * unit)))


def workload_circle(api):
    """16384 x 8192, TRT 3/16, circle D = 512 at (4096, 4096), u = 0.03, Re = 200, from the
    uniform stream."""
    u, d, re = 0.03, 512.0, 200.0
    return api.ChannelSpec(nx=16384, ny=8192, inflow=api.INFLOW_UNIFORM_OPEN, u=u, nu=u * d / re,
                           body=api.circle(4096.0, 4096.0, 256.0), collision=api.TRT, magic=3.0 / 16.0, ramp_steps=0)


WORKLOADS = {
    # name: (spec builder, reference length, storage precision, default steps, description)
    "notched_8192x4096": (lambda api, scale: workload_notched(api, nx=8192 // scale, ny=4096 // scale, d=256.0 / scale), 256.0, "f32", 20000,
           "D2Q9 TRT 3/16 {nx}x{ny}, notched square D={d:g}, Re=1000, u=0.04, {prec}-deviation storage"),
    "wake_662x302": (lambda api, scale: workload_wake(api), 20.0, "f64", 35660,
           "D2Q9 BGK 662x302 wake, circle D=20 at (160,151), Re=150, u=0.05, {prec} storage"),
    "circle_16384x8192": (lambda api, scale: workload_circle(api), 512.0, "f16", 50000,
           "D2Q9 TRT 3/16 16384x8192, circle D=512, Re=200, u=0.03, {prec}-deviation storage (x2^10)"),
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def run_gpu(vl, api, wl, precision, steps, warmup, scale, keep=False, kernel_timing=True):
    builder, d, _, _, _ = WORKLOADS[wl]
    spec = builder(api, scale)
    d = d / scale if wl == "notched_8192x4096" else d
    t0 = time.time()
    ch = vl.make_channel(spec, precision)
    info = ch.sim.info()
    log(f"[bench] {PREC_NAME[precision]}: built {info.nx}x{info.ny}, {info.fluid_nodes} fluid nodes, "
        f"{info.link_count} links in {time.time() - t0:.1f} s")
    # timed region: `steps` steps through vl_channel_run, in pieces of up to 10000 steps, each
    # followed by a progress line.  Whole step windows (30 steps at 5 steps per pass, 32 at 4)
    # replay captured graphs; the whole groups of K steps left over replay one graph too, which
    # vl_sim_prepare instantiates (host-side set-up: it runs nothing) for the state the timed run
    # starts from, before the warmup steps, so that the warmup runs right before the clock starts
    # and the GPU is not left idle in between.  The fixed cost of a call (control table upload,
    # force history readback) stays in the timed region, so a run of a few groups still reads a
    # little below the full-length rate.
    ch.sim.prepare(min(10000, steps), after=warmup)
    coef_w = ch.run(warmup, d) if warmup else np.zeros((0, 2))
    pieces, done = [], 0
    t0 = time.perf_counter()
    while done < steps:
        n = min(10000, steps - done)
        pieces.append(ch.run(n, d))
        done += n
        if steps > 10000:
            log(f"[bench] {PREC_NAME[precision]}: {done}/{steps} steps, {time.perf_counter() - t0:.1f} s")
    secs = time.perf_counter() - t0
    coef = np.concatenate(pieces)
    state_hash = ch.sim.state_hash() if wl == "wake_662x302" else None
    # kernel pass: the same step launched directly with HIP events around every bulk kernel on
    # the stream it runs on (events cannot sit inside a replayed graph)
    bulk_ms, bulk_n, total_ms, tsteps = 0.0, 0, 0.0, 0
    if kernel_timing:
        ch.sim.set_timing(True)
        ch.run(min(300, steps), d)
        bulk_ms, bulk_n, total_ms, tsteps = ch.sim.timing()
        ch.sim.set_timing(False)
    info = ch.sim.info()
    res = dict(precision=precision, fluid=int(info.fluid_nodes), bnodes=int(info.bnode_count), secs=secs,
               glups=info.fluid_nodes * steps / secs / 1e9, bulk_ms=bulk_ms / max(bulk_n, 1),
               step_ms=total_ms / max(tsteps, 1), coef=np.concatenate([coef_w, coef]), d=d, dt=spec.u / d,
               split=bool(info.split), steps_total=int(info.t), nx=int(info.nx), ny=int(info.ny), spec=spec,
               state_hash=state_hash, pass_steps=int(info.pass_steps), pass_nodes=int(info.pass_nodes),
               pass_tasks=int(info.pass_tasks), listed=int(info.listed_nodes), pass_rows=int(info.pass_rows),
               pass_far_nodes=int(info.pass_far_nodes), pass_far_tasks=int(info.pass_far_tasks))
    kname = (f"{res['pass_steps']}-step pass kernel" if res["pass_steps"] > 1
             else ("bulk kernel" if res["split"] else "step kernel"))
    log(f"[bench] {PREC_NAME[precision]}: {steps} steps in {secs:.2f} s = {res['glups']:.2f} GLUPS"
        + (f", {kname} {res['bulk_ms'] * 1e3:.1f} us, whole step {res['step_ms'] * 1e3:.1f} us" if kernel_timing else ""))
    if keep:
        res["channel"] = ch
    else:
        ch.close()
    return res


def dominant_kernel(r):
    """(name, units, algorithmic bytes per launch) of the kernel that dominates a step.
    k_pass takes its covered interior nodes K steps per launch and moves each node's nine
    populations in and out of HBM once per launch.  The timed launch is the far tasks' pass
    (the near tasks' pass runs beside it on another stream), so its units are the far tasks'
    nodes.  k_bulk moves them once per step for every interior node; fused, k_step does so for
    every fluid node."""
    p, b = PREC_NAME[r["precision"]], BYTES_PER_NODE[r["precision"]]
    if r["pass_steps"] > 1:
        n = r["pass_far_nodes"] if r["pass_far_tasks"] > 0 else r["pass_nodes"]
        return f"k_pass<{p},K={r['pass_steps']}>", n, b * n
    if r["split"]:
        return f"k_bulk<{p}>", r["fluid"] - r["bnodes"], b * (r["fluid"] - r["bnodes"])
    return f"k_step<{p}>", r["fluid"], b * r["fluid"]


def coef_stats(vl, r):
    cd, cl = r["coef"][:, 0], r["coef"][:, 1]
    return vl.series_stats(cd, cl, r["dt"], 0.5 * len(cd) * r["dt"])


DCD_BOUND = 0.005  # |dCd| < 0.5 % against FP64


def precision_gate(vl, api, wl, prec, scale):
    """The workload's full run (its default step count, from t = 0, no warmup) in `prec` and in
    FP64: whole-run GLUPS of both, and the force statistics of `prec` against FP64 over the
    second half of the run (Series::stats, src/cases/channel.rs:166-189)."""
    steps = WORKLOADS[wl][3]
    keep = wl == "circle_16384x8192"
    lo = run_gpu(vl, api, wl, prec, steps, 0, scale, keep=keep, kernel_timing=False)
    hi = run_gpu(vl, api, wl, api.F64, steps, 0, scale, keep=keep, kernel_timing=False)
    s_lo, s_hi = coef_stats(vl, lo), coef_stats(vl, hi)
    dcd = abs(s_lo.cd_mean / s_hi.cd_mean - 1.0)
    gate = {"steps": steps, "storage": PREC_NAME[prec], "glups": round(lo["glups"], 3),
            "fp64_glups": round(hi["glups"], 3), "speedup_vs_fp64": round(lo["glups"] / hi["glups"], 3),
            "stats_window": f"steps {steps // 2}-{steps} (second half)",
            "cd_mean": s_lo.cd_mean, "cd_mean_fp64": s_hi.cd_mean, "dcd_rel": dcd,
            "cl_rms": s_lo.cl_rms, "cl_rms_fp64": s_hi.cl_rms,
            "dcl_rms_rel": abs(s_lo.cl_rms / s_hi.cl_rms - 1.0) if s_hi.cl_rms != 0 else None,
            "st": s_lo.st, "st_fp64": s_hi.st, "bound_dcd_rel": DCD_BOUND, "ok": bool(dcd < DCD_BOUND)}
    if keep:
        gate["velocity_l2_error_vs_fp64"] = velocity_error(lo["channel"], hi["channel"])
        lo["channel"].close()
        hi["channel"].close()
    log(f"[gate] {steps} steps: {PREC_NAME[prec]} {lo['glups']:.1f} GLUPS, f64 {hi['glups']:.1f} GLUPS "
        f"({gate['speedup_vs_fp64']:.2f}x); |dCd| {dcd:.2e} (bound {DCD_BOUND}) -> {'ok' if gate['ok'] else 'FAIL'}")
    return gate


def velocity_error(a, b):
    """Per-node velocity L2 error of run a against run b (fluid nodes; solid nodes are 0 in both)."""
    _, ax, ay = a.sim.macros_all()
    _, bx, by = b.sim.macros_all()
    num = np.sqrt(np.sum((ax - bx) ** 2 + (ay - by) ** 2))
    return float(num / np.sqrt(np.sum(bx ** 2 + by ** 2)))


def host_cpus():
    """(cores this process may use, what nproc reports, the CPU model).  The usable count is the
    affinity mask, further limited by a cgroup CPU quota when one is set."""
    nproc = os.cpu_count() or 1
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else nproc
    try:
        with open("/sys/fs/cgroup/cpu.max") as fh:
            quota, period = fh.read().split()[:2]
        if quota != "max":
            usable = min(usable, max(1, int(int(quota) // int(period))))
    except (OSError, ValueError):
        pass
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return usable, nproc, model


def cpu_baseline(api, wl, seconds_target, scale, steps_full, gpu_hash=None):
    """The oracle on the same domain, on every core this process may use, a bounded number of
    steps (the whole run when it fits in `seconds_target`)."""
    import oracle
    be = oracle.backend()
    threads, nproc, model = host_cpus()
    builder, d, _, _, _ = WORKLOADS[wl]
    spec = builder(api, scale)
    d = d / scale if wl == "notched_8192x4096" else d
    t0 = time.time()
    ch = api.Channel(be, spec, threads)
    fluid = ch.sim.fluid_nodes
    log(f"[cpu] oracle built in {time.time() - t0:.1f} s, {threads} threads")
    t0 = time.perf_counter()
    ch.run(2, d)
    per = (time.perf_counter() - t0) / 2
    steps = max(2, min(steps_full - 2, int(seconds_target / max(per, 1e-6))))
    t0 = time.perf_counter()
    ch.run(steps, d)
    secs = time.perf_counter() - t0
    match = None
    if gpu_hash is not None and steps + 2 == steps_full:
        match = ch.sim.state_hash() == gpu_hash
    ch.close()
    g = fluid * steps / secs / 1e9
    log(f"[cpu] {steps} steps in {secs:.1f} s = {g * 1e3:.1f} MLUPS" + ("" if match is None else f"; state hash "
                                                                         f"matches GPU: {match}"))
    out = {"value": round(g, 5), "unit": "GLUPS", "cores": threads, "kind": "port",
           "sample": f"oracle (C++ restatement of the reference solver), {steps + 2} steps of the "
                     f"{spec.nx}x{spec.ny} domain from t=0, {threads} threads, {secs:.1f} s timed",
           "host_nproc": nproc, "host_cpu_model": model,
           "host_cores_usable": threads}
    if match is not None:
        out["full_run_state_hash_matches_gpu"] = match
    # one thread on the same domain (a few seconds), and the reference's own benchmark (cylinder
    # rig n = 40, 3000 steps, one thread) against its published per-core rate and hash
    ch = api.Channel(be, spec, 1)
    ch.run(1, d)
    t0 = time.perf_counter()
    n1 = 0
    while time.perf_counter() - t0 < 3.0 and n1 + 1 < steps_full:
        ch.run(1, d)
        n1 += 1
    g1 = fluid * n1 / (time.perf_counter() - t0) / 1e9
    ch.close()
    h, fl, _, _, s40 = oracle.cylinder_bench(40, 3000, 1)
    out["value_1thread"] = round(g1, 5)
    out["bench_n40_mlups_1thread"] = round(fl * 3000 / s40 / 1e6, 1)
    out["bench_n40_state_hash_is_published"] = h == 0x6B6814B976BBBAFF
    out["reference_published_mlups_per_core"] = 130.0
    log(f"[cpu] 1 thread: {g1 * 1e3:.1f} MLUPS on the same domain; cylinder bench n=40 (3000 steps, 1 thread): "
        f"{out['bench_n40_mlups_1thread']} MLUPS (published 130 per Apple M5 core), hash published: "
        f"{out['bench_n40_state_hash_is_published']}")
    return out


def batch_bench(vl, api, ncases, seed, sample):
    """The batched sweep: `ncases` seeded towing-tank cases (random star polygons, grid
    200-1400 x 80-600, D 10-60, Re 40-250, BGK/TRT, u = 0.05, 20 D/U each) in FP32 storage, one
    batch on one GPU; the same batch in FP64, for the drag of every case that finishes; then
    `sample` of the cases run one at a time (single-case path, FP32) for the sequential rate."""
    from vortexlab_amd import workloads
    cases = workloads.sweep_cases(seed, ncases)
    b, t_build, t_run = workloads.run_batch(vl.backend(), cases, api.F32)
    info = b.info()
    lu = int(info.lattice_updates)  # updates performed (a retired case counts the steps it ran)
    results = [b.result(k) for k in range(ncases)]
    fluid = [b.case_fluid_nodes(k) for k in range(ncases)]
    setup = {"on_device": bool(info.setup_on_device),
             "ms": dict(zip(("cases_and_tables", "pass_plans", "packing", "autotune"),
                            [round(v, 1) for v in info.setup_ms]))}
    b.close()
    # the same cases' set-up on host threads (the host builder), build only
    os.environ["VL_BATCH_SETUP"] = "host"
    try:
        bh, t_build_host = workloads.build_batch(vl.backend(), cases, api.F32)
        setup["host_builder_seconds"] = round(t_build_host, 3)
        setup["host_builder_ms"] = dict(zip(setup["ms"], [round(v, 1) for v in bh.info().setup_ms]))
        bh.close()
    finally:
        del os.environ["VL_BATCH_SETUP"]
    log(f"[batch] set-up {t_build:.2f} s on the device {setup['ms']}; host builder {t_build_host:.2f} s "
        f"{setup['host_builder_ms']}")
    finished = [k for k in range(ncases) if results[k][2] == 0]
    diverged = ncases - len(finished)
    log(f"[batch] {ncases} cases, {info.fluid_nodes / 1e6:.1f} M fluid nodes, {lu / 1e12:.3f} T updates performed; "
        f"set-up {t_build:.1f} s; tiles {info.tile_w}x{info.tile_h}, grid {info.grid_bulk}")
    log(f"[batch] run {t_run:.2f} s: {len(finished)} finished = {len(finished) / t_run:.2f} cases/s, "
        f"{lu / t_run / 1e9:.1f} GLUPS; {diverged} diverged (retired)")
    # the same cases in FP64: the drag of every finished case against its FP64 run (an FP64
    # batch case is bit-identical to the case run alone in FP64)
    b64, t_build64, t_run64 = workloads.run_batch(vl.backend(), cases, api.F64)
    lu64 = int(b64.info().lattice_updates)
    res64 = [b64.result(k) for k in range(ncases)]
    b64.close()
    dcd, fp64_diverged = {}, 0
    for k in finished:
        if res64[k][2] != 0:
            fp64_diverged += 1
            continue
        h = cases[k].steps // 2
        dcd[k] = abs(results[k][0][h:, 0].mean() / res64[k][0][h:, 0].mean() - 1.0)
    within = [k for k in finished if k in dcd and dcd[k] < DCD_BOUND]
    vals = np.array(list(dcd.values())) if dcd else np.array([np.nan])
    log(f"[batch] FP64 batch {t_run64:.2f} s ({lu64 / t_run64 / 1e9:.1f} GLUPS); |dCd| FP32 vs FP64 over "
        f"{len(dcd)} finished cases: max {np.nanmax(vals):.2e}, median {np.nanmedian(vals):.2e}; "
        f"{len(within)} within {DCD_BOUND}; {fp64_diverged} finished in FP32 but diverged in FP64")
    rng = np.random.default_rng(seed + 1)
    pick = sorted(rng.choice(ncases, size=min(sample, ncases), replace=False).tolist())
    seq_secs, rows, times = 0.0, [], []
    for k in pick:
        c = cases[k]
        ch = vl.make_channel(c.spec, api.F32)
        t0 = time.perf_counter()
        try:
            ch.run(c.steps, c.d)
        except api.Diverged:
            pass
        dt = time.perf_counter() - t0
        seq_secs += dt
        done = int(ch.sim.info().t)
        rows.append((done, done * ch.sim.fluid_nodes))
        times.append(dt)
        ch.close()
    seq_rate = len(pick) / seq_secs
    # one case alone costs a per-step launch time plus a per-update time: fitted on the sample,
    # summed over every case of the batch (with the steps each ran in the batch)
    coeff = np.linalg.lstsq(np.array(rows, dtype=float), np.array(times), rcond=None)[0]
    seq_total = sum(coeff[0] * results[k][1] + coeff[1] * results[k][1] * fluid[k] for k in range(ncases))
    seq_rate_all = len(finished) / seq_total
    log(f"[batch] sequential single-case runs: {len(pick)} cases in {seq_secs:.2f} s = {seq_rate:.2f} cases/s; "
        f"model {coeff[0] * 1e6:.2f} us/step + {1e-9 / coeff[1]:.1f} GLUPS -> {seq_rate_all:.2f} finished cases/s "
        f"over the batch; batched / sequential = {seq_total / t_run:.2f}")
    # The same cases driven the way the reference's case loop drives a Channel (tunnel.rs:76-85:
    # one step, then the force coefficients, per call): 1000 steps each on 8 of the sampled cases,
    # fitted and summed over the batch in the same way.
    rows1, times1 = [], []
    for k in pick[:8]:
        c = cases[k]
        ch = vl.make_channel(c.spec, api.F32)
        n = min(1000, c.steps)
        t0 = time.perf_counter()
        for _ in range(n):
            ch.run(1, c.d)
        times1.append(time.perf_counter() - t0)
        rows1.append((n, n * ch.sim.fluid_nodes))
        ch.close()
    coeff1 = np.linalg.lstsq(np.array(rows1, dtype=float), np.array(times1), rcond=None)[0]
    seq1_total = sum(coeff1[0] * results[k][1] + coeff1[1] * results[k][1] * fluid[k] for k in range(ncases))
    log(f"[batch] driven one step per call: {coeff1[0] * 1e6:.2f} us/step; batched / that = {seq1_total / t_run:.1f}")
    return {"batched_cases_per_s": round(len(finished) / t_run, 3),
            "batched_solves_per_s_at_gate": round(len(within) / t_run, 3),
            "batched_glups": round(lu / t_run / 1e9, 2), "batched_lattice_updates": lu,
            "batched_cases": ncases, "batched_finished": len(finished), "batched_diverged": diverged,
            "batched_seconds": round(t_run, 2), "batched_setup_seconds": round(t_build, 3), "batched_setup": setup,
            "batched_tile": f"{info.tile_w}x{info.tile_h}",
            "batched_pass_steps": int(info.pass_steps), "batched_pass_nodes": int(info.pass_nodes),
            "batched_listed_nodes": int(info.listed_nodes),
            "batched_dcd_checked": len(dcd), "batched_within_gate": len(within),
            "batched_max_dcd_vs_fp64": float(np.nanmax(vals)), "batched_median_dcd_vs_fp64": float(np.nanmedian(vals)),
            "batched_fp64_diverged_only": fp64_diverged, "batched_fp64_seconds": round(t_run64, 2),
            "batched_fp64_glups": round(lu64 / t_run64 / 1e9, 2),
            "sequential_cases_per_s": round(seq_rate_all, 3), "sequential_sample": len(pick),
            "sequential_sample_cases_per_s": round(seq_rate, 3),
            "sequential_step_us": round(coeff[0] * 1e6, 2), "sequential_update_glups": round(1e-9 / coeff[1], 1),
            "batched_vs_sequential": round(seq_total / t_run, 2),
            "sequential_one_step_per_call_step_us": round(coeff1[0] * 1e6, 2),
            "batched_vs_sequential_one_step_per_call": round(seq1_total / t_run, 2)}


def surrogate_setup(vl, api, n=9216, seed=5):
    """Set-up time (build only) of the surrogate's solver batch: n towing-tank cases of 24 cells
    per diameter (8192 training + 1024 held-out shapes), FP32."""
    from vortexlab_amd import surrogate, workloads
    cases, _, _ = surrogate.sample_cases(seed, n)
    b, secs = workloads.build_batch(vl.backend(), cases, api.F32)
    info = b.info()
    out = {"cases": n, "seconds": round(secs, 3), "on_device": bool(info.setup_on_device),
           "ms": dict(zip(("cases_and_tables", "pass_plans", "packing", "autotune"), [round(v, 1) for v in info.setup_ms])),
           "fluid_nodes": int(info.fluid_nodes)}
    b.close()
    log(f"[batch] set-up of {n} surrogate cases: {secs:.2f} s {out['ms']}")
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="notched_8192x4096")
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default: the workload's full run)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps first (default 100; 0 for the wake case)")
    ap.add_argument("--precision", choices=["f32", "f16", "f64"], default=None)
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison runs (and the gate)")
    ap.add_argument("--no-gate", action="store_true",
                    help="skip the full-length FP32 / FP64 runs of the precision gate")
    ap.add_argument("--cpu-seconds", type=float, default=None,
                    help="CPU-baseline sample length (default 15 s; 90 s for the wake case, which covers its whole run; 0 = skip)")
    ap.add_argument("--scale", type=int, default=1, help="divide the notched-square domain by this (testing only)")
    ap.add_argument("--batch-cases", type=int, default=None, help="cases of the batched sweep (0 = skip)")
    ap.add_argument("--batch-sample", type=int, default=32, help="cases re-run one at a time for the comparison")
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
    wl = args.workload
    _, _, prec_default, steps_default, desc = WORKLOADS[wl]
    steps = args.steps if args.steps is not None else steps_default
    warmup = args.warmup if args.warmup is not None else (0 if wl == "wake_662x302" else 100)
    batch_cases = args.batch_cases if args.batch_cases is not None else (512 if wl == "notched_8192x4096" else 0)
    pname = args.precision or prec_default
    prec = {"f64": api.F64, "f32": api.F32, "f16": api.F16}[pname]
    compare = not args.no_fp64 and prec != api.F64
    main_run = run_gpu(vl, api, wl, prec, steps, warmup, args.scale)
    extra = {}
    if compare:
        # FP64 over the same timed steps: the speed ratio at this run length
        ref = run_gpu(vl, api, wl, api.F64, steps, warmup, args.scale)
        extra = {
            "fp64_glups": round(ref["glups"], 3),
            "speedup_vs_fp64": round(main_run["glups"] / ref["glups"], 3),
            "fp64_kernel": dominant_kernel(ref)[0],
            "fp64_kernel_frac_hbm": round(dominant_kernel(ref)[2] / (ref["bulk_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "fp64_glups_frac_hbm": round(ref["glups"] * BYTES_PER_NODE[0] / HBM_PEAK_GBS, 4),
        }
        # the accuracy gate needs the developed force history: the workload's full length in both
        # precisions, whatever --steps is
        if not args.no_gate:
            gate = precision_gate(vl, api, wl, prec, args.scale)
            extra["precision_gate"] = gate
            for key in ("dcd_rel", "dcl_rms_rel", "cd_mean", "cd_mean_fp64", "cl_rms", "cl_rms_fp64", "st", "st_fp64"):
                extra[key] = gate[key]
            extra["gate_ok"] = gate["ok"]
    kname, units, traffic_alg = dominant_kernel(main_run)
    achieved = traffic_alg / (main_run["bulk_ms"] * 1e-3) / 1e9
    traffic = None
    tname = kname[:kname.index("<")] + (f"{main_run['pass_steps']}" if main_run["pass_steps"] > 1 else "")
    tf = os.path.join(ROOT, "profiles", f"{tname}_traffic_{wl}_{PREC_NAME[prec]}.json")
    valu = None
    if os.path.exists(tf) and args.scale == 1:
        with open(tf) as fh:
            prof = json.load(fh)
        traffic = prof.get("hbm_bytes_per_launch")
        if prof.get("valu_insts_per_launch"):
            # vector instructions per launch (SQ_INSTS_VALU of the committed counter run) over this
            # run's launch time, against one wave64 VALU instruction per CU per clock at 2.4 GHz
            rate = prof["valu_insts_per_launch"] / (main_run["bulk_ms"] * 1e-3) / 1e9
            valu = {"bound": "valu", "achieved": round(rate, 1), "peak": VALU_PEAK_G, "unit": "G wave64 instructions/s",
                    "frac": round(rate / VALU_PEAK_G, 4), "insts_per_launch": prof["valu_insts_per_launch"]}
    import ctypes
    copy = ctypes.c_double()
    vl.backend().check(vl.backend().device_copy_bandwidth(1 << 30, 20, ctypes.byref(copy)))
    log(f"[bench] plain device copy (1 GiB): {copy.value:.0f} GB/s")
    if batch_cases > 0:
        extra.update(batch_bench(vl, api, batch_cases, args.batch_seed, args.batch_sample))
        extra["batched_setup_9216_surrogate_cases"] = surrogate_setup(vl, api)
    # the CPU run covers the whole wake job when it fits its time budget: then its final state
    # hash is compared with the GPU run's
    gpu_hash = main_run["state_hash"] if prec == api.F64 and warmup == 0 else None
    cpu_seconds = args.cpu_seconds if args.cpu_seconds is not None else (90.0 if wl == "wake_662x302" else 15.0)
    cpu = (cpu_baseline(api, wl, cpu_seconds, args.scale, steps + warmup, gpu_hash)
           if cpu_seconds > 0 else None)
    out = {
        "metric": METRIC, "value": round(main_run["glups"], 3), "unit": "GLUPS", "n_gpus": 1,
        "steps": steps, "warmup": warmup, "ms_per_step": round(main_run["secs"] / steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": PREC_NAME[prec],
        "data": "synthetic (seeded geometry; flow from t=0)",
        "config": {"workload": desc.format(nx=main_run["nx"], ny=main_run["ny"], d=256.0 / args.scale,
                                           prec=PREC_NAME[prec]),
                   "nx": main_run["nx"], "ny": main_run["ny"], "fluid_nodes": main_run["fluid"],
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": kname, "kernel_us": round(main_run["bulk_ms"] * 1e3, 2),
                     "algorithmic_bytes_per_launch": traffic_alg, "steps_per_launch": max(main_run["pass_steps"], 1),
                     "copy_gbs": round(copy.value, 1), "frac_of_copy": round(achieved / copy.value, 4),
                     "valu": valu},
        # the metric's "%HBM peak in GLUPS": GLUPS x the single-step bytes per update over the
        # peak; above 1 when a pass carries more than one step per HBM round trip
        "glups_frac_hbm": round(main_run["glups"] * BYTES_PER_NODE[prec] / HBM_PEAK_GBS, 4),
        "pass": {"steps": main_run["pass_steps"], "covered_nodes": main_run["pass_nodes"],
                 "far_nodes": main_run["pass_far_nodes"], "listed_nodes": main_run["listed"],
                 "tasks": main_run["pass_tasks"], "far_tasks": main_run["pass_far_tasks"],
                 "rows_per_task": main_run["pass_rows"]},
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
