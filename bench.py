#!/usr/bin/env python3
"""ISPH projection-step benchmark on one MI355X.

Headline workload: the urban street canyon at 2M particles -- domain
2.0 x 1.0 x 0.5 m at dx = 0.008 m, 4x2 cuboid buildings (0.1 x 0.1 x 0.15 m,
0.2 m spacing) carried by IBM markers, log-law inflow (u* = 0.3 m/s,
z0 = 0.001 m), MOST ground layer with Obukhov length L = +50 m, Smagorinsky
LES (Cs = 0.12), dt = 5e-5 s, BiCGSTAB PPE at relative tolerance 1e-6.  The
same steps run in FP64 (the reference's Jacobi-preconditioned BiCGSTAB)
and in MIXED precision with the multigrid-preconditioned inner solve;
``value`` is the MIXED throughput.  A third leg runs MIXED with the
reference's Jacobi preconditioner, which separates what the precision buys
from what the preconditioner buys.

particle-steps/s = n * K / wall time of K timed steps (after W warm-up steps,
the first of which runs the neighbour search and the next four of which time
the autotuner's candidate workgroup sizes), state resident in HBM.

Two more legs ride along: a batched ensemble (64 canyon cases of 2e4..3e5
particles stepped as one launch sequence, against the same cases stepped one
at a time) and a serving run (a seeded Poisson stream of canyon requests
through the continuous-batching server under a 240 GB budget), which gives
the p50/p99 case latency.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for sub in ("urban-sph_a100_amd", "oracle"):
    sys.path.insert(0, os.path.join(ROOT, sub))

import numpy as np  # noqa: E402

METRIC = "ISPH particle-steps/s @2M (mixed vs FP64, same PPE tol); p50/p99 case latency"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


GRAPHS = True   # --no-graphs: launch every BiCGSTAB kernel directly (for counter runs)
INNER_TOL = None   # --inner-tol: floor of the MIXED inner solves' relative tolerance (library default if None)
PARAM_SET = {}     # --set name=value: further sph_params of the headline legs (library defaults otherwise)


def run_precision(cases, abi, urbansph, P, nrm, geom, precision, steps, warmup, dt, tol, precond=0):
    prm = cases.canyon_params(geom, urbansph.default_params(), dt=dt, tol=tol, precision=precision)
    prm.update(ppe_precond=precond)
    if INNER_TOL is not None:
        prm.update(ppe_inner_tol=INNER_TOL)
    prm.update(**PARAM_SET)
    case = urbansph.Case(prm, len(P))
    if not GRAPHS:
        case.options(graphs=False, autotune=True)
    case.upload(P, None, nrm)
    log(f"[bench]   uploaded n={len(P)}")
    count = 0
    t_setup = time.perf_counter()
    first = None
    for _ in range(warmup):
        case.step(count, dt)
        case.sync()
        if first is None:   # the first step from the deck: the per-step precision check
            first = {f: case.field(f) for f in ("PRES", "UX", "UY", "UZ")}
        count += 1
    t_setup = time.perf_counter() - t_setup
    lat = []
    iters = []
    t0 = time.perf_counter()
    for _ in range(steps):
        a = time.perf_counter()
        case.step(count, dt)
        case.sync()
        lat.append((time.perf_counter() - a) * 1e3)
        iters.append(case.ppe_report().iterations)
        count += 1
    wall = time.perf_counter() - t0
    # one more pass with per-kernel HIP-event timing (not part of the timed loop)
    case.profile(True)
    prof_steps = 2
    for _ in range(prof_steps):
        case.step(count, dt)
        count += 1
    case.sync()
    prof = case.profile_report()
    case.profile(False)
    rep = case.ppe_report()
    out = dict(value=len(P) * steps / wall, ms_per_step=wall / steps * 1e3, lat=lat, iters=iters,
               prof=prof, prof_steps=prof_steps, rel_res=rep.rel_residual, refinements=rep.refinements,
               precond=rep.precond, mg_levels=rep.mg_levels, mg_coarsest=rep.mg_coarsest, mg_status=rep.mg_status,
               warmup_s=t_setup, dev_bytes=case.device_bytes(), pres=case.field("PRES"), ux=case.field("UX"),
               first=first, tuning=case.tuning(),
               ppe=dict(rows=rep.unknowns, nnz=rep.nnz, slots=rep.slots, ordering=["cell", "lattice"][rep.ordering],
                        index16=rep.index16, stencil_slices=rep.stencil_slices, stencil_nnz=rep.stencil_nnz,
                        slices=(rep.unknowns + 63) // 64))
    case.close()
    return out


def run_first_step(cases, abi, urbansph, P, nrm, geom, dt, tol, precision=None):
    """Pressure and velocity after the first step from the deck (FP64 unless
    another precision is given; MIXED runs the multigrid)."""
    precision = abi.FP64 if precision is None else precision
    prm = cases.canyon_params(geom, urbansph.default_params(), dt=dt, tol=tol, precision=precision)
    if precision == abi.MIXED:
        prm.update(ppe_precond=abi.PRECOND_MULTIGRID)
    case = urbansph.Case(prm, len(P))
    case.upload(P, None, nrm)
    case.step(0, dt)
    case.sync()
    out = {f: case.field(f) for f in ("PRES", "UX", "UY", "UZ")}
    case.close()
    return out


def precision_report(first_fp64, first_mixed, first_conv, floor_mixed):
    """Relative L2 distances after the first step from the deck: MIXED and
    FP64 at the configured PPE tolerance against FP64 converged to 1e-4 of
    it, and the float floor (MIXED at that tight tolerance), for the pressure
    and the velocity vector."""
    def rel(a, b):
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))

    def vec(d):
        return np.concatenate([d["UX"], d["UY"], d["UZ"]])
    out = {}
    for name, d in (("mixed", first_mixed), ("fp64", first_fp64), ("mixed_float_floor", floor_mixed)):
        out[name] = {"pres": rel(d["PRES"], first_conv["PRES"]), "u": rel(vec(d), vec(first_conv))}
    return out


def measured_traffic(precision):
    """HBM bytes per SpMV launch from the committed rocprofv3 counter
    summary (FETCH_SIZE x2.0 + WRITE_SIZE x1.0, factors from a 1 GiB copy);
    None when the summary is absent."""
    import re
    path = os.path.join(ROOT, "profiles", "spmv_canyon2m.md")
    if not os.path.exists(path):
        return None
    key = "mixed" if precision == "mixed" else "fp64"
    m = re.search(r"traffic_bytes_per_launch_%s_spmv: (\d+)" % key, open(path).read())
    return int(m.group(1)) if m else None


FP32_VECTOR_PEAK_TFLOPS = 157.3
GATHER_PASSES = ("ppe_assemble", "advection", "prep", "pressure_force", "neumann")


def interaction_rooflines(prof):
    """The gather passes against their bounds: FLOP per launch and memory-
    fabric bytes per launch (FETCH_SIZE x2 + WRITE_SIZE) from the committed
    rocprofv3 summary, divided by this run's average launch time (HIP events
    on the engine's stream)."""
    import re
    path = os.path.join(ROOT, "profiles", "interaction_canyon2m.md")
    if not os.path.exists(path):
        return None
    text = open(path).read()
    out = {}
    for k in GATHER_PASSES:
        m = re.search(r"interaction_%s: fetch_bytes (\d+) write_bytes (\d+) flop (\d+)" % k, text)
        if not m or k not in prof:
            continue
        launches, ms, _ = prof[k]
        avg_s = ms / launches / 1e3
        fab, flop = int(m.group(1)) + int(m.group(2)), int(m.group(3))
        out[k] = {"avg_us": round(avg_s * 1e6, 1), "flop_per_launch": flop,
                  "tflops": round(flop / avg_s / 1e12, 3),
                  "frac_fp32_vector_peak": round(flop / avg_s / 1e12 / FP32_VECTOR_PEAK_TFLOPS, 4),
                  "fabric_bytes_per_launch": fab, "fabric_GBps": round(fab / avg_s / 1e9, 1),
                  "frac_hbm_peak": round(fab / avg_s / 1e9 / HBM_PEAK_GBS, 3)}
        v = re.search(r"interaction_%s: .* valu_instructions (\d+)" % k, text)
        if v:
            # wave64 VALU instructions x 4 cycles over 1024 SIMDs at the 2.4 GHz peak clock
            out[k]["valu_wave_instructions"] = int(v.group(1))
            out[k]["valu_issue_share"] = round(int(v.group(1)) * 4 / (1024 * 2.4e9 * avg_s), 3)
    return out or None


def run_batched(cases, abi, urbansph, steps, warmup, dt, tol, n_cases=64):
    """The 64-case ensemble (counts log-uniform in [2e4, 3e5], inflow scale
    U(0.5, 2), building height U(0.05, 0.2), seed 7) stepped as one batch,
    then the same cases stepped one at a time with sph_step."""
    ens = cases.ensemble(n_cases=n_cases)
    ntot = sum(len(e[0]) for e in ens)
    log(f"[bench] batched: {n_cases} cases, {ntot} particles")
    cs = []
    for P, nrm, geom, info in ens:
        prm = cases.canyon_params(geom, urbansph.default_params(), dt=dt, tol=tol, precision=abi.MIXED)
        prm.update(ppe_precond=abi.PRECOND_MULTIGRID)
        c = urbansph.Case(prm, len(P))
        c.upload(P, None, nrm)
        cs.append(c)
    b = urbansph.Batch()
    k = len(cs)
    count = 0
    for _ in range(warmup):
        b.step(cs, [count] * k, [dt] * k)
        b.sync()
        count += 1
    t0 = time.perf_counter()
    for _ in range(steps):
        b.step(cs, [count] * k, [dt] * k)
        count += 1
    b.sync()
    wall = time.perf_counter() - t0
    iters = [c.ppe_report().iterations for c in cs]
    out = dict(value=ntot * steps / wall, ms_per_step=wall / steps * 1e3, particles=ntot, cases=k,
               tuning=b.tuning(), ppe_iters_max=max(iters), ppe_iters_min=min(iters),
               dev_bytes=sum(c.device_bytes() for c in cs))
    log(f"[bench] batched: {out['value'] / 1e6:.2f} M particle-steps/s, {out['ms_per_step']:.1f} ms/step")
    # the same cases one at a time (their own engines; block size 256)
    seq_steps = 2
    for c in cs:
        c.options(graphs=True, autotune=False)
    for c in cs:
        c.step(count, dt)
    for c in cs:
        c.sync()
    count += 1
    t0 = time.perf_counter()
    for _ in range(seq_steps):
        for c in cs:
            c.step(count, dt)
            c.sync()
        count += 1
    wall = time.perf_counter() - t0
    out["sequential_value"] = ntot * seq_steps / wall
    out["speedup_vs_sequential"] = out["value"] / out["sequential_value"]
    log(f"[bench] sequential: {out['sequential_value'] / 1e6:.2f} M particle-steps/s")
    b.close()
    for c in cs:
        c.close()
    return out


def run_apr(cases, abi, urbansph, steps, warmup, dt, tol, r1=0.05, r2=0.0125):
    """The multi-resolution canyon: the 2M canyon refined by APR to level 1
    (dx 0.004 m) within r1 and level 2 (dx 0.002 m) within r2 of a building
    wall or roof (the MOST layer keeps its resolution), mixed precision.
    Throughput counts live particles.  The steps are the first ones after
    refinement: level-2 particles beside coarse boundary particles let the
    state drift within a few steps (the FP64 oracle does the same), so the
    leg times few steps and reports whether each solve converged."""
    P, nrm, geom = cases.canyon(dx=0.008)
    t_deck = cases.apr_targets(P["x"], P["y"], P["z"], P["p_type"], P["i_type"], geom["buildings"], r1, r2,
                               most=P["MOST_buffer"])
    cap = len(P) + 8 * int((t_deck == 1).sum()) + 72 * int((t_deck == 2).sum()) + 1024
    prm = cases.canyon_params(geom, urbansph.default_params(), dt=dt, tol=tol, precision=abi.MIXED)
    prm.update(ppe_precond=abi.PRECOND_MULTIGRID)   # the solve reports Jacobi where rows have tails
    c = urbansph.Case(prm, cap)
    c.upload(P, None, nrm)
    c.step(0, dt)
    c.sync()
    t = cases.apr_targets(c.field("X"), c.field("Y"), c.field("Z"), c.ifield("P_TYPE"), c.ifield("I_TYPE"),
                          geom["buildings"], r1, r2, most=c.ifield("MOST_BUFFER"))
    a = time.perf_counter()
    n = c.apr_refine(t, max_passes=200)
    t_apr = time.perf_counter() - a
    live = int((c.ifield("I_TYPE") != 3).sum())
    h = c.field("H")[c.ifield("I_TYPE") != 3]
    log(f"[bench] apr: {len(P)} -> {live} live particles ({n} with parents) in {t_apr:.1f} s; "
        f"h levels {sorted(set(np.round(h / h.max(), 3)))}")
    count = 1
    for k in range(warmup):
        a = time.perf_counter()
        c.step(count, dt)
        c.sync()
        log(f"[bench] apr: warmup step {k + 1}/{warmup} {1e3 * (time.perf_counter() - a):.0f} ms, "
            f"PPE iters {c.ppe_report().iterations}")
        count += 1
    t0 = time.perf_counter()
    iters = []
    for k in range(steps):
        c.step(count, dt)
        c.sync()
        iters.append(c.ppe_report().iterations)
        count += 1
        log(f"[bench] apr: step {k + 1}/{steps} at {time.perf_counter() - t0:.1f} s, PPE iters {iters[-1]}")
    wall = time.perf_counter() - t0
    rep = c.ppe_report()
    c.profile(True)
    c.step(count, dt)
    c.sync()
    prof = c.profile_report()
    c.profile(False)
    out = dict(value=live * steps / wall, ms_per_step=wall / steps * 1e3, particles_live=live, particles_total=n,
               coarse=len(P), refine_s=t_apr, iters=iters, rel_res=rep.rel_residual, converged=rep.converged,
               ppe_rows=rep.unknowns, ppe_nnz=rep.nnz, ppe_slots=rep.slots, ppe_ordering=rep.ordering,
               ppe_index16=rep.index16, ppe_tail_rows=rep.tail_rows, heavy_lists=rep.heavy_lists,
               ppe_precond="multigrid" if rep.precond == abi.PRECOND_MULTIGRID else "jacobi", mg_status=rep.mg_status,
               kernel_ms=dict(sorted(((k, round(v[1], 2)) for k, v in prof.items()), key=lambda kv: -kv[1])[:12]))
    log(f"[bench] apr: {out['value'] / 1e6:.2f} M particle-steps/s, {out['ms_per_step']:.1f} ms/step, iters {iters}")
    log(f"[bench] apr: kernel ms of one step {out['kernel_ms']}")
    c.close()
    return out


def run_ale(cases, abi, urbansph, P, nrm, geom, steps, warmup, dt, tol):
    """The 2M canyon in Lagrangian/ALE mode (eli = 1 for fluid): the
    corrector moves the particles and every step runs the neighbour search
    and rebuilds the PPE structure, mixed precision."""
    P = P.copy()
    P["eli"] = np.where(P["p_type"] == 1, 1.0, 0.0).astype(np.float32)
    prm = cases.canyon_params(geom, urbansph.default_params(), dt=dt, tol=tol, precision=abi.MIXED)
    prm.update(ale_positions=1, ppe_precond=abi.PRECOND_MULTIGRID)   # hierarchy rebuilt every step
    prm.update(**PARAM_SET)
    c = urbansph.Case(prm, len(P))
    if not GRAPHS:
        c.options(graphs=False, autotune=True)
    c.upload(P, None, nrm)
    count = 0
    for k in range(warmup):
        a = time.perf_counter()
        c.step(count, dt)
        c.sync()
        log(f"[bench] ale: warmup step {k + 1}/{warmup} {1e3 * (time.perf_counter() - a):.0f} ms, "
            f"PPE iters {c.ppe_report().iterations}")
        count += 1
    t0 = time.perf_counter()
    iters = []
    for k in range(steps):
        c.step(count, dt)
        c.sync()
        iters.append(c.ppe_report().iterations)
        count += 1
        log(f"[bench] ale: step {k + 1}/{steps} at {time.perf_counter() - t0:.1f} s, PPE iters {iters[-1]}")
    wall = time.perf_counter() - t0
    rep = c.ppe_report()
    c.profile(True)
    c.step(count, dt)
    c.sync()
    prof = c.profile_report()
    c.profile(False)
    out = dict(value=len(P) * steps / wall, ms_per_step=wall / steps * 1e3, iters=iters, rel_res=rep.rel_residual,
               kernel_ms_sum_one_step=round(sum(v[1] for v in prof.values()), 2),
               converged=rep.converged, ppe_precond="multigrid" if rep.precond == abi.PRECOND_MULTIGRID else "jacobi",
               kernel_ms=dict(sorted(((k, round(v[1], 2)) for k, v in prof.items()), key=lambda kv: -kv[1])[:20]))
    log(f"[bench] ale: {out['value'] / 1e6:.2f} M particle-steps/s, {out['ms_per_step']:.1f} ms/step, iters {iters}")
    log(f"[bench] ale: kernel ms of one step {out['kernel_ms']}")
    c.close()
    return out


def run_serving(cases, abi, urbansph, dt, tol, budget=240e9, n_req=16, rate=4.0, lo=1e4, hi=2e6, steps_lo=50,
                steps_hi=500):
    """A seeded Poisson stream of canyon requests through the server: wall
    clock, requests submitted at their arrival times, one co-batched step per
    tick.  Latency = completion - arrival."""
    reqs = cases.serving_requests(n_req=n_req, rate=rate, lo=lo, hi=hi, steps_lo=steps_lo, steps_hi=steps_hi)
    decks = []
    for r in reqs:
        P, nrm, geom = cases.sized_canyon(r["target"], r["u_scale"], r["bh"])
        prm = cases.canyon_params(geom, urbansph.default_params(), dt=dt, tol=tol, precision=abi.MIXED, L=r["L"])
        prm.update(ppe_precond=abi.PRECOND_MULTIGRID)
        decks.append((P, nrm, prm))
    log(f"[bench] serving: {n_req} requests at {rate}/s, sizes {min(len(d[0]) for d in decks)}.."
        f"{max(len(d[0]) for d in decks)}, steps {min(r['steps'] for r in reqs)}..{max(r['steps'] for r in reqs)}")
    srv = urbansph.Server(budget_bytes=int(budget), max_cases=256)
    t0 = srv.clock()
    ids = []
    nxt = 0
    max_running = 0
    while nxt < len(reqs) or srv.status().running or srv.status().queued:
        now = srv.clock() - t0
        while nxt < len(reqs) and reqs[nxt]["t_arrival"] <= now:
            P, nrm, prm = decks[nxt]
            ids.append(srv.submit(prm, P, reqs[nxt]["steps"], dt, normals=nrm, t_arrival=t0 + reqs[nxt]["t_arrival"]))
            nxt += 1
        st = srv.status()
        if not st.running and not st.queued:
            time.sleep(max(0.0, reqs[nxt]["t_arrival"] - (srv.clock() - t0)))
            continue
        srv.tick()
        max_running = max(max_running, srv.status().running)
    infos = [srv.request(i) for i in ids]
    lat = np.array([i.t_finish - i.t_arrival for i in infos])
    wait = np.array([i.t_admit - i.t_arrival for i in infos])
    st = srv.status()
    span = max(i.t_finish for i in infos) - (t0 + reqs[0]["t_arrival"])
    offered = sum(len(d[0]) * r["steps"] for d, r in zip(decks, reqs)) / max(reqs[-1]["t_arrival"], 1e-9)
    rc_name = {abi.E_NOCONV: "no PPE convergence", abi.E_NONFINITE: "state went non-finite"}
    failed = [dict(request=k, particles=i.n, steps_done=i.steps_done, steps=i.steps,
                   why=rc_name.get(i.rc, f"rc {i.rc}"), L=reqs[k]["L"], u_scale=round(reqs[k]["u_scale"], 3),
                   building_height=round(reqs[k]["bh"], 3))
              for k, i in enumerate(infos) if i.state != abi.REQ_DONE]
    for f in failed:
        log(f"[bench] serving: request {f['request']} (n={f['particles']}) failed after {f['steps_done']} of "
            f"{f['steps']} steps: {f['why']}")
    out = dict(requests=n_req, rate_per_s=rate, failed=len(failed), failed_requests=failed,
               offered_particle_steps_per_s=offered, span_s=span, lo=lo, hi=hi, steps_lo=steps_lo, steps_hi=steps_hi,
               latencies_s=sorted(round(float(x), 3) for x in lat),
               p50_ms=float(np.percentile(lat, 50) * 1e3), p99_ms=float(np.percentile(lat, 99) * 1e3),
               admit_wait_p99_ms=float(np.percentile(wait, 99) * 1e3),
               value=st.particle_steps / span, ticks=st.ticks, max_running=max_running,
               budget_bytes=int(budget))
    log(f"[bench] serving: p50 {out['p50_ms']:.0f} ms, p99 {out['p99_ms']:.0f} ms, "
        f"{out['value'] / 1e6:.2f} M particle-steps/s over {span:.1f} s, max {max_running} running")
    srv.close()
    return out


def host_cpu():
    """Model name and physical core count of this host (lscpu), and the
    CPUs this process may run on."""
    import subprocess
    info = {}
    try:
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        kv = dict(l.split(":", 1) for l in txt.splitlines() if ":" in l)
        kv = {k.strip(): v.strip() for k, v in kv.items()}
        info["model"] = kv.get("Model name")
        cps, sk = kv.get("Core(s) per socket"), kv.get("Socket(s)")
        if cps and sk and cps.isdigit() and sk.isdigit():
            info["physical_cores"] = int(cps) * int(sk)
        info["logical_cpus"] = int(kv["CPU(s)"]) if kv.get("CPU(s)", "").isdigit() else None
    except (OSError, subprocess.SubprocessError):
        pass
    info["cpus_available_to_process"] = len(os.sched_getaffinity(0))
    return info


def cpu_baseline(cases, abi, P, nrm, geom, dt, tol, max_seconds=20.0):
    """The FP64 oracle (OpenMP, one thread per available CPU) on the same
    2.13 M-particle canyon: the searching first step untimed, then steps
    while the bound allows (at least one)."""
    import oracle
    threads = oracle.set_threads(int(os.environ.get("OMP_NUM_THREADS", "0") or 0))
    prm = cases.canyon_params(geom, oracle.default_params(), dt=dt, tol=tol)
    o = oracle.Oracle(prm, len(P))
    o.upload(P, None, nrm)
    a = time.perf_counter()
    o.step(0, dt)                       # neighbour search + first solve (not timed)
    t_first = time.perf_counter() - a
    times = []
    count = 1
    t_all = time.perf_counter()
    while count == 1 or (count <= 3 and time.perf_counter() - t_all + times[-1] < max_seconds):
        a = time.perf_counter()
        o.step(count, dt)
        times.append(time.perf_counter() - a)
        count += 1
    o.close()
    med = float(np.median(times))
    return dict(value=len(P) / med, unit="particle-steps/s", cores=threads, kind="port",
                host=host_cpu(), seconds_per_step=[round(t, 2) for t in times], first_step_s=round(t_first, 2),
                sample=f"the same urban canyon, n={len(P)}, FP64 BiCGSTAB (Jacobi) tol {tol:g}, "
                       f"median of {len(times)} step(s) after the searching first step, {threads} OpenMP threads")


def run_box_cfg0(cases, abi, urbansph, oracle_too=True, gpu_steps=20):
    """The reference's CPU-runnable case: the 32^3 FP64 box (dx = 0.01 m,
    jitter 0.05 dx, Wendland C2, h = 1.5 dx, 3 wall layers, dt = 1e-4 s,
    PPE tolerance 1e-8), FP64 on the GPU and the oracle on 1 host thread and
    on all of them (median of 3 steps after the searching first step)."""
    P, nrm = cases.box(nf=32, walls=3)
    prm = cases.box_params(urbansph.default_params(), tol=1e-8)
    dt = 1e-4
    g = urbansph.Case(prm, len(P))
    g.upload(P, None, nrm)
    for c in range(3):
        g.step(c, dt)
    g.sync()
    t0 = time.perf_counter()
    for c in range(3, 3 + gpu_steps):
        g.step(c, dt)
    g.sync()
    gpu = len(P) * gpu_steps / (time.perf_counter() - t0)
    g.close()
    out = {"workload": "32^3 FP64 box, dx=0.01 m, jitter 0.05 dx, Wendland C2, 3 wall layers, dt=1e-4 s, "
                       "BiCGSTAB tol 1e-8", "particles": len(P), "gpu_fp64": round(gpu, 1),
           "unit": "particle-steps/s"}
    if oracle_too:
        import oracle
        oprm = cases.box_params(oracle.default_params(), tol=1e-8)
        every = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
        for label, threads in (("cpu_all_threads", every), ("cpu_1_thread", 1)):
            used = oracle.set_threads(threads)
            o = oracle.Oracle(oprm, len(P))
            o.upload(P, None, nrm)
            o.step(0, dt)
            times = []
            for c in range(1, 4):
                a = time.perf_counter()
                o.step(c, dt)
                times.append(time.perf_counter() - a)
            o.close()
            out[label] = {"value": round(len(P) / float(np.median(times)), 1), "threads": used}
        oracle.set_threads(every)
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=6)
    ap.add_argument("--dx", type=float, default=0.008)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--dt", type=float, default=5e-5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-ale", action="store_true", help="skip the Lagrangian/ALE leg")
    ap.add_argument("--no-graphs", action="store_true",
                    help="headline legs without HIP graphs (each Krylov kernel launched directly)")
    ap.add_argument("--no-apr", action="store_true")
    ap.add_argument("--batch-steps", type=int, default=5)
    ap.add_argument("--serve-requests", type=int, default=16)
    ap.add_argument("--only", choices=["fp64", "mixed", "mixed_jacobi", "fp64_multigrid"], default=None)
    ap.add_argument("--inner-tol", type=float, default=None,
                    help="floor of the MIXED inner solves' relative tolerance (sph_params.ppe_inner_tol)")
    ap.add_argument("--set", action="append", default=[], metavar="NAME=VALUE",
                    help="an sph_params member of the headline legs (int or float value), repeatable")
    ap.add_argument("--copy-calib", action="store_true",
                    help="run only a 1 GiB device-to-device copy kernel (for scaling HBM counters)")
    args = ap.parse_args()
    if args.copy_calib:
        # 1 GiB device-to-device copy: the reference for scaling the HBM
        # byte counters, and the achievable copy bandwidth of this HBM
        import torch
        a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(10):
            b.copy_(a)
        e1.record()
        torch.cuda.synchronize()
        sec = e0.elapsed_time(e1) / 1e3 / 10
        # read-only rate: a sum over 4 GiB (reads dominate, one float written)
        big = torch.ones(1 << 30, dtype=torch.float32, device="cuda")
        for _ in range(3):
            big.sum()
        torch.cuda.synchronize()
        e0.record()
        for _ in range(10):
            big.sum()
        e1.record()
        torch.cuda.synchronize()
        rsec = e0.elapsed_time(e1) / 1e3 / 10
        print(json.dumps({"copy_calib_bytes_read": a.numel() * 4, "copy_calib_bytes_written": a.numel() * 4,
                          "copy_GBps_read_plus_write": round(2 * a.numel() * 4 / sec / 1e9, 1),
                          "read_only_GBps": round(big.numel() * 4 / rsec / 1e9, 1)}))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported in this build "
              f"(multi-GPU halo exchange is design-only)", file=sys.stderr)
        sys.exit(2)
    args.warmup = max(1, args.warmup)
    global GRAPHS, INNER_TOL
    GRAPHS = not args.no_graphs
    INNER_TOL = args.inner_tol
    for kv in args.set:
        k, v = kv.split("=", 1)
        try:
            PARAM_SET[k] = int(v, 0)
        except ValueError:
            PARAM_SET[k] = float(v)

    import abi
    import cases
    import urbansph
    urbansph.lib()

    t = time.perf_counter()
    P, nrm, geom = cases.canyon(dx=args.dx)
    n = len(P)
    log(f"[bench] canyon n={n} (fluid {geom['n_fluid']}, wall {geom['n_wall']}, lid {geom['n_lid']}, "
        f"markers {geom['n_marker']}) built in {time.perf_counter() - t:.1f}s")

    res = {}
    legs = (("fp64", abi.FP64, abi.PRECOND_JACOBI), ("mixed", abi.MIXED, abi.PRECOND_MULTIGRID),
            ("mixed_jacobi", abi.MIXED, abi.PRECOND_JACOBI), ("fp64_multigrid", abi.FP64, abi.PRECOND_MULTIGRID))
    for name, prec, pc in legs:
        if args.only and args.only != name:
            continue
        log(f"[bench] {name}: warmup {args.warmup}, timed {args.steps}")
        res[name] = run_precision(cases, abi, urbansph, P, nrm, geom, prec, args.steps, args.warmup, args.dt,
                                  args.tol, precond=pc)
        r = res[name]
        log(f"[bench] {name}: {r['value'] / 1e6:.2f} M particle-steps/s, {r['ms_per_step']:.1f} ms/step, "
            f"PPE iters {r['iters']}, refinements {r['refinements']}, rel res {r['rel_res']:.2e}, "
            f"preconditioner {'multigrid' if r['precond'] == abi.PRECOND_MULTIGRID else 'jacobi'} "
            f"(levels {r['mg_levels']}, coarsest {r['mg_coarsest']})")

    main_name = "mixed" if "mixed" in res else ("mixed_jacobi" if "mixed_jacobi" in res else "fp64")
    m = res[main_name]
    # dominant kernel: the SpMV of the PPE solve
    sp = m["prof"].get("spmv")
    roof = None
    if sp:
        launches, ms, byts = sp
        avg_s = ms / launches / 1e3
        achieved = byts / launches / avg_s / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic("fp64" if main_name == "fp64" else "mixed"),
                "kernel": "spmv", "avg_us": round(avg_s * 1e6, 2),
                "alg_bytes_per_launch": round(byts / launches)}
    breakdown = {k: round(v[1] / m["prof_steps"], 3) for k, v in sorted(m["prof"].items(), key=lambda kv: -kv[1][1])}
    lat = np.array(m["lat"])
    line = {
        "metric": METRIC,
        "value": round(m["value"], 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(m["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair sums + f32 inner Krylov, f64 accumulation/dots/refinement" if main_name != "fp64" else "f64",
        "data": "synthetic (generated street-canyon deck, no dataset)",
        "config": {"workload": f"urban street canyon 2.0x1.0x0.5 m, dx={args.dx} m, 4x2 IBM buildings, "
                               f"MOST L=+50 m, SSM Cs=0.12, dt={args.dt} s, BiCGSTAB tol {args.tol:g}",
                   "particles": n, "precision": "fp64" if main_name == "fp64" else "mixed",
                   "ppe_precond": "multigrid" if m["precond"] == abi.PRECOND_MULTIGRID else "jacobi",
                   "mg_levels": m["mg_levels"], "mg_coarsest": m["mg_coarsest"], "mg_status": m["mg_status"],
                   "ppe": m["ppe"]},
        "step_latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                            "p99": round(float(np.percentile(lat, 99)), 3),
                            "what": "per-step latency of the resident 2M case"},
        "tuning": m["tuning"],
        "ppe_iters_per_step": m["iters"],
        "kernel_ms_per_step": breakdown,
        "roofline": roof,
        "interaction_rooflines": interaction_rooflines(m["prof"]),
    }
    if "fp64" in res:
        line["fp64"] = {"value": round(res["fp64"]["value"], 1), "ms_per_step": round(res["fp64"]["ms_per_step"], 3),
                        "ppe_iters_per_step": res["fp64"]["iters"]}
        if "mixed_jacobi" in res:
            line["mixed_jacobi"] = {
                "value": round(res["mixed_jacobi"]["value"], 1),
                "ms_per_step": round(res["mixed_jacobi"]["ms_per_step"], 3),
                "ppe_iters_per_step": res["mixed_jacobi"]["iters"],
                "over_fp64": round(res["mixed_jacobi"]["value"] / res["fp64"]["value"], 3),
                "what": "MIXED with the reference's Jacobi preconditioner: the precision's share of the speed-up"}
        if "fp64_multigrid" in res:
            fm = res["fp64_multigrid"]
            line["fp64_multigrid"] = {
                "value": round(fm["value"], 1), "ms_per_step": round(fm["ms_per_step"], 3),
                "ppe_iters_per_step": fm["iters"], "mg_levels": fm["mg_levels"],
                "ppe_precond": "multigrid" if fm["precond"] == abi.PRECOND_MULTIGRID else "jacobi",
                "what": "FP64 BiCGSTAB with the same multigrid preconditioner as the headline MIXED leg"}
            if "mixed" in res:
                line["fp64_multigrid"]["mixed_over_this"] = round(res["mixed"]["value"] / fm["value"], 3)
                line["fp64_multigrid"]["note"] = ("mixed_over_this: the precision's share of the speed-up at the same "
                                                  "preconditioner; fp64_multigrid over fp64: the preconditioner's share "
                                                  "in FP64")
                line["fp64_multigrid"]["over_fp64"] = round(fm["value"] / res["fp64"]["value"], 3)
        if "mixed" in res:
            line["mixed_over_fp64"] = round(res["mixed"]["value"] / res["fp64"]["value"], 3)

            def rel(a, b):
                return float(np.linalg.norm(a - b) / np.linalg.norm(b))
            # against FP64 converged to 1e-4 of the tolerance, one step from
            # the deck; the float floor is MIXED at that tight tolerance
            conv = run_first_step(cases, abi, urbansph, P, nrm, geom, args.dt, args.tol * 1e-4)
            floor = run_first_step(cases, abi, urbansph, P, nrm, geom, args.dt, args.tol * 1e-4, precision=abi.MIXED)
            pr = precision_report(res["fp64"]["first"], res["mixed"]["first"], conv, floor)
            bound = (pr["mixed"]["pres"] <= 2 * pr["fp64"]["pres"] + 2 * pr["mixed_float_floor"]["pres"]
                     and pr["mixed"]["u"] <= 2 * pr["fp64"]["u"] + 2 * pr["mixed_float_floor"]["u"]
                     and pr["mixed_float_floor"]["pres"] <= 5e-6 and pr["mixed_float_floor"]["u"] <= 5e-6)
            line["mixed_vs_fp64_rel_l2"] = {
                "vs_converged_fp64_one_step": pr,
                "bound_met": bool(bound),
                "mixed_within_1e-5_of_converged": bool(max(pr["mixed"].values()) <= 1e-5),
                "mixed_vs_fp64_same_tol_one_step": {
                    "pres": rel(res["mixed"]["first"]["PRES"], res["fp64"]["first"]["PRES"]),
                    "ux": rel(res["mixed"]["first"]["UX"], res["fp64"]["first"]["UX"])},
                "after_run": {"pres": rel(res["mixed"]["pres"], res["fp64"]["pres"]),
                              "ux": rel(res["mixed"]["ux"], res["fp64"]["ux"])},
                "note": f"vs_converged_fp64_one_step: the first step from the deck, MIXED (multigrid) and FP64 (Jacobi) "
                        f"at PPE tol {args.tol:g} and MIXED at {args.tol * 1e-4:g} (the float floor), each against "
                        f"FP64 at {args.tol * 1e-4:g}; bound: MIXED <= 2 x FP64 at the same tol + 2 x float floor, "
                        f"float floor <= 5e-6 (pressure and velocity vector); after_run: the two timed runs after "
                        f"warmup+steps+2 steps, where tolerance differences have been carried through the dynamics"}
            log(f"[bench] precision vs converged FP64: {pr}")
    if not args.no_batch:
        bt = run_batched(cases, abi, urbansph, args.batch_steps, 5, args.dt, args.tol)
        line["batched"] = {
            "workload": "64 canyon cases, particle counts log-uniform in [2e4, 3e5] (seed 7), inflow scale U(0.5, 2), "
                        "building height U(0.05, 0.2), dx=0.008 m, mixed precision with the multigrid preconditioner, "
                        "one launch sequence per step",
            "particles": bt["particles"], "value": round(bt["value"], 1), "unit": "particle-steps/s",
            "ms_per_step": round(bt["ms_per_step"], 3), "steps": args.batch_steps, "warmup": 5,
            "sequential_value": round(bt["sequential_value"], 1),
            "speedup_vs_sequential": round(bt["speedup_vs_sequential"], 3),
            "ppe_iters_range": [bt["ppe_iters_min"], bt["ppe_iters_max"]], "tuning": bt["tuning"],
            "device_bytes": bt["dev_bytes"]}
    if not args.no_apr:
        ar = run_apr(cases, abi, urbansph, 3, 2, args.dt, args.tol)
        line["apr"] = {
            "workload": "street canyon refined by APR: dx 0.008 m far field, 0.004 m within 0.05 m and 0.002 m within "
                        "0.0125 m of building walls and roofs (MOST layer unrefined), mixed precision, "
                        "first steps after refinement",
            "particles_live": ar["particles_live"], "particles_with_parents": ar["particles_total"],
            "value": round(ar["value"], 1), "unit": "particle-steps/s (live particles)",
            "ms_per_step": round(ar["ms_per_step"], 3), "steps": 3, "warmup": 2,
            "refine_seconds": round(ar["refine_s"], 2), "ppe_iters_per_step": ar["iters"],
            "converged": ar["converged"], "kernel_ms_one_step": ar["kernel_ms"], "ppe_precond": ar["ppe_precond"],
            "mg_status": ar["mg_status"],
            "ppe": {"rows": ar["ppe_rows"], "nnz": ar["ppe_nnz"], "sell_slots": ar["ppe_slots"],
                    "ordering": ["cell", "lattice"][ar["ppe_ordering"]], "index16": ar["ppe_index16"], "tail_rows": ar["ppe_tail_rows"]},
            "heavy_lists": ar["heavy_lists"]}
    if not args.no_ale:
        al = run_ale(cases, abi, urbansph, P, nrm, geom, 3, 2, args.dt, args.tol)
        line["ale"] = {
            "workload": "the same 2M canyon in Lagrangian/ALE mode: fluid particles move (eli = 1), neighbour search, "
                        "PPE structure and multigrid hierarchy rebuilt every step, mixed precision",
            "particles": n, "value": round(al["value"], 1), "unit": "particle-steps/s",
            "ms_per_step": round(al["ms_per_step"], 3), "steps": 3, "warmup": 2, "ppe_iters_per_step": al["iters"],
            "converged": al["converged"], "kernel_ms_one_step": al["kernel_ms"], "ppe_precond": al["ppe_precond"],
            "kernel_ms_sum_one_step": al["kernel_ms_sum_one_step"],
            "note": "kernel_ms_sum_one_step: the timed launches of one profiled step (passes, search, PPE structure, multigrid set-up; HIP events); ms_per_step minus it is host time and the untimed small launches"}
    if not args.no_serve:
        sv = run_serving(cases, abi, urbansph, args.dt, args.tol, n_req=args.serve_requests)
        line["latency_ms"] = {"p50": round(sv["p50_ms"], 1), "p99": round(sv["p99_ms"], 1),
                              "what": "serving: completion - arrival per canyon request"}
        line["serving"] = {
            "workload": f"{sv['requests']} canyon requests, Poisson arrivals {sv['rate_per_s']}/s, counts log-uniform "
                        f"in [{sv['lo']:g}, {sv['hi']:g}], {sv['steps_lo']}..{sv['steps_hi']} steps, "
                        "L in {-50, +50, inf} m, mixed precision with the multigrid preconditioner, 240 GB budget, "
                        "state resident, steps of concurrent requests co-batched",
            "value": round(sv["value"], 1), "unit": "particle-steps/s", "p50_ms": round(sv["p50_ms"], 1),
            "p99_ms": round(sv["p99_ms"], 1), "admit_wait_p99_ms": round(sv["admit_wait_p99_ms"], 1),
            "offered_particle_steps_per_s": round(sv["offered_particle_steps_per_s"], 1),
            "span_s": round(sv["span_s"], 2), "latencies_s": sv["latencies_s"],
            "ticks": sv["ticks"], "max_running": sv["max_running"], "failed": sv["failed"],
            "failed_requests": sv["failed_requests"],
            "note": "the offered load (particle-steps per second of arrivals) exceeds one GPU's throughput, so "
                    "latency includes the time requests share the device with those admitted before them; a "
                    "request whose state goes non-finite is retired as failed at that step while the others go on"}
    if not args.no_cpu:
        log("[bench] cpu baseline (oracle) ...")
        line["cpu_baseline"] = cpu_baseline(cases, abi, P, nrm, geom, args.dt, args.tol)
        log(f"[bench] cpu baseline {line['cpu_baseline']['value'] / 1e3:.1f} k particle-steps/s "
            f"on {line['cpu_baseline']['cores']} threads")
        log("[bench] 32^3 box (the reference's CPU case): GPU FP64 and the oracle on 1 and all threads ...")
        line["box_cfg0"] = run_box_cfg0(cases, abi, urbansph)
        b0 = line["box_cfg0"]
        log(f"[bench] box: GPU FP64 {b0['gpu_fp64'] / 1e6:.2f} M, oracle {b0['cpu_all_threads']['value'] / 1e3:.1f} k "
            f"({b0['cpu_all_threads']['threads']} threads), {b0['cpu_1_thread']['value'] / 1e3:.1f} k (1 thread) "
            f"particle-steps/s")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
