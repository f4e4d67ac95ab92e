#!/usr/bin/env python3
"""Benchmark of the SOFI2D time loop on one MI355X.

Workload (headline configuration): viscoelastic isotropic, L = 3 relaxation
mechanisms (tau = 0.05 at 15 Hz), FD order 8 (Holberg 0.1 %), 4000 x 2000
grid, DH = 2.5 m, DT = 2.5e-4 s, NT = 8000, free surface + 20-point CPML,
12 layers with a seeded 5 % Gaussian perturbation, 15 Hz Ricker explosive
source, 1000 receivers (gpi_project_amd.workloads.config2).

A "step" is one time step of the whole loop (velocity update, stress update,
source, free surface, receiver sampling).  W warm-up steps run on a shot that
is then reset; the K timed steps are steps 1..K of a fresh shot, so with the
default K = NT the timed region is the complete time loop.  Prints one JSON
line with GLUPS (NX*NY*K / t / 1e9), the HBM roofline of the dominant kernel
(the stress update, HIP events on the handle's stream) and the reference's
CPU time loop on this host's cores as a baseline.
"""
from __future__ import annotations

import argparse
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "GLUPS & %HBM roofline (8th-order visco); batched shots/s, p50/p99 shot latency"


def algorithmic_bytes(L: int):
    """Per cell and step, two-sweep leapfrog, perfect stencil reuse:
    velocity kernel reads sxx syy sxy vx vy rip rjp, writes vx vy (9 floats);
    stress kernel reads vx vy sxx syy sxy, r p q (3L), f g fipjp and three
    defect bases, writes sxx syy sxy and r p q (20 + 6L floats)."""
    vel = 4 * (7 + 2)
    stress = 4 * (5 + 3 * L + 6 + 3 + 3 * L)
    return vel, stress


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(case_dir: str, json_name: str, nx: int, ny: int, steps: int, cores: int):
    """The reference built from its own sources (oracle/_ref/sofi2D_ref) run
    with MPI on `cores` host cores for `steps` time steps of the same
    workload; falls back to the FP32 oracle on one core if it cannot run."""
    ref = os.path.join(ROOT, "oracle", "_ref", "sofi2D_ref")
    mpiexec = shutil.which("mpiexec") or "/opt/conda/bin/mpiexec"
    px = cores
    while px > 1 and (nx % px or cores % px or ny % (cores // px)):
        px -= 1
    py = cores // px
    txt = open(os.path.join(case_dir, json_name)).read()
    dt = 2.5e-4
    txt = txt.replace('"NPROCX" : "1"', f'"NPROCX" : "{px}"').replace('"NPROCY" : "1"', f'"NPROCY" : "{py}"')
    import re
    txt = re.sub(r'"TIME" : "[^"]*"', f'"TIME" : "{steps * dt!r}"', txt)
    js = os.path.join(case_dir, "cpu_" + json_name)
    open(js, "w").write(txt)
    if os.path.exists(ref) and os.path.exists(mpiexec):
        try:
            log(f"cpu baseline: reference on {px * py} ranks ({px}x{py}), {steps} steps")
            r = subprocess.run([mpiexec, "-np", str(px * py), ref, os.path.basename(js)], cwd=case_dir,
                               capture_output=True, text=True, timeout=600)
            t = [float(l.split()[2]) for l in r.stdout.splitlines() if l.startswith("REF_TIME_LOOP")]
            if r.returncode == 0 and t:
                return {"value": nx * ny * steps / t[0] / 1e9, "unit": "GLUPS", "cores": px * py, "kind": "reference",
                        "sample": f"{steps} time steps of the same {nx}x{ny} workload, reference time_loop wall time "
                                  f"{t[0]:.2f} s on {px}x{py} MPI ranks"}
            log("reference run failed: " + (r.stdout[-500:] + r.stderr[-500:]))
        except Exception as e:  # noqa: BLE001
            log(f"reference run failed: {e}")
    # fallback: the CPU oracle on one core
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from gpi_project_amd.sofi import Problem
    cwd = os.getcwd()
    os.chdir(case_dir)
    try:
        P = Problem(json_name)
        o = pyoracle.Oracle(P, "f32")
        n = max(2, steps // 10)
        t0 = time.perf_counter()
        o.run(1, n)
        t = time.perf_counter() - t0
        o.close()
        P.close()
    finally:
        os.chdir(cwd)
    return {"value": nx * ny * n / t / 1e9, "unit": "GLUPS", "cores": 1, "kind": "port",
            "sample": f"{n} time steps of the same workload on the FP32 CPU oracle"}


CAL_BYTES = 512 << 20   # calibration copy for the counter passes


def pmc_child(case_dir: str, launch: str, steps: int):
    """Run under rocprofv3 --pmc: one device-to-device copy of CAL_BYTES (a
    plain copy kernel of known size), then `steps` time steps of the workload
    in case_dir."""
    import ctypes as C
    import glob
    from gpi_project_amd.sofi import Problem, Simulation
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(CAL_BYTES)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(CAL_BYTES)) == 0
    assert hip.hipMemset(a, 0, C.c_size_t(CAL_BYTES)) == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(CAL_BYTES), 3) == 0      # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    os.chdir(case_dir)
    js = [f for f in glob.glob("*.json") if not f.startswith("cpu_")][0]
    P = Problem(js)
    sim = Simulation(P)
    sim.set_launch(*launch.split(":"))
    sim.load_shot(nsnap=0)
    sim.run(1, steps)
    sim.close()
    P.close()


def pmc_traffic(case_dir: str, launch: str, steps: int = 20):
    """HBM bytes per step of the velocity and stress kernels from two
    rocprofv3 counter passes (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch),
    each scaled by what the same counter reads for the calibration copy
    (CAL_BYTES read and written).  "velocity" sums the k_velocity* kernels,
    "stress" the k_stress* and k_tti_frame* kernels (the TTI stress
    half-step is a fused interior kernel and frame passes), over the `steps`
    steps of the one-shot run.  Returns ({kind: bytes per step}, factors,
    {kernel: {counter: bytes per dispatch, "dispatches": n}}) or
    (None, reason, None)."""
    import csv
    import glob
    import re
    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(prof):
        return None, "rocprofv3 not found", None
    env = dict(os.environ, TMPDIR="/tmp")
    per = {}
    factors = {}
    for cnt in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="sofi_pmc_", dir="/tmp")
        cmd = ["timeout", "-s", "KILL", "150", prof, "--pmc", cnt, "-f", "csv", "-d", d, "-o", "pmc", "--",
               sys.executable, "-u", os.path.abspath(__file__), "--pmc-child", case_dir, "--launch", launch,
               "--pmc-steps", str(steps)]
        log(f"counter pass {cnt}")
        r = subprocess.run(cmd, capture_output=True, text=True, cwd="/tmp", env=env)
        files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
        if r.returncode != 0 or not files:
            shutil.rmtree(d, ignore_errors=True)
            return None, f"rocprofv3 --pmc {cnt} failed (rc {r.returncode}): {(r.stdout + r.stderr)[-300:]}", None
        vals = {}
        for row in csv.DictReader(open(files[0])):
            name = re.sub(r"\(anonymous namespace\)::", "", row["Kernel_Name"]).replace("void ", "").split("(")[0]
            vals.setdefault(name, []).append(float(row["Counter_Value"]))
        shutil.rmtree(d, ignore_errors=True)
        cal = max(vals.get("__amd_rocclr_copyBuffer", [0.0]))
        if cal <= 0:
            return None, f"no calibration copy in the {cnt} pass", None
        factors[cnt] = (CAL_BYTES / 1024) / cal
        for name, v in vals.items():
            if name.startswith("k_"):
                per.setdefault(name, {})[cnt] = sum(v) / len(v) * factors[cnt] * 1024
                per[name]["dispatches"] = len(v)
                log(f"  {cnt} {name}: {sum(v) / len(v):.1f} KiB per dispatch ({len(v)} dispatches)")
        log(f"  {cnt} calibration copy of {CAL_BYTES >> 20} MiB: {cal:.1f} KiB -> factor {factors[cnt]:.4f}")
    out = {}
    for name, v in per.items():
        if "FETCH_SIZE" not in v or "WRITE_SIZE" not in v:
            continue
        kind = ("velocity" if name.startswith("k_velocity") else
                "stress" if name.startswith("k_stress") or name.startswith("k_tti_frame") else None)
        if kind:
            out[kind] = out.get(kind, 0.0) + (v["FETCH_SIZE"] + v["WRITE_SIZE"]) * v["dispatches"] / steps
    return out, factors, per


def bench_batch(args, weq: str):
    """Batched anisotropic shots on one medium (workloads.config3): the
    whole batch advanced by the same launches vs the same shots one after the
    other on a one-shot handle.  Prints one JSON line (shots/s)."""
    import numpy as np
    from gpi_project_amd import workloads
    from gpi_project_amd.sofi import Batch, Problem, Simulation
    td = tempfile.mkdtemp(prefix="sofi_bench_")
    gnx, gny = (int(v) for v in args.grid.lower().split("x"))
    js = workloads.config3(td, weq, nshots=args.shots, nx=gnx, ny=gny, nt=args.nt)
    cwd = os.getcwd()
    os.chdir(td)
    P = Problem(os.path.basename(js))
    d = P.desc
    cells = d.nx * d.ny
    P_ntr = P.receivers()[0].size
    log(f"{weq}: {d.nx}x{d.ny}, NT={d.nt}, FD{d.fdorder}, {P.nshots} shots, {P_ntr} receivers")
    # one at a time (first args.single shots, full NT each)
    sim = Simulation(P)
    sim.set_launch(*args.launch.split(":"))
    sim.load_shot(nsnap=0)
    sim.run(1, min(50, d.nt))               # warm-up
    t_one = []
    ref = None
    for ishot in range(1, args.single + 1):
        P.set_shot(ishot)
        sim.load_shot(nsnap=0)
        t_one.append(sim.run())
        if ishot == 1:
            ref = sim.section(0)
    sim.close()
    ms_one = float(np.mean(t_one))
    log(f"one at a time: {ms_one:.1f} ms per shot ({cells * d.nt / (ms_one * 1e-3) / 1e9:.2f} GLUPS)")
    B = Batch(P, P.nshots)
    B.set_resident(args.resident)
    B.set_streams(args.streams)
    if args.block >= 0:
        B.set_block(args.block)
    B.load_shots(list(range(1, P.nshots + 1)))
    ms_b = B.run()                          # the autotuner's trial steps are part of the timed run
    cfg = B.config
    grp, spt = cfg["slots_per_launch"], cfg["slots_per_thread"]
    res = cfg["resident"]
    s0 = B.section_of(0, 0)
    same = np.array_equal(s0, ref)
    err0 = float(np.linalg.norm(s0.astype(np.float64) - ref) / (np.linalg.norm(ref.astype(np.float64)) + 1e-300))
    B.close()
    P.close()
    os.chdir(cwd)
    traffic = None
    if not args.no_traffic:
        # HBM bytes per step of one shot on a one-shot handle (the first shot of the workload)
        traffic, info, kern = pmc_traffic(td, args.launch)
        log(f"one-shot traffic per step {traffic}, counter factors {info}, kernels {kern}")
    shutil.rmtree(td, ignore_errors=True)
    nb = args.shots
    glups_b = nb * cells * d.nt / (ms_b * 1e-3) / 1e9
    p_s = {"EL_ISO": 3, "EL_VTI": 4}.get(weq, 8)
    share = cfg["slot_block"] if cfg["slot_block"] > 0 else spt      # slots that share one read of a cell's medium
    bytes_b = 4 * (15 + (2 + p_s) / share)
    log(f"batched: {ms_b:.1f} ms for {nb} shots ({glups_b:.2f} GLUPS), launch {cfg}, "
        f"slot 0 vs one-shot run: bitwise {same}, rel-L2 {err0:.1e}")
    out = {
        "metric": "batched shots/s", "value": round(nb / (ms_b * 1e-3), 4), "unit": "shots/s", "n_gpus": 1,
        "steps": d.nt, "warmup": 50, "ms_per_step": round(ms_b / d.nt, 5), "higher_is_better": True,
        "scaling": "strong", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic: layered model with smooth random Thomsen fields (workloads.config3)",
        "config": {"workload": f"{nb} batched {weq} shots, {d.nx}x{d.ny} FD6, NT={d.nt}, {P_ntr} receivers each",
                   "nx": d.nx, "ny": d.ny, "nt": d.nt, "fdorder": d.fdorder, "weq": weq, "shots": nb,
                   "launch": cfg},
        "one_at_a_time_shots_per_s": round(1e3 / ms_one, 4),
        "speedup_vs_one_at_a_time": round((nb / ms_b) / (1.0 / ms_one), 3),
        "batched_glups": round(glups_b, 3),
        "roofline": {"bound": "hbm", "achieved": round(bytes_b * nb * cells * d.nt / (ms_b * 1e-3) / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(bytes_b * nb * cells * d.nt / (ms_b * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": None, "kernel": "whole step (all kernels)",
                     "algorithmic_bytes_per_shot_cell": round(bytes_b, 2)},
        "slot0_bitwise_equal_to_one_shot_run": bool(same), "slot0_rel_l2_vs_one_shot_run": err0,
        "one_shot": {"glups": round(cells * d.nt / (ms_one * 1e-3) / 1e9, 3),
                     "algorithmic_bytes_per_cell": 4 * (15 + 2 + p_s),
                     "achieved_gbs": round(4 * (15 + 2 + p_s) * cells * d.nt / (ms_one * 1e-3) / 1e9, 1),
                     "traffic_bytes_per_step": ({k: round(v) for k, v in traffic.items()} if traffic else None),
                     "traffic_ratio": (round(sum(traffic.values()) / (4 * (15 + 2 + p_s) * cells), 3)
                                       if traffic else None)},
    }
    print(json.dumps(out), flush=True)


def bench_serve(args):
    """Served shot requests (workloads.config4_requests): each request is a
    complete run directory (parameter file, model files, sources) of its own
    size, order, rheology and NT.  Three passes over the same requests:
      one at a time  every request set up, run and read back alone, in order:
                     the service times and the one-at-a-time capacity
      served         Poisson arrivals at 1.5x that capacity through the
                     serving loop (admission by footprint, concurrent
                     handles): shots/s and p50/p99 latency
      saturated      every request waiting at t = 0: the serving loop's
                     throughput ceiling
    The one-at-a-time latencies under the same arrivals follow from the
    measured service times (a FIFO single server).  Every served result is
    compared bit for bit with the request's one-at-a-time result."""
    import numpy as np
    from gpi_project_amd import serve, workloads
    td = tempfile.mkdtemp(prefix="sofi_serve_")
    q4 = workloads.config4_requests(args.requests)
    t0 = time.time()
    paths = [workloads.config4_write(os.path.join(td, f"r{q.rid:03d}"), q) for q in q4]
    cell_steps = sum(q.nx * q.ny * q.nt for q in q4)
    byts = sum((80 if q.weq == "EL_ISO" else 164) * q.nx * q.ny * q.nt for q in q4)
    log(f"{len(q4)} requests written in {time.time() - t0:.1f} s ({cell_steps / 1e9:.1f} G cell-steps)")
    mk = lambda: [serve.Request(q.rid, p) for q, p in zip(q4, paths)]
    # warm-up: the smallest request once (code objects, allocator)
    w = min(q4, key=lambda q: q.nx * q.ny * q.nt)
    serve.run_sequential([serve.Request(w.rid, paths[w.rid])], args.launch)
    alone = mk()
    seq = serve.run_sequential(alone, args.launch)
    svc = np.diff([0.0] + [r.end for r in alone])          # all waiting at t = 0: back to back
    mean_svc = seq["span_s"] / len(alone)
    log(f"one at a time: {seq['span_s']:.2f} s, {seq['shots_per_s']:.2f} shots/s, {seq['glups']:.1f} GLUPS")
    gaps = np.array([q.gap for q in q4]) * mean_svc / 1.5
    arrivals = np.cumsum(gaps) - gaps[0]
    fifo_end, t = [], 0.0
    for a, s in zip(arrivals, svc):
        t = max(t, a) + s
        fifo_end.append(t)
    lat_seq = np.array(fifo_end) - arrivals

    def served(arr, worklist=False):
        reqs = mk()
        for r, a in zip(reqs, arr):
            r.arrival = float(a)
        srv = serve.Server(gpu_workers=args.workers, launch=args.launch, worklist=worklist)
        st = srv.serve(reqs)
        same = all(r.error is None and set(r.sections) == set(a.sections) and
                   all(np.array_equal(r.sections[s][c], a.sections[s][c]) for s in a.sections for c in a.sections[s])
                   for r, a in zip(reqs, alone))
        st["peak_hbm_bytes"] = srv.peak_used
        st["lane_requests"] = srv.lane_requests
        return st, same

    st, same = served(arrivals)
    log(f"served (Poisson, 1.5x capacity): {st['span_s']:.2f} s, {st['shots_per_s']:.2f} shots/s, "
        f"p50 {st['p50_s']:.3f} s p99 {st['p99_s']:.3f} s, identical to alone: {same}")
    sat, same2 = served(np.zeros(len(q4)))
    log(f"saturated: {sat['span_s']:.2f} s, {sat['shots_per_s']:.2f} shots/s, {sat['glups']:.1f} GLUPS, "
        f"identical to alone: {same2}")
    wl = wl_sat = None
    same3 = same4 = True
    if args.serve_worklist:
        wl, same3 = served(arrivals, worklist=True)
        log(f"served with work-list lanes: {wl['span_s']:.2f} s, {wl['shots_per_s']:.2f} shots/s, p50 "
            f"{wl['p50_s']:.3f} s p99 {wl['p99_s']:.3f} s, {wl['lane_requests']} requests in lanes, identical "
            f"to alone: {same3}")
        wl_sat, same4 = served(np.zeros(len(q4)), worklist=True)
        log(f"saturated with work-list lanes: {wl_sat['span_s']:.2f} s, {wl_sat['shots_per_s']:.2f} shots/s, "
            f"{wl_sat['glups']:.1f} GLUPS, identical to alone: {same4}")
    shutil.rmtree(td, ignore_errors=True)
    ach = byts / sat["span_s"] / 1e9
    out = {
        "metric": "served shots/s", "value": round(st["shots_per_s"], 4), "unit": "shots/s", "n_gpus": 1,
        "steps": int(sum(q.nt for q in q4)), "warmup": 1, "ms_per_step": None, "higher_is_better": True,
        "scaling": "strong", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic: per-request seeded layered models (workloads.config4_requests, seed 5)",
        "config": {"workload": f"{len(q4)} served requests, sides log-U[256,4096], FD4/8, NT U{{1000..8000}}, "
                               "EL_ISO/VEL_ISO, Poisson arrivals at 1.5x the one-at-a-time capacity",
                   "requests": len(q4), "gpu_workers": args.workers, "cell_steps": cell_steps},
        "p50_latency_s": round(st["p50_s"], 4), "p99_latency_s": round(st["p99_s"], 4),
        "one_at_a_time_shots_per_s": round(seq["shots_per_s"], 4),
        "one_at_a_time_p50_latency_s": round(float(np.percentile(lat_seq, 50)), 4),
        "one_at_a_time_p99_latency_s": round(float(np.percentile(lat_seq, 99)), 4),
        "saturated_shots_per_s": round(sat["shots_per_s"], 4),
        "speedup_vs_one_at_a_time": round(st["shots_per_s"] / seq["shots_per_s"], 3),
        "saturated_speedup_vs_one_at_a_time": round(sat["shots_per_s"] / seq["shots_per_s"], 3),
        "served_glups": round(st["glups"], 3), "saturated_glups": round(sat["glups"], 3),
        "one_at_a_time_glups": round(seq["glups"], 3),
        "one_at_a_time_device_s": round(sum(r.gpu_ms for r in alone) * 1e-3, 3),
        "peak_admitted_hbm_bytes": st["peak_hbm_bytes"],
        "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "whole saturated job (80 B per EL_ISO, 164 B per VEL_ISO cell-step)"},
        "results_identical_alone_and_in_mix": bool(same and same2 and same3 and same4),
    }
    if wl is not None:
        out["worklist_lanes"] = {
            "served_shots_per_s": round(wl["shots_per_s"], 4), "p50_latency_s": round(wl["p50_s"], 4),
            "p99_latency_s": round(wl["p99_s"], 4), "saturated_shots_per_s": round(wl_sat["shots_per_s"], 4),
            "saturated_glups": round(wl_sat["glups"], 3), "requests_in_lanes": wl["lane_requests"],
            "speedup_vs_one_at_a_time": round(wl["shots_per_s"] / seq["shots_per_s"], 3)}
    print(json.dumps(out), flush=True)


def bench_many(args):
    """Shots on grids of uneven size advanced together (sofi_run_many): the
    served request mix (workloads.config4_requests) restricted to grids that
    do not fill the GPU alone (sides log-U[256, 1024], NT U{1000..3000}).
    All handles are opened and loaded first; the timed region is the time
    loops of all requests, run
      one at a time     sofi_run per request, default launch shapes
      concurrently      --workers handles at once, each on its own stream
      work-list         one sofi_run_many per FD order and rheology (the four
                        groups at once), each step's updates of a group in
                        one persistent launch over its (handle, tile) items
    Every work-list result is compared bit for bit with the request's own
    sofi_run in the one-thread-per-cell shape."""
    import numpy as np
    from concurrent.futures import ThreadPoolExecutor
    from gpi_project_amd import workloads
    from gpi_project_amd.sofi import Problem, Simulation, many_config, run_many
    td = tempfile.mkdtemp(prefix="sofi_many_")
    q4 = workloads.config4_requests(args.requests, side=(256, 1024), nt=(1000, 3000))
    if args.many_weq:
        q4 = [q for q in q4 if q.weq == args.many_weq]
    passes = set(args.many_passes.split(","))
    paths = [workloads.config4_write(os.path.join(td, f"r{q.rid:03d}"), q) for q in q4]
    cwd = os.getcwd()
    probs = []
    for pth in paths:
        os.chdir(os.path.dirname(pth))
        probs.append(Problem(os.path.basename(pth)))
    os.chdir(cwd)
    cell_steps = sum(q.nx * q.ny * q.nt for q in q4)
    byts = sum((80 if q.weq == "EL_ISO" else 164) * q.nx * q.ny * q.nt for q in q4)
    groups = {}
    for k, q in enumerate(q4):
        groups.setdefault((q.fdorder, q.weq), []).append(k)
    log(f"{len(q4)} requests, {cell_steps / 1e9:.1f} G cell-steps, groups "
        f"{ {f'FD{f} {w}': len(v) for (f, w), v in groups.items()} }")

    def opened(launch):
        sims = []
        for P in probs:
            s = Simulation(P)
            s.set_launch(launch)
            s.load_shot(nsnap=0)
            sims.append(s)
        return sims

    def sections(sims):
        return [{c: s.section(c) for c in P.seismo_comps()} for s, P in zip(sims, probs)]

    def close(sims):
        for s in sims:
            s.close()

    res = {}
    if "alone" in passes or "concurrent" in passes:
        # warm-up: code objects of both shapes
        for launch in ("auto", "cell"):
            s = Simulation(probs[0])
            s.set_launch(launch)
            s.load_shot(nsnap=0)
            s.run(1, 50)
            s.close()
    if "alone" in passes:
        sims = opened("auto")
        t0 = time.perf_counter()
        for s in sims:
            s.run()
        res["one at a time"] = time.perf_counter() - t0
        close(sims)
    if "concurrent" in passes:
        sims = opened("auto")
        t0 = time.perf_counter()
        with ThreadPoolExecutor(args.workers) as ex:
            list(ex.map(lambda s: s.run(), sims))
        res["concurrently"] = time.perf_counter() - t0
        close(sims)
    sims = opened("cell")
    cfg = {}
    t0 = time.perf_counter()

    def group_run(key):
        idx = groups[key]
        run_many([sims[k] for k in idx])
        cfg[f"FD{key[0]} {key[1]}"] = many_config()

    with ThreadPoolExecutor(args.many_groups or len(groups)) as ex:
        list(ex.map(group_run, list(groups)))
    res["work-list"] = time.perf_counter() - t0
    got = sections(sims)
    close(sims)
    same = None
    if "alone" in passes:
        sims = opened("cell")
        for s in sims:
            s.run()
        ref = sections(sims)
        close(sims)
        same = all(np.array_equal(a[c], b[c]) for a, b in zip(got, ref) for c in b)
    for P in probs:
        P.close()
    shutil.rmtree(td, ignore_errors=True)
    rate = {k: len(q4) / v for k, v in res.items()}
    for k, v in res.items():
        log(f"{k}: {v:.3f} s, {rate[k]:.2f} shots/s, {cell_steps / v / 1e9:.1f} GLUPS")
    log(f"work-list launches {cfg}; identical to one-shot cell runs: {same}")
    ach = byts / res["work-list"] / 1e9
    out = {
        "metric": "shots/s of shots on grids of uneven size", "value": round(rate["work-list"], 4),
        "unit": "shots/s", "n_gpus": 1, "steps": int(sum(q.nt for q in q4)), "warmup": 1, "ms_per_step": None,
        "higher_is_better": True, "scaling": "strong", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic: per-request seeded layered models (workloads.config4_requests, seed 5)",
        "config": {"workload": f"{len(q4)} requests, sides log-U[256,1024], FD4/8, NT U{{1000..3000}}, "
                               f"{args.many_weq or 'EL_ISO/VEL_ISO'}, all waiting; one sofi_run_many per FD "
                               "order and rheology",
                   "requests": len(q4), "cell_steps": cell_steps, "workers": args.workers},
        "one_at_a_time_shots_per_s": round(rate["one at a time"], 4) if "one at a time" in rate else None,
        "concurrent_shots_per_s": round(rate["concurrently"], 4) if "concurrently" in rate else None,
        "speedup_vs_one_at_a_time":
            round(rate["work-list"] / rate["one at a time"], 3) if "one at a time" in rate else None,
        "speedup_vs_concurrent": round(rate["work-list"] / rate["concurrently"], 3) if "concurrently" in rate else None,
        "glups": round(cell_steps / res["work-list"] / 1e9, 3),
        "one_at_a_time_glups": round(cell_steps / res["one at a time"] / 1e9, 3) if "one at a time" in rate else None,
        "launch": {k: {"rows": v[0], "blocks": v[1], "waves": v[2], "tuned": v[3]} for k, v in cfg.items()},
        "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "whole work-list job (80 B per EL_ISO, 164 B per VEL_ISO cell-step)"},
        "results_identical_to_one_shot_cell_runs": same,
    }
    print(json.dumps(out), flush=True)


def bench_small(args):
    """The reference's small CPU-runnable case (workloads.config1: EL_ISO
    400 x 400 FD4, NT = 2000, 100 receivers): a launch-bound loop whose
    working set (~7 MB) stays in the caches.  The whole loop step by step and
    from captured blocks of 64 steps; prints one JSON line (GLUPS of the
    captured run)."""
    import numpy as np
    from gpi_project_amd import workloads
    from gpi_project_amd.sofi import Problem, Simulation
    td = tempfile.mkdtemp(prefix="sofi_bench_")
    js = workloads.config1(td)
    cwd = os.getcwd()
    os.chdir(td)
    P = Problem(os.path.basename(js))
    d = P.desc
    res = {}
    for mode in ("off", "on"):
        sim = Simulation(P)
        sim.set_graph(mode, 64)
        sim.load_shot(nsnap=0)
        sim.run(1, min(args.warmup, d.nt))
        best = None
        for _ in range(3):
            sim.load_shot(nsnap=0)
            ms = sim.run()
            best = ms if best is None else min(best, ms)
        res[mode] = (best, sim.section(1), sim.graph_blocks)
        sim.close()
    P.close()
    os.chdir(cwd)
    shutil.rmtree(td, ignore_errors=True)
    cells = d.nx * d.ny
    glups = {m: cells * d.nt / (res[m][0] * 1e-3) / 1e9 for m in res}
    same = bool(np.array_equal(res["on"][1], res["off"][1]))
    log(f"step by step {res['off'][0]:.2f} ms ({glups['off']:.2f} GLUPS), captured blocks {res['on'][0]:.2f} ms "
        f"({glups['on']:.2f} GLUPS, {res['on'][2]} blocks), identical: {same}")
    out = {
        "metric": "GLUPS", "value": round(glups["on"], 3), "unit": "GLUPS", "n_gpus": 1, "steps": d.nt,
        "warmup": args.warmup, "ms_per_step": round(res["on"][0] / d.nt, 5), "higher_is_better": True,
        "scaling": "strong", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic: homogeneous model (workloads.config1)",
        "config": {"workload": "elastic FD4 400x400 NT=2000, CPML, 100 receivers", "nx": d.nx, "ny": d.ny,
                   "nt": d.nt, "fdorder": d.fdorder, "weq": "EL_ISO"},
        "step_by_step_glups": round(glups["off"], 3), "step_by_step_ms_per_step": round(res["off"][0] / d.nt, 5),
        "graph_speedup": round(glups["on"] / glups["off"], 3), "results_identical": same,
    }
    print(json.dumps(out), flush=True)


def bench_fwi(args):
    """Elastic FWI gradient (workloads.config5: 1500 x 500 FD8, NT = 3000,
    32 shots, 300 receivers, DTINV = 5): the measured data come from forward
    runs on the true model; then one gradient iteration on the background
    model, per shot a forward pass storing every 5th step, L2 residuals
    read from the data files, an adjoint pass correlating with the store and
    the gradient.  Shot 1 is repeated with DTINV = 1 (the reference's full
    rate; a 45 GB store) for the direction of the DTINV = 5 gradient.
    Prints one JSON line (gradient shots/s)."""
    import numpy as np
    from gpi_project_amd import fwi, workloads
    from gpi_project_amd.sofi import Problem, Simulation
    td = tempfile.mkdtemp(prefix="sofi_fwi_")
    cwd = os.getcwd()
    os.chdir(td)
    js = workloads.config5(td, "true", nshots=args.fwi_shots)
    P = Problem(os.path.basename(js))
    d = P.desc
    sim = Simulation(P)
    t0 = time.perf_counter()
    for ishot in range(1, P.nshots + 1):
        P.set_shot(ishot)
        sim.load_shot(nsnap=0)
        sim.run()
        for c in P.seismo_comps():
            P.write_section(c, sim.section(c))
    sim.close()
    P.close()
    log(f"measured data: {d.nx}x{d.ny} NT={d.nt}, {args.fwi_shots} shots in {time.perf_counter() - t0:.1f} s")
    js = workloads.config5(td, "init", nshots=args.fwi_shots)
    P = Problem(os.path.basename(js))
    d = P.desc
    sim = Simulation(P)
    sim.fwi_set_model(P.model(0), P.model(1), P.model(2))
    fwi.shot_gradient(P, sim, 1, write=False)              # warm-up (code objects, store pages)
    fp = sim.footprint
    sim.close()
    # one shot after the other on one handle, then args.fwi_workers shots at once
    t0 = time.perf_counter()
    rs = fwi.gradient_json(os.path.basename(js), verbose=False, write=False, workers=1, sections=False)
    t_seq = time.perf_counter() - t0
    t0 = time.perf_counter()
    rc = fwi.gradient_json(os.path.basename(js), verbose=False, write=False, workers=args.fwi_workers,
                           sections=False)
    t_it = time.perf_counter() - t0
    same = all(np.array_equal(a, b) for a, b in zip(rs["sum"], rc["sum"]))
    # groups of args.fwi_batch shots as the slots of batch handles sharing the medium
    t_b, cos_b, dev_ms_b = None, None, None
    if args.fwi_batch > 1:
        t0 = time.perf_counter()
        rb = fwi.gradient_json(os.path.basename(js), verbose=False, write=False, batch=args.fwi_batch,
                               workers=args.fwi_batch_workers, sections=False)
        t_b = time.perf_counter() - t0
        cos_b = min(float(np.dot(a.ravel().astype(np.float64), b.ravel()) /
                          (np.linalg.norm(a.astype(np.float64)) * np.linalg.norm(b.astype(np.float64)) + 1e-300))
                    for a, b in zip(rb["sum"], rs["sum"]))
        dev_ms_b = sum(r["ms_forward"] + r["ms_adjoint"] for r in rb["shots"].values())
        log(f"batches of {args.fwi_batch} shots on {args.fwi_batch_workers} handle(s): {t_b:.2f} s "
            f"({P.nshots / t_b:.2f} shots/s, device {dev_ms_b / 1e3:.2f} s), sums vs one handle: cosine {cos_b:.8f}")
    dev_ms = sum(r["ms_forward"] + r["ms_adjoint"] for r in rs["shots"].values())
    l2 = sum(r["l2"] for r in rs["shots"].values())
    gsum = rc["sum"]
    g5 = rs["shots"][1]["gradient"]
    finite = all(np.isfinite(g).all() for g in gsum)
    log(f"one handle: {P.nshots} shots in {t_seq:.2f} s ({P.nshots / t_seq:.2f} shots/s, device {dev_ms / 1e3:.2f} s); "
        f"{args.fwi_workers} handles at once: {t_it:.2f} s ({P.nshots / t_it:.2f} shots/s), sums identical: {same}")
    # shot 1 at the full rate
    d1 = d.copy()
    d1.dtinv = 1
    sim1 = Simulation(P, desc=d1)
    sim1.fwi_set_model(P.model(0), P.model(1), P.model(2))
    r1 = fwi.shot_gradient(P, sim1, 1, write=False)
    sim1.close()
    cos = [float(np.dot(a.ravel().astype(np.float64), b.ravel()) /
                 (np.linalg.norm(a.astype(np.float64)) * np.linalg.norm(b.astype(np.float64)) + 1e-300))
           for a, b in zip(g5, r1["gradient"])]
    log(f"L2 {l2:.4e}, finite {finite}; DTINV=5 vs 1 cosine vp/vs/rho {cos}, DTINV=1 shot "
        f"{(r1['ms_forward'] + r1['ms_adjoint']):.0f} ms")
    cells = d.nx * d.ny
    # CPU: the FP32 oracle's forward (storing) and adjoint (correlating) steps on the same shot, 40 each
    cpu = None
    if args.cpu_steps > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        n = 40
        ds = d.copy()
        ds.nt = n
        P.set_shot(1)
        o = pyoracle.Oracle(P, "f32", desc=ds)
        o.fwi_set_model(P.model(0), P.model(1), P.model(2))
        ri, rj = P.receivers()
        o.fwi_set_adjoint(ri, rj, np.zeros((len(ri), n), np.float32), np.zeros((len(ri), n), np.float32))
        tc = time.perf_counter()
        o.reset(0)
        o.run()
        o.reset(1)
        o.run()
        tc = time.perf_counter() - tc
        o.close()
        cpu = {"value": round(1.0 / (tc / n * d.nt), 6), "unit": "gradient shots/s", "cores": 1, "kind": "port",
               "sample": f"{n} forward (storing) + {n} adjoint (correlating) steps of shot 1 on the FP32 CPU oracle, "
                         f"scaled to NT = {d.nt}"}
    P.close()
    os.chdir(cwd)
    shutil.rmtree(td, ignore_errors=True)
    # per cell and step: forward 80 B + store 20 B every DTINV steps; adjoint 80 B + read 20 B and
    # gradient accumulators (6 floats read + written, the correlation inputs) every DTINV steps
    byts = cells * d.nt * (2 * 80 + (20 + 20 + 48) / d.dtinv) * P.nshots
    # the value: the faster of the concurrent one-shot handles and the batches
    use_b = t_b is not None and t_b < t_it
    t_val = t_b if use_b else t_it
    mode = (f"batches of {args.fwi_batch} shots on {args.fwi_batch_workers} handle(s)" if use_b else
            f"{args.fwi_workers} one-shot handles at once")
    # device-time roofline: the batched run's device time when batches ran on one handle, else one handle's
    dev_roof = dev_ms_b if (t_b is not None and args.fwi_batch_workers == 1) else dev_ms
    out = {
        "metric": "FWI gradient shots/s", "value": round(P.nshots / t_val, 4), "unit": "shots/s", "n_gpus": 1,
        "steps": 2 * d.nt * P.nshots, "warmup": 1, "ms_per_step": round(t_val * 1e3 / (2 * d.nt * P.nshots), 5),
        "higher_is_better": True, "scaling": "strong", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic: linear-gradient background with a seeded Gaussian anomaly (workloads.config5)",
        "config": {"workload": f"elastic FWI gradient, {d.nx}x{d.ny} FD8, NT={d.nt}, {P.nshots} shots, "
                               f"300 receivers, DTINV={d.dtinv}", "nx": d.nx, "ny": d.ny, "nt": d.nt,
                   "shots": P.nshots, "dtinv": d.dtinv},
        "mode": mode, "device_s": round(dev_ms / 1e3, 3), "iteration_s": round(t_val, 3),
        "workers": args.fwi_workers, "concurrent_handles_shots_per_s": round(P.nshots / t_it, 4),
        "batched_shots_per_s": round(P.nshots / t_b, 4) if t_b else None,
        "batched_device_s": round(dev_ms_b / 1e3, 3) if dev_ms_b else None,
        "batched_sums_cosine_vs_one_handle": round(cos_b, 8) if cos_b is not None else None,
        "one_handle_shots_per_s": round(P.nshots / t_seq, 4), "sums_identical_to_one_handle": bool(same),
        "glups_equivalent": round(2 * cells * d.nt * P.nshots / t_val / 1e9, 3),
        "footprint_bytes": fp, "cosine_dtinv5_vs_dtinv1": [round(c, 6) for c in cos],
        "dtinv1_shot_ms": round(r1["ms_forward"] + r1["ms_adjoint"], 1), "gradient_finite": finite,
        "roofline": {"bound": "hbm", "achieved": round(byts / (dev_roof * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(byts / (dev_roof * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": None,
                     "kernel": ("forward + adjoint passes (all kernels, device time of the batched run)"
                                if dev_roof is dev_ms_b else "forward + adjoint passes (all kernels, device time)")},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8000, help="timed time steps (default: the whole NT = 8000 loop)")
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--nx", type=int, default=4000)
    ap.add_argument("--ny", type=int, default=2000)
    ap.add_argument("--nt", type=int, default=8000)
    ap.add_argument("--time-every", type=int, default=8, help="time every n-th step's kernels with HIP events")
    ap.add_argument("--cpu-steps", type=int, default=160, help="time steps of the CPU baseline sample (0: skip)")
    ap.add_argument("--cpu-cores", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--launch", default="auto",
                    help="shape of the update kernels: auto, cell, vec2, vec4, or VEL:STRESS for each "
                         "(default: chosen by grid size)")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 counter passes")
    ap.add_argument("--workload", default="c2",
                    choices=("c1", "c2", "c3-vti", "c3-tti", "c3-iso", "c4", "c4-many", "c5"),
                    help="c2: the headline single-shot run (default); c1: the small elastic case; c3-*: 64 "
                         "batched anisotropic shots; c4: served requests of mixed sizes; c4-many: small requests "
                         "of uneven size in work-list launches; c5: an elastic FWI gradient")
    ap.add_argument("--shots", type=int, default=64, help="shots of the batched workloads")
    ap.add_argument("--single", type=int, default=4, help="shots timed one at a time for the comparison")
    ap.add_argument("--grid", default="2000x1000", help="NXxNY of the batched workloads")
    ap.add_argument("--requests", type=int, default=48, help="requests of the served workload")
    ap.add_argument("--resident", type=int, default=0, help="slots per resident group of a batch (0: automatic)")
    ap.add_argument("--streams", type=int, default=0, help="streams of a batch's resident groups (0: automatic)")
    ap.add_argument("--block", type=int, default=-1,
                    help="a batch's slot blocks: 0 slot-loop kernels, 1/2/4 slots per thread (default: autotuned)")
    ap.add_argument("--fwi-shots", type=int, default=32, help="shots of the FWI workload")
    ap.add_argument("--fwi-workers", type=int, default=8, help="FWI shots computed at once (one handle each)")
    ap.add_argument("--fwi-batch", type=int, default=8, help="FWI shots per batch handle (slots sharing the medium)")
    ap.add_argument("--fwi-batch-workers", type=int, default=1, help="FWI batch handles at once")
    ap.add_argument("--workers", type=int, default=4, help="concurrent handles of the served workload")
    ap.add_argument("--serve-worklist", action="store_true",
                    help="c4: also serve with work-list lanes for the small one-thread-per-cell requests")
    ap.add_argument("--many-weq", default="", help="c4-many: only the requests of this rheology")
    ap.add_argument("--many-passes", default="alone,concurrent,worklist",
                    help="c4-many: the passes to run (the work-list pass always runs)")
    ap.add_argument("--many-groups", type=int, default=0,
                    help="c4-many: work-list groups (FD order, rheology) run at once (0: all)")
    ap.add_argument("--pmc-child", metavar="CASE_DIR", help=argparse.SUPPRESS)
    ap.add_argument("--pmc-steps", type=int, default=20, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.pmc_child:
        pmc_child(args.pmc_child, args.launch, args.pmc_steps)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported: shots are independent and would "
              "shard across GPUs with no communication during a shot", file=sys.stderr)
        sys.exit(2)
    if args.workload.startswith("c3"):
        if "--nt" not in sys.argv:
            args.nt = 4000
        bench_batch(args, {"c3-vti": "EL_VTI", "c3-tti": "EL_TTI", "c3-iso": "EL_ISO"}[args.workload])
        return
    if args.workload == "c4":
        bench_serve(args)
        return
    if args.workload == "c4-many":
        bench_many(args)
        return
    if args.workload == "c1":
        bench_small(args)
        return
    if args.workload == "c5":
        bench_fwi(args)
        return
    if args.steps > args.nt:
        args.nt = args.steps

    from gpi_project_amd import workloads
    from gpi_project_amd.sofi import Problem, Simulation

    td = tempfile.mkdtemp(prefix="sofi_bench_")
    t0 = time.time()
    js = workloads.config2(td, nx=args.nx, ny=args.ny, nt=args.nt)
    log(f"workload written in {time.time() - t0:.1f} s: {js}")
    cwd = os.getcwd()
    os.chdir(td)
    t0 = time.time()
    P = Problem(os.path.basename(js))
    d = P.desc
    log(f"host setup {time.time() - t0:.1f} s: {d.nx}x{d.ny}, NT={d.nt}, FD{d.fdorder}, L={d.L}, {P.receivers()[0].size} receivers")
    sim = Simulation(P)
    sim.set_launch(*args.launch.split(":"))
    log(f"device footprint {sim.footprint / 2**30:.2f} GiB, launch shape {args.launch}")
    # warm-up on a shot that is then discarded
    sim.load_shot(nsnap=0)
    if args.warmup > 0:
        sim.run(1, min(args.warmup, d.nt))
    sim.load_shot(nsnap=0)
    nmax = args.steps // max(args.time_every, 1) + 1
    sim.set_kernel_timing(args.time_every, nmax)
    log(f"timed region: {args.steps} steps")
    t_wall0 = time.perf_counter()
    ms = sim.run(1, args.steps)          # returns after the handle's stream has drained
    t_wall = time.perf_counter() - t_wall0
    kt = sim.kernel_timing()
    log(f"time loop {ms:.1f} ms (wall {t_wall * 1e3:.1f} ms); kernels {kt}")
    cells = d.nx * d.ny
    glups = cells * args.steps / (ms * 1e-3) / 1e9
    b_vel, b_str = algorithmic_bytes(d.L)
    achieved = b_str * cells / (kt["stress_ms"] * 1e-3) / 1e9 if kt["stress_ms"] > 0 else None
    step_gbs = (b_vel + b_str) * cells * args.steps / (ms * 1e-3) / 1e9
    vel_gbs = b_vel * cells / (kt["velocity_ms"] * 1e-3) / 1e9 if kt["velocity_ms"] > 0 else None
    # a short check that the run stayed finite
    import numpy as np
    secs = sim.section(0)
    assert np.isfinite(secs).all()
    sim.close()
    P.close()
    os.chdir(cwd)
    traffic = None
    if not args.no_traffic:
        traffic, info, _ = pmc_traffic(td, args.launch)
        log(f"traffic per step {traffic}, counter factors {info}")
    cpu = None
    if args.cpu_steps > 0:
        cpu = cpu_baseline(td, os.path.basename(js), d.nx, d.ny, args.cpu_steps, args.cpu_cores)
    shutil.rmtree(td, ignore_errors=True)
    out = {
        "metric": METRIC,
        "value": round(glups, 3),
        "unit": "GLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms / args.steps, 5),
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded layered model with a 5 % Gaussian perturbation (workloads.config2)",
        "config": {"workload": "viscoelastic L=3 FD8 4000x2000 NT=8000, free surface + CPML, 1000 receivers",
                   "nx": d.nx, "ny": d.ny, "nt": d.nt, "fdorder": d.fdorder, "L": d.L, "weq": "VEL_ISO",
                   "timed_steps": args.steps},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1) if achieved else None, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
                     "traffic": round(traffic["stress"]) if traffic and "stress" in traffic else None,
                     "kernel": "stress update (k_stress VEL_ISO)",
                     "algorithmic_bytes_per_cell": b_str, "kernel_ms": round(kt["stress_ms"], 5),
                     "velocity_kernel_gbs": round(vel_gbs, 1) if vel_gbs else None,
                     "traffic_velocity": round(traffic["velocity"]) if traffic and "velocity" in traffic else None,
                     "algorithmic_bytes_per_launch": b_str * cells,
                     "whole_step_gbs": round(step_gbs, 1), "whole_step_frac": round(step_gbs / HBM_PEAK_GBS, 4)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
