#!/usr/bin/env python3
"""Throughput of the 2D DPD step on one MI355X.

Workload: the plain method at scale -- N = 1,048,576 beads at rho = 3
(L = 591.2067), 50% W, 20% precursor P, 30% H-T-T lipid beads (104,858
lipids), precursor chemistry on every step (P -> T next to a lipid tail
p = 0.005, T-T bond form p = 0.01 / break p = 0.001), dt = 0.01, harmonic
bonds, no warm-up ramp; synthetic random soup, all state resident in HBM.

One step = integrate + cell sort + chemistry + pair forces + bonded forces
for every bead.  W untimed steps, then K timed steps replayed from a hipGraph,
with a device synchronisation on both sides of the timed region.  The pair
kernel is then timed alone with HIP events on the library's stream for the
roofline; the reference engine (compiled from its source) is timed on the host
cores on a bounded sample of the same physics as the CPU baseline.

Other workloads (--workload): ensemble (256 replicas in one engine against a
sequential loop), adjoint (window gradient against plain steps), fit (batched
parameter fit), refchem (the reference's own default parameters with its
chemistry -- catalysts, decay, capped feed -- at N = 1M, against the
reference CPU engine on the same parameters).
"""
import argparse
import ctypes as C
import importlib
import json
import math
import os
import re
import shutil
import sys
import tempfile
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "2D DPD particle-steps/s, N=1M ρ=3; batched-ensemble & adjoint vs plain"
HBM_PEAK_GBS = 8000.0
# the pair arithmetic is FP32; positions, velocities, force sums and reductions are FP64
DTYPE = "fp32 pair arithmetic, fp64 state"
TRAFFIC_FILE = os.path.join(REPO, "profiles", "k_pair_N1M_traffic.json")


def measured_traffic(N):
    """HBM bytes per k_pair launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE measurement of this workload (calibrated against a copy), or
    None when there is none for this N."""
    try:
        with open(TRAFFIC_FILE) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    return d.get("traffic_bytes_per_launch") if d.get("N") == N and d.get("kernel") == "k_pair" else None


def copy_probe():
    """Calibration kernel for the HBM counters: five device-to-device copies of
    1 GiB (torch copy_, one __amd_rocclr_copyBuffer dispatch each)."""
    import torch
    n = 256 * 1024 * 1024
    a = torch.empty(n, dtype=torch.float32, device="cuda")
    a.fill_(1.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_per_dispatch": n * 4}), flush=True)


def config2_params(pm, N, nsub, chem=True):
    p = pm.Params()
    p.update({
        "L": repr(pm.L_for_N(N)), "rho": "3", "dt": "0.01", "lambda": "0.65", "gamma": "4.5", "kT": "1",
        "k_bond": "100", "r0": "0.5", "k_angle": "0", "lip_nH": "1", "lip_nT": "2", "f_lip": "0.3", "c_N": "0.2",
        "f_Ka": "0", "f_Kb": "0", "a_W_W": "25", "a_H_H": "25", "a_T_T": "25", "a_W_H": "25", "a_W_T": "80",
        "a_H_T": "80", "a_N_W": "25", "a_N_H": "25", "a_N_N": "25", "a_N_T": "50", "warmup": "0", "chem": "0",
        "pchem": "1" if chem else "0", "n_chem": "1", "p_conv": "0.005", "r_conv": "0.7", "p_form": "0.01",
        "r_form": "0.6", "p_break": "0.001", "nsub": str(nsub), "seed": "1",
    })
    return p.finalize()


class HipEvents:
    """hipEvent timing through the HIP runtime (the library's stream)."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so")
        self.hip.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [C.c_void_p]
        self.hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.hip.hipEventDestroy.argtypes = [C.c_void_p]

    def new(self):
        e = C.c_void_p()
        rc = self.hip.hipEventCreate(C.byref(e))
        if rc:
            raise RuntimeError(f"hipEventCreate failed ({rc})")
        return e

    def ms(self, a, b):
        self.hip.hipEventSynchronize(b)
        t = C.c_float()
        self.hip.hipEventElapsedTime(C.byref(t), a, b)
        return t.value


def time_pair_kernel(e, reps=200):
    """Average duration (s) of one pair-kernel launch, from HIP events recorded
    on the stream the kernel runs on around one replay of reps back-to-back
    launches (a graph, so no host gaps between them)."""
    ev = HipEvents()
    a, b = ev.new(), ev.new()
    stream = C.c_void_p(e.stream)
    e.lib.oc_launch_pair(e.h, reps)  # builds the graph of reps launches, untimed
    ev.hip.hipEventRecord(a, stream)
    rc = e.lib.oc_launch_pair(e.h, reps)
    ev.hip.hipEventRecord(b, stream)
    if rc:
        raise RuntimeError(e.lib.oc_last_error().decode())
    ms = ev.ms(a, b)
    ev.hip.hipEventDestroy(a)
    ev.hip.hipEventDestroy(b)
    return ms * 1e-3 / reps


def host_info():
    """Host CPU of this run: logical CPUs, the ones this process may use, the
    model name and the OpenMP thread limit set in the environment."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS", "")
    return {"nproc": os.cpu_count() or 1, "usable_cores": usable, "cpu_model": model,
            "omp_num_threads_env": int(omp) if omp.isdigit() else None}


def cpu_threads(L, host):
    """OpenMP threads of the reference CPU run: the reference's own rule
    T = min(nproc, floor(L)/2) (at most floor(L)/2 column strips run at once,
    cel/engine/oercel.c:750-756), with nproc the cores this process may use,
    further capped by OMP_NUM_THREADS when the environment sets it."""
    t = min(host["usable_cores"], int(math.floor(L)) // 2)
    if host["omp_num_threads_env"]:
        t = min(t, host["omp_num_threads_env"])
    return max(1, t)


def cpu_baseline(pm, N, nsub, target_s=12.0):
    """The reference engine (oracle/_ref, built from cel/engine/oercel.c with its
    own -O3 flags) on the host cores, same physics with P as its type N and
    chemistry off (it has no precursor chemistry).  Falls back to the oracle
    restatement when the reference binaries were not built."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle
    p = config2_params(pm, N, nsub, chem=False)
    host = host_info()
    threads = cpu_threads(p.L, host)
    binp = oracle.ref_binary(fast=True)
    if os.path.exists(binp):
        ov = {"L": repr(p.L), "dt": "0.01", "chem": "0", "f_lip": "0.3", "lip_nH": "1", "lip_nT": "2",
              "f_Ka": "0", "f_Kb": "0", "c_N": "0.2", "a_W_H": "25", "a_H_T": "80", "a_W_T": "80", "a_N_T": "50",
              "a_N_W": "25", "a_N_H": "25", "k_angle": "0", "warmup": "0", "snap_every": "0", "stats_every": "0",
              "ckpt_every": "0", "log_every": "0", "events": "0", "threads": str(threads), "nsub": str(nsub),
              "seed": "1"}
        tmp = tempfile.mkdtemp(prefix="ocref_")
        try:
            def run(steps):
                out = oracle.run_ref_cli(os.path.join(tmp, f"r{steps}"), dict(ov, steps=str(steps)), fast=True)
                m = re.search(r"timing: (\d+) steps, ([0-9.]+) ms/step", out)
                return float(m.group(2))
            ms = run(10)
            steps = int(max(20, min(3000, target_s / (ms * 1e-3))))
            ms = run(steps)
        finally:
            shutil.rmtree(tmp, ignore_errors=True)
        return {"value": N / (ms * 1e-3), "unit": "particle-steps/s", "cores": threads, "kind": "reference",
                "sample": f"reference CLI {os.path.basename(binp)} (cel/engine/oercel.c, -O3), N={N}, same physics "
                          f"with P as type N, chemistry off, {steps} steps after setup, {threads} OpenMP threads "
                          f"(min of usable cores, floor(L)/2 strips and OMP_NUM_THREADS), {ms:.3f} ms/step",
                "host": host}
    # restatement fallback
    cp = p.to_c()
    sysm = oracle.System(cp, N)
    sysm.start(oracle.soup(cp, N))
    t0 = time.perf_counter()
    sysm.step(5)
    dt = (time.perf_counter() - t0) / 5
    return {"value": N / dt, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"oracle restatement, N={N}, chemistry off, 5 steps", "host": host}


def ensemble_params(pm, nrep, seed=0):
    """Batched-ensemble workload: N log-uniform in [2048, 65536] at rho = 3
    (L = sqrt(N/3)), lipid fraction U[0.05, 0.4], a_WT U[60, 100], distinct
    seeds, the physics and chemistry of the plain workload."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for r in range(nrep):
        N = int(round(math.exp(rng.uniform(math.log(2048), math.log(65536)))))
        p = config2_params(pm, N, 4)
        p.update({"f_lip": repr(float(rng.uniform(0.05, 0.4))), "a_W_T": repr(float(rng.uniform(60, 100))),
                  "seed": str(1000 + r)})
        out.append(p.finalize())
    return out


def ensemble_measure(pm, eng, replicas, steps, equil, tune=True):
    """Throughput of R replicas stepped together by one engine (one launch per
    phase for all replicas, hipGraph replay) against the same replicas stepped
    one after another by single-replica engines (also graph replay), both
    timed over the same number of steps from the same equilibrated states:
    the batched engine runs `equil` untimed steps from the soups, and every
    single engine is loaded with its replica's state in slot order."""
    ps = ensemble_params(pm, replicas)
    t_tune = 0.0
    if tune:
        at = importlib.import_module("welcome-youtube_amd.autotune")
        t0 = time.perf_counter()
        ps, blk, tb = at.tune_ensemble(ps, verbose=True)
        t_tune = time.perf_counter() - t0
    else:
        blk, tb = 0, 0
    Ntot = sum(p.N for p in ps)
    e = eng.Engine(ps, pair_block=blk, tile_beads=tb)
    e.start()
    e.step(equil)
    states = [(e.get_state(r), e.cells(r)[1], e.step_count(r)) for r in range(len(ps))]
    e.step(2)  # graph warm
    e.step_count()
    t0 = time.perf_counter()
    e.step(steps)
    e.step_count()
    batched = time.perf_counter() - t0
    T_end = [e.stats(r)["T_kin"] for r in range(len(ps))]
    e.close()
    singles = []
    for p, (s, order, st) in zip(ps, states):
        q = eng.Engine(p)
        q.set_state(s, step=st, slot_gid=order)
        q.step(2)  # graph warm
        singles.append(q)
    for q in singles:
        q.step_count()
    t0 = time.perf_counter()
    for q in singles:
        q.step(steps)
    for q in singles:
        q.step_count()
    seq = time.perf_counter() - t0
    for q in singles:
        q.close()
    return {
        "value": Ntot * steps / batched, "unit": "particle-steps/s", "steps": steps, "equil": equil,
        "ms_per_step": 1e3 * batched / steps,
        "sequential_value": Ntot * steps / seq, "sequential_ms_per_step": 1e3 * seq / steps,
        "speedup_vs_sequential": seq / batched,
        "workload": f"batched ensemble, {replicas} replicas, N log-uniform [2048, 65536], rho=3, f_lip U[0.05,0.4], "
                    "a_WT U[60,100], distinct seeds, precursor chemistry on, hipGraph replay",
        "N_total": Ntot, "N_min": min(p.N for p in ps), "N_max": max(p.N for p in ps),
        "autotuned": bool(tune), "tune_s": t_tune, "pair_block": blk, "tile_beads": tb,
        "nsub": sorted({p.nsub for p in ps}),
        "T_kin_end_mean": sum(T_end) / len(T_end),
    }


def bench_ensemble(a, pm, eng):
    d = ensemble_measure(pm, eng, a.replicas, a.steps, a.equil, tune=not a.no_tune)
    out = {
        "metric": "batched-ensemble particle-steps/s (vs a sequential loop of single-replica engines)",
        "value": d["value"], "unit": "particle-steps/s", "n_gpus": 1, "steps": a.steps, "warmup": a.equil,
        "ms_per_step": d["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": DTYPE, "data": "synthetic", "config": {"workload": d["workload"]}, "ensemble": d,
    }
    print(json.dumps(out), flush=True)


def adjoint_params(pm, N=65536, seed=1, **over):
    """Adjoint window workload: N = 65,536 at rho = 3 (L = 147.8017), 30% H-T-T
    lipid beads, 70% W, harmonic bonds only, chemistry off, dt = 0.01."""
    p = config2_params(pm, N, 4, chem=False)
    p.update({"c_N": "0", "seed": str(seed)})
    p.update({k: str(v) for k, v in over.items()})
    return p.finalize()


def adjoint_measure(pm, eng, nwin, equil):
    """Cost of a window gradient (forward with checkpoints every 20 steps +
    the adjoint sweep with recomputation) against `nwin` plain graph-replayed
    steps of the same system from the same state, N = 65,536."""
    abi = importlib.import_module("welcome-youtube_amd.abi")
    spec = abi.rdf_spec(2, 2, 64, 3.0)
    # target: the tail-tail RDF of a seed run at a_TT = 20, a_WT = 90
    pt = adjoint_params(pm, seed=2, a_T_T=20, a_W_T=90)
    with eng.Engine(pt) as et:
        et.start()
        et.step(equil)
        target = et.window_forward(nwin, spec)[0]
    p = adjoint_params(pm)
    e = eng.Engine(p)
    e.start()
    e.step(equil)
    st0 = e.get_state()
    s0 = e.step_count()
    order = e.cells()[1]
    e.step(nwin)  # warm the graph
    e.set_state(st0, step=s0, slot_gid=order)
    e.step_count()
    t0 = time.perf_counter()
    e.step(nwin)
    e.step_count()
    t_fwd = time.perf_counter() - t0
    e.set_state(st0, step=s0, slot_gid=order)
    e.window_forward(nwin, spec)  # warm buffers
    e.set_state(st0, step=s0, slot_gid=order)
    t0 = time.perf_counter()
    gbar = e.window_forward(nwin, spec)
    t1 = time.perf_counter()
    dl = 2.0 * (gbar - target[None, :])
    grad = e.window_backward(dl)[0]
    t2 = time.perf_counter()
    loss = float(((gbar[0] - target) ** 2).sum())
    e.close()
    return {
        "value": (t2 - t0) / t_fwd, "unit": "x forward", "higher_is_better": False, "window": nwin, "equil": equil,
        "workload": "adjoint window: N=65,536, rho=3, 30% H-T-T lipids, k_angle=0, chemistry off, "
                    f"{nwin}-step window, checkpoint every 20, tail-tail RDF 64 bins on [0,3), target from a seed "
                    "run at a_TT=20, a_WT=90",
        "N": p.N, "forward_ms": 1e3 * t_fwd, "window_forward_ms": 1e3 * (t1 - t0),
        "window_backward_ms": 1e3 * (t2 - t1), "loss": loss,
        "grad": {n: float(g) for n, g in zip(eng.Engine.THETA_NAMES, grad)},
    }


def bench_adjoint(a, pm, eng):
    d = adjoint_measure(pm, eng, a.window, a.equil)
    out = {
        "metric": "adjoint window: gradient cost vs plain forward steps",
        "value": d["value"], "unit": "x forward", "n_gpus": 1, "steps": a.window, "warmup": a.equil,
        "ms_per_step": (d["window_forward_ms"] + d["window_backward_ms"]) / a.window, "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": DTYPE, "data": "synthetic",
        "config": {"workload": d["workload"], "N": d["N"]}, "adjoint": d,
    }
    print(json.dumps(out), flush=True)


def bench_fit(a, pm, eng):
    """Batched parameter fit: R replicas of N beads (config-4 physics, distinct
    seeds) run adjoint windows together; the replica-mean gradient of the RDF
    loss drives Adam on the six a_ij toward the RDF of a seed run at a_TT = 20,
    a_WT = 90.  value = replica particle-steps per second through the window
    gradient (forward window + adjoint sweep) of one Adam step."""
    import numpy as np
    abi = importlib.import_module("welcome-youtube_amd.abi")
    fit = importlib.import_module("welcome-youtube_amd.fit")
    spec = abi.rdf_spec(2, 2, 64, 3.0)
    R, N, nwin = a.replicas, a.N, a.window

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    pt = adjoint_params(pm, N=N, seed=2, a_T_T=20, a_W_T=90)
    with eng.Engine(pt) as et:
        et.start()
        et.step(a.equil)
        target = et.window_forward(nwin, spec)[0]
    log(f"target RDF ready (N={N}, a_TT=20, a_WT=90)")
    ps = [adjoint_params(pm, N=N, seed=1000 + r) for r in range(R)]
    e = eng.Engine(ps)
    e.start()
    e.step(a.equil)
    e.step_count()
    t0 = time.perf_counter()
    e.step(nwin)
    e.step_count()
    t_plain = time.perf_counter() - t0
    log(f"{R} x {N} replicas equilibrated; {nwin} plain steps {1e3 * t_plain:.1f} ms")

    def cb(rec):
        th = rec["theta"]
        log(f"adam {rec['step']:3d}  loss {rec['loss']:.6f}  a_WT {th[2]:.3f}  a_TT {th[5]:.3f}  "
            f"window {1e3 * rec['window_s']:.1f} ms")

    t0 = time.perf_counter()
    hist = fit.fit_rdf(e, spec, target, steps=a.steps, lr=a.lr, window=nwin, rewind=not a.no_rewind, callback=cb)
    total = time.perf_counter() - t0
    e.close()
    tw = float(np.median([h["window_s"] for h in hist]))
    names = eng.Engine.THETA_NAMES
    out = {
        "metric": "batched parameter fit: replica particle-steps/s through the window gradient",
        "value": R * N * nwin / tw, "unit": "particle-steps/s", "n_gpus": 1, "steps": a.steps, "warmup": a.equil,
        "ms_per_step": 1e3 * total / a.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": DTYPE, "data": "synthetic",
        "config": {"workload": f"batched parameter fit: {R} replicas x {N:,} beads (rho=3, 30% H-T-T lipids, "
                               f"k_angle=0, chemistry off, distinct seeds), {nwin}-step adjoint windows, checkpoint "
                               f"every 20, tail-tail RDF 64 bins on [0,3), replica-mean gradient, {a.steps} Adam "
                               f"steps (lr {a.lr}) on the six a_ij toward the RDF of a seed run at a_TT=20, a_WT=90, "
                               + ("trajectories continued" if a.no_rewind else
                                  "every window from the equilibrated state (same age as the target)"),
                   "replicas": R, "N": N},
        "gradient_cost_vs_plain": tw / t_plain,
        "loss_first": hist[0]["loss"], "loss_last": hist[-1]["loss"],
        "loss_min": min(h["loss"] for h in hist),
        "theta_first": {n: float(v) for n, v in zip(names, hist[0]["theta"])},
        "theta_last": {n: float(v) for n, v in zip(names, hist[-1]["theta"])},
        "loss_trace": [round(h["loss"], 6) for h in hist],
    }
    print(json.dumps(out), flush=True)


def refchem_params(pm, N, nsub, seed=1):
    """The reference's default parameters -- chem = 1 (catalyst synthesis and
    replication, decay, food supply with its 0.12 cap, a check every 10
    steps), 10% H-T-T lipids, 0.5% Ka, 12% N, dt 0.03, the warm-up ramp -- at
    N beads (L from N at rho = 3)."""
    return pm.load(overrides={"L": repr(pm.L_for_N(N)), "nsub": str(nsub), "seed": str(seed)}).check_engine()


def bench_refchem(a, pm, eng):
    """The reference engine's own default workload (physics and chemistry) on
    the GPU, and the reference CPU engine on the same parameters."""
    p = refchem_params(pm, a.N, a.nsub)
    N = p.N
    e = eng.Engine(p, pair_block=a.pair_block, tile_beads=a.tile_beads, pair_variant=a.pair_variant)
    e.start()
    e.step(a.warmup, graph=not a.no_graph)
    e.events()
    e.step_count()
    t0 = time.perf_counter()
    e.step(a.steps, graph=not a.no_graph)
    sec = time.perf_counter() - t0
    ev = e.events()
    st = e.stats()
    e.close()
    names = {3: "dlip", 4: "dN", 5: "feed", 6: "syn", 7: "rep", 8: "mut", 9: "dK"}
    kinds = {names[k]: int((ev["kind"] == k).sum()) for k in names}
    cpu = None
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle
    binp = oracle.ref_binary(fast=True)
    if not a.no_cpu and os.path.exists(binp):
        host = host_info()
        threads = cpu_threads(p.L, host)
        ov = {"L": repr(p.L), "nsub": str(a.nsub), "seed": "1", "snap_every": "0", "stats_every": "0",
              "ckpt_every": "0", "log_every": "0", "events": "0", "threads": str(threads)}
        tmp = tempfile.mkdtemp(prefix="ocref_")
        try:
            def run(steps):
                out = oracle.run_ref_cli(os.path.join(tmp, f"r{steps}"), dict(ov, steps=str(steps)), fast=True)
                return float(re.search(r"timing: (\d+) steps, ([0-9.]+) ms/step", out).group(2))
            ms = run(10)
            steps = int(max(20, min(3000, 12.0 / (ms * 1e-3))))
            steps = (steps + 9) // 10 * 10
            ms = run(steps)
        finally:
            shutil.rmtree(tmp, ignore_errors=True)
        cpu = {"value": N / (ms * 1e-3), "unit": "particle-steps/s", "cores": threads, "kind": "reference",
               "sample": f"reference CLI {os.path.basename(binp)} (cel/engine/oercel.c, -O3), N={N}, its default "
                         f"parameters (chemistry on), {steps} steps from a fresh soup, {threads} OpenMP threads, "
                         f"{ms:.3f} ms/step", "host": host}
    out = {
        "metric": "2D DPD particle-steps/s with the reference's default chemistry, N=1M",
        "value": N * a.steps / sec, "unit": "particle-steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": 1e3 * sec / a.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": DTYPE, "data": "synthetic",
        "config": {"workload": f"the reference's default parameters at N={N:,} (L={p.L:.4f}): chem=1 with "
                               "catalysts (synthesis, replication), decay and the capped food supply every 10 "
                               "steps, 10% H-T-T lipids, dt=0.03, warm-up ramp",
                   "N": N, "nsub": a.nsub, "graph": not a.no_graph, "events_in_timed_region": kinds,
                   "T_kin_end": st["T_kin"]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default 500; fit: 50 Adam steps)")
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--N", type=int, default=None, help="beads (default 1,048,576; fit: 262,144 per replica)")
    ap.add_argument("--nsub", type=int, default=4)
    ap.add_argument("--pair-block", type=int, default=0)
    ap.add_argument("--tile-beads", type=int, default=0)
    ap.add_argument("--pair-variant", type=int, default=None,
                    help="pair kernel: 0 single pass, 1 per-lane two phases, 2 cross-lane pair queue "
                         "(default: the library's)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--workload", choices=("plain", "ensemble", "adjoint", "fit", "refchem", "copyprobe"),
                    default="plain")
    ap.add_argument("--window", type=int, default=200)
    ap.add_argument("--equil", type=int, default=1000)
    ap.add_argument("--replicas", type=int, default=None, help="default 256 (ensemble), 32 (fit)")
    ap.add_argument("--lr", type=float, default=0.5, help="Adam step size of the fit workload")
    ap.add_argument("--no-rewind", action="store_true",
                    help="fit workload: continue the trajectories from window to window instead of restarting every "
                         "window from the equilibrated state")
    ap.add_argument("--no-tune", action="store_true", help="ensemble: library launch shape, no autotuner")
    ap.add_argument("--no-sub", action="store_true",
                    help="plain workload: skip the ensemble and adjoint sub-measurements")
    ap.add_argument("--ens-steps", type=int, default=200, help="timed steps of the ensemble sub-measurement")
    ap.add_argument("--ens-equil", type=int, default=500, help="untimed steps of the ensemble sub-measurement")
    a = ap.parse_args()
    fitw = a.workload == "fit"
    a.steps = a.steps if a.steps is not None else (50 if fitw else (200 if a.workload == "ensemble" else 500))
    a.N = a.N if a.N is not None else (262144 if fitw else 1048576)
    a.replicas = a.replicas if a.replicas is not None else (32 if fitw else 256)
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build runs the step on a single MI355X "
              "(multi-GPU domain decomposition is a design note, not implemented).", file=sys.stderr)
        sys.exit(2)

    if a.workload == "copyprobe":
        return copy_probe()
    pm = importlib.import_module("welcome-youtube_amd.params")
    eng = importlib.import_module("welcome-youtube_amd.engine")
    if a.workload == "ensemble":
        return bench_ensemble(a, pm, eng)
    if a.workload == "adjoint":
        return bench_adjoint(a, pm, eng)
    if a.workload == "fit":
        return bench_fit(a, pm, eng)
    if a.workload == "refchem":
        return bench_refchem(a, pm, eng)
    p = config2_params(pm, a.N, a.nsub)
    N = p.N

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    e = eng.Engine(p, pair_block=a.pair_block, tile_beads=a.tile_beads, pair_variant=a.pair_variant)
    e.start()
    # untimed equilibration from the random soup (no warm-up ramp), so the
    # timed steps run at the steady state (kT ~ 1), then the user's warm-up.
    # The event drain and the temperature read (host work, the device idle)
    # come before the equilibration's last 100 steps, which run in one call
    # with the warm-up: the device is busy without a pause up to the timed
    # steps (a pause lets the clocks drop for the first milliseconds after it)
    run_in = min(100, a.equil)
    e.step(a.equil - run_in, graph=not a.no_graph)
    e.events()  # drain
    st0 = e.stats()  # T_kin 100 steps before the warm-up steps
    e.step(run_in + a.warmup, graph=not a.no_graph)
    s0 = e.step_count()  # device synchronised; events of steps > s0 are the timed region's
    log(f"plain N={N}: {a.equil} + {a.warmup} untimed steps, T_kin {st0['T_kin']:.4f}")
    t0 = time.perf_counter()
    e.step(a.steps, graph=not a.no_graph)  # returns after a stream synchronisation
    t1 = time.perf_counter()
    sec = t1 - t0
    st = e.stats()
    ev = e.events()
    ev = ev[ev["step"] > s0]
    log(f"plain: {1e3 * sec / a.steps:.4f} ms/step, T_kin {st['T_kin']:.4f}")
    pair_s = time_pair_kernel(e)
    ncell = e.cells()[0].size - 1
    # algorithmic bytes of one force launch, per bead at the compulsory
    # layout: the pair force (FP32 cell-relative position + predicted velocity
    # 16, gid|type 4, force 8, cell starts ~1.3), the bonded terms of the 30%
    # lipid beads (~11) and the next step's cell key (4)
    bytes_alg = int(N * (16 + 4 + 8 + 11 + 4)) + 4 * (ncell + 1)
    # bytes the kernel must move in this build's layout: staged xy (FP64 16),
    # vt (8), gid|type (4), links mp (1), lnk (4), prv (4), the half-step
    # velocity the next-key epilogue integrates (16), force (16) and key (4)
    bytes_layout = N * (16 + 8 + 4 + 1 + 4 + 4 + 16 + 16 + 4) + 4 * (ncell + 1)
    achieved = bytes_alg / pair_s / 1e9
    e.close()
    sub = {}
    if not a.no_sub:
        log("ensemble sub-workload ...")
        sub["ensemble"] = ensemble_measure(pm, eng, 256, a.ens_steps, a.ens_equil, tune=not a.no_tune)
        log(f"ensemble: {sub['ensemble']['value']:.3e} particle-steps/s, "
            f"{sub['ensemble']['speedup_vs_sequential']:.2f}x sequential")
        log("adjoint sub-workload ...")
        sub["adjoint"] = adjoint_measure(pm, eng, a.window, 1000)
        log(f"adjoint: {sub['adjoint']['value']:.3f}x forward")
    cpu = None if a.no_cpu else cpu_baseline(pm, N, a.nsub)
    out = {
        "metric": METRIC,
        "value": N * a.steps / sec,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * sec / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE,
        "data": "synthetic",
        "config": {
            "workload": f"plain 2D DPD step at scale: N={N:,}, rho=3, L={p.L:.4f}, 50% W / 20% P / 30% H-T-T "
                        "lipid beads, precursor chemistry every step, dt=0.01, k_angle=0, no warm-up ramp, "
                        f"{a.equil} untimed equilibration steps from the random soup before the warm-up",
            "N": N, "L": p.L, "nsub": a.nsub, "graph": not a.no_graph, "equil": a.equil,
            "state": "fp64 positions/velocities/forces, fp32 pair arithmetic",
            "T_kin_start": st0["T_kin"], "T_kin_end": st["T_kin"],  # start: before the warm-up steps
            "events_in_timed_region": int(len(ev)),
            "momentum_end": [st["px"], st["py"]],
        },
        "roofline": {
            "bound": "hbm", "kernel": "k_pair", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(N),
            "avg_launch_us": pair_s * 1e6, "algorithmic_bytes_per_launch": bytes_alg,
            "layout_bytes_per_launch": bytes_layout,
            "limiter": "issue and latency of the per-candidate FP32 pair arithmetic and hash on LDS-staged "
                       "candidates, not HBM",
        },
        "cpu_baseline": cpu,
    }
    out.update(sub)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
