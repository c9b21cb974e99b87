"""Benchmark of the InteraxFF HIP platform on one MI355X.

The headline: the fused dispersion + exchange + short-range pair evaluation of the 98304-atom water-like box
(9.94 nm, cutoff 1.0 nm, long range correction), FP32 arithmetic with FP64 positions and 64-bit fixed-point forces.
A step is one force+energy evaluation on coordinates perturbed by N(0, 0.005 nm); the coordinates of every step are
resident in HBM before the timed region starts.  The default run adds, in the same JSON line, the other measured
forms: "triplet" (the 32768-site ATM fluid), "batched" (1024 clusters in one graph-replayed evaluation) and "served"
(Poisson request streams with p50/p99 latency).  One JSON line is printed.

  python bench.py [--gpus 1] [--steps 20] [--warmup 5] [--workload all|pair|triplet|batch|serve|lambda|copy]
"""
import argparse
import ctypes
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (os.path.join(ROOT, "openmm-interax-plugin_amd"), os.path.join(ROOT, "oracle")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np

PEAK_FP32_TFLOPS = 157.3     # MI355X FP32 vector peak (packed), TFLOP/s
PEAK_HBM_GBS = 8000.0
FLOPS_PER_PAIR = 240         # fused dispersion + exchange + short range, FMA = 2
TIMER_STRIDE = 10            # pair kernel timed with HIP events on every 10th step of the timed loop
FLOPS_PER_TRIPLE = 130       # undamped ATM triple


def log(*args):
    print(*args, file=sys.stderr, flush=True)


def cpu_threads():
    """Threads of the CPU baseline: the cores this process may run on (its affinity), capped by OMP_NUM_THREADS, the
    share of a multi-GPU host that one GPU's jobs are given."""
    try:
        available = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        available = os.cpu_count() or 1
    cap = os.environ.get("OMP_NUM_THREADS", "")
    if cap.isdigit() and int(cap) > 0:
        available = min(available, int(cap))
    return max(1, available)


def host_description():
    """The host the CPU baseline ran on: CPU model, logical CPUs, the affinity of this process, SIMD width."""
    flags = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    flags = line
                    break
    except OSError:
        pass
    simd = "AVX-512" if "avx512f" in flags else "AVX2" if "avx2" in flags else "unknown"
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = None
    return {"cpu_model": cpu_description(), "nproc": os.cpu_count(), "affinity": affinity,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS"), "simd": simd}


def time_cpu_platform(evaluate, budget_s=8.0, max_runs=20):
    """Seconds per call of evaluate() (the multi-threaded CPU-platform restatement), over repeated calls until the
    budget is spent (at least one call)."""
    runs = 0
    start = time.perf_counter()
    while True:
        evaluate()
        runs += 1
        spent = time.perf_counter()-start
        if spent >= budget_s or runs >= max_runs:
            return spent/runs, runs


def cpu_description():
    model = platform.processor() or "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model


def traffic_from_profiles(kernel: str):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE, calibrated), if recorded."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            data = json.load(f)
        return data.get(kernel, {}).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


class Frames:
    """One device state (posq, correction, force buffer, energy) per step's coordinates."""

    def __init__(self, ih, num_particles, positions_list, box, precision):
        self.ih = ih
        self.lib = ih.load_library()
        self.states = []
        self.buffers = []
        for pos in positions_list:
            s = ctypes.c_void_p()
            ih._check(self.lib.ixh_state_create(0, num_particles, precision, ctypes.byref(s)))
            p = np.ascontiguousarray(pos.reshape(-1))
            ih._check(self.lib.ixh_state_set_positions(s, ih._dptr(p)))
            b = np.ascontiguousarray(np.asarray(box, dtype=np.float64).reshape(-1))
            ih._check(self.lib.ixh_state_set_box(s, ih._dptr(b)))
            buf = ih.Buffers()
            ih._check(self.lib.ixh_state_get_buffers(s, ctypes.byref(buf)))
            self.states.append(s)
            self.buffers.append(buf)

    def forces(self, k, n):
        out = np.zeros((n, 3))
        self.ih._check(self.lib.ixh_state_get_forces(self.states[k], self.ih._dptr(out)))
        return out

    def energy(self, k):
        e = ctypes.c_double()
        self.ih._check(self.lib.ixh_state_get_energy(self.states[k], ctypes.byref(e)))
        return e.value

    def close(self):
        for s in self.states:
            self.lib.ixh_state_destroy(s)


def bench_pair(args):
    import interax_hip as ih
    import workloads
    lib = ih.load_library()
    t0 = time.time()
    system, base = workloads.config1()
    n = system.getNumParticles()
    box = system.getDefaultPeriodicBoxVectors()
    total = args.warmup+args.steps
    frames_pos = [workloads.perturbed(base, k) for k in range(total)]
    log(f"[bench] workload built in {time.time()-t0:.1f} s: {n} atoms, {total} coordinate sets")
    frames = Frames(ih, n, frames_pos, box, ih.IXH_MIXED)
    forces = [f for f in system.getForces() if isinstance(f, ih.InteraxPairForce)]
    t0 = time.time()
    group = ih._PairGroup(0, system, forces, ih.IXH_MIXED)
    log(f"[bench] kernel handle created (long range correction on the device) in {time.time()-t0:.1f} s")
    flags = ih.IXH_INCLUDE_FORCES | ih.IXH_INCLUDE_ENERGY
    terms = ih.IXH_TERM_DISPERSION | ih.IXH_TERM_EXCHANGE | ih.IXH_TERM_SHORT_RANGE
    # Every step on one stream, as a platform runs a Context's kernels: the neighbour-list check of a step runs on
    # the device ahead of its tiles and its answer is read while the tiles run; a step whose list turned out stale is
    # evaluated again with a rebuilt list (and counted in the timed region).
    stream = frames.buffers[0].stream
    for k, st in enumerate(frames.states):
        ih._check(lib.ixh_state_set_stream(st, stream))
        ih._check(lib.ixh_state_get_buffers(st, ctypes.byref(frames.buffers[k])))
    reruns = [0]
    # launch shape of the tile kernel: tiles per work item, tuned on the first coordinate set
    t0 = time.time()
    tiles_per_item = ctypes.c_int()
    ih._check(lib.ixh_pair_autotune(group.handle, ctypes.byref(frames.buffers[0]), terms, ctypes.byref(tiles_per_item)))
    log(f"[bench] tile kernel tuned in {time.time()-t0:.2f} s: {tiles_per_item.value} tiles per work item")

    def step(k):
        ih._check(lib.ixh_state_clear(frames.states[k]))
        ih._check(lib.ixh_pair_execute(group.handle, ctypes.byref(frames.buffers[k]), terms, flags | ih.IXH_DEFER_LIST_CHECK))
        valid = ctypes.c_int()
        ih._check(lib.ixh_pair_list_was_valid(group.handle, ctypes.byref(valid)))
        if not valid.value:
            reruns[0] += 1
            ih._check(lib.ixh_state_clear(frames.states[k]))
            ih._check(lib.ixh_pair_execute(group.handle, ctypes.byref(frames.buffers[k]), terms, flags))

    for k in range(args.warmup):
        step(k)
    for k in range(args.warmup):
        ih._check(lib.ixh_state_synchronize(frames.states[k]))
    ih._check(lib.ixh_pair_reset_kernel_time(group.handle))
    reruns[0] = 0
    # The pair kernel's duration comes from HIP events recorded around it on every TIMER_STRIDE-th step of the timed
    # loop: each recorded event leaves the queue idle for a few microseconds, which timing every step would add to
    # the throughput being measured.
    start = time.perf_counter()
    for k in range(args.warmup, total):
        timed = (k-args.warmup) % TIMER_STRIDE == 0
        if timed:
            ih._check(lib.ixh_set_profiling(1))
        step(k)
        if timed:
            ih._check(lib.ixh_set_profiling(0))
    for k in range(args.warmup, total):
        ih._check(lib.ixh_state_synchronize(frames.states[k]))
    elapsed = time.perf_counter()-start
    kernel_ms = ctypes.c_double()
    launches = ctypes.c_longlong()
    ih._check(lib.ixh_pair_kernel_time(group.handle, ctypes.byref(kernel_ms), ctypes.byref(launches)))
    pairs = ctypes.c_longlong()
    ih._check(lib.ixh_pair_last_pair_count(group.handle, ctypes.byref(pairs)))
    pairs = pairs.value
    tiles, candidates = ctypes.c_longlong(), ctypes.c_longlong()
    ih._check(lib.ixh_pair_list_statistics(group.handle, ctypes.byref(tiles), ctypes.byref(candidates)))
    value = pairs*args.steps/elapsed
    avg_kernel_s = kernel_ms.value/1e3/max(1, launches.value)
    achieved = FLOPS_PER_PAIR*pairs/avg_kernel_s/1e12
    log(f"[bench] {args.steps} steps in {elapsed*1e3:.1f} ms; {pairs} pairs per step; pair kernel {avg_kernel_s*1e6:.1f} us")

    result = {
        "metric": "pair-interactions/s", "value": value, "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": elapsed*1e3/args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "plain pair kernel at scale: 32768 water-like molecules (98304 atoms), 9.94 nm periodic box, "
                               "cutoff 1.0 nm, long range correction, dispersion+exchange+short range fused, coordinates "
                               "perturbed by N(0, 0.005 nm) per step",
                   "atoms": n, "pairs_per_step": pairs, "tiles": tiles.value, "candidate_pairs": candidates.value,
                   "stale_list_reevaluations": reruns[0], "tiles_per_work_item": tiles_per_item.value,
                   "precision": "mixed (FP64 positions, FP32 arithmetic, "
                                                                     "64-bit fixed-point forces, FP64 energy)"},
        "autotune": {"tiles_per_work_item": tiles_per_item.value, "searched": [1, 2, 3, 4, 6]},
        "roofline": {"bound": "valu", "note": "compute-bound: FP32 VALU (no MFMA on this path: nothing here is a dense "
                                              "contraction) against the MI355X FP32 peak; 240 algorithmic flops per "
                                              "interacting pair",
                     "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": achieved/PEAK_FP32_TFLOPS,
                     "traffic": traffic_from_profiles("k_pair_queue"), "kernel": "k_pair_queue",
                     "kernel_avg_us": avg_kernel_s*1e6, "launches": launches.value,
                     "timed_launches": f"HIP events recorded around every {TIMER_STRIDE}th launch of the timed loop"},
    }

    # accuracy of the first and the last timed step against the FP64 oracle (the long range correction, a constant
    # of the parameters, is left out of both sides: the tests check it), and the oracle's own speed
    if not args.skip_oracle:
        import oracle
        disp = forces[0]
        disp.setUseLongRangeCorrection(False)
        checks = []
        cpu_s = 0.0
        for k in (args.warmup, total-1):
            gpu_forces = frames.forces(k, n)
            gpu_energy = frames.energy(k)
            t0 = time.perf_counter()
            ref_energy = 0.0
            ref_forces = np.zeros((n, 3))
            for f in forces:
                e, fr = oracle.pair_force(f, frames_pos[k], box, 1.0, double_precision=True)
                ref_energy += e
                ref_forces += fr
            cpu_s = time.perf_counter()-t0
            err = gpu_forces-ref_forces
            checks.append({"step": k-args.warmup, "force_rms_rel": float(np.linalg.norm(err)/np.linalg.norm(ref_forces)),
                           "force_max_rel_to_rms": float(np.abs(err).max()/np.sqrt(np.mean(ref_forces**2))),
                           "energy_rel": abs(gpu_energy-ref_energy)/abs(ref_energy)})
        disp.setUseLongRangeCorrection(True)
        result["accuracy"] = {"force_rms_rel": max(c["force_rms_rel"] for c in checks),
                              "force_max_rel_to_rms": max(c["force_max_rel_to_rms"] for c in checks),
                              "energy_rel": max(c["energy_rel"] for c in checks), "steps": checks,
                              "against": "FP64 oracle (restated Reference platform), first and last timed step; "
                                         "worst of the two"}
        threads = cpu_threads()
        k = args.warmup
        per_call, runs = time_cpu_platform(
            lambda: [oracle.cpu_pair_force(f, frames_pos[k], box, 1.0, threads) for f in forces])
        result["cpu_baseline"] = {"value": pairs/per_call, "unit": "pairs/s", "cores": threads, "kind": "port",
                                  "sample": f"{runs} evaluations of the first timed step (the three pair forces, "
                                            f"{pairs} interacting pairs) by the FP32 multi-threaded restatement of the "
                                            f"reference CPU platform (oracle/cpu_platform.c), {threads} threads",
                                  "host": host_description(), "seconds_per_evaluation": per_call,
                                  "fp64_oracle": {"value": pairs/cpu_s, "unit": "pairs/s", "cores": 1, "seconds": cpu_s,
                                                  "sample": "one evaluation by the FP64 oracle, single-threaded"}}
    frames.close()
    group.close()
    return result


def bench_triplet(args):
    """The ATM triplet kernel on the 32768-site fluid (number density 21 nm^-3, minimum separation 0.3 nm, cutoff
    0.7 nm, C9 the geometric mean of per-site c_i ~ U(1e-6, 4e-6)), one force+energy evaluation per step on
    coordinates perturbed by N(0, 0.005 nm)."""
    import interax_hip as ih
    import workloads
    lib = ih.load_library()
    t0 = time.time()
    system, base = workloads.atm_fluid()
    n = system.getNumParticles()
    box = system.getDefaultPeriodicBoxVectors()
    total = args.warmup+args.steps
    frames_pos = [workloads.perturbed(base, k) for k in range(total)]
    log(f"[bench] workload built in {time.time()-t0:.1f} s: {n} sites, {total} coordinate sets")
    frames = Frames(ih, n, frames_pos, box, ih.IXH_MIXED)
    force = [f for f in system.getForces() if isinstance(f, ih.InteraxThreeBodyForce)][0]
    tb = ih._ThreeBody(0, system, force, ih.IXH_MIXED)
    flags = ih.IXH_INCLUDE_FORCES | ih.IXH_INCLUDE_ENERGY

    # every step on one stream, the list check deferred as for the pair forces: a step whose lists turned out stale is
    # evaluated again with rebuilt lists (and counted in the timed region)
    stream = frames.buffers[0].stream
    for k, st in enumerate(frames.states):
        ih._check(lib.ixh_state_set_stream(st, stream))
        ih._check(lib.ixh_state_get_buffers(st, ctypes.byref(frames.buffers[k])))
    reruns = [0]

    def step(k):
        ih._check(lib.ixh_state_clear(frames.states[k]))
        ih._check(lib.ixh_threebody_execute(tb.handle, ctypes.byref(frames.buffers[k]), flags | ih.IXH_DEFER_LIST_CHECK))
        valid = ctypes.c_int()
        ih._check(lib.ixh_threebody_list_was_valid(tb.handle, ctypes.byref(valid)))
        if not valid.value:
            reruns[0] += 1
            ih._check(lib.ixh_state_clear(frames.states[k]))
            ih._check(lib.ixh_threebody_execute(tb.handle, ctypes.byref(frames.buffers[k]), flags))

    for k in range(args.warmup):
        step(k)
    for k in range(args.warmup):
        ih._check(lib.ixh_state_synchronize(frames.states[k]))
    ih._check(lib.ixh_threebody_reset_kernel_time(tb.handle))
    reruns[0] = 0
    # the triple kernel is timed with HIP events on every TIMER_STRIDE-th step (as the pair kernel)
    start = time.perf_counter()
    for k in range(args.warmup, total):
        timed = (k-args.warmup) % TIMER_STRIDE == 0
        if timed:
            ih._check(lib.ixh_set_profiling(1))
        step(k)
        if timed:
            ih._check(lib.ixh_set_profiling(0))
    for k in range(args.warmup, total):
        ih._check(lib.ixh_state_synchronize(frames.states[k]))
    elapsed = time.perf_counter()-start
    kernel_ms = ctypes.c_double()
    launches = ctypes.c_longlong()
    ih._check(lib.ixh_threebody_kernel_time(tb.handle, ctypes.byref(kernel_ms), ctypes.byref(launches)))
    triples = ctypes.c_longlong()
    ih._check(lib.ixh_threebody_last_triple_count(tb.handle, ctypes.byref(triples)))
    triples = triples.value
    avg_kernel_s = kernel_ms.value/1e3/max(1, launches.value)
    achieved = FLOPS_PER_TRIPLE*triples/avg_kernel_s/1e12
    log(f"[bench] {args.steps} steps in {elapsed*1e3:.1f} ms; {triples} triples per step; triple kernel {avg_kernel_s*1e6:.1f} us")
    result = {
        "metric": "triple-interactions/s", "value": triples*args.steps/elapsed, "unit": "triples/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": elapsed*1e3/args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "ATM triplet kernel: 32768 single-site atoms, number density 21 nm^-3 (RSA, minimum "
                               "separation 0.3 nm), periodic cube, cutoff 0.7 nm, no damping, geometric-mean C9, "
                               "coordinates perturbed by N(0, 0.005 nm) per step",
                   "sites": n, "triples_per_step": triples, "stale_list_reevaluations": reruns[0],
                   "neighbour_lists": "cutoff + 0.1 nm skin, rebuilt when a site has moved 0.05 nm; checked on the "
                                      "device every step, the answer read while the triples run",
                   "precision": "mixed (FP64 site centroids, FP32 arithmetic, 64-bit fixed-point forces, FP64 energy)"},
        "roofline": {"bound": "valu", "note": "FP32 VALU (no MFMA on this path) against the MI355X FP32 peak; 130 "
                                              "algorithmic flops per interacting triple, over the triple kernel's "
                                              "HIP-event time (ms_per_step covers the whole evaluation: centroids, "
                                              "neighbour search, triples, site forces)",
                     "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": achieved/PEAK_FP32_TFLOPS,
                     "traffic": traffic_from_profiles("k_triples_packed"), "kernel": "k_triples_packed",
                     "kernel_avg_us": avg_kernel_s*1e6, "launches": launches.value,
                     "timed_launches": f"HIP events recorded around every {TIMER_STRIDE}th launch of the timed loop"},
    }
    if not args.skip_oracle:
        import oracle
        checks = []
        for k in (args.warmup, total-1):
            gpu_forces = frames.forces(k, n)
            gpu_energy = frames.energy(k)
            t0 = time.perf_counter()
            ref_energy, ref_forces = oracle.threebody_force(force, frames_pos[k], box, 1.0)
            cpu_s = time.perf_counter()-t0
            err = gpu_forces-ref_forces
            checks.append({"step": k-args.warmup, "force_rms_rel": float(np.linalg.norm(err)/np.linalg.norm(ref_forces)),
                           "force_max_rel_to_rms": float(np.abs(err).max()/np.sqrt(np.mean(ref_forces**2))),
                           "energy_rel": abs(gpu_energy-ref_energy)/abs(ref_energy)})
        result["accuracy"] = {"force_rms_rel": max(c["force_rms_rel"] for c in checks),
                              "force_max_rel_to_rms": max(c["force_max_rel_to_rms"] for c in checks),
                              "energy_rel": max(c["energy_rel"] for c in checks), "steps": checks,
                              "against": "FP64 oracle (restated Reference platform), first and last timed step"}
        k = args.warmup
        threads = cpu_threads()
        per_call, runs = time_cpu_platform(lambda: oracle.cpu_threebody_force(force, frames_pos[k], box, 1.0, threads))
        result["cpu_baseline"] = {"value": triples/per_call, "unit": "triples/s", "cores": threads, "kind": "port",
                                  "sample": f"{runs} three-body evaluations of step 0 ({triples} triples, the all-pairs "
                                            f"neighbour search of the CPU platform included) by the multi-threaded "
                                            f"restatement of the reference CPU platform (oracle/cpu_platform.c), "
                                            f"{threads} threads",
                                  "host": host_description(), "seconds_per_evaluation": per_call,
                                  "fp64_oracle": {"value": triples/cpu_s, "unit": "triples/s", "cores": 1, "seconds": cpu_s,
                                                  "sample": "the same evaluation by the FP64 oracle (cell lists), "
                                                            "single-threaded"}}
    frames.close()
    tb.close()
    return result


def bench_batch(args, sample_every=16):
    """The batched configuration: 1024 independent water-like clusters (molecule counts log-uniform in [2, 256],
    RSA-packed at 33.4 nm^-3), all three pair forces without cutoff plus a molecular ATM force on the oxygens (no
    damping, cutoff 50 nm, i.e. every triple).  A unit is one cluster evaluation.  The batched form (one segmented
    evaluation of all clusters, graph-replayed, launch shape tuned) is timed against 1024 sequential plain
    per-cluster evaluations (one pair and one three-body handle per cluster) on the same device."""
    import interax_hip as ih
    import workloads
    lib = ih.load_library()
    t0 = time.time()
    sizes = workloads.cluster_sizes(1024)
    built = [workloads.water_cluster(m, seed=k) for k, m in enumerate(sizes)]
    systems = [b[0] for b in built]
    positions = [b[1] for b in built]
    atoms = sum(len(p) for p in positions)
    m = np.array(sizes, dtype=np.int64)
    excluded = 3*m      # three intramolecular pairs per molecule
    pairs = int(((3*m)*(3*m-1)//2-excluded).sum())
    triples = int((m*(m-1)*(m-2)//6).sum())
    log(f"[bench] {len(systems)} clusters built in {time.time()-t0:.1f} s: {atoms} atoms, {pairs} pairs, {triples} triples")
    flags = ih.IXH_INCLUDE_FORCES | ih.IXH_INCLUDE_ENERGY

    # plain: one Context-like set of handles per cluster, evaluated one after another
    t0 = time.time()
    plain = []
    for system, pos in zip(systems, positions):
        n = system.getNumParticles()
        st = ctypes.c_void_p()
        ih._check(lib.ixh_state_create(0, n, ih.IXH_MIXED, ctypes.byref(st)))
        ih._check(lib.ixh_state_set_positions(st, ih._dptr(np.ascontiguousarray(pos.reshape(-1)))))
        buf = ih.Buffers()
        ih._check(lib.ixh_state_get_buffers(st, ctypes.byref(buf)))
        pair = ih._PairGroup(0, system, [f for f in system.getForces() if isinstance(f, ih.InteraxPairForce)], ih.IXH_MIXED)
        tb = ih._ThreeBody(0, system, [f for f in system.getForces() if isinstance(f, ih.InteraxThreeBodyForce)][0], ih.IXH_MIXED)
        plain.append((st, buf, pair, tb))
    terms = ih.IXH_TERM_DISPERSION | ih.IXH_TERM_EXCHANGE | ih.IXH_TERM_SHORT_RANGE

    def plain_pass():
        for st, buf, pair, tb in plain:
            ih._check(lib.ixh_state_clear(st))
            ih._check(lib.ixh_pair_execute(pair.handle, ctypes.byref(buf), terms, flags))
            ih._check(lib.ixh_threebody_execute(tb.handle, ctypes.byref(buf), flags))
        for st, _, _, _ in plain:
            ih._check(lib.ixh_state_synchronize(st))

    plain_pass()
    log(f"[bench] plain handles created and warmed up in {time.time()-t0:.1f} s")
    plain_steps = max(1, min(args.steps, 3))
    start = time.perf_counter()
    for _ in range(plain_steps):
        plain_pass()
    plain_s = (time.perf_counter()-start)/plain_steps
    plain_rate = len(systems)/plain_s
    log(f"[bench] plain: {plain_s*1e3:.1f} ms per 1024 clusters ({plain_rate:.0f} cluster evaluations/s)")
    for st, _, pair, tb in plain:
        pair.close()
        tb.close()
        lib.ixh_state_destroy(st)

    # batched
    t0 = time.time()
    batch = ih.Batch(systems, "mixed")
    batch.setPositions(positions)
    default_shape = batch.shape()
    for _ in range(max(1, args.warmup)):
        batch.execute()
    batch.energies()
    start = time.perf_counter()
    for _ in range(args.steps):
        batch.execute()
    batch.energies()
    default_s = (time.perf_counter()-start)/args.steps
    chosen = batch.autotune()
    log(f"[bench] batch created and tuned in {time.time()-t0:.1f} s: {default_s*1e3:.3f} ms per batch with the default "
        f"shape {default_shape}, chosen {chosen}")
    for _ in range(args.warmup):
        batch.execute()
    batch.energies()
    start = time.perf_counter()
    for _ in range(args.steps):
        batch.execute()
    energies = batch.energies()
    batch_s = (time.perf_counter()-start)/args.steps
    batch_rate = len(systems)/batch_s
    log(f"[bench] batched: {batch_s*1e3:.2f} ms per 1024 clusters ({batch_rate:.0f} cluster evaluations/s, "
        f"{batch_rate/plain_rate:.1f}x plain)")
    # kernel times of eager, profiled evaluations (HIP events around the pair-tile and triple kernels)
    ih._check(lib.ixh_batch_reset_kernel_time(batch.handle))
    ih._check(lib.ixh_set_profiling(1))
    for _ in range(3):
        batch.execute()
    batch.energies()
    ih._check(lib.ixh_set_profiling(0))
    pair_ms, triple_ms = ctypes.c_double(), ctypes.c_double()
    pair_n, triple_n = ctypes.c_longlong(), ctypes.c_longlong()
    ih._check(lib.ixh_batch_kernel_time(batch.handle, ctypes.byref(pair_ms), ctypes.byref(pair_n), ctypes.byref(triple_ms),
                                        ctypes.byref(triple_n)))
    triple_s = triple_ms.value/1e3/max(1, triple_n.value)
    pair_kernel_s = pair_ms.value/1e3/max(1, pair_n.value)
    forces = batch.forces()
    flops = 240*pairs+130*triples
    result = {
        "metric": "batched cluster evaluations/s", "value": batch_rate, "unit": "cluster evals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": batch_s*1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "1024 independent water-like clusters, molecule counts log-uniform in [2, 256], no cutoff, "
                               "dispersion+exchange+short range plus ATM on the oxygens (cutoff 50 nm)",
                   "clusters": len(systems), "atoms": atoms, "pairs_per_batch": pairs, "triples_per_batch": triples,
                   "launch_shape": chosen, "graph_replay": True},
        "autotune": {"chosen": chosen, "default": default_shape, "ms_per_step_default_shape": default_s*1e3,
                     "searched": "pair tile launch 256-16384 workgroups, then triple kernel 1/2/4/8 wavefronts x "
                                 "8192-65536 neighbour pairs per work item; every shape gives the same results bit for bit"},
        "plain": {"value": plain_rate, "unit": "cluster evals/s", "ms_per_1024": plain_s*1e3,
                  "form": "1024 sequential evaluations, one pair and one three-body handle per cluster"},
        "speedup_vs_plain": batch_rate/plain_rate,
        "roofline": {"bound": "valu", "note": "the dominant kernel, k_triple_groups (ATM triples of the batch): 130 "
                                              "algorithmic FP32 flops per triple over its HIP-event duration (eager "
                                              "evaluations); no MFMA on this path",
                     "achieved": 130*triples/triple_s/1e12, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": 130*triples/triple_s/1e12/PEAK_FP32_TFLOPS,
                     "traffic": traffic_from_profiles("k_triple_groups_batch"), "kernel": "k_triple_groups",
                     "kernel_avg_us": triple_s*1e6, "launches": triple_n.value},
        "roofline_evaluation": {"bound": "valu", "note": "240 flops per pair and 130 per triple over the whole batched "
                                                         "evaluation (all kernels, graph replay)",
                                "achieved": flops/batch_s/1e12, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                                "frac": flops/batch_s/1e12/PEAK_FP32_TFLOPS},
        "roofline_pair_tiles": {"bound": "valu", "kernel": "k_pair_rotate", "kernel_avg_us": pair_kernel_s*1e6,
                                "note": "the pair half's tile launch, event-timed while it shares the device with the "
                                        "triple kernel (it starts after the three-body neighbour search)",
                                "achieved": 240*pairs/pair_kernel_s/1e12, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                                "frac": 240*pairs/pair_kernel_s/1e12/PEAK_FP32_TFLOPS},
    }
    if not args.skip_oracle:
        import oracle
        sample = list(range(0, len(systems), sample_every))
        t0 = time.perf_counter()
        err2 = ref2 = 0.0
        max_rel = energy_rel = 0.0
        for k in sample:
            e_ref, f_ref = oracle.evaluate(systems[k], positions[k])
            d = forces[k]-f_ref
            err2 += float((d**2).sum())
            ref2 += float((f_ref**2).sum())
            max_rel = max(max_rel, float(np.abs(d).max()/np.sqrt(np.mean(f_ref**2))))
            energy_rel = max(energy_rel, abs(energies[k]-e_ref)/abs(e_ref))
        cpu_s = time.perf_counter()-t0
        result["accuracy"] = {"force_rms_rel": float(np.sqrt(err2/ref2)), "force_max_rel_to_rms": max_rel,
                              "energy_rel_max": energy_rel,
                              "against": f"FP64 oracle on every {sample_every}th cluster ({len(sample)} clusters)"}
        threads = cpu_threads()

        def cpu_pass():
            for k in range(len(systems)):
                for f in systems[k].getForces():
                    if isinstance(f, ih.InteraxPairForce):
                        oracle.cpu_pair_force(f, positions[k], None, 1.0, threads)
                    else:
                        oracle.cpu_threebody_force(f, positions[k], None, 1.0, threads)
        per_call, runs = time_cpu_platform(cpu_pass, budget_s=10.0, max_runs=3)
        result["cpu_baseline"] = {"value": len(systems)/per_call, "unit": "cluster evals/s", "cores": threads, "kind": "port",
                                  "sample": f"{runs} passes over the 1024 clusters, one after another, by the "
                                            f"multi-threaded restatement of the reference CPU platform "
                                            f"(oracle/cpu_platform.c: FP32 pair forces, FP64 three-body), {threads} "
                                            f"threads",
                                  "host": host_description(), "seconds_per_pass": per_call,
                                  "fp64_oracle": {"value": len(sample)/cpu_s, "unit": "cluster evals/s", "cores": 1,
                                                  "seconds": cpu_s, "sample": f"{len(sample)} of the 1024 clusters (every "
                                                  f"{sample_every}th) by the FP64 oracle, single-threaded"}}
    batch.close()
    return result


def bench_serve(args, catalog_size=256, requests_per_rate=2000, rates=(1000, 2000, 5000, 10000, 20000, 50000)):
    """The served configuration: Poisson request streams at several rates, mixing cluster evaluations (a catalog of
    water-like clusters with 2-1000 molecules, log-uniform, all pair forces without cutoff plus the molecular ATM force)
    and alchemical lambda windows (a 12-atom alchemical solute in 1000 water-like molecules, periodic 3.106 nm box,
    cutoff 0.9 nm, soft-core radius 0 or 0.3 nm: the energies of one configuration at the 16 values of lambda_sterics
    in linspace(0, 1, 16), one request, by the fused window evaluation); 3 in 4 requests are clusters.  Coordinates of
    every request are the registered ones perturbed by N(0, 0.005 nm)."""
    import interax_hip as ih
    import workloads
    t0 = time.time()
    sizes = workloads.cluster_sizes(catalog_size, seed=7, low=2, high=1000)
    catalog = [workloads.water_cluster(m, seed=10000+k) for k, m in enumerate(sizes)]
    boxes = [workloads.alchemical_box(0.0, seed=1), workloads.alchemical_box(0.3, seed=2)]
    server = ih.Server("mixed", max_batch=64)
    entries = [server.add_system(s) for s, _ in catalog]
    alch = [server.add_system(s, p) for s, p in boxes]
    server.start()
    base = [p for _, p in catalog]+[p for _, p in boxes]
    log(f"[bench] served catalog: {len(catalog)} clusters ({sum(len(p) for _, p in catalog)} atoms) and 2 alchemical "
        f"boxes of {len(boxes[0][1])} atoms, built in {time.time()-t0:.1f} s")
    # warm up every system once
    for e, p in zip(entries+alch, base):
        server.wait(server.submit(e, p, 0.5), e)
    for e, p in zip(alch, base[len(entries):]):
        server.wait_windows(server.submit_windows(e, p, np.linspace(0, 1, 16)))
    rng = np.random.default_rng(11)
    lambdas_grid = np.linspace(0, 1, 16)
    table = []
    for rate in rates:
        n = requests_per_rate
        arrival = np.cumsum(rng.exponential(1.0/rate, size=n))
        is_cluster = rng.random(n) < 0.75
        which = np.where(is_cluster, rng.integers(0, len(entries), size=n), len(entries)+rng.integers(0, 2, size=n))
        lam = np.ones(n)
        windows = [None if c else lambdas_grid for c in is_cluster]
        positions = [base[k]+rng.normal(0, 0.005, size=base[k].shape) for k in which]
        b0, r0 = server.statistics()
        latency, _, makespan = server.replay(arrival, np.array([(entries+alch)[k] for k in which]), positions, lam, windows)
        b1, r1 = server.statistics()
        row = {"rate": rate, "requests": n, "throughput": n/makespan, "p50_ms": float(np.percentile(latency, 50)),
               "p99_ms": float(np.percentile(latency, 99)),
               "cluster_p50_ms": float(np.percentile(latency[is_cluster], 50)),
               "cluster_p99_ms": float(np.percentile(latency[is_cluster], 99)),
               "alchemical_p50_ms": float(np.percentile(latency[~is_cluster], 50)),
               "alchemical_p99_ms": float(np.percentile(latency[~is_cluster], 99)),
               "mean_cluster_batch": (r1-r0)/max(1, b1-b0)}
        table.append(row)
        log(f"[bench] rate {rate}/s: throughput {row['throughput']:.0f}/s, p50 {row['p50_ms']:.2f} ms, p99 {row['p99_ms']:.2f} ms, "
            f"clusters p50/p99 {row['cluster_p50_ms']:.2f}/{row['cluster_p99_ms']:.2f} ms, alchemical p50/p99 "
            f"{row['alchemical_p50_ms']:.2f}/{row['alchemical_p99_ms']:.2f} ms, mean batch {row['mean_cluster_batch']:.1f}")
    accuracy = None
    if not args.skip_oracle:
        # sampled requests (clusters small enough for the oracle to finish in seconds, and both alchemical boxes at
        # two lambda values) against the FP64 oracle
        import oracle
        small = [k for k, m in enumerate(sizes) if m <= 120][:6]
        err2 = ref2 = worst = energy_rel = 0.0
        for k in small:
            system, pos = catalog[k]
            moved = pos+rng.normal(0, 0.005, size=pos.shape)
            e, f, _ = server.wait(server.submit(entries[k], moved), entries[k])
            e_ref, f_ref = oracle.evaluate(system, moved)
            err2 += float(((f-f_ref)**2).sum())
            ref2 += float((f_ref**2).sum())
            worst = max(worst, float(np.abs(f-f_ref).max()/np.sqrt(np.mean(f_ref**2))))
            energy_rel = max(energy_rel, abs(e-e_ref)/abs(e_ref))
        window_rel = 0.0
        for b, (system, pos) in enumerate(boxes):
            moved = pos+rng.normal(0, 0.005, size=pos.shape)
            values = np.array([0.2, 0.8])
            energies, _ = server.wait_windows(server.submit_windows(alch[b], moved, values))
            for v, e in zip(values, energies):
                e_ref = oracle.evaluate(system, moved, system.getDefaultPeriodicBoxVectors(),
                                        parameters={"lambda_sterics": float(v)})[0]
                window_rel = max(window_rel, abs(e-e_ref)/abs(e_ref))
        accuracy = {"force_rms_rel": float(np.sqrt(err2/ref2)), "force_max_rel_to_rms": worst,
                    "energy_rel_max": max(energy_rel, window_rel), "window_energy_rel_max": window_rel,
                    "against": f"FP64 oracle on {len(small)} served cluster requests (at most 120 molecules) and 2 "
                               f"window requests (2 lambda values each, both alchemical boxes)"}
    memory = server.memory()
    server.close()
    best = max(table, key=lambda r: r["throughput"])
    return {"metric": "served requests/s with p50/p99 latency", "value": best["throughput"], "unit": "requests/s",
            "n_gpus": 1, "steps": requests_per_rate, "warmup": 1, "ms_per_step": None, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
            "config": {"workload": "Poisson request streams: 3/4 cluster evaluations from a catalog of 256 water-like "
                                   "clusters (2-1000 molecules, log-uniform, pair forces without cutoff + ATM), 1/4 "
                                   "alchemical lambda-window requests (one configuration of a 12-atom solute in 1000 "
                                   "waters, 3.106 nm box, cutoff 0.9 nm, soft-core radius 0 or 0.3 nm, energies at 16 "
                                   "lambda values by the fused window evaluation)",
                       "max_batch": 64, "lanes": 2},
            "rates": table, "accuracy": accuracy, "memory": memory}


def bench_lambda(args, windows=16):
    """Lambda-window fusion on the alchemical job of the served configuration (a 12-atom alchemical solute in 1000
    water-like molecules, periodic 3.106 nm box, cutoff 0.9 nm, soft-core radius 0.3 nm): the energies of one
    configuration at 16 values of lambda_sterics, as 16 evaluations (one getState per value) and as one fused call
    (one evaluation plus the alchemical pairs at every value).  A unit is one window energy."""
    import interax_hip as ih
    import workloads
    system, pos = workloads.alchemical_box(0.3, seed=2, use_lrc=True)
    values = np.linspace(0.0, 1.0, windows)
    context = ih.Context(system, {"Precision": "mixed"})
    context.setPositions(pos)
    for _ in range(args.warmup):
        context.getEnergiesAtLambdas(values)
    start = time.perf_counter()
    for _ in range(args.steps):
        fused = context.getEnergiesAtLambdas(values)
    fused_s = (time.perf_counter()-start)/args.steps

    def separate():
        out = []
        for v in values:
            context.setParameter("lambda_sterics", float(v))
            out.append(context.getState(getForces=False).getPotentialEnergy())
        return np.array(out)
    separate()
    start = time.perf_counter()
    for _ in range(args.steps):
        reference = separate()
    separate_s = (time.perf_counter()-start)/args.steps
    context.close()
    log(f"[bench] {windows} window energies: fused {fused_s*1e3:.3f} ms, separate {separate_s*1e3:.3f} ms "
        f"({separate_s/fused_s:.1f}x); max difference {np.abs(fused-reference).max():.3e} kJ/mol")
    return {"metric": "lambda-window energies/s", "value": windows/fused_s, "unit": "window energies/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": fused_s*1e3, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
            "config": {"workload": f"energies of one configuration of the alchemical job (12-atom alchemical solute in "
                                   f"1000 water-like molecules, 3012 atoms, periodic 3.106 nm, cutoff 0.9 nm, soft-core "
                                   f"radius 0.3 nm, long range correction) at {windows} values of lambda_sterics",
                       "atoms": system.getNumParticles(), "windows": windows},
            "separate": {"value": windows/separate_s, "unit": "window energies/s", "ms_per_16": separate_s*1e3,
                         "form": "one evaluation (getState, energy only) per lambda value"},
            "speedup_vs_separate": separate_s/fused_s,
            "max_difference_kj_mol": float(np.abs(fused-reference).max())}


def copy_calibration(args):
    """A device-to-device copy of a known size: the HBM counters of its blit kernel calibrate FETCH_SIZE and
    WRITE_SIZE (bytes per counted unit) for the traffic figures of the other kernels."""
    hip = ctypes.CDLL("libamdhip64.so")
    size = 256 << 20
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(src), ctypes.c_size_t(size)) == 0
    assert hip.hipMalloc(ctypes.byref(dst), ctypes.c_size_t(size)) == 0
    assert hip.hipMemset(src, 1, ctypes.c_size_t(size)) == 0
    for _ in range(args.steps):
        assert hip.hipMemcpy(dst, src, ctypes.c_size_t(size), 3) == 0   # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)
    return {"workload": "copy", "bytes_per_copy": size, "copies": args.steps}


def main():
    parser = argparse.ArgumentParser()
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=20)
    parser.add_argument("--warmup", type=int, default=5)
    parser.add_argument("--workload", default="all",
                        choices=["all", "pair", "triplet", "batch", "serve", "lambda", "copy"])
    parser.add_argument("--skip-oracle", action="store_true", help="no accuracy check and no CPU baseline")
    args = parser.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build runs on a single MI355X and has no multi-GPU path "
              f"(--gpus {args.gpus} requested)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "copy":
        result = copy_calibration(args)
    elif args.workload == "triplet":
        result = bench_triplet(args)
    elif args.workload == "batch":
        result = bench_batch(args)
    elif args.workload == "serve":
        result = bench_serve(args)
    elif args.workload == "lambda":
        result = bench_lambda(args)
    else:
        result = bench_pair(args)
        if args.workload == "all":
            # the other forms of the metric, measured in the same run: config 2 (ATM fluid), config 3 (batched
            # clusters), config 4 (served requests with p50/p99)
            for key, fn in (("triplet", bench_triplet), ("batched", bench_batch), ("served", bench_serve)):
                t0 = time.time()
                result[key] = fn(args)
                log(f"[bench] {key} measured in {time.time()-t0:.1f} s")
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
