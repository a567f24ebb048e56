"""Throughput of the batched mass-spring step on one MI355X.

Workload (default): 512 voxel robots on a 9x9x9 lattice (8^3 voxels, 729
masses, 7448 springs each including air springs), materials drawn per voxel
from a seed with p(empty) = 0.2, land environment (gravity, ground, static /
kinetic friction), FP32, dt = 1e-4, one "step" = one 1 s evaluation of the
whole population (the reference's float time loop: 10000 substeps), launches
captured in a hipGraph.  Metric: spring-updates/s = sum over robots of all
springs (air included, as optimize/benchmark.cpp:104-109 counts them) x
substeps / wall seconds of the simulate calls.

    python bench.py [--steps K] [--warmup W] [--solver force|xpbd]

Prints one JSON line.  --gpus N with N > 1 exits: this build measures one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

import mercury_amd as mh  # noqa: E402

LAND = dict(gravity=9.81, ground_k=1.0e5, mu_s=1.0, mu_k=0.8, drag=0.0, damping=0.99)
WATER = dict(gravity=0.0, ground_k=0.0, mu_s=0.0, mu_k=0.0, drag=1.0, damping=0.99)
# streaming kernel with terms in global memory, HBM bytes per spring-update:
# record 9 B + term write 16 B + two gathers of (entry 4 B + term 16 B)
STREAM_BYTES_PER_SPRING = 65.0
# chunked streaming kernel (terms in LDS chunks), HBM bytes per
# spring-update: record 9 B + two u16 gather entries 4 B
CHUNK_HBM_BYTES_PER_SPRING = 13.0
# mass-parallel streaming kernel: each endpoint reads its record of the
# spring (other endpoint and material, rest length) every substep, 8 B, or
# 6 B in the compact format (mh_sim_info.stream_record_bytes)
MSTREAM_RECORD_ENDPOINTS = 2
# the spring record alone (two u16 endpoints, f32 rest length, u8 material):
# the least any kernel that streams the springs reads per spring-update
SPRING_RECORD_BYTES = 9.0
HBM_PEAK_GBS = 8000.0
# LDS bytes per spring-update of the LDS-resident kernel, per robot size:
# 57 B per spring (two endpoint positions 24 B, spring record 9 B, result
# write 12 B and gather read 12 B) + 48 B per mass (pos/vel read + write).
LDS_BYTES_PER_SPRING = 57.0
LDS_BYTES_PER_MASS = 48.0
# the force solver's stored term is {D, f}, 16 B instead of 12: 4 B more on
# the write and on the gather read
FORCE_TERM_EXTRA_BYTES = 8.0
# mass-parallel kernel: both endpoints evaluate a spring, each reading the
# other endpoint's position (12 B) and its (rest length, material) pair's
# actuated length and stiffness (8 B); per mass its own position read and
# write.  The gather flavour: the first endpoint reads 20 B and writes the
# 12 B term, the second endpoint reads the term.
MASS_LDS_BYTES_PER_SPRING = 40.0
MASSG_LDS_BYTES_PER_SPRING = 44.0
MASS_LDS_BYTES_PER_MASS = 32.0
# MI355X nominal LDS bandwidth: 256 CUs x 128 B/clk x 2.4 GHz
LDS_NOMINAL_GBS = 256 * 128 * 2.4
# VALU issue peak for plain (not packed) FP32 instructions: one wave64
# instruction per CU per cycle, 256 CUs at 2.4 GHz -- half the 157.3
# TFLOP/s FP32 vector peak, which counts packed FMA (2 lanes x 2 flops)
MI355X_CUS = 256
VALU_PEAK_WAVE_INSTR = MI355X_CUS * 2.4e9


# --kernel choices: MH_KERNEL_* family bits (robots over the LDS budget may
# always stream)
KERNEL_FAMILIES = {"auto": 0, "mass": 4 | 2 | 16, "mass-gather": 8 | 2 | 16, "lds": 1 | 2 | 16, "stream": 2,
                   "mass-stream": 16}


def kernel_label(info):
    return ("+".join(n for bit, n in ((8, "mass-gather"), (4, "mass"), (1, "lds"), (16, "mass-stream"),
                                      (2, "stream"))
                     if info.kernel & bit) + (" (chunked)" if info.stream_chunk > 0 else ""))


def float_loop_steps(duration, dt):
    r, n = np.float32(duration), 0
    while r > np.float32(0):
        r = np.float32(r - np.float32(dt))
        n += 1
    return n


def cpu_threads():
    """Host threads for the CPU baseline: every core in the process's
    affinity mask, unless OMP_NUM_THREADS sets the host's share lower (a
    GPU host that allots a fixed number of cores to each GPU sets it)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    cap = os.environ.get("OMP_NUM_THREADS")
    if cap and cap.isdigit() and int(cap) > 0:
        n = min(n, int(cap))
    return max(1, n)


def host_description():
    """nproc, the affinity mask's size, the thread cap and the CPU model."""
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = None
    return {"nproc": os.cpu_count(), "affinity_cores": affinity,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS"), "cpu_model": model}


def cpu_baseline(els, solver, dt, steps, threads, env_kw, f64=False):
    """The oracle's FP32 mirror (plain C, -O3, no fast-math) on host cores;
    f64: its FP64 twin instead."""
    import oracle
    from mercury_amd.simulator import pack
    a = pack(els)
    by_id, _, _ = oracle.composite_tables(oracle.make_base(mh.materials.WORKLOAD_ROWS))
    mt = oracle.MaterialTable(by_id["k"], by_id["dL0"], by_id["omega"], by_id["phi"])
    pop = oracle.Population(a["mass_offset"], a["spring_offset"], a["mass"], a["pairs"],
                            a["mean_length"], a["spring_material"])
    env = oracle.make_env(dt, **env_kw)

    def run(n, th):
        t = time.perf_counter()
        if solver == mh.MH_SOLVER_FORCE:
            fn = oracle.force_f64 if f64 else oracle.force_f32
            fn(pop, mt, env, 0.0, n, a["pos"], a["vel"], nthreads=th)
        else:
            fn = oracle.xpbd_f64 if f64 else oracle.xpbd_f32
            fn(pop, mt, dt, 0.0, n, a["pos"], a["vel"], nthreads=th)
        return time.perf_counter() - t

    run(max(1, steps // 20), threads)  # warm-up
    secs = run(steps, threads)
    springs = len(a["mean_length"])
    return springs * steps / secs, secs


def measured_counters(workload, kernel):
    """Per-launch hardware counts of the step kernel from the committed
    rocprofv3 summary (profiles/counters.json): VALU and MFMA wave
    instructions, LDS bank-conflict cycles, for this workload and kernel."""
    path = os.path.join(REPO, "profiles", "counters.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh).get(f"{workload}/{kernel}")


def measured_traffic(workload, kernel, substeps_per_launch):
    """HBM bytes per step-kernel launch from the committed rocprofv3 counter
    summary (FETCH_SIZE + WRITE_SIZE per dispatch, each scaled by what the
    counter reads for a copy kernel of known size) for this workload and
    kernel, if one exists.  Kernels that stream their records every substep
    have an entry per substep, scaled to this launch's substeps; the others
    read and write their state once per launch."""
    path = os.path.join(REPO, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        entry = json.load(fh).get(f"{workload}/{kernel}")
    if not entry:
        return None
    per_launch = (entry["bytes_per_substep"] * substeps_per_launch if "bytes_per_substep" in entry
                  else entry["bytes_per_launch"])
    return {"bytes_per_launch": per_launch, "source": "profiles/traffic.json: " + entry["summary"]}


def roofline_head(bound, achieved, peak, unit, extra):
    """bound / achieved / peak / unit / frac of the dominant kernel.  With
    hardware counters for it, two rooflines are compared: VALU issue (wave64
    VALU instructions per second against one per CU-cycle at 2.4 GHz;
    independent FMA chains sustain 0.92 of that on MI355X) and the kernel's
    byte model (LDS or HBM).  The bound is the one the kernel reaches the
    larger fraction of; the other is reported beside it.  frac_survey is
    the fraction on the algorithmic LDS model of 57 B per spring plus 48 B
    per mass per substep."""
    v = extra.get("valu_issue")
    byte_frac = (achieved / peak) if achieved else None
    if v and (byte_frac is None or v["frac"] > byte_frac):
        out = {"bound": "valu", "achieved": v["achieved"], "peak": v["peak"], "unit": v["unit"],
               "frac": v["frac"], "byte_model_bound": bound, "byte_model_achieved": achieved,
               "byte_model_peak": peak, "byte_model_unit": unit, "byte_model_frac": byte_frac}
    else:
        out = {"bound": bound, "achieved": achieved, "peak": peak, "unit": unit, "frac": byte_frac}
        if v:
            out.update({"valu_achieved": v["achieved"], "valu_peak": v["peak"], "valu_frac": v["frac"]})
    if "survey_model" in extra:
        out["frac_survey"] = extra["survey_model"]["frac"]
    return out


BF16_DENSE_PEAK = 2.5e15  # MI355X dense bf16 matrix FLOP/s


def surrogate_epoch():
    """One training epoch of each surrogate on synthetic data: the MLP
    64-256-256-1 on 200 000 samples (batch 512) and the convolutional
    network on 20 000 genomes of 8^3 voxels (batch 256).  GEMM FLOPs per
    sample: forward 2 n_in n_out per layer, the weight gradient the same,
    the input gradient for every layer but the first."""
    from mercury_amd import evolve, surrogate
    rng = np.random.default_rng(0)
    out = {}
    sizes = [64, 256, 256, 1]
    mlp = surrogate.MLP(sizes, seed=1)
    n, batch = 200000, 512
    X = rng.normal(size=(n, 64)).astype(np.float32)
    y = rng.normal(size=(n, 1)).astype(np.float32)
    mlp.train(X[:4096], y[:4096], epochs=1, batch=batch)  # warm-up
    t = time.perf_counter()
    mlp.train(X, y, epochs=1, batch=batch)
    sec = time.perf_counter() - t
    layer = [2 * a * b for a, b in zip(sizes[:-1], sizes[1:])]
    flop = (n // batch) * batch * (2 * sum(layer) + sum(layer[1:]))
    out["mlp"] = {"samples": n, "batch": batch, "seconds": sec, "gemm_flop": flop,
                  "tflops_wall": flop / sec / 1e12, "frac_bf16_peak_wall": flop / sec / BF16_DENSE_PEAK}
    cnn = surrogate.CNN(seed=1)
    n, batch = 20000, 256
    G = evolve.random_genomes(rng, n, 8)
    t_c = rng.normal(size=n).astype(np.float32)
    cnn.train(G[:512], t_c[:512], epochs=1, batch=batch)
    t = time.perf_counter()
    cnn.train(G, t_c, epochs=1, batch=batch)
    sec = time.perf_counter() - t
    conv = [2 * 512 * 135 * 32, 2 * 512 * 864 * 32, 2 * 64 * 864 * 64]
    dense = [2 * 64 * 256, 2 * 256 * 256, 2 * 256 * 1]
    per = 2 * sum(conv) + sum(conv[1:]) + 2 * sum(dense) + sum(dense)
    flop = (n // batch) * batch * per
    out["cnn"] = {"genomes": n, "batch": batch, "seconds": sec, "gemm_flop": flop,
                  "tflops_wall": flop / sec / 1e12, "frac_bf16_peak_wall": flop / sec / BF16_DENSE_PEAK}
    print(json.dumps({"surrogate_epoch": out, "dtype": "bf16 operands, f32 accumulate",
                      "peak_flops": BF16_DENSE_PEAK}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="timed 1 s evaluations")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["voxel", "nn"], default="voxel",
                    help="voxel: 512 8^3 voxel robots on land; nn: 1024 uneven KNN robots in water")
    ap.add_argument("--robots", type=int, default=0, help="0 = the workload's count")
    ap.add_argument("--voxels", type=int, default=8)
    ap.add_argument("--duration", type=float, default=1.0)
    ap.add_argument("--dt", type=float, default=1e-4)
    ap.add_argument("--solver", choices=["force", "xpbd"], default="force")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--substeps", type=int, default=0, help="substeps per launch (0 = tuned)")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--kernel", choices=list(KERNEL_FAMILIES), default="auto",
                    help="restrict LDS-resident robots to one kernel family (larger robots stream)")
    ap.add_argument("--cpu-steps", type=int, default=2000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-xpbd", action="store_true", help="skip the XPBD side measurement")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run the known-size copy kernel (counter calibration)")
    ap.add_argument("--calibrate-valu", action="store_true",
                    help="only run the VALU issue-rate kernels (FMA, MUL and reciprocal chains)")
    ap.add_argument("--surrogate-epoch", action="store_true",
                    help="only time one training epoch of each surrogate (matrix-core GEMMs)")
    args = ap.parse_args()

    if args.calibrate_copy:
        import ctypes
        ms = ctypes.c_float()
        mh._lib.check(mh.load().mh_calibrate_copy(0, args.calibrate_copy, 5, ctypes.byref(ms)))
        print(json.dumps({"copy_bytes": args.calibrate_copy, "avg_ms": ms.value,
                          "GB/s": 2 * args.calibrate_copy / (ms.value * 1e-3) / 1e9}))
        return

    if args.calibrate_valu:
        import ctypes
        rates = {}
        for kind, name in enumerate(("fma", "mul", "fma_16_chains", "rcp", "rcp_fma")):
            g = ctypes.c_float()
            mh._lib.check(mh.load().mh_calibrate_valu(0, kind, ctypes.byref(g)))
            rates[name] = g.value * 1e9
        print(json.dumps({"valu_wave_instr_per_s": rates}))
        return

    if args.surrogate_epoch:
        surrogate_epoch()
        return

    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported by this build "
              "(robots are independent: a multi-GPU run shards the population by robot).",
              file=sys.stderr)
        sys.exit(2)

    solver = mh.MH_SOLVER_FORCE if args.solver == "force" else mh.MH_SOLVER_XPBD_REF
    ms = mh.workload_materials()
    t_build = time.perf_counter()
    if args.workload == "voxel":
        args.robots = args.robots or 512
        env_kw = LAND
        els = mh.voxel_population(ms, args.robots, voxels_per_axis=args.voxels, p_empty=0.2,
                                  seed=args.seed, mass=0.001)
        metric = f"spring-updates/s ({args.robots}x{args.voxels}^3 voxel robots)"
        workload = (f"{args.robots} voxel robots, {args.voxels}^3 voxels ({els[0].n_masses} masses, "
                    f"{els[0].n_springs} springs each), {args.solver} solver, land, dt={args.dt}, "
                    f"{args.duration} s evaluation")
    else:
        args.robots = args.robots or 1024
        env_kw = WATER
        els = mh.robots.nn_population(ms, args.robots, mass_range=(200, 2000), K=25,
                                      seed=args.seed, mass=0.01)
        metric = f"spring-updates/s ({args.robots} uneven NN robots, water)"
        workload = (f"{args.robots} NN robots, masses U[200,2000], K=25 nearest-neighbour springs, "
                    f"MLP 3-25-25-4 materials, {args.solver} solver, water drag 1.0, dt={args.dt}, "
                    f"{args.duration} s evaluation")
    t_build = time.perf_counter() - t_build
    n_springs = sum(e.n_springs for e in els)
    n_masses = sum(e.n_masses for e in els)
    n_active = int(sum((e.spring_material != 0).sum() for e in els))
    n_sub = float_loop_steps(args.duration, args.dt)

    def make_sim(slv):
        cfg = mh.SimConfig(solver=slv, time_step=args.dt, substeps_per_launch=args.substeps,
                           use_graph=1, kernel=KERNEL_FAMILIES[args.kernel], **env_kw)
        sim = mh.Simulator(ms, cfg)
        t = time.perf_counter()
        sim.SetElements(els)
        t_up = time.perf_counter() - t
        if not args.no_autotune:
            sim.autotune(2000)  # long enough for sustained clocks
        return sim, t_up

    def timed(sim, k, warm):
        for _ in range(warm):
            sim.Reset()
            sim.Simulate(args.duration)
        sim.synchronize()
        sim.set_timing(True)
        t = time.perf_counter()
        for _ in range(k):
            sim.Reset()
            sim.Simulate(args.duration)
        sim.synchronize()
        secs = time.perf_counter() - t
        kms, launches = sim.last_kernel_ms()
        sim.set_timing(False)
        return secs, kms, launches

    sim, t_upload = make_sim(solver)
    secs, kernel_ms, launches = timed(sim, args.steps, args.warmup)
    info = sim.info()
    updates = n_springs * n_sub * args.steps
    value = updates / secs

    # roofline of the dominant (step) kernel from its average launch time;
    # launches counts the launches of one simulate call (one step)
    import ctypes
    lds_peak = ctypes.c_float()
    mh._lib.check(mh.load().mh_calibrate_lds(0, ctypes.byref(lds_peak)))
    lds_peak = float(lds_peak.value)
    steps_per_launch = n_sub / max(launches, 1)
    hbm_extra = None
    if info.kernel & mh.MH_KERNEL_MASS_STREAM and info.n_stream_springs > 0:
        # the streamed robots carry the work: HBM bytes of their records
        bound, peak, unit = "hbm", HBM_PEAK_GBS, "GB/s"
        per_unit = MSTREAM_RECORD_ENDPOINTS * float(info.stream_record_bytes)
        per_launch = per_unit * info.n_stream_springs * steps_per_launch
        if kernel_ms > 0:
            rec = SPRING_RECORD_BYTES * info.n_stream_springs * steps_per_launch / (kernel_ms * 1e-3) / 1e9
            hbm_extra = {"record_model": {"bytes_per_spring_update": SPRING_RECORD_BYTES,
                                          "achieved": rec, "frac": rec / HBM_PEAK_GBS}}
    elif info.kernel & (mh.MH_KERNEL_MASS | mh.MH_KERNEL_MASS_GATHER) and not info.kernel & mh.MH_KERNEL_STREAM:
        bound, peak, unit = "lds", lds_peak, "GB/s"
        bps = MASSG_LDS_BYTES_PER_SPRING if info.kernel & mh.MH_KERNEL_MASS_GATHER else MASS_LDS_BYTES_PER_SPRING
        per_launch = (bps * n_springs + MASS_LDS_BYTES_PER_MASS * n_masses) * steps_per_launch
        per_unit = bps + MASS_LDS_BYTES_PER_MASS * n_masses / n_springs
    elif info.kernel == 1 or info.stream_chunk > 0:
        # LDS-resident kernel, or chunked streaming kernel whose terms also
        # live in LDS: LDS bytes over all springs
        bound, peak, unit = "lds", lds_peak, "GB/s"
        bps = LDS_BYTES_PER_SPRING + (FORCE_TERM_EXTRA_BYTES if solver == mh.MH_SOLVER_FORCE else 0.0)
        per_launch = (bps * n_springs + LDS_BYTES_PER_MASS * n_masses) * steps_per_launch
        per_unit = bps + LDS_BYTES_PER_MASS * n_masses / n_springs
        if info.stream_chunk > 0 and kernel_ms > 0:
            hb = CHUNK_HBM_BYTES_PER_SPRING * info.n_stream_springs * steps_per_launch
            hbm_extra = {"hbm_bytes_per_spring_update": CHUNK_HBM_BYTES_PER_SPRING,
                         "hbm_achieved_gbs": hb / (kernel_ms * 1e-3) / 1e9,
                         "hbm_frac": hb / (kernel_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                         "stream_chunk_springs": info.stream_chunk}
    else:  # the global-memory streaming kernel carries the big robots
        bound, peak, unit = "hbm", HBM_PEAK_GBS, "GB/s"
        per_launch = STREAM_BYTES_PER_SPRING * info.n_stream_springs * steps_per_launch
        per_unit = STREAM_BYTES_PER_SPRING
    achieved = per_launch / (kernel_ms * 1e-3) / 1e9 if kernel_ms > 0 else None
    kernel_name = kernel_label(info)
    # the dominant kernel's family names its committed profiles
    dominant = "mass-stream" if info.kernel & mh.MH_KERNEL_MASS_STREAM and info.n_stream_springs > 0 else kernel_name
    traffic = measured_traffic(args.workload, dominant, steps_per_launch)
    counters = measured_counters(args.workload, dominant)
    extra = {}
    if bound == "lds" and kernel_ms > 0:
        # the algorithmic LDS model: 57 B per spring (two endpoint
        # positions, the 9 B record, one term write and one gather read)
        # + 48 B per mass, whatever the kernel reads
        survey = (LDS_BYTES_PER_SPRING * n_springs + LDS_BYTES_PER_MASS * n_masses) * steps_per_launch
        extra["survey_model"] = {
            "bytes_per_spring_update": LDS_BYTES_PER_SPRING + LDS_BYTES_PER_MASS * n_masses / n_springs,
            "achieved": survey / (kernel_ms * 1e-3) / 1e9,
            "frac": survey / (kernel_ms * 1e-3) / 1e9 / lds_peak,
            "frac_nominal_peak": survey / (kernel_ms * 1e-3) / 1e9 / LDS_NOMINAL_GBS}
    if counters and kernel_ms > 0:
        # counters of a profiled launch (profiles/counters.json), scaled to
        # this launch's substeps; gpu_cycles_per_launch is GRBM_GUI_ACTIVE
        # per XCD, so instructions per CU-cycle do not depend on the clock
        scale = steps_per_launch / counters["substeps_per_launch"]
        valu_launch = counters["valu_per_launch"] * scale
        valu = valu_launch / (kernel_ms * 1e-3)
        per_cu_cycle = counters["valu_per_launch"] / (MI355X_CUS * counters["gpu_cycles_per_launch"])
        extra["valu_issue"] = {
            "wave_instr_per_launch": valu_launch,
            "achieved": valu, "peak": VALU_PEAK_WAVE_INSTR, "unit": "wave64 instr/s",
            "frac": valu / VALU_PEAK_WAVE_INSTR,
            "per_cu_cycle": per_cu_cycle,
            # the clock the profiled cycles imply for this launch's time
            "clock_ghz": counters["gpu_cycles_per_launch"] * scale / (kernel_ms * 1e-3) / 1e9,
            # issue cycles: a transcendental holds the VALU two cycles
            # (SQ_ACTIVE_INST_VALU = SQ_INSTS_VALU + SQ_INSTS_VALU_TRANS_F32)
            "busy_per_cu_cycle": (counters["active_valu_per_launch"] /
                                  (MI355X_CUS * counters["gpu_cycles_per_launch"])
                                  if counters.get("active_valu_per_launch") else None),
            "fma_chains_per_cu_cycle": counters.get("fma_chains_per_cu_cycle"),
            "transcendental_share": counters.get("trans_per_launch", 0) / counters["valu_per_launch"],
            "lane_instr_per_active_spring_update": valu_launch * 64 / (n_active * steps_per_launch)}
        extra["mfma_per_launch"] = counters.get("mfma_per_launch")
        extra["mfma_busy_cycles_per_launch"] = counters.get("mfma_busy_cycles_per_launch")
        extra["counters_source"] = "profiles/counters.json: " + counters["summary"]

    out = {
        "metric": metric,
        "value": value,
        "unit": "spring-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": secs / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": ("synthetic (seeded voxel materials)" if args.workload == "voxel" else
                 "synthetic (seeded NN robots: random positions, random-init material MLP)"),
        "config": {
            "workload": workload,
            "robots": args.robots, "springs_total": n_springs, "active_springs": n_active,
            "masses_total": n_masses, "substeps_per_eval": n_sub, "solver": args.solver,
            "kernel": kernel_name, "block": info.block_size,
            "blocks": info.n_blocks, "lds_bytes_per_block": info.lds_bytes,
            "substeps_per_launch": info.substeps_per_launch, "hipgraph": True,
            "lds_robots": info.n_lds_robots, "stream_robots": info.n_stream_robots,
            "robots_per_block": info.robots_per_block,
        },
        "active_spring_updates_per_s": n_active * n_sub * args.steps / secs,
        "roofline": {
            **roofline_head(bound, achieved, peak, unit, extra),
            "traffic": traffic["bytes_per_launch"] if traffic else None,
            **({"traffic_source": traffic["source"]} if traffic else {}),
            "peak_kind": ("measured ds_read_b128 bandwidth (mh_calibrate_lds); nominal "
                          f"{LDS_NOMINAL_GBS:.0f}") if bound == "lds" else "MI355X HBM3E",
            "kernel_ms_per_launch": kernel_ms,
            "launches_per_step": launches,
            "algorithmic_bytes_per_spring_update": per_unit,
            **({"streamed_part": hbm_extra} if hbm_extra else {}),
            **extra,
        },
        "host_seconds": {"topology_build": t_build, "upload": t_upload},
    }
    if not args.no_xpbd and solver == mh.MH_SOLVER_FORCE and args.workload == "voxel":
        sx, _ = make_sim(mh.MH_SOLVER_XPBD_REF)
        xs, xkms, _ = timed(sx, 1, 1)
        out["xpbd_ref_spring_updates_per_s"] = n_springs * n_sub / xs
        sx.close()
    if not args.no_cpu:
        th = cpu_threads()
        v_all, s_all = cpu_baseline(els, solver, args.dt, args.cpu_steps, th, env_kw)
        sub = els[:32]
        v_one, s_one = cpu_baseline(sub, solver, args.dt, max(50, args.cpu_steps // 8), 1, env_kw)
        v64, _ = cpu_baseline(sub, solver, args.dt, max(50, args.cpu_steps // 8), th, env_kw, f64=True)
        out["cpu_baseline"] = {
            "value": v_all, "unit": "spring-updates/s", "cores": th, "kind": "port",
            "host": host_description(), "per_core_value": v_all / th,
            "sample": f"oracle FP32 mirror, same {args.robots} robots, {args.cpu_steps} substeps "
                      f"({s_all:.1f} s wall on {th} threads)",
            "single_core_value": v_one,
            "single_core_sample": f"32 robots, {max(50, args.cpu_steps // 8)} substeps, 1 thread",
            "fp64_oracle_value": v64,
            "fp64_oracle_sample": f"FP64 oracle, 32 robots, {max(50, args.cpu_steps // 8)} substeps, "
                                  f"{th} threads",
        }
        out["gpu_vs_cpu"] = value / v_all
    print(json.dumps(out))


if __name__ == "__main__":
    main()
