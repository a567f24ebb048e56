"""GPU parity of the step kernels against the oracle (through the C-ABI).

Bar: bit-exact against the FP32 mirror (oracle/step.c) for positions,
velocities and stresses; FP64 oracle agreement for the force solver within
1e-5 relative over 1000 steps; identical rankings.
"""
import numpy as np
import pytest

import oracle
import mercury_amd as mh
from helpers import (LAND, WATER, bits, gpu_sim, mirror, reference_case, rel_err,
                     uneven_case, workload_case)

pytestmark = pytest.mark.gpu

XPBD, FORCE = mh.MH_SOLVER_XPBD_REF, mh.MH_SOLVER_FORCE


def run_gpu(case, solver, n, dt, env_kw=None, stress=False, **cfg):
    sim = gpu_sim(case, solver, dt, env_kw, **cfg)
    sim.Step(n, trackStresses=stress)
    pos, vel, st = sim.collect_arrays()
    return sim, pos, vel, st


@pytest.mark.parametrize("dt", [1e-3, 5e-3])
def test_xpbd_reference_materials_bitexact(dt):
    print(f"xpbd reference materials dt={dt}", flush=True)
    case = reference_case(n_robots=6, voxels=4, seed=11)
    _, pos, vel, st = run_gpu(case, XPBD, 300, dt, stress=True)
    p, v, s = mirror(case, XPBD, 300, dt, stress=True)
    assert np.array_equal(bits(pos), bits(p))
    assert np.array_equal(bits(vel), bits(v))
    assert np.array_equal(bits(st), bits(s))
    assert np.abs(pos - case.pos0).max() > 1e-4  # robots actually moved


def test_xpbd_config2_robot_bitexact():
    print("xpbd 8^3 workload robots", flush=True)
    case = workload_case(n_robots=3, voxels=8, seed=5)
    _, pos, vel, st = run_gpu(case, XPBD, 200, 1e-4, stress=True)
    p, v, s = mirror(case, XPBD, 200, 1e-4, stress=True)
    assert np.array_equal(bits(pos), bits(p))
    assert np.array_equal(bits(vel), bits(v))
    assert np.array_equal(bits(st), bits(s))


@pytest.mark.parametrize("env", ["land", "water"])
def test_force_bitexact(env):
    print(f"force {env}", flush=True)
    kw = LAND if env == "land" else WATER
    case = workload_case(n_robots=4, voxels=8, seed=21)
    _, pos, vel, _ = run_gpu(case, FORCE, 500, 1e-4, kw)
    p, v, _ = mirror(case, FORCE, 500, 1e-4, kw)
    assert np.array_equal(bits(pos), bits(p))
    assert np.array_equal(bits(vel), bits(v))
    assert np.abs(pos - case.pos0).max() > 1e-5


def test_uneven_population_bitexact():
    print("uneven population", flush=True)
    case = uneven_case()
    _, pos, vel, _ = run_gpu(case, FORCE, 300, 1e-4, LAND)
    p, v, _ = mirror(case, FORCE, 300, 1e-4, LAND)
    assert np.array_equal(bits(pos), bits(p))
    assert np.array_equal(bits(vel), bits(v))


@pytest.mark.parametrize("cfg", [dict(substeps_per_launch=1), dict(substeps_per_launch=7),
                                 dict(use_graph=0, substeps_per_launch=13),
                                 dict(block_size=256), dict(block_size=512),
                                 dict(robots_per_block=1), dict(kernel=mh.MH_KERNEL_STREAM),
                                 dict(kernel=mh.MH_KERNEL_STREAM, block_size=256),
                                 dict(kernel=mh.MH_KERNEL_STREAM, block_size=512),
                                 dict(kernel=mh.MH_KERNEL_LDS | mh.MH_KERNEL_STREAM),
                                 dict(kernel=mh.MH_KERNEL_LDS | mh.MH_KERNEL_STREAM, block_size=1024),
                                 dict(kernel=mh.MH_KERNEL_MASS | mh.MH_KERNEL_STREAM, robots_per_block=1),
                                 dict(kernel=mh.MH_KERNEL_MASS | mh.MH_KERNEL_STREAM, block_size=1024),
                                 dict(kernel=mh.MH_KERNEL_MASS_GATHER | mh.MH_KERNEL_STREAM),
                                 dict(kernel=mh.MH_KERNEL_MASS_GATHER | mh.MH_KERNEL_STREAM,
                                      robots_per_block=1),
                                 dict(kernel=mh.MH_KERNEL_MASS_GATHER | mh.MH_KERNEL_STREAM, block_size=256),
                                 dict(block_size=768)],
                         ids=lambda c: "-".join(f"{k}{v}" for k, v in c.items()))
def test_launch_plan_does_not_change_results(cfg):
    print(f"launch plan {cfg}", flush=True)
    case = uneven_case(seed=9)
    _, pos_a, vel_a, _ = run_gpu(case, FORCE, 120, 1e-4, LAND)
    _, pos_b, vel_b, _ = run_gpu(case, FORCE, 120, 1e-4, LAND, **cfg)
    assert np.array_equal(bits(pos_a), bits(pos_b))
    assert np.array_equal(bits(vel_a), bits(vel_b))
    case2 = reference_case(n_robots=5, voxels=3, seed=2)
    _, pa, va, sa = run_gpu(case2, XPBD, 60, 1e-3, stress=True)
    _, pb, vb, sb = run_gpu(case2, XPBD, 60, 1e-3, stress=True, **cfg)
    assert np.array_equal(bits(pa), bits(pb))
    assert np.array_equal(bits(sa), bits(sb))


@pytest.mark.parametrize("cfg", [dict(kernel=mh.MH_KERNEL_MASS_STREAM),
                                 dict(kernel=mh.MH_KERNEL_MASS_STREAM, block_size=512, substeps_per_launch=7),
                                 dict(kernel=mh.MH_KERNEL_MASS | mh.MH_KERNEL_MASS_STREAM, use_graph=0)],
                         ids=lambda c: "-".join(f"{k}{v}" for k, v in c.items()))
def test_mass_stream_kernel_on_land_bitexact(cfg):
    """The mass-parallel streaming kernel serves the force solver only: on
    an uneven voxel batch on land (ground contact, friction) it gives the
    default plan's trajectory bit for bit."""
    print(f"mass-parallel streaming kernel, land, {cfg}", flush=True)
    case = uneven_case(seed=9)
    _, pos_a, vel_a, _ = run_gpu(case, FORCE, 120, 1e-4, LAND)
    sim, pos_b, vel_b, _ = run_gpu(case, FORCE, 120, 1e-4, LAND, **cfg)
    assert sim.info().kernel & mh.MH_KERNEL_MASS_STREAM or cfg["kernel"] != mh.MH_KERNEL_MASS_STREAM
    assert np.array_equal(bits(pos_a), bits(pos_b))
    assert np.array_equal(bits(vel_a), bits(vel_b))


def test_graph_replays_match_eager():
    print("graph replay", flush=True)
    case = workload_case(n_robots=5, voxels=4, seed=4)
    out = []
    for use_graph in (1, 0):
        sim = gpu_sim(case, FORCE, 1e-4, LAND, use_graph=use_graph, substeps_per_launch=16)
        traj = []
        for rep in range(3):
            sim.Reset()  # same t0 each call -> the cached graph is replayed
            sim.Simulate(0.005)
            sim.synchronize()
            traj.append(sim.collect_arrays()[0])
        out.append(traj)
    for a, b in zip(*out):
        assert np.array_equal(bits(a), bits(b))
    assert not np.array_equal(bits(out[0][0]), bits(out[0][1]))  # state advanced


def float_loop_steps(duration, dt):
    """Simulator.cu:428-449: while (remaining > 0) remaining -= dt (float)."""
    r, n = np.float32(duration), 0
    while r > np.float32(0):
        r = np.float32(r - np.float32(dt))
        n += 1
    return n


def test_simulate_follows_float_time_loop():
    print("time loop", flush=True)
    case = workload_case(n_robots=2, voxels=2, seed=1)
    sim = gpu_sim(case, FORCE, 1e-4, LAND)
    total, t = 0, np.float32(0)
    for dur in (0.0123, 0.05):
        n = float_loop_steps(dur, 1e-4)
        assert n == len(oracle.step_times(np.float32(0), np.float32(1e-4), np.float32(dur)))
        sim.Simulate(dur)
        total += n
        for _ in range(n):
            t = np.float32(t + np.float32(1e-4))
        assert np.float32(sim.getTotalTime()) == t
    sim.synchronize()
    p, v, _ = mirror(case, FORCE, total, 1e-4, LAND)
    assert np.array_equal(bits(sim.collect_arrays()[0]), bits(p))


def test_fp64_oracle_tolerance_force_1k_steps():
    print("fp64 tolerance force", flush=True)
    for kw in (LAND, WATER):
        case = workload_case(n_robots=6, voxels=8, seed=33)
        _, pos, _, _ = run_gpu(case, FORCE, 1000, 1e-4, kw)
        p64, _, _ = mirror(case, FORCE, 1000, 1e-4, kw, f64=True)
        err = rel_err(pos, p64, case.a["mass_offset"])
        assert err.max() <= 1e-5, err


def test_fp64_oracle_xpbd_error_bound():
    """XPBD's v = 0.99 (x' - x)/dt amplifies FP32 position rounding by 1/dt,
    so the FP32 step (the reference's own arithmetic) drifts from FP64 by an
    amount that shrinks with dt; the bound asserted is the one measured for
    the FP32 mirror (dt = 5e-3: 3.6e-5 over 1k steps)."""
    print("fp64 bound xpbd", flush=True)
    case = reference_case(n_robots=4, voxels=8, seed=2)
    _, pos, _, _ = run_gpu(case, XPBD, 1000, 5e-3)
    p64, _, _ = mirror(case, XPBD, 1000, 5e-3, f64=True)
    assert rel_err(pos, p64, case.a["mass_offset"]).max() <= 1e-4


def test_fitness_ranking_matches_fp64():
    print("ranking", flush=True)
    case = workload_case(n_robots=24, voxels=8, seed=77)
    sim = gpu_sim(case, FORCE, 1e-4, LAND)
    sim.Step(2000)
    disp = sim.fitness(mh.MH_FITNESS_DISPLACEMENT)
    fit = sim.fitness(mh.MH_FITNESS_REFERENCE)
    p64, _, _ = mirror(case, FORCE, 2000, 1e-4, LAND, f64=True)
    mo = case.a["mass_offset"]
    d64 = []
    for r in range(len(mo) - 1):
        air = case.a["mass_material"][mo[r]:mo[r + 1]] == 0
        c = p64[mo[r]:mo[r + 1]][~air].mean(0) if (~air).any() else np.zeros(3)
        c0 = case.a["pos"][mo[r]:mo[r + 1]][~air].astype(np.float64).mean(0) if (~air).any() else np.zeros(3)
        d64.append(c[0] - c0[0])
    d64 = np.array(d64)
    gaps = np.diff(np.sort(d64))
    print("min gap", gaps.min(), "max |disp32 - disp64|", np.abs(disp - d64).max(), flush=True)
    assert np.array_equal(np.argsort(disp, kind="stable"), np.argsort(d64, kind="stable"))
    assert np.array_equal(fit, np.maximum(disp / 1.0, 0).astype(np.float32))


def test_device_com_matches_oracle():
    print("device com", flush=True)
    case = workload_case(n_robots=7, voxels=5, seed=8)
    sim = gpu_sim(case, FORCE, 1e-4, LAND)
    sim.Step(50)
    com = sim.com()
    pos = sim.collect_arrays()[0]
    mo = case.a["mass_offset"]
    for r in range(len(mo) - 1):
        air = (case.a["mass_material"][mo[r]:mo[r + 1]] == 0).astype(np.uint8)
        want = oracle.com(pos[mo[r]:mo[r + 1]], air)
        assert np.array_equal(bits(com[r]), bits(want))


def test_edge_cases():
    print("edge cases", flush=True)
    ms = mh.workload_materials()
    # empty population: simulate is a no-op that advances time
    sim = mh.Simulator(ms, mh.SimConfig(solver=FORCE))
    sim.SetElements([])
    sim.Simulate(0.001)
    assert sim.getTotalTime() > 0
    # an all-air robot next to a normal one
    air = mh.voxel_build(ms, 3, 10.0, np.ones(27, np.uint32), mass=0.001)
    ok = mh.voxel_build(ms, 3, 10.0, np.full(27, 2, np.uint32), mass=0.001)
    from helpers import Case
    case = Case([air, ok, air], ms, mh.materials.WORKLOAD_ROWS)
    _, pos, vel, _ = run_gpu(case, FORCE, 100, 1e-4, LAND)
    p, v, _ = mirror(case, FORCE, 100, 1e-4, LAND)
    assert np.array_equal(bits(pos), bits(p))
    # invalid inputs are rejected with an error, not a fault
    bad = ok.copy()
    bad.pairs = bad.pairs.copy()
    bad.pairs[0, 1] = 60000
    with pytest.raises(mh.MhError):
        mh.Simulator(ms, mh.SimConfig(solver=FORCE)).SetElements([bad])
    bad2 = ok.copy()
    bad2.mass = np.zeros(27, np.float32)
    with pytest.raises(mh.MhError):
        mh.Simulator(ms, mh.SimConfig(solver=FORCE)).SetElements([bad2])


def ring_robot(ms, rng, n, k, spacing=0.01):
    """n masses on a jittered grid, spring (i, i + j) for j = 1..k: every
    spring has its own rest length, so the robot has n*k distinct (rest
    length, material) pairs."""
    from mercury_amd.robots import Element
    side = int(np.ceil(n ** (1 / 3)))
    g = np.stack(np.unravel_index(np.arange(n), (side, side, side)), 1).astype(np.float32) * spacing
    pos = (g + rng.uniform(-0.2, 0.2, g.shape).astype(np.float32) * spacing + np.float32(0.002)).astype(np.float32)
    vox = mh.voxel_build(ms, 3, 10.0, np.full(27, 2, np.uint32), mass=0.001)
    mats = np.unique(vox.spring_material[vox.spring_material != 0])
    pairs = np.array([(i, (i + j) % n) for i in range(n) for j in range(1, k + 1)], np.uint16)
    lengths = mh.robots.rest_lengths(pos, pairs) * rng.uniform(0.9, 1.1, len(pairs)).astype(np.float32)
    return Element(pos, np.full(n, vox.mass_material[0], np.uint8), pairs, lengths.astype(np.float32),
                   rng.choice(mats, len(pairs)).astype(np.uint8), mass=np.full(n, 0.001, np.float32))


@pytest.mark.parametrize("family,solver", [(mh.MH_KERNEL_MASS, FORCE), (mh.MH_KERNEL_MASS, XPBD),
                                           (mh.MH_KERNEL_MASS_GATHER, XPBD)])
def test_mass_pair_table_limits(family, solver):
    """Robots of 160 distinct (rest length, material) pairs: packs of the
    mass-parallel kernels close when the union of their pairs would pass
    511, and a robot with 600 pairs goes to the streaming kernel."""
    print(f"pair table limits family={family} solver={solver}", flush=True)
    from helpers import Case
    ms = mh.workload_materials()
    rng = np.random.default_rng(17)
    els = [ring_robot(ms, rng, 40, 4) for _ in range(9)]
    els.insert(4, ring_robot(ms, rng, 150, 4))
    case = Case(els, ms, mh.materials.WORKLOAD_ROWS)
    dt = 1e-4 if solver == FORCE else 1e-3
    kw = LAND if solver == FORCE else None
    sim, pos, vel, st = run_gpu(case, solver, 200, dt, kw, stress=solver == XPBD,
                                kernel=family | mh.MH_KERNEL_STREAM, block_size=512)
    info = sim.info()
    assert info.n_stream_robots == 1 and info.n_lds_robots == 9
    # packs take consecutive robots, at most 3 of 160 pairs: {0, 1, 2}, {3}
    # (the 600-pair robot ends the run), {5, 6, 7}, {8, 9}; plus the
    # 600-pair robot's streaming workgroup
    assert info.n_blocks == 4 + 1
    p, v, s = mirror(case, solver, 200, dt, kw, stress=solver == XPBD)
    assert np.array_equal(bits(pos), bits(p))
    assert np.array_equal(bits(vel), bits(v))
    if solver == XPBD:
        assert np.array_equal(bits(st), bits(s))
    assert np.abs(pos - case.pos0).max() > 1e-6


def grid_robot(ms, dims, spacing=0.01):
    """A box of masses with springs to the +x, +y and +z neighbours; the rest
    lengths are three constants (one per axis, 5 % off the spacing, so the
    box moves), so the robot has at most 3 x (materials) pairs."""
    from mercury_amd.robots import Element
    nx, ny, nz = dims
    idx = np.arange(nx * ny * nz)
    x, y, z = idx % nx, (idx // nx) % ny, idx // (nx * ny)
    pos = (np.stack([x, y, z], 1) * spacing + 0.002).astype(np.float32)
    pairs, lengths = [], []
    for step, ok, L in ((1, x < nx - 1, 0.95), (nx, y < ny - 1, 1.05), (nx * ny, z < nz - 1, 1.0)):
        a = idx[ok]
        pairs.append(np.stack([a, a + step], 1))
        lengths.append(np.full(len(a), L * spacing, np.float32))
    pairs = np.concatenate(pairs).astype(np.uint16)
    vox = mh.voxel_build(ms, 3, 10.0, np.full(27, 2, np.uint32), mass=0.001)
    mats = np.unique(vox.spring_material[vox.spring_material != 0])
    smat = mats[np.arange(len(pairs)) % len(mats)].astype(np.uint8)
    return Element(pos, np.full(len(pos), vox.mass_material[0], np.uint8), pairs, np.concatenate(lengths),
                   smat, mass=np.full(len(pos), 0.001, np.float32))


@pytest.mark.parametrize("dims,block,family,fits,solver", [
    ((8, 8, 12), 768, mh.MH_KERNEL_MASS, True, FORCE),          # 768 masses: a full 768-lane workgroup
    ((8, 8, 16), 1024, mh.MH_KERNEL_MASS, True, FORCE),         # 1024 masses: a full 1024-lane workgroup
    ((8, 8, 12), 768, mh.MH_KERNEL_MASS_GATHER, True, XPBD),
    ((7, 10, 11), 768, mh.MH_KERNEL_MASS, False, FORCE),        # 770 masses: over the 768-lane kernel
    ((7, 10, 11), 0, mh.MH_KERNEL_MASS, True, FORCE),           # ... which the 1024-lane one takes
])
def test_mass_kernel_capacity_edges(dims, block, family, fits, solver):
    print(f"mass kernel capacity dims={dims} block={block} family={family} solver={solver}", flush=True)
    from helpers import Case
    ms = mh.workload_materials()
    case = Case([grid_robot(ms, dims)], ms, mh.materials.WORKLOAD_ROWS)
    n, dt, env = (300, 1e-4, LAND) if solver == FORCE else (60, 1e-3, None)
    if not fits:
        with pytest.raises(mh.MhError):
            run_gpu(case, solver, 10, dt, env, kernel=family, block_size=block)
        return
    sim, pos, vel, _ = run_gpu(case, solver, n, dt, env, kernel=family, block_size=block)
    info = sim.info()
    assert info.n_lds_robots == 1 and info.n_stream_robots == 0
    assert info.block_size == (block or 1024)
    p, v, _ = mirror(case, solver, n, dt, env)
    assert np.array_equal(bits(pos), bits(p))
    assert np.array_equal(bits(vel), bits(v))
    assert np.abs(pos - case.pos0).max() > 1e-6


def test_graph_replays_for_any_clock():
    """Continuing-clock calls (no Reset between them) replay one captured
    graph: launch start times come from the device clock table, so the
    graph does not depend on t0.  Results stay bit-exact with the mirror
    run over the whole span, and with the eager path."""
    print("graph replay across the clock", flush=True)
    case = workload_case(n_robots=5, voxels=4, seed=14)
    n = float_loop_steps(0.004, 1e-4)
    out = []
    for use_graph in (1, 0):
        sim = gpu_sim(case, FORCE, 1e-4, LAND, use_graph=use_graph, substeps_per_launch=16)
        for rep in range(3):
            sim.Simulate(0.004)
        sim.synchronize()
        if use_graph:
            assert sim.info().graph_builds == 1
        out.append(sim.collect_arrays()[:2])
    p, v, _ = mirror(case, FORCE, 3 * n, 1e-4, LAND)
    for pos, vel in out:
        assert np.array_equal(bits(pos), bits(p))
        assert np.array_equal(bits(vel), bits(v))


def test_tuned_substeps_and_cached_plan_respect_the_config():
    """The tuner's choice (including substeps per launch) never changes
    results, is applied to a later population of the same shape at upload,
    and a simulator whose config excludes the cached kernel family tunes
    its own plan."""
    print("autotune cache and substeps", flush=True)
    case = workload_case(n_robots=6, voxels=4, seed=15)
    sim = gpu_sim(case, FORCE, 1e-4, LAND)
    info = sim.autotune(400)
    assert info.substeps_per_launch in (0, 250)
    sim.Step(300)
    p, v, _ = mirror(case, FORCE, 300, 1e-4, LAND)
    assert np.array_equal(bits(sim.collect_arrays()[0]), bits(p))
    # same shape, fresh simulator: the cached plan at upload
    again = gpu_sim(case, FORCE, 1e-4, LAND)
    i2 = again.info()
    assert (i2.kernel, i2.block_size, i2.substeps_per_launch) == (info.kernel, info.block_size,
                                                                 info.substeps_per_launch)
    # a simulator restricted to streaming must not take the cached plan
    stream = gpu_sim(case, FORCE, 1e-4, LAND, kernel=mh.MH_KERNEL_STREAM)
    assert stream.info().kernel == mh.MH_KERNEL_STREAM
    assert stream.autotune(200).kernel == mh.MH_KERNEL_STREAM
    stream.Step(300)
    assert np.array_equal(bits(stream.collect_arrays()[0]), bits(p))
