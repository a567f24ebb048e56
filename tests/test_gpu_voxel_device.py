"""Voxel populations built on the device (mh_sim_set_voxel_population)
against the host build (mh_voxel_build_batch + mh_sim_set_population):
the same materials after Strip, and bit-identical simulation, collect,
stress, com and fitness."""
import numpy as np
import pytest

import mercury_amd as mh
from mercury_amd import evolve
from mercury_amd.robots import last_layer_air, voxel_build, voxel_build_batch
from helpers import LAND, bits

pytestmark = pytest.mark.gpu

FORCE, XPBD = mh.MH_SOLVER_FORCE, mh.MH_SOLVER_XPBD_REF


def genomes(seed, count, voxels, p_empty):
    return evolve.random_genomes(np.random.default_rng(seed), count, voxels, p_empty=p_empty)


def composite_encodings(seed, count, n, nb=22, p_air=0.2):
    """Reference-material lattices with one- and two-material voxels."""
    rng = np.random.default_rng(seed)
    a = rng.integers(1, nb + 1, (count, n ** 3)).astype(np.uint32)
    b = rng.integers(1, nb + 1, (count, n ** 3)).astype(np.uint32)
    enc = (np.uint32(1) << a) | np.where(rng.random((count, n ** 3)) < 0.5, np.uint32(1) << b, np.uint32(0))
    enc[rng.random((count, n ** 3)) < p_air] = 1
    return np.stack([last_layer_air(n, e) for e in enc])


def both_paths(ms, enc, n, res, mass, solver, env=None, **cfg):
    kw = dict(env or {})
    dev = mh.Simulator(ms, mh.SimConfig(solver=solver, time_step=1e-4 if solver == FORCE else 1e-3, **kw, **cfg))
    dev.set_voxel_population(enc, n, res, mass=mass)
    host = mh.Simulator(ms, mh.SimConfig(solver=solver, time_step=1e-4 if solver == FORCE else 1e-3, **kw, **cfg))
    pop = voxel_build_batch(ms, n, res, enc, mass=mass)
    host.set_population(pop)
    return dev, host, pop


@pytest.mark.parametrize("case", ["workload_8", "sparse_8", "small_3", "reference_composite_5"])
def test_device_build_materials_match_host_build(case):
    print(f"device voxel build materials: {case}", flush=True)
    if case == "reference_composite_5":
        ms, n = mh.MaterialSet.reference(), 5
        enc = composite_encodings(3, 40, n)
    else:
        ms = mh.workload_materials()
        v, p_empty, count = {"workload_8": (8, 0.2, 64), "sparse_8": (8, 0.65, 64), "small_3": (3, 0.3, 50)}[case]
        n = v + 1
        enc = evolve.genome_to_lattice(genomes(5, count, v, p_empty), v)
    enc = np.concatenate([enc, np.ones((1, n ** 3), np.uint32)])  # an all-air robot
    sim = mh.Simulator(ms, mh.SimConfig(solver=FORCE, **LAND))
    sim.set_voxel_population(enc, n, 10.0, mass=0.001)
    mid, sid, stripped = sim.voxel_materials()
    S = len(sid) // len(enc)
    for r, e in enumerate(enc):
        el = voxel_build(ms, n, 10.0, e, mass=0.001)
        assert el.n_springs == S
        assert np.array_equal(mid[r * n ** 3:(r + 1) * n ** 3], el.mass_material)
        assert np.array_equal(stripped[r * n ** 3:(r + 1) * n ** 3], el.stripped_encoding)
        assert np.array_equal(sid[r * S:(r + 1) * S], el.spring_material)
    if case == "sparse_8":  # Strip removed fragments in some robots
        kept = stripped.reshape(len(enc), -1) != 1
        assert (kept != (enc != 1)).any()


@pytest.mark.parametrize("solver", [FORCE, XPBD])
def test_device_population_steps_bitexact(solver):
    print(f"device voxel population vs host upload, solver {solver}", flush=True)
    if solver == FORCE:
        ms, n, env = mh.workload_materials(), 9, LAND
        enc = evolve.genome_to_lattice(genomes(7, 24, 8, 0.2), 8)
        mass, steps = 0.001, 600
    else:
        ms, n, env = mh.MaterialSet.reference(), 5, None
        enc = composite_encodings(9, 20, n)
        mass, steps = 1.0, 300
    dev, host, pop = both_paths(ms, enc, n, 10.0, mass, solver, env)
    i = dev.info()
    assert i.kernel == mh.MH_KERNEL_MASS and i.n_blocks == len(enc)
    assert i.n_active_springs == host.info().n_active_springs
    for sim in (dev, host):
        sim.Step(steps, trackStresses=solver == XPBD)
    pd, vd, sd = dev.collect_arrays()
    ph, vh, sh = host.collect_arrays()
    assert np.array_equal(bits(pd), bits(ph)) and np.array_equal(bits(vd), bits(vh))
    assert np.array_equal(bits(sd), bits(sh))
    assert np.abs(pd - pop["pos"]).max() > 1e-5
    for mode in (mh.MH_FITNESS_REFERENCE, mh.MH_FITNESS_DISPLACEMENT):
        assert np.array_equal(bits(dev.fitness(mode)), bits(host.fitness(mode)))
    assert np.array_equal(bits(dev.com()), bits(host.com()))
    if solver == XPBD:
        assert np.abs(sd).max() > 0
        # set_stress / collect round trip in the device layout
        dev.set_stress(sd * 2)
        assert np.array_equal(bits(dev.collect_arrays()[2]), bits(sd * 2))


def test_device_population_full_size_one_second():
    """512 robots of 8^3 voxels, 1 s on land: the device-built population
    and the host upload give the same state and displacements, the tuned
    substeps included."""
    print("device voxel population, 512 x 8^3, 1 s", flush=True)
    ms = mh.workload_materials()
    enc = evolve.genome_to_lattice(genomes(0, 512, 8, 0.2), 8)
    dev, host, _ = both_paths(ms, enc, 9, 10.0, 0.001, FORCE, LAND)
    dev.autotune(2000)
    for sim in (dev, host):
        sim.Reset()
        sim.Simulate(1.0)
    pd, vd, _ = dev.collect_arrays()
    ph, vh, _ = host.collect_arrays()
    assert np.array_equal(bits(pd), bits(ph)) and np.array_equal(bits(vd), bits(vh))
    d = dev.fitness(mh.MH_FITNESS_DISPLACEMENT)
    assert np.array_equal(bits(d), bits(host.fitness(mh.MH_FITNESS_DISPLACEMENT)))
    # a second population of the same shape replaces the first
    enc2 = evolve.genome_to_lattice(genomes(1, 300, 8, 0.2), 8)
    dev.set_voxel_population(enc2, 9, 10.0, mass=0.001)
    host.set_population(voxel_build_batch(ms, 9, 10.0, enc2, mass=0.001))
    for sim in (dev, host):
        sim.Simulate(0.05)
    assert np.array_equal(bits(dev.collect_arrays()[0]), bits(host.collect_arrays()[0]))


def test_device_population_errors():
    print("device voxel population errors", flush=True)
    ms = mh.workload_materials()
    sim = mh.Simulator(ms, mh.SimConfig(solver=FORCE, **LAND))
    bad = np.ones((2, 27), np.uint32)
    bad[1, 0] = 1 << 9  # no such base material in a 4-material set
    with pytest.raises(mh.MhError):
        sim.set_voxel_population(bad, 3, 10.0, mass=0.001)
    with pytest.raises(mh.MhError):  # 11^3 lattice points
        sim.set_voxel_population(np.ones((1, 11 ** 3), np.uint32), 11, 10.0)
    stream_only = mh.Simulator(ms, mh.SimConfig(solver=FORCE, kernel=mh.MH_KERNEL_STREAM, **LAND))
    with pytest.raises(mh.MhError):
        stream_only.set_voxel_population(np.ones((1, 27), np.uint32), 3, 10.0)
    # the simulator still works after a refused upload
    sim.set_voxel_population(np.full((3, 27), 4, np.uint32), 3, 10.0, mass=0.001)
    sim.Step(10)
    assert np.isfinite(sim.collect_arrays()[0]).all()


def test_evolution_simulator_device_build_matches_host_build():
    print("BatchSimulator device vs host build", flush=True)
    g = genomes(3, 200, 8, 0.2)
    a = evolve.BatchSimulator(voxels=8, duration=0.05, device_build=True).fitness(g)
    b = evolve.BatchSimulator(voxels=8, duration=0.05, device_build=False).fitness(g)
    assert np.array_equal(a, b) and np.abs(a).max() > 0
