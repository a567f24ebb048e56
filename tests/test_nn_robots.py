"""NN-robot producer: MLP materials, KNN rows (GPU) and KNN springs."""
import numpy as np
import pytest

import oracle
import mercury_amd as mh
from mercury_amd import robots
from helpers import WATER, Case, bits, gpu_sim, mirror


def mlp_materials_np(weights, pos):
    """Independent float32 restatement: sequential dot products, ReLU,
    first maximum."""
    a = pos.astype(np.float32).T  # (3, n)
    for li, w in enumerate(weights):
        s = np.zeros((w.shape[0], a.shape[1]), np.float32)
        for k in range(w.shape[1]):
            s = (s + (w[:, k:k + 1] * a[k:k + 1, :]).astype(np.float32)).astype(np.float32)
        a = np.where(s < 0, np.float32(0), s) if li < len(weights) - 1 else s
    return np.argmax(a, axis=0).astype(np.uint8)


def test_nn_materials_match_restatement():
    rng = np.random.default_rng(1)
    sizes = [3, 25, 25, 4]
    W = [rng.uniform(-1, 1, (o, i)).astype(np.float32) for i, o in zip(sizes[:-1], sizes[1:])]
    pos = (rng.random((700, 3)) - 0.5).astype(np.float32)
    got = robots.nn_materials(W, sizes, pos)
    assert np.array_equal(got, mlp_materials_np(W, pos))
    assert len(np.unique(got)) > 1


def test_knn_springs_match_restatement():
    rng = np.random.default_rng(2)
    ms = mh.workload_materials()
    _, _, by_enc = oracle.composite_tables(oracle.make_base(mh.materials.WORKLOAD_ROWS))
    pos = (rng.random((80, 3)) * 0.3).astype(np.float32)
    pos[5] = pos[6] + np.float32([5e-4, 0, 0])  # closer than min_length: skipped
    nbr, dist = oracle.knn(pos, 12)
    enc = (1 << rng.integers(1, 5, 80)).astype(np.uint32)
    pairs, ln, sid = robots.knn_springs(ms, nbr, dist, enc, min_length=1e-3)
    exp = []
    for i in range(80):
        for r in range(12):
            d = dist[i, r]
            if d >= np.float32(1e-3) and np.isfinite(d):
                e = int(enc[i] | enc[nbr[i, r]])
                exp.append((i, nbr[i, r], d, by_enc[oracle.encoded_idx(e, 11)].id))
    assert len(pairs) == len(exp) < 80 * 12
    assert np.array_equal(pairs, np.array([(a, b) for a, b, _, _ in exp], np.uint16))
    assert np.array_equal(bits(ln), bits(np.array([d for _, _, d, _ in exp], np.float32)))
    assert np.array_equal(sid, np.array([m for *_, m in exp], np.uint8))


@pytest.mark.gpu
def test_gpu_knn_matches_oracle():
    print("gpu knn", flush=True)
    rng = np.random.default_rng(3)
    sizes = [40, 1300, 257, 700]
    pos = [(rng.random((n, 3)) * 0.5).astype(np.float32) for n in sizes]
    pos[0][3] = pos[0][9]  # coincident masses
    mo = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint32)
    nbr, dist = robots.knn_batch(mo, np.concatenate(pos), 25)
    for r, p in enumerate(pos):
        on, od = oracle.knn(p, 25)
        assert np.array_equal(nbr[mo[r]:mo[r + 1]], on)
        assert np.array_equal(bits(dist[mo[r]:mo[r + 1]]), bits(od))


@pytest.mark.gpu
def test_nn_population_mixed_plan_bitexact():
    print("nn population mixed plan", flush=True)
    ms = mh.workload_materials()
    els = robots.nn_population(ms, 6, mass_range=(60, 700), K=25, seed=0)  # 3 LDS, 3 streamed
    case = Case(els, ms, mh.materials.WORKLOAD_ROWS)
    sim = gpu_sim(case, mh.MH_SOLVER_FORCE, 1e-4, WATER)
    info = sim.info()
    assert info.n_stream_robots > 0 and info.n_lds_robots > 0, (info.n_lds_robots, info.n_stream_robots)
    sim.Step(200)
    pos, vel, _ = sim.collect_arrays()
    p, v, _ = mirror(case, mh.MH_SOLVER_FORCE, 200, 1e-4, WATER)
    assert np.array_equal(bits(pos), bits(p))
    assert np.array_equal(bits(vel), bits(v))
    # autotuned plan gives the same trajectory
    sim2 = gpu_sim(case, mh.MH_SOLVER_FORCE, 1e-4, WATER)
    sim2.autotune(20)
    sim2.Step(200)
    assert np.array_equal(bits(sim2.collect_arrays()[0]), bits(p))


@pytest.mark.gpu
def test_large_nn_robots_many_chunks_bitexact():
    """Robots of 1500-2000 masses (37-50 k springs) take several LDS chunks per
    substep in the chunked kernel, whose next chunk is loaded while the
    current one is worked on; the trajectory equals the mirror's."""
    print("nn 2 large robots, chunked kernel, 100 steps", flush=True)
    ms = mh.workload_materials()
    els = robots.nn_population(ms, 2, mass_range=(1500, 2000), K=25, seed=3)
    case = Case(els, ms, mh.materials.WORKLOAD_ROWS)
    sim = gpu_sim(case, mh.MH_SOLVER_FORCE, 1e-4, WATER, kernel=mh.MH_KERNEL_STREAM)
    info = sim.info()
    assert info.stream_chunk > 0 and max(e.n_springs for e in els) > 4 * info.stream_chunk
    sim.Step(100)
    pos, vel, _ = sim.collect_arrays()
    p, v, _ = mirror(case, mh.MH_SOLVER_FORCE, 100, 1e-4, WATER)
    assert np.array_equal(bits(pos), bits(p))
    assert np.array_equal(bits(vel), bits(v))


# six base materials: 6 single and 15 two-material composites, over the 15
# that the 6-byte record format holds
SIX_ROWS = mh.materials.WORKLOAD_ROWS + [(3.0e3, 0.1, mh.materials.ACT_OMEGA, 1.5),
                                         (7.0e3, 0.0, mh.materials.ACT_OMEGA, 0.0)]


@pytest.mark.gpu
@pytest.mark.parametrize("block,n_base", [(1024, 4), (512, 4), (1024, 6)])
def test_nn_robots_mass_stream_kernel_bitexact(block, n_base):
    """step_mstream (one lane per mass, each evaluating all of its springs,
    records streamed every substep): robots of 60-2000 masses, so threads
    serve one to four masses, give the mirror's trajectory bit for bit, as
    does the chunked kernel on the same robots.  Four base materials give
    at most 10 compact materials and the 6-byte records; six give more than
    15 and the 8-byte records."""
    print(f"nn robots on the mass-parallel streaming kernel, block {block}, {n_base} base materials", flush=True)
    rows = mh.materials.WORKLOAD_ROWS if n_base == 4 else SIX_ROWS
    ms = mh.MaterialSet.custom(rows)
    els = robots.nn_population(ms, 4, mass_range=(60, 2000), K=25, n_materials=n_base, seed=5)
    assert max(e.n_masses for e in els) > 1024
    if n_base == 6:  # every composite of the six (ids 1-21) on some springs
        rng = np.random.default_rng(6)
        for e in els:
            e.spring_material = rng.integers(1, ms.count(), e.n_springs).astype(np.uint8)
    n_comp = len(np.unique(np.concatenate([e.spring_material for e in els])))
    assert (n_comp <= 15) == (n_base == 4)
    case = Case(els, ms, rows)
    sim = gpu_sim(case, mh.MH_SOLVER_FORCE, 1e-4, WATER, kernel=mh.MH_KERNEL_MASS_STREAM, block_size=block)
    info = sim.info()
    assert info.kernel == mh.MH_KERNEL_MASS_STREAM and info.n_stream_robots == 4
    assert info.stream_block_size == block
    assert info.stream_record_bytes == (6 if n_base == 4 else 8)
    pos0 = sim.collect_arrays()[0]
    sim.Step(150)
    pos, vel, _ = sim.collect_arrays()
    p, v, _ = mirror(case, mh.MH_SOLVER_FORCE, 150, 1e-4, WATER)
    assert np.array_equal(bits(pos), bits(p))
    assert np.array_equal(bits(vel), bits(v))
    assert np.abs(pos - pos0).max() > 1e-6
    chunk = gpu_sim(case, mh.MH_SOLVER_FORCE, 1e-4, WATER, kernel=mh.MH_KERNEL_STREAM)
    assert chunk.info().stream_chunk > 0
    chunk.Step(150)
    assert np.array_equal(bits(chunk.collect_arrays()[0]), bits(pos))
    # the variant the autotuner picks among the mass-parallel streaming ones
    tuned = gpu_sim(case, mh.MH_SOLVER_FORCE, 1e-4, WATER, kernel=mh.MH_KERNEL_MASS_STREAM)
    tuned.autotune(40)
    tuned.Step(150)
    assert np.array_equal(bits(tuned.collect_arrays()[0]), bits(pos))
