"""Evolution loop pieces on the CPU: genome encoding, variation operators,
ALPS layer bookkeeping and the surrogate fallback switch (driven by a
stand-in fitness so no GPU is needed), plus one short GPU run."""
import numpy as np
import pytest

import mercury_amd as mh
from mercury_amd import evolve, surrogate


def test_genome_to_lattice_matches_voxel_encoding_layout():
    rng = np.random.default_rng(0)
    g = evolve.random_genomes(rng, 3, 4)
    enc = evolve.genome_to_lattice(g, 4)
    n = 5
    assert enc.shape == (3, n ** 3)
    # the last lattice layer in every axis is air
    assert np.array_equal(mh.robots.last_layer_air(n, enc[0]), enc[0])
    # voxel (x, y, z) sits at lattice x + n*(y + n*z) with one bit per material
    for r in range(3):
        for v in range(64):
            x, y, z = v % 4, (v // 4) % 4, v // 16
            e = enc[r, x + n * (y + n * z)]
            assert e == (1 if g[r, v] == 0 else 1 << int(g[r, v]))


def test_variation_operators():
    rng = np.random.default_rng(1)
    a = evolve.random_genomes(rng, 200, 4)
    b = evolve.random_genomes(rng, 200, 4)
    assert np.array_equal(evolve.swap_crossover(rng, a, b, 0.0), a)
    c = evolve.swap_crossover(rng, a, b, 1.0)
    assert ((c == a) | (c == b)).all()
    frac_b = ((c == b) & (a != b)).sum() / (a != b).sum()
    assert 0.45 < frac_b < 0.55
    assert np.array_equal(evolve.mutate(rng, a, 0.0), a)
    m = evolve.mutate(rng, a, 1.0)
    assert m.min() >= 0 and m.max() <= 4
    changed = (evolve.mutate(rng, a, 0.05) != a).mean()
    assert 0.02 < changed < 0.06  # 0.05 x P(new draw differs) = 0.05 x 4/5


def test_age_limits_polynomial():
    lim = evolve.age_limits(4, 5)
    assert lim[:3] == [5, 10, 20] and lim[3] > 10 ** 12


def test_vectorized_features_agree_with_element_features():
    """The genome-array features and the built-robot features describe the
    same material layout when Strip removes nothing."""
    ms = mh.workload_materials()
    g = np.full((1, 27), 1, np.uint8)
    g[0, ::2] = 3
    g[0, 1::4] = 4
    enc = evolve.genome_to_lattice(g, 3)
    el = mh.robots.voxel_build(ms, 4, 10.0, enc[0], mass=0.001)
    fa = evolve.genome_features(g, 3)[0]
    fb = surrogate.genome_features([el], 3, mh.materials.WORKLOAD_ROWS)[0]
    np.testing.assert_array_equal(fa, fb)


class _FakeSim:
    """Fitness = share of phase-0 actuated voxels minus share of empty ones."""

    def __init__(self):
        self.evals = 0

    def fitness(self, g):
        self.evals += len(g)
        return (g == 3).mean(1) - (g == 0).mean(1)


def test_plain_run_improves_and_counts_evaluations():
    sim = _FakeSim()
    r = evolve.run(seed=0, pop=256, voxels=4, layers=4, max_generations=6, sim=sim)
    assert r.generations == 6
    # initial population + pop per generation + a fresh layer every age_gap
    assert sim.evals == 256 + 6 * 256 + (6 // 5) * (256 // 4)
    best = [h[2] for h in r.history]
    assert all(b1 >= b0 for b0, b1 in zip(best, best[1:])) and best[-1] > best[0]


def test_run_stops_at_threshold():
    sim = _FakeSim()
    r = evolve.run(seed=0, pop=256, voxels=4, layers=4, threshold=0.25, max_generations=50,
                   sim=sim)
    assert r.best >= 0.25 and r.evals_to_threshold == sim.evals


class _FakeSurrogate:
    """Predicts fitness with a fixed sign: +1 informative, -1 anti-correlated."""
    sign = 1.0

    def __init__(self, seed=0, **kw):
        pass

    def fit(self, X, y, **kw):
        return 0.0

    def predict(self, X):
        # share of material 3 minus the empty share, from the per-layer
        # shares (4 layers x 4 materials x (share, balance) at 4^3 voxels)
        share = X[:, :32:2].reshape(len(X), 4, 4).mean(1)
        return self.sign * (share[:, 2] - (1.0 - share.sum(1)))


def test_surrogate_screening_and_rho_fallback(monkeypatch):
    monkeypatch.setattr(evolve, "Surrogate", _FakeSurrogate)
    _FakeSurrogate.sign = 1.0
    good = evolve.run(seed=0, pop=256, voxels=4, layers=4, max_generations=4, sim=_FakeSim(),
                      use_surrogate=True)
    assert good.fallback_generations == 0 and min(good.rho_history) > 0.6
    plain = evolve.run(seed=0, pop=256, voxels=4, layers=4, max_generations=4, sim=_FakeSim())
    assert good.best >= plain.best
    _FakeSurrogate.sign = -1.0
    bad = evolve.run(seed=0, pop=256, voxels=4, layers=4, max_generations=4, sim=_FakeSim(),
                     use_surrogate=True, fallback_retrain=False)
    # first screened generation shows rho < 0.6; the following ones are unscreened
    assert bad.rho_history[0] < 0.6 and bad.fallback_generations >= 3


@pytest.mark.gpu
def test_short_gpu_evolution_run():
    print("evolve: 2 generations, pop 256, 4^3 voxels, 0.05 s", flush=True)
    sim = evolve.BatchSimulator(voxels=4, duration=0.05)
    r = evolve.run(seed=0, pop=256, voxels=4, layers=4, max_generations=2, sim=sim)
    assert sim.evals == 256 * 3 and np.isfinite(r.best)
    r2 = evolve.run(seed=0, pop=256, voxels=4, layers=4, max_generations=2, use_surrogate=True,
                    train_epochs=2, sim=sim)
    assert len(r2.rho_history) == 2 and np.isfinite(r2.best)


def test_evals_to_and_report():
    h = [(4096, 1.0, 0.05), (8192, 2.0, 0.07), (12288, 3.0, 0.09)]
    assert evolve.evals_to(h, 0.06) == (8192, 2.0)
    assert evolve.evals_to(h, 0.2) == (-1, -1.0)
    rows = [dict(surrogate=False, history=h), dict(surrogate=True, history=h[1:]),
            dict(surrogate=False, history=h[:1])]
    t = evolve.report(rows, [0.07, 0.09])
    assert t["plain"][0.07] == (8192.0, 1, 2) and t["surrogate"][0.09] == (12288.0, 1, 1)


@pytest.mark.gpu
def test_short_gpu_evolution_run_cnn_surrogate():
    print("evolve: cnn surrogate, 2 generations, pop 256, 8^3 voxels, 0.05 s", flush=True)
    sim = evolve.BatchSimulator(voxels=8, duration=0.05)
    r = evolve.run(seed=1, pop=256, voxels=8, layers=4, max_generations=2, use_surrogate=True,
                   surrogate_kind="cnn", train_epochs=1, sim=sim)
    assert len(r.rho_history) == 2 and np.isfinite(r.best)
    with pytest.raises(ValueError):
        evolve.GenomeModel("cnn", 4)


def test_warm_start_keeps_one_surrogate(monkeypatch):
    made = []

    class Counting(_FakeSurrogate):
        def __init__(self, seed=0, **kw):
            made.append(seed)

    monkeypatch.setattr(evolve, "Surrogate", Counting)
    Counting.sign = 1.0
    evolve.run(seed=0, pop=256, voxels=4, layers=4, max_generations=11, sim=_FakeSim(),
               use_surrogate=True, warm_start=True)
    assert len(made) == 1
    made.clear()
    evolve.run(seed=0, pop=256, voxels=4, layers=4, max_generations=11, sim=_FakeSim(),
               use_surrogate=True)
    assert len(made) == 3  # generation 1, then every 5 generations


def test_fallback_retrain_retrains_while_ranking_is_poor(monkeypatch):
    """With an anti-correlated surrogate every generation falls back; the
    fallback_retrain rule retrains it each generation (generation 1 and
    2..6), the default every 5 generations (1 and 5)."""
    made = []

    class Counting(_FakeSurrogate):
        def __init__(self, seed=0, **kw):
            made.append(seed)

    monkeypatch.setattr(evolve, "Surrogate", Counting)
    Counting.sign = -1.0
    evolve.run(seed=0, pop=256, voxels=4, layers=4, max_generations=6, sim=_FakeSim(),
               use_surrogate=True, fallback_retrain=True)
    assert len(made) == 6
    made.clear()
    evolve.run(seed=0, pop=256, voxels=4, layers=4, max_generations=6, sim=_FakeSim(),
               use_surrogate=True, fallback_retrain=False)
    assert len(made) == 2
    # an informative surrogate never falls back, so the rule adds nothing
    made.clear()
    Counting.sign = 1.0
    evolve.run(seed=0, pop=256, voxels=4, layers=4, max_generations=6, sim=_FakeSim(),
               use_surrogate=True, fallback_retrain=True)
    assert len(made) == 2


def test_mirror_augmentation_doubles_training_data(monkeypatch):
    seen = []

    class Recording(_FakeSurrogate):
        def fit(self, X, y, **kw):
            seen.append((X.copy(), y.copy()))
            return 0.0

        def predict(self, X):
            return X[:, 0].astype(np.float64)

    monkeypatch.setattr(evolve, "Surrogate", Recording)
    rng = np.random.default_rng(2)
    g = evolve.random_genomes(rng, 10, 4)
    m = evolve.mirror_z(g, 4)
    assert np.array_equal(evolve.mirror_z(m, 4), g)
    # voxel (x, y, z) of the mirror is voxel (x, y, 3 - z) of the original
    assert np.array_equal(m.reshape(-1, 4, 4, 4)[:, 0], g.reshape(-1, 4, 4, 4)[:, 3])
    y = rng.normal(size=10).astype(np.float32)
    model = evolve.GenomeModel("mlp", 4, mirror=True)
    model.fit(g, y)
    X, yy = seen[-1]
    assert len(X) == 20 and np.array_equal(yy, np.concatenate([y, y]))
    assert np.allclose(X[10:], evolve.genome_features(m, 4), atol=1e-6)  # batched float32 sums
    # predictions average the two views
    want = 0.5 * (evolve.genome_features(g, 4)[:, 0] + evolve.genome_features(m, 4)[:, 0])
    assert np.allclose(model.predict(g), want)


def test_mirrored_robots_travel_alike_in_the_oracle():
    """The basis of mirror augmentation: a robot mirrored across the lateral
    z axis moves its centre of mass along x about as far as the original
    (FP32 mirror of the force step, land, 16 random 8^3 robots, 1 s)."""
    from scipy.stats import spearmanr
    import oracle
    from helpers import LAND, Case, mirror
    ms = mh.workload_materials()
    g = evolve.random_genomes(np.random.default_rng(4), 16, 8)

    def travel(gen):
        enc = evolve.genome_to_lattice(gen, 8)
        case = Case([mh.voxel_build(ms, 9, 10.0, e, mass=0.001) for e in enc], ms, mh.materials.WORKLOAD_ROWS)
        p, _, _ = mirror(case, mh.MH_SOLVER_FORCE, 10000, 1e-4, LAND, nthreads=4)
        mo, mm = case.a["mass_offset"], case.a["mass_material"]
        out = []
        for r in range(len(enc)):
            air = (mm[mo[r]:mo[r + 1]] == 0).astype(np.uint8)
            out.append(oracle.com(p[mo[r]:mo[r + 1]], air)[0] - oracle.com(case.pos0[mo[r]:mo[r + 1]], air)[0])
        return np.array(out)

    a, b = travel(g), travel(evolve.mirror_z(g, 8))
    assert np.abs(a).max() > 1e-4                      # the robots do move
    assert spearmanr(a, b).correlation > 0.95


class _RecordingSim(evolve.BatchSimulator):
    """The simulator of the evolution loop, remembering every batch it scored."""

    def __init__(self, **kw):
        super().__init__(**kw)
        self.batches = []

    def fitness(self, genomes):
        self.batches.append(np.array(genomes, copy=True))
        return super().fitness(genomes)


def _check_screen(entry, per, n_audit, layers):
    """The simulated offspring of a screened generation are, per layer, the
    best (per - n_audit) by prediction and n_audit audits drawn from the rest."""
    pred, lay, keep, aud = entry["pred"], entry["layer_of"], entry["keep"], entry["audited"]
    assert len(keep) == per * layers and aud.sum() == n_audit * layers
    for i in range(layers):
        mine = np.flatnonzero(lay == i)
        k = keep[lay[keep] == i]
        a = aud[lay[keep] == i]
        top = mine[np.argsort(-pred[mine], kind="stable")[:per - n_audit]]
        assert np.array_equal(np.sort(k[~a]), np.sort(top))
        rest = np.setdiff1d(mine, top)
        assert len(np.unique(k[a])) == n_audit and np.isin(k[a], rest).all()


@pytest.mark.gpu
def test_surrogate_screened_generation_full_size(monkeypatch):
    """One generation of the screened protocol at full size: population
    4096 robots of 8^3 voxels, ALPS with 8 layers, elitism 0.1, swap
    crossover 0.5, mutation 0.05; the MLP surrogate scores 4 x 4096 =
    16384 offspring and the best of each layer plus a 10 % audit are
    simulated for 1 s -- 4096 evaluations per generation, counted one per
    simulated robot as Evaluator_impl.h:103-106 counts them.  Then the same
    with a sign-flipped surrogate: its audit ranking falls below 0.6, so
    the next generation is simulated without screening."""
    print("screened generation: pop 4096, 8^3, 1 s, 16384 scored", flush=True)
    pop, layers = 4096, 8
    per, n_audit = pop // layers, int(round(0.1 * (pop // layers)))
    sim = _RecordingSim(voxels=8, duration=1.0)
    log = []
    kw = dict(pop=pop, voxels=8, layers=layers, crossover=0.5, mutation=0.05, elitism=0.1,
              use_surrogate=True, screen_factor=4, fallback_retrain=False, surrogate_kind="mlp",
              duration=1.0, sim=sim)
    r = evolve.run(seed=0, max_generations=1, log=log, **kw)
    assert [len(b) for b in sim.batches] == [pop, pop] and sim.evals == 2 * pop
    assert [h[0] for h in r.history] == [pop, 2 * pop]
    (e,) = log
    assert e["screened"] and e["scored"] == 4 * pop and e["simulated"] == pop
    _check_screen(e, per, n_audit, layers)
    assert np.array_equal(sim.batches[1], e["genomes"][e["keep"]])
    print(f"audit rho {e['rho']:.3f}", flush=True)

    class Flipped(evolve.GenomeModel):
        def predict(self, genomes):
            return -super().predict(genomes)

    monkeypatch.setattr(evolve, "GenomeModel", Flipped)
    sim.batches.clear()
    log2 = []
    r2 = evolve.run(seed=1, max_generations=2, log=log2, **kw)
    first, second = log2
    assert first["screened"] and first["rho"] < 0.6
    _check_screen(first, per, n_audit, layers)
    assert not second["screened"] and second["simulated"] == pop
    assert r2.fallback_generations == 1 and len(r2.rho_history) == 2
    assert [len(b) for b in sim.batches] == [pop, pop, pop]
