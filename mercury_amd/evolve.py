"""Evolution loop over voxel genomes, with optional surrogate pre-screening.

The caller of the evaluate path (optimize/optimizer/optimizer_impl.h:113-449,
Evaluator_impl.h:39-107), rebuilt for the surrogate-screened workload:
  * genomes: one material per voxel of an n^3 voxel grid (0 = empty);
  * variation: swap crossover (uniform per-voxel swap between two parents)
    with probability `crossover`, then per-voxel mutation with probability
    `mutation`;
  * ALPS (Hornby 2006, age-layered population structure): `layers` age
    layers with polynomial age limits; layer 0 is re-seeded with random
    genomes every `age_gap` generations; parents come from a layer and the
    one below; each layer keeps its best `elitism` fraction and fills the
    rest by fitness; individuals past their layer's age limit move up;
    (the reference parses ALPS but never runs it, optimizer_impl.h:421-428);
  * evaluation: every simulated robot counts one evaluation, as
    Evaluator_impl.h:103-106 counts them; fitness is the COM-x displacement
    after `duration` s on land;
  * surrogate: scores `screen_factor` x offspring (8 by default); of each
    layer's share the best (1 - audit) by prediction and a uniform `audit`
    sample of the rest are simulated; retrained every `retrain_every`
    generations on all simulated pairs; when its Spearman correlation on
    the audit sample falls below `min_rho`, either the surrogate is
    retrained at once on the grown data and screens the next generation
    (`fallback_retrain`, the default), or the next generations are
    simulated without screening (still scored, to watch the correlation)
    until it recovers or the periodic retrain brings it back.
"""
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .materials import WORKLOAD_ROWS, workload_materials
from .robots import last_layer_air, voxel_build, voxel_build_batch
from .simulator import SimConfig, Simulator
from .surrogate import GenomeSurrogate, Surrogate, layered_features, spearman

LAND = dict(gravity=9.81, ground_k=1.0e5, mu_s=1.0, mu_k=0.8, drag=0.0, damping=0.99)


def random_genomes(rng, n, voxels, n_materials=4, p_empty=0.2):
    """(n, voxels^3) material indices: 0 = empty with p_empty, else 1..n_materials."""
    g = rng.integers(1, n_materials + 1, (n, voxels ** 3)).astype(np.uint8)
    g[rng.random((n, voxels ** 3)) < p_empty] = 0
    return g


def mutate(rng, g, rate, n_materials=4):
    g = g.copy()
    hit = rng.random(g.shape) < rate
    g[hit] = rng.integers(0, n_materials + 1, int(hit.sum())).astype(np.uint8)
    return g


def swap_crossover(rng, a, b, rate):
    """Per child: with probability `rate`, each voxel from the second parent
    with probability 1/2; otherwise a copy of the first parent."""
    do = rng.random(len(a)) < rate
    swap = (rng.random(a.shape) < 0.5) & do[:, None]
    return np.where(swap, b, a)


def genome_to_lattice(g, voxels):
    """Voxel materials -> encodings on the (voxels+1)^3 lattice (last layer air)."""
    n = voxels + 1
    idx = np.arange(n ** 3)
    x, y, z = idx % n, (idx // n) % n, idx // (n * n)
    inside = (x < voxels) & (y < voxels) & (z < voxels)
    enc = np.ones((len(g), n ** 3), np.uint32)
    vox_id = x[inside] + voxels * (y[inside] + voxels * z[inside])
    mats = g[:, vox_id].astype(np.uint32)
    enc[:, inside] = np.where(mats == 0, 1, np.uint32(1) << mats)
    return enc


def genome_features(g, voxels, rows=WORKLOAD_ROWS):
    """The surrogate's 64 inputs of genomes g[robot, voxel]
    (surrogate.layered_features of the genome's own materials)."""
    return layered_features(g, voxels, len(rows))


class BatchSimulator:
    """Builds and simulates genome batches; one evaluation per robot.
    device_build: the robots are built on the device from their voxel
    encodings (Simulator.set_voxel_population; only the encodings cross
    PCIe), else on host threads and uploaded (voxel_build_batch +
    set_population); both give the same fitness."""

    def __init__(self, voxels=8, duration=1.0, dt=1e-4, spacing=0.1, mass=0.001, device=0,
                 build_threads=0, device_build=True):
        self.voxels, self.duration = voxels, duration
        self.build_threads = build_threads
        self.device_build = device_build
        self.ms = workload_materials()
        self.res = 1.0 / spacing
        self.mass = mass
        self.sim = Simulator(self.ms, SimConfig(solver=_lib.MH_SOLVER_FORCE, time_step=dt,
                                                device=device, **LAND))
        self.evals = 0
        self.sim_seconds = 0.0
        self.tuned = False

    def fitness(self, genomes):
        enc = genome_to_lattice(genomes, self.voxels)
        t0 = time.perf_counter()
        if self.device_build:
            self.sim.set_voxel_population(enc, self.voxels + 1, self.res, mass=self.mass)
        else:
            pop = voxel_build_batch(self.ms, self.voxels + 1, self.res, enc, mass=self.mass,
                                    threads=self.build_threads)
            t0 = time.perf_counter()
            self.sim.set_population(pop)
        if not self.tuned:
            self.sim.autotune(100)
            self.tuned = True
        self.sim.Reset()
        self.sim.Simulate(self.duration)
        f = self.sim.fitness(_lib.MH_FITNESS_DISPLACEMENT)
        self.sim_seconds += time.perf_counter() - t0
        self.evals += len(genomes)
        return f.astype(np.float64)


def mirror_z(g, voxels):
    """Genomes mirrored across the lateral axis (voxel z -> voxels - 1 - z).
    Gravity is along -y and ground friction is isotropic in x-z, so a
    mirrored robot travels the same COM-x distance up to the chaotic
    divergence of the two runs (rank correlation 0.99 on random 8^3
    genomes)."""
    return np.ascontiguousarray(g.reshape(len(g), voxels, voxels, voxels)[:, ::-1]).reshape(len(g), -1)


class GenomeModel:
    """Surrogate on genomes: kind "mlp" feeds genome_features to the MLP,
    kind "cnn" feeds the voxel grid to the convolutional network.  With
    mirror=True every simulated genome also trains the model as its
    z-mirror (same fitness), and predictions average the two views."""

    def __init__(self, kind, voxels, seed=0, epochs=None, mirror=False):
        if kind not in ("mlp", "cnn"):
            raise ValueError("surrogate kind must be 'mlp' or 'cnn'")
        if kind == "cnn" and voxels != 8:
            raise ValueError("the convolutional surrogate takes 8^3 genomes")
        self.kind, self.voxels, self.mirror = kind, voxels, mirror
        self.epochs = epochs if epochs is not None else (30 if kind == "mlp" else 8)
        self.model = Surrogate(seed=seed) if kind == "mlp" else GenomeSurrogate(seed=seed)

    def fit(self, genomes, y, seed=0):
        if self.mirror:
            genomes = np.concatenate([genomes, mirror_z(genomes, self.voxels)])
            y = np.concatenate([y, y])
        if self.kind == "mlp":
            return self.model.fit(genome_features(genomes, self.voxels), y, epochs=self.epochs, batch=512,
                                  lr=1e-3, seed=seed)
        return self.model.fit(genomes, y, epochs=self.epochs, batch=256, lr=1e-3, seed=seed)

    def _predict(self, genomes):
        if self.kind == "mlp":
            return self.model.predict(genome_features(genomes, self.voxels))
        return self.model.predict(genomes)

    def predict(self, genomes):
        p = self._predict(genomes)
        if self.mirror:
            p = 0.5 * (p + self._predict(mirror_z(genomes, self.voxels)))
        return p


@dataclass
class Layer:
    genomes: np.ndarray
    fitness: np.ndarray
    age: np.ndarray


@dataclass
class RunResult:
    evals_to_threshold: int = -1
    seconds_to_threshold: float = -1.0
    generations: int = 0
    best: float = -np.inf
    history: list = field(default_factory=list)
    rho_history: list = field(default_factory=list)
    fallback_generations: int = 0


def age_limits(layers, age_gap):
    """Polynomial ALPS age limits: age_gap * (1, 2, 4, 6, 9, 12, 16, ...), last unbounded."""
    poly = [1, 2, 4, 6, 9, 12, 16, 20, 25, 30]
    lim = [age_gap * poly[min(i, len(poly) - 1)] for i in range(layers - 1)]
    return lim + [np.iinfo(np.int64).max]


def run(seed=0, pop=4096, voxels=8, layers=8, crossover=0.5, mutation=0.05, elitism=0.1,
        age_gap=5, threshold=None, max_generations=40, use_surrogate=False, screen_factor=8,
        retrain_every=5, min_rho=0.6, train_epochs=None, duration=1.0, sim=None, verbose=False,
        audit=0.1, surrogate_kind="mlp", max_train=200000, warm_start=False,
        fallback_retrain=True, mirror=False, log=None):
    """One evolution run; stops when the best fitness reaches `threshold` or
    after max_generations.  Returns a RunResult.  warm_start: each retrain
    continues from the previous network's weights instead of a fresh one.
    fallback_retrain: when the surrogate ranks below min_rho, retrain it
    before the next generation (which it then screens) instead of
    simulating unscreened generations until the periodic retrain.
    mirror: train the surrogate on every genome and its z-mirror.
    log: a list that receives one dict per generation: whether the
    surrogate screened it, how many offspring were scored and simulated,
    and for a screened generation the scored genomes, their predictions
    and layers, the indices simulated and which of them are audits."""
    rng = np.random.default_rng(seed)
    sim = sim or BatchSimulator(voxels=voxels, duration=duration)
    evals0, t_start = sim.evals, time.perf_counter()
    res = RunResult()
    per = pop // layers
    limits = age_limits(layers, age_gap)
    data_x, data_y = [], []
    sur = None
    screening = use_surrogate

    g0 = random_genomes(rng, pop, voxels)
    f0 = sim.fitness(g0)
    data_x.append(g0)
    data_y.append(f0)
    order = np.argsort(-f0, kind="stable")
    L = [Layer(g0[order[i * per:(i + 1) * per]], f0[order[i * per:(i + 1) * per]],
               np.zeros(per, np.int64)) for i in range(layers)]

    def record():
        best = max(float(l.fitness.max()) for l in L if len(l.fitness))
        res.best = max(res.best, best)
        res.history.append((sim.evals - evals0, time.perf_counter() - t_start, res.best))
        if threshold is not None and res.best >= threshold and res.evals_to_threshold < 0:
            res.evals_to_threshold = sim.evals - evals0
            res.seconds_to_threshold = time.perf_counter() - t_start
        return res.evals_to_threshold >= 0

    if record():
        return res
    for gen in range(1, max_generations + 1):
        res.generations = gen
        if use_surrogate and (sur is None or gen % retrain_every == 0 or (fallback_retrain and not screening)):
            if sur is None or not warm_start:
                sur = GenomeModel(surrogate_kind, voxels, seed=seed + gen, epochs=train_epochs, mirror=mirror)
            X = np.concatenate(data_x)[-max_train:]
            Y = np.concatenate(data_y)[-max_train:]
            sur.fit(X, Y, seed=seed + gen)
            screening = True
        # offspring per layer: parents from the layer and the one below
        n_child = pop * (screen_factor if (use_surrogate and screening) else 1)
        kids, kid_age = [], []
        for i in range(layers):
            pool_g = np.concatenate([L[j].genomes for j in (i - 1, i) if j >= 0])
            pool_f = np.concatenate([L[j].fitness for j in (i - 1, i) if j >= 0])
            pool_a = np.concatenate([L[j].age for j in (i - 1, i) if j >= 0])
            m = n_child // layers

            def tournament():
                a = rng.integers(0, len(pool_g), (m, 2))
                return np.where(pool_f[a[:, 0]] >= pool_f[a[:, 1]], a[:, 0], a[:, 1])
            pa, pb = tournament(), tournament()
            child = mutate(rng, swap_crossover(rng, pool_g[pa], pool_g[pb], crossover), mutation)
            kids.append(child)
            kid_age.append(np.maximum(pool_a[pa], pool_a[pb]) + 1)
        kids = np.concatenate(kids)
        kid_age = np.concatenate(kid_age)
        layer_of = np.repeat(np.arange(layers), n_child // layers)
        entry = dict(generation=gen, screened=bool(use_surrogate and screening), scored=0)
        if use_surrogate and screening:
            pred = sur.predict(kids)
            entry.update(scored=len(kids), genomes=kids, pred=pred, layer_of=layer_of)
            keep, audited = [], []
            n_audit = int(round(audit * per))
            for i in range(layers):
                # the top share of every layer's offspring by prediction, plus
                # a uniform sample of the rest that measures the surrogate
                sel = np.where(layer_of == i)[0]
                order = sel[np.argsort(-pred[sel], kind="stable")]
                top = order[:per - n_audit]
                rest = order[per - n_audit:]
                aud = rng.choice(rest, size=min(n_audit, len(rest)), replace=False) if n_audit else rest[:0]
                keep += [top, aud]
                audited.append(np.ones(len(aud), bool))
                audited.insert(-1, np.zeros(len(top), bool))
            keep = np.concatenate(keep)
            audited = np.concatenate(audited)
            entry.update(keep=keep, audited=audited)
            kids, kid_age, layer_of, pred = kids[keep], kid_age[keep], layer_of[keep], pred[keep]
        else:
            res.fallback_generations += int(use_surrogate)
        fit = sim.fitness(kids)
        data_x.append(kids)
        data_y.append(fit)
        entry["simulated"] = len(kids)
        if log is not None:
            log.append(entry)
        if use_surrogate and screening:
            # with an audit sample, rank quality is measured on it alone (the
            # screened top share is range-restricted)
            rho = spearman(pred[audited], fit[audited]) if audited.sum() >= 8 else spearman(pred, fit)
            res.rho_history.append(rho)
            entry["rho"] = rho
            if rho < min_rho:
                screening = False
        elif use_surrogate:
            rho = spearman(sur.predict(kids), fit)
            res.rho_history.append(rho)
            screening = rho >= min_rho
        # replacement: merge, move over-age individuals up, keep elites + best
        for i in range(layers):
            sel = layer_of == i
            g = np.concatenate([L[i].genomes, kids[sel]])
            f = np.concatenate([L[i].fitness, fit[sel]])
            a = np.concatenate([L[i].age + 1, kid_age[sel]])
            L[i] = Layer(g, f, a)
        for i in range(layers - 1):
            old = L[i].age > limits[i]
            if old.any():
                up = L[i + 1]
                L[i + 1] = Layer(np.concatenate([up.genomes, L[i].genomes[old]]),
                                 np.concatenate([up.fitness, L[i].fitness[old]]),
                                 np.concatenate([up.age, L[i].age[old]]))
                L[i] = Layer(L[i].genomes[~old], L[i].fitness[~old], L[i].age[~old])
        for i in range(layers):
            o = np.argsort(-L[i].fitness, kind="stable")
            n_elite = max(1, int(round(elitism * per)))
            keep = o[:max(per, n_elite)]
            L[i] = Layer(L[i].genomes[keep], L[i].fitness[keep], L[i].age[keep])
        if gen % age_gap == 0:  # ALPS: fresh random genomes into the youngest layer
            g = random_genomes(rng, per, voxels)
            f = sim.fitness(g)
            data_x.append(g)
            data_y.append(f)
            L[0] = Layer(g, f, np.zeros(per, np.int64))
        done = record()
        if verbose:
            print(f"gen {gen} evals {sim.evals - evals0} best {res.best:.5f} "
                  f"{'rho %.3f' % res.rho_history[-1] if res.rho_history else ''}", flush=True)
        if done:
            break
    return res


def surrogate_quality(n_train=200000, n_test=20000, voxels=8, duration=1.0, seed=0,
                      kinds=("mlp", "cnn"), epochs=None, chunk=4096, sim=None, verbose=True):
    """Simulate n_train + n_test random genomes, train each surrogate kind on
    the first n_train and report Spearman rho and MAE on the held-out rest."""
    rng = np.random.default_rng(seed)
    sim = sim or BatchSimulator(voxels=voxels, duration=duration)
    n = n_train + n_test
    G = np.zeros((n, voxels ** 3), np.uint8)
    Y = np.zeros(n, np.float64)
    t0 = time.perf_counter()
    for a in range(0, n, chunk):
        g = random_genomes(rng, min(chunk, n - a), voxels)
        Y[a:a + len(g)] = sim.fitness(g)
        G[a:a + len(g)] = g
        if verbose:
            print(f"simulated {a + len(g)}/{n}", flush=True)
    out = dict(n_train=n_train, n_test=n_test, fitness_mean=float(Y.mean()), fitness_std=float(Y.std()),
               data_seconds=time.perf_counter() - t0, sim_seconds=sim.sim_seconds)
    from .surrogate import mae
    for kind in kinds:
        t0 = time.perf_counter()
        sur = GenomeModel(kind, voxels, seed=seed, epochs=epochs)
        loss = sur.fit(G[:n_train], Y[:n_train], seed=seed)
        t_train = time.perf_counter() - t0
        pred = sur.predict(G[n_train:])
        out[kind] = dict(spearman=spearman(pred, Y[n_train:]), mae=mae(pred, Y[n_train:]),
                         final_train_loss=loss, train_seconds=t_train, epochs=sur.epochs)
        if verbose:
            print(kind, out[kind], flush=True)
    return out


def evals_to(history, threshold):
    """(evaluations, seconds) at which a run's best fitness first reached
    threshold, from RunResult.history; (-1, -1) if it never did."""
    for ev, sec, best in history:
        if best >= threshold:
            return ev, sec
    return -1, -1.0


def compare(seeds=5, threshold=None, pop=4096, max_generations=40, voxels=8, duration=1.0,
            verbose=True, seed_start=0, modes=(False, True), **kw):
    """Plain GA vs surrogate-screened GA over the given seeds, with the same
    seeds and evaluation counting.  With a threshold each run stops when its
    best fitness reaches it; without one every run goes max_generations and
    evaluations-to-threshold follow from the histories (evals_to)."""
    sim = BatchSimulator(voxels=voxels, duration=duration)
    out = []
    for s in range(seed_start, seed_start + seeds):
        for use in modes:
            r = run(seed=s, pop=pop, voxels=voxels, threshold=threshold, use_surrogate=use,
                    max_generations=max_generations, sim=sim, verbose=verbose, **kw)
            row = dict(seed=s, surrogate=use, evals_to_threshold=r.evals_to_threshold,
                       seconds_to_threshold=r.seconds_to_threshold, generations=r.generations,
                       best=r.best, fallback_generations=r.fallback_generations,
                       rho=[round(x, 3) for x in r.rho_history],
                       history=[(int(e), round(t, 2), float(b)) for e, t, b in r.history])
            out.append(row)
            if verbose:
                import json
                print(json.dumps(row), flush=True)
    return out


def report(rows, thresholds):
    """Mean evaluations (over the seeds that reached it) to each threshold,
    per loop, from compare() rows; returns {label: {threshold: (mean, reached, seeds)}}."""
    out = {}
    for row in rows:
        label = row.get("label") or ("surrogate" if row["surrogate"] else "plain")
        out.setdefault(label, []).append(row)
    table = {}
    for label, rs in out.items():
        table[label] = {}
        for thr in thresholds:
            ev = [evals_to(r["history"], thr)[0] for r in rs]
            hit = [e for e in ev if e >= 0]
            table[label][thr] = (float(np.mean(hit)) if hit else float("nan"), len(hit), len(rs))
    return table


def main(argv=None):
    import argparse
    import json
    ap = argparse.ArgumentParser(description="voxel-robot evolution with surrogate screening")
    sub = ap.add_subparsers(dest="cmd", required=True)
    q = sub.add_parser("quality", help="surrogate Spearman/MAE on held-out simulated genomes")
    q.add_argument("--train", type=int, default=200000)
    q.add_argument("--test", type=int, default=20000)
    q.add_argument("--epochs", type=int, default=None)
    q.add_argument("--kinds", default="mlp,cnn")
    q.add_argument("--seed", type=int, default=0)
    c = sub.add_parser("compare", help="evaluations to threshold with and without the surrogate")
    c.add_argument("--seeds", type=int, default=5)
    c.add_argument("--threshold", type=float, default=None)
    c.add_argument("--seed-start", type=int, default=0)
    c.add_argument("--pop", type=int, default=4096)
    c.add_argument("--generations", type=int, default=40)
    c.add_argument("--screen-factor", type=int, default=8, help="offspring scored per simulated one")
    c.add_argument("--epochs", type=int, default=None)
    c.add_argument("--surrogate", choices=("mlp", "cnn"), default="mlp")
    c.add_argument("--only", choices=("both", "plain", "surrogate"), default="both")
    c.add_argument("--audit", type=float, default=0.1, help="share of each layer's simulated "
                   "offspring drawn at random to measure the surrogate")
    c.add_argument("--warm-start", action="store_true", help="retrain from the previous weights")
    c.add_argument("--fallback-retrain", action=argparse.BooleanOptionalAction, default=True,
                   help="retrain every generation while the surrogate ranks below the fallback rho")
    c.add_argument("--mirror", action="store_true",
                   help="train the surrogate on every genome and its z-mirror, predict from both")
    c.add_argument("--max-train", type=int, default=200000,
                   help="each retrain uses the most recent this many simulated genomes")
    rp = sub.add_parser("report", help="evaluations-to-threshold table from compare logs")
    rp.add_argument("logs", nargs="+", help="[LABEL=]path of a compare log (JSON lines); rows are "
                    "labelled LABEL:plain / LABEL:surrogate")
    rp.add_argument("--thresholds", default="0.07,0.075,0.08,0.085,0.09,0.095,0.1")
    r = sub.add_parser("run", help="one evolution run")
    r.add_argument("--seed", type=int, default=0)
    r.add_argument("--pop", type=int, default=4096)
    r.add_argument("--generations", type=int, default=10)
    r.add_argument("--surrogate", action="store_true")
    for p in (q, c, r):
        p.add_argument("--voxels", type=int, default=8)
        p.add_argument("--duration", type=float, default=1.0)
    a = ap.parse_args(argv)
    if a.cmd == "report":
        rows = []
        for spec in a.logs:
            label, path = spec.split("=", 1) if "=" in spec else (None, spec)
            with open(path) as fh:
                for line in fh:
                    if line.startswith('{"seed"'):
                        row = json.loads(line)
                        kind = "surrogate" if row["surrogate"] else "plain"
                        row["label"] = f"{label}:{kind}" if label else kind
                        rows.append(row)
        thr = [float(x) for x in a.thresholds.split(",")]
        table = report(rows, thr)
        labels = list(table)
        print("| threshold | " + " | ".join(f"{l}: mean evals (seeds reached)" for l in labels) + " |")
        print("|---" * (len(labels) + 1) + "|")
        for t in thr:
            cells = [f"{table[l][t][0]:.0f} ({table[l][t][1]}/{table[l][t][2]})" for l in labels]
            print(f"| {t} | " + " | ".join(cells) + " |")
        return
    if a.cmd == "quality":
        res = surrogate_quality(a.train, a.test, voxels=a.voxels, duration=a.duration,
                                seed=a.seed, kinds=tuple(a.kinds.split(",")), epochs=a.epochs)
        print(json.dumps(res), flush=True)
    elif a.cmd == "compare":
        rows = compare(a.seeds, a.threshold, pop=a.pop, max_generations=a.generations,
                       voxels=a.voxels, duration=a.duration, screen_factor=a.screen_factor,
                       train_epochs=a.epochs, seed_start=a.seed_start, surrogate_kind=a.surrogate,
                       audit=a.audit, warm_start=a.warm_start, fallback_retrain=a.fallback_retrain,
                       mirror=a.mirror, max_train=a.max_train,
                       modes={"both": (False, True), "plain": (False,), "surrogate": (True,)}[a.only])
        for use in (False, True):
            e = [x["evals_to_threshold"] for x in rows if x["surrogate"] == use]
            print(json.dumps(dict(surrogate=use, evals_to_threshold=e,
                                  reached=sum(v >= 0 for v in e))), flush=True)
    else:
        res = run(seed=a.seed, pop=a.pop, voxels=a.voxels, duration=a.duration,
                  max_generations=a.generations, use_surrogate=a.surrogate, verbose=True)
        print(json.dumps(dict(best=res.best, history=res.history, rho=res.rho_history)), flush=True)


if __name__ == "__main__":
    main()
