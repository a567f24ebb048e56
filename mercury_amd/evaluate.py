"""Evaluator: the reference's batch evaluation over the simulator.

Follows Evaluator<T>::BatchEvaluate (optimize/optimizer/Evaluator_impl.h:39-107):
Reset -> SetElements -> [devo cycles] -> Simulate(baselineTime) -> Collect ->
Update + Reset -> SetElements -> Simulate(evaluationTime) -> Collect ->
updateFitness, counting one evaluation per solution (lines 103-106).

Every solution is Reset before its first upload (SoftBody::Reset,
SoftBody.cpp:68-80): masses at their protoPos with zero velocity and every
spring's mean length recomputed from protoPos (for a built voxel robot this
moves about a quarter of the lengths by one ulp, since the builder measures
them before ShiftX/ShiftY).  With no devo cycles the first phase cannot
change the result: each solution's Update takes the collected masses and
springs, and the following Reset restores exactly what the first
SetElements uploaded.  The population is therefore uploaded once and the
discarded phase is not simulated.

With devo cycles (Evaluator_impl.h:66-70), each cycle simulates devo_time
seconds tracking stress and then replaces every robot's most stressed
springs (mercury_amd.devo), re-uploading the population with the current
state, the simulated time and the accumulated stress carried over.  The springs found this way are
what the robots keep: Update takes the collected springs and
SoftBody::Reset recomputes their lengths from the proto positions, so the
evaluation phase starts from the rest configuration with the new
topology.  The baseline phase still changes nothing and is skipped.
"""
import numpy as np

from . import _lib
from .devo import MaterialLookup, devo_robot, rest_from_proto
from .simulator import SimConfig, Simulator


class Evaluator:
    """Evaluator_impl.h:13-36 defaults: baselineTime 5 s, evaluationTime 10 s,
    devo_time 1 s, no devo cycles (config.h:24-25); 128 replaced springs per
    robot (config.h:18)."""

    def __init__(self, material_set, config=None, baseline_time=5.0, evaluation_time=10.0,
                 devo_cycles=0, devo_time=1.0, replaced_springs=128, seed=0):
        self.sim = Simulator(material_set, config or SimConfig())
        self.lookup = MaterialLookup(material_set)
        self.baseline_time = float(baseline_time)
        self.evaluation_time = float(evaluation_time)
        self.devo_cycles = int(devo_cycles)
        self.devo_time = float(devo_time)
        self.replaced_springs = int(replaced_springs)
        self.seed = int(seed)
        self.eval_count = 0
        self._batches = 0

    def devo(self, robots):
        """The devo cycles on a batch of robots; returns the robots with their
        replaced springs, lengths reset to the proto positions."""
        sim = self.sim
        sim.Reset()
        sim.SetElements(robots)
        current = list(robots)
        for c in range(self.devo_cycles):
            sim.Simulate(self.devo_time, True)
            t = sim.getTotalTime()
            state = sim.Collect()
            new = []
            for r, (el, st) in enumerate(zip(current, state)):
                rng = np.random.default_rng([self.seed, self._batches, c, r])
                nxt, _ = devo_robot(el, st.stress, st.pos, t, self.replaced_springs, rng, self.lookup)
                nxt.pos, nxt.vel = st.pos, st.vel
                new.append(nxt)
            sim.SetElements(new)
            sim.setTotalTime(t)
            # the stress keeps accumulating over the cycles, as the
            # reference's buffer does (cleared only by SetElements); a
            # replaced spring keeps its slot's stress
            sim.set_stress(np.concatenate([st.stress for st in state]) if state else np.zeros(0, np.float32))
            current = new
        return [rest_from_proto(el, src.pos) for el, src in zip(current, robots)]

    def batch_evaluate(self, robots, valid=None, mode=_lib.MH_FITNESS_REFERENCE):
        """Fitness of every robot after evaluation_time seconds, from its
        rest configuration.  Invalid robots (valid[i] False) are not
        simulated and score 0, as SoftBody::updateFitness gives them; every
        robot counts one evaluation."""
        n = len(robots)
        fit = np.zeros(n, np.float32)
        if n == 0:
            return fit
        valid = np.ones(n, bool) if valid is None else np.asarray(valid, bool)
        idx = np.flatnonzero(valid)
        if len(idx):
            batch = [rest_from_proto(robots[i], robots[i].pos) for i in idx]
            if self.devo_cycles:
                batch = self.devo(batch)
                self.last_devo = batch
            self.sim.Reset()
            self.sim.SetElements(batch)
            self.sim.Simulate(self.evaluation_time)
            fit[idx] = self.sim.fitness(mode)
        self.eval_count += n
        self._batches += 1
        return fit

    def collect(self):
        """Robots as they stand after the last batch_evaluate."""
        return self.sim.Collect()
