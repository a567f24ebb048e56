"""Evaluator: the reference's batch evaluation over the simulator.

Follows Evaluator<T>::BatchEvaluate (optimize/optimizer/Evaluator_impl.h:39-107):
Reset -> SetElements -> [devo cycles] -> Simulate(baselineTime) -> Collect ->
Update + Reset -> SetElements -> Simulate(evaluationTime) -> Collect ->
updateFitness, counting one evaluation per solution (lines 103-106).

With no devo cycles the first phase cannot change the result: each
solution's Update takes the collected masses and springs, and the following
SoftBody::Reset (SoftBody.cpp:68-80) puts every mass back at its protoPos
with zero velocity and recomputes the mean lengths from protoPos, so the
second SetElements uploads exactly what the first one did.  The population
is therefore uploaded once and the discarded phase is not simulated.
Devo (stress-sorted spring replacement) is not part of this library, so
devo_cycles > 0 is rejected.
"""
import numpy as np

from . import _lib
from .simulator import SimConfig, Simulator


class Evaluator:
    """Evaluator_impl.h:13-36 defaults: baselineTime 5 s, evaluationTime 10 s."""

    def __init__(self, material_set, config=None, baseline_time=5.0, evaluation_time=10.0,
                 devo_cycles=0):
        if devo_cycles:
            raise NotImplementedError("devo cycles (spring replacement) are not supported")
        self.sim = Simulator(material_set, config or SimConfig())
        self.baseline_time = float(baseline_time)
        self.evaluation_time = float(evaluation_time)
        self.eval_count = 0

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
            self.sim.Reset()
            self.sim.SetElements([robots[i] for i in idx])
            self.sim.Simulate(self.evaluation_time)
            fit[idx] = self.sim.fitness(mode)
        self.eval_count += n
        return fit

    def collect(self):
        """Robots as they stand after the last batch_evaluate."""
        return self.sim.Collect()
