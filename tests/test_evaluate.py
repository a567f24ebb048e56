"""Evaluator (Evaluator_impl.h:39-107) over the simulator."""
import numpy as np
import pytest

import mercury_amd as mh
from mercury_amd.evaluate import Evaluator


def test_devo_cycles_rejected():
    with pytest.raises(NotImplementedError):
        Evaluator(None, devo_cycles=1)


@pytest.mark.gpu
def test_batch_evaluate_equals_two_phase_reference_flow():
    """The reference runs a baseline phase, collects, resets every robot to
    its protoPos and uploads again; the evaluator uploads once.  Both give
    the same fitness, bit for bit."""
    print("evaluator: 6 robots 4^3, baseline 0.03 s, evaluation 0.05 s", flush=True)
    ms = mh.workload_materials()
    robots = mh.voxel_population(ms, 6, voxels_per_axis=4, seed=5)
    cfg = mh.SimConfig(solver=mh.MH_SOLVER_FORCE)
    ev = Evaluator(ms, cfg, baseline_time=0.03, evaluation_time=0.05)
    valid = np.array([1, 1, 0, 1, 1, 1], bool)
    fit = ev.batch_evaluate(robots, valid)
    assert ev.eval_count == 6 and fit[2] == 0

    sim = mh.Simulator(ms, cfg)
    live = [r for r, v in zip(robots, valid) if v]
    sim.Reset()
    sim.SetElements(live)
    sim.Simulate(0.03)
    phase1 = sim.Collect()
    again = []
    for r, c in zip(live, phase1):
        e = r.copy()                       # Update(c) then Reset(): back to protoPos,
        e.pairs, e.spring_material = c.pairs, c.spring_material  # springs as collected
        again.append(e)
    sim.Reset()
    sim.SetElements(again)
    sim.Simulate(0.05)
    ref = sim.fitness(mh.MH_FITNESS_REFERENCE)
    assert np.array_equal(fit[valid], ref)
    assert (fit >= 0).all()
