"""Evaluator (Evaluator_impl.h:39-107) over the simulator."""
import numpy as np
import pytest

import mercury_amd as mh
from mercury_amd.devo import rest_from_proto
from mercury_amd.evaluate import Evaluator


def test_evaluator_uploads_reset_robots():
    """BatchEvaluate resets every solution before uploading it
    (Evaluator_impl.h:52-53): its springs' lengths are recomputed from the
    proto positions, which moves some of the built lengths by an ulp."""
    ms = mh.workload_materials()
    r = mh.voxel_population(ms, 1, voxels_per_axis=4, seed=5)[0]
    reset = rest_from_proto(r, r.pos)
    assert np.array_equal(reset.mean_length.view(np.uint32),
                          mh.robots.rest_lengths(r.pos, r.pairs).view(np.uint32))
    assert np.abs(reset.mean_length - r.mean_length).max() <= 1e-6


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
    live = [rest_from_proto(r, r.pos) for r, v in zip(robots, valid) if v]   # R.Reset()
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
