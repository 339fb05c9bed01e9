"""Batched frames and the tile autotuner.

Graph replay: a plan with max_batch = B replays one captured hipGraph per
B frames; a plan with max_batch = 1 launches every frame kernel by kernel.
The graph runs the sort and spread of frame f + 1 beside the y and x passes
of frame f, with the spread output alternating between two buffers, so B = 3
(odd: the first frame of the next replay reuses the buffer of the last
frame of this one) and B = 4 are both checked.
The kernels and their order are the same; the fine grids agree bit for bit
and only the order of the FP64 LDS additions inside a workgroup's shell
accumulators may differ, so the accumulators must agree to 1e-12 relative (a
dropped, repeated or stale frame changes them at the 1e-1 level).  The graph
is replayed several times with eager work and synchronisations in between,
and every replay is checked.

Autotuner: a plan with tune on picks its tiles on the first frame; its
tables must match the untuned plan (the summation order inside a tile
changes with the tile, so to rounding) and the FP64 oracle.
"""

import numpy as np
import pytest

from helpers import oracle_run, random_case, rel_dev

pytestmark = pytest.mark.gpu


def _plan(cell, origin, types, precision, **kw):
    from sqcalc_amd import HipStructureFactor
    return HipStructureFactor(cell, origin, types, {1: "Si", 2: "O"}, weight="neutron",
                              norm="mean", qmax=5.0, nq=60, precision=precision,
                              eps=1e-10 if precision == "double" else None, **kw)


@pytest.mark.parametrize("precision,batch", [("double", 4), ("single", 4), ("single", 3)])
def test_graph_replay_matches_eager(hip_device, precision, batch):
    import torch
    cell, origin, types, frames = random_case(natoms=1500, length=24.0, nframes=10, seed=31)
    pos = torch.tensor(frames, dtype=torch.float64, device=hip_device).contiguous()
    print("case graph replay", precision, batch, flush=True)
    eager = _plan(cell, origin, types, precision, max_batch=1)
    eager.accumulate_device(pos)
    ref_num, ref_part, _, ref_n = eager.fetch()
    assert ref_n == 10
    graph = _plan(cell, origin, types, precision, max_batch=batch)
    for rep in range(3):
        graph.reset()
        graph.accumulate_device(pos)          # 10 // batch replays, the rest launched directly
        # eager work of the program between replays
        junk = (pos * 2.0).sum()
        torch.cuda.synchronize()
        num, part, _, n = graph.fetch()
        assert n == 10
        np.testing.assert_allclose(num, ref_num, rtol=1e-12, atol=0, err_msg=str(rep))
        np.testing.assert_allclose(part, ref_part, rtol=1e-12, atol=1e-12 * np.abs(ref_part).max(),
                                   err_msg=str(rep))
        assert float(junk) != 0.0
    # host positions: uploaded into the staging buffer, replayed without a copy
    graph.reset()
    graph.accumulate(frames[:8])
    num, part, _, n = graph.fetch()
    eager.reset()
    eager.accumulate(frames[:8])
    e_num, e_part, _, _ = eager.fetch()
    assert n == 8
    np.testing.assert_allclose(num, e_num, rtol=1e-12, atol=0)
    np.testing.assert_allclose(part, e_part, rtol=1e-12, atol=1e-12 * np.abs(e_part).max())
    eager.close()
    graph.close()


@pytest.mark.parametrize("precision", ["double", "single"])
def test_autotuned_plan_matches(hip_device, precision):
    cell, origin, types, frames = random_case(natoms=1500, length=24.0, nframes=3, seed=32)
    print("case tune", precision, flush=True)
    base = _plan(cell, origin, types, precision)
    base.accumulate(frames)
    tuned = _plan(cell, origin, types, precision, tune=True)
    tuned.accumulate(frames)
    t = list(tuned.info.tile)
    assert (t[0], t[1]) in [(16, 2), (16, 1), (8, 2), (8, 1), (4, 2), (4, 1), (2, 1)]
    assert t[2] in (2, 4, 8, 16)
    a, b = base.table(), tuned.table()
    tol_plans = 1e-12 if precision == "double" else 1e-6
    assert rel_dev(a[:, 1:], b[:, 1:]) < tol_plans
    run = oracle_run(cell, origin, types, {1: "Si", 2: "O"}, "neutron", "mean", 0.0, 5.0, 60,
                     frames)
    ref = run.table()
    tol = 1e-9 if precision == "double" else 1e-6
    dev = rel_dev(ref[:, 1:], b[:, 1:])
    print("  tiles %s, max deviation from the oracle %.2e" % (t, dev))
    assert dev < tol
    base.close()
    tuned.close()
