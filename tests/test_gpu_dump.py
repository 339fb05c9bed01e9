"""A LAMMPS dump streamed through the native reader into the device path.

``sqcalc_amd.dump.feed`` parses batch k + 1 on a second thread while the
library runs batch k.  Its accumulators must equal those of the same frames
handed to ``accumulate`` directly, to 1e-12 relative: the fine grids are
bit-identical, only the order of the FP64 LDS additions into a workgroup's
shell sums may differ, while a dropped, repeated or stale frame moves them at
the 1e-1 level.  The S(q) table must match the FP64 oracle on the frames the
reader returned.
"""

import numpy as np
import pytest

from helpers import oracle_run, rel_dev

pytestmark = pytest.mark.gpu


def _write_dump(path, cell, types, frames):
    n = len(types)
    with open(path, "w") as fh:
        for f, pos in enumerate(frames):
            fh.write("ITEM: TIMESTEP\n%d\nITEM: NUMBER OF ATOMS\n%d\n" % (f, n))
            fh.write("ITEM: BOX BOUNDS pp pp pp\n")
            fh.write("".join("0.0 %.10f\n" % cell[d, d] for d in range(3)))
            fh.write("ITEM: ATOMS id type x y z\n")
            w = np.mod(pos, np.diag(cell))
            fh.write("".join("%d %d %.10f %.10f %.10f\n" % (i + 1, types[i], *w[i])
                             for i in range(n)))


@pytest.mark.parametrize("precision", ["double", "single"])
def test_feed_matches_accumulate_and_oracle(hip_device, tmp_path, precision):
    from sqcalc_amd import HipStructureFactor
    from sqcalc_amd.dump import DumpReader, feed
    rng = np.random.default_rng(17)
    L, n, nf = 20.0, 1000, 10
    cell = np.diag([L, L, L])
    types = 1 + (np.arange(n) % 3 == 0).astype(int)
    pos = rng.random((n, 3)) * L
    frames = []
    for _ in range(nf):
        frames.append(pos.copy())
        pos = pos + rng.normal(0.0, 0.4, pos.shape)
    path = tmp_path / "traj.dump"
    _write_dump(path, cell, types, frames)
    with DumpReader(path, 4) as rd:
        read = list(rd)
    assert len(read) == nf
    print("case dump feed", precision, flush=True)

    def plan():
        return HipStructureFactor(read[0].cell, read[0].origin, read[0].type_id,
                                  {1: "Si", 2: "O"}, weight="neutron", norm="mean", qmax=6.0,
                                  nq=80, precision=precision,
                                  eps=1e-10 if precision == "double" else None, max_batch=4)
    fed = plan()
    assert feed(fed, path, batch=4, nthreads=4) == nf
    direct = plan()
    direct.accumulate(np.stack([fr.pos for fr in read]))
    a, b = fed.fetch(), direct.fetch()
    assert a[3] == b[3] == nf
    np.testing.assert_allclose(a[0], b[0], rtol=1e-12, atol=0)
    np.testing.assert_allclose(a[1], b[1], rtol=1e-12, atol=1e-12 * np.abs(b[1]).max())
    run = oracle_run(read[0].cell, read[0].origin, read[0].type_id, {1: "Si", 2: "O"}, "neutron",
                     "mean", 0.0, 6.0, 80, [fr.pos for fr in read])
    tol = 1e-9 if precision == "double" else 1e-5
    assert rel_dev(run.table(), fed.table()) < tol
    fed.close()
    direct.close()
    # the plan from the first frame, the whole file through accumulate_dump
    from sqcalc_amd import HipStructureFactor as H
    sf = H.from_frame(read[0], {1: "Si", 2: "O"}, weight="neutron", norm="mean", qmax=6.0,
                      nq=80, precision=precision, eps=1e-10 if precision == "double" else None,
                      max_batch=4)
    assert sf.accumulate_dump(path, batch=3) == nf
    c = sf.fetch()
    assert c[3] == nf
    np.testing.assert_allclose(c[0], b[0], rtol=1e-12, atol=0)
    sf.close()
