"""Shared Debye cases: the configurations of the reference's
test/test_neighbour_list.f90 (orthogonal, triclinic, slab, cluster, small box
with several images of a pair, skin reuse) and a few more, as trajectories."""

import numpy as np

from oracle import oracle as O


def make_cell(lengths, tilt=(0.0, 0.0, 0.0)):
    """Columns a1 = (lx, 0, 0), a2 = (xy, ly, 0), a3 = (xz, yz, lz)."""
    lx, ly, lz = lengths
    xy, xz, yz = tilt
    return np.array([[lx, xy, xz], [0.0, ly, yz], [0.0, 0.0, lz]])


def trajectory(rng, cell, origin, n, nframes, sigma, pbc=(1, 1, 1), wrap=True):
    """Uniform atoms, rattled by sigma per frame; along non-periodic axes the
    atoms stay within [0.1, 0.9] of the cell."""
    s = rng.uniform(0.0, 1.0, (n, 3))
    for k in range(3):
        if not pbc[k]:
            s[:, k] = rng.uniform(0.1, 0.9, n)
    frames = []
    for _ in range(nframes):
        f = s.copy()
        if wrap:
            for k in range(3):
                if pbc[k]:
                    f[:, k] -= np.floor(f[:, k])
        frames.append(f @ cell.T + origin)
        step = rng.normal(0.0, sigma, (n, 3)) @ np.linalg.inv(cell).T
        for k in range(3):
            if not pbc[k]:
                step[:, k] = 0.0
        s = s + step
    return np.array(frames)


CASES = {
    # name: lengths, tilt, origin, pbc, natoms, rmax, skin, frames, sigma
    "orthogonal": ((20.0, 20.0, 20.0), (0, 0, 0), (0, 0, 0), (1, 1, 1), 300, 5.0, 1.0, 3, 0.3),
    "triclinic": ((18.0, 17.0, 21.0), (2.5, -1.5, 3.0), (1.5, -2.0, 0.25), (1, 1, 1), 300, 5.0,
                  1.0, 3, 0.3),
    "slab": ((18.0, 18.0, 14.0), (0, 0, 0), (0, 0, 0), (1, 1, 0), 300, 4.0, 0.0, 2, 0.2),
    "cluster": ((24.0, 24.0, 24.0), (0, 0, 0), (0, 0, 0), (0, 0, 0), 200, 6.0, 0.0, 2, 0.0),
    "small_box": ((8.0, 8.0, 8.0), (0, 0, 0), (0, 0, 0), (1, 1, 1), 120, 5.0, 0.5, 3, 0.4),
    "skin_reuse": ((20.0, 20.0, 20.0), (0, 0, 0), (0, 0, 0), (1, 1, 1), 400, 5.0, 1.0, 20, 0.02),
    "wide_cut": ((20.0, 21.0, 22.0), (3.0, 0, -2.0), (0, 0, 0), (1, 1, 1), 400, 9.9, 0.0, 2, 1.5),
    # a general cell (no zero entries): the triclinic one turned about (1, 2, 2) / 3
    "rotated": ((18.0, 17.0, 21.0), (2.5, -1.5, 3.0), (0.5, 0.5, -1.0), (1, 1, 1), 300, 6.0, 1.0,
                3, 0.3),
}


def rotation(axis, angle):
    axis = np.asarray(axis, dtype=np.float64) / np.linalg.norm(axis)
    k = np.array([[0.0, -axis[2], axis[1]], [axis[2], 0.0, -axis[0]], [-axis[1], axis[0], 0.0]])
    return np.eye(3) + np.sin(angle) * k + (1.0 - np.cos(angle)) * (k @ k)


def run_case(name, seed=3, ntypes=3):
    lengths, tilt, origin, pbc, n, rmax, skin, nf, sigma = CASES[name]
    rng = np.random.default_rng(seed)
    cell = make_cell(lengths, tilt)
    if name == "rotated":
        cell = rotation((1.0, 2.0, 2.0), 0.7) @ cell
    origin = np.asarray(origin, dtype=np.float64)
    frames = trajectory(rng, cell, origin, n, nf, sigma, pbc)
    types = rng.integers(1, ntypes + 1, n).astype(np.int32)
    return cell, origin, pbc, types, ntypes, rmax, skin, frames


def oracle_hist(cell, origin, pbc, types, ntypes, rmax, dr, skin, frames, method=1):
    run = O.DebyeRun(cell, origin, types, ntypes, pbc=pbc, rmax=rmax, dr=dr, skin=skin)
    for f in frames:
        run.add_frame(f, method)
    return run


def lattice(nside, spacing):
    g = np.arange(nside) * spacing
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    return np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
