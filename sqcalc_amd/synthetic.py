"""Seeded synthetic trajectories of the benchmark configurations.

There are no datasets on the box, so every workload is generated from a
seed with numpy's PCG64.  The recipes follow the configurations the path is
measured on:

  c1  cube L=20, N=1000, types 1:2 (333/667), uniform, 10 frames, q_max 10
  c2  cube L=113, N=96,000 (32,000 Si + 64,000 O) on a seeded subset of the
      46^3 simple-cubic lattice + N(0, 0.3^2) jitter per frame, 100 frames,
      neutron weights, partials, q_max 10
  c3  cube L=80, N=32,000 (50/50), uniform start, random walk sigma 0.05 per
      frame, 1000 frames, unit weights, q_max 8
  c4  cube L=44, N=8,000 of 3 types (40/40/20), uniform; target from the
      jittered 20^3 lattice of a second seed; q_max 6
  c5  triclinic a=b=c=250, xy=10, xz=5, N=10^6 of 4 types (Si, O, Na, Ca at
      0.25/0.55/0.12/0.08), jittered 100^3 fractional lattice sigma 0.4,
      20 frames, xray weights, q_max 4
"""

import numpy as np


def type_assignment(natoms, fractions, rng):
    """Type ids 1..len(fractions), rounded counts, last type takes the rest."""
    counts = [int(round(f * natoms)) for f in fractions[:-1]]
    counts.append(natoms - sum(counts))
    types = np.repeat(np.arange(1, len(fractions) + 1), counts)
    rng.shuffle(types)
    return types


def lattice_sites(cell, side):
    g = (np.arange(side) + 0.5) / side
    s = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    return s @ np.asarray(cell).T


class Workload:
    def __init__(self, name, cell, origin, types, mapping, weight, norm, qmax, nq=500,
                 qmin=0.0, frames=None, frame_fn=None, nframes=0, precision="single"):
        self.name, self.cell, self.origin = name, np.asarray(cell, float), np.asarray(origin, float)
        self.types, self.mapping, self.weight, self.norm = types, mapping, weight, norm
        self.qmin, self.qmax, self.nq = qmin, qmax, nq
        self._frames, self._frame_fn, self.nframes = frames, frame_fn, nframes
        self.precision = precision

    @property
    def natoms(self):
        return len(self.types)

    def frames(self, start=0, count=None):
        count = self.nframes - start if count is None else count
        if self._frames is not None:
            return self._frames[start:start + count]
        return np.stack([self._frame_fn(k) for k in range(start, start + count)])


def config(name, seed=2024):
    rng = np.random.default_rng(seed)
    if name == "c1":
        L, N = 20.0, 1000
        cell = np.diag([L] * 3)
        types = type_assignment(N, [1 / 3, 2 / 3], rng)
        frames = rng.random((10, N, 3)) * L
        return Workload(name, cell, np.zeros(3), types, {1: "Si", 2: "O"}, "unit", "mean", 10.0,
                        frames=frames, nframes=10, precision="double")
    if name == "c2":
        L, N, side, sig = 113.0, 96000, 46, 0.3
        cell = np.diag([L] * 3)
        sites = lattice_sites(cell, side)
        base = sites[rng.permutation(len(sites))[:N]]
        types = np.array([1] * 32000 + [2] * 64000)
        rng.shuffle(types)

        def frame(k):
            r = np.random.default_rng([seed, 2, k])
            return base + r.normal(0.0, sig, base.shape)
        return Workload(name, cell, np.zeros(3), types, {1: "Si", 2: "O"}, "neutron", "mean",
                        10.0, frame_fn=frame, nframes=100)
    if name == "c3":
        L, N, sig = 80.0, 32000, 0.05
        cell = np.diag([L] * 3)
        types = type_assignment(N, [0.5, 0.5], rng)
        start = rng.random((N, 3)) * L
        steps = np.random.default_rng([seed, 3])

        cache = {}

        def frame(k):
            # random walk: frame k = frame k-1 + N(0, sig^2), built incrementally
            if not cache:
                cache[0] = start.copy()
                cache["last"] = 0
            while cache["last"] < k:
                cache[cache["last"] + 1] = cache[cache["last"]] + steps.normal(0.0, sig, (N, 3))
                cache.pop(cache["last"] - 1, None)
                cache["last"] += 1
            return cache[k] % L
        return Workload(name, cell, np.zeros(3), types, None, "unit", "mean", 8.0,
                        frame_fn=frame, nframes=1000)
    if name == "c4":
        L, N = 44.0, 8000
        cell = np.diag([L] * 3)
        types = type_assignment(N, [0.4, 0.4, 0.2], rng)
        frames = rng.random((1, N, 3)) * L
        return Workload(name, cell, np.zeros(3), types, {1: "Si", 2: "O", 3: "Na"}, "unit",
                        "mean", 6.0, frames=frames, nframes=1)
    if name == "c4_target":
        L, N = 44.0, 8000
        cell = np.diag([L] * 3)
        r2 = np.random.default_rng(seed + 1)
        types = type_assignment(N, [0.4, 0.4, 0.2], r2)
        frames = (lattice_sites(cell, 20) + r2.normal(0.0, 0.3, (N, 3)))[None]
        return Workload(name, cell, np.zeros(3), types, {1: "Si", 2: "O", 3: "Na"}, "unit",
                        "mean", 6.0, frames=frames, nframes=1)
    if name == "c5":
        a = 250.0
        cell = np.array([[a, 10.0, 5.0], [0.0, a, 0.0], [0.0, 0.0, a]])
        N, side, sig = 1000000, 100, 0.4
        types = type_assignment(N, [0.25, 0.55, 0.12, 0.08], rng)
        base = lattice_sites(cell, side)

        def frame(k):
            r = np.random.default_rng([seed, 5, k])
            return base + r.normal(0.0, sig, base.shape)
        return Workload(name, cell, np.zeros(3), types, {1: "Si", 2: "O", 3: "Na", 4: "Ca"},
                        "xray", "mean", 4.0, frame_fn=frame, nframes=20)
    raise KeyError(name)
