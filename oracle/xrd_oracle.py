"""Powder XRD restatement for the oracle (test infrastructure only).

Restates, per mode in plain loops, the reference's
  - two_theta_deg and lorentz_polarization (src/sqc_structure_factor.f90:906-923),
  - sf_configure_xrd: bin width from the box, bin of every kept mode and its
    Lorentz-polarization weight (:846-904, cell widths src/sqc_cell.f90:102-118),
  - sf_accumulate_values' XRD sum: xrd_num(bin) += value * weight (:595-643),
  - sf_accumulate_xrd_partials: xrd_partial(a, b, bin) += w_a w_b Re(rho_a
    rho_b*) weight, a <= b (:538-588),
  - and the per-atom, per-frame normalisation of sf_write_xrd /
    sf_xrd_partial_value (:937-1024).
It is driven by the oracle's direct sum rho_a(q) (oracle.direct_rho).
Pinned in tests/test_oracle.py by the reference's own known answers
(test/test_xrd.f90) and by the pair sum rule I = sum_a I_aa + 2 sum_{a<b} I_ab.
"""

import math

import numpy as np

from . import oracle as O


def two_theta_deg(q, lam):
    s = q * lam / (4.0 * math.pi)
    s = max(-1.0, min(1.0, s))
    return 2.0 * math.asin(s) * 180.0 / math.pi


def lorentz_polarization(tth_deg):
    tth = tth_deg * math.pi / 180.0
    th = 0.5 * tth
    return (1.0 + math.cos(tth) ** 2) / (math.sin(th) ** 2 * math.cos(th))


def _widths(cell):
    a = [[cell[i][j] for j in range(3)] for i in range(3)]
    col = [[a[0][k], a[1][k], a[2][k]] for k in range(3)]

    def cross(u, v):
        return [u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]]

    c = cross(col[1], col[2])
    vol = abs(col[0][0] * c[0] + col[0][1] * c[1] + col[0][2] * c[2])
    out = []
    for i in range(3):
        n = cross(col[(i + 1) % 3], col[(i + 2) % 3])
        out.append(vol / math.sqrt(n[0] ** 2 + n[1] ** 2 + n[2] ** 2))
    return out


class Xrd:
    """XRD accumulation over an oracle Run's modes."""

    def __init__(self, run, lam, tth_min=1.0, tth_max=179.0, step=0.0, lp=True,
                 pbc=(True, True, True)):
        self.run, self.lam, self.lp = run, lam, lp
        self.tth_min, self.tth_max = tth_min, tth_max
        if step <= 0.0:
            w = _widths(np.asarray(run.cell).tolist())
            bound = min(w[i] for i in range(3) if pbc[i])
            theta_max = min(0.5 * math.pi / 180.0 * tth_max, 60.0 * math.pi / 180.0)
            step = (lam / (bound * math.cos(theta_max))) / (math.pi / 180.0)
        self.bins = max(1, int(math.ceil((tth_max - tth_min) / step)))
        self.step = (tth_max - tth_min) / self.bins
        qlen = run.md["qlen"]
        self.bin = []
        self.weight = []
        self.count = [0] * self.bins
        for q in qlen:
            tth = two_theta_deg(float(q), lam)
            b = int((tth - tth_min) / self.step) + 1
            b = min(max(b, 1), self.bins)
            self.bin.append(b - 1)
            self.count[b - 1] += 1
            self.weight.append(lorentz_polarization(tth) if lp else 1.0)
        S = len(run.species_type)
        self.num = [0.0] * self.bins
        self.partial = np.zeros((S, S, self.bins))
        self.nframes = 0

    def add_frame(self, pos):
        run = self.run
        rho = O.direct_rho(pos, run.cell, run.origin, run.species_of, len(run.species_type),
                           run.md)
        S, K = rho.shape
        amps = [O.amplitude(run.weight, run.amp, run.xcoef, a, run.md["qlen"]) for a in range(S)]
        for m in range(K):
            t = sum(amps[a][m] * rho[a, m] for a in range(S))
            value = t.real * t.real + t.imag * t.imag
            self.num[self.bin[m]] += value * self.weight[m]
        for a in range(S):
            for b in range(a, S):
                prod = (rho[a] * np.conj(rho[b])).real
                for m in range(K):
                    self.partial[a, b, self.bin[m]] += (amps[a][m] * amps[b][m] * prod[m]
                                                        * self.weight[m])
        self.nframes += 1

    def table(self):
        """(two_theta, I, [pair columns, a <= b]) per atom and frame."""
        f = float(max(self.nframes, 1)) * self.run.natoms
        tth = [self.tth_min + (b + 0.5) * self.step for b in range(self.bins)]
        S = self.partial.shape[0]
        pairs = [self.partial[a, b] / f for a in range(S) for b in range(a, S)]
        return np.array(tth), np.array(self.num) / f, pairs
