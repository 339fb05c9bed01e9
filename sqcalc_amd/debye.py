"""Debye method on one HIP device: pair-distance histograms and what the
reference evaluates from them.

The ordered pair counts histogram(a, b, k) per frame come from the library
(sqc_hip_debye_* in include/sqc_hip.h); this module mirrors the rest of
debye_structure_factor_t (src/sqc_debye.f90:45-93) on the host:

* set-up as debye_setup (:95-177): ntypes from the mapping and the type ids,
  default rmax half the smallest periodic width (the configuration's span
  when no axis is periodic), nbins = max(1, ceiling(rmax / dr));
* ``structure_factor`` as debye_prepare_output (:236-345) with
  sf_shell_value / sf_partial_value (src/sqc_structure_factor.f90:657-719):
  the pair sum over bin centres, the r = 0 self term, debyer's cut-off
  correction (shared by the partials in proportion to x_a x_b);
* ``rdf`` as debye_write_rdf / debye_fill_g_partial (:347-452);
* ``pair_entropy`` as debye_write_pair_entropy (:454-560).
Columns and labels follow sf_write_results (present type pairs t <= u).
"""

import ctypes
import math

import numpy as np

from . import _lib, elements


def shell_sums(hist, nhist, counts, volume, pbc, rmax, dr, amplitude, norm="mean", partials=True,
               correct_cutoff=True, qmin=0.0, qmax=20.0, nq=500):
    """(q, num, den, partial_num[a][b][nq]) of debye_prepare_output
    (src/sqc_debye.f90:236-345) from nhist frames of ordered pair counts
    hist[a][b][k]; amplitude(t, q) is the weight of type index t.  The
    table is num / (nframes den) (sf_shell_value), partial_num / (nframes N)
    (sf_partial_value with no lattice vectors per shell)."""
    if nhist <= 0:
        raise RuntimeError("no frames were accumulated")
    hist = np.asarray(hist, dtype=np.float64)
    T, _, nb = hist.shape
    counts = np.asarray(counts)
    n = float(np.sum(counts))
    dq = (qmax - qmin) / nq
    q = qmin + (np.arange(1, nq + 1) - 0.5) * dq
    r = (np.arange(1, nb + 1) - 0.5) * dr
    four_pi = 4.0 * math.acos(-1.0)
    rho0 = n / volume
    cut = correct_cutoff and any(pbc)
    live = counts > 0
    avg_h = hist / float(nhist)
    num, den = np.zeros(nq), np.zeros(nq)
    part = np.zeros((T, T, nq))
    for s in range(nq):
        qs = float(q[s])
        w = np.array([amplitude(t, qs) for t in range(T)])
        sinc = np.sin(qs * r) / (qs * r)
        # sum_k n_ab[k] sinc(q r_k) per type pair (empty bins add nothing)
        pair = avg_h @ sinc
        pair_sum = float(np.sum((w[:, None] * w[None, :] * pair)[np.ix_(live, live)]))
        self_sum = float(np.sum(counts[live] * w[live] * w[live]))
        weight_sum = float(np.sum(counts * w))
        if norm == "n":
            den[s] = n
        elif norm == "mean":
            den[s] = weight_sum * weight_sum / n
        else:
            den[s] = self_sum
        shape = 0.0
        if cut:
            qc = qs * rmax
            shape = four_pi / (qs * qs) * (rmax * math.cos(qc) - math.sin(qc) / qs)
        if partials:
            for a in range(T):
                if not live[a]:
                    continue
                for b in range(a, T):
                    if not live[b]:
                        continue
                    ps = float(pair[a, b])
                    if a == b:
                        ps += float(counts[a])
                    # the cut-off correction shared in proportion to x_a x_b
                    ps += (counts[a] / n) * (counts[b] / n) * (rho0 * shape) * n
                    part[a, b, s] = ps * nhist
        corr = (weight_sum / n) ** 2 * rho0 * shape if cut else 0.0
        num[s] = (pair_sum + self_sum + corr * n) * nhist if den[s] > 0.0 else 0.0
    return q, num, den, part


class HipDebye:
    """Pair histograms of a trajectory on one HIP device.

    cell: 3x3 with a1, a2, a3 as columns; types: LAMMPS type id per atom
    (1-based); mapping: {type_id: species label} for neutron / xray weights;
    pbc: periodicity per cell vector; rmax 0 (or None) takes the reference's
    default, which needs ``first_frame`` when no axis is periodic.
    """

    def __init__(self, cell, origin, types, mapping=None, pbc=(True, True, True), rmax=None,
                 dr=0.01, skin=1.0, weight="unit", norm="mean", partials=True,
                 faber_ziman=False, correct_cutoff=True, first_frame=None, device=0):
        self.lib = _lib.load()
        self.cell = np.ascontiguousarray(cell, dtype=np.float64).reshape(3, 3)
        self.origin = np.ascontiguousarray(origin, dtype=np.float64).reshape(3)
        types = np.ascontiguousarray(types, dtype=np.int32).reshape(-1)
        self.types = types
        self.natoms = int(types.size)
        self.mapping = {int(t): elements.canonical(lab) for t, lab in dict(mapping or {}).items()}
        self.ntypes = max(max(self.mapping) if self.mapping else 1, int(types.max()))
        self.counts = np.array([np.count_nonzero(types == t) for t in range(1, self.ntypes + 1)])
        self.pbc = tuple(bool(p) for p in pbc)
        self.volume = abs(float(np.linalg.det(self.cell)))
        self.weight, self.norm = weight, norm
        self.partials, self.faber_ziman = bool(partials), bool(faber_ziman)
        self.correct_cutoff = bool(correct_cutoff)
        if weight != "unit":
            present = [t for t in range(1, self.ntypes + 1) if self.counts[t - 1] > 0]
            missing = [t for t in present if t not in self.mapping]
            if missing:
                raise ValueError("weights %s need a species for type(s) %s (-m)" % (weight, missing))
        self.amp = np.ones(self.ntypes)
        self.xcoef = np.zeros((self.ntypes, 9))
        for t in range(1, self.ntypes + 1):
            if t not in self.mapping or self.counts[t - 1] == 0:
                continue
            if weight == "neutron":
                self.amp[t - 1] = elements.neutron_b(self.mapping[t])
            elif weight == "xray":
                self.xcoef[t - 1] = elements.xray_coefficients(self.mapping[t])
        self.rmax = self._default_rmax(rmax, first_frame)
        self.dr, self.skin = float(dr), float(skin)
        cfg = _lib.DebyeConfig()
        cfg.abi_version = _lib.ABI_VERSION
        cfg.device = device
        cfg.natoms = self.natoms
        cfg.ntypes = self.ntypes
        for k in range(3):
            cfg.pbc[k] = 1 if self.pbc[k] else 0
            cfg.origin[k] = float(self.origin[k])
        cfg.type_id = types.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        flat = self.cell.ravel(order="F")
        for i in range(9):
            cfg.cell[i] = float(flat[i])
        cfg.rmax, cfg.dr, cfg.skin = self.rmax, self.dr, self.skin
        plan = ctypes.c_void_p()
        _lib.check(self.lib.sqc_hip_debye_plan_create(ctypes.byref(plan), ctypes.byref(cfg)),
                   "sqc_hip_debye_plan_create")
        self.plan = plan
        self.nbins = int(self.info.nbins)

    def _default_rmax(self, rmax, first_frame):
        if rmax is not None and rmax > 0.0:
            return float(rmax)
        bound = math.inf
        a = self.cell
        vol = self.volume
        for i in range(3):
            n = np.cross(a[:, (i + 1) % 3], a[:, (i + 2) % 3])
            if self.pbc[i]:
                bound = min(bound, vol / math.sqrt(float(np.sum(n * n))))
        if bound < math.inf:
            return 0.5 * bound
        if first_frame is None:
            raise ValueError("rmax must be given (or first_frame) when no axis is periodic")
        pos = np.asarray(first_frame, dtype=np.float64).reshape(-1, 3)
        span = pos.max(axis=0) - pos.min(axis=0)
        r = math.sqrt(float(np.sum(span * span)))
        if r <= 0.0:
            r = math.sqrt(float(np.sum((a @ np.ones(3)) ** 2)))
        return r

    # -- plumbing ---------------------------------------------------------
    @property
    def info(self):
        info = _lib.DebyeInfo()
        _lib.check(self.lib.sqc_hip_debye_plan_info(self.plan, ctypes.byref(info)), "debye_plan_info")
        return info

    def close(self):
        if getattr(self, "plan", None) and self.plan.value:
            self.lib.sqc_hip_debye_destroy(self.plan)
            self.plan = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def accumulate(self, positions):
        """positions: (natoms, 3) or (nframes, natoms, 3) host array."""
        pos = np.ascontiguousarray(positions, dtype=np.float64)
        nf = 1 if pos.ndim == 2 else pos.shape[0]
        if pos.size != nf * self.natoms * 3:
            raise ValueError("positions do not match natoms")
        _lib.check(self.lib.sqc_hip_debye_accumulate(self.plan, pos.ctypes.data, nf),
                   "debye_accumulate")

    def accumulate_device(self, pos_tensor):
        """pos_tensor: float64 CUDA tensor (natoms, 3) or (nframes, natoms, 3)."""
        t = pos_tensor
        if t.dtype.itemsize != 8 or not t.is_cuda or not t.is_contiguous():
            raise ValueError("need a contiguous float64 device tensor")
        nf = 1 if t.dim() == 2 else t.shape[0]
        if t.numel() != nf * self.natoms * 3:
            raise ValueError("positions do not match natoms")
        import torch
        stream = torch.cuda.current_stream(t.device).cuda_stream
        _lib.check(self.lib.sqc_hip_debye_accumulate_device(self.plan, t.data_ptr(), nf, stream),
                   "debye_accumulate_device")

    def profile(self, enable=True):
        _lib.check(self.lib.sqc_hip_debye_profile(self.plan, 1 if enable else 0), "debye_profile")

    def stage_times(self):
        """{stage: (total ms, launches)} since the last call: positions, pairs, row_sum."""
        ms = (ctypes.c_double * 3)()
        n = (ctypes.c_int64 * 3)()
        _lib.check(self.lib.sqc_hip_debye_stage_times(self.plan, ms, n, 3), "debye_stage_times")
        return {k: (ms[i], n[i]) for i, k in enumerate(("positions", "pairs", "row_sum"))}

    def histogram(self):
        """(hist[a][b][k] ordered pair counts summed over the frames, nframes)."""
        h = np.zeros((self.ntypes, self.ntypes, self.nbins))
        nf = ctypes.c_int64(0)
        _lib.check(self.lib.sqc_hip_debye_fetch(self.plan, h.ctypes.data, ctypes.byref(nf)),
                   "debye_fetch")
        return h, int(nf.value)

    def reset(self):
        _lib.check(self.lib.sqc_hip_debye_reset(self.plan), "debye_reset")

    # -- evaluation -------------------------------------------------------
    def amplitude(self, t, q):
        """scheme%amplitude of type index t (0-based) at |q| (src/sqc_weights.f90:136-154)."""
        if self.weight == "xray":
            c = self.xcoef[t]
            st = (q / (4.0 * math.acos(-1.0))) ** 2
            return c[8] + sum(c[k] * math.exp(-c[4 + k] * st) for k in range(4))
        return float(self.amp[t])

    def _den(self, w):
        if self.norm == "n":
            return float(self.natoms)
        live = self.counts > 0
        if self.norm == "mean":
            sw = float(np.sum(self.counts[live] * w[live]))
            return sw * sw / self.natoms
        return float(np.sum(self.counts[live] * w[live] * w[live]))

    def shell_sums(self, qmin, qmax, nq):
        """(q, num, den, partial_num[a][b][nq], nframes) of prepare_output."""
        hist, nhist = self.histogram()
        q, num, den, part = shell_sums(hist, nhist, self.counts, self.volume, self.pbc, self.rmax,
                                       self.dr, self.amplitude, self.norm, self.partials,
                                       self.correct_cutoff, qmin, qmax, nq)
        return q, num, den, part, nhist

    def structure_factor(self, qmin=0.0, qmax=20.0, nq=500):
        """Rows (q, S(q), S_ab columns...) as sf_write_results writes them."""
        q, num, den, part, nframes = self.shell_sums(qmin, qmax, nq)
        frames = float(max(nframes, 1))
        s = np.where(den > 0, num / (frames * np.where(den > 0, den, 1.0)), 0.0)
        cols = [q, s]
        if self.partials:
            for a, b in self._pairs():
                v = part[a, b] / (frames * self.natoms)
                if self.faber_ziman:
                    xa, xb = self.counts[a] / self.natoms, self.counts[b] / self.natoms
                    v = (v - (xa if a == b else 0.0)) / (xa * xb) + 1.0
                cols.append(v)
        return np.stack(cols, axis=1)

    def _pairs(self):
        live = [t for t in range(self.ntypes) if self.counts[t] > 0]
        return [(a, b) for a in live for b in live if b >= a]

    def pair_label(self, a, b):
        ta, tb = a + 1, b + 1
        if ta in self.mapping and tb in self.mapping:
            return "%s-%s" % (self.mapping[ta], self.mapping[tb])
        return "%d-%d" % (ta, tb)

    def column_labels(self):
        if not self.partials:
            return []
        pre = "A(" if self.faber_ziman else "S("
        return [pre + self.pair_label(a, b) + ")" for a, b in self._pairs()]

    def format_table(self, qmin=0.0, qmax=20.0, nq=500):
        rows = self.structure_factor(qmin, qmax, nq)
        labels = self.column_labels()
        head = "# q S(q)" + "".join(" " + lab.ljust(24) for lab in labels) if labels else "# q S(q)"
        lines = [head]
        for row in rows:
            line = "%14.6f  %20.12E" % (row[0], row[1])
            for v in row[2:]:
                line += "  %20.12E" % v
            lines.append(line)
        return "\n".join(lines) + "\n"

    def g_partial(self, exact_shell=False):
        """(r, g[a][b][k]) of debye_fill_g_partial."""
        hist, nhist = self.histogram()
        r = (np.arange(1, self.nbins + 1) - 0.5) * self.dr
        if exact_shell:
            rlo = np.maximum(0.0, r - 0.5 * self.dr)
            shell = 4.0 * math.acos(-1.0) / 3.0 * ((r + 0.5 * self.dr) ** 3 - rlo ** 3)
        else:
            shell = 4.0 * math.acos(-1.0) * r ** 2 * self.dr
        g = np.zeros_like(hist)
        for a in range(self.ntypes):
            for b in range(self.ntypes):
                if self.counts[a] == 0 or self.counts[b] == 0:
                    continue
                rho_b = self.counts[b] / self.volume
                g[a, b] = (hist[a, b] / nhist) / (self.counts[a] * shell * rho_b)
        return r, g

    def rdf(self):
        """(r, g_total, g[a][b]) of debye_write_rdf: the total weighted with the
        q -> 0 amplitudes."""
        r, g = self.g_partial(False)
        w = np.array([self.amplitude(t, 0.0) for t in range(self.ntypes)])
        c = self.counts / self.natoms
        f_mean = float(np.sum(c * w))
        g_tot = np.zeros(self.nbins)
        if abs(f_mean) > 0.0:
            live = self.counts > 0
            cw = (c * w)[live]
            g_tot = np.einsum("a,b,abk->k", cw, cw, g[np.ix_(live, live)]) / f_mean ** 2
        return r, g_tot, g

    def pair_entropy(self):
        """(total S2, {pair label: S2_ab}, r, S2 curves per pair) of
        debye_write_pair_entropy, in kB per particle."""
        r, g = self.g_partial(True)
        _, nhist = self.histogram()
        T = self.ntypes
        for a in range(T):
            for b in range(a + 1, T):
                gf = 0.5 * (g[a, b] + g[b, a])
                g[a, b] = gf
                g[b, a] = gf
        rho = self.natoms / self.volume
        rlo = np.maximum(0.0, r - 0.5 * self.dr)
        shell = 4.0 * math.acos(-1.0) / 3.0 * ((r + 0.5 * self.dr) ** 3 - rlo ** 3)
        total, parts, curves = 0.0, {}, {}
        for a in range(T):
            for b in range(a, T):
                lab = self.pair_label(a, b)
                ca, cb = int(self.counts[a]), int(self.counts[b])
                if (a == b and ca <= 1) or (a != b and (ca == 0 or cb == 0)):
                    parts[lab], curves[lab] = 0.0, np.zeros(self.nbins)
                    continue
                xa, xb = ca / self.natoms, cb / self.natoms
                gf = g[a, b]
                pos = gf > 0.0
                integrand = np.ones(self.nbins)
                integrand[pos] = gf[pos] * np.log(gf[pos]) - gf[pos] + 1.0
                bias = self.volume / (2.0 * nhist * ca * cb)
                cum = np.cumsum(shell * integrand - bias)
                curve = -0.5 * rho * xa * xb * cum
                parts[lab], curves[lab] = float(curve[-1]), curve
                total += parts[lab] * (1.0 if a == b else 2.0)
        return total, parts, r, curves
