"""Host-side mirror of the reference's static structure-factor plug-in.

``HipStructureFactor`` plays the role of ``structure_factor_t`` with the GPU
plug-in attached (src/sqc_structure_factor.f90:71-192, src/sqc_gpu.f90:42-71):
it classifies the atoms of the first frame, builds the plan (mode set,
shells, denominators) in libsqc_hip, accumulates frames and produces the
``# q S(q) S(a-b)...`` table of ``sf_write_results``
(src/sqc_structure_factor.f90:771-827).  All per-frame work runs in the HIP
library; this module only moves arguments across the C ABI.
"""

import ctypes

import numpy as np

from . import _lib, elements


def cell_from_bounds(lo, hi, tilt=(0.0, 0.0, 0.0)):
    """Restricted triclinic cell from LAMMPS bounds (src/sqc_cell.f90:39-63).

    Returns (A, origin) with the lattice vectors in the columns of A.
    """
    xy, xz, yz = (float(t) for t in tilt)
    xlo = lo[0] - min(0.0, xy, xz, xy + xz)
    xhi = hi[0] - max(0.0, xy, xz, xy + xz)
    ylo = lo[1] - min(0.0, yz)
    yhi = hi[1] - max(0.0, yz)
    a = np.zeros((3, 3))
    a[:, 0] = [xhi - xlo, 0.0, 0.0]
    a[:, 1] = [xy, yhi - ylo, 0.0]
    a[:, 2] = [xz, yz, hi[2] - lo[2]]
    return a, np.array([xlo, ylo, lo[2]], dtype=np.float64)


class HipStructureFactor:
    """S(q) of a trajectory on one HIP device.

    cell: 3x3 with a1, a2, a3 as columns; types: LAMMPS type id per atom;
    mapping: {type_id: species label} (needed for neutron / xray weights).
    precision: "double" (plain FP64 path) or "single" (FP32 grid, FP64 sums).
    """

    def __init__(self, cell, origin, types, mapping=None, weight="unit", norm="mean",
                 qmin=0.0, qmax=20.0, nq=500, eps=None, precision="double", partials=True,
                 faber_ziman=False, want_grid=False, device=0, upsampfac=0.0, max_batch=0,
                 tune=False, mode_partials=False):
        self.lib = _lib.load()
        self.cell = np.ascontiguousarray(cell, dtype=np.float64).reshape(3, 3)
        self.origin = np.ascontiguousarray(origin, dtype=np.float64).reshape(3)
        types = np.asarray(types, dtype=np.int64)
        self.types = types
        self.natoms = int(types.size)
        self.mapping = dict(mapping or {})
        self.ntypes = max(int(types.max()), max(self.mapping) if self.mapping else 1)
        # species = present type ids, ascending (src/sqc_structure_factor.f90:1127-1144)
        self.species_type = np.unique(types)
        self.species_of = np.ascontiguousarray(
            np.searchsorted(self.species_type, types), dtype=np.int32)
        self.counts = np.array([np.count_nonzero(types == t) for t in self.species_type])
        self.weight = weight
        self.norm = norm
        self.partials = bool(partials)
        self.faber_ziman = bool(faber_ziman)
        self.want_grid = bool(want_grid)
        self.qmin, self.qmax, self.nq = float(qmin), float(qmax), int(nq)
        S = len(self.species_type)
        # every label of the mapping resolves to a species, whatever the
        # weights (src/sqc_weights.f90:49-93); labels print canonically
        self.mapping = {int(t): elements.canonical(lab) for t, lab in self.mapping.items()}
        if weight != "unit":
            missing = [int(t) for t in self.species_type if int(t) not in self.mapping]
            if missing:
                raise ValueError("weights %s need a species for type(s) %s (-m)" % (weight, missing))
        labels = [self.mapping.get(int(t), str(int(t))) for t in self.species_type]
        self.labels = labels
        amp = np.ones(S)
        xcoef = np.zeros((S, 9))
        if weight == "neutron":
            amp = np.array([elements.neutron_b(l) for l in labels], dtype=np.float64)
        elif weight == "xray":
            xcoef = np.array([elements.xray_coefficients(l) for l in labels], dtype=np.float64)
        self.amp = np.ascontiguousarray(amp)
        self.xcoef = np.ascontiguousarray(xcoef)
        prec = {"double": _lib.F64, "single": _lib.F32}[precision]
        self.precision = precision
        if eps is None:
            # single precision defaults to 1e-5, as the reference does for
            # --precision single without --eps (app/main.f90:224-230)
            eps = 1e-5 if prec == _lib.F32 else 1e-9
        cfg = _lib.Config()
        cfg.abi_version = _lib.ABI_VERSION
        cfg.device = device
        cfg.precision = prec
        cfg.weight_kind = _lib.WEIGHT[weight]
        cfg.norm = _lib.NORM[norm]
        cfg.partials = 1 if self.partials else 0
        cfg.want_grid = 1 if want_grid else 0
        cfg.mode_partials = 1 if (mode_partials and self.partials) else 0
        self.mode_partials = bool(cfg.mode_partials)
        cfg.nq = self.nq
        cfg.qmin, cfg.qmax, cfg.eps = self.qmin, self.qmax, float(eps)
        cfg.upsampfac = float(upsampfac)
        for i in range(9):
            cfg.cell[i] = float(self.cell.ravel(order="F")[i])
        for i in range(3):
            cfg.origin[i] = float(self.origin[i])
        cfg.natoms = self.natoms
        cfg.nspecies = S
        cfg.species_of = self.species_of.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        cfg.amp_const = self.amp.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        cfg.xray_coef = self.xcoef.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        cfg.max_batch = int(max_batch)
        cfg.tune = 1 if tune else 0
        plan = ctypes.c_void_p()
        rc = self.lib.sqc_hip_plan_create(ctypes.byref(plan), ctypes.byref(cfg))
        _lib.check(rc, "sqc_hip_plan_create")
        self.plan = plan
        self.den = np.zeros(self.nq)
        self.shell_count = np.zeros(self.nq, dtype=np.int64)
        _lib.check(self.lib.sqc_hip_shells(self.plan, self.den.ctypes.data,
                                           self.shell_count.ctypes.data), "shells")

    # -- plumbing ---------------------------------------------------------
    @property
    def info(self):
        """Plan geometry (queried each time: the autotuner may change the tiles
        on the first frame)."""
        info = _lib.Info()
        _lib.check(self.lib.sqc_hip_plan_info(self.plan, ctypes.byref(info)), "plan_info")
        return info

    def close(self):
        if getattr(self, "plan", None):
            self.lib.sqc_hip_destroy(self.plan)
            self.plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def npairs(self):
        S = len(self.species_type)
        return S * (S + 1) // 2 if self.partials else 0

    # -- accumulation (accumulate_frame) -------------------------------------
    @classmethod
    def from_frame(cls, frame, mapping=None, **kw):
        """Plan for the cell, origin and atom types of a dump frame
        (sqcalc_amd.dump.Frame), as method_setup takes the first frame
        (src/sqc_structure_factor.f90:172-182)."""
        return cls(frame.cell, frame.origin, frame.type_id, mapping, **kw)

    def accumulate_dump(self, path, batch=8, nthreads=0, max_frames=None):
        """Every frame of a LAMMPS dump (sqcalc_amd.dump.feed: the next batch
        is parsed while the device runs this one); returns the frames added."""
        from .dump import feed
        return feed(self, path, cell=self.cell, batch=batch, nthreads=nthreads,
                    max_frames=max_frames)

    def accumulate(self, positions):
        """Add frames; positions (N, 3) or (F, N, 3) Cartesian, host memory."""
        pos = np.ascontiguousarray(positions, dtype=np.float64)
        nf = 1 if pos.ndim == 2 else pos.shape[0]
        if pos.reshape(-1, 3).shape[0] != nf * self.natoms:
            raise ValueError("positions do not match the atom count of the plan")
        _lib.check(self.lib.sqc_hip_accumulate(self.plan, pos.ctypes.data, nf), "accumulate")

    def accumulate_device(self, pos_tensor):
        """Add frames already resident on the device (torch float64, (F, N, 3))."""
        t = pos_tensor
        if t.dtype.itemsize != 8 or not t.is_contiguous() or t.numel() % (3 * self.natoms):
            raise ValueError("need a contiguous float64 tensor of (F, N, 3) positions")
        nf = t.numel() // (3 * self.natoms)
        _lib.check(self.lib.sqc_hip_accumulate_device(self.plan, t.data_ptr(), nf),
                   "accumulate_device")

    def fetch(self):
        num = np.zeros(self.nq)
        S = len(self.species_type)
        part = np.zeros((S, S, self.nq))
        gnum = np.zeros(self.info.nmodes) if self.want_grid else None
        nfr = ctypes.c_int64(0)
        _lib.check(self.lib.sqc_hip_fetch(self.plan, num.ctypes.data, part.ctypes.data,
                                          gnum.ctypes.data if gnum is not None else None,
                                          ctypes.byref(nfr)), "fetch")
        return num, part, gnum, nfr.value

    def reset(self):
        _lib.check(self.lib.sqc_hip_reset(self.plan), "reset")

    def set_tiles(self, spread_x, spread_y, y_chunk, x_lines):
        """Pin the launch shapes the autotuner would choose (see ``info.tile``
        and ``info.x_lines``); the plan then skips tuning.  Raises for shapes
        the plan cannot run (the cuFINUFFT gpu_binsize options of
        src/sqc_cufinufft.f90:45-53)."""
        _lib.check(self.lib.sqc_hip_set_tiles(self.plan, int(spread_x), int(spread_y),
                                              int(y_chunk), int(x_lines)), "set_tiles")

    def profile(self, enable=True):
        """Record HIP events around every stage launch (see stage_times)."""
        _lib.check(self.lib.sqc_hip_profile(self.plan, 1 if enable else 0), "profile")

    def stage_times(self):
        """{stage: (total ms, launches)} since the last call (synchronises)."""
        n = len(_lib.STAGES)
        ms = (ctypes.c_double * n)()
        cnt = (ctypes.c_int64 * n)()
        _lib.check(self.lib.sqc_hip_stage_times(self.plan, ms, cnt, n), "stage_times")
        return {name: (ms[i], cnt[i]) for i, name in enumerate(_lib.STAGES) if cnt[i]}

    def stage_bytes(self):
        """Algorithmic HBM bytes per launch of the transform stages (one frame):
        each stage reads its input once and writes its output once.

          spread_z      W1 = F (M3+1) n2 n1 b written; per atom the stencil
                        corner (16 B), its sort index (4 B) and 3 ns kernel
                        values read
          dft_y         W1 read, W2 = F L n1 b written (L canonical-half lines)
          dft_x_bin     W2 read, the line table (20 B per line) and the shell
                        index of every candidate mode (2 B) read
          adj_x         rho (F C b, C candidate modes) and the shell table read,
                        W2 written
          adj_y         W2 read, W1 written
          adj_z_gather  W1 and per atom 16 + 4 B plus 6 ns kernel values read,
                        3 doubles per (atom, tile) slot written
        with b = 8 (FP32 complex) or 16 (FP64 complex).
        """
        i = self.info
        vs = 8 if self.precision == "single" else 16
        st = vs // 2
        n1, n2 = i.fine[0], i.fine[1]
        M3 = (i.modes[2] - 1) // 2
        F = i.nfields
        ns = i.nspread
        w1 = F * (M3 + 1) * n2 * n1 * vs
        w2 = F * i.nlines * n1 * vs
        C = i.rho_len
        atoms = self.natoms * (16 + 4 + 3 * ns * st)
        slots = self.natoms * self._slots_per_atom() * 3 * 8
        return {"spread_z": w1 + atoms, "dft_y": w1 + w2,
                "dft_x_bin": w2 + 20 * i.nlines + 2 * C,
                "adj_x": F * C * vs + 2 * C + 20 * i.nlines + w2, "adj_y": w1 + w2,
                "adj_z_gather": w1 + atoms + self.natoms * 3 * ns * st + slots}

    def _slots_per_atom(self):
        t = self.info.tile
        ns = self.info.nspread
        fine = self.info.fine
        sx = min(-(-fine[0] // t[0]), (ns - 1 + t[0] - 1) // t[0] + 2)
        sy = min(-(-fine[1] // t[1]), (ns - 1 + t[1] - 1) // t[1] + 2)
        return sx * sy

    # -- output (sf_shell_value, sf_partial_value, sf_write_results) ---------
    def table(self):
        num, part, _, nframes = self.fetch()
        frames = float(max(nframes, 1))
        dq = (self.qmax - self.qmin) / self.nq
        q = self.qmin + (np.arange(1, self.nq + 1) - 0.5) * dq
        safe = np.where(self.den > 0, self.den, 1.0)
        s = np.where(self.den > 0, num / (frames * safe), 0.0)
        cols = [q, s]
        if self.partials:
            S = len(self.species_type)
            for a in range(S):
                for b in range(a, S):
                    v = part[a, b] / (frames * self.natoms * np.maximum(self.shell_count, 1))
                    if self.faber_ziman:
                        xa, xb = self.counts[a] / self.natoms, self.counts[b] / self.natoms
                        v = (v - (xa if a == b else 0.0)) / (xa * xb) + 1.0
                    cols.append(v)
        return np.stack(cols, axis=1)

    def column_labels(self):
        out = []
        if not self.partials:
            return out
        S = len(self.species_type)
        for a in range(S):
            for b in range(a, S):
                ta, tb = int(self.species_type[a]), int(self.species_type[b])
                if ta in self.mapping and tb in self.mapping:
                    pair = "%s-%s" % (self.mapping[ta], self.mapping[tb])
                else:
                    pair = "%d-%d" % (ta, tb)
                out.append(("A(" if self.faber_ziman else "S(") + pair + ")")
        return out

    def format_table(self):
        rows = self.table()
        labels = self.column_labels()
        lines = ["# q S(q)" + "".join(" " + lab.ljust(24) for lab in labels)]
        for r in rows:
            line = "%14.6f  %20.12E" % (r[0], r[1])
            for v in r[2:]:
                line += "  %20.12E" % v
            lines.append(line)
        return "\n".join(lines) + "\n"

    def grid_table(self):
        """(qvec, S) per kept mode, src/sqc_structure_factor.f90:815-823."""
        _, _, gnum, nframes = self.fetch()
        K = int(self.info.nmodes)
        qvec = np.zeros((K, 3))
        hkl = np.zeros((K, 3), dtype=np.int32)
        _lib.check(self.lib.sqc_hip_modes(self.plan, hkl.ctypes.data, qvec.ctypes.data, None),
                   "modes")
        qlen = np.linalg.norm(qvec, axis=1)
        gden = self._mode_den(qlen)
        frames = float(max(nframes, 1))
        s = np.where(gden > 0, gnum / (frames * np.where(gden > 0, gden, 1.0)), 0.0)
        return qvec, s

    def mode_qlen(self):
        """|q| of every kept mode, in the reference's mode order."""
        K = int(self.info.nmodes)
        qvec = np.zeros((K, 3))
        hkl = np.zeros((K, 3), dtype=np.int32)
        _lib.check(self.lib.sqc_hip_modes(self.plan, hkl.ctypes.data, qvec.ctypes.data, None),
                   "modes")
        return np.linalg.norm(qvec, axis=1)

    def fetch_mode_partials(self):
        """Frame sums of Re(rho_a rho_b*) per kept mode, [npairs][K]."""
        S = len(self.species_type)
        out = np.zeros((S * (S + 1) // 2, int(self.info.nmodes)))
        _lib.check(self.lib.sqc_hip_fetch_mode_partials(self.plan, out.ctypes.data),
                   "fetch_mode_partials")
        return out

    def _amplitudes(self, qlen):
        """w_a(|q|) per species and mode (src/sqc_weights.f90:136-154)."""
        out = np.zeros((len(self.species_type), qlen.size))
        for i in range(len(self.species_type)):
            if self.weight == "xray":
                c = self.xcoef[i]
                st = (qlen / (4.0 * np.pi)) ** 2
                out[i] = c[8] + sum(c[k] * np.exp(-c[4 + k] * st) for k in range(4))
            else:
                out[i] = self.amp[i]
        return out

    def xrd(self, lam, tth_min=1.0, tth_max=179.0, step=0.0, lp=True, pbc=(True, True, True)):
        """Powder pattern of the accumulated frames (needs want_grid; the pair
        columns need mode_partials).  The plan's q window should be
        xrd.q_range(lam, tth_min, tth_max), as the reference sets it."""
        from . import xrd as X
        _, _, gnum, nframes = self.fetch()
        if gnum is None:
            raise ValueError("the XRD pattern needs a plan made with want_grid")
        qlen = self.mode_qlen()
        single = len(self.species_type) == 1
        gpart = self.fetch_mode_partials() if (self.mode_partials and not single) else None
        pat = X.pattern(gnum, gpart, self._amplitudes(qlen), qlen, nframes, self.natoms, lam,
                        tth_min, tth_max, step, lp, self.cell, pbc)
        if self.mode_partials and single:
            # one species: its pair column is the total (w_a^2 |rho_a|^2)
            pat["pairs"] = [pat["intensity"].copy()]
        pat["nframes"] = nframes
        return pat

    def format_xrd(self, lam, tth_min=1.0, tth_max=179.0, step=0.0, lp=True,
                   pbc=(True, True, True)):
        """The --xrd text table (src/sqc_structure_factor.f90:964-1024)."""
        from . import xrd as X
        pat = self.xrd(lam, tth_min, tth_max, step, lp, pbc)
        labels = [lab[2:-1] for lab in self.column_labels()] if pat["pairs"] else []
        return X.format_pattern(pat, lam, tth_min, tth_max, pat["nframes"], lp, labels)

    def _mode_den(self, qlen):
        if self.norm == "n":
            return np.full_like(qlen, float(self.natoms))
        sw = np.zeros_like(qlen)
        sw2 = np.zeros_like(qlen)
        for i, n in enumerate(self.counts):
            if self.weight == "xray":
                c = self.xcoef[i]
                st = (qlen / (4.0 * np.pi)) ** 2
                w = c[8] + sum(c[k] * np.exp(-c[4 + k] * st) for k in range(4))
            else:
                w = np.full_like(qlen, self.amp[i])
            sw += n * w
            sw2 += n * w * w
        return sw * sw / self.natoms if self.norm == "mean" else sw2
