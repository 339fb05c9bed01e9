"""FP64 CPU oracle for the reciprocal-space S(q) path (Python side).

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of ``bench.py`` import this module, and only as the reference
results to compare with or the timed CPU restatement; the product
(``sqcalc_amd``) never does.

It wraps ``libsqc_oracle.so`` (``sqc_oracle.c``, ``sqc_nufft_cpu.c``) and
restates the host-side bookkeeping of the reference in numpy:

* ``shell_table``   S(q) and partial columns, src/sqc_structure_factor.f90:657-719
* ``format_table``  the ``# q S(q) S(a-b)...`` text, src/sqc_structure_factor.f90:771-827
* ``denominators``  mode_denominator summed per shell, :389-451
* ``phihat``        Fourier coefficients of the ES kernel by Gauss-Legendre
                    quadrature (the role of external/finufft/src/cuda/common.cu:27-66)

How each function is pinned is listed in the header of ``sqc_oracle.c`` and in
tests/test_oracle.py.  ``nufft_frame`` (the CPU restatement of the reference
FINUFFT path) is pinned by the direct sum.
"""

import ctypes
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

# IT92 X-ray coefficients (a1..a4, b1..b4, c) and bound coherent neutron
# lengths b [fm]: the reference's own tables (src/sqc_element_data.f90; IT92
# 6.1.1.4, Neutron News 3 (1992) 29-37) as the fixture
# tests/golden/elements/elements.json holds them (tests/golden/make_elements_golden.py).
_TABLES = None


def element_tables():
    """(elements, species) rows of the fixture: elements by symbol, species
    by canonical label."""
    global _TABLES
    if _TABLES is None:
        import json
        path = os.path.join(HERE, "..", "tests", "golden", "elements", "elements.json")
        with open(path) as fh:
            d = json.load(fh)
        _TABLES = ({e["symbol"]: e for e in d["elements"]},
                   {sp["label"]: sp for sp in d["species"]}, d)
    return _TABLES


WEIGHT_KIND = {"unit": 0, "neutron": 1, "xray": 2}
NORM_KIND = {"mean": 1, "self": 2, "n": 3}

_d = ctypes.POINTER(ctypes.c_double)
_i32 = ctypes.POINTER(ctypes.c_int32)
_i64 = ctypes.POINTER(ctypes.c_int64)


def lib():
    """Load (building on first use when the compiler is at hand) the oracle."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = os.path.join(HERE, "libsqc_oracle.so")
    if not os.path.exists(path):
        import subprocess
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    L = ctypes.CDLL(path)
    L.sqo_cell_setup.restype = ctypes.c_double
    L.sqo_cell_setup.argtypes = [_d, _d]
    L.sqo_cell_from_bounds.argtypes = [_d, _d, _d, _d, _d]
    L.sqo_mode_extent.argtypes = [_d, ctypes.c_double, _i64]
    L.sqo_mode_count.restype = ctypes.c_int64
    L.sqo_mode_count.argtypes = [_d, ctypes.c_double, ctypes.c_double]
    L.sqo_modes.restype = ctypes.c_int64
    L.sqo_modes.argtypes = [_d, ctypes.c_double, ctypes.c_double, ctypes.c_int, _i32, _d, _d,
                            _i64, _i32]
    L.sqo_denominators.argtypes = [ctypes.c_int64, _d, _i32, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_int, _i64, _d, _d, ctypes.c_int64,
                                   _d, _i64]
    L.sqo_direct_rho.restype = ctypes.c_int
    L.sqo_direct_rho.argtypes = [ctypes.c_int64, _d, _d, _d, _i32, ctypes.c_int, ctypes.c_int64,
                                 _i32, _d, ctypes.c_int, _d]
    L.sqo_shell_sums.argtypes = [ctypes.c_int, ctypes.c_int64, _d, _d, _i32, ctypes.c_int,
                                 ctypes.c_int, _d, _d, _d, _d]
    L.sqo_direct_grad.restype = ctypes.c_int
    L.sqo_direct_grad.argtypes = [ctypes.c_int64, _d, _d, _d, _i32, ctypes.c_int, ctypes.c_int64,
                                  _d, _d, _i32, _d, ctypes.c_int, _d, _d, _d, _d, _d, _i64,
                                  ctypes.c_int, _d]
    L.sqo_nufft_spread.restype = ctypes.c_int
    L.sqo_nufft_spread.argtypes = [ctypes.c_int64, _d, _d, _d, _i32, _d, ctypes.c_int, _i64,
                                   ctypes.c_int, ctypes.c_double, ctypes.c_int64, _d]
    L.sqo_nufft_bin.argtypes = [ctypes.c_int, _d, _i64, ctypes.c_int64, _i32, _d, _i32,
                                ctypes.c_int, _d, _d, _d, ctypes.c_int, _d, _d, _d, _d]
    L.sqo_num_threads.restype = ctypes.c_int
    L.sqo_dyn_rho.restype = ctypes.c_int
    L.sqo_dyn_rho.argtypes = [ctypes.c_int64, _d, _i32, ctypes.c_int, _d, _d, ctypes.c_int,
                              ctypes.c_int64, _i32, ctypes.c_int, _d]
    _u64 = ctypes.POINTER(ctypes.c_uint64)
    L.sqo_cell_widths.argtypes = [_d, _d]
    L.sqo_debye_cells.argtypes = [_d, ctypes.c_double, ctypes.c_double, _i32, _d]
    L.sqo_debye_positions.argtypes = [ctypes.c_int64, _d, _d, _d, ctypes.c_int, _d, _d]
    L.sqo_debye_hist_rows.restype = ctypes.c_int
    L.sqo_debye_hist_rows.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int,
                                      _i32, _d, _i32, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_double, _d, ctypes.c_int, _u64]
    L.sqo_debye_hist.restype = ctypes.c_int
    L.sqo_debye_hist.argtypes = [ctypes.c_int64, ctypes.c_int, _i32, _d, _i32, ctypes.c_double,
                                 ctypes.c_double, ctypes.c_double, _d, ctypes.c_int, _u64]
    _LIB = L
    return L


def _p(a, t):
    return a.ctypes.data_as(t)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# --------------------------------------------------------------------------
# geometry and mode set
# --------------------------------------------------------------------------

def cell_from_bounds(lo, hi, tilt=(0.0, 0.0, 0.0)):
    """(A, origin): A holds a1, a2, a3 as columns (src/sqc_cell.f90:39-63)."""
    a = np.zeros(9)
    o = np.zeros(3)
    lib().sqo_cell_from_bounds(_p(_f64(lo), _d), _p(_f64(hi), _d), _p(_f64(tilt), _d),
                               _p(a, _d), _p(o, _d))
    return a.reshape(3, 3, order="F"), o


def reciprocal(cell):
    """B with b1, b2, b3 as columns, B = 2 pi A^-T (src/sqc_cell.f90:76-86)."""
    a = _f64(np.asarray(cell).reshape(3, 3).ravel(order="F"))
    b = np.zeros(9)
    lib().sqo_cell_setup(_p(a, _d), _p(b, _d))
    return b.reshape(3, 3, order="F")


def modes(cell, qmin, qmax, nq):
    """Kept reciprocal-lattice modes of sf_configure (src/sqc_structure_factor.f90:327-399)."""
    a = _f64(np.asarray(cell).reshape(3, 3).ravel(order="F"))
    L = lib()
    m = np.zeros(3, dtype=np.int64)
    L.sqo_mode_extent(_p(a, _d), qmax, _p(m, _i64))
    k = int(L.sqo_mode_count(_p(a, _d), qmin, qmax))
    hkl = np.zeros((k, 3), dtype=np.int32)
    qvec = np.zeros((k, 3))
    qlen = np.zeros(k)
    gidx = np.zeros(k, dtype=np.int64)
    shell = np.zeros(k, dtype=np.int32)
    L.sqo_modes(_p(a, _d), qmin, qmax, nq, _p(hkl, _i32), _p(qvec, _d), _p(qlen, _d),
                _p(gidx, _i64), _p(shell, _i32))
    return {"m": m, "hkl": hkl, "qvec": qvec, "qlen": qlen, "gidx": gidx, "shell": shell,
            "qmin": float(qmin), "qmax": float(qmax), "nq": int(nq)}


def weight_params(kind, labels):
    """Per-species constants (unit/neutron) or IT92 coefficients (xray)."""
    S = len(labels)
    amp = np.ones(S)
    xcoef = np.zeros((S, 9))
    elements, species, d = element_tables()
    for i, lab in enumerate(labels):
        if kind == "neutron":
            if lab in species:
                e = d["elements"][species[lab]["element"] - 1]
            else:
                e = elements[lab]
            assert e["has_neutron"], lab
            amp[i] = e["neutron_b"]
        elif kind == "xray":
            xcoef[i] = species[lab]["coef"]
    return amp, xcoef


def amplitude(kind, amp, xcoef, sp, q):
    q = np.asarray(q, dtype=np.float64)
    if kind == "unit":
        return np.ones_like(q)
    if kind == "neutron":
        return np.full_like(q, amp[sp])
    c = xcoef[sp]
    st = (q / (4.0 * math.acos(-1.0))) ** 2
    f = np.full_like(q, c[8])
    for i in range(4):
        f = f + c[i] * np.exp(-c[4 + i] * st)
    return f


def denominators(md, norm, kind, counts, amp, xcoef, natoms):
    """den[nq], shell_count[nq] (src/sqc_structure_factor.f90:389-451)."""
    nq = md["nq"]
    den = np.zeros(nq)
    cnt = np.zeros(nq, dtype=np.int64)
    counts = np.ascontiguousarray(counts, dtype=np.int64)
    lib().sqo_denominators(len(md["qlen"]), _p(md["qlen"], _d), _p(md["shell"], _i32), nq,
                           NORM_KIND[norm], WEIGHT_KIND[kind], len(counts), _p(counts, _i64),
                           _p(_f64(amp), _d), _p(_f64(xcoef), _d), natoms, _p(den, _d),
                           _p(cnt, _i64))
    return den, cnt


# --------------------------------------------------------------------------
# direct sum
# --------------------------------------------------------------------------

def direct_rho(pos, cell, origin, species_of, nspecies, md, method=1):
    """rho_a(q_m), complex [S, K] (src/sqc_structure_factor.f90:1398-1475)."""
    a = _f64(np.asarray(cell).reshape(3, 3).ravel(order="F"))
    pos = _f64(pos)
    sp = np.ascontiguousarray(species_of, dtype=np.int32)
    K = len(md["qlen"])
    rho = np.zeros((nspecies, K, 2))
    rc = lib().sqo_direct_rho(len(sp), _p(pos, _d), _p(a, _d), _p(_f64(origin), _d),
                              _p(sp, _i32), nspecies, K, _p(md["hkl"], _i32),
                              _p(_f64(md["qvec"]), _d), method, _p(rho, _d))
    if rc:
        raise MemoryError("oracle direct sum")
    return rho[..., 0] + 1j * rho[..., 1]


def shell_sums(rho, md, kind, amp, xcoef, partials=True):
    """num[nq] and partial[S, S, nq] of one frame, in mode order."""
    S, K = rho.shape
    nq = md["nq"]
    r = np.ascontiguousarray(np.stack([rho.real, rho.imag], axis=-1))
    num = np.zeros(nq)
    part = np.zeros((S, S, nq)) if partials else None
    lib().sqo_shell_sums(S, K, _p(r, _d), _p(md["qlen"], _d), _p(md["shell"], _i32), nq,
                         WEIGHT_KIND[kind], _p(_f64(amp), _d), _p(_f64(xcoef), _d),
                         _p(num, _d), _p(part, _d) if partials else None)
    return num, part


def gradient(pos, cell, origin, species_of, md, rho, kind, amp, xcoef, gS, den, shell_count,
             gP=None):
    """Analytic FP64 dL/dr (3 per atom) for upstream gradients on the shell values."""
    a = _f64(np.asarray(cell).reshape(3, 3).ravel(order="F"))
    pos = _f64(pos)
    sp = np.ascontiguousarray(species_of, dtype=np.int32)
    S, K = rho.shape
    r = np.ascontiguousarray(np.stack([rho.real, rho.imag], axis=-1))
    grad = np.zeros((len(sp), 3))
    gP_arr = _f64(gP) if gP is not None else None
    rc = lib().sqo_direct_grad(len(sp), _p(pos, _d), _p(a, _d), _p(_f64(origin), _d),
                               _p(sp, _i32), S, K, _p(_f64(md["qvec"]), _d),
                               _p(md["qlen"], _d), _p(md["shell"], _i32), _p(r, _d),
                               WEIGHT_KIND[kind], _p(_f64(amp), _d), _p(_f64(xcoef), _d),
                               _p(_f64(gS), _d), _p(_f64(den), _d),
                               _p(gP_arr, _d) if gP is not None else None,
                               _p(np.ascontiguousarray(shell_count, dtype=np.int64), _i64),
                               md["nq"], _p(grad, _d))
    if rc:
        raise MemoryError("oracle gradient")
    return grad


# --------------------------------------------------------------------------
# whole-run helpers (shell table, text format)
# --------------------------------------------------------------------------

class Run:
    """Accumulates frames like structure_factor_t (static subset).

    types: LAMMPS type id per atom; labels: {type_id: species label} or None.
    Species are the present type ids in ascending order
    (src/sqc_structure_factor.f90:1109-1163).
    """

    def __init__(self, cell, origin, types, labels=None, weight="unit", norm="mean",
                 qmin=0.0, qmax=20.0, nq=500, partials=True, faber_ziman=False):
        self.cell = np.asarray(cell, dtype=np.float64).reshape(3, 3)
        self.origin = np.asarray(origin, dtype=np.float64)
        types = np.asarray(types, dtype=np.int64)
        self.types = types
        self.ntypes = max(int(types.max()), max(labels) if labels else 1)
        self.species_type = np.unique(types)
        self.species_of = np.searchsorted(self.species_type, types).astype(np.int32)
        self.counts = np.array([np.count_nonzero(types == t) for t in self.species_type])
        self.labels = labels or {}
        self.weight = weight
        self.norm = norm
        self.partials = partials
        self.faber_ziman = faber_ziman
        names = [self.labels.get(int(t), str(int(t))) for t in self.species_type]
        self.amp, self.xcoef = (weight_params(weight, names) if weight != "unit"
                                else (np.ones(len(names)), np.zeros((len(names), 9))))
        self.md = modes(self.cell, qmin, qmax, nq)
        self.natoms = len(types)
        self.den, self.shell_count = denominators(self.md, norm, weight, self.counts, self.amp,
                                                  self.xcoef, self.natoms)
        S = len(self.species_type)
        self.num = np.zeros(nq)
        self.partial = np.zeros((S, S, nq))
        self.nframes = 0

    def add_frame(self, pos, method=1):
        rho = direct_rho(pos, self.cell, self.origin, self.species_of, len(self.species_type),
                         self.md, method=method)
        num, part = shell_sums(rho, self.md, self.weight, self.amp, self.xcoef, True)
        self.num += num
        self.partial += part
        self.nframes += 1
        return rho

    def table(self):
        return shell_table(self.md, self.num, self.den, self.partial if self.partials else None,
                           self.shell_count, self.nframes, self.natoms, self.counts,
                           self.faber_ziman)

    def column_labels(self):
        out = []
        for i in range(len(self.species_type)):
            for j in range(i, len(self.species_type)):
                ti, tj = int(self.species_type[i]), int(self.species_type[j])
                if ti in self.labels and tj in self.labels:
                    pair = "%s-%s" % (self.labels[ti], self.labels[tj])
                else:
                    pair = "%d-%d" % (ti, tj)
                out.append(("A(" if self.faber_ziman else "S(") + pair + ")")
        return out


def shell_table(md, num, den, partial, shell_count, nframes, natoms, counts, faber_ziman=False):
    """Rows (q, S, partial columns...), src/sqc_structure_factor.f90:657-719."""
    nq = md["nq"]
    dq = (md["qmax"] - md["qmin"]) / nq
    frames = float(max(nframes, 1))
    q = md["qmin"] + (np.arange(1, nq + 1) - 0.5) * dq
    s = np.where(den > 0.0, num / (frames * np.where(den > 0, den, 1.0)), 0.0)
    cols = [q, s]
    if partial is not None:
        S = partial.shape[0]
        for a in range(S):
            for b in range(a, S):
                v = partial[a, b] / (frames * natoms * np.maximum(shell_count, 1))
                if faber_ziman:
                    xa, xb = counts[a] / natoms, counts[b] / natoms
                    v = (v - (xa if a == b else 0.0)) / (xa * xb) + 1.0
                cols.append(v)
    return np.stack(cols, axis=1)


def format_table(rows, labels):
    """Text of the shell table as sf_write_results writes it."""
    out = []
    if labels:
        out.append("# q S(q)" + "".join(" " + lab.ljust(24) for lab in labels))
    else:
        out.append("# q S(q)")
    for r in rows:
        line = "%14.6f  %20.12E" % (r[0], r[1])
        for v in r[2:]:
            line += "  %20.12E" % v
        out.append(line)
    return "\n".join(out) + "\n"


# --------------------------------------------------------------------------
# NUFFT pieces shared by the CPU restatement
# --------------------------------------------------------------------------

def next235even(n):
    """Smallest even 2,3,5-smooth integer >= n (external/finufft/src/finufft_core.cpp:87-90)."""
    n = max(int(n), 2)
    if n % 2:
        n += 1
    while True:
        k = n
        for p in (2, 3, 5):
            while k % p == 0:
                k //= p
        if k == 1:
            return n
        n += 2


def es_params(eps, upsampfac):
    """(ns, beta) of the ES kernel, rule of external/finufft/src/cuda/spreadinterp.cpp:61-82."""
    if upsampfac == 2.0:
        ns = math.ceil(-math.log10(eps / 10.0))
    else:
        ns = math.ceil(-math.log(eps) / (math.pi * math.sqrt(1.0 - 1.0 / upsampfac)))
    ns = max(2, min(ns, 16))
    bon = 2.30
    if ns == 2:
        bon = 2.20
    if ns == 3:
        bon = 2.26
    if ns == 4:
        bon = 2.38
    if upsampfac != 2.0:
        bon = 0.97 * math.pi * (1.0 - 1.0 / (2.0 * upsampfac))
    return ns, bon * ns


def phihat(nf, hmax, ns, beta, nquad=None):
    """phihat(h) = int phi(2u/ns) cos(2 pi h u / nf) du, h = 0..hmax (Gauss-Legendre)."""
    nquad = nquad or max(4 * ns, 64)
    x, w = np.polynomial.legendre.leggauss(nquad)
    u = 0.5 * ns * x
    wu = 0.5 * ns * w
    ker = np.exp(beta * (np.sqrt(1.0 - x * x) - 1.0))
    h = np.arange(hmax + 1)[:, None]
    return (wu * ker * np.cos(2.0 * np.pi * h * u[None, :] / nf)).sum(axis=1)


def nufft_frame(run, pos, eps=1e-9, upsampfac=1.25, nthreads=None, fields=None):
    """One frame through the CPU restatement of the reference FINUFFT path.

    Returns (num, partial) for the frame.  The FFT is scipy.fft (pocketfft),
    sign +1, unnormalised, with ``workers=nthreads``.
    """
    import scipy.fft
    L = lib()
    md = run.md
    ns, beta = es_params(eps, upsampfac)
    n = np.array([next235even(max(math.ceil(upsampfac * m), 2 * ns)) for m in md["m"]],
                 dtype=np.int64)
    S = len(run.species_type)
    per_species = run.partials or run.weight == "xray"
    nf = S if per_species else 1
    field_of = run.species_of if per_species else np.zeros(run.natoms, dtype=np.int32)
    weight = np.ones(run.natoms) if per_species else run.amp[run.species_of]
    a = _f64(run.cell.ravel(order="F"))
    grid = np.zeros((nf, n[2], n[1], n[0], 2))
    rc = L.sqo_nufft_spread(run.natoms, _p(_f64(pos), _d), _p(a, _d), _p(run.origin, _d),
                            _p(np.ascontiguousarray(field_of, dtype=np.int32), _i32),
                            _p(_f64(weight), _d), nf, _p(n, _i64), ns, beta, 1000,
                            _p(grid, _d))
    if rc:
        raise MemoryError("cpu nufft spread")
    g = grid[..., 0] + 1j * grid[..., 1]
    del grid
    ghat = scipy.fft.ifftn(g, axes=(1, 2, 3), norm="forward",
                           workers=nthreads or os.cpu_count())
    del g
    gh = np.ascontiguousarray(np.stack([ghat.real, ghat.imag], axis=-1))
    del ghat
    hm = [int(np.abs(md["hkl"][:, d]).max()) for d in range(3)]
    ph = [phihat(int(n[d]), hm[d], ns, beta) for d in range(3)]
    num = np.zeros(md["nq"])
    part = np.zeros((nf, nf, md["nq"]))
    famp = run.amp if per_species else np.ones(1)
    kind = 2 if run.weight == "xray" else 0
    L.sqo_nufft_bin(nf, _p(gh, _d), _p(n, _i64), len(md["qlen"]), _p(md["hkl"], _i32),
                    _p(md["qlen"], _d), _p(md["shell"], _i32), md["nq"], _p(ph[0], _d),
                    _p(ph[1], _d), _p(ph[2], _d), kind, _p(_f64(famp), _d),
                    _p(_f64(run.xcoef), _d), _p(num, _d),
                    _p(part, _d) if run.partials else None)
    return num, (part if run.partials else None), {"n": n, "ns": ns, "beta": beta}


# --------------------------------------------------------------------------
# dyn: density amplitudes on a list of Miller indices
# --------------------------------------------------------------------------

def dyn_rho(pos, species_of, nspecies, recip, hkl, base=None, method=1):
    """rho[a][m] complex of exp(i (n_m . b) . r_j [+ i base . r_j]) summed over
    the atoms of species a: method 1 the reference's separable tables, 0 its
    direct cos/sin branch (sqo_dyn_rho, src/sqc_dynamics.f90:826-873)."""
    pos = _f64(np.asarray(pos).reshape(-1, 3))
    sp = np.ascontiguousarray(species_of, dtype=np.int32)
    hkl = np.ascontiguousarray(hkl, dtype=np.int32).reshape(-1, 3)
    recip = _f64(np.asarray(recip).reshape(3, 3).ravel(order="F"))
    bv = _f64(np.zeros(3) if base is None else base)
    out = np.zeros((nspecies, hkl.shape[0], 2))
    rc = lib().sqo_dyn_rho(len(sp), _p(pos, _d), _p(sp, _i32), nspecies, _p(recip, _d),
                           _p(bv, _d), 0 if base is None else 1, hkl.shape[0], _p(hkl, _i32),
                           method, _p(out, _d))
    if rc:
        raise MemoryError("oracle dyn_rho")
    return out[..., 0] + 1j * out[..., 1]


# --------------------------------------------------------------------------
# Debye pair histograms and the S(q) / g(r) evaluated from them
# --------------------------------------------------------------------------

def cell_widths(cell):
    a = _f64(np.asarray(cell).reshape(3, 3).ravel(order="F"))
    w = np.zeros(3)
    lib().sqo_cell_widths(_p(a, _d), _p(w, _d))
    return w


def debye_cells(cell, rmax, skin):
    """(div[3], bins (3, 3)) of the reference's cell list (nl_configure)."""
    a = _f64(np.asarray(cell).reshape(3, 3).ravel(order="F"))
    div = np.zeros(3, dtype=np.int32)
    bins = np.zeros(9)
    lib().sqo_debye_cells(_p(a, _d), rmax, skin, _p(div, _i32), _p(bins, _d))
    return div, bins.reshape(3, 3, order="F")


class DebyeRun:
    """Ordered pair counts of a trajectory as the reference's Debye method
    forms them (sqo_debye_positions + sqo_debye_hist per frame):
    hist[a][b][k], 0-based types, uint64, summed over the frames."""

    def __init__(self, cell, origin, types, ntypes, pbc=(1, 1, 1), rmax=None, dr=0.01, skin=1.0):
        self.cell = np.asarray(cell, dtype=np.float64).reshape(3, 3)
        self.a = _f64(self.cell.ravel(order="F"))
        self.origin = _f64(origin)
        self.types = np.ascontiguousarray(types, dtype=np.int32)
        self.ntypes = int(ntypes)
        self.pbc = np.ascontiguousarray([1 if p else 0 for p in pbc], dtype=np.int32)
        self.rmax, self.dr, self.skin = float(rmax), float(dr), float(skin)
        self.nbins = max(1, int(math.ceil(self.rmax / self.dr)))
        n = len(self.types)
        self.pos = np.zeros((n, 3))
        self.posw = np.zeros((n, 3))
        self.nframes = 0
        self.hist = np.zeros((self.ntypes, self.ntypes, self.nbins), dtype=np.uint64)

    def rows(self, r, i0, i1, method=1):
        """Counts of the pairs whose first atom is in [i0, i1) of a first
        frame r (the list positions of that frame), not added to the run."""
        r = _f64(np.asarray(r).reshape(-1, 3))
        pos, posw = np.zeros_like(r), np.zeros_like(r)
        L = lib()
        L.sqo_debye_positions(len(self.types), _p(self.a, _d), _p(self.origin, _d), _p(r, _d), 1,
                              _p(pos, _d), _p(posw, _d))
        h = np.zeros_like(self.hist)
        rc = L.sqo_debye_hist_rows(len(self.types), i0, i1, self.ntypes, _p(self.types, _i32),
                                   _p(self.a, _d), _p(self.pbc, _i32), self.rmax, self.dr,
                                   self.skin, _p(pos, _d), method,
                                   h.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
        if rc:
            raise MemoryError("oracle debye_hist")
        return h

    def add_frame(self, r, method=1):
        r = _f64(np.asarray(r).reshape(-1, 3))
        L = lib()
        L.sqo_debye_positions(len(self.types), _p(self.a, _d), _p(self.origin, _d), _p(r, _d),
                              1 if self.nframes == 0 else 0, _p(self.pos, _d), _p(self.posw, _d))
        rc = L.sqo_debye_hist(len(self.types), self.ntypes, _p(self.types, _i32), _p(self.a, _d),
                              _p(self.pbc, _i32), self.rmax, self.dr, self.skin, _p(self.pos, _d),
                              method, self.hist.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
        if rc:
            raise MemoryError("oracle debye_hist")
        self.nframes += 1


def debye_sq(hist, nhist, counts, natoms, volume, pbc, rmax, dr, qmin, qmax, nq, kind, amp,
             xcoef, norm="mean", partials=True, correct_cutoff=True):
    """(num[nq], den[nq], partial_num[T][T][nq]) of debye_prepare_output
    (src/sqc_debye.f90:236-345) for nhist frames of ordered pair counts
    hist[a][b][k] of types with amplitudes amp / xcoef: the pair sum over
    bin centres r_k = (k - 1/2) dr, the r = 0 self term, debyer's cut-off
    correction; the table is num / (nframes den) as sf_shell_value."""
    hist = np.asarray(hist, dtype=np.float64)
    T, _, nbins = hist.shape
    counts = np.asarray(counts, dtype=np.int64)
    dq = (qmax - qmin) / nq
    four_pi = 4.0 * math.acos(-1.0)
    r_bin = (np.arange(1, nbins + 1) - 0.5) * dr
    num, den = np.zeros(nq), np.zeros(nq)
    part = np.zeros((T, T, nq)) if partials else None
    for s in range(nq):
        q = qmin + (s + 1 - 0.5) * dq
        w = np.array([float(amplitude(kind, amp, xcoef, t, q)) for t in range(T)])
        sinc = np.sin(q * r_bin) / (q * r_bin)
        pair_sum = 0.0
        for a in range(T):
            if counts[a] == 0:
                continue
            for b in range(T):
                if counts[b] == 0:
                    continue
                h = hist[a, b]
                keep = h > 0.0
                pair_sum += float(np.sum(w[a] * w[b] * (h[keep] / nhist) * sinc[keep]))
        live = counts > 0
        self_sum = float(np.sum(counts[live] * w[live] * w[live]))
        weight_sum = float(np.sum(counts * w))
        if norm == "n":
            den[s] = float(natoms)
        elif norm == "mean":
            sw = float(np.sum(counts[live] * w[live]))
            den[s] = sw * sw / natoms
        else:
            den[s] = self_sum
        cut = correct_cutoff and any(pbc)
        if partials:
            pcorr = 0.0
            if cut:
                qc = q * rmax
                pcorr = four_pi * (natoms / volume) / (q * q) * (rmax * math.cos(qc) - math.sin(qc) / q)
            for a in range(T):
                if counts[a] == 0:
                    continue
                for b in range(a, T):
                    if counts[b] == 0:
                        continue
                    h = hist[a, b]
                    keep = h > 0.0
                    ps = float(np.sum((h[keep] / nhist) * sinc[keep]))
                    if a == b:
                        ps += float(counts[a])
                    xa, xb = counts[a] / natoms, counts[b] / natoms
                    ps += xa * xb * pcorr * natoms
                    part[a, b, s] = ps * nhist
        corr = 0.0
        if cut:
            rho0 = natoms / volume
            avg = weight_sum / natoms
            qc = q * rmax
            corr = avg ** 2 * four_pi * rho0 / (q * q) * (rmax * math.cos(qc) - math.sin(qc) / q)
        num[s] = (pair_sum + self_sum + corr * natoms) * nhist if den[s] > 0.0 else 0.0
    return num, den, part


def debye_g(hist, nhist, counts, volume, dr, exact_shell=False):
    """(r[nbins], g[T][T][nbins]) of debye_fill_g_partial (src/sqc_debye.f90:422-452)."""
    hist = np.asarray(hist, dtype=np.float64)
    T, _, nbins = hist.shape
    r = (np.arange(1, nbins + 1) - 0.5) * dr
    if exact_shell:
        rlo = np.maximum(0.0, r - 0.5 * dr)
        shell = 4.0 * math.acos(-1.0) / 3.0 * ((r + 0.5 * dr) ** 3 - rlo ** 3)
    else:
        shell = 4.0 * math.acos(-1.0) * r ** 2 * dr
    g = np.zeros((T, T, nbins))
    for a in range(T):
        for b in range(T):
            if counts[a] == 0 or counts[b] == 0:
                continue
            rho_b = counts[b] / volume
            g[a, b] = (hist[a, b] / nhist) / (counts[a] * shell * rho_b)
    return r, g
