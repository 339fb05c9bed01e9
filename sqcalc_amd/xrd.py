"""Powder XRD pattern I(2theta) from the device's per-mode sums.

The reference bins every kept mode's |rho(q)|^2, times a Lorentz-polarization
factor, by its two-theta angle (src/sqc_structure_factor.f90:595-643,
:837-904), and bins the amplitude-weighted pair sums the same way for the
pair columns (:538-588).  The per-frame factors (bin, Lorentz-polarization,
amplitudes f_a(|q|)) do not change between frames, so the pattern of a
trajectory follows exactly from the per-mode frame sums the device keeps:
gnum (|sum_a w_a rho_a|^2, --grid) and gpart (Re rho_a rho_b*, mode_partials).
The binning here runs once, when the pattern is asked for.
"""

import math

import numpy as np

DEG = math.pi / 180.0


def two_theta_deg(q, lam):
    """Two-theta [deg] of momentum transfer q at wavelength lam, the sine
    clipped to [-1, 1] (src/sqc_structure_factor.f90:906-913)."""
    s = np.clip(np.asarray(q, dtype=np.float64) * lam / (4.0 * math.pi), -1.0, 1.0)
    return 2.0 * np.arcsin(s) * 180.0 / math.pi


def lorentz_polarization(tth_deg):
    """(1 + cos^2 2theta) / (sin^2 theta cos theta) (:915-923)."""
    tth = np.asarray(tth_deg, dtype=np.float64) * DEG
    th = 0.5 * tth
    return (1.0 + np.cos(tth) ** 2) / (np.sin(th) ** 2 * np.cos(th))


def q_range(lam, tth_min, tth_max):
    """The q window of a two-theta range, q = 4 pi sin(theta) / lambda
    (src/sqc_options.f90:650-651)."""
    return (4.0 * math.pi * math.sin(0.5 * DEG * tth_min) / lam,
            4.0 * math.pi * math.sin(0.5 * DEG * tth_max) / lam)


def cell_widths(cell):
    """Distances between opposite faces of the cell (src/sqc_cell.f90:102-118)."""
    a = np.asarray(cell, dtype=np.float64).reshape(3, 3)
    vol = abs(np.linalg.det(a))
    if vol <= 0.0:
        return np.zeros(3)
    out = np.zeros(3)
    for i in range(3):
        j, k = (i + 1) % 3, (i + 2) % 3
        out[i] = vol / np.linalg.norm(np.cross(a[:, j], a[:, k]))
    return out


def configure(qlen, lam, tth_min=1.0, tth_max=179.0, step=0.0, lp=True, cell=None,
              pbc=(True, True, True)):
    """Bin map of the kept modes (src/sqc_structure_factor.f90:846-904):
    returns (bin per mode, 0-based; weight per mode; modes per bin; step; bins).
    step <= 0 takes the reciprocal-lattice spacing of the box at
    min(theta_max, 60 deg)."""
    if step <= 0.0:
        w = cell_widths(cell)
        per = [w[i] for i in range(3) if pbc[i]]
        if not per or min(per) <= 0.0:
            raise ValueError("cannot derive the XRD bin width from a degenerate box; pass step")
        theta_max = min(0.5 * DEG * tth_max, 60.0 * DEG)
        step = (lam / (min(per) * math.cos(theta_max))) / DEG
    nbins = max(1, int(math.ceil((tth_max - tth_min) / step)))
    step = (tth_max - tth_min) / nbins
    tth = two_theta_deg(qlen, lam)
    b = np.clip(((tth - tth_min) / step).astype(np.int64), 0, nbins - 1)
    weight = lorentz_polarization(tth) if lp else np.ones_like(tth)
    count = np.bincount(b, minlength=nbins)
    return b, weight, count, step, nbins


def pattern(gnum, gpart, amps, qlen, nframes, natoms, lam, tth_min=1.0, tth_max=179.0, step=0.0,
            lp=True, cell=None, pbc=(True, True, True)):
    """I(2theta) per atom and frame, and the pair columns.

    gnum[K]: frame sums of the weighted |rho|^2 per mode; gpart[npairs][K] or
    None: frame sums of Re rho_a rho_b* per mode, pairs a <= b; amps[S][K]:
    the species amplitudes w_a(|q|) per mode.
    """
    b, w, count, step, nbins = configure(qlen, lam, tth_min, tth_max, step, lp, cell, pbc)
    norm = 1.0 / (max(nframes, 1) * natoms) if natoms > 0 else 0.0
    total = np.bincount(b, weights=gnum * w, minlength=nbins) * norm
    pairs = []
    if gpart is not None:
        S = amps.shape[0]
        k = 0
        for i in range(S):
            for j in range(i, S):
                pairs.append(np.bincount(b, weights=amps[i] * amps[j] * gpart[k] * w,
                                         minlength=nbins) * norm)
                k += 1
    centers = tth_min + (np.arange(1, nbins + 1) - 0.5) * step
    return {"two_theta": centers, "intensity": total, "pairs": pairs, "count": count,
            "step": step, "bins": nbins}


def _f0(v, d):
    """Fortran f0.d: no leading zero before the point."""
    s = "%.*f" % (d, v)
    if s.startswith("0."):
        s = s[1:]
    elif s.startswith("-0."):
        s = "-" + s[2:]
    return s


def format_pattern(pat, lam, tth_min, tth_max, nframes, lp, pair_labels):
    """The text of sf_write_xrd (src/sqc_structure_factor.f90:964-1024)."""
    frames = "1 frame" if nframes == 1 else "%d frames" % nframes
    lines = ["# xrd: lambda %s A, 2theta %s to %s deg, step %s deg, %d bins, %s"
             % (_f0(lam, 6), _f0(tth_min, 4), _f0(tth_max, 4), _f0(pat["step"], 5),
                pat["bins"], frames)]
    if lp:
        lines.append("# I(2theta) = sum over the reciprocal lattice points in each bin of "
                     "|rho(q)|^2 times the Lorentz-polarization factor, per atom")
    else:
        lines.append("# I(2theta) = sum over the reciprocal lattice points in each bin of "
                     "|rho(q)|^2, without the Lorentz-polarization factor, per atom")
    has_pairs = bool(pat["pairs"])
    if has_pairs:
        lines.append("# the pair columns add up to the total: I = sum_a I(a-a) + 2 sum_{a<b} I(a-b)")
    head = "# 2theta[deg] I"
    if has_pairs:
        head += "".join(" I(%s)" % lab for lab in pair_labels)
    lines.append(head)
    for k in range(pat["bins"]):
        row = "%12.4f  %20.12E" % (pat["two_theta"][k], pat["intensity"][k])
        for col in pat["pairs"]:
            row += "  %20.12E" % col[k]
        lines.append(row)
    return "\n".join(lines) + "\n"
