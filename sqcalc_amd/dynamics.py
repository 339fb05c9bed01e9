"""Density amplitudes of the reference's dynamic method on the HIP device.

``HipDynAmplitudes`` computes, per frame and species,
    rho_a(n) = sum_{j in a} exp(i (n1 b1 + n2 b2 + n3 b3) . r_j [+ i base . r_j])
for a list of Miller indices n: the rho block of dyn_accumulate
(src/sqc_dynamics.f90:826-873, phases of dyn_atom_phase :1094-1115 and
src/sqc_phase.f90:67-113) as FP64 matrix products on the matrix cores
(sqc_hip_dyn_* in include/sqc_hip.h).  ``grid_modes`` gives the
reciprocal-lattice mode set of `--qpoints grid:QMAX`; ``line_modes`` the q
line of `--qpoints line`.  The time correlations built from the amplitudes
(F(q, t), S(q, w)) stay with the caller.
"""

import ctypes

import numpy as np

from . import _lib


def grid_modes(cell, qmax):
    """Miller indices of every reciprocal-lattice vector with 0 < |q| <= qmax,
    one of each +- pair (the first non-zero index positive), ordered by |q|
    and then by Miller index: the mode set of modes_build without a budget
    (modes_fill, modes_canonical, modes_order of src/sqc_modes.f90:484-596).
    Vectors whose |q| differ by less than the reference's 1e-6 shell
    tolerance keep the |q| order here, where the reference orders them by
    Miller index; the set is the same."""
    a = np.asarray(cell, dtype=np.float64).reshape(3, 3)
    b = 2.0 * np.pi * np.linalg.inv(a).T
    bound = [int(np.ceil(qmax * np.linalg.norm(a[:, i]) / (2.0 * np.pi))) + 1 for i in range(3)]
    rng = [np.arange(-bound[i], bound[i] + 1) for i in range(3)]
    h3, h2, h1 = np.meshgrid(rng[2], rng[1], rng[0], indexing="ij")
    hkl = np.stack([h1.ravel(), h2.ravel(), h3.ravel()], axis=1)
    q = hkl @ b.T
    ql = np.sqrt(np.sum(q * q, axis=1))
    first = np.where(hkl[:, 0] != 0, hkl[:, 0], np.where(hkl[:, 1] != 0, hkl[:, 1], hkl[:, 2]))
    keep = (ql <= qmax) & (first > 0)
    hkl, ql = hkl[keep], ql[keep]
    order = np.lexsort((hkl[:, 2], hkl[:, 1], hkl[:, 0], ql))
    return np.ascontiguousarray(hkl[order], dtype=np.int32), ql[order]


def line_modes(nintervals, s0, s1, direction):
    """The q line of `--qpoints line:NINT,S0,S1,DX,DY,DZ`
    (src/sqc_dynamics.f90:293-327): modes (i, 0, 0), i = 0..NINT, on the basis
    b1 = dq u with base s0 u.  Returns (hkl, recip, base)."""
    u = np.asarray(direction, dtype=np.float64)
    u = u / np.sqrt(np.sum(u ** 2))
    dq = (s1 - s0) / nintervals
    hkl = np.zeros((nintervals + 1, 3), dtype=np.int32)
    hkl[:, 0] = np.arange(nintervals + 1)
    recip = np.zeros((3, 3))
    recip[:, 0] = dq * u
    base = s0 * u if abs(s0) > 0 else None
    return hkl, recip, base


class HipDynAmplitudes:
    """rho_a(n, t) of a trajectory on one HIP device.

    recip: 3x3 with b1, b2, b3 in the columns (the reciprocal basis of the
    cell for lattice modes); types: LAMMPS type id per atom (species = the
    present ids, ascending); hkl: (nmodes, 3) Miller indices; base: optional
    3-vector added to every phase (a q line's start)."""

    def __init__(self, recip, types, hkl, base=None, device=0):
        self.lib = _lib.load()
        types = np.asarray(types, dtype=np.int64)
        self.natoms = int(types.size)
        self.species_type = np.unique(types)
        self.species_of = np.ascontiguousarray(np.searchsorted(self.species_type, types),
                                               dtype=np.int32)
        self.hkl = np.ascontiguousarray(hkl, dtype=np.int32).reshape(-1, 3)
        self.nmodes = int(self.hkl.shape[0])
        self.nspecies = len(self.species_type)
        recip = np.asarray(recip, dtype=np.float64).reshape(3, 3)
        cfg = _lib.DynConfig()
        cfg.abi_version = _lib.ABI_VERSION
        cfg.device = device
        cfg.natoms = self.natoms
        cfg.nspecies = self.nspecies
        cfg.base_on = 0 if base is None else 1
        cfg.species_of = self.species_of.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        for i in range(9):
            cfg.recip[i] = float(recip.ravel(order="F")[i])
        if base is not None:
            for i in range(3):
                cfg.base[i] = float(base[i])
        cfg.nmodes = self.nmodes
        cfg.hkl = self.hkl.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        plan = ctypes.c_void_p()
        _lib.check(self.lib.sqc_hip_dyn_plan_create(ctypes.byref(plan), ctypes.byref(cfg)),
                   "sqc_hip_dyn_plan_create", warn_ok=False)
        self.plan = plan

    @property
    def info(self):
        info = _lib.DynInfo()
        _lib.check(self.lib.sqc_hip_dyn_plan_info(self.plan, ctypes.byref(info)), "dyn_plan_info")
        return info

    def rho(self, positions):
        """(F, S, nmodes) complex128 for host positions (N, 3) or (F, N, 3)."""
        pos = np.ascontiguousarray(positions, dtype=np.float64)
        nf = 1 if pos.ndim == 2 else pos.shape[0]
        if pos.reshape(-1, 3).shape[0] != nf * self.natoms:
            raise ValueError("positions do not match the atom count of the plan")
        out = np.zeros((nf, self.nspecies, self.nmodes), dtype=np.complex128)
        _lib.check(self.lib.sqc_hip_dyn_rho(self.plan, pos.ctypes.data, nf, out.ctypes.data),
                   "sqc_hip_dyn_rho", warn_ok=False)
        return out

    def rho_device(self, pos_tensor, out=None):
        """Device positions (torch float64, (F, N, 3)) -> (F, S, nmodes)
        complex128 on the device, ordered on torch's current stream."""
        import torch
        t = pos_tensor
        if t.dtype != torch.float64 or not t.is_contiguous() or t.numel() % (3 * self.natoms):
            raise ValueError("need a contiguous float64 tensor of (F, N, 3) positions")
        nf = t.numel() // (3 * self.natoms)
        if out is None:
            out = torch.empty((nf, self.nspecies, self.nmodes), dtype=torch.complex128,
                              device=t.device)
        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        _lib.check(self.lib.sqc_hip_dyn_rho_device(self.plan, t.data_ptr(), nf, out.data_ptr(),
                                                   stream), "sqc_hip_dyn_rho_device", warn_ok=False)
        return out

    def profile(self, enable=True):
        _lib.check(self.lib.sqc_hip_dyn_profile(self.plan, 1 if enable else 0), "dyn_profile")

    def stage_times(self):
        """{stage: (total ms, launches)} since the last call (synchronises)."""
        names = ["dyn_phase", "dyn_mfma", "dyn_gather"]
        ms = (ctypes.c_double * 3)()
        cnt = (ctypes.c_int64 * 3)()
        _lib.check(self.lib.sqc_hip_dyn_stage_times(self.plan, ms, cnt, 3), "dyn_stage_times")
        return {n: (ms[i], cnt[i]) for i, n in enumerate(names) if cnt[i]}

    def close(self):
        if getattr(self, "plan", None):
            self.lib.sqc_hip_dyn_destroy(self.plan)
            self.plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
