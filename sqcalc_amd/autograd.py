"""torch.autograd entry point: S(q) of one frame and its adjoint on the GPU.

``structure_factor(pos, sf)`` returns (S, P) for positions ``pos`` (N, 3)
float64 on the device: S[nq] = num / den (one frame, the normalisation of
src/sqc_structure_factor.f90:657-665) and P[npairs, nq] the pair columns
pair / (N count) (:696-719, OVITO convention).  Backward runs the
hand-written adjoint of libsqc_hip (type-2 transform of the shell weights
with the kernel derivative); only the species amplitudes at the kept modes
are saved between the passes, the fine grid is recomputed.  There is no
counterpart in the reference, which has no gradient.
"""

import torch

from . import _lib


def _stream():
    return ctypes_void(torch.cuda.current_stream().cuda_stream)


def ctypes_void(v):
    import ctypes
    return ctypes.c_void_p(v)


def _norms(sf, device):
    """Per-shell factors 1/den (0 where den = 0) and 1/(N count) on the
    device, built once per plan and device: a host-to-device copy per frame
    would synchronise the stream."""
    cache = sf.__dict__.setdefault("_device_norms", {})
    key = str(device)
    if key not in cache:
        import numpy as np
        den = np.asarray(sf.den, dtype=np.float64)
        inv_den = np.where(den > 0, 1.0 / np.where(den > 0, den, 1.0), 0.0)
        inv_pair = 1.0 / (sf.natoms * np.maximum(np.asarray(sf.shell_count, np.float64), 1.0))
        cache[key] = (torch.tensor(inv_den, device=device), torch.tensor(inv_pair, device=device))
    return cache[key]


def _pair_rows(sf, device):
    """Row a * S + b of every pair a <= b in the order of the P rows, as a
    device index tensor built once per plan and device."""
    cache = sf.__dict__.setdefault("_device_pair_rows", {})
    key = str(device)
    if key not in cache:
        S = len(sf.species_type)
        rows = [a * S + b for a in range(S) for b in range(a, S)]
        cache[key] = torch.tensor(rows, dtype=torch.int64, device=device)
    return cache[key]


class _SqFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos, sf):
        if pos.dtype != torch.float64 or not pos.is_cuda:
            raise TypeError("positions must be a float64 tensor on the HIP device")
        pos = pos.contiguous()
        info = sf.info
        nq = sf.nq
        out = torch.empty((info.ncomp, nq), dtype=torch.float64, device=pos.device)
        cdt = torch.complex64 if sf.precision == "single" else torch.complex128
        rho = torch.empty((info.nfields, info.rho_len), dtype=cdt, device=pos.device)
        rc = sf.lib.sqc_hip_forward_device(sf.plan, pos.data_ptr(), 1, out.data_ptr(),
                                           rho.data_ptr(), _stream())
        _lib.check(rc, "sqc_hip_forward_device")
        inv_den, inv_pair = _norms(sf, pos.device)
        S = out[0] * inv_den
        P = out[1:] * inv_pair
        ctx.sf = sf
        # an output the loss does not use comes back as None rather than as
        # zeros, so a loss on S alone skips the pair terms of the adjoint
        ctx.set_materialize_grads(False)
        ctx.save_for_backward(pos, rho)
        return S, P

    @staticmethod
    def backward(ctx, gS, gP):
        pos, rho = ctx.saved_tensors
        sf = ctx.sf
        nq = sf.nq
        gS = (gS if gS is not None else torch.zeros(nq, dtype=torch.float64,
                                                    device=pos.device)).contiguous()
        gP_full = None
        if gP is not None and sf.partials:
            # the pair rows go to their (a, b <= a) places on the device, with
            # no look at their values: a zero row adds nothing, and a host
            # test of the values would synchronise (and break graph capture)
            S = len(sf.species_type)
            gP_full = torch.zeros((S * S, nq), dtype=torch.float64, device=pos.device)
            gP_full.index_copy_(0, _pair_rows(sf, pos.device), gP.contiguous())
            gP_full = gP_full.view(S, S, nq)
        grad = torch.empty_like(pos)
        rc = sf.lib.sqc_hip_adjoint_device(sf.plan, pos.data_ptr(), rho.data_ptr(),
                                           gS.data_ptr(),
                                           gP_full.data_ptr() if gP_full is not None else None,
                                           grad.data_ptr(), _stream())
        _lib.check(rc, "sqc_hip_adjoint_device")
        return grad, None


def structure_factor(pos, sf):
    """(S, P) of one frame; differentiable with respect to ``pos``."""
    return _SqFunction.apply(pos, sf)


def sq_matching_loss(pos, sf, target, wq=None):
    """L = sum_s w_s (S_s - T_s)^2 for one frame (the loss of the gradient workload)."""
    S, _ = structure_factor(pos, sf)
    d2 = (S - target) ** 2
    # unit weights: no weight tensor, two fewer kernels per step and per backward
    return torch.sum(d2) if wq is None else torch.sum(wq * d2)
