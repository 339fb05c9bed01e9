"""ctypes binding of libsqc_hip.so (declarations in include/sqc_hip.h).

The product path always goes through this library; there is no CPU
fallback.  ``load()`` raises when the shared object is missing.
"""

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# SQC_HIP_LIBRARY selects another build of the library (A/B measurements)
LIB_PATH = os.environ.get("SQC_HIP_LIBRARY", os.path.join(HERE, "libsqc_hip.so"))

ABI_VERSION = 2
F64, F32 = 0, 1
WEIGHT = {"unit": 0, "neutron": 1, "xray": 2}
STAGES = ["atom_sort", "spread_z", "dft_y", "dft_x_bin", "finalize", "adj_x", "adj_y",
          "adj_z_gather", "adj_sum"]
NORM = {"mean": 1, "self": 2, "n": 3}

_c = ctypes
_vp = _c.c_void_p


class Config(_c.Structure):
    _fields_ = [
        ("abi_version", _c.c_int32), ("device", _c.c_int32), ("precision", _c.c_int32),
        ("weight_kind", _c.c_int32), ("norm", _c.c_int32), ("partials", _c.c_int32),
        ("want_grid", _c.c_int32), ("nq", _c.c_int32),
        ("qmin", _c.c_double), ("qmax", _c.c_double), ("eps", _c.c_double),
        ("upsampfac", _c.c_double), ("cell", _c.c_double * 9), ("origin", _c.c_double * 3),
        ("natoms", _c.c_int64), ("nspecies", _c.c_int32), ("mode_partials", _c.c_int32),
        ("species_of", _c.POINTER(_c.c_int32)), ("amp_const", _c.POINTER(_c.c_double)),
        ("xray_coef", _c.POINTER(_c.c_double)), ("tune", _c.c_int32),
        ("max_batch", _c.c_int32),
    ]


class Info(_c.Structure):
    _fields_ = [
        ("modes", _c.c_int64 * 3), ("nmodes", _c.c_int64), ("nhalf", _c.c_int64),
        ("fine", _c.c_int64 * 3), ("nspread", _c.c_int32), ("nfields", _c.c_int32),
        ("beta", _c.c_double), ("upsampfac", _c.c_double), ("device_bytes", _c.c_int64),
        ("nlines", _c.c_int64), ("tile", _c.c_int32 * 3), ("ncomp", _c.c_int32),
        ("rho_len", _c.c_int64), ("shell_sums_in_hbm", _c.c_int32),
        ("x_lines", _c.c_int32),
    ]


class DynConfig(_c.Structure):
    _fields_ = [
        ("abi_version", _c.c_int32), ("device", _c.c_int32), ("natoms", _c.c_int64),
        ("nspecies", _c.c_int32), ("base_on", _c.c_int32),
        ("species_of", _c.POINTER(_c.c_int32)), ("recip", _c.c_double * 9),
        ("base", _c.c_double * 3), ("nmodes", _c.c_int64), ("hkl", _c.POINTER(_c.c_int32)),
    ]


class DynInfo(_c.Structure):
    _fields_ = [
        ("nmodes", _c.c_int64), ("macs_per_frame", _c.c_int64), ("computed_pairs", _c.c_int64),
        ("device_bytes", _c.c_int64), ("nitems", _c.c_int32), ("nchunks", _c.c_int32),
        ("chunk_atoms", _c.c_int32), ("column_axis", _c.c_int32),
    ]


class DebyeConfig(_c.Structure):
    _fields_ = [
        ("abi_version", _c.c_int32), ("device", _c.c_int32), ("natoms", _c.c_int64),
        ("ntypes", _c.c_int32), ("pbc", _c.c_int32 * 3), ("type_id", _c.POINTER(_c.c_int32)),
        ("cell", _c.c_double * 9), ("origin", _c.c_double * 3), ("rmax", _c.c_double),
        ("dr", _c.c_double), ("skin", _c.c_double),
    ]


class DebyeInfo(_c.Structure):
    _fields_ = [
        ("nbins", _c.c_int32), ("ntypes", _c.c_int32), ("div", _c.c_int32 * 3),
        ("hist_in_lds", _c.c_int32), ("workgroups", _c.c_int32), ("tile", _c.c_int32),
        ("pairs_per_frame", _c.c_int64), ("device_bytes", _c.c_int64), ("nframes", _c.c_int64),
    ]


class SpeciesInfo(_c.Structure):
    _fields_ = [
        ("label", _c.c_char * 16), ("symbol", _c.c_char * 4), ("z", _c.c_int32),
        ("charge", _c.c_int32), ("valence", _c.c_int32), ("has_neutron", _c.c_int32),
        ("has_xray", _c.c_int32), ("reserved", _c.c_int32), ("neutron_b", _c.c_double),
        ("xray_coef", _c.c_double * 9),
    ]


class DumpFrame(_c.Structure):
    _fields_ = [
        ("natoms", _c.c_int64), ("timestep", _c.c_int64), ("cell", _c.c_double * 9),
        ("origin", _c.c_double * 3), ("pbc", _c.c_int32 * 3), ("triclinic", _c.c_int32),
        ("scaled", _c.c_int32), ("has_id", _c.c_int32), ("unwrapped", _c.c_int32),
        ("unwrapped_from_dump", _c.c_int32), ("n_ambiguous", _c.c_int64),
        ("lineno", _c.c_int64), ("frames", _c.c_int64),
    ]


# Every entry point of include/sqc_hip.h with its ctypes signature.
SIGNATURES = {
    "sqc_hip_device_count": (_c.c_int, [_c.POINTER(_c.c_int)]),
    "sqc_hip_plan_create": (_c.c_int, [_c.POINTER(_vp), _c.POINTER(Config)]),
    "sqc_hip_plan_info": (_c.c_int, [_vp, _c.POINTER(Info)]),
    "sqc_hip_set_tiles": (_c.c_int, [_vp, _c.c_int, _c.c_int, _c.c_int, _c.c_int]),
    "sqc_hip_shells": (_c.c_int, [_vp, _vp, _vp]),
    "sqc_hip_modes": (_c.c_int, [_vp, _vp, _vp, _vp]),
    "sqc_hip_accumulate": (_c.c_int, [_vp, _vp, _c.c_int64]),
    "sqc_hip_accumulate_device": (_c.c_int, [_vp, _vp, _c.c_int64]),
    "sqc_hip_fetch": (_c.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "sqc_hip_fetch_mode_partials": (_c.c_int, [_vp, _vp]),
    "sqc_hip_reset": (_c.c_int, [_vp]),
    "sqc_hip_forward_device": (_c.c_int, [_vp, _vp, _c.c_int64, _vp, _vp, _vp]),
    "sqc_hip_adjoint_device": (_c.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "sqc_hip_adjoint_device_noalloc": (_c.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "sqc_hip_gradient": (_c.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "sqc_hip_profile": (_c.c_int, [_vp, _c.c_int]),
    "sqc_hip_stage_times": (_c.c_int, [_vp, _vp, _vp, _c.c_int]),
    "sqc_hip_destroy": (_c.c_int, [_vp]),
    "sqc_hip_error_string": (_c.c_char_p, [_c.c_int]),
    "sqc_hip_last_error": (_c.c_char_p, []),
    "sqc_hip_config_size": (_c.c_size_t, []),
    "sqc_hip_dyn_plan_create": (_c.c_int, [_c.POINTER(_vp), _c.POINTER(DynConfig)]),
    "sqc_hip_dyn_plan_info": (_c.c_int, [_vp, _c.POINTER(DynInfo)]),
    "sqc_hip_dyn_rho_device": (_c.c_int, [_vp, _vp, _c.c_int64, _vp, _vp]),
    "sqc_hip_dyn_rho": (_c.c_int, [_vp, _vp, _c.c_int64, _vp]),
    "sqc_hip_dyn_profile": (_c.c_int, [_vp, _c.c_int]),
    "sqc_hip_dyn_stage_times": (_c.c_int, [_vp, _vp, _vp, _c.c_int]),
    "sqc_hip_dyn_destroy": (_c.c_int, [_vp]),
    "sqc_hip_debye_plan_create": (_c.c_int, [_c.POINTER(_vp), _c.POINTER(DebyeConfig)]),
    "sqc_hip_debye_plan_info": (_c.c_int, [_vp, _c.POINTER(DebyeInfo)]),
    "sqc_hip_debye_accumulate_device": (_c.c_int, [_vp, _vp, _c.c_int64, _vp]),
    "sqc_hip_debye_profile": (_c.c_int, [_vp, _c.c_int]),
    "sqc_hip_debye_stage_times": (_c.c_int, [_vp, _vp, _vp, _c.c_int]),
    "sqc_hip_debye_accumulate": (_c.c_int, [_vp, _vp, _c.c_int64]),
    "sqc_hip_debye_fetch": (_c.c_int, [_vp, _vp, _c.POINTER(_c.c_int64)]),
    "sqc_hip_debye_reset": (_c.c_int, [_vp]),
    "sqc_hip_debye_destroy": (_c.c_int, [_vp]),
    "sqc_hip_species_lookup": (_c.c_int, [_c.c_char_p, _c.POINTER(SpeciesInfo)]),
    "sqc_hip_species_table": (_c.c_int, [_c.c_int32, _c.c_int32, _c.POINTER(SpeciesInfo),
                                         _c.POINTER(_c.c_int32)]),
    "sqc_hip_species_error": (_c.c_char_p, []),
    "sqc_dump_open": (_c.c_int, [_c.POINTER(_vp), _c.c_char_p, _c.c_int]),
    "sqc_dump_next": (_c.c_int, [_vp, _c.POINTER(DumpFrame)]),
    "sqc_dump_copy": (_c.c_int, [_vp, _vp, _vp, _vp]),
    "sqc_dump_read_frames": (_c.c_int, [_vp, _c.c_int64, _c.c_int64, _vp, _vp,
                                        _c.POINTER(_c.c_int64), _c.POINTER(DumpFrame)]),
    "sqc_dump_read_frames_begin": (_c.c_int, [_vp, _c.c_int64, _c.c_int64, _vp, _vp]),
    "sqc_dump_read_frames_end": (_c.c_int, [_vp, _c.POINTER(_c.c_int64)]),
    "sqc_dump_close": (_c.c_int, [_vp]),
    "sqc_dump_error_string": (_c.c_char_p, [_c.c_int]),
    "sqc_dump_message": (_c.c_char_p, [_vp]),
}

_LIB = None


class SqcHipError(RuntimeError):
    def __init__(self, code, where):
        lib = load()
        msg = lib.sqc_hip_last_error().decode() or lib.sqc_hip_error_string(code).decode()
        super().__init__("%s failed (%d): %s" % (where, code, msg))
        self.code = code


def load():
    """Load libsqc_hip.so; raise if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libsqc_hip.so is not built (%s); run __graft_entry__.build() or "
                           "python -m sqcalc_amd.build" % LIB_PATH)
    lib = _c.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.sqc_hip_config_size() != _c.sizeof(Config):
        raise RuntimeError("sqc_hip_config layout mismatch: library %d bytes, binding %d"
                           % (lib.sqc_hip_config_size(), _c.sizeof(Config)))
    _LIB = lib
    return lib


def check(code, where, warn_ok=True):
    if code == 0 or (warn_ok and code == 1):
        return code
    raise SqcHipError(code, where)


def device_count():
    n = _c.c_int(0)
    rc = load().sqc_hip_device_count(_c.byref(n))
    return n.value if rc == 0 else 0
