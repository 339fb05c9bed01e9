"""Species labels of the -m TYPE:LABEL mapping and their scattering data.

A thin layer over the library's lookup (sqc_hip_species_lookup in
include/sqc_hip.h), which restates the reference's rules for labels
(src/sqc_elements.f90:241-411: aliases, case, ions "O2-" / "Na1+" / "Fe+",
valence states "Cval") over the data of src/sqc_element_data.f90: 104
elements with their bound coherent neutron lengths b [fm] (Neutron News 3
(1992) 29-37) and 213 X-ray species with IT92 form-factor coefficients
(International Tables C (1992) 6.1.1.4, f(q) = c + sum_i a_i exp(-b_i
(q / 4 pi)^2)).  An ion or valence label takes its element's neutron length
(src/sqc_weights.f90:136-170).
"""

import ctypes

from . import _lib


class SpeciesError(ValueError):
    """An unresolvable label; code 34 unknown element, 35 no such charge state."""

    def __init__(self, code, message):
        super().__init__(message)
        self.code = code


def lookup(label):
    """dict(label, symbol, z, charge, valence, has_neutron, has_xray,
    neutron_b, xray_coef) of a species label, or SpeciesError."""
    lib = _lib.load()
    info = _lib.SpeciesInfo()
    rc = lib.sqc_hip_species_lookup(str(label).encode(), info)
    if rc != 0:
        raise SpeciesError(rc, lib.sqc_hip_species_error().decode())
    return _as_dict(info)


def _as_dict(info):
    return {"label": info.label.decode(), "symbol": info.symbol.decode(), "z": info.z,
            "charge": info.charge, "valence": bool(info.valence),
            "has_neutron": bool(info.has_neutron), "has_xray": bool(info.has_xray),
            "neutron_b": info.neutron_b, "xray_coef": tuple(info.xray_coef)}


def table(kind):
    """Every row of the element table ("elements": each element with its
    neutral X-ray row) or of the X-ray species table ("species")."""
    lib = _lib.load()
    k = {"elements": 0, "species": 1}[kind]
    n = ctypes.c_int32(0)
    _lib.check(lib.sqc_hip_species_table(k, 0, None, ctypes.byref(n)), "species_table")
    out = []
    for i in range(n.value):
        info = _lib.SpeciesInfo()
        _lib.check(lib.sqc_hip_species_table(k, i, info, None), "species_table")
        out.append(_as_dict(info))
    return out


def canonical(label):
    return lookup(label)["label"]


def neutron_b(label):
    s = lookup(label)
    if not s["has_neutron"]:
        raise SpeciesError(15, "no neutron scattering length for %s; use --weight xray or unit"
                           % s["label"])
    return s["neutron_b"]


def xray_coefficients(label):
    s = lookup(label)
    if not s["has_xray"]:
        raise SpeciesError(15, "no X-ray form factor for %s; use --weight neutron or unit"
                           % s["label"])
    return s["xray_coef"]
