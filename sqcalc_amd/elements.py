"""Scattering data for the species the path is run with.

Neutron bound coherent lengths b [fm] (Neutron News 3 (1992) 29-37) per
element and IT92 X-ray form-factor coefficients (International Tables C
(1992) 6.1.1.4: f(q) = c + sum_i a_i exp(-b_i (q/4pi)^2)) per species label.
These are the published values the reference tabulates in
src/sqc_element_data.f90 (amplitude lookup: src/sqc_weights.f90:136-154,
src/sqc_elements.f90:181-239).  An ion or valence label ("O2-", "Sival")
uses the neutron length of its element.
"""

import re

NEUTRON_B = {
    "H": -3.739, "C": 6.646, "O": 5.803, "Na": 3.63, "Al": 3.449, "Si": 4.1491,
    "Cl": 9.577, "Ca": 4.7, "Ni": 10.3,
}

# label: (a1..a4, b1..b4, c)
XRAY = {
    "H": (0.493002, 0.322912, 0.140191, 0.04081, 10.5109, 26.1257, 3.14236, 57.7997, 0.003038),
    "C": (2.31, 1.02, 1.5886, 0.865, 20.8439, 10.2075, 0.5687, 51.6512, 0.2156),
    "O": (3.0485, 2.2868, 1.5463, 0.867, 13.2771, 5.7011, 0.3239, 32.9089, 0.2508),
    "Na": (4.7626, 3.1736, 1.2674, 1.1128, 3.285, 8.8422, 0.3136, 129.424, 0.676),
    "Al": (6.4202, 1.9002, 1.5936, 1.9646, 3.0387, 0.7426, 31.5472, 85.0886, 1.1151),
    "Si": (6.2915, 3.0353, 1.9891, 1.541, 2.4386, 32.3337, 0.6785, 81.6937, 1.1407),
    "Cl": (11.4604, 7.1964, 6.2556, 1.6455, 0.0104, 1.1662, 18.5194, 47.7784, -9.5574),
    "Ca": (8.6266, 7.3873, 1.5899, 1.0211, 10.4421, 0.6599, 85.7484, 178.437, 1.3751),
    "Ni": (12.8376, 7.292, 4.4438, 2.38, 3.8785, 0.2565, 12.1763, 66.3421, 1.0341),
    "O2-": (3.7504, 2.84294, 1.54298, 1.62091, 16.5151, 6.59203, 0.319201, 43.3486, 0.24206),
    "Si4+": (4.43918, 3.20345, 1.19453, 0.41653, 1.64167, 3.43757, 0.2149, 6.65365, 0.746297),
    "Na1+": (3.2565, 3.9362, 1.3998, 1.0032, 2.6671, 6.1153, 0.2001, 14.039, 0.404),
    "Ca2+": (15.6348, 7.9518, 8.4372, 0.8537, -0.0074, 0.6089, 10.3116, 25.9905, -14.875),
    "Ni2+": (11.4166, 7.4005, 5.3442, 0.9773, 3.6766, 0.2449, 8.873, 22.1626, 0.8614),
    "Al3+": (4.17448, 3.3876, 1.20296, 0.528137, 1.93816, 4.14553, 0.228753, 8.28524, 0.706786),
}


def element_of(label):
    m = re.match(r"[A-Z][a-z]?", label)
    if not m:
        raise ValueError("not a species label: %r" % label)
    return m.group(0)


def neutron_b(label):
    elem = element_of(label)
    if elem not in NEUTRON_B:
        raise KeyError("no neutron scattering length for %s (available: %s)"
                       % (elem, ", ".join(sorted(NEUTRON_B))))
    return NEUTRON_B[elem]


def xray_coefficients(label):
    if label not in XRAY:
        elem = element_of(label)
        alt = sorted(k for k in XRAY if element_of(k) == elem)
        raise KeyError("no X-ray form factor for %s (available: %s)"
                       % (label, ", ".join(alt) if alt else "none"))
    return XRAY[label]
