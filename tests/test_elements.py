"""The library's species lookup (sqc_hip_species_lookup) and its tables
against the reference's element data, on the CPU.

tests/golden/elements/elements.json holds the reference's tables as data
(tests/golden/make_elements_golden.py reads src/sqc_element_data.f90): 104
elements with their neutron lengths and 213 X-ray species with IT92
coefficients.  Every row must come back exactly, every label must resolve by
the reference's rules (src/sqc_elements.f90:241-411: aliases, any case, ion
and valence suffixes, the neutral row of a bare symbol) and the two failures
must carry the reference's messages.  Independently of the fixture, the
electron-count identity f(0) = sum a_i + c = Z - charge holds within 0.06
for every neutral atom and ion (the audit bound the reference's data module
states), which a swapped or shifted coefficient column would break.
"""

import json
import os

import pytest

from helpers import GOLDEN


@pytest.fixture(scope="module")
def data():
    with open(os.path.join(GOLDEN, "elements", "elements.json")) as fh:
        return json.load(fh)


def test_tables_match_reference_rows(data):
    from sqcalc_amd import elements as E
    els, sps = E.table("elements"), E.table("species")
    assert len(els) == len(data["elements"]) == 104
    assert len(sps) == len(data["species"]) == 213
    for got, ref in zip(els, data["elements"]):
        assert (got["symbol"], got["z"], got["has_neutron"]) == (ref["symbol"], ref["z"],
                                                                  ref["has_neutron"])
        assert got["neutron_b"] == ref["neutron_b"]
        ns = ref["neutral_species"]
        assert got["has_xray"] == (ns > 0)
        if ns:
            assert list(got["xray_coef"]) == data["species"][ns - 1]["coef"]
    for got, ref in zip(sps, data["species"]):
        e = data["elements"][ref["element"] - 1]
        assert got["label"] == ref["label"] and got["symbol"] == e["symbol"]
        assert (got["charge"], got["valence"]) == (ref["charge"], ref["valence"])
        assert list(got["xray_coef"]) == ref["coef"]
        assert got["neutron_b"] == e["neutron_b"] and got["has_neutron"] == e["has_neutron"]


def test_every_label_resolves(data):
    from sqcalc_amd import elements as E
    for ref in data["species"]:
        for text in (ref["label"], ref["label"].upper(), ref["label"].lower(),
                     "  " + ref["label"] + " "):
            got = E.lookup(text)
            assert got["label"] == ref["label"], text
            assert list(got["xray_coef"]) == ref["coef"], text
    for e in data["elements"]:
        got = E.lookup(e["symbol"].lower())
        assert got["symbol"] == e["symbol"] and got["label"] == e["symbol"]
        assert got["has_xray"] == (e["neutral_species"] > 0)
    for src, dst in data["aliases"].items():
        assert E.lookup(src)["label"] == dst and E.lookup(src.upper())["label"] == dst


def test_suffix_grammar():
    from sqcalc_amd import elements as E
    assert E.lookup("Na+")["label"] == "Na1+"        # a bare sign means 1
    assert E.lookup("Cl-")["label"] == "Cl1-"
    assert E.lookup("CVAL")["label"] == "Cval"       # C + val, not a "Cv" element
    assert E.lookup("O 2-")["label"] == "O2-"        # blanks before the suffix
    assert E.lookup("Fe3+")["charge"] == 3 and E.lookup("O2-")["charge"] == -2
    assert E.lookup("Si4+")["neutron_b"] == E.lookup("Si")["neutron_b"]   # element's b


@pytest.mark.parametrize("text,code,message", [
    ("Xx", 34, 'unknown element symbol "Xx" in the element mapping'),
    ("0+", 34, 'unknown element symbol "0+" in the element mapping'),
    ("Si0+", 34, 'unknown element symbol "Si0+" in the element mapping'),
    ("Fe7+", 35, 'element "Fe" has no charge state "Fe7+"; available: Fe, Fe2+, Fe3+'),
    ("C+", 35, 'element "C" has no charge state "C+"; available: C, Cval'),
])
def test_failures_carry_the_reference_messages(text, code, message):
    from sqcalc_amd import elements as E
    with pytest.raises(E.SpeciesError) as ex:
        E.lookup(text)
    assert ex.value.code == code and str(ex.value) == message


def test_electron_count_identity(data):
    worst = 0.0
    for sp in data["species"]:
        if sp["valence"]:
            continue
        z = data["elements"][sp["element"] - 1]["z"]
        f0 = sum(sp["coef"][:4]) + sp["coef"][8]
        worst = max(worst, abs(f0 - (z - sp["charge"])))
    assert worst < 0.06, worst


def test_weights_of_the_host():
    from sqcalc_amd import elements as E
    with pytest.raises(E.SpeciesError) as ex:
        E.neutron_b("Fm")
    assert "no neutron scattering length for Fm" in str(ex.value)
    with pytest.raises(E.SpeciesError) as ex:
        E.xray_coefficients("Es")
    assert "no X-ray form factor for Es" in str(ex.value)
    assert E.neutron_b("fe2+") == 9.45
