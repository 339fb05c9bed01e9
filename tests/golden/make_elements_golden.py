"""Write tests/golden/elements/elements.json: the element and X-ray species tables
of the reference, as data, read from its generated data module
src/sqc_element_data.f90 (the published values it tabulates: CRC masses,
International Tables C (1992) 6.1.1.4 IT92 coefficients, Neutron News 3
(1992) bound coherent lengths).

The fixture pins the library's own table and lookup (sqc_hip_species_lookup)
and gives the oracle its weights.  Run in the container that holds the
reference:

  python tests/golden/make_elements_golden.py [/root/reference/src/sqc_element_data.f90]
"""

import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/src/sqc_element_data.f90"


def arrays(text):
    """name -> list of values of every `name(...) = [...]` or
    `name(...) = reshape([...], ...)` array constructor in the module."""
    out = {}
    pat = re.compile(r"::\s*(\w+)\([^)]*\)\s*=\s*(reshape\(\s*)?\[(.*?)\]", re.S)
    for m in pat.finditer(text):
        body = m.group(3).replace("&", " ")
        vals = []
        for tok in re.findall(r'"[^"]*"|\.true\.|\.false\.|[-+]?[0-9.]+(?:[dDeE][-+]?\d+)?', body):
            if tok.startswith('"'):
                vals.append(tok[1:-1].strip())
            elif tok in (".true.", ".false."):
                vals.append(tok == ".true.")
            elif re.search(r"[.dDeE]", tok):
                vals.append(float(tok.replace("d", "e").replace("D", "e")))
            else:
                vals.append(int(tok))
        out[m.group(1)] = vals
    return out


def main():
    src = sys.argv[1] if len(sys.argv) > 1 else SRC
    text = open(src).read()
    a = arrays(text)
    ne, ns = len(a["element_z"]), len(a["species_label"])
    assert ne == int(re.search(r"n_elements\s*=\s*(\d+)", text).group(1))
    assert ns == int(re.search(r"n_species\s*=\s*(\d+)", text).group(1))
    xa, xb = a["species_xray_a"], a["species_xray_b"]
    assert len(xa) == len(xb) == 4 * ns
    elements = [{"z": a["element_z"][i], "symbol": a["element_symbol"][i],
                 "neutron_b": float(a["element_neutron_b"][i]),
                 "has_neutron": a["element_has_neutron"][i],
                 "neutral_species": a["element_neutral_species"][i]} for i in range(ne)]
    species = [{"label": a["species_label"][k], "element": a["species_element"][k],
                "charge": a["species_charge"][k], "valence": a["species_valence"][k],
                "coef": [float(x) for x in xa[4 * k:4 * k + 4] + xb[4 * k:4 * k + 4]]
                + [float(a["species_xray_c"][k])]} for k in range(ns)]
    aliases = dict(zip(a["species_alias_from"], a["species_alias_to"]))
    out = {"source": "src/sqc_element_data.f90", "elements": elements, "species": species,
           "aliases": aliases}
    path = os.path.join(HERE, "elements", "elements.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=0)
    print("wrote %s: %d elements, %d species, %d aliases" % (path, ne, ns, len(aliases)))


if __name__ == "__main__":
    main()
