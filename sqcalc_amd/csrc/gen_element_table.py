"""Generate sqc_element_table.inc, the element and X-ray species data of
the library's species lookup, from the reference's data module
src/sqc_element_data.f90 (published values: International Tables C (1992)
6.1.1.4 IT92 form-factor coefficients, Neutron News 3 (1992) bound coherent
neutron lengths).  A maintainer step, run where the reference lies; the
output is committed:

  python sqcalc_amd/csrc/gen_element_table.py /root/reference/src/sqc_element_data.f90
"""

import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def read_constructors(path):
    """Every parameter array of the module: name -> list of value tokens,
    collected line by line from the `= [` (or `= reshape([`) line to the
    closing `]`."""
    out, name, buf = {}, None, []
    for raw in open(path):
        line = raw.split("!", 1)[0].rstrip()
        if name is None:
            if "parameter" in line and "::" in line and "[" in line:
                head = line.split("::", 1)[1]
                name = head.split("(", 1)[0].strip()
                buf = [line.split("[", 1)[1]]
            else:
                continue
        else:
            buf.append(line)
        joined = " ".join(buf)
        if "]" in joined:
            body = joined.split("]", 1)[0].replace("&", " ")
            out[name] = [t.strip() for t in body.split(",") if t.strip()]
            name = None
    return out


def num(tok):
    return float(tok.lower().replace("d", "e"))


def main():
    src = sys.argv[1]
    a = read_constructors(src)
    sym = [t.strip('"').strip() for t in a["element_symbol"]]
    z = [int(t) for t in a["element_z"]]
    nb = [num(t) for t in a["element_neutron_b"]]
    hasn = [t == ".true." for t in a["element_has_neutron"]]
    lab = [t.strip('"').strip() for t in a["species_label"]]
    sel = [int(t) for t in a["species_element"]]
    chg = [int(t) for t in a["species_charge"]]
    val = [t == ".true." for t in a["species_valence"]]
    xa = [num(t) for t in a["species_xray_a"]]
    xb = [num(t) for t in a["species_xray_b"]]
    xc = [num(t) for t in a["species_xray_c"]]
    afrom = [t.strip('"').strip() for t in a["species_alias_from"]]
    ato = [t.strip('"').strip() for t in a["species_alias_to"]]
    ne, ns = len(sym), len(lab)
    assert len(z) == len(nb) == len(hasn) == ne
    assert len(sel) == len(chg) == len(val) == len(xc) == ns and len(xa) == len(xb) == 4 * ns
    lines = ["// Generated by gen_element_table.py from the reference's element data",
             "// (src/sqc_element_data.f90): published values, do not edit by hand.",
             "// Elements: symbol, Z, has a neutron length, b [fm].",
             "constexpr int kNumElements = %d;" % ne,
             "constexpr ElementRow kElements[kNumElements] = {"]
    for i in range(ne):
        lines.append('    {"%s", %d, %d, %r},' % (sym[i], z[i], 1 if hasn[i] else 0, nb[i]))
    lines += ["};",
              "// X-ray species: label, element (0-based row above), charge, valence",
              "// state, IT92 a1..a4, b1..b4, c.",
              "constexpr int kNumSpecies = %d;" % ns,
              "constexpr SpeciesRow kSpecies[kNumSpecies] = {"]
    for k in range(ns):
        coef = xa[4 * k:4 * k + 4] + xb[4 * k:4 * k + 4] + [xc[k]]
        lines.append('    {"%s", %d, %d, %d, {%s}},' % (lab[k], sel[k] - 1, chg[k],
                                                      1 if val[k] else 0,
                                                      ", ".join(repr(c) for c in coef)))
    lines += ["};", "// Labels that name another row.",
              "constexpr int kNumAliases = %d;" % len(afrom),
              "constexpr AliasRow kAliases[kNumAliases] = {"]
    for f, t in zip(afrom, ato):
        lines.append('    {"%s", "%s"},' % (f, t))
    lines.append("};")
    out = os.path.join(HERE, "sqc_element_table.inc")
    with open(out, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    print("wrote %s: %d elements, %d species, %d aliases" % (out, ne, ns, len(afrom)))


if __name__ == "__main__":
    main()
