// sqc_elements.cpp -- species labels of the -m TYPE:LABEL mapping and their
// scattering data: the bound coherent neutron length of the element and the
// IT92 X-ray form-factor coefficients of the label (neutral atom, ion or
// valence state).
//
// The lookup restates the reference's rules (src/sqc_elements.f90): the
// aliases of the data module, case-insensitive labels, the element as the
// longest one- or two-letter symbol that leaves a valid suffix, suffixes
// "", "val", "N+" / "N-" / "+" / "-" (split_species_name and
// parse_charge_suffix, :293-394), the canonical label of an ion with its
// digits ("Na1+", species_name :396-411), the neutral row of a bare symbol
// (element_neutral_species) and the two failures of table_find_species
// (:241-291) with their messages.  The data rows are generated from the
// reference's data module into sqc_element_table.inc.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>

#include "sqc_hip.h"

namespace {

struct ElementRow {
  const char *symbol;
  int z;
  int has_neutron;
  double b;
};
struct SpeciesRow {
  const char *label;
  int element;  // row in kElements
  int charge;
  int valence;
  double coef[9];
};
struct AliasRow {
  const char *from, *to;
};

#include "sqc_element_table.inc"

thread_local std::string g_species_error;

std::string lower(const std::string &s) {
  std::string o(s);
  for (char &c : o)
    if (c >= 'A' && c <= 'Z') c = (char)(c + 32);
  return o;
}

// trim(adjustl(s)): no leading or trailing blanks
std::string strip(const std::string &s) {
  size_t a = 0, b = s.size();
  while (a < b && s[a] == ' ') ++a;
  while (b > a && s[b - 1] == ' ') --b;
  return s.substr(a, b - a);
}

// "SI" and "si" both become "Si" (the first one or two characters)
std::string normalise(const std::string &s, size_t n) {
  std::string o = s.substr(0, std::min(n, s.size()));
  for (size_t i = 0; i < o.size(); ++i) {
    char &c = o[i];
    if (i == 0 && c >= 'a' && c <= 'z') c = (char)(c - 32);
    if (i > 0 && c >= 'A' && c <= 'Z') c = (char)(c + 32);
  }
  return strip(o);
}

int element_index(const std::string &sym) {
  if (sym.empty()) return -1;
  for (int i = 0; i < kNumElements; ++i)
    if (sym == kElements[i].symbol) return i;
  return -1;
}

// Suffix of a label: "", "val", "2-", "1+" or "+" (parse_charge_suffix).
bool parse_suffix(const std::string &tok, int *charge, bool *valence) {
  *charge = 0;
  *valence = false;
  const size_t n = tok.size();
  if (n == 0) return true;
  if (tok == "val") {
    *valence = true;
    return true;
  }
  long digits = 0;
  size_t i = 0;
  while (i < n && tok[i] >= '0' && tok[i] <= '9') {
    digits = digits * 10 + (tok[i] - '0');
    if (digits > 1000000) return false;
    ++i;
  }
  if (i + 1 != n) return false;  // exactly one sign character after the digits
  if (tok[i] != '+' && tok[i] != '-') return false;
  if (digits == 0 && i > 0) return false;
  if (digits == 0) digits = 1;
  *charge = tok[i] == '-' ? -(int)digits : (int)digits;
  return true;
}

// Element, charge and valence flag of a label (split_species_name).
bool split(const std::string &text, int *element, int *charge, bool *valence) {
  if (text.empty()) return false;
  if (text.size() >= 2) {
    const int e = element_index(normalise(text, 2));
    if (e >= 0 && parse_suffix(lower(strip(text.substr(2))), charge, valence)) {
      *element = e;
      return true;
    }
  }
  const int e = element_index(normalise(text, 1));
  if (e < 0) return false;
  if (!parse_suffix(lower(strip(text.substr(1))), charge, valence)) return false;
  *element = e;
  return true;
}

void fill_element(int e, sqc_hip_species_info *info) {
  std::memset(info, 0, sizeof(*info));
  std::snprintf(info->symbol, sizeof(info->symbol), "%s", kElements[e].symbol);
  std::snprintf(info->label, sizeof(info->label), "%s", kElements[e].symbol);
  info->z = kElements[e].z;
  info->has_neutron = kElements[e].has_neutron;
  info->neutron_b = kElements[e].b;
}

void fill_species(int k, sqc_hip_species_info *info) {
  fill_element(kSpecies[k].element, info);
  std::snprintf(info->label, sizeof(info->label), "%s", kSpecies[k].label);
  info->charge = kSpecies[k].charge;
  info->valence = kSpecies[k].valence;
  info->has_xray = 1;
  for (int i = 0; i < 9; ++i) info->xray_coef[i] = kSpecies[k].coef[i];
}

// The neutral row of an element: charge 0 and no valence flag.
int neutral_species(int e) {
  for (int k = 0; k < kNumSpecies; ++k)
    if (kSpecies[k].element == e && kSpecies[k].charge == 0 && !kSpecies[k].valence) return k;
  return -1;
}

}  // namespace

extern "C" {

int sqc_hip_species_lookup(const char *text, sqc_hip_species_info *info) {
  if (!text || !info) {
    g_species_error = "null argument";
    return SQC_HIP_ERR_BAD_ARG;
  }
  const std::string given = strip(text);
  std::string wanted = given;
  for (int i = 0; i < kNumAliases; ++i)
    if (lower(wanted) == lower(kAliases[i].from)) wanted = kAliases[i].to;
  int e = -1, charge = 0;
  bool valence = false;
  if (!split(wanted, &e, &charge, &valence)) {
    g_species_error = "unknown element symbol \"" + given + "\" in the element mapping";
    return SQC_HIP_ERR_UNKNOWN_ELEMENT;
  }
  if (charge == 0 && !valence) {
    const int k = neutral_species(e);
    if (k >= 0)
      fill_species(k, info);
    else
      fill_element(e, info);  // no X-ray row: has_xray = 0
    std::snprintf(info->label, sizeof(info->label), "%s", kElements[e].symbol);
    return SQC_HIP_OK;
  }
  std::string canonical = kElements[e].symbol;
  if (valence)
    canonical += "val";
  else
    canonical += std::to_string(charge < 0 ? -charge : charge) + (charge > 0 ? "+" : "-");
  for (int k = 0; k < kNumSpecies; ++k)
    if (lower(canonical) == lower(kSpecies[k].label)) {
      fill_species(k, info);
      return SQC_HIP_OK;
    }
  std::string states;
  for (int k = 0; k < kNumSpecies; ++k)
    if (kSpecies[k].element == e) states += (states.empty() ? "" : ", ") + std::string(kSpecies[k].label);
  g_species_error = "element \"" + std::string(kElements[e].symbol) + "\" has no charge state \"" +
                    given + "\"; available: " + states;
  fill_element(e, info);
  return SQC_HIP_ERR_NO_CHARGE_STATE;
}

int sqc_hip_species_table(int32_t kind, int32_t index, sqc_hip_species_info *info,
                          int32_t *count) {
  const int n = kind == 0 ? kNumElements : kind == 1 ? kNumSpecies : -1;
  if (n < 0) {
    g_species_error = "kind must be 0 (elements) or 1 (X-ray species)";
    return SQC_HIP_ERR_BAD_ARG;
  }
  if (count) *count = n;
  if (!info) return SQC_HIP_OK;
  if (index < 0 || index >= n) {
    g_species_error = "table index out of range";
    return SQC_HIP_ERR_BAD_ARG;
  }
  if (kind == 0) {
    const int k = neutral_species(index);
    if (k >= 0)
      fill_species(k, info);
    else
      fill_element(index, info);
    std::snprintf(info->label, sizeof(info->label), "%s", kElements[index].symbol);
  } else {
    fill_species(index, info);
  }
  return SQC_HIP_OK;
}

const char *sqc_hip_species_error(void) { return g_species_error.c_str(); }

}  // extern "C"
