/*
 * sqc_oracle.c -- FP64 CPU oracle for the reciprocal-space S(q) path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (sqcalc_amd/, include/,
 * the Fortran host) links, loads or calls this file; only tests/, the smoke
 * check and the CPU-baseline leg of bench.py do, and only as the reference
 * results to compare with (or, in bench.py, as the timed CPU path).
 *
 * It restates, in plain C, the reference's exact evaluation of the density
 * amplitudes and of the shell bookkeeping around them:
 *
 *   cell set-up        src/sqc_cell.f90:39-98      (A, B = 2 pi A^-T, fractional)
 *   mode enumeration   src/sqc_structure_factor.f90:313-399  (sf_configure)
 *   denominators       src/sqc_structure_factor.f90:425-451  (mode_denominator)
 *   direct sum         src/sqc_structure_factor.f90:1398-1515 (direct_accumulate)
 *   partial sums       src/sqc_structure_factor.f90:481-521   (sf_accumulate_partials)
 *   shell sums         src/sqc_structure_factor.f90:595-643   (sf_accumulate_values)
 *
 * plus the analytic FP64 gradient of a shell-sum loss, which has no
 * counterpart in the reference (derivation in the function header).
 *
 * Every floating-point expression is evaluated in the order the reference
 * writes it and the file is compiled with -ffp-contract=off, so that the
 * mode set and shell indices agree bit for bit with the reference rules.
 *
 * How each function is pinned (so that a plausible slip fails a check):
 *   sqo_cell_setup      a_i . b_j = 2 pi delta_ij invariant (tests/test_oracle.py)
 *   sqo_modes           brute-force numpy enumeration of the same rule, golden
 *                       tables written by the reference's own numpy reference
 *                       (tests/golden/), Bragg-shell closed form S = N
 *   sqo_direct          golden tables of the reference numpy reference (unit,
 *                       neutron, xray, triclinic); closed forms: single species
 *                       S = S_aa, sum rule S = S_aa + 2 S_ab + S_bb, perfect
 *                       lattice S = N, B2 antiphase +-N/4; separable and
 *                       cos/sin variants agree to 1e-12
 *   sqo_direct_grad     central finite differences of sqo_direct (h = 1e-4 A)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static const double SQO_PI = 3.14159265358979323846;

/* ---------------------------------------------------------------------- */
/* Cell: A holds the lattice vectors in its columns, column-major (a1 =     */
/* A[0..2]).  B likewise holds b1, b2, b3 in its columns.                   */
/* ---------------------------------------------------------------------- */

#define M(a, i, j) ((a)[(i) + 3 * (j)])

static double det3(const double *m) {
  return M(m, 0, 0) * (M(m, 1, 1) * M(m, 2, 2) - M(m, 1, 2) * M(m, 2, 1)) -
         M(m, 0, 1) * (M(m, 1, 0) * M(m, 2, 2) - M(m, 1, 2) * M(m, 2, 0)) +
         M(m, 0, 2) * (M(m, 1, 0) * M(m, 2, 1) - M(m, 1, 1) * M(m, 2, 0));
}

/* Explicit cofactor inverse, src/sqc_cell.f90:135-151. */
void sqo_inverse3(const double *m, double *inv) {
  double det = det3(m);
  M(inv, 0, 0) = (M(m, 1, 1) * M(m, 2, 2) - M(m, 1, 2) * M(m, 2, 1));
  M(inv, 1, 0) = -(M(m, 1, 0) * M(m, 2, 2) - M(m, 1, 2) * M(m, 2, 0));
  M(inv, 2, 0) = (M(m, 1, 0) * M(m, 2, 1) - M(m, 1, 1) * M(m, 2, 0));
  M(inv, 0, 1) = -(M(m, 0, 1) * M(m, 2, 2) - M(m, 0, 2) * M(m, 2, 1));
  M(inv, 1, 1) = (M(m, 0, 0) * M(m, 2, 2) - M(m, 0, 2) * M(m, 2, 0));
  M(inv, 2, 1) = -(M(m, 0, 0) * M(m, 2, 1) - M(m, 0, 1) * M(m, 2, 0));
  M(inv, 0, 2) = (M(m, 0, 1) * M(m, 1, 2) - M(m, 0, 2) * M(m, 1, 1));
  M(inv, 1, 2) = -(M(m, 0, 0) * M(m, 1, 2) - M(m, 0, 2) * M(m, 1, 0));
  M(inv, 2, 2) = (M(m, 0, 0) * M(m, 1, 1) - M(m, 0, 1) * M(m, 1, 0));
  for (int k = 0; k < 9; ++k) inv[k] = inv[k] / det;
}

/* b_j = 2 pi (row j of A^-1), src/sqc_cell.f90:76-86.  Returns the volume. */
double sqo_cell_setup(const double *a, double *b) {
  double inv[9];
  const double two_pi = 2.0 * acos(-1.0);
  sqo_inverse3(a, inv);
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i) M(b, i, j) = two_pi * M(inv, j, i);
  return fabs(det3(a));
}

/* Restricted triclinic cell from LAMMPS bounding limits,
 * src/sqc_cell.f90:39-63.  tilt = (xy, xz, yz). */
void sqo_cell_from_bounds(const double *lo, const double *hi, const double *tilt,
                          double *a, double *origin) {
  double xy = tilt[0], xz = tilt[1], yz = tilt[2];
  double mn = 0.0, mx = 0.0;
  double c[4] = {0.0, xy, xz, xy + xz};
  for (int k = 0; k < 4; ++k) {
    if (c[k] < mn) mn = c[k];
    if (c[k] > mx) mx = c[k];
  }
  double xlo = lo[0] - mn, xhi = hi[0] - mx;
  double ylo = lo[1] - (yz < 0.0 ? yz : 0.0), yhi = hi[1] - (yz > 0.0 ? yz : 0.0);
  double zlo = lo[2], zhi = hi[2];
  origin[0] = xlo; origin[1] = ylo; origin[2] = zlo;
  memset(a, 0, 9 * sizeof(double));
  M(a, 0, 0) = xhi - xlo;
  M(a, 0, 1) = xy; M(a, 1, 1) = yhi - ylo;
  M(a, 0, 2) = xz; M(a, 1, 2) = yz; M(a, 2, 2) = zhi - zlo;
}

/* Wrapped fractional coordinates s - floor(s) of one Cartesian point,
 * src/sqc_cell.f90:89-98 and src/sqc_structure_factor.f90:1415-1419. */
static void frac_wrap(const double *inv, const double *origin, const double *r, double *s) {
  double d0 = r[0] - origin[0], d1 = r[1] - origin[1], d2 = r[2] - origin[2];
  for (int i = 0; i < 3; ++i) {
    double v = M(inv, i, 0) * d0 + M(inv, i, 1) * d1 + M(inv, i, 2) * d2;
    s[i] = v - floor(v);
  }
}

/* ---------------------------------------------------------------------- */
/* Mode set, src/sqc_structure_factor.f90:327-399                           */
/* ---------------------------------------------------------------------- */

/* Grid extent per axis: m_i = max(3, 2 ceil(qmax |a_i| / 2 pi) + 1). */
void sqo_mode_extent(const double *a, double qmax, int64_t *m) {
  const double two_pi = 2.0 * acos(-1.0);
  for (int i = 0; i < 3; ++i) {
    double len = sqrt(M(a, 0, i) * M(a, 0, i) + M(a, 1, i) * M(a, 1, i) + M(a, 2, i) * M(a, 2, i));
    int64_t mi = 2 * (int64_t)ceil(qmax * len / two_pi) + 1;
    m[i] = mi < 3 ? 3 : mi;
  }
}

/* |q| of h, evaluated as q = (b1 h1 + b2 h2) + b3 h3, |q| = sqrt(q.q). */
static double qlen_of(const double *b, int h1, int h2, int h3, double *q) {
  double r1 = (double)h1, r2 = (double)h2, r3 = (double)h3;
  for (int i = 0; i < 3; ++i) q[i] = M(b, i, 0) * r1 + M(b, i, 1) * r2 + M(b, i, 2) * r3;
  return sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
}

/* Count the kept modes (first enumeration pass). */
int64_t sqo_mode_count(const double *a, double qmin, double qmax) {
  double b[9], q[3];
  int64_t m[3], n = 0;
  sqo_cell_setup(a, b);
  sqo_mode_extent(a, qmax, m);
  for (int64_t p3 = 1; p3 <= m[2]; ++p3) {
    int h3 = (int)(p3 - (m[2] + 1) / 2);
    for (int64_t p2 = 1; p2 <= m[1]; ++p2) {
      int h2 = (int)(p2 - (m[1] + 1) / 2);
      for (int64_t p1 = 1; p1 <= m[0]; ++p1) {
        int h1 = (int)(p1 - (m[0] + 1) / 2);
        if (h1 == 0 && h2 == 0 && h3 == 0) continue;
        double ql = qlen_of(b, h1, h2, h3, q);
        if (ql > qmax || ql < qmin) continue;
        ++n;
      }
    }
  }
  return n;
}

/* Fill the kept modes (second pass).  hkl[3K], qvec[3K], qlen[K], gidx[K]
 * (1-based flat index, x fastest), shell[K] (1-based, clamped to 1..nq). */
int64_t sqo_modes(const double *a, double qmin, double qmax, int nq, int32_t *hkl,
                  double *qvec, double *qlen, int64_t *gidx, int32_t *shell) {
  double b[9], q[3];
  int64_t m[3], n = 0;
  sqo_cell_setup(a, b);
  sqo_mode_extent(a, qmax, m);
  double dq = (qmax - qmin) / (double)nq;
  for (int64_t p3 = 1; p3 <= m[2]; ++p3) {
    int h3 = (int)(p3 - (m[2] + 1) / 2);
    for (int64_t p2 = 1; p2 <= m[1]; ++p2) {
      int h2 = (int)(p2 - (m[1] + 1) / 2);
      for (int64_t p1 = 1; p1 <= m[0]; ++p1) {
        int h1 = (int)(p1 - (m[0] + 1) / 2);
        if (h1 == 0 && h2 == 0 && h3 == 0) continue;
        double ql = qlen_of(b, h1, h2, h3, q);
        if (ql > qmax || ql < qmin) continue;
        hkl[3 * n] = h1; hkl[3 * n + 1] = h2; hkl[3 * n + 2] = h3;
        qvec[3 * n] = q[0]; qvec[3 * n + 1] = q[1]; qvec[3 * n + 2] = q[2];
        qlen[n] = ql;
        gidx[n] = p1 + m[0] * (p2 - 1 + m[1] * (p3 - 1));
        int s = (int)((ql - qmin) / dq) + 1;
        if (s < 1) s = 1;
        if (s > nq) s = nq;
        shell[n] = s;
        ++n;
      }
    }
  }
  return n;
}

/* ---------------------------------------------------------------------- */
/* Weights, src/sqc_weights.f90:136-154, src/sqc_elements.f90:224-239      */
/* kind 0 unit, 1 neutron (const per species), 2 xray (IT92 coefficients,   */
/* nine per species: a1..a4, b1..b4, c).                                    */
/* ---------------------------------------------------------------------- */

double sqo_amplitude(int kind, const double *amp_const, const double *xcoef, int sp, double q) {
  if (kind == 1) return amp_const[sp];
  if (kind == 2) {
    const double *c = xcoef + 9 * sp;
    double stol2 = q / (4.0 * acos(-1.0));
    stol2 = stol2 * stol2;
    double f = c[8];
    for (int i = 0; i < 4; ++i) f = f + c[i] * exp(-c[4 + i] * stol2);
    return f;
  }
  return 1.0;
}

/* Per-mode denominators summed per shell, src/sqc_structure_factor.f90:389-451.
 * norm: 1 mean, 2 self, 3 natom.  counts[S] atoms per species. */
void sqo_denominators(int64_t nmodes, const double *qlen, const int32_t *shell, int nq,
                      int norm, int kind, int nspecies, const int64_t *counts,
                      const double *amp_const, const double *xcoef, int64_t natoms,
                      double *den, int64_t *shell_count) {
  for (int s = 0; s < nq; ++s) { den[s] = 0.0; shell_count[s] = 0; }
  for (int64_t i = 0; i < nmodes; ++i) {
    double d;
    if (norm == 3) {
      d = (double)natoms;
    } else {
      double sw = 0.0, sw2 = 0.0;
      for (int t = 0; t < nspecies; ++t) {
        if (counts[t] <= 0) continue;
        double w = sqo_amplitude(kind, amp_const, xcoef, t, qlen[i]);
        sw = sw + (double)counts[t] * w;
        sw2 = sw2 + (double)counts[t] * w * w;
      }
      d = (norm == 1) ? sw * sw / (double)natoms : sw2;
    }
    den[shell[i] - 1] += d;
    shell_count[shell[i] - 1] += 1;
  }
}

/* ---------------------------------------------------------------------- */
/* Direct sum, src/sqc_structure_factor.f90:1398-1515                       */
/* ---------------------------------------------------------------------- */

/* Species amplitudes rho_a(q_m) = sum_{j in a} exp(i q_m . r_j) for every kept
 * mode, r_j = A (s_j - floor s_j).  method 0 evaluates cos/sin of the
 * Cartesian phase q . r (the reference's expression); method 1 uses the
 * separable tables exp(2 pi i h_d s_d), each entry from its own sin/cos
 * (src/sqc_phase.f90:6-30 is the reference's use of the same factorisation).
 * rho is [S][K] complex, interleaved (re, im). */
int sqo_direct_rho(int64_t natoms, const double *pos, const double *a, const double *origin,
                   const int32_t *species_of, int nspecies, int64_t nmodes,
                   const int32_t *hkl, const double *qvec, int method, double *rho) {
  double inv[9];
  sqo_inverse3(a, inv);
  double *s = (double *)malloc(sizeof(double) * 3 * (size_t)natoms);
  double *cart = (double *)malloc(sizeof(double) * 3 * (size_t)natoms);
  if (!s || !cart) { free(s); free(cart); return 1; }
  for (int64_t j = 0; j < natoms; ++j) {
    frac_wrap(inv, origin, pos + 3 * j, s + 3 * j);
    for (int i = 0; i < 3; ++i)
      cart[3 * j + i] = M(a, i, 0) * s[3 * j] + M(a, i, 1) * s[3 * j + 1] + M(a, i, 2) * s[3 * j + 2];
  }
  memset(rho, 0, sizeof(double) * 2 * (size_t)nspecies * (size_t)nmodes);
  if (method == 0) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t m = 0; m < nmodes; ++m) {
      double qx = qvec[3 * m], qy = qvec[3 * m + 1], qz = qvec[3 * m + 2];
      for (int64_t j = 0; j < natoms; ++j) {
        double ph = qx * cart[3 * j] + qy * cart[3 * j + 1] + qz * cart[3 * j + 2];
        int sp = species_of[j];
        rho[2 * ((size_t)sp * nmodes + m)] += cos(ph);
        rho[2 * ((size_t)sp * nmodes + m) + 1] += sin(ph);
      }
    }
  } else {
    /* per-atom tables over the h range present in hkl */
    int hmax[3] = {0, 0, 0};
    for (int64_t m = 0; m < nmodes; ++m)
      for (int d = 0; d < 3; ++d) {
        int h = abs(hkl[3 * m + d]);
        if (h > hmax[d]) hmax[d] = h;
      }
    int w[3] = {2 * hmax[0] + 1, 2 * hmax[1] + 1, 2 * hmax[2] + 1};
    size_t per = (size_t)(w[0] + w[1] + w[2]);
    double *tab = (double *)malloc(sizeof(double) * 2 * per * (size_t)natoms);
    if (!tab) { free(s); free(cart); return 1; }
    const double two_pi = 2.0 * SQO_PI;
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < natoms; ++j) {
      double *t = tab + 2 * per * (size_t)j;
      int off = 0;
      for (int d = 0; d < 3; ++d) {
        for (int h = -hmax[d]; h <= hmax[d]; ++h) {
          /* reduce h*s mod 1 exactly enough before the trig call */
          double x = (double)h * s[3 * j + d];
          x = x - floor(x);
          t[2 * (off + h + hmax[d])] = cos(two_pi * x);
          t[2 * (off + h + hmax[d]) + 1] = sin(two_pi * x);
        }
        off += w[d];
      }
    }
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t m = 0; m < nmodes; ++m) {
      int i1 = hkl[3 * m] + hmax[0];
      int i2 = w[0] + hkl[3 * m + 1] + hmax[1];
      int i3 = w[0] + w[1] + hkl[3 * m + 2] + hmax[2];
      for (int64_t j = 0; j < natoms; ++j) {
        const double *t = tab + 2 * per * (size_t)j;
        double ar = t[2 * i1], ai = t[2 * i1 + 1];
        double br = t[2 * i2], bi = t[2 * i2 + 1];
        double cr = t[2 * i3], ci = t[2 * i3 + 1];
        double er = ar * br - ai * bi, ei = ar * bi + ai * br;
        double fr = er * cr - ei * ci, fi = er * ci + ei * cr;
        int sp = species_of[j];
        rho[2 * ((size_t)sp * nmodes + m)] += fr;
        rho[2 * ((size_t)sp * nmodes + m) + 1] += fi;
      }
    }
    free(tab);
  }
  free(s);
  free(cart);
  return 0;
}

/* One frame's shell sums from species amplitudes:
 *   num[s]        += sum_{m in s} |sum_a w_a(|q_m|) rho_a(q_m)|^2
 *   partial[a,b,s] += sum_{m in s} Re(rho_a rho_b*)      (a <= b, [S][S][nq])
 * in mode order, as src/sqc_structure_factor.f90:481-521 and :595-643. */
void sqo_shell_sums(int nspecies, int64_t nmodes, const double *rho, const double *qlen,
                    const int32_t *shell, int nq, int kind, const double *amp_const,
                    const double *xcoef, double *num, double *partial) {
  for (int64_t m = 0; m < nmodes; ++m) {
    double tr = 0.0, ti = 0.0;
    for (int a = 0; a < nspecies; ++a) {
      double w = sqo_amplitude(kind, amp_const, xcoef, a, qlen[m]);
      tr += w * rho[2 * ((size_t)a * nmodes + m)];
      ti += w * rho[2 * ((size_t)a * nmodes + m) + 1];
    }
    int s = shell[m] - 1;
    num[s] += tr * tr + ti * ti;
    if (partial) {
      for (int a = 0; a < nspecies; ++a) {
        double ar = rho[2 * ((size_t)a * nmodes + m)], ai = rho[2 * ((size_t)a * nmodes + m) + 1];
        for (int b = a; b < nspecies; ++b) {
          double br = rho[2 * ((size_t)b * nmodes + m)], bi = rho[2 * ((size_t)b * nmodes + m) + 1];
          partial[((size_t)a * nspecies + b) * nq + s] += ar * br + ai * bi;
        }
      }
    }
  }
}

/* ---------------------------------------------------------------------- */
/* Analytic FP64 gradient of a loss on one frame's shell values.            */
/*                                                                          */
/* For S_s = num_s / den_s and P_ab,s = partial_ab,s / (N count_s), with    */
/* upstream gradients gS[s] = dL/dS_s and gP[a][b][s] = dL/dP_ab,s (a<=b):  */
/*   G_m   = gS[s(m)] / den[s(m)]                                            */
/*   H_abm = gP[a][b][s(m)] / (N count[s(m)])                                */
/*   C_cm  = 2 G_m w_c(|q_m|) conj(T_m)                                      */
/*         + sum_{a<=b} H_abm (delta_ca conj(rho_b) + delta_cb conj(rho_a))   */
/*   dL/dr_j = Re( sum_m i q_m C_{c(j) m} exp(i q_m . r_j) )                 */
/* which follows from d rho_c / d r_j = i q exp(i q . r_j) for j in c.       */
/* grad is (3, N).  gP may be NULL.                                          */
/* ---------------------------------------------------------------------- */
int sqo_direct_grad(int64_t natoms, const double *pos, const double *a, const double *origin,
                    const int32_t *species_of, int nspecies, int64_t nmodes,
                    const double *qvec, const double *qlen, const int32_t *shell,
                    const double *rho, int kind, const double *amp_const, const double *xcoef,
                    const double *gS, const double *den, const double *gP,
                    const int64_t *shell_count, int nq, double *grad) {
  double inv[9];
  sqo_inverse3(a, inv);
  double *coef = (double *)malloc(sizeof(double) * 2 * (size_t)nspecies * (size_t)nmodes);
  double *cart = (double *)malloc(sizeof(double) * 3 * (size_t)natoms);
  if (!coef || !cart) { free(coef); free(cart); return 1; }
  for (int64_t j = 0; j < natoms; ++j) {
    double s[3];
    frac_wrap(inv, origin, pos + 3 * j, s);
    for (int i = 0; i < 3; ++i)
      cart[3 * j + i] = M(a, i, 0) * s[0] + M(a, i, 1) * s[1] + M(a, i, 2) * s[2];
  }
  for (int64_t m = 0; m < nmodes; ++m) {
    int s = shell[m] - 1;
    double tr = 0.0, ti = 0.0;
    for (int c = 0; c < nspecies; ++c) {
      double w = sqo_amplitude(kind, amp_const, xcoef, c, qlen[m]);
      tr += w * rho[2 * ((size_t)c * nmodes + m)];
      ti += w * rho[2 * ((size_t)c * nmodes + m) + 1];
    }
    double G = (den[s] > 0.0) ? gS[s] / den[s] : 0.0;
    for (int c = 0; c < nspecies; ++c) {
      double w = sqo_amplitude(kind, amp_const, xcoef, c, qlen[m]);
      double cr = 2.0 * G * w * tr, ci = -2.0 * G * w * ti;
      if (gP) {
        double nc = (double)natoms * (double)(shell_count[s] > 0 ? shell_count[s] : 1);
        for (int aa = 0; aa < nspecies; ++aa)
          for (int bb = aa; bb < nspecies; ++bb) {
            double H = gP[((size_t)aa * nspecies + bb) * nq + s] / nc;
            if (H == 0.0) continue;
            if (c == aa) {
              cr += H * rho[2 * ((size_t)bb * nmodes + m)];
              ci -= H * rho[2 * ((size_t)bb * nmodes + m) + 1];
            }
            if (c == bb) {
              cr += H * rho[2 * ((size_t)aa * nmodes + m)];
              ci -= H * rho[2 * ((size_t)aa * nmodes + m) + 1];
            }
          }
      }
      coef[2 * ((size_t)c * nmodes + m)] = cr;
      coef[2 * ((size_t)c * nmodes + m) + 1] = ci;
    }
  }
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t j = 0; j < natoms; ++j) {
    int c = species_of[j];
    double g0 = 0.0, g1 = 0.0, g2 = 0.0;
    for (int64_t m = 0; m < nmodes; ++m) {
      double qx = qvec[3 * m], qy = qvec[3 * m + 1], qz = qvec[3 * m + 2];
      double ph = qx * cart[3 * j] + qy * cart[3 * j + 1] + qz * cart[3 * j + 2];
      double er = cos(ph), ei = sin(ph);
      double cr = coef[2 * ((size_t)c * nmodes + m)], ci = coef[2 * ((size_t)c * nmodes + m) + 1];
      /* Re(i z) = -Im(z), z = C e^{i ph} */
      double im = cr * ei + ci * er;
      g0 -= qx * im;
      g1 -= qy * im;
      g2 -= qz * im;
    }
    grad[3 * j] = g0;
    grad[3 * j + 1] = g1;
    grad[3 * j + 2] = g2;
  }
  free(coef);
  free(cart);
  return 0;
}

int sqo_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ---------------------------------------------------------------------- */
/* dyn: density amplitudes on a list of Miller indices                      */
/* ---------------------------------------------------------------------- */

typedef struct {
  double re, im;
} sqo_cplx;

static sqo_cplx cmul_(sqo_cplx a, sqo_cplx b) {
  sqo_cplx r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
  return r;
}

/* One atom's three factor tables, t[j][n + nmax[j]] = base_j exp(i n phi_j)
 * for n = -nmax_j .. nmax_j (phase_tables, src/sqc_phase.f90:67-99): the
 * positive half by the recurrence prod *= u, the negative half by its own
 * recurrence with conj(u) when the axis has a base, else as the conjugate of
 * the positive half.  width = 2 max(nmax) + 1 entries per axis. */
static void dyn_tables(sqo_cplx *t, int width, const double *phi, const int *nmax,
                       const sqo_cplx *base, int base_on) {
  for (int j = 0; j < 3; ++j) {
    sqo_cplx *tj = t + (size_t)j * width;
    const int off = nmax[j] > 0 ? nmax[j] : 0;
    const sqo_cplx u = {cos(phi[j]), sin(phi[j])};
    const sqo_cplx c0 = base_on ? base[j] : (sqo_cplx){1.0, 0.0};
    sqo_cplx prod = c0;
    tj[off] = c0;
    for (int n = 1; n <= off; ++n) {
      prod = cmul_(prod, u);
      tj[off + n] = prod;
    }
    if (base_on) {
      const sqo_cplx uc = {u.re, -u.im};
      prod = c0;
      for (int n = 1; n <= off; ++n) {
        prod = cmul_(prod, uc);
        tj[off - n] = prod;
      }
    } else {
      for (int n = 1; n <= off; ++n) {
        tj[off - n].re = tj[off + n].re;
        tj[off - n].im = -tj[off + n].im;
      }
    }
  }
}

/* rho[a][m] (re, im) = sum over the atoms of species a of exp(i q_m . r_j),
 * q_m = hkl_m1 b1 + hkl_m2 b2 + hkl_m3 b3 (b in the columns of recip), plus
 * base . r when base_on (a q line's start).  Restates the rho block of
 * dyn_accumulate (src/sqc_dynamics.f90:826-873): method 1 the separable
 * branch (per atom: phi_k = recip(1,k) r1 + recip(2,k) r2 + recip(3,k) r3,
 * dyn_atom_phase :1094-1115, the tables above, and rho += t1 t2 t3 in that
 * order, phase_factor src/sqc_phase.f90:103-113); method 0 the direct
 * branch (:853-873, one cos/sin of q . r per mode and atom, with
 * q = n1 b1 + n2 b2 + n3 b3 and the base added).  Atoms are split over the
 * threads in order and the per-thread sums added in thread order, as the
 * reference reduces rho_thread. */
int sqo_dyn_rho(int64_t natoms, const double *pos, const int32_t *species_of, int nspecies,
                const double *recip, const double *base_vec, int base_on, int64_t nmodes,
                const int32_t *hkl, int method, double *rho) {
  int nmax[3] = {0, 0, 0};
  for (int64_t m = 0; m < nmodes; ++m)
    for (int j = 0; j < 3; ++j) {
      const int v = hkl[3 * m + j] < 0 ? -hkl[3 * m + j] : hkl[3 * m + j];
      if (v > nmax[j]) nmax[j] = v;
    }
  int mx = nmax[0];
  if (nmax[1] > mx) mx = nmax[1];
  if (nmax[2] > mx) mx = nmax[2];
  const int width = 2 * mx + 1;
  const size_t nr = (size_t)nspecies * nmodes;
  int nthr = 1;
#ifdef _OPENMP
  nthr = omp_get_max_threads();
#endif
  double *acc = (double *)calloc((size_t)nthr * nr * 2, sizeof(double));
  if (!acc) return 1;
  int fail = 0;
#pragma omp parallel num_threads(nthr)
  {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    double *mine = acc + (size_t)tid * nr * 2;
    sqo_cplx *t = (sqo_cplx *)malloc(sizeof(sqo_cplx) * 3 * width);
    if (!t) {
#pragma omp atomic write
      fail = 1;
    }
#pragma omp for schedule(static)
    for (int64_t i = 0; i < natoms; ++i) {
      if (!t) continue;
      const double *r = pos + 3 * i;
      const int a = species_of[i];
      double phi[3];
      for (int k = 0; k < 3; ++k)
        phi[k] = recip[0 + 3 * k] * r[0] + recip[1 + 3 * k] * r[1] + recip[2 + 3 * k] * r[2];
      double d = 0.0;
      if (base_on) d = base_vec[0] * r[0] + base_vec[1] * r[1] + base_vec[2] * r[2];
      if (method == 1) {
        sqo_cplx base[3] = {{cos(d), sin(d)}, {1.0, 0.0}, {1.0, 0.0}};
        dyn_tables(t, width, phi, nmax, base, base_on);
        for (int64_t m = 0; m < nmodes; ++m) {
          const int32_t *n = hkl + 3 * m;
          const sqo_cplx v = cmul_(cmul_(t[n[0] + nmax[0]], t[width + n[1] + nmax[1]]),
                                   t[2 * width + n[2] + nmax[2]]);
          double *o = mine + 2 * ((size_t)a * nmodes + m);
          o[0] += v.re;
          o[1] += v.im;
        }
      } else {
        for (int64_t m = 0; m < nmodes; ++m) {
          const int32_t *n = hkl + 3 * m;
          double q[3];
          for (int c = 0; c < 3; ++c)
            q[c] = recip[c] * n[0] + recip[c + 3] * n[1] + recip[c + 6] * n[2];
          double ph = q[0] * r[0] + q[1] * r[1] + q[2] * r[2];
          if (base_on) ph += d;
          double *o = mine + 2 * ((size_t)a * nmodes + m);
          o[0] += cos(ph);
          o[1] += sin(ph);
        }
      }
    }
    free(t);
  }
  if (!fail) {
    for (size_t k = 0; k < 2 * nr; ++k) {
      double s = 0.0;
      for (int th = 0; th < nthr; ++th) s += acc[(size_t)th * nr * 2 + k];
      rho[k] = s;
    }
  }
  free(acc);
  return fail;
}

/* ---------------------------------------------------------------------- */
/* Debye: ordered pair-distance histograms per type pair                    */
/* ---------------------------------------------------------------------- */

/* Perpendicular widths volume / |a_j x a_k|, src/sqc_cell.f90:100-117. */
void sqo_cell_widths(const double *a, double *w) {
  const double vol = fabs(det3(a));
  for (int i = 0; i < 3; ++i) {
    const int j = (i + 1) % 3, k = (j + 1) % 3;
    const double *u = a + 3 * j, *v = a + 3 * k;
    const double n0 = u[1] * v[2] - u[2] * v[1];
    const double n1 = u[2] * v[0] - u[0] * v[2];
    const double n2 = u[0] * v[1] - u[1] * v[0];
    w[i] = vol / sqrt(n0 * n0 + n1 * n1 + n2 * n2);
  }
}

/* The cell list's decomposition, nl_configure (src/sqc_neighbour_list.f90:
 * 109-123): div_i = max(1, int(width_i / (rmax + max(skin, 0)))) and
 * bins(:, i) = a_i / div_i (column-major). */
void sqo_debye_cells(const double *a, double rmax, double skin, int32_t *div, double *bins) {
  double w[3];
  sqo_cell_widths(a, w);
  const double reach = rmax + (skin > 0.0 ? skin : 0.0);
  for (int i = 0; i < 3; ++i) {
    const int d = (int)(w[i] / reach);
    div[i] = d > 1 ? d : 1;
    for (int c = 0; c < 3; ++c) M(bins, c, i) = M(a, c, i) / (double)div[i];
  }
}

/* The list's positions of one frame, nl_begin_frame (src/sqc_neighbour_list.f90
 * :211-229).  s = wrap(A^-1 (r - o)), r_new = A s; the first frame takes
 * pos = posw = r_new, later frames pos += A (ds - anint(ds)) with
 * ds = A^-1 (r_new - posw) and posw = r_new. */
void sqo_debye_positions(int64_t natoms, const double *a, const double *origin, const double *r,
                         int first, double *pos, double *posw) {
  double inv[9];
  sqo_inverse3(a, inv);
  for (int64_t i = 0; i < natoms; ++i) {
    double s[3], rn[3];
    frac_wrap(inv, origin, r + 3 * i, s);
    for (int c = 0; c < 3; ++c) rn[c] = M(a, c, 0) * s[0] + M(a, c, 1) * s[1] + M(a, c, 2) * s[2];
    double *p = pos + 3 * i, *pw = posw + 3 * i;
    if (first) {
      for (int c = 0; c < 3; ++c) p[c] = pw[c] = rn[c];
      continue;
    }
    double dr[3], ds[3];
    for (int c = 0; c < 3; ++c) dr[c] = rn[c] - pw[c];
    for (int c = 0; c < 3; ++c) ds[c] = M(inv, c, 0) * dr[0] + M(inv, c, 1) * dr[1] + M(inv, c, 2) * dr[2];
    for (int c = 0; c < 3; ++c) ds[c] = ds[c] - round(ds[c]);
    for (int c = 0; c < 3; ++c) dr[c] = M(a, c, 0) * ds[0] + M(a, c, 1) * ds[1] + M(a, c, 2) * ds[2];
    for (int c = 0; c < 3; ++c) {
      p[c] = p[c] + dr[c];
      pw[c] = rn[c];
    }
  }
}

/* Ordered pair counts of one frame of list positions, the pair loop of
 * debye_accumulate (src/sqc_debye.f90:208-226) over every (i, j, image m)
 * the cell list holds: d = (pos_j - pos_i) + bins . (div * m), counted when
 * d . d <= rmax^2 in bin int(sqrt(d . d) / dr) + 1 <= nbins, i == j only
 * with m != 0.  The list keeps exactly the images within reach of each
 * pair (nl_build :263-330), and |d| <= rmax bounds the fractional
 * displacement, |s_k + m_k| <= rmax / width_k, so the images are taken from
 * that bound: method 1 per pair from s = A^-1 pos_j - A^-1 pos_i, method 0
 * from a box of +-(ceil(rmax / width_k) + 1) around the minimum image (brute
 * force).  hist[a][b][k] (0-based, ntypes x ntypes x nbins) gains the counts;
 * type_id is 1-based. */
int sqo_debye_hist_rows(int64_t natoms, int64_t i0, int64_t i1, int ntypes,
                        const int32_t *type_id, const double *a, const int32_t *pbc, double rmax,
                        double dr, double skin, const double *pos, int method, uint64_t *hist) {
  double inv[9], w[3], bins[9];
  int32_t div[3];
  sqo_inverse3(a, inv);
  sqo_cell_widths(a, w);
  sqo_debye_cells(a, rmax, skin, div, bins);
  const double inv_dr = 1.0 / dr;
  const double rr = rmax * rmax;
  const int nbins = (int)ceil(rmax / dr) > 1 ? (int)ceil(rmax / dr) : 1;
  const size_t nh = (size_t)ntypes * ntypes * nbins;
  double *frac = (double *)malloc(sizeof(double) * 3 * (size_t)natoms);
  const int nthr = sqo_num_threads();
  uint64_t *acc = (uint64_t *)calloc(nh * (size_t)nthr, sizeof(uint64_t));
  if (!frac || !acc) {
    free(frac);
    free(acc);
    return 1;
  }
  for (int64_t i = 0; i < natoms; ++i)
    for (int c = 0; c < 3; ++c)
      frac[3 * i + c] = M(inv, c, 0) * pos[3 * i] + M(inv, c, 1) * pos[3 * i + 1] +
                        M(inv, c, 2) * pos[3 * i + 2];
  double reach[3];
  int wide[3];
  for (int k = 0; k < 3; ++k) {
    reach[k] = rmax / w[k] + 1e-9;
    wide[k] = (int)ceil(rmax / w[k]) + 1;
  }
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t i = i0; i < i1; ++i) {
#ifdef _OPENMP
    uint64_t *mine = acc + nh * (size_t)omp_get_thread_num();
#else
    uint64_t *mine = acc;
#endif
    const int ta = type_id[i] - 1;
    for (int64_t j = 0; j < natoms; ++j) {
      int lo[3], hi[3];
      for (int k = 0; k < 3; ++k) {
        const double s = frac[3 * j + k] - frac[3 * i + k];
        if (!pbc[k]) {
          lo[k] = hi[k] = 0;
        } else if (method == 1) {
          lo[k] = (int)ceil(-s - reach[k]);
          hi[k] = (int)floor(-s + reach[k]);
        } else {
          const int c0 = -(int)round(s);
          lo[k] = c0 - wide[k];
          hi[k] = c0 + wide[k];
        }
      }
      const double e0 = pos[3 * j] - pos[3 * i], e1 = pos[3 * j + 1] - pos[3 * i + 1],
                   e2 = pos[3 * j + 2] - pos[3 * i + 2];
      const int tb = type_id[j] - 1;
      for (int m2 = lo[2]; m2 <= hi[2]; ++m2)
        for (int m1 = lo[1]; m1 <= hi[1]; ++m1)
          for (int m0 = lo[0]; m0 <= hi[0]; ++m0) {
            if (i == j && m0 == 0 && m1 == 0 && m2 == 0) continue;
            const double c0 = (double)(div[0] * m0), c1 = (double)(div[1] * m1),
                         c2 = (double)(div[2] * m2);
            const double v0 = M(bins, 0, 0) * c0 + M(bins, 0, 1) * c1 + M(bins, 0, 2) * c2;
            const double v1 = M(bins, 1, 0) * c0 + M(bins, 1, 1) * c1 + M(bins, 1, 2) * c2;
            const double v2 = M(bins, 2, 0) * c0 + M(bins, 2, 1) * c1 + M(bins, 2, 2) * c2;
            const double d0 = e0 + v0, d1 = e1 + v1, d2 = e2 + v2;
            const double dd = d0 * d0 + d1 * d1 + d2 * d2;
            if (dd > rr) continue;
            const int bin = (int)(sqrt(dd) * inv_dr) + 1;
            if (bin < 1 || bin > nbins) continue;
            mine[((size_t)ta * ntypes + tb) * nbins + (bin - 1)] += 1;
          }
    }
  }
  for (size_t k = 0; k < nh; ++k) {
    uint64_t s = 0;
    for (int t = 0; t < nthr; ++t) s += acc[nh * (size_t)t + k];
    hist[k] += s;
  }
  free(acc);
  free(frac);
  return 0;
}

/* All first atoms i of the frame (sqo_debye_hist_rows takes the pairs whose
 * first atom is in [i0, i1), which the CPU timing samples). */
int sqo_debye_hist(int64_t natoms, int ntypes, const int32_t *type_id, const double *a,
                   const int32_t *pbc, double rmax, double dr, double skin, const double *pos,
                   int method, uint64_t *hist) {
  return sqo_debye_hist_rows(natoms, 0, natoms, ntypes, type_id, a, pbc, rmax, dr, skin, pos,
                             method, hist);
}
