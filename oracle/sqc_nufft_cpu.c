/*
 * sqc_nufft_cpu.c -- CPU restatement of the reference's FINUFFT type-1 path.
 *
 * TEST INFRASTRUCTURE ONLY: this is the CPU baseline that bench.py times next
 * to the GPU path (cpu_baseline.kind = "port"), and a second, independent
 * NUFFT that the tests check against the direct sum.  The product never links
 * it.
 *
 * It follows the reference CPU flow of one frame
 * (src/sqc_structure_factor.f90:1165-1319):
 *   wrap to the cell, s - floor(s)                 (:1182-1188)
 *   sort points into 16x4x4 bins, split the sorted list into sub-problems,
 *   spread each sub-problem into a private padded box and add the box into
 *   the periodic fine grid          (external/finufft/src/spreadinterp.cpp:1517-1720,
 *                                     :1073-1151 kernel loop, :1199-1260 wrapped add)
 *   3D FFT, sign +1                  (external/finufft/src/fft.cpp:22-121; done by the
 *                                     caller, see oracle/oracle.py)
 *   deconvolve by phihat_1 phihat_2 phihat_3 and pick the kept modes
 *                                    (external/finufft/src/finufft_core.cpp:372-414)
 *   serial species-pair accumulation and shell sums
 *                                    (src/sqc_structure_factor.f90:481-521, :595-643)
 *
 * The spreading kernel is the exponential of semicircle
 * phi(z) = exp(beta (sqrt(1 - z^2) - 1)), z = 2 u / ns, the kernel of the
 * reference GPU library (external/finufft/include/cufinufft/spreadinterp.h:64-85),
 * evaluated directly (no Horner table).  Grid coordinates are u = s n, i.e.
 * FINUFFT's fold_rescale of x = 2 pi s - pi; the resulting (-1)^(h1+h2+h3)
 * phase of the reference convention is dropped, as it cancels in every
 * product this path forms.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define M(a, i, j) ((a)[(i) + 3 * (j)])

void sqo_inverse3(const double *m, double *inv);

static double es_kernel(double u, double ns, double beta) {
  double z = 2.0 * u / ns;
  if (fabs(z) >= 1.0) return 0.0;
  return exp(beta * (sqrt(1.0 - z * z) - 1.0));
}

typedef struct {
  int64_t idx;
  int64_t bin;
} sort_item;

static int cmp_item(const void *pa, const void *pb) {
  const sort_item *a = (const sort_item *)pa, *b = (const sort_item *)pb;
  if (a->bin != b->bin) return a->bin < b->bin ? -1 : 1;
  return a->idx < b->idx ? -1 : (a->idx > b->idx);
}

/* Spread every atom with field_of[j] >= 0 into grid[field] (complex128,
 * interleaved, layout [field][z][y][x], x fastest; the caller zeroes it).
 * weight[j] is the real strength of atom j. */
int sqo_nufft_spread(int64_t natoms, const double *pos, const double *a, const double *origin,
                     const int32_t *field_of, const double *weight, int nfields,
                     const int64_t *n, int ns, double beta, int64_t max_subprob,
                     double *grid) {
  double inv[9];
  sqo_inverse3(a, inv);
  double *u = (double *)malloc(sizeof(double) * 3 * (size_t)natoms);
  sort_item *items = (sort_item *)malloc(sizeof(sort_item) * (size_t)natoms);
  if (!u || !items) { free(u); free(items); return 1; }
  const int64_t bx = 16, by = 4, bz = 4;
  int64_t nbx = (n[0] + bx - 1) / bx, nby = (n[1] + by - 1) / by;
  for (int64_t j = 0; j < natoms; ++j) {
    const double *r = pos + 3 * j;
    double d0 = r[0] - origin[0], d1 = r[1] - origin[1], d2 = r[2] - origin[2];
    for (int i = 0; i < 3; ++i) {
      double v = M(inv, i, 0) * d0 + M(inv, i, 1) * d1 + M(inv, i, 2) * d2;
      v = v - floor(v);
      double g = v * (double)n[i];
      if (g >= (double)n[i]) g -= (double)n[i];
      u[3 * j + i] = g;
    }
    int64_t ib0 = (int64_t)(u[3 * j] / bx), ib1 = (int64_t)(u[3 * j + 1] / by),
            ib2 = (int64_t)(u[3 * j + 2] / bz);
    int f = field_of[j] < 0 ? nfields : field_of[j];
    items[j].idx = j;
    items[j].bin = (((int64_t)f * ((n[2] + bz - 1) / bz) + ib2) * nby + ib1) * nbx + ib0;
  }
  qsort(items, (size_t)natoms, sizeof(sort_item), cmp_item);
  /* sub-problems: runs of at most max_subprob points of one field */
  int64_t nsub = 0;
  int64_t *sub_start = (int64_t *)malloc(sizeof(int64_t) * ((size_t)natoms + 1));
  if (!sub_start) { free(u); free(items); return 1; }
  for (int64_t k = 0; k < natoms;) {
    int f = field_of[items[k].idx];
    if (f < 0) break;
    int64_t e = k;
    while (e < natoms && e - k < max_subprob && field_of[items[e].idx] == f) ++e;
    sub_start[nsub++] = k;
    k = e;
  }
  int64_t last = 0;
  while (last < natoms && field_of[items[last].idx] >= 0) ++last;
  sub_start[nsub] = last;
  const int64_t plane = n[0] * n[1];
  const int64_t vol = plane * n[2];
  const double half = 0.5 * (double)ns;
#pragma omp parallel
  {
    double *ker = (double *)malloc(sizeof(double) * 3 * (size_t)ns);
#pragma omp for schedule(dynamic, 1)
    for (int64_t sp = 0; sp < nsub; ++sp) {
      int64_t k0 = sub_start[sp], k1 = sub_start[sp + 1];
      int f = field_of[items[k0].idx];
      /* bounding box of the stencils, in unwrapped grid indices */
      int64_t lo[3], hi[3];
      for (int d = 0; d < 3; ++d) { lo[d] = INT64_MAX; hi[d] = INT64_MIN; }
      for (int64_t k = k0; k < k1; ++k) {
        int64_t j = items[k].idx;
        for (int d = 0; d < 3; ++d) {
          int64_t i0 = (int64_t)ceil(u[3 * j + d] - half);
          if (i0 < lo[d]) lo[d] = i0;
          if (i0 + ns - 1 > hi[d]) hi[d] = i0 + ns - 1;
        }
      }
      int64_t w0 = hi[0] - lo[0] + 1, w1 = hi[1] - lo[1] + 1, w2 = hi[2] - lo[2] + 1;
      double *box = (double *)calloc((size_t)(2 * w0 * w1 * w2), sizeof(double));
      for (int64_t k = k0; k < k1; ++k) {
        int64_t j = items[k].idx;
        int64_t i0[3];
        for (int d = 0; d < 3; ++d) {
          i0[d] = (int64_t)ceil(u[3 * j + d] - half);
          for (int t = 0; t < ns; ++t)
            ker[d * ns + t] = es_kernel((double)(i0[d] + t) - u[3 * j + d], (double)ns, beta);
        }
        double c = weight[j];
        for (int t2 = 0; t2 < ns; ++t2) {
          double k2 = c * ker[2 * ns + t2];
          for (int t1 = 0; t1 < ns; ++t1) {
            double k12 = k2 * ker[ns + t1];
            double *row = box + 2 * (((i0[2] - lo[2] + t2) * w1 + (i0[1] - lo[1] + t1)) * w0 +
                                     (i0[0] - lo[0]));
            for (int t0 = 0; t0 < ns; ++t0) row[2 * t0] += k12 * ker[t0];
          }
        }
      }
      double *g = grid + 2 * (size_t)f * (size_t)vol;
      for (int64_t z = 0; z < w2; ++z) {
        int64_t gz = ((lo[2] + z) % n[2] + n[2]) % n[2];
        for (int64_t y = 0; y < w1; ++y) {
          int64_t gy = ((lo[1] + y) % n[1] + n[1]) % n[1];
          for (int64_t x = 0; x < w0; ++x) {
            double v = box[2 * ((z * w1 + y) * w0 + x)];
            if (v == 0.0) continue;
            int64_t gx = ((lo[0] + x) % n[0] + n[0]) % n[0];
#pragma omp atomic
            g[2 * (gz * plane + gy * n[0] + gx)] += v;
          }
        }
      }
      free(box);
    }
    free(ker);
  }
  free(sub_start);
  free(items);
  free(u);
  return 0;
}

/* Deconvolve the transformed grids at the kept modes and accumulate one
 * frame.  ghat is [nfields][n3][n2][n1] complex128 after the +1 FFT;
 * phihat_d[|h|] are the kernel Fourier coefficients per axis.
 * famp[f] is the amplitude applied to field f when forming the total
 * (all 1 for a single pre-weighted field), or NULL with kind 2 (x-ray).
 * partial (S x S x nq) is accumulated when non-NULL and nfields > 1. */
void sqo_nufft_bin(int nfields, const double *ghat, const int64_t *n, int64_t nmodes,
                   const int32_t *hkl, const double *qlen, const int32_t *shell, int nq,
                   const double *ph1, const double *ph2, const double *ph3, int kind,
                   const double *famp, const double *xcoef, double *num, double *partial) {
  const int64_t vol = n[0] * n[1] * n[2];
  double *rho = (double *)malloc(sizeof(double) * 2 * (size_t)nfields * (size_t)nmodes);
#pragma omp parallel for schedule(static)
  for (int64_t m = 0; m < nmodes; ++m) {
    int h1 = hkl[3 * m], h2 = hkl[3 * m + 1], h3 = hkl[3 * m + 2];
    int64_t i1 = ((h1 % n[0]) + n[0]) % n[0], i2 = ((h2 % n[1]) + n[1]) % n[1],
            i3 = ((h3 % n[2]) + n[2]) % n[2];
    double dec = 1.0 / (ph1[abs(h1)] * ph2[abs(h2)] * ph3[abs(h3)]);
    for (int f = 0; f < nfields; ++f) {
      const double *v = ghat + 2 * ((size_t)f * vol + (size_t)((i3 * n[1] + i2) * n[0] + i1));
      rho[2 * ((size_t)f * nmodes + m)] = v[0] * dec;
      rho[2 * ((size_t)f * nmodes + m) + 1] = v[1] * dec;
    }
  }
  /* species pairs, serial over pairs x modes as the reference does */
  if (partial && nfields > 1)
    for (int a = 0; a < nfields; ++a)
      for (int b = a; b < nfields; ++b)
        for (int64_t m = 0; m < nmodes; ++m) {
          double v = rho[2 * ((size_t)a * nmodes + m)] * rho[2 * ((size_t)b * nmodes + m)] +
                     rho[2 * ((size_t)a * nmodes + m) + 1] * rho[2 * ((size_t)b * nmodes + m) + 1];
          partial[((size_t)a * nfields + b) * nq + shell[m] - 1] += v;
        }
  for (int64_t m = 0; m < nmodes; ++m) {
    double tr = 0.0, ti = 0.0;
    for (int f = 0; f < nfields; ++f) {
      double w;
      if (kind == 2) {
        const double *c = xcoef + 9 * f;
        double st = qlen[m] / (4.0 * acos(-1.0));
        st = st * st;
        w = c[8];
        for (int i = 0; i < 4; ++i) w = w + c[i] * exp(-c[4 + i] * st);
      } else {
        w = famp[f];
      }
      tr += w * rho[2 * ((size_t)f * nmodes + m)];
      ti += w * rho[2 * ((size_t)f * nmodes + m) + 1];
    }
    num[shell[m] - 1] += tr * tr + ti * ti;
  }
  free(rho);
}
