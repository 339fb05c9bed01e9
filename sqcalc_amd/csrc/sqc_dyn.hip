// sqc_dyn.hip -- density amplitudes of the dyn path on gfx950 FP64 matrix cores.
//
// The reference's dynamic method sums, per frame and species, one phase per
// (mode, atom) for a set of Miller indices n = (n1, n2, n3) on a basis
// b1, b2, b3 (the reciprocal lattice, or dq u for a q line with its start
// s0 u as a mode-independent base):
//   rho_a(n) = sum_{j in a} exp(i (n1 phi1_j + n2 phi2_j + n3 phi3_j + phi0_j)),
//   phi_k = b_k . r_j, phi0 = base . r_j
// (dyn_accumulate src/sqc_dynamics.f90:826-873 with dyn_atom_phase :1094-1115
// and the separable tables of src/sqc_phase.f90:67-113).
//
// The phase separates, exp(i(R + C)) = exp(iR) exp(iC) with R from the two
// "row" indices and C from the "column" index, so for the modes grouped by
// their row pair p = (n_r1, n_r2) and column index l the amplitudes are a
// complex matrix product over the atoms of a species:
//   rho[p][l] = sum_j A[p][j] B[j][l],  A = exp(i(n_r1 phi_r1 + n_r2 phi_r2 + phi0)),
//                                        B = exp(i l phi_c).
// Rows are sorted by their column range and taken 16 at a time (a row
// block); each block's columns are covered by items of 128.  One workgroup
// computes one item (16 x 128) over a chunk of one species' atoms: each of
// its four waves runs v_mfma_f64_16x16x4 over a quarter of the chunk (four
// atoms a step, real and imaginary parts as four real products per 16 x 16
// tile), the waves' sums are added in a fixed order in LDS and written as
// the chunk's partial.  A gather kernel sums the chunk partials of each
// mode's species in chunk order and writes rho in the caller's mode order,
// so results are reproducible bit for bit.
//
// Operands are formed in registers from the atom phases (an FP64 sine and
// cosine for A and for the first 16 columns of B, then the recurrence
// exp(i(l + 16) phi_c) = exp(i l phi_c) exp(i 16 phi_c) across the item's
// column tiles, exp(i 16 phi_c) computed once per atom).  Tiles past the
// last column of a row block are skipped.  Everything is FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "sqc_hip.h"
#include "sqc_internal.hpp"

typedef double dyn_double4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int kRows = 16;       // rows (p) of a block: the MFMA's M
constexpr int kTiles = 8;       // 16-column tiles of an item
constexpr int kCols = 16 * kTiles;
constexpr int kItem = kRows * kCols;  // complex values of an item's partial
constexpr int kThreadsDyn = 256;
constexpr int kNumDynStages = 3;  // 0 atom phases, 1 matrix product, 2 gather

}  // namespace

struct sqc_hip_dyn_plan {
  sqc_hip_dyn_config cfg{};
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  int64_t natoms = 0, nmodes = 0;
  int S = 0;
  int axis[3] = {0, 1, 2};  // row axes r1, r2 and the column axis
  double recip[9];
  double base[3];
  int base_on = 0;
  int nitems = 0, nchunks = 0, chunk_atoms = 0;
  int64_t macs = 0;         // complex multiply-adds per frame (live tiles x 256 x atoms)
  std::vector<int32_t> chunk_first, chunk_count;   // per species: its first chunk, count
  int32_t *d_perm = nullptr;     // atoms in species order
  int2 *d_prow = nullptr;        // (n_r1, n_r2) of every sorted row
  int4 *d_item = nullptr;        // (first row, first column, live tiles, -) of every item
  int4 *d_chunk = nullptr;       // (first sorted atom, atom count, species, -) per chunk
  int32_t *d_cfirst = nullptr, *d_ccount = nullptr;  // per species: first chunk, chunks
  double *d_phase = nullptr;     // [natoms][8] phases (and the 16-column step) in species order
  double2 *d_part = nullptr;     // [nchunks][nitems][16][128] chunk partials
  int32_t *d_mode_item = nullptr, *d_mode_rc = nullptr;  // per mode: item, row * 128 + col
  double *d_pos = nullptr;       // host-input staging, one frame
  double2 *d_rho = nullptr;      // host-output staging, one frame
  int64_t bytes = 0;
  int profiling = 0;
  struct Pending {
    int stage;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  double stage_ms[kNumDynStages] = {};
  int64_t stage_n[kNumDynStages] = {};
};

namespace {

// sin and cos of x: Cody-Waite reduction x = k pi/2 + r with pi/2 in three
// FMA parts (P1 the double nearest pi/2, P2 and P3 the next two doubles of
// pi/2 - P1, from Machin's formula for pi), |r| <= pi/4, then the Taylor
// series of sin (to r^17) and cos (to r^16), whose truncation is below
// 1e-16 there, and the quadrant.  The reduction's error is about 1e-16
// absolute for |x| < 2^30 (phases q . r up to about 10^9 rad).
__device__ __forceinline__ void dyn_sincos(double x, double *sn, double *cs) {
  const double P1 = 1.5707963267948966, P2 = 6.123233995736766e-17,
               P3 = -1.4973849048591698e-33;
  const double k = rint(x * 0.6366197723675814);
  const double r = fma(-k, P3, fma(-k, P2, fma(-k, P1, x)));
  const double r2 = r * r;
  // 1 / n! of the odd and even terms
  constexpr double f3 = 1.0 / 6.0, f5 = 1.0 / 120.0, f7 = 1.0 / 5040.0, f9 = 1.0 / 362880.0,
                   f11 = 1.0 / 39916800.0, f13 = 1.0 / 6227020800.0,
                   f15 = 1.0 / 1307674368000.0, f17 = 1.0 / 355687428096000.0;
  constexpr double f2 = 1.0 / 2.0, f4 = 1.0 / 24.0, f6 = 1.0 / 720.0, f8 = 1.0 / 40320.0,
                   f10 = 1.0 / 3628800.0, f12 = 1.0 / 479001600.0,
                   f14 = 1.0 / 87178291200.0, f16 = 1.0 / 20922789888000.0;
  double ps = fma(r2, -f17, f15);
  ps = fma(r2, ps, -f13);
  ps = fma(r2, ps, f11);
  ps = fma(r2, ps, -f9);
  ps = fma(r2, ps, f7);
  ps = fma(r2, ps, -f5);
  ps = fma(r2, ps, f3);
  const double sr = fma(-r * r2, ps, r);  // r - r^3 (1/3! - r^2 (1/5! - ...))
  double pc = fma(r2, f16, -f14);
  pc = fma(r2, pc, f12);
  pc = fma(r2, pc, -f10);
  pc = fma(r2, pc, f8);
  pc = fma(r2, pc, -f6);
  pc = fma(r2, pc, f4);
  pc = fma(r2, pc, -f2);
  const double cr = fma(r2, pc, 1.0);
  const int q = (int)((long long)k & 3);
  double a = sr, b = cr;
  if (q & 1) {
    a = cr;
    b = -sr;
  }
  if (q & 2) {
    a = -a;
    b = -b;
  }
  *sn = a;
  *cs = b;
}

#define DYN_TRY(call, what)                                                          \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess)                                                            \
      return sqc::set_error(SQC_HIP_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Phases of the atoms in species order: (phi_r1, phi_r2, phi_c, phi0), each
// phi_k = b_k . r in the reference's term order (dyn_atom_phase).
__global__ void k_dyn_phase(int64_t natoms, const double *__restrict__ pos,
                            const int32_t *__restrict__ perm, double b00, double b10, double b20,
                            double b01, double b11, double b21, double b02, double b12,
                            double b22, double q0, double q1, double q2, int base_on,
                            double *__restrict__ phase) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= natoms) return;
  const int64_t j = perm[t];
  const double r0 = pos[3 * j], r1 = pos[3 * j + 1], r2 = pos[3 * j + 2];
  double4 o, w;
  o.x = b00 * r0 + b10 * r1 + b20 * r2;
  o.y = b01 * r0 + b11 * r1 + b21 * r2;
  o.z = b02 * r0 + b12 * r1 + b22 * r2;
  o.w = base_on ? q0 * r0 + q1 * r1 + q2 * r2 : 0.0;
  // exp(i 16 phi_c): the step between an item's column tiles
  dyn_sincos(16.0 * o.z, &w.y, &w.x);
  w.z = w.w = 0.0;
  reinterpret_cast<double4 *>(phase)[2 * t] = o;
  reinterpret_cast<double4 *>(phase)[2 * t + 1] = w;
}

// The atom loop of one wave over NT live column tiles (a compile-time count
// keeps the accumulators in place across the loop): per step of four atoms
// each lane forms A for its (row, atom) and B for its (atom, column) in
// every tile, then the tiles' real products are issued, the first ones of
// every accumulator before the second ones.
template <int NT>
__device__ __forceinline__ void dyn_steps(const double4 *__restrict__ ph, int a0, int a1,
                                          int sub, double hr1, double hr2, double lc,
                                          dyn_double4 *accr, dyn_double4 *acci) {
  for (int s = a0; s < a1; s += 4) {
    const int j = s + sub;
    double ar = 0.0, ai = 0.0, br[NT], bi[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) br[t] = bi[t] = 0.0;
    if (j < a1) {  // atoms past the end contribute A = 0
      const double4 f = ph[2 * j], w = ph[2 * j + 1];
      dyn_sincos(hr1 * f.x + hr2 * f.y + f.w, &ai, &ar);
      dyn_sincos(lc * f.z, &bi[0], &br[0]);
#pragma unroll
      for (int t = 1; t < NT; ++t) {
        br[t] = br[t - 1] * w.x - bi[t - 1] * w.y;
        bi[t] = br[t - 1] * w.y + bi[t - 1] * w.x;
      }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      accr[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, br[t], accr[t], 0, 0, 0);
      acci[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, bi[t], acci[t], 0, 0, 0);
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      accr[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai, bi[t], accr[t], 0, 0, 0);
      acci[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, br[t], acci[t], 0, 0, 0);
    }
  }
}

// One item (16 rows x 128 columns) over one chunk of one species' atoms.
// grid (nitems, nchunks); wave w takes atoms [w q, (w + 1) q) of the chunk,
// q = ceil(count / 4) rounded up to a multiple of 4.
__global__ void __launch_bounds__(kThreadsDyn) k_dyn_rho(const int2 *__restrict__ prow,
                                                         const int4 *__restrict__ item,
                                                         const int4 *__restrict__ chunk,
                                                         const double *__restrict__ phase,
                                                         int nitems, double2 *__restrict__ part) {
  __shared__ double red[kTiles][2][4][64];  // one wave's sums: (tile, re/im, reg, lane)
  const int it = blockIdx.x, ch = blockIdx.y;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int4 im = item[it];
  const int4 cj = chunk[ch];
  const int row = lane & 15, sub = lane >> 4;  // A row / B column, and the atom of the step
  const int2 pr = prow[im.x + row];
  const double hr1 = (double)pr.x, hr2 = (double)pr.y;
  const double lc = (double)(im.y + row);  // this lane's column in tile 0
  int q = (cj.y + 3) / 4;
  q = (q + 3) & ~3;
  const int a0 = min(cj.y, wave * q), a1 = min(cj.y, a0 + q);
  const double4 *ph = reinterpret_cast<const double4 *>(phase) + 2 * (size_t)cj.x;
  dyn_double4 accr[kTiles], acci[kTiles];
#pragma unroll
  for (int t = 0; t < kTiles; ++t) {
    accr[t] = dyn_double4{0.0, 0.0, 0.0, 0.0};
    acci[t] = dyn_double4{0.0, 0.0, 0.0, 0.0};
  }
  // tiles that hold columns of the block (wave-uniform)
  switch (im.z) {
    case 1: dyn_steps<1>(ph, a0, a1, sub, hr1, hr2, lc, accr, acci); break;
    case 2: dyn_steps<2>(ph, a0, a1, sub, hr1, hr2, lc, accr, acci); break;
    case 3: dyn_steps<3>(ph, a0, a1, sub, hr1, hr2, lc, accr, acci); break;
    case 4: dyn_steps<4>(ph, a0, a1, sub, hr1, hr2, lc, accr, acci); break;
    case 5: dyn_steps<5>(ph, a0, a1, sub, hr1, hr2, lc, accr, acci); break;
    case 6: dyn_steps<6>(ph, a0, a1, sub, hr1, hr2, lc, accr, acci); break;
    case 7: dyn_steps<7>(ph, a0, a1, sub, hr1, hr2, lc, accr, acci); break;
    default: dyn_steps<8>(ph, a0, a1, sub, hr1, hr2, lc, accr, acci); break;
  }
  // the four waves' sums in a fixed order, ((w3 + w2) + w1) + w0, through
  // one slot of LDS
  auto put = [&]() {
#pragma unroll
    for (int t = 0; t < kTiles; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        red[t][0][r][lane] = accr[t][r];
        red[t][1][r][lane] = acci[t][r];
      }
  };
  auto add = [&]() {
#pragma unroll
    for (int t = 0; t < kTiles; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        accr[t][r] = red[t][0][r][lane] + accr[t][r];
        acci[t][r] = red[t][1][r][lane] + acci[t][r];
      }
  };
  for (int w = 3; w >= 1; --w) {
    if (wave == w) {
      if (w < 3) add();
      put();
    }
    __syncthreads();
  }
  if (wave != 0) return;
  add();
  double2 *out = part + ((size_t)ch * nitems + it) * kItem;
#pragma unroll
  for (int t = 0; t < kTiles; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      // D layout of v_mfma_f64_16x16x4: lane L, register r hold row
      // 4 r + L / 16, column L % 16
      const int i = 4 * r + (lane >> 4), c = 16 * t + (lane & 15);
      out[i * kCols + c] = double2{accr[t][r], acci[t][r]};
    }
}

// rho[a][m]: the chunk partials of mode m's item summed in chunk order over
// the chunks of species a.
__global__ void k_dyn_gather(int64_t nmodes, int S, const int32_t *__restrict__ mode_item,
                             const int32_t *__restrict__ mode_rc,
                             const int32_t *__restrict__ chunk_first,
                             const int32_t *__restrict__ chunk_count, int nitems,
                             const double2 *__restrict__ part, double2 *__restrict__ rho) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nmodes * S) return;
  const int a = (int)(t / nmodes);
  const int64_t m = t - (int64_t)a * nmodes;
  const int it = mode_item[m], rc = mode_rc[m];
  double sr = 0.0, si = 0.0;
  const int c0 = chunk_first[a], c1 = c0 + chunk_count[a];
  for (int c = c0; c < c1; ++c) {
    const double2 v = part[((size_t)c * nitems + it) * kItem + rc];
    sr += v.x;
    si += v.y;
  }
  rho[t] = double2{sr, si};
}

template <typename T>
int dev_alloc(T **p, size_t bytes, int64_t &total) {
  const hipError_t e = hipMalloc((void **)p, bytes ? bytes : 8);
  if (e != hipSuccess)
    return sqc::set_error(SQC_HIP_ERR_ALLOC, std::string("hipMalloc: ") + hipGetErrorString(e));
  total += (int64_t)(bytes ? bytes : 8);
  return 0;
}

template <typename T>
int dev_upload(T **p, const void *src, size_t bytes, int64_t &total) {
  if (int rc = dev_alloc(p, bytes, total)) return rc;
  if (bytes) DYN_TRY(hipMemcpy(*p, src, bytes, hipMemcpyHostToDevice), "hipMemcpy");
  return 0;
}

void free_dyn(sqc_hip_dyn_plan *p) {
  void *bufs[] = {p->d_perm,      p->d_prow,    p->d_item,    p->d_chunk,
                  p->d_cfirst,    p->d_ccount,  p->d_phase,   p->d_part,
                  p->d_mode_item, p->d_mode_rc, p->d_pos,     p->d_rho};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  for (auto &q : p->pending) {
    (void)hipEventDestroy(q.a);
    (void)hipEventDestroy(q.b);
  }
  if (p->ev_in) (void)hipEventDestroy(p->ev_in);
  if (p->ev_out) (void)hipEventDestroy(p->ev_out);
  if (p->stream) (void)hipStreamDestroy(p->stream);
  delete p;
}

// Rows of the modes for one choice of the column axis: the distinct row
// pairs sorted by (first column, last column descending), blocks of 16
// rows, and their items of 128 columns.
struct Layout {
  int axis[3];
  std::vector<int32_t> row_of;          // per mode: sorted row index
  std::vector<int2> rows;               // (n_r1, n_r2) per sorted row
  std::vector<int4> items;              // (first row, first column, live tiles, -)
  std::vector<int32_t> item_base;       // per block: its first item
  std::vector<int32_t> block_lo;        // per block: first column
};

Layout make_layout(const int32_t *hkl, int64_t nm, int cax) {
  Layout L;
  const int r1 = cax == 0 ? 1 : 0, r2 = cax == 2 ? 1 : 2;
  L.axis[0] = r1;
  L.axis[1] = r2;
  L.axis[2] = cax;
  struct Row {
    int a, b, lo, hi;
  };
  std::vector<Row> rows;
  std::vector<int32_t> idx(nm);
  {
    // distinct (n_r1, n_r2) pairs with their column range
    std::vector<int64_t> order(nm);
    for (int64_t m = 0; m < nm; ++m) order[m] = m;
    std::sort(order.begin(), order.end(), [&](int64_t x, int64_t y) {
      const int xa = hkl[3 * x + r1], ya = hkl[3 * y + r1];
      if (xa != ya) return xa < ya;
      return hkl[3 * x + r2] < hkl[3 * y + r2];
    });
    for (int64_t k = 0; k < nm; ++k) {
      const int64_t m = order[k];
      const int a = hkl[3 * m + r1], b = hkl[3 * m + r2], c = hkl[3 * m + cax];
      if (rows.empty() || rows.back().a != a || rows.back().b != b) rows.push_back({a, b, c, c});
      Row &r = rows.back();
      r.lo = std::min(r.lo, c);
      r.hi = std::max(r.hi, c);
      idx[m] = (int32_t)rows.size() - 1;
    }
  }
  std::vector<int32_t> sorted(rows.size());
  for (size_t i = 0; i < rows.size(); ++i) sorted[i] = (int32_t)i;
  std::stable_sort(sorted.begin(), sorted.end(), [&](int32_t x, int32_t y) {
    if (rows[x].lo != rows[y].lo) return rows[x].lo < rows[y].lo;
    return rows[x].hi > rows[y].hi;
  });
  std::vector<int32_t> rank(rows.size());
  for (size_t i = 0; i < sorted.size(); ++i) rank[sorted[i]] = (int32_t)i;
  const int nrows = (int)rows.size();
  const int nblk = (nrows + kRows - 1) / kRows;
  L.rows.resize((size_t)nblk * kRows, int2{0, 0});
  for (int i = 0; i < nrows; ++i) L.rows[i] = int2{rows[sorted[i]].a, rows[sorted[i]].b};
  L.row_of.resize(nm);
  for (int64_t m = 0; m < nm; ++m) L.row_of[m] = rank[idx[m]];
  for (int b = 0; b < nblk; ++b) {
    int lo = 1 << 30, hi = -(1 << 30);
    for (int i = b * kRows; i < std::min(nrows, (b + 1) * kRows); ++i) {
      lo = std::min(lo, rows[sorted[i]].lo);
      hi = std::max(hi, rows[sorted[i]].hi);
    }
    L.item_base.push_back((int32_t)L.items.size());
    L.block_lo.push_back(lo);
    for (int c = lo; c <= hi; c += kCols)
      L.items.push_back(int4{b * kRows, c, std::min(kTiles, (hi - c) / 16 + 1), 0});
  }
  return L;
}

struct DynTimer {
  sqc_hip_dyn_plan *p;
  int stage;
  hipEvent_t a = nullptr, b = nullptr;
  DynTimer(sqc_hip_dyn_plan *pl, int st) : p(pl), stage(st) {
    if (!p->profiling) return;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
      if (a) (void)hipEventDestroy(a);
      a = b = nullptr;
      return;
    }
    (void)hipEventRecord(a, p->stream);
  }
  ~DynTimer() {
    if (!a) return;
    (void)hipEventRecord(b, p->stream);
    p->pending.push_back({stage, a, b});
  }
};

int collect_dyn(sqc_hip_dyn_plan *p) {
  if (p->pending.empty()) return 0;
  const hipError_t e = hipStreamSynchronize(p->stream);
  for (auto &q : p->pending) {
    float ms = 0.f;
    if (e == hipSuccess && hipEventElapsedTime(&ms, q.a, q.b) == hipSuccess) {
      p->stage_ms[q.stage] += ms;
      p->stage_n[q.stage] += 1;
    }
    (void)hipEventDestroy(q.a);
    (void)hipEventDestroy(q.b);
  }
  p->pending.clear();
  if (e != hipSuccess)
    return sqc::set_error(SQC_HIP_ERR_HIP, std::string("dyn timing: ") + hipGetErrorString(e));
  return 0;
}

// The launch sequence of one frame on the plan stream.
int run_frame(sqc_hip_dyn_plan *p, const double *pos, double2 *rho) {
  const double *b = p->recip;
  const int *ax = p->axis;
  {
    DynTimer t(p, 0);
    const unsigned g = (unsigned)((p->natoms + 255) / 256);
    // basis columns in the roles r1, r2, c: column k of recip is b_k
    hipLaunchKernelGGL(k_dyn_phase, dim3(g), dim3(256), 0, p->stream, p->natoms, pos, p->d_perm,
                       b[3 * ax[0]], b[3 * ax[0] + 1], b[3 * ax[0] + 2], b[3 * ax[1]],
                       b[3 * ax[1] + 1], b[3 * ax[1] + 2], b[3 * ax[2]], b[3 * ax[2] + 1],
                       b[3 * ax[2] + 2], p->base[0], p->base[1], p->base[2], p->base_on,
                       p->d_phase);
  }
  {
    DynTimer t(p, 1);
    hipLaunchKernelGGL(k_dyn_rho, dim3(p->nitems, p->nchunks), dim3(kThreadsDyn), 0, p->stream,
                       p->d_prow, p->d_item, p->d_chunk, p->d_phase, p->nitems, p->d_part);
  }
  {
    DynTimer t(p, 2);
    const int64_t n = p->nmodes * p->S;
    hipLaunchKernelGGL(k_dyn_gather, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, p->stream,
                       p->nmodes, p->S, p->d_mode_item, p->d_mode_rc,
                       p->d_cfirst, p->d_ccount, p->nitems, p->d_part, rho);
  }
  DYN_TRY(hipGetLastError(), "dyn launch");
  return 0;
}

}  // namespace

extern "C" {

int sqc_hip_dyn_plan_create(sqc_hip_dyn_plan **plan, const sqc_hip_dyn_config *cfg) {
  if (!plan || !cfg) return sqc::set_error(SQC_HIP_ERR_BAD_ARG, "null argument");
  *plan = nullptr;
  if (cfg->abi_version != SQC_HIP_ABI_VERSION)
    return sqc::set_error(SQC_HIP_ERR_ABI, "sqc_hip_dyn_config from ABI version " +
                                               std::to_string(cfg->abi_version) + ", library has " +
                                               std::to_string(SQC_HIP_ABI_VERSION));
  if (cfg->natoms < 1 || cfg->natoms > 2000000000LL || !cfg->species_of)
    return sqc::set_error(SQC_HIP_ERR_BAD_ARG, "natoms must be in 1..2e9 with species_of given");
  if (cfg->nspecies < 1 || cfg->nspecies > 4096)
    return sqc::set_error(SQC_HIP_ERR_BAD_ARG, "nspecies must be in 1..4096");
  if (cfg->nmodes < 1 || cfg->nmodes > 2000000000LL || !cfg->hkl)
    return sqc::set_error(SQC_HIP_ERR_BAD_ARG, "nmodes must be in 1..2e9 with hkl given");
  for (int64_t m = 0; m < 3 * cfg->nmodes; ++m)
    if (cfg->hkl[m] < -(1 << 20) || cfg->hkl[m] > (1 << 20))
      return sqc::set_error(SQC_HIP_ERR_BAD_ARG, "Miller index beyond 2^20");
  std::vector<int64_t> counts(cfg->nspecies, 0);
  for (int64_t j = 0; j < cfg->natoms; ++j) {
    const int a = cfg->species_of[j];
    if (a < 0 || a >= cfg->nspecies)
      return sqc::set_error(SQC_HIP_ERR_BAD_ARG, "species_of out of range");
    ++counts[a];
  }
  int ndev = 0;
  DYN_TRY(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
  if (cfg->device < 0 || cfg->device >= ndev)
    return sqc::set_error(SQC_HIP_ERR_NO_DEVICE, "HIP device " + std::to_string(cfg->device) +
                                                     " requested but " + std::to_string(ndev) +
                                                     " device(s) available");
  auto *p = new sqc_hip_dyn_plan();
  auto bail = [&](int rc) {
    free_dyn(p);
    return rc;
  };
  p->cfg = *cfg;
  p->device = cfg->device;
  p->natoms = cfg->natoms;
  p->nmodes = cfg->nmodes;
  p->S = cfg->nspecies;
  std::memcpy(p->recip, cfg->recip, sizeof(p->recip));
  std::memcpy(p->base, cfg->base, sizeof(p->base));
  p->base_on = cfg->base_on ? 1 : 0;
  if (hipSetDevice(p->device) != hipSuccess) return bail(sqc::set_error(SQC_HIP_ERR_HIP, "hipSetDevice"));
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, p->device) != hipSuccess)
    return bail(sqc::set_error(SQC_HIP_ERR_HIP, "hipGetDeviceProperties"));
  const int ncu = std::max(1, prop.multiProcessorCount);
  // the column axis that needs the fewest items
  Layout best;
  auto live_tiles = [](const Layout &L) {
    int64_t n = 0;
    for (const int4 &i : L.items) n += i.z;
    return n;
  };
  for (int cax = 0; cax < 3; ++cax) {
    Layout L = make_layout(cfg->hkl, cfg->nmodes, cax);
    if (cax == 0 || live_tiles(L) < live_tiles(best)) best = std::move(L);
  }
  const int64_t tiles = live_tiles(best);
  for (int i = 0; i < 3; ++i) p->axis[i] = best.axis[i];
  p->nitems = (int)best.items.size();
  if (p->nitems >= (1 << 30)) return bail(sqc::set_error(SQC_HIP_ERR_MAXNALLOC, "too many items"));
  // atoms in species order, in chunks sized so that the grid fills the device
  std::vector<int32_t> perm;
  perm.reserve(cfg->natoms);
  for (int a = 0; a < p->S; ++a)
    for (int64_t j = 0; j < cfg->natoms; ++j)
      if (cfg->species_of[j] == a) perm.push_back((int32_t)j);
  int ch = 8192;
  auto nch = [&](int c) {
    int64_t n = 0;
    for (int a = 0; a < p->S; ++a) n += (counts[a] + c - 1) / c;
    return n;
  };
  // at least 16 workgroups a CU (four rounds of the four it holds), so the
  // last round's idle CUs cost little
  while (ch > 256 && (int64_t)p->nitems * nch(ch) < 16LL * ncu) ch /= 2;
  p->chunk_atoms = ch;
  std::vector<int4> chunks;
  p->chunk_first.assign(p->S, 0);
  p->chunk_count.assign(p->S, 0);
  {
    int64_t first = 0;
    for (int a = 0; a < p->S; ++a) {
      p->chunk_first[a] = (int32_t)chunks.size();
      for (int64_t o = 0; o < counts[a]; o += ch)
        chunks.push_back(int4{(int)(first + o), (int)std::min<int64_t>(ch, counts[a] - o), a, 0});
      p->chunk_count[a] = (int32_t)chunks.size() - p->chunk_first[a];
      first += counts[a];
    }
  }
  p->nchunks = (int)chunks.size();
  if (p->nchunks >= 65536)
    return bail(sqc::set_error(SQC_HIP_ERR_MAXNALLOC, "more than 65535 atom chunks"));
  const double part_bytes = (double)p->nchunks * p->nitems * kItem * sizeof(double2);
  if (part_bytes > 16e9)
    return bail(sqc::set_error(SQC_HIP_ERR_MAXNALLOC,
                               "chunk partials would take " + std::to_string(part_bytes / 1e9) +
                                   " GB; fewer modes or atoms per plan"));
  p->macs = tiles * 256 * cfg->natoms;
  // mode map
  std::vector<int32_t> mitem(cfg->nmodes), mrc(cfg->nmodes);
  for (int64_t m = 0; m < cfg->nmodes; ++m) {
    const int r = best.row_of[m], b = r / kRows;
    const int c = cfg->hkl[3 * m + p->axis[2]] - best.block_lo[b];
    mitem[m] = best.item_base[b] + c / kCols;
    mrc[m] = (r % kRows) * kCols + c % kCols;
  }
  int rc;
  if ((rc = dev_upload(&p->d_perm, perm.data(), perm.size() * 4, p->bytes))) return bail(rc);
  if ((rc = dev_upload(&p->d_prow, best.rows.data(), best.rows.size() * sizeof(int2), p->bytes)))
    return bail(rc);
  if ((rc = dev_upload(&p->d_item, best.items.data(), best.items.size() * sizeof(int4), p->bytes)))
    return bail(rc);
  if ((rc = dev_upload(&p->d_chunk, chunks.data(), chunks.size() * sizeof(int4), p->bytes)))
    return bail(rc);
  if ((rc = dev_upload(&p->d_cfirst, p->chunk_first.data(), p->S * 4, p->bytes))) return bail(rc);
  if ((rc = dev_upload(&p->d_ccount, p->chunk_count.data(), p->S * 4, p->bytes))) return bail(rc);
  if ((rc = dev_upload(&p->d_mode_item, mitem.data(), mitem.size() * 4, p->bytes))) return bail(rc);
  if ((rc = dev_upload(&p->d_mode_rc, mrc.data(), mrc.size() * 4, p->bytes))) return bail(rc);
  if ((rc = dev_alloc(&p->d_phase, sizeof(double) * 8 * cfg->natoms, p->bytes))) return bail(rc);
  if ((rc = dev_alloc(&p->d_part, (size_t)part_bytes, p->bytes))) return bail(rc);
  if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&p->ev_in, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&p->ev_out, hipEventDisableTiming) != hipSuccess)
    return bail(sqc::set_error(SQC_HIP_ERR_HIP, "stream or event creation"));
  if (hipDeviceSynchronize() != hipSuccess) return bail(sqc::set_error(SQC_HIP_ERR_HIP, "set-up"));
  *plan = p;
  return SQC_HIP_OK;
}

int sqc_hip_dyn_plan_info(const sqc_hip_dyn_plan *p, sqc_hip_dyn_info *info) {
  if (!p || !info) return sqc::set_error(SQC_HIP_ERR_BAD_ARG, "null argument");
  std::memset(info, 0, sizeof(*info));
  info->nmodes = p->nmodes;
  info->nitems = p->nitems;
  info->nchunks = p->nchunks;
  info->chunk_atoms = p->chunk_atoms;
  info->column_axis = p->axis[2];
  info->macs_per_frame = p->macs;
  info->computed_pairs = p->macs / p->natoms;
  info->device_bytes = p->bytes;
  return SQC_HIP_OK;
}

int sqc_hip_dyn_rho_device(sqc_hip_dyn_plan *p, const double *pos_device, int64_t nframes,
                           double *rho_device, void *stream) {
  if (!p || !pos_device || !rho_device || nframes < 0)
    return sqc::set_error(SQC_HIP_ERR_BAD_ARG, "bad dyn arguments");
  DYN_TRY(hipSetDevice(p->device), "hipSetDevice");
  const hipStream_t user = (hipStream_t)stream;
  DYN_TRY(hipEventRecord(p->ev_in, user), "hipEventRecord");
  DYN_TRY(hipStreamWaitEvent(p->stream, p->ev_in, 0), "hipStreamWaitEvent");
  int rc = 0;
  for (int64_t f = 0; f < nframes && !rc; ++f)
    rc = run_frame(p, pos_device + f * 3 * p->natoms,
                   reinterpret_cast<double2 *>(rho_device) + f * p->S * p->nmodes);
  DYN_TRY(hipEventRecord(p->ev_out, p->stream), "hipEventRecord");
  DYN_TRY(hipStreamWaitEvent(user, p->ev_out, 0), "hipStreamWaitEvent");
  return rc;
}

int sqc_hip_dyn_rho(sqc_hip_dyn_plan *p, const double *pos, int64_t nframes, double *rho) {
  if (!p || !pos || !rho || nframes < 0) return sqc::set_error(SQC_HIP_ERR_BAD_ARG, "bad dyn arguments");
  DYN_TRY(hipSetDevice(p->device), "hipSetDevice");
  if (!p->d_pos) {
    if (int rc = dev_alloc(&p->d_pos, sizeof(double) * 3 * p->natoms, p->bytes)) return rc;
    if (int rc = dev_alloc(&p->d_rho, sizeof(double2) * p->S * p->nmodes, p->bytes)) return rc;
  }
  const size_t nr = sizeof(double2) * p->S * p->nmodes;
  for (int64_t f = 0; f < nframes; ++f) {
    DYN_TRY(hipMemcpyAsync(p->d_pos, pos + f * 3 * p->natoms, sizeof(double) * 3 * p->natoms,
                           hipMemcpyHostToDevice, p->stream),
            "upload positions");
    if (int rc = run_frame(p, p->d_pos, p->d_rho)) return rc;
    DYN_TRY(hipMemcpyAsync(reinterpret_cast<char *>(rho) + f * nr, p->d_rho, nr,
                           hipMemcpyDeviceToHost, p->stream),
            "download amplitudes");
    DYN_TRY(hipStreamSynchronize(p->stream), "hipStreamSynchronize");
  }
  return SQC_HIP_OK;
}

int sqc_hip_dyn_profile(sqc_hip_dyn_plan *p, int enable) {
  if (!p) return sqc::set_error(SQC_HIP_ERR_BAD_ARG, "null plan");
  DYN_TRY(hipSetDevice(p->device), "hipSetDevice");
  if (int rc = collect_dyn(p)) return rc;
  p->profiling = enable ? 1 : 0;
  return SQC_HIP_OK;
}

int sqc_hip_dyn_stage_times(sqc_hip_dyn_plan *p, double *ms, int64_t *launches, int nstage) {
  if (!p) return sqc::set_error(SQC_HIP_ERR_BAD_ARG, "null plan");
  DYN_TRY(hipSetDevice(p->device), "hipSetDevice");
  if (int rc = collect_dyn(p)) return rc;
  for (int i = 0; i < nstage && i < kNumDynStages; ++i) {
    if (ms) ms[i] = p->stage_ms[i];
    if (launches) launches[i] = p->stage_n[i];
    p->stage_ms[i] = 0.0;
    p->stage_n[i] = 0;
  }
  return SQC_HIP_OK;
}

int sqc_hip_dyn_destroy(sqc_hip_dyn_plan *p) {
  if (!p) return SQC_HIP_OK;
  (void)hipSetDevice(p->device);
  (void)collect_dyn(p);
  free_dyn(p);
  return SQC_HIP_OK;
}

}  // extern "C"
