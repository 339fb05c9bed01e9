// sqc_debye.hip -- pair-distance histograms of the Debye method on gfx950.
//
// The reference's Debye method (src/sqc_debye.f90) counts, per frame, every
// ordered atom pair (i, j) within the cut-off by its distance bin and type
// pair, histogram(type_i, type_j, int(|d| / dr) + 1) += 1 (debye_accumulate
// :179-235), and evaluates S(q) and g(r) from the histograms afterwards.
// The pair displacement is that of its neighbour list (src/sqc_neighbour_list.f90):
//   d = pos_j - pos_i + bins . shift,
// pos the list's positions (wrapped into the cell on the first frame, then
// carried on by the smallest-image step of each frame, nl_begin_frame
// :190-247), bins(:, k) = a_k / div_k the cells of its decomposition
// (div_k = max(1, int(width_k / (rmax + skin)))) and shift = div * m the
// image m of j in cell units.  Every expression here follows that order and
// the file is compiled with FP contraction off, so each pair lands in the
// reference's bin.
//
// On the device the pairs are counted unordered: d(j, i) = -d(i, j) exactly
// (IEEE subtraction and products are sign-symmetric), so one evaluation adds
// the pair to (t_i, t_j) and (t_j, t_i).  Atoms sit in slots sorted by type,
// each type padded to whole tiles of 1024 (or 256) slots, so a pair of tiles
// is one type pair.  Persistent workgroups walk contiguous runs of the upper
// triangle of tile pairs; the j tile sits in LDS, each thread takes one i.
// A pair first tries its nearest image m = -rint(s) (s the fractional
// displacement); only when a second image can be within reach (rmax above
// half a cell width) are all images m with |s_k + m_k| within rmax / width_k
// tried, so cut-offs up to the smallest width are counted like the cell
// list counts them.  Counts go to one u32 histogram row in LDS (integer
// adds: the result does not depend on their order), added to the
// workgroup's own row in HBM when the type pair changes, or straight to HBM
// rows when a row does not fit the LDS; the rows are summed into the run's
// u64 histogram per frame.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "sqc_hip.h"
#include "sqc_internal.hpp"

#pragma clang fp contract(off)

namespace {

// Atoms per tile (and threads per workgroup): 1024 when the upper triangle
// of 1024-atom tiles keeps every CU busy, else 256.
constexpr int kTileBig = 1024, kTileSmall = 256;
constexpr size_t kHistLds = 96 * 1024;  // LDS for the histogram of a workgroup
// Fractional coordinates at or beyond this are not paired (and reported).
constexpr double kFarCells = 1048576.0;

struct DebyeGeo {
  double a[9], inv[9], bins[9], origin[3];
  double div[3];
  double reach_frac[3];  // rmax / width_k: the largest |s_k| of a pair within rmax
  double thr[3];         // 1 - reach_frac - 1e-9 (+inf along non-periodic axes)
  int pbc[3];
  double rmax2;          // rmax * rmax
  double inv_dr;
  int nbins, ntypes, npu;  // npu: unordered type pairs
  int64_t natoms;
  int64_t nslots;        // atoms and padding: each type's slots fill whole tiles
  int tile;              // slots per tile (the launch's block size)
  int ntile;
  int64_t ntilepairs;
};

}  // namespace

struct sqc_hip_debye_plan {
  sqc_hip_debye_config cfg{};
  DebyeGeo g{};
  int device = 0;
  hipStream_t stream = nullptr;
  int started = 0;
  int nblk = 0;
  int nrows = 0;      // rows of d_part
  int hist_in_lds = 1;
  int64_t nframes = 0;
  int64_t bytes = 0;
  int cell_kind = 0;             // 0 general, 1 upper triangular, 2 diagonal bins, 3 and div 1
  // per slot (atoms sorted by type): the list's positions, the pair kernel's
  // packed copy {pos, f1} {f2, f3}, the atom and its 0-based type; per tile
  // its single type or -1
  double *d_pos = nullptr, *d_posw = nullptr;
  double4 *d_pa = nullptr;
  double2 *d_pb = nullptr;
  int32_t *d_perm = nullptr, *d_tile_type = nullptr;  // slot -> atom (-1: padding)
  int *d_flag = nullptr;         // position checks of the frames (k_debye_prep)
  uint32_t *d_part = nullptr;    // [nrows][npu][nbins] per-frame counts
  unsigned long long *d_hist = nullptr;  // [npu][nbins] run histogram
  double *d_stage = nullptr;     // host-input staging, one frame
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  // per-stage HIP-event times (0 positions, 1 pairs, 2 row sum) when enabled
  int profiling = 0;
  struct Pending {
    int stage;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  double stage_ms[3] = {};
  int64_t stage_n[3] = {};
};

namespace {

#define DEB_TRY(call, what)                                                                   \
  do {                                                                                        \
    hipError_t e_ = (call);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      return sqc::set_error(SQC_HIP_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e_)); \
  } while (0)

// matmul(m, v) of a 3x3 column-major matrix, summed k = 1, 2, 3
__device__ __host__ inline void mat3v(const double *m, const double *v, double *o) {
  for (int i = 0; i < 3; ++i) o[i] = m[i] * v[0] + m[i + 3] * v[1] + m[i + 6] * v[2];
}

// The neighbour list's positions of one frame (nl_begin_frame): the
// fractional coordinate s = A^-1 (r - o) folded into [0, 1) and r_new = A s;
// on the first frame pos = r_new, afterwards pos += A (ds - anint(ds)) with
// ds = A^-1 (r_new - r_prev).  Atoms are held in slots sorted by type
// (perm: slot -> atom), so that most tiles hold a single type.  The pair
// kernel's copy packs pos with f = A^-1 pos, which picks the pair images;
// f is 0 along non-periodic axes (no images there), and an atom whose f is
// not finite or 2^20 cells or more away gets pos = NaN (no pairs).
// flag bits: 1 such an atom, 2 an atom outside the box along a
// non-periodic axis
__global__ void k_debye_prep(DebyeGeo g, const double *__restrict__ r,
                             const int32_t *__restrict__ perm, int first,
                             double *__restrict__ pos, double *__restrict__ posw,
                             double4 *__restrict__ pa, double2 *__restrict__ pb,
                             int *__restrict__ flag) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= g.nslots) return;
  const int64_t i = perm[k];
  if (i < 0) return;  // padding slot
  double rel[3], s[3], rn[3];
  for (int c = 0; c < 3; ++c) rel[c] = r[3 * i + c] - g.origin[c];
  mat3v(g.inv, rel, s);
  for (int c = 0; c < 3; ++c) s[c] = s[c] - floor(s[c]);
  mat3v(g.a, s, rn);
  double p[3];
  if (first) {
    for (int c = 0; c < 3; ++c) p[c] = rn[c];
  } else {
    double dr[3], ds[3];
    for (int c = 0; c < 3; ++c) dr[c] = rn[c] - posw[3 * k + c];
    mat3v(g.inv, dr, ds);
    for (int c = 0; c < 3; ++c) ds[c] = ds[c] - round(ds[c]);  // anint: halves away from 0
    mat3v(g.a, ds, dr);
    for (int c = 0; c < 3; ++c) p[c] = pos[3 * k + c] + dr[c];
  }
  double f[3];
  mat3v(g.inv, p, f);
  int bad = 0;
  for (int c = 0; c < 3; ++c) {
    pos[3 * k + c] = p[c];
    posw[3 * k + c] = rn[c];
    if (!(fabs(f[c]) < kFarCells)) bad |= 1;
    else if (!g.pbc[c] && (f[c] < 0.0 || f[c] >= 1.0)) bad |= 2;
    if (!g.pbc[c]) f[c] = 0.0;
  }
  if (bad & 1) {
    const double nan = __builtin_nan("");
    p[0] = p[1] = p[2] = nan;
    f[0] = f[1] = f[2] = 0.0;
  }
  pa[k] = make_double4(p[0], p[1], p[2], f[0]);
  pb[k] = make_double2(f[1], f[2]);
  if (bad) atomicOr(flag, bad);
}

// Unordered type-pair index of (a, b), a <= b.
__device__ __forceinline__ int upair(int a, int b, int T) {
  if (a > b) {
    const int t = a;
    a = b;
    b = t;
  }
  return a * T - (a * (a - 1)) / 2 + (b - a);
}

// bins . c summed k = 1, 2, 3 (matmul order).  Cell 1: bins upper
// triangular (the restricted triclinic cell of a LAMMPS box), 2 and 3:
// diagonal (3: one list cell per axis, shift = m);
// the products with the zero entries are +-0 and leave every sum, and so
// every squared distance, as it is.
template <int Cell>
__device__ __forceinline__ void bins_times(const DebyeGeo &g, double c0, double c1, double c2,
                                           double &v0, double &v1, double &v2) {
  if (Cell >= 2) {
    v0 = g.bins[0] * c0;
    v1 = g.bins[4] * c1;
    v2 = g.bins[8] * c2;
  } else if (Cell == 1) {
    v0 = g.bins[0] * c0 + g.bins[3] * c1 + g.bins[6] * c2;
    v1 = g.bins[4] * c1 + g.bins[7] * c2;
    v2 = g.bins[8] * c2;
  } else {
    v0 = g.bins[0] * c0 + g.bins[3] * c1 + g.bins[6] * c2;
    v1 = g.bins[1] * c0 + g.bins[4] * c1 + g.bins[7] * c2;
    v2 = g.bins[2] * c0 + g.bins[5] * c1 + g.bins[8] * c2;
  }
}

// One (pair, image): d = (pos_j - pos_i) + bins . (div * m), counted in bin
// int(sqrt(d . d) / dr) when d . d <= rmax^2.  The bin comes from one
// Newton step on the hardware reciprocal square root (relative error below
// 1e-14, against 5e-8 for the bare instruction); only where that lands
// within 1e-12 t of a bin edge (and for d . d = 0) is the correctly rounded
// sqrt(d . d) / dr taken, so every pair gets the reference's bin.
template <int Cell>
__device__ __forceinline__ void count_image(const DebyeGeo &g, double e0, double e1, double e2,
                                            double c0, double c1, double c2, uint32_t *hist,
                                            int row) {
  double v0, v1, v2;
  bins_times<Cell>(g, c0, c1, c2, v0, v1, v2);
  const double d0 = e0 + v0, d1 = e1 + v1, d2 = e2 + v2;
  const double dd = d0 * d0 + d1 * d1 + d2 * d2;
  if (!(dd <= g.rmax2)) return;
  const double r0 = __builtin_amdgcn_rsq(dd);
  const double y0 = dd * r0, h0 = 0.5 * r0;
  const double y1 = __builtin_fma(y0, __builtin_fma(-y0, h0, 0.5), y0);
  const double t = y1 * g.inv_dr;
  const double fl = floor(t), tol = 1e-12 * t;
  int bin;
  if (t - fl > tol && fl + 1.0 - t > tol)
    bin = (int)fl;
  else
    bin = (int)(sqrt(dd) * g.inv_dr);
  if (bin < g.nbins) atomicAdd(&hist[row + bin], 1u);
}

// The pairs of atom i (registers) with the j tile's atoms j0..nj-1 (LDS),
// counted into one histogram row.
template <int Cell>
__device__ __forceinline__ void pair_row(const DebyeGeo &g, double4 ai, double2 bi,
                                         const double4 *sa, const double2 *sb, int j0, int nj,
                                         double thr0, double thr1, double thr2, uint32_t *hist) {
  for (int jl = j0; jl < nj; ++jl) {
    const double4 aj = sa[jl];
    const double2 bj = sb[jl];
    const double s0 = aj.w - ai.w, s1 = bj.x - bi.x, s2 = bj.y - bi.y;
    const double r0 = rint(s0), r1 = rint(s1), r2 = rint(s2);
    const double e0 = aj.x - ai.x, e1 = aj.y - ai.y, e2 = aj.z - ai.z;
    if (fabs(s0 - r0) < thr0 && fabs(s1 - r1) < thr1 && fabs(s2 - r2) < thr2) {
      // shift = div * m with m = -rint(s): exact small integers (m itself
      // when every div is 1)
      if (Cell == 3)
        count_image<Cell>(g, e0, e1, e2, -r0, -r1, -r2, hist, 0);
      else
        count_image<Cell>(g, e0, e1, e2, g.div[0] * -r0, g.div[1] * -r1, g.div[2] * -r2, hist, 0);
      continue;
    }
    const double rb0 = g.reach_frac[0] + 1e-9, rb1 = g.reach_frac[1] + 1e-9,
                 rb2 = g.reach_frac[2] + 1e-9;
    int lo0 = 0, hi0 = 0, lo1 = 0, hi1 = 0, lo2 = 0, hi2 = 0;
    if (g.pbc[0]) { lo0 = (int)ceil(-s0 - rb0); hi0 = (int)floor(-s0 + rb0); }
    if (g.pbc[1]) { lo1 = (int)ceil(-s1 - rb1); hi1 = (int)floor(-s1 + rb1); }
    if (g.pbc[2]) { lo2 = (int)ceil(-s2 - rb2); hi2 = (int)floor(-s2 + rb2); }
    for (int m2 = lo2; m2 <= hi2; ++m2)
      for (int m1 = lo1; m1 <= hi1; ++m1)
        for (int m0 = lo0; m0 <= hi0; ++m0)
          count_image<Cell>(g, e0, e1, e2, g.div[0] * m0, g.div[1] * m1, g.div[2] * m2, hist, 0);
  }
}

// Persistent workgroups, each over a contiguous run of the upper triangle
// of tiles: the j tile in LDS, one i per thread.  Per pair the nearest image
// m = -rint(s) of the fractional displacement s; a second image per axis is
// within reach only when |s + m| >= 1 - rmax / width (rmax above half a
// width), which takes the general loop over every image in reach.  Each
// type's slots fill whole tiles, so a tile pair counts into one row (type
// pair): Lds keeps that row in LDS and adds it to the workgroup's own rows
// in HBM when the row changes and at the end; otherwise the atomics go to
// rows in HBM shared by workgroups.  Padding slots have NaN positions and
// never count.
// Diagonal cells fit 8 waves per SIMD (two 1024-thread workgroups per CU)
// without spills; the others keep their registers.
template <int kTile, bool Lds, int Cell>
__global__ void __launch_bounds__(kTile, Cell >= 2 ? 8 : 1)
    k_debye_pairs(DebyeGeo g, const double4 *__restrict__ pa, const double2 *__restrict__ pb,
                  const int32_t *__restrict__ tile_type, uint32_t *__restrict__ part, int nrows) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double4 sa[kTile];
  __shared__ double2 sb[kTile];
  const int tid = threadIdx.x;
  const int nb = g.nbins;
  uint32_t *rows = part + (size_t)(blockIdx.x % nrows) * g.npu * nb;
  uint32_t *lrow = reinterpret_cast<uint32_t *>(smem);
  const double thr0 = g.thr[0], thr1 = g.thr[1], thr2 = g.thr[2];
  const int64_t chunk = (g.ntilepairs + gridDim.x - 1) / gridDim.x;
  const int64_t i0 = (int64_t)blockIdx.x * chunk;
  const int64_t i1 = i0 + chunk < g.ntilepairs ? i0 + chunk : g.ntilepairs;
  int cur = -1;  // the row held in LDS
  if (Lds)
    for (int k = tid; k < nb; k += kTile) lrow[k] = 0u;
  for (int64_t idx = i0; idx < i1; ++idx) {
    // upper-triangle tile pair (ti <= tj) of index idx = tj (tj + 1) / 2 + ti
    int64_t tj = (int64_t)((sqrt(8.0 * (double)idx + 1.0) - 1.0) * 0.5);
    while ((tj + 1) * (tj + 2) / 2 <= idx) ++tj;
    while (tj * (tj + 1) / 2 > idx) --tj;
    const int64_t ti = idx - tj * (tj + 1) / 2;
    const int row = upair(tile_type[ti], tile_type[tj], g.ntypes);
    __syncthreads();  // the previous tile pair's reads of sa/sb and counts are done
    if (Lds && row != cur) {
      if (cur >= 0)
        for (int k = tid; k < nb; k += kTile) {
          const uint32_t v = lrow[k];
          if (v) rows[(size_t)cur * nb + k] += v;
          lrow[k] = 0u;
        }
      cur = row;
    }
    sa[tid] = pa[tj * kTile + tid];
    sb[tid] = pb[tj * kTile + tid];
    __syncthreads();
    const int64_t i = ti * kTile + tid;
    pair_row<Cell>(g, pa[i], pb[i], sa, sb, ti == tj ? tid + 1 : 0, kTile, thr0, thr1, thr2,
                   Lds ? lrow : rows + (size_t)row * nb);
  }
  if (!Lds || cur < 0) return;
  __syncthreads();
  for (int k = tid; k < nb; k += kTile) {
    const uint32_t v = lrow[k];
    if (v) rows[(size_t)cur * nb + k] += v;
  }
}

__global__ void k_debye_zero(size_t n, uint32_t *__restrict__ part) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) part[k] = 0u;
}

// The workgroups' rows of one frame added into the run histogram.
// (blockIdx.y: a run of `per` rows; integer sums, so any order gives the same counts)
__global__ void k_debye_reduce(int nrows, int nh, int per, const uint32_t *__restrict__ part,
                               unsigned long long *__restrict__ hist) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nh) return;
  const int b0 = blockIdx.y * per, b1 = min(nrows, b0 + per);
  unsigned long long s = 0;
  for (int b = b0; b < b1; ++b) s += part[(size_t)b * nh + k];
  if (s) atomicAdd(&hist[k], s);
}

double det3h(const double *m) {
  auto M = [&](int i, int j) { return m[i + 3 * j]; };
  return M(0, 0) * (M(1, 1) * M(2, 2) - M(1, 2) * M(2, 1)) -
         M(0, 1) * (M(1, 0) * M(2, 2) - M(1, 2) * M(2, 0)) +
         M(0, 2) * (M(1, 0) * M(2, 1) - M(1, 1) * M(2, 0));
}

// The reference's cofactor inverse (src/sqc_cell.f90:135-151 term order).
void inverse3h(const double *m, double *inv) {
  auto M = [&](int i, int j) { return m[i + 3 * j]; };
  auto I = [&](int i, int j) -> double & { return inv[i + 3 * j]; };
  const double det = det3h(m);
  I(0, 0) = (M(1, 1) * M(2, 2) - M(1, 2) * M(2, 1));
  I(1, 0) = -(M(1, 0) * M(2, 2) - M(1, 2) * M(2, 0));
  I(2, 0) = (M(1, 0) * M(2, 1) - M(1, 1) * M(2, 0));
  I(0, 1) = -(M(0, 1) * M(2, 2) - M(0, 2) * M(2, 1));
  I(1, 1) = (M(0, 0) * M(2, 2) - M(0, 2) * M(2, 0));
  I(2, 1) = -(M(0, 0) * M(2, 1) - M(0, 1) * M(2, 0));
  I(0, 2) = (M(0, 1) * M(1, 2) - M(0, 2) * M(1, 1));
  I(1, 2) = -(M(0, 0) * M(1, 2) - M(0, 2) * M(1, 0));
  I(2, 2) = (M(0, 0) * M(1, 1) - M(0, 1) * M(1, 0));
  for (int k = 0; k < 9; ++k) inv[k] = inv[k] / det;
}

// Perpendicular widths of the cell (cell_widths, src/sqc_cell.f90:100-117).
void widths3(const double *a, double *w) {
  const double vol = std::fabs(det3h(a));
  for (int i = 0; i < 3; ++i) {
    const int j = (i + 1) % 3, k = (j + 1) % 3;
    const double *u = a + 3 * j, *v = a + 3 * k;
    const double n0 = u[1] * v[2] - u[2] * v[1];
    const double n1 = u[2] * v[0] - u[0] * v[2];
    const double n2 = u[0] * v[1] - u[1] * v[0];
    w[i] = vol / std::sqrt(n0 * n0 + n1 * n1 + n2 * n2);
  }
}

void free_debye(sqc_hip_debye_plan *p) {
  void *bufs[] = {p->d_pos,  p->d_posw, p->d_pa,   p->d_pb,        p->d_perm,
                  p->d_part, p->d_hist, p->d_stage, p->d_flag, p->d_tile_type};
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

struct DebTimer {
  sqc_hip_debye_plan *p;
  int stage;
  hipEvent_t a = nullptr, b = nullptr;
  DebTimer(sqc_hip_debye_plan *pl, int st) : p(pl), stage(st) {
    if (!p->profiling) return;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
      if (a) (void)hipEventDestroy(a);
      a = b = nullptr;
      return;
    }
    (void)hipEventRecord(a, p->stream);
  }
  ~DebTimer() {
    if (!a) return;
    (void)hipEventRecord(b, p->stream);
    p->pending.push_back({stage, a, b});
  }
};

int collect_debye(sqc_hip_debye_plan *p) {
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
    return sqc::set_error(SQC_HIP_ERR_HIP, std::string("debye timing: ") + hipGetErrorString(e));
  return 0;
}

// The pair kernel of the plan's tile, histogram place and cell kind.
template <int kTile, bool Lds, int Cell>
void launch_pairs_t(sqc_hip_debye_plan *p) {
  const size_t lds = Lds ? (size_t)p->g.nbins * 4 : 0;
  hipLaunchKernelGGL((k_debye_pairs<kTile, Lds, Cell>), dim3(p->nblk), dim3(kTile), lds,
                     p->stream, p->g, p->d_pa, p->d_pb, p->d_tile_type, p->d_part, p->nrows);
}

template <int kTile, bool Lds>
void launch_pairs_c(sqc_hip_debye_plan *p) {
  if (p->cell_kind == 3)
    launch_pairs_t<kTile, Lds, 3>(p);
  else if (p->cell_kind == 2)
    launch_pairs_t<kTile, Lds, 2>(p);
  else if (p->cell_kind == 1)
    launch_pairs_t<kTile, Lds, 1>(p);
  else
    launch_pairs_t<kTile, Lds, 0>(p);
}

void launch_pairs(sqc_hip_debye_plan *p) {
  if (p->g.tile == kTileBig) {
    if (p->hist_in_lds)
      launch_pairs_c<kTileBig, true>(p);
    else
      launch_pairs_c<kTileBig, false>(p);
  } else {
    if (p->hist_in_lds)
      launch_pairs_c<kTileSmall, true>(p);
    else
      launch_pairs_c<kTileSmall, false>(p);
  }
}

// Address of the pair kernel that launch_pairs runs (for its LDS attribute).
template <int kTile, bool Lds>
const void *pairs_fn_c(int cell_kind) {
  if (cell_kind == 3) return (const void *)k_debye_pairs<kTile, Lds, 3>;
  if (cell_kind == 2) return (const void *)k_debye_pairs<kTile, Lds, 2>;
  if (cell_kind == 1) return (const void *)k_debye_pairs<kTile, Lds, 1>;
  return (const void *)k_debye_pairs<kTile, Lds, 0>;
}

int debye_frame(sqc_hip_debye_plan *p, const double *pos_dev) {
  DebyeGeo &g = p->g;
  const unsigned ga = (unsigned)((g.nslots + 255) / 256);
  {
    DebTimer t(p, 0);
    hipLaunchKernelGGL(k_debye_prep, dim3(ga), dim3(256), 0, p->stream, g, pos_dev, p->d_perm,
                       p->started ? 0 : 1, p->d_pos, p->d_posw, p->d_pa, p->d_pb, p->d_flag);
  }
  p->started = 1;
  const int nh = g.npu * g.nbins;
  {
    DebTimer t(p, 1);
    const size_t n = (size_t)p->nrows * nh;
    hipLaunchKernelGGL(k_debye_zero, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, p->stream, n,
                       p->d_part);
    launch_pairs(p);
  }
  {
    DebTimer t(p, 2);
    const int per = 16;
    hipLaunchKernelGGL(k_debye_reduce, dim3((nh + 255) / 256, (p->nrows + per - 1) / per),
                       dim3(256), 0, p->stream, p->nrows, nh, per, p->d_part, p->d_hist);
  }
  DEB_TRY(hipGetLastError(), "debye launch");
  ++p->nframes;
  return 0;
}

// The position flags of the frames so far (the stream is synchronised).
int check_flag(sqc_hip_debye_plan *p) {
  int flag = 0;
  DEB_TRY(hipMemcpy(&flag, p->d_flag, sizeof(int), hipMemcpyDeviceToHost), "read position flags");
  if (flag & 2)
    return sqc::set_error(SQC_HIP_ERR_POSITIONS, "an atom lies outside a non-periodic box");
  if (flag & 1)
    return sqc::set_error(SQC_HIP_ERR_POSITIONS,
                          "an atom is not finite or moved 2^20 cells or more from its first image");
  return SQC_HIP_OK;
}

template <typename T>
int dalloc(T **ptr, size_t bytes, int64_t &total) {
  const hipError_t e = hipMalloc((void **)ptr, bytes ? bytes : 8);
  if (e != hipSuccess)
    return sqc::set_error(SQC_HIP_ERR_ALLOC, std::string("hipMalloc: ") + hipGetErrorString(e));
  total += (int64_t)(bytes ? bytes : 8);
  return 0;
}

}  // namespace

extern "C" {

int sqc_hip_debye_plan_create(sqc_hip_debye_plan **plan, const sqc_hip_debye_config *cfg) {
  if (!plan || !cfg) return sqc::set_error(SQC_HIP_ERR_BAD_ARG, "null argument");
  *plan = nullptr;
  if (cfg->abi_version != SQC_HIP_ABI_VERSION)
    return sqc::set_error(SQC_HIP_ERR_ABI, "sqc_hip_debye_config from another ABI version");
  if (cfg->natoms < 2 || cfg->natoms > 2000000000LL || !cfg->type_id)
    return sqc::set_error(SQC_HIP_ERR_BAD_ARG, "natoms must be in 2..2e9 with type_id given");
  if (cfg->ntypes < 1 || cfg->ntypes > 64)
    return sqc::set_error(SQC_HIP_ERR_BAD_ARG, "ntypes must be in 1..64");
  for (int64_t j = 0; j < cfg->natoms; ++j)
    if (cfg->type_id[j] < 1 || cfg->type_id[j] > cfg->ntypes)
      return sqc::set_error(SQC_HIP_ERR_BAD_ARG, "atom type id " + std::to_string(cfg->type_id[j]) +
                                                     " is outside 1.." + std::to_string(cfg->ntypes));
  if (!(cfg->dr > 0.0)) return sqc::set_error(SQC_HIP_ERR_BAD_ARG, "the radial bin width must be positive");
  if (!(cfg->rmax > 0.0)) return sqc::set_error(SQC_HIP_ERR_BAD_ARG, "the pair cutoff must be positive");
  auto *p = new sqc_hip_debye_plan();
  auto bail = [&](int rc) {
    free_debye(p);
    return rc;
  };
  p->cfg = *cfg;
  DebyeGeo &g = p->g;
  std::memcpy(g.a, cfg->cell, sizeof(g.a));
  std::memcpy(g.origin, cfg->origin, sizeof(g.origin));
  inverse3h(g.a, g.inv);
  double w[3];
  widths3(g.a, w);
  double bound = 1e300;
  for (int k = 0; k < 3; ++k) {
    g.pbc[k] = cfg->pbc[k] ? 1 : 0;
    if (!(w[k] > 0.0)) return bail(sqc::set_error(SQC_HIP_ERR_BAD_ARG, "degenerate simulation cell"));
    if (g.pbc[k]) bound = std::min(bound, w[k]);
  }
  if (bound < 1e300 && cfg->rmax >= bound)
    return bail(sqc::set_error(SQC_HIP_ERR_BAD_ARG,
                               "the pair cutoff must stay below the smallest periodic box side"));
  const double reach = cfg->rmax + std::max(cfg->skin, 0.0);
  for (int k = 0; k < 3; ++k) {
    const int dv = std::max(1, (int)(w[k] / reach));
    g.div[k] = (double)dv;
    for (int c = 0; c < 3; ++c) g.bins[c + 3 * k] = g.a[c + 3 * k] / (double)dv;
    g.reach_frac[k] = cfg->rmax / w[k];
    g.thr[k] = g.pbc[k] ? 1.0 - (g.reach_frac[k] + 1e-9) : HUGE_VAL;
  }
  {
    const double *A = g.a;
    const bool lower0 = A[1] == 0.0 && A[2] == 0.0 && A[5] == 0.0;
    const bool upper0 = A[3] == 0.0 && A[6] == 0.0 && A[7] == 0.0;
    p->cell_kind = lower0 ? (upper0 ? 2 : 1) : 0;
    if (p->cell_kind == 2 && g.div[0] == 1.0 && g.div[1] == 1.0 && g.div[2] == 1.0)
      p->cell_kind = 3;
  }
  g.rmax2 = cfg->rmax * cfg->rmax;
  g.inv_dr = 1.0 / cfg->dr;
  g.nbins = std::max(1, (int)std::ceil(cfg->rmax / cfg->dr));
  g.ntypes = cfg->ntypes;
  g.npu = cfg->ntypes * (cfg->ntypes + 1) / 2;
  if ((double)g.nbins * cfg->ntypes * cfg->ntypes > 2e8)
    return bail(sqc::set_error(SQC_HIP_ERR_MAXNALLOC,
                               "the pair histogram would need too many bins x type pairs; "
                               "increase dr or reduce rmax"));
  g.natoms = cfg->natoms;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || cfg->device < 0 || cfg->device >= ndev)
    return bail(sqc::set_error(SQC_HIP_ERR_NO_DEVICE, "HIP device " + std::to_string(cfg->device) +
                                                          " not available"));
  p->device = cfg->device;
  if (hipSetDevice(p->device) != hipSuccess) return bail(sqc::set_error(SQC_HIP_ERR_HIP, "hipSetDevice"));
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, p->device) != hipSuccess)
    return bail(sqc::set_error(SQC_HIP_ERR_HIP, "hipGetDeviceProperties"));
  const int ncu = std::max(1, prop.multiProcessorCount);
  // slots: each type's atoms (in atom order) padded to whole tiles of the
  // launch's block size, 1024 when its upper triangle keeps every CU busy
  const int64_t n = cfg->natoms;
  std::vector<int64_t> count(cfg->ntypes, 0);
  for (int64_t j = 0; j < n; ++j) ++count[cfg->type_id[j] - 1];
  auto slots_for = [&](int tile) {
    int64_t s = 0;
    for (int t = 0; t < cfg->ntypes; ++t) s += (count[t] + tile - 1) / tile * tile;
    return s;
  };
  {
    const int64_t nt = slots_for(kTileBig) / kTileBig;
    g.tile = nt * (nt + 1) / 2 >= 4 * (int64_t)ncu ? kTileBig : kTileSmall;
  }
  g.nslots = slots_for(g.tile);
  g.ntile = (int)(g.nslots / g.tile);
  g.ntilepairs = (int64_t)g.ntile * (g.ntile + 1) / 2;
  std::vector<int32_t> perm(g.nslots, -1), ttype(g.ntile);
  {
    std::vector<int64_t> cur(cfg->ntypes, 0);
    int64_t base = 0;
    for (int t = 0; t < cfg->ntypes; ++t) {
      cur[t] = base;
      for (int64_t k = base / g.tile; k < (base + (count[t] + g.tile - 1) / g.tile * g.tile) / g.tile; ++k)
        ttype[k] = t;
      base += (count[t] + g.tile - 1) / g.tile * g.tile;
    }
    for (int64_t j = 0; j < n; ++j) perm[cur[cfg->type_id[j] - 1]++] = (int32_t)j;
  }
  // one histogram row (type pair) in LDS; workgroups per CU as the pair
  // kernel's registers and LDS allow
  const size_t hb = (size_t)g.npu * g.nbins * 4;
  const size_t lrow = (size_t)g.nbins * 4;
  p->hist_in_lds = lrow <= kHistLds ? 1 : 0;
  const void *fn = g.tile == kTileBig
                       ? (p->hist_in_lds ? pairs_fn_c<kTileBig, true>(p->cell_kind)
                                         : pairs_fn_c<kTileBig, false>(p->cell_kind))
                       : (p->hist_in_lds ? pairs_fn_c<kTileSmall, true>(p->cell_kind)
                                         : pairs_fn_c<kTileSmall, false>(p->cell_kind));
  const size_t dyn = p->hist_in_lds ? lrow : 0;
  if (p->hist_in_lds &&
      hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn) != hipSuccess)
    return bail(sqc::set_error(SQC_HIP_ERR_HIP, "setting the histogram LDS size"));
  int per_cu = 1;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, g.tile, dyn) != hipSuccess)
    return bail(sqc::set_error(SQC_HIP_ERR_HIP, "pair kernel occupancy"));
  per_cu = std::max(1, per_cu);
  p->nblk = (int)std::max<int64_t>(1, std::min<int64_t>(g.ntilepairs, (int64_t)per_cu * ncu));
  // the workgroups' own rows with the LDS row; else rows shared by the
  // workgroups (at most 1 GiB) take the HBM atomics
  p->nrows = p->hist_in_lds ? p->nblk
                            : (int)std::max<size_t>(1, std::min<size_t>((size_t)p->nblk,
                                                                         ((size_t)1 << 30) / hb));
  // padding slots: NaN positions (no pair counts), fractions 0
  std::vector<double4> pa0(g.nslots);
  std::vector<double2> pb0(g.nslots, make_double2(0.0, 0.0));
  for (int64_t k = 0; k < g.nslots; ++k) {
    const double nan = std::nan("");
    pa0[k] = make_double4(nan, nan, nan, 0.0);
  }
  int rc;
  const size_t n3 = sizeof(double) * 3 * g.nslots;
  if ((rc = dalloc(&p->d_pos, n3, p->bytes)) || (rc = dalloc(&p->d_posw, n3, p->bytes)) ||
      (rc = dalloc(&p->d_pa, sizeof(double4) * g.nslots, p->bytes)) ||
      (rc = dalloc(&p->d_pb, sizeof(double2) * g.nslots, p->bytes)) ||
      (rc = dalloc(&p->d_perm, sizeof(int32_t) * g.nslots, p->bytes)) ||
      (rc = dalloc(&p->d_tile_type, sizeof(int32_t) * g.ntile, p->bytes)) ||
      (rc = dalloc(&p->d_part, (size_t)p->nrows * hb, p->bytes)) ||
      (rc = dalloc(&p->d_hist, sizeof(unsigned long long) * g.npu * g.nbins, p->bytes)) ||
      (rc = dalloc(&p->d_flag, sizeof(int), p->bytes)))
    return bail(rc);
  if (hipMemcpy(p->d_pa, pa0.data(), sizeof(double4) * g.nslots, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(p->d_pb, pb0.data(), sizeof(double2) * g.nslots, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(p->d_perm, perm.data(), sizeof(int32_t) * g.nslots, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(p->d_tile_type, ttype.data(), sizeof(int32_t) * g.ntile, hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemset(p->d_hist, 0, sizeof(unsigned long long) * g.npu * g.nbins) != hipSuccess ||
      hipMemset(p->d_flag, 0, sizeof(int)) != hipSuccess ||
      hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&p->ev_in, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&p->ev_out, hipEventDisableTiming) != hipSuccess ||
      hipDeviceSynchronize() != hipSuccess)
    return bail(sqc::set_error(SQC_HIP_ERR_HIP, "debye plan set-up"));
  *plan = p;
  return SQC_HIP_OK;
}

int sqc_hip_debye_plan_info(const sqc_hip_debye_plan *p, sqc_hip_debye_info *info) {
  if (!p || !info) return sqc::set_error(SQC_HIP_ERR_BAD_ARG, "null argument");
  std::memset(info, 0, sizeof(*info));
  info->nbins = p->g.nbins;
  info->ntypes = p->g.ntypes;
  for (int k = 0; k < 3; ++k) info->div[k] = (int32_t)p->g.div[k];
  info->hist_in_lds = p->hist_in_lds;
  info->workgroups = p->nblk;
  info->tile = p->g.tile;
  info->pairs_per_frame = (int64_t)p->g.natoms * (p->g.natoms - 1) / 2;
  info->device_bytes = p->bytes;
  info->nframes = p->nframes;
  return SQC_HIP_OK;
}

int sqc_hip_debye_accumulate_device(sqc_hip_debye_plan *p, const double *pos_device,
                                    int64_t nframes, void *stream) {
  if (!p || !pos_device || nframes < 0) return sqc::set_error(SQC_HIP_ERR_BAD_ARG, "bad arguments");
  DEB_TRY(hipSetDevice(p->device), "hipSetDevice");
  const hipStream_t user = (hipStream_t)stream;
  DEB_TRY(hipEventRecord(p->ev_in, user), "hipEventRecord");
  DEB_TRY(hipStreamWaitEvent(p->stream, p->ev_in, 0), "hipStreamWaitEvent");
  int rc = 0;
  for (int64_t f = 0; f < nframes && !rc; ++f) rc = debye_frame(p, pos_device + f * 3 * p->g.natoms);
  DEB_TRY(hipEventRecord(p->ev_out, p->stream), "hipEventRecord");
  DEB_TRY(hipStreamWaitEvent(user, p->ev_out, 0), "hipStreamWaitEvent");
  return rc;
}

int sqc_hip_debye_profile(sqc_hip_debye_plan *p, int enable) {
  if (!p) return sqc::set_error(SQC_HIP_ERR_BAD_ARG, "null plan");
  DEB_TRY(hipSetDevice(p->device), "hipSetDevice");
  if (int rc = collect_debye(p)) return rc;
  p->profiling = enable ? 1 : 0;
  return SQC_HIP_OK;
}

int sqc_hip_debye_stage_times(sqc_hip_debye_plan *p, double *ms, int64_t *launches, int nstage) {
  if (!p) return sqc::set_error(SQC_HIP_ERR_BAD_ARG, "null plan");
  DEB_TRY(hipSetDevice(p->device), "hipSetDevice");
  if (int rc = collect_debye(p)) return rc;
  for (int i = 0; i < nstage && i < 3; ++i) {
    if (ms) ms[i] = p->stage_ms[i];
    if (launches) launches[i] = p->stage_n[i];
    p->stage_ms[i] = 0.0;
    p->stage_n[i] = 0;
  }
  return SQC_HIP_OK;
}

int sqc_hip_debye_accumulate(sqc_hip_debye_plan *p, const double *pos, int64_t nframes) {
  if (!p || !pos || nframes < 0) return sqc::set_error(SQC_HIP_ERR_BAD_ARG, "bad arguments");
  DEB_TRY(hipSetDevice(p->device), "hipSetDevice");
  const size_t n3 = sizeof(double) * 3 * p->g.natoms;
  if (!p->d_stage)
    if (int rc = dalloc(&p->d_stage, n3, p->bytes)) return rc;
  for (int64_t f = 0; f < nframes; ++f) {
    DEB_TRY(hipStreamSynchronize(p->stream), "hipStreamSynchronize");
    DEB_TRY(hipMemcpyAsync(p->d_stage, pos + f * 3 * p->g.natoms, n3, hipMemcpyHostToDevice, p->stream),
            "upload positions");
    if (int rc = debye_frame(p, p->d_stage)) return rc;
  }
  DEB_TRY(hipStreamSynchronize(p->stream), "hipStreamSynchronize");
  return check_flag(p);
}

int sqc_hip_debye_fetch(sqc_hip_debye_plan *p, double *histogram, int64_t *nframes) {
  if (!p) return sqc::set_error(SQC_HIP_ERR_BAD_ARG, "null plan");
  DEB_TRY(hipSetDevice(p->device), "hipSetDevice");
  DEB_TRY(hipStreamSynchronize(p->stream), "hipStreamSynchronize");
  const DebyeGeo &g = p->g;
  if (histogram) {
    std::vector<unsigned long long> h((size_t)g.npu * g.nbins);
    DEB_TRY(hipMemcpy(h.data(), p->d_hist, sizeof(unsigned long long) * h.size(), hipMemcpyDeviceToHost),
            "download histogram");
    const int T = g.ntypes;
    for (int a = 0; a < T; ++a)
      for (int b = 0; b < T; ++b) {
        const int lo = std::min(a, b), hi = std::max(a, b);
        const int u = lo * T - (lo * (lo - 1)) / 2 + (hi - lo);
        const double f = a == b ? 2.0 : 1.0;  // ordered pairs: both orders of a like pair
        for (int k = 0; k < g.nbins; ++k)
          histogram[((size_t)a * T + b) * g.nbins + k] = f * (double)h[(size_t)u * g.nbins + k];
      }
  }
  if (nframes) *nframes = p->nframes;
  return check_flag(p);
}

int sqc_hip_debye_reset(sqc_hip_debye_plan *p) {
  if (!p) return sqc::set_error(SQC_HIP_ERR_BAD_ARG, "null plan");
  DEB_TRY(hipSetDevice(p->device), "hipSetDevice");
  DEB_TRY(hipStreamSynchronize(p->stream), "hipStreamSynchronize");
  DEB_TRY(hipMemset(p->d_hist, 0, sizeof(unsigned long long) * p->g.npu * p->g.nbins), "hipMemset");
  DEB_TRY(hipMemset(p->d_flag, 0, sizeof(int)), "hipMemset");
  p->nframes = 0;
  return SQC_HIP_OK;
}

int sqc_hip_debye_destroy(sqc_hip_debye_plan *p) {
  if (!p) return SQC_HIP_OK;
  (void)hipSetDevice(p->device);
  free_debye(p);
  return SQC_HIP_OK;
}

}  // extern "C"
