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
// the pair to (t_i, t_j) and (t_j, t_i).  Persistent workgroups walk the
// upper triangle of 256 x 256 atom tiles; the j tile sits in LDS, each
// thread takes one i.  For every pair the images m with |s_k + m_k| within
// rmax / width_k (s the fractional displacement) are tried, so cut-offs up
// to the smallest width are counted like the cell list counts them.  Counts
// go to a u32 histogram in LDS (integer adds: the result does not depend
// on their order), or straight to the workgroup's row in HBM when it does
// not fit; the rows are summed into the run's u64 histogram per frame.
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
  int pbc[3];
  double rmax2;          // rmax * rmax
  double inv_dr;
  int nbins, ntypes, npu;  // npu: unordered type pairs
  int64_t natoms;
  int tile;              // atoms per tile (the launch's block size)
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
  double *d_pos = nullptr, *d_posw = nullptr, *d_frac = nullptr;
  int32_t *d_type = nullptr;     // 0-based type per atom
  int *d_flag = nullptr;         // position checks of the frames (k_debye_prep)
  uint32_t *d_part = nullptr;    // [nrows][npu][nbins] per-frame counts
  unsigned long long *d_hist = nullptr;  // [npu][nbins] run histogram
  double *d_stage = nullptr;     // host-input staging, one frame
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
// ds = A^-1 (r_new - r_prev).  frac = A^-1 pos picks the pair images.
// flag bits: 1 an atom drifted 2^20 cells or more from its first image (or
// is not finite), 2 an atom lies outside the box along a non-periodic axis
__global__ void k_debye_prep(DebyeGeo g, const double *__restrict__ r, int first,
                             double *__restrict__ pos, double *__restrict__ posw,
                             double *__restrict__ frac, int *__restrict__ flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.natoms) return;
  double rel[3], s[3], rn[3];
  for (int k = 0; k < 3; ++k) rel[k] = r[3 * i + k] - g.origin[k];
  mat3v(g.inv, rel, s);
  for (int k = 0; k < 3; ++k) s[k] = s[k] - floor(s[k]);
  mat3v(g.a, s, rn);
  double p[3];
  if (first) {
    for (int k = 0; k < 3; ++k) p[k] = rn[k];
  } else {
    double dr[3], ds[3];
    for (int k = 0; k < 3; ++k) dr[k] = rn[k] - posw[3 * i + k];
    mat3v(g.inv, dr, ds);
    for (int k = 0; k < 3; ++k) ds[k] = ds[k] - round(ds[k]);  // anint: halves away from 0
    mat3v(g.a, ds, dr);
    for (int k = 0; k < 3; ++k) p[k] = pos[3 * i + k] + dr[k];
  }
  double f[3];
  mat3v(g.inv, p, f);
  int bad = 0;
  for (int k = 0; k < 3; ++k) {
    pos[3 * i + k] = p[k];
    posw[3 * i + k] = rn[k];
    frac[3 * i + k] = f[k];
    if (!(fabs(f[k]) < kFarCells)) bad |= 1;
    else if (!g.pbc[k] && (f[k] < 0.0 || f[k] >= 1.0)) bad |= 2;
  }
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

template <int kTile, bool Lds>
__global__ void __launch_bounds__(kTile) k_debye_pairs(DebyeGeo g, const double *__restrict__ pos,
                                                       const double *__restrict__ frac,
                                                       const int32_t *__restrict__ type,
                                                       uint32_t *__restrict__ part, int nrows) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double sp[kTile][3], sf[kTile][3];
  __shared__ int st[kTile];
  const int tid = threadIdx.x;
  const int nh = g.npu * g.nbins;
  // Lds: the workgroup's own row, written once at the end; otherwise rows
  // shared by workgroups (zeroed before the launch) take the atomics
  uint32_t *row = part + (size_t)(blockIdx.x % nrows) * nh;
  uint32_t *hist = Lds ? reinterpret_cast<uint32_t *>(smem) : row;
  if (Lds)
    for (int k = tid; k < nh; k += kTile) hist[k] = 0u;
  const double rb0 = g.reach_frac[0] + 1e-9, rb1 = g.reach_frac[1] + 1e-9,
               rb2 = g.reach_frac[2] + 1e-9;
  for (int64_t idx = blockIdx.x; idx < g.ntilepairs; idx += gridDim.x) {
    // upper-triangle tile pair (ti <= tj) of index idx = tj (tj + 1) / 2 + ti
    int64_t tj = (int64_t)((sqrt(8.0 * (double)idx + 1.0) - 1.0) * 0.5);
    while ((tj + 1) * (tj + 2) / 2 <= idx) ++tj;
    while (tj * (tj + 1) / 2 > idx) --tj;
    const int64_t ti = idx - tj * (tj + 1) / 2;
    __syncthreads();  // the previous tile's reads of sp/sf/st are done
    {
      const int64_t j = tj * kTile + tid;
      if (j < g.natoms) {
        for (int k = 0; k < 3; ++k) {
          sp[tid][k] = pos[3 * j + k];
          sf[tid][k] = frac[3 * j + k];
        }
        st[tid] = type[j];
      }
    }
    __syncthreads();
    const int64_t i = ti * kTile + tid;
    if (i >= g.natoms) continue;
    const double pi0 = pos[3 * i], pi1 = pos[3 * i + 1], pi2 = pos[3 * i + 2];
    const double fi0 = frac[3 * i], fi1 = frac[3 * i + 1], fi2 = frac[3 * i + 2];
    const int ta = type[i];
    const int nj = (int)std::min<int64_t>(kTile, g.natoms - tj * kTile);
    const int j0 = ti == tj ? tid + 1 : 0;
    for (int jl = j0; jl < nj; ++jl) {
      // fractional displacement and the images within reach, per axis
      const double s0 = sf[jl][0] - fi0, s1 = sf[jl][1] - fi1, s2 = sf[jl][2] - fi2;
      if (!(fabs(s0) < 2.0 * kFarCells && fabs(s1) < 2.0 * kFarCells &&
            fabs(s2) < 2.0 * kFarCells))
        continue;  // reported by the position pass: keeps the image loops bounded
      int lo0 = 0, hi0 = 0, lo1 = 0, hi1 = 0, lo2 = 0, hi2 = 0;
      if (g.pbc[0]) { lo0 = (int)ceil(-s0 - rb0); hi0 = (int)floor(-s0 + rb0); }
      if (g.pbc[1]) { lo1 = (int)ceil(-s1 - rb1); hi1 = (int)floor(-s1 + rb1); }
      if (g.pbc[2]) { lo2 = (int)ceil(-s2 - rb2); hi2 = (int)floor(-s2 + rb2); }
      const double e0 = sp[jl][0] - pi0, e1 = sp[jl][1] - pi1, e2 = sp[jl][2] - pi2;
      for (int m2 = lo2; m2 <= hi2; ++m2)
        for (int m1 = lo1; m1 <= hi1; ++m1)
          for (int m0 = lo0; m0 <= hi0; ++m0) {
            // bins . shift with shift = div * m, summed k = 1, 2, 3
            const double c0 = g.div[0] * m0, c1 = g.div[1] * m1, c2 = g.div[2] * m2;
            const double v0 = g.bins[0] * c0 + g.bins[3] * c1 + g.bins[6] * c2;
            const double v1 = g.bins[1] * c0 + g.bins[4] * c1 + g.bins[7] * c2;
            const double v2 = g.bins[2] * c0 + g.bins[5] * c1 + g.bins[8] * c2;
            const double d0 = e0 + v0, d1 = e1 + v1, d2 = e2 + v2;
            const double dd = d0 * d0 + d1 * d1 + d2 * d2;
            if (!(dd <= g.rmax2)) continue;
            const int bin = (int)(sqrt(dd) * g.inv_dr);
            if (bin >= g.nbins) continue;
            const int k = upair(ta, st[jl], g.ntypes) * g.nbins + bin;
            if (Lds)
              atomicAdd(&hist[k], 1u);
            else
              atomicAdd(&row[k], 1u);
          }
    }
  }
  if (!Lds) return;
  __syncthreads();
  for (int k = tid; k < nh; k += kTile) row[k] = hist[k];
}

__global__ void k_debye_zero(size_t n, uint32_t *__restrict__ part) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) part[k] = 0u;
}

// The workgroups' rows of one frame added into the run histogram.
__global__ void k_debye_reduce(int nblk, int nh, const uint32_t *__restrict__ part,
                               unsigned long long *__restrict__ hist) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nh) return;
  unsigned long long s = 0;
  for (int b = 0; b < nblk; ++b) s += part[(size_t)b * nh + k];
  hist[k] += s;
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
  void *bufs[] = {p->d_pos,  p->d_posw, p->d_frac,  p->d_type,
                  p->d_part, p->d_hist, p->d_stage, p->d_flag};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  if (p->stream) {
    (void)hipStreamSynchronize(p->stream);
    (void)hipStreamDestroy(p->stream);
  }
  delete p;
}

int debye_frame(sqc_hip_debye_plan *p, const double *pos_dev) {
  DebyeGeo &g = p->g;
  const unsigned ga = (unsigned)((g.natoms + 255) / 256);
  hipLaunchKernelGGL(k_debye_prep, dim3(ga), dim3(256), 0, p->stream, g, pos_dev,
                     p->started ? 0 : 1, p->d_pos, p->d_posw, p->d_frac, p->d_flag);
  p->started = 1;
  const int nh = g.npu * g.nbins;
  const dim3 grid(p->nblk), block(g.tile);
  if (p->hist_in_lds) {
    const size_t lds = (size_t)nh * 4;
    if (g.tile == kTileBig)
      hipLaunchKernelGGL((k_debye_pairs<kTileBig, true>), grid, block, lds, p->stream, g, p->d_pos,
                         p->d_frac, p->d_type, p->d_part, p->nrows);
    else
      hipLaunchKernelGGL((k_debye_pairs<kTileSmall, true>), grid, block, lds, p->stream, g,
                         p->d_pos, p->d_frac, p->d_type, p->d_part, p->nrows);
  } else {
    const size_t n = (size_t)p->nrows * nh;
    hipLaunchKernelGGL(k_debye_zero, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, p->stream, n,
                       p->d_part);
    if (g.tile == kTileBig)
      hipLaunchKernelGGL((k_debye_pairs<kTileBig, false>), grid, block, 0, p->stream, g, p->d_pos,
                         p->d_frac, p->d_type, p->d_part, p->nrows);
    else
      hipLaunchKernelGGL((k_debye_pairs<kTileSmall, false>), grid, block, 0, p->stream, g,
                         p->d_pos, p->d_frac, p->d_type, p->d_part, p->nrows);
  }
  hipLaunchKernelGGL(k_debye_reduce, dim3((nh + 255) / 256), dim3(256), 0, p->stream, p->nrows,
                     nh, p->d_part, p->d_hist);
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
  {
    const int64_t nt = (cfg->natoms + kTileBig - 1) / kTileBig;
    g.tile = nt * (nt + 1) / 2 >= 4 * (int64_t)ncu ? kTileBig : kTileSmall;
  }
  g.ntile = (int)((cfg->natoms + g.tile - 1) / g.tile);
  g.ntilepairs = (int64_t)g.ntile * (g.ntile + 1) / 2;
  const size_t hb = (size_t)g.npu * g.nbins * 4;
  p->hist_in_lds = hb <= kHistLds ? 1 : 0;
  // workgroups per CU: LDS (the j tile's positions, fractions and types, plus
  // the histogram) and waves (at most 28 of the pair kernel's per CU)
  const size_t lds_static = (size_t)g.tile * (6 * sizeof(double) + sizeof(int));
  const size_t lds_wg = lds_static + (p->hist_in_lds ? hb : 0);
  const int by_waves = std::max(1, 28 / (g.tile / 64));
  const int per_cu = std::max(1, std::min(by_waves, (int)((160 * 1024) / lds_wg)));
  p->nblk = (int)std::max<int64_t>(1, std::min<int64_t>(g.ntilepairs, (int64_t)per_cu * ncu));
  // without the LDS histogram, rows shared by the workgroups (at most 1 GiB)
  p->nrows = p->hist_in_lds ? p->nblk
                            : (int)std::max<size_t>(1, std::min<size_t>((size_t)p->nblk,
                                                                         ((size_t)1 << 30) / hb));
  if (p->hist_in_lds) {
    const void *fn = g.tile == kTileBig ? (const void *)k_debye_pairs<kTileBig, true>
                                        : (const void *)k_debye_pairs<kTileSmall, true>;
    if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)hb) != hipSuccess)
      return bail(sqc::set_error(SQC_HIP_ERR_HIP, "setting the histogram LDS size"));
  }
  std::vector<int32_t> t0(cfg->natoms);
  for (int64_t j = 0; j < cfg->natoms; ++j) t0[j] = cfg->type_id[j] - 1;
  int rc;
  const size_t n3 = sizeof(double) * 3 * cfg->natoms;
  if ((rc = dalloc(&p->d_pos, n3, p->bytes)) || (rc = dalloc(&p->d_posw, n3, p->bytes)) ||
      (rc = dalloc(&p->d_frac, n3, p->bytes)) ||
      (rc = dalloc(&p->d_type, sizeof(int32_t) * cfg->natoms, p->bytes)) ||
      (rc = dalloc(&p->d_part, (size_t)p->nrows * hb, p->bytes)) ||
      (rc = dalloc(&p->d_hist, sizeof(unsigned long long) * g.npu * g.nbins, p->bytes)) ||
      (rc = dalloc(&p->d_flag, sizeof(int), p->bytes)))
    return bail(rc);
  if (hipMemcpy(p->d_type, t0.data(), sizeof(int32_t) * cfg->natoms, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(p->d_hist, 0, sizeof(unsigned long long) * g.npu * g.nbins) != hipSuccess ||
      hipMemset(p->d_flag, 0, sizeof(int)) != hipSuccess ||
      hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess ||
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
  info->pairs_per_frame = (int64_t)p->g.natoms * (p->g.natoms - 1) / 2;
  info->device_bytes = p->bytes;
  info->nframes = p->nframes;
  return SQC_HIP_OK;
}

int sqc_hip_debye_accumulate_device(sqc_hip_debye_plan *p, const double *pos_device,
                                    int64_t nframes) {
  if (!p || !pos_device || nframes < 0) return sqc::set_error(SQC_HIP_ERR_BAD_ARG, "bad arguments");
  DEB_TRY(hipSetDevice(p->device), "hipSetDevice");
  for (int64_t f = 0; f < nframes; ++f)
    if (int rc = debye_frame(p, pos_device + f * 3 * p->g.natoms)) return rc;
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
