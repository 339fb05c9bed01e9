// sqc_kernels.hip -- gfx950 kernels of the reciprocal-space S(q) path.
//
// Forward, per frame (one launch each):
//   k_prep        wrap the atoms into the cell (FP64): stencil start, the
//                 3 x ns kernel values and the tile bin of every atom
//   k_scan        exclusive scan of the bin counts (one workgroup)
//   k_scatter     counting-sort scatter of atom indices by bin
//   k_binsort     order each bin by atom index (makes the spread deterministic)
//   k_spread_z    per (x,y) tile: stage the atoms that reach the tile in LDS,
//                 spread them into LDS columns, z-DFT, write h3 >= 0
//   k_fft_y       per (h3, x-chunk): y-DFT in LDS, write the canonical lines
//   k_fft_x_bin   per group of lines: x-DFT in LDS, deconvolve, combine species,
//                 shell sums in LDS, per-workgroup partials
//   k_finalize    ordered, compensated sum of the workgroup partials
// Adjoint (type-2 with the kernel derivative): k_adj_x, k_adj_y,
// k_adj_z_gather, k_adj_sum.
//
// All DFTs carry the +i sign of the reference transform
// (src/sqc_structure_factor.f90:1069-1071, iflag = +1).  A line of n points
// is transformed by one wavefront, in place in LDS: mixed-radix Stockham
// stages whose butterflies are staged through registers; the lanes of a wave
// execute in lockstep and LDS serves one wave in order, so a stage needs no
// workgroup barrier.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <algorithm>
#include <type_traits>

#include "sqc_dft.hpp"
#include "sqc_internal.hpp"

namespace sqc {

constexpr int NT = kThreads;  // threads per workgroup for every LDS kernel
constexpr int NW = NT / 64;

// Keeps the compiler from moving LDS accesses across this point; the wave
// itself already executes its LDS instructions in order.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One Stockham stage of a line of n points (contiguous in LDS) by one wave.
template <typename T, int R>
__device__ __forceinline__ void wave_stage(typename C2<T>::t *line, int n, int Ns, uint32_t magic,
                                           const typename C2<T>::t *tw, int lane) {
  using V = typename C2<T>::t;
  constexpr int MAXB = (kMaxLine + R * 64 - 1) / (R * 64);
  const int nb = n / R;
  const int tstep = n / (Ns * R);
  V v[MAXB][R];
  int qq[MAXB], kk[MAXB];
#pragma unroll
  for (int it = 0; it < MAXB; ++it) {
    const int j = lane + 64 * it;
    if (j < nb) {
      const int q = magic ? (int)__umulhi((unsigned)j, magic) : j;
      const int k = j - (int)__umul24(q, Ns);
      qq[it] = q;
      kk[it] = k;
#pragma unroll
      for (int r = 0; r < R; ++r) v[it][r] = line[(j + r * nb)];
      if (Ns > 1) {
        const int kt = (int)__umul24(k, tstep);
        int ti = kt;
#pragma unroll
        for (int r = 1; r < R; ++r, ti += kt) v[it][r] = cmul(v[it][r], tw[ti]);
      }
    }
  }
  wave_sync();
#pragma unroll
  for (int it = 0; it < MAXB; ++it) {
    const int j = lane + 64 * it;
    if (j < nb) {
      Dft<T, R>::run(v[it]);
      const int d0 = (int)__umul24(qq[it], Ns * R) + kk[it];
#pragma unroll
      for (int r = 0; r < R; ++r) line[(d0 + r * Ns)] = v[it][r];
    }
  }
  wave_sync();
}

template <typename T>
__device__ void wave_fft_any(typename C2<T>::t *line, const AxisPlan &ax,
                             const typename C2<T>::t *tw, int lane) {
  int Ns = 1;
  for (int s = 0; s < ax.nst; ++s) {
    const int R = ax.rad[s];
    switch (R) {
      case 2: wave_stage<T, 2>(line, ax.n, Ns, ax.magic[s], tw, lane); break;
      case 3: wave_stage<T, 3>(line, ax.n, Ns, ax.magic[s], tw, lane); break;
      case 4: wave_stage<T, 4>(line, ax.n, Ns, ax.magic[s], tw, lane); break;
      case 5: wave_stage<T, 5>(line, ax.n, Ns, ax.magic[s], tw, lane); break;
      default: wave_stage<T, 8>(line, ax.n, Ns, ax.magic[s], tw, lane); break;
    }
    Ns *= R;
  }
}

// Lines longer than kMaxLine hold too many points for the in-place stages'
// registers.  Their stages run out of place between the
// line and a scratch line of the wave's own, one butterfly per lane at a
// time; the result is copied back when it ends in the scratch.
template <typename T, int R>
__device__ void long_stage(const typename C2<T>::t *src, typename C2<T>::t *dst, int n, int Ns,
                           uint32_t magic, const typename C2<T>::t *tw, int lane) {
  using V = typename C2<T>::t;
  const int nb = n / R;
  const int tstep = n / (Ns * R);
  for (int j = lane; j < nb; j += 64) {
    const int q = magic ? (int)__umulhi((unsigned)j, magic) : j;
    const int k = j - q * Ns;
    V v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = src[j + r * nb];
    if (Ns > 1) {
      const int kt = k * tstep;
      int ti = kt;
#pragma unroll
      for (int r = 1; r < R; ++r, ti += kt) v[r] = cmul(v[r], tw[ti]);
    }
    Dft<T, R>::run(v);
    const int d0 = q * Ns * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) dst[d0 + r * Ns] = v[r];
  }
  wave_sync();
}

template <typename T>
__device__ void wave_fft_long(typename C2<T>::t *line, typename C2<T>::t *scratch,
                              const AxisPlan &ax, const typename C2<T>::t *tw, int lane) {
  using V = typename C2<T>::t;
  V *a = line, *b = scratch;
  int Ns = 1;
  wave_sync();
  for (int s = 0; s < ax.nst; ++s) {
    const int R = ax.rad[s];
    switch (R) {
      case 2: long_stage<T, 2>(a, b, ax.n, Ns, ax.magic[s], tw, lane); break;
      case 3: long_stage<T, 3>(a, b, ax.n, Ns, ax.magic[s], tw, lane); break;
      case 4: long_stage<T, 4>(a, b, ax.n, Ns, ax.magic[s], tw, lane); break;
      case 5: long_stage<T, 5>(a, b, ax.n, Ns, ax.magic[s], tw, lane); break;
      default: long_stage<T, 8>(a, b, ax.n, Ns, ax.magic[s], tw, lane); break;
    }
    V *t = a;
    a = b;
    b = t;
    Ns *= R;
  }
  if (a != line) {
    for (int i = lane; i < ax.n; i += 64) line[i] = a[i];
    wave_sync();
  }
}

// Compile-time line DFT for the fine-grid sizes the benchmark workloads use:
// the radix plan, strides and divisors are constants, so the LDS accesses of
// a butterfly become one address plus immediate offsets.
struct RadixPlan {
  int nst;
  int rad[kMaxStages];
};

constexpr RadixPlan radix_plan(int n) {
  RadixPlan p{0, {}};
  const int order[5] = {8, 5, 4, 3, 2};
  for (int i = 0; i < 5; ++i)
    while (n % order[i] == 0) {
      p.rad[p.nst++] = order[i];
      n /= order[i];
    }
  return p;
}

// Two lines per wave: when N / 15 is a multiple of 32, a first radix of 15
// has 32 butterflies per line, so one pass of the wave (lanes 0-31 on the
// first line, 32-63 on the second) does both lines' first stage; the rest of
// N follows the greedy plan.  For N = 480 that is 15, 8, 4: three stages
// instead of 8, 5, 4, 3, about a fifth fewer vector and a quarter fewer LDS
// instructions per line.
constexpr RadixPlan radix_plan_lines(int n, int lines) {
  if (lines != 2 || n % 15 != 0 || (n / 15) % 32 != 0) return radix_plan(n);
  const RadixPlan r = radix_plan(n / 15);
  RadixPlan p{r.nst + 1, {}};
  p.rad[0] = 15;
  for (int i = 0; i < r.nst; ++i) p.rad[i + 1] = r.rad[i];
  return p;
}

// The first stage writes element e = R j + r (stride R across the lanes),
// which falls on few LDS banks.  When R divides 32 and the second stage reads
// with a stride that is a multiple of 32, the two stages exchange the data in
// a padded layout, e -> e + e / 32: both index maps stay a per-lane base plus
// compile-time offsets, so the padding costs no instructions.  Line buffers
// hold N + N / 32 + 1 elements.
template <int N, int L> struct StagePad {
  static constexpr RadixPlan P = radix_plan_lines(N, L);
  static constexpr bool on =
      P.nst >= 2 && 32 % P.rad[0] == 0 && (N / P.rad[1]) % 32 == 0;
};

// A first stage whose butterflies of all L lines fill one pass of the wave
// exactly (L * N / R == 64): lane l * N / R + j does butterfly j of line l.
template <typename T, int N, int R, int L>
__device__ __forceinline__ void split_first_stage(typename C2<T>::t *const *ln, int lane) {
  using V = typename C2<T>::t;
  constexpr int nb = N / R;
  static_assert(L * nb == 64, "one pass of the wave");
  const int l = lane / nb, j = lane % nb;
  V *line = ln[0];
#pragma unroll
  for (int i = 1; i < L; ++i)
    if (l == i) line = ln[i];
  V v[R];
#pragma unroll
  for (int r = 0; r < R; ++r) v[r] = line[j + r * nb];
  wave_sync();
  Dft<T, R>::run(v);
#pragma unroll
  for (int r = 0; r < R; ++r) line[j * R + r] = v[r];
  wave_sync();
}

// One stage over L lines of the same length at once: the lines' LDS reads
// are issued together and each twiddle is read once for all of them.
template <typename T, int N, int R, int Ns, bool InPad, bool OutPad, int L>
__device__ __forceinline__ void fixed_stage(typename C2<T>::t *const *ln,
                                            const typename C2<T>::t *tw, int lane) {
  using V = typename C2<T>::t;
  constexpr int nb = N / R;
  constexpr int MAXB = (nb + 63) / 64;
  constexpr int tstep = N / (Ns * R);
  constexpr int rstride = InPad ? nb + nb / 32 : nb;
  static_assert(!InPad || nb % 32 == 0, "padded reads need a stride that is a multiple of 32");
  static_assert(!OutPad || (Ns == 1 && 32 % R == 0), "padded writes are for a first stage");
  V v[L][MAXB][R];
#pragma unroll
  for (int it = 0; it < MAXB; ++it) {
    const int j = lane + 64 * it;
    if (it + 1 < MAXB || nb % 64 == 0 || j < nb) {
#pragma unroll
      for (int l = 0; l < L; ++l) {
        const V *src = ln[l] + j + (InPad ? (j >> 5) : 0);
#pragma unroll
        for (int r = 0; r < R; ++r) v[l][it][r] = src[r * rstride];
      }
      if constexpr (Ns > 1) {
        const int k = j % Ns;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          const V w = tw[k * (r * tstep)];
#pragma unroll
          for (int l = 0; l < L; ++l) v[l][it][r] = cmul(v[l][it][r], w);
        }
      }
    }
  }
  wave_sync();
#pragma unroll
  for (int it = 0; it < MAXB; ++it) {
    const int j = lane + 64 * it;
    if (it + 1 < MAXB || nb % 64 == 0 || j < nb) {
#pragma unroll
      for (int l = 0; l < L; ++l) {
        Dft<T, R>::run(v[l][it]);
        if constexpr (OutPad) {
          V *dst = ln[l] + j * R + ((j * R) >> 5);
#pragma unroll
          for (int r = 0; r < R; ++r) dst[r] = v[l][it][r];
        } else {
          V *dst = ln[l] + (j / Ns) * (Ns * R) + j % Ns;
#pragma unroll
          for (int r = 0; r < R; ++r) dst[r * Ns] = v[l][it][r];
        }
      }
    }
  }
  wave_sync();
}

template <typename T, int N, int S, int Ns, int L>
__device__ __forceinline__ void fixed_stages(typename C2<T>::t *const *ln,
                                             const typename C2<T>::t *tw, int lane) {
  constexpr RadixPlan P = radix_plan_lines(N, L);
  if constexpr (S < P.nst) {
    constexpr int R = P.rad[S];
    constexpr bool pad = StagePad<N, L>::on;
    if constexpr (S == 0 && L > 1 && L * (N / R) == 64)
      split_first_stage<T, N, R, L>(ln, lane);
    else
      fixed_stage<T, N, R, Ns, pad && S == 1, pad && S == 0, L>(ln, tw, lane);
    fixed_stages<T, N, S + 1, Ns * R, L>(ln, tw, lane);
  }
}

// Line DFT of the plan's length: compile-time when N > 0 (N == ax.n),
// otherwise the runtime-planned stages; lines longer than kMaxLine go
// through `scratch` (a line of the wave's own).
template <typename T, int N>
__device__ __forceinline__ void wave_fft(typename C2<T>::t *line, const AxisPlan &ax,
                                         const typename C2<T>::t *tw, int lane,
                                         typename C2<T>::t *scratch) {
  if constexpr (N > 0) {
    typename C2<T>::t *ln[1] = {line};
    fixed_stages<T, N, 0, 1, 1>(ln, tw, lane);
  } else {
    if (ax.n > kMaxLine)
      wave_fft_long<T>(line, scratch, ax, tw, lane);
    else
      wave_fft_any<T>(line, ax, tw, lane);
  }
}

// The DFTs of lines first, first + NW, ... < nlines (lines at base + i *
// stride) by one wave.  FP32 lines of a compile-time length go two at a time.
template <typename T, int N>
__device__ __forceinline__ void wave_fft_lines(typename C2<T>::t *base, int stride, int first,
                                               int nlines, const AxisPlan &ax,
                                               const typename C2<T>::t *tw, int lane,
                                               typename C2<T>::t *scratch) {
  int i = first;
  if constexpr (N > 0 && sizeof(T) == 4) {
    for (; i + NW < nlines; i += 2 * NW) {
      typename C2<T>::t *ln[2] = {base + (size_t)i * stride, base + (size_t)(i + NW) * stride};
      fixed_stages<T, N, 0, 1, 2>(ln, tw, lane);
    }
  }
  for (; i < nlines; i += NW) wave_fft<T, N>(base + (size_t)i * stride, ax, tw, lane, scratch);
}

// ---------------------------------------------------------------------------
// ES kernel phi(u) = exp(beta (sqrt(1 - z^2) - 1)), z = 2 u / ns, and dphi/du.
// The argument is formed in FP64 (u comes from FP64 grid coordinates); the
// exponential runs in the path precision.
// ---------------------------------------------------------------------------
template <typename T> __device__ __forceinline__ T es_exp(double a);
template <> __device__ __forceinline__ float es_exp<float>(double a) { return __expf((float)a); }
template <> __device__ __forceinline__ double es_exp<double>(double a) { return exp(a); }

// dphi bounds sqrt(1 - z^2) below by smin, which caps the derivative's
// singular term at the support edge.
template <typename T>
__device__ __forceinline__ T es_phi(double u, double inv_half_ns, double beta, double smin,
                                    T *dphi) {
  const double z = u * inv_half_ns;  // 2 u / ns
  const double q = 1.0 - z * z;
  if (q <= 0.0) {
    if (dphi) *dphi = T(0);
    return T(0);
  }
  const double sq = sqrt(q);
  const T ph = es_exp<T>(beta * (sq - 1.0));
  if (dphi) *dphi = (T)((double)ph * (-beta * z / fmax(sq, smin)) * inv_half_ns);
  return ph;
}

__device__ __forceinline__ int wrapi(int i, int n) {
  i %= n;
  return i < 0 ? i + n : i;
}

// Overlap of the stencil [i0, i0 + ns) (periodic in n) with [t0, t0 + tw):
// returns the count and sets the first stencil offset and the first local index.
__device__ __forceinline__ int overlap(int i0, int ns, int n, int t0, int tw, int *k_lo, int *l_lo) {
#pragma unroll
  for (int sh = -1; sh <= 1; ++sh) {
    const int a = i0 + sh * n;
    const int lo = max(a, t0), hi = min(a + ns, t0 + tw);
    if (lo < hi) {
      *k_lo = lo - a;
      *l_lo = lo - t0;
      return hi - lo;
    }
  }
  return 0;
}

// Copy total elements with U loads in flight per thread: load(idx) reads
// global memory, store(idx, v) writes LDS.  Loading every element before
// storing any keeps the loads of a thread independent of each other.
template <int U, typename V, typename Load, typename Store>
__device__ __forceinline__ void batched_copy(int total, Load load, Store store) {
  // the loads of a batch are unconditional (the index is clamped to a valid
  // element) so that they issue back to back; only the stores are guarded
  if (total <= 0) return;
  for (int base = 0; base < total; base += U * NT) {
    V tmp[U];
#pragma unroll
    for (int u = 0; u < U; ++u) tmp[u] = load(min(base + u * NT + (int)threadIdx.x, total - 1));
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int idx = base + u * NT + (int)threadIdx.x;
      if (idx < total) store(idx, tmp[u]);
    }
  }
}

// The hardware deals consecutive workgroups to the XCDs in turn.  Renumber
// them so that every XCD works through one contiguous range of logical
// blocks: neighbouring x chunks of the y pass, which read the two halves of
// the same cache lines, then share one L2 (without it the y pass fetched
// 2.5x its bytes from HBM).  A bijection of [0, nb) for any nb and XCD count.
__device__ __forceinline__ int xcd_block(int i, int nb, int nxcd) {
  const int xcd = i % nxcd, k = i / nxcd;
  const int q = nb / nxcd, r = nb % nxcd;
  return xcd < r ? xcd * (q + 1) + k : r * (q + 1) + (xcd - r) * q + k;
}

// q = c / d for small non-negative c, d through a float reciprocal; the 0.5
// offset keeps c / d at least 1 / (2 d) away from an integer.
__device__ __forceinline__ int small_div(int c, float inv_d) { return (int)(((float)c + 0.5f) * inv_d); }

// c / D for a non-negative c: by the constant D when it is known at compile
// time (D > 0), else through the float reciprocal inv_d of the run-time d.
template <int D> __device__ __forceinline__ int div_by(int c, float inv_d) {
  if constexpr (D > 0)
    return (int)((unsigned)c / (unsigned)D);
  else
    return small_div(c, inv_d);
}

// ---------------------------------------------------------------------------
// Atom preparation and counting sort
// ---------------------------------------------------------------------------
// One thread per kernel value (atom j, axis i, point k): consecutive threads
// write consecutive values of the per-atom [3][ns] table.  The thread of
// point 0 of axis 0 also writes the atom's stencil corner and bin key.
template <typename T>
__global__ void k_prep(Geo g, const double *__restrict__ pos, const int32_t *__restrict__ species,
                       int4 *__restrict__ i0o, T *__restrict__ phi, T *__restrict__ dphi,
                       int32_t *__restrict__ key, int32_t *__restrict__ count) {
  const int ns = g.ns, per = 3 * ns;
  const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;  // < natoms * 3 ns < 2^31
  // j = t / per and ia = v / ns by reciprocals: the FP64 quotient of
  // t < 2^31 is off by at most one and corrected; v < 48 divides exactly
  // in FP32
  unsigned j = (unsigned)((double)t * g.inv_per);
  if (j * (unsigned)per > t) --j;
  else if ((j + 1) * (unsigned)per <= t) ++j;
  if (j >= (unsigned)g.natoms) return;
  const int v = (int)(t - j * per);
  const int ia = small_div(v, (float)g.inv_half_ns * 0.5f), k = v - ia * ns;
  const double d0 = pos[3 * (size_t)j] - g.origin[0];
  const double d1 = pos[3 * (size_t)j + 1] - g.origin[1];
  const double d2 = pos[3 * (size_t)j + 2] - g.origin[2];
  // fractional coordinate, wrapped, and grid coordinate of axis i
  auto grid_coord = [&](int i) {
    double s = g.ainv[i] * d0 + g.ainv[i + 3] * d1 + g.ainv[i + 6] * d2;
    s = s - floor(s);
    double u = s * (double)g.n[i];
    if (u >= (double)g.n[i]) u -= (double)g.n[i];
    return u;
  };
  const double u = grid_coord(ia);
  const int i0 = (int)ceil(u - g.half_ns);
  T d;
  phi[t] = es_phi<T>((double)(i0 + k) - u, g.inv_half_ns, g.beta, g.dphi_smin,
                     dphi ? &d : (T *)nullptr);
  if (dphi) dphi[t] = d;
  if (v == 0) {
    double x[3];
    int c[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      x[i] = grid_coord(i);
      c[i] = (int)ceil(x[i] - g.half_ns);
    }
    i0o[j] = int4{c[0], c[1], c[2], 0};
    const int tx = min((int)x[0] / g.txr, g.ntx - 1);
    const int ty = min((int)x[1] / g.ty, g.nty - 1);
    const int f = (g.F == 1) ? 0 : species[j];
    const int b = (f * g.nty + ty) * g.ntx + tx;
    key[j] = b;
    atomicAdd(&count[b], 1);
  }
}

// Exclusive scan of the nb bin counts by one workgroup: each wave owns a
// contiguous run of 4-count groups, read coalesced (one group per lane per
// round); wave totals are combined through LDS, then every wave scans its
// run with lane shuffles.  The counts are cleared for the next frame.
__global__ void __launch_bounds__(1024) k_scan(int32_t *__restrict__ count, int32_t *__restrict__ start,
                                               int32_t *__restrict__ cursor, int nb) {
  __shared__ int wsum[16];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int ng = (nb + 3) >> 2;
  const int seg = (ng + 15) >> 4;
  const int g0 = min(ng, wave * seg), g1 = min(ng, g0 + seg);
  auto load4 = [&](int gi) {
    if (4 * gi + 3 < nb) return reinterpret_cast<const int4 *>(count)[gi];
    int4 c{0, 0, 0, 0};
    const int b = 4 * gi;
    if (b < nb) c.x = count[b];
    if (b + 1 < nb) c.y = count[b + 1];
    if (b + 2 < nb) c.z = count[b + 2];
    return c;
  };
  int acc = 0;
  for (int gi = g0 + lane; gi < g1; gi += 64) {
    const int4 c = load4(gi);
    acc += c.x + c.y + c.z + c.w;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
  if (lane == 0) wsum[wave] = acc;
  __syncthreads();
  int carry = 0, total = 0;
#pragma unroll
  for (int w = 0; w < 16; ++w) {
    carry += w < wave ? wsum[w] : 0;
    total += wsum[w];
  }
  if (tid == 0) start[nb] = total;
  for (int base = g0; base < g1; base += 64) {
    const int gi = base + lane;
    const int4 c = gi < g1 ? load4(gi) : int4{0, 0, 0, 0};
    const int sum4 = c.x + c.y + c.z + c.w;
    int incl = sum4;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int y = __shfl_up(incl, off);
      if (lane >= off) incl += y;
    }
    const int o = carry + incl - sum4;
    const int4 r{o, o + c.x, o + c.x + c.y, o + c.x + c.y + c.z};
    if (gi < g1) {
      const int b = 4 * gi;
      if (b + 3 < nb) {
        reinterpret_cast<int4 *>(start)[gi] = r;
        reinterpret_cast<int4 *>(cursor)[gi] = r;
        reinterpret_cast<int4 *>(count)[gi] = int4{0, 0, 0, 0};  // ready for the next frame
      } else {
        const int rv[3] = {r.x, r.y, r.z};
        for (int e = 0; e < 3 && b + e < nb; ++e) {
          start[b + e] = rv[e];
          cursor[b + e] = rv[e];
          count[b + e] = 0;
        }
      }
    }
    carry += __shfl(incl, 63);
  }
}

__global__ void k_scatter(int64_t natoms, const int32_t *__restrict__ key, int32_t *__restrict__ cursor,
                          int32_t *__restrict__ sidx) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= natoms) return;
  const int p = atomicAdd(&cursor[key[j]], 1);
  sidx[p] = (int32_t)j;
}

// Order every bin by atom index: one wave per bin; bins of up to 64 atoms
// are ranked in registers, larger ones fall back to an insertion sort.
__global__ void k_binsort(int nb, const int32_t *__restrict__ start, int32_t *__restrict__ sidx,
                          const int4 *__restrict__ i0, int4 *__restrict__ i0s) {
  const int b = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (b >= nb) return;
  const int k0 = start[b], k1 = start[b + 1];
  const int n = k1 - k0;
  if (n == 0) return;
  if (n <= 64) {
    const int v = lane < n ? sidx[k0 + lane] : 0x7fffffff;
    int rank = 0;
    for (int i = 0; i < n; ++i) {
      const int w = __shfl(v, i);
      rank += (w < v) ? 1 : 0;
    }
    if (lane < n) {
      sidx[k0 + rank] = v;
      int4 c = i0[v];  // the stencil corners in sorted order, .w the atom
      c.w = v;
      i0s[k0 + rank] = c;
    }
    return;
  }
  // larger bins: the same ranks, counted 64 candidates at a time, with the
  // corners written to i0s in sorted order (.w holds the atom index, which
  // then goes back to sidx), so no element is overwritten before it is read
  for (int e0 = 0; e0 < n; e0 += 64) {
    const int e = e0 + lane;
    const int v = e < n ? sidx[k0 + e] : 0x7fffffff;
    int rank = 0;
    for (int s0 = 0; s0 < n; s0 += 64) {
      const int w = s0 + lane < n ? sidx[k0 + s0 + lane] : 0x7fffffff;
      const int m = min(64, n - s0);
      for (int i = 0; i < m; ++i) rank += (__shfl(w, i) < v) ? 1 : 0;
    }
    if (e < n) {
      int4 c = i0[v];
      c.w = v;
      i0s[k0 + rank] = c;
    }
  }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  for (int k = k0 + lane; k < k1; k += 64) sidx[k] = i0s[k].w;
}

// Bins to visit around a tile, without visiting a bin twice on small grids.
__device__ __forceinline__ void bin_window(int t, int h, int nt, int *first, int *count) {
  if (2 * h + 1 >= nt) {
    *first = 0;
    *count = nt;
  } else {
    *first = t - h;
    *count = 2 * h + 1;
  }
}

// ---------------------------------------------------------------------------
// Candidate staging shared by the spreader and the gather.  The atoms whose
// stencil reaches the tile are compacted into LDS in bin order (a block scan
// of the overlap flags), so every round is deterministic.
//   rec.x = atom, rec.y = kx | xl << 8 | ox << 16, rec.z = ky | yl << 8 | oy << 16,
//   rec.w = first z point of the stencil, wrapped into [0, n3), | species << 16
//           (single-field spreads, which weight each atom) | z-slab mask << 24
//           (the spreader)
// ---------------------------------------------------------------------------
constexpr int kMaxWindowBins = 128;

struct Stage {
  int nbin;                           // bins in the window
  int total;                          // candidate atoms in the window
  int bin_first[kMaxWindowBins];      // sorted index of each bin's first atom
  int bin_pref[kMaxWindowBins + 1];   // candidates before each bin
  int wsum[NW];
  int nstaged;
};

// The bins of a tile's window: thread t < nb holds bin t's first sorted
// index and its atom count.
struct Window {
  int nb, s0, cnt;
};
__device__ __forceinline__ Window window_load(const Geo &g, int tx, int ty, int f,
                                              const int32_t *__restrict__ start) {
  int bxs, bxn, bys, byn;
  bin_window(tx, g.hx, g.ntx, &bxs, &bxn);
  bin_window(ty, g.hy, g.nty, &bys, &byn);
  Window w{bxn * byn, 0, 0};
  const int t = threadIdx.x;
  if (t < w.nb) {
    const int by = wrapi(bys + t / bxn, g.nty);
    const int bx = wrapi(bxs + t % bxn, g.ntx);
    const int bin = (f * g.nty + by) * g.ntx + bx;
    w.s0 = start[bin];
    w.cnt = start[bin + 1] - w.s0;
  }
  return w;
}
__device__ void stage_window(Stage &st, const Window &w) {
  const int nb = w.nb;
  const int t = threadIdx.x;
  if (t < nb) {
    st.bin_first[t] = w.s0;
    st.bin_pref[t + 1] = w.cnt;
  }
  __syncthreads();
  if (t < 64) {
    // inclusive scan of the counts by one wave, two bins per lane
    // (nb <= kMaxWindowBins = 128)
    const int b0 = 2 * t, b1 = b0 + 1;
    const int c0 = b0 < nb ? st.bin_pref[b0 + 1] : 0;
    const int c1 = b1 < nb ? st.bin_pref[b1 + 1] : 0;
    int sum = c0 + c1;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int y = __shfl_up(sum, off);
      if (t >= off) sum += y;
    }
    const int excl = sum - c0 - c1;
    if (b0 < nb) st.bin_pref[b0 + 1] = excl + c0;
    if (b1 < nb) st.bin_pref[b1 + 1] = excl + c0 + c1;
    if (t == 63) st.total = sum;
    if (t == 0) {
      st.bin_pref[0] = 0;
      st.nbin = nb;
      st.nstaged = 0;
    }
  }
  __syncthreads();
}
__device__ void stage_window(Stage &st, const Geo &g, int tx, int ty, int f,
                             const int32_t *__restrict__ start) {
  stage_window(st, window_load(g, tx, ty, f, start));
}

// Stage candidates [base, base + NT) of the window behind the ones already
// staged; returns (block-uniform) the number staged so far.
__device__ int stage_chunk(Stage &st, int4 *rec, int base, const Geo &g, int x0, int txw, int y0,
                           int tyw, const int32_t *__restrict__ sidx,
                           const int4 *__restrict__ i0s,
                           const int32_t *__restrict__ species = nullptr, int slabs = 0) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int t = base + tid;
  int flag = 0;
  int4 r = int4{0, 0, 0, 0};
  if (t < st.total) {
    // the last bin whose first candidate is at or before t
    int b = 0, hi = st.nbin - 1;
    while (b < hi) {
      const int mid = (b + hi + 1) >> 1;
      if (st.bin_pref[mid] <= t) b = mid;
      else hi = mid - 1;
    }
    const int k = st.bin_first[b] + (t - st.bin_pref[b]);
    const int j = sidx[k];
    const int4 i0 = i0s[k];  // sorted copy: independent of the sidx load
    int kx = 0, xl = 0, ky = 0, yl = 0;
    const int ox = overlap(i0.x, g.ns, g.n[0], x0, txw, &kx, &xl);
    const int oy = ox ? overlap(i0.y, g.ns, g.n[1], y0, tyw, &ky, &yl) : 0;
    if (ox && oy) {
      flag = 1;
      // rec.w: first z point (wrapped, < 2^16) | species << 16 when asked
      // for | mask of the z slabs the stencil reaches << 24 when asked for
      const int sp = species ? species[j] : 0;
      const int n3 = g.n[2], ns = g.ns;
      const int a = wrapi(i0.z, n3);
      int mask = 0;
      for (int w = 0; w < slabs; ++w) {
        const int z0 = (n3 * w) / slabs, z1 = (n3 * (w + 1)) / slabs;
        const bool hit = (max(a, z0) < min(a + ns, z1)) || (a + ns > n3 && z0 < a + ns - n3);
        mask |= hit ? (1 << w) : 0;
      }
      r = int4{j, kx | (xl << 8) | (ox << 16), ky | (yl << 8) | (oy << 16),
               a | (sp << 16) | (mask << 24)};
    }
  }
  const unsigned long long m = __ballot(flag);
  const int before = __popcll(m & ((1ULL << lane) - 1ULL));
  if (lane == 0) st.wsum[wave] = __popcll(m);
  __syncthreads();
  int off = st.nstaged;
  for (int w = 0; w < wave; ++w) off += st.wsum[w];
  if (flag) rec[off + before] = r;
  int total = st.nstaged;
  for (int w = 0; w < NW; ++w) total += st.wsum[w];
  __syncthreads();
  if (tid == 0) st.nstaged = total;
  return total;
}


// Add the z points [z0, z0 + pc) of one atom's stencil into the tile: cell
// c = xy * pc + kz, two cells per lane per round with both reads issued
// before the writes (the cells of one atom are distinct).  fz of point kz is
// lane fzl + kz of ph; fxy and the column offset of xy cell xy are lane xy
// of fxy and cofs.
template <typename T>
__device__ __forceinline__ void spread_piece(T *tflat, T ph, T fxy, int cofs, int oxy, int fzl,
                                             int z0, int pc, int lane) {
  if (pc <= 0) return;
  const int cells = oxy * pc;
  const float inv_pc = __builtin_amdgcn_rcpf((float)pc);
  for (int c0 = 0; c0 < cells; c0 += 128) {
    const int ca = c0 + lane, cb = ca + 64;
    const int va = ca < cells ? ca : 0, vb = cb < cells ? cb : 0;
    const int xa = small_div(va, inv_pc), xb = small_div(vb, inv_pc);
    const int za = va - xa * pc, zb = vb - xb * pc;
    const T fa = __shfl(fxy, xa) * __shfl(ph, fzl + za);
    const T fb = __shfl(fxy, xb) * __shfl(ph, fzl + zb);
    const int oa = __shfl(cofs, xa) + 2 * (z0 + za);
    const int ob = __shfl(cofs, xb) + 2 * (z0 + zb);
    const bool wa = ca < cells, wb = cb < cells;
    T ra = T(0), rb = T(0);
    if (wa) ra = tflat[oa];
    if (wb) rb = tflat[ob];
    if (wa) tflat[oa] = ra + fa;
    if (wb) tflat[ob] = rb + fb;
  }
}

// Rotation by R lanes within each row of 16 lanes (DPP row_ror), and the
// sum over a row that it gives every lane of the row.
template <int R> __device__ __forceinline__ double row_ror(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, 0x120 + R, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x120 + R, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double row16_sum(double v) {
  v += row_ror<8>(v);
  v += row_ror<4>(v);
  v += row_ror<2>(v);
  v += row_ror<1>(v);
  return v;
}

// Lane K of each row of 16 lanes, broadcast to the row (DPP row_newbcast).
template <int K> __device__ __forceinline__ float row_bcast(float v) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x150 + K, 0xF, 0xF, false));
}
template <int K> __device__ __forceinline__ double row_bcast(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, 0x150 + K, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x150 + K, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// f(integral_constant<int, K>) for K = B .. E - 1, in order.
template <int B, int E, typename F> __device__ __forceinline__ void static_for(F &&f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>());
    static_for<B + 1, E>(f);
  }
}

// Lane shuffle through ds_bpermute with a byte address (4 * source lane)
// computed once by the caller.
__device__ __forceinline__ float bperm(int addr4, float v) {
  return __int_as_float(__builtin_amdgcn_ds_bpermute(addr4, __float_as_int(v)));
}
__device__ __forceinline__ double bperm(int addr4, double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_ds_bpermute(addr4, (int)b);
  const int hi = __builtin_amdgcn_ds_bpermute(addr4, (int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Cells of a whole stencil (all ns z points in the wave's slab) go through a
// map fixed per lane: cell c = lane + 64 h is xy cell c / ns, z point c % ns.
// Up to kWholeRounds rounds of 64 cells (ox * oy * ns <= 192).
constexpr int kWholeRounds = 3;
// Registers of kernel values a spreading wave loads at once.
constexpr int kPhiRegs = 4;

struct WholeMap {
  int xy[kWholeRounds];    // xy cell
  int fz4[kWholeRounds];   // 4 * (2 ns + z point) (bpermute address of fz in ph)
  int z2[kWholeRounds];    // 2 * z point (offset in the interleaved tile)
};

__device__ __forceinline__ WholeMap whole_map(int ns, int lane) {
  WholeMap m;
  const float inv_ns = __builtin_amdgcn_rcpf((float)ns);
#pragma unroll
  for (int h = 0; h < kWholeRounds; ++h) {
    const int c = lane + 64 * h;
    const int xy = small_div(c, inv_ns), z = c - xy * ns;
    m.xy[h] = xy;
    m.fz4[h] = 4 * (2 * ns + z);
    m.z2[h] = 2 * z;
  }
  return m;
}

// Add a whole stencil (ox x oy columns from tile column (xll, yll), kernel
// points kxl, kyl on, z points a .. a + ns - 1) into the tile.  Each lane
// finds its cell's column and LDS offset arithmetically and picks its three
// kernel values out of ph (lanes po ..) itself, so the LDS reads of a round
// depend on nothing but the record.  The product is formed as
// ((w fx) fy) fz, the order of the partial-stencil path.
template <typename T>
__device__ __forceinline__ void spread_whole(T *tflat, T ph, T w, int ox, int kxl, int xll,
                                             int kyl, int yll, int ns, int ccx, int cs, int cells,
                                             int a, int po, const WholeMap &m, int lane) {
  const float inv_ox = __builtin_amdgcn_rcpf((float)ox);
  const int kx4 = 4 * (po + kxl), ky4 = 4 * (po + ns + kyl), po4 = 4 * po, a2 = 2 * a;
#pragma unroll
  for (int h = 0; h < kWholeRounds; ++h) {
    if (64 * h < cells) {
      const int xy = m.xy[h];
      const int iy = small_div(xy, inv_ox), ix = xy - iy * ox;
      const int xl = xll + ix, yl = yll + iy;
      const int o = 2 * (yl * ccx + (xl >> 1)) * cs + (xl & 1) + a2 + m.z2[h];
      const bool on = lane + 64 * h < cells;
      T r = T(0);
      if (on) r = tflat[o];
      const T f = w * bperm(kx4 + 4 * ix, ph) * bperm(ky4 + 4 * iy, ph) * bperm(m.fz4[h] + po4, ph);
      if (on) tflat[o] = r + f;
    }
  }
}

// ---------------------------------------------------------------------------
// Spread + z-DFT.  One workgroup per (x,y) tile of txr x ty real columns and
// one field.  The tile is held in LDS as txr/2 x ty complex columns of n3
// points (stride cs): real column 2p in .x, 2p+1 in .y, so one complex DFT
// transforms two real columns.  Wave w owns the z slab [w n3/4, (w+1) n3/4)
// and walks the staged atoms in order, so every LDS cell is updated by one
// wave, in atom order: no atomics, and the grid is bit-reproducible.
// ---------------------------------------------------------------------------
template <typename T, int N, int TT>
__global__ void __launch_bounds__(NT) k_spread_z(Geo g, AxisPlan az, const int4 *__restrict__ i0s,
                                                 const T *__restrict__ phi,
                                                 const int32_t *__restrict__ sidx,
                                                 const int32_t *__restrict__ start,
                                                 const int32_t *__restrict__ species,
                                                 const typename C2<T>::t *__restrict__ twz,
                                                 typename C2<T>::t *__restrict__ W1,
                                                 const double *__restrict__ wtab) {
  using V = typename C2<T>::t;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ Stage st;
  const int n1 = g.n[0], n2 = g.n[1], n3 = g.n[2];
  // the column stride is a constant for the compiled line lengths
  const int cs = N > 0 ? spread_col_stride(N) : g.cs;
  // the tile shape: compiled in for the common 8 x 2 tile (TT = 1), read
  // at run time otherwise
  constexpr bool kt = TT == 1;
  const int txr = kt ? 8 : g.txr, tty = kt ? 2 : g.ty;
  const int ltxr = kt ? 3 : g.ltxr, ltyy = kt ? 1 : g.lty;
  const int ccx = txr / 2, cc = ccx * tty;
  V *tile = reinterpret_cast<V *>(smem);
  int4 *rec = reinterpret_cast<int4 *>(tile + (size_t)cc * cs);
  T *tflat = reinterpret_cast<T *>(tile);
  // grid (tiles, fields) in hardware order: measured faster here than the
  // XCD-contiguous renumbering the y pass uses
  const int f = blockIdx.y, tb = blockIdx.x;
  const int tx = tb % g.ntx, ty = tb / g.ntx;
  const int x0 = tx * txr, y0 = ty * tty;
  const int txw = min(txr, n1 - x0), tyw = min(tty, n2 - y0);
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  // the window's bin loads are issued before the tile clear, which they
  // overlap
  const Window win = window_load(g, tx, ty, f, start);
  {
    // clear the tile 16 bytes a store (cc * cs * sizeof(V) is a multiple of
    // 16: cs is even)
    float4 *t4 = reinterpret_cast<float4 *>(tile);
    const int n4 = cc * cs * (int)sizeof(V) / 16;
    for (int i = tid; i < n4; i += NT) t4[i] = float4{0.f, 0.f, 0.f, 0.f};
  }
  stage_window(st, win);

  const int zs0 = (n3 * wave) / NW, zs1 = (n3 * (wave + 1)) / NW;
  const int ns = g.ns;
  const WholeMap wm = whole_map(ns, lane);
  // spread one staged atom into this wave's slab; ph: lane l < 3 ns holds
  // its kernel value l (x, then y, then z)
  // (lanes po .. po + 3 ns - 1 of ph: po is 0 or 32)
  auto spread_atom = [&](const int4 r, const T ph, const int po) {
    const int kxl = r.y & 255, xll = (r.y >> 8) & 255, ox = r.y >> 16;
    const int kyl = r.z & 255, yll = (r.z >> 8) & 255, oy = r.z >> 16;
    const int a = r.w & 0xFFFF;
    // lane l < ox*oy holds the xy weight and the LDS offset of xy cell l
    // (ox <= txr and oy <= ty with txr * ty <= 32, so ox*oy <= 64)
    T w = T(1);
    if (g.F == 1) {
      const int sp = (r.w >> 16) & 0xFF;
      w = g.S > kGeoSpecies ? (T)wtab[sp * kWtab] : (T)g.amp[sp];
    }
    const int cells = ox * oy * ns;
    if (a >= zs0 && a + ns <= zs1 && cells <= 64 * kWholeRounds) {
      spread_whole<T>(tflat, ph, w, ox, kxl, xll, kyl, yll, ns, ccx, cs, cells, a, po, wm, lane);
      return;
    }
    T fxy;
    int cofs;
    {
      const int oxy = ox * oy;
      const int l = lane < oxy ? lane : 0;
      const int iy = small_div(l, __builtin_amdgcn_rcpf((float)ox));
      const int ix = l - iy * ox;
      fxy = w * __shfl(ph, po + kxl + ix) * __shfl(ph, po + ns + kyl + iy);
      const int xl = xll + ix, yl = yll + iy;
      cofs = 2 * (yl * ccx + (xl >> 1)) * cs + (xl & 1);
    }
    // the stencil's z range inside this wave's slab: one piece, or two
    // when it wraps through z = 0
    const int e1 = min(a + ns, n3);
    spread_piece<T>(tflat, ph, fxy, cofs, ox * oy, po + 2 * ns + max(a, zs0) - a, max(a, zs0),
                    min(e1, zs1) - max(a, zs0), lane);
    if (a + ns > n3) {
      const int lo2 = zs0, hi2 = min(a + ns - n3, zs1);
      spread_piece<T>(tflat, ph, fxy, cofs, ox * oy, po + 2 * ns + n3 - a + lo2, lo2, hi2 - lo2,
                      lane);
    }
  };

  for (int base = 0; base < st.total; base += NT) {
    const int nst = stage_chunk(st, rec, base, g, x0, txw, y0, tyw, sidx, i0s,
                                g.F == 1 ? species : nullptr, NW);
    // this wave's atoms in staging order: a ballot on its slab bit per 64
    // records, then the set bits in groups whose kernel values are loaded
    // together (A atoms per register, at lanes 0 and 32 when A = 2)
    auto walk = [&](auto per_reg) {
      constexpr int A = decltype(per_reg)::value;
      // registers of kernel values per group: eight atoms a group either way
      constexpr int PR = A == 2 ? kPhiRegs : 2 * kPhiRegs;
      for (int b0 = 0; b0 < nst; b0 += 64) {
        // every lane holds one record of the block; an atom's fields are
        // read from its lane (no LDS round trip per atom)
        const int idx = b0 + lane;
        const int4 mr = idx < nst ? rec[idx] : int4{0, 0, 0, 0};
        unsigned long long m = __ballot((mr.w >> (24 + wave)) & 1);
        // the next group's atoms (lanes of mr) and their kernel values: a
        // group's loads are issued before the previous group is spread
        auto take = [&](int (&ub)[A * PR], T (&pb)[PR]) {
#pragma unroll
          for (int i = 0; i < A * PR; ++i) {
            ub[i] = m ? __ffsll((long long)m) - 1 : -1;
            m &= m - 1;
          }
#pragma unroll
          for (int q = 0; q < PR; ++q) {
            const int u0 = ub[A * q], u1 = ub[A * q + A - 1];
            const int j0 = u0 >= 0 ? __builtin_amdgcn_readlane(mr.x, u0) : 0;
            const int j1 = u1 >= 0 ? __builtin_amdgcn_readlane(mr.x, u1) : 0;
            const bool hi = A == 2 && lane >= 32;
            const int u = hi ? u1 : u0, j = hi ? j1 : j0;
            const int kv = A == 2 ? (lane & 31) : lane;
            pb[q] = T(0);
            // 32-bit offsets: natoms * 3 ns < 2^31 (checked by the plan)
            if (u >= 0 && kv < 3 * ns) pb[q] = phi[(unsigned)(j * 3 * ns + kv)];
          }
        };
        if (!m) continue;
        int ub[A * PR], un[A * PR];
        T pb[PR], pn[PR];
        take(ub, pb);
        while (true) {
          const bool more = m != 0;
          if (more) take(un, pn);
          // the taken atoms come first in ub, the rest are -1
#pragma unroll
          for (int i = 0; i < A * PR; ++i) {
            const int u = ub[i];
            if (u >= 0) {
              const int4 r{0, __builtin_amdgcn_readlane(mr.y, u),
                           __builtin_amdgcn_readlane(mr.z, u), __builtin_amdgcn_readlane(mr.w, u)};
              spread_atom(r, pb[i / A], (i % A) * 32);
            }
          }
          if (!more) break;
          for (int i = 0; i < A * PR; ++i) ub[i] = un[i];
          for (int q = 0; q < PR; ++q) pb[q] = pn[q];
        }
      }
    };
    if (3 * ns <= 32)
      walk(std::integral_constant<int, 2>());
    else
      walk(std::integral_constant<int, 1>());
    __syncthreads();  // the next chunk overwrites rec from its start
    if (tid == 0) st.nstaged = 0;
    __syncthreads();
  }
  __syncthreads();
  // long lines: the wave's scratch line lies past the tile (the staged
  // records are not in use during the DFT)
  wave_fft_lines<T, N>(tile, cs, wave, cc, az, twz, lane,
                       tile + (size_t)cc * cs + (size_t)wave * n3);
  __syncthreads();
  // split the packed pairs and write h3 = 0..M3: a thread keeps one complex
  // column (two real columns, adjacent in x) and steps h3, writing both real
  // columns' values with one store.  (txr / 2) * ty is a power of two that
  // divides NT, and txw is even (n1 and x0 are).
  const int M3 = g.M[2];
  const int lper = ltxr - 1 + ltyy;
  const int xp = tid & (ccx - 1), yl = (tid >> (ltxr - 1)) & (tty - 1);
  if (2 * xp >= txw || yl >= tyw) return;
  const int kstep = NT >> lper;
  const V *colp = tile + (yl * ccx + xp) * cs;
  const size_t plane = (size_t)n2 * n1;
  V *out = W1 + ((size_t)f * (M3 + 1) * n2 + (y0 + yl)) * n1 + x0 + 2 * xp;
  const T h = T(0.5);
  for (int k3 = tid >> lper; k3 <= M3; k3 += kstep) {
    const V Z = colp[k3];
    const V Zm = colp[k3 ? n3 - k3 : 0];
    // even column (Z + conj Zm) / 2, odd column (Z - conj Zm) / (2i)
    const V e{h * (Z.x + Zm.x), h * (Z.y - Zm.y)};
    const V o{h * (Z.y + Zm.y), h * (Zm.x - Z.x)};
    V *dst = out + (size_t)k3 * plane;
    // W1 is written once here and read once by the y pass; it is larger
    // than the L2s and the Infinity Cache together, so the stores bypass
    // them (non-temporal): 0.401 -> 0.373 ms for this kernel, and the y
    // pass that follows it 0.256 -> 0.248 ms
    if constexpr (sizeof(T) == 4) {
      // 16-byte aligned: the element offset is even
      typedef float nf4 __attribute__((ext_vector_type(4)));
      __builtin_nontemporal_store(nf4{e.x, e.y, o.x, o.y}, reinterpret_cast<nf4 *>(dst));
    } else {
      typedef double nd2 __attribute__((ext_vector_type(2)));
      __builtin_nontemporal_store(nd2{e.x, e.y}, reinterpret_cast<nd2 *>(dst));
      __builtin_nontemporal_store(nd2{o.x, o.y}, reinterpret_cast<nd2 *>(dst + 1));
    }
  }
}

// ---------------------------------------------------------------------------
// y-DFT.  One workgroup per (x-chunk, h3, field); lines are transposed into
// contiguous LDS rows (stride n2 + 1) and transformed one per wave.
// ---------------------------------------------------------------------------
// TX2: the column chunk at compile time (the tuner picks it among 2, 4, 8
// and 16), so that a thread's column and row steps are constants.

// A non-temporal load of one complex value, for data read once and larger
// than the caches.
template <typename V> __device__ __forceinline__ V nt_load(const V *p) {
  V v;
  if constexpr (sizeof(V) == 8) {
    const unsigned long long b =
        __builtin_nontemporal_load(reinterpret_cast<const unsigned long long *>(p));
    __builtin_memcpy(&v, &b, 8);
  } else {
    typedef double nd2 __attribute__((ext_vector_type(2)));
    const nd2 b = __builtin_nontemporal_load(reinterpret_cast<const nd2 *>(p));
    v.x = b.x;
    v.y = b.y;
  }
  return v;
}
template <typename T, int N, int TX2>
__global__ void __launch_bounds__(NT) k_fft_y(Geo g, AxisPlan ay,
                                              const typename C2<T>::t *__restrict__ W1,
                                              const typename C2<T>::t *__restrict__ twy,
                                              const int32_t *__restrict__ lo2,
                                              const int32_t *__restrict__ hi2,
                                              const int32_t *__restrict__ loff,
                                              typename C2<T>::t *__restrict__ W2) {
  using V = typename C2<T>::t;
  extern __shared__ __align__(16) unsigned char smem[];
  const int n1 = g.n[0], n2 = g.n[1], M3 = g.M[2], ls = g.ls2;
  V *lines = reinterpret_cast<V *>(smem);
  constexpr int tx2 = TX2, lt = TX2 == 16 ? 4 : TX2 == 8 ? 3 : TX2 == 4 ? 2 : 1, tm = TX2 - 1;
  static_assert(TX2 == 2 || TX2 == 4 || TX2 == 8 || TX2 == 16, "column chunk");
  const int nbx = (g.n[0] + tx2 - 1) / tx2;
  const int blk = xcd_block(blockIdx.x, gridDim.x, g.nxcd);
  const int x0 = (blk % nbx) * tx2, k3 = (blk / nbx) % (g.M[2] + 1), f = blk / (nbx * (g.M[2] + 1));
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int lo = lo2[k3], hi = hi2[k3];
  if (hi < lo) return;
  const V *src = W1 + ((size_t)f * (M3 + 1) + k3) * n2 * n1;
  batched_copy<8, V>(
      n2 * tx2,
      [&](int idx) {
        const int x = min(x0 + (idx & tm), n1 - 1), y = idx >> lt;
        return src[(unsigned)(y * n1 + x)];
      },
      [&](int idx, V v) { lines[(idx & tm) * ls + (idx >> lt)] = v; });
  __syncthreads();
  wave_fft_lines<T, N>(lines, ls, wave, tx2, ay, twy, lane,
                       lines + (size_t)tx2 * ls + (size_t)wave * g.n[1]);
  __syncthreads();
  const int nl = hi - lo + 1;
  V *dst = W2 + ((size_t)f * g.nlines + loff[k3]) * n1;
  for (int idx = tid; idx < nl * tx2; idx += NT) {
    const int xl = idx & tm, r = idx >> lt;
    const int x = x0 + xl;
    if (x >= n1) continue;
    const int h2 = lo + r;
    const int yi = h2 < 0 ? h2 + n2 : h2;
    dst[(size_t)r * n1 + x] = lines[xl * ls + yi];
  }
}

// ---------------------------------------------------------------------------
// x-DFT, deconvolution and shell sums.  Persistent workgroups walk groups of
// lb lines; every field of a line is transformed together so that the pair
// products of one mode are formed in registers.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double xray_f(const double *c, double q) {
  const double st = q * (1.0 / (4.0 * 3.14159265358979323846));
  const double st2 = st * st;
  double f = c[8];
#pragma unroll
  for (int i = 0; i < 4; ++i) f += c[i] * exp(-c[4 + i] * st2);
  return f;
}

__device__ __forceinline__ double qlen_of(const Geo &g, int h1, int h2, int h3) {
  const double h1d = h1, h2d = h2, h3d = h3;
  const double qx = g.b[0] * h1d + g.b[3] * h2d + g.b[6] * h3d;
  const double qy = g.b[1] * h1d + g.b[4] * h2d + g.b[7] * h3d;
  const double qz = g.b[2] * h1d + g.b[5] * h2d + g.b[8] * h3d;
  return sqrt(qx * qx + qy * qy + qz * qz);
}

// Per-group line tables staged in LDS by the x-pass kernels.
// Shell indices a thread of the x pass prefetches per line group (groups
// hold up to kPreShell * NT candidates before the rest is read directly).
constexpr int kPreShell = 8;

struct LineGroup {
  int off[kMaxLb + 1];     // candidate offset of each line (and the end)
  int h[kMaxLb][3];        // h2, h3, lo1
  double d23[kMaxLb];      // 1 / (phihat_2(|h2|) phihat_3(h3))
};

// The line table of a group, read by the first nl + 1 threads: issue() starts
// the loads, put() stores them in LDS once they are needed, so the loads can
// be in flight together with the rows'.
struct GroupLoad {
  int4 info;
  int off;
  double d23;
  __device__ __forceinline__ void issue(int l0, int nl, const int4 *__restrict__ line_info,
                                        const int32_t *__restrict__ line_off,
                                        const double *__restrict__ line_d23) {
    const int tid = threadIdx.x;
    if (tid <= nl) off = line_off[l0 + tid];
    if (tid < nl) {
      info = line_info[l0 + tid];
      d23 = line_d23[l0 + tid];
    }
  }
  __device__ __forceinline__ void put(LineGroup &lg, int nl) const {
    const int tid = threadIdx.x;
    if (tid <= nl) lg.off[tid] = off;
    if (tid < nl) {
      lg.h[tid][0] = info.x;
      lg.h[tid][1] = info.y;
      lg.h[tid][2] = info.z;
      lg.d23[tid] = d23;
    }
  }
};

__device__ __forceinline__ void load_group(LineGroup &lg, int l0, int nl,
                                           const int4 *__restrict__ line_info,
                                           const int32_t *__restrict__ line_off,
                                           const double *__restrict__ line_d23) {
  GroupLoad gl;
  gl.issue(l0, nl, line_info, line_off, line_d23);
  gl.put(lg, nl);
}

// Exact shell sums.  Every value v added into a shell accumulator becomes a
// 64-bit integer count of the plan's quantum q (a power of two): hi =
// rint(v / q).  FP64 plans keep a second word for the remainder, lo =
// rint((v / q - hi) 2^k).  The host picks q from a bound on |v| and on the
// number of values per shell, and k from the latter, so that no sum of his
// or of los of one shell can leave 63 bits: integer additions are exact, so
// a shell's sums do not depend on the order in which lanes, waves or
// workgroups add them (Demmel and Nguyen's reproducible summation rounds to
// a common quantum the same way).  A value is then carried to q / 2 (one
// word) or q 2^-k / 2 (two words).  The rounding to an integer is one FP64
// add: for |x| < 2^51, x + 1.5 2^52 lies in [2^52, 2^53), where the spacing
// is 1, so it is rint(x) plus the constant, and its bits minus the
// constant's bits are rint(x) as an integer.  v / q, the remainder and its
// scaling are exact (powers of two, |v / q - hi| <= 1/2); contraction is off
// so that each rounding happens as written.
struct Parts {
  long long hi, lo;
};
__device__ __forceinline__ Parts exact_parts(double v, double inv_q, double lo_scale, bool two) {
#pragma clang fp contract(off)
  constexpr double kMagic = 6755399441055744.0;  // 1.5 * 2^52
  const long long mbits = __double_as_longlong(kMagic);
  const double x = v * inv_q;
  const double t = x + kMagic;
  Parts p{__double_as_longlong(t) - mbits, 0};
  if (two) {
    const double y = (x - (t - kMagic)) * lo_scale;
    p.lo = __double_as_longlong(y + kMagic) - mbits;
  }
  return p;
}

// FC: the number of fields when the epilogue is compiled for it (0: g.F at
// run time, at most kGeoSpecies, with the species amplitudes of a mode in
// registers; -1: any g.F, the amplitudes re-read from the LDS rows and the
// weights from the plan's table).  With more than one field, fields and
// species coincide.  LB: the lines per group when compiled in (2, the common
// tuned value; 0: lb_).
template <typename T, int N, int FC, int LB>
__global__ void __launch_bounds__(NT) k_fft_x_bin(Geo g, AxisPlan ax1, int lb_, int ngroups,
                                                  const typename C2<T>::t *__restrict__ W2,
                                                  const typename C2<T>::t *__restrict__ twx,
                                                  const int4 *__restrict__ line_info,
                                                  const int32_t *__restrict__ line_off,
                                                  const uint16_t *__restrict__ stab,
                                                  const double *__restrict__ ph1,
                                                  const double *__restrict__ line_d23,
                                                  unsigned long long *__restrict__ wgpart,
                                                  typename C2<T>::t *__restrict__ rho_out,
                                                  int64_t ncand, double *__restrict__ gnum,
                                                  double *__restrict__ gpart,
                                                  const double *__restrict__ wtab) {
  using V = typename C2<T>::t;
  const int lb = LB > 0 ? LB : lb_;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ LineGroup lg;
  const int n1 = g.n[0], F = FC > 0 ? FC : g.F, nq = g.nq, ncomp = g.ncomp, M1 = g.M[0];
  constexpr int FM = FC > 0 ? FC : kGeoSpecies;
  const int rs = g.rs1;
  V *rows = reinterpret_cast<V *>(smem);
  V *tw = rows + (size_t)lb * F * rs;
  // exact shell sums (exact_parts): the hi words of the ncomp x nq sums,
  // then (FP64 plans) their lo words, in LDS, or, when they do not fit
  // there, in this workgroup's own row of wgpart (zeroed here; the group
  // loop's barrier orders that before the first add)
  // (two pointers, not one selected at run time, so that the LDS adds stay
  // LDS instructions)
  // only the run-time field-count instance keeps sums in HBM (the host
  // routes every such plan there), so the compiled instances keep pure LDS adds
  const bool gacc = FC <= 0 && g.acc_global != 0;
  const int ntot = ncomp * nq, nw = g.xwords * ntot;
  unsigned long long *acc = reinterpret_cast<unsigned long long *>(tw + n1);
  unsigned long long *accg = wgpart + (size_t)blockIdx.x * nw;
  // 1 / phihat_1(|h1|)
  double *inv1 = reinterpret_cast<double *>(gacc ? acc : acc + nw);
  // scratch lines of the long-line DFT, past the 1 / phihat_1 table
  V *lscr = reinterpret_cast<V *>(
      smem + ((reinterpret_cast<unsigned char *>(inv1 + M1 + 1) - smem + 15) & ~(ptrdiff_t)15));
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  if (gacc)
    for (int i = tid; i < nw; i += NT) accg[i] = 0ull;
  else
    for (int i = tid; i < nw; i += NT) acc[i] = 0ull;
  // address spaces spelled out: the two adds must not merge into one generic
  // (flat) atomic on a selected pointer
  typedef __attribute__((address_space(3))) unsigned long long lds_u64;
  typedef __attribute__((address_space(1))) unsigned long long hbm_u64;
  // quantum of the total (component 0) and of the pair sums
  const double iq0 = g.xq_inv[0], iq1 = g.xq_inv[1], los = g.xq_lo;
  const bool two = g.xwords == 2;
  auto add = [&](int i, double v) {
    const Parts pt = exact_parts(v, i < nq ? iq0 : iq1, los, two);
    if (gacc) {
      __hip_atomic_fetch_add((hbm_u64 *)(accg + i), (unsigned long long)pt.hi, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
      if (two)
        __hip_atomic_fetch_add((hbm_u64 *)(accg + ntot + i), (unsigned long long)pt.lo,
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      __hip_atomic_fetch_add((lds_u64 *)(acc + i), (unsigned long long)pt.hi, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_WORKGROUP);
      if (two)
        __hip_atomic_fetch_add((lds_u64 *)(acc + ntot + i), (unsigned long long)pt.lo,
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
  };
  for (int i = tid; i < n1; i += NT) tw[i] = twx[i];
  for (int i = tid; i <= M1; i += NT) inv1[i] = 1.0 / ph1[i];
  for (int grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
    const int l0 = grp * lb;
    const int nl = min(lb, g.nlines - l0);
    __syncthreads();
    GroupLoad gl;
    gl.issue(l0, nl, line_info, line_off, line_d23);
    const float inv_n1 = 1.0f / (float)n1, inv_f = 1.0f / (float)F;
    batched_copy<8, V>(
        nl * F * n1,
        [&](int idx) {
          const int r = div_by<N>(idx, inv_n1), x = idx - r * n1;
          const int li = div_by<FC>(r, inv_f), f = r - li * F;
          // W2 is read once (non-temporal: 0.210 -> 0.208 ms)
          return nt_load(W2 + (size_t)(f * g.nlines + l0 + li) * n1 + x);
        },
        [&](int idx, V v) {
          const int r = div_by<N>(idx, inv_n1);
          rows[r * rs + (idx - r * n1)] = v;
        });
    gl.put(lg, nl);
    __syncthreads();
    const int c0 = lg.off[0], c1 = lg.off[nl];
    // shell indices of this thread's candidates, in flight during the DFT
    uint16_t sv[kPreShell];
    if (c1 > c0) {
#pragma unroll
      for (int k = 0; k < kPreShell; ++k) sv[k] = stab[min(c0 + tid + k * NT, c1 - 1)];
    }
    // one line at a time: pairs raise the register count past three waves per SIMD
    for (int r = wave; r < nl * F; r += NW)
      wave_fft<T, N>(rows + (size_t)r * rs, ax1, tw, lane, lscr + (size_t)wave * n1);
    __syncthreads();
    // candidate c of line li, kept in shell s
    // the factor on the pair sums (1 whenever there are two or more fields)
    const double ps = FC == 2 ? 1.0 : g.pair_scale;
    // any number of fields: the amplitudes are read from the rows again for
    // the pair sums, and the weights come from the plan's table
    auto bin_mode_any = [&](int c, int li, unsigned s) {
      const int h2 = lg.h[li][0], h3 = lg.h[li][1];
      const int h1 = lg.h[li][2] + (c - lg.off[li]);
      const double d = lg.d23[li] * inv1[abs(h1)];
      const int ix = h1 < 0 ? h1 + n1 : h1;
      const V *rp = rows + (size_t)li * F * rs + ix;
      const bool xr = g.weight_kind == 2;
      const double q = xr ? qlen_of(g, h1, h2, h3) : 0.0;
      double tr = 0.0, ti = 0.0;
      for (int a = 0; a < F; ++a) {
        const V v = rp[(size_t)a * rs];
        const double ar = (double)v.x * d, ai = (double)v.y * d;
        if (rho_out) rho_out[(size_t)a * ncand + c] = V{(T)ar, (T)ai};
        const double w = xr ? xray_f(wtab + (size_t)a * kWtab + 1, q) : wtab[(size_t)a * kWtab];
        tr += w * ar;
        ti += w * ai;
      }
      const double val = tr * tr + ti * ti;
      add(s, 2.0 * val);
      if (gnum) gnum[c] += val;
      if (g.partials || gpart) {
        int pidx = 0;
        for (int a = 0; a < F; ++a) {
          const V va = rp[(size_t)a * rs];
          const double ar = (double)va.x * d, ai = (double)va.y * d;
          for (int b = a; b < F; ++b, ++pidx) {
            const V vb = rp[(size_t)b * rs];
            const double pv = g.pair_scale * (ar * ((double)vb.x * d) + ai * ((double)vb.y * d));
            if (g.partials) add((pidx + 1) * nq + s, 2.0 * pv);
            if (gpart) gpart[(size_t)pidx * ncand + c] += pv;
          }
        }
      }
    };
    auto bin_mode = [&](int c, int li, unsigned s) {
      if constexpr (FC < 0) {
        bin_mode_any(c, li, s);
        return;
      }
      const int h2 = lg.h[li][0], h3 = lg.h[li][1];
      const int h1 = lg.h[li][2] + (c - lg.off[li]);
      const double d = lg.d23[li] * inv1[abs(h1)];
      const int ix = h1 < 0 ? h1 + n1 : h1;
      double rr[FM], ri[FM];
#pragma unroll
      for (int f = 0; f < FM && f < F; ++f) {
        const V v = rows[((size_t)li * F + f) * rs + (ix)];
        rr[f] = (double)v.x * d;
        ri[f] = (double)v.y * d;
        if (rho_out) rho_out[(size_t)f * ncand + c] = V{(T)rr[f], (T)ri[f]};
      }
      double tr, ti;
      if (F == 1) {
        // one field: neutron and unit weights ride in the atom strengths; a
        // single species under X-ray weights carries strength 1 and takes
        // f(|q|) here
        const double w0 =
            (FC != 1 && g.weight_kind == 2) ? xray_f(g.xcoef[0], qlen_of(g, h1, h2, h3)) : 1.0;
        tr = w0 * rr[0];
        ti = w0 * ri[0];
      } else {
        const double q = g.weight_kind == 2 ? qlen_of(g, h1, h2, h3) : 0.0;
        tr = 0.0;
        ti = 0.0;
#pragma unroll
        for (int a = 0; a < FM && a < F; ++a) {
          const double w = (g.weight_kind == 2) ? xray_f(g.xcoef[a], q) : g.amp[a];
          tr += w * rr[a];
          ti += w * ri[a];
        }
      }
      const double val = tr * tr + ti * ti;
      add(s, 2.0 * val);
      if (gnum) gnum[c] += val;
      if (FC != 1 && g.partials) {
        // pairs a <= b of the per-species fields (one species: its own pair)
        int pidx = 1;
#pragma unroll
        for (int a = 0; a < FM && a < F; ++a)
#pragma unroll
          for (int b = a; b < FM && b < F; ++b, ++pidx)
            add(pidx * nq + s, 2.0 * ps * (rr[a] * rr[b] + ri[a] * ri[b]));
      }
      // per-mode pair sums (one thread owns candidate c within a frame), kept
      // out of the loop above so the plain path compiles as before
      if (FC != 1 && gpart) {
        int q = 0;
#pragma unroll
        for (int a = 0; a < FM && a < F; ++a)
#pragma unroll
          for (int b = a; b < FM && b < F; ++b, ++q)
            gpart[(size_t)q * ncand + c] += ps * (rr[a] * rr[b] + ri[a] * ri[b]);
      }
    };
    // one copy of the body: the prefetched index of round k is picked by a
    // compare-select chain, so sv stays in registers.  A thread's candidates
    // ascend, so its line search resumes where the previous one stopped.
    int li = 0;
#pragma unroll 1
    for (int k = 0, c = c0 + tid; c < c1; ++k, c += NT) {
      unsigned sh = sv[0];
#pragma unroll
      for (int u = 1; u < kPreShell; ++u) sh = (k == u) ? sv[u] : sh;
      const unsigned s = k < kPreShell ? sh : (unsigned)stab[c];
      while (li + 1 < nl && lg.off[li + 1] <= c) ++li;
      if (s != kNoShell) bin_mode(c, li, s);
    }
  }
  __syncthreads();
  if (gacc) return;  // the sums are already in the row
  for (int i = tid; i < nw; i += NT) accg[i] = acc[i];
}

// Sum of the workgroup rows of one frame: the integer words of an entry are
// summed exactly (exact_parts bounds them), so the order of the rows does
// not matter; the frame's value is hi q + lo q 2^-k.  The frame is then
// added into the run's accumulators with Neumaier's compensation, frame
// after frame.
constexpr int kFinParts = 16;
constexpr int kFinEntries = 256 / kFinParts;

__device__ __forceinline__ void neumaier(double &s, double &c, double v) {
  const double t = s + v;
  c += (fabs(s) >= fabs(v)) ? (s - t) + v : (v - t) + s;
  s = t;
}

__global__ void __launch_bounds__(256) k_finalize(const unsigned long long *__restrict__ wgpart,
                                                  int nblk, int ntot, int words, int nq,
                                                  double q0, double q1, double lo_q,
                                                  double *__restrict__ acc,
                                                  double *__restrict__ frame_out, int accumulate) {
  __shared__ unsigned long long ph[kFinParts][kFinEntries], pl[kFinParts][kFinEntries];
  const int e = threadIdx.x % kFinEntries, part = threadIdx.x / kFinEntries;
  const int i = blockIdx.x * kFinEntries + e;
  const size_t row = (size_t)words * ntot;
  unsigned long long sh = 0, sl = 0;
  if (i < ntot) {
    int b = part;
    for (; b + 3 * kFinParts < nblk; b += 4 * kFinParts) {
      unsigned long long h[4], l[4] = {0, 0, 0, 0};
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        h[u] = wgpart[(size_t)(b + u * kFinParts) * row + i];
        if (words == 2) l[u] = wgpart[(size_t)(b + u * kFinParts) * row + ntot + i];
      }
      sh += (h[0] + h[1]) + (h[2] + h[3]);
      sl += (l[0] + l[1]) + (l[2] + l[3]);
    }
    for (; b < nblk; b += kFinParts) {
      sh += wgpart[(size_t)b * row + i];
      if (words == 2) sl += wgpart[(size_t)b * row + ntot + i];
    }
  }
  ph[part][e] = sh;
  pl[part][e] = sl;
  __syncthreads();
  if (part != 0 || i >= ntot) return;
  unsigned long long th = 0, tl = 0;
  for (int k = 0; k < kFinParts; ++k) {
    th += ph[k][e];
    tl += pl[k][e];
  }
  const double q = i < nq ? q0 : q1;
  const double total = (double)(long long)th * q + (double)(long long)tl * (q * lo_q);
  if (frame_out) frame_out[i] = total;
  if (accumulate) {
    double a = acc[i], cc = acc[ntot + i];
    neumaier(a, cc, total);
    acc[i] = a;
    acc[ntot + i] = cc;
  }
}

// ---------------------------------------------------------------------------
// Adjoint.  V_c(x) = sum_m C_cm exp(i h_m . x) is real; its gradient at the
// atoms is interpolated with the kernel derivative.
// ---------------------------------------------------------------------------
template <typename T, int N, int FC>
__global__ void __launch_bounds__(NT) k_adj_x(Geo g, AxisPlan ax1, int lb, int ngroups, int Fa_,
                                              const typename C2<T>::t *__restrict__ rho,
                                              int64_t ncand,
                                              const typename C2<T>::t *__restrict__ twx,
                                              const int4 *__restrict__ line_info,
                                              const int32_t *__restrict__ line_off,
                                              const uint16_t *__restrict__ stab,
                                              const double *__restrict__ ph1,
                                              const double *__restrict__ line_d23,
                                              const double *__restrict__ den,
                                              const double *__restrict__ cnt,
                                              const double *__restrict__ gS,
                                              const double *__restrict__ gP,
                                              typename C2<T>::t *__restrict__ W2,
                                              const double *__restrict__ wtab) {
  using V = typename C2<T>::t;
  extern __shared__ __align__(16) unsigned char smem[];
  // fields, species and adjoint fields coincide when there is more than one
  const int n1 = g.n[0], F = FC > 0 ? FC : g.F, nq = g.nq, Fa = FC > 0 ? FC : Fa_;
  const int S = F > 1 ? F : g.S;
  constexpr int FM = FC > 0 ? FC : kGeoSpecies;
  __shared__ LineGroup lg;
  const int rs = g.rs1;
  V *rows = reinterpret_cast<V *>(smem);
  V *tw = rows + (size_t)lb * Fa * rs;
  double *inv1 = reinterpret_cast<double *>(tw + n1);
  // per shell: G = gS / den (0 for an empty shell) and, with pair gradients,
  // 1 / (N max(count, 1)), in the LDS the forward pass keeps its shell sums
  // in (two rows of nq)
  double *Gsh = inv1 + g.M[0] + 1, *inc = Gsh + nq;
  // scratch lines of the long-line DFT, past the shell tables
  V *lscr = reinterpret_cast<V *>(
      smem + ((reinterpret_cast<unsigned char *>(inc + nq) - smem + 15) & ~(ptrdiff_t)15));
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  for (int i = tid; i < n1; i += NT) tw[i] = twx[i];
  for (int i = tid; i <= g.M[0]; i += NT) inv1[i] = 1.0 / ph1[i];
  for (int i = tid; i < nq; i += NT) {
    Gsh[i] = den[i] > 0.0 ? gS[i] / den[i] : 0.0;
    if (gP) inc[i] = 1.0 / ((double)g.natoms * fmax(cnt[i], 1.0));
  }
  for (int grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
    const int l0 = grp * lb;
    const int nl = min(lb, g.nlines - l0);
    __syncthreads();
    load_group(lg, l0, nl, line_info, line_off, line_d23);
    for (int idx = tid; idx < nl * Fa * rs; idx += NT) rows[idx] = V{T(0), T(0)};
    __syncthreads();
    const int c0 = lg.off[0], c1 = lg.off[nl];
    // shell indices and field-0 amplitudes of this thread's first KP
    // candidates, all in flight at once
    constexpr int KP = sizeof(T) == 4 ? kPreShell : kPreShell / 2;
    uint16_t sv[KP];
    V rv[KP];
    if (c1 > c0) {
#pragma unroll
      for (int k = 0; k < KP; ++k) {
        const int c = min(c0 + tid + k * NT, c1 - 1);
        sv[k] = stab[c];
        rv[k] = rho[c];
      }
    }
    // candidate c of line li, kept in shell s
    // the factor on the pair sums (1 whenever there are two or more fields)
    const double ps = FC == 2 ? 1.0 : g.pair_scale;
    // any number of fields (F = Fa > kGeoSpecies): the amplitudes are read
    // from rho per use and the weights from the plan's table
    auto amplify_any = [&](int c, int li, unsigned s) {
      const int h2 = lg.h[li][0], h3 = lg.h[li][1];
      const int h1 = lg.h[li][2] + (c - lg.off[li]);
      const double d = lg.d23[li] * inv1[abs(h1)];
      const int ix = h1 < 0 ? h1 + n1 : h1;
      const bool xr = g.weight_kind == 2;
      const double q = xr ? qlen_of(g, h1, h2, h3) : 0.0;
      auto wt = [&](int a) {
        return xr ? xray_f(wtab + (size_t)a * kWtab + 1, q) : wtab[(size_t)a * kWtab];
      };
      double tr = 0.0, ti = 0.0;
      for (int a = 0; a < F; ++a) {
        const V v = rho[(size_t)a * ncand + c];
        const double w = wt(a);
        tr += w * (double)v.x;
        ti += w * (double)v.y;
      }
      const double G = Gsh[s];
      const double ic = gP ? inc[s] : 0.0;
      for (int cf = 0; cf < Fa; ++cf) {
        const double wc = wt(cf);
        double cr = 2.0 * G * wc * tr, ci = -2.0 * G * wc * ti;
        if (gP) {
          // the pairs (a, b) with a = cf or b = cf
          for (int o = 0; o < F; ++o) {
            const int a = o < cf ? o : cf, b = o < cf ? cf : o;
            const V vo = rho[(size_t)o * ncand + c];
            double H = gP[((size_t)a * S + b) * nq + s] * ic * g.pair_scale;
            if (o == cf) H *= 2.0;  // the pair (cf, cf) counts both factors
            cr += H * (double)vo.x;
            ci -= H * (double)vo.y;
          }
        }
        rows[((size_t)li * Fa + cf) * rs + (ix)] = V{(T)(cr * d), (T)(ci * d)};
      }
    };
    auto amplify = [&](int c, int li, unsigned s, V r0) {
      if constexpr (FC < 0) {
        amplify_any(c, li, s);
        return;
      }
      const int h2 = lg.h[li][0], h3 = lg.h[li][1];
      const int h1 = lg.h[li][2] + (c - lg.off[li]);
      const double d = lg.d23[li] * inv1[abs(h1)];
      const int ix = h1 < 0 ? h1 + n1 : h1;
      double rr[FM], ri[FM], w[FM];
      rr[0] = r0.x;
      ri[0] = r0.y;
#pragma unroll
      for (int f = 1; f < FM && f < F; ++f) {
        const V v = rho[(size_t)f * ncand + c];
        rr[f] = v.x;
        ri[f] = v.y;
      }
      double tr = 0.0, ti = 0.0;
      if (F == 1) {
        // as in the forward epilogue: f(|q|) of a single X-ray species
        w[0] = (FC != 1 && g.weight_kind == 2) ? xray_f(g.xcoef[0], qlen_of(g, h1, h2, h3)) : 1.0;
        tr = w[0] * rr[0];
        ti = w[0] * ri[0];
      } else {
        const double q = g.weight_kind == 2 ? qlen_of(g, h1, h2, h3) : 0.0;
#pragma unroll
        for (int a = 0; a < FM && a < F; ++a) {
          w[a] = (g.weight_kind == 2) ? xray_f(g.xcoef[a], q) : g.amp[a];
          tr += w[a] * rr[a];
          ti += w[a] * ri[a];
        }
      }
      const double G = Gsh[s];
#pragma unroll
      for (int cf = 0; cf < FM && cf < Fa; ++cf) {
        const double wc = w[cf];
        double cr = 2.0 * G * wc * tr, ci = -2.0 * G * wc * ti;
        if (FC != 1 && gP) {
          // pair gradients exist only with one field per species
          const double ic = inc[s];
#pragma unroll
          for (int a = 0; a < FM && a < F; ++a)
#pragma unroll
            for (int b = a; b < FM && b < F; ++b) {
              const double H = gP[((size_t)a * S + b) * nq + s] * ic * ps;
              if (cf == a) { cr += H * rr[b]; ci -= H * ri[b]; }
              if (cf == b) { cr += H * rr[a]; ci -= H * ri[a]; }
            }
        }
        rows[((size_t)li * Fa + cf) * rs + (ix)] = V{(T)(cr * d), (T)(ci * d)};
      }
    };
    int li = 0;  // a thread's candidates ascend: the line search resumes
#pragma unroll 1
    for (int k = 0, c = c0 + tid; c < c1; ++k, c += NT) {
      unsigned sh = sv[0];
      V r0 = rv[0];
#pragma unroll
      for (int u = 1; u < KP; ++u) {
        sh = (k == u) ? sv[u] : sh;
        r0 = (k == u) ? rv[u] : r0;
      }
      if (k >= KP) {
        sh = stab[c];
        r0 = rho[c];
      }
      while (li + 1 < nl && lg.off[li + 1] <= c) ++li;
      if (sh != kNoShell) amplify(c, li, sh, r0);
    }
    __syncthreads();
    for (int r = wave; r < nl * Fa; r += NW)
      wave_fft<T, N>(rows + (size_t)r * rs, ax1, tw, lane, lscr + (size_t)wave * n1);
    __syncthreads();
    const float inv_n1 = 1.0f / (float)n1, inv_fa = 1.0f / (float)Fa;
    for (int idx = tid; idx < nl * Fa * n1; idx += NT) {
      const int r = div_by<N>(idx, inv_n1), x = idx - r * n1;
      const int li = div_by<FC>(r, inv_fa), cf = r - li * Fa;
      W2[(size_t)(cf * g.nlines + l0 + li) * n1 + x] = rows[r * rs + x];
    }
  }
}

template <typename T, int N, int TX2>
__global__ void __launch_bounds__(NT) k_adj_y(Geo g, AxisPlan ay,
                                              const typename C2<T>::t *__restrict__ W2,
                                              const typename C2<T>::t *__restrict__ twy,
                                              const int32_t *__restrict__ lo2,
                                              const int32_t *__restrict__ hi2,
                                              const int32_t *__restrict__ loff,
                                              typename C2<T>::t *__restrict__ W1) {
  using V = typename C2<T>::t;
  extern __shared__ __align__(16) unsigned char smem[];
  const int n1 = g.n[0], n2 = g.n[1], M3 = g.M[2], ls = g.ls2;
  V *lines = reinterpret_cast<V *>(smem);
  constexpr int tx2 = TX2, lt = TX2 == 16 ? 4 : TX2 == 8 ? 3 : TX2 == 4 ? 2 : 1, tm = TX2 - 1;
  static_assert(TX2 == 2 || TX2 == 4 || TX2 == 8 || TX2 == 16, "column chunk");
  const int nbx = (g.n[0] + tx2 - 1) / tx2;
  const int blk = xcd_block(blockIdx.x, gridDim.x, g.nxcd);
  const int x0 = (blk % nbx) * tx2, k3 = (blk / nbx) % (g.M[2] + 1), cf = blk / (nbx * (g.M[2] + 1));
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int lo = lo2[k3], hi = hi2[k3];
  for (int i = tid; i < tx2 * ls; i += NT) lines[i] = V{T(0), T(0)};
  __syncthreads();
  const V *src = W2 + ((size_t)cf * g.nlines + (hi >= lo ? loff[k3] : 0)) * n1;
  batched_copy<8, V>(
      (hi - lo + 1) * tx2,
      [&](int idx) {
        const int x = x0 + (idx & tm);
        return x < n1 ? src[(size_t)(idx >> lt) * n1 + x] : V{T(0), T(0)};
      },
      [&](int idx, V v) {
        const int h2 = lo + (idx >> lt);
        const int yi = h2 < 0 ? h2 + n2 : h2;
        if (x0 + (idx & tm) < n1) lines[(idx & tm) * ls + yi] = v;
      });
  __syncthreads();
  wave_fft_lines<T, N>(lines, ls, wave, tx2, ay, twy, lane,
                       lines + (size_t)tx2 * ls + (size_t)wave * g.n[1]);
  __syncthreads();
  V *dst = W1 + ((size_t)cf * (M3 + 1) + k3) * n2 * n1;
  for (int idx = tid; idx < n2 * tx2; idx += NT) {
    const int xl = idx & tm, y = idx >> lt;
    const int x = x0 + xl;
    if (x >= n1) continue;
    V v = lines[xl * ls + y];
    // the h3 = 0 plane holds only the canonical half of (h1, h2): the full
    // plane sum is that half plus its conjugate
    if (k3 == 0) v = V{T(2) * v.x, T(0)};
    dst[(size_t)y * n1 + x] = v;
  }
}

template <typename T, int N, int TT>
__global__ void __launch_bounds__(NT) k_adj_z_gather(Geo g, AxisPlan az,
                                                     const typename C2<T>::t *__restrict__ W1,
                                                     const typename C2<T>::t *__restrict__ twz,
                                                     const int4 *__restrict__ i0s,
                                                     const T *__restrict__ phi,
                                                     const T *__restrict__ dphi,
                                                     const int32_t *__restrict__ sidx,
                                                     const int32_t *__restrict__ start,
                                                     double *__restrict__ slots) {
  using V = typename C2<T>::t;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ Stage st;
  const int n1 = g.n[0], n2 = g.n[1], n3 = g.n[2], M3 = g.M[2];
  const int cs = N > 0 ? spread_col_stride(N) : g.cs;
  // the tile shape: compiled in for the common 8 x 2 tile (TT = 1), read
  // at run time otherwise
  constexpr bool kt = TT == 1;
  const int txr = kt ? 8 : g.txr, tty = kt ? 2 : g.ty;
  const int ccx = txr / 2, cc = ccx * tty;
  V *tile = reinterpret_cast<V *>(smem);
  int4 *rec = reinterpret_cast<int4 *>(tile + (size_t)cc * cs);
  const T *tflat = reinterpret_cast<const T *>(tile);
  // grid (tiles, fields), renumbered so that every XCD walks a contiguous
  // range of tiles: neighbouring tiles read the two halves of the same W1
  // cache lines, and then share one L2
  const int cf = blockIdx.y, tb = xcd_block(blockIdx.x, gridDim.x, g.nxcd);
  const int tx = tb % g.ntx, ty = tb / g.ntx;
  const int x0 = tx * txr, y0 = ty * tty;
  const int txw = min(txr, n1 - x0), tyw = min(tty, n2 - y0);
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  {
    // clear the tile 16 bytes a store (cc * cs * sizeof(V) is a multiple of
    // 16: cs is even)
    float4 *t4 = reinterpret_cast<float4 *>(tile);
    const int n4 = cc * cs * (int)sizeof(V) / 16;
    for (int i = tid; i < n4; i += NT) t4[i] = float4{0.f, 0.f, 0.f, 0.f};
  }
  __syncthreads();
  const int pw = txw / 2;
  const V *src = W1 + (size_t)cf * (M3 + 1) * n2 * n1;
  struct VV {
    V a, b;
  };
  const float inv_pw = 1.0f / (float)pw, inv_ty = 1.0f / (float)tyw;
  batched_copy<4, VV>(
      (M3 + 1) * tyw * pw,
      [&](int idx) {
        const int r = small_div(idx, inv_pw), p = idx - r * pw;
        const int k3 = small_div(r, inv_ty), yl = r - k3 * tyw;
        const V *q = src + (size_t)(k3 * n2 + y0 + yl) * n1 + x0 + 2 * p;
        return VV{q[0], q[1]};
      },
      [&](int idx, VV v) {
        const int r = small_div(idx, inv_pw), p = idx - r * pw;
        const int k3 = small_div(r, inv_ty), yl = r - k3 * tyw;
        const int col = yl * ccx + p;
        const V A = v.a, B = v.b;
        tile[(size_t)col * cs + (k3)] = V{A.x - B.y, A.y + B.x};
        if (k3 > 0) tile[(size_t)col * cs + (n3 - k3)] = V{A.x + B.y, B.x - A.y};
      });
  __syncthreads();
  // long lines: the wave's scratch line lies past the tile (the staged
  // records are not in use during the DFT)
  wave_fft_lines<T, N>(tile, cs, wave, cc, az, twz, lane,
                       tile + (size_t)cc * cs + (size_t)wave * n3);
  stage_window(st, g, tx, ty, cf, start);

  const int ns = g.ns;
  const int S2 = g.sx * g.sy;
  // Gather: the tile is only read, so one wave serves several atoms at once:
  // lpa lanes per atom (one per xy column of its stencil in the tile, at
  // least 8), ga = 64 / lpa atoms per round.  Each lane sums its column over
  // z in the path precision, and the three derivative sums are reduced in
  // FP64 over the atom's lanes.
  const int lpa = max(8, txr * tty), ga = 64 / lpa;
  const int grp = lane / lpa, l = lane - grp * lpa;
  // the (atom, tile) pair's slot: tiles counted from the one that holds the
  // first stencil point, which the record places at tile column xll - kxl
  // (row yll - kyl), modulo the grid
  auto put_slot = [&](const int4 r, double gx, double gy, double gz) {
    const int kxl = r.y & 255, xll = (r.y >> 8) & 255;
    const int kyl = r.z & 255, yll = (r.z >> 8) & 255;
    const int xs = wrapi(x0 + xll - kxl, n1), ys = wrapi(y0 + yll - kyl, n2);
    const int dx = wrapi(tx - xs / txr, g.ntx);
    const int dy = wrapi(ty - ys / tty, g.nty);
    double *o = slots + ((size_t)r.x * S2 + dx * g.sy + dy) * 3;
    o[0] = gx;
    o[1] = gy;
    o[2] = gz;
  };
  // The compiled 8 x 2 tile: 16 lanes per atom, one DPP row.  A round's
  // kernel values stay in registers: lane i of an atom's row holds value i
  // of each of the six tables (phi, then dphi, along x, y, z), and the next
  // round's are read while this one computes.  A lane takes its column's x
  // and y values by a shuffle within the row and the z values by row
  // broadcasts, so a round touches the LDS only for the record and the
  // tile column.
  // (the record of the lane's atom comes with them)
  auto fetch_rows = [&](T (&dst)[6], int4 &r, int s0, int na) {
    r = rec[s0 + min(grp, na - 1)];
    const bool ok = grp < na && l < ns;
    const size_t b = (size_t)(ok ? r.x : 0) * 3 * ns + l;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      dst[q] = T(0);
      if (ok) dst[q] = q < 3 ? phi[b + q * ns] : dphi[b + (q - 3) * ns];
    }
  };
  auto gather_rows = [&](int nst) {
    T cur[6], nxt[6];
    int4 r{0, 0, 0, 0}, rn{0, 0, 0, 0};
    if (wave * ga < nst) fetch_rows(cur, r, wave * ga, min(ga, nst - wave * ga));
    for (int s0 = wave * ga; s0 < nst; s0 += NW * ga) {
      const int na = min(ga, nst - s0);
      const int s1 = s0 + NW * ga;
      if (s1 < nst) fetch_rows(nxt, rn, s1, min(ga, nst - s1));
      const bool act = grp < na;
      const int kxl = r.y & 255, xll = (r.y >> 8) & 255, ox = r.y >> 16;
      const int kyl = r.z & 255, yll = (r.z >> 8) & 255, oy = r.z >> 16;
      const int a = r.w & 0xFFFF;
      const bool cell = act && l < ox * oy;
      const int lc = cell ? l : 0;
      const int iy = small_div(lc, __builtin_amdgcn_rcpf((float)ox));
      const int ix = lc - iy * ox;
      // every lane takes part in the shuffles
      const int ax = 4 * (16 * grp + kxl + ix), ay = 4 * (16 * grp + kyl + iy);
      const T fx = bperm(ax, cur[0]), fy = bperm(ay, cur[1]);
      const T dfx = bperm(ax, cur[3]), dfy = bperm(ay, cur[4]);
      const int xl = xll + ix, yl = yll + iy;
      // idle lanes read column 0 (in the tile, unused)
      const T *colv = tflat + (cell ? 2 * (yl * ccx + (xl >> 1)) * cs + (xl & 1) : 0);
      T sv = T(0), sd = T(0);
      static_for<0, 16>([&](auto kc) {
        constexpr int kz = decltype(kc)::value;
        if (kz < ns) {
          const T pz = row_bcast<kz>(cur[2]), dz = row_bcast<kz>(cur[5]);
          const int z = a + kz < n3 ? a + kz : a + kz - n3;
          const T v = colv[2 * z];
          sv += v * pz;
          sd += v * dz;
        }
      });
      double gx = 0.0, gy = 0.0, gz = 0.0;
      if (cell) {
        gx = (double)(dfx * fy) * (double)sv;
        gy = (double)(fx * dfy) * (double)sv;
        gz = (double)(fx * fy) * (double)sd;
      }
      gx = row16_sum(gx);
      gy = row16_sum(gy);
      gz = row16_sum(gz);
      if (act && l == 0) put_slot(r, gx, gy, gz);
#pragma unroll
      for (int q = 0; q < 6; ++q) cur[q] = nxt[q];
      r = rn;
    }
  };
  // Other tiles: the kernel values of a round's atoms go to a per-wave
  // scratch, [atom][phi 3 ns, dphi 3 ns]: value k of the round is read by
  // lane k % 64 into register k / 64.  When they fit kGV registers, the next
  // round's are read while this one computes.
  T *scr = reinterpret_cast<T *>(rec + kStageCap) + (size_t)wave * ga * 6 * ns;
  constexpr int kGV = 6;
  const int nval = 6 * ns;
  const float inv_nval = __builtin_amdgcn_rcpf((float)nval);
  const bool ahead = ga * nval <= kGV * 64;
  T nv[kGV];
  auto fetch = [&](int s0, int na) {
#pragma unroll
    for (int u = 0; u < kGV; ++u) {
      const int k = lane + 64 * u;
      nv[u] = T(0);
      if (k < na * nval) {
        const int aa = small_div(k, inv_nval), kk = k - aa * nval;
        const int j = rec[s0 + aa].x;
        nv[u] = kk < 3 * ns ? phi[(size_t)j * 3 * ns + kk] : dphi[(size_t)j * 3 * ns + kk - 3 * ns];
      }
    }
  };
  auto gather_scratch = [&](int nst) {
    if (ahead && wave * ga < nst) fetch(wave * ga, min(ga, nst - wave * ga));
    for (int s0 = wave * ga; s0 < nst; s0 += NW * ga) {
      const int na = min(ga, nst - s0);
      if (ahead) {
#pragma unroll
        for (int u = 0; u < kGV; ++u)
          if (lane + 64 * u < na * nval) scr[lane + 64 * u] = nv[u];
        const int s1 = s0 + NW * ga;
        if (s1 < nst) fetch(s1, min(ga, nst - s1));
      } else {
        for (int k = lane; k < na * nval; k += 64) {
          const int aa = small_div(k, inv_nval), kk = k - aa * nval;
          const int j = rec[s0 + aa].x;
          scr[k] = kk < 3 * ns ? phi[(size_t)j * 3 * ns + kk] : dphi[(size_t)j * 3 * ns + kk - 3 * ns];
        }
      }
      wave_sync();
      const bool act = grp < na;
      const int4 r = rec[s0 + min(grp, na - 1)];
      const int kxl = r.y & 255, xll = (r.y >> 8) & 255, ox = r.y >> 16;
      const int kyl = r.z & 255, yll = (r.z >> 8) & 255, oy = r.z >> 16;
      const int a = r.w & 0xFFFF;
      const bool cell = act && l < ox * oy;
      const int lc = cell ? l : 0;
      const int iy = small_div(lc, __builtin_amdgcn_rcpf((float)ox));
      const int ix = lc - iy * ox;
      const T *ps = scr + grp * 6 * ns;
      double gx = 0.0, gy = 0.0, gz = 0.0;
      if (cell) {
        const T fx = ps[kxl + ix], fy = ps[ns + kyl + iy];
        const T dfx = ps[3 * ns + kxl + ix], dfy = ps[4 * ns + kyl + iy];
        const int xl = xll + ix, yl = yll + iy;
        const T *colv = tflat + 2 * (yl * ccx + (xl >> 1)) * cs + (xl & 1);
        T sv = T(0), sd = T(0);
        // z points a .. a + ns - 1: the run before the wrap, then the rest
        const int n0 = min(ns, n3 - a);
        const T *pz = ps + 2 * ns, *dz = ps + 5 * ns;
        for (int kz = 0; kz < n0; ++kz) {
          const T v = colv[2 * (a + kz)];
          sv += v * pz[kz];
          sd += v * dz[kz];
        }
        for (int kz = n0; kz < ns; ++kz) {
          const T v = colv[2 * (kz - n0)];
          sv += v * pz[kz];
          sd += v * dz[kz];
        }
        gx = (double)(dfx * fy) * (double)sv;
        gy = (double)(fx * dfy) * (double)sv;
        gz = (double)(fx * fy) * (double)sd;
      }
      if (lpa == 16) {
        // an atom's lanes are one DPP row: rotations within the row
        gx = row16_sum(gx);
        gy = row16_sum(gy);
        gz = row16_sum(gz);
      } else {
        for (int off = lpa >> 1; off > 0; off >>= 1) {
          gx += __shfl_xor(gx, off);
          gy += __shfl_xor(gy, off);
          gz += __shfl_xor(gz, off);
        }
      }
      if (act && l == 0) put_slot(r, gx, gy, gz);
      wave_sync();  // the next round overwrites the scratch
    }
  };
  for (int base = 0; base < st.total; base += NT) {
    const int nst = stage_chunk(st, rec, base, g, x0, txw, y0, tyw, sidx, i0s);
    const bool last = base + NT >= st.total;
    if (!last && nst <= kStageCap - NT) continue;
    if constexpr (kt)
      gather_rows(nst);
    else
      gather_scratch(nst);
    __syncthreads();
    if (tid == 0) st.nstaged = 0;
    __syncthreads();
  }
}

// Sum each atom's slots in a fixed order and map grid derivatives to dL/dr:
// dL/dr = sum_i b_i (n_i / 2 pi) dV/dX_i, dV/dX_i = -sum_g V phi'(u_i) prod phi.
__global__ void k_adj_sum(Geo g, const int4 *__restrict__ i0a,
                          const int32_t *__restrict__ species, const double *__restrict__ slots,
                          int Fa, double *__restrict__ grad, const double *__restrict__ wtab) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= g.natoms) return;
  const int4 i0 = i0a[j];
  const int xs = wrapi(i0.x, g.n[0]), ys = wrapi(i0.y, g.n[1]);
  const int tx0 = xs / g.txr, ty0 = ys / g.ty;
  const int S2 = g.sx * g.sy;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int dx = 0; dx < g.sx; ++dx) {
    const int tx = wrapi(tx0 + dx, g.ntx);
    if (dx > 0 && tx == tx0) break;
    const int x0 = tx * g.txr, txw = min(g.txr, g.n[0] - x0);
    int a, b;
    if (overlap(i0.x, g.ns, g.n[0], x0, txw, &a, &b) == 0) continue;
    for (int dy = 0; dy < g.sy; ++dy) {
      const int ty = wrapi(ty0 + dy, g.nty);
      if (dy > 0 && ty == ty0) break;
      const int y0 = ty * g.ty, tyw = min(g.ty, g.n[1] - y0);
      if (overlap(i0.y, g.ns, g.n[1], y0, tyw, &a, &b) == 0) continue;
      const double *o = slots + ((size_t)j * S2 + dx * g.sy + dy) * 3;
      s0 += o[0];
      s1 += o[1];
      s2 += o[2];
    }
  }
  double w = 1.0;
  if (Fa == 1) w = g.S > kGeoSpecies ? wtab[species[j] * kWtab] : g.amp[species[j]];
  const double inv_two_pi = 1.0 / (2.0 * 3.14159265358979323846);
  const double d0 = -s0 * g.n[0] * inv_two_pi * w;
  const double d1 = -s1 * g.n[1] * inv_two_pi * w;
  const double d2 = -s2 * g.n[2] * inv_two_pi * w;
  // columns of b are b_1, b_2, b_3
#pragma unroll
  for (int c = 0; c < 3; ++c) grad[3 * j + c] = g.b[c] * d0 + g.b[c + 3] * d1 + g.b[c + 6] * d2;
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
// Records a pair of events around one stage when the plan is profiling.
struct StageTimer {
  Plan &p;
  int stage;
  hipEvent_t a = nullptr, b = nullptr;
  StageTimer(Plan &pl, int st) : p(pl), stage(st) {
    if (!p.profiling) return;
    a = take();
    b = take();
    if (!a || !b) {
      if (a) p.ev_free.push_back(a);
      if (b) p.ev_free.push_back(b);
      a = b = nullptr;
      return;
    }
    (void)hipEventRecord(a, p.stream);
  }
  hipEvent_t take() {
    hipEvent_t e = nullptr;
    if (!p.ev_free.empty()) {
      e = p.ev_free.back();
      p.ev_free.pop_back();
    } else if (hipEventCreate(&e) != hipSuccess) {
      e = nullptr;
    }
    return e;
  }
  ~StageTimer() {
    if (!a) return;
    (void)hipEventRecord(b, p.stream);
    p.pending.push_back({stage, a, b});
  }
};

int collect_stage_times(Plan &p) {
  if (p.pending.empty()) return 0;
  hipError_t e = hipStreamSynchronize(p.stream);
  for (auto &q : p.pending) {
    float ms = 0.f;
    if (e == hipSuccess && hipEventElapsedTime(&ms, q.a, q.b) == hipSuccess) {
      p.stage_ms[q.stage] += ms;
      p.stage_n[q.stage] += 1;
    }
    p.ev_free.push_back(q.a);
    p.ev_free.push_back(q.b);
  }
  p.pending.clear();
  return (int)e;
}

#define SQC_CHECK(x)                         \
  do {                                       \
    hipError_t e_ = (x);                     \
    if (e_ != hipSuccess) return (int)e_;    \
  } while (0)


// Sizes with a compile-time line DFT: the fine grids of the benchmark
// workloads (sigma 1.25 in FP32, sigma 2 in FP64); every other size runs the
// runtime-planned DFT (N = 0).
#define SQC_FIXED_F32(X) X(90) X(120) X(270) X(432) X(480)
#define SQC_FIXED_F64(X) X(144) X(180) X(432) X(648) X(750)

template <typename T> struct FixedSizes;
template <> struct FixedSizes<float> {
  template <typename F> static void pick(int n, F &&f) {
#define SQC_CASE(N) if (n == N) return f(std::integral_constant<int, N>{});
    SQC_FIXED_F32(SQC_CASE)
#undef SQC_CASE
    f(std::integral_constant<int, 0>{});
  }
};
template <> struct FixedSizes<double> {
  template <typename F> static void pick(int n, F &&f) {
#define SQC_CASE(N) if (n == N) return f(std::integral_constant<int, N>{});
    SQC_FIXED_F64(SQC_CASE)
#undef SQC_CASE
    f(std::integral_constant<int, 0>{});
  }
};

// The y pass's column chunk as a compile-time constant (2, 4, 8 or 16).
template <typename Fn> static void pick_chunk(int tx2, Fn &&fn) {
  if (tx2 == 16) return fn(std::integral_constant<int, 16>{});
  if (tx2 == 8) return fn(std::integral_constant<int, 8>{});
  if (tx2 == 4) return fn(std::integral_constant<int, 4>{});
  fn(std::integral_constant<int, 2>{});
}

// Field counts the x-pass epilogue is compiled for; others run its loops
// with the count read at run time.  The one-field instance serves one
// weighted field with no pair sums (neutron or unit weights, partials off):
// a single X-ray species, whose f(|q|) the epilogue applies, and a single
// species with its pair column take the run-time instance.
template <typename Fn> static void pick_fields(const Geo &g, Fn &&fn) {
  const int F = g.F;
  if (F > kGeoSpecies) return fn(std::integral_constant<int, -1>{});
  if (g.acc_global) return fn(std::integral_constant<int, 0>{});  // shell sums in HBM
  if (F == 1 && g.weight_kind != 2 && !g.partials) return fn(std::integral_constant<int, 1>{});
  if (F == 2) return fn(std::integral_constant<int, 2>{});
  fn(std::integral_constant<int, 0>{});
}

template <typename T>
static void sort_atoms(Plan &p, const double *d_pos, bool with_dphi, hipStream_t st) {
  Geo &g = p.geo;
  const int nbins_tot = g.F * g.nbins;
  const int blk = 256;
  const unsigned ga = (unsigned)((g.natoms + blk - 1) / blk);
  StageTimer t(p, 0);
  const unsigned gv = (unsigned)((g.natoms * 3 * g.ns + blk - 1) / blk);
  hipLaunchKernelGGL(k_prep<T>, dim3(gv), dim3(blk), 0, st, g, d_pos, p.d_species, p.d_i0,
                     (T *)p.d_phi, with_dphi ? (T *)p.d_dphi : (T *)nullptr, p.d_key,
                     p.d_count);
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, st, p.d_count, p.d_start, p.d_cursor,
                     nbins_tot);
  hipLaunchKernelGGL(k_scatter, dim3(ga), dim3(blk), 0, st, (int64_t)g.natoms, p.d_key,
                     p.d_cursor, p.d_sidx);
  hipLaunchKernelGGL(k_binsort, dim3((nbins_tot + 3) / 4), dim3(256), 0, st, nbins_tot, p.d_start,
                     p.d_sidx, p.d_i0, p.d_i0s);
}

// The forward pass in three parts: the atom sort and the spread write W1;
// the y pass reads W1 and writes W2; the x pass and the finalize read W2.  A
// captured batch runs the sort and spread of frame i + 1 beside the y and x
// passes of frame i (two streams, W1 double-buffered).
template <typename T>
static int forward_front_t(Plan &p, const double *d_pos, void *W1, hipStream_t st) {
  using V = typename C2<T>::t;
  Geo &g = p.geo;
  sort_atoms<T>(p, d_pos, false, st);
  {
    StageTimer t(p, 1);
    FixedSizes<T>::pick(g.n[2], [&](auto n) {
      auto go = [&](auto tt) {
        hipLaunchKernelGGL((k_spread_z<T, decltype(n)::value, decltype(tt)::value>),
                           dim3(g.ntx * g.nty, g.F), dim3(NT), p.lds1, st, g, p.ax[2], p.d_i0s,
                           (const T *)p.d_phi, p.d_sidx, p.d_start, p.d_species,
                           (const V *)p.d_tw[2], (V *)W1, p.d_wtab);
      };
      if (g.txr == 8 && g.ty == 2)
        go(std::integral_constant<int, 1>{});
      else
        go(std::integral_constant<int, 0>{});
    });
  }
  SQC_CHECK(hipGetLastError());
  return 0;
}

template <typename T>
static int forward_y_t(Plan &p, const void *W1, void *W2, hipStream_t st) {
  using V = typename C2<T>::t;
  Geo &g = p.geo;
  {
    StageTimer t(p, 2);
    FixedSizes<T>::pick(g.n[1], [&](auto n) {
      pick_chunk(p.tx2, [&](auto c) {
        hipLaunchKernelGGL((k_fft_y<T, decltype(n)::value, decltype(c)::value>),
                           dim3((g.n[0] + p.tx2 - 1) / p.tx2 * (g.M[2] + 1) * g.F), dim3(NT),
                           p.lds2, st, g, p.ax[1], (const V *)W1, (const V *)p.d_tw[1], p.d_lo2,
                           p.d_hi2, p.d_loff, (V *)W2);
      });
    });
  }
  SQC_CHECK(hipGetLastError());
  return 0;
}

template <typename T>
static int forward_x_t(Plan &p, const void *W2, hipStream_t st, int mode, double *frame_out,
                       void *rho_out) {
  using V = typename C2<T>::t;
  Geo &g = p.geo;
  const int ngroups = (g.nlines + p.lb - 1) / p.lb;
  {
    StageTimer t(p, 3);
    FixedSizes<T>::pick(g.n[0], [&](auto n) {
      pick_fields(g, [&](auto fc) {
        auto go = [&](auto lbc) {
          hipLaunchKernelGGL(
              (k_fft_x_bin<T, decltype(n)::value, decltype(fc)::value, decltype(lbc)::value>),
              dim3(p.nblk3), dim3(NT), p.lds3, st, g, p.ax[0], p.lb, ngroups, (const V *)W2,
              (const V *)p.d_tw[0], (const int4 *)p.d_line_info, p.d_line_off, p.d_stab,
              p.d_ph[0], p.d_line_d23, p.d_wgpart, (V *)rho_out, (int64_t)p.ncand,
              mode == 0 ? p.d_gnum : nullptr, mode == 0 ? p.d_gpart : nullptr, p.d_wtab);
        };
        if (p.lb == 2)
          go(std::integral_constant<int, 2>{});
        else
          go(std::integral_constant<int, 0>{});
      });
    });
  }
  SQC_CHECK(hipGetLastError());
  return launch_finalize(p, frame_out, mode == 0, st);
}

int launch_finalize(Plan &p, double *frame_out, int accumulate, hipStream_t st) {
  const int ntot = p.geo.ncomp * p.geo.nq;
  StageTimer t(p, 4);
  const Geo &g = p.geo;
  hipLaunchKernelGGL(k_finalize, dim3((ntot + kFinEntries - 1) / kFinEntries), dim3(256), 0,
                     st, p.d_wgpart, p.nblk3, ntot, g.xwords, g.nq, 1.0 / g.xq_inv[0],
                     1.0 / g.xq_inv[1], 1.0 / g.xq_lo, p.d_acc, frame_out, accumulate);
  SQC_CHECK(hipGetLastError());
  return 0;
}

int launch_forward_front(Plan &p, const double *d_pos, void *W1, hipStream_t st) {
  return p.prec == 1 ? forward_front_t<float>(p, d_pos, W1, st)
                     : forward_front_t<double>(p, d_pos, W1, st);
}

int launch_forward_y(Plan &p, const void *W1, void *W2, hipStream_t st) {
  return p.prec == 1 ? forward_y_t<float>(p, W1, W2, st) : forward_y_t<double>(p, W1, W2, st);
}

int launch_forward_x(Plan &p, const void *W2, hipStream_t st, int mode, double *frame_out,
                     void *rho_out) {
  return p.prec == 1 ? forward_x_t<float>(p, W2, st, mode, frame_out, rho_out)
                     : forward_x_t<double>(p, W2, st, mode, frame_out, rho_out);
}

int launch_forward(Plan &p, const double *d_pos, int mode, double *frame_out, void *rho_out) {
  if (int rc = launch_forward_front(p, d_pos, p.d_W1, p.stream)) return rc;
  if (int rc = launch_forward_y(p, p.d_W1, p.d_W2, p.stream)) return rc;
  return launch_forward_x(p, p.d_W2, p.stream, mode, frame_out, rho_out);
}

template <typename T>
static int adjoint_t(Plan &p, const double *d_pos, const void *rho, const double *gS,
                     const double *gP, double *grad, double *slots) {
  using V = typename C2<T>::t;
  Geo &g = p.geo;
  hipStream_t st = p.stream;
  const int blk = 256;
  const unsigned ga = (unsigned)((g.natoms + blk - 1) / blk);
  const int Fa = g.F;
  // the atoms of this frame, sorted exactly as in the forward pass
  sort_atoms<T>(p, d_pos, true, st);
  const int ngroups = (g.nlines + p.lb - 1) / p.lb;
  {
    StageTimer t(p, 5);
    FixedSizes<T>::pick(g.n[0], [&](auto n) {
      pick_fields(g, [&](auto fc) {
        hipLaunchKernelGGL((k_adj_x<T, decltype(n)::value, decltype(fc)::value>), dim3(p.nblk3),
                           dim3(NT), p.lds3, st, g, p.ax[0], p.lb, ngroups, Fa, (const V *)rho,
                           (int64_t)p.ncand, (const V *)p.d_tw[0], (const int4 *)p.d_line_info,
                           p.d_line_off, p.d_stab, p.d_ph[0], p.d_line_d23, p.d_den, p.d_cnt,
                           gS, gP, (V *)p.d_W2, p.d_wtab);
      });
    });
  }
  {
    StageTimer t(p, 6);
    FixedSizes<T>::pick(g.n[1], [&](auto n) {
      pick_chunk(p.tx2, [&](auto c) {
        hipLaunchKernelGGL((k_adj_y<T, decltype(n)::value, decltype(c)::value>),
                           dim3((g.n[0] + p.tx2 - 1) / p.tx2 * (g.M[2] + 1) * Fa), dim3(NT),
                           p.lds2, st, g, p.ax[1], (const V *)p.d_W2, (const V *)p.d_tw[1],
                           p.d_lo2, p.d_hi2, p.d_loff, (V *)p.d_W1);
      });
    });
  }
  {
    StageTimer t(p, 7);
    FixedSizes<T>::pick(g.n[2], [&](auto n) {
      auto go = [&](auto tt) {
        hipLaunchKernelGGL((k_adj_z_gather<T, decltype(n)::value, decltype(tt)::value>),
                           dim3(g.ntx * g.nty, Fa), dim3(NT), p.lds1g, st, g, p.ax[2],
                           (const V *)p.d_W1, (const V *)p.d_tw[2], p.d_i0s,
                           (const T *)p.d_phi, (const T *)p.d_dphi, p.d_sidx, p.d_start,
                           slots);
      };
      if (g.txr == 8 && g.ty == 2)
        go(std::integral_constant<int, 1>{});
      else
        go(std::integral_constant<int, 0>{});
    });
  }
  {
    StageTimer t(p, 8);
    hipLaunchKernelGGL(k_adj_sum, dim3(ga), dim3(blk), 0, st, g, p.d_i0, p.d_species, slots,
                       Fa, grad, p.d_wtab);
  }
  SQC_CHECK(hipGetLastError());
  return 0;
}

// slots: 3 doubles per (atom, tile) pair, [natoms][sx][sy][3]; every slot
// that k_adj_sum reads is written by k_adj_z_gather in the same launch
// sequence, so the buffer needs no clearing.
int launch_adjoint(Plan &p, const double *d_pos, const void *rho, const double *gS,
                   const double *gP, double *grad, double *slots) {
  return p.prec == 1 ? adjoint_t<float>(p, d_pos, rho, gS, gP, grad, slots)
                     : adjoint_t<double>(p, d_pos, rho, gS, gP, grad, slots);
}

// Raise the dynamic-LDS ceiling of each kernel to what the plan needs.
template <typename T>
static int lds_limits_t(Plan &p) {
  int err = 0;
  auto set = [&](const void *f, size_t bytes) {
    if (err) return;
    const hipError_t e =
        hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      err = (int)e;
    }
  };
  FixedSizes<T>::pick(p.geo.n[2], [&](auto n) {
    set((const void *)k_spread_z<T, decltype(n)::value, 0>, p.lds1);
    set((const void *)k_spread_z<T, decltype(n)::value, 1>, p.lds1);
    set((const void *)k_adj_z_gather<T, decltype(n)::value, 0>, p.lds1g);
    set((const void *)k_adj_z_gather<T, decltype(n)::value, 1>, p.lds1g);
  });
  FixedSizes<T>::pick(p.geo.n[1], [&](auto n) {
    pick_chunk(p.tx2, [&](auto c) {
      set((const void *)k_fft_y<T, decltype(n)::value, decltype(c)::value>, p.lds2);
    });
    pick_chunk(p.tx2, [&](auto c) {
      set((const void *)k_adj_y<T, decltype(n)::value, decltype(c)::value>, p.lds2);
    });
  });
  FixedSizes<T>::pick(p.geo.n[0], [&](auto n) {
    pick_fields(p.geo, [&](auto fc) {
      set((const void *)k_fft_x_bin<T, decltype(n)::value, decltype(fc)::value, 0>, p.lds3);
      set((const void *)k_fft_x_bin<T, decltype(n)::value, decltype(fc)::value, 2>, p.lds3);
    });
    pick_fields(p.geo, [&](auto fc) {
      set((const void *)k_adj_x<T, decltype(n)::value, decltype(fc)::value>, p.lds3);
    });
  });
  return err;
}

// Workgroups of the persistent x pass: what the CUs hold at once by registers
// and LDS together (the occupancy calculator of the instance that runs),
// never more than there is work for.
template <typename T>
static void persistent_grids_t(Plan &p) {
  const Geo &g = p.geo;
  int nx = 0;
  FixedSizes<T>::pick(g.n[0], [&](auto n) {
    pick_fields(g, [&](auto fc) {
      auto query = [&](auto lbc) {
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &nx, k_fft_x_bin<T, decltype(n)::value, decltype(fc)::value, decltype(lbc)::value>,
                NT, p.lds3) != hipSuccess)
          nx = 0;
      };
      if (p.lb == 2)
        query(std::integral_constant<int, 2>{});
      else
        query(std::integral_constant<int, 0>{});
    });
  });
  (void)hipGetLastError();
  if (nx > 0) {
    const int ngroups = (g.nlines + p.lb - 1) / p.lb;
    p.nblk3 = std::max(1, std::min(ngroups, nx * p.ncu));
    cap_wgpart_rows(p);
  }
}

void persistent_grids(Plan &p) {
  if (p.prec == 1)
    persistent_grids_t<float>(p);
  else
    persistent_grids_t<double>(p);
}

size_t static_lds_spread() { return sizeof(Stage); }
size_t static_lds_xpass() { return sizeof(LineGroup); }

int kernel_lds_limits(Plan &p) {
  return p.prec == 1 ? lds_limits_t<float>(p) : lds_limits_t<double>(p);
}

}  // namespace sqc
