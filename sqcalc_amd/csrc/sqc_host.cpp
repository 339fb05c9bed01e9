// sqc_host.cpp -- plan construction and the C ABI of libsqc_hip.
//
// Plan creation restates, for the GPU library, what the reference computes on
// the host before the first frame:
//   cell set-up, B = 2 pi A^-T              src/sqc_cell.f90:76-98
//   m_i, kept modes, shells, den, counts    src/sqc_structure_factor.f90:327-451
//   kernel width / shape, fine grid         external/finufft/src/cuda/spreadinterp.cpp:61-82,
//                                           external/finufft/src/finufft_core.cpp:81-99
//   kernel Fourier coefficients             external/finufft/src/cuda/common.cu:27-66
// and then lays the kept modes out for the device: x-lines of the canonical
// half spectrum, a u16 shell per candidate mode, and the tile shapes of the
// three DFT passes.  Every |q| and shell index is evaluated with the
// reference's expression order and FP contraction off, so the shell of each
// mode is the reference's, bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "sqc_internal.hpp"

#pragma clang fp contract(off)

struct sqc_hip_plan {
  sqc::Plan p;
  hipStream_t own_stream = nullptr;
  // scratch of the host-memory convenience entry points
  double *g_pos = nullptr, *g_out = nullptr, *g_gs = nullptr, *g_grad = nullptr;
  void *g_rho = nullptr;
};

// Device scratch of sqc_hip_adjoint_device_noalloc: the adjoint's gradient
// slots (3 doubles per (atom, tile) pair) for callers that cannot allocate.
// Each GPU the library runs on holds one copy until the process ends, and
// every call on that GPU shares it.  In host code the name is only a
// stand-in: kernels get the device address from hipGetSymbolAddress.
static __device__ __align__(256) double scratch[SQC_HIP_NOALLOC_SCRATCH_BYTES / sizeof(double)];

namespace {

thread_local std::string g_last_error;

// Device address of `scratch` per device ordinal, looked up once per device
// when the first plan on that device is made (the device is current then).
std::mutex g_scratch_mu;
std::vector<double *> g_scratch_ptr;

int fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

int hip_fail(hipError_t e, const char *what) {
  g_last_error = std::string(what) + ": " + hipGetErrorString(e);
  return SQC_HIP_ERR_HIP;
}

#define HIP_TRY(call, what)                          \
  do {                                               \
    hipError_t e_ = (call);                          \
    if (e_ != hipSuccess) return hip_fail(e_, what); \
  } while (0)

// Look up (once per device, with the device current) where this device holds
// the no-allocation scratch.
int scratch_setup(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize(device + 1, nullptr);
  if (g_scratch_ptr[device]) return 0;
  void *ptr = nullptr;
  HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
  if (!ptr) return fail(SQC_HIP_ERR_HIP, "hipGetSymbolAddress(scratch) returned no address");
  g_scratch_ptr[device] = static_cast<double *>(ptr);
  return 0;
}

double *scratch_of(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return device < (int)g_scratch_ptr.size() ? g_scratch_ptr[device] : nullptr;
}

double det3(const double *m) {
  auto M = [&](int i, int j) { return m[i + 3 * j]; };
  return M(0, 0) * (M(1, 1) * M(2, 2) - M(1, 2) * M(2, 1)) -
         M(0, 1) * (M(1, 0) * M(2, 2) - M(1, 2) * M(2, 0)) +
         M(0, 2) * (M(1, 0) * M(2, 1) - M(1, 1) * M(2, 0));
}

// Cofactor inverse in the reference's term order (src/sqc_cell.f90:135-151).
void inverse3(const double *m, double *inv) {
  auto M = [&](int i, int j) { return m[i + 3 * j]; };
  auto I = [&](int i, int j) -> double & { return inv[i + 3 * j]; };
  const double det = det3(m);
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

struct ModeGeom {
  double b[9];
  double qmin, qmax, dq;
  int nq;
  // |q| of h with the reference's order: q_i = (b1_i h1 + b2_i h2) + b3_i h3
  double qlen(int h1, int h2, int h3, double *q = nullptr) const {
    const double r1 = h1, r2 = h2, r3 = h3;
    double v[3];
    for (int i = 0; i < 3; ++i) v[i] = b[i] * r1 + b[i + 3] * r2 + b[i + 6] * r3;
    if (q) {
      q[0] = v[0];
      q[1] = v[1];
      q[2] = v[2];
    }
    return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  }
  bool kept(double ql) const { return !(ql > qmax || ql < qmin); }
  int shell(double ql) const {  // 0-based
    int s = (int)((ql - qmin) / dq) + 1;
    if (s < 1) s = 1;
    if (s > nq) s = nq;
    return s - 1;
  }
};

// Smallest even 2,3,5-smooth integer >= n (external/finufft/src/finufft_core.cpp:81-99).
int64_t next235even(int64_t n) {
  if (n <= 2) return 2;
  if (n % 2) ++n;
  for (;; n += 2) {
    int64_t k = n;
    for (int p : {2, 3, 5})
      while (k % p == 0) k /= p;
    if (k == 1) return n;
  }
}

bool factor(int n, sqc::AxisPlan &ax) {
  ax.n = n;
  ax.nst = 0;
  for (int r : {8, 5, 4, 3, 2}) {
    while (n % r == 0) {
      if (ax.nst >= sqc::kMaxStages) return false;
      ax.rad[ax.nst++] = r;
      n /= r;
    }
  }
  // magic multipliers for j / Ns at every stage
  uint64_t Ns = 1;
  for (int s = 0; s < ax.nst; ++s) {
    ax.magic[s] = Ns == 1 ? 0u : (uint32_t)(((1ULL << 32) + Ns - 1) / Ns);
    Ns *= (uint64_t)ax.rad[s];
  }
  return n == 1;
}

// Gauss-Legendre nodes/weights on [-1, 1] by Newton iteration on P_q.
void gauss_legendre(int q, std::vector<double> &x, std::vector<double> &w) {
  x.assign(q, 0.0);
  w.assign(q, 0.0);
  const double pi = 3.14159265358979323846;
  for (int i = 0; i < (q + 1) / 2; ++i) {
    double z = std::cos(pi * (i + 0.75) / (q + 0.5));
    double dp = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = z;
      for (int k = 2; k <= q; ++k) {
        const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      dp = q * (z * p1 - p0) / (z * z - 1.0);
      const double dz = p1 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    x[i] = -z;
    x[q - 1 - i] = z;
    w[i] = w[q - 1 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
  }
}

// phihat(h) = int phi(2u/ns) cos(2 pi h u / nf) du for h = 0..hmax.
std::vector<double> phihat(int nf, int hmax, int ns, double beta) {
  std::vector<double> x, w;
  gauss_legendre(std::max(4 * ns, 64), x, w);
  const double pi = 3.14159265358979323846;
  std::vector<double> out(hmax + 1, 0.0);
  for (int h = 0; h <= hmax; ++h) {
    double s = 0.0;
    for (size_t i = 0; i < x.size(); ++i) {
      const double u = 0.5 * ns * x[i];
      const double ker = std::exp(beta * (std::sqrt(1.0 - x[i] * x[i]) - 1.0));
      s += 0.5 * ns * w[i] * ker * std::cos(2.0 * pi * h * u / nf);
    }
    out[h] = s;
  }
  return out;
}

// exp(+2 pi i t / n) for t = 0..n-1, using the symmetry of the circle so that
// every entry comes from an angle in [0, pi/4].
void twiddles(int n, std::vector<double> &re, std::vector<double> &im) {
  re.resize(n);
  im.resize(n);
  const double pi = 3.14159265358979323846;
  for (int t = 0; t < n; ++t) {
    // reduce t/n to an octant: angle = 2 pi t / n
    const long long k8 = 8LL * t;  // octant index * n
    const int oct = (int)(k8 / n);
    const double rem = (double)(k8 - (long long)oct * n) / n;  // in [0,1)
    double a = rem * pi / 4.0;
    double c, s;
    // angle = oct * pi/4 + a
    switch (oct) {
      case 0: c = std::cos(a); s = std::sin(a); break;
      case 1: c = std::sin(pi / 4.0 - a); s = std::cos(pi / 4.0 - a); break;
      case 2: c = -std::sin(a); s = std::cos(a); break;
      case 3: c = -std::cos(pi / 4.0 - a); s = std::sin(pi / 4.0 - a); break;
      case 4: c = -std::cos(a); s = -std::sin(a); break;
      case 5: c = -std::sin(pi / 4.0 - a); s = -std::cos(pi / 4.0 - a); break;
      case 6: c = std::sin(a); s = -std::cos(a); break;
      default: c = std::cos(pi / 4.0 - a); s = -std::sin(pi / 4.0 - a); break;
    }
    re[t] = c;
    im[t] = s;
  }
}

template <typename T>
int upload(T **dst, const void *src, size_t bytes, int64_t &total) {
  const size_t nb = bytes ? bytes : 8;
  hipError_t e = hipMalloc((void **)dst, nb);
  if (e != hipSuccess) return hip_fail(e, "hipMalloc");
  total += (int64_t)nb;
  if (src && bytes) {
    e = hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpy");
  }
  return 0;
}

template <typename T>
int alloc(T **dst, size_t bytes, int64_t &total) {
  return upload(dst, nullptr, bytes, total);
}

int upload_twiddles(sqc::Plan &p, int axis) {
  std::vector<double> re, im;
  twiddles(p.geo.n[axis], re, im);
  const int n = p.geo.n[axis];
  if (p.prec == 1) {
    std::vector<float> buf(2 * n);
    for (int t = 0; t < n; ++t) {
      buf[2 * t] = (float)re[t];
      buf[2 * t + 1] = (float)im[t];
    }
    return upload((float **)&p.d_tw[axis], buf.data(), buf.size() * sizeof(float), p.bytes);
  }
  std::vector<double> buf(2 * n);
  for (int t = 0; t < n; ++t) {
    buf[2 * t] = re[t];
    buf[2 * t + 1] = im[t];
  }
  return upload((double **)&p.d_tw[axis], buf.data(), buf.size() * sizeof(double), p.bytes);
}

double amplitude(const sqc::Plan &p, int sp, double q) {
  const double *w = p.wtab.data() + (size_t)sp * sqc::kWtab;
  if (p.geo.weight_kind == SQC_HIP_WEIGHT_XRAY) {
    const double *c = w + 1;
    double st = q / (4.0 * std::acos(-1.0));
    st = st * st;
    double f = c[8];
    for (int i = 0; i < 4; ++i) f = f + c[i] * std::exp(-c[4 + i] * st);
    return f;
  }
  return w[0];
}

// Quanta of the exact shell sums (the x pass's exact_parts).  A value added
// into a shell is 2 |sum_a w_a rho_a|^2 (the total) or 2 ps Re(rho_a conj
// rho_b) (a pair), and |rho_a| <= N_a for a field of unit strengths
// (|rho| <= sum_a |w_a| N_a for one weighted field), so |v| <= vmax below,
// with a factor 4 for the transform's own error; a shell takes at most cmax
// values per component in a frame.  q = 2^e with cmax vmax / q <= 2^62 keeps
// every sum of hi words (each at most vmax / q + 1/2) below 2^63, and
// vmax / q <= 2^50 keeps each within the range of the rounding step; the lo
// words are at most 2^(k-1) each, and 2^k <= 2^62 / 2^ceil(log2(cmax + 1))
// (and <= 2^51) keeps their sums below 2^61.  One word carries a value to
// q / 2, about
// 1e-8 of a typical mode at 96,000 atoms, where vmax is the Bragg bound N^2
// times the weights: FP32 plans (whose transform is good to 1e-7) keep one,
// FP64 plans two (q 2^-k / 2, far below the FP64 rounding of the sums).
int ceil_log2(double x) { return (int)std::ceil(std::log2(x)); }

void set_shell_quanta(sqc::Plan &p, double qmax) {
  sqc::Geo &g = p.geo;
  const int S = g.S;
  const double smax = qmax / (4.0 * std::acos(-1.0));
  double A = 0.0, nn = 0.0;
  for (int a = 0; a < S; ++a) {
    double w;
    const double *wa = p.wtab.data() + (size_t)a * sqc::kWtab;
    if (g.weight_kind == SQC_HIP_WEIGHT_XRAY) {
      // |f_a(q)| <= sum_i |a_i| max(1, exp(-b_i s^2)) + |c| over the q range
      const double *c = wa + 1;
      w = std::fabs(c[8]);
      for (int i = 0; i < 4; ++i)
        w += std::fabs(c[i]) * std::max(1.0, std::exp(-c[4 + i] * smax * smax));
    } else {
      w = std::fabs(wa[0]);
    }
    A += w * (double)p.counts[a];
    for (int b = a; b < S; ++b) nn = std::max(nn, (double)p.counts[a] * (double)p.counts[b]);
  }
  int64_t cmax = 1;
  for (int64_t c : p.shell_count) cmax = std::max(cmax, c);
  const double tiny = std::ldexp(1.0, -900);
  const double vmax[2] = {4.0 * 2.0 * A * A, 4.0 * 2.0 * nn};
  for (int c = 0; c < 2; ++c)
    g.xq_inv[c] = std::ldexp(1.0, std::min(62 - ceil_log2(std::max((double)cmax * vmax[c], tiny)),
                                           50 - ceil_log2(std::max(vmax[c], tiny))));
  g.xq_lo = std::ldexp(1.0, std::min(51, 62 - ceil_log2((double)cmax + 1.0)));
  g.xwords = p.prec == 1 ? 1 : 2;
}

// Spread tiles (txr real x-columns, ty y-columns), largest first.  Every
// tile has txr * ty <= 32, which keeps a stencil's xy footprint inside a
// tile at 64 cells or fewer (the spreader's per-lane xy table relies on it),
// and txr even (two real columns per complex DFT line).
constexpr int kTiles[][2] = {{16, 2}, {8, 2}, {16, 1}, {4, 2}, {8, 1}, {4, 1}, {2, 1}};
constexpr int kNumTiles = 7;
constexpr size_t kLdsCap = 160 * 1024;
// Default LDS budgets when the tuner is off.  FP32: at least three spread and
// x-pass workgroups and four y-pass workgroups resident per CU, so that the
// global loads of some overlap the DFTs of others.  FP64 lines are twice as
// large; there two spread/x-pass workgroups and a 64 KB y pass do better.
size_t lds_pref(const sqc::Plan &p) { return p.prec == 1 ? 56 * 1024 : 80 * 1024; }
size_t lds_pref_y(const sqc::Plan &p) { return p.prec == 1 ? 40 * 1024 : 64 * 1024; }

size_t vsize(const sqc::Plan &p) { return p.prec == 1 ? 8 : 16; }

// Scratch lines of the long-line DFT: one line per wave for axes longer
// than the in-place DFT takes.
size_t long_scratch(const sqc::Plan &p, int n) {
  return n > sqc::kMaxLine ? (size_t)(sqc::kThreads / 64) * n * vsize(p) : 0;
}

size_t spread_lds(const sqc::Plan &p, int txr, int ty) {
  const sqc::Geo &g = p.geo;
  // the tile and the staged atom records (or, during the DFT, the long-line
  // scratch in their place); the DFT twiddles are read from global memory
  // (L1-resident), which leaves room for more workgroups
  return (size_t)(txr / 2) * ty * g.cs * vsize(p) +
         std::max<size_t>(sqc::kStageCap * 16, long_scratch(p, g.n[2]));
}

// The gather adds a per-wave scratch of the kernel values of the atoms of
// one round: (64 / lanes per atom) atoms x 6 ns values.
size_t gather_lds(const sqc::Plan &p, int txr, int ty) {
  const sqc::Geo &g = p.geo;
  const int lpa = std::max(8, txr * ty);
  const size_t after = sqc::kStageCap * 16 +
                       (size_t)(sqc::kThreads / 64) * (64 / lpa) * 6 * g.ns * (vsize(p) / 2);
  return (size_t)(txr / 2) * ty * g.cs * vsize(p) + std::max(after, long_scratch(p, g.n[2]));
}

// LDS a CU offers the dynamic part of an x-pass workgroup.
size_t xpass_cap() { return kLdsCap - sqc::static_lds_xpass(); }

// y-line stride: the smallest odd stride that holds a line (n2 + n2/32 + 1
// points, the DFT's padded stage layout).  An odd stride spreads the
// transposed copies (lane (line, y) at line * ls2 + y) over the banks and
// keeps five 8-line workgroups per CU at n2 = 480; strides tuned to the bank
// count measured no faster.
int ystride(const sqc::Plan &p, int tx2) {
  (void)tx2;
  const int cap = p.geo.n[1] + p.geo.n[1] / 32 + 1;
  return cap | 1;
}

size_t ypass_lds(const sqc::Plan &p, int tx2) {
  return (size_t)tx2 * ystride(p, tx2) * vsize(p) + long_scratch(p, p.geo.n[1]);
}

int ilog2i(int v) {
  int l = 0;
  while ((1 << l) < v) ++l;
  return l;
}

// Whether a spread tile is usable: LDS within `limit`, at least two tiles per
// axis, and a bin window of at most 128 bins.
bool tile_fits(const sqc::Plan &p, int txr, int ty, size_t limit) {
  const sqc::Geo &g = p.geo;
  const int n1 = g.n[0], n2 = g.n[1], ns = g.ns;
  if (spread_lds(p, txr, ty) + sqc::static_lds_spread() > limit) return false;
  // the gather holds the same tile plus its kernel-value scratch
  if (gather_lds(p, txr, ty) + sqc::static_lds_spread() > kLdsCap) return false;
  if (2 * txr > n1 || 2 * ty > n2) return false;
  const int ntx = (n1 + txr - 1) / txr, nty = (n2 + ty - 1) / ty;
  const int hx = (int)std::ceil((0.5 * ns + 1.0) / txr) + (n1 % txr ? 1 : 0);
  const int hy = (int)std::ceil((0.5 * ns + 1.0) / ty) + (n2 % ty ? 1 : 0);
  return std::min(2 * hx + 1, ntx) * std::min(2 * hy + 1, nty) <= 128;
}

// Everything that follows from the spread tile and the y-pass column chunk.
void set_tiles(sqc::Plan &p, int txr, int ty, int tx2) {
  sqc::Geo &g = p.geo;
  const int n1 = g.n[0], n2 = g.n[1], ns = g.ns;
  g.txr = txr;
  g.ty = ty;
  p.lds1 = spread_lds(p, txr, ty);
  p.lds1g = gather_lds(p, txr, ty);
  g.ltxr = ilog2i(g.txr);
  g.lty = ilog2i(g.ty);
  g.ntx = (n1 + g.txr - 1) / g.txr;
  g.nty = (n2 + g.ty - 1) / g.ty;
  g.nbins = g.ntx * g.nty;
  // a stencil that wraps past a narrow last tile can come from one bin further
  g.hx = (int)std::ceil((0.5 * ns + 1.0) / g.txr) + (n1 % g.txr ? 1 : 0);
  g.hy = (int)std::ceil((0.5 * ns + 1.0) / g.ty) + (n2 % g.ty ? 1 : 0);
  g.sx = std::min(g.ntx, (ns - 1 + g.txr - 1) / g.txr + 2);
  g.sy = std::min(g.nty, (ns - 1 + g.ty - 1) / g.ty + 2);
  p.tx2 = tx2;
  g.ls2 = ystride(p, tx2);
  p.lds2 = ypass_lds(p, tx2);
}

// Bin counters of the atom sort (sized by the tile grid); the count array
// starts at zero and the scan leaves it at zero.
int alloc_bins(sqc::Plan &p) {
  const int nbt = p.geo.F * p.geo.nbins;
  for (int32_t **b : {&p.d_count, &p.d_start, &p.d_cursor}) {
    if (*b) {
      (void)hipFree(*b);
      p.bytes -= (int64_t)sizeof(int32_t) * (p.nbt_alloc + 1);
      *b = nullptr;
    }
  }
  int rc;
  if ((rc = alloc(&p.d_count, sizeof(int32_t) * (nbt + 1), p.bytes))) return rc;
  if ((rc = alloc(&p.d_start, sizeof(int32_t) * (nbt + 1), p.bytes))) return rc;
  if ((rc = alloc(&p.d_cursor, sizeof(int32_t) * (nbt + 1), p.bytes))) return rc;
  p.nbt_alloc = nbt;
  HIP_TRY(hipMemset(p.d_count, 0, sizeof(int32_t) * (nbt + 1)), "hipMemset");
  return 0;
}

// Shell sums kept in HBM rows (many species with partials) are at most
// 512 MB together: fewer persistent workgroups beyond that.
void cap_rows(sqc::Plan &p) {
  if (!p.geo.acc_global) return;
  const size_t row = (size_t)p.geo.xwords * p.geo.ncomp * p.geo.nq * sizeof(uint64_t);
  const int cap = (int)std::max<size_t>(1, ((size_t)512 << 20) / row);
  p.nblk3 = std::min(p.nblk3, std::max(cap, std::min(p.ncu, p.nblk3)));
}

// x pass: lb lines x F fields of n1 points, the twiddles, the shell
// accumulators and the 1 / phihat_1 table.
size_t xpass_lds(const sqc::Plan &p, int lb) {
  const sqc::Geo &g = p.geo;
  // shell sums (the integer words of ncomp rows of nq) unless they live in
  // HBM; the adjoint x pass keeps two rows of nq there
  const size_t rows = g.acc_global ? 2 : std::max<size_t>(2, (size_t)g.xwords * g.ncomp);
  const size_t accb = (rows * g.nq + g.M[0] + 1) * sizeof(double);
  const size_t ls = long_scratch(p, g.n[0]);
  return ((size_t)lb * g.F * g.rs1 + g.n[0]) * vsize(p) + accb + (ls ? 16 + ls : 0);
}

// Lines per group and the persistent grid: as many workgroups as the CUs
// hold at once (LDS-limited, at most 8 of 4 waves per CU).
void set_xpass(sqc::Plan &p, int lb) {
  p.lb = lb;
  p.lds3 = xpass_lds(p, lb);
  const int resident = (int)std::max<size_t>(
      1, std::min<size_t>(8, kLdsCap / (p.lds3 + sqc::static_lds_xpass())));
  const int ngroups = (p.geo.nlines + lb - 1) / lb;
  p.nblk3 = std::max(1, std::min(ngroups, resident * p.ncu));
  cap_rows(p);
}

// Per-workgroup shell rows of the x pass (one row per persistent workgroup).
int alloc_wgpart(sqc::Plan &p) {
  const size_t row = (size_t)p.geo.xwords * p.geo.ncomp * p.geo.nq;
  const size_t n = row * p.nblk3;
  if (p.d_wgpart && p.wgpart_rows == p.nblk3) return 0;
  if (p.d_wgpart) {
    (void)hipFree(p.d_wgpart);
    p.bytes -= (int64_t)(sizeof(uint64_t) * row * p.wgpart_rows);
    p.d_wgpart = nullptr;
  }
  if (int rc = alloc(&p.d_wgpart, sizeof(uint64_t) * n, p.bytes)) return rc;
  p.wgpart_rows = p.nblk3;
  return 0;
}

void drop_graph(sqc::Plan &p) {
  if (p.gexec) (void)hipGraphExecDestroy(p.gexec);
  p.gexec = nullptr;
}

void free_plan(sqc_hip_plan *h) {
  sqc::Plan &p = h->p;
  void *bufs[] = {p.d_pos, p.d_i0, p.d_i0s, p.d_phi, p.d_dphi, p.d_key, p.d_count, p.d_start, p.d_cursor, p.d_sidx,
                  p.d_species, p.d_W1, p.d_W1b, p.d_W2, p.d_tw[0], p.d_tw[1], p.d_tw[2], p.d_ph[0],
                  p.d_ph[1], p.d_ph[2], p.d_lo2, p.d_hi2, p.d_loff, p.d_line_info,
                  p.d_line_off, p.d_line_d23, p.d_stab, p.d_den, p.d_cnt, p.d_wtab, p.d_wgpart, p.d_acc,
                  p.d_frame,
                  p.d_gnum, p.d_gpart, p.d_slots, p.d_gs, p.d_rho, h->g_pos, h->g_out, h->g_gs,
                  h->g_grad, h->g_rho};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  if (!p.pending.empty()) (void)hipStreamSynchronize(p.stream);
  for (auto &q : p.pending) {
    (void)hipEventDestroy(q.a);
    (void)hipEventDestroy(q.b);
  }
  for (hipEvent_t e : p.ev_free) (void)hipEventDestroy(e);
  drop_graph(p);
  for (hipEvent_t e : p.gev) (void)hipEventDestroy(e);
  if (p.stream2) (void)hipStreamDestroy(p.stream2);
  if (p.ev_in) (void)hipEventDestroy(p.ev_in);
  if (p.ev_out) (void)hipEventDestroy(p.ev_out);
  if (h->own_stream) (void)hipStreamDestroy(h->own_stream);
  delete h;
}

int build_plan(sqc_hip_plan *h, const sqc_hip_config *cfg) {
  sqc::Plan &p = h->p;
  sqc::Geo &g = p.geo;
  p.cfg = *cfg;
  p.prec = cfg->precision;
  const int S = cfg->nspecies;
  g.S = S;
  g.natoms = cfg->natoms;
  g.nq = cfg->nq;
  g.weight_kind = cfg->weight_kind;
  g.partials = cfg->partials ? 1 : 0;
  p.wtab.assign((size_t)S * sqc::kWtab, 0.0);
  for (int a = 0; a < S; ++a) {
    // strengths: 1 under unit and X-ray weights (the X-ray amplitude is
    // f_a(|q|), applied in the x-pass epilogue), b_a under neutron weights
    double *w = p.wtab.data() + (size_t)a * sqc::kWtab;
    w[0] = cfg->weight_kind == SQC_HIP_WEIGHT_NEUTRON ? cfg->amp_const[a] : 1.0;
    for (int i = 0; i < 9; ++i) w[1 + i] = cfg->xray_coef ? cfg->xray_coef[9 * a + i] : 0.0;
    if (a < sqc::kGeoSpecies) {
      g.amp[a] = w[0];
      for (int i = 0; i < 9; ++i) g.xcoef[a][i] = w[1 + i];
    }
  }
  p.species_of.assign(cfg->species_of, cfg->species_of + cfg->natoms);
  p.counts.assign(S, 0);
  for (int64_t j = 0; j < cfg->natoms; ++j) {
    const int s = p.species_of[j];
    if (s < 0 || s >= S) return fail(SQC_HIP_ERR_BAD_ARG, "species_of out of range");
    ++p.counts[s];
  }
  // fields: one weighted transform unless the pairs or q-dependent weights
  // need each species on its own (src/sqc_gpu.f90:193-197)
  const bool per_species = cfg->partials || cfg->weight_kind == SQC_HIP_WEIGHT_XRAY;
  g.F = per_species ? S : 1;
  g.ncomp = 1 + ((cfg->partials && S > 0) ? S * (S + 1) / 2 : 0);
  // a single species under neutron or unit weights is spread with its weight
  // as the strength; its pair column is of the unweighted amplitude
  g.pair_scale = 1.0;
  if (g.F == 1 && cfg->weight_kind != SQC_HIP_WEIGHT_XRAY)
    g.pair_scale = p.wtab[0] != 0.0 ? 1.0 / (p.wtab[0] * p.wtab[0]) : 0.0;

  // --- cell and modes ---
  const double two_pi = 2.0 * std::acos(-1.0);
  inverse3(cfg->cell, g.ainv);
  ModeGeom mg{};
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i) mg.b[i + 3 * j] = two_pi * g.ainv[j + 3 * i];
  std::memcpy(g.b, mg.b, sizeof(g.b));
  std::memcpy(g.origin, cfg->origin, sizeof(g.origin));
  mg.qmin = cfg->qmin;
  mg.qmax = cfg->qmax;
  mg.nq = cfg->nq;
  mg.dq = (cfg->qmax - cfg->qmin) / (double)cfg->nq;
  if (!(mg.dq > 0.0)) return fail(SQC_HIP_ERR_BAD_ARG, "q range must be positive");
  p.dq = mg.dq;
  for (int i = 0; i < 3; ++i) {
    const double *a = cfg->cell + 3 * i;
    const double len = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    if (!(len > 0.0)) return fail(SQC_HIP_ERR_BAD_ARG, "degenerate simulation cell");
    int64_t mi = 2 * (int64_t)std::ceil(cfg->qmax * len / two_pi) + 1;
    p.m[i] = std::max<int64_t>(mi, 3);
    g.M[i] = (int)((p.m[i] - 1) / 2);
  }
  if (p.m[0] * p.m[1] * p.m[2] > 400000000LL)
    return fail(SQC_HIP_ERR_MAXNALLOC, "reciprocal grid exceeds 4e8 points; reduce qmax");

  // denominators and counts over every kept mode, reference order
  p.den.assign(cfg->nq, 0.0);
  p.shell_count.assign(cfg->nq, 0);
  p.nmodes = 0;
  for (int64_t p3 = 1; p3 <= p.m[2]; ++p3) {
    const int h3 = (int)(p3 - (p.m[2] + 1) / 2);
    for (int64_t p2 = 1; p2 <= p.m[1]; ++p2) {
      const int h2 = (int)(p2 - (p.m[1] + 1) / 2);
      for (int64_t p1 = 1; p1 <= p.m[0]; ++p1) {
        const int h1 = (int)(p1 - (p.m[0] + 1) / 2);
        if (h1 == 0 && h2 == 0 && h3 == 0) continue;
        const double ql = mg.qlen(h1, h2, h3);
        if (!mg.kept(ql)) continue;
        ++p.nmodes;
        const int s = mg.shell(ql);
        double d;
        if (cfg->norm == SQC_HIP_NORM_NATOM) {
          d = (double)cfg->natoms;
        } else {
          double sw = 0.0, sw2 = 0.0;
          for (int t = 0; t < S; ++t) {
            if (p.counts[t] <= 0) continue;
            const double w = amplitude(p, t, ql);
            sw = sw + (double)p.counts[t] * w;
            sw2 = sw2 + (double)p.counts[t] * w * w;
          }
          d = cfg->norm == SQC_HIP_NORM_MEAN ? sw * sw / (double)cfg->natoms : sw2;
        }
        p.den[s] += d;
        p.shell_count[s] += 1;
      }
    }
  }
  if (p.nmodes == 0)
    return fail(SQC_HIP_ERR_NO_MODES, "no reciprocal lattice points fall inside the requested q range");
  set_shell_quanta(p, cfg->qmax);

  // canonical half-space lines: h3 > 0; h3 = 0, h2 > 0; h3 = h2 = 0, h1 > 0
  const int M1 = g.M[0], M2 = g.M[1], M3 = g.M[2];
  p.lo2.assign(M3 + 1, 0);
  p.hi2.assign(M3 + 1, -1);
  p.loff.assign(M3 + 1, 0);
  p.line_info.clear();
  p.line_off.clear();
  std::vector<uint16_t> stab;
  p.nhalf = 0;
  int nlines = 0;
  for (int h3 = 0; h3 <= M3; ++h3) {
    const int h2a = (h3 == 0) ? 0 : -M2;
    // first pass: which h2 have kept modes
    int lo = 1 << 30, hi = -(1 << 30);
    std::vector<int> l1(2 * M2 + 1, 1 << 30), h1v(2 * M2 + 1, -(1 << 30));
    for (int h2 = h2a; h2 <= M2; ++h2) {
      const int h1a = (h3 == 0 && h2 == 0) ? 1 : -M1;
      for (int h1 = h1a; h1 <= M1; ++h1) {
        if (mg.kept(mg.qlen(h1, h2, h3))) {
          l1[h2 + M2] = std::min(l1[h2 + M2], h1);
          h1v[h2 + M2] = std::max(h1v[h2 + M2], h1);
        }
      }
      if (h1v[h2 + M2] >= l1[h2 + M2]) {
        lo = std::min(lo, h2);
        hi = std::max(hi, h2);
      }
    }
    p.loff[h3] = nlines;
    if (hi < lo) {
      p.lo2[h3] = 0;
      p.hi2[h3] = -1;
      continue;
    }
    p.lo2[h3] = lo;
    p.hi2[h3] = hi;
    for (int h2 = lo; h2 <= hi; ++h2) {
      int a = l1[h2 + M2], b = h1v[h2 + M2];
      if (b < a) { a = 0; b = -1; }
      p.line_info.insert(p.line_info.end(), {h2, h3, a, b});
      p.line_off.push_back((int32_t)stab.size());
      for (int h1 = a; h1 <= b; ++h1) {
        const double ql = mg.qlen(h1, h2, h3);
        const bool kp = mg.kept(ql) && !(h1 == 0 && h2 == 0 && h3 == 0);
        stab.push_back(kp ? (uint16_t)mg.shell(ql) : sqc::kNoShell);
        if (kp) ++p.nhalf;
      }
      if (stab.size() > 2000000000ULL)
        return fail(SQC_HIP_ERR_MAXNALLOC, "too many candidate modes");
      ++nlines;
    }
  }
  p.line_off.push_back((int32_t)stab.size());
  p.ncand = (int64_t)stab.size();
  g.nlines = nlines;
  if (2 * p.nhalf != p.nmodes)
    return fail(SQC_HIP_ERR_BAD_ARG, "kept-mode set is not symmetric under h -> -h");

  // --- kernel and fine grid ---
  double eps = cfg->eps > 0 ? cfg->eps : (p.prec == 1 ? 1e-6 : 1e-9);
  double sigma = cfg->upsampfac > 0 ? cfg->upsampfac : (p.prec == 1 ? 1.25 : 2.0);
  if (!(sigma > 1.0)) return fail(SQC_HIP_ERR_BAD_ARG, "upsampfac must exceed 1");
  p.sigma = sigma;
  int ret = SQC_HIP_OK;
  // FP32 forward sums stop gaining below about 1e-7, but the adjoint's
  // kernel-derivative aliasing keeps falling with eps (down to the ns = 16
  // cap), so FP32 plans take eps down to 1e-10 as well
  const double emin = p.prec == 1 ? 1e-10 : 2.3e-16;
  if (eps < emin) {
    eps = emin;
    ret = SQC_HIP_WARN_EPS_TOO_SMALL;
  }
  int ns;
  if (sigma == 2.0)
    ns = (int)std::ceil(-std::log10(eps / 10.0));
  else
    ns = (int)std::ceil(-std::log(eps) / (3.14159265358979323846 * std::sqrt(1.0 - 1.0 / sigma)));
  ns = std::max(2, ns);
  if (ns > 16) {
    ns = 16;
    ret = SQC_HIP_WARN_EPS_TOO_SMALL;
  }
  double bon = 2.30;
  if (ns == 2) bon = 2.20;
  if (ns == 3) bon = 2.26;
  if (ns == 4) bon = 2.38;
  if (sigma != 2.0) bon = 0.97 * 3.14159265358979323846 * (1.0 - 1.0 / (2.0 * sigma));
  g.ns = ns;
  g.beta = bon * ns;
  g.half_ns = 0.5 * ns;
  g.inv_half_ns = 2.0 / ns;
  g.inv_per = 1.0 / (3.0 * ns);
  // dphi/du = -phi beta z / sqrt(1 - z^2) (2 / ns) has an integrable
  // singularity at the support edge |z| = 1, where phi = e^-beta.  An atom a
  // few ulps inside the edge (atoms on grid points or half-points of a
  // lattice-aligned cell) would get a spike of e^-beta beta / sqrt(1 - z^2)
  // there.  Bounding the square root below by 2 e^-beta / eps keeps that
  // term under eps beta / ns, the transform's own accuracy relative to the
  // derivative's scale; elsewhere dphi is unchanged.
  g.dphi_smin = std::min(0.1, std::max(1e-6, 2.0 * std::exp(-g.beta) / eps));
  for (int i = 0; i < 3; ++i) {
    const int64_t nf = next235even(std::max<int64_t>((int64_t)std::ceil(sigma * p.m[i]), 2 * ns));
    if (nf > 16384) return fail(SQC_HIP_ERR_MAXNALLOC, "fine grid exceeds 16384 points per axis");
    g.n[i] = (int)nf;
    if (!factor(g.n[i], p.ax[i])) return fail(SQC_HIP_ERR_BAD_ARG, "fine grid not 2,3,5-smooth");
  }

  // --- launch shapes ---
  const size_t vs = p.prec == 1 ? 8 : 16;
  const size_t lds_pref = ::lds_pref(p);
  const int n1 = g.n[0], n2 = g.n[1], n3 = g.n[2];
  for (int i = 0; i < 3; ++i)
    if (g.n[i] > sqc::kMaxLongLine)
      return fail(SQC_HIP_ERR_MAXNALLOC, "fine grid of " + std::to_string(g.n[i]) +
                                             " points per axis exceeds the LDS DFT limit of " +
                                             std::to_string(sqc::kMaxLongLine));
  // LDS z column stride of the spread tile: 24 words mod 64 spreads the
  // neighbouring columns of a spread update over the banks
  // (line buffers hold n + n/32 + 1 points: the DFT's padded stage layout)
  g.cs = sqc::spread_col_stride(n3);
  g.rs1 = n1 + n1 / 32 + 1;
  {
    int pick = -1;
    for (int pass = 0; pass < 2 && pick < 0; ++pass)
      for (int c = 0; c < kNumTiles; ++c)
        if (tile_fits(p, kTiles[c][0], kTiles[c][1], pass == 0 ? lds_pref : kLdsCap)) {
          pick = c;
          break;
        }
    if (pick < 0) return fail(SQC_HIP_ERR_MAXNALLOC, "no spread tile fits the LDS");
    int tx2 = 16;
    while (tx2 > 2 && ypass_lds(p, tx2) > lds_pref_y(p)) tx2 /= 2;  // the y pass is compiled for 2..16
    set_tiles(p, kTiles[pick][0], kTiles[pick][1], tx2);
  }
  // shell sums in HBM when even one x line per group leaves them no room
  g.acc_global = 0;
  if (xpass_lds(p, 1) > xpass_cap()) g.acc_global = 1;
  int lb0 = sqc::kMaxLb;
  while (lb0 > 1 && xpass_lds(p, lb0) > lds_pref) --lb0;
  if (xpass_lds(p, lb0) > xpass_cap())
    return fail(SQC_HIP_ERR_MAXNALLOC, "x line x species too large for the LDS DFT");
  int dev_count = 0;
  HIP_TRY(hipGetDeviceCount(&dev_count), "hipGetDeviceCount");
  if (cfg->device < 0 || cfg->device >= dev_count)
    return fail(SQC_HIP_ERR_NO_DEVICE, "HIP device " + std::to_string(cfg->device) +
                                           " requested but " + std::to_string(dev_count) +
                                           " device(s) available");
  p.device = cfg->device;
  HIP_TRY(hipSetDevice(p.device), "hipSetDevice");
  if (int rc = scratch_setup(p.device)) return rc;
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, p.device), "hipGetDeviceProperties");
  p.ncu = prop.multiProcessorCount;
  {
    int nx = 1;
    if (hipDeviceGetAttribute(&nx, hipDeviceAttributeNumberOfXccs, p.device) != hipSuccess ||
        nx < 1) {
      (void)hipGetLastError();
      nx = 1;
    }
    g.nxcd = nx;
  }
  set_xpass(p, lb0);
  HIP_TRY(hipStreamCreateWithFlags(&h->own_stream, hipStreamNonBlocking), "hipStreamCreate");
  p.stream = h->own_stream;
  HIP_TRY(hipEventCreateWithFlags(&p.ev_in, hipEventDisableTiming), "hipEventCreate");
  HIP_TRY(hipEventCreateWithFlags(&p.ev_out, hipEventDisableTiming), "hipEventCreate");

  // --- device buffers ---
  const int64_t N = cfg->natoms;
  int rc;
  p.batch_cap = cfg->max_batch > 0 ? cfg->max_batch : 8;
  p.use_graph = p.batch_cap > 1;
  // the atom preparation runs one 32-bit thread index per kernel value
  if ((double)N * 3 * ns >= 2147483647.0)
    return fail(SQC_HIP_ERR_MAXNALLOC, "natoms x 3 ns exceeds 2^31 kernel values");
  if ((rc = alloc(&p.d_pos, sizeof(double) * 3 * N * p.batch_cap, p.bytes))) return rc;
  if ((rc = alloc(&p.d_i0, sizeof(int4) * N, p.bytes))) return rc;
  if ((rc = alloc(&p.d_i0s, sizeof(int4) * N, p.bytes))) return rc;
  if ((rc = alloc((char **)&p.d_phi, (vs / 2) * 3 * ns * N, p.bytes))) return rc;
  if ((rc = alloc((char **)&p.d_dphi, (vs / 2) * 3 * ns * N, p.bytes))) return rc;
  if ((rc = alloc(&p.d_key, sizeof(int32_t) * N, p.bytes))) return rc;
  if ((rc = alloc(&p.d_sidx, sizeof(int32_t) * N, p.bytes))) return rc;
  if ((rc = upload(&p.d_species, p.species_of.data(), sizeof(int32_t) * N, p.bytes))) return rc;
  if ((rc = alloc_bins(p))) return rc;
  // the passes index W1 and W2 with 64-bit offsets past a field, a line
  // group or a plane; the offsets within one stay below 2^31
  if ((double)(M3 + 1) * n2 * n1 >= 2147483647.0 ||
      (double)g.F * std::max(g.nlines, 1) >= 2147483647.0)
    return fail(SQC_HIP_ERR_MAXNALLOC, "transform buffers exceed 2^31 elements per field");
  const size_t w1 = (size_t)g.F * (M3 + 1) * n2 * n1 * vs;
  const size_t w2 = (size_t)g.F * std::max(g.nlines, 1) * n1 * vs;
  if ((rc = alloc((char **)&p.d_W1, w1, p.bytes))) return rc;
  if ((rc = alloc((char **)&p.d_W2, w2, p.bytes))) return rc;
  if (p.use_graph) {
    // captured batches overlap consecutive frames: a second W1 and stream
    if ((rc = alloc((char **)&p.d_W1b, w1, p.bytes))) return rc;
    HIP_TRY(hipStreamCreateWithFlags(&p.stream2, hipStreamNonBlocking), "hipStreamCreate");
  }
  for (int i = 0; i < 3; ++i)
    if ((rc = upload_twiddles(p, i))) return rc;
  std::vector<double> phv[3];
  for (int i = 0; i < 3; ++i) {
    phv[i] = phihat(g.n[i], g.M[i], ns, g.beta);
    if ((rc = upload(&p.d_ph[i], phv[i].data(), phv[i].size() * sizeof(double), p.bytes)))
      return rc;
  }
  {
    // deconvolution factor of every x line, 1 / (phihat_2(|h2|) phihat_3(h3))
    std::vector<double> d23(std::max(g.nlines, 1));
    for (int l = 0; l < g.nlines; ++l) {
      const int h2 = p.line_info[4 * l], h3 = p.line_info[4 * l + 1];
      d23[l] = 1.0 / (phv[1][std::abs(h2)] * phv[2][h3]);
    }
    if ((rc = upload(&p.d_line_d23, d23.data(), d23.size() * sizeof(double), p.bytes))) return rc;
  }
  if ((rc = upload(&p.d_lo2, p.lo2.data(), p.lo2.size() * 4, p.bytes))) return rc;
  if ((rc = upload(&p.d_hi2, p.hi2.data(), p.hi2.size() * 4, p.bytes))) return rc;
  if ((rc = upload(&p.d_loff, p.loff.data(), p.loff.size() * 4, p.bytes))) return rc;
  if ((rc = upload(&p.d_line_info, p.line_info.data(), p.line_info.size() * 4, p.bytes))) return rc;
  if ((rc = upload(&p.d_line_off, p.line_off.data(), p.line_off.size() * 4, p.bytes))) return rc;
  if ((rc = upload(&p.d_stab, stab.data(), stab.size() * 2, p.bytes))) return rc;
  if ((rc = upload(&p.d_den, p.den.data(), p.den.size() * 8, p.bytes))) return rc;
  if ((rc = upload(&p.d_wtab, p.wtab.data(), p.wtab.size() * 8, p.bytes))) return rc;
  {
    std::vector<double> cnt(p.shell_count.begin(), p.shell_count.end());
    if ((rc = upload(&p.d_cnt, cnt.data(), cnt.size() * 8, p.bytes))) return rc;
  }
  const size_t ntot = (size_t)g.ncomp * g.nq;
  if ((rc = alloc_wgpart(p))) return rc;
  if ((rc = alloc(&p.d_acc, sizeof(double) * ntot * 2, p.bytes))) return rc;
  HIP_TRY(hipMemset(p.d_acc, 0, sizeof(double) * ntot * 2), "hipMemset");
  if (cfg->want_grid) {
    if ((rc = alloc(&p.d_gnum, sizeof(double) * p.ncand, p.bytes))) return rc;
    HIP_TRY(hipMemset(p.d_gnum, 0, sizeof(double) * p.ncand), "hipMemset");
  }
  if (cfg->mode_partials && g.partials && g.ncomp > 1) {
    const size_t n = sizeof(double) * (size_t)(g.ncomp - 1) * p.ncand;
    if ((rc = alloc(&p.d_gpart, n, p.bytes))) return rc;
    HIP_TRY(hipMemset(p.d_gpart, 0, n), "hipMemset");
  }
  if (int e = sqc::kernel_lds_limits(p))
    return hip_fail((hipError_t)e, "setting the LDS size of the kernels");
  sqc::persistent_grids(p);
  if ((rc = alloc_wgpart(p))) return rc;
  HIP_TRY(hipDeviceSynchronize(), "plan set-up");
  return ret;
}

int check_plan(const sqc_hip_plan *plan) {
  if (!plan) return fail(SQC_HIP_ERR_BAD_ARG, "null plan");
  return 0;
}

// Orders the plan stream after the caller's stream on entry and the caller's
// stream after the plan stream on exit.  The caller's stream may be the
// legacy default stream (handle 0), which is what torch uses by default.
struct OrderScope {
  sqc::Plan &p;
  hipStream_t user;
  OrderScope(sqc::Plan &pl, void *s) : p(pl), user((hipStream_t)s) {
    (void)hipEventRecord(p.ev_in, user);
    (void)hipStreamWaitEvent(p.stream, p.ev_in, 0);
  }
  ~OrderScope() {
    (void)hipEventRecord(p.ev_out, p.stream);
    (void)hipStreamWaitEvent(user, p.ev_out, 0);
  }
};

// Forward time of one frame (ms, best of two after a warm-up), nothing
// accumulated.
// Per-stage forward times of one frame (ms, best of three after a warm-up),
// nothing accumulated: stage[0] atom sort, [1] spread + z, [2] y, [3] x + bin.
int time_stages(sqc::Plan &p, const double *pos_dev, double stage[4]) {
  for (int i = 0; i < 4; ++i) stage[i] = 1e30;
  const int saved = p.profiling;
  (void)sqc::collect_stage_times(p);
  double keep_ms[sqc::kNumStages];
  int64_t keep_n[sqc::kNumStages];
  for (int i = 0; i < sqc::kNumStages; ++i) {
    keep_ms[i] = p.stage_ms[i];
    keep_n[i] = p.stage_n[i];
  }
  p.profiling = 1;
  int rc = 0;
  for (int it = 0; it < 6 && !rc; ++it) {
    for (int i = 0; i < sqc::kNumStages; ++i) {
      p.stage_ms[i] = 0.0;
      p.stage_n[i] = 0;
    }
    rc = sqc::launch_forward(p, pos_dev, 2, nullptr, nullptr);
    if (!rc) rc = sqc::collect_stage_times(p);
    if (!rc && it > 0)
      for (int i = 0; i < 4; ++i) stage[i] = std::min(stage[i], p.stage_ms[i]);
  }
  // the caller's own stage statistics are left as they were
  for (int i = 0; i < sqc::kNumStages; ++i) {
    p.stage_ms[i] = keep_ms[i];
    p.stage_n[i] = keep_n[i];
  }
  p.profiling = saved;
  return rc ? hip_fail((hipError_t)rc, "tuning launch") : 0;
}

// Re-derive everything that depends on the tiles and resize what they size.
int apply_tiles(sqc::Plan &p, int txr, int ty, int tx2, int lb) {
  set_tiles(p, txr, ty, tx2);
  set_xpass(p, lb);
  if (int rc = alloc_bins(p)) return rc;
  if (int rc = alloc_wgpart(p)) return rc;
  if (p.d_slots) {  // sized by the tiles an atom can touch; the adjoint re-allocates it
    (void)hipFree(p.d_slots);
    p.bytes -= p.slots_bytes;
    p.d_slots = nullptr;
  }
  if (int e = sqc::kernel_lds_limits(p)) return hip_fail((hipError_t)e, "setting LDS sizes");
  sqc::persistent_grids(p);
  if (int rc = alloc_wgpart(p)) return rc;
  drop_graph(p);
  return 0;
}

// Autotuner (config.tune): on the first frame, time every spread tile that
// fits the LDS (atom sort + spread stages), then every y-pass column chunk (y
// stage) and every x-pass line-group size (x stage) with the best choices so
// far, and keep the fastest.  Nothing is accumulated while tuning.
int tune(sqc::Plan &p, const double *pos_dev) {
  p.tuned = 1;
  if (!p.cfg.tune) return 0;
  // each choice is judged on the stages it changes, and a candidate replaces
  // the current choice only when it is more than 1% faster
  constexpr double kGain = 0.99;
  int btxr = p.geo.txr, bty = p.geo.ty, btx2 = p.tx2, blb = p.lb;
  double st[4];
  if (int rc = time_stages(p, pos_dev, st)) return rc;
  // spread tiles, judged on the sort, the spread and the y pass (which reads
  // the tile's write pattern): the best and the runner-up of a first sweep
  // are timed again, in turn, before one is kept (neighbouring tiles often
  // differ by a few percent, within the spread of a single timing)
  auto tile_cost = [&]() { return st[0] + st[1] + st[2]; };
  double best = tile_cost();
  int rtxr = 0, rty = 0;
  double second = 1e30;
  for (int c = 0; c < kNumTiles; ++c) {
    const int txr = kTiles[c][0], ty = kTiles[c][1];
    if ((txr == btxr && ty == bty) || !tile_fits(p, txr, ty, kLdsCap)) continue;
    if (int rc = apply_tiles(p, txr, ty, btx2, blb)) return rc;
    if (int rc = time_stages(p, pos_dev, st)) return rc;
    const double v = tile_cost();
    if (v < kGain * best) {
      second = best;
      rtxr = btxr;
      rty = bty;
      best = v;
      btxr = txr;
      bty = ty;
    } else if (v < second) {
      second = v;
      rtxr = txr;
      rty = ty;
    }
  }
  if (rtxr > 0) {
    double tb = 1e30, tr = 1e30;
    for (int rep = 0; rep < 2; ++rep) {
      if (int rc = apply_tiles(p, rtxr, rty, btx2, blb)) return rc;
      if (int rc = time_stages(p, pos_dev, st)) return rc;
      tr = std::min(tr, tile_cost());
      if (int rc = apply_tiles(p, btxr, bty, btx2, blb)) return rc;
      if (int rc = time_stages(p, pos_dev, st)) return rc;
      tb = std::min(tb, tile_cost());
    }
    if (tr < kGain * tb) {
      btxr = rtxr;
      bty = rty;
    }
  }
  if (int rc = apply_tiles(p, btxr, bty, btx2, blb)) return rc;
  if (int rc = time_stages(p, pos_dev, st)) return rc;
  best = st[2];
  for (int tx2 = 16; tx2 >= 2; tx2 /= 2) {
    if (ypass_lds(p, tx2) > kLdsCap || tx2 == btx2) continue;
    if (int rc = apply_tiles(p, btxr, bty, tx2, blb)) return rc;
    if (int rc = time_stages(p, pos_dev, st)) return rc;
    if (st[2] < kGain * best) {
      best = st[2];
      btx2 = tx2;
    }
  }
  if (int rc = apply_tiles(p, btxr, bty, btx2, blb)) return rc;
  if (int rc = time_stages(p, pos_dev, st)) return rc;
  best = st[3];
  for (int lb = sqc::kMaxLb; lb >= 1; lb /= 2) {
    if (xpass_lds(p, lb) > xpass_cap() || lb == blb) continue;
    if (int rc = apply_tiles(p, btxr, bty, btx2, lb)) return rc;
    if (int rc = time_stages(p, pos_dev, st)) return rc;
    if (st[3] < kGain * best) {
      best = st[3];
      blb = lb;
    }
  }
  return apply_tiles(p, btxr, bty, btx2, blb);
}

// Capture batch_cap forward frames reading p.d_pos into one graph, as a
// two-stage pipeline: frame f's atom sort and spread (stream) write W1 or
// W1b by the parity of f; its y pass, x pass and finalize (stream2) follow
// them.  So the sort and spread of frame f + 1 run beside the y and x passes
// of frame f; the spread of frame f + 2 waits for the y pass of frame f,
// which reads the same W1.  The finalizes keep their order, so the sums are
// those of frames launched one by one.
int capture_batch(sqc::Plan &p) {
  const int64_t stride = 3 * p.geo.natoms;
  const int64_t B = p.batch_cap;
  // events: [0] the fork, then per frame S (spread done), Y (y pass done)
  // and B (x pass done); every index is checked against the vector, so a
  // slip in this bookkeeping is an error return, not a stray handle
  const size_t nev = 3 * (size_t)B + 1;
  while (p.gev.size() < nev) {
    hipEvent_t e = nullptr;
    HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
    p.gev.push_back(e);
  }
  auto ev = [&](int kind, int64_t f) -> hipEvent_t {
    const size_t i = kind < 0 ? 0 : 1 + (size_t)kind * (size_t)B + (size_t)f;
    return (f >= 0 && f < B && i < p.gev.size()) ? p.gev[i] : nullptr;
  };
  enum { kS = 0, kY = 1, kB = 2 };
  if (!p.stream || !p.stream2 || !p.d_W1 || !p.d_W1b)
    return fail(SQC_HIP_ERR_BAD_ARG, "frame batch capture without its second stream and buffer");
  hipGraph_t graph = nullptr;
  HIP_TRY(hipStreamBeginCapture(p.stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
  int rc = 0;
  // run one HIP call unless an earlier one failed
  auto step = [&](auto &&call) {
    if (rc) return;
    const hipError_t e = call();
    if (e != hipSuccess) rc = (int)e;
  };
  // an event that does not exist ends the capture with an error
  auto record = [&](int kind, int64_t f, hipStream_t s) {
    step([&] {
      hipEvent_t e = ev(kind, f);
      return e ? hipEventRecord(e, s) : hipErrorInvalidResourceHandle;
    });
  };
  auto wait = [&](hipStream_t s, int kind, int64_t f) {
    step([&] {
      hipEvent_t e = ev(kind, f);
      return e ? hipStreamWaitEvent(s, e, 0) : hipErrorInvalidResourceHandle;
    });
  };
  step([&] { return hipEventRecord(ev(-1, 0), p.stream); });
  step([&] { return hipStreamWaitEvent(p.stream2, ev(-1, 0), 0); });
  for (int64_t f = 0; f < B && !rc; ++f) {
    void *W1 = (f & 1) ? p.d_W1b : p.d_W1;
    if (f >= 2) wait(p.stream, kY, f - 2);  // W1 reuse: the y pass of frame f - 2 has read it
    if (!rc) rc = sqc::launch_forward_front(p, p.d_pos + f * stride, W1, p.stream);
    record(kS, f, p.stream);
    wait(p.stream2, kS, f);
    if (!rc) rc = sqc::launch_forward_y(p, W1, p.d_W2, p.stream2);
    record(kY, f, p.stream2);
    if (!rc) rc = sqc::launch_forward_x(p, p.d_W2, p.stream2, 0, nullptr, nullptr);
    record(kB, f, p.stream2);
  }
  // join the side stream back into the capturing stream
  wait(p.stream, kB, B - 1);
  const hipError_t e = hipStreamEndCapture(p.stream, &graph);
  if (rc || e != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return hip_fail(rc ? (hipError_t)rc : e, "capturing the frame batch");
  }
  const hipError_t ei = hipGraphInstantiate(&p.gexec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ei != hipSuccess) {
    p.gexec = nullptr;
    return hip_fail(ei, "hipGraphInstantiate");
  }
  return 0;
}

// Accumulate frames whose positions are in device memory.  Full batches of
// batch_cap frames replay the captured graph (positions copied into d_pos
// first unless they are already there); a short tail, and every frame while
// the stage timers are on, is launched kernel by kernel.
int run_frames(sqc::Plan &p, const double *pos_dev, int64_t nframes) {
  const int64_t stride = 3 * p.geo.natoms;
  if (nframes > 0 && !p.tuned)
    if (int rc = tune(p, pos_dev)) return rc;
  for (int64_t f0 = 0; f0 < nframes; f0 += p.batch_cap) {
    const int64_t nb = std::min<int64_t>(p.batch_cap, nframes - f0);
    const double *src = pos_dev + f0 * stride;
    if (p.use_graph && !p.profiling && nb == p.batch_cap) {
      if (src != p.d_pos)
        HIP_TRY(hipMemcpyAsync(p.d_pos, src, sizeof(double) * stride * nb,
                               hipMemcpyDeviceToDevice, p.stream),
                "stage positions");
      if (!p.gexec)
        if (int rc = capture_batch(p)) return rc;
      HIP_TRY(hipGraphLaunch(p.gexec, p.stream), "hipGraphLaunch");
    } else {
      for (int64_t f = 0; f < nb; ++f) {
        const int rc = sqc::launch_forward(p, src + f * stride, 0, nullptr, nullptr);
        if (rc) return hip_fail((hipError_t)rc, "forward launch");
      }
    }
    p.nframes += nb;
  }
  return 0;
}

}  // namespace

namespace sqc {
void cap_wgpart_rows(Plan &p) { cap_rows(p); }
// The thread's error message for sqc_hip_last_error, set by the other
// translation units of the library.
int set_error(int code, const std::string &msg) { return fail(code, msg); }
}  // namespace sqc

extern "C" {

int sqc_hip_device_count(int *count) {
  if (!count) return fail(SQC_HIP_ERR_BAD_ARG, "null count");
  *count = 0;
  hipError_t e = hipGetDeviceCount(count);
  if (e != hipSuccess) {
    *count = 0;
    return hip_fail(e, "hipGetDeviceCount");
  }
  return SQC_HIP_OK;
}

size_t sqc_hip_config_size(void) { return sizeof(sqc_hip_config); }

int sqc_hip_plan_create(sqc_hip_plan **plan, const sqc_hip_config *cfg) {
  if (!plan || !cfg) return fail(SQC_HIP_ERR_BAD_ARG, "null argument");
  *plan = nullptr;
  if (cfg->abi_version != SQC_HIP_ABI_VERSION)
    return fail(SQC_HIP_ERR_ABI, "sqc_hip_config from ABI version " +
                                     std::to_string(cfg->abi_version) + ", library has " +
                                     std::to_string(SQC_HIP_ABI_VERSION));
  if (cfg->precision != SQC_HIP_F64 && cfg->precision != SQC_HIP_F32)
    return fail(SQC_HIP_ERR_BAD_ARG, "precision must be SQC_HIP_F64 or SQC_HIP_F32");
  if (cfg->natoms <= 0 || cfg->natoms > 2000000000LL || !cfg->species_of)
    return fail(SQC_HIP_ERR_BAD_ARG, "natoms must be in 1..2e9 with species_of given");
  if (cfg->nspecies < 1 || cfg->nspecies > sqc::kMaxSpecies)
    return fail(SQC_HIP_ERR_BAD_ARG, "nspecies must be in 1..8");
  if (cfg->nq < 1 || cfg->nq > 65535)
    return fail(SQC_HIP_ERR_BAD_ARG, "nq must be in 1..65535");
  if (cfg->weight_kind < 0 || cfg->weight_kind > 2)
    return fail(SQC_HIP_ERR_BAD_ARG, "unknown weight kind");
  if (cfg->norm < 1 || cfg->norm > 3) return fail(SQC_HIP_ERR_BAD_ARG, "unknown norm");
  if (cfg->weight_kind == SQC_HIP_WEIGHT_NEUTRON && !cfg->amp_const)
    return fail(SQC_HIP_ERR_BAD_ARG, "neutron weights need amp_const");
  if (cfg->weight_kind == SQC_HIP_WEIGHT_XRAY && !cfg->xray_coef)
    return fail(SQC_HIP_ERR_BAD_ARG, "x-ray weights need xray_coef");
  if (!(cfg->qmax > cfg->qmin) || cfg->qmin < 0.0)
    return fail(SQC_HIP_ERR_BAD_ARG, "q range must be positive");
  auto *h = new sqc_hip_plan();
  const int rc = build_plan(h, cfg);
  if (rc > SQC_HIP_WARN_EPS_TOO_SMALL) {
    free_plan(h);
    return rc;
  }
  *plan = h;
  return rc;
}

int sqc_hip_plan_info(const sqc_hip_plan *plan, sqc_hip_info *info) {
  if (int rc = check_plan(plan)) return rc;
  if (!info) return fail(SQC_HIP_ERR_BAD_ARG, "null info");
  const sqc::Plan &p = plan->p;
  std::memset(info, 0, sizeof(*info));
  for (int i = 0; i < 3; ++i) {
    info->modes[i] = p.m[i];
    info->fine[i] = p.geo.n[i];
  }
  info->nmodes = p.nmodes;
  info->nhalf = p.nhalf;
  info->nspread = p.geo.ns;
  info->nfields = p.geo.F;
  info->beta = p.geo.beta;
  info->upsampfac = p.sigma;
  info->device_bytes = p.bytes;
  info->nlines = p.geo.nlines;
  info->tile[0] = p.geo.txr;
  info->tile[1] = p.geo.ty;
  info->tile[2] = p.tx2;
  info->ncomp = p.geo.ncomp;
  info->rho_len = p.ncand;
  info->shell_sums_in_hbm = p.geo.acc_global;
  info->x_lines = p.lb;
  return SQC_HIP_OK;
}

int sqc_hip_set_tiles(sqc_hip_plan *plan, int spread_x, int spread_y, int y_chunk,
                      int x_lines) {
  if (int rc = check_plan(plan)) return rc;
  sqc::Plan &p = plan->p;
  bool listed = false;
  for (int c = 0; c < kNumTiles; ++c) listed |= kTiles[c][0] == spread_x && kTiles[c][1] == spread_y;
  if (!listed || !tile_fits(p, spread_x, spread_y, kLdsCap))
    return fail(SQC_HIP_ERR_BAD_ARG, "spread tile " + std::to_string(spread_x) + " x " +
                                         std::to_string(spread_y) + " is not usable for this plan");
  if ((y_chunk != 2 && y_chunk != 4 && y_chunk != 8 && y_chunk != 16) ||
      ypass_lds(p, y_chunk) > kLdsCap)
    return fail(SQC_HIP_ERR_BAD_ARG, "y-pass column chunk " + std::to_string(y_chunk) +
                                         " is not usable for this plan");
  if (x_lines < 1 || x_lines > sqc::kMaxLb || xpass_lds(p, x_lines) > xpass_cap())
    return fail(SQC_HIP_ERR_BAD_ARG, "x-pass line group " + std::to_string(x_lines) +
                                         " is not usable for this plan");
  HIP_TRY(hipSetDevice(p.device), "hipSetDevice");
  // the buffers the tiles size are reallocated: nothing queued may use them
  HIP_TRY(hipStreamSynchronize(p.stream), "hipStreamSynchronize");
  if (p.stream2) HIP_TRY(hipStreamSynchronize(p.stream2), "hipStreamSynchronize");
  if (int rc = apply_tiles(p, spread_x, spread_y, y_chunk, x_lines)) return rc;
  p.tuned = 1;  // pinned: the tuner does not run on the next frame
  return SQC_HIP_OK;
}

int sqc_hip_shells(const sqc_hip_plan *plan, double *den, int64_t *shell_count) {
  if (int rc = check_plan(plan)) return rc;
  const sqc::Plan &p = plan->p;
  for (int s = 0; s < p.geo.nq; ++s) {
    if (den) den[s] = p.den[s];
    if (shell_count) shell_count[s] = p.shell_count[s];
  }
  return SQC_HIP_OK;
}

int sqc_hip_modes(const sqc_hip_plan *plan, int32_t *hkl, double *qvec, int64_t *gidx) {
  if (int rc = check_plan(plan)) return rc;
  const sqc::Plan &p = plan->p;
  ModeGeom mg{};
  std::memcpy(mg.b, p.geo.b, sizeof(mg.b));
  mg.qmin = p.cfg.qmin;
  mg.qmax = p.cfg.qmax;
  mg.nq = p.cfg.nq;
  mg.dq = p.dq;
  int64_t k = 0;
  for (int64_t p3 = 1; p3 <= p.m[2]; ++p3) {
    const int h3 = (int)(p3 - (p.m[2] + 1) / 2);
    for (int64_t p2 = 1; p2 <= p.m[1]; ++p2) {
      const int h2 = (int)(p2 - (p.m[1] + 1) / 2);
      for (int64_t p1 = 1; p1 <= p.m[0]; ++p1) {
        const int h1 = (int)(p1 - (p.m[0] + 1) / 2);
        if (h1 == 0 && h2 == 0 && h3 == 0) continue;
        double q[3];
        const double ql = mg.qlen(h1, h2, h3, q);
        if (!mg.kept(ql)) continue;
        if (hkl) {
          hkl[3 * k] = h1;
          hkl[3 * k + 1] = h2;
          hkl[3 * k + 2] = h3;
        }
        if (qvec) {
          qvec[3 * k] = q[0];
          qvec[3 * k + 1] = q[1];
          qvec[3 * k + 2] = q[2];
        }
        if (gidx) gidx[k] = p1 + p.m[0] * (p2 - 1 + p.m[1] * (p3 - 1));
        ++k;
      }
    }
  }
  return SQC_HIP_OK;
}

int sqc_hip_accumulate_device(sqc_hip_plan *plan, const double *pos_device, int64_t nframes) {
  if (int rc = check_plan(plan)) return rc;
  if (!pos_device || nframes < 0) return fail(SQC_HIP_ERR_BAD_ARG, "bad positions");
  sqc::Plan &p = plan->p;
  HIP_TRY(hipSetDevice(p.device), "hipSetDevice");
  if (int rc = run_frames(p, pos_device, nframes)) return rc;
  return SQC_HIP_OK;
}

int sqc_hip_accumulate(sqc_hip_plan *plan, const double *pos, int64_t nframes) {
  if (int rc = check_plan(plan)) return rc;
  if (!pos || nframes < 0) return fail(SQC_HIP_ERR_BAD_ARG, "bad positions");
  sqc::Plan &p = plan->p;
  HIP_TRY(hipSetDevice(p.device), "hipSetDevice");
  const int64_t stride = 3 * p.geo.natoms;
  for (int64_t f0 = 0; f0 < nframes; f0 += p.batch_cap) {
    const int64_t nb = std::min<int64_t>(p.batch_cap, nframes - f0);
    HIP_TRY(hipStreamSynchronize(p.stream), "hipStreamSynchronize");
    HIP_TRY(hipMemcpyAsync(p.d_pos, pos + f0 * stride, sizeof(double) * stride * nb,
                           hipMemcpyHostToDevice, p.stream),
            "upload positions");
    if (int rc = sqc_hip_accumulate_device(plan, p.d_pos, nb)) return rc;
  }
  HIP_TRY(hipStreamSynchronize(p.stream), "hipStreamSynchronize");
  return SQC_HIP_OK;
}

// Candidate slot of every kept mode, in the reference's enumeration order
// (src/sqc_structure_factor.f90:340-399): a mode outside the canonical half
// space reads the slot of -h, whose |rho|^2 and Re(rho_a rho_b*) are equal.
static std::vector<int64_t> mode_candidates(const sqc::Plan &p) {
  const sqc::Geo &g = p.geo;
  ModeGeom mg{};
  std::memcpy(mg.b, g.b, sizeof(mg.b));
  mg.qmin = p.cfg.qmin;
  mg.qmax = p.cfg.qmax;
  mg.nq = p.cfg.nq;
  mg.dq = p.dq;
  std::vector<int64_t> out;
  out.reserve((size_t)p.nmodes);
  for (int64_t p3 = 1; p3 <= p.m[2]; ++p3) {
    const int h3 = (int)(p3 - (p.m[2] + 1) / 2);
    for (int64_t p2 = 1; p2 <= p.m[1]; ++p2) {
      const int h2 = (int)(p2 - (p.m[1] + 1) / 2);
      for (int64_t p1 = 1; p1 <= p.m[0]; ++p1) {
        const int h1 = (int)(p1 - (p.m[0] + 1) / 2);
        if (h1 == 0 && h2 == 0 && h3 == 0) continue;
        if (!mg.kept(mg.qlen(h1, h2, h3))) continue;
        int c1 = h1, c2 = h2, c3 = h3;
        const bool canon = c3 > 0 || (c3 == 0 && (c2 > 0 || (c2 == 0 && c1 > 0)));
        if (!canon) {
          c1 = -c1;
          c2 = -c2;
          c3 = -c3;
        }
        const int line = p.loff[c3] + (c2 - p.lo2[c3]);
        out.push_back(p.line_off[line] + (c1 - p.line_info[4 * line + 2]));
      }
    }
  }
  return out;
}

int sqc_hip_fetch_mode_partials(sqc_hip_plan *plan, double *gpart) {
  if (int rc = check_plan(plan)) return rc;
  sqc::Plan &p = plan->p;
  if (!gpart) return fail(SQC_HIP_ERR_BAD_ARG, "null gpart");
  if (!p.d_gpart)
    return fail(SQC_HIP_ERR_BAD_ARG, "plan was created without mode_partials (and partials)");
  HIP_TRY(hipSetDevice(p.device), "hipSetDevice");
  HIP_TRY(hipStreamSynchronize(p.stream), "hipStreamSynchronize");
  const int npairs = p.geo.ncomp - 1;
  std::vector<double> gc((size_t)npairs * p.ncand);
  HIP_TRY(hipMemcpy(gc.data(), p.d_gpart, sizeof(double) * gc.size(), hipMemcpyDeviceToHost),
          "download mode partials");
  const std::vector<int64_t> cand = mode_candidates(p);
  const size_t K = cand.size();
  for (int q = 0; q < npairs; ++q)
    for (size_t k = 0; k < K; ++k) gpart[(size_t)q * K + k] = gc[(size_t)q * p.ncand + cand[k]];
  return SQC_HIP_OK;
}

int sqc_hip_fetch(sqc_hip_plan *plan, double *num, double *partial, double *gnum,
                  int64_t *nframes) {
  if (int rc = check_plan(plan)) return rc;
  sqc::Plan &p = plan->p;
  const sqc::Geo &g = p.geo;
  HIP_TRY(hipSetDevice(p.device), "hipSetDevice");
  HIP_TRY(hipStreamSynchronize(p.stream), "hipStreamSynchronize");
  const size_t ntot = (size_t)g.ncomp * g.nq;
  std::vector<double> acc(2 * ntot);
  HIP_TRY(hipMemcpy(acc.data(), p.d_acc, sizeof(double) * 2 * ntot, hipMemcpyDeviceToHost),
          "download accumulators");
  if (num)
    for (int s = 0; s < g.nq; ++s) num[s] = acc[s] + acc[ntot + s];
  if (partial) {
    const int S = g.S;
    std::fill(partial, partial + (size_t)S * S * g.nq, 0.0);
    if (g.ncomp > 1) {
      int pidx = 1;
      for (int a = 0; a < S; ++a)
        for (int b = a; b < S; ++b, ++pidx)
          for (int s = 0; s < g.nq; ++s)
            partial[((size_t)a * S + b) * g.nq + s] =
                acc[(size_t)pidx * g.nq + s] + acc[ntot + (size_t)pidx * g.nq + s];
    }
  }
  if (gnum) {
    if (!p.d_gnum) return fail(SQC_HIP_ERR_BAD_ARG, "plan was created without want_grid");
    std::vector<double> gc(p.ncand);
    HIP_TRY(hipMemcpy(gc.data(), p.d_gnum, sizeof(double) * p.ncand, hipMemcpyDeviceToHost),
            "download grid");
    const std::vector<int64_t> cand = mode_candidates(p);
    for (size_t k = 0; k < cand.size(); ++k) gnum[k] = gc[cand[k]];
  }
  if (nframes) *nframes = p.nframes;
  return SQC_HIP_OK;
}

int sqc_hip_reset(sqc_hip_plan *plan) {
  if (int rc = check_plan(plan)) return rc;
  sqc::Plan &p = plan->p;
  HIP_TRY(hipSetDevice(p.device), "hipSetDevice");
  HIP_TRY(hipStreamSynchronize(p.stream), "hipStreamSynchronize");
  HIP_TRY(hipMemset(p.d_acc, 0, sizeof(double) * 2 * p.geo.ncomp * p.geo.nq), "hipMemset");
  if (p.d_gnum) HIP_TRY(hipMemset(p.d_gnum, 0, sizeof(double) * p.ncand), "hipMemset");
  if (p.d_gpart)
    HIP_TRY(hipMemset(p.d_gpart, 0, sizeof(double) * (p.geo.ncomp - 1) * p.ncand), "hipMemset");
  p.nframes = 0;
  return SQC_HIP_OK;
}

int sqc_hip_forward_device(sqc_hip_plan *plan, const double *pos_device, int64_t nframes,
                           double *out_device, void *rho_device, void *stream) {
  if (int rc = check_plan(plan)) return rc;
  if (!pos_device || !out_device || nframes < 0)
    return fail(SQC_HIP_ERR_BAD_ARG, "bad forward arguments");
  sqc::Plan &p = plan->p;
  HIP_TRY(hipSetDevice(p.device), "hipSetDevice");
  OrderScope scope(p, stream);
  if (nframes > 0 && !p.tuned)
    if (int rc = tune(p, pos_device)) return rc;
  const int64_t stride = 3 * p.geo.natoms;
  const size_t vs = p.prec == 1 ? 8 : 16;
  const size_t rstride = (size_t)p.geo.F * p.ncand * vs;
  for (int64_t f = 0; f < nframes; ++f) {
    void *rho = rho_device ? (void *)((char *)rho_device + f * rstride) : nullptr;
    const int rc = sqc::launch_forward(p, pos_device + f * stride, 1,
                                       out_device + f * p.geo.ncomp * p.geo.nq, rho);
    if (rc) return hip_fail((hipError_t)rc, "forward launch");
  }
  return SQC_HIP_OK;
}

int sqc_hip_adjoint_device(sqc_hip_plan *plan, const double *pos_device, const void *rho_device,
                           const double *gS_device, const double *gP_device,
                           double *grad_device, void *stream) {
  if (int rc = check_plan(plan)) return rc;
  if (!pos_device || !rho_device || !gS_device || !grad_device)
    return fail(SQC_HIP_ERR_BAD_ARG, "bad adjoint arguments");
  sqc::Plan &p = plan->p;
  if (gP_device && p.geo.ncomp == 1)
    return fail(SQC_HIP_ERR_BAD_ARG, "pair gradients need a plan created with partials");
  HIP_TRY(hipSetDevice(p.device), "hipSetDevice");
  if (!p.d_slots) {
    p.slots_bytes = (int64_t)sizeof(double) * 3 * p.geo.natoms * p.geo.sx * p.geo.sy;
    int rc = alloc(&p.d_slots, (size_t)p.slots_bytes, p.bytes);
    if (rc) return rc;
  }
  OrderScope scope(p, stream);
  const int rc = sqc::launch_adjoint(p, pos_device, rho_device, gS_device, gP_device, grad_device,
                                     p.d_slots);
  if (rc) return hip_fail((hipError_t)rc, "adjoint launch");
  return SQC_HIP_OK;
}

int sqc_hip_adjoint_device_noalloc(sqc_hip_plan *plan, const double *pos_device,
                                   const void *rho_device, const double *gS_device,
                                   const double *gP_device, double *grad_device, void *stream) {
  if (int rc = check_plan(plan)) return rc;
  if (!pos_device || !rho_device || !gS_device || !grad_device)
    return fail(SQC_HIP_ERR_BAD_ARG, "bad adjoint arguments");
  sqc::Plan &p = plan->p;
  if (gP_device && p.geo.ncomp == 1)
    return fail(SQC_HIP_ERR_BAD_ARG, "pair gradients need a plan created with partials");
  const size_t need = sizeof(double) * 3 * (size_t)p.geo.natoms * p.geo.sx * p.geo.sy;
  if (need > SQC_HIP_NOALLOC_SCRATCH_BYTES)
    return fail(SQC_HIP_ERR_MAXNALLOC,
                "the adjoint's gradient slots need " + std::to_string(need) +
                    " bytes, more than the " + std::to_string(SQC_HIP_NOALLOC_SCRATCH_BYTES) +
                    "-byte no-allocation scratch (" + std::to_string(p.geo.natoms) +
                    " atoms x " + std::to_string(p.geo.sx * p.geo.sy) +
                    " tiles x 24 B); use sqc_hip_adjoint_device");
  HIP_TRY(hipSetDevice(p.device), "hipSetDevice");
  double *slots = scratch_of(p.device);
  if (!slots) return fail(SQC_HIP_ERR_HIP, "no scratch address for this device");
  OrderScope scope(p, stream);
  const int rc =
      sqc::launch_adjoint(p, pos_device, rho_device, gS_device, gP_device, grad_device, slots);
  if (rc) return hip_fail((hipError_t)rc, "adjoint launch");
  return SQC_HIP_OK;
}

int sqc_hip_gradient(sqc_hip_plan *plan, const double *pos, const double *target,
                     const double *wq, double *loss, double *grad) {
  if (int rc = check_plan(plan)) return rc;
  if (!pos || !target || !grad) return fail(SQC_HIP_ERR_BAD_ARG, "bad gradient arguments");
  sqc::Plan &p = plan->p;
  const sqc::Geo &g = p.geo;
  HIP_TRY(hipSetDevice(p.device), "hipSetDevice");
  const size_t vs = p.prec == 1 ? 8 : 16;
  int rc;
  if (!plan->g_pos) {
    if ((rc = alloc(&plan->g_pos, sizeof(double) * 3 * g.natoms, p.bytes))) return rc;
    if ((rc = alloc(&plan->g_out, sizeof(double) * g.ncomp * g.nq, p.bytes))) return rc;
    if ((rc = alloc(&plan->g_gs, sizeof(double) * g.nq, p.bytes))) return rc;
    if ((rc = alloc(&plan->g_grad, sizeof(double) * 3 * g.natoms, p.bytes))) return rc;
    if ((rc = alloc((char **)&plan->g_rho, g.F * p.ncand * vs, p.bytes))) return rc;
  }
  HIP_TRY(hipMemcpyAsync(plan->g_pos, pos, sizeof(double) * 3 * g.natoms, hipMemcpyHostToDevice,
                         p.stream),
          "upload positions");
  if ((rc = sqc_hip_forward_device(plan, plan->g_pos, 1, plan->g_out, plan->g_rho, p.stream)))
    return rc;
  std::vector<double> num(g.nq), gs(g.nq);
  HIP_TRY(hipMemcpyAsync(num.data(), plan->g_out, sizeof(double) * g.nq, hipMemcpyDeviceToHost,
                         p.stream),
          "download shells");
  HIP_TRY(hipStreamSynchronize(p.stream), "hipStreamSynchronize");
  double L = 0.0;
  for (int s = 0; s < g.nq; ++s) {
    const double S = p.den[s] > 0.0 ? num[s] / p.den[s] : 0.0;
    const double w = wq ? wq[s] : 1.0;
    const double r = S - target[s];
    L += w * r * r;
    gs[s] = 2.0 * w * r;
  }
  if (loss) *loss = L;
  HIP_TRY(hipMemcpyAsync(plan->g_gs, gs.data(), sizeof(double) * g.nq, hipMemcpyHostToDevice,
                         p.stream),
          "upload gS");
  if ((rc = sqc_hip_adjoint_device(plan, plan->g_pos, plan->g_rho, plan->g_gs, nullptr,
                                   plan->g_grad, p.stream)))
    return rc;
  HIP_TRY(hipMemcpyAsync(grad, plan->g_grad, sizeof(double) * 3 * g.natoms,
                         hipMemcpyDeviceToHost, p.stream),
          "download gradient");
  HIP_TRY(hipStreamSynchronize(p.stream), "hipStreamSynchronize");
  return SQC_HIP_OK;
}

int sqc_hip_profile(sqc_hip_plan *plan, int enable) {
  if (int rc = check_plan(plan)) return rc;
  sqc::Plan &p = plan->p;
  HIP_TRY(hipSetDevice(p.device), "hipSetDevice");
  const int e = sqc::collect_stage_times(p);
  if (e) return hip_fail((hipError_t)e, "stage timing");
  p.profiling = enable ? 1 : 0;
  return SQC_HIP_OK;
}

int sqc_hip_stage_times(sqc_hip_plan *plan, double *ms, int64_t *launches, int nstage) {
  if (int rc = check_plan(plan)) return rc;
  sqc::Plan &p = plan->p;
  HIP_TRY(hipSetDevice(p.device), "hipSetDevice");
  const int e = sqc::collect_stage_times(p);
  if (e) return hip_fail((hipError_t)e, "stage timing");
  for (int i = 0; i < nstage && i < sqc::kNumStages; ++i) {
    if (ms) ms[i] = p.stage_ms[i];
    if (launches) launches[i] = p.stage_n[i];
    p.stage_ms[i] = 0.0;
    p.stage_n[i] = 0;
  }
  return SQC_HIP_OK;
}

int sqc_hip_destroy(sqc_hip_plan *plan) {
  if (!plan) return SQC_HIP_OK;
  (void)hipSetDevice(plan->p.device);
  (void)hipStreamSynchronize(plan->p.stream);
  (void)sqc::collect_stage_times(plan->p);
  free_plan(plan);
  return SQC_HIP_OK;
}

const char *sqc_hip_error_string(int code) {
  switch (code) {
    case SQC_HIP_OK: return "success";
    case SQC_HIP_WARN_EPS_TOO_SMALL: return "requested tolerance too small; clipped";
    case SQC_HIP_ERR_MAXNALLOC: return "grid too large";
    case SQC_HIP_ERR_ALLOC: return "allocation failed";
    case SQC_HIP_ERR_BAD_ARG: return "invalid argument";
    case SQC_HIP_ERR_ABI: return "sqc_hip_config ABI mismatch";
    case SQC_HIP_ERR_NO_DEVICE: return "no usable HIP device";
    case SQC_HIP_ERR_HIP: return "HIP runtime error";
    case SQC_HIP_ERR_NO_MODES: return "no reciprocal lattice points in the q range";
    case SQC_HIP_ERR_POSITIONS: return "atom positions outside what the pair list holds";
    default: return "unknown error";
  }
}

const char *sqc_hip_last_error(void) { return g_last_error.c_str(); }

}  // extern "C"
