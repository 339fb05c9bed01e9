// sqc_internal.hpp -- plan layout shared by the host code and the gfx950 kernels.
//
// Data layout in HBM (F = fields, i.e. species transforms or one weighted
// transform; T2 = float2 or double2; n = fine grid, M = (m-1)/2):
//   pos      [batch][natoms][3]       f64   Cartesian input, frame_t%pos order
//   W1       [F][M3+1][n2][n1]        T2    after spread + z-DFT, h3 >= 0 only
//   W2       [F][nlines][n1]          T2    after the y-DFT, canonical (h2,h3) lines
//   stab     [ncand]                  u16   shell of each candidate mode (0xFFFF: not kept)
//   wgpart   [nblk][words][ncomp][nq] i64   per-workgroup shell partials of one frame
//                                           (integer words of the exact sums)
//   acc      [2][ncomp][nq]           f64   run accumulators (sum, compensation)
// Real strengths make F[g] real, so only the half spectrum h3 > 0, or h3 = 0
// and h2 > 0, or h3 = h2 = 0 and h1 > 0 is transformed; its modes count twice.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "sqc_hip.h"

namespace sqc {

constexpr int kMaxStages = 12;
// Species a plan takes, and the species whose weights ride in Geo itself
// (plans with more read them from the plan's weight table in HBM, and the
// x passes run their any-field-count instance).
constexpr int kMaxSpecies = 32;
constexpr int kGeoSpecies = 8;
// Weight table: per species its constant amplitude and IT92 a1..a4 b1..b4 c.
constexpr int kWtab = 10;
constexpr uint16_t kNoShell = 0xFFFF;
// timed stages: 0 atom sort, 1 spread+z, 2 y, 3 x+bin, 4 finalize,
// 5 adjoint x, 6 adjoint y, 7 adjoint z+gather, 8 adjoint slot sum
constexpr int kNumStages = 9;

// Mixed-radix factorisation of one axis.
struct AxisPlan {
  int n;
  int nst;
  int rad[kMaxStages];
  // ceil(2^32 / Ns) of each stage (0 when Ns = 1): j / Ns = umulhi(j, magic)
  uint32_t magic[kMaxStages];
};

// Longest line whose DFT stages run in place in registers (fine-grid points
// per axis); longer lines, up to kMaxLongLine, run out of place through a
// scratch line per wave.
constexpr int kMaxLine = 1024;
constexpr int kMaxLongLine = 2048;
// Most lines an x-pass workgroup transforms per step.
constexpr int kMaxLb = 8;
// Atoms staged into LDS per round of the spreader / gather.
constexpr int kStageCap = 256;
constexpr int kThreads = 256;   // threads per workgroup of the LDS kernels

// LDS stride of a z column of the spread tile: room for the DFT's padded
// stage layout (n + n/32 + 1 points), rounded up to 12 words mod 32 so that
// the neighbouring columns of a stencil update fall on different banks.
constexpr int spread_col_stride(int n3) {
  const int cap = n3 + n3 / 32 + 1;
  return cap + ((12 - cap) % 32 + 32) % 32;
}

// Everything a kernel needs to know about the geometry, passed by value.
struct Geo {
  double ainv[9];   // A^-1, column-major
  double origin[3];
  double b[9];      // reciprocal basis, columns b1 b2 b3
  int n[3];         // fine grid
  int M[3];         // (m - 1) / 2
  int ns;
  double beta;
  double half_ns;
  double inv_half_ns;  // 2 / ns
  double inv_per;      // 1 / (3 ns): atom of a kernel-value thread
  double dphi_smin;    // lower bound on sqrt(1 - z^2) in the kernel derivative
  // spread tiles (real x-columns x rows) and the bins that sort atoms by tile
  int txr, ty, ntx, nty, nbins;
  int ltxr, lty;     // log2 of txr and ty
  int cs;           // LDS stride of a z column in the spread tile (padded)
  int ls2;          // LDS stride of a y line (padded)
  int rs1;          // LDS stride of an x row (padded)
  int hx, hy;       // bin halo each side
  int sx, sy;       // adjoint gradient slots per atom per axis
  int F;            // fields
  int S;            // species
  int ncomp;        // 1 + number of species pairs (when partials)
  int nq;
  int weight_kind;
  int nlines;
  int partials;
  int64_t natoms;
  double amp[kGeoSpecies];
  // factor on the pair sums: 1, or 1 / amp[0]^2 for a one-field plan under
  // neutron or unit weights, whose single field carries the species weight
  // in its strength (the pair columns are of the unweighted amplitudes)
  double pair_scale;
  // 1: the x pass keeps its shell sums in its own row of the per-workgroup
  // partials in HBM (FP64 atomics) because ncomp * nq doubles exceed the LDS
  int acc_global;
  // exact shell sums: 1 / q of the total and of the pair sums (q a power of
  // two), 2^k of the remainder word, and the words per sum (1 or 2)
  double xq_inv[2];
  double xq_lo;
  int xwords;
  double xcoef[kGeoSpecies][9];
  int nxcd;         // XCDs the workgroups are dealt to
};

struct Plan {
  sqc_hip_config cfg{};
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  int prec = 0;               // 0 f64, 1 f32
  Geo geo{};
  int64_t m[3]{};
  double sigma = 2.0;
  double dq = 0.0;
  int64_t nmodes = 0;         // K
  int64_t nhalf = 0;          // canonical kept modes
  int64_t ncand = 0;          // candidate slots in the line table
  std::vector<int32_t> species_of;
  std::vector<int64_t> counts;  // atoms per species
  std::vector<double> wtab;     // [S][kWtab] weight table (host copy of d_wtab)
  std::vector<double> den;
  std::vector<int64_t> shell_count;
  // per h3 = 0..M3: first line and the h2 interval
  std::vector<int32_t> lo2, hi2, loff;
  // per line: (h2, h3, lo1, hi1) and the candidate offset
  std::vector<int32_t> line_info;  // 4 per line
  std::vector<int32_t> line_off;
  AxisPlan ax[3]{};
  // launch shapes
  int tx2 = 16;               // y-pass x-columns per workgroup
  int lb = 8;                 // x-pass lines per workgroup step
  int nblk3 = 0;              // x-pass persistent workgroups
  int wgpart_rows = 0;        // rows d_wgpart was sized for
  int ncu = 1;                // compute units of the device
  size_t lds1 = 0, lds2 = 0, lds3 = 0;
  size_t lds1g = 0;           // gather: lds1 plus its per-wave kernel-value scratch
  int64_t nframes = 0;
  int64_t batch_cap = 1;
  int nbt_alloc = 0;           // bins the sort counters were sized for
  int64_t slots_bytes = 0;     // size of d_slots
  int tuned = 0;               // autotuner has run (or was not asked for)
  // batched frames: a captured graph of batch_cap frames read from d_pos
  hipGraphExec_t gexec = nullptr;
  int use_graph = 1;
  // the graph runs the y and x passes of frame i on stream2 while the atom
  // sort and spread of frame i + 1 run on stream, W1 alternating with W1b
  hipStream_t stream2 = nullptr;
  void *d_W1b = nullptr;
  std::vector<hipEvent_t> gev;   // fork/join events of the capture
  // device buffers
  double *d_pos = nullptr;
  int4 *d_i0 = nullptr;        // stencil start per atom (i0x, i0y, i0z, -)
  int4 *d_i0s = nullptr;       // the same in sorted (bin) order
  void *d_phi = nullptr;       // kernel values per atom [3][ns], path precision
  void *d_dphi = nullptr;      // kernel derivatives (adjoint)
  int32_t *d_key = nullptr, *d_count = nullptr, *d_start = nullptr, *d_cursor = nullptr,
          *d_sidx = nullptr, *d_species = nullptr;
  void *d_W1 = nullptr, *d_W2 = nullptr, *d_tw[3] = {nullptr, nullptr, nullptr};
  double *d_ph[3] = {nullptr, nullptr, nullptr};
  int32_t *d_lo2 = nullptr, *d_hi2 = nullptr, *d_loff = nullptr, *d_line_info = nullptr,
          *d_line_off = nullptr;
  uint16_t *d_stab = nullptr;
  double *d_line_d23 = nullptr;  // per x line: 1 / (phihat_2 phihat_3)
  double *d_den = nullptr, *d_cnt = nullptr;
  double *d_wtab = nullptr;      // [S][kWtab]
  unsigned long long *d_wgpart = nullptr;
  double *d_acc = nullptr, *d_frame = nullptr;
  double *d_gnum = nullptr;
  double *d_gpart = nullptr;     // [npairs][ncand] per-mode pair sums (mode_partials)
  double *d_slots = nullptr;
  double *d_gs = nullptr;
  void *d_rho = nullptr;
  int64_t bytes = 0;
  // per-stage timing (events recorded on the launch stream when enabled)
  int profiling = 0;
  struct Pending {
    int stage;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> ev_free;  // recycled timing events
  double stage_ms[kNumStages] = {};
  int64_t stage_n[kNumStages] = {};
};

// Launchers (sqc_kernels.hip).  All run on p.stream.
int launch_forward(Plan &p, const double *d_pos_frame, int mode, double *frame_out,
                   void *rho_out);
int launch_adjoint(Plan &p, const double *d_pos_frame, const void *rho, const double *gS,
                   const double *gP, double *grad, double *slots);
int launch_forward_front(Plan &p, const double *d_pos_frame, void *W1, hipStream_t st);
int launch_forward_y(Plan &p, const void *W1, void *W2, hipStream_t st);
int launch_forward_x(Plan &p, const void *W2, hipStream_t st, int mode, double *frame_out,
                     void *rho_out);
int launch_finalize(Plan &p, double *frame_out, int accumulate, hipStream_t st);
// Record msg as the thread's sqc_hip_last_error() text; returns code.
int set_error(int code, const std::string &msg);
int kernel_lds_limits(Plan &p);
// Static LDS (bytes) of the spread and gather kernels (their staging block)
// and of the x passes (their line-group table): a launch's dynamic LDS must
// leave room for it within the 160 KB of a CU.
size_t static_lds_spread();
size_t static_lds_xpass();
void persistent_grids(Plan &p);
// Fewer persistent x-pass workgroups when their HBM shell rows would exceed
// 512 MB (sqc_host.cpp).
void cap_wgpart_rows(Plan &p);
int collect_stage_times(Plan &p);

}  // namespace sqc
