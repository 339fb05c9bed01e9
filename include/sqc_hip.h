/*
 * sqc_hip.h -- C ABI of the MI355X (gfx950) reciprocal-space S(q) library.
 *
 * This library is the drop-in replacement for the reference's GPU plug-in
 * (src/sqc_gpu.f90, cufinufft_structure_factor_t) and the cuFINUFFT binding
 * below it (src/sqc_cufinufft.f90).  Instead of handing a transform plan and
 * raw device buffers to the host, it owns the whole per-frame path: wrap and
 * bin the atoms, spread onto the upsampled grid, the 3D FFT, deconvolution,
 * species combination and the FP64 shell reduction.  The host downloads only
 * the per-shell sums.
 *
 * Conventions
 *   - All arrays are host memory unless a name says "device".
 *   - cell[9] is column-major with the lattice vectors a1, a2, a3 in its
 *     columns, exactly the layout of cell_t%a (src/sqc_cell.f90:17-31).
 *   - Species are the LAMMPS type ids present in the first frame, ascending
 *     (src/sqc_structure_factor.f90:1127-1144); species_of[] is 0-based.
 *   - Shell arrays are [nq], 0-based shell s = reference shell s+1.
 *   - Pair arrays are [S][S][nq]; only a <= b is filled, matching
 *     partial_num(t_a, t_b, :) of src/sqc_structure_factor.f90:481-521.
 *   - Return value: SQC_HIP_OK or an error code; sqc_hip_error_string()
 *     gives the text and sqc_hip_last_error() a message with detail.
 *     The library never exits the process.
 *   - One plan is driven by one host thread at a time.  All device work of a
 *     plan runs on the plan's own stream; the fetch and gradient calls are
 *     the synchronisation points.
 */
#ifndef SQC_HIP_H
#define SQC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SQC_HIP_ABI_VERSION 2

/* Return codes.  Values 1..20 keep the meaning of FINUFFT's codes where one
 * applies (external/finufft/include/finufft_errors.h:5-31). */
enum {
  SQC_HIP_OK = 0,
  SQC_HIP_WARN_EPS_TOO_SMALL = 1,   /* eps clipped to what the kernel width allows */
  SQC_HIP_ERR_MAXNALLOC = 2,        /* grid too large */
  SQC_HIP_ERR_ALLOC = 11,           /* device or host allocation failed */
  SQC_HIP_ERR_BAD_ARG = 20,         /* invalid argument (message has detail) */
  SQC_HIP_ERR_ABI = 30,             /* config struct from another ABI version */
  SQC_HIP_ERR_NO_DEVICE = 31,       /* no usable HIP device */
  SQC_HIP_ERR_HIP = 32,             /* a HIP runtime call failed */
  SQC_HIP_ERR_NO_MODES = 33,        /* no reciprocal lattice point in [qmin, qmax] */
  SQC_HIP_ERR_UNKNOWN_ELEMENT = 34, /* species label names no element */
  SQC_HIP_ERR_NO_CHARGE_STATE = 35, /* element known, charge or valence state not tabulated */
  SQC_HIP_ERR_POSITIONS = 36        /* Debye: an atom outside a non-periodic box, or adrift */
};

enum { SQC_HIP_F64 = 0, SQC_HIP_F32 = 1 };
enum { SQC_HIP_WEIGHT_UNIT = 0, SQC_HIP_WEIGHT_NEUTRON = 1, SQC_HIP_WEIGHT_XRAY = 2 };
enum { SQC_HIP_NORM_MEAN = 1, SQC_HIP_NORM_SELF = 2, SQC_HIP_NORM_NATOM = 3 };

/* Plan configuration (plain data; the pointers are read during
 * sqc_hip_plan_create only). */
typedef struct sqc_hip_config {
  int32_t abi_version;      /* must be SQC_HIP_ABI_VERSION */
  int32_t device;           /* HIP device ordinal (--gpu-id) */
  int32_t precision;        /* SQC_HIP_F64 or SQC_HIP_F32 (--precision) */
  int32_t weight_kind;      /* SQC_HIP_WEIGHT_* (-w) */
  int32_t norm;             /* SQC_HIP_NORM_* (--norm) */
  int32_t partials;         /* accumulate Re(rho_a rho_b*) per shell (--partials) */
  int32_t want_grid;        /* accumulate |rho|^2 per kept mode (--grid) */
  int32_t nq;               /* number of shells (--nq) */
  double qmin, qmax;        /* shell range (--qmin, --qmax) */
  double eps;               /* NUFFT tolerance (--eps) */
  double upsampfac;         /* 0 = automatic (2.0 for f64, 1.25 for f32) */
  double cell[9];           /* lattice vectors in columns */
  double origin[3];         /* lower corner of the cell */
  int64_t natoms;
  int32_t nspecies;         /* species present, 1..32 */
  int32_t mode_partials;    /* with partials: also accumulate Re(rho_a rho_b*) per
                               kept mode (the pair sums behind the --xrd pair
                               columns, src/sqc_structure_factor.f90:538-588) */
  const int32_t *species_of;  /* [natoms], 0-based species of each atom */
  const double *amp_const;    /* [nspecies] neutron amplitudes (ignored, and may be
                                 NULL, under unit and X-ray weights: the X-ray
                                 amplitude is f_a(|q|) from xray_coef alone) */
  const double *xray_coef;    /* [nspecies][9] IT92 a1..a4, b1..b4, c */
  int32_t tune;             /* 1 = time the spread tiles, y-pass chunks and x-pass
                               line groups on the first frame, keep the fastest */
  int32_t max_batch;        /* frames per captured hipGraph replay (0 = 8); 1 =
                               launch every frame kernel by kernel.  Above 1 the
                               plan holds a second spread buffer: a replay spreads
                               a frame while the previous one is transformed */
} sqc_hip_config;

typedef struct sqc_hip_plan sqc_hip_plan;

/* Plan geometry and sizes, for reports and for callers that size buffers. */
typedef struct sqc_hip_info {
  int64_t modes[3];         /* m_i, src/sqc_structure_factor.f90:329-338 */
  int64_t nmodes;           /* kept modes K (both halves) */
  int64_t nhalf;            /* kept modes with h in the canonical half space */
  int64_t fine[3];          /* upsampled grid n_i */
  int32_t nspread;          /* kernel width ns */
  int32_t nfields;          /* transforms per frame (species, or 1 weighted) */
  double beta;              /* ES kernel shape */
  double upsampfac;         /* sigma actually used */
  int64_t device_bytes;     /* device memory held by the plan */
  int64_t nlines;           /* x-lines of the pruned half spectrum */
  int32_t tile[3];          /* spread tile (x, y) and y-pass column chunk */
  int32_t ncomp;            /* 1 + species pairs: rows of a per-frame output */
  int64_t rho_len;          /* complex values per field in a saved rho buffer */
  int32_t shell_sums_in_hbm; /* 1: the x pass adds its shell sums in its own row of
                               HBM (ncomp * nq doubles do not fit the LDS beside one
                               x line); 0: in LDS */
  int32_t x_lines;          /* x lines per workgroup group of the x pass */
} sqc_hip_info;

/* Number of visible HIP devices (cudaGetDeviceCount in src/sqc_cufinufft.f90:143-174). */
int sqc_hip_device_count(int *count);

/* Create a plan: sizes, mode tables, denominators and device buffers
 * (replaces cufinufft_makeplan, src/sqc_gpu.f90:185-209, and the mode/shell
 * set-up of sf_configure, src/sqc_structure_factor.f90:313-422). */
int sqc_hip_plan_create(sqc_hip_plan **plan, const sqc_hip_config *cfg);

int sqc_hip_plan_info(const sqc_hip_plan *plan, sqc_hip_info *info);

/* Pin the launch shapes the autotuner chooses (sqc_hip_info.tile and
 * x_lines): spread tile spread_x x spread_y real columns (16x2, 8x2, 16x1,
 * 4x2, 8x1, 4x1 or 2x1, as the plan's LDS and bin window allow), y-pass
 * column chunk 2, 4, 8 or 16, x-pass lines per group 1..8.  The plan then
 * skips tuning.  Shapes the plan cannot run return SQC_HIP_ERR_BAD_ARG and
 * leave it unchanged.  Waits for the plan's queued work.  Counterpart of
 * the cuFINUFFT spreader options gpu_binsizex/y/z and gpu_maxsubprobsize
 * the reference passes with its plan (src/sqc_cufinufft.f90:45-53), which
 * it leaves at their defaults (:281-289). */
int sqc_hip_set_tiles(sqc_hip_plan *plan, int spread_x, int spread_y, int y_chunk, int x_lines);

/* Shell bookkeeping of the plan (den and shell_count of
 * src/sqc_structure_factor.f90:408-416). */
int sqc_hip_shells(const sqc_hip_plan *plan, double *den, int64_t *shell_count);

/* Kept modes in the reference's enumeration order (hkl[3K] int32, qvec[3K],
 * gidx[K] 1-based), for the --grid table (src/sqc_structure_factor.f90:815-823). */
int sqc_hip_modes(const sqc_hip_plan *plan, int32_t *hkl, double *qvec, int64_t *gidx);

/* Accumulate nframes frames.  pos holds Cartesian positions, (3, natoms)
 * column-major per frame, frames consecutive (frame_t%pos of
 * src/sqc_dump.f90:22-45).  Host memory; copied before the call returns.
 * Replaces gpu_accumulate (src/sqc_gpu.f90:213-362). */
int sqc_hip_accumulate(sqc_hip_plan *plan, const double *pos, int64_t nframes);

/* Same with the positions already in device memory (HBM-resident input). */
int sqc_hip_accumulate_device(sqc_hip_plan *plan, const double *pos_device, int64_t nframes);

/* Synchronise and copy out the accumulators: num[nq], partial[S][S][nq]
 * (may be NULL), gnum[K] in reference mode order (may be NULL), nframes. */
int sqc_hip_fetch(sqc_hip_plan *plan, double *num, double *partial, double *gnum,
                  int64_t *nframes);

/* The per-mode pair sums of a plan made with mode_partials: gpart[npairs][K],
 * pairs (a, b), a <= b, in species order, modes in the order of gnum.  The
 * per-mode analogue of partial[] (src/sqc_structure_factor.f90:538-588 bins
 * them by two-theta with the amplitudes and Lorentz-polarization factor). */
int sqc_hip_fetch_mode_partials(sqc_hip_plan *plan, double *gpart);

/* Zero the accumulators. */
int sqc_hip_reset(sqc_hip_plan *plan);

/* Per-frame forward on device memory, for the autograd layer.  For each of
 * nframes frames writes, on the plan's stream:
 *   out_device[f][1 + npairs][nq]  num, then the a<=b pair sums
 *   rho_device (optional) the species amplitudes at the canonical kept modes,
 *              [f][nfields][rho_len] complex of the plan precision.
 * stream: the caller's hipStream_t.  The work runs on the plan's stream,
 * ordered after everything already queued on `stream` and before anything
 * queued on it later; NULL is the legacy default stream. */
int sqc_hip_forward_device(sqc_hip_plan *plan, const double *pos_device, int64_t nframes,
                           double *out_device, void *rho_device, void *stream);

/* Adjoint on device memory: given the upstream gradients of one frame's
 * normalised shell values,
 *   gS_device[nq]            dL/dS_s,      S_s  = num_s / den_s
 *   gP_device[S][S][nq]      dL/dP_ab,s,   P_ab = pair_ab,s / (N count_s)  (may be NULL)
 * and the rho saved by sqc_hip_forward_device for the same positions, writes
 * grad_device[3*natoms] = dL/dr (FP64).  No counterpart in the reference. */
int sqc_hip_adjoint_device(sqc_hip_plan *plan, const double *pos_device, const void *rho_device,
                           const double *gS_device, const double *gP_device,
                           double *grad_device, void *stream);

/* sqc_hip_adjoint_device for callers that cannot allocate device memory:
 * code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc.  No counterpart in the reference.  Same arguments and result;
 * allocates nothing.  The kernels' gradient slots (3 doubles per (atom,
 * tile) pair: natoms x sx x sy x 24 bytes, sx x sy = 28 at the 8 x 2 spread
 * tile and ns = 10) live in a static device buffer of
 * SQC_HIP_NOALLOC_SCRATCH_BYTES that each GPU holds until the process ends,
 * instead of the buffer sqc_hip_adjoint_device allocates on its first call.
 * A plan whose slots exceed it (about 150,000 atoms at that tile) gets
 * SQC_HIP_ERR_MAXNALLOC with the sizes in the message.  Every plan on a GPU
 * shares the buffer and it keeps the last call's contents (nothing in it
 * needs to start at zero): calls that overlap on one GPU, from other host
 * threads or on streams not ordered with each other, race. */
#define SQC_HIP_NOALLOC_SCRATCH_BYTES ((size_t)96 << 20)
int sqc_hip_adjoint_device_noalloc(sqc_hip_plan *plan, const double *pos_device,
                                   const void *rho_device, const double *gS_device,
                                   const double *gP_device, double *grad_device, void *stream);

/* Convenience: loss and gradient of L = sum_s w_s (S_s - T_s)^2 for one
 * frame, host memory in and out (target[nq], wq[nq], grad[3*natoms]). */
int sqc_hip_gradient(sqc_hip_plan *plan, const double *pos, const double *target,
                     const double *wq, double *loss, double *grad);

/* Per-stage timing with HIP events on the launch stream (no counterpart in
 * the reference, whose only timing is the driver's wall clock,
 * app/main.f90:218-220).  Stages: 0 atom wrap+sort, 1 spread+z-DFT, 2 y-DFT,
 * 3 x-DFT+shell sums, 4 finalize, 5 adjoint x, 6 adjoint y, 7 adjoint
 * z+gather, 8 adjoint slot sum.  sqc_hip_stage_times synchronises, returns
 * the milliseconds and launch counts accumulated since the last call, and
 * clears them. */
int sqc_hip_profile(sqc_hip_plan *plan, int enable);
int sqc_hip_stage_times(sqc_hip_plan *plan, double *ms, int64_t *launches, int nstage);

int sqc_hip_destroy(sqc_hip_plan *plan);

const char *sqc_hip_error_string(int code);
const char *sqc_hip_last_error(void);

/* sizeof(sqc_hip_config), for the host-side layout check that mirrors the
 * reference's opts ABI check (src/sqc_cufinufft.f90:277-290). */
size_t sqc_hip_config_size(void);

/* ---- dyn: density amplitudes on a set of Miller indices (FP64, MFMA) ----
 *
 * The amplitudes of the reference's dynamic method, per frame and species,
 *   rho_a(n) = sum_{j in a} exp(i (n1 b1 + n2 b2 + n3 b3) . r_j + i base . r_j)
 * for modes n = (n1, n2, n3): the reciprocal lattice (b = cell%b, no base)
 * or a q line (b1 = dq u, base = s0 u).  Replaces the rho block of
 * dyn_accumulate (src/sqc_dynamics.f90:826-873, with dyn_atom_phase :1094-1115
 * and phase_tables / phase_factor of src/sqc_phase.f90:67-113); the time
 * correlations of the amplitudes stay on the host.  Positions are Cartesian,
 * unwrapped as the reference requires, (3, natoms) per frame. */
typedef struct sqc_hip_dyn_config {
  int32_t abi_version;      /* SQC_HIP_ABI_VERSION */
  int32_t device;
  int64_t natoms;
  int32_t nspecies;
  int32_t base_on;          /* 1: add base . r to every phase (q line) */
  const int32_t *species_of; /* [natoms], 0-based */
  double recip[9];          /* b1, b2, b3 in the columns (recip of dyn_adopt_modes) */
  double base[3];
  int64_t nmodes;
  const int32_t *hkl;       /* [nmodes][3] Miller indices (mode_index) */
} sqc_hip_dyn_config;

typedef struct sqc_hip_dyn_plan sqc_hip_dyn_plan;

typedef struct sqc_hip_dyn_info {
  int64_t nmodes;
  int64_t macs_per_frame;   /* complex multiply-adds the matrix cores run per frame */
  int64_t computed_pairs;   /* (row, column) products per atom: live 16 x 16 tiles x 256 */
  int64_t device_bytes;
  int32_t nitems;           /* 16 x 128 blocks of the mode product */
  int32_t nchunks;          /* atom chunks (per species) */
  int32_t chunk_atoms;
  int32_t column_axis;      /* Miller index that runs along the product's columns */
} sqc_hip_dyn_info;

int sqc_hip_dyn_plan_create(sqc_hip_dyn_plan **plan, const sqc_hip_dyn_config *cfg);
int sqc_hip_dyn_plan_info(const sqc_hip_dyn_plan *plan, sqc_hip_dyn_info *info);
/* rho_device[f][nspecies][nmodes] complex (re, im) FP64 for nframes frames of
 * device positions, ordered on `stream` like sqc_hip_forward_device. */
int sqc_hip_dyn_rho_device(sqc_hip_dyn_plan *plan, const double *pos_device, int64_t nframes,
                           double *rho_device, void *stream);
/* The same from and to host memory (synchronous). */
int sqc_hip_dyn_rho(sqc_hip_dyn_plan *plan, const double *pos, int64_t nframes, double *rho);
/* HIP-event times on the plan stream: 0 atom phases, 1 matrix product,
 * 2 chunk sum and gather (as sqc_hip_profile / sqc_hip_stage_times). */
int sqc_hip_dyn_profile(sqc_hip_dyn_plan *plan, int enable);
int sqc_hip_dyn_stage_times(sqc_hip_dyn_plan *plan, double *ms, int64_t *launches, int nstage);
int sqc_hip_dyn_destroy(sqc_hip_dyn_plan *plan);

/* ---- Debye: ordered pair-distance histograms per type pair (FP64) ----
 *
 * The pair counts of the reference's Debye method, histogram(a, b, k) = the
 * number of ordered pairs (i, j), i of type a, j of type b, whose distance
 * falls in bin k = int(|d| / dr) + 1 <= nbins = max(1, ceiling(rmax / dr)),
 * |d| <= rmax, summed over the frames.  Replaces debye_accumulate
 * (src/sqc_debye.f90:179-235) with the displacement of its cell list
 * (nl_begin_frame / nl_build, src/sqc_neighbour_list.f90:190-330): every pair
 * lands in the bin the reference puts it in.  Set-up follows debye_setup
 * (src/sqc_debye.f90:95-177): the same checks and messages, rmax below the
 * smallest periodic width.  The cell and origin are those of the first frame
 * (the reference's list keeps them); positions are Cartesian, (3, natoms)
 * per frame, in any wrapping. */
typedef struct sqc_hip_debye_config {
  int32_t abi_version;      /* SQC_HIP_ABI_VERSION */
  int32_t device;
  int64_t natoms;
  int32_t ntypes;           /* type ids run 1..ntypes */
  int32_t pbc[3];           /* periodicity of each cell vector */
  const int32_t *type_id;   /* [natoms], 1-based as in the dump */
  double cell[9];           /* lattice vectors a1, a2, a3 in the columns */
  double origin[3];
  double rmax;              /* pair cutoff (> 0; the caller resolves the default) */
  double dr;                /* bin width */
  double skin;              /* Verlet skin of the reference's list: sets its cells */
} sqc_hip_debye_config;

typedef struct sqc_hip_debye_plan sqc_hip_debye_plan;

typedef struct sqc_hip_debye_info {
  int32_t nbins;
  int32_t ntypes;
  int32_t div[3];           /* the reference list's cells per cell vector */
  int32_t hist_in_lds;      /* 1: each workgroup counts into an LDS histogram */
  int32_t workgroups;
  int32_t tile;             /* atoms per tile = threads per workgroup (256 or 1024) */
  int64_t pairs_per_frame;  /* unordered atom pairs evaluated per frame */
  int64_t device_bytes;
  int64_t nframes;          /* frames counted since create / reset */
} sqc_hip_debye_info;

int sqc_hip_debye_plan_create(sqc_hip_debye_plan **plan, const sqc_hip_debye_config *cfg);
int sqc_hip_debye_plan_info(const sqc_hip_debye_plan *plan, sqc_hip_debye_info *info);
/* Count nframes frames of device positions ([f][natoms][3]), ordered on
 * `stream` like sqc_hip_forward_device: the plan's work starts after what
 * was queued on it and work queued on it afterwards waits for the counts. */
int sqc_hip_debye_accumulate_device(sqc_hip_debye_plan *plan, const double *pos_device,
                                    int64_t nframes, void *stream);
/* The same from host memory (synchronous). */
int sqc_hip_debye_accumulate(sqc_hip_debye_plan *plan, const double *pos, int64_t nframes);
/* histogram[a][b][k] (ntypes x ntypes x nbins, C order, 0-based) of ordered
 * pair counts summed over the frames, as doubles like the reference's
 * histogram before its division by nhist; *nframes the frames counted.
 * Either pointer may be NULL. */
int sqc_hip_debye_fetch(sqc_hip_debye_plan *plan, double *histogram, int64_t *nframes);
/* SQC_HIP_ERR_POSITIONS from fetch (and from the synchronous accumulate):
 * some frame had an atom outside the box along a non-periodic axis (the
 * reference's "an atom lies outside a non-periodic box", which it reports
 * on the frames that rebuild its list) or 2^20 cells away from its first
 * image; such an atom's pairs are not counted along that axis. */
/* Clear the counts, the frame count and the position flags (the list's
 * unwrapped positions carry on, as the reference's do). */
int sqc_hip_debye_reset(sqc_hip_debye_plan *plan);
/* HIP-event times on the plan stream: 0 list positions, 1 pair counts,
 * 2 row sum (as sqc_hip_profile / sqc_hip_stage_times). */
int sqc_hip_debye_profile(sqc_hip_debye_plan *plan, int enable);
int sqc_hip_debye_stage_times(sqc_hip_debye_plan *plan, double *ms, int64_t *launches,
                              int nstage);
int sqc_hip_debye_destroy(sqc_hip_debye_plan *plan);

/* ---- Species labels and their scattering data (host side, no device) ----
 *
 * The -m TYPE:LABEL mapping resolves each label to an element and, for the
 * X-ray weights, to a row of the IT92 table: a neutral atom ("Si"), an ion
 * ("O2-", "Na1+", "Fe+") or a valence state ("Cval").  Replaces
 * element_table_t%find_species (src/sqc_elements.f90:241-291) with the data
 * of src/sqc_element_data.f90 (104 elements, 213 X-ray species, 2 aliases);
 * the amplitudes follow weight_scheme_t%amplitude (src/sqc_weights.f90:
 * 136-154): b of the element under neutron weights, f(q) = c + sum_i a_i
 * exp(-b_i (q / 4 pi)^2) of the label under X-ray weights. */
typedef struct sqc_hip_species_info {
  char label[16];           /* canonical label ("Si", "O2-", "Na1+", "Cval") */
  char symbol[4];           /* element symbol */
  int32_t z;                /* atomic number */
  int32_t charge;           /* formal charge (0 for neutral atoms and valence states) */
  int32_t valence;          /* 1 for a valence-state row */
  int32_t has_neutron;      /* 1: the element's bound coherent length is tabulated */
  int32_t has_xray;         /* 1: xray_coef holds the label's IT92 row */
  int32_t reserved;
  double neutron_b;         /* fm, of the element (ions take their element's) */
  double xray_coef[9];      /* a1..a4, b1..b4, c (config xray_coef layout) */
} sqc_hip_species_info;

/* Resolve a label as table_find_species does: SQC_HIP_OK (a bare symbol of
 * an element without an X-ray row resolves with has_xray = 0),
 * SQC_HIP_ERR_UNKNOWN_ELEMENT or SQC_HIP_ERR_NO_CHARGE_STATE, with the
 * reference's message (listing the tabulated states) in
 * sqc_hip_species_error(). */
int sqc_hip_species_lookup(const char *label, sqc_hip_species_info *info);
/* Row `index` of the element table (kind 0: each element with its neutral
 * X-ray row when it has one) or of the X-ray species table (kind 1);
 * *count (may be NULL) gets the table's length, info may be NULL. */
int sqc_hip_species_table(int32_t kind, int32_t index, sqc_hip_species_info *info,
                          int32_t *count);
const char *sqc_hip_species_error(void);

/* ---- LAMMPS dump reader (host side, no device needed) ----
 *
 * Replaces the reference's trajectory reader lammps_dump_reader_t
 * (src/sqc_dump.f90:72-283: open, next_frame, close) and the driver's
 * per-frame checks (app/main.f90:240-247), and feeds sqc_hip_accumulate
 * directly: sqc_dump_read_frames writes a batch of frames in its pos layout.
 * Frames follow the reference reader: header records, BOX BOUNDS to the
 * restricted triclinic cell (src/sqc_cell.f90:39-63), positions unwrapped
 * across frames unless the dump has xu yu zu (src/sqc_dump.f90:300-364).
 * The scaled columns xs ys zs / xsu ysu zsu are accepted as well.  The file
 * is memory-mapped and each atom block is parsed by nthreads threads. */

/* Status codes: 0 and 1..11 are the reference reader's ierr values. */
enum {
  SQC_DUMP_OK = 0,
  SQC_DUMP_END = 1,                /* clean end of the trajectory */
  SQC_DUMP_ERR_OPEN = 2,
  SQC_DUMP_ERR_CLOSED = 3,
  SQC_DUMP_ERR_HEADER = 4,         /* no TIMESTEP where a frame starts, bad step, early end */
  SQC_DUMP_ERR_NATOMS = 5,         /* NUMBER OF ATOMS not a positive integer */
  SQC_DUMP_ERR_BOX = 6,            /* BOX BOUNDS row with fewer than two numbers */
  SQC_DUMP_ERR_NO_ATOMS = 7,       /* frame without BOX BOUNDS */
  SQC_DUMP_ERR_ATOM_LINE = 8,      /* atom line missing or not all-numeric */
  SQC_DUMP_ERR_COLUMNS = 9,        /* ATOMS line without type or coordinates */
  SQC_DUMP_ERR_ORDER = 10,         /* ids change order between frames */
  SQC_DUMP_ERR_COUNT = 11,         /* atom count changes (unwrapping state) */
  SQC_DUMP_ERR_NATOMS_CHANGED = 12, /* app/main.f90:240-243 (stop 8) */
  SQC_DUMP_ERR_BOX_CHANGED = 13,   /* app/main.f90:244-247 (stop 9) */
  SQC_DUMP_ERR_ARG = 14
};

typedef struct sqc_dump_reader sqc_dump_reader;

/* One frame's header: frame_t of src/sqc_dump.f90:22-45. */
typedef struct sqc_dump_frame {
  int64_t natoms;
  int64_t timestep;
  double cell[9];            /* lattice vectors in the columns (cell_t%a) */
  double origin[3];
  int32_t pbc[3];            /* 1 where the BOX BOUNDS flag starts with 'p' */
  int32_t triclinic;         /* the BOX BOUNDS line named xy */
  int32_t scaled;            /* coordinates came from xs ys zs (or xsu ysu zsu) */
  int32_t has_id;
  int32_t unwrapped;         /* always 1: positions are continuous */
  int32_t unwrapped_from_dump;
  int64_t n_ambiguous;       /* atom steps above 0.4 of a cell vector so far */
  int64_t lineno;            /* lines read so far */
  int64_t frames;            /* frames read so far */
} sqc_dump_frame;

/* nthreads > 0: that many parse threads; nthreads <= 0: the host's threads
 * (OMP_NUM_THREADS when set, else the hardware count) less -nthreads, so -1
 * leaves a core to a thread that drives the device. */
int sqc_dump_open(sqc_dump_reader **reader, const char *path, int nthreads);
/* Next frame into the reader's arrays; SQC_DUMP_END at the end.  After an
 * error the reader returns that error from then on. */
int sqc_dump_next(sqc_dump_reader *reader, sqc_dump_frame *frame);
/* The last sqc_dump_next frame: pos[natoms][3], id[natoms], type_id[natoms]
 * (any may be NULL). */
int sqc_dump_copy(const sqc_dump_reader *reader, double *pos, int64_t *id, int32_t *type_id);
/* Up to maxframes frames of natoms atoms into pos[maxframes][natoms][3],
 * the layout of sqc_hip_accumulate.  *nread < maxframes means the end of the
 * file.  Each frame must keep natoms atoms and, when cell is non-null, that
 * cell (app/main.f90:240-247).  last (may be NULL) gets the last frame's
 * header. */
int sqc_dump_read_frames(sqc_dump_reader *reader, int64_t maxframes, int64_t natoms,
                         const double cell[9], double *pos, int64_t *nread,
                         sqc_dump_frame *last);
/* sqc_dump_read_frames on a thread of its own: _begin returns at once, the
 * batch is written into pos while the caller works (for instance runs the
 * previous batch through sqc_hip_accumulate); _end waits for it and returns
 * its status and frame count.  One batch in flight per reader; the reader
 * must not be used otherwise until _end.  No counterpart in the reference,
 * whose reader is synchronous. */
int sqc_dump_read_frames_begin(sqc_dump_reader *reader, int64_t maxframes, int64_t natoms,
                               const double cell[9], double *pos);
int sqc_dump_read_frames_end(sqc_dump_reader *reader, int64_t *nread);
int sqc_dump_close(sqc_dump_reader *reader);
const char *sqc_dump_error_string(int code);
/* Detail of the error that ended the reader, with its line number. */
const char *sqc_dump_message(const sqc_dump_reader *reader);

#ifdef __cplusplus
}
#endif

#endif /* SQC_HIP_H */
