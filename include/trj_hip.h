/*
 * trj_hip.h -- C ABI of the MI355X structural hot path of trj_analysis:
 * the species-resolved pair histogram g_ab(r) and the static structure
 * factor S(Q), as hand-written HIP kernels for gfx950.
 *
 * The reference has no plugin registry: its "operator API" is the set of
 * Fortran module procedures of mod_rdf / mod_sq, called by the driver
 * (src/main/trj_analysis.f90:193-217) and by src/core/util.f90:188-198,
 * 257-289, with all data passed through module globals (src/core/common.cuf).
 * Each entry point below names the reference procedure it replaces.
 *
 * Conventions
 *  - Every function returns TRJ_OK (0) or a negative TRJ_E* code; the library
 *    never exits.  trj_last_error() returns the message of the last failure
 *    on a context (the reference prints and stops, rdf.cuf:241-242).
 *  - Indices follow the reference: species types are 1..nsp; histograms are
 *    laid out as histomix_d(lsmax, nit) (bin fastest, pair index ij slowest);
 *    S(Q) shells are 0..nqmax.
 *  - One context per host thread; contexts are independent.  All device work
 *    of a context runs on its own HIP stream; calls on one context are
 *    asynchronous until a call that returns results (those synchronise).
 *  - Plain pointers and sizes only.  Host pointers are borrowed for the
 *    duration of the call.
 */
#ifndef TRJ_HIP_H
#define TRJ_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TRJ_ABI_VERSION 1

/* Return codes. */
#define TRJ_OK 0
#define TRJ_EINVAL (-1)   /* invalid argument or parameter */
#define TRJ_EDEVICE (-2)  /* HIP runtime error */
#define TRJ_ENOMEM (-3)   /* device or pinned allocation failed */
#define TRJ_ESTATE (-4)   /* call out of order (e.g. frame before init) */
#define TRJ_ERANGE (-5)   /* coordinates outside the range the kernels accept */
#define TRJ_ELIMIT (-6)   /* size above a documented limit */

/* Position dtype / layout of trj_frame inputs. */
#define TRJ_F32 0
#define TRJ_F64 1
#define TRJ_AOS 0 /* x0 y0 (z0) x1 y1 (z1) ...: the reference's r(ndim,N), stride ndim
                      (common.cuf:52, netcdf.f90:651) */
#define TRJ_SOA 1 /* x[0..N) y[0..N) (z[0..N)) */

/* What a frame computes (RDFcomp and/or SQcalc, trj_analysis.f90:211,217). */
#define TRJ_WANT_RDF 1u
#define TRJ_WANT_SQ 2u

/* q-vector sets. */
#define TRJ_QSET_REF 0      /* the reference set: first-octant lattice vectors with
                               |q| <= qmin (sqfnD, sq.cuf:523-591) plus the three axis
                               directions for shells nqmin+1..nqmax (sqfact,
                               sq.cuf:372-454) */
#define TRJ_QSET_HALFSPACE 1 /* dense grid nx,ny in [-gx/2, gx/2), nz in [0, gz/2),
                                origin excluded; shells nint(|q|/dq) */

/* S(Q) contraction precision (phases always reduced in FP64). */
#define TRJ_PREC_FP64 0   /* FP64 MFMA, FP64 accumulation */
#define TRJ_PREC_FP32 1   /* FP32 MFMA operands and accumulation within a K block */
#define TRJ_PREC_FP16X3 2 /* FP16 hi+lo split, 3 products, FP32 accumulation */
#define TRJ_PREC_BF16X3 3 /* BF16 hi+lo split, 3 products, FP32 accumulation */

/* S(Q) compatibility flags. */
#define TRJ_SQ_NQ_CONSISTENT 1u /* normalise each shell by the number of vectors
                                   actually summed, instead of the reference's host
                                   count (sq.cuf:328-342), which misses the on-sphere
                                   vectors of shell nqmin */
#define TRJ_SQ_CHARGE 2u        /* also sum |sum_j q_j e^{iq.r_j}|^2 (S_qq; sq.cuf:424-432,
                                   563-582 with ex_qc); frames come through trj_frame_q */
#define TRJ_SQ_CENTRES 4u       /* sqfnDcl (sq.cuf:798-839, the S(Q) of cluster centres
                                   passed as the frame's points): each vector's |rho|^2
                                   divided by the frame's point count, octant set only
                                   (no axis vectors); sqcl.dat is sqf(i)/(Nconf nq(i)),
                                   i = 1..nqmin (log.f90:637-642) */
#define TRJ_SQ_NO_AUTOTUNE 8u   /* keep the default K-split and workgroup order of
                                   the FP64 contraction instead of timing candidates
                                   on the first frame (trj_sq_tuning) */

/* RDF flags. */
#define TRJ_RDF_CHARGE 1u /* also accumulate the charge-weighted histogram g_qq
                             (rdf.cuf:466-472, 285-287); frames come through
                             trj_frame_q */
#define TRJ_RDF_FLOOR_BINS 2u /* bin k = int(r/deltar) + 1, i.e. [(k-1)dr, k dr),
                                 as histogram tools bin (freud's RegularAxis);
                                 the reference's rule is nint(r/deltar).  The
                                 printrdf normalisations assume the reference's
                                 rule. */
#define TRJ_RDF_NO_AUTOTUNE 4u /* keep the default launch shape of the pair
                                  kernels instead of timing the candidates on
                                  the first frame (trj_rdf_tuning) */

typedef struct trj_ctx trj_ctx;

/* Library and context lifetime.  trj_create selects the HIP device
 * (util.f90:54-80 gpu_and_header / cudaSetDevice) and creates the context's
 * stream.  trj_destroy frees everything (clean_memory, util.f90:200-237). */
int trj_abi_version(void);
int trj_create(trj_ctx** out, int device);
void trj_destroy(trj_ctx* ctx);
const char* trj_last_error(const trj_ctx* ctx);

/* /INPUT_RDF/ and the /INPUT/ fields the path uses (input.f90:64-69). */
typedef struct {
  int32_t ndim;        /* 2 or 3 */
  int32_t nsp;         /* 1..10 (nspmax, common.cuf:41) */
  double deltar;       /* bin width, > 0 */
  double rcrdf;        /* cutoff; <= 0 means half the shortest side (rdf.cuf:55-58) */
  int32_t periodic[3]; /* /INPUT/ periodic: min image on 2.5*L for false axes
                          (netcdf.f90:603-609) */
  uint32_t flags;      /* TRJ_RDF_*; 0 for the plain histogram */
} trj_rdf_params;

/* RDF_init (rdf.cuf:51-122).  box: first frame's cell lengths; the cutoff is
 * first capped to half the shortest side (netcdf.f90:610-612).  Returns lsmax
 * and nit = nsp(nsp+1)/2.  Zeroes the accumulated counts.  Where the reference
 * refuses a histogram larger than the shared memory (rdf.cuf:61-64), the pair
 * kernels keep it in LDS in tiles of whole pair-type rows, one pass per tile,
 * as many rows per tile as the device's LDS holds (TRJ_ELIMIT only when one
 * row does not fit).  The first frame after init times the candidate launch
 * shapes (rows per tile, wavefronts per home cell) and keeps the fastest. */
int trj_rdf_init(trj_ctx* ctx, const trj_rdf_params* p, const double box[3],
                 int32_t* lsmax_out, int32_t* nit_out);

/* The RDF autotuner's choice (no counterpart in the reference, whose launch
 * shape is fixed, rdf.cuf:208-214): histogram rows per LDS tile, passes per
 * frame, wavefronts per home cell (per x-row part on cells of half the
 * cutoff), the chosen shape's time in ms on the tuning frame (its grid's cell
 * sort and its pair kernel) and the number of shapes timed (0 before the
 * first frame, or with TRJ_RDF_NO_AUTOTUNE).  The cell side it chose is
 * reported by trj_rdf_cell_grid.  Any argument may be NULL. */
int trj_rdf_tuning(trj_ctx* ctx, int32_t* rows_per_pass, int32_t* passes, int32_t* waves_per_cell, double* ms,
                   int32_t* tried);

/* The cell grid of the last RDF frame (no counterpart in the reference, whose
 * rdf_sh loops over all pairs, rdf.cuf:426-487): cells per axis (0 when the
 * frame took the all-pairs kernel) and the cell side as a multiple of the
 * cutoff the autotuner chose (0.5, 1, 1.25 or 1.5; cells of one cutoff
 * without the autotuner).  Half-cutoff cells need every periodic axis to hold
 * at least 7 of them and no sampling points (trj_rdf_sampling).  With the
 * autotuner on, the environment variable TRJ_RDF_CELL_SCALE (0.5, or 1 to 4),
 * read by trj_rdf_init, fixes that multiple instead.  Any argument may be
 * NULL. */
int trj_rdf_cell_grid(trj_ctx* ctx, int32_t cells[3], double* side_over_cutoff);

/* /INPUT_SQ/ (input.f90:70) plus the build's options. */
typedef struct {
  int32_t ndim;
  int32_t nsp;
  double qmax;          /* > 0 */
  double qmin;          /* the reference has no default; <= 0 means qmax */
  double bsc[10];       /* scattering lengths per species (default 1.0, input.f90:106-107) */
  int32_t qset;         /* TRJ_QSET_* */
  int32_t grid[3];      /* TRJ_QSET_HALFSPACE: full grid extents (even) */
  int32_t precision;    /* TRJ_PREC_* */
  uint32_t flags;       /* TRJ_SQ_* */
} trj_sq_params;

/* Geometry of the q-set, as sq_init computes it (sq.cuf:156-200). */
typedef struct {
  double dq;          /* FP32 value of min 2*pi/L (sq.cuf:159) */
  double fk[3];       /* FP32 values of 2*pi/L per axis */
  double qmin_eff;    /* qmin after the lattice adjust (sq.cuf:193) */
  int32_t nqmax;      /* nint(qmax/dq) */
  int32_t nqmin;      /* min lmax after the mod-4/mod-16 adjust */
  int32_t lmax[3];
  int64_t nvec_grid;  /* vectors in the dense (octant or half-space) part */
  int64_t nvec_axis;  /* axis-direction vectors */
  int64_t box_ext[3]; /* extent of the dense box computed per axis */
} trj_sq_info;

/* sq_init (sq.cuf:150-369).  box: first frame's cell lengths; the q-grid is
 * fixed from here on (fk is set once, sq.cuf:189).  Zeroes the accumulators. */
int trj_sq_init(trj_ctx* ctx, const trj_sq_params* p, const double box[3], trj_sq_info* info);

/* Host-only planning (no context, no GPU): the q-set geometry and the shell
 * counts nq(0:nqmax) the given parameters produce (sq.cuf:156-200, 303-343).
 * nq may be NULL; otherwise it has info->nqmax+1 entries. */
int trj_sq_plan(const trj_sq_params* p, const double box[3], trj_sq_info* info, int64_t* nq);

/* Host-only: lsmax for the given parameters (rdf.cuf:55-58 after the cap). */
int trj_rdf_plan(const trj_rdf_params* p, const double box[3], int32_t* lsmax, double* rc_eff);

/* One frame: transfer_cpu_gpu (common.cuf:368-378) + RDFcomp (rdf.cuf:184-353)
 * + SQcalc (sq.cuf:58-126).  pos: host positions (dtype TRJ_F32/F64, layout
 * TRJ_AOS/SOA), itype: host species 1..nsp, box: this frame's cell lengths.
 * Periodic axes are wrapped into [0,L) (floor) on the device.  Asynchronous:
 * host buffers are staged before return. */
int trj_frame(trj_ctx* ctx, const void* pos, int32_t dtype, int32_t layout,
              const int32_t* itype, int64_t n, const double box[3], uint32_t what);

/* trj_frame with per-atom charges (the NetCDF "q" variable, netcdf.f90:
 * 404-408; ex_qc): charge[n] on the host, in the order of pos.  Required when
 * the context was initialised with TRJ_RDF_CHARGE or TRJ_SQ_CHARGE (trj_frame
 * then fails with TRJ_EINVAL).  The pair kernel adds each in-range pair's
 * q_i q_j, rounded once to a per-frame fixed point (2^(30-e), max q^2 < 2^e),
 * as exact int64 sums; the frame's sums are then divided by avq2 = sum(q^2)/N
 * (rdf.cuf:285-287) into FP64 accumulators.  S(Q) adds a charge-weighted
 * channel to the contraction and the axis sums. */
int trj_frame_q(trj_ctx* ctx, const void* pos, int32_t dtype, int32_t layout, const int32_t* itype,
                const double* charge, int64_t n, const double box[3], uint32_t what);

/* trj_frame_device_q: trj_frame_q with device-resident positions, types and
 * charges (charge_dev[n] FP64, in the order of pos_dev); asynchronous like
 * trj_frame_device.  The fixed-point scale and avq2 are reduced on the device
 * in a fixed order (1024 contiguous runs, then a pairwise tree), so avq2 can
 * differ from trj_frame_q's index-order host sum in the last bit; the int64
 * pair sums are the same.  A non-finite charge is reported at the next
 * synchronising call. */
int trj_frame_device_q(trj_ctx* ctx, const void* pos_dev, int32_t dtype, int32_t layout,
                       const int32_t* itype_dev, const double* charge_dev, int64_t n, const double box[3],
                       uint32_t what);

/* Same as trj_frame for inputs already in device memory of this context's
 * device (HBM-resident frames). */
int trj_frame_device(trj_ctx* ctx, const void* pos_dev, int32_t dtype, int32_t layout,
                     const int32_t* itype_dev, int64_t n, const double box[3], uint32_t what);

/* trj_frame_device for an RDF-only frame (what = TRJ_WANT_RDF) that allocates
 * nothing: no hipMalloc, hipFree or allocating library call.  It has no
 * counterpart in the reference; it is for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid it).
 * The frame's per-frame buffers (wrapped positions, species blocks, cell
 * list, per-workgroup histograms) come from a static device array of 64 MiB
 * that each GPU the library runs on holds until the process ends; the
 * context's own buffers (bin edges, totals) come from trj_rdf_init, which
 * does allocate.  Frames whose buffers do not fit (about 0.75M atoms with 3
 * species and lsmax 501) are refused with TRJ_ELIMIT, as are contexts with
 * charges, slices or sampling points (TRJ_EINVAL).  Counts accumulate into
 * the context as with trj_frame_device.  All calls on one GPU share the one
 * array, and it is in use until the frame's work on the context stream has
 * finished: calls that overlap on one GPU (other contexts or threads before
 * the previous frame completed) race. */
int trj_frame_device_noalloc(trj_ctx* ctx, const void* pos_dev, int32_t dtype, int32_t layout,
                             const int32_t* itype_dev, int64_t n, const double box[3], uint32_t what);

/* Wait for the context's queued work; reports deferred device errors
 * (e.g. TRJ_ERANGE from the coordinate check). */
int trj_sync(trj_ctx* ctx);

/* Accumulated pair counts (the reference's histomix after RDFcomp's host
 * reduction, rdf.cuf:301-310, kept as exact int64): counts[ij-1][k-1],
 * nit x lsmax.  Synchronises. */
int trj_rdf_counts(trj_ctx* ctx, int64_t* counts);

/* printrdf normalisation (log.f90:484-518): rows i=1..lsmax-2, each
 * r, g_11, g_12, ..., g_1n, g_22, ... (1+nit doubles).  nconf <= 0 uses the
 * number of RDF frames processed.  Volume and species counts are the last
 * frame's.  Synchronises. */
int trj_rdf_gmix(trj_ctx* ctx, int64_t nconf, double* out);

/* Charge-weighted histogram (TRJ_RDF_CHARGE): acc[k-1], k = 1..lsmax, the
 * reference's gqq (sum over frames of sum q_i q_j / avq2).  Synchronises. */
int trj_rdf_qq(trj_ctx* ctx, double* acc);

/* printrdf's g_qq column (log.f90:501-503): for i = 1..lsmax-2,
 * 2 V gqq(i) / (N^2 dV Nconf) with the last frame's V and N.  nconf <= 0
 * uses the number of RDF frames processed.  Synchronises. */
int trj_rdf_gqq(trj_ctx* ctx, int64_t nconf, double* out);

/* rdf.cuf:313-344: first minimum of the running total g(r) with parabolic
 * interpolation (find_first_minimum_interp, rdf.cuf:755-822). */
int trj_rdf_first_minimum(trj_ctx* ctx, double* rmin);

/* sq_transfer_gpu_cpu (sq.cuf:128-148): raw shell sums sqf[0..nqmax],
 * sqfp[s][0..nqmax] (FP64) and the shell counts nq[0..nqmax] used for
 * normalisation.  Any pointer may be NULL.  Synchronises. */
int trj_sq_shells(trj_ctx* ctx, double* sqf, double* sqfp, int64_t* nq);

/* printSQ normalisation (log.f90:317-401).  cols: nqmax rows of ncols doubles
 * (Q, S_NN, n) or, for two species with all bsc == 1, (Q, S_NN, S_cc, S_11,
 * S_22, S_12, n); mix: nqmax rows of (Q, S_11, ..., S_nn).  nconf <= 0 uses
 * the number of S(Q) frames processed.  Synchronises. */
int trj_sq_columns(trj_ctx* ctx, int64_t nconf, double* cols, int32_t* ncols, double* mix);

/* Charge-weighted shell sums (TRJ_SQ_CHARGE): sqfq[0..nqmax], the
 * reference's sqfq_d after the frames.  Synchronises. */
int trj_sq_qq(trj_ctx* ctx, double* sqfq);

/* printSQ's S_qq column (log.f90:371-376): for i = 1..nqmax,
 * (sqfq(i) / avq2) / (N Nconf nq(i)), avq2 = sum(q^2)/N of the last S(Q)
 * frame (sq.cuf:143-145); 0 where nq(i) = 0 or avq2 = 0.  nconf <= 0 uses the
 * number of S(Q) frames processed.  Synchronises. */
int trj_sq_qq_column(trj_ctx* ctx, int64_t nconf, double* out);

/* rho_s(q) of the last frame for nvec integer vectors kv[3*v..] that belong
 * to the dense q-set: rho[(v*nsp + s)*2 + {0,1}].  Debug/parity access to the
 * contraction output.  Synchronises. */
int trj_sq_rho(trj_ctx* ctx, int64_t nvec, const int32_t* kv, double* rho);

/* The S(Q) autotuner's choice (no counterpart in the reference, whose sqfnD
 * launch shape is fixed, sq.cuf:175-179): K-splits of the last frame's
 * contraction, the largest atoms per split the FP64 autotuner chose (0 for
 * the FP32-accumulating precisions, whose split length their accuracy bound
 * sets, and without the autotuner), whether the contraction's workgroups run
 * XCD-grouped (the tiles of one split on one XCD; always for the FP32 and
 * 16-bit kernels, the autotuner's choice for FP64), the chosen shape's time
 * in ms (contraction and shell pass) on the tuning frame and the number of
 * shapes timed (0 before the first frame or with TRJ_SQ_NO_AUTOTUNE).  Any
 * argument may be NULL. */
int trj_sq_tuning(trj_ctx* ctx, int32_t* splits, int32_t* atoms_per_split, int32_t* xcd_grouped, double* ms,
                  int32_t* tried);

/* rdf_clear / sq_clear (rdf.cuf:124, sq.cuf:774): zero the accumulators and
 * frame counters, keep the plans. */
int trj_reset(trj_ctx* ctx);

/* Timing of the last frame's kernels on the context stream (hipEvents),
 * milliseconds: [0] RDF (cell build + pairs + merge), [1] S(Q) contraction,
 * [2] S(Q) shells + axis vectors, [3] ingest (wrap, sort).  (gptime,
 * common.cuf:328-336, without the host sync per module.) */
int trj_last_timings(trj_ctx* ctx, double* ms4);

/* Kernel time summed over all frames harvested since the last reset, same
 * four columns as trj_last_timings; nframes = frames harvested.  Synchronises. */
int trj_timing_totals(trj_ctx* ctx, double* ms4, int64_t* nframes);
int trj_timing_reset(trj_ctx* ctx);

/* Device memory the context holds, in bytes. */
int64_t trj_device_bytes(const trj_ctx* ctx);

/* Frame pipeline of a context.  TRJ_PIPE_SERIAL (the default): one stream; a
 * frame runs ingest, the RDF, then S(Q), as the reference runs RDFcomp then
 * SQcalc per frame (trj_analysis.f90:173-242).  TRJ_PIPE_OVERLAP: for frames
 * that ask for both, the RDF runs on a second stream of the context,
 * concurrently with the S(Q) contraction; the frame completes when both have.
 * Results are identical in both modes.  Synchronises. */
#define TRJ_PIPE_SERIAL 0
#define TRJ_PIPE_OVERLAP 1
int trj_set_pipeline(trj_ctx* ctx, int mode);

/* Accumulated results of a context -- RDF counts, S(Q) shell sums, frame counts,
 * and the last frame's volume and species counts that trj_rdf_gmix /
 * trj_sq_columns normalise with -- as a byte blob, for checkpoint and resume
 * and for merging runs over different frames (the reference keeps these in
 * module globals, histomix / sqf / sqfp, rdf.cuf:301-310, sq.cuf:128-148,
 * and has no save).  trj_state_bytes gives the size; trj_state_load checks
 * that the blob belongs to the same RDF / S(Q) plans and either replaces the
 * context's accumulators (TRJ_STATE_REPLACE) or adds to them
 * (TRJ_STATE_ADD).  Both synchronise. */
#define TRJ_STATE_REPLACE 0
#define TRJ_STATE_ADD 1
int64_t trj_state_bytes(const trj_ctx* ctx);
int trj_state_save(trj_ctx* ctx, void* buf, int64_t bytes);
int trj_state_load(trj_ctx* ctx, const void* buf, int64_t bytes, int mode);

/* Measurement probe, no counterpart in the reference: the device's LDS-atomic
 * rate (uint32 adds per second, all CUs, 64-lane waves adding to random bins
 * of a 4096-bin histogram), the ceiling the RDF's per-pair histogram update
 * is priced against.  Runs a short kernel on the default stream of `device`. */
int trj_lds_atomic_rate(int device, double* per_s);
/* Measured dense-MFMA rate of the device, for a roofline against what the
 * instruction reaches rather than the nominal peak: two wavefronts per SIMD on
 * every CU each run independent accumulations of one shape (TRJ_PREC_FP64:
 * v_mfma_f64_16x16x4f64; TRJ_PREC_FP16X3 / TRJ_PREC_BF16X3: the 16x16x32
 * FP16 / BF16 MFMA with FP32 accumulation).  *flops_per_s: dense flops per
 * second; *clock_mhz (may be NULL): sustained shader clock over the loop.
 * No counterpart in the reference. */
int trj_mfma_rate(int device, int precision, double* flops_per_s, double* clock_mhz);
/* Measured co-issue of 16-bit MFMAs (16x16x32 FP16, FP32 accumulation) and
 * FP32 VALU FMAs on one SIMD, the bound of the 16-bit split contraction: one
 * 512-thread workgroup per CU, valu_per_mfma (1..4) independent FMAs per MFMA.
 * mode 0: MFMAs only; 1: FMAs only; 2: wavefronts 0-3 MFMAs and 4-7 FMAs (one
 * of each per SIMD); 3: every wavefront interleaves its FMAs after each MFMA.
 * *cycles_per_mfma: the SIMD's shader cycles for the work divided by its MFMA
 * count (about 16 for MFMAs alone; the sum of the two parts when they do not
 * overlap).  No counterpart in the reference. */
int trj_coissue_rate(int device, int valu_per_mfma, int mode, double* cycles_per_mfma);

/* ------------------------------------------------------------------------
 * NetCDF trajectory input (host only): LAMMPS `dump netcdf` files in the
 * classic NetCDF formats (CDF-1, CDF-2 64-bit offset, CDF-5 64-bit data),
 * read without the NetCDF library.  Replaces the nf90_* reads of read_nc_cfg
 * (netcdf.f90:173-550) for coordinates(frame,atom,spatial), type(frame,atom),
 * cell_origin and cell_lengths(frame,cell_spatial).  netCDF-4 (HDF5) files
 * are refused; `nccopy -k cdf5` converts them. */
typedef struct trj_nc trj_nc;
typedef struct {
  int64_t nframes, natoms;
  int32_t version;      /* 1, 2 or 5 */
  int32_t coord_dtype;  /* TRJ_F32 / TRJ_F64, as stored */
  int32_t has_type, has_cell_origin, has_cell_lengths, has_time;
  int32_t has_charge;   /* q(frame, atom) present (the reference's ex_qc) */
} trj_nc_file_info;
/* On failure *out is NULL and errbuf (errlen bytes, may be NULL) holds the reason. */
int trj_nc_open(trj_nc** out, const char* path, char* errbuf, int32_t errlen);
void trj_nc_close(trj_nc* h);
const char* trj_nc_last_error(const trj_nc* h);
int trj_nc_info(const trj_nc* h, trj_nc_file_info* info);
/* Frame `frame` (0-based) as stored: pos AoS [natoms][3], type[natoms] (may
 * be NULL), cell origin and lengths (zeros when the file has none). */
int trj_nc_read(trj_nc* h, int64_t frame, double* pos, int32_t* type, double origin[3], double lengths[3]);
/* time(frame) of frame `frame` (0-based) as stored; the reference reads it
 * into the integer nstep and skips frames whose nstep is 0 (netcdf.f90:432-436,
 * trj_analysis.f90:153-157, 196-198). */
int trj_nc_time(trj_nc* h, int64_t frame, double* t);
/* Frame `frame` as read_nc_cfg / select_ncdfinput hand it on (netcdf.f90:
 * 317-357, 456-531, 582-650): the LAMMPS types sp_types[0..nsp-1] are the
 * species, numbered 1..nsp in order of first appearance in the first frame
 * read (the map is then fixed); atoms of other types are dropped.  periodic
 * is in/out like mod_common's: in, the /INPUT/ periodic flags -- those axes
 * get cell_origin subtracted; out, cleared where the cell length is zero,
 * which gives the axis the length span + 8.  Non-periodic coordinates are
 * translated to sit centred in [0, L) (to start at 4 for span + 8).  Output:
 * pos_soa x[n] y[n] z[n] and itype[n] (capacity natoms each), n, box. */
int trj_nc_frame_ref(trj_nc* h, int64_t frame, int32_t nsp, const int32_t* sp_types, double* pos_soa,
                     int32_t* itype, int64_t* n_out, double box[3], int32_t periodic[3]);
/* keep != 0: trj_nc_frame_ref hands coordinates of non-periodic axes on as
 * read, without the centring (the reference's behaviour; needed by the
 * z-slices, whose centres are absolute z values).  Default 0. */
int trj_nc_keep_nonperiodic(trj_nc* h, int32_t keep);
/* trj_nc_frame_ref plus the charges of the same atoms, charge[n] (capacity
 * natoms; may be NULL): q(frame, atom) as stored, converted to FP64
 * (netcdf.f90:404-408, 657-658).  Fails with TRJ_EINVAL when charge is given
 * and the file has no q variable. */
int trj_nc_frame_ref_q(trj_nc* h, int64_t frame, int32_t nsp, const int32_t* sp_types, double* pos_soa,
                       int32_t* itype, double* charge, int64_t* n_out, double box[3], int32_t periodic[3]);

/* ------------------------------------------------------------------------
 * F(Q,t) and Fs(Q,t) (FqtnD, dynamics.cuf:421-508, with the time origins of
 * rtcorr, dynamics.cuf:157-245; fkt.dat / fskt.dat of print_rtcor, :246-330).
 * Frames ask for it with TRJ_WANT_FQT.  The q-set is the reference's S(Q,w)
 * set (sq.cuf:234-282): vectors (i, j, k) >= 0 of the box given here with
 * qw(l) - dq/2 < |q| < qw(l) + dq/2, FP32.  Buffer b (0-based) starts at frame
 * b*jump + 1, restarts itlimit frames later (its sums then go to the totals
 * and count_org grows); lag indx = frame - origin + 1 <= ndyn is accumulated:
 *   F(q, indx)  += Re(rho_b(q, t0) conj(rho_b(q, t))), rho_b = sum bsc_s e^{iq.r}
 *   Fs(q, indx) += sum_j cos(q.(r_j(t) - r_j(t0)))
 * summed over the vectors of each qw(l), FP64 phases, fixed-order sums. */
#define TRJ_WANT_FQT 4u
typedef struct {
  int32_t ndim, nsp;
  int32_t nqw;          /* target |q| values (/INPUT_SQW/ qw), 1..64 */
  int32_t nbuffer;      /* time origins (/INPUT_DYN/ nbuffer), 1..64 */
  int32_t jump;         /* frames between origins (/INPUT_DYN/ jump) */
  int32_t ndyn;         /* lags kept: Nconf - (nbuffer-1) jump (dyn_init) */
  int32_t itlimit;      /* frames after which an origin restarts (rtcorr: ndyn when tlimit < 0) */
  int32_t pad;
  const double* qw;     /* [nqw] */
  const double* bsc;    /* [nsp] scattering lengths */
} trj_fqt_params;
/* nvec: vectors of the q-set; nw[nqw]: vectors per qw (may be NULL). */
int trj_fqt_init(trj_ctx* ctx, const trj_fqt_params* p, const double box[3], int32_t* nvec, int64_t* nw);
/* fkt, fskt [nqw][ndyn]: the totals over completed origins divided by
 * (N count_org nw(l)) (print_rtcor), raw when no origin has completed; either
 * may be NULL. */
int trj_fqt_results(trj_ctx* ctx, double* fkt, double* fskt, int64_t* count_org);

/* ------------------------------------------------------------------------
 * Number fluctuations (nrandom > 0 of /INPUT_RDF/; s2D3_sh / s2D2_sh,
 * rdf.cuf:489-615, s2n.dat of log.f90:449, 482, 526-533): around nrandom
 * points fixed for the run, each frame adds N_k (atoms with nint(r/dr) <= k
 * and r^2 < side2, the RDF's rule) and N_k^2 to int64 sums.  The reference
 * draws its points with CUDA's device generator (seed blockidx + 12345); this
 * library draws them with SplitMix64 (Steele, Lea and Flood 2014) from `seed`
 * on the host, so the points differ from the reference's but are returned by
 * trj_rdf_sampling_points.  Call after trj_rdf_init (which turns sampling
 * off); nrandom = 0 turns it off; the sums restart at 0.  Not with slices;
 * the state blob refuses a context that samples. */
int trj_rdf_sampling(trj_ctx* ctx, int32_t nrandom, uint64_t seed);
/* The unit points, pts[nrandom][3] in (0, 1]; the frame's point is box * pts. */
int trj_rdf_sampling_points(trj_ctx* ctx, double* pts);
/* gi[k], g2i[k], k = 0..lsmax: sums of N_k and N_k^2 over points and frames. */
int trj_rdf_s2n_sums(trj_ctx* ctx, int64_t* gi, int64_t* g2i);
/* s2n.dat column for r_i, i = 1..lsmax-2: g2i/(Nconf nrandom) - (gi/(Nconf nrandom))^2. */
int trj_rdf_s2n(trj_ctx* ctx, int64_t nconf, double* out);

/* ------------------------------------------------------------------------
 * Confined slices: the reference's twoDstruc_3D analysis (input.f90:166-191:
 * rdf_sq_cl_dyn_sqw_conf_ord(6) with RDF or S(Q); /INPUT_CONF/ zslice, zgrid).
 * Atoms are assigned to the first nearest zslice when |zslice - z| <= zgrid/2
 * (FP32, as the reference; rdf.cuf:692-696); z is the coordinate as given,
 * never wrapped or range-checked: a caller with a periodic z axis wraps it
 * first, as select_ncdfinput does (r - L int(r/L), netcdf.f90:679-682).  Call trj_set_slices before
 * trj_rdf_init / trj_sq_init; nslice = 0 turns slices off.  With slices:
 *  - the RDF is rdfxy (rdf.cuf:672-736): pairs of one slice, xy distance with
 *    the image rule on x and y (both must be periodic), bin int(r/dr) + 1;
 *    each frame's slice histogram is divided by countp_j * countp_l, the
 *    slice's species counts without the last atom of the species-sorted order
 *    (rdf.cuf:284-296, 688); charges add g_qq per slice, divided by
 *    count^2 times the slice's mean q^2;
 *  - S(Q) is sqf3xy / sqf2xy (sq.cuf:596-771): the 2D q-set in the xy plane
 *    (the plan of ndim = 2 on the frame's Lx, Ly), per slice, each vector's
 *    |rho|^2 divided by the slice's (species') atom count; S_qq per slice is
 *    divided at output by the slice's mean q^2 of the last frame
 *    (sq.cuf:135-138);
 *  - the unsliced result calls and the state blob return TRJ_ESTATE.
 * Arrays are slice-major: hxy/counts [nslice][nit][lsmax]. */
#define TRJ_MAX_SLICES 16
int trj_set_slices(trj_ctx* ctx, int32_t nslice, const float* zslice, float zgrid);
/* The last frame's slice counts (rdfxy rule): count[nslice], countp[nslice][nsp]. */
int trj_slice_counts(trj_ctx* ctx, int64_t* count, int64_t* countp);
/* histomix_xy (normalised per frame) and the raw pair counts; either may be NULL. */
int trj_rdf_xy(trj_ctx* ctx, double* hxy, int64_t* counts);
/* gxy_j-l.dat tables (log.f90:494-551): out[pair][i-1][slice] for r_i = i dr,
 * i = 1..lsmax-2, pairs in ind_ij order; nconf <= 0: frames processed. */
int trj_rdf_xy_gmix(trj_ctx* ctx, int64_t nconf, double* out);
/* Charge-weighted slice accumulators gqqxy [nslice][lsmax] (TRJ_RDF_CHARGE). */
int trj_rdf_xy_qq(trj_ctx* ctx, double* acc);
/* gxy_cc.dat table: out[i-1][slice] = V (zgrid/Lz) gqqxy / (dV Nconf). */
int trj_rdf_xy_gcc(trj_ctx* ctx, int64_t nconf, double* out);
/* Slice shell sums: sqf [nslice][nqmax+1], sqfp [nslice][nsp][nqmax+1],
 * sqfq [nslice][nqmax+1] (TRJ_SQ_CHARGE); any may be NULL. */
int trj_sq_xy(trj_ctx* ctx, double* sqf, double* sqfp, double* sqfq);
/* sqxy.dat / sqpxy_j-j.dat / sqxy_qq.dat tables (log.f90:333-378): rows
 * i = 1..nqmax of (Q, one value per slice): sqxy [nqmax][1+nslice], sqpxy
 * [nsp][nqmax][1+nslice], sqqxy [nqmax][1+nslice] (may be NULL). */
int trj_sq_xy_columns(trj_ctx* ctx, int64_t nconf, double* sqxy, double* sqpxy, double* sqqxy);

/* ------------------------------------------------------------------------
 * Batched frames: many independent frames of uneven size and species count,
 * each with its own pair histogram (the frame loop of trj_analysis.f90:173-242
 * with RDFcomp per frame, rdf.cuf:184-353), run as groups of frames in
 * fixed-grid persistent kernels.  With use_graph the per-group launch
 * sequence is captured once into a HIP graph and replayed for every group.
 * Inputs are concatenated in device memory: positions AoS (stride ndim,
 * TRJ_F32 or TRJ_F64) frame after frame, species 1..nsp_f likewise. */
typedef struct trj_batch trj_batch;

typedef struct {
  int64_t n;            /* atoms of this frame */
  int32_t nsp;          /* species of this frame, 1..10 */
  int32_t periodic[3];
  double box[3];
} trj_batch_frame;

typedef struct {
  int32_t ndim;               /* 2 or 3, same for all frames */
  double deltar, rcrdf;       /* /INPUT_RDF/, same for all frames */
  int32_t frames_per_launch;  /* frames per group; 0 = 64 */
  int32_t waves_per_cell;     /* wavefronts per home cell; 0 = from the mean occupancy */
  int32_t workgroups;         /* persistent grid; 0 = 1024 */
  int32_t use_graph;          /* 1: capture the group launch sequence once and replay it */
} trj_batch_params;

int trj_batch_create(trj_batch** out, int device, const trj_batch_params* p);
void trj_batch_destroy(trj_batch* b);
const char* trj_batch_last_error(const trj_batch* b);

/* Host-only: lsmax and nit of each frame and the offset of its histogram
 * (int64, nit x lsmax, histomix_d(lsmax,nit) order) in the output; returns
 * the total number of int64 values through *total. */
int trj_batch_layout(const trj_batch_params* p, int32_t nframes, const trj_batch_frame* fr,
                     int32_t* lsmax, int32_t* nit, int64_t* hist_off, int64_t* total);

/* Run all frames; hist_host receives every frame's counts at hist_off.
 * gpu_ms (may be NULL) receives the device time of the whole run. */
int trj_batch_run(trj_batch* b, int32_t nframes, const trj_batch_frame* fr, const void* pos_dev,
                  int32_t dtype, const int32_t* itype_dev, int64_t* hist_host, double* gpu_ms);

/* Project autotuner: times candidate (frames_per_launch, workgroups,
 * waves_per_cell) settings on the first frames given (at most 128), keeps
 * the fastest in the batch object and reports it through chosen. */
int trj_batch_autotune(trj_batch* b, int32_t nframes, const trj_batch_frame* fr, const void* pos_dev,
                       int32_t dtype, const int32_t* itype_dev, trj_batch_params* chosen);

/* Batched S(Q): the SQcalc of every frame (sq.cuf:58-126, called per frame by
 * trj_analysis.f90:173-242), each frame on the half-space grid of its own box
 * (TRJ_QSET_HALFSPACE: nx, ny in [-g/2, g/2), nz in [0, gz/2); q = 2 pi n / L
 * with the frame's L; shells nint(|q| / dq), dq = min 2 pi / L, 1..nqmax with
 * nqmax = min(nint(qmax / dq), the grid's largest shell)).  Frames are grouped
 * (at most frames_per_launch, 4 GiB of partial sums and 2^30 atoms a group);
 * a group's frames are sorted by species into one set of turn planes and run
 * as channels of one contraction launch (splits of at most 32768 atoms in
 * FP64, 8192 otherwise, as the per-frame path), then one shell pass.  With
 * use_graph the group sequence is captured once and replayed per group.
 * Every axis is wrapped (as trj_frame does for S(Q)-only frames); the frames'
 * periodic flags are not used.  This is the per-frame path's S(Q) in one pass
 * over many frames: a frame's sums equal trj_sq_shells after trj_sq_init on
 * its box and one trj_frame, up to the order of FP64 additions. */
typedef struct {
  int32_t grid[3];    /* half-space grid extents (even, 2..4096) */
  int32_t precision;  /* TRJ_PREC_* */
  double qmax;        /* > 0 */
} trj_batch_sq_params;

/* Host-only: nqmax of each frame and the offsets of its outputs: sqf and nq
 * hold nqmax+1 values at sqf_off[f], sqfp holds nsp x (nqmax+1) at
 * sqfp_off[f]; totals through *total_sqf, *total_sqfp.  Any output may be NULL. */
int trj_batch_sq_layout(const trj_batch_params* p, const trj_batch_sq_params* q, int32_t nframes,
                        const trj_batch_frame* fr, int32_t* nqmax, int64_t* sqf_off, int64_t* sqfp_off,
                        int64_t* total_sqf, int64_t* total_sqfp);

/* Run all frames.  bsc: [nframes][10] scattering lengths (NULL: all 1).
 * sqf: sum over a shell's vectors of |sum_s b_s rho_s|^2; sqfp: of |rho_s|^2;
 * nq: vectors per shell (the raw sums of trj_sq_shells for each frame, at the
 * trj_batch_sq_layout offsets).  Any output may be NULL; gpu_ms (may be NULL)
 * receives the device time of the whole run.  Synchronises. */
int trj_batch_sq_run(trj_batch* b, const trj_batch_sq_params* q, int32_t nframes, const trj_batch_frame* fr,
                     const void* pos_dev, int32_t dtype, const int32_t* itype_dev, const double* bsc,
                     double* sqf, double* sqfp, int64_t* nq, double* gpu_ms);

/* ------------------------------------------------------------------------
 * Request server: continuous batching of independent analysis requests
 * (RDF and/or S(Q) of one frame each).  Each slot is a context with its own
 * stream and its own worker thread.  A request waits in a FIFO until a slot
 * is free and the device memory it may need fits the HBM budget next to what
 * every slot holds: a running slot its context's retained buffers plus its
 * request's footprint, an idle slot its retained buffers.  When the next
 * request does not fit, idle slots free their buffers.  A request whose
 * footprint alone exceeds the budget is refused (TRJ_ELIMIT).  Its host
 * preparation and device work overlap with the other slots'.  Host arrays of
 * a request must stay valid until trj_srv_poll reports it done.
 * There is no single counterpart in the reference, which processes one frame
 * at a time from its driver loop (src/main/trj_analysis.f90:173-242). */
typedef struct trj_srv trj_srv;

typedef struct {
  int64_t n;
  const void* pos;      /* host, AoS stride ndim (rdf.ndim or sq.ndim) */
  int32_t dtype;        /* TRJ_F32 / TRJ_F64 */
  const int32_t* itype; /* host, 1..nsp */
  double box[3];
  uint32_t what;        /* TRJ_WANT_RDF | TRJ_WANT_SQ */
  trj_rdf_params rdf;
  trj_sq_params sq;
} trj_request;

int trj_srv_create(trj_srv** out, int device, int32_t slots, int64_t hbm_budget_bytes);
void trj_srv_destroy(trj_srv* s);
const char* trj_srv_last_error(const trj_srv* s);
/* Host-only: an upper bound on the device bytes a new context allocates for
 * the request (admission). */
int64_t trj_request_footprint(const trj_request* r);
/* Queue a request; *ticket identifies it.  Admits queued requests whose turn has come. */
int trj_srv_submit(trj_srv* s, const trj_request* r, int64_t* ticket);
/* Admit what fits; report completion of ticket without blocking (*done = 0/1). */
int trj_srv_poll(trj_srv* s, int64_t ticket, int32_t* done);
/* Results of a completed request (any pointer may be NULL); releases its slot.
 * counts: nit x lsmax (RDF); sqf, sqfp, nq as trj_sq_shells (S(Q)). */
int trj_srv_result(trj_srv* s, int64_t ticket, int64_t* counts, double* sqf, double* sqfp, int64_t* nq);
/* Requests queued, running, and device bytes the slots hold now (as trj_srv_memory). */
int trj_srv_stats(const trj_srv* s, int32_t* queued, int32_t* running, int64_t* reserved_bytes);
/* Device memory accounting: bytes the slots hold now (running: retained +
 * footprint; idle: retained), the largest such total so far (never above the
 * budget), and the number of requests whose context grew past what was held
 * for it (0 while trj_request_footprint is an upper bound). */
int trj_srv_memory(const trj_srv* s, int64_t* reserved, int64_t* peak, int64_t* misses);

#ifdef __cplusplus
}
#endif
#endif
