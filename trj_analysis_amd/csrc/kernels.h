// kernels.h -- launchers of the trj_hip device kernels (kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "trj_internal.h"

namespace trj {

// Error bits raised by device kernels, read back at synchronisation points.
constexpr uint32_t kFlagRange = 1u;  // non-periodic coordinate out of range / non-finite
constexpr uint32_t kFlagType = 2u;   // species type outside 1..nsp

// Confined slices (twoDstruc_3D): centres and width in FP32, as the reference.
struct SliceSet {
  int n;
  float z[kMaxSlices];
  float zgrid;
};

struct IngestArgs {
  const void* raw;       // positions as given
  int dtype;             // TRJ_F32 / TRJ_F64
  int layout;            // TRJ_AOS (stride ndim) / TRJ_SOA
  const int32_t* itype;  // 1..nsp
  int64_t n;
  int ndim, nsp;
  int periodic[3];
  double box[3];
  double* pos;           // out: wrapped FP64 SoA, 3*n
  uint32_t* flags;
  int free_axis;         // non-periodic axis without the range check (-1: none)
};
void launch_ingest(const IngestArgs& a, hipStream_t s);

// Stable partition of atoms by key: per-block counts (key-major), a scan,
// then a scatter of the S(Q) turns u = x / Lq (hi + lo) and of the unit steps
// e^{2 pi i u} into key order.  The key is the species (key NULL; nkeys =
// nsp) or slice * nsp + species (nkeys = nslice * nsp); atoms without a key
// have no entry.  u holds 12 planes of `stride` doubles.  With charges q (may
// be NULL) the M sorted entries are written again at M + offset, weighted by
// q_i in a 13th plane (1 for the first copy): the charge channel of each
// group of gsize keys; sp_off then has nkeys + 1 + ngroups entries.
void launch_species_count(const int32_t* itype, const int32_t* key, int64_t n, int nkeys, int32_t* blk, int nblk,
                          hipStream_t s);
void launch_exclusive_scan(const int32_t* in, int32_t* out, int64_t count, hipStream_t s);
// The same with the count read from device memory when the kernel starts.
void launch_exclusive_scan_dev(const int32_t* in, int32_t* out, const int32_t* count, hipStream_t s);
void launch_species_offsets(const int32_t* blkoff, const int32_t* blk, int nblk, int nkeys, int gsize, int ngroups,
                            int32_t* sp_off, hipStream_t s);
void launch_species_scatter(const int32_t* itype, const int32_t* key, int64_t n, int nkeys, const int32_t* blkoff,
                            int nblk, const double* pos, const double lq[3], int ndim, const double* q,
                            const int32_t* sp_off, int64_t stride, double* u, hipStream_t s);

// Slices: keys from z (FP32 rule of the reference) and the species-sorted
// last atom; per-slice counts (without that atom; q: sum of q^2 and count
// over all members); 2D cells per slice; the per-frame normalisation.
void launch_slice_keys(const double* pos, int64_t n, const int32_t* itype, int nsp, const SliceSet& ss,
                       int32_t* key, unsigned long long* last, hipStream_t s);
void launch_slice_counts(const int32_t* key, int64_t n, const unsigned long long* last, int nsp, int32_t* count,
                         int32_t* countp, const double* q, double* qsum, int32_t* qcnt, hipStream_t s);

// Cell list.
struct CellGrid {
  int nc[3];
  double inv_h[3];
  int periodic[3];
  // Stencil reach of the RDF pair kernel in cells: 1 (0 taken as 1) for cells
  // of side >= the cutoff; 2 for cells of side >= half the cutoff, walked as
  // home chunks along x-rows of cells (k_rdf_cells).
  int reach;
};
void launch_cell_count(const double* pos, int64_t n, int ndim, const CellGrid& g,
                       int32_t* cellid, int32_t* count, hipStream_t s);
void launch_cell_count_slices(const double* pos, const int32_t* key, int64_t n, int nsp, const CellGrid& g,
                              int nslice, int32_t* cellid, int32_t* count, hipStream_t s);
// q (may be NULL): per-atom charges, scattered into cell order as sq.
void launch_cell_scatter(const double* pos, const int32_t* itype, const int32_t* cellid,
                         int64_t n, const int32_t* start, int32_t* cursor, double* spos,
                         int32_t* stype, const double* q, double* sq, hipStream_t s);

// Cell sort without global atomics (k_cell_hist, k_cell_colscan, the
// exclusive scan, k_cell_place): `blocks` workgroups of 1024 threads take
// `chunk` atoms each and count them per cell in lds_bytes of LDS.  lds is
// false where the counters do not fit the device's LDS or a chunk would pass
// 65535 atoms; the global-atomic count and scatter above serve then.
constexpr int kCellSortBlocks = 256;
struct CellSort {
  bool lds = false;
  int blocks = 0;
  int64_t chunk = 0;
  size_t lds_bytes = 0;
};
CellSort cell_sort_plan(int64_t n, int64_t ncell, size_t lds_max);
// rank: n entries; blkcnt: blocks * ncell; count and start: ncell + 1 (count
// needs no clearing).  key (may be NULL): the slice keys, as for
// launch_cell_count_slices, with nslice slices and ncell = nslice * ncx * ncy
// + 1; else the cells of a frame in ndim dimensions.
void launch_cell_sort(const CellSort& p, const double* pos, const int32_t* key, int nsp, int nslice, int64_t n,
                      int ndim, const CellGrid& g, int64_t ncell, int32_t* rank, int32_t* blkcnt, int32_t* count,
                      int32_t* start, const int32_t* itype, double* spos, int32_t* stype, const double* q,
                      double* sq, hipStream_t s);

struct PairArgs {
  int ndim, nsp, nit, lsmax;
  // Histogram tile of one pass: pair-type rows row0 .. row0+nrows-1 of the
  // nit x lsmax histogram sit in LDS; pairs of other rows are counted by the
  // other passes (the launchers loop over them).  To a launcher, nrows is the
  // rows per pass (0: all nit rows in one pass).
  int row0, nrows;
  // Axis k takes the image shift of the neighbour cell (no per-pair test):
  // periodic with >= 5 cells, where every pair between two cells has the same
  // nearest image and the FP64 rule d - L*nint(d/L) picks exactly that shift;
  // or non-periodic, where neighbour cells never need the image (shift 0).
  int cellshift[3];
  double lmi[3], half[3];
  double side2;
  // FP32 prefilter of the cell-list kernel when every axis takes the image
  // shift, set by launch_rdf_cells from a bound on the FP32 rounding: the
  // FP32 r2 of a pair (from coordinates relative to the home cell's corner)
  // is within pf_delta of its FP64 r2, and side2_pf >= side2 + pf_delta.
  float side2_pf;
  double pf_delta;
  float inv_dr;
  const double* thr;  // lsmax+3 edges
  // Host side of the prefilter, read by launch_rdf_cells only: the edges thr
  // on the host, the bin width and the rule (1: int(r/deltar) + 1, else
  // nint(r/deltar)); NULL thr_host: no entry is decided in FP32.
  const double* thr_host;
  double deltar;
  int floor_bins;
  // Set by launch_rdf_cells: bin k's edges in FP32 registers,
  // t_k = ((k + pf_off)^2) * pf_dr2, moved inwards by pf_margin (checked on
  // the host against thr for every k); the bin estimate
  // int(sqrt(r2) * inv_dr + pf_est).
  float pf_dr2, pf_off, pf_est, pf_margin;
  // Charge-weighted histogram (NULL q: off): q in the pair kernel's atom
  // order, each in-range pair adds rint(q_i q_j qp[0]) to gq[k-1] (int64,
  // two's complement in unsigned 64-bit atomics).  qp = {scale, avq2} of
  // the frame, in device memory.
  const double* q;
  unsigned long long* gq;
  const double* qp;
};
// LDS of the all-pairs kernel for a histogram tile of `rows` pair-type rows.
size_t pair_lds_bytes(int rows, int lsmax, bool qq = false);
// LDS of the cell-list pair kernel: pair_lds_bytes plus its hit rings.
size_t cells_lds_bytes(int rows, int lsmax, bool qq = false, bool pf = true);
// Wavefronts per workgroup of the cell-list pair kernel: eight share one
// histogram tile (LDS per wavefront, and so residency, as for its VGPRs).
constexpr int kCellWaves = 8;
// Cell-list pairs (half stencil); every periodic axis needs nc >= 3.
void launch_rdf_cells(const PairArgs& a, const CellGrid& g, const int32_t* start,
                      const double* spos, const int32_t* stype, int64_t n, int waves_per_cell,
                      int nblocks, uint32_t* partial, hipStream_t s);
// All pairs i<j (small boxes) of n atoms in SoA planes of `stride` (0: n);
// range (device, may be NULL): the atoms [range[0], range[1]) instead.
void launch_rdf_allpairs(const PairArgs& a, const double* pos, const int32_t* itype, int64_t n,
                         int nblocks, uint32_t* partial, hipStream_t s, int64_t stride = 0,
                         const int32_t* range = nullptr);
// F(Q,t) and Fs(Q,t) over time origins (FqtnD).
constexpr int kFqtChunks = 32;   // atom chunks per q-vector (fixed-order partials)
constexpr int kFqtMaxBuffers = 64;
struct FqtBsc {
  double b[kNspMax];
};
struct FqtArgs {
  const int32_t* kv;      // [nvec][3], sorted by |q| bin
  const int32_t* bin_off; // [nqw+1] vector ranges of the bins
  int nvec, nqw, ndyn, nbuffer, ndim, nsp;
  double invl[3];         // 1/L of the q-grid box
  const double* pos;      // current atoms, SoA (stride n)
  const int32_t* itype;
  int64_t n;
  FqtBsc bsc;
  const double* r0;       // [nbuffer][3][n] origin positions
  double* rho0;           // [nbuffer][nvec][2] rho(q, t0)
  int nact;                         // active origins this frame
  int32_t act[kFqtMaxBuffers];      // their buffers
  int32_t act_indx[kFqtMaxBuffers]; // their lag index 1..ndyn
  int32_t start_now[kFqtMaxBuffers];// 1 where the origin is this frame
  double* part;           // [nvec][nact+1][kFqtChunks][2]
  double* vsum;           // [nvec][nact+1][2]
  double* fkt;            // [nbuffer][nqw][ndyn]
  double* fskt;
};
void launch_fqt(const FqtArgs& a, hipStream_t s);
void launch_fqt_collect(int b, int nqw, int ndyn, double* fkt, double* fskt, double* tot, double* tots, hipStream_t s);

// Number fluctuations around fixed random points (s2D3_sh / s2D2_sh).
struct S2nArgs {
  const double* pts;    // unit points, [3][np], in (0, 1]
  int np, ndim, lsmax;
  double box[3], lmi[3], half[3], side2;
  float inv_dr;
  const double* thr;    // the RDF's nint-rule edges
  const double* pos;    // atoms, SoA of stride n (cell-sorted when start is given)
  int64_t n;
  const int32_t* start; // cell starts of the RDF's cell list, or NULL: all atoms
  CellGrid g;
  unsigned long long* gi;   // [lsmax+1] sums of N_k over points
  unsigned long long* g2i;  // [lsmax+1] sums of N_k^2
};
size_t s2n_lds_bytes(int lsmax);
void launch_s2n(const S2nArgs& a, hipStream_t s);

// Slice histograms of the frame into the accumulators, then cleared:
// counts += frame; hxy += frame / (countp_j countp_l); gq (may be NULL):
// gqacc += (gq / qp[0]) / (count^2 qsum/qcnt).
void launch_xy_finish(int nslice, int nsp, int lsmax, int64_t* frame, const int32_t* count, const int32_t* countp,
                      int64_t* counts, double* hxy, unsigned long long* gq, const double* qp, const double* qsum,
                      const int32_t* qcnt, double* gqacc, hipStream_t s);
// counts[b] += sum over slices of partial[slice][b]; with qw (species
// charges) also gq[k] += qw->w[ij] * that sum, b = ij * lsmax + k.
struct QqWeights {
  long long w[kNspMax * (kNspMax + 1) / 2];  // rint(q_i q_j scale) per species pair (ind_ij order)
};
void launch_rdf_merge(const uint32_t* partial, int nslices, int nbins, int64_t* counts,
                      hipStream_t s, const QqWeights* qw = nullptr, int lsmax = 0,
                      unsigned long long* gq = nullptr);
// acc[k] += ((double)gq[k] / qp[0]) / qp[1], then gq[k] = 0 (k < lsmax).
void launch_qq_finish(unsigned long long* gq, int lsmax, const double* qp, double* acc, hipStream_t s);
// qp = {2^(30-e) with max q^2 < 2^e, sum(q^2)/n} from device charges q[n], in
// a fixed order (one workgroup); a non-finite charge sets kFlagRange.
void launch_qq_params(const double* q, int64_t n, double* qp, uint32_t* flags, hipStream_t s);

// S(Q) contraction: rho_s(q) over the active tiles of the dense q-box.
struct ContractArgs {
  const double* u;        // species-sorted turns x/L as hi[3][n], lo[3][n], then
                          // e^{2 pi i u} as cos[3][n], sin[3][n] (n = plane stride)
  const double* w;        // per-entry weights (plane 12 of u) or NULL: all 1
  const int32_t* sp_off;  // nch+1
  const int32_t* tiles;   // 4 ints per tile: kx0, ky0, kz0, pad
  int ntiles, ksplit, nsp, ndim;
  int nch;                // channels: nsp species, + 1 charge channel
  int64_t n;
  double* rho;            // [split][channel][tile][tile vectors][2]: FP64
                          // (k_sq_contract_f64) or FP32 (the FP32 and 16-bit
                          // kernels, whose split sums are FP32)
  // Work items (batched frames; NULL for one frame): ksplit is then the
  // number of items (the grid is ntiles x ksplit), item k = items[k] =
  // {c0, c1, split, splits} takes split `split` of `splits` of each channel
  // c0 .. c1-1 and writes channel c at rho plane item_plane[k] + c - c0
  // ([plane][tile][tile vectors][2]).  Items k >= *nitems exit at once (a
  // fixed grid replayed for groups of fewer items).
  const int4* items;
  const int32_t* item_plane;
  const int32_t* nitems;
  // Workgroup order: 0 in (split, tile) order; 1 XCD-grouped, the tiles of
  // one split on one XCD (its L2 shares the split's turns).  For one frame
  // (items NULL) the FP32 and 16-bit kernels always take the grouped order
  // and the FP64 kernel takes this field (the autotuner's choice); batches
  // keep (split, tile) order.
  int xcd_order;
  // FP64 kernel, one frame: the phase anchors e^{2 pi i n0 u} of its 8-value
  // segments (n0 = anc_base[axis] + 8 m) as a table, row anc_row[axis] + m,
  // entry j at anc[row * n + j] (cos, sin); NULL: the kernel reduces and
  // evaluates them itself.  The values are the same either way.
  const double2* anc;
  int anc_row[3], anc_base[3];
  // 16-bit kernels, one frame: the same anchors in FP32 (e^{2 pi i f} of the
  // FP64-reduced turns f, as the kernel evaluates them), rows as above, then
  // row anc_nrows + axis: the unit step e^{2 pi i u} of each axis in FP32.
  const float2* anc32;
  int anc_nrows;
};
void launch_sq_contract_f64(const ContractArgs& a, hipStream_t s);
// The anchor table of the FP64 contraction (ContractArgs::anc): rows[axis]
// rows of anchors n0 = base[axis] + 8 m for every entry of the turn planes u
// (plane stride n), axes in order x, y, z.
// live_off, live_nch (both NULL, or both device pointers): only entries below
// live_off[*live_nch] are evaluated (a batch group's entries in a buffer of
// stride n).
void launch_sq_anchors(const double* u, int64_t n, const int rows[3], const int base[3], double2* anc,
                       hipStream_t s, const int32_t* live_off = nullptr, const int32_t* live_nch = nullptr);
// The FP32 anchor table of the 16-bit contraction (ContractArgs::anc32).
void launch_sq_anchors32(const double* u, int64_t n, const int rows[3], const int base[3], float2* anc,
                         hipStream_t s, const int32_t* live_off = nullptr, const int32_t* live_nch = nullptr);
// Bytes of one (split, channel, vector) partial of rho: FP64 complex from the
// FP64 contraction, FP32 complex from the others.
inline size_t rho_entry_bytes(int precision) { return precision == TRJ_PREC_FP64 ? 16 : 8; }
// FP32 MFMA, or FP16 / BF16 hi+lo split with three products (precision =
// TRJ_PREC_FP32 / FP16X3 / BF16X3); FP32 accumulation folded into FP64.
void launch_sq_contract_lp(const ContractArgs& a, int precision, hipStream_t s);
// Sum the splits, square, and bin into shells (FP64 atomics).  Channels are
// nslice groups of nsp species, then (nch > nslice * nsp) one charge channel
// per slice; outputs sqf [nslice][nqmax+1], sqfp [nslice][nsp][nqmax+1],
// sqfq [nslice][nqmax+1].  norm (the channel offsets, may be NULL) divides
// each vector's values by the atom counts (slices).
// qsp (may be NULL): charges constant within each species; the charge sums
// are then sum_s q_s rho_s of the species channels (no charge channel).
// rho_f32: the partials are FP32 complex (precision other than FP64).
void launch_sq_shells(const void* rho, bool rho_f32, const int16_t* tile_shell, int ntiles, int tile_vec,
                      int ksplit, int nsp, int nslice, int nch, const double* bsc, int nqmax, const int32_t* norm,
                      const double* qsp, double* sqf, double* sqfp, double* sqfq, hipStream_t s);
// Explicit q-vectors (axis directions): direct FP64 phase sums over atom
// blocks, combined in a fixed order; part needs sq_list_scratch_bytes().
size_t sq_list_scratch_bytes(int nvec, int nch);
void launch_sq_list(const double* u, const double* w, const int32_t* sp_off, int64_t stride, int nsp, int nslice,
                    int nch, int ndim, const int32_t* kv, const int32_t* shell, int nvec, const double* bsc,
                    int nqmax, bool norm, const double* qsp, double* part, double* sqf, double* sqfp, double* sqfq,
                    hipStream_t s);

// LDS-atomic throughput probe (random bins of a 4096-bin uint32 histogram);
// returns the number of atomics the launch performs; *out gets their sum.
int64_t launch_lds_atomics(int nblocks, int iters, uint32_t* out, hipStream_t s);
// Dense-MFMA probe (one shape per precision: FP64 16x16x4, FP16/BF16 16x16x32):
// returns the flops one launch performs; clk[0] += shader-clock cycles and
// clk[1] += 100 MHz real-time ticks summed over wavefronts.
double launch_mfma_rate(int precision, int nblocks, int iters, float* out, unsigned long long* clk, hipStream_t s);
double launch_coissue(int kv, int mode, int ncu, int trips, size_t lds_bytes, float* out, unsigned long long* clk,
                      hipStream_t s);

// ------------------------------------------------------------------ batches
// Frames of one launch group, concatenated.  Device-side descriptors drive
// fixed-grid kernels, so one captured launch sequence serves every group.
struct BatchFrameDesc {
  int64_t atom_off;   // first atom in the caller's concatenated input
  int64_t local_off;  // first atom in the group's arrays
  int64_t task_off;   // first pair task in the group
  int64_t hist_off;   // first int64 of this frame's histogram in the output
  int32_t n, nsp, nit, lsmax;
  int32_t cell_off, ncell, W, pad;
  int32_t nc[3], periodic[3], cellshift[3], pad2;
  double box[3], inv_h[3], lmi[3], half[3];
  double side2;
  float inv_dr;
  float pad3;
};
struct BatchHeader {
  int32_t nframes, ncells;
  int64_t natoms, ntasks;
};
struct BatchDev {
  const BatchHeader* hdr;
  const BatchFrameDesc* fr;
  const void* pos_in;      // caller's positions (AoS, stride ndim), dtype below
  const int32_t* itype_in;
  int dtype, ndim;
  double* pos;             // group-local FP64 SoA, 3 * max atoms
  double* spos;            // cell-sorted
  int32_t* stype;
  int32_t* cellid;
  int32_t* count;          // ncells + 1
  int32_t* start;          // ncells + 1
  int32_t* cursor;         // ncells
  int32_t* blocksum;       // scan scratch, nblocks + 1
  int64_t cap_atoms;       // array capacities (checked on the host)
  const double* thr;       // bin edges, lsmax_max + 3
  int64_t* hist;           // all frames' histograms
  uint32_t* flags;
};
// Launch the whole per-group sequence on stream s with grid nblocks (all
// kernels are grid-stride loops; nothing depends on the group's sizes).
// pair_lds: the largest pair_lds_bytes of the group's frames (the pair kernel
// adds its hit rings, cells_lds_bytes - pair_lds_bytes).
void launch_batch_group(const BatchDev& d, int nblocks, size_t pair_lds, hipStream_t s);

// S(Q) of batched frames.  A group's frames are sorted by species into one
// set of turn planes (frame-major, then species, then input order); frame f's
// species are channels chan0 .. chan0 + nsp - 1 of one contraction launch
// over the box-independent half-space tiles, and its shells come from its own
// box.  Count blocks of 256 atoms never straddle two frames.
struct BatchSqFrame {
  int64_t atom_off;    // first atom in the caller's concatenated input
  int64_t out_off;     // sqf and nq of this frame at out_off (nqmax + 1 values)
  int64_t outp_off;    // sqfp at outp_off ([nsp][nqmax + 1])
  int32_t n, nsp;
  int32_t blk0, nblk;  // count blocks [blk0, blk0 + nblk) of the group
  int32_t moff;        // its [nsp][nblk] block counts in the group's flat count array
  int32_t chan0;       // first channel
  int32_t plane0, ks;  // channel s, split k at rho plane plane0 + k * nsp + s
  int32_t nqmax;
  int32_t vpu;         // tile vectors per unit of the shell kernel
  int32_t unit0, pad;  // its first unit
  double box[3];
  double fk[3], dq;    // 2 pi / L per axis (FP64) and their minimum
  double bsc[kNspMax];
};
struct BatchSqHeader {
  int32_t nframes, nblocks;  // frames and count blocks of the group
  int32_t ncounts, nch;      // flat block counts, channels
  int32_t nitems, nunits;    // contraction work items, shell-kernel units
};
struct BatchSqDev {
  const BatchSqHeader* hdr;
  const BatchSqFrame* fr;
  const void* pos_in;  // caller's positions (AoS, stride ndim), dtype below
  const int32_t* itype_in;
  int dtype, ndim;
  int32_t* blk;        // flat block counts
  int32_t* blkoff;     // their exclusive scan
  int32_t* sp_off;     // channel offsets, nch + 1
  double* u;           // turn planes, 12 x stride
  int64_t stride;
  const int4* items;   // contraction work items and their first rho planes
  const int32_t* item_plane;
  int cap_items;       // grid of the contraction: ntiles x cap_items
  const int32_t* tiles;
  int ntiles, tx, ty, tz;
  int kmax[3];
  double* rho;         // [plane][tile][tile vectors][2], FP64 or (rho_f32) FP32
  int rho_f32;
  double* sqf;         // all frames' outputs (BatchSqFrame offsets)
  double* sqfp;
  unsigned long long* nq;
  uint32_t* flags;
  int nblocks;         // fixed grid of the count, scatter and shell kernels
  int shell_lds;       // doubles of LDS accumulators per workgroup (0: global atomics)
  void* anc;           // anchor table (stride entries per row) or NULL
  int anc_rows[3], anc_base[3];
};
void launch_batch_sq_group(const BatchSqDev& d, int precision, hipStream_t s);

}  // namespace trj
