// trj_ctx.hip -- context, frame pipeline and C ABI of the trj_hip library
// (declared in include/trj_hip.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/trj_hip.h"
#include "kernels.h"
#include "trj_internal.h"

using namespace trj;

namespace {

struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
};

// Per-frame timing events: [0] start, [1] ingest done, [2] S(Q) start (after
// the RDF when serial), [3] contraction done, [4] frame done; with the RDF on
// its own stream, [5] / [6] bracket it there.
struct EventSet {
  hipEvent_t ev[7] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  bool pending = false;
  bool has_rdf = false, has_sq = false, overlapped = false;
};

constexpr int kEventRing = 8;
constexpr size_t kPairLdsLimit = 64 * 1024;

}  // namespace

struct trj_ctx {
  int device = 0;
  int ncu = 0;  // compute units of the device (0: unknown)
  hipStream_t stream = nullptr;
  // TRJ_PIPE_OVERLAP: the RDF of frames that also ask for S(Q) runs on
  // rdf_stream; the frame's last event on `stream` waits for it.
  int pipe = TRJ_PIPE_SERIAL;
  hipStream_t rdf_stream = nullptr;
  std::string err;
  // device buffers
  // (cellid: per-atom cell ids for the global-atomic cell count, or each
  // atom's rank in its workgroup's cell count for the LDS cell sort, whose
  // per-workgroup cell counts are cellblk; cursor: the atomic scatter's
  // per-cell cursors.)
  DevBuf raw, itype, pos, spos, stype, cellid, cellcount, cellstart, cursor, cellblk, partial, counts, thr, blk,
      blkoff, sp_off, u, rho, tiles, tile_shell, sqf, sqfp, axis_kv, axis_shell, listpart, flags, tunebuf, anc;
  // Charges (TRJ_RDF_CHARGE / TRJ_SQ_CHARGE): the frame's charges, their
  // cell-sorted copy, the frame's fixed-point g_qq sums, the g_qq and S_qq
  // accumulators.
  DevBuf qraw, qsort, gqf, gqacc, sqfq;
  // qparm: the frame's {fixed-point scale, avq2}, then avq2 of the last S(Q)
  // frame with charges (the S_qq column divides by it).
  DevBuf qparm;
  bool rdf_charge = false, sq_charge = false;
  bool sq_centres = false;  // TRJ_SQ_CENTRES: per-frame 1/N, octant set only (sqfnDcl)
  int sq_nch = 1;           // channels of the last frame's phase sums (rho layout)
  // Confined slices (twoDstruc_3D; off when slices.n == 0): per-atom keys,
  // the species-sorted last atom, per-slice counts and q^2 sums of the frame
  // (and of the last S(Q) frame), the slice pair histograms (frame, raw
  // totals, normalised accumulators; charges: fixed-point frame sums and
  // their accumulator), and the floor-rule bin edges.
  SliceSet slices{};
  // F(Q,t) (TRJ_WANT_FQT): the S(Q,w) q-set sorted by target |q|, origin
  // bookkeeping of rtcorr, origin positions and rho(q,t0), per-origin sums
  // and their totals over completed origins.
  bool fqt_ready = false;
  int fqt_nqw = 0, fqt_nbuf = 0, fqt_jump = 1, fqt_ndyn = 0, fqt_itlimit = 0, fqt_nvec = 0, fqt_ndim = 3,
      fqt_nsp = 1;
  double fqt_invl[3] = {1.0, 1.0, 1.0};
  FqtBsc fqt_bsc{};
  std::vector<int64_t> fqt_nw;
  std::vector<int> fqt_origin;
  int64_t fqt_frames = 0, fqt_count_org = 0, fqt_n = 0;
  DevBuf fqkv, fqbin, fqr0, fqrho0, fqpart, fqvsum, fqfkt, fqfskt, fqtot, fqtots;
  // Number fluctuations (nrandom > 0): fixed unit points and the int64 sums
  // of N_k and N_k^2 over points and frames.
  int nrand = 0;
  std::vector<double> rpts_host;  // [nrand][3]
  DevBuf rpts, s2gi, s2g2i;
  DevBuf slkey, sllast, slcount, slcountp, slqsum, slqcnt, sqqsum, sqqcnt, xyframe, xycounts, xyacc, xygqf,
      xygqacc, thr_xy;
  // pinned staging for host inputs (double-buffered)
  void* pinned[2] = {nullptr, nullptr};
  size_t pinned_cap[2] = {0, 0};
  hipEvent_t pinned_done[2] = {nullptr, nullptr};
  int pinned_next = 0;
  // RDF
  bool rdf_ready = false;
  RdfPlan rdf;
  int64_t rdf_frames = 0;
  // S(Q)
  bool sq_ready = false;
  SqPlan sq;
  int64_t sq_frames = 0;
  int ksplit = 1;
  std::unordered_map<int64_t, int> tile_index;  // packed (kx0,ky0,kz0) -> active tile
  // last frame
  int64_t last_n = 0;
  double last_volume = 0.0;
  double last_box[3] = {0.0, 0.0, 0.0};
  int last_nsp = 0;
  // timing
  EventSet evs[kEventRing];
  int ev_next = 0;
  double last_ms[4] = {0, 0, 0, 0};
  double tot_ms[4] = {0, 0, 0, 0};
  int64_t tot_frames = 0;
  // Set while trj_frame_device_noalloc runs: the context may then use only
  // buffers that already have their size (ensure() refuses to allocate).
  bool noalloc = false;
  // Set by trj_frame_device_noalloc where the LDS cell sort's per-workgroup
  // counts do not fit its scratch: the frame then sorts its cells with the
  // global-atomic count and scatter.
  bool atomic_cells = false;
  // Largest LDS allocation of one workgroup on this device.
  size_t lds_max = 65536;
  // Launch shape of the RDF pair kernels picked on the first frame after
  // trj_rdf_init by timing the candidates (cell side as a multiple of the
  // cutoff, histogram rows per LDS tile, wavefronts per home cell); kept
  // until the next trj_rdf_init.
  struct {
    bool enabled = true, done = false;
    int rows = 0, W = 0;
    double scale = 1.0;
    double ms = 0.0;
    int tried = 0;
  } rdf_tune;
  // The cell grid of the last RDF frame (cells per axis, side / cutoff).
  int rdf_nc[3] = {0, 0, 0};
  double rdf_cell_scale = 0.0;
  // TRJ_RDF_CELL_SCALE at trj_rdf_init (>= 1; 0: unset): cells of that side
  // over the cutoff, and the autotuner tries no other.
  double rdf_fixed_scale = 0.0;
  // K-split of the FP64 contraction picked on the first S(Q) frame after
  // trj_sq_init by timing candidate split lengths (atoms per split); kept
  // until the next trj_sq_init.  The FP32-accumulating kernels keep the split
  // their accuracy bound sets.
  struct {
    bool enabled = true, done = false;
    int span = 0, order = 0;
    double ms = 0.0;
    int tried = 0;
  } sq_tune;
  hipEvent_t tune_ev[2] = {nullptr, nullptr};
  // FP64 S(Q) frames precompute the contraction's phase anchors as a table;
  // TRJ_SQ_ANCHOR_TABLE=0 at trj_sq_init has the kernel evaluate them itself
  // (the same values: a check of the table).
  bool sq_anchor_table = true;
};

namespace {

// Every device buffer a context owns (trj_destroy frees them, trj_device_bytes
// sums their capacities, the server's release path frees them).
std::vector<DevBuf*> dev_bufs(trj_ctx* c) {
  return {&c->raw,       &c->itype,     &c->pos,      &c->spos,       &c->stype,    &c->cellid,   &c->cellcount,
          &c->cellstart, &c->cursor,    &c->partial,  &c->counts,     &c->thr,      &c->blk,      &c->blkoff,
          &c->sp_off,    &c->u,         &c->rho,      &c->tiles,      &c->tile_shell, &c->sqf,    &c->sqfp,
          &c->axis_kv,   &c->axis_shell, &c->listpart, &c->flags,     &c->qraw,     &c->qsort,    &c->gqf,
          &c->gqacc,     &c->sqfq,      &c->qparm,    &c->slkey,      &c->sllast,   &c->slcount,  &c->slcountp,
          &c->slqsum,    &c->slqcnt,    &c->sqqsum,   &c->sqqcnt,     &c->xyframe,  &c->xycounts, &c->xyacc,
          &c->xygqf,     &c->xygqacc,   &c->thr_xy,   &c->rpts,       &c->s2gi,     &c->s2g2i,    &c->fqkv,
          &c->fqbin,     &c->fqr0,      &c->fqrho0,   &c->fqpart,     &c->fqvsum,   &c->fqfkt,    &c->fqfskt,
          &c->fqtot,     &c->fqtots,    &c->cellblk,  &c->tunebuf,    &c->anc};
}

}  // namespace

// Scratch of the allocation-free entry point (trj_frame_device_noalloc): the
// per-frame buffers of one RDF frame live here instead of in buffers the
// context allocates.  Every device the library runs on holds its own copy
// until the process ends.  In host code the name `scratch` is not a device
// address: the address is taken once per device with hipGetSymbolAddress
// (resolve_scratch, from trj_create with the device current) and only that
// address is passed to kernels and copies.
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
hipError_t g_scratch_err[kMaxDevices] = {};

void resolve_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  g_scratch_err[device] = e;
  if (e == hipSuccess) g_scratch_ptr[device] = p;
}

// This device's scratch address, or NULL with the reason in *why.
void* scratch_address(int device, std::string* why) {
  if (device < 0 || device >= kMaxDevices) {
    *why = "device index above the allocation-free entry point's limit of " + std::to_string(kMaxDevices);
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device])
    *why = std::string("hipGetSymbolAddress of the scratch array failed: ") + hipGetErrorString(g_scratch_err[device]);
  return g_scratch_ptr[device];
}


int fail(trj_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPC(ctx, call)                                                                   \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return fail((ctx), TRJ_EDEVICE, std::string(#call " failed: ") + hipGetErrorString(e_)); \
  } while (0)

int ensure(trj_ctx* c, DevBuf& b, size_t bytes) {
  if (bytes == 0) bytes = 16;
  if (b.cap >= bytes) return TRJ_OK;
  if (c->noalloc)
    return fail(c, TRJ_ELIMIT, "allocation-free frame: a buffer of " + std::to_string(bytes) +
                                   " bytes is not provided (the library refuses to allocate here)");
  if (b.p) {
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(c, TRJ_EDEVICE, hipGetErrorString(e));
    hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
  }
  size_t want = bytes + bytes / 8;
  hipError_t e = hipMalloc(&b.p, want);
  if (e != hipSuccess) {
    b.p = nullptr;
    return fail(c, TRJ_ENOMEM, "device allocation of " + std::to_string(want) + " bytes failed");
  }
  b.cap = want;
  return TRJ_OK;
}

template <class T>
T* P(DevBuf& b) {
  return static_cast<T*>(b.p);
}

void harvest(trj_ctx* c, EventSet& s) {
  if (!s.pending) return;
  hipEventSynchronize(s.ev[4]);
  float ms[4] = {0, 0, 0, 0};
  hipEventElapsedTime(&ms[3], s.ev[0], s.ev[1]);  // ingest
  if (s.overlapped)
    hipEventElapsedTime(&ms[0], s.ev[5], s.ev[6]);  // RDF on its stream
  else
    hipEventElapsedTime(&ms[0], s.ev[1], s.ev[2]);  // RDF
  hipEventElapsedTime(&ms[1], s.ev[2], s.ev[3]);  // S(Q) contraction
  hipEventElapsedTime(&ms[2], s.ev[3], s.ev[4]);  // S(Q) shells + axis vectors
  for (int k = 0; k < 4; ++k) {
    c->last_ms[k] = ms[k];
    c->tot_ms[k] += ms[k];
  }
  c->tot_frames += 1;
  s.pending = false;
}

void harvest_all(trj_ctx* c) {
  // oldest first so that last_ms ends with the newest frame
  for (int k = 0; k < kEventRing; ++k) harvest(c, c->evs[(c->ev_next + k) % kEventRing]);
}

int check_flags(trj_ctx* c) {
  uint32_t f = 0;
  if (!c->flags.p) return TRJ_OK;
  HIPC(c, hipMemcpyAsync(&f, c->flags.p, sizeof(f), hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  if (f) {
    HIPC(c, hipMemsetAsync(c->flags.p, 0, sizeof(uint32_t), c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    if (f & kFlagType) return fail(c, TRJ_EINVAL, "a frame had species types outside 1..nsp");
    return fail(c, TRJ_ERANGE,
                "a frame had non-finite coordinates, or coordinates of a non-periodic axis "
                "outside [-0.25 L, 1.25 L)");
  }
  return TRJ_OK;
}

int sync_all(trj_ctx* c) {
  HIPC(c, hipStreamSynchronize(c->stream));
  harvest_all(c);
  return check_flags(c);
}

int64_t pack_tile(int x, int y, int z) {
  return ((static_cast<int64_t>(x) + 4096) << 40) | ((static_cast<int64_t>(y) + 4096) << 20) |
         (static_cast<int64_t>(z) + 4096);
}

int validate_box(trj_ctx* c, int ndim, const double box[3]) {
  if (!box) return fail(c, TRJ_EINVAL, "box is NULL");
  for (int k = 0; k < ndim; ++k)
    if (!(box[k] > 0.0) || !std::isfinite(box[k])) return fail(c, TRJ_EINVAL, "box lengths must be positive");
  return TRJ_OK;
}

// RDF of a frame with slices (rdfxy, rdf.cuf:672-736): the slice members in
// 2D cells of each slice (or, when a slice's box has fewer than 3 cells per
// axis, all pairs of the slice), one pair launch per slice into that slice's
// frame histogram, then the per-frame normalisation (k_xy_finish).
int rdf_slices(trj_ctx* c, PairArgs pa, const FrameGeom& g, const double box[3], int64_t n,
               const int32_t* itype_dev, const double* q_dev, hipStream_t rs) {
  const RdfPlan& R = c->rdf;
  const int nsl = c->slices.n, nsp = R.nsp;
  const int nbins = R.nit * R.lsmax;
  int r;
  pa.ndim = 2;  // x and y: the reference's image rule on both (rdf.cuf:712-715)
  pa.row0 = 0;
  pa.nrows = R.rows;
  pa.thr = P<double>(c->thr_xy);
  pa.thr_host = R.thr_xy.size() == R.thr.size() ? R.thr_xy.data() : nullptr;
  pa.floor_bins = 1;
  pa.cellshift[2] = 0;
  const double rc = std::sqrt(g.side2);
  CellGrid cg{};
  bool cells_ok = true;
  for (int k = 0; k < 2; ++k) {
    int nc = static_cast<int>(std::floor(box[k] / (rc * (1.0 + 1e-6))));
    nc = std::max(1, std::min(nc, 2048));
    cg.nc[k] = nc;
    cg.inv_h[k] = nc / box[k];
    cg.periodic[k] = 1;
    if (nc < 3) cells_ok = false;
    pa.cellshift[k] = nc >= 5 ? 1 : 0;
  }
  if (!cells_ok)
    for (int k = 0; k < 2; ++k) {
      cg.nc[k] = 1;
      cg.inv_h[k] = 0.0;
    }
  cg.nc[2] = 1;
  cg.inv_h[2] = 0.0;
  cg.periodic[2] = 0;
  const int64_t ncell2 = static_cast<int64_t>(cg.nc[0]) * cg.nc[1];
  const int64_t ncell = nsl * ncell2 + 1;  // + the cell of atoms outside every slice
  if ((r = ensure(c, c->cellid, sizeof(int32_t) * n))) return r;
  if ((r = ensure(c, c->cellcount, sizeof(int32_t) * (ncell + 1)))) return r;
  if ((r = ensure(c, c->cellstart, sizeof(int32_t) * (ncell + 1)))) return r;
  if ((r = ensure(c, c->spos, sizeof(double) * 3 * n))) return r;
  if ((r = ensure(c, c->stype, sizeof(int32_t) * n))) return r;
  if (q_dev && (r = ensure(c, c->qsort, sizeof(double) * n))) return r;
  const CellSort cs = cell_sort_plan(n, ncell, c->lds_max);
  if (cs.lds) {
    if ((r = ensure(c, c->cellblk, sizeof(int32_t) * cs.blocks * static_cast<size_t>(ncell)))) return r;
    launch_cell_sort(cs, P<double>(c->pos), P<int32_t>(c->slkey), nsp, nsl, n, 2, cg, ncell, P<int32_t>(c->cellid),
                     P<int32_t>(c->cellblk), P<int32_t>(c->cellcount), P<int32_t>(c->cellstart), itype_dev,
                     P<double>(c->spos), P<int32_t>(c->stype), q_dev, q_dev ? P<double>(c->qsort) : nullptr, rs);
  } else {
    if ((r = ensure(c, c->cursor, sizeof(int32_t) * ncell))) return r;
    HIPC(c, hipMemsetAsync(c->cellcount.p, 0, sizeof(int32_t) * (ncell + 1), rs));
    HIPC(c, hipMemsetAsync(c->cursor.p, 0, sizeof(int32_t) * ncell, rs));
    launch_cell_count_slices(P<double>(c->pos), P<int32_t>(c->slkey), n, nsp, cg, nsl, P<int32_t>(c->cellid),
                             P<int32_t>(c->cellcount), rs);
    launch_exclusive_scan(P<int32_t>(c->cellcount), P<int32_t>(c->cellstart), ncell + 1, rs);
    launch_cell_scatter(P<double>(c->pos), itype_dev, P<int32_t>(c->cellid), n, P<int32_t>(c->cellstart),
                        P<int32_t>(c->cursor), P<double>(c->spos), P<int32_t>(c->stype), q_dev,
                        q_dev ? P<double>(c->qsort) : nullptr, rs);
  }
  if (q_dev) {
    pa.q = P<double>(c->qsort);
    pa.qp = P<double>(c->qparm);
  }
  int nblocks, W = 1;
  if (cells_ok) {
    const double mean = static_cast<double>(n) / static_cast<double>(nsl * ncell2);
    W = static_cast<int>(std::lround(mean / 64.0));
    W = std::max(1, std::min(W, 8));
    nblocks = static_cast<int>(std::min<int64_t>((ncell2 * W + kCellWaves - 1) / kCellWaves, 2048));
  } else {
    const int64_t nb = (n + 63) / 64;
    nblocks = static_cast<int>(std::min<int64_t>((nb * nb + 3) / 4, 1024));
  }
  if ((r = ensure(c, c->partial, sizeof(uint32_t) * nbins * static_cast<size_t>(nblocks)))) return r;
  for (int sl = 0; sl < nsl; ++sl) {
    if (q_dev) pa.gq = P<unsigned long long>(c->xygqf) + static_cast<int64_t>(sl) * R.lsmax;
    if (cells_ok)
      launch_rdf_cells(pa, cg, P<int32_t>(c->cellstart) + sl * ncell2, P<double>(c->spos), P<int32_t>(c->stype), n,
                       W, nblocks, P<uint32_t>(c->partial), rs);
    else
      launch_rdf_allpairs(pa, P<double>(c->spos), P<int32_t>(c->stype), n, nblocks, P<uint32_t>(c->partial), rs, n,
                          P<int32_t>(c->cellstart) + sl);
    launch_rdf_merge(P<uint32_t>(c->partial), nblocks, nbins, P<int64_t>(c->xyframe) + static_cast<int64_t>(sl) * nbins,
                     rs);
  }
  launch_xy_finish(nsl, nsp, R.lsmax, P<int64_t>(c->xyframe), P<int32_t>(c->slcount), P<int32_t>(c->slcountp),
                   P<int64_t>(c->xycounts), P<double>(c->xyacc), q_dev ? P<unsigned long long>(c->xygqf) : nullptr,
                   P<double>(c->qparm), P<double>(c->slqsum), P<int32_t>(c->slqcnt), P<double>(c->xygqacc), rs);
  return TRJ_OK;
}

// One frame of F(Q,t) (rtcorr's origin bookkeeping, dynamics.cuf:157-245, and
// FqtnD): origins of buffer b at frames (b-1) jump + 1, restarted after
// itlimit frames; lags beyond ndyn are not accumulated.  Everything the host
// decides comes from its own counters: no synchronisation.
int fqt_step(trj_ctx* c, int64_t n, const int32_t* itype_dev) {
  int r;
  const int iconf = static_cast<int>(++c->fqt_frames);
  const int nb = c->fqt_nbuf;
  if (iconf == 1) {
    c->fqt_n = n;
    for (int b = 0; b < nb; ++b) c->fqt_origin[b] = b * c->fqt_jump + 1;
    if ((r = ensure(c, c->fqr0, sizeof(double) * 3 * n * nb))) return r;
  }
  FqtArgs a{};
  a.kv = P<int32_t>(c->fqkv);
  a.bin_off = P<int32_t>(c->fqbin);
  a.nvec = c->fqt_nvec;
  a.nqw = c->fqt_nqw;
  a.ndyn = c->fqt_ndyn;
  a.nbuffer = nb;
  a.ndim = c->fqt_ndim;
  a.nsp = c->fqt_nsp;
  for (int k = 0; k < 3; ++k) a.invl[k] = c->fqt_invl[k];
  a.pos = P<double>(c->pos);
  a.itype = itype_dev;
  a.n = n;
  a.bsc = c->fqt_bsc;
  a.r0 = P<double>(c->fqr0);
  a.rho0 = P<double>(c->fqrho0);
  for (int b = 0; b < nb; ++b) {
    if (iconf == c->fqt_origin[b]) {  // r0 = r at the origin frame (dynamics.cuf:202-206)
      HIPC(c, hipMemcpyAsync(P<double>(c->fqr0) + static_cast<int64_t>(b) * 3 * n, c->pos.p, sizeof(double) * 3 * n,
                             hipMemcpyDeviceToDevice, c->stream));
      a.start_now[b] = 1;
    }
    const int ix = iconf - c->fqt_origin[b] + 1;
    if (iconf >= c->fqt_origin[b] && ix <= c->fqt_ndyn) {
      a.act[a.nact] = b;
      a.act_indx[a.nact] = ix;
      ++a.nact;
    }
  }
  const size_t np = static_cast<size_t>(c->fqt_nvec) * (a.nact + 1);
  if ((r = ensure(c, c->fqpart, sizeof(double) * 2 * kFqtChunks * std::max<size_t>(np, 1)))) return r;
  if ((r = ensure(c, c->fqvsum, sizeof(double) * 2 * std::max<size_t>(np, 1)))) return r;
  a.part = P<double>(c->fqpart);
  a.vsum = P<double>(c->fqvsum);
  a.fkt = P<double>(c->fqfkt);
  a.fskt = P<double>(c->fqfskt);
  launch_fqt(a, c->stream);
  // Origins that reached itlimit restart at the next frame; their sums go to
  // the totals (dynamics.cuf:220-238).
  for (int b = 0; b < nb; ++b)
    if (iconf - c->fqt_origin[b] == c->fqt_itlimit) {
      c->fqt_origin[b] = iconf + 1;
      launch_fqt_collect(b, c->fqt_nqw, c->fqt_ndyn, P<double>(c->fqfkt), P<double>(c->fqfskt), P<double>(c->fqtot),
                         P<double>(c->fqtots), c->stream);
      ++c->fqt_count_org;
    }
  HIPC(c, hipGetLastError());
  return TRJ_OK;
}

// Cell grid and launch shape of the RDF pair kernels for one frame: cells of
// side >= the cutoff on every axis (at most 160 per axis in 3D, 2048 in 2D);
// a periodic axis with fewer than 3 cells sends the frame to the all-pairs
// kernel.  The image shift replaces the per-pair minimum image on periodic
// axes with at least 5 cells.  W wavefronts share a home cell.
// Smallest relative gain for which an autotuner leaves its default for a
// candidate that costs more memory traffic or work per pair.
constexpr double kTuneGain = 0.02;

struct RdfCells {
  CellGrid cg{};
  bool cells_ok = true;
  int64_t ncell = 1;
  int cellshift[3] = {0, 0, 0};
  int W = 1;
  int nblocks = 1;
  double scale = 1.0;  // cell side >= scale * cutoff
  int64_t units = 1;   // what W wavefronts share: cells, or with reach 2 x-rows of cells
};

RdfCells rdf_cell_plan(int ndim, const double box[3], const int periodic[3], double side2, int64_t n,
                       int W_override = 0, double scale = 1.0) {
  RdfCells p;
  // cells of side >= scale * cutoff: scale >= 1, or 0.5 (reach 2: cells of
  // half the cutoff, walked in x-rows; only where every axis takes the image
  // shift with at least 7 cells on a periodic axis, else the plan is not ok)
  p.scale = scale == 0.5 ? 0.5 : std::max(1.0, scale);
  p.cg.reach = p.scale == 0.5 ? 2 : 1;
  const double rc = std::sqrt(side2) * p.scale;
  for (int k = 0; k < 3; ++k) {
    if (k < ndim) {
      int nc = static_cast<int>(std::floor(box[k] / (rc * (1.0 + 1e-6))));
      nc = std::max(1, std::min(nc, ndim == 3 ? 160 : 2048));
      p.cg.nc[k] = nc;
      p.cg.inv_h[k] = nc / box[k];
      p.cg.periodic[k] = periodic[k];
      if (periodic[k] && nc < 3) p.cells_ok = false;
      // A periodic axis takes the neighbour cell's image shift with at least 5
      // cells.  A non-periodic axis has shift 0 (its minimum image on 2.5 L
      // only changes pairs more than 5L/4 apart, beyond the cutoff <= L/2
      // either way, and neighbour cells are never that far apart).
      p.cellshift[k] = periodic[k] ? (nc >= 5 ? 1 : 0) : 1;
    } else {
      p.cg.nc[k] = 1;
      p.cg.inv_h[k] = 0.0;
      p.cg.periodic[k] = 0;
    }
    p.ncell *= p.cg.nc[k];
  }
  if (p.cg.reach == 2) {
    for (int k = 0; k < ndim; ++k)
      if (periodic[k] && p.cg.nc[k] < 7) p.cells_ok = false;
    if (p.cg.nc[0] < 7) p.cells_ok = false;  // a row's window of 7 cells holds no cell twice
  }
  if (p.cells_ok) {
    // home chunks: 16 atoms with the image shift on every axis (the
    // prefiltered kernel), else 64; W wavefronts share a cell's chunks, or
    // with reach 2 split an x-row of cells into W parts (about 4 chunks each)
    bool all_shift = true;
    for (int k = 0; k < ndim; ++k) all_shift = all_shift && p.cellshift[k];
    p.units = p.cg.reach == 2 ? p.ncell / p.cg.nc[0] : p.ncell;
    const double chunk = p.cg.reach == 2 ? 64.0 : (all_shift ? 16.0 : 64.0);
    const double mean = static_cast<double>(n) / static_cast<double>(p.units);
    p.W = std::max(1, std::min(static_cast<int>(std::lround(mean / chunk)), 16));
    if (W_override > 0) p.W = W_override;
    p.nblocks = static_cast<int>(std::min<int64_t>((p.units * p.W + kCellWaves - 1) / kCellWaves, 2048));
  } else {
    const int64_t nb = (n + 63) / 64;
    p.nblocks = static_cast<int>(std::min<int64_t>((nb * nb + 3) / 4, 1024));
  }
  return p;
}

// Largest number of pair-type rows (<= nit) whose histogram tile fits
// `budget` bytes of LDS in the cell-list kernel (0: not even one row).
int rows_fitting(int nit, int lsmax, bool qq, size_t budget) {
  int r = nit;
  while (r > 0 && cells_lds_bytes(r, lsmax, qq) > budget) --r;
  return r;
}

// The RDF autotuner: on the first frame after trj_rdf_init, times the pair
// kernel (all its passes) for each candidate launch shape and keeps the
// fastest in c->rdf_tune.  Candidates on the frame's cell list (cells of side
// >= the cutoff, already sorted by the caller): histogram tiles that fit the
// device's LDS at 1, 2, 3 and 4 workgroups per CU (fewer rows per tile means
// more passes but more resident workgroups), and the default waves per home
// cell, one fewer and one more, half and twice.  Then, with the best tile,
// cells of side >= 1.25 and 1.5 cutoffs (sort(p) sorts the frame into plan
// p's cells): fewer, fuller cells, more candidates per in-range pair; they
// win where cells of one cutoff hold few atoms, and are kept only when at
// least kTuneGain faster than the best shape on cells of one cutoff.  Every
// candidate counts the same pairs, so the choice changes speed only.  The
// trials leave the cells of the last grid tried; the caller sorts again into
// the chosen one.
int tune_pairs(trj_ctx* c, const PairArgs& pa0, const RdfCells& base, int64_t n, hipStream_t rs,
               const std::function<RdfCells(double)>& plan, const std::function<int(const RdfCells&)>& sort) {
  auto& tu = c->rdf_tune;
  const RdfPlan& R = c->rdf;
  std::vector<int> rows;
  for (int per_cu = 1; per_cu <= 4; ++per_cu) {
    const int r = rows_fitting(R.nit, R.lsmax, false, c->lds_max / per_cu);
    if (r >= 1 && std::find(rows.begin(), rows.end(), r) == rows.end()) rows.push_back(r);
  }
  if (rows.empty()) rows.push_back(R.rows);
  auto waves = [](int W0) {
    std::vector<int> ws = {W0};
    for (int w : {W0 / 2, W0 - 1, W0 + 1, 2 * W0})
      if (w >= 1 && w <= 16 && std::find(ws.begin(), ws.end(), w) == ws.end()) ws.push_back(w);
    return ws;
  };
  tu.rows = R.rows;
  tu.W = base.W;
  tu.scale = base.scale;
  tu.tried = 0;
  const size_t nbins = static_cast<size_t>(R.nit) * R.lsmax;
  int r;
  if ((r = ensure(c, c->partial, sizeof(uint32_t) * nbins * 2048))) return r;
  uint32_t* part = P<uint32_t>(c->partial);
  PairArgs pa = pa0;
  double best = 1e300;
  // A grid other than the cells of one cutoff is kept only when its sort and
  // pair kernel together are at least kTuneGain faster (timing noise must not
  // trade a known grid for another).
  double need = 1.0, grid_ms = 0.0;
  auto trial = [&](const RdfCells& p, int rw, int W, bool timed) -> int {
    // the sorted arrays as they are now: sort() may have grown (moved) them
    const int32_t* start = P<int32_t>(c->cellstart);
    const double* spos = P<double>(c->spos);
    const int32_t* stype = P<int32_t>(c->stype);
    const int nb = static_cast<int>(std::min<int64_t>((p.units * W + kCellWaves - 1) / kCellWaves, 2048));
    pa.nrows = rw;
    for (int k = 0; k < 3; ++k) pa.cellshift[k] = p.cellshift[k];
    if (!timed) {
      launch_rdf_cells(pa, p.cg, start, spos, stype, n, W, nb, part, rs);
      HIPC(c, hipGetLastError());
      return TRJ_OK;
    }
    // the faster of two runs (first-frame timings are noisy at the few-% level)
    float ms = 0.f;
    for (int rep = 0; rep < 2; ++rep) {
      HIPC(c, hipEventRecord(c->tune_ev[0], rs));
      launch_rdf_cells(pa, p.cg, start, spos, stype, n, W, nb, part, rs);
      HIPC(c, hipEventRecord(c->tune_ev[1], rs));
      HIPC(c, hipEventSynchronize(c->tune_ev[1]));
      float m = 0.f;
      HIPC(c, hipEventElapsedTime(&m, c->tune_ev[0], c->tune_ev[1]));
      ms = rep == 0 ? m : std::min(ms, m);
    }
    ++tu.tried;
    if ((ms + grid_ms) * need < best) {
      best = (ms + grid_ms) * need;
      tu.rows = rw;
      tu.W = W;
      tu.scale = p.scale;
    }
    return TRJ_OK;
  };
  // the cell sort of a grid, timed (a grid's cost is its sort plus its pair kernel)
  auto timed_sort = [&](const RdfCells& p, double* ms_out) -> int {
    int rr;
    float ms = 0.f;
    for (int rep = 0; rep < 2; ++rep) {
      HIPC(c, hipEventRecord(c->tune_ev[0], rs));
      if ((rr = sort(p))) return rr;
      HIPC(c, hipEventRecord(c->tune_ev[1], rs));
      HIPC(c, hipEventSynchronize(c->tune_ev[1]));
      float m = 0.f;
      HIPC(c, hipEventElapsedTime(&m, c->tune_ev[0], c->tune_ev[1]));
      ms = rep == 0 ? m : std::min(ms, m);
    }
    *ms_out = ms;
    return TRJ_OK;
  };
  double sort_ms = 0.0;
  if ((r = timed_sort(base, &sort_ms))) return r;
  // one untimed run (code object and caches warm)
  if ((r = trial(base, rows[0], base.W, false))) return r;
  for (int rw : rows)
    for (int W : waves(base.W))
      if ((r = trial(base, rw, W, true))) return r;
  const int rbest = tu.rows;
  // other grids: compared by sort + pair time
  best += sort_ms;
  need = 1.0 + kTuneGain;
  for (double f : {0.5, 1.25, 1.5}) {
    if (c->rdf_fixed_scale > 0.0) break;
    if (f < 1.0 && c->nrand > 0) continue;  // the sampling spheres read cells of one cutoff
    const RdfCells p = plan(f);
    if (!p.cells_ok || p.ncell == base.ncell) continue;
    if ((r = timed_sort(p, &sort_ms))) return r;
    grid_ms = sort_ms;
    for (int W : waves(p.W))
      if ((r = trial(p, rbest, W, true))) return r;
  }
  grid_ms = 0.0;
  tu.ms = tu.scale == base.scale ? best : best / need;  // with the grid's sort
  tu.done = true;
  return TRJ_OK;
}

// The S(Q) autotuner (FP64 contraction): on the first S(Q) frame after
// trj_sq_init, times the contraction of this frame's turns and the shell pass
// that reads its partial sums (run(ks, order) launches both for ks splits in
// that workgroup order, the shells into scratch sums) and keeps the fastest
// (span, order) in c->sq_tune.  Candidates: K-splits of at most 32768 (the
// default) and 16384 atoms, each in (split, tile) order and XCD-grouped
// (contract_work_index).  Short splits keep the turns the running workgroups
// share in L2 but write and read more partial sums, so a 16384 span is kept
// only when at least kTuneGain faster.  Longer splits are not tried: a
// split's compensated FP64 sums lose accuracy with its length on coherent
// (Bragg-like) vectors.  The choice changes the FP64 summation order only.
int tune_contraction(trj_ctx* c, int ntiles, int64_t n, int nch, const std::function<void(int, int)>& run) {
  auto& st = c->sq_tune;
  const SqPlan& S = c->sq;
  struct Cand {
    int span, ks, order;
  };
  std::vector<Cand> cand;
  std::vector<int> splits;
  for (int span : {32768, 16384}) {
    const int ks = contraction_splits(ntiles, n, S.p.precision, nch, span);
    if (std::find(splits.begin(), splits.end(), ks) != splits.end()) continue;
    splits.push_back(ks);
    for (int order : {0, 1}) cand.push_back({span, ks, order});
  }
  st.span = cand[0].span;
  st.order = 0;
  st.tried = 0;
  st.done = true;
  const int kmax = *std::max_element(splits.begin(), splits.end());
  int r;
  if ((r = ensure(c, c->rho, rho_entry_bytes(S.p.precision) * static_cast<size_t>(kmax) * nch * ntiles *
                                 S.tk.vec())))
    return r;
  run(cand[0].ks, 0);  // untimed: code objects and caches warm
  double best = 1e300, best_ms = 0.0;
  for (const Cand& k : cand) {
    HIPC(c, hipEventRecord(c->tune_ev[0], c->stream));
    run(k.ks, k.order);
    HIPC(c, hipEventRecord(c->tune_ev[1], c->stream));
    HIPC(c, hipEventSynchronize(c->tune_ev[1]));
    float ms = 0.f;
    HIPC(c, hipEventElapsedTime(&ms, c->tune_ev[0], c->tune_ev[1]));
    ++st.tried;
    const double score = k.span == cand[0].span ? ms : ms * (1.0 + kTuneGain);
    if (score < best) {
      best = score;
      best_ms = ms;
      st.span = k.span;
      st.order = k.order;
    }
  }
  HIPC(c, hipGetLastError());
  st.ms = best_ms;
  return TRJ_OK;
}

// The dimensionality and species count a frame with these `what` bits is read
// with: the RDF's when it is asked for, else the S(Q)'s, else F(Q,t)'s.  Every
// module the frame asks for must be initialised with the same ndim and nsp.
// Host and device entry points both call this before touching any input, so
// the bytes copied out of the caller's r(ndim,N) array and the bytes the
// ingest kernel reads are sized by the same rule.
int frame_shape(trj_ctx* c, uint32_t what, int* ndim_out, int* nsp_out) {
  const bool do_rdf = (what & TRJ_WANT_RDF) != 0;
  const bool do_sq = (what & TRJ_WANT_SQ) != 0;
  const bool do_fqt = (what & TRJ_WANT_FQT) != 0;
  if (do_rdf && !c->rdf_ready) return fail(c, TRJ_ESTATE, "frame asks for RDF before trj_rdf_init");
  if (do_sq && !c->sq_ready) return fail(c, TRJ_ESTATE, "frame asks for S(Q) before trj_sq_init");
  if (do_fqt && !c->fqt_ready) return fail(c, TRJ_ESTATE, "frame asks for F(Q,t) before trj_fqt_init");
  if (!do_rdf && !do_sq && !do_fqt) return fail(c, TRJ_EINVAL, "frame asks for nothing");
  const int ndim = do_rdf ? c->rdf.ndim : (do_sq ? c->sq.p.ndim : c->fqt_ndim);
  const int nsp = do_rdf ? c->rdf.nsp : (do_sq ? c->sq.p.nsp : c->fqt_nsp);
  if (do_rdf && do_sq && (c->rdf.ndim != c->sq.p.ndim || c->rdf.nsp != c->sq.p.nsp))
    return fail(c, TRJ_EINVAL, "RDF and S(Q) were initialised with different ndim/nsp");
  if (do_fqt && (c->fqt_ndim != ndim || c->fqt_nsp != nsp))
    return fail(c, TRJ_EINVAL, "F(Q,t) was initialised with a different ndim/nsp");
  *ndim_out = ndim;
  *nsp_out = nsp;
  return TRJ_OK;
}

// The frame pipeline, device inputs.
// q_dev: the frame's charges on the device (NULL without); the caller has
// queued the frame's {scale, avq2} into qparm[0..1] on the context stream.
// qspecies (host, nsp values, may be NULL): the charges when they are constant
// within each species; S(Q) then forms the charge sums from the species
// channels instead of contracting a charge channel, and the RDF takes its
// fixed-point charge sums from the species histogram (qscale_host: the
// frame's scale).
int run_frame(trj_ctx* c, const void* pos_dev, int dtype, int layout, const int32_t* itype_dev, int64_t n,
              const double box_in[3], uint32_t what, const double* q_dev = nullptr,
              const double* qspecies = nullptr, double qscale_host = 0.0) {
  const bool do_rdf = (what & TRJ_WANT_RDF) != 0;
  const bool do_sq = (what & TRJ_WANT_SQ) != 0;
  const bool do_fqt = (what & TRJ_WANT_FQT) != 0;
  int ndim = 0, nsp = 0;
  if (int r0 = frame_shape(c, what, &ndim, &nsp)) return r0;
  const bool rdf_q = do_rdf && c->rdf_charge, sq_q = do_sq && c->sq_charge;
  const bool sq_ch = sq_q && !qspecies;  // a charge channel in the contraction
  if ((rdf_q || sq_q) && !q_dev)
    return fail(c, TRJ_EINVAL, "charges are enabled (TRJ_RDF_CHARGE / TRJ_SQ_CHARGE): frames go through trj_frame_q");
  if (do_fqt && c->fqt_frames > 0 && n != c->fqt_n)
    return fail(c, TRJ_EINVAL, "F(Q,t): the number of atoms changed between frames");
  if (n < 1) return fail(c, TRJ_EINVAL, "n must be positive");
  if (n > (int64_t(1) << 31) - 1024) return fail(c, TRJ_ELIMIT, "n above 2^31 is not supported");
  // The S(Q) kernels address the turn planes with 32-bit offsets up to 3 x the
  // plane stride (n, or 2n with the charge channel).
  if (do_sq && n * (sq_ch ? 2 : 1) > int64_t(0xFFFFFFFFu) / 3)
    return fail(c, TRJ_ELIMIT, "S(Q) with n above 1431655765 (715827882 with charges) is not supported");
  if (dtype != TRJ_F32 && dtype != TRJ_F64) return fail(c, TRJ_EINVAL, "bad dtype");
  if (layout != TRJ_AOS && layout != TRJ_SOA) return fail(c, TRJ_EINVAL, "bad layout");
  if (int r = validate_box(c, ndim, box_in)) return r;
  double box[3] = {box_in[0], box_in[1], ndim == 3 ? box_in[2] : 1.0};
  int periodic[3] = {1, 1, 1};
  if (do_rdf)
    for (int k = 0; k < 3; ++k) periodic[k] = c->rdf.periodic[k];
  else
    for (int k = 0; k < 3; ++k) periodic[k] = 1;

  EventSet& es = c->evs[c->ev_next];
  harvest(c, es);
  c->ev_next = (c->ev_next + 1) % kEventRing;
  HIPC(c, hipEventRecord(es.ev[0], c->stream));

  // Ingest: wrapped FP64 SoA, species offsets.
  int r;
  if ((r = ensure(c, c->pos, sizeof(double) * 3 * n))) return r;
  if ((r = ensure(c, c->flags, sizeof(uint32_t)))) return r;
  IngestArgs ia{};
  ia.raw = pos_dev;
  ia.dtype = dtype;
  ia.layout = layout;
  ia.itype = itype_dev;
  ia.n = n;
  ia.ndim = ndim;
  ia.nsp = nsp;
  for (int k = 0; k < 3; ++k) {
    ia.periodic[k] = periodic[k];
    ia.box[k] = box[k];
  }
  ia.pos = P<double>(c->pos);
  ia.flags = P<uint32_t>(c->flags);
  // With slices z only assigns atoms to slices: it is taken as given (no
  // wrap, no range check), as select_ncdfinput leaves it (netcdf.f90:317-331,
  // 679-682); S(Q)-only contexts have no periodic flags to wrap it by.
  const int nsl = c->slices.n;
  ia.free_axis = nsl > 0 ? 2 : -1;
  if (nsl > 0) ia.periodic[2] = 0;
  launch_ingest(ia, c->stream);
  if (nsl > 0) {
    // Slice keys, the species-sorted last atom, and the slice counts.
    if ((r = ensure(c, c->slkey, sizeof(int32_t) * n))) return r;
    HIPC(c, hipMemsetAsync(c->sllast.p, 0, sizeof(unsigned long long), c->stream));
    launch_slice_keys(P<double>(c->pos), n, itype_dev, nsp, c->slices, P<int32_t>(c->slkey),
                      P<unsigned long long>(c->sllast), c->stream);
    HIPC(c, hipMemsetAsync(c->slcount.p, 0, sizeof(int32_t) * nsl, c->stream));
    HIPC(c, hipMemsetAsync(c->slcountp.p, 0, sizeof(int32_t) * nsl * nsp, c->stream));
    HIPC(c, hipMemsetAsync(c->slqsum.p, 0, sizeof(double) * nsl, c->stream));
    HIPC(c, hipMemsetAsync(c->slqcnt.p, 0, sizeof(int32_t) * nsl, c->stream));
    launch_slice_counts(P<int32_t>(c->slkey), n, P<unsigned long long>(c->sllast), nsp, P<int32_t>(c->slcount),
                        P<int32_t>(c->slcountp), (rdf_q || sq_q) ? q_dev : nullptr, P<double>(c->slqsum),
                        P<int32_t>(c->slqcnt), c->stream);
  }
  // Sort keys: the species, or slice * nsp + species with slices; the charge
  // channels repeat each slice's (or all) sorted entries.
  const int nkeys = nsl > 0 ? nsl * nsp : nsp;
  const int ngroups = sq_ch ? (nsl > 0 ? nsl : 1) : 0;
  const int32_t* keys = nsl > 0 ? P<int32_t>(c->slkey) : nullptr;
  const int nblk = static_cast<int>((n + 255) / 256);
  if ((r = ensure(c, c->blk, sizeof(int32_t) * nkeys * nblk))) return r;
  if ((r = ensure(c, c->blkoff, sizeof(int32_t) * nkeys * nblk))) return r;
  if ((r = ensure(c, c->sp_off, sizeof(int32_t) * (nkeys + 2 + nsl)))) return r;
  // Plane stride of u: n, or 2n with charge channels (the sorted entries
  // again, weighted by their charges).
  const int64_t ustride = sq_ch ? 2 * n : n;
  launch_species_count(itype_dev, keys, n, nkeys, P<int32_t>(c->blk), nblk, c->stream);
  launch_exclusive_scan(P<int32_t>(c->blk), P<int32_t>(c->blkoff), static_cast<int64_t>(nkeys) * nblk, c->stream);
  launch_species_offsets(P<int32_t>(c->blkoff), P<int32_t>(c->blk), nblk, nkeys, nsp, ngroups,
                         P<int32_t>(c->sp_off), c->stream);
  if (do_sq) {
    // Key-contiguous turns u = x / Lq for the contraction (Lq: the box the
    // q-grid was built on, sq.cuf:189); slices use x and y only.
    if ((r = ensure(c, c->u, sizeof(double) * (sq_ch ? 13 : 12) * ustride))) return r;
    const double lq[3] = {c->sq.box[0], c->sq.box[1], c->sq.box[2]};
    launch_species_scatter(itype_dev, keys, n, nkeys, P<int32_t>(c->blkoff), nblk, P<double>(c->pos), lq,
                           nsl > 0 ? 2 : ndim, sq_ch ? q_dev : nullptr, P<int32_t>(c->sp_off), ustride,
                           P<double>(c->u), c->stream);
  }
  HIPC(c, hipGetLastError());
  HIPC(c, hipEventRecord(es.ev[1], c->stream));

  // RDF (RDFcomp, rdf.cuf:184-353), on its own stream when overlapped with
  // the S(Q) contraction (it reads pos, itype and its own buffers only).
  const bool overlap = c->pipe == TRJ_PIPE_OVERLAP && do_rdf && do_sq;
  hipStream_t rs = overlap ? c->rdf_stream : c->stream;
  if (overlap) {
    HIPC(c, hipStreamWaitEvent(rs, es.ev[1], 0));
    HIPC(c, hipEventRecord(es.ev[5], rs));
  }
  if (do_rdf) {
    RdfPlan& R = c->rdf;
    FrameGeom g = frame_geometry(ndim, box, R.periodic, &R.rc);
    PairArgs pa{};
    pa.ndim = ndim;
    pa.nsp = nsp;
    pa.nit = R.nit;
    pa.lsmax = R.lsmax;
    for (int k = 0; k < 3; ++k) {
      pa.lmi[k] = g.lmi[k];
      pa.half[k] = g.half[k];
    }
    pa.side2 = g.side2;
    pa.inv_dr = static_cast<float>(1.0 / R.deltar);
    pa.thr = P<double>(c->thr);
    pa.thr_host = R.thr.data();
    pa.deltar = R.deltar;
    pa.floor_bins = R.floor_bins ? 1 : 0;
    // Species charges: the merge adds the fixed-point sums from the species
    // histogram; the pair kernels then run without their charge path.
    const bool qq_species = rdf_q && qspecies && nsl == 0;
    QqWeights qw{};
    if (qq_species) {
      int ij = 0;
      for (int i = 0; i < nsp; ++i)
        for (int j = i; j < nsp; ++j, ++ij)
          qw.w[ij] = static_cast<long long>(std::rint(qspecies[i] * qspecies[j] * qscale_host));
    }
    if (rdf_q) {
      pa.gq = P<unsigned long long>(c->gqf);
      pa.qp = P<double>(c->qparm);
    }
    const int nbins = R.nit * R.lsmax;
    if (nsl > 0) {
      if ((r = rdf_slices(c, pa, g, box, n, itype_dev, rdf_q ? q_dev : nullptr, rs))) return r;
    } else {
      // Cell grid: side >= cutoff (times the autotuner's scale) on every axis.
      auto& tu = c->rdf_tune;
      auto plan = [&](double scale) {
        return rdf_cell_plan(ndim, box, periodic, g.side2, n, tu.done ? tu.W : 0, scale);
      };
      RdfCells rcl = plan(tu.done ? tu.scale : (c->rdf_fixed_scale > 0.0 ? c->rdf_fixed_scale : 1.0));
      // half-cutoff cells only where the plan allows them and no sampling
      // spheres read the cells
      if (rcl.cg.reach == 2 && (!rcl.cells_ok || c->nrand > 0)) rcl = plan(1.0);
      const bool cells_ok = rcl.cells_ok;
      for (int k = 0; k < 3; ++k) pa.cellshift[k] = rcl.cellshift[k];
      pa.row0 = 0;
      pa.nrows = tu.done ? tu.rows : R.rows;
      int nblocks = rcl.nblocks;
      if (cells_ok) {
        if ((r = ensure(c, c->spos, sizeof(double) * 3 * n))) return r;
        if ((r = ensure(c, c->stype, sizeof(int32_t) * n))) return r;
        const bool qq_pairs = rdf_q && !qq_species;
        if (qq_pairs && (r = ensure(c, c->qsort, sizeof(double) * n))) return r;
        // Counting sort of the frame's atoms into the cells of plan p:
        // cell starts, cell-sorted positions, species (and charges).
        auto sort_cells = [&](const RdfCells& p) -> int {
          const int64_t ncell = p.ncell;
          int rr;
          if ((rr = ensure(c, c->cellid, sizeof(int32_t) * n))) return rr;
          if ((rr = ensure(c, c->cellcount, sizeof(int32_t) * (ncell + 1)))) return rr;
          if ((rr = ensure(c, c->cellstart, sizeof(int32_t) * (ncell + 1)))) return rr;
          const CellSort cs = c->atomic_cells ? CellSort{} : cell_sort_plan(n, ncell, c->lds_max);
          if (cs.lds) {
            if ((rr = ensure(c, c->cellblk, sizeof(int32_t) * cs.blocks * static_cast<size_t>(ncell)))) return rr;
            launch_cell_sort(cs, P<double>(c->pos), nullptr, 0, 0, n, ndim, p.cg, ncell, P<int32_t>(c->cellid),
                             P<int32_t>(c->cellblk), P<int32_t>(c->cellcount), P<int32_t>(c->cellstart), itype_dev,
                             P<double>(c->spos), P<int32_t>(c->stype), qq_pairs ? q_dev : nullptr,
                             qq_pairs ? P<double>(c->qsort) : nullptr, rs);
          } else {
            if ((rr = ensure(c, c->cursor, sizeof(int32_t) * ncell))) return rr;
            HIPC(c, hipMemsetAsync(c->cellcount.p, 0, sizeof(int32_t) * (ncell + 1), rs));
            HIPC(c, hipMemsetAsync(c->cursor.p, 0, sizeof(int32_t) * ncell, rs));
            launch_cell_count(P<double>(c->pos), n, ndim, p.cg, P<int32_t>(c->cellid), P<int32_t>(c->cellcount),
                              rs);
            launch_exclusive_scan(P<int32_t>(c->cellcount), P<int32_t>(c->cellstart), ncell + 1, rs);
            launch_cell_scatter(P<double>(c->pos), itype_dev, P<int32_t>(c->cellid), n, P<int32_t>(c->cellstart),
                                P<int32_t>(c->cursor), P<double>(c->spos), P<int32_t>(c->stype),
                                qq_pairs ? q_dev : nullptr, qq_pairs ? P<double>(c->qsort) : nullptr, rs);
          }
          HIPC(c, hipGetLastError());
          return TRJ_OK;
        };
        if ((r = sort_cells(rcl))) return r;
        if (qq_pairs) pa.q = P<double>(c->qsort);
        if (!tu.done && tu.enabled && !c->noalloc && !qq_pairs) {
          if ((r = tune_pairs(c, pa, rcl, n, rs, plan, sort_cells))) return r;
          rcl = plan(tu.scale);
          // the trials leave the cells of the last grid tried
          if ((r = sort_cells(rcl))) return r;
          for (int k = 0; k < 3; ++k) pa.cellshift[k] = rcl.cellshift[k];
          pa.nrows = tu.rows;
          nblocks = rcl.nblocks;
        }
        const CellGrid& cg = rcl.cg;
        for (int k = 0; k < 3; ++k) c->rdf_nc[k] = cg.nc[k];
        c->rdf_cell_scale = rcl.scale;
        const int W = rcl.W;
        if ((r = ensure(c, c->partial, sizeof(uint32_t) * nbins * static_cast<size_t>(nblocks)))) return r;
        launch_rdf_cells(pa, cg, P<int32_t>(c->cellstart), P<double>(c->spos), P<int32_t>(c->stype), n, W,
                         nblocks, P<uint32_t>(c->partial), rs);
      } else {
        if ((r = ensure(c, c->partial, sizeof(uint32_t) * nbins * static_cast<size_t>(nblocks)))) return r;
        if (rdf_q && !qq_species) pa.q = q_dev;  // input order, like pos and itype here
        launch_rdf_allpairs(pa, P<double>(c->pos), itype_dev, n, nblocks, P<uint32_t>(c->partial), rs);
      }
      launch_rdf_merge(P<uint32_t>(c->partial), nblocks, nbins, P<int64_t>(c->counts), rs, qq_species ? &qw : nullptr,
                       R.lsmax, P<unsigned long long>(c->gqf));
      if (rdf_q) launch_qq_finish(P<unsigned long long>(c->gqf), R.lsmax, P<double>(c->qparm), P<double>(c->gqacc), rs);
      if (c->nrand > 0) {
        // Number fluctuations around the fixed random points (rdf.cuf:243-245,
        // 251-253), on this frame's cell list when there is one.
        S2nArgs sa{};
        sa.pts = P<double>(c->rpts);
        sa.np = c->nrand;
        sa.ndim = ndim;
        sa.lsmax = R.lsmax;
        for (int k = 0; k < 3; ++k) {
          sa.box[k] = box[k];
          sa.lmi[k] = g.lmi[k];
          sa.half[k] = g.half[k];
        }
        sa.side2 = g.side2;
        sa.inv_dr = pa.inv_dr;
        sa.thr = P<double>(c->thr);
        sa.pos = cells_ok ? P<double>(c->spos) : P<double>(c->pos);
        sa.n = n;
        sa.start = cells_ok ? P<int32_t>(c->cellstart) : nullptr;
        sa.g = rcl.cg;
        sa.gi = P<unsigned long long>(c->s2gi);
        sa.g2i = P<unsigned long long>(c->s2g2i);
        launch_s2n(sa, rs);
      }
    }
    HIPC(c, hipGetLastError());
    c->rdf_frames += 1;
    c->last_volume = g.volume;
  } else {
    double v = 1.0;
    for (int k = 0; k < ndim; ++k) v *= box[k];
    c->last_volume = v;
  }
  if (overlap) HIPC(c, hipEventRecord(es.ev[6], rs));
  HIPC(c, hipEventRecord(es.ev[2], c->stream));

  // S(Q) (SQcalc, sq.cuf:58-126).
  if (do_sq) {
    SqPlan& S = c->sq;
    const int ntiles = static_cast<int>(S.tiles.size());
    const int nch = nkeys + ngroups;  // (slice,) species channels, then charge channels
    c->sq_nch = nch;
    const int sdim = nsl > 0 ? 2 : ndim;
    // With slices every vector's values are divided by the slice's (species')
    // atom count (sqf3xy / sqf2xy, sq.cuf:653-670, 741-765).
    const int32_t* norm = (nsl > 0 || c->sq_centres) ? P<int32_t>(c->sp_off) : nullptr;
    const int ngs = nsl > 0 ? nsl : 1;
    if (ntiles > 0) {
      ContractArgs ca{};
      ca.u = P<double>(c->u);
      ca.w = sq_ch ? P<double>(c->u) + 12 * ustride : nullptr;
      ca.sp_off = P<int32_t>(c->sp_off);
      ca.tiles = P<int32_t>(c->tiles);
      ca.ntiles = ntiles;
      ca.nsp = nsp;
      ca.nch = nch;
      ca.ndim = sdim;
      ca.n = ustride;
      // FP64 and 16-bit: the phase anchors of the tiles' segments as a
      // table, computed once per frame (for lattice_262k in FP64 82 rows x
      // 262144 entries, 344 MB, against 4.7e9 anchor evaluations inside the
      // contraction).  Not for the allocation-free entry point.
      const int64_t anc_bytes = sq_anchor_bytes(S, ustride);
      if (c->sq_anchor_table && !c->noalloc && anc_bytes > 0) {
        if ((r = ensure(c, c->anc, static_cast<size_t>(anc_bytes)))) return r;
        if (S.p.precision == TRJ_PREC_FP64) {
          launch_sq_anchors(P<double>(c->u), ustride, S.anc_rows, S.anc_base, P<double2>(c->anc), c->stream);
          ca.anc = P<double2>(c->anc);
        } else {
          launch_sq_anchors32(P<double>(c->u), ustride, S.anc_rows, S.anc_base, P<float2>(c->anc), c->stream);
          ca.anc32 = P<float2>(c->anc);
          ca.anc_nrows = S.anc_rows[0] + S.anc_rows[1] + S.anc_rows[2];
        }
        ca.anc_row[0] = 0;
        ca.anc_row[1] = S.anc_rows[0];
        ca.anc_row[2] = S.anc_rows[0] + S.anc_rows[1];
        for (int k = 0; k < 3; ++k) ca.anc_base[k] = S.anc_base[k];
      }
      auto& st = c->sq_tune;
      if (S.p.precision == TRJ_PREC_FP64 && st.enabled && !st.done && !c->noalloc) {
        // trials: the contraction and a shell pass into scratch sums
        const size_t ns = static_cast<size_t>(S.nqmax + 1) * ngs;
        if ((r = ensure(c, c->tunebuf, sizeof(double) * ns * (2 + nsp)))) return r;
        double* tsqf = P<double>(c->tunebuf);
        auto run = [&](int ks, int order) {
          ca.ksplit = ks;
          ca.xcd_order = order;
          ca.rho = P<double>(c->rho);
          launch_sq_contract_f64(ca, c->stream);
          launch_sq_shells(c->rho.p, false, P<int16_t>(c->tile_shell), ntiles, S.tk.vec(), ks, nsp, ngs, nch,
                           S.p.bsc, S.nqmax, norm, sq_q ? qspecies : nullptr, tsqf, tsqf + ns, tsqf + ns * (1 + nsp),
                           c->stream);
        };
        if ((r = tune_contraction(c, ntiles, n, nch, run))) return r;
      }
      const int ks = contraction_splits(ntiles, n, S.p.precision, nch, st.done ? st.span : 0);
      c->ksplit = ks;
      const size_t rho_bytes =
          rho_entry_bytes(S.p.precision) * static_cast<size_t>(ks) * nch * ntiles * S.tk.vec();
      if ((r = ensure(c, c->rho, rho_bytes))) return r;
      ca.ksplit = ks;
      ca.xcd_order = st.done ? st.order : 0;
      ca.rho = P<double>(c->rho);
      if (S.p.precision == TRJ_PREC_FP64)
        launch_sq_contract_f64(ca, c->stream);
      else
        launch_sq_contract_lp(ca, S.p.precision, c->stream);
    }
    HIPC(c, hipGetLastError());
    HIPC(c, hipEventRecord(es.ev[3], c->stream));
    if (ntiles > 0)
      launch_sq_shells(c->rho.p, S.p.precision != TRJ_PREC_FP64, P<int16_t>(c->tile_shell), ntiles, S.tk.vec(),
                       c->ksplit, nsp, ngs, nch,
                       S.p.bsc, S.nqmax, norm, sq_q ? qspecies : nullptr, P<double>(c->sqf), P<double>(c->sqfp),
                       P<double>(c->sqfq), c->stream);
    const int nax = static_cast<int>(S.axis_shell.size());
    if ((r = ensure(c, c->listpart, sq_list_scratch_bytes(nax, nch)))) return r;
    launch_sq_list(P<double>(c->u), sq_ch ? P<double>(c->u) + 12 * ustride : nullptr, P<int32_t>(c->sp_off), ustride,
                   nsp, ngs, nch, sdim, P<int32_t>(c->axis_kv), P<int32_t>(c->axis_shell), nax, S.p.bsc, S.nqmax,
                   nsl > 0 || c->sq_centres, sq_q ? qspecies : nullptr, P<double>(c->listpart), P<double>(c->sqf),
                   P<double>(c->sqfp), P<double>(c->sqfq), c->stream);
    // sq.cuf:143-145 divides by avq2 of the last frame: keep it in qparm[2];
    // with slices, by each slice's mean q^2 of the last frame (sq.cuf:135-138).
    if (sq_q && nsl == 0)
      HIPC(c, hipMemcpyAsync(P<double>(c->qparm) + 2, P<double>(c->qparm) + 1, sizeof(double),
                             hipMemcpyDeviceToDevice, c->stream));
    if (sq_q && nsl > 0) {
      HIPC(c, hipMemcpyAsync(c->sqqsum.p, c->slqsum.p, sizeof(double) * nsl, hipMemcpyDeviceToDevice, c->stream));
      HIPC(c, hipMemcpyAsync(c->sqqcnt.p, c->slqcnt.p, sizeof(int32_t) * nsl, hipMemcpyDeviceToDevice, c->stream));
    }
    HIPC(c, hipGetLastError());
    c->sq_frames += 1;
  } else {
    HIPC(c, hipEventRecord(es.ev[3], c->stream));
  }
  if (overlap) HIPC(c, hipStreamWaitEvent(c->stream, es.ev[6], 0));
  HIPC(c, hipEventRecord(es.ev[4], c->stream));
  es.pending = true;
  es.has_rdf = do_rdf;
  es.has_sq = do_sq;
  es.overlapped = overlap;
  if (do_fqt && (r = fqt_step(c, n, itype_dev))) return r;
  c->last_n = n;
  c->last_nsp = nsp;
  for (int k = 0; k < 3; ++k) c->last_box[k] = box[k];
  return TRJ_OK;
}

}  // namespace

namespace trj {
hipStream_t context_stream(trj_ctx* c) { return c->stream; }

// K-splits of the contraction (the splits are summed in FP64 by the shells
// kernel): at least 2048 workgroups when the atoms allow it; for FP64, atom
// ranges of at most fp64_span (0: 32768, the autotuner may pick another) so
// that the workgroups running at one time share
// their turns (96 B per atom) in L2, with the partial rho buffer (one per
// split) kept within 4 GiB; for the FP32-accumulating kernels at most 8192
// atoms per workgroup (their accuracy bound, not capped).  The 16-bit kernel
// holds a whole CU per workgroup and its workgroups take equal time, so the
// split count is raised (to at most twice) until the workgroups fill whole
// rounds of kWgRound = 256, the CU count of the MI355X.  The count, and with it
// the FP32 accumulation length that sets the 16-bit error, depends on the
// frame only, never on the CU count the device reports.
int contraction_splits(int ntiles, int64_t n, int precision, int nsp, int64_t fp64_span) {
  constexpr int64_t kWgRound = 256;
  int ks = static_cast<int>((2048 + ntiles - 1) / ntiles);
  ks = std::max(1, std::min<int>(ks, static_cast<int>(std::max<int64_t>(1, n / 8192))));
  const int64_t span = precision == TRJ_PREC_FP64 ? (fp64_span > 0 ? fp64_span : 32768) : 8192;
  ks = std::max<int>(ks, static_cast<int>((n + span - 1) / span));
  // Small frames on few tiles: enough splits to fill one round of kWgRound
  // workgroups while each split keeps at least 1024 atoms (32 chunks), e.g. 9
  // splits instead of 2 for 1e4 atoms on the 8 tiles of the 64^3 half-space.
  ks = std::max<int>(ks, static_cast<int>(std::min<int64_t>((kWgRound + ntiles - 1) / ntiles, n / 1024)));
  if ((precision == TRJ_PREC_FP16X3 || precision == TRJ_PREC_BF16X3) && ntiles > 0) {
    const int64_t wgs = static_cast<int64_t>(ntiles) * ks;
    const int64_t rounds = (wgs + kWgRound - 1) / kWgRound;
    // the smallest split count whose workgroups fill `rounds` rounds, if any
    const int64_t want = (rounds * kWgRound) / ntiles;
    if (want * ntiles == rounds * kWgRound && want >= ks && want <= 2 * static_cast<int64_t>(ks))
      ks = static_cast<int>(want);
  }
  if (precision != TRJ_PREC_FP64) return ks;
  const int64_t per_split = int64_t(16) * std::max(1, nsp) * std::max(1, ntiles) * tile_shape(precision).vec();
  const int64_t cap = std::max<int64_t>(1, (int64_t(4) << 30) / per_split);
  return static_cast<int>(std::min<int64_t>(ks, cap));
}

// Frees every device buffer of an idle context and leaves it as a new one:
// no module initialised, no slices, no sampling points.  The streams, events
// and pinned staging stay.
int ctx_release(trj_ctx* c) {
  hipSetDevice(c->device);
  if (int r = sync_all(c)) return r;
  if (c->rdf_stream) HIPC(c, hipStreamSynchronize(c->rdf_stream));
  for (DevBuf* b : dev_bufs(c)) {
    if (b->p) hipFree(b->p);
    b->p = nullptr;
    b->cap = 0;
  }
  c->rdf_ready = c->sq_ready = c->fqt_ready = false;
  c->rdf_charge = c->sq_charge = c->sq_centres = false;
  c->slices = SliceSet{};
  c->nrand = 0;
  c->rpts_host.clear();
  c->rdf_frames = c->sq_frames = 0;
  c->fqt_frames = c->fqt_count_org = 0;
  // The error flags word is read after every frame and must start at zero,
  // as trj_create leaves it.
  if (int r = ensure(c, c->flags, sizeof(uint32_t))) return r;
  HIPC(c, hipMemsetAsync(c->flags.p, 0, sizeof(uint32_t), c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  return TRJ_OK;
}
}  // namespace trj

namespace {

int read_ntype(trj_ctx* c, std::vector<int64_t>& ntype) {
  const int nsp = c->last_nsp;
  std::vector<int32_t> off(nsp + 1, 0);
  HIPC(c, hipMemcpyAsync(off.data(), c->sp_off.p, sizeof(int32_t) * (nsp + 1), hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  ntype.assign(nsp, 0);
  for (int s = 0; s < nsp; ++s) ntype[s] = off[s + 1] - off[s];
  return TRJ_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
extern "C" {

int trj_abi_version(void) { return TRJ_ABI_VERSION; }

int trj_create(trj_ctx** out, int device) {
  if (!out) return TRJ_EINVAL;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return TRJ_EDEVICE;
  if (device < 0 || device >= ndev) return TRJ_EINVAL;
  if (hipSetDevice(device) != hipSuccess) return TRJ_EDEVICE;
  trj_ctx* c = new trj_ctx();
  c->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->ncu = prop.multiProcessorCount;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return TRJ_EDEVICE;
  }
  for (auto& s : c->evs)
    for (auto& e : s.ev)
      if (hipEventCreate(&e) != hipSuccess) {
        trj_destroy(c);
        return TRJ_EDEVICE;
      }
  for (auto& e : c->pinned_done)
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
      trj_destroy(c);
      return TRJ_EDEVICE;
    }
  for (auto& e : c->tune_ev)
    if (hipEventCreate(&e) != hipSuccess) {
      trj_destroy(c);
      return TRJ_EDEVICE;
    }
  int lds = 0;
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && lds > 0)
    c->lds_max = static_cast<size_t>(lds);
  // The device address of this device's allocation-free scratch array (a
  // failure is reported by trj_frame_device_noalloc, the only user).
  resolve_scratch(device);
  if (ensure(c, c->flags, sizeof(uint32_t)) != TRJ_OK ||
      hipMemsetAsync(c->flags.p, 0, sizeof(uint32_t), c->stream) != hipSuccess) {
    trj_destroy(c);
    return TRJ_ENOMEM;
  }
  *out = c;
  return TRJ_OK;
}

void trj_destroy(trj_ctx* c) {
  if (!c) return;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  for (DevBuf* b : dev_bufs(c))
    if (b->p) hipFree(b->p);
  for (int k = 0; k < 2; ++k) {
    if (c->pinned[k]) hipHostFree(c->pinned[k]);
    if (c->pinned_done[k]) hipEventDestroy(c->pinned_done[k]);
  }
  for (auto& s : c->evs)
    for (auto& e : s.ev)
      if (e) hipEventDestroy(e);
  for (auto& e : c->tune_ev)
    if (e) hipEventDestroy(e);
  if (c->rdf_stream) hipStreamDestroy(c->rdf_stream);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
}

const char* trj_last_error(const trj_ctx* c) { return c ? c->err.c_str() : "null context"; }

int trj_set_pipeline(trj_ctx* c, int mode) {
  if (!c) return TRJ_EINVAL;
  if (mode != TRJ_PIPE_SERIAL && mode != TRJ_PIPE_OVERLAP) return fail(c, TRJ_EINVAL, "unknown pipeline mode");
  hipSetDevice(c->device);
  int r = sync_all(c);
  if (r) return r;
  if (mode == TRJ_PIPE_OVERLAP && !c->rdf_stream)
    HIPC(c, hipStreamCreateWithFlags(&c->rdf_stream, hipStreamNonBlocking));
  c->pipe = mode;
  return TRJ_OK;
}

int trj_set_slices(trj_ctx* c, int32_t nslice, const float* zslice, float zgrid) {
  if (!c) return TRJ_EINVAL;
  if (c->rdf_ready || c->sq_ready) return fail(c, TRJ_ESTATE, "trj_set_slices comes before trj_rdf_init / trj_sq_init");
  if (nslice < 0 || nslice > kMaxSlices) return fail(c, TRJ_EINVAL, "nslice must be in 0..16");
  if (nslice > 0 && !zslice) return fail(c, TRJ_EINVAL, "NULL zslice");
  if (nslice > 0 && !(zgrid > 0.0f && std::isfinite(zgrid))) return fail(c, TRJ_EINVAL, "zgrid must be positive");
  for (int k = 0; k < nslice; ++k)
    if (!std::isfinite(zslice[k])) return fail(c, TRJ_EINVAL, "zslice values must be finite");
  hipSetDevice(c->device);
  int r;
  c->slices = SliceSet{};
  if (nslice == 0) return TRJ_OK;
  if ((r = ensure(c, c->sllast, sizeof(unsigned long long)))) return r;
  if ((r = ensure(c, c->slcount, sizeof(int32_t) * nslice))) return r;
  if ((r = ensure(c, c->slcountp, sizeof(int32_t) * nslice * kNspMax))) return r;
  if ((r = ensure(c, c->slqsum, sizeof(double) * nslice))) return r;
  if ((r = ensure(c, c->slqcnt, sizeof(int32_t) * nslice))) return r;
  if ((r = ensure(c, c->sqqsum, sizeof(double) * nslice))) return r;
  if ((r = ensure(c, c->sqqcnt, sizeof(int32_t) * nslice))) return r;
  HIPC(c, hipMemsetAsync(c->sqqsum.p, 0, sizeof(double) * nslice, c->stream));
  HIPC(c, hipMemsetAsync(c->sqqcnt.p, 0, sizeof(int32_t) * nslice, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  c->slices.n = nslice;
  for (int k = 0; k < nslice; ++k) c->slices.z[k] = zslice[k];
  c->slices.zgrid = zgrid;
  return TRJ_OK;
}

int trj_rdf_init(trj_ctx* c, const trj_rdf_params* p, const double box[3], int32_t* lsmax_out,
                 int32_t* nit_out) {
  if (!c || !p) return TRJ_EINVAL;
  if (p->ndim != 2 && p->ndim != 3) return fail(c, TRJ_EINVAL, "rdf: ndim must be 2 or 3");
  if (p->nsp < 1 || p->nsp > kNspMax) return fail(c, TRJ_EINVAL, "rdf: nsp must be in 1..10");
  if (!(p->deltar > 0.0)) return fail(c, TRJ_EINVAL, "rdf: deltar must be positive");
  if (int r = validate_box(c, p->ndim, box)) return r;
  hipSetDevice(c->device);
  int r = sync_all(c);
  if (r) return r;
  RdfPlan R;
  R.ndim = p->ndim;
  R.nsp = p->nsp;
  R.nit = p->nsp * (p->nsp + 1) / 2;
  R.deltar = p->deltar;
  for (int k = 0; k < 3; ++k) R.periodic[k] = (k < p->ndim) ? (p->periodic[k] != 0) : 1;
  R.lsmax = rdf_lsmax(p->ndim, box, p->rcrdf, p->deltar, &R.rc);
  if (R.lsmax < 1) return fail(c, TRJ_EINVAL, "rdf: lsmax < 1");
  if (p->flags & ~(TRJ_RDF_CHARGE | TRJ_RDF_FLOOR_BINS | TRJ_RDF_NO_AUTOTUNE))
    return fail(c, TRJ_EINVAL, "rdf: unknown flags");
  const bool qq = (p->flags & TRJ_RDF_CHARGE) != 0;
  // The histogram sits in LDS in tiles of whole pair-type rows (rdf.cuf:61-64
  // refuses a grid larger than the shared memory instead): as many rows as
  // the device's LDS holds per pass, at least one.
  R.rows = rows_fitting(R.nit, R.lsmax, qq, c->lds_max);
  if (R.rows < 1)
    return fail(c, TRJ_ELIMIT,
                "rdf: one row of the g(r) grid needs " + std::to_string(cells_lds_bytes(1, R.lsmax, qq)) +
                    " bytes of LDS, above the device's " + std::to_string(c->lds_max) +
                    "; increase deltar or reduce rcrdf (rdf.cuf:61-64)");
  c->rdf_tune = {};
  c->rdf_tune.enabled = (p->flags & TRJ_RDF_NO_AUTOTUNE) == 0;
  c->rdf_fixed_scale = 0.0;
  const char* e = c->rdf_tune.enabled ? std::getenv("TRJ_RDF_CELL_SCALE") : nullptr;
  if (e) {
    const double f = std::atof(e);
    if (f == 0.5 || (f >= 1.0 && f <= 4.0)) c->rdf_fixed_scale = f;
  }
  const int nsl = c->slices.n;
  if (nsl > 0) {
    if (R.ndim != 3) return fail(c, TRJ_EINVAL, "rdf: slices need ndim = 3");
    if (!R.periodic[0] || !R.periodic[1]) return fail(c, TRJ_EINVAL, "rdf: slices need periodic x and y");
    std::vector<double> tx = rdf_bin_thresholds_floor(R.deltar, R.lsmax);
    tx.push_back(INFINITY);
    tx[0] = 0.0;
    R.thr_xy = tx;
    const size_t nb = static_cast<size_t>(nsl) * R.nit * R.lsmax;
    if ((r = ensure(c, c->thr_xy, sizeof(double) * tx.size()))) return r;
    if ((r = ensure(c, c->xyframe, sizeof(int64_t) * nb))) return r;
    if ((r = ensure(c, c->xycounts, sizeof(int64_t) * nb))) return r;
    if ((r = ensure(c, c->xyacc, sizeof(double) * nb))) return r;
    HIPC(c, hipMemcpyAsync(c->thr_xy.p, tx.data(), sizeof(double) * tx.size(), hipMemcpyHostToDevice, c->stream));
    HIPC(c, hipMemsetAsync(c->xyframe.p, 0, sizeof(int64_t) * nb, c->stream));
    HIPC(c, hipMemsetAsync(c->xycounts.p, 0, sizeof(int64_t) * nb, c->stream));
    HIPC(c, hipMemsetAsync(c->xyacc.p, 0, sizeof(double) * nb, c->stream));
    if (qq) {
      const size_t nq = static_cast<size_t>(nsl) * R.lsmax;
      if ((r = ensure(c, c->xygqf, sizeof(unsigned long long) * nq))) return r;
      if ((r = ensure(c, c->xygqacc, sizeof(double) * nq))) return r;
      HIPC(c, hipMemsetAsync(c->xygqf.p, 0, sizeof(unsigned long long) * nq, c->stream));
      HIPC(c, hipMemsetAsync(c->xygqacc.p, 0, sizeof(double) * nq, c->stream));
    }
    // Keep the host copy alive until the copy above has completed.
    HIPC(c, hipStreamSynchronize(c->stream));
  }
  std::vector<double> thr = (p->flags & TRJ_RDF_FLOOR_BINS) ? rdf_bin_thresholds_floor(R.deltar, R.lsmax)
                                                            : rdf_bin_thresholds(R.deltar, R.lsmax);
  thr.push_back(INFINITY);  // thr[lsmax+2]
  thr[0] = 0.0;
  R.thr = thr;
  R.floor_bins = (p->flags & TRJ_RDF_FLOOR_BINS) != 0;
  if ((r = ensure(c, c->thr, sizeof(double) * thr.size()))) return r;
  if ((r = ensure(c, c->counts, sizeof(int64_t) * R.nit * R.lsmax))) return r;
  // Species offsets of the last frame, at their largest size: the result
  // calls read them, so they stay in the context (also after an
  // allocation-free frame).
  if ((r = ensure(c, c->sp_off, sizeof(int32_t) * (kKeyMax + 2 + kMaxSlices)))) return r;
  HIPC(c, hipMemcpyAsync(c->thr.p, thr.data(), sizeof(double) * thr.size(), hipMemcpyHostToDevice, c->stream));
  HIPC(c, hipMemsetAsync(c->counts.p, 0, sizeof(int64_t) * R.nit * R.lsmax, c->stream));
  if (qq) {
    if ((r = ensure(c, c->gqf, sizeof(unsigned long long) * R.lsmax))) return r;
    if ((r = ensure(c, c->gqacc, sizeof(double) * R.lsmax))) return r;
    if ((r = ensure(c, c->qparm, 4 * sizeof(double)))) return r;
    HIPC(c, hipMemsetAsync(c->gqf.p, 0, sizeof(unsigned long long) * R.lsmax, c->stream));
    HIPC(c, hipMemsetAsync(c->gqacc.p, 0, sizeof(double) * R.lsmax, c->stream));
  }
  HIPC(c, hipStreamSynchronize(c->stream));
  c->rdf = R;
  c->rdf_charge = qq;
  c->rdf_ready = true;
  c->nrand = 0;  // sampling points follow the plan: trj_rdf_sampling again
  c->rpts_host.clear();
  c->rdf_frames = 0;
  if (lsmax_out) *lsmax_out = R.lsmax;
  if (nit_out) *nit_out = R.nit;
  return TRJ_OK;
}

int trj_sq_init(trj_ctx* c, const trj_sq_params* p, const double box[3], trj_sq_info* info) {
  if (!c || !p) return TRJ_EINVAL;
  hipSetDevice(c->device);
  int r = sync_all(c);
  if (r) return r;
  SqPlan S;
  std::string err;
  const int nsl = c->slices.n;
  if (nsl > 0) {
    // Slices: the 2D q-set in the xy plane of each slice (sq.cuf:156-200 with
    // twoDstruc_3D), axis vectors from nqmin + 1 (sqf2xy, sq.cuf:91-92,
    // 114-116; from 1 when the 16-adjust leaves nqmin below 1).
    if (p->ndim != 3) return fail(c, TRJ_EINVAL, "sq: slices need ndim = 3");
    if (p->qset != TRJ_QSET_REF) return fail(c, TRJ_EINVAL, "sq: slices use the reference q-set");
    trj_sq_params p2 = *p;
    p2.ndim = 2;
    r = sq_plan(p2, box, &S, &err);
    if (r) return fail(c, r, err);
    S.p.ndim = 3;
    S.axis_kv.clear();
    S.axis_shell.clear();
    std::fill(S.nq_summed.begin(), S.nq_summed.end(), 0);
    for (size_t e = 0; e < S.tile_shell.size(); ++e)
      if (S.tile_shell[e] >= 0) S.nq_summed[S.tile_shell[e]] += 1;
    for (int k = std::max(1, S.nqmin + 1); k <= S.nqmax; ++k)
      for (int d = 0; d < 2; ++d) {
        int v[3] = {0, 0, 0};
        v[d] = k;
        S.axis_kv.insert(S.axis_kv.end(), v, v + 3);
        S.axis_shell.push_back(k);
        S.nq_summed[k] += 1;
      }
    S.info.nvec_axis = static_cast<int64_t>(S.axis_shell.size());
    if (p->flags & TRJ_SQ_NQ_CONSISTENT) S.nq = S.nq_summed;
  } else {
    r = sq_plan(*p, box, &S, &err);
    if (r) return fail(c, r, err);
    if (p->flags & TRJ_SQ_CENTRES) {
      // sqfnDcl (sq.cuf:798-839) sums the octant set only.
      if (p->qset != TRJ_QSET_REF) return fail(c, TRJ_EINVAL, "sq: TRJ_SQ_CENTRES uses the reference q-set");
      for (int k : S.axis_shell) S.nq_summed[k] -= 1;
      S.axis_kv.clear();
      S.axis_shell.clear();
      S.info.nvec_axis = 0;
    }
  }
  c->sq_tune = {};
  c->sq_tune.enabled = (p->flags & TRJ_SQ_NO_AUTOTUNE) == 0;
  {
    const char* e = std::getenv("TRJ_SQ_ANCHOR_TABLE");
    c->sq_anchor_table = !(e && std::strcmp(e, "0") == 0);
  }
  const int ns = S.nqmax + 1;
  const int ngs = nsl > 0 ? nsl : 1;
  if ((r = ensure(c, c->sqf, sizeof(double) * ns * ngs))) return r;
  if ((r = ensure(c, c->sqfp, sizeof(double) * ns * S.p.nsp * ngs))) return r;
  if ((r = ensure(c, c->tiles, sizeof(SqTile) * std::max<size_t>(1, S.tiles.size())))) return r;
  if ((r = ensure(c, c->tile_shell, sizeof(int16_t) * std::max<size_t>(1, S.tile_shell.size())))) return r;
  if ((r = ensure(c, c->axis_kv, sizeof(int32_t) * std::max<size_t>(1, S.axis_kv.size())))) return r;
  if ((r = ensure(c, c->axis_shell, sizeof(int32_t) * std::max<size_t>(1, S.axis_shell.size())))) return r;
  HIPC(c, hipMemsetAsync(c->sqf.p, 0, sizeof(double) * ns * ngs, c->stream));
  HIPC(c, hipMemsetAsync(c->sqfp.p, 0, sizeof(double) * ns * S.p.nsp * ngs, c->stream));
  const bool sqq = (p->flags & TRJ_SQ_CHARGE) != 0;
  if (sqq) {
    if ((r = ensure(c, c->sqfq, sizeof(double) * ns * ngs))) return r;
    if ((r = ensure(c, c->qparm, 4 * sizeof(double)))) return r;
    HIPC(c, hipMemsetAsync(c->sqfq.p, 0, sizeof(double) * ns * ngs, c->stream));
    HIPC(c, hipMemsetAsync(c->qparm.p, 0, 4 * sizeof(double), c->stream));
  }
  if (!S.tiles.empty()) {
    HIPC(c, hipMemcpyAsync(c->tiles.p, S.tiles.data(), sizeof(SqTile) * S.tiles.size(), hipMemcpyHostToDevice,
                           c->stream));
    HIPC(c, hipMemcpyAsync(c->tile_shell.p, S.tile_shell.data(), sizeof(int16_t) * S.tile_shell.size(),
                           hipMemcpyHostToDevice, c->stream));
  }
  if (!S.axis_shell.empty()) {
    HIPC(c, hipMemcpyAsync(c->axis_kv.p, S.axis_kv.data(), sizeof(int32_t) * S.axis_kv.size(),
                           hipMemcpyHostToDevice, c->stream));
    HIPC(c, hipMemcpyAsync(c->axis_shell.p, S.axis_shell.data(), sizeof(int32_t) * S.axis_shell.size(),
                           hipMemcpyHostToDevice, c->stream));
  }
  HIPC(c, hipStreamSynchronize(c->stream));
  c->tile_index.clear();
  for (size_t t = 0; t < S.tiles.size(); ++t)
    c->tile_index[pack_tile(S.tiles[t].kx0, S.tiles[t].ky0, S.tiles[t].kz0)] = static_cast<int>(t);
  c->sq = std::move(S);
  c->sq_charge = sqq;
  c->sq_centres = (p->flags & TRJ_SQ_CENTRES) != 0;
  c->sq_ready = true;
  c->sq_frames = 0;
  if (info) *info = c->sq.info;
  return TRJ_OK;
}

int trj_frame_device(trj_ctx* c, const void* pos_dev, int32_t dtype, int32_t layout, const int32_t* itype_dev,
                     int64_t n, const double box[3], uint32_t what) {
  if (!c) return TRJ_EINVAL;
  if (!pos_dev || !itype_dev) return fail(c, TRJ_EINVAL, "NULL input");
  hipSetDevice(c->device);
  return run_frame(c, pos_dev, dtype, layout, itype_dev, n, box, what);
}

int trj_rdf_tuning(trj_ctx* c, int32_t* rows_per_pass, int32_t* passes, int32_t* waves_per_cell, double* ms,
                   int32_t* tried) {
  if (!c) return TRJ_EINVAL;
  if (!c->rdf_ready) return fail(c, TRJ_ESTATE, "trj_rdf_tuning before trj_rdf_init");
  const auto& tu = c->rdf_tune;
  const int rows = tu.done ? tu.rows : c->rdf.rows;
  if (rows_per_pass) *rows_per_pass = rows;
  if (passes) *passes = (c->rdf.nit + rows - 1) / rows;
  if (waves_per_cell) *waves_per_cell = tu.done ? tu.W : 0;
  if (ms) *ms = tu.ms;
  if (tried) *tried = tu.tried;
  return TRJ_OK;
}

int trj_rdf_cell_grid(trj_ctx* c, int32_t cells[3], double* side_over_cutoff) {
  if (!c) return TRJ_EINVAL;
  if (!c->rdf_ready) return fail(c, TRJ_ESTATE, "trj_rdf_cell_grid before trj_rdf_init");
  if (cells)
    for (int k = 0; k < 3; ++k) cells[k] = c->rdf_nc[k];
  if (side_over_cutoff) *side_over_cutoff = c->rdf_cell_scale;
  return TRJ_OK;
}

int trj_sq_tuning(trj_ctx* c, int32_t* splits, int32_t* atoms_per_split, int32_t* xcd_grouped, double* ms,
                  int32_t* tried) {
  if (!c) return TRJ_EINVAL;
  if (!c->sq_ready) return fail(c, TRJ_ESTATE, "trj_sq_tuning before trj_sq_init");
  const auto& st = c->sq_tune;
  if (splits) *splits = c->sq_frames > 0 ? c->ksplit : 0;
  if (atoms_per_split) *atoms_per_split = st.done ? st.span : 0;
  if (xcd_grouped) *xcd_grouped = c->sq.p.precision == TRJ_PREC_FP64 ? (st.done ? st.order : 0) : 1;
  if (ms) *ms = st.ms;
  if (tried) *tried = st.tried;
  return TRJ_OK;
}

int trj_frame_device_noalloc(trj_ctx* c, const void* pos_dev, int32_t dtype, int32_t layout,
                             const int32_t* itype_dev, int64_t n, const double box_in[3], uint32_t what) {
  if (!c) return TRJ_EINVAL;
  if (!pos_dev || !itype_dev) return fail(c, TRJ_EINVAL, "NULL input");
  if (what != TRJ_WANT_RDF) return fail(c, TRJ_EINVAL, "allocation-free frames compute the RDF only (what = TRJ_WANT_RDF)");
  if (!c->rdf_ready) return fail(c, TRJ_ESTATE, "frame asks for RDF before trj_rdf_init");
  if (c->rdf_charge || c->slices.n > 0 || c->nrand > 0)
    return fail(c, TRJ_EINVAL, "allocation-free frames take no charges, slices or sampling points");
  if (n < 1) return fail(c, TRJ_EINVAL, "n must be positive");
  const RdfPlan& R = c->rdf;
  const int ndim = R.ndim;
  if (int r = validate_box(c, ndim, box_in)) return r;
  hipSetDevice(c->device);
  std::string why;
  char* base = static_cast<char*>(scratch_address(c->device, &why));
  if (!base) return fail(c, TRJ_EDEVICE, why);
  // The buffers run_frame asks for on this frame, at the sizes it asks
  // (same geometry, same cell plan), each at a 256-byte offset.
  const double box[3] = {box_in[0], box_in[1], ndim == 3 ? box_in[2] : 1.0};
  double rc = R.rc;
  const FrameGeom g = frame_geometry(ndim, box, R.periodic, &rc);
  RdfCells cp = rdf_cell_plan(ndim, box, R.periodic, g.side2, n, c->rdf_tune.done ? c->rdf_tune.W : 0,
                              c->rdf_tune.done ? c->rdf_tune.scale
                                               : (c->rdf_fixed_scale > 0.0 ? c->rdf_fixed_scale : 1.0));
  if (cp.cg.reach == 2 && (!cp.cells_ok || c->nrand > 0))
    cp = rdf_cell_plan(ndim, box, R.periodic, g.side2, n, c->rdf_tune.done ? c->rdf_tune.W : 0, 1.0);
  const size_t nblk = static_cast<size_t>((n + 255) / 256);
  const size_t un = static_cast<size_t>(n), nc = static_cast<size_t>(cp.ncell);
  const size_t cells = cp.cells_ok ? 1 : 0;
  // The cell sort's own buffers: the LDS sort's per-workgroup cell counts, or
  // the atomic scatter's per-cell cursors where those do not fit.
  const CellSort cs = cp.cells_ok ? cell_sort_plan(n, cp.ncell, c->lds_max) : CellSort{};
  DevBuf* sw[] = {&c->pos, &c->blk, &c->blkoff, &c->cellid, &c->cellcount, &c->cellstart, &c->spos,
                  &c->stype, &c->partial, &c->cellblk, &c->cursor};
  size_t len[] = {24 * un, 4 * R.nsp * nblk, 4 * R.nsp * nblk, cells * 4 * un, cells * 4 * (nc + 1),
                  cells * 4 * (nc + 1), cells * 24 * un, cells * 4 * un,
                  4 * static_cast<size_t>(R.nit) * R.lsmax * static_cast<size_t>(cp.nblocks),
                  cs.lds ? 4 * static_cast<size_t>(cs.blocks) * nc : 0, cs.lds ? 0 : cells * 4 * nc};
  constexpr int nb = sizeof(len) / sizeof(len[0]);
  size_t off[nb], total = 0;
  auto lay_out = [&]() {
    total = 0;
    for (int k = 0; k < nb; ++k) {
      off[k] = total;
      total += (len[k] + 255) & ~size_t(255);
    }
  };
  lay_out();
  // 64 KiB stay free at the end of the array as a guard.
  bool atomic_cells = false;
  if (cs.lds && total + (size_t(64) << 10) > kScratchBytes) {
    len[nb - 2] = 0;
    len[nb - 1] = cells * 4 * nc;
    atomic_cells = true;
    lay_out();
  }
  if (total + (size_t(64) << 10) > kScratchBytes)
    return fail(c, TRJ_ELIMIT, "allocation-free frame needs " + std::to_string(total) +
                                   " bytes of scratch; the library holds " + std::to_string(kScratchBytes) +
                                   " (64 KiB of it kept free)");
  DevBuf saved[nb];
  for (int k = 0; k < nb; ++k) {
    saved[k] = *sw[k];
    sw[k]->p = base + off[k];
    sw[k]->cap = len[k];
  }
  c->noalloc = true;
  c->atomic_cells = atomic_cells;
  const int r = run_frame(c, pos_dev, dtype, layout, itype_dev, n, box_in, TRJ_WANT_RDF);
  c->noalloc = false;
  c->atomic_cells = false;
  for (int k = 0; k < nb; ++k) *sw[k] = saved[k];
  return r;
}

}  // extern "C"

namespace {

// Host-input frame: stage positions, types and (optional) charges through
// pinned memory, then run the frame from the device copies.
int frame_host(trj_ctx* c, const void* pos, int32_t dtype, int32_t layout, const int32_t* itype,
               const double* charge, int64_t n, const double box[3], uint32_t what) {
  if (!c) return TRJ_EINVAL;
  if (!pos || !itype) return fail(c, TRJ_EINVAL, "NULL input");
  if (n < 1) return fail(c, TRJ_EINVAL, "n must be positive");
  if (dtype != TRJ_F32 && dtype != TRJ_F64) return fail(c, TRJ_EINVAL, "bad dtype");
  hipSetDevice(c->device);
  int ndim = 0, nspf = 0;
  if (int r0 = frame_shape(c, what, &ndim, &nspf)) return r0;
  const size_t pbytes = static_cast<size_t>(n) * ndim * (dtype == TRJ_F32 ? 4 : 8);
  const size_t tbytes = static_cast<size_t>(n) * sizeof(int32_t);
  const size_t qbytes = charge ? static_cast<size_t>(n) * sizeof(double) : 0;
  const size_t pbytes_al = (pbytes + 255) & ~size_t(255);
  const size_t tbytes_al = (tbytes + 255) & ~size_t(255);
  const size_t qbytes_al = (qbytes + 255) & ~size_t(255);
  const size_t total = pbytes_al + tbytes_al + qbytes_al + (charge ? 2 * sizeof(double) : 0);
  int r;
  if ((r = ensure(c, c->raw, pbytes))) return r;
  if ((r = ensure(c, c->itype, tbytes))) return r;
  if (charge && (r = ensure(c, c->qraw, qbytes))) return r;
  if (charge && (r = ensure(c, c->qparm, 4 * sizeof(double)))) return r;
  // Fixed-point scale of q_i q_j (2^(30-e), max q^2 < 2^e) and avq2 =
  // sum(q^2)/n, both from the host copy in index order.
  double qscale = 1.0, avq2 = 0.0;
  // Charges constant within each species (per-type charges): S(Q) then needs
  // no charge channel, rho_q = sum_s q_s rho_s.
  double qsp[kNspMax];
  bool by_species = charge != nullptr && nspf > 0;
  bool seen[kNspMax] = {false};
  if (charge) {
    double m = 0.0, s2 = 0.0;
    for (int64_t i = 0; i < n; ++i) {
      const double v = charge[i] * charge[i];
      if (!std::isfinite(v)) return fail(c, TRJ_EINVAL, "non-finite charge");
      if (v > m) m = v;
      s2 += v;
      if (by_species) {
        const int t = itype[i] - 1;
        if (t < 0 || t >= nspf) {
          by_species = false;
        } else if (!seen[t]) {
          seen[t] = true;
          qsp[t] = charge[i];
        } else if (charge[i] != qsp[t]) {
          by_species = false;
        }
      }
    }
    if (by_species)
      for (int t = 0; t < nspf; ++t)
        if (!seen[t]) qsp[t] = 0.0;
    int e = 0;
    if (m > 0.0) std::frexp(m, &e);
    qscale = std::ldexp(1.0, 30 - e);
    avq2 = s2 / static_cast<double>(n);
  }
  // Stage through pinned memory; wait until the previous copy out of this
  // staging buffer has completed.
  const int k = c->pinned_next;
  c->pinned_next ^= 1;
  HIPC(c, hipEventSynchronize(c->pinned_done[k]));
  if (c->pinned_cap[k] < total) {
    if (c->pinned[k]) hipHostFree(c->pinned[k]);
    c->pinned[k] = nullptr;
    c->pinned_cap[k] = 0;
    size_t want = total + total / 8;
    if (hipHostMalloc(&c->pinned[k], want, hipHostMallocDefault) != hipSuccess)
      return fail(c, TRJ_ENOMEM, "pinned staging allocation failed");
    c->pinned_cap[k] = want;
  }
  char* st = static_cast<char*>(c->pinned[k]);
  std::memcpy(st, pos, pbytes);
  std::memcpy(st + pbytes_al, itype, tbytes);
  if (charge) {
    std::memcpy(st + pbytes_al + tbytes_al, charge, qbytes);
    const double qp[2] = {qscale, avq2};
    std::memcpy(st + pbytes_al + tbytes_al + qbytes_al, qp, sizeof(qp));
  }
  HIPC(c, hipMemcpyAsync(c->raw.p, st, pbytes, hipMemcpyHostToDevice, c->stream));
  HIPC(c, hipMemcpyAsync(c->itype.p, st + pbytes_al, tbytes, hipMemcpyHostToDevice, c->stream));
  if (charge) {
    HIPC(c, hipMemcpyAsync(c->qraw.p, st + pbytes_al + tbytes_al, qbytes, hipMemcpyHostToDevice, c->stream));
    HIPC(c, hipMemcpyAsync(c->qparm.p, st + pbytes_al + tbytes_al + qbytes_al, 2 * sizeof(double),
                           hipMemcpyHostToDevice, c->stream));
  }
  HIPC(c, hipEventRecord(c->pinned_done[k], c->stream));
  return run_frame(c, c->raw.p, dtype, layout, P<int32_t>(c->itype), n, box, what,
                   charge ? P<double>(c->qraw) : nullptr, by_species ? qsp : nullptr, qscale);
}

}  // namespace

extern "C" {

int trj_frame(trj_ctx* c, const void* pos, int32_t dtype, int32_t layout, const int32_t* itype, int64_t n,
              const double box[3], uint32_t what) {
  return frame_host(c, pos, dtype, layout, itype, nullptr, n, box, what);
}

int trj_frame_q(trj_ctx* c, const void* pos, int32_t dtype, int32_t layout, const int32_t* itype,
                const double* charge, int64_t n, const double box[3], uint32_t what) {
  if (!c) return TRJ_EINVAL;
  if (!charge) return fail(c, TRJ_EINVAL, "NULL charges");
  return frame_host(c, pos, dtype, layout, itype, charge, n, box, what);
}

int trj_frame_device_q(trj_ctx* c, const void* pos_dev, int32_t dtype, int32_t layout, const int32_t* itype_dev,
                       const double* charge_dev, int64_t n, const double box[3], uint32_t what) {
  if (!c) return TRJ_EINVAL;
  if (!pos_dev || !itype_dev || !charge_dev) return fail(c, TRJ_EINVAL, "NULL input");
  if (n < 1) return fail(c, TRJ_EINVAL, "n must be positive");
  hipSetDevice(c->device);
  int r;
  if ((r = ensure(c, c->qparm, 4 * sizeof(double)))) return r;
  if ((r = ensure(c, c->flags, sizeof(uint32_t)))) return r;
  launch_qq_params(charge_dev, n, P<double>(c->qparm), P<uint32_t>(c->flags), c->stream);
  HIPC(c, hipGetLastError());
  return run_frame(c, pos_dev, dtype, layout, itype_dev, n, box, what, charge_dev);
}

int trj_sync(trj_ctx* c) {
  if (!c) return TRJ_EINVAL;
  hipSetDevice(c->device);
  return sync_all(c);
}

int trj_rdf_counts(trj_ctx* c, int64_t* counts) {
  if (!c || !counts) return TRJ_EINVAL;
  if (c->slices.n > 0) return fail(c, TRJ_ESTATE, "trj_rdf_counts: the context analyses z-slices (trj_set_slices)");
  if (!c->rdf_ready) return fail(c, TRJ_ESTATE, "RDF not initialised");
  hipSetDevice(c->device);
  int r = sync_all(c);
  if (r) return r;
  HIPC(c, hipMemcpy(counts, c->counts.p, sizeof(int64_t) * c->rdf.nit * c->rdf.lsmax, hipMemcpyDeviceToHost));
  return TRJ_OK;
}

int trj_rdf_gmix(trj_ctx* c, int64_t nconf, double* out) {
  if (!c || !out) return TRJ_EINVAL;
  if (c->slices.n > 0) return fail(c, TRJ_ESTATE, "trj_rdf_gmix: the context analyses z-slices (trj_set_slices)");
  if (!c->rdf_ready) return fail(c, TRJ_ESTATE, "RDF not initialised");
  if (c->rdf_frames == 0) return fail(c, TRJ_ESTATE, "no RDF frame processed");
  const RdfPlan& R = c->rdf;
  std::vector<int64_t> counts(static_cast<size_t>(R.nit) * R.lsmax);
  int r = trj_rdf_counts(c, counts.data());
  if (r) return r;
  std::vector<int64_t> ntype;
  if ((r = read_ntype(c, ntype))) return r;
  if (nconf <= 0) nconf = c->rdf_frames;
  // log.f90:484-518
  const double pi = 3.14159265358979323846;
  const int ncol = 1 + R.nit;
  for (int i = 1; i <= R.lsmax - 2; ++i) {
    const double ri = i * R.deltar, a = ri + R.deltar / 2, b = ri - R.deltar / 2;
    const double dv = (R.ndim == 3) ? 4.0 * pi * (a * a * a - b * b * b) / 3.0 : pi * (a * a - b * b);
    double* row = out + static_cast<size_t>(i - 1) * ncol;
    row[0] = ri;
    int col = 1;
    for (int j = 1; j <= R.nsp; ++j)
      for (int l = j; l <= R.nsp; ++l) {
        const int ij = (j - 1) * (R.nsp - 1) - (j - 1) * (j - 2) / 2 + l;
        const double h = static_cast<double>(counts[static_cast<size_t>(ij - 1) * R.lsmax + (i - 1)]);
        const double fac = (j == l) ? 2.0 : 1.0;  // (j/l + 1), integer division
        row[col++] = fac * c->last_volume * h /
                     (dv * static_cast<double>(ntype[l - 1]) * static_cast<double>(ntype[j - 1]) *
                      static_cast<double>(nconf));
      }
  }
  return TRJ_OK;
}

int trj_rdf_qq(trj_ctx* c, double* acc) {
  if (!c || !acc) return TRJ_EINVAL;
  if (c->slices.n > 0) return fail(c, TRJ_ESTATE, "trj_rdf_qq: the context analyses z-slices (trj_set_slices)");
  if (!c->rdf_ready || !c->rdf_charge) return fail(c, TRJ_ESTATE, "RDF not initialised with TRJ_RDF_CHARGE");
  hipSetDevice(c->device);
  int r = sync_all(c);
  if (r) return r;
  HIPC(c, hipMemcpy(acc, c->gqacc.p, sizeof(double) * c->rdf.lsmax, hipMemcpyDeviceToHost));
  return TRJ_OK;
}

int trj_rdf_gqq(trj_ctx* c, int64_t nconf, double* out) {
  if (!c || !out) return TRJ_EINVAL;
  if (c->slices.n > 0) return fail(c, TRJ_ESTATE, "trj_rdf_gqq: the context analyses z-slices (trj_set_slices)");
  if (!c->rdf_ready || !c->rdf_charge) return fail(c, TRJ_ESTATE, "RDF not initialised with TRJ_RDF_CHARGE");
  if (c->rdf_frames == 0) return fail(c, TRJ_ESTATE, "no RDF frame processed");
  const RdfPlan& R = c->rdf;
  std::vector<double> acc(R.lsmax);
  int r = trj_rdf_qq(c, acc.data());
  if (r) return r;
  if (nconf <= 0) nconf = c->rdf_frames;
  // log.f90:501-503: gchch = 2 V gqq(i) / (N^2 dV Nconf)
  const double pi = 3.14159265358979323846;
  const double N = static_cast<double>(c->last_n);
  for (int i = 1; i <= R.lsmax - 2; ++i) {
    const double ri = i * R.deltar, a = ri + R.deltar / 2, b = ri - R.deltar / 2;
    const double dv = (R.ndim == 3) ? 4.0 * pi * (a * a * a - b * b * b) / 3.0 : pi * (a * a - b * b);
    out[i - 1] = 2.0 * c->last_volume * acc[i - 1] / (N * N * dv * static_cast<double>(nconf));
  }
  return TRJ_OK;
}

int trj_rdf_first_minimum(trj_ctx* c, double* rmin) {
  if (!c || !rmin) return TRJ_EINVAL;
  if (c->slices.n > 0) return fail(c, TRJ_ESTATE, "trj_rdf_first_minimum: the context analyses z-slices (trj_set_slices)");
  if (!c->rdf_ready || c->rdf_frames == 0) return fail(c, TRJ_ESTATE, "no RDF frame processed");
  const RdfPlan& R = c->rdf;
  std::vector<int64_t> counts(static_cast<size_t>(R.nit) * R.lsmax);
  int r = trj_rdf_counts(c, counts.data());
  if (r) return r;
  // rdf.cuf:313-342: running total g over all species pairs, natms^2 normalisation.
  const double pi = 3.14159265358979323846;
  const int nrdf = R.lsmax;
  std::vector<double> g(nrdf + 1, 0.0);
  const double nat = static_cast<double>(c->last_n);
  for (int i = 1; i <= nrdf; ++i) {
    const double ri = i * R.deltar, a = ri + R.deltar / 2, b = ri - R.deltar / 2;
    const double dv = (R.ndim == 3) ? 4.0 * pi * (a * a * a - b * b * b) / 3.0 : pi * (a * a - b * b);
    double s = 0.0;
    for (int ij = 0; ij < R.nit; ++ij) s += static_cast<double>(counts[static_cast<size_t>(ij) * R.lsmax + i - 1]);
    g[i] = 2.0 * c->last_volume * s / (dv * nat * nat);
  }
  // rdf.cuf:755-822
  int imax = 1;
  for (int i = 2; i <= nrdf - 1; ++i)
    if (g[i] > 0.1 && g[i] > g[i - 1] && g[i] > g[i + 1]) {
      imax = i;
      break;
    }
  if (imax == 1) {
    double gmax = g[1];
    for (int i = 2; i <= nrdf; ++i)
      if (g[i] > gmax) {
        gmax = g[i];
        imax = i;
      }
  }
  int imin = imax + 1;
  for (int i = imax + 1; i <= nrdf - 1; ++i)
    if (g[i] < g[i - 1] && g[i] < g[i + 1]) {
      imin = i;
      break;
    }
  if (imin > 0 && imin < nrdf) {
    const double x1 = R.deltar * (imin - 1), x2 = R.deltar * imin, x3 = R.deltar * (imin + 1);
    const double y1 = g[imin - 1], y2 = g[imin], y3 = g[imin + 1];
    const double denom = y1 - 2.0 * y2 + y3;
    *rmin = (std::fabs(denom) > 1e-12) ? x2 + 0.5 * (y1 - y3) / denom * (x3 - x1) / 2.0 : x2;
  } else {
    *rmin = R.deltar * imin;
  }
  return TRJ_OK;
}

int trj_sq_shells(trj_ctx* c, double* sqf, double* sqfp, int64_t* nq) {
  if (!c) return TRJ_EINVAL;
  if (c->slices.n > 0) return fail(c, TRJ_ESTATE, "trj_sq_shells: the context analyses z-slices (trj_set_slices)");
  if (!c->sq_ready) return fail(c, TRJ_ESTATE, "S(Q) not initialised");
  hipSetDevice(c->device);
  int r = sync_all(c);
  if (r) return r;
  const int ns = c->sq.nqmax + 1;
  if (sqf) HIPC(c, hipMemcpy(sqf, c->sqf.p, sizeof(double) * ns, hipMemcpyDeviceToHost));
  if (sqfp) HIPC(c, hipMemcpy(sqfp, c->sqfp.p, sizeof(double) * ns * c->sq.p.nsp, hipMemcpyDeviceToHost));
  if (nq)
    for (int s = 0; s < ns; ++s) nq[s] = c->sq.nq[s];
  return TRJ_OK;
}

int trj_sq_qq(trj_ctx* c, double* sqfq) {
  if (!c || !sqfq) return TRJ_EINVAL;
  if (c->slices.n > 0) return fail(c, TRJ_ESTATE, "trj_sq_qq: the context analyses z-slices (trj_set_slices)");
  if (!c->sq_ready || !c->sq_charge) return fail(c, TRJ_ESTATE, "S(Q) not initialised with TRJ_SQ_CHARGE");
  hipSetDevice(c->device);
  int r = sync_all(c);
  if (r) return r;
  HIPC(c, hipMemcpy(sqfq, c->sqfq.p, sizeof(double) * (c->sq.nqmax + 1), hipMemcpyDeviceToHost));
  return TRJ_OK;
}

int trj_sq_qq_column(trj_ctx* c, int64_t nconf, double* out) {
  if (!c || !out) return TRJ_EINVAL;
  if (c->slices.n > 0) return fail(c, TRJ_ESTATE, "trj_sq_qq_column: the context analyses z-slices (trj_set_slices)");
  if (!c->sq_ready || !c->sq_charge) return fail(c, TRJ_ESTATE, "S(Q) not initialised with TRJ_SQ_CHARGE");
  if (c->sq_frames == 0) return fail(c, TRJ_ESTATE, "no S(Q) frame processed");
  const SqPlan& S = c->sq;
  std::vector<double> q(S.nqmax + 1);
  int r = trj_sq_qq(c, q.data());
  if (r) return r;
  if (nconf <= 0) nconf = c->sq_frames;
  // sq.cuf:143-145 divides the accumulated sums by avq2 of the last frame
  // read; log.f90:371-376 then writes sqfq(i) / (Nsites Nconf nq(i)).
  double av = 0.0;
  HIPC(c, hipMemcpy(&av, P<double>(c->qparm) + 2, sizeof(double), hipMemcpyDeviceToHost));
  const double N = static_cast<double>(c->last_n);
  for (int i = 1; i <= S.nqmax; ++i) {
    const double cn = static_cast<double>(nconf) * static_cast<double>(S.nq[i]);
    out[i - 1] = (cn > 0.0 && av > 0.0) ? (q[i] / av) / (N * cn) : 0.0;
  }
  return TRJ_OK;
}

int trj_sq_columns(trj_ctx* c, int64_t nconf, double* cols, int32_t* ncols_out, double* mix) {
  if (!c || !cols) return TRJ_EINVAL;
  if (c->slices.n > 0) return fail(c, TRJ_ESTATE, "trj_sq_columns: the context analyses z-slices (trj_set_slices)");
  if (!c->sq_ready || c->sq_frames == 0) return fail(c, TRJ_ESTATE, "no S(Q) frame processed");
  const SqPlan& S = c->sq;
  const int nsp = S.p.nsp, ns = S.nqmax + 1;
  std::vector<double> sqf(ns), sqfp(static_cast<size_t>(ns) * nsp);
  int r = trj_sq_shells(c, sqf.data(), sqfp.data(), nullptr);
  if (r) return r;
  std::vector<int64_t> ntype;
  if ((r = read_ntype(c, ntype))) return r;
  if (nconf <= 0) nconf = c->sq_frames;
  // log.f90:317-401
  const double N = static_cast<double>(c->last_n);
  double baver = 0.0;
  bool bsc_one = true;
  for (int s = 0; s < nsp; ++s) {
    baver += static_cast<double>(ntype[s]) * S.p.bsc[s];
    if (std::fabs(S.p.bsc[s] - 1.0) > 1.0e-4) bsc_one = false;
  }
  baver /= N;
  const bool binary = (nsp == 2 && bsc_one);
  const int ncols = binary ? 7 : 3;
  if (ncols_out) *ncols_out = ncols;
  const double x1 = binary ? ntype[0] / N : 0.0, x2 = binary ? ntype[1] / N : 0.0;
  for (int i = 1; i <= S.nqmax; ++i) {
    double* row = cols + static_cast<size_t>(i - 1) * ncols;
    const double nqi = static_cast<double>(S.nq[i]);
    const double cn = static_cast<double>(nconf) * nqi;
    const double snn = (sqf[i] / (baver * baver)) / (N * cn);
    row[0] = static_cast<double>(static_cast<float>(i) * S.dq);
    row[1] = snn;
    if (binary) {
      const double s11 = x1 * sqfp[i] / (ntype[0] * cn);
      const double s22 = x2 * sqfp[ns + i] / (ntype[1] * cn);
      const double s12 = 0.5 * (snn - s11 - s22);
      row[2] = x2 * x2 * s11 + x1 * x1 * s22 - 2.0 * x1 * x2 * s12;
      row[3] = s11;
      row[4] = s22;
      row[5] = s12;
      row[6] = nqi;
    } else {
      row[2] = nqi;
    }
    if (mix) {
      double* m = mix + static_cast<size_t>(i - 1) * (1 + nsp);
      m[0] = row[0];
      for (int s = 0; s < nsp; ++s) m[1 + s] = sqfp[static_cast<size_t>(s) * ns + i] / (ntype[s] * cn);
    }
  }
  return TRJ_OK;
}

int trj_sq_rho(trj_ctx* c, int64_t nvec, const int32_t* kv, double* rho) {
  if (!c || (nvec > 0 && (!kv || !rho))) return TRJ_EINVAL;
  if (c->slices.n > 0) return fail(c, TRJ_ESTATE, "trj_sq_rho: the context analyses z-slices (trj_set_slices)");
  if (!c->sq_ready || c->sq_frames == 0) return fail(c, TRJ_ESTATE, "no S(Q) frame processed");
  hipSetDevice(c->device);
  int r = sync_all(c);
  if (r) return r;
  const SqPlan& S = c->sq;
  const int nsp = S.p.nsp, ntiles = static_cast<int>(S.tiles.size());
  const int nch = c->sq_nch;  // channels of the last frame's phase sums
  // Many vectors: one copy of the frame's phase sums instead of one per value.
  // The partials are FP64 complex (FP64 contraction) or FP32 complex.
  const bool f32 = S.p.precision != TRJ_PREC_FP64;
  const size_t esz = f32 ? sizeof(float) : sizeof(double);
  std::vector<unsigned char> all;
  if (nvec > 16) {
    all.resize(esz * 2 * c->ksplit * nch * ntiles * S.tk.vec());
    HIPC(c, hipMemcpy(all.data(), c->rho.p, all.size(), hipMemcpyDeviceToHost));
  }
  for (int64_t v = 0; v < nvec; ++v) {
    const int x = kv[3 * v], y = kv[3 * v + 1], z = kv[3 * v + 2];
    if (x < S.kmin[0] || y < S.kmin[1] || z < S.kmin[2])
      return fail(c, TRJ_EINVAL, "vector outside the dense q-box");
    const int TKX = S.tk.tx, TKY = S.tk.ty, TKZ = S.tk.tz, TILE_VEC = S.tk.vec();
    const int x0 = S.kmin[0] + ((x - S.kmin[0]) / TKX) * TKX;
    const int y0 = S.kmin[1] + ((y - S.kmin[1]) / TKY) * TKY;
    const int z0 = S.kmin[2] + ((z - S.kmin[2]) / TKZ) * TKZ;
    auto it = c->tile_index.find(pack_tile(x0, y0, z0));
    if (it == c->tile_index.end()) return fail(c, TRJ_EINVAL, "vector not in the dense q-set");
    const int tile = it->second;
    const int vi = ((x - x0) * TKY + (y - y0)) * TKZ + (z - z0);
    if (S.tile_shell[static_cast<size_t>(tile) * TILE_VEC + vi] < 0)
      return fail(c, TRJ_EINVAL, "vector not in the dense q-set");
    for (int s = 0; s < nsp; ++s) {
      // Neumaier's compensated sum over the splits, as the shells kernels add them
      double acc[2] = {0.0, 0.0}, cmp[2] = {0.0, 0.0};
      for (int sp = 0; sp < c->ksplit; ++sp) {
        unsigned char raw[2 * sizeof(double)];
        const size_t idx = (((static_cast<size_t>(sp) * nch + s) * ntiles + tile) * TILE_VEC + vi) * 2;
        if (!all.empty()) {
          std::memcpy(raw, all.data() + idx * esz, 2 * esz);
        } else {
          HIPC(c, hipMemcpy(raw, static_cast<const unsigned char*>(c->rho.p) + idx * esz, 2 * esz,
                            hipMemcpyDeviceToHost));
        }
        double val[2];
        if (f32) {
          float fv[2];
          std::memcpy(fv, raw, sizeof(fv));
          val[0] = fv[0];
          val[1] = fv[1];
        } else {
          std::memcpy(val, raw, sizeof(val));
        }
        for (int k = 0; k < 2; ++k) {
          const double t = acc[k] + val[k];
          cmp[k] += (std::fabs(acc[k]) >= std::fabs(val[k])) ? (acc[k] - t) + val[k] : (val[k] - t) + acc[k];
          acc[k] = t;
        }
      }
      rho[(v * nsp + s) * 2] = acc[0] + cmp[0];
      rho[(v * nsp + s) * 2 + 1] = acc[1] + cmp[1];
    }
  }
  return TRJ_OK;
}

// ------------------------------------------------------------------ F(Q,t)

int trj_fqt_init(trj_ctx* c, const trj_fqt_params* p, const double box[3], int32_t* nvec_out, int64_t* nw_out) {
  if (!c || !p || !box) return TRJ_EINVAL;
  if (p->ndim != 2 && p->ndim != 3) return fail(c, TRJ_EINVAL, "fqt: ndim must be 2 or 3");
  if (p->nsp < 1 || p->nsp > kNspMax) return fail(c, TRJ_EINVAL, "fqt: nsp must be in 1..10");
  if (p->nqw < 1 || p->nqw > 64 || !p->qw) return fail(c, TRJ_EINVAL, "fqt: nqw must be in 1..64");
  if (p->nbuffer < 1 || p->nbuffer > kFqtMaxBuffers) return fail(c, TRJ_EINVAL, "fqt: nbuffer must be in 1..64");
  if (p->jump < 1 || p->ndyn < 1 || p->itlimit < 1) return fail(c, TRJ_EINVAL, "fqt: jump, ndyn, itlimit must be >= 1");
  if (int r = validate_box(c, p->ndim, box)) return r;
  hipSetDevice(c->device);
  int r = sync_all(c);
  if (r) return r;
  // The S(Q,w) q-set (sq.cuf:234-282) in FP32: (i, j, k) >= 0, not all 0, up
  // to nint(qmm/fk) per axis, assigned to the first qw(l) with
  // qw(l) - dq/2 < |q| < qw(l) + dq/2.  The reference's first (counting)
  // loop keeps k = 0 unless confined while its second (storing) loop runs k
  // over 0..nint(qmm/fk3); the storing loop's set is used here.
  const int ndim = p->ndim;
  float fk[3], side[3];
  for (int k = 0; k < 3; ++k) {
    side[k] = static_cast<float>(k < ndim ? box[k] : 1.0);
    fk[k] = (2.0f * 3.1415926535898f) / side[k];
  }
  float dq = std::fmin(fk[0], fk[1]);
  if (ndim == 3) dq = std::fmin(dq, fk[2]);
  std::vector<float> qw(p->nqw);
  float qmm = 0.0f;
  for (int l = 0; l < p->nqw; ++l) {
    qw[l] = static_cast<float>(p->qw[l]);
    qmm = std::fmax(qmm, qw[l]);
  }
  const int imax = static_cast<int>(std::lround(qmm / fk[0])), jmax = static_cast<int>(std::lround(qmm / fk[1]));
  const int kmax = ndim == 3 ? static_cast<int>(std::lround(qmm / fk[2])) : 0;
  std::vector<std::vector<int32_t>> bins(p->nqw);
  for (int i = 0; i <= imax; ++i)
    for (int j = 0; j <= jmax; ++j)
      for (int k = 0; k <= kmax; ++k) {
        if (i + j + k == 0) continue;
        const float a = i * fk[0], b = j * fk[1], cc = k * fk[2];
        const float rr = std::sqrt(a * a + b * b + cc * cc);
        for (int l = 0; l < p->nqw; ++l)
          if (qw[l] - dq / 2 < rr && qw[l] + dq / 2 > rr) {
            bins[l].insert(bins[l].end(), {i, j, k});
            break;
          }
      }
  std::vector<int32_t> kv, off(p->nqw + 1, 0);
  c->fqt_nw.assign(p->nqw, 0);
  for (int l = 0; l < p->nqw; ++l) {
    kv.insert(kv.end(), bins[l].begin(), bins[l].end());
    c->fqt_nw[l] = static_cast<int64_t>(bins[l].size() / 3);
    off[l + 1] = off[l] + static_cast<int32_t>(bins[l].size() / 3);
  }
  const int nvec = off[p->nqw];
  const size_t nacc = static_cast<size_t>(p->nbuffer) * p->nqw * p->ndyn;
  if ((r = ensure(c, c->fqkv, sizeof(int32_t) * std::max<size_t>(kv.size(), 3)))) return r;
  if ((r = ensure(c, c->fqbin, sizeof(int32_t) * off.size()))) return r;
  if ((r = ensure(c, c->fqrho0, sizeof(double) * 2 * std::max<size_t>(1, static_cast<size_t>(nvec) * p->nbuffer)))) return r;
  if ((r = ensure(c, c->fqfkt, sizeof(double) * nacc))) return r;
  if ((r = ensure(c, c->fqfskt, sizeof(double) * nacc))) return r;
  if ((r = ensure(c, c->fqtot, sizeof(double) * p->nqw * p->ndyn))) return r;
  if ((r = ensure(c, c->fqtots, sizeof(double) * p->nqw * p->ndyn))) return r;
  if (!kv.empty()) HIPC(c, hipMemcpy(c->fqkv.p, kv.data(), sizeof(int32_t) * kv.size(), hipMemcpyHostToDevice));
  HIPC(c, hipMemcpy(c->fqbin.p, off.data(), sizeof(int32_t) * off.size(), hipMemcpyHostToDevice));
  HIPC(c, hipMemset(c->fqfkt.p, 0, sizeof(double) * nacc));
  HIPC(c, hipMemset(c->fqfskt.p, 0, sizeof(double) * nacc));
  HIPC(c, hipMemset(c->fqtot.p, 0, sizeof(double) * p->nqw * p->ndyn));
  HIPC(c, hipMemset(c->fqtots.p, 0, sizeof(double) * p->nqw * p->ndyn));
  c->fqt_nqw = p->nqw;
  c->fqt_nbuf = p->nbuffer;
  c->fqt_jump = p->jump;
  c->fqt_ndyn = p->ndyn;
  c->fqt_itlimit = p->itlimit;
  c->fqt_nvec = nvec;
  c->fqt_ndim = ndim;
  c->fqt_nsp = p->nsp;
  for (int k = 0; k < 3; ++k) c->fqt_invl[k] = 1.0 / (k < ndim ? box[k] : 1.0);
  for (int s = 0; s < kNspMax; ++s) c->fqt_bsc.b[s] = s < p->nsp ? p->bsc[s] : 0.0;
  c->fqt_origin.assign(p->nbuffer, 0);
  c->fqt_frames = 0;
  c->fqt_count_org = 0;
  c->fqt_ready = true;
  if (nvec_out) *nvec_out = nvec;
  if (nw_out)
    for (int l = 0; l < p->nqw; ++l) nw_out[l] = c->fqt_nw[l];
  return TRJ_OK;
}

int trj_fqt_results(trj_ctx* c, double* fkt, double* fskt, int64_t* count_org) {
  if (!c) return TRJ_EINVAL;
  if (!c->fqt_ready) return fail(c, TRJ_ESTATE, "F(Q,t) not initialised");
  hipSetDevice(c->device);
  int r = sync_all(c);
  if (r) return r;
  const size_t m = static_cast<size_t>(c->fqt_nqw) * c->fqt_ndyn;
  std::vector<double> a(m), b(m);
  HIPC(c, hipMemcpy(a.data(), c->fqtot.p, sizeof(double) * m, hipMemcpyDeviceToHost));
  HIPC(c, hipMemcpy(b.data(), c->fqtots.p, sizeof(double) * m, hipMemcpyDeviceToHost));
  if (count_org) *count_org = c->fqt_count_org;
  // print_rtcor (dynamics.cuf:268-272): / (Nsites count_org nw(i)); raw sums
  // when no origin has completed.
  const double N = static_cast<double>(c->fqt_n);
  for (int l = 0; l < c->fqt_nqw; ++l) {
    const double d = c->fqt_count_org > 0 && c->fqt_nw[l] > 0
                         ? N * static_cast<double>(c->fqt_count_org) * static_cast<double>(c->fqt_nw[l])
                         : 1.0;
    for (int t = 0; t < c->fqt_ndyn; ++t) {
      const size_t i = static_cast<size_t>(l) * c->fqt_ndyn + t;
      if (fkt) fkt[i] = a[i] / d;
      if (fskt) fskt[i] = b[i] / d;
    }
  }
  return TRJ_OK;
}

// ------------------------------------------------------ number fluctuations

namespace {
// SplitMix64 (Steele, Lea and Flood, OOPSLA 2014): the state advances by the
// golden-ratio increment and each output is its variant-13 mix.
uint64_t splitmix64(uint64_t& x) {
  uint64_t z = (x += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
}  // namespace

int trj_rdf_sampling(trj_ctx* c, int32_t nrandom, uint64_t seed) {
  if (!c) return TRJ_EINVAL;
  if (!c->rdf_ready) return fail(c, TRJ_ESTATE, "trj_rdf_sampling comes after trj_rdf_init");
  if (c->slices.n > 0) return fail(c, TRJ_EINVAL, "number fluctuations are not sampled with slices (rdf.cuf:226-245)");
  if (nrandom < 0 || nrandom > (1 << 20)) return fail(c, TRJ_EINVAL, "nrandom must be in 0..1048576");
  hipSetDevice(c->device);
  int r = sync_all(c);
  if (r) return r;
  c->nrand = 0;
  c->rpts_host.clear();
  if (nrandom == 0) return TRJ_OK;
  // Unit points in (0, 1], fixed for the run as the reference's per-block
  // seeds are (rdf.cuf:507-518); [3][nrandom] on the device.
  std::vector<double> dev(static_cast<size_t>(3) * nrandom);
  c->rpts_host.resize(static_cast<size_t>(3) * nrandom);
  uint64_t x = seed;
  for (int p = 0; p < nrandom; ++p)
    for (int k = 0; k < 3; ++k) {
      const double u = static_cast<double>((splitmix64(x) >> 11) + 1) * 0x1.0p-53;
      c->rpts_host[static_cast<size_t>(p) * 3 + k] = u;
      dev[static_cast<size_t>(k) * nrandom + p] = u;
    }
  const size_t nb = static_cast<size_t>(c->rdf.lsmax) + 1;
  if ((r = ensure(c, c->rpts, sizeof(double) * dev.size()))) return r;
  if ((r = ensure(c, c->s2gi, sizeof(unsigned long long) * nb))) return r;
  if ((r = ensure(c, c->s2g2i, sizeof(unsigned long long) * nb))) return r;
  if (s2n_lds_bytes(c->rdf.lsmax) > kPairLdsLimit) return fail(c, TRJ_ELIMIT, "rdf: lsmax too large for the sampling");
  HIPC(c, hipMemcpy(c->rpts.p, dev.data(), sizeof(double) * dev.size(), hipMemcpyHostToDevice));
  HIPC(c, hipMemset(c->s2gi.p, 0, sizeof(unsigned long long) * nb));
  HIPC(c, hipMemset(c->s2g2i.p, 0, sizeof(unsigned long long) * nb));
  c->nrand = nrandom;
  return TRJ_OK;
}

int trj_rdf_sampling_points(trj_ctx* c, double* pts) {
  if (!c || !pts) return TRJ_EINVAL;
  if (c->nrand == 0) return fail(c, TRJ_ESTATE, "no sampling points (trj_rdf_sampling)");
  std::memcpy(pts, c->rpts_host.data(), sizeof(double) * c->rpts_host.size());
  return TRJ_OK;
}

int trj_rdf_s2n_sums(trj_ctx* c, int64_t* gi, int64_t* g2i) {
  if (!c) return TRJ_EINVAL;
  if (c->nrand == 0) return fail(c, TRJ_ESTATE, "no sampling points (trj_rdf_sampling)");
  hipSetDevice(c->device);
  int r = sync_all(c);
  if (r) return r;
  const size_t nb = static_cast<size_t>(c->rdf.lsmax) + 1;
  if (gi) HIPC(c, hipMemcpy(gi, c->s2gi.p, sizeof(int64_t) * nb, hipMemcpyDeviceToHost));
  if (g2i) HIPC(c, hipMemcpy(g2i, c->s2g2i.p, sizeof(int64_t) * nb, hipMemcpyDeviceToHost));
  return TRJ_OK;
}

int trj_rdf_s2n(trj_ctx* c, int64_t nconf, double* out) {
  if (!c || !out) return TRJ_EINVAL;
  const int ls = c->rdf.lsmax;
  std::vector<int64_t> gi(static_cast<size_t>(ls) + 1), g2i(static_cast<size_t>(ls) + 1);
  int r = trj_rdf_s2n_sums(c, gi.data(), g2i.data());
  if (r) return r;
  if (c->rdf_frames == 0) return fail(c, TRJ_ESTATE, "no RDF frame processed");
  if (nconf <= 0) nconf = c->rdf_frames;
  // log.f90:526-533: g2i/(Nconf nrandom) - (gi/(Nconf nrandom))^2, i = 1..lsmax-2
  const double m = static_cast<double>(nconf) * static_cast<double>(c->nrand);
  for (int i = 1; i <= ls - 2; ++i) {
    const double a = static_cast<double>(gi[i]) / m;
    out[i - 1] = static_cast<double>(g2i[i]) / m - a * a;
  }
  return TRJ_OK;
}

// ---------------------------------------------------------------- slices

int trj_slice_counts(trj_ctx* c, int64_t* count, int64_t* countp) {
  if (!c) return TRJ_EINVAL;
  if (c->slices.n == 0) return fail(c, TRJ_ESTATE, "no slices (trj_set_slices)");
  if (c->rdf_frames == 0 && c->sq_frames == 0) return fail(c, TRJ_ESTATE, "no frame processed");
  hipSetDevice(c->device);
  int r = sync_all(c);
  if (r) return r;
  const int nsl = c->slices.n, nsp = c->last_nsp;
  std::vector<int32_t> a(nsl), b(static_cast<size_t>(nsl) * nsp);
  HIPC(c, hipMemcpy(a.data(), c->slcount.p, sizeof(int32_t) * nsl, hipMemcpyDeviceToHost));
  HIPC(c, hipMemcpy(b.data(), c->slcountp.p, sizeof(int32_t) * b.size(), hipMemcpyDeviceToHost));
  if (count)
    for (int k = 0; k < nsl; ++k) count[k] = a[k];
  if (countp)
    for (size_t k = 0; k < b.size(); ++k) countp[k] = b[k];
  return TRJ_OK;
}

int trj_rdf_xy(trj_ctx* c, double* hxy, int64_t* counts) {
  if (!c) return TRJ_EINVAL;
  if (!c->rdf_ready || c->slices.n == 0) return fail(c, TRJ_ESTATE, "RDF not initialised with slices");
  hipSetDevice(c->device);
  int r = sync_all(c);
  if (r) return r;
  const size_t nb = static_cast<size_t>(c->slices.n) * c->rdf.nit * c->rdf.lsmax;
  if (hxy) HIPC(c, hipMemcpy(hxy, c->xyacc.p, sizeof(double) * nb, hipMemcpyDeviceToHost));
  if (counts) HIPC(c, hipMemcpy(counts, c->xycounts.p, sizeof(int64_t) * nb, hipMemcpyDeviceToHost));
  return TRJ_OK;
}

int trj_rdf_xy_qq(trj_ctx* c, double* acc) {
  if (!c || !acc) return TRJ_EINVAL;
  if (!c->rdf_ready || c->slices.n == 0 || !c->rdf_charge)
    return fail(c, TRJ_ESTATE, "RDF not initialised with slices and TRJ_RDF_CHARGE");
  hipSetDevice(c->device);
  int r = sync_all(c);
  if (r) return r;
  HIPC(c, hipMemcpy(acc, c->xygqacc.p, sizeof(double) * c->slices.n * c->rdf.lsmax, hipMemcpyDeviceToHost));
  return TRJ_OK;
}

namespace {
// log.f90:494-496: cylindrical shell of r_i = i deltar times zgrid.
double xy_shell(double deltar, int i, double zgrid) {
  const double pi = 3.14159265358979323846;
  const double ri = i * deltar, a = ri + deltar / 2, b = ri - deltar / 2;
  return pi * (a * a - b * b) * zgrid;
}
}  // namespace

int trj_rdf_xy_gmix(trj_ctx* c, int64_t nconf, double* out) {
  if (!c || !out) return TRJ_EINVAL;
  if (!c->rdf_ready || c->slices.n == 0) return fail(c, TRJ_ESTATE, "RDF not initialised with slices");
  if (c->rdf_frames == 0) return fail(c, TRJ_ESTATE, "no RDF frame processed");
  const RdfPlan& R = c->rdf;
  const int nsl = c->slices.n, nsp = R.nsp;
  std::vector<double> h(static_cast<size_t>(nsl) * R.nit * R.lsmax);
  int r = trj_rdf_xy(c, h.data(), nullptr);
  if (r) return r;
  if (nconf <= 0) nconf = c->rdf_frames;
  // log.f90:511-514: (j/l + 1) V (zgrid/Lz) histomix_xy / (dV Nconf); the
  // integer j/l + 1 is 2 for j = l, 1 otherwise.
  const double zg = c->slices.zgrid, V = c->last_volume, fz = zg / c->last_box[2];
  int pair = 0;
  for (int j = 0; j < nsp; ++j)
    for (int l = j; l < nsp; ++l, ++pair)
      for (int i = 1; i <= R.lsmax - 2; ++i) {
        const double dv = xy_shell(R.deltar, i, zg);
        const double f = (j == l) ? 2.0 : 1.0;
        for (int sl = 0; sl < nsl; ++sl)
          out[(static_cast<int64_t>(pair) * (R.lsmax - 2) + i - 1) * nsl + sl] =
              f * V * fz * h[(static_cast<size_t>(sl) * R.nit + pair) * R.lsmax + i - 1] /
              (dv * static_cast<double>(nconf));
      }
  return TRJ_OK;
}

int trj_rdf_xy_gcc(trj_ctx* c, int64_t nconf, double* out) {
  if (!c || !out) return TRJ_EINVAL;
  const RdfPlan& R = c->rdf;
  const int nsl = c->slices.n;
  std::vector<double> g(static_cast<size_t>(std::max(nsl, 1)) * std::max(R.lsmax, 1));
  int r = trj_rdf_xy_qq(c, g.data());
  if (r) return r;
  if (c->rdf_frames == 0) return fail(c, TRJ_ESTATE, "no RDF frame processed");
  if (nconf <= 0) nconf = c->rdf_frames;
  // log.f90:520-524: V (zgrid/Lz) gqqxy / (dV Nconf)
  const double zg = c->slices.zgrid, V = c->last_volume, fz = zg / c->last_box[2];
  for (int i = 1; i <= R.lsmax - 2; ++i) {
    const double dv = xy_shell(R.deltar, i, zg);
    for (int sl = 0; sl < nsl; ++sl)
      out[static_cast<int64_t>(i - 1) * nsl + sl] =
          V * fz * g[static_cast<size_t>(sl) * R.lsmax + i - 1] / (dv * static_cast<double>(nconf));
  }
  return TRJ_OK;
}

int trj_sq_xy(trj_ctx* c, double* sqf, double* sqfp, double* sqfq) {
  if (!c) return TRJ_EINVAL;
  if (!c->sq_ready || c->slices.n == 0) return fail(c, TRJ_ESTATE, "S(Q) not initialised with slices");
  if (sqfq && !c->sq_charge) return fail(c, TRJ_ESTATE, "S(Q) not initialised with TRJ_SQ_CHARGE");
  hipSetDevice(c->device);
  int r = sync_all(c);
  if (r) return r;
  const size_t ns = static_cast<size_t>(c->sq.nqmax + 1) * c->slices.n;
  if (sqf) HIPC(c, hipMemcpy(sqf, c->sqf.p, sizeof(double) * ns, hipMemcpyDeviceToHost));
  if (sqfp) HIPC(c, hipMemcpy(sqfp, c->sqfp.p, sizeof(double) * ns * c->sq.p.nsp, hipMemcpyDeviceToHost));
  if (sqfq) HIPC(c, hipMemcpy(sqfq, c->sqfq.p, sizeof(double) * ns, hipMemcpyDeviceToHost));
  return TRJ_OK;
}

int trj_sq_xy_columns(trj_ctx* c, int64_t nconf, double* sqxy, double* sqpxy, double* sqqxy) {
  if (!c) return TRJ_EINVAL;
  if (!c->sq_ready || c->slices.n == 0) return fail(c, TRJ_ESTATE, "S(Q) not initialised with slices");
  if (c->sq_frames == 0) return fail(c, TRJ_ESTATE, "no S(Q) frame processed");
  if (sqqxy && !c->sq_charge) return fail(c, TRJ_ESTATE, "S(Q) not initialised with TRJ_SQ_CHARGE");
  const SqPlan& S = c->sq;
  const int nsl = c->slices.n, nsp = S.p.nsp, ns = S.nqmax + 1;
  std::vector<double> f(static_cast<size_t>(ns) * nsl), fp(static_cast<size_t>(ns) * nsl * nsp),
      fq(sqqxy ? static_cast<size_t>(ns) * nsl : 0);
  int r = trj_sq_xy(c, f.data(), fp.data(), sqqxy ? fq.data() : nullptr);
  if (r) return r;
  if (nconf <= 0) nconf = c->sq_frames;
  // sq.cuf:135-138: the charge sums divided by each slice's mean q^2 of the
  // last frame (q2slice, common.cuf:218-232).
  std::vector<double> q2(nsl, 0.0);
  if (sqqxy) {
    std::vector<double> qs(nsl);
    std::vector<int32_t> qc(nsl);
    HIPC(c, hipMemcpy(qs.data(), c->sqqsum.p, sizeof(double) * nsl, hipMemcpyDeviceToHost));
    HIPC(c, hipMemcpy(qc.data(), c->sqqcnt.p, sizeof(int32_t) * nsl, hipMemcpyDeviceToHost));
    for (int k = 0; k < nsl; ++k) q2[k] = qc[k] > 0 ? qs[k] / qc[k] : 0.0;
  }
  // log.f90:366-378: Q = i dq (FP32), values / (Nconf nq(i)); no 1/N (the
  // kernels divided each vector by the slice counts).
  for (int i = 1; i <= S.nqmax; ++i) {
    const double cn = static_cast<double>(nconf) * static_cast<double>(S.nq[i]);
    const double qv = static_cast<double>(static_cast<float>(i) * S.dq);
    if (sqxy) {
      double* row = sqxy + static_cast<int64_t>(i - 1) * (1 + nsl);
      row[0] = qv;
      for (int sl = 0; sl < nsl; ++sl) row[1 + sl] = cn > 0.0 ? f[static_cast<size_t>(sl) * ns + i] / cn : 0.0;
    }
    if (sqpxy)
      for (int s = 0; s < nsp; ++s) {
        double* row = sqpxy + (static_cast<int64_t>(s) * S.nqmax + i - 1) * (1 + nsl);
        row[0] = qv;
        for (int sl = 0; sl < nsl; ++sl)
          row[1 + sl] = cn > 0.0 ? fp[(static_cast<size_t>(sl) * nsp + s) * ns + i] / cn : 0.0;
      }
    if (sqqxy) {
      double* row = sqqxy + static_cast<int64_t>(i - 1) * (1 + nsl);
      row[0] = qv;
      for (int sl = 0; sl < nsl; ++sl)
        row[1 + sl] = (cn > 0.0 && q2[sl] > 0.0) ? (fq[static_cast<size_t>(sl) * ns + i] / q2[sl]) / cn : 0.0;
    }
  }
  return TRJ_OK;
}

int trj_reset(trj_ctx* c) {
  if (!c) return TRJ_EINVAL;
  hipSetDevice(c->device);
  int r = sync_all(c);
  if (r) return r;
  if (c->rdf_ready)
    HIPC(c, hipMemsetAsync(c->counts.p, 0, sizeof(int64_t) * c->rdf.nit * c->rdf.lsmax, c->stream));
  if (c->rdf_ready && c->rdf_charge) {
    HIPC(c, hipMemsetAsync(c->gqf.p, 0, sizeof(unsigned long long) * c->rdf.lsmax, c->stream));
    HIPC(c, hipMemsetAsync(c->gqacc.p, 0, sizeof(double) * c->rdf.lsmax, c->stream));
  }
  const int ngs = c->slices.n > 0 ? c->slices.n : 1;
  if (c->sq_ready) {
    const int ns = (c->sq.nqmax + 1) * ngs;
    HIPC(c, hipMemsetAsync(c->sqf.p, 0, sizeof(double) * ns, c->stream));
    HIPC(c, hipMemsetAsync(c->sqfp.p, 0, sizeof(double) * ns * c->sq.p.nsp, c->stream));
    if (c->sq_charge) HIPC(c, hipMemsetAsync(c->sqfq.p, 0, sizeof(double) * ns, c->stream));
  }
  if (c->fqt_ready) {
    const size_t m = static_cast<size_t>(c->fqt_nqw) * c->fqt_ndyn;
    HIPC(c, hipMemsetAsync(c->fqfkt.p, 0, sizeof(double) * m * c->fqt_nbuf, c->stream));
    HIPC(c, hipMemsetAsync(c->fqfskt.p, 0, sizeof(double) * m * c->fqt_nbuf, c->stream));
    HIPC(c, hipMemsetAsync(c->fqtot.p, 0, sizeof(double) * m, c->stream));
    HIPC(c, hipMemsetAsync(c->fqtots.p, 0, sizeof(double) * m, c->stream));
    c->fqt_frames = 0;
    c->fqt_count_org = 0;
  }
  if (c->nrand > 0) {
    HIPC(c, hipMemsetAsync(c->s2gi.p, 0, sizeof(unsigned long long) * (c->rdf.lsmax + 1), c->stream));
    HIPC(c, hipMemsetAsync(c->s2g2i.p, 0, sizeof(unsigned long long) * (c->rdf.lsmax + 1), c->stream));
  }
  if (c->rdf_ready && c->slices.n > 0) {
    const size_t nb = static_cast<size_t>(ngs) * c->rdf.nit * c->rdf.lsmax;
    HIPC(c, hipMemsetAsync(c->xycounts.p, 0, sizeof(int64_t) * nb, c->stream));
    HIPC(c, hipMemsetAsync(c->xyacc.p, 0, sizeof(double) * nb, c->stream));
    if (c->rdf_charge) HIPC(c, hipMemsetAsync(c->xygqacc.p, 0, sizeof(double) * ngs * c->rdf.lsmax, c->stream));
  }
  HIPC(c, hipStreamSynchronize(c->stream));
  c->rdf_frames = 0;
  c->sq_frames = 0;
  return TRJ_OK;
}

// ---------------------------------------------------------------- state blob
// Accumulated results of a context: RDF counts, S(Q) shell sums, frame counts
// and the last frame's volume and species counts (what the normalisations
// read), behind a header that identifies the plans they belong to.

namespace {
struct StateHeader {
  char magic[8];
  int32_t version, has_rdf, has_sq, nsp;
  int32_t rdf_ndim, nit, lsmax, rdf_periodic[3];
  double deltar, rc;
  int32_t sq_ndim, nqmax, qset, sq_charge;
  double dq, qbox[3];
  int64_t nvec_grid, nvec_axis;
  int64_t rdf_frames, sq_frames, last_n;
  double last_volume;
  int32_t last_nsp, rdf_charge;
  double sq_avq2;
  int64_t ntype[kNspMax];
};
constexpr char kStateMagic[8] = {'T', 'R', 'J', 'S', 'T', 'A', 'T', 'E'};

int64_t state_payload(const trj_ctx* c) {
  int64_t b = 0;
  if (c->rdf_ready) b += int64_t(8) * c->rdf.nit * c->rdf.lsmax;
  if (c->rdf_ready && c->rdf_charge) b += int64_t(8) * c->rdf.lsmax;
  if (c->sq_ready) b += int64_t(8) * (c->sq.nqmax + 1) * (1 + c->sq.p.nsp);
  if (c->sq_ready && c->sq_charge) b += int64_t(8) * (c->sq.nqmax + 1);
  return b;
}

StateHeader state_header(const trj_ctx* c) {
  StateHeader h{};
  std::memcpy(h.magic, kStateMagic, 8);
  h.version = 1;
  h.has_rdf = c->rdf_ready;
  h.has_sq = c->sq_ready;
  h.nsp = c->rdf_ready ? c->rdf.nsp : (c->sq_ready ? c->sq.p.nsp : 0);
  if (c->rdf_ready) {
    h.rdf_ndim = c->rdf.ndim;
    h.nit = c->rdf.nit;
    h.lsmax = c->rdf.lsmax;
    for (int k = 0; k < 3; ++k) h.rdf_periodic[k] = c->rdf.periodic[k];
    h.deltar = c->rdf.deltar;
    h.rc = c->rdf.rc;
    h.rdf_charge = c->rdf_charge;
  }
  if (c->sq_ready) {
    h.sq_ndim = c->sq.p.ndim;
    h.nqmax = c->sq.nqmax;
    h.qset = c->sq.p.qset;
    h.dq = c->sq.info.dq;
    for (int k = 0; k < 3; ++k) h.qbox[k] = c->sq.box[k];
    h.nvec_grid = c->sq.info.nvec_grid;
    h.nvec_axis = c->sq.info.nvec_axis;
    h.sq_charge = c->sq_charge;
  }
  h.rdf_frames = c->rdf_frames;
  h.sq_frames = c->sq_frames;
  h.last_n = c->last_n;
  h.last_volume = c->last_volume;
  h.last_nsp = c->last_nsp;
  return h;
}
}  // namespace

int64_t trj_state_bytes(const trj_ctx* c) {
  if (!c) return TRJ_EINVAL;
  if (c->slices.n > 0 || c->nrand > 0) return TRJ_ESTATE;
  return static_cast<int64_t>(sizeof(StateHeader)) + state_payload(c);
}

int trj_state_save(trj_ctx* c, void* buf, int64_t bytes) {
  if (!c || !buf) return TRJ_EINVAL;
  if (c->slices.n > 0) return fail(c, TRJ_ESTATE, "trj_state_save: the context analyses z-slices (trj_set_slices)");
  if (c->nrand > 0) return fail(c, TRJ_ESTATE, "trj_state_save: the state blob does not carry the number-fluctuation sums");
  if (bytes < trj_state_bytes(c)) return fail(c, TRJ_EINVAL, "state buffer too small");
  hipSetDevice(c->device);
  int r = sync_all(c);
  if (r) return r;
  StateHeader h = state_header(c);
  if (c->sq_ready && c->sq_charge)
    HIPC(c, hipMemcpy(&h.sq_avq2, P<double>(c->qparm) + 2, sizeof(double), hipMemcpyDeviceToHost));
  if (c->last_nsp > 0 && (c->rdf_frames > 0 || c->sq_frames > 0)) {
    std::vector<int64_t> nt;
    if ((r = read_ntype(c, nt))) return r;
    for (int k = 0; k < c->last_nsp; ++k) h.ntype[k] = nt[k];
  }
  unsigned char* p = static_cast<unsigned char*>(buf);
  std::memcpy(p, &h, sizeof(h));
  p += sizeof(h);
  if (c->rdf_ready) {
    const size_t b = sizeof(int64_t) * c->rdf.nit * c->rdf.lsmax;
    HIPC(c, hipMemcpy(p, c->counts.p, b, hipMemcpyDeviceToHost));
    p += b;
    if (c->rdf_charge) {
      HIPC(c, hipMemcpy(p, c->gqacc.p, sizeof(double) * c->rdf.lsmax, hipMemcpyDeviceToHost));
      p += sizeof(double) * c->rdf.lsmax;
    }
  }
  if (c->sq_ready) {
    const size_t ns = c->sq.nqmax + 1;
    HIPC(c, hipMemcpy(p, c->sqf.p, sizeof(double) * ns, hipMemcpyDeviceToHost));
    p += sizeof(double) * ns;
    HIPC(c, hipMemcpy(p, c->sqfp.p, sizeof(double) * ns * c->sq.p.nsp, hipMemcpyDeviceToHost));
    p += sizeof(double) * ns * c->sq.p.nsp;
    if (c->sq_charge) HIPC(c, hipMemcpy(p, c->sqfq.p, sizeof(double) * ns, hipMemcpyDeviceToHost));
  }
  return TRJ_OK;
}

int trj_state_load(trj_ctx* c, const void* buf, int64_t bytes, int mode) {
  if (!c || !buf) return TRJ_EINVAL;
  if (c->slices.n > 0) return fail(c, TRJ_ESTATE, "trj_state_load: the context analyses z-slices (trj_set_slices)");
  if (c->nrand > 0) return fail(c, TRJ_ESTATE, "trj_state_load: the state blob does not carry the number-fluctuation sums");
  if (mode != TRJ_STATE_REPLACE && mode != TRJ_STATE_ADD) return fail(c, TRJ_EINVAL, "unknown state load mode");
  if (bytes < static_cast<int64_t>(sizeof(StateHeader))) return fail(c, TRJ_EINVAL, "state blob too short");
  StateHeader h;
  std::memcpy(&h, buf, sizeof(h));
  if (std::memcmp(h.magic, kStateMagic, 8) != 0 || h.version != 1)
    return fail(c, TRJ_EINVAL, "not a trj_hip state blob (or another version)");
  const StateHeader m = state_header(c);
  const bool same = h.has_rdf == m.has_rdf && h.has_sq == m.has_sq && h.nsp == m.nsp && h.rdf_ndim == m.rdf_ndim &&
                    h.nit == m.nit && h.lsmax == m.lsmax && h.deltar == m.deltar && h.rc == m.rc &&
                    h.rdf_periodic[0] == m.rdf_periodic[0] && h.rdf_periodic[1] == m.rdf_periodic[1] &&
                    h.rdf_periodic[2] == m.rdf_periodic[2] && h.sq_ndim == m.sq_ndim && h.nqmax == m.nqmax &&
                    h.qset == m.qset && h.dq == m.dq && h.qbox[0] == m.qbox[0] && h.qbox[1] == m.qbox[1] &&
                    h.qbox[2] == m.qbox[2] && h.nvec_grid == m.nvec_grid && h.nvec_axis == m.nvec_axis &&
                    h.rdf_charge == m.rdf_charge && h.sq_charge == m.sq_charge;
  if (!same) return fail(c, TRJ_EINVAL, "state blob belongs to other RDF / S(Q) plans than this context's");
  if (bytes < trj_state_bytes(c)) return fail(c, TRJ_EINVAL, "state blob too short");
  if (h.last_nsp < 0 || h.last_nsp > kNspMax) return fail(c, TRJ_EINVAL, "corrupt state blob");
  hipSetDevice(c->device);
  int r = sync_all(c);
  if (r) return r;
  const unsigned char* p = static_cast<const unsigned char*>(buf) + sizeof(h);
  const bool add = mode == TRJ_STATE_ADD;
  if (c->rdf_ready) {
    const size_t nb = static_cast<size_t>(c->rdf.nit) * c->rdf.lsmax;
    std::vector<int64_t> v(nb), cur(add ? nb : 0);
    std::memcpy(v.data(), p, sizeof(int64_t) * nb);
    p += sizeof(int64_t) * nb;
    if (add) {
      HIPC(c, hipMemcpy(cur.data(), c->counts.p, sizeof(int64_t) * nb, hipMemcpyDeviceToHost));
      for (size_t k = 0; k < nb; ++k) v[k] += cur[k];
    }
    HIPC(c, hipMemcpy(c->counts.p, v.data(), sizeof(int64_t) * nb, hipMemcpyHostToDevice));
    if (c->rdf_charge) {
      const size_t nl = c->rdf.lsmax;
      std::vector<double> g(nl), gc(add ? nl : 0);
      std::memcpy(g.data(), p, sizeof(double) * nl);
      p += sizeof(double) * nl;
      if (add) {
        HIPC(c, hipMemcpy(gc.data(), c->gqacc.p, sizeof(double) * nl, hipMemcpyDeviceToHost));
        for (size_t k = 0; k < nl; ++k) g[k] += gc[k];
      }
      HIPC(c, hipMemcpy(c->gqacc.p, g.data(), sizeof(double) * nl, hipMemcpyHostToDevice));
    }
  }
  if (c->sq_ready) {
    const size_t nv = static_cast<size_t>(c->sq.nqmax + 1) * (1 + c->sq.p.nsp);
    std::vector<double> v(nv), cur(add ? nv : 0);
    std::memcpy(v.data(), p, sizeof(double) * nv);
    if (add) {
      const size_t ns = c->sq.nqmax + 1;
      HIPC(c, hipMemcpy(cur.data(), c->sqf.p, sizeof(double) * ns, hipMemcpyDeviceToHost));
      HIPC(c, hipMemcpy(cur.data() + ns, c->sqfp.p, sizeof(double) * (nv - ns), hipMemcpyDeviceToHost));
      for (size_t k = 0; k < nv; ++k) v[k] += cur[k];
    }
    const size_t ns = c->sq.nqmax + 1;
    HIPC(c, hipMemcpy(c->sqf.p, v.data(), sizeof(double) * ns, hipMemcpyHostToDevice));
    HIPC(c, hipMemcpy(c->sqfp.p, v.data() + ns, sizeof(double) * (nv - ns), hipMemcpyHostToDevice));
    if (c->sq_charge) {
      p += sizeof(double) * nv;
      std::vector<double> g(ns), gc(add ? ns : 0);
      std::memcpy(g.data(), p, sizeof(double) * ns);
      if (add) {
        HIPC(c, hipMemcpy(gc.data(), c->sqfq.p, sizeof(double) * ns, hipMemcpyDeviceToHost));
        for (size_t k = 0; k < ns; ++k) g[k] += gc[k];
      }
      HIPC(c, hipMemcpy(c->sqfq.p, g.data(), sizeof(double) * ns, hipMemcpyHostToDevice));
    }
  }
  // The normalisation inputs: the loaded ones unless this context has
  // processed frames of its own and is adding.
  const bool own = add && (c->rdf_frames > 0 || c->sq_frames > 0);
  c->rdf_frames = (add ? c->rdf_frames : 0) + h.rdf_frames;
  c->sq_frames = (add ? c->sq_frames : 0) + h.sq_frames;
  if (!own && h.last_nsp > 0) {
    c->last_n = h.last_n;
    c->last_volume = h.last_volume;
    c->last_nsp = h.last_nsp;
    if (c->sq_ready && c->sq_charge)
      HIPC(c, hipMemcpy(P<double>(c->qparm) + 2, &h.sq_avq2, sizeof(double), hipMemcpyHostToDevice));
    std::vector<int32_t> off(h.last_nsp + 1, 0);
    for (int k = 0; k < h.last_nsp; ++k) off[k + 1] = off[k] + static_cast<int32_t>(h.ntype[k]);
    if ((r = ensure(c, c->sp_off, sizeof(int32_t) * off.size()))) return r;
    HIPC(c, hipMemcpy(c->sp_off.p, off.data(), sizeof(int32_t) * off.size(), hipMemcpyHostToDevice));
  }
  return TRJ_OK;
}

int trj_lds_atomic_rate(int device, double* per_s) {
  if (!per_s) return TRJ_EINVAL;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return TRJ_EDEVICE;
  if (hipSetDevice(device) != hipSuccess) return TRJ_EDEVICE;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return TRJ_EDEVICE;
  uint32_t* out = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = TRJ_EDEVICE;
  if (hipMalloc(&out, sizeof(uint32_t)) == hipSuccess && hipEventCreate(&e0) == hipSuccess &&
      hipEventCreate(&e1) == hipSuccess) {
    const int nblocks = prop.multiProcessorCount * 8;
    const int iters = 2048;
    hipMemset(out, 0, sizeof(uint32_t));
    launch_lds_atomics(nblocks, 64, out, nullptr);  // warm-up
    hipEventRecord(e0, nullptr);
    const int64_t total = launch_lds_atomics(nblocks, iters, out, nullptr);
    hipEventRecord(e1, nullptr);
    if (hipEventSynchronize(e1) == hipSuccess) {
      float ms = 0.f;
      hipEventElapsedTime(&ms, e0, e1);
      if (ms > 0.f) {
        *per_s = static_cast<double>(total) / (ms * 1e-3);
        rc = TRJ_OK;
      }
    }
  }
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  if (out) hipFree(out);
  return rc;
}

int trj_coissue_rate(int device, int valu_per_mfma, int mode, double* cycles_per_mfma) {
  if (!cycles_per_mfma || valu_per_mfma < 1 || valu_per_mfma > 4 || mode < 0 || mode > 3) return TRJ_EINVAL;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return TRJ_EDEVICE;
  if (hipSetDevice(device) != hipSuccess) return TRJ_EDEVICE;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return TRJ_EDEVICE;
  const int ncu = prop.multiProcessorCount;
  // More than half of a CU's LDS, so that each CU holds one workgroup.
  const size_t lds = std::min<size_t>(prop.sharedMemPerBlock, static_cast<size_t>(96) << 10);
  const int trips = 512;
  float* out = nullptr;
  unsigned long long* clk = nullptr;
  int rc = TRJ_EDEVICE;
  if (hipMalloc(&out, sizeof(float) * ncu * 512) == hipSuccess &&
      hipMalloc(&clk, sizeof(unsigned long long)) == hipSuccess) {
    double m = launch_coissue(valu_per_mfma, mode, ncu, trips, lds, out, clk, nullptr);  // warm-up
    unsigned long long cyc = 0;
    if (m > 0.0 && hipMemset(clk, 0, sizeof(unsigned long long)) == hipSuccess) {
      m = launch_coissue(valu_per_mfma, mode, ncu, trips, lds, out, clk, nullptr);
      if (m > 0.0 && hipDeviceSynchronize() == hipSuccess &&
          hipMemcpy(&cyc, clk, sizeof(cyc), hipMemcpyDeviceToHost) == hipSuccess && cyc > 0) {
        *cycles_per_mfma = static_cast<double>(cyc) / ncu / m;
        rc = TRJ_OK;
      }
    }
  }
  if (out) hipFree(out);
  if (clk) hipFree(clk);
  return rc;
}

int trj_mfma_rate(int device, int precision, double* flops_per_s, double* clock_mhz) {
  if (!flops_per_s) return TRJ_EINVAL;
  if (precision != TRJ_PREC_FP64 && precision != TRJ_PREC_FP16X3 && precision != TRJ_PREC_BF16X3)
    return TRJ_EINVAL;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return TRJ_EDEVICE;
  if (hipSetDevice(device) != hipSuccess) return TRJ_EDEVICE;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return TRJ_EDEVICE;
  const int nblocks = prop.multiProcessorCount * 2;  // two wavefronts per SIMD
  const int iters = precision == TRJ_PREC_FP64 ? 4000 : 16000;
  float* out = nullptr;
  unsigned long long* clk = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = TRJ_EDEVICE;
  if (hipMalloc(&out, sizeof(float) * nblocks * 256) == hipSuccess &&
      hipMalloc(&clk, sizeof(unsigned long long) * 2) == hipSuccess && hipEventCreate(&e0) == hipSuccess &&
      hipEventCreate(&e1) == hipSuccess) {
    launch_mfma_rate(precision, nblocks, iters / 16, out, clk, nullptr);  // warm-up
    hipMemset(clk, 0, sizeof(unsigned long long) * 2);
    hipEventRecord(e0, nullptr);
    const double flops = launch_mfma_rate(precision, nblocks, iters, out, clk, nullptr);
    hipEventRecord(e1, nullptr);
    unsigned long long c2[2] = {0, 0};
    if (hipEventSynchronize(e1) == hipSuccess &&
        hipMemcpy(c2, clk, sizeof(c2), hipMemcpyDeviceToHost) == hipSuccess) {
      float ms = 0.f;
      hipEventElapsedTime(&ms, e0, e1);
      if (ms > 0.f) {
        *flops_per_s = flops / (ms * 1e-3);
        if (clock_mhz) *clock_mhz = c2[1] ? 100.0 * static_cast<double>(c2[0]) / static_cast<double>(c2[1]) : 0.0;
        rc = TRJ_OK;
      }
    }
  }
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  if (out) hipFree(out);
  if (clk) hipFree(clk);
  return rc;
}

int trj_last_timings(trj_ctx* c, double* ms4) {
  if (!c || !ms4) return TRJ_EINVAL;
  hipSetDevice(c->device);
  int r = sync_all(c);
  if (r) return r;
  for (int k = 0; k < 4; ++k) ms4[k] = c->last_ms[k];
  return TRJ_OK;
}

int trj_timing_totals(trj_ctx* c, double* ms4, int64_t* nframes) {
  if (!c || !ms4) return TRJ_EINVAL;
  hipSetDevice(c->device);
  int r = sync_all(c);
  if (r) return r;
  for (int k = 0; k < 4; ++k) ms4[k] = c->tot_ms[k];
  if (nframes) *nframes = c->tot_frames;
  return TRJ_OK;
}

int trj_timing_reset(trj_ctx* c) {
  if (!c) return TRJ_EINVAL;
  hipSetDevice(c->device);
  int r = sync_all(c);
  if (r) return r;
  for (int k = 0; k < 4; ++k) c->tot_ms[k] = 0.0;
  c->tot_frames = 0;
  return TRJ_OK;
}

int64_t trj_device_bytes(const trj_ctx* c) {
  if (!c) return 0;
  int64_t s = 0;
  for (const DevBuf* b : dev_bufs(const_cast<trj_ctx*>(c))) s += static_cast<int64_t>(b->cap);
  return s;
}

}  // extern "C"
