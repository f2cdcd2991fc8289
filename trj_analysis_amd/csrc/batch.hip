// batch.hip -- the batched-frames API of the trj_hip library (trj_batch_*):
// per-frame planning, grouping into launch groups, device descriptors, and
// graph capture / replay of the fixed-grid per-group launch sequence.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/trj_hip.h"
#include "kernels.h"
#include "trj_internal.h"

using namespace trj;

namespace {

struct Buf {
  void* p = nullptr;
  size_t cap = 0;
};

struct FramePlan {
  int lsmax, nit, ncell, W;
  int64_t ntasks, hist_off;
  BatchFrameDesc d;
};

}  // namespace

struct trj_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  trj_batch_params p{};
  std::string err;
  Buf hdr, fr, pos, spos, stype, cellid, count, start, cursor, blocksum, thr, hist, flags;
  void* pin[2] = {nullptr, nullptr};
  size_t pin_cap[2] = {0, 0};
  hipEvent_t pin_done[2] = {nullptr, nullptr};
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int thr_lsmax = -1;
  double thr_deltar = 0.0;
  // captured group sequence
  hipGraphExec_t gexec = nullptr;
  BatchDev gdev{};
  int g_nblocks = 0;
  size_t g_lds = 0;
  // batched S(Q) (trj_batch_sq_run): group descriptors, sort and contraction
  // buffers, all frames' shell sums, and its own captured sequence
  Buf sq_hdr, sq_fr, sq_items, sq_plane, sq_blk, sq_blkoff, sq_spoff, sq_u, sq_rho, sq_tiles, sq_sqf, sq_sqfp, sq_nq,
      sq_anc;
  void* sq_pin[2] = {nullptr, nullptr};
  size_t sq_pin_cap[2] = {0, 0};
  hipEvent_t sq_pin_done[2] = {nullptr, nullptr};
  hipGraphExec_t sq_gexec = nullptr;
  BatchSqDev sq_gdev{};
  int sq_gprec = -1;
};

namespace {

int bfail(trj_batch* b, int code, const std::string& m) {
  if (b) b->err = m;
  return code;
}

#define BHIP(b, call)                                                                          \
  do {                                                                                         \
    hipError_t e_ = (call);                                                                    \
    if (e_ != hipSuccess) return bfail((b), TRJ_EDEVICE, std::string(#call ": ") + hipGetErrorString(e_)); \
  } while (0)

int bensure(trj_batch* b, Buf& x, size_t bytes) {
  if (bytes == 0) bytes = 16;
  if (x.cap >= bytes) return TRJ_OK;
  BHIP(b, hipStreamSynchronize(b->stream));
  if (x.p) (void)hipFree(x.p);
  x.p = nullptr;
  x.cap = 0;
  const size_t want = bytes + bytes / 4;
  if (hipMalloc(&x.p, want) != hipSuccess) return bfail(b, TRJ_ENOMEM, "batch: device allocation failed");
  x.cap = want;
  if (b->gexec) {  // pointers changed: the captured sequences are stale
    (void)hipGraphExecDestroy(b->gexec);
    b->gexec = nullptr;
  }
  if (b->sq_gexec) {
    (void)hipGraphExecDestroy(b->sq_gexec);
    b->sq_gexec = nullptr;
  }
  return TRJ_OK;
}

int plan_frames(const trj_batch_params& p, int nframes, const trj_batch_frame* fr, std::vector<FramePlan>& out,
                int64_t* total, std::string* err) {
  if (p.ndim != 2 && p.ndim != 3) {
    *err = "batch: ndim must be 2 or 3";
    return TRJ_EINVAL;
  }
  if (!(p.deltar > 0.0)) {
    *err = "batch: deltar must be positive";
    return TRJ_EINVAL;
  }
  out.assign(nframes, FramePlan{});
  int64_t off = 0;
  for (int f = 0; f < nframes; ++f) {
    const trj_batch_frame& F = fr[f];
    FramePlan& P = out[f];
    if (F.n < 1 || F.nsp < 1 || F.nsp > kNspMax) {
      *err = "batch: frame " + std::to_string(f) + " has n < 1 or nsp outside 1..10";
      return TRJ_EINVAL;
    }
    double box[3] = {F.box[0], F.box[1], p.ndim == 3 ? F.box[2] : 1.0};
    for (int k = 0; k < p.ndim; ++k)
      if (!(box[k] > 0.0)) {
        *err = "batch: frame " + std::to_string(f) + " has a non-positive box length";
        return TRJ_EINVAL;
      }
    double rc = 0.0;
    P.lsmax = rdf_lsmax(p.ndim, box, p.rcrdf, p.deltar, &rc);
    P.nit = F.nsp * (F.nsp + 1) / 2;
    if (cells_lds_bytes(P.nit, P.lsmax) > 64 * 1024) {
      *err = "batch: frame " + std::to_string(f) + " histogram exceeds 64 KiB of LDS";
      return TRJ_ELIMIT;
    }
    int per[3] = {F.periodic[0] != 0, F.periodic[1] != 0, F.periodic[2] != 0};
    double rc_sticky = p.rcrdf;  // frames are independent: the cap applies per frame
    FrameGeom g = frame_geometry(p.ndim, box, per, &rc_sticky);
    BatchFrameDesc& D = P.d;
    std::memset(&D, 0, sizeof(D));
    const double rcut = std::sqrt(g.side2);
    P.ncell = 1;
    for (int k = 0; k < 3; ++k) {
      if (k < p.ndim) {
        int nc = static_cast<int>(std::floor(box[k] / (rcut * (1.0 + 1e-6))));
        nc = std::max(1, std::min(nc, p.ndim == 3 ? 160 : 2048));
        if (per[k] && nc < 3) {
          *err = "batch: frame " + std::to_string(f) +
                 " has fewer than 3 cells on a periodic axis; use the per-frame path (trj_frame)";
          return TRJ_EINVAL;
        }
        D.nc[k] = nc;
        D.inv_h[k] = nc / box[k];
        D.periodic[k] = per[k];
        D.cellshift[k] = (per[k] && nc >= 5) ? 1 : 0;
      } else {
        D.nc[k] = 1;
        D.inv_h[k] = 0.0;
      }
      D.box[k] = box[k];
      D.lmi[k] = g.lmi[k];
      D.half[k] = g.half[k];
      P.ncell *= D.nc[k];
    }
    const double mean = static_cast<double>(F.n) / P.ncell;
    P.W = p.waves_per_cell > 0 ? p.waves_per_cell
                               : std::max(1, std::min(8, static_cast<int>(std::lround(mean / 64.0))));
    P.ntasks = static_cast<int64_t>(P.ncell) * P.W;
    D.n = static_cast<int32_t>(F.n);
    D.nsp = F.nsp;
    D.nit = P.nit;
    D.lsmax = P.lsmax;
    D.ncell = P.ncell;
    D.W = P.W;
    D.side2 = g.side2;
    D.inv_dr = static_cast<float>(1.0 / p.deltar);
    P.hist_off = off;
    D.hist_off = off;
    off += static_cast<int64_t>(P.nit) * P.lsmax;
  }
  *total = off;
  return TRJ_OK;
}

int run_groups(trj_batch* b, const trj_batch_params& p, int nframes, const trj_batch_frame* fr,
               const void* pos_dev, int dtype, const int32_t* itype_dev, int64_t* hist_host, double* gpu_ms) {
  std::vector<FramePlan> plan;
  int64_t total = 0;
  int r = plan_frames(p, nframes, fr, plan, &total, &b->err);
  if (r) return r;
  if (dtype != TRJ_F32 && dtype != TRJ_F64) return bfail(b, TRJ_EINVAL, "batch: bad dtype");
  const int F = p.frames_per_launch > 0 ? p.frames_per_launch : 64;
  const int G = p.workgroups > 0 ? p.workgroups : 1024;
  // capacities over all groups
  std::vector<int64_t> atom_off(nframes + 1, 0);
  for (int f = 0; f < nframes; ++f) atom_off[f + 1] = atom_off[f] + fr[f].n;
  int64_t max_atoms = 1, max_cells = 1;
  int max_lsmax = 1;
  size_t lds = 0;
  for (int g0 = 0; g0 < nframes; g0 += F) {
    const int g1 = std::min(nframes, g0 + F);
    int64_t na = 0, nc = 0;
    for (int f = g0; f < g1; ++f) {
      na += fr[f].n;
      nc += plan[f].ncell;
      max_lsmax = std::max(max_lsmax, plan[f].lsmax);
      lds = std::max(lds, pair_lds_bytes(plan[f].nit, plan[f].lsmax));
    }
    max_atoms = std::max(max_atoms, na);
    max_cells = std::max(max_cells, nc);
  }
  if (max_atoms > (int64_t(1) << 31) - 1024 || max_cells > (int64_t(1) << 30))
    return bfail(b, TRJ_ELIMIT, "batch: a launch group exceeds 2^31 atoms; lower frames_per_launch");
  if ((r = bensure(b, b->hdr, sizeof(BatchHeader)))) return r;
  if ((r = bensure(b, b->fr, sizeof(BatchFrameDesc) * F))) return r;
  if ((r = bensure(b, b->pos, sizeof(double) * 3 * max_atoms))) return r;
  if ((r = bensure(b, b->spos, sizeof(double) * 3 * max_atoms))) return r;
  if ((r = bensure(b, b->stype, sizeof(int32_t) * max_atoms))) return r;
  if ((r = bensure(b, b->cellid, sizeof(int32_t) * max_atoms))) return r;
  if ((r = bensure(b, b->count, sizeof(int32_t) * (max_cells + 1)))) return r;
  if ((r = bensure(b, b->start, sizeof(int32_t) * (max_cells + 1)))) return r;
  if ((r = bensure(b, b->cursor, sizeof(int32_t) * max_cells))) return r;
  if ((r = bensure(b, b->blocksum, sizeof(int32_t) * (G + 1)))) return r;
  if ((r = bensure(b, b->hist, sizeof(int64_t) * std::max<int64_t>(1, total)))) return r;
  if ((r = bensure(b, b->flags, sizeof(uint32_t)))) return r;
  if (b->thr_lsmax < max_lsmax || b->thr_deltar != p.deltar) {
    std::vector<double> thr = rdf_bin_thresholds(p.deltar, max_lsmax);
    thr.push_back(INFINITY);
    thr[0] = 0.0;
    if ((r = bensure(b, b->thr, sizeof(double) * thr.size()))) return r;
    BHIP(b, hipMemcpyAsync(b->thr.p, thr.data(), sizeof(double) * thr.size(), hipMemcpyHostToDevice, b->stream));
    b->thr_lsmax = max_lsmax;
    b->thr_deltar = p.deltar;
  }
  // Frames of different lsmax share the edges: thr[m] depends on m and deltar only.
  BatchDev d{};
  d.hdr = static_cast<const BatchHeader*>(b->hdr.p);
  d.fr = static_cast<const BatchFrameDesc*>(b->fr.p);
  d.pos_in = pos_dev;
  d.itype_in = itype_dev;
  d.dtype = dtype;
  d.ndim = p.ndim;
  d.pos = static_cast<double*>(b->pos.p);
  d.spos = static_cast<double*>(b->spos.p);
  d.stype = static_cast<int32_t*>(b->stype.p);
  d.cellid = static_cast<int32_t*>(b->cellid.p);
  d.count = static_cast<int32_t*>(b->count.p);
  d.start = static_cast<int32_t*>(b->start.p);
  d.cursor = static_cast<int32_t*>(b->cursor.p);
  d.blocksum = static_cast<int32_t*>(b->blocksum.p);
  d.cap_atoms = max_atoms;
  d.thr = static_cast<const double*>(b->thr.p);
  d.hist = static_cast<int64_t*>(b->hist.p);
  d.flags = static_cast<uint32_t*>(b->flags.p);
  BHIP(b, hipMemsetAsync(b->hist.p, 0, sizeof(int64_t) * std::max<int64_t>(1, total), b->stream));
  BHIP(b, hipMemsetAsync(b->flags.p, 0, sizeof(uint32_t), b->stream));
  const bool graph = p.use_graph != 0;
  const bool same = b->gexec && std::memcmp(&b->gdev, &d, sizeof(d)) == 0 && b->g_nblocks == G && b->g_lds >= lds;
  if (graph && !same) {
    if (b->gexec) {
      (void)hipGraphExecDestroy(b->gexec);
      b->gexec = nullptr;
    }
    hipGraph_t gr = nullptr;
    BHIP(b, hipStreamSynchronize(b->stream));
    BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    launch_batch_group(d, G, lds, b->stream);
    BHIP(b, hipStreamEndCapture(b->stream, &gr));
    hipError_t e = hipGraphInstantiate(&b->gexec, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    if (e != hipSuccess) return bfail(b, TRJ_EDEVICE, std::string("batch graph: ") + hipGetErrorString(e));
    b->gdev = d;
    b->g_nblocks = G;
    b->g_lds = lds;
  }
  BHIP(b, hipEventRecord(b->ev0, b->stream));
  int k = 0;
  for (int g0 = 0; g0 < nframes; g0 += F, k ^= 1) {
    const int g1 = std::min(nframes, g0 + F);
    const int nf = g1 - g0;
    const size_t bytes = sizeof(BatchHeader) + sizeof(BatchFrameDesc) * nf;
    BHIP(b, hipEventSynchronize(b->pin_done[k]));
    if (b->pin_cap[k] < bytes) {
      if (b->pin[k]) (void)hipHostFree(b->pin[k]);
      b->pin[k] = nullptr;
      if (hipHostMalloc(&b->pin[k], sizeof(BatchHeader) + sizeof(BatchFrameDesc) * F, hipHostMallocDefault) !=
          hipSuccess)
        return bfail(b, TRJ_ENOMEM, "batch: pinned allocation failed");
      b->pin_cap[k] = sizeof(BatchHeader) + sizeof(BatchFrameDesc) * F;
    }
    BatchHeader* H = static_cast<BatchHeader*>(b->pin[k]);
    BatchFrameDesc* D = reinterpret_cast<BatchFrameDesc*>(static_cast<char*>(b->pin[k]) + sizeof(BatchHeader));
    int64_t la = 0, lt = 0;
    int lc = 0;
    for (int f = g0; f < g1; ++f) {
      BatchFrameDesc x = plan[f].d;
      x.atom_off = atom_off[f];
      x.local_off = la;
      x.task_off = lt;
      x.cell_off = lc;
      D[f - g0] = x;
      la += fr[f].n;
      lt += plan[f].ntasks;
      lc += plan[f].ncell;
    }
    H->nframes = nf;
    H->ncells = lc;
    H->natoms = la;
    H->ntasks = lt;
    BHIP(b, hipMemcpyAsync(b->hdr.p, H, sizeof(BatchHeader), hipMemcpyHostToDevice, b->stream));
    BHIP(b, hipMemcpyAsync(b->fr.p, D, sizeof(BatchFrameDesc) * nf, hipMemcpyHostToDevice, b->stream));
    BHIP(b, hipEventRecord(b->pin_done[k], b->stream));
    if (graph)
      BHIP(b, hipGraphLaunch(b->gexec, b->stream));
    else
      launch_batch_group(d, G, lds, b->stream);
    BHIP(b, hipGetLastError());
  }
  BHIP(b, hipEventRecord(b->ev1, b->stream));
  uint32_t flags = 0;
  BHIP(b, hipMemcpyAsync(&flags, b->flags.p, sizeof(flags), hipMemcpyDeviceToHost, b->stream));
  if (hist_host)
    BHIP(b, hipMemcpyAsync(hist_host, b->hist.p, sizeof(int64_t) * total, hipMemcpyDeviceToHost, b->stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  if (gpu_ms) {
    float ms = 0.f;
    BHIP(b, hipEventElapsedTime(&ms, b->ev0, b->ev1));
    *gpu_ms = ms;
  }
  if (flags & kFlagType) return bfail(b, TRJ_EINVAL, "batch: species types outside 1..nsp");
  if (flags & kFlagRange) return bfail(b, TRJ_ERANGE, "batch: non-finite or out-of-range coordinates");
  return TRJ_OK;
}

// ------------------------------------------------------------- batched S(Q)
// Per-frame S(Q) plan on the half-space grid of the frame's own box (the
// rules of sq_plan for TRJ_QSET_HALFSPACE, plan.cpp): 2 pi / L per axis,
// dq their minimum, nqmax = min(nint(qmax / dq), the largest shell on the
// grid).  The largest shell is the grid corner's (|kx| = gx/2, |ky| = gy/2,
// kz = gz/2 - 1): every operation of the shell rule is monotone in |k|.
// Splits: at most 32768 atoms (FP64) or 8192 (FP32 accumulation) per
// workgroup, the per-frame path's bound.
struct SqFramePlan {
  int nqmax, ks, nblk;
  double fk[3], dq;
  int64_t out_off, outp_off;
};

constexpr int64_t kSqRhoCap = int64_t(4) << 30;  // rho bytes of one group
constexpr int64_t kSqAtomCap = int64_t(1) << 30;  // atoms of one group

int plan_sq_frames(const trj_batch_params& p, const trj_batch_sq_params& q, int nframes, const trj_batch_frame* fr,
                   std::vector<SqFramePlan>& out, int64_t* tot_sqf, int64_t* tot_sqfp, std::string* err) {
  if (p.ndim != 2 && p.ndim != 3) {
    *err = "batch sq: ndim must be 2 or 3";
    return TRJ_EINVAL;
  }
  for (int k = 0; k < p.ndim; ++k)
    if (q.grid[k] < 2 || (q.grid[k] % 2) != 0 || q.grid[k] > 4096) {
      *err = "batch sq: half-space grid extents must be even, 2..4096";
      return TRJ_EINVAL;
    }
  if (!(q.qmax > 0.0)) {
    *err = "batch sq: qmax must be positive";
    return TRJ_EINVAL;
  }
  if (q.precision < TRJ_PREC_FP64 || q.precision > TRJ_PREC_BF16X3) {
    *err = "batch sq: unknown precision";
    return TRJ_EINVAL;
  }
  const int64_t span = q.precision == TRJ_PREC_FP64 ? 32768 : 8192;
  out.assign(nframes, SqFramePlan{});
  int64_t o = 0, op = 0;
  for (int f = 0; f < nframes; ++f) {
    const trj_batch_frame& F = fr[f];
    SqFramePlan& P = out[f];
    if (F.n < 1 || F.n > kSqAtomCap || F.nsp < 1 || F.nsp > kNspMax) {
      *err = "batch sq: frame " + std::to_string(f) + " has n outside 1..2^30 or nsp outside 1..10";
      return TRJ_EINVAL;
    }
    double box[3] = {1.0, 1.0, 1.0};
    for (int k = 0; k < p.ndim; ++k) {
      if (!(F.box[k] > 0.0) || !std::isfinite(F.box[k])) {
        *err = "batch sq: frame " + std::to_string(f) + " has a non-positive box length";
        return TRJ_EINVAL;
      }
      box[k] = F.box[k];
    }
    for (int k = 0; k < 3; ++k) P.fk[k] = 2.0 * M_PI / box[k];
    P.dq = std::fmin(P.fk[0], P.fk[1]);
    if (p.ndim == 3) P.dq = std::fmin(P.dq, P.fk[2]);
    const double qa = (-q.grid[0] / 2) * P.fk[0], qb = (-q.grid[1] / 2) * P.fk[1];
    const double qc = p.ndim == 3 ? (q.grid[2] / 2 - 1) * P.fk[2] : 0.0;
    const long maxshell = std::lround(std::sqrt(qa * qa + qb * qb + qc * qc) / P.dq);
    const double qcap = std::floor(q.qmax / P.dq + 0.5);
    P.nqmax = qcap < static_cast<double>(maxshell) ? static_cast<int>(qcap) : static_cast<int>(maxshell);
    if (P.nqmax > 32767) {
      *err = "batch sq: frame " + std::to_string(f) + " has nqmax above 32767";
      return TRJ_ELIMIT;
    }
    P.ks = static_cast<int>((F.n + span - 1) / span);
    P.nblk = static_cast<int>((F.n + 255) / 256);
    P.out_off = o;
    P.outp_off = op;
    o += P.nqmax + 1;
    op += static_cast<int64_t>(F.nsp) * (P.nqmax + 1);
  }
  *tot_sqf = o;
  *tot_sqfp = op;
  return TRJ_OK;
}

// The half-space tiles of the contraction kernels for `precision`: every
// tile of the dense box, all blocks marked (each frame's shells pick their
// vectors later).
std::vector<int32_t> halfspace_tiles(int ndim, const int32_t grid[3], int precision, int kmax[3]) {
  const TileShape tk = tile_shape(precision);
  const int kmin[3] = {-grid[0] / 2, -grid[1] / 2, 0};
  kmax[0] = grid[0] / 2 - 1;
  kmax[1] = grid[1] / 2 - 1;
  kmax[2] = ndim == 3 ? grid[2] / 2 - 1 : 0;
  std::vector<int32_t> t;
  for (int x0 = kmin[0]; x0 <= kmax[0]; x0 += tk.tx)
    for (int y0 = kmin[1]; y0 <= kmax[1]; y0 += tk.ty)
      for (int z0 = kmin[2]; z0 <= kmax[2]; z0 += tk.tz) {
        t.push_back(x0);
        t.push_back(y0);
        t.push_back(z0);
        t.push_back(-1);
      }
  return t;
}

int run_sq_groups(trj_batch* b, const trj_batch_sq_params& q, int nframes, const trj_batch_frame* fr,
                  const void* pos_dev, int dtype, const int32_t* itype_dev, const double* bsc, double* sqf_host,
                  double* sqfp_host, int64_t* nq_host, double* gpu_ms) {
  const trj_batch_params& p = b->p;
  std::vector<SqFramePlan> plan;
  int64_t tot_sqf = 0, tot_sqfp = 0;
  int r = plan_sq_frames(p, q, nframes, fr, plan, &tot_sqf, &tot_sqfp, &b->err);
  if (r) return r;
  if (dtype != TRJ_F32 && dtype != TRJ_F64) return bfail(b, TRJ_EINVAL, "batch sq: bad dtype");
  int kmax[3];
  const std::vector<int32_t> tiles = halfspace_tiles(p.ndim, q.grid, q.precision, kmax);
  const int ntiles = static_cast<int>(tiles.size() / 4);
  const TileShape tk = tile_shape(q.precision);
  const int64_t tvec_total = int64_t(ntiles) * tk.vec();
  const int64_t plane_bytes = static_cast<int64_t>(rho_entry_bytes(q.precision)) * tvec_total;
  constexpr int64_t kUnitLoads = 256 * 48;
  // Groups: consecutive frames, at most frames_per_launch (0: 64), their rho
  // planes within kSqRhoCap and their atoms within kSqAtomCap.
  const int F = p.frames_per_launch > 0 ? p.frames_per_launch : 64;
  std::vector<int> gstart;
  int64_t cap_atoms = 1, cap_counts = 1, cap_planes = 1;
  int cap_items = 1, cap_frames = 1, cap_nch = 1, cap_blocks = 1;
  {
    int f = 0;
    while (f < nframes) {
      const int64_t planes0 = int64_t(plan[f].ks) * fr[f].nsp;
      if (planes0 * plane_bytes > kSqRhoCap)
        return bfail(b, TRJ_ELIMIT, "batch sq: frame " + std::to_string(f) +
                                        " alone needs more than 4 GiB of partial sums; use the per-frame path");
      gstart.push_back(f);
      int64_t atoms = 0, counts = 0, planes = 0;
      int items = 0, nch = 0, blocks = 0, nf = 0;
      while (f < nframes && nf < F) {
        const int64_t pl = int64_t(plan[f].ks) * fr[f].nsp;
        if (nf > 0 && ((planes + pl) * plane_bytes > kSqRhoCap || atoms + fr[f].n > kSqAtomCap)) break;
        atoms += fr[f].n;
        counts += int64_t(fr[f].nsp) * plan[f].nblk;
        planes += pl;
        items += plan[f].ks;
        nch += fr[f].nsp;
        blocks += plan[f].nblk;
        ++nf;
        ++f;
      }
      cap_atoms = std::max(cap_atoms, atoms);
      cap_counts = std::max(cap_counts, counts);
      cap_planes = std::max(cap_planes, planes);
      cap_items = std::max(cap_items, items);
      cap_frames = std::max(cap_frames, nf);
      cap_nch = std::max(cap_nch, nch);
      cap_blocks = std::max(cap_blocks, blocks);
    }
  }
  gstart.push_back(nframes);
  const int ngroups = static_cast<int>(gstart.size()) - 1;
  // Shell kernel: LDS accumulators for the largest frame's shells when they fit.
  int max_acc = 0;
  for (int f = 0; f < nframes; ++f) max_acc = std::max(max_acc, (2 + fr[f].nsp) * (plan[f].nqmax + 1));
  const int shell_lds = max_acc * 8 <= 48 * 1024 ? max_acc : 0;
  if ((r = bensure(b, b->sq_hdr, sizeof(BatchSqHeader)))) return r;
  if ((r = bensure(b, b->sq_fr, sizeof(BatchSqFrame) * cap_frames))) return r;
  if ((r = bensure(b, b->sq_items, sizeof(int4) * cap_items))) return r;
  if ((r = bensure(b, b->sq_plane, sizeof(int32_t) * cap_items))) return r;
  if ((r = bensure(b, b->sq_blk, sizeof(int32_t) * cap_counts))) return r;
  if ((r = bensure(b, b->sq_blkoff, sizeof(int32_t) * cap_counts))) return r;
  if ((r = bensure(b, b->sq_spoff, sizeof(int32_t) * (cap_nch + 1)))) return r;
  if ((r = bensure(b, b->sq_u, sizeof(double) * 12 * cap_atoms))) return r;
  if ((r = bensure(b, b->sq_rho, static_cast<size_t>(plane_bytes) * cap_planes))) return r;
  if ((r = bensure(b, b->sq_tiles, sizeof(int32_t) * tiles.size()))) return r;
  if ((r = bensure(b, b->sq_sqf, sizeof(double) * tot_sqf))) return r;
  if ((r = bensure(b, b->sq_sqfp, sizeof(double) * tot_sqfp))) return r;
  if ((r = bensure(b, b->sq_nq, sizeof(int64_t) * tot_sqf))) return r;
  if ((r = bensure(b, b->flags, sizeof(uint32_t)))) return r;
  // Anchor table of the contraction (as for one frame: the same values the
  // kernel would evaluate), for the largest group's entries; off with
  // TRJ_SQ_ANCHOR_TABLE=0 or above 8 GiB.
  int anc_rows[3], anc_base[3];
  anchor_layout(tiles.data(), 4, tiles.size() / 4, tk, anc_rows, anc_base);
  const char* anc_env = std::getenv("TRJ_SQ_ANCHOR_TABLE");
  const int64_t anc_bytes =
      (anc_env && std::strcmp(anc_env, "0") == 0) ? 0 : anchor_table_bytes(q.precision, anc_rows, cap_atoms);
  if (anc_bytes > 0 && (r = bensure(b, b->sq_anc, static_cast<size_t>(anc_bytes)))) return r;
  BHIP(b, hipMemcpyAsync(b->sq_tiles.p, tiles.data(), sizeof(int32_t) * tiles.size(), hipMemcpyHostToDevice,
                         b->stream));
  BHIP(b, hipMemsetAsync(b->sq_sqf.p, 0, sizeof(double) * tot_sqf, b->stream));
  BHIP(b, hipMemsetAsync(b->sq_sqfp.p, 0, sizeof(double) * tot_sqfp, b->stream));
  BHIP(b, hipMemsetAsync(b->sq_nq.p, 0, sizeof(int64_t) * tot_sqf, b->stream));
  BHIP(b, hipMemsetAsync(b->flags.p, 0, sizeof(uint32_t), b->stream));
  BatchSqDev d{};
  d.hdr = static_cast<const BatchSqHeader*>(b->sq_hdr.p);
  d.fr = static_cast<const BatchSqFrame*>(b->sq_fr.p);
  d.pos_in = pos_dev;
  d.itype_in = itype_dev;
  d.dtype = dtype;
  d.ndim = p.ndim;
  d.blk = static_cast<int32_t*>(b->sq_blk.p);
  d.blkoff = static_cast<int32_t*>(b->sq_blkoff.p);
  d.sp_off = static_cast<int32_t*>(b->sq_spoff.p);
  d.u = static_cast<double*>(b->sq_u.p);
  d.stride = cap_atoms;
  d.items = static_cast<const int4*>(b->sq_items.p);
  d.item_plane = static_cast<const int32_t*>(b->sq_plane.p);
  d.cap_items = cap_items;
  d.tiles = static_cast<const int32_t*>(b->sq_tiles.p);
  d.ntiles = ntiles;
  d.tx = tk.tx;
  d.ty = tk.ty;
  d.tz = tk.tz;
  for (int k = 0; k < 3; ++k) d.kmax[k] = kmax[k];
  d.rho = static_cast<double*>(b->sq_rho.p);
  d.rho_f32 = q.precision != TRJ_PREC_FP64;
  d.anc = anc_bytes > 0 ? b->sq_anc.p : nullptr;
  for (int k = 0; k < 3; ++k) {
    d.anc_rows[k] = anc_rows[k];
    d.anc_base[k] = anc_base[k];
  }
  d.sqf = static_cast<double*>(b->sq_sqf.p);
  d.sqfp = static_cast<double*>(b->sq_sqfp.p);
  d.nq = static_cast<unsigned long long*>(b->sq_nq.p);
  d.flags = static_cast<uint32_t*>(b->flags.p);
  d.nblocks = p.workgroups > 0 ? p.workgroups : 1024;
  d.shell_lds = shell_lds;
  const bool graph = p.use_graph != 0;
  const bool same = b->sq_gexec && std::memcmp(&b->sq_gdev, &d, sizeof(d)) == 0 && b->sq_gprec == q.precision;
  if (graph && !same) {
    if (b->sq_gexec) {
      (void)hipGraphExecDestroy(b->sq_gexec);
      b->sq_gexec = nullptr;
    }
    hipGraph_t gr = nullptr;
    BHIP(b, hipStreamSynchronize(b->stream));
    BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    launch_batch_sq_group(d, q.precision, b->stream);
    BHIP(b, hipStreamEndCapture(b->stream, &gr));
    hipError_t e = hipGraphInstantiate(&b->sq_gexec, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    if (e != hipSuccess) return bfail(b, TRJ_EDEVICE, std::string("batch sq graph: ") + hipGetErrorString(e));
    b->sq_gdev = d;
    b->sq_gprec = q.precision;
  }
  const size_t pin_bytes = sizeof(BatchSqHeader) + sizeof(BatchSqFrame) * cap_frames +
                           (sizeof(int4) + sizeof(int32_t)) * cap_items;
  std::vector<int> order;
  std::vector<int64_t> aoff(nframes + 1, 0);
  for (int f = 0; f < nframes; ++f) aoff[f + 1] = aoff[f] + fr[f].n;
  BHIP(b, hipEventRecord(b->ev0, b->stream));
  int k = 0;
  for (int g = 0; g < ngroups; ++g, k ^= 1) {
    const int g0 = gstart[g], g1 = gstart[g + 1], nf = g1 - g0;
    BHIP(b, hipEventSynchronize(b->sq_pin_done[k]));
    if (b->sq_pin_cap[k] < pin_bytes) {
      if (b->sq_pin[k]) (void)hipHostFree(b->sq_pin[k]);
      b->sq_pin[k] = nullptr;
      b->sq_pin_cap[k] = 0;
      if (hipHostMalloc(&b->sq_pin[k], pin_bytes, hipHostMallocDefault) != hipSuccess)
        return bfail(b, TRJ_ENOMEM, "batch sq: pinned allocation failed");
      b->sq_pin_cap[k] = pin_bytes;
    }
    char* pin = static_cast<char*>(b->sq_pin[k]);
    BatchSqHeader* H = reinterpret_cast<BatchSqHeader*>(pin);
    BatchSqFrame* D = reinterpret_cast<BatchSqFrame*>(pin + sizeof(BatchSqHeader));
    int4* items = reinterpret_cast<int4*>(pin + sizeof(BatchSqHeader) + sizeof(BatchSqFrame) * cap_frames);
    int32_t* iplane = reinterpret_cast<int32_t*>(items + cap_items);
    int blk0 = 0, moff = 0, chan0 = 0, plane0 = 0, units = 0;
    for (int f = g0; f < g1; ++f) {
      BatchSqFrame& x = D[f - g0];
      std::memset(&x, 0, sizeof(x));
      x.atom_off = aoff[f];
      x.out_off = plan[f].out_off;
      x.outp_off = plan[f].outp_off;
      x.n = static_cast<int32_t>(fr[f].n);
      x.nsp = fr[f].nsp;
      x.blk0 = blk0;
      x.nblk = plan[f].nblk;
      x.moff = moff;
      x.chan0 = chan0;
      x.plane0 = plane0;
      x.ks = plan[f].ks;
      x.nqmax = plan[f].nqmax;
      // Shell units: about kUnitLoads split partials each (one vector per
      // thread at least, at most 32).
      const int64_t per_vec = int64_t(x.ks) * x.nsp;
      const int64_t vpu = std::max<int64_t>(256, std::min<int64_t>(32 * 256, (kUnitLoads / per_vec) / 256 * 256));
      x.vpu = static_cast<int32_t>(vpu);
      x.unit0 = units;
      units += static_cast<int>((tvec_total + vpu - 1) / vpu);
      for (int a = 0; a < 3; ++a) {
        x.box[a] = a < p.ndim ? fr[f].box[a] : 1.0;
        x.fk[a] = plan[f].fk[a];
      }
      x.dq = plan[f].dq;
      for (int s = 0; s < kNspMax; ++s) x.bsc[s] = s < fr[f].nsp ? (bsc ? bsc[int64_t(f) * kNspMax + s] : 1.0) : 0.0;
      blk0 += x.nblk;
      moff += x.nsp * x.nblk;
      chan0 += x.nsp;
      plane0 += x.ks * x.nsp;
    }
    // Work items, the largest frames' first (the grid is dispatched in order).
    order.resize(nf);
    for (int j = 0; j < nf; ++j) order[j] = g0 + j;
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return fr[a].n > fr[c].n; });
    int ni = 0;
    for (int f : order) {
      const BatchSqFrame& x = D[f - g0];
      for (int s = 0; s < x.ks; ++s) {
        items[ni] = make_int4(x.chan0, x.chan0 + x.nsp, s, x.ks);
        iplane[ni] = x.plane0 + s * x.nsp;
        ++ni;
      }
    }
    H->nframes = nf;
    H->nblocks = blk0;
    H->ncounts = moff;
    H->nch = chan0;
    H->nitems = ni;
    H->nunits = units;
    BHIP(b, hipMemcpyAsync(b->sq_hdr.p, H, sizeof(BatchSqHeader), hipMemcpyHostToDevice, b->stream));
    BHIP(b, hipMemcpyAsync(b->sq_fr.p, D, sizeof(BatchSqFrame) * nf, hipMemcpyHostToDevice, b->stream));
    BHIP(b, hipMemcpyAsync(b->sq_items.p, items, sizeof(int4) * ni, hipMemcpyHostToDevice, b->stream));
    BHIP(b, hipMemcpyAsync(b->sq_plane.p, iplane, sizeof(int32_t) * ni, hipMemcpyHostToDevice, b->stream));
    BHIP(b, hipEventRecord(b->sq_pin_done[k], b->stream));
    if (graph)
      BHIP(b, hipGraphLaunch(b->sq_gexec, b->stream));
    else
      launch_batch_sq_group(d, q.precision, b->stream);
    BHIP(b, hipGetLastError());
  }
  BHIP(b, hipEventRecord(b->ev1, b->stream));
  uint32_t flags = 0;
  BHIP(b, hipMemcpyAsync(&flags, b->flags.p, sizeof(flags), hipMemcpyDeviceToHost, b->stream));
  if (sqf_host)
    BHIP(b, hipMemcpyAsync(sqf_host, b->sq_sqf.p, sizeof(double) * tot_sqf, hipMemcpyDeviceToHost, b->stream));
  if (sqfp_host)
    BHIP(b, hipMemcpyAsync(sqfp_host, b->sq_sqfp.p, sizeof(double) * tot_sqfp, hipMemcpyDeviceToHost, b->stream));
  if (nq_host)
    BHIP(b, hipMemcpyAsync(nq_host, b->sq_nq.p, sizeof(int64_t) * tot_sqf, hipMemcpyDeviceToHost, b->stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  if (gpu_ms) {
    float ms = 0.f;
    BHIP(b, hipEventElapsedTime(&ms, b->ev0, b->ev1));
    *gpu_ms = ms;
  }
  if (flags & kFlagType) return bfail(b, TRJ_EINVAL, "batch sq: species types outside 1..nsp");
  if (flags & kFlagRange) return bfail(b, TRJ_ERANGE, "batch sq: non-finite coordinates");
  return TRJ_OK;
}

}  // namespace

extern "C" {

int trj_batch_create(trj_batch** out, int device, const trj_batch_params* p) {
  if (!out || !p) return TRJ_EINVAL;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return TRJ_EDEVICE;
  if (device < 0 || device >= ndev) return TRJ_EINVAL;
  if (hipSetDevice(device) != hipSuccess) return TRJ_EDEVICE;
  trj_batch* b = new trj_batch();
  b->device = device;
  b->p = *p;
  bool ok = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) == hipSuccess;
  ok = ok && hipEventCreate(&b->ev0) == hipSuccess && hipEventCreate(&b->ev1) == hipSuccess;
  for (auto& e : b->pin_done) ok = ok && hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess;
  for (auto& e : b->sq_pin_done) ok = ok && hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess;
  if (!ok) {
    trj_batch_destroy(b);
    return TRJ_EDEVICE;
  }
  *out = b;
  return TRJ_OK;
}

void trj_batch_destroy(trj_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->sq_gexec) (void)hipGraphExecDestroy(b->sq_gexec);
  Buf* bufs[] = {&b->hdr,    &b->fr,     &b->pos,       &b->spos,      &b->stype,    &b->cellid,   &b->count,
                 &b->start,  &b->cursor, &b->blocksum,  &b->thr,       &b->hist,     &b->flags,    &b->sq_hdr,
                 &b->sq_fr,  &b->sq_items, &b->sq_plane, &b->sq_blk,   &b->sq_blkoff, &b->sq_spoff, &b->sq_u,
                 &b->sq_rho, &b->sq_tiles, &b->sq_sqf,  &b->sq_sqfp,   &b->sq_nq,     &b->sq_anc};
  for (Buf* x : bufs)
    if (x->p) (void)hipFree(x->p);
  for (int k = 0; k < 2; ++k) {
    if (b->pin[k]) (void)hipHostFree(b->pin[k]);
    if (b->pin_done[k]) (void)hipEventDestroy(b->pin_done[k]);
    if (b->sq_pin[k]) (void)hipHostFree(b->sq_pin[k]);
    if (b->sq_pin_done[k]) (void)hipEventDestroy(b->sq_pin_done[k]);
  }
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

const char* trj_batch_last_error(const trj_batch* b) { return b ? b->err.c_str() : "null batch"; }

int trj_batch_layout(const trj_batch_params* p, int32_t nframes, const trj_batch_frame* fr, int32_t* lsmax,
                     int32_t* nit, int64_t* hist_off, int64_t* total) {
  if (!p || !fr || nframes < 0) return TRJ_EINVAL;
  std::vector<FramePlan> plan;
  int64_t tot = 0;
  std::string err;
  int r = plan_frames(*p, nframes, fr, plan, &tot, &err);
  if (r) return r;
  for (int f = 0; f < nframes; ++f) {
    if (lsmax) lsmax[f] = plan[f].lsmax;
    if (nit) nit[f] = plan[f].nit;
    if (hist_off) hist_off[f] = plan[f].hist_off;
  }
  if (total) *total = tot;
  return TRJ_OK;
}

int trj_batch_run(trj_batch* b, int32_t nframes, const trj_batch_frame* fr, const void* pos_dev, int32_t dtype,
                  const int32_t* itype_dev, int64_t* hist_host, double* gpu_ms) {
  if (!b) return TRJ_EINVAL;
  if (nframes < 1 || !fr || !pos_dev || !itype_dev) return bfail(b, TRJ_EINVAL, "batch: empty or NULL input");
  (void)hipSetDevice(b->device);
  return run_groups(b, b->p, nframes, fr, pos_dev, dtype, itype_dev, hist_host, gpu_ms);
}

int trj_batch_autotune(trj_batch* b, int32_t nframes, const trj_batch_frame* fr, const void* pos_dev,
                       int32_t dtype, const int32_t* itype_dev, trj_batch_params* chosen) {
  if (!b) return TRJ_EINVAL;
  if (nframes < 1 || !fr || !pos_dev || !itype_dev) return bfail(b, TRJ_EINVAL, "batch: empty or NULL input");
  (void)hipSetDevice(b->device);
  const int ns = std::min<int>(nframes, 128);
  const int fpl[] = {16, 64, 128};
  const int wgs[] = {1024, 2048};
  const int wpc[] = {0, 1, 2};
  double best = 1e300;
  trj_batch_params bp = b->p;
  for (int f : fpl) {
    if (f > ns && f != fpl[0]) continue;
    for (int g : wgs)
      for (int w : wpc) {
        trj_batch_params t = b->p;
        t.frames_per_launch = f;
        t.workgroups = g;
        t.waves_per_cell = w;
        double ms = 0.0;
        int r = run_groups(b, t, ns, fr, pos_dev, dtype, itype_dev, nullptr, &ms);  // warm (captures)
        if (r) return r;
        r = run_groups(b, t, ns, fr, pos_dev, dtype, itype_dev, nullptr, &ms);
        if (r) return r;
        if (ms < best) {
          best = ms;
          bp = t;
        }
      }
  }
  b->p = bp;
  if (chosen) *chosen = bp;
  return TRJ_OK;
}

int trj_batch_sq_layout(const trj_batch_params* p, const trj_batch_sq_params* q, int32_t nframes,
                        const trj_batch_frame* fr, int32_t* nqmax, int64_t* sqf_off, int64_t* sqfp_off,
                        int64_t* total_sqf, int64_t* total_sqfp) {
  if (!p || !q || !fr || nframes < 0) return TRJ_EINVAL;
  std::vector<SqFramePlan> plan;
  int64_t a = 0, c = 0;
  std::string err;
  int r = plan_sq_frames(*p, *q, nframes, fr, plan, &a, &c, &err);
  if (r) return r;
  for (int f = 0; f < nframes; ++f) {
    if (nqmax) nqmax[f] = plan[f].nqmax;
    if (sqf_off) sqf_off[f] = plan[f].out_off;
    if (sqfp_off) sqfp_off[f] = plan[f].outp_off;
  }
  if (total_sqf) *total_sqf = a;
  if (total_sqfp) *total_sqfp = c;
  return TRJ_OK;
}

int trj_batch_sq_run(trj_batch* b, const trj_batch_sq_params* q, int32_t nframes, const trj_batch_frame* fr,
                     const void* pos_dev, int32_t dtype, const int32_t* itype_dev, const double* bsc, double* sqf,
                     double* sqfp, int64_t* nq, double* gpu_ms) {
  if (!b) return TRJ_EINVAL;
  if (!q || nframes < 1 || !fr || !pos_dev || !itype_dev) return bfail(b, TRJ_EINVAL, "batch sq: empty or NULL input");
  (void)hipSetDevice(b->device);
  return run_sq_groups(b, *q, nframes, fr, pos_dev, dtype, itype_dev, bsc, sqf, sqfp, nq, gpu_ms);
}

}  // extern "C"
