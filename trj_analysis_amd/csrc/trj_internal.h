// trj_internal.h -- shared declarations of the trj_hip library (host planning
// and kernel launchers).  Not part of the public ABI.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/trj_hip.h"

namespace trj {

constexpr int kNspMax = 10;  // nspmax, src/core/common.cuf:41
constexpr int kMaxSlices = 16;              // z-slices of the confined analysis
constexpr int kKeyMax = kMaxSlices * kNspMax;  // sort keys: slice * nsp + species

// ---------------------------------------------------------------- RDF plan
struct RdfPlan {
  int ndim = 3, nsp = 1, nit = 1, lsmax = 0;
  int rows = 0;  // pair-type rows of the histogram per LDS tile (pass) of the pair kernels
  double deltar = 0.0;
  double rc = 0.0;       // sticky capped cutoff (netcdf.f90:610-612)
  int periodic[3] = {1, 1, 1};
  // Exact bin edges in r^2: thr[m] = smallest double r2 with
  // nint(sqrt(r2)/deltar) >= m, m = 1..lsmax+1 (thr[0] unused).
  std::vector<double> thr;
  bool floor_bins = false;  // thr holds the edges of int(r/deltar) + 1 (TRJ_RDF_FLOOR_BINS)
  std::vector<double> thr_xy;  // edges of the slice rule int(r/deltar) + 1, when slices are set
};

// rdf.cuf:55-58 after the cap of netcdf.f90:610-612.
int rdf_lsmax(int ndim, const double box[3], double rcrdf, double deltar, double* rc_eff);
// Thresholds so that bin(r2) = max{m : thr[m] <= r2} equals nint(sqrt(r2)/deltar).
std::vector<double> rdf_bin_thresholds(double deltar, int lsmax);
// Edges of the slice rule k = int(r/deltar) + 1, same layout.
std::vector<double> rdf_bin_thresholds_floor(double deltar, int lsmax);
// Smallest d >= 0 with nint(d/L) >= 1 in FP64 (minimum-image switch point).
double minimage_threshold(double L);

// Per-frame geometry (netcdf.f90:602-622).
struct FrameGeom {
  double box[3];
  double lmi[3];   // minimum-image lengths (sidelv)
  double half[3];  // minimage_threshold(lmi)
  double side2;
  double volume;
};
FrameGeom frame_geometry(int ndim, const double box[3], const int periodic[3], double* rc_sticky);

// ---------------------------------------------------------------- S(Q) plan
// Tile of the dense q-box computed by one workgroup of the contraction:
// kx in [kx0, kx0+tx), ky in [ky0, ky0+ty), kz in [kz0, kz0+tz); vectors are
// numbered (i*ty + j)*tz + k inside a tile (kz fastest).
struct TileShape {
  int tx, ty, tz;
  int vec() const { return tx * ty * tz; }
};
// FP64: M = 64 (kx,ky) pairs x N = 32 kz per workgroup.  FP32 / split 16-bit:
// M = 64 kx x N = 128 (ky,kz) pairs, so the phase products ey*ez form the
// operand shared by all four wavefronts.
inline TileShape tile_shape(int precision) {
  if (precision == TRJ_PREC_FP64) return TileShape{8, 8, 32};
  return precision == TRJ_PREC_FP32 ? TileShape{64, 4, 32} : TileShape{64, 8, 32};
}

struct SqTile {
  // blocks: bit (rb * (tz / 16) + cb) is set when row block rb (16 rows of the
  // tile's (kx, ky) rows, kx major) and 16-wide kz column block cb hold at
  // least one summed vector.
  int32_t kx0, ky0, kz0, blocks;
};

struct SqPlan {
  trj_sq_params p{};
  trj_sq_info info{};
  float fk[3], dq, qminp, qmin2;
  int l[3], kext[3];
  int nqmax = 0, nqmin = 0;
  double box[3];                     // box the q-grid was built on
  std::vector<int64_t> nq;           // normalisation counts (0..nqmax)
  std::vector<int64_t> nq_summed;    // vectors actually summed per shell
  TileShape tk{8, 8, 32};            // tile shape of the contraction kernel
  std::vector<SqTile> tiles;         // active tiles of the dense box
  std::vector<int16_t> tile_shell;   // [tile][tk.vec()], -1 = not summed
  std::vector<int32_t> axis_kv;      // axis vectors (x,y,z) triples
  std::vector<int32_t> axis_shell;   // their shells
  int kmin[3], kmax[3];              // inclusive index range of the dense box
  // FP64 and 16-bit tiles: their 8-value phase segments start at
  // anc_base[a] + 8 m, m < anc_rows[a] (FP64: x kx0, y ky0, z kz0 + 0, 8, 16,
  // 24; 16-bit: x kx0 + 0 .. 56, y ky0, z kz0 + 0 .. 24); zero for FP32.
  int anc_rows[3] = {0, 0, 0}, anc_base[3] = {0, 0, 0};
};

// Anchor rows of the FP64 and 16-bit contractions' tiles (tile origins kx0,
// ky0, kz0 at org[t * stride + 0..2]): the 8-value phase segments of the
// tiles start at base[a] + 8 m, m < rows[a].
void anchor_layout(const int32_t* org, int stride, size_t ntiles, const TileShape& tk, int rows[3], int base[3]);
// Bytes of an anchor table with the given rows for n turn entries (FP64:
// 16 B per anchor; 16-bit: 8 B, plus three rows of FP32 unit steps); 0 above
// 8 GiB (the kernel then evaluates the anchors itself).
inline int64_t anchor_table_bytes(int precision, const int rows[3], int64_t n) {
  const int64_t nr = rows[0] + rows[1] + rows[2];
  if (nr == 0 || precision == TRJ_PREC_FP32) return 0;
  const int64_t b = precision == TRJ_PREC_FP64 ? int64_t(16) * nr * n : int64_t(8) * (nr + 3) * n;
  return b <= (int64_t(8) << 30) ? b : 0;
}
// Bytes of the contraction's anchor table for n turn entries: FP64 anchors
// (16 B), or FP32 anchors and three FP32 unit-step rows (8 B each) for the
// 16-bit kernels; 0 when the frame goes without one (FP32, or a table above
// 8 GiB: the kernel then evaluates the anchors itself).
inline int64_t sq_anchor_bytes(const SqPlan& S, int64_t n) { return anchor_table_bytes(S.p.precision, S.anc_rows, n); }

// sq.cuf:156-200 (REF) or the half-space grid; fills everything above.
int sq_plan(const trj_sq_params& p, const double box[3], SqPlan* out, std::string* err);

// K-splits of the S(Q) contraction for ntiles tiles and n atoms.
int contraction_splits(int ntiles, int64_t n, int precision, int nsp, int64_t fp64_span = 0);

// Frees an idle context's device buffers and leaves it uninitialised
// (the request server's way of returning memory to its HBM budget).
int ctx_release(trj_ctx* c);

// ---------------------------------------------------------------- kernels
struct DevFrame;  // defined in trj_ctx.hip

}  // namespace trj
