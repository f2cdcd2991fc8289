// kernels.hip -- gfx950 kernels of the trj_hip library.
//
// RDF: positions are wrapped in FP64 on the device, binned into a cell list
// (counting sort), and each wavefront takes 64 atoms of one home cell and
// walks the half stencil with wave-uniform j loads.  The pair rule of
// rdf.cuf:452-463 is evaluated in FP64 without division or sqrt: the minimum
// image switch point and the bin edges in r^2 are located exactly on the host
// (plan.cpp), so every decision equals the FP64 oracle's.  Counts go to an
// LDS histogram per workgroup (uint32), are written once per workgroup, and
// merged into int64 totals.
//
// S(Q): the dense q-box is tiled 8 (kx) x 8 (ky) x 32 (kz); per tile
// rho_s(q) = sum_j ex_j(kx) ey_j(ky) ez_j(kz) is a complex GEMM with
// M = 64 (kx,ky) pairs, N = 32 kz values and K = atoms of species s, run on
// v_mfma_f64_16x16x4_f64 (4 real products per complex product).  Phases come
// from an exact FP64 reduction of n*u (u = x/L in turns, with the FMA error
// term) and sincospi, extended by an 8-step angle-addition recurrence.
// MFMA accumulators are folded every 128 atoms into FP64 sums with TwoSum
// compensation, so the rounding error of a 1e5-atom sum stays near that of
// one 128-term block even when long runs of atoms add coherently.
//
// Compiled with -ffp-contract=off: every FP64 operation of the RDF pair rule
// rounds once, in the order written, like the oracle.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <utility>

#include "../../include/trj_hip.h"
#include "kernels.h"
#include "trj_internal.h"

namespace trj {

namespace {

constexpr int kWave = 64;

__device__ inline double wrap_periodic(double x, double L) {
  double f = floor(x / L);
  double t = f * L;
  double w = x - t;
  if (w < 0.0) w = w + L;
  if (w >= L) w = w - L;
  return w;
}

__global__ __launch_bounds__(256) void k_ingest(IngestArgs a) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  double c[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k < a.ndim; ++k) {
    const int64_t idx = (a.layout == TRJ_SOA) ? k * a.n + i : i * a.ndim + k;
    c[k] = (a.dtype == TRJ_F32) ? static_cast<double>(static_cast<const float*>(a.raw)[idx])
                                : static_cast<const double*>(a.raw)[idx];
  }
  uint32_t bad = 0;
  for (int k = 0; k < a.ndim; ++k) {
    if (!isfinite(c[k])) {
      bad |= kFlagRange;
    } else if (a.periodic[k]) {
      c[k] = wrap_periodic(c[k], a.box[k]);
    } else if (k != a.free_axis && (c[k] < -0.25 * a.box[k] || c[k] >= 1.25 * a.box[k])) {
      bad |= kFlagRange;
    }
  }
  const int t = a.itype[i];
  if (t < 1 || t > a.nsp) bad |= kFlagType;
  if (bad) atomicOr(a.flags, bad);
  a.pos[i] = c[0];
  a.pos[a.n + i] = c[1];
  a.pos[2 * a.n + i] = c[2];
}

// The sort key of atom i: key[i] when given (slice * nsp + species, -1 for
// atoms outside every slice), else its species 0..nsp-1 (-1 when invalid).
__device__ inline int sort_key(const int32_t* itype, const int32_t* key, int64_t i, int64_t n, int nkeys) {
  if (i >= n) return -1;
  int s = key ? key[i] : itype[i] - 1;
  return (s < 0 || s >= nkeys) ? -1 : s;
}

__global__ __launch_bounds__(256) void k_species_count(const int32_t* __restrict__ itype,
                                                       const int32_t* __restrict__ key, int64_t n,
                                                       int nsp, int32_t* __restrict__ blk, int nblk) {
  __shared__ int cnt[4][kKeyMax];
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  const int s = sort_key(itype, key, i, n, nsp);
  for (int sp = 0; sp < nsp; ++sp) {
    const unsigned long long m = __ballot(s == sp);
    if (lane == 0) cnt[w][sp] = __popcll(m);
  }
  __syncthreads();
  for (int sp = threadIdx.x; sp < nsp; sp += 256) {
    blk[static_cast<int64_t>(sp) * nblk + blockIdx.x] = cnt[0][sp] + cnt[1][sp] + cnt[2][sp] + cnt[3][sp];
  }
}

// Single-workgroup exclusive scan (count up to a few 1e5 entries): 4096
// entries a step, four consecutive ones per thread (a serial scan in
// registers), then the 64 lanes of each wavefront by shuffles and the 16
// wavefront totals through LDS.
__global__ __launch_bounds__(1024) void k_exclusive_scan(const int32_t* __restrict__ in,
                                                         int32_t* __restrict__ out, int64_t count,
                                                         const int32_t* __restrict__ count_dev) {
  __shared__ int32_t wsum[16];
  if (count_dev) count = *count_dev;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const bool vec = ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 15) == 0;
  int32_t carry = 0;
  for (int64_t base = 0; base < count; base += 4096) {
    const int64_t i = base + 4 * static_cast<int64_t>(t);
    const bool whole = vec && i + 3 < count;
    int32_t v[4];
    if (whole) {
      const int4 q = *reinterpret_cast<const int4*>(in + i);
      v[0] = q.x;
      v[1] = q.y;
      v[2] = q.z;
      v[3] = q.w;
    } else {
      for (int k = 0; k < 4; ++k) v[k] = i + k < count ? in[i + k] : 0;
    }
    const int32_t s1 = v[0], s2 = s1 + v[1], s3 = s2 + v[2], s4 = s3 + v[3];
    int32_t x = s4;
    for (int off = 1; off < 64; off <<= 1) {
      const int32_t y = __shfl_up(x, off);
      if (lane >= off) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    int32_t wb = 0, tot = 0;
    for (int k = 0; k < 16; ++k) {
      const int32_t sk = wsum[k];
      wb += k < w ? sk : 0;
      tot += sk;
    }
    const int32_t ex = carry + wb + x - s4;
    if (whole) {
      *reinterpret_cast<int4*>(out + i) = make_int4(ex, ex + s1, ex + s2, ex + s3);
    } else {
      const int32_t e[4] = {ex, ex + s1, ex + s2, ex + s3};
      for (int k = 0; k < 4; ++k)
        if (i + k < count) out[i + k] = e[k];
    }
    carry += tot;
    __syncthreads();
  }
}

// sp_off[0..nkeys]: key ranges of the sorted entries (M = sp_off[nkeys]
// members).  With ngroups > 0 (charges) the charge channels follow: group g
// is keys [g*gsize, (g+1)*gsize) again at M + their offsets, so
// sp_off[nkeys + 1 + g] = M + sp_off[(g+1)*gsize].
__global__ void k_species_offsets(const int32_t* __restrict__ blkoff, const int32_t* __restrict__ blk,
                                  int nblk, int nkeys, int gsize, int ngroups, int32_t* __restrict__ sp_off) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int s = 0; s < nkeys; ++s) sp_off[s] = blkoff[static_cast<int64_t>(s) * nblk];
  const int64_t last = static_cast<int64_t>(nkeys) * nblk - 1;
  const int32_t m = blkoff[last] + blk[last];
  sp_off[nkeys] = m;
  for (int g = 0; g < ngroups; ++g) sp_off[nkeys + 1 + g] = m + ((g + 1) * gsize < nkeys ? sp_off[(g + 1) * gsize] : m);
}

// Phase in turns, frac(n*(uh + ul)), reduced exactly: the product's rounding
// error (FMA) and the low part of u are added after the integer part is gone.
__device__ inline double frac_turns(int nk, double uh, double ul) {
  const double dn = static_cast<double>(nk);
  const double hi = dn * uh;
  const double lo = fma(dn, uh, -hi) + dn * ul;
  return (hi - rint(hi)) + lo;
}

__global__ __launch_bounds__(256) void k_species_scatter(const int32_t* __restrict__ itype,
                                                         const int32_t* __restrict__ key, int64_t n,
                                                         int nsp, const int32_t* __restrict__ blkoff,
                                                         int nblk, const double* __restrict__ pos,
                                                         double lq0, double lq1, double lq2, int ndim,
                                                         const double* __restrict__ q,
                                                         const int32_t* __restrict__ sp_off, int64_t stride,
                                                         double* __restrict__ u) {
  __shared__ int cnt[4][kKeyMax];
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  const int s = sort_key(itype, key, i, n, nsp);
  unsigned long long mine = 0;
  for (int sp = 0; sp < nsp; ++sp) {
    const unsigned long long m = __ballot(s == sp);
    if (lane == 0) cnt[w][sp] = __popcll(m);
    if (s == sp) mine = m;
  }
  __syncthreads();
  // An atom without a key (type outside 1..nsp, flagged by the ingest, or
  // outside every slice) has no entry.
  if (s < 0) return;
  int off = blkoff[static_cast<int64_t>(s) * nblk + blockIdx.x];
  for (int ww = 0; ww < w; ++ww) off += cnt[ww][s];
  off += __popcll(mine & ((1ull << lane) - 1ull));
  // Turns u = x/L as an unevaluated sum hi + lo: hi rounded, lo = (x - hi*L)/L
  // with the remainder exact through the FMA; then the unit step e^{2 pi i u}
  // of the phase recurrences as cos, sin.  With charges the same values also
  // go to entry M + off of the charge channels (M = members), and plane 12
  // holds the weights (1, and q_i for the charge entry).
  const int64_t qo = q ? static_cast<int64_t>(sp_off[nsp]) + off : 0;
  const double lq[3] = {lq0, lq1, lq2};
  for (int k = 0; k < 3; ++k) {
    const double x = (k < ndim) ? pos[k * n + i] : 0.0;
    const double hi = x / lq[k];
    const double lo = fma(-hi, lq[k], x) / lq[k];
    double sn, cs;
    sincospi(2.0 * frac_turns(1, hi, lo), &sn, &cs);
    u[k * stride + off] = hi;
    u[(3 + k) * stride + off] = lo;
    u[(6 + k) * stride + off] = cs;
    u[(9 + k) * stride + off] = sn;
    if (q) {
      u[k * stride + qo] = hi;
      u[(3 + k) * stride + qo] = lo;
      u[(6 + k) * stride + qo] = cs;
      u[(9 + k) * stride + qo] = sn;
    }
  }
  if (q) {
    u[12 * stride + off] = 1.0;
    u[12 * stride + qo] = q[i];
  }
}

// ------------------------------------------------------------------ slices
// Slice of each atom (rdf.cuf:692-696, sq.cuf:640-643) from its ingested z in
// FP32, as the reference: the first nearest zslice (minloc), member when
// |zslice - z| <= zgrid/2.  key = slice * nsp + species - 1, or -1.  last
// gets max(species << 32 | i) over valid atoms: the last atom of the
// species-sorted order, which rdfxy's thread loop leaves out (rdf.cuf:688).
__global__ __launch_bounds__(256) void k_slice_keys(const double* __restrict__ pos, int64_t n,
                                                    const int32_t* __restrict__ itype, int nsp, SliceSet ss,
                                                    int32_t* __restrict__ key,
                                                    unsigned long long* __restrict__ last) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  const int t = itype[i];
  const float z = static_cast<float>(pos[2 * n + i]);
  int best = 0;
  float bd = fabsf(ss.z[0] - z);
  for (int s = 1; s < ss.n; ++s) {
    const float d = fabsf(ss.z[s] - z);
    if (d < bd) {
      bd = d;
      best = s;
    }
  }
  const bool ok = t >= 1 && t <= nsp;
  key[i] = (ok && bd <= ss.zgrid / 2.0f) ? best * nsp + (t - 1) : -1;
  if (ok) atomicMax(last, (static_cast<unsigned long long>(t) << 32) | static_cast<unsigned long long>(i));
}

// Per-slice counts of the pair normalisation: members per slice and species
// without the species-sorted last atom (rdf.cuf:699-700); qsum/qcnt (q not
// NULL): sum of q^2 and the count over all members (common.cuf:218-232).
__global__ __launch_bounds__(256) void k_slice_counts(const int32_t* __restrict__ key, int64_t n,
                                                      const unsigned long long* __restrict__ last, int nsp,
                                                      int32_t* __restrict__ count, int32_t* __restrict__ countp,
                                                      const double* __restrict__ q, double* __restrict__ qsum,
                                                      int32_t* __restrict__ qcnt) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  const int k = key[i];
  if (k < 0) return;
  const int sl = k / nsp;
  if (static_cast<int64_t>(*last & 0xffffffffull) != i) {
    atomicAdd(&count[sl], 1);
    atomicAdd(&countp[k], 1);
  }
  if (q) {
    atomicAdd(&qsum[sl], q[i] * q[i]);
    atomicAdd(&qcnt[sl], 1);
  }
}

// Cell of each atom for the slice pairs: 2D cells of slice sl numbered
// sl * ncell2 + cy * ncx + cx; atoms outside every slice go to cell
// nslice * ncell2 (never visited).
__device__ __forceinline__ int slice_cell(const double* __restrict__ pos, const int32_t* __restrict__ key,
                                          int64_t n, int64_t i, int nsp, const CellGrid& g, int nslice) {
  const int ncell2 = g.nc[0] * g.nc[1];
  const int k = key[i];
  if (k < 0) return nslice * ncell2;
  int c[2];
  for (int d = 0; d < 2; ++d) {
    const int ck = static_cast<int>(floor(pos[d * n + i] * g.inv_h[d]));
    c[d] = ck < 0 ? 0 : (ck >= g.nc[d] ? g.nc[d] - 1 : ck);
  }
  return (k / nsp) * ncell2 + c[1] * g.nc[0] + c[0];
}

// Cell of each atom of a frame: (cz * ncy + cy) * ncx + cx, coordinates
// clamped into the grid.
__device__ __forceinline__ int frame_cell(const double* __restrict__ pos, int64_t n, int64_t i, int ndim,
                                          const CellGrid& g) {
  int c[3] = {0, 0, 0};
  for (int k = 0; k < ndim; ++k) {
    const int ck = static_cast<int>(floor(pos[k * n + i] * g.inv_h[k]));
    c[k] = ck < 0 ? 0 : (ck >= g.nc[k] ? g.nc[k] - 1 : ck);
  }
  return (c[2] * g.nc[1] + c[1]) * g.nc[0] + c[0];
}

__global__ __launch_bounds__(256) void k_cell_count_slices(const double* __restrict__ pos,
                                                           const int32_t* __restrict__ key, int64_t n, int nsp,
                                                           CellGrid g, int nslice,
                                                           int32_t* __restrict__ cellid,
                                                           int32_t* __restrict__ count) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  const int id = slice_cell(pos, key, n, i, nsp, g, nslice);
  cellid[i] = id;
  atomicAdd(&count[id], 1);
}

// The frame's slice histograms into the accumulators (rdf.cuf:284-296):
// counts += frame, hxy += frame / (countp_j countp_l) when both are positive;
// with charges (gq not NULL) gqq += (fixed / scale) / (count^2 q2) with q2
// the slice's mean q^2 (rdf.cuf:262-266 divides by countslice^2 * avq2).  The
// frame buffers are cleared for the next frame.
__global__ __launch_bounds__(256) void k_xy_finish(int nslice, int nsp, int lsmax,
                                                   int64_t* __restrict__ frame, const int32_t* __restrict__ count,
                                                   const int32_t* __restrict__ countp,
                                                   int64_t* __restrict__ counts, double* __restrict__ hxy,
                                                   unsigned long long* __restrict__ gq, const double* __restrict__ qp,
                                                   const double* __restrict__ qsum, const int32_t* __restrict__ qcnt,
                                                   double* __restrict__ gqacc) {
  const int nit = nsp * (nsp + 1) / 2;
  const int64_t total = static_cast<int64_t>(nslice) * nit * lsmax;
  const int64_t e = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (e < total) {
    const int sl = static_cast<int>(e / (static_cast<int64_t>(nit) * lsmax));
    const int ij = static_cast<int>((e / lsmax) % nit);
    int j = 0, rem = ij;
    while (rem >= nsp - j) {
      rem -= nsp - j;
      ++j;
    }
    const int l = j + rem;
    const int64_t f = frame[e];
    counts[e] += f;
    const double cj = countp[sl * nsp + j], cl = countp[sl * nsp + l];
    if (cj > 0.0 && cl > 0.0) hxy[e] += static_cast<double>(f) / (cj * cl);
    frame[e] = 0;
  }
  if (gq && e < static_cast<int64_t>(nslice) * lsmax) {
    const int sl = static_cast<int>(e / lsmax);
    const long long f = static_cast<long long>(gq[e]);
    const double c = count[sl];
    const double q2 = qcnt[sl] > 0 ? qsum[sl] / qcnt[sl] : 0.0;
    if (c > 0.0 && q2 > 0.0) gqacc[e] += (static_cast<double>(f) / qp[0]) / (c * c * q2);
    gq[e] = 0ull;
  }
}

__global__ __launch_bounds__(256) void k_cell_count(const double* __restrict__ pos, int64_t n, int ndim,
                                                    CellGrid g, int32_t* __restrict__ cellid,
                                                    int32_t* __restrict__ count) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  const int id = frame_cell(pos, n, i, ndim, g);
  cellid[i] = id;
  atomicAdd(&count[id], 1);
}

__global__ __launch_bounds__(256) void k_cell_scatter(const double* __restrict__ pos,
                                                      const int32_t* __restrict__ itype,
                                                      const int32_t* __restrict__ cellid, int64_t n,
                                                      const int32_t* __restrict__ start,
                                                      int32_t* __restrict__ cursor,
                                                      double* __restrict__ spos,
                                                      int32_t* __restrict__ stype,
                                                      const double* __restrict__ q,
                                                      double* __restrict__ sq) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  const int c = cellid[i];
  const int slot = start[c] + atomicAdd(&cursor[c], 1);
  const double x = pos[i], y = pos[n + i], z = pos[2 * n + i];
  const int t = itype[i];
  spos[slot] = x;
  spos[n + slot] = y;
  spos[2 * n + slot] = z;
  stype[slot] = t;
  if (q) sq[slot] = q[i];
}

// Cell sort without global atomics.  Workgroup b takes atoms
// [b * chunk, (b + 1) * chunk) and counts them per cell in LDS, two 16-bit
// counters to a word (chunk <= 65535, so a counter never carries into its
// neighbour); the value an atom's LDS atomic returns is its rank among the
// chunk's atoms in its cell.  blkcnt[b * ncell + c] receives the chunk's
// count in cell c.  With key (slice pairs) the cells are slice_cell's, else
// frame_cell's.
__global__ __launch_bounds__(1024) void k_cell_hist(const double* __restrict__ pos,
                                                    const int32_t* __restrict__ key, int nsp, int nslice,
                                                    int64_t n, int ndim, CellGrid g, int ncell, int64_t chunk,
                                                    int32_t* __restrict__ rank, int32_t* __restrict__ blkcnt) {
  extern __shared__ uint32_t hcell[];
  const int nw = (ncell + 1) >> 1;
  for (int k = threadIdx.x; k < nw; k += 1024) hcell[k] = 0u;
  __syncthreads();
  const int64_t i0 = static_cast<int64_t>(blockIdx.x) * chunk;
  const int64_t i1 = i0 + chunk < n ? i0 + chunk : n;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += 1024) {
    const int id = key ? slice_cell(pos, key, n, i, nsp, g, nslice) : frame_cell(pos, n, i, ndim, g);
    const int sh = (id & 1) << 4;
    const uint32_t old = atomicAdd(&hcell[id >> 1], 1u << sh);
    rank[i] = static_cast<int32_t>((old >> sh) & 0xffffu);
  }
  __syncthreads();
  int32_t* row = blkcnt + static_cast<int64_t>(blockIdx.x) * ncell;
  for (int c = threadIdx.x; c < ncell; c += 1024) row[c] = static_cast<int32_t>((hcell[c >> 1] >> ((c & 1) << 4)) & 0xffffu);
}

// Per cell c (a lane each, 64 cells a workgroup): the exclusive prefix over
// the workgroups of blkcnt[. * ncell + c], written back in place, and the
// cell's total to count[c]; count[ncell] = 0 so that the exclusive scan of
// count[0..ncell] ends at n.  Each of the 16 wavefronts takes up to 16 of
// the <= 256 workgroup rows.
__global__ __launch_bounds__(1024) void k_cell_colscan(int32_t* __restrict__ blkcnt, int nblk, int ncell,
                                                       int32_t* __restrict__ count) {
  __shared__ int32_t part[16][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int c = blockIdx.x * 64 + lane;
  const int rpw = (nblk + 15) >> 4;
  const int b0 = w * rpw;
  int32_t v[16];
  int32_t s = 0;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int b = b0 + k;
    v[k] = (k < rpw && b < nblk && c < ncell) ? blkcnt[static_cast<int64_t>(b) * ncell + c] : 0;
    s += v[k];
  }
  part[w][lane] = s;
  __syncthreads();
  int32_t base = 0, tot = 0;
  for (int k = 0; k < 16; ++k) {
    const int32_t p = part[k][lane];
    base += k < w ? p : 0;
    tot += p;
  }
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int b = b0 + k;
    if (k < rpw && b < nblk && c < ncell) blkcnt[static_cast<int64_t>(b) * ncell + c] = base;
    base += v[k];
  }
  if (w == 0 && c <= ncell) count[c] = c < ncell ? tot : 0;
}

// Atom i (of counting workgroup b = i / chunk) goes to slot start[c] +
// blkcnt[b * ncell + c] + rank[i] of its cell c (recomputed from its
// position, as k_cell_hist found it): positions, types and charges in cell
// order.  One thread per atom over the whole grid: the counting workgroups
// are few when the cells are many (their LDS counters fill a CU), and the
// placement's scattered writes want every CU.
__global__ __launch_bounds__(256) void k_cell_place(const double* __restrict__ pos,
                                                    const int32_t* __restrict__ key, int nsp, int nslice,
                                                    int64_t n, int ndim, CellGrid g, int ncell, int64_t chunk,
                                                    const int32_t* __restrict__ rank,
                                                    const int32_t* __restrict__ blkcnt,
                                                    const int32_t* __restrict__ start,
                                                    const int32_t* __restrict__ itype, double* __restrict__ spos,
                                                    int32_t* __restrict__ stype, const double* __restrict__ q,
                                                    double* __restrict__ sq) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * 256) {
    const int c = key ? slice_cell(pos, key, n, i, nsp, g, nslice) : frame_cell(pos, n, i, ndim, g);
    const int32_t* base = blkcnt + (i / chunk) * ncell;
    const int slot = start[c] + base[c] + rank[i];
    const double x = pos[i], y = pos[n + i], z = pos[2 * n + i];
    spos[slot] = x;
    spos[n + slot] = y;
    spos[2 * n + slot] = z;
    stype[slot] = itype[i];
    if (q) sq[slot] = q[i];
  }
}

// rdf.cuf:748-752 for i <= j (callers order the pair).
__device__ inline int pair_index(int a, int b, int nsp) {
  return (a - 1) * (nsp - 1) - (a - 1) * (a - 2) / 2 + b;
}

// nint(sqrt(r2)/deltar) through the exact edges thr[0..lsmax+2]
// (thr[0] = 0, thr[lsmax+2] = +inf).  The float estimate is within one bin of
// the exact one (its relative error is ~1e-7 and r/deltar < 2^15), so one
// comparison against each neighbouring edge corrects it.
__device__ inline int bin_of(double r2, const double* thr, float inv_dr, int lsmax) {
  int k = static_cast<int>(sqrtf(static_cast<float>(r2)) * inv_dr + 0.5f);
  k = k > lsmax + 1 ? lsmax + 1 : k;
  const double lo = thr[k], hi = thr[k + 1];
  return k + (r2 >= hi ? 1 : 0) - (r2 < lo ? 1 : 0);
}

__device__ inline double min_image(double d, double L, double h) {
  return (d >= h) ? d - L : ((d <= -h) ? d + L : d);
}

struct Stencil {
  int n;
  int off[14][3];
};

// Number fluctuations around fixed random points (s2D3_sh / s2D2_sh,
// rdf.cuf:489-615): block p counts the atoms in bins k = nint(r/dr) (r^2 <
// side2, the RDF's rule and edges) around point p, then adds the cumulative
// counts N_k = sum_{b <= k} h_b and N_k^2 to gi / g2i (int64, exact).  With
// start (cell-sorted atoms) only the 3^ndim cells around the point are read.
__global__ __launch_bounds__(256) void k_s2n(S2nArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  uint32_t* hist = reinterpret_cast<uint32_t*>(smem);
  uint32_t* part = hist + ((a.lsmax + 1 + 3) & ~3);
  const int nb = a.lsmax + 1;
  for (int b = threadIdx.x; b < nb; b += 256) hist[b] = 0u;
  __syncthreads();
  const int p = blockIdx.x;
  double xp[3];
  for (int k = 0; k < 3; ++k) xp[k] = k < a.ndim ? a.box[k] * a.pts[static_cast<int64_t>(k) * a.np + p] : 0.0;
  const double* __restrict__ px = a.pos;
  const double* __restrict__ py = a.pos + a.n;
  const double* __restrict__ pz = a.pos + 2 * a.n;
  auto visit = [&](int64_t j0, int64_t j1) {
    for (int64_t j = j0 + threadIdx.x; j < j1; j += 256) {
      const double dx = min_image(xp[0] - px[j], a.lmi[0], a.half[0]);
      const double dy = min_image(xp[1] - py[j], a.lmi[1], a.half[1]);
      double r2 = dx * dx + dy * dy;
      if (a.ndim == 3) {
        const double dz = min_image(xp[2] - pz[j], a.lmi[2], a.half[2]);
        r2 += dz * dz;
      }
      if (r2 < a.side2) {
        const int k = bin_of(r2, a.thr, a.inv_dr, a.lsmax);
        if (k <= a.lsmax) atomicAdd(&hist[k], 1u);
      }
    }
  };
  if (a.start) {
    int c[3] = {0, 0, 0};
    for (int k = 0; k < a.ndim; ++k) {
      int ck = static_cast<int>(floor(xp[k] * a.g.inv_h[k]));
      c[k] = ck < 0 ? 0 : (ck >= a.g.nc[k] ? a.g.nc[k] - 1 : ck);
    }
    const int zr = a.ndim == 3 ? 1 : 0;
    for (int dz = -zr; dz <= zr; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          int nx = c[0] + dx, ny = c[1] + dy, nz = c[2] + dz;
          if (a.g.periodic[0]) nx = (nx + a.g.nc[0]) % a.g.nc[0]; else if (nx < 0 || nx >= a.g.nc[0]) continue;
          if (a.g.periodic[1]) ny = (ny + a.g.nc[1]) % a.g.nc[1]; else if (ny < 0 || ny >= a.g.nc[1]) continue;
          if (a.ndim == 3) {
            if (a.g.periodic[2]) nz = (nz + a.g.nc[2]) % a.g.nc[2]; else if (nz < 0 || nz >= a.g.nc[2]) continue;
          }
          const int cell = (nz * a.g.nc[1] + ny) * a.g.nc[0] + nx;
          visit(a.start[cell], a.start[cell + 1]);
        }
  } else {
    visit(0, a.n);
  }
  __syncthreads();
  // Cumulative counts: each thread sums a contiguous run of bins, thread 0
  // scans the 256 run totals, then each thread finishes its run.
  const int per = (nb + 255) / 256;
  const int b0 = threadIdx.x * per, b1 = min(nb, b0 + per);
  uint32_t run = 0;
  for (int b = b0; b < b1; ++b) run += hist[b];
  part[threadIdx.x] = run;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t acc = 0;
    for (int t = 0; t < 256; ++t) {
      const uint32_t v = part[t];
      part[t] = acc;
      acc += v;
    }
  }
  __syncthreads();
  unsigned long long nk = part[threadIdx.x];
  for (int b = b0; b < b1; ++b) {
    nk += hist[b];
    atomicAdd(&a.gi[b], nk);
    atomicAdd(&a.g2i[b], nk * nk);
  }
}

// ------------------------------------------------------------------ F(Q,t)
// FqtnD (dynamics.cuf:421-508): per q-vector v and chunk of atoms, block
// z = 0 sums the current bsc-weighted phases (rho(q,t)), block z = 1 + a sums
// cos(q.(r(t) - r(t0))) against active origin a's stored positions.  Phases
// in FP64: turns t = k.r/L reduced by rint before sincospi.
__device__ inline double fqt_turns(const int* k, const double* x, int64_t j, int64_t stride, const double* invl,
                                   int ndim) {
  double t = k[0] * (x[j] * invl[0]) + k[1] * (x[stride + j] * invl[1]);
  if (ndim == 3) t += k[2] * (x[2 * stride + j] * invl[2]);
  return t - rint(t);
}

__global__ __launch_bounds__(256) void k_fqt_partial(FqtArgs a) {
  __shared__ double red[2][256];
  const int v = blockIdx.x, chunk = blockIdx.y, z = blockIdx.z;
  const int k[3] = {a.kv[3 * v], a.kv[3 * v + 1], a.kv[3 * v + 2]};
  const int64_t j0 = a.n * chunk / kFqtChunks, j1 = a.n * (chunk + 1) / kFqtChunks;
  double pr = 0.0, pi = 0.0;
  if (z == 0) {
    for (int64_t j = j0 + threadIdx.x; j < j1; j += 256) {
      const int t = a.itype[j];
      const double w = (t >= 1 && t <= a.nsp) ? a.bsc.b[t - 1] : 0.0;
      double sn, cs;
      sincospi(2.0 * fqt_turns(k, a.pos, j, a.n, a.invl, a.ndim), &sn, &cs);
      pr += w * cs;
      pi += w * sn;
    }
  } else {
    const double* r0 = a.r0 + static_cast<int64_t>(a.act[z - 1]) * 3 * a.n;
    for (int64_t j = j0 + threadIdx.x; j < j1; j += 256) {
      const double f = fqt_turns(k, a.pos, j, a.n, a.invl, a.ndim) - fqt_turns(k, r0, j, a.n, a.invl, a.ndim);
      pr += cospi(2.0 * f);
    }
  }
  red[0][threadIdx.x] = pr;
  red[1][threadIdx.x] = pi;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (static_cast<int>(threadIdx.x) < off) {
      red[0][threadIdx.x] += red[0][threadIdx.x + off];
      red[1][threadIdx.x] += red[1][threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double* o = a.part + ((static_cast<int64_t>(v) * gridDim.z + z) * kFqtChunks + chunk) * 2;
    o[0] = red[0][0];
    o[1] = red[1][0];
  }
}

// Per vector: rho(q,t) and the self sums from the chunks in a fixed order;
// rho0 of origins starting now is set to rho(q,t) first (the reference stores
// r0 = r at the origin frame, dynamics.cuf:202-206).
__global__ void k_fqt_vectors(FqtArgs a) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= a.nvec) return;
  const int nz = a.nact + 1;
  double* o = a.vsum + static_cast<int64_t>(v) * nz * 2;
  for (int z = 0; z < nz; ++z) {
    double sr = 0.0, si = 0.0;
    for (int c = 0; c < kFqtChunks; ++c) {
      const double* p = a.part + ((static_cast<int64_t>(v) * nz + z) * kFqtChunks + c) * 2;
      sr += p[0];
      si += p[1];
    }
    o[2 * z] = sr;
    o[2 * z + 1] = si;
  }
  for (int b = 0; b < a.nbuffer; ++b)
    if (a.start_now[b]) {
      a.rho0[(static_cast<int64_t>(b) * a.nvec + v) * 2] = o[0];
      a.rho0[(static_cast<int64_t>(b) * a.nvec + v) * 2 + 1] = o[1];
    }
}

// Per (|q| bin l, active origin a): fkt[b][l][indx] += sum over the bin's
// vectors of Re(rho0 conj(rho)) and fskt[b][l][indx] += their self sums
// (dynamics.cuf:500-501), vectors in a fixed order.
__global__ void k_fqt_bins(FqtArgs a) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= a.nqw * a.nact) return;
  const int l = e % a.nqw, ai = e / a.nqw;
  const int b = a.act[ai], indx = a.act_indx[ai];
  const int nz = a.nact + 1;
  double f = 0.0, fs = 0.0;
  for (int v = a.bin_off[l]; v < a.bin_off[l + 1]; ++v) {
    const double* o = a.vsum + static_cast<int64_t>(v) * nz * 2;
    const double* r0 = a.rho0 + (static_cast<int64_t>(b) * a.nvec + v) * 2;
    f += r0[0] * o[0] + r0[1] * o[1];
    fs += o[2 * (ai + 1)];
  }
  const int64_t idx = (static_cast<int64_t>(b) * a.nqw + l) * a.ndyn + (indx - 1);
  a.fkt[idx] += f;
  a.fskt[idx] += fs;
}

// A restarting origin's sums into the totals, then cleared (dynamics.cuf:
// 225-238).
__global__ void k_fqt_collect(int b, int nqw, int ndyn, double* fkt, double* fskt, double* tot, double* tots) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nqw * ndyn) return;
  const int64_t i = static_cast<int64_t>(b) * nqw * ndyn + e;
  tot[e] += fkt[i];
  tots[e] += fskt[i];
  fkt[i] = 0.0;
  fskt[i] = 0.0;
}

// Returns the bin the pair was added to (0: none).
__device__ inline int pair_accumulate(const PairArgs& a, const double* thr, uint32_t* hist,
                                      bool ok, double xi, double yi, double zi, int ti, double xj,
                                      double yj, double zj, int tj) {
  const double dx = min_image(xj - xi, a.lmi[0], a.half[0]);
  const double dy = min_image(yj - yi, a.lmi[1], a.half[1]);
  double r2;
  if (a.ndim == 3) {
    const double dz = min_image(zj - zi, a.lmi[2], a.half[2]);
    r2 = dx * dx + dy * dy + dz * dz;
  } else {
    r2 = dx * dx + dy * dy;
  }
  if (ok && r2 < a.side2) {
    const int k = bin_of(r2, thr, a.inv_dr, a.lsmax);
    if (k >= 1 && k <= a.lsmax) {
      const int lo = ti < tj ? ti : tj, hi = ti < tj ? tj : ti;
      const int row = pair_index(lo, hi, a.nsp) - 1 - a.row0;
      if (static_cast<unsigned>(row) < static_cast<unsigned>(a.nrows)) atomicAdd(&hist[row * a.lsmax + (k - 1)], 1u);
      return k;
    }
  }
  return 0;
}

// q_i q_j rounded once to the frame's fixed point (exact power-of-two scale),
// as the two's complement bits of an int64 (added with unsigned atomics).
__device__ inline unsigned long long qq_fixed(double qq, double scale) {
  return static_cast<unsigned long long>(static_cast<long long>(rint(qq * scale)));
}

__device__ inline void hist_setup(const PairArgs& a, uint32_t* hist, double* thr) {
  const int nbins = a.nrows * a.lsmax;
  for (int b = threadIdx.x; b < nbins; b += blockDim.x) hist[b] = 0u;
  for (int b = threadIdx.x; b < a.lsmax + 3; b += blockDim.x) thr[b] = a.thr[b];
  __syncthreads();
}

// The block's tile (rows row0 .. row0+nrows-1) into its nit x lsmax slice.
__device__ inline void hist_flush(const PairArgs& a, const uint32_t* hist, uint32_t* partial) {
  __syncthreads();
  const int nbins = a.nrows * a.lsmax;
  uint32_t* out = partial + static_cast<int64_t>(blockIdx.x) * a.nit * a.lsmax + static_cast<int64_t>(a.row0) * a.lsmax;
  for (int b = threadIdx.x; b < nbins; b += blockDim.x) out[b] = hist[b];
}

// Broadcast lane l's value to the whole wave (result is wave-uniform, held in
// scalar registers).
__device__ inline double readlane_d(double v, int l) {
  const unsigned long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane(static_cast<int>(b & 0xffffffffull), l);
  const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), l);
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

// In-range pairs are not binned where they are found: a wavefront appends
// (r2, types) to its own ring in LDS (ballot + mbcnt) and bins 64 of them at
// a time with every lane busy, so the binning and the LDS atomic cost one
// pass per 64 hits instead of a masked pass per candidate.
constexpr int kHitQ = 128;  // ring entries per wavefront (holds < 64 + 64)


// Bin entries head .. head+cnt-1 (cnt <= 64) of a wavefront's ring, one per
// lane.  QQ: the ring also holds q_i q_j (qqr), added to the block's
// fixed-point charge histogram gqs.
template <bool QQ = false>
__device__ inline void bin_hits(const PairArgs& a, const double* thr, uint32_t* hist, const double* qr2,
                                const int* qtt, uint32_t head, int cnt, int lane, const double* qqr = nullptr,
                                unsigned long long* gqs = nullptr, double qsc = 0.0) {
  if (lane < cnt) {
    const uint32_t e = (head + lane) & (kHitQ - 1);
    const double r2 = qr2[e];
    const int tt = qtt[e];
    const int ti = tt >> 8, tj = tt & 255;
    const int k = bin_of(r2, thr, a.inv_dr, a.lsmax);
    if (k >= 1 && k <= a.lsmax) {
      const int lo = ti < tj ? ti : tj, hi = ti < tj ? tj : ti;
      const int row = pair_index(lo, hi, a.nsp) - 1 - a.row0;
      if (static_cast<unsigned>(row) < static_cast<unsigned>(a.nrows)) atomicAdd(&hist[row * a.lsmax + (k - 1)], 1u);
      // the charge sums are per bin, for every pair type: the first pass adds them
      if constexpr (QQ)
        if (a.row0 == 0) atomicAdd(&gqs[k - 1], qq_fixed(qqr[e], qsc));
    }
  }
}

// Candidates of one neighbour cell for the 64 home atoms of a wavefront.
// Atoms outside the cell ranges or with types outside 1..nsp carry NaN
// coordinates, so r2 < side2 fails for them without a separate test.
template <bool SHIFT, bool D3, bool SELF, bool IMG, bool QQ = false>
__device__ inline void rdf_neighbours(const PairArgs& a, const double* thr, uint32_t* hist, double* qr2, int* qtt,
                                      uint32_t& head, uint32_t& tail, int lane, int js, int je, int gi,
                                      double xi, double yi, double zi, int ti8, double shx, double shy,
                                      double shz, const double* __restrict__ sx, const double* __restrict__ sy,
                                      const double* __restrict__ sz, const int32_t* __restrict__ stype,
                                      double qi = 0.0, const double* __restrict__ sq = nullptr,
                                      double* qqr = nullptr, unsigned long long* gqs = nullptr,
                                      double qsc = 0.0) {
  const double nan = __builtin_nan("");
  for (int jb = js; jb < je; jb += kWave) {
    const int jl = jb + lane;
    int tl = jl < je ? stype[jl] : 0;
    const bool vl = static_cast<unsigned>(tl - 1) < static_cast<unsigned>(a.nsp);
    tl = vl ? tl : 0;
    const double xl = vl ? sx[jl] : nan;
    const double yl = vl ? sy[jl] : nan;
    const double zl = (D3 && vl) ? sz[jl] : nan;
    double ql = 0.0;
    if constexpr (QQ) ql = vl ? sq[jl] : 0.0;
    const int cnt = (je - jb) < kWave ? (je - jb) : kWave;
    for (int jj = 0; jj < cnt; ++jj) {
      double r2;
      if constexpr (SHIFT) {
        // d - L*nint(d/L) == d - s for every pair of these two cells
        // (s = -shift, exact: L*1 = L); no per-pair test.
        // (IMG false: all shifts are 0 and d - 0 == d exactly.)
        double dx = readlane_d(xl, jj) - xi, dy = readlane_d(yl, jj) - yi;
        if constexpr (IMG) {
          dx = dx - shx;
          dy = dy - shy;
        }
        r2 = dx * dx + dy * dy;
        if constexpr (D3) {
          double dz = readlane_d(zl, jj) - zi;
          if constexpr (IMG) dz = dz - shz;
          r2 = r2 + dz * dz;
        }
      } else {
        const double dx = min_image(readlane_d(xl, jj) - xi, a.lmi[0], a.half[0]);
        const double dy = min_image(readlane_d(yl, jj) - yi, a.lmi[1], a.half[1]);
        r2 = dx * dx + dy * dy;
        if constexpr (D3) {
          const double dz = min_image(readlane_d(zl, jj) - zi, a.lmi[2], a.half[2]);
          r2 = r2 + dz * dz;
        }
      }
      bool hit = r2 < a.side2;
      if constexpr (SELF) hit = hit && (jb + jj > gi);
      const unsigned long long m = __ballot(hit);
      double qj = 0.0;
      if constexpr (QQ) qj = readlane_d(ql, jj);
      if (hit) {
        const uint32_t below = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32),
                                                         __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u));
        const uint32_t e = (tail + below) & (kHitQ - 1);
        qr2[e] = r2;
        qtt[e] = ti8 | __builtin_amdgcn_readlane(tl, jj);
        if constexpr (QQ) qqr[e] = qi * qj;
      }
      tail += static_cast<uint32_t>(__popcll(m));
      if (tail - head >= kWave) {
        __builtin_amdgcn_wave_barrier();
        bin_hits<QQ>(a, thr, hist, qr2, qtt, head, kWave, lane, qqr, gqs, qsc);
        __builtin_amdgcn_wave_barrier();
        head += kWave;
      }
    }
  }
}

// One chunk of up to 64 atoms of home cell cellL (cells numbered from
// cell_off in `start`) against the half stencil; hits go to the wavefront's
// ring (qr2, qtt, head, tail).
template <bool SHIFT, bool D3, bool QQ = false>
__device__ inline void rdf_home_chunk(const PairArgs& a, const double* thr, uint32_t* hist, double* qr2, int* qtt,
                                      uint32_t& head, uint32_t& tail, int lane, const Stencil& st, const int nc[3],
                                      const int periodic[3], int cell_off, int cellL, int ch, int cs, int ce,
                                      const int32_t* __restrict__ start, const double* __restrict__ sx,
                                      const double* __restrict__ sy, const double* __restrict__ sz,
                                      const int32_t* __restrict__ stype, const double* __restrict__ sq = nullptr,
                                      double* qqr = nullptr, unsigned long long* gqs = nullptr,
                                      double qsc = 0.0) {
  const double nan = __builtin_nan("");
  const int ncx = nc[0], ncy = nc[1], ncz = nc[2];
  const int cx = cellL % ncx, cy = (cellL / ncx) % ncy, cz = cellL / (ncx * ncy);
  const int gi = cs + ch * kWave + lane;
  const int ti = gi < ce ? stype[gi] : 0;
  const bool oki = static_cast<unsigned>(ti - 1) < static_cast<unsigned>(a.nsp);
  const double xi = oki ? sx[gi] : nan;
  const double yi = oki ? sy[gi] : nan;
  const double zi = (D3 && oki) ? sz[gi] : nan;
  const int ti8 = (oki ? ti : 0) << 8;
  double qi = 0.0;
  if constexpr (QQ) qi = oki ? sq[gi] : 0.0;
  for (int o = 0; o < st.n; ++o) {
    int nx = cx + st.off[o][0], ny = cy + st.off[o][1], nz = cz + st.off[o][2];
    // Image shift of a wrapped neighbour, the L*nint(d/L) the FP64 rule
    // subtracts: a neighbour below cell 0 is the top cell (d ~ +L, subtract
    // +L); one above the top cell is cell 0 (d ~ -L, subtract -L).
    const double shx = nx < 0 ? a.lmi[0] : (nx >= ncx ? -a.lmi[0] : 0.0);
    const double shy = ny < 0 ? a.lmi[1] : (ny >= ncy ? -a.lmi[1] : 0.0);
    const double shz = D3 ? (nz < 0 ? a.lmi[2] : (nz >= ncz ? -a.lmi[2] : 0.0)) : 0.0;
    if (periodic[0]) nx = (nx + ncx) % ncx; else if (nx < 0 || nx >= ncx) continue;
    if (periodic[1]) ny = (ny + ncy) % ncy; else if (ny < 0 || ny >= ncy) continue;
    if (D3) {
      if (periodic[2]) nz = (nz + ncz) % ncz; else if (nz < 0 || nz >= ncz) continue;
    }
    const int nb = cell_off + (nz * ncy + ny) * ncx + nx;
    // 64 neighbours per pass: one coalesced load per lane, then a broadcast
    // per pair through the scalar registers.
    if (o == 0) {
      rdf_neighbours<SHIFT, D3, true, false, QQ>(a, thr, hist, qr2, qtt, head, tail, lane, cs + ch * kWave + 1,
                                                 ce, gi, xi, yi, zi, ti8, 0.0, 0.0, 0.0, sx, sy, sz, stype, qi, sq,
                                                 qqr, gqs, qsc);
    } else {
      const int js = __builtin_amdgcn_readfirstlane(start[nb]);
      const int je = __builtin_amdgcn_readfirstlane(start[nb + 1]);
      if (shx != 0.0 || shy != 0.0 || shz != 0.0)
        rdf_neighbours<SHIFT, D3, false, true, QQ>(a, thr, hist, qr2, qtt, head, tail, lane, js, je, gi, xi, yi,
                                                   zi, ti8, shx, shy, shz, sx, sy, sz, stype, qi, sq, qqr, gqs,
                                                   qsc);
      else
        rdf_neighbours<SHIFT, D3, false, false, QQ>(a, thr, hist, qr2, qtt, head, tail, lane, js, je, gi, xi, yi,
                                                    zi, ti8, 0.0, 0.0, 0.0, sx, sy, sz, stype, qi, sq, qqr, gqs,
                                                    qsc);
    }
  }
}

// ------------------------------------------------------------------------
// Prefiltered pair loop (every axis takes its neighbour cell's image shift).
// A wavefront takes 16 home atoms, the same 16 in each of its four 16-lane
// rows, and a batch of 64 neighbours, one per lane (neighbour 16 r + k of
// the batch in lane k of row r).  Coordinates are FP32 relative to the home
// cell's corner: x - corner for the home atoms, x - (corner + s) for the
// neighbours (s the neighbour cell's image shift), each difference in FP64
// and rounded once.  Step k broadcasts lane k of every row to its row (DPP
// row_newbcast, a modifier of the subtraction), so each step tests 64
// (home, neighbour) pairs in FP32 with no register or memory traffic for the
// broadcast.  |FP32 r2 - FP64 r2| < pf_delta (a bound on the rounding, from
// the host), so pairs with FP32 r2 below side2_pf >= side2 + pf_delta cover
// every pair in range.  They go to the wavefront's ring as (home index,
// neighbour index, FP32 r2, the two species); the ring is binned 64 entries
// at a time, every lane busy.  An entry whose FP32 r2 +- pf_delta lies
// inside one bin and below side2 has that bin under the exact FP64 rule
// (the bin's edges computed in FP32 registers and moved inwards by a margin
// the host proves for every bin: at least pf_delta beyond the exact edges);
// the others (near an edge) go through the FP64 rule itself:
// d = x_j - x_i, minimum image (on these axes the neighbour cell's image
// shift), r2 < side2, the bin edges.
constexpr int kRow = 16;

template <int K>
__device__ inline float row_bcast(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x150 + K, 0xF, 0xF,
                                                               false));
}

// row_bcast for an integer; bound_ctrl (no old value: the source lane of a
// row broadcast always exists).
template <int K>
__device__ inline int row_bcast_i(int v) {
  return __builtin_amdgcn_update_dpp(0, v, 0x150 + K, 0xF, 0xF, true);
}

struct PfTables {
  float cut;  // cut <= side2 - pf_delta
};

constexpr int kSlowQ = 64;  // near-edge entries per wavefront waiting for the FP64 rule

struct PfWave {
  int4* ring;
  uint32_t tail, fill;  // wave-uniform: entries written so far, entries not yet binned
  int2* slow;           // near-edge entries (home, neighbour), kSlowQ
  uint32_t nslow;       // wave-uniform: entries in slow
};

// Adds one binned entry (bin k of the FP64 rule) to the histogram tile and
// (QQ) its charge product to the block's charge histogram.
template <bool QQ>
__device__ inline void add_hit_pf(const PairArgs& a, uint32_t* hist, const int4& en, int k,
                                  const double* __restrict__ sq, unsigned long long* gqs, double qsc) {
  if (k >= 1 && k <= a.lsmax) {
    const int ti = en.w & 255, tj = en.w >> 8;  // the pair's species, stored at the push
    const int tlo = ti < tj ? ti : tj, thi = ti < tj ? tj : ti;
    const int row = pair_index(tlo, thi, a.nsp) - 1 - a.row0;
    if (static_cast<unsigned>(row) < static_cast<unsigned>(a.nrows)) atomicAdd(&hist[row * a.lsmax + (k - 1)], 1u);
    if constexpr (QQ)
      if (a.row0 == 0) atomicAdd(&gqs[k - 1], qq_fixed(sq[en.x] * sq[en.y], qsc));
  }
}

// The FP64 rule for the cnt entries of the slow queue, one per lane: d =
// x_j - x_i, minimum image, r2 < side2, the exact edges.
template <bool D3, bool QQ>
__device__ inline void bin_slow_pf(const PairArgs& a, const double* thr, uint32_t* hist, const int2* slow, int cnt,
                                   int lane, const double* __restrict__ sx, const double* __restrict__ sy,
                                   const double* __restrict__ sz, const int32_t* __restrict__ stype,
                                   const double* __restrict__ sq, unsigned long long* gqs, double qsc) {
  if (lane < cnt) {
    const int2 e = slow[lane];
    const int gi = e.x, gj = e.y;
    const int4 en = make_int4(gi, gj, 0, stype[gi] | (stype[gj] << 8));
    const double dx = min_image(sx[gj] - sx[gi], a.lmi[0], a.half[0]);
    const double dy = min_image(sy[gj] - sy[gi], a.lmi[1], a.half[1]);
    double r2 = dx * dx + dy * dy;
    if constexpr (D3) {
      const double dz = min_image(sz[gj] - sz[gi], a.lmi[2], a.half[2]);
      r2 = r2 + dz * dz;
    }
    add_hit_pf<QQ>(a, hist, en, r2 < a.side2 ? bin_of(r2, thr, a.inv_dr, a.lsmax) : 0, sq, gqs, qsc);
  }
}

// Bins cnt ring entries from head, one per lane.  An entry whose FP32 r2
// decides its bin is added at once; a near-edge one joins the slow queue,
// which takes the FP64 rule 64 entries at a time (its loads and FP64 work
// then run for a full wavefront of such entries, not for one lane of a
// wavefront of sure ones).
template <bool D3, bool QQ>
__device__ inline void bin_hits_pf(const PairArgs& a, const double* thr, const PfTables& tb, uint32_t* hist,
                                   PfWave& w, uint32_t head, int cnt, int lane, const double* __restrict__ sx,
                                   const double* __restrict__ sy, const double* __restrict__ sz,
                                   const int32_t* __restrict__ stype, const double* __restrict__ sq,
                                   unsigned long long* gqs, double qsc) {
  const bool act = lane < cnt;
  int4 en = make_int4(0, 0, 0, 0);
  int k = 0;
  bool sure = true;
  if (act) {
    en = w.ring[(head + lane) & (kHitQ - 1)];
    const float r2f = __int_as_float(en.z);
    // an estimate (the hardware square root): a wrong bin fails the test below
    k = static_cast<int>(__builtin_amdgcn_sqrtf(r2f) * a.inv_dr + a.pf_est);
    k = k > a.lsmax + 1 ? a.lsmax + 1 : k;
    // bin k's edges moved inwards, in FP32 registers: the host checked for
    // every k that they lie beyond the exact edges +- pf_delta
    const float x0 = static_cast<float>(k) + a.pf_off, x1 = x0 + 1.0f;
    const float elo = x0 * x0 * a.pf_dr2 + a.pf_margin;
    const float ehi = x1 * x1 * a.pf_dr2 - a.pf_margin;
    sure = r2f < tb.cut && r2f >= elo && r2f < ehi;
  }
  const unsigned long long ms = __ballot(act && !sure);
  if (ms) {
    const uint32_t ns = static_cast<uint32_t>(__popcll(ms));
    if (w.nslow + ns > static_cast<uint32_t>(kSlowQ)) {
      __builtin_amdgcn_wave_barrier();
      bin_slow_pf<D3, QQ>(a, thr, hist, w.slow, static_cast<int>(w.nslow), lane, sx, sy, sz, stype, sq, gqs, qsc);
      __builtin_amdgcn_wave_barrier();
      w.nslow = 0;
    }
    if (act && !sure) {
      const uint32_t below = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(ms >> 32),
                                                       __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(ms), 0u));
      w.slow[w.nslow + below] = make_int2(en.x, en.y);
    }
    w.nslow += ns;
  }
  if (act && sure) add_hit_pf<QQ>(a, hist, en, k, sq, gqs, qsc);
}

// Step K of a batch: home lanes against neighbour 16 r + K of the batch.
// With CHECK the step is skipped when no row holds a neighbour at K (only a
// batch of fewer than 16 neighbours needs the test).
// PACKED: the batch's neighbours come from several runs (rdf_packed_pf) and
// jrow is this lane's own neighbour index, broadcast per step; otherwise the
// neighbours are contiguous and jrow is the first of this lane's row of 16.
template <int K, bool D3, bool SELF, bool QQ, bool CHECK, bool PACKED = false>
__device__ inline void pf_step(const PairArgs& a, const double* thr, const PfTables& tb, uint32_t* hist, PfWave& w,
                               int lane, int jb, int je, int gi, int ti, int jrow, float hx, float hy, float hz,
                               const float4& q, const double* __restrict__ sx, const double* __restrict__ sy,
                               const double* __restrict__ sz, const int32_t* __restrict__ stype,
                               const double* __restrict__ sq, unsigned long long* gqs, double qsc) {
  if constexpr (CHECK)
    if (jb + K >= je) return;  // no row holds a neighbour at K (row 0 fills first)
  const float dx = row_bcast<K>(q.x) - hx, dy = row_bcast<K>(q.y) - hy;
  float r2 = fmaf(dy, dy, dx * dx);
  if constexpr (D3) {
    const float dz = row_bcast<K>(q.z) - hz;
    r2 = fmaf(dz, dz, r2);
  }
  const int gj = PACKED ? row_bcast_i<K>(jrow) : jrow + K;
  // the neighbour's species (broadcast with every lane active: a DPP read of
  // a lane outside EXEC would not return its value)
  const int tj = row_bcast_i<K>(__builtin_bit_cast(int, q.w));
  bool hit = r2 < a.side2_pf;
  if constexpr (SELF) hit = hit && (gj > gi);
  const unsigned long long m = __ballot(hit);
  if (hit) {
    const uint32_t below = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32),
                                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u));
    w.ring[(w.tail + below) & (kHitQ - 1)] = make_int4(gi, gj, __float_as_int(r2), ti | (tj << 8));
  }
  const uint32_t c = static_cast<uint32_t>(__popcll(m));
  w.tail += c;
  w.fill += c;
  if (w.fill >= static_cast<uint32_t>(kWave)) {
    __builtin_amdgcn_wave_barrier();
    bin_hits_pf<D3, QQ>(a, thr, tb, hist, w, w.tail - w.fill, kWave, lane, sx, sy, sz, stype, sq, gqs, qsc);
    __builtin_amdgcn_wave_barrier();
    w.fill -= kWave;
  }
}

template <bool D3, bool SELF, bool QQ, bool CHECK, bool PACKED, int... K>
__device__ inline void pf_batch(const PairArgs& a, const double* thr, const PfTables& tb, uint32_t* hist, PfWave& w,
                                int lane, int jb, int je, int gi, int ti, int jrow, float hx, float hy, float hz,
                                const float4& q, const double* __restrict__ sx, const double* __restrict__ sy,
                                const double* __restrict__ sz, const int32_t* __restrict__ stype,
                                const double* __restrict__ sq, unsigned long long* gqs, double qsc,
                                std::integer_sequence<int, K...>) {
  (pf_step<K, D3, SELF, QQ, CHECK, PACKED>(a, thr, tb, hist, w, lane, jb, je, gi, ti, jrow, hx, hy, hz, q, sx, sy,
                                           sz, stype, sq, gqs, qsc),
   ...);
}

// Neighbours [js, je) of one cell; (ox, oy, oz) = corner + image shift.
template <bool D3, bool SELF, bool QQ>
__device__ inline void rdf_neighbours_pf(const PairArgs& a, const double* thr, const PfTables& tb, uint32_t* hist,
                                         PfWave& w, int lane, int js, int je, int gi, int ti, float hx, float hy,
                                         float hz,
                                         double ox, double oy, double oz, const double* __restrict__ sx,
                                         const double* __restrict__ sy, const double* __restrict__ sz,
                                         const int32_t* __restrict__ stype, const double* __restrict__ sq,
                                         unsigned long long* gqs, double qsc) {
  const float fnan = __builtin_nanf("");
  const int r16 = (lane >> 4) * kRow;
  for (int jb = js; jb < je; jb += kWave) {
    // lane r*16 + k holds neighbour jb + r*16 + k (NaN past the cell or for a
    // type outside 1..nsp: no FP32 distance to it passes)
    const int jl = jb + r16 + (lane & (kRow - 1));
    const int tl = jl < je ? stype[jl] : 0;
    const bool vl = static_cast<unsigned>(tl - 1) < static_cast<unsigned>(a.nsp);
    float4 q;
    q.x = vl ? static_cast<float>(sx[jl] - ox) : fnan;
    q.y = vl ? static_cast<float>(sy[jl] - oy) : fnan;
    q.z = (D3 && vl) ? static_cast<float>(sz[jl] - oz) : 0.f;
    q.w = __builtin_bit_cast(float, tl);
    if (je - jb >= kRow)  // every step has a neighbour in row 0
      pf_batch<D3, SELF, QQ, false, false>(a, thr, tb, hist, w, lane, jb, je, gi, ti, jb + r16, hx, hy, hz, q, sx, sy,
                                           sz, stype, sq, gqs, qsc, std::make_integer_sequence<int, kRow>{});
    else
      pf_batch<D3, SELF, QQ, true, false>(a, thr, tb, hist, w, lane, jb, je, gi, ti, jb + r16, hx, hy, hz, q, sx, sy,
                                          sz, stype, sq, gqs, qsc, std::make_integer_sequence<int, kRow>{});
  }
}

template <bool D3, bool QQ>
__device__ inline void rdf_home_chunk_pf(const PairArgs& a, const double* thr, const PfTables& tb, uint32_t* hist,
                                         PfWave& w, int lane, const Stencil& st, const CellGrid& g, int cellL, int ch,
                                         int cs, int ce, const int32_t* __restrict__ start,
                                         const double* __restrict__ sx, const double* __restrict__ sy,
                                         const double* __restrict__ sz, const int32_t* __restrict__ stype,
                                         const double* __restrict__ sq, unsigned long long* gqs, double qsc) {
  const double nan = __builtin_nan("");
  const int ncx = g.nc[0], ncy = g.nc[1], ncz = g.nc[2];
  const int cx = cellL % ncx, cy = (cellL / ncx) % ncy, cz = cellL / (ncx * ncy);
  // the home cell's corner: the reference point of the FP32 coordinates
  const double ox = static_cast<double>(cx) / g.inv_h[0];
  const double oy = static_cast<double>(cy) / g.inv_h[1];
  const double oz = D3 ? static_cast<double>(cz) / g.inv_h[2] : 0.0;
  const int h0 = cs + ch * kRow;
  const int gi = h0 + (lane & (kRow - 1));  // the same 16 home atoms in every row
  const int ti = gi < ce ? stype[gi] : 0;
  const bool oki = static_cast<unsigned>(ti - 1) < static_cast<unsigned>(a.nsp);
  const float hx = static_cast<float>(oki ? sx[gi] - ox : nan);
  const float hy = static_cast<float>(oki ? sy[gi] - oy : nan);
  const float hz = D3 ? static_cast<float>(oki ? sz[gi] - oz : nan) : 0.f;
  // The forward half stencil as runs of x-adjacent cells, which are adjacent
  // in the cell order and so share one range of the cell-sorted arrays and
  // one image shift: the home cell with its +x neighbour (from the chunk's
  // first atom; j > i also holds for the next cell, whose atoms all come
  // later), then the rows (dy, dz) = (+1, 0) and, in 3D, (-1..+1, +1), each
  // over x = cx-1 .. cx+1.  A cell across a periodic x boundary is a run of
  // its own with the x shift; across a non-periodic one it is skipped.
  // Shift s = +1 below cell 0 (d ~ +L: subtract +L), -1 above the top cell;
  // neighbours are taken relative to corner + s: x_j - (corner + s).
  {
    // the home run: own cell (j > i) and, without a wrap, the +x cell
    const int je = __builtin_amdgcn_readfirstlane(cx + 1 < ncx ? start[cellL + 2] : ce);
    rdf_neighbours_pf<D3, true, QQ>(a, thr, tb, hist, w, lane, h0, je, gi, ti, hx, hy, hz, ox, oy, oz, sx, sy, sz,
                                    stype, sq, gqs, qsc);
  }
  // The other runs, one call site: q = 0 the wrapped +x cell of the home row;
  // q = 1 + 3 r + k for row r (r = 0: (dy, dz) = (+1, 0); r = 1..3: (r - 2,
  // +1)) and k = 0 the cell across the -x boundary, 1 the cells inside, 2 the
  // cell across the +x boundary.
  const int nq = D3 ? 13 : 4;
  for (int q = 0; q < nq; ++q) {
    int c0, c1, sxi, syi = 0, szi = 0;
    if (q == 0) {
      if (!(cx + 1 >= ncx && g.periodic[0])) continue;
      c0 = c1 = cellL - cx;
      sxi = -1;
    } else {
      const int r = (q - 1) / 3, k = (q - 1) % 3;
      int ny = cy + (r == 0 ? 1 : r - 2), nz = cz + (r == 0 ? 0 : 1);
      syi = ny < 0 ? 1 : (ny >= ncy ? -1 : 0);
      szi = D3 ? (nz < 0 ? 1 : (nz >= ncz ? -1 : 0)) : 0;
      if (g.periodic[1]) ny = (ny + ncy) % ncy; else if (ny < 0 || ny >= ncy) continue;
      if (D3) {
        if (g.periodic[2]) nz = (nz + ncz) % ncz; else if (nz < 0 || nz >= ncz) continue;
      }
      const int base = (nz * ncy + ny) * ncx;
      if (k == 0) {
        if (!(cx - 1 < 0 && g.periodic[0])) continue;
        c0 = c1 = base + ncx - 1;
        sxi = 1;
      } else if (k == 1) {
        c0 = base + (cx - 1 < 0 ? 0 : cx - 1);
        c1 = base + (cx + 1 >= ncx ? ncx - 1 : cx + 1);
        sxi = 0;
      } else {
        if (!(cx + 1 >= ncx && g.periodic[0])) continue;
        c0 = c1 = base;
        sxi = -1;
      }
    }
    const double px = ox + static_cast<double>(sxi) * a.lmi[0];
    const double py = oy + static_cast<double>(syi) * a.lmi[1];
    const double pz = D3 ? oz + static_cast<double>(szi) * a.lmi[2] : 0.0;
    const int js = __builtin_amdgcn_readfirstlane(start[c0]);
    const int je = __builtin_amdgcn_readfirstlane(start[c1 + 1]);
    rdf_neighbours_pf<D3, false, QQ>(a, thr, tb, hist, w, lane, js, je, gi, ti, hx, hy, hz, px, py, pz, sx, sy, sz,
                                     stype, sq, gqs, qsc);
  }
}

// A wavefront's table of neighbour runs for one home chunk: run k (k < 64,
// at most 37 here) in lane k of three registers; its first packed index
// (lane n: the total),
// its first atom, and its image shift as (sx + 1) + 3 (sy + 1) + 9 (sz + 1),
// each shift in -1..1 (subtract s L, as the unpacked calls' corner + s L).
struct RunTable {
  int cum = 0, js = 0, code = 0;
  int n = 0, total = 0;  // wave-uniform: runs, packed length
};

__device__ inline void run_add(RunTable& t, int lane, int js, int je, int sxi, int syi, int szi) {
  if (je <= js) return;
  const bool me = lane == t.n;
  t.cum = me ? t.total : t.cum;
  t.js = me ? js : t.js;
  t.code = me ? (sxi + 1) + 3 * (syi + 1) + 9 * (szi + 1) : t.code;
  t.total += je - js;
  ++t.n;
}

// The runs of t against one home chunk as one packed sequence: batches of
// 64 neighbours straddle runs (lane 16 r + k holds packed index P + 16 r + k,
// from the run whose first packed index is the largest at or below it), so
// that short runs do not leave batches mostly empty.  Same steps as
// rdf_neighbours_pf, with the neighbour index broadcast per step.
template <bool D3, bool QQ>
__device__ inline void rdf_packed_pf(const PairArgs& a, const double* thr, const PfTables& tb, uint32_t* hist,
                                     PfWave& w, int lane, RunTable& t, int gi, int ti, float hx, float hy, float hz,
                                     double ox, double oy, double oz, const double* __restrict__ sx,
                                     const double* __restrict__ sy, const double* __restrict__ sz,
                                     const int32_t* __restrict__ stype, const double* __restrict__ sq,
                                     unsigned long long* gqs, double qsc) {
  if (t.n == 0) return;
  t.cum = lane == t.n ? t.total : t.cum;  // sentinel
  const float fnan = __builtin_nanf("");
  int r0 = 0;  // wave-uniform: the run holding packed index P
  for (int P = 0; P < t.total; P += kWave) {
    const int v = P + lane;
    int r = r0;
    for (int k = r0 + 1; k < t.n; ++k) {
      const int ck = __builtin_amdgcn_readlane(t.cum, k);
      if (ck >= P + kWave) break;
      r += ck <= v ? 1 : 0;
    }
    const int cr = __builtin_amdgcn_ds_bpermute(r << 2, t.cum);
    const int jr = __builtin_amdgcn_ds_bpermute(r << 2, t.js);
    const int code = __builtin_amdgcn_ds_bpermute(r << 2, t.code);
    const int jl = jr + (v - cr);
    const int tl = v < t.total ? stype[jl] : 0;
    const bool vl = static_cast<unsigned>(tl - 1) < static_cast<unsigned>(a.nsp);
    const double px = ox + static_cast<double>(code % 3 - 1) * a.lmi[0];
    const double py = oy + static_cast<double>((code / 3) % 3 - 1) * a.lmi[1];
    const double pz = D3 ? oz + static_cast<double>(code / 9 - 1) * a.lmi[2] : 0.0;
    float4 q;
    q.x = vl ? static_cast<float>(sx[jl] - px) : fnan;
    q.y = vl ? static_cast<float>(sy[jl] - py) : fnan;
    q.z = (D3 && vl) ? static_cast<float>(sz[jl] - pz) : 0.f;
    q.w = __builtin_bit_cast(float, tl);
    if (t.total - P >= kRow)
      pf_batch<D3, false, QQ, false, true>(a, thr, tb, hist, w, lane, P, t.total, gi, ti, jl, hx, hy, hz, q, sx, sy,
                                           sz, stype, sq, gqs, qsc, std::make_integer_sequence<int, kRow>{});
    else
      pf_batch<D3, false, QQ, true, true>(a, thr, tb, hist, w, lane, P, t.total, gi, ti, jl, hx, hy, hz, q, sx, sy,
                                          sz, stype, sq, gqs, qsc, std::make_integer_sequence<int, kRow>{});
    while (r0 + 1 < t.n && __builtin_amdgcn_readlane(t.cum, r0 + 1) <= P + kWave) ++r0;
  }
}

// Cells of side >= half the cutoff (reach 2; prefilter path only, every
// periodic axis with >= 7 cells): one wavefront walks cells [c0, c1) of the
// x-row `row` (cy, cz) in home chunks of up to 16 consecutive atoms that span
// at most 3 cells, so that cells of ~14 atoms still fill the 16 home lanes.
// A chunk whose atoms lie in cells cxa..cxb meets, in each row of the forward
// half stencil ((dy, dz) = (+1, 0), (+2, 0), (-2..+2, +1), (-2..+2, +2)), the
// cells cxa - 2 .. cxb + 2: one range of the cell-sorted arrays, split where
// it wraps across a periodic x boundary (at most 7 cells <= ncx, so no cell
// twice).  Its own row runs from the chunk's first atom (j > i) to cell
// cxb + 2.  Every neighbour within the cutoff of a home atom in cell cx lies
// in cells cx - 2 .. cx + 2 of these rows (2 h >= rc); cells further out only
// add candidates that the cutoff test rejects.  Each pair i < j is counted
// once: from i's chunk, in j's row when that row is forward of i's, else
// (same row) by j > i.
template <bool D3, bool QQ>
__device__ inline void rdf_row_pf(const PairArgs& a, const double* thr, const PfTables& tb, uint32_t* hist,
                                  PfWave& w, int lane, const CellGrid& g, int row, int c0, int c1,
                                  const int32_t* __restrict__ start, const double* __restrict__ sx,
                                  const double* __restrict__ sy, const double* __restrict__ sz,
                                  const int32_t* __restrict__ stype, const double* __restrict__ sq,
                                  unsigned long long* gqs, double qsc) {
  const double nan = __builtin_nan("");
  const int ncx = g.nc[0], ncy = g.nc[1], ncz = g.nc[2];
  const int cy = row % ncy, cz = row / ncy;
  const int rb = row * ncx;  // the row's first cell
  const double oy = static_cast<double>(cy) / g.inv_h[1];
  const double oz = D3 ? static_cast<double>(cz) / g.inv_h[2] : 0.0;
  auto st = [&](int c) { return __builtin_amdgcn_readfirstlane(start[c]); };
  const int hend = st(rb + c1);
  int cx = c0;
  int h0 = st(rb + c0);
  while (h0 < hend) {
    while (st(rb + cx + 1) <= h0) ++cx;  // the cell of the chunk's first atom
    const int cmax = (cx + 2 < c1 - 1) ? cx + 2 : c1 - 1;
    const int lim = st(rb + cmax + 1);
    const int he = (h0 + kRow < lim) ? h0 + kRow : lim;
    int cxb = cx;  // the cell of its last atom
    while (st(rb + cxb + 1) <= he - 1) ++cxb;
    // the chunk's first cell's corner: the reference point of the FP32 coordinates
    const double ox = static_cast<double>(cx) / g.inv_h[0];
    const int gi = h0 + (lane & (kRow - 1));  // the same 16 home atoms in every row
    const int ti = gi < he ? stype[gi] : 0;
    const bool oki = static_cast<unsigned>(ti - 1) < static_cast<unsigned>(a.nsp);
    const float hx = static_cast<float>(oki ? sx[gi] - ox : nan);
    const float hy = static_cast<float>(oki ? sy[gi] - oy : nan);
    const float hz = D3 ? static_cast<float>(oki ? sz[gi] - oz : nan) : 0.f;
    // its own row: j > i up to cell cxb + 2, then the cells across a periodic +x boundary
    {
      const int xe = (cxb + 2 < ncx - 1) ? cxb + 2 : ncx - 1;
      rdf_neighbours_pf<D3, true, QQ>(a, thr, tb, hist, w, lane, h0, st(rb + xe + 1), gi, ti, hx, hy, hz, ox, oy,
                                      oz, sx, sy, sz, stype, sq, gqs, qsc);
    }
    // the other runs, packed into full batches
    RunTable rt;
    if (g.periodic[0] && cxb + 2 >= ncx) run_add(rt, lane, st(rb), st(rb + cxb + 2 - ncx + 1), -1, 0, 0);
    const int nrow = D3 ? 12 : 2;
    for (int q = 0; q < nrow; ++q) {
      const int dz = q < 2 ? 0 : 1 + (q - 2) / 5;
      const int dy = q < 2 ? q + 1 : (q - 2) % 5 - 2;
      int ny = cy + dy, nz = cz + dz;
      const int syi = ny < 0 ? 1 : (ny >= ncy ? -1 : 0);
      const int szi = D3 ? (nz < 0 ? 1 : (nz >= ncz ? -1 : 0)) : 0;
      if (g.periodic[1]) ny = (ny + ncy) % ncy; else if (ny < 0 || ny >= ncy) continue;
      if (D3) {
        if (g.periodic[2]) nz = (nz + ncz) % ncz; else if (nz < 0 || nz >= ncz) continue;
      }
      const int base = (nz * ncy + ny) * ncx;
      const int lo = cx - 2, hi = cxb + 2;
      // cells inside the row, then (periodic x) those across the -x and +x boundaries
      const int m0 = lo < 0 ? 0 : lo, m1 = hi > ncx - 1 ? ncx - 1 : hi;
      run_add(rt, lane, st(base + m0), st(base + m1 + 1), 0, syi, szi);
      if (g.periodic[0] && lo < 0) run_add(rt, lane, st(base + ncx + lo), st(base + ncx), 1, syi, szi);
      if (g.periodic[0] && hi >= ncx) run_add(rt, lane, st(base), st(base + hi - ncx + 1), -1, syi, szi);
    }
    rdf_packed_pf<D3, QQ>(a, thr, tb, hist, w, lane, rt, gi, ti, hx, hy, hz, ox, oy, oz, sx, sy, sz, stype, sq, gqs,
                          qsc);
    h0 = he;
  }
}

// ROWS: cells of half the cutoff walked in x-rows (rdf_row_pf; SHIFT only).
template <bool SHIFT, bool D3, bool QQ, bool ROWS = false>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(8))) void k_rdf_cells(PairArgs a, CellGrid g, Stencil st,
                                                   const int32_t* __restrict__ start,
                                                   const double* __restrict__ spos,
                                                   const int32_t* __restrict__ stype, int64_t n,
                                                   int W, int64_t ntasks,
                                                   uint32_t* __restrict__ partial) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int nbins = a.nrows * a.lsmax;  // this pass's tile
  uint32_t* hist = reinterpret_cast<uint32_t*>(smem);
  double* thr = reinterpret_cast<double*>(smem + ((static_cast<size_t>(nbins) * 4 + 15) & ~size_t(15)));
  // Same layout as cells_lds_bytes: hist, thr, the rings, then (QQ) the
  // charge-product rings and the block's fixed-point charge histogram.
  const size_t qoff = (((static_cast<size_t>(nbins) * 4 + 15) & ~size_t(15)) +
                       static_cast<size_t>(a.lsmax + 3) * 8 + 15) & ~size_t(15);
  const size_t qqoff = qoff + kCellWaves * kHitQ * (sizeof(double) + sizeof(int));
  double* qqr = nullptr;
  unsigned long long* gqs = nullptr;
  double qsc = 0.0;  // the frame's fixed-point scale (qp[0])
  if constexpr (QQ) {
    qsc = a.qp[0];
    gqs = reinterpret_cast<unsigned long long*>(smem + qqoff + kCellWaves * kHitQ * sizeof(double));
    for (int b = threadIdx.x; b < a.lsmax; b += blockDim.x) gqs[b] = 0ull;
  }
  hist_setup(a, hist, thr);
  // wave-uniform (so are the task loop and the ring counters it carries)
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave), lane = threadIdx.x & (kWave - 1);
  double* qr2 = reinterpret_cast<double*>(smem + qoff) + wave * kHitQ;
  int* qtt = reinterpret_cast<int*>(smem + qoff + kCellWaves * kHitQ * sizeof(double)) + wave * kHitQ;
  if constexpr (QQ) qqr = reinterpret_cast<double*>(smem + qqoff) + wave * kHitQ;
  uint32_t head = 0, tail = 0;  // wave-uniform ring counters
  const double* __restrict__ sx = spos;
  const double* __restrict__ sy = spos + n;
  const double* __restrict__ sz = spos + 2 * n;
  // Prefilter path: (home, neighbour, FP32 r2, species) entries, 16 B each,
  // in the space of the (r2, type) rings and the (QQ) charge rings; home
  // chunks of 16 atoms.
  // The slow queues after the rings and the (QQ) charge histogram, as
  // cells_lds_bytes lays them out.
  const size_t soff = qoff + kCellWaves * kHitQ * (QQ ? 20 : 16) +
                      (QQ ? (static_cast<size_t>(a.lsmax) * 8 + 15) & ~size_t(15) : 0);
  PfWave pw{reinterpret_cast<int4*>(smem + qoff) + wave * kHitQ, 0u, 0u,
            reinterpret_cast<int2*>(smem + soff) + wave * kSlowQ, 0u};
  PfTables tb{};
  if constexpr (SHIFT) {
    // the FP32 rounding of side2 - 2 delta stays below side2 - delta (delta
    // >= 2^-21 side2 exceeds that rounding)
    tb.cut = static_cast<float>(a.side2 - 2.0 * a.pf_delta);
  }
  const int chunk = SHIFT ? kRow : kWave;
  if constexpr (SHIFT && ROWS) {
    // tasks (row, part): part p of W walks cells [p ncx / W, (p + 1) ncx / W) of the row
    for (int64_t task = static_cast<int64_t>(blockIdx.x) * kCellWaves + wave; task < ntasks;
         task += static_cast<int64_t>(gridDim.x) * kCellWaves) {
      const int row = static_cast<int>(task / W), part = static_cast<int>(task % W);
      const int c0 = static_cast<int>(static_cast<int64_t>(part) * g.nc[0] / W);
      const int c1 = static_cast<int>(static_cast<int64_t>(part + 1) * g.nc[0] / W);
      if (c0 < c1)
        rdf_row_pf<D3, QQ>(a, thr, tb, hist, pw, lane, g, row, c0, c1, start, sx, sy, sz, stype, a.q, gqs, qsc);
    }
  }
  for (int64_t task = static_cast<int64_t>(blockIdx.x) * kCellWaves + wave; task < ntasks && !(SHIFT && ROWS);
       task += static_cast<int64_t>(gridDim.x) * kCellWaves) {
    const int cell = static_cast<int>(task / W);
    const int ch0 = static_cast<int>(task % W);
    const int cs = __builtin_amdgcn_readfirstlane(start[cell]);
    const int ce = __builtin_amdgcn_readfirstlane(start[cell + 1]);
    for (int ch = ch0; cs + ch * chunk < ce; ch += W) {
      if constexpr (SHIFT)
        rdf_home_chunk_pf<D3, QQ>(a, thr, tb, hist, pw, lane, st, g, cell, ch, cs, ce, start, sx, sy, sz, stype, a.q,
                                  gqs, qsc);
      else
        rdf_home_chunk<SHIFT, D3, QQ>(a, thr, hist, qr2, qtt, head, tail, lane, st, g.nc, g.periodic, 0, cell, ch,
                                      cs, ce, start, sx, sy, sz, stype, a.q, qqr, gqs, qsc);
    }
  }
  __builtin_amdgcn_wave_barrier();
  if constexpr (SHIFT) {
    bin_hits_pf<D3, QQ>(a, thr, tb, hist, pw, pw.tail - pw.fill, static_cast<int>(pw.fill), lane, sx, sy, sz, stype,
                        a.q, gqs, qsc);
    __builtin_amdgcn_wave_barrier();
    bin_slow_pf<D3, QQ>(a, thr, hist, pw.slow, static_cast<int>(pw.nslow), lane, sx, sy, sz, stype, a.q, gqs, qsc);
  } else
    bin_hits<QQ>(a, thr, hist, qr2, qtt, head, static_cast<int>(tail - head), lane, qqr, gqs, qsc);
  hist_flush(a, hist, partial);  // begins with a barrier: every bin_hits is done
  if constexpr (QQ) {
    for (int b = threadIdx.x; b < a.lsmax; b += blockDim.x) {
      const unsigned long long v = gqs[b];
      if (v) atomicAdd(&a.gq[b], v);
    }
  }
}

// All pairs i < j of atoms [0, n) of SoA arrays with plane stride `stride`;
// with range (device, may be NULL) the atoms [range[0], range[1]) instead.
__global__ __launch_bounds__(256) void k_rdf_allpairs(PairArgs a, const double* __restrict__ pos,
                                                      const int32_t* __restrict__ itype, int64_t n,
                                                      int64_t stride, const int32_t* __restrict__ range,
                                                      uint32_t* __restrict__ partial) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int nbins = a.nrows * a.lsmax;  // this pass's tile
  uint32_t* hist = reinterpret_cast<uint32_t*>(smem);
  double* thr = reinterpret_cast<double*>(smem + ((static_cast<size_t>(nbins) * 4 + 15) & ~size_t(15)));
  // Same layout as pair_lds_bytes: hist, thr, then (a.q) the block's
  // fixed-point charge histogram.
  unsigned long long* gqs = reinterpret_cast<unsigned long long*>(
      smem + ((((static_cast<size_t>(nbins) * 4 + 15) & ~size_t(15)) + static_cast<size_t>(a.lsmax + 3) * 8 + 15) &
              ~size_t(15)));
  if (a.q)
    for (int b = threadIdx.x; b < a.lsmax; b += blockDim.x) gqs[b] = 0ull;
  const double qsc = a.q ? a.qp[0] : 0.0;  // the frame's fixed-point scale
  hist_setup(a, hist, thr);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave), lane = threadIdx.x & (kWave - 1);
  const int64_t base = range ? range[0] : 0;
  if (range) n = range[1] - range[0];
  const int nb = static_cast<int>((n + kWave - 1) / kWave);
  const double* __restrict__ px = pos + base;
  const double* __restrict__ py = pos + stride + base;
  const double* __restrict__ pz = pos + 2 * stride + base;
  itype += base;
  const double* __restrict__ qv = a.q ? a.q + base : nullptr;
  const int64_t ntasks = static_cast<int64_t>(nb) * nb;
  for (int64_t task = static_cast<int64_t>(blockIdx.x) * 4 + wave; task < ntasks;
       task += static_cast<int64_t>(gridDim.x) * 4) {
    const int ib = static_cast<int>(task / nb), jb = static_cast<int>(task % nb);
    if (jb < ib) continue;
    const int64_t gi = static_cast<int64_t>(ib) * kWave + lane;
    const bool vi = gi < n;
    const double xi = vi ? px[gi] : 0.0;
    const double yi = vi ? py[gi] : 0.0;
    const double zi = vi ? pz[gi] : 0.0;
    const int ti = vi ? itype[gi] : 0;
    const bool oki = vi && static_cast<unsigned>(ti - 1) < static_cast<unsigned>(a.nsp);
    const double qi = (qv && vi) ? qv[gi] : 0.0;
    const int64_t js = static_cast<int64_t>(jb) * kWave;
    const int64_t je = js + kWave < n ? js + kWave : n;
    for (int64_t j = js; j < je; ++j) {
      const int tj = itype[j];
      const bool ok = oki && (jb > ib || j > gi) &&
                      static_cast<unsigned>(tj - 1) < static_cast<unsigned>(a.nsp);
      const int k = pair_accumulate(a, thr, hist, ok, xi, yi, zi, ti, px[j], py[j], pz[j], tj);
      if (qv && k && a.row0 == 0) atomicAdd(&gqs[k - 1], qq_fixed(qi * qv[j], qsc));
    }
  }
  hist_flush(a, hist, partial);  // begins with a barrier
  if (a.q) {
    for (int b = threadIdx.x; b < a.lsmax; b += blockDim.x) {
      const unsigned long long v = gqs[b];
      if (v) atomicAdd(&a.gq[b], v);
    }
  }
}

// The frame's charge sums into the FP64 accumulators (rdf.cuf:270-274:
// gqq += gqq_frame / avq2), then cleared for the next frame.  qp = {scale,
// avq2}; the scale is a power of two, so f / scale = f * (1 / scale) exactly.
__global__ __launch_bounds__(256) void k_qq_finish(unsigned long long* __restrict__ gq, int lsmax,
                                                   const double* __restrict__ qp, double* __restrict__ acc) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= lsmax) return;
  const double inv_scale = 1.0 / qp[0], avq2 = qp[1];
  const long long f = static_cast<long long>(gq[k]);
  if (avq2 > 0.0) acc[k] += (static_cast<double>(f) * inv_scale) / avq2;
  gq[k] = 0ull;
}

// The frame's charge parameters from device-resident charges, in one
// workgroup with a fixed reduction order: qp[0] = 2^(30-e) with max q^2 < 2^e
// (2^30 when every charge is 0), qp[1] = sum(q^2)/n.  Thread t sums a
// contiguous run of q^2 in index order, then the 1024 partial sums meet in a
// pairwise tree.  A non-finite charge raises kFlagRange.
__global__ __launch_bounds__(1024) void k_qq_params(const double* __restrict__ q, int64_t n, double* __restrict__ qp,
                                                    uint32_t* __restrict__ flags) {
  __shared__ double ssum[1024];
  __shared__ double smax[1024];
  const int t = threadIdx.x;
  const int64_t per = (n + 1023) / 1024;
  const int64_t b = t * per, e = b + per < n ? b + per : n;
  double s = 0.0, m = 0.0;
  bool bad = false;
  for (int64_t i = b; i < e; ++i) {
    const double v = q[i] * q[i];
    bad |= !isfinite(v);
    s += v;
    m = v > m ? v : m;
  }
  ssum[t] = s;
  smax[t] = m;
  if (bad) atomicOr(flags, kFlagRange);
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if (t < w) {
      ssum[t] = ssum[t] + ssum[t + w];
      smax[t] = smax[t + w] > smax[t] ? smax[t + w] : smax[t];
    }
    __syncthreads();
  }
  if (t == 0) {
    int ex = 0;
    if (smax[0] > 0.0) frexp(smax[0], &ex);
    qp[0] = ldexp(1.0, 30 - ex);
    qp[1] = ssum[0] / static_cast<double>(n);
  }
}

// Slice group blockIdx.y sums its share of the per-workgroup slices, then adds
// into the int64 totals (integer atomics: the result is order independent).
constexpr int kMergeGroups = 32;

// With species charges (qw not NULL): every pair of species pair ij adds
// the same rint(q_i q_j scale) = qw->w[ij] to the fixed-point g_qq sums, so
// the frame's sums are sum_ij w_ij count_ij[k] (gq[k], int64 two's
// complement), exactly what the pair kernel's charge path adds pair by pair.
__global__ __launch_bounds__(256) void k_rdf_merge(const uint32_t* __restrict__ partial, int nslices,
                                                   int nbins, int64_t* __restrict__ counts, QqWeights qw,
                                                   int lsmax, unsigned long long* __restrict__ gq) {
  const int b = blockIdx.x * 256 + threadIdx.x;
  if (b >= nbins) return;
  const int g = blockIdx.y;
  const int s0 = static_cast<int>(static_cast<int64_t>(nslices) * g / kMergeGroups);
  const int s1 = static_cast<int>(static_cast<int64_t>(nslices) * (g + 1) / kMergeGroups);
  uint64_t s = 0;
  for (int sl = s0; sl < s1; ++sl) s += partial[static_cast<int64_t>(sl) * nbins + b];
  if (s) {
    atomicAdd(reinterpret_cast<unsigned long long*>(&counts[b]), static_cast<unsigned long long>(s));
    if (gq)
      atomicAdd(&gq[b % lsmax],
                static_cast<unsigned long long>(qw.w[b / lsmax] * static_cast<long long>(s)));
  }
}

// ------------------------------------------------------------------ S(Q)

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int KC = 32;            // atoms staged per chunk
constexpr int kFoldChunks = 8;    // fold MFMA accumulators every 8 chunks (256 atoms)
// FP64 tile: 8 kx x 8 ky x 32 kz (M = 64 (kx,ky) pairs, N = 32 kz).
constexpr int TKX = 8, TKY = 8, TKZ = 32, TILE_VEC = TKX * TKY * TKZ;
// LDS record of one atom's phases: x at 0..7, y at 8..15, z at 16..47.  The
// stride of 49 slots (16 B each) puts consecutive atoms one slot apart in the
// bank cycle, so staging writes (lane = atom) and operand reads are
// conflict-free.
constexpr int kRec = 49;

// m e^{2 pi i n u} for n = n0 .. n0+7 into dst from the anchor e^{2 pi i n0 u}
// = c + i s (times the channel weight m), then angle-addition steps by
// e^{2 pi i u} = cs + i ss.
__device__ inline void phase8_from(double c, double s, double cs, double ss, double2* dst, double m) {
  double re = c * m, im = s * m;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    dst[k] = make_double2(re, im);
    const double nr = fma(re, cs, -(im * ss));
    const double ni = fma(re, ss, im * cs);
    re = nr;
    im = ni;
  }
}
// The same with the anchor from the turns u = uh + ul: exact reduction of
// n0 u, then sincospi.
__device__ inline void phase8(double uh, double ul, int n0, double cs, double ss, double2* dst, double m) {
  double s, c;
  sincospi(2.0 * frac_turns(n0, uh, ul), &s, &c);
  phase8_from(c, s, cs, ss, dst, m);
}

// Anchor table of the FP64 contraction: row row0 + m of an axis holds
// e^{2 pi i n0 u_j} with n0 = base + 8 m (m < rows) for every entry j,
// computed as phase8 would (bit-identical values).  One thread per entry
// and axis (blockIdx.y): the turns are read once, each row written
// coalesced.
struct AnchorAxes {
  int rows[3], base[3], row0[3];
};
__global__ __launch_bounds__(256) void k_sq_anchors(const double* __restrict__ u, int64_t n, AnchorAxes ax,
                                                    double2* __restrict__ anc, const int32_t* live_off,
                                                    const int32_t* live_nch) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (j >= (live_off ? static_cast<int64_t>(live_off[*live_nch]) : n)) return;
  const int a = blockIdx.y;
  const double uh = u[a * n + j], ul = u[(3 + a) * n + j];
  double2* dst = anc + static_cast<int64_t>(ax.row0[a]) * n + j;
  for (int m = 0; m < ax.rows[a]; ++m) {
    double s, c;
    sincospi(2.0 * frac_turns(ax.base[a] + 8 * m, uh, ul), &s, &c);
    dst[m * n] = make_double2(c, s);
  }
}

// s + c += x with Knuth's TwoSum: s keeps the rounded sum, c the error, kept
// in FP32 (it only has to hold the few low-order bits lost per fold).
__device__ inline void two_sum_acc(double& s, float& c, double x) {
  const double t = s + x;
  const double bp = t - s;
  const double e = (s - (t - bp)) + (x - bp);
  s = t;
  c += static_cast<float>(e);
}

// One 32-atom chunk of MFMAs for one wave: rows w*16.. of the tile, the
// 16-wide kz column blocks selected by CB (bit 0: kz 0..15, bit 1: 16..31).
// Complex products in three real ones: t1 = Ar Br, t2 = Ai Bi,
// t3 = (Ar + Ai)(Br + Bi); Re = t1 - t2, Im = t3 - t1 - t2.
template <int CB>
__device__ inline void contract_chunk_f64(const double2 (*ph)[kRec], int kq, int mx, int my, int i16,
                                          d4 (&t1)[2], d4 (&t2)[2], d4 (&t3)[2]) {
#pragma unroll
  for (int kk = 0; kk < KC / 4; ++kk) {
    const double2* rec = ph[kk * 4 + kq];
    const double2 ex = rec[mx];
    const double2 ey = rec[8 + my];
    const double ar = fma(ex.x, ey.x, -(ex.y * ey.y));
    const double ai = fma(ex.x, ey.y, ex.y * ey.x);
    const double as = ar + ai;
#pragma unroll
    for (int nb = 0; nb < 2; ++nb) {
      if (!((CB >> nb) & 1)) continue;
      const double2 bz = rec[16 + 16 * nb + i16];
      t1[nb] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, bz.x, t1[nb], 0, 0, 0);
      t2[nb] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, bz.y, t2[nb], 0, 0, 0);
      t3[nb] = __builtin_amdgcn_mfma_f64_16x16x4f64(as, bz.x + bz.y, t3[nb], 0, 0, 0);
    }
  }
}

// Channels, split and first rho plane of work item `item` (blockIdx.x /
// ntiles): every channel's split `item` of ksplit for one frame, or the
// batched frames' item table.
struct ItemRange {
  int c0, c1, split, splits;
  size_t plane0;
  bool skip;
};
__device__ inline ItemRange item_range(const ContractArgs& a, int item) {
  ItemRange r;
  if (!a.items) {
    r.c0 = 0;
    r.c1 = a.nch;
    r.split = item;
    r.splits = a.ksplit;
    r.plane0 = static_cast<size_t>(item) * a.nch;
    r.skip = false;
    return r;
  }
  r.skip = item >= *a.nitems;
  const int4 it = a.items[r.skip ? 0 : item];
  r.c0 = it.x;
  r.c1 = it.y;
  r.split = it.z;
  r.splits = it.w;
  r.plane0 = static_cast<size_t>(a.item_plane[r.skip ? 0 : item]);
  return r;
}

// Work index (item * ntiles + tile) of this workgroup in the contraction
// kernels.  Grouped (one frame, whose splits all cost the same; a batch's
// grid is sized for its largest group and its items differ, so that
// contiguous ranges per XCD would load the XCDs unevenly): the dispatcher
// deals workgroups to the 8 XCDs in turn
// (workgroup b to XCD b % 8); workgroup b = 8 r + x takes index x * per + r,
// per = ceil(nwork / 8), so each XCD walks a contiguous range of (split,
// tile) pairs in order: the tiles of one split run on one XCD at about the
// same time and read the split's turns through that XCD's L2 instead of each
// XCD reading them from HBM (measured: 8x fewer HBM reads for the 16-bit
// kernel on a 1M-atom frame, 6x for the FP64 one on 262k atoms).  Otherwise
// index b.  The grid is 8 * per workgroups; indices >= nwork return at once.
// Both mappings are bijections whatever the dispatch order.
__device__ inline int64_t contract_work_index(const ContractArgs& a, bool grouped) {
  if (!grouped) return blockIdx.x;
  const int64_t nwork = static_cast<int64_t>(a.ntiles) * a.ksplit;
  const int64_t per = (nwork + 7) / 8;
  return static_cast<int64_t>(blockIdx.x % 8) * per + blockIdx.x / 8;
}

// rho_s(q) for one tile and one atom split: a complex GEMM over atoms whose
// operands are generated per 32-atom chunk in LDS (double-buffered: the next
// chunk is staged while this one is contracted).  Waves 0-2 stage one
// 8-value segment per lane (lane = atom; wave 0: x | y, waves 1-2: z); the
// next chunk's turns are prefetched into registers a chunk ahead.
template <bool TABLE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_sq_contract_f64(
    ContractArgs a) {
  __shared__ double2 ph[2][KC][kRec];
  const int64_t wk = contract_work_index(a, a.xcd_order != 0);
  if (wk >= static_cast<int64_t>(a.ntiles) * a.ksplit) return;
  const int tile = static_cast<int>(wk % a.ntiles);
  const ItemRange wi = item_range(a, static_cast<int>(wk / a.ntiles));
  if (wi.skip) return;
  const int kx0 = a.tiles[4 * tile], ky0 = a.tiles[4 * tile + 1], kz0 = a.tiles[4 * tile + 2];
  const int t = threadIdx.x, w = t / kWave, lane = t & (kWave - 1);
  const int i16 = lane & 15, kq = lane >> 4;
  const int m = w * 16 + i16, mx = m / TKY, my = m % TKY;
  // Column blocks of this wave's 16 rows that hold summed vectors (uniform).
  const int cb = (a.tiles[4 * tile + 3] >> (2 * w)) & 3;
  // Staging role: atom st_at, segment st_seg (0: x, 1: y, 2-5: z 8-blocks).
  const bool stager = t < KC * 6;
  const int st_at = t % KC, st_seg = t / KC;
  const int st_ax = st_seg < 2 ? st_seg : 2;
  const int st_n0 = st_seg == 0 ? kx0 : (st_seg == 1 ? ky0 : kz0 + 8 * (st_seg - 2));
  const int st_slot = st_seg < 2 ? 8 * st_seg : 16 + 8 * (st_seg - 2);
  // Turn planes as uniform bases plus one 32-bit lane offset (3 n < 2^32 is
  // checked by the launcher).
  const double* u_hi = a.u;
  const double* u_lo = a.u + 3 * a.n;
  const double* u_c = a.u + 6 * a.n;
  const double* u_s = a.u + 9 * a.n;
  const uint32_t st_base = static_cast<uint32_t>(st_ax) * static_cast<uint32_t>(a.n);
  // With the anchor table, the segment's anchors (cos, sin) replace the
  // turns hi, lo: one 16-byte load instead of the FP64 reduction and
  // sincospi per segment and chunk.
  const double2* anc = (TABLE && stager)
                           ? a.anc + static_cast<int64_t>(a.anc_row[st_ax] + (st_n0 - a.anc_base[st_ax]) / 8) * a.n
                           : a.anc;
  // The x segment carries the channel weight (1, or the charge of the entry).
  const bool st_w = a.w != nullptr && st_seg == 0;
  for (int s = wi.c0; s < wi.c1; ++s) {
    const int a0 = a.sp_off[s], a1 = a.sp_off[s + 1];
    const int64_t len = a1 - a0;
    const int b0 = a0 + static_cast<int>(len * wi.split / wi.splits);
    const int b1 = a0 + static_cast<int>(len * (wi.split + 1) / wi.splits);
    d4 t1[2], t2[2], t3[2];
    double sumR[2][4], sumI[2][4];
    float cmpR[2][4], cmpI[2][4];
    for (int nb = 0; nb < 2; ++nb) {
      t1[nb] = d4{0.0, 0.0, 0.0, 0.0};
      t2[nb] = t1[nb];
      t3[nb] = t1[nb];
      for (int r = 0; r < 4; ++r) {
        sumR[nb][r] = sumI[nb][r] = 0.0;
        cmpR[nb][r] = cmpI[nb][r] = 0.0f;
      }
    }
    double uh = 0.0, ul = 0.0, uc = 1.0, us = 0.0, uw = 1.0;
    if (stager && b0 + st_at < b1) {
      const uint32_t o = st_base + static_cast<uint32_t>(b0 + st_at);
      if (TABLE) {
        const double2 an = anc[b0 + st_at];
        uh = an.x;
        ul = an.y;
      } else {
        uh = u_hi[o];
        ul = u_lo[o];
      }
      uc = u_c[o];
      us = u_s[o];
      if (st_w) uw = a.w[b0 + st_at];
    }
    // Write chunk c0's phases into buffer buf from the prefetched turns, then
    // prefetch the turns of chunk c0 + KC.
    auto stage = [&](int buf, int c0) {
      if (!stager) return;
      double2* dst = ph[buf][st_at] + st_slot;
      if (c0 + st_at < b1) {
        if (TABLE)
          phase8_from(uh, ul, uc, us, dst, uw);
        else
          phase8(uh, ul, st_n0, uc, us, dst, uw);
      } else {
#pragma unroll
        for (int k = 0; k < 8; ++k) dst[k] = make_double2(0.0, 0.0);
      }
      const int nx = c0 + KC + st_at;
      if (nx < b1) {
        const uint32_t o = st_base + static_cast<uint32_t>(nx);
        if (TABLE) {
          const double2 an = anc[nx];
          uh = an.x;
          ul = an.y;
        } else {
          uh = u_hi[o];
          ul = u_lo[o];
        }
        uc = u_c[o];
        us = u_s[o];
        if (st_w) uw = a.w[nx];
      }
    };
    __syncthreads();  // the previous species' readers are done with both buffers
    if (b0 < b1) stage(0, b0);
    __syncthreads();
    int it = 0;
    for (int g0 = b0; g0 < b1; g0 += KC * kFoldChunks) {
      const int g1 = min(b1, g0 + KC * kFoldChunks);
      for (int c0 = g0; c0 < g1; c0 += KC, ++it) {
        const int cur = it & 1;
        if (c0 + KC < b1) stage(cur ^ 1, c0 + KC);
        if (cb == 3)
          contract_chunk_f64<3>(ph[cur], kq, mx, my, i16, t1, t2, t3);
        else if (cb == 1)
          contract_chunk_f64<1>(ph[cur], kq, mx, my, i16, t1, t2, t3);
        else if (cb == 2)
          contract_chunk_f64<2>(ph[cur], kq, mx, my, i16, t1, t2, t3);
        __syncthreads();
      }
      for (int nb = 0; nb < 2; ++nb) {
        if (!((cb >> nb) & 1)) continue;
        for (int r = 0; r < 4; ++r) {
          two_sum_acc(sumR[nb][r], cmpR[nb][r], t1[nb][r] - t2[nb][r]);
          two_sum_acc(sumI[nb][r], cmpI[nb][r], t3[nb][r] - t1[nb][r] - t2[nb][r]);
        }
        t1[nb] = d4{0.0, 0.0, 0.0, 0.0};
        t2[nb] = t1[nb];
        t3[nb] = t1[nb];
      }
    }
    // D layout of v_mfma_f64_16x16x4_f64: register r of a lane holds row
    // 4*r + lane/16, column lane%16 (the FP32/16-bit 16x16 forms hold row
    // 4*(lane/16) + r instead).
    // Vector index within the tile: row mrow = w*16 + 4r + kq (= kx*8 + ky),
    // column nb*16 + i16 (kz), i.e. mrow * TKZ + nb * 16 + i16.
    double2* out = reinterpret_cast<double2*>(a.rho) +
                   ((wi.plane0 + (s - wi.c0)) * a.ntiles + tile) * TILE_VEC +
                   ((w * 16 + kq) * TKZ + i16);
#pragma unroll
    for (int nb = 0; nb < 2; ++nb) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        out[4 * r * TKZ + nb * 16] = make_double2(sumR[nb][r] + static_cast<double>(cmpR[nb][r]),
                                                  sumI[nb][r] + static_cast<double>(cmpI[nb][r]));
      }
    }
  }
}

// ---------------------------------------------------------- reduced precision
// FP32 tile 64 kx x 4 ky x 32 kz: M = 64 kx (one 16-row block per wavefront),
// N = 128 (ky,kz) pairs (eight 16-column blocks), K = 32 atoms per chunk.
// The phase products ey*ez are formed once per workgroup (the B operand all
// four wavefronts share); ex needs no product.  (The 16-bit kernel further
// down uses 64 x 8 x 32 tiles and forms B per wavefront.)
//
// FP32 kernel (v_mfma_f32_16x16x4_f32, MFMA-bound): phases by exact FP64
// reduction of the anchor n0*u and of the step u, FP32 sincospi, FP32 angle
// addition; FP32 operand tables in LDS.
constexpr int LX = 64, LY = 4, LZ = 32, LN = LY * LZ;
constexpr int LKC = 32;           // atoms per chunk (K of one 16x16x32 MFMA)
constexpr int LKF = LKC + 1;      // padded row length of the FP32 operand tables

typedef float f4 __attribute__((ext_vector_type(4)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));

template <int PREC>
struct LpTypes;
template <>
struct LpTypes<TRJ_PREC_FP16X3> {
  typedef _Float16 T;
  typedef h8 V;
  __device__ static f4 mfma(V a, V b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }
};
template <>
struct LpTypes<TRJ_PREC_BF16X3> {
  typedef __bf16 T;
  typedef b8 V;
  __device__ static f4 mfma(V a, V b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
};

// e^{2 pi i n u}, n = n0 .. n0+cnt-1, in FP32 (cnt <= 8).
__device__ inline void phase_f32(double uh, double ul, int n0, int cnt, float* re, float* im) {
  float s, c, ss, cs;
  sincospif(static_cast<float>(2.0 * frac_turns(n0, uh, ul)), &s, &c);
  sincospif(static_cast<float>(2.0 * frac_turns(1, uh, ul)), &ss, &cs);
  float r = c, i = s;
  for (int k = 0; k < cnt; ++k) {
    re[k] = r;
    im[k] = i;
    const float nr = fmaf(r, cs, -(i * ss));
    const float ni = fmaf(r, ss, i * cs);
    r = nr;
    i = ni;
  }
}

struct LpShared32 {
  float A[2][LX][LKF];   // re, im of ex
  float B[2][LN][LKF];   // re, im of ey*ez
  float2 ey[LKC][LY];
  float2 ez[LKC][LZ];
};

__global__ __launch_bounds__(256) void k_sq_contract_f32(ContractArgs a) {
  __shared__ LpShared32 sh;
  const int64_t wk = contract_work_index(a, a.items == nullptr);
  if (wk >= static_cast<int64_t>(a.ntiles) * a.ksplit) return;
  const int tile = static_cast<int>(wk % a.ntiles);
  const ItemRange wi = item_range(a, static_cast<int>(wk / a.ntiles));
  if (wi.skip) return;
  const int kx0 = a.tiles[4 * tile], ky0 = a.tiles[4 * tile + 1], kz0 = a.tiles[4 * tile + 2];
  const int t = threadIdx.x, w = t / kWave, lane = t & (kWave - 1);
  const int i16 = lane & 15, kq = lane >> 4;
  const double* __restrict__ ux = a.u;
  const double* __restrict__ uy = a.u + a.n;
  const double* __restrict__ uz = a.u + 2 * a.n;
  const double* __restrict__ lx = a.u + 3 * a.n;
  const double* __restrict__ ly = a.u + 4 * a.n;
  const double* __restrict__ lz = a.u + 5 * a.n;
  for (int s = wi.c0; s < wi.c1; ++s) {
    const int a0 = a.sp_off[s], a1 = a.sp_off[s + 1];
    const int64_t len = a1 - a0;
    const int b0 = a0 + static_cast<int>(len * wi.split / wi.splits);
    const int b1 = a0 + static_cast<int>(len * (wi.split + 1) / wi.splits);
    // FP32 accumulation over this workgroup's atoms (at most 8192 per species,
    // set by the host through ksplit); each split's sums are stored as they
    // are (FP32 complex) and the K-splits are summed in FP64.
    f4 accR[8], accI[8];
    for (int nb = 0; nb < 8; ++nb) {
      accR[nb] = f4{0.f, 0.f, 0.f, 0.f};
      accI[nb] = accR[nb];
    }
    for (int c0 = b0; c0 < b1; c0 += LKC) {
      __syncthreads();
      if (t < 64) {
        // ex: 8 kx x 4 atoms per thread (8 segments x 8 atom quads).
        const int seg = t & 7, aq = (t >> 3) * 4;
        float re[4][8], im[4][8];
        for (int q = 0; q < 4; ++q) {
          const int ai = c0 + aq + q;
          if (ai < b1) {
            phase_f32(ux[ai], lx[ai], kx0 + 8 * seg, 8, re[q], im[q]);
            if (a.w) {  // channel weight on ex
              const float wq = static_cast<float>(a.w[ai]);
              for (int k = 0; k < 8; ++k) {
                re[q][k] *= wq;
                im[q][k] *= wq;
              }
            }
          } else {
            for (int k = 0; k < 8; ++k) re[q][k] = im[q][k] = 0.f;
          }
        }
        for (int k = 0; k < 8; ++k) {
          const int row = 8 * seg + k;
          for (int q = 0; q < 4; ++q) {
            sh.A[0][row][aq + q] = re[q][k];
            sh.A[1][row][aq + q] = im[q][k];
          }
        }
      } else if (t < 64 + LKC * 5) {
        // ey (4 values) and ez (4 segments of 8) per atom.
        const int tt = t - 64, at = tt / 5, sub = tt % 5, ai = c0 + at;
        float re[8], im[8];
        const int cnt = (sub == 0) ? LY : 8;
        if (ai < b1) {
          if (sub == 0)
            phase_f32(uy[ai], ly[ai], ky0, LY, re, im);
          else
            phase_f32(uz[ai], lz[ai], kz0 + 8 * (sub - 1), 8, re, im);
        } else {
          for (int k = 0; k < 8; ++k) re[k] = im[k] = 0.f;
        }
        for (int k = 0; k < cnt; ++k) {
          if (sub == 0)
            sh.ey[at][k] = make_float2(re[k], im[k]);
          else
            sh.ez[at][8 * (sub - 1) + k] = make_float2(re[k], im[k]);
        }
      }
      __syncthreads();
      {  // B = ey*ez: one (ky,kz) row x 16 atoms per thread, stored 8 atoms at a time.
        const int n = t & (LN - 1), ab = (t >> 7) * 16;
        const int ky = n / LZ, kz = n % LZ;
        for (int h = 0; h < 2; ++h) {
          float re[8], im[8];
          for (int k = 0; k < 8; ++k) {
            const float2 y = sh.ey[ab + 8 * h + k][ky], z = sh.ez[ab + 8 * h + k][kz];
            re[k] = fmaf(y.x, z.x, -(y.y * z.y));
            im[k] = fmaf(y.x, z.y, y.y * z.x);
          }
          for (int k = 0; k < 8; ++k) {
            sh.B[0][n][ab + 8 * h + k] = re[k];
            sh.B[1][n][ab + 8 * h + k] = im[k];
          }
        }
      }
      __syncthreads();
      const int row = 16 * w + i16;
#pragma unroll
      for (int kk = 0; kk < LKC / 4; ++kk) {
        const int at = 4 * kk + kq;
        const float ar = sh.A[0][row][at], ai = sh.A[1][row][at];
#pragma unroll
        for (int nb = 0; nb < 8; ++nb) {
          const float br = sh.B[0][16 * nb + i16][at], bi = sh.B[1][16 * nb + i16][at];
          accR[nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, br, accR[nb], 0, 0, 0);
          accI[nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, bi, accI[nb], 0, 0, 0);
          accR[nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(-ai, bi, accR[nb], 0, 0, 0);
          accI[nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(ai, br, accI[nb], 0, 0, 0);
        }
      }
    }
    const size_t base = ((wi.plane0 + (s - wi.c0)) * a.ntiles + tile) * (LX * LN);
#pragma unroll
    for (int nb = 0; nb < 8; ++nb) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int kxl = 16 * w + 4 * kq + r;
        const int n = 16 * nb + i16;
        const int v = (kxl * LY + n / LZ) * LZ + (n % LZ);
        reinterpret_cast<float2*>(a.rho)[base + v] = make_float2(accR[nb][r], accI[nb][r]);
      }
    }
  }
}

// ---------------------------------------------------- 16-bit split kernels
// Same tile and GEMM as k_sq_contract_f32 (M = 64 kx, one 16-row block per
// wavefront; N = 128 (ky,kz) pairs; K = 32 atoms per chunk), organised so
// that the VALU work and the MFMAs of different wavefronts overlap:
//  * each chunk's phases are staged once as FP32 complex records in LDS,
//    value-major (row = kx, kz or ky value, column = atom); every lane runs
//    the same code: one 16-value segment (FP64 reduction of the anchor, FP32
//    polynomial, FP32 angle addition), branch-free (atoms past the end give
//    zero ex), so the staging of chunk c+1 is scheduled among the MFMAs of c;
//  * the B operand ey*ez (shared by the four wavefronts) is formed and split
//    into hi + lo once per chunk, stored 8 atoms per 16-byte entry so that the
//    MFMA operand loads are contiguous (no padding); each wavefront splits its
//    own 16 rows of ex straight into registers;
//  * the -Im*Im products accumulate separately (no operand negation).
typedef float f8 __attribute__((ext_vector_type(8)));

// hi = round(x), lo = round(x - hi), eight values at a time (packed converts).
template <class V>
__device__ inline void split8(const f8& x, V& hi, V& lo) {
  hi = __builtin_convertvector(x, V);
  const f8 back = __builtin_convertvector(hi, f8);
  lo = __builtin_convertvector(x - back, V);
}
// FP16: lo = round(x - hi) in one v_fma_mix{lo,hi}_f16 per value (x - hi is
// exact in FP32 and rounded once to FP16, as the generic form does with a
// widening convert, a subtraction and a narrowing convert).
__device__ inline uint32_t lo_pair_f16(float x0, float x1, uint32_t h) {
  uint32_t d;
  asm("v_fma_mixlo_f16 %0, %1, 1.0, -%2 op_sel_hi:[0,0,1]\n\t"
      "v_fma_mixhi_f16 %0, %3, 1.0, -%2 op_sel:[0,0,1] op_sel_hi:[0,0,1]"
      : "=&v"(d)
      : "v"(x0), "v"(h), "v"(x1));
  return d;
}
template <>
__device__ inline void split8<h8>(const f8& x, h8& hi, h8& lo) {
  hi = __builtin_convertvector(x, h8);
  uint32_t ph[4], pl[4];
  __builtin_memcpy(ph, &hi, sizeof(ph));
#pragma unroll
  for (int k = 0; k < 4; ++k) pl[k] = lo_pair_f16(x[2 * k], x[2 * k + 1], ph[k]);
  __builtin_memcpy(&lo, pl, sizeof(pl));
}
// BF16: the FP32 value of a packed pair's halves is a bit shift or mask, so hi
// is widened by bit operations instead of a second rounding per element.
template <>
__device__ inline void split8<b8>(const f8& x, b8& hi, b8& lo) {
  typedef __bf16 b2 __attribute__((ext_vector_type(2)));
  typedef float f2 __attribute__((ext_vector_type(2)));
  uint32_t ph[4], pl[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const b2 h = __builtin_convertvector(f2{x[2 * k], x[2 * k + 1]}, b2);
    const uint32_t p = __builtin_bit_cast(uint32_t, h);
    const float b0 = __uint_as_float(p << 16), b1 = __uint_as_float(p & 0xffff0000u);
    const b2 l = __builtin_convertvector(f2{x[2 * k] - b0, x[2 * k + 1] - b1}, b2);
    ph[k] = p;
    pl[k] = __builtin_bit_cast(uint32_t, l);
  }
  hi = __builtin_bit_cast(b8, ph);
  lo = __builtin_bit_cast(b8, pl);
}

// e^{2 pi i f} in FP32 for a turn fraction f (|f| <= 1/2, FP64): exact
// quarter-turn reduction in FP64, Taylor polynomials of sin and cos on
// [-pi/4, pi/4] (truncation below 3e-8), rotation by the quarter turns.
__device__ inline float2 cis_turns_f32(double f) {
  const double t = 4.0 * f;  // quarter turns (exact), |t| <= 2
  const double q = rint(t);
  const float x = static_cast<float>(t - q) * 1.57079632679489662f;  // |x| <= pi/4
  const float x2 = x * x;
  // sin x = x - x^3/3! + x^5/5! - x^7/7! + x^9/9!
  float sp = fmaf(x2, 2.75573192e-6f, -1.98412698e-4f);
  sp = fmaf(x2, sp, 8.33333333e-3f);
  sp = fmaf(x2, sp, -1.66666667e-1f);
  const float sn = fmaf(x * x2, sp, x);
  // cos x = 1 - x^2/2! + x^4/4! - x^6/6! + x^8/8!
  float cp = fmaf(x2, 2.48015873e-5f, -1.38888889e-3f);
  cp = fmaf(x2, cp, 4.16666667e-2f);
  cp = fmaf(x2, cp, -0.5f);
  const float cs = fmaf(x2, cp, 1.0f);
  // times i^q: q = 1 -> (-sn, cs), 2 -> (-cs, -sn), 3 -> (sn, -cs)
  const int qi = static_cast<int>(q) & 3;
  const float re = (qi & 1) ? sn : cs, im = (qi & 1) ? cs : sn;
  return make_float2(((qi + 1) & 2) ? -re : re, (qi & 2) ? -im : im);
}

// FP32 anchor table of the 16-bit contraction: rows as k_sq_anchors
// (cis_turns_f32 of the exactly reduced turns, the values the kernel would
// compute), then row nrow + axis: the FP32 unit step e^{2 pi i u} of the axis.
__global__ __launch_bounds__(256) void k_sq_anchors32(const double* __restrict__ u, int64_t n, AnchorAxes ax,
                                                      int nrow, float2* __restrict__ anc, const int32_t* live_off,
                                                      const int32_t* live_nch) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (j >= (live_off ? static_cast<int64_t>(live_off[*live_nch]) : n)) return;
  const int a = blockIdx.y;
  const double uh = u[a * n + j], ul = u[(3 + a) * n + j];
  anc[static_cast<int64_t>(nrow + a) * n + j] =
      make_float2(static_cast<float>(u[(6 + a) * n + j]), static_cast<float>(u[(9 + a) * n + j]));
  float2* dst = anc + static_cast<int64_t>(ax.row0[a]) * n + j;
  for (int m = 0; m < ax.rows[a]; ++m) dst[m * n] = cis_turns_f32(frac_turns(ax.base[a] + 8 * m, uh, ul));
}

// Phase records of the 16-bit kernel: e^{2 pi i n u} in FP32 complex, value-
// major (row = one kx, kz or ky value, column = atom), row stride kRS float2
// (32 atoms + 2: lane-per-atom writes and the operand reads are free of bank
// conflicts).  Each segment is an anchor from the FP64-reduced turns and then
// steps by the FP64 unit step rounded to FP32 (8 steps add at most ~1e-6
// relative error, below the 16-bit operands' own).
constexpr int kRS = 34;

// 16-bit contraction (FP16 / BF16 hi+lo splits, three products per real
// product) on tiles of 64 kx x 8 ky x 32 kz with eight wavefronts, two per
// SIMD.  Wavefront w owns the 32 columns of ky = ky0 + w against all 64 rows:
// it forms and splits its B operand ey * ez in registers, and only the A
// operand (ex, split once per workgroup) goes through LDS, 16 KB per
// wavefront and chunk (a shared B would be 32 KB).  Per 32-atom chunk i:
//   S1(i): the phases of the chunk into rec[i % 3]: 13 segments of 8 values
//          per atom (x 8, z 4, y 1; lanes past 416 repeat segments 0-2,
//          writing the same values to the same places);
//   S2(i): A split into As[i % 2] (4 atoms per lane);
//   S3(i): B of the wavefront's columns into registers, then 96 MFMAs
//          (v_mfma_f32_16x16x32_{f16,bf16}; the -Im*Im products in their own
//          accumulators, no operand negation).
// The two wavefronts of a SIMD alternate between S3's MFMAs and the VALU work
// (B of S3, S2, S1), two barriers per chunk; the LDS buffers (three phase
// buffers, two A buffers) cover the half-step offset between them.  The turns
// of a chunk are loaded one chunk step before they are staged.  FP32
// accumulation runs over at most 8192 atoms per workgroup and channel (the
// K-splits, stored as FP32 complex partials, are summed in FP64).
constexpr int WX = 64, WY = 8, WZ = 32, WN = WY * WZ, WROWS = 104;
template <int PREC, bool TABLE>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2))) void k_sq_contract_lp16w(
    ContractArgs a) {
  typedef LpTypes<PREC> LT;
  typedef typename LT::V V;
  __shared__ __attribute__((aligned(16))) float2 rec[3][WROWS][kRS];  // x 0..63, z 64..95, y 96..103
  __shared__ __attribute__((aligned(16))) V As[2][4][LKC / 8][WX];     // re_hi, re_lo, im_hi, im_lo of ex
  const int64_t wk = contract_work_index(a, a.items == nullptr);
  if (wk >= static_cast<int64_t>(a.ntiles) * a.ksplit) return;
  const int tile = static_cast<int>(wk % a.ntiles);
  const ItemRange wi = item_range(a, static_cast<int>(wk / a.ntiles));
  if (wi.skip) return;
  const int kx0 = a.tiles[4 * tile], ky0 = a.tiles[4 * tile + 1], kz0 = a.tiles[4 * tile + 2];
  const int t = threadIdx.x, w = t / kWave, lane = t & (kWave - 1);
  const int i16 = lane & 15, kq = lane >> 4;
  // Staging: segment sg of atom sa.
  const int sa = t % LKC;
  const int sg = (t / LKC) % 13;
  const int axis = sg < 8 ? 0 : (sg < 12 ? 2 : 1);
  const int sn0 = sg < 8 ? kx0 + 8 * sg : (sg < 12 ? kz0 + 8 * (sg - 8) : ky0);
  const int drow = sg < 8 ? 8 * sg : (sg < 12 ? 64 + 8 * (sg - 8) : 96);
  const uint32_t n32 = static_cast<uint32_t>(a.n);
  const double* u_hi = a.u + axis * n32;
  const double* u_lo = a.u + (3 + axis) * n32;
  const double* u_c = a.u + (6 + axis) * n32;
  const double* u_s = a.u + (9 + axis) * n32;
  const double* wsrc = a.w ? a.w : u_hi;
  // With the anchor table: the segment's FP32 anchors and the axis' FP32
  // unit step replace the turns hi, lo, cos, sin (no FP64 reduction here).
  const float2* t_anc =
      TABLE ? a.anc32 + static_cast<int64_t>(a.anc_row[axis] + (sn0 - a.anc_base[axis]) / 8) * a.n : nullptr;
  const float2* t_step = TABLE ? a.anc32 + static_cast<int64_t>(a.anc_nrows + axis) * a.n : nullptr;
  // A split: row ar, atoms 4 aq .. 4 aq + 3 (half of one 8-atom entry).
  const int ar = t & 63, aq = t >> 6;
  for (int s = wi.c0; s < wi.c1; ++s) {
    const int a0 = a.sp_off[s], a1 = a.sp_off[s + 1];
    const int64_t len = a1 - a0;
    const int b0 = a0 + static_cast<int>(len * wi.split / wi.splits);
    const int b1 = a0 + static_cast<int>(len * (wi.split + 1) / wi.splits);
    if (b0 >= b1) {
      const size_t base = ((wi.plane0 + (s - wi.c0)) * a.ntiles + tile) * (WX * WN);
      for (int v = t; v < WX * WN; v += 512) reinterpret_cast<float2*>(a.rho)[base + v] = make_float2(0.f, 0.f);
      continue;
    }
    f4 accR[4][2], accI[4][2], accM[4][2];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        accR[mb][e] = f4{0.f, 0.f, 0.f, 0.f};
        accI[mb][e] = accR[mb][e];
        accM[mb][e] = accR[mb][e];
      }
    // Turns of one chunk, loaded one chunk step before they are staged.
    struct Turns {
      double h, l, c, s, w;
      float2 an, st;
      float fw, f1;
    } F;
    auto fetch = [&](Turns& f, int c0) {
      const int ai = c0 + sa;
      const uint32_t o = static_cast<uint32_t>(ai < b1 ? ai : b1 - 1);
      if (TABLE) {
        f.an = t_anc[o];
        f.st = t_step[o];
      } else {
        f.h = u_hi[o];
        f.l = u_lo[o];
        f.c = u_c[o];
        f.s = u_s[o];
      }
      // The weight is loaded and used unconditionally (from the turns when
      // there are none; times 0 then) so that no branch splits the loop body:
      // the factor is w (ex, weights), 1 (ex without weights, ey, ez), 0 (past b1).
      f.w = wsrc[o];
      const bool xin = axis == 0 && ai < b1;
      f.fw = (xin && a.w) ? 1.f : 0.f;
      f.f1 = (axis != 0 || (xin && !a.w)) ? 1.f : 0.f;
    };
    // Phases of one segment in pieces: the anchor, then steps k0 .. k1-1
    // (the rotation state carries over between pieces).
    float pre, pim, psc, pss;
    auto stage_begin = [&](const Turns& f) {
      const float2 an = TABLE ? f.an : cis_turns_f32(frac_turns(sn0, f.h, f.l));
      const float m = fmaf(static_cast<float>(f.w), f.fw, f.f1);
      psc = TABLE ? f.st.x : static_cast<float>(f.c);
      pss = TABLE ? f.st.y : static_cast<float>(f.s);
      pre = an.x * m;
      pim = an.y * m;
    };
    auto stage_steps = [&](int buf, int k0, int k1) {
      float2* dst = &rec[buf][drow][sa];
#pragma unroll
      for (int k = k0; k < k1; ++k) {
        dst[k * kRS] = make_float2(pre, pim);
        const float nr = fmaf(pre, psc, -(pim * pss));
        const float ni = fmaf(pre, pss, pim * psc);
        pre = nr;
        pim = ni;
      }
    };
    auto split_a = [&](int rb, int buf) {
      const float4* er = reinterpret_cast<const float4*>(&rec[rb][ar][4 * aq]);
      const float4 e0 = er[0], e1 = er[1];
      // re of the 4 atoms in halves 0-3, im in halves 4-7
      const f8 x = {e0.x, e0.z, e1.x, e1.z, e0.y, e0.w, e1.y, e1.w};
      V hi, lo;
      split8(x, hi, lo);
      typedef uint32_t u2 __attribute__((ext_vector_type(2)));
      const uint4 ph = __builtin_bit_cast(uint4, hi), pl = __builtin_bit_cast(uint4, lo);
      u2* dst = reinterpret_cast<u2*>(&As[buf][0][aq >> 1][ar]) + (aq & 1);
      constexpr int kPart = (LKC / 8) * WX * 2;  // one part, in 8-byte units
      dst[0] = u2{ph.x, ph.y};
      dst[kPart] = u2{pl.x, pl.y};
      dst[2 * kPart] = u2{ph.z, ph.w};
      dst[3 * kPart] = u2{pl.z, pl.w};
    };
    auto gen_b = [&](int buf, V (&B)[2][4]) {
      const float4* yr = reinterpret_cast<const float4*>(&rec[buf][96 + w][8 * kq]);
      float4 yv[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) yv[k] = yr[k];
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const float4* zr = reinterpret_cast<const float4*>(&rec[buf][64 + 16 * e + i16][8 * kq]);
        f8 re, im;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const float4 zv = zr[k];
          re[2 * k] = fmaf(yv[k].x, zv.x, -(yv[k].y * zv.y));
          im[2 * k] = fmaf(yv[k].x, zv.y, yv[k].y * zv.x);
          re[2 * k + 1] = fmaf(yv[k].z, zv.z, -(yv[k].w * zv.w));
          im[2 * k + 1] = fmaf(yv[k].z, zv.w, yv[k].w * zv.z);
        }
        split8(re, B[e][0], B[e][1]);
        split8(im, B[e][2], B[e][3]);
      }
    };
    auto mma = [&](const V (&Af)[4], const V (&B)[2][4], int mb) {
      {
        const V arh = Af[0], arl = Af[1], aih = Af[2], ail = Af[3];
        f4 r0 = accR[mb][0], i0 = accI[mb][0], r1 = accR[mb][1], i1 = accI[mb][1];
        f4 m0 = accM[mb][0], m1 = accM[mb][1];
        r0 = LT::mfma(arh, B[0][0], r0);
        i0 = LT::mfma(arh, B[0][2], i0);
        r1 = LT::mfma(arh, B[1][0], r1);
        i1 = LT::mfma(arh, B[1][2], i1);
        r0 = LT::mfma(arh, B[0][1], r0);
        i0 = LT::mfma(arh, B[0][3], i0);
        r1 = LT::mfma(arh, B[1][1], r1);
        i1 = LT::mfma(arh, B[1][3], i1);
        r0 = LT::mfma(arl, B[0][0], r0);
        i0 = LT::mfma(arl, B[0][2], i0);
        r1 = LT::mfma(arl, B[1][0], r1);
        i1 = LT::mfma(arl, B[1][2], i1);
        m0 = LT::mfma(aih, B[0][2], m0);
        i0 = LT::mfma(aih, B[0][0], i0);
        m1 = LT::mfma(aih, B[1][2], m1);
        i1 = LT::mfma(aih, B[1][0], i1);
        m0 = LT::mfma(aih, B[0][3], m0);
        i0 = LT::mfma(aih, B[0][1], i0);
        m1 = LT::mfma(aih, B[1][3], m1);
        i1 = LT::mfma(aih, B[1][1], i1);
        m0 = LT::mfma(ail, B[0][2], m0);
        i0 = LT::mfma(ail, B[0][0], i0);
        m1 = LT::mfma(ail, B[1][2], m1);
        i1 = LT::mfma(ail, B[1][0], i1);
        accR[mb][0] = r0;
        accI[mb][0] = i0;
        accR[mb][1] = r1;
        accI[mb][1] = i1;
        accM[mb][0] = m0;
        accM[mb][1] = m1;
      }
    };
    V B0[2][4];
    __syncthreads();  // the previous channel's readers are done with rec and As
    fetch(F, b0);
    stage_begin(F);
    stage_steps(0, 0, 8);
    fetch(F, b0 + LKC);
    __syncthreads();
    split_a(0, 0);
    stage_begin(F);
    stage_steps(1, 0, 8);
    fetch(F, b0 + 2 * LKC);
    __syncthreads();
    // Two groups of wavefronts, one of each per SIMD (wavefront w runs on
    // SIMD w % 4), take turns on the matrix cores: while one group issues
    // its 96 MFMAs of chunk c, the other does its VALU work, and they swap at
    // a barrier.  Per chunk c (phases of chunk c in rec[rb], A in As[ab]):
    //   H1: group 0 MFMAs of c; group 1 B of c, its half of the A split of
    //       c+1 (rec[rb+1] -> As[1-ab]) and its segments of the phases of c+2
    //       (-> rec[rb+2], mod 3);
    //   H2: group 1 MFMAs of c; group 0 B of c+1, its half of the A split of
    //       c+1 and its segments of the phases of c+2.
    // As[ab] is read by both groups' MFMAs and rewritten (for c+2) only after
    // H2; rec[rb+2] held chunk c-1, whose last reader ran in H1 of c-1.  The
    // turns are refilled (for c+3) right after they are staged.  Past b1 the
    // stages run on zero-weight atoms.
    const bool g0 = __builtin_amdgcn_readfirstlane(w) < 4;
    auto valu = [&](int rgen, int rnext, int anext, int rstage, int cnext) {
      gen_b(rgen, B0);
      split_a(rnext, anext);
      stage_begin(F);
      stage_steps(rstage, 0, 8);
      fetch(F, cnext);
    };
    auto mfma_all = [&](int abuf) {
#pragma unroll
      for (int mb = 0; mb < 4; ++mb) {
        V Af[4];
#pragma unroll
        for (int p = 0; p < 4; ++p) Af[p] = As[abuf][p][kq][16 * mb + i16];
        mma(Af, B0, mb);
      }
    };
    // Each group runs its own loop (the same number of barriers in both), so
    // that no register lives across the two schedules.  The scheduler may not
    // move MFMAs (register-only work) across a barrier.
    auto half_barrier = [] {
      __builtin_amdgcn_sched_barrier(0);
      __syncthreads();
      __builtin_amdgcn_sched_barrier(0);
    };
    int rb = 0;
    if (g0) {
      gen_b(0, B0);
      for (int c0 = b0;;) {
        const int r1 = rb == 2 ? 0 : rb + 1, r2 = rb == 0 ? 2 : rb - 1;
        mfma_all(0);
        half_barrier();
        valu(r1, r1, 1, r2, c0 + 3 * LKC);
        half_barrier();
        c0 += LKC;
        rb = r1;
        if (c0 >= b1) break;
        const int q1 = rb == 2 ? 0 : rb + 1, q2 = rb == 0 ? 2 : rb - 1;
        mfma_all(1);
        half_barrier();
        valu(q1, q1, 0, q2, c0 + 3 * LKC);
        half_barrier();
        c0 += LKC;
        rb = q1;
        if (c0 >= b1) break;
      }
    } else {
      for (int c0 = b0;;) {
        const int r1 = rb == 2 ? 0 : rb + 1, r2 = rb == 0 ? 2 : rb - 1;
        valu(rb, r1, 1, r2, c0 + 3 * LKC);
        half_barrier();
        mfma_all(0);
        half_barrier();
        c0 += LKC;
        rb = r1;
        if (c0 >= b1) break;
        const int q1 = rb == 2 ? 0 : rb + 1, q2 = rb == 0 ? 2 : rb - 1;
        valu(rb, q1, 0, q2, c0 + 3 * LKC);
        half_barrier();
        mfma_all(1);
        half_barrier();
        c0 += LKC;
        rb = q1;
        if (c0 >= b1) break;
      }
    }
    // D layout: register r holds row 4*(lane/16) + r, column lane%16.
    const size_t base = ((wi.plane0 + (s - wi.c0)) * a.ntiles + tile) * (WX * WN);
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
      for (int e = 0; e < 2; ++e)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int kxl = 16 * mb + 4 * kq + r;
          const int v = (kxl * WY + w) * WZ + 16 * e + i16;
          // Re = r - m rounded once to FP32 (|error| <= 2^-24 |Re| per split,
          // far below the FP32 accumulation error of the split itself)
          reinterpret_cast<float2*>(a.rho)[base + v] = make_float2(accR[mb][e][r] - accM[mb][e][r], accI[mb][e][r]);
        }
  }
}

struct Bsc {
  double b[kNspMax];
};

// Neumaier's compensated sum: s + c carries the sum of the added values with
// the rounding error of each addition kept in c (the split partials of a
// coherent vector can be much larger than their sum).
struct CSum {
  double s = 0.0, c = 0.0;
  __device__ inline void add(double x) {
    const double t = s + x;
    c += (fabs(s) >= fabs(x)) ? (s - t) + x : (x - t) + s;
    s = t;
  }
  __device__ inline double value() const { return s + c; }
};

// |rho|^2 of one q-vector into the shell sums.  Channels: slice sl, species
// s at sl * nsp + s; the charge channel of slice sl (nch > nslice * nsp) at
// nslice * nsp + sl.  Outputs (add(kind, index, value)): kind 0 the bsc-
// weighted total of slice sl, kind 1 the partial of channel sl * nsp + s,
// kind 2 the charge channel of slice sl.  With norm (the channel offsets,
// sqf3xy/sqf2xy, sq.cuf:653-670) each value is divided by the slice's (or
// channel's) atom count and skipped when that count is 0; without, nslice = 1
// and values are added as they are.
// With qsp (charges constant within each species) the charge channel is not
// contracted: rho_q = sum_s q_s rho_s of the slice's species channels.
template <class Fetch, class Add>
__device__ inline void shell_vector(int nsp, int nslice, int nch, const Bsc& bsc, const int32_t* __restrict__ norm,
                                    const Bsc* qsp, Fetch fetch, Add add) {
  for (int sl = 0; sl < nslice; ++sl) {
    double cnt = 1.0;
    if (norm) cnt = static_cast<double>(norm[(sl + 1) * nsp] - norm[sl * nsp]);
    double tr = 0.0, ti = 0.0, qr = 0.0, qi = 0.0;
    for (int s = 0; s < nsp; ++s) {
      const int c = sl * nsp + s;
      double rr, ri;
      fetch(c, rr, ri);
      const double p = rr * rr + ri * ri;
      if (norm) {
        const double cs = static_cast<double>(norm[c + 1] - norm[c]);
        if (cs > 0.0) add(1, c, p / cs);
      } else {
        add(1, c, p);
      }
      tr += bsc.b[s] * rr;
      ti += bsc.b[s] * ri;
      if (qsp) {
        qr += qsp->b[s] * rr;
        qi += qsp->b[s] * ri;
      }
    }
    if (cnt > 0.0) {
      add(0, sl, norm ? (tr * tr + ti * ti) / cnt : tr * tr + ti * ti);
      if (qsp) {
        add(2, sl, norm ? (qr * qr + qi * qi) / cnt : qr * qr + qi * qi);
      } else if (nch > nslice * nsp) {
        double rr, ri;
        fetch(nslice * nsp + sl, rr, ri);
        add(2, sl, norm ? (rr * rr + ri * ri) / cnt : rr * rr + ri * ri);
      }
    }
  }
}

// R2 is the type of one split partial: double2 (the FP64 contraction) or
// float2 (the FP32 and 16-bit ones).
template <class R2>
__global__ __launch_bounds__(256) void k_sq_shells(const R2* __restrict__ rho,
                                                   const int16_t* __restrict__ tile_shell, int ntiles,
                                                   int tile_vec, int ksplit, int nsp, int nslice, int nch, Bsc bsc,
                                                   int nqmax, const int32_t* __restrict__ norm, Bsc qsp, int qsp_on,
                                                   double* __restrict__ sqf, double* __restrict__ sqfp,
                                                   double* __restrict__ sqfq, int use_lds) {
  extern __shared__ __align__(16) unsigned char smem[];
  // LDS accumulators: sqf [nslice][ns], sqfp [nslice*nsp][ns], sqfq [nslice][ns]
  double* acc = reinterpret_cast<double*>(smem);
  const int ns = nqmax + 1;
  const int o1 = nslice * ns, o2 = nslice * (1 + nsp) * ns;
  const int nacc = o2 + ((nch > nslice * nsp || qsp_on) ? nslice * ns : 0);
  if (use_lds) {
    for (int b = threadIdx.x; b < nacc; b += 256) acc[b] = 0.0;
    __syncthreads();
  }
  const int64_t total = static_cast<int64_t>(ntiles) * tile_vec;
  const R2* r2 = rho;
  for (int64_t e = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; e < total;
       e += static_cast<int64_t>(gridDim.x) * 256) {
    const int sh = tile_shell[e];
    if (sh < 0) continue;
    auto fetch = [&](int c, double& rr, double& ri) {
      CSum sr, si;
      for (int sp = 0; sp < ksplit; ++sp) {
        const R2 v = r2[(static_cast<int64_t>(sp) * nch + c) * total + e];
        sr.add(static_cast<double>(v.x));
        si.add(static_cast<double>(v.y));
      }
      rr = sr.value();
      ri = si.value();
    };
    auto add = [&](int kind, int idx, double v) {
      const int off = (kind == 0 ? 0 : (kind == 1 ? o1 : o2)) + idx * ns + sh;
      if (use_lds)
        atomicAdd(&acc[off], v);
      else if (kind == 0)
        atomicAdd(&sqf[idx * ns + sh], v);
      else if (kind == 1)
        atomicAdd(&sqfp[static_cast<int64_t>(idx) * ns + sh], v);
      else
        atomicAdd(&sqfq[idx * ns + sh], v);
    };
    shell_vector(nsp, nslice, nch, bsc, norm, qsp_on ? &qsp : nullptr, fetch, add);
  }
  if (use_lds) {
    __syncthreads();
    for (int b = threadIdx.x; b < nacc; b += 256) {
      const double v = acc[b];
      if (v != 0.0) {
        if (b < o1)
          atomicAdd(&sqf[b], v);
        else if (b < o2)
          atomicAdd(&sqfp[b - o1], v);
        else
          atomicAdd(&sqfq[b - o2], v);
      }
    }
  }
}

// The same sums with the K-splits of each vector spread over the four
// wavefronts of a workgroup (one vector per lane, 64 vectors per workgroup
// step): wavefront w adds splits w, w + 4, ... (compensated); the four sums
// and their compensations are combined in that order in LDS and wavefront 0
// bins the vector.  Sixteen times the loads in flight of one thread per
// vector over all its splits.  LDS: the shell accumulators (use_lds), 4 x 64
// partial sums and 4 x 64 compensations, then the vector's channel sums
// [nch][64].
template <class R2>
__global__ __launch_bounds__(256) void k_sq_shells4(const R2* __restrict__ rho,
                                                    const int16_t* __restrict__ tile_shell, int ntiles,
                                                    int tile_vec, int ksplit, int nsp, int nslice, int nch, Bsc bsc,
                                                    int nqmax, const int32_t* __restrict__ norm, Bsc qsp, int qsp_on,
                                                    double* __restrict__ sqf, double* __restrict__ sqfp,
                                                    double* __restrict__ sqfq, int use_lds) {
  extern __shared__ __align__(16) unsigned char smem[];
  double* acc = reinterpret_cast<double*>(smem);
  const int ns = nqmax + 1;
  const int o1 = nslice * ns, o2 = nslice * (1 + nsp) * ns;
  const int nacc = o2 + ((nch > nslice * nsp || qsp_on) ? nslice * ns : 0);
  double2* part = reinterpret_cast<double2*>(smem + (use_lds ? ((static_cast<size_t>(nacc) * 8 + 15) & ~size_t(15)) : 0));
  double2* chan = part + 8 * kWave;
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  if (use_lds) {
    for (int b = threadIdx.x; b < nacc; b += 256) acc[b] = 0.0;
    __syncthreads();
  }
  const int64_t total = static_cast<int64_t>(ntiles) * tile_vec;
  for (int64_t base = static_cast<int64_t>(blockIdx.x) * kWave; base < total;
       base += static_cast<int64_t>(gridDim.x) * kWave) {
    const int64_t e = base + lane;
    const int sh = e < total ? tile_shell[e] : -1;
    for (int c = 0; c < nch; ++c) {
      CSum sr, si;
      if (sh >= 0) {
        const R2* src = rho + static_cast<int64_t>(c) * total + e;
        const int64_t step = static_cast<int64_t>(nch) * total;
#pragma unroll 4
        for (int sp = w; sp < ksplit; sp += 4) {
          const R2 v = src[sp * step];
          sr.add(static_cast<double>(v.x));
          si.add(static_cast<double>(v.y));
        }
      }
      // each wavefront's sum and its compensation
      part[w * kWave + lane] = make_double2(sr.s, si.s);
      part[(4 + w) * kWave + lane] = make_double2(sr.c, si.c);
      __syncthreads();
      if (w == 0) {
        CSum tr, ti;
        for (int k = 0; k < 4; ++k) {
          tr.add(part[k * kWave + lane].x);
          ti.add(part[k * kWave + lane].y);
        }
        double cr = 0.0, ci = 0.0;
        for (int k = 0; k < 4; ++k) {
          cr += part[(4 + k) * kWave + lane].x;
          ci += part[(4 + k) * kWave + lane].y;
        }
        chan[c * kWave + lane] = make_double2(tr.s + (tr.c + cr), ti.s + (ti.c + ci));
      }
      __syncthreads();
    }
    if (w == 0 && sh >= 0) {
      auto fetch = [&](int c, double& rr, double& ri) {
        const double2 t = chan[c * kWave + lane];
        rr = t.x;
        ri = t.y;
      };
      auto add = [&](int kind, int idx, double v) {
        const int off = (kind == 0 ? 0 : (kind == 1 ? o1 : o2)) + idx * ns + sh;
        if (use_lds)
          atomicAdd(&acc[off], v);
        else if (kind == 0)
          atomicAdd(&sqf[idx * ns + sh], v);
        else if (kind == 1)
          atomicAdd(&sqfp[static_cast<int64_t>(idx) * ns + sh], v);
        else
          atomicAdd(&sqfq[idx * ns + sh], v);
      };
      shell_vector(nsp, nslice, nch, bsc, norm, qsp_on ? &qsp : nullptr, fetch, add);
    }
  }
  if (use_lds) {
    __syncthreads();
    for (int b = threadIdx.x; b < nacc; b += 256) {
      const double v = acc[b];
      if (v != 0.0) {
        if (b < o1)
          atomicAdd(&sqf[b], v);
        else if (b < o2)
          atomicAdd(&sqfp[b - o1], v);
        else
          atomicAdd(&sqfq[b - o2], v);
      }
    }
  }
}

// Explicit q-vectors: block (v, b) sums the phases of the atoms of every
// species in its 1/nb share; partial sums are combined in a fixed order by
// k_sq_list_finish (deterministic), which then bins |rho|^2 into the shells.
constexpr int kListBlocks = 64;

__global__ __launch_bounds__(256) void k_sq_list(const double* __restrict__ u, const double* __restrict__ wt,
                                                 const int32_t* __restrict__ sp_off, int64_t n, int nch,
                                                 int ndim, const int32_t* __restrict__ kv,
                                                 double* __restrict__ part) {
  __shared__ double red[2][256];
  const int v = blockIdx.x, b = blockIdx.y;
  const int kx = kv[3 * v], ky = kv[3 * v + 1], kz = kv[3 * v + 2];
  for (int s = 0; s < nch; ++s) {
    const int64_t a0 = sp_off[s], a1 = sp_off[s + 1], len = a1 - a0;
    const int64_t j0 = a0 + len * b / kListBlocks, j1 = a0 + len * (b + 1) / kListBlocks;
    double pr = 0.0, pi = 0.0;
    for (int64_t j = j0 + threadIdx.x; j < j1; j += 256) {
      double f = frac_turns(kx, u[j], u[3 * n + j]) + frac_turns(ky, u[n + j], u[4 * n + j]);
      if (ndim == 3) f += frac_turns(kz, u[2 * n + j], u[5 * n + j]);
      f -= rint(f);
      double sn, cn;
      sincospi(2.0 * f, &sn, &cn);
      const double m = wt ? wt[j] : 1.0;  // channel weight
      pr += m * cn;
      pi += m * sn;
    }
    red[0][threadIdx.x] = pr;
    red[1][threadIdx.x] = pi;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
      if (static_cast<int>(threadIdx.x) < off) {
        red[0][threadIdx.x] += red[0][threadIdx.x + off];
        red[1][threadIdx.x] += red[1][threadIdx.x + off];
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      double* o = part + ((static_cast<int64_t>(v) * kListBlocks + b) * nch + s) * 2;
      o[0] = red[0][0];
      o[1] = red[1][0];
    }
    __syncthreads();
  }
}

__global__ void k_sq_list_finish(const double* __restrict__ part, const int32_t* __restrict__ shell, int nvec,
                                 int nsp, int nslice, int nch, Bsc bsc, int nqmax, const int32_t* __restrict__ norm,
                                 Bsc qsp, int qsp_on, double* __restrict__ sqf, double* __restrict__ sqfp,
                                 double* __restrict__ sqfq) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nvec) return;
  const int sh = shell[v], ns = nqmax + 1;
  auto fetch = [&](int c, double& rr, double& ri) {
    rr = 0.0;
    ri = 0.0;
    for (int b = 0; b < kListBlocks; ++b) {
      const double* o = part + ((static_cast<int64_t>(v) * kListBlocks + b) * nch + c) * 2;
      rr += o[0];
      ri += o[1];
    }
  };
  auto add = [&](int kind, int idx, double val) {
    if (kind == 0)
      atomicAdd(&sqf[idx * ns + sh], val);
    else if (kind == 1)
      atomicAdd(&sqfp[static_cast<int64_t>(idx) * ns + sh], val);
    else
      atomicAdd(&sqfq[idx * ns + sh], val);
  };
  shell_vector(nsp, nslice, nch, bsc, norm, qsp_on ? &qsp : nullptr, fetch, add);
}

// ------------------------------------------------------------------ batches

__device__ inline int frame_of_atom(const BatchDev& d, int nframes, int64_t i) {
  int lo = 0, hi = nframes - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (d.fr[mid].local_off <= i) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__device__ inline int frame_of_task(const BatchFrameDesc* fr, int nframes, int64_t t) {
  int lo = 0, hi = nframes - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (fr[mid].task_off <= t) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Zero the cell counters and cursors (a fill kernel: graph-safe).
__global__ __launch_bounds__(256) void bk_zero(BatchDev d) {
  const int64_t nc = d.hdr->ncells;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; i <= nc;
       i += static_cast<int64_t>(gridDim.x) * 256) {
    d.count[i] = 0;
    if (i < nc) d.cursor[i] = 0;
  }
}

// Wrap (FP64, as k_ingest) and count cells, all frames of the group.
__global__ __launch_bounds__(256) void bk_ingest(BatchDev d) {
  const int nframes = d.hdr->nframes;
  const int64_t na = d.hdr->natoms;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; i < na;
       i += static_cast<int64_t>(gridDim.x) * 256) {
    const int f = frame_of_atom(d, nframes, i);
    const BatchFrameDesc& F = d.fr[f];
    const int64_t src = F.atom_off + (i - F.local_off);
    double c[3] = {0.0, 0.0, 0.0};
    uint32_t bad = 0;
    int cc[3] = {0, 0, 0};
    for (int k = 0; k < d.ndim; ++k) {
      const int64_t idx = src * d.ndim + k;
      c[k] = (d.dtype == TRJ_F32) ? static_cast<double>(static_cast<const float*>(d.pos_in)[idx])
                                  : static_cast<const double*>(d.pos_in)[idx];
      if (!isfinite(c[k])) {
        bad |= kFlagRange;
        c[k] = 0.0;
      } else if (F.periodic[k]) {
        c[k] = wrap_periodic(c[k], F.box[k]);
      } else if (c[k] < -0.25 * F.box[k] || c[k] >= 1.25 * F.box[k]) {
        bad |= kFlagRange;
      }
      int ck = static_cast<int>(floor(c[k] * F.inv_h[k]));
      cc[k] = ck < 0 ? 0 : (ck >= F.nc[k] ? F.nc[k] - 1 : ck);
    }
    const int t = d.itype_in[src];
    if (t < 1 || t > F.nsp) bad |= kFlagType;
    if (bad) atomicOr(d.flags, bad);
    d.pos[i] = c[0];
    d.pos[d.cap_atoms + i] = c[1];
    d.pos[2 * d.cap_atoms + i] = c[2];
    const int id = F.cell_off + (cc[2] * F.nc[1] + cc[1]) * F.nc[0] + cc[0];
    d.cellid[i] = id;
    atomicAdd(&d.count[id], 1);
  }
}

// Exclusive scan of count[0..ncells] into start, three phases over a fixed
// grid: per-block sums, one block scanning those, per-block rescans.
__device__ inline void block_chunk(int64_t total, int b, int nb, int64_t* lo, int64_t* hi) {
  *lo = total * b / nb;
  *hi = total * (b + 1) / nb;
}

__global__ __launch_bounds__(256) void bk_scan_reduce(BatchDev d) {
  __shared__ int32_t red[256];
  int64_t lo, hi;
  block_chunk(static_cast<int64_t>(d.hdr->ncells) + 1, blockIdx.x, gridDim.x, &lo, &hi);
  int32_t s = 0;
  for (int64_t i = lo + threadIdx.x; i < hi; i += 256) s += d.count[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (static_cast<int>(threadIdx.x) < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) d.blocksum[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(1024) void bk_scan_blocks(BatchDev d, int nb) {
  __shared__ int32_t tmp[1024];
  __shared__ int32_t carry;
  const int t = threadIdx.x;
  if (t == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < nb; base += 1024) {
    const int i = base + t;
    const int32_t v = i < nb ? d.blocksum[i] : 0;
    tmp[t] = v;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      const int32_t x = t >= off ? tmp[t - off] : 0;
      __syncthreads();
      tmp[t] += x;
      __syncthreads();
    }
    if (i < nb) d.blocksum[i] = carry + tmp[t] - v;
    __syncthreads();
    if (t == 1023) carry += tmp[1023];
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void bk_scan_apply(BatchDev d) {
  __shared__ int32_t tmp[256];
  __shared__ int32_t carry;
  int64_t lo, hi;
  block_chunk(static_cast<int64_t>(d.hdr->ncells) + 1, blockIdx.x, gridDim.x, &lo, &hi);
  const int t = threadIdx.x;
  if (t == 0) carry = d.blocksum[blockIdx.x];
  __syncthreads();
  for (int64_t base = lo; base < hi; base += 256) {
    const int64_t i = base + t;
    const int32_t v = i < hi ? d.count[i] : 0;
    tmp[t] = v;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
      const int32_t x = t >= off ? tmp[t - off] : 0;
      __syncthreads();
      tmp[t] += x;
      __syncthreads();
    }
    if (i < hi) d.start[i] = carry + tmp[t] - v;
    __syncthreads();
    if (t == 255) carry += tmp[255];
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void bk_scatter(BatchDev d) {
  const int64_t na = d.hdr->natoms;
  const int nframes = d.hdr->nframes;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; i < na;
       i += static_cast<int64_t>(gridDim.x) * 256) {
    const int c = d.cellid[i];
    const int slot = d.start[c] + atomicAdd(&d.cursor[c], 1);
    const int f = frame_of_atom(d, nframes, i);
    const BatchFrameDesc& F = d.fr[f];
    d.spos[slot] = d.pos[i];
    d.spos[d.cap_atoms + slot] = d.pos[d.cap_atoms + i];
    d.spos[2 * d.cap_atoms + slot] = d.pos[2 * d.cap_atoms + i];
    d.stype[slot] = d.itype_in[F.atom_off + (i - F.local_off)];
  }
}

// Pair tasks of all frames: workgroup b owns the contiguous task range
// [b*T/G, (b+1)*T/G) and walks it frame segment by frame segment, flushing
// its LDS histogram into the frame's int64 histogram at each segment end.
// Pairs of every frame of the group.  LDS: the four wavefront hit rings
// first (fixed size), then the frame's histogram and bin edges.
constexpr size_t kRingBytes = static_cast<size_t>(4) * kHitQ * (sizeof(double) + sizeof(int));

__global__ __launch_bounds__(256) void bk_pairs(BatchDev d) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int nframes = d.hdr->nframes;
  const int64_t T = d.hdr->ntasks;
  int64_t t0, t1;
  block_chunk(T, blockIdx.x, gridDim.x, &t0, &t1);
  if (t0 >= t1) return;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave), lane = threadIdx.x & (kWave - 1);
  double* qr2 = reinterpret_cast<double*>(smem) + wave * kHitQ;
  int* qtt = reinterpret_cast<int*>(smem + 4 * kHitQ * sizeof(double)) + wave * kHitQ;
  unsigned char* hsm = smem + kRingBytes;
  const double* __restrict__ sx = d.spos;
  const double* __restrict__ sy = d.spos + d.cap_atoms;
  const double* __restrict__ sz = d.spos + 2 * d.cap_atoms;
  int f = frame_of_task(d.fr, nframes, t0);
  Stencil st{};
  {
    int k = 1;
    const int zr = (d.ndim == 3) ? 1 : 0;
    for (int dz = -zr; dz <= zr; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx)
          if ((dz > 0) || (dz == 0 && dy > 0) || (dz == 0 && dy == 0 && dx > 0)) {
            st.off[k][0] = dx;
            st.off[k][1] = dy;
            st.off[k][2] = dz;
            ++k;
          }
    st.n = k;
  }
  for (int64_t seg0 = t0; seg0 < t1; ++f) {
    const BatchFrameDesc& F = d.fr[f];
    const int64_t seg1 = (f + 1 < nframes && d.fr[f + 1].task_off < t1) ? d.fr[f + 1].task_off : t1;
    PairArgs a{};
    a.ndim = d.ndim;
    a.nsp = F.nsp;
    a.nit = F.nit;
    a.lsmax = F.lsmax;
    for (int k = 0; k < 3; ++k) {
      a.lmi[k] = F.lmi[k];
      a.half[k] = F.half[k];
      a.cellshift[k] = F.cellshift[k];
    }
    a.side2 = F.side2;
    a.inv_dr = F.inv_dr;
    a.row0 = 0;
    a.nrows = a.nit;
    const int nbins = a.nit * a.lsmax;
    uint32_t* hist = reinterpret_cast<uint32_t*>(hsm);
    double* thr = reinterpret_cast<double*>(hsm + ((static_cast<size_t>(nbins) * 4 + 15) & ~size_t(15)));
    a.thr = d.thr;
    __syncthreads();
    hist_setup(a, hist, thr);
    bool shift = true;
    for (int k = 0; k < d.ndim; ++k) shift = shift && F.cellshift[k];
    uint32_t head = 0, tail = 0;
    for (int64_t task = seg0 + wave; task < seg1; task += 4) {
      const int64_t lt = task - F.task_off;
      const int cellL = static_cast<int>(lt / F.W);
      const int ch0 = static_cast<int>(lt % F.W);
      const int cell = F.cell_off + cellL;
      const int cs = __builtin_amdgcn_readfirstlane(d.start[cell]);
      const int ce = __builtin_amdgcn_readfirstlane(d.start[cell + 1]);
      for (int ch = ch0; cs + ch * kWave < ce; ch += F.W) {
        if (d.ndim == 3) {
          if (shift)
            rdf_home_chunk<true, true>(a, thr, hist, qr2, qtt, head, tail, lane, st, F.nc, F.periodic, F.cell_off,
                                       cellL, ch, cs, ce, d.start, sx, sy, sz, d.stype);
          else
            rdf_home_chunk<false, true>(a, thr, hist, qr2, qtt, head, tail, lane, st, F.nc, F.periodic, F.cell_off,
                                        cellL, ch, cs, ce, d.start, sx, sy, sz, d.stype);
        } else {
          if (shift)
            rdf_home_chunk<true, false>(a, thr, hist, qr2, qtt, head, tail, lane, st, F.nc, F.periodic, F.cell_off,
                                        cellL, ch, cs, ce, d.start, sx, sy, sz, d.stype);
          else
            rdf_home_chunk<false, false>(a, thr, hist, qr2, qtt, head, tail, lane, st, F.nc, F.periodic,
                                         F.cell_off, cellL, ch, cs, ce, d.start, sx, sy, sz, d.stype);
        }
      }
    }
    __builtin_amdgcn_wave_barrier();
    bin_hits(a, thr, hist, qr2, qtt, head, static_cast<int>(tail - head), lane);
    __syncthreads();
    int64_t* out = d.hist + F.hist_off;
    for (int b = threadIdx.x; b < nbins; b += blockDim.x) {
      const uint32_t v = hist[b];
      if (v) atomicAdd(reinterpret_cast<unsigned long long*>(&out[b]), static_cast<unsigned long long>(v));
    }
    seg0 = seg1;
  }
}

// LDS-atomic throughput probe: every lane adds 1 to 16 pseudo-random bins of
// a 4096-bin uint32 histogram in LDS, `iters` times over (the bins are drawn
// once, so the loop issues almost nothing but ds_add_u32).
constexpr int kProbeBins = 4096;

__global__ __launch_bounds__(256) void k_lds_atomics(int iters, uint32_t* __restrict__ out) {
  __shared__ uint32_t h[kProbeBins];
  for (int b = threadIdx.x; b < kProbeBins; b += 256) h[b] = 0u;
  __syncthreads();
  uint32_t x = (blockIdx.x * 256u + threadIdx.x) * 2654435761u + 0x9e3779b9u;
  uint32_t off[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    off[k] = x & (kProbeBins - 1);
  }
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int k = 0; k < 16; ++k) atomicAdd(&h[off[k]], 1u);
  }
  __syncthreads();
  uint32_t sum = 0;
  for (int b = threadIdx.x; b < kProbeBins; b += 256) sum += h[b];
  atomicAdd(out, sum);
}

// Dense-MFMA issue-rate probe: every wavefront runs iters x 8 independent
// accumulations of one MFMA shape (FP64 16x16x4, or 16-bit 16x16x32 with
// FP32 accumulation); no memory traffic in the loop.  The shader-clock and
// 100 MHz real-time counters bracket the loop (lane 0 of each wavefront adds
// the differences to clk[0], clk[1]) to give the sustained clock.
// Eight independent accumulation chains of one MFMA shape; a non-inlined
// function keeps the accumulators in place across iterations.
template <int PREC>
__device__ __noinline__ float mfma_chains(int iters) {
  typedef LpTypes<PREC> LT;
  typedef typename LT::V V;
  typedef typename LT::T T;
  V a, b;
  for (int k = 0; k < 8; ++k) {
    a[k] = static_cast<T>(1e-3f * k);
    b[k] = static_cast<T>(1e-3f * (threadIdx.x & 15));
  }
  f4 c0 = {0.f, 0.f, 0.f, 0.f}, c1 = c0, c2 = c0, c3 = c0, c4 = c0, c5 = c0, c6 = c0, c7 = c0;
  for (int i = 0; i < iters; ++i) {
    c0 = LT::mfma(a, b, c0);
    c1 = LT::mfma(a, b, c1);
    c2 = LT::mfma(a, b, c2);
    c3 = LT::mfma(a, b, c3);
    c4 = LT::mfma(a, b, c4);
    c5 = LT::mfma(a, b, c5);
    c6 = LT::mfma(a, b, c6);
    c7 = LT::mfma(a, b, c7);
  }
  const f4 t = c0 + c1 + c2 + c3 + c4 + c5 + c6 + c7;
  return t[0] + t[1] + t[2] + t[3];
}
__device__ __noinline__ float mfma_chains_f64(int iters) {
  typedef double d4 __attribute__((ext_vector_type(4)));
  const double a = 1e-3 * (threadIdx.x & 7), b = 1e-3 * (threadIdx.x >> 6);
  d4 c0 = {0.0, 0.0, 0.0, 0.0}, c1 = c0, c2 = c0, c3 = c0, c4 = c0, c5 = c0, c6 = c0, c7 = c0;
  for (int i = 0; i < iters; ++i) {
    c0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c0, 0, 0, 0);
    c1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c1, 0, 0, 0);
    c2 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c2, 0, 0, 0);
    c3 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c3, 0, 0, 0);
    c4 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c4, 0, 0, 0);
    c5 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c5, 0, 0, 0);
    c6 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c6, 0, 0, 0);
    c7 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c7, 0, 0, 0);
  }
  const d4 t = c0 + c1 + c2 + c3 + c4 + c5 + c6 + c7;
  return static_cast<float>(t[0] + t[1] + t[2] + t[3]);
}

template <int PREC>
__global__ __launch_bounds__(512) void k_mfma_rate(int iters, float* __restrict__ out,
                                                    unsigned long long* __restrict__ clk) {
  const uint64_t r0 = __builtin_amdgcn_s_memrealtime(), m0 = __builtin_amdgcn_s_memtime();
  const float s = PREC == TRJ_PREC_FP64 ? mfma_chains_f64(iters) : mfma_chains<PREC == TRJ_PREC_FP64 ? TRJ_PREC_BF16X3 : PREC>(iters);
  const uint64_t r1 = __builtin_amdgcn_s_memrealtime(), m1 = __builtin_amdgcn_s_memtime();
  out[blockIdx.x * 256 + threadIdx.x] = s;
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(&clk[0], static_cast<unsigned long long>(m1 - m0));
    atomicAdd(&clk[1], static_cast<unsigned long long>(r1 - r0));
  }
}

// Co-issue of 16-bit MFMAs and FP32 VALU work on one SIMD.  One 512-thread
// workgroup per CU (the launch reserves LDS so that no second one joins), two
// wavefronts per SIMD; per SIMD M = 32 * trips MFMAs (16x16x32 FP16, eight
// independent accumulations) and KV * M independent FP32 FMAs (16 chains):
//   mode 0: both wavefronts MFMAs only (M / 2 each);
//   mode 1: both wavefronts FMAs only;
//   mode 2: wavefronts 0-3 MFMAs only, 4-7 FMAs only (across wavefronts);
//   mode 3: both wavefronts interleave KV FMAs after each MFMA (in-wave).
// clk[0] sums, over workgroups, wavefront 0's shader cycles from the start to
// the closing barrier (the SIMD's time for its work).
template <int KV, bool DM, bool DV>
__device__ __forceinline__ void coissue_loop(int n, f4 (&acc)[8], float (&x)[16], h8 a, h8 b, float y, float z) {
  for (int t = 0; t < n; ++t) {
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        if (DM) acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, acc[j], 0, 0, 0);
        if (DV) {
#pragma unroll
          for (int r = 0; r < KV; ++r) x[(j * KV + r) & 15] = fmaf(x[(j * KV + r) & 15], y, z);
        }
      }
  }
}

template <int KV, int MODE>
__global__ __launch_bounds__(512) void k_coissue(int trips, float* __restrict__ out,
                                                 unsigned long long* __restrict__ clk) {
  const int w = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / kWave);
  const int lane = threadIdx.x & (kWave - 1);
  h8 a, b;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    a[i] = static_cast<_Float16>(0.001f * (lane + i));
    b[i] = static_cast<_Float16>(0.002f * (lane - i));
  }
  f4 acc[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) acc[j] = f4{0.f, 0.f, 0.f, 0.f};
  float x[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) x[j] = lane * 0.01f + j;
  const float y = 0.999f + 1e-7f * lane, z = 0.001f;
  __syncthreads();
  const uint64_t t0 = __builtin_amdgcn_s_memtime();
  if (MODE == 0) coissue_loop<KV, true, false>(trips / 2, acc, x, a, b, y, z);
  else if (MODE == 1) coissue_loop<KV, false, true>(trips / 2, acc, x, a, b, y, z);
  else if (MODE == 2) {
    if (w < 4) coissue_loop<KV, true, false>(trips, acc, x, a, b, y, z);
    else coissue_loop<KV, false, true>(trips, acc, x, a, b, y, z);
  } else coissue_loop<KV, true, true>(trips / 2, acc, x, a, b, y, z);
  __syncthreads();
  const uint64_t t1 = __builtin_amdgcn_s_memtime();
  float s = 0.f;
#pragma unroll
  for (int j = 0; j < 8; ++j) s += acc[j][0] + acc[j][1] + acc[j][2] + acc[j][3];
#pragma unroll
  for (int j = 0; j < 16; ++j) s += x[j];
  out[blockIdx.x * 512 + threadIdx.x] = s;
  if (threadIdx.x == 0) atomicAdd(&clk[0], static_cast<unsigned long long>(t1 - t0));
}

inline unsigned grid1d(int64_t n, int bs) { return static_cast<unsigned>((n + bs - 1) / bs); }

}  // namespace

void launch_ingest(const IngestArgs& a, hipStream_t s) {
  if (a.n <= 0) return;
  hipLaunchKernelGGL(k_ingest, dim3(grid1d(a.n, 256)), dim3(256), 0, s, a);
}

void launch_species_count(const int32_t* itype, const int32_t* key, int64_t n, int nkeys, int32_t* blk, int nblk,
                          hipStream_t s) {
  hipLaunchKernelGGL(k_species_count, dim3(nblk), dim3(256), 0, s, itype, key, n, nkeys, blk, nblk);
}

void launch_exclusive_scan(const int32_t* in, int32_t* out, int64_t count, hipStream_t s) {
  hipLaunchKernelGGL(k_exclusive_scan, dim3(1), dim3(1024), 0, s, in, out, count,
                     static_cast<const int32_t*>(nullptr));
}

void launch_exclusive_scan_dev(const int32_t* in, int32_t* out, const int32_t* count, hipStream_t s) {
  hipLaunchKernelGGL(k_exclusive_scan, dim3(1), dim3(1024), 0, s, in, out, int64_t(0), count);
}

void launch_species_offsets(const int32_t* blkoff, const int32_t* blk, int nblk, int nkeys, int gsize, int ngroups,
                            int32_t* sp_off, hipStream_t s) {
  hipLaunchKernelGGL(k_species_offsets, dim3(1), dim3(64), 0, s, blkoff, blk, nblk, nkeys, gsize, ngroups, sp_off);
}

void launch_species_scatter(const int32_t* itype, const int32_t* key, int64_t n, int nkeys, const int32_t* blkoff,
                            int nblk, const double* pos, const double lq[3], int ndim, const double* q,
                            const int32_t* sp_off, int64_t stride, double* u, hipStream_t s) {
  hipLaunchKernelGGL(k_species_scatter, dim3(nblk), dim3(256), 0, s, itype, key, n, nkeys, blkoff, nblk, pos,
                     lq[0], lq[1], lq[2], ndim, q, sp_off, stride, u);
}

void launch_cell_count(const double* pos, int64_t n, int ndim, const CellGrid& g, int32_t* cellid,
                       int32_t* count, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_cell_count, dim3(grid1d(n, 256)), dim3(256), 0, s, pos, n, ndim, g, cellid, count);
}

void launch_cell_scatter(const double* pos, const int32_t* itype, const int32_t* cellid, int64_t n,
                         const int32_t* start, int32_t* cursor, double* spos, int32_t* stype, const double* q,
                         double* sq, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_cell_scatter, dim3(grid1d(n, 256)), dim3(256), 0, s, pos, itype, cellid, n, start,
                     cursor, spos, stype, q, sq);
}

CellSort cell_sort_plan(int64_t n, int64_t ncell, size_t lds_max) {
  CellSort p;
  const int64_t words = (ncell + 1) / 2;
  if (n <= 0 || ncell <= 0 || ncell > (int64_t(1) << 30) || static_cast<size_t>(words) * 4 > lds_max ||
      n > int64_t(kCellSortBlocks) * 65535)
    return p;
  // at most 4 n per-workgroup counts, at most 65535 atoms a workgroup
  int64_t b = std::min<int64_t>(kCellSortBlocks, std::max<int64_t>(1, 4 * n / ncell));
  b = std::max<int64_t>(b, (n + 65534) / 65535);
  p.lds = true;
  p.blocks = static_cast<int>(b);
  p.chunk = (n + b - 1) / b;
  p.lds_bytes = static_cast<size_t>(words) * 4;
  return p;
}

void launch_cell_sort(const CellSort& p, const double* pos, const int32_t* key, int nsp, int nslice, int64_t n,
                      int ndim, const CellGrid& g, int64_t ncell, int32_t* rank, int32_t* blkcnt, int32_t* count,
                      int32_t* start, const int32_t* itype, double* spos, int32_t* stype, const double* q,
                      double* sq, hipStream_t s) {
  const int nc = static_cast<int>(ncell);
  if (p.lds_bytes > 65536)
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_cell_hist), hipFuncAttributeMaxDynamicSharedMemorySize,
                              static_cast<int>(p.lds_bytes));
  hipLaunchKernelGGL(k_cell_hist, dim3(p.blocks), dim3(1024), p.lds_bytes, s, pos, key, nsp, nslice, n, ndim, g, nc,
                     p.chunk, rank, blkcnt);
  hipLaunchKernelGGL(k_cell_colscan, dim3(static_cast<unsigned>((ncell + 1 + 63) / 64)), dim3(1024), 0, s, blkcnt,
                     p.blocks, nc, count);
  launch_exclusive_scan(count, start, ncell + 1, s);
  hipLaunchKernelGGL(k_cell_place, dim3(static_cast<unsigned>(std::min<int64_t>((n + 255) / 256, 8192))), dim3(256),
                     0, s, pos, key, nsp, nslice, n, ndim, g, nc, p.chunk, rank, blkcnt, start, itype, spos, stype, q,
                     sq);
}

size_t pair_lds_bytes(int rows, int lsmax, bool qq) {
  const size_t b = ((static_cast<size_t>(rows) * lsmax * 4 + 15) & ~size_t(15)) + static_cast<size_t>(lsmax + 3) * 8;
  return qq ? ((b + 15) & ~size_t(15)) + static_cast<size_t>(lsmax) * 8 : b;
}

// Layout of the cell-list kernel's LDS: the histogram tile and the edges
// (pair_lds_bytes), the per-wavefront rings (16 B per entry: (r2, types) and
// (QQ) charge products, or the prefilter's (home, neighbour, r2, species)),
// the (QQ) block's charge histogram, then (pf: the prefilter) the slow
// queues of near-edge entries.
size_t cells_lds_bytes(int rows, int lsmax, bool qq, bool pf) {
  const size_t base = (pair_lds_bytes(rows, lsmax) + 15) & ~size_t(15);
  // 16 B per entry; 20 B with the charge products of the per-pair path
  const size_t rings = static_cast<size_t>(kCellWaves) * kHitQ * (qq ? 20 : 16);
  const size_t slow = pf ? static_cast<size_t>(kCellWaves) * kSlowQ * 8 : 0;
  return base + rings + (qq ? (static_cast<size_t>(lsmax) * 8 + 15) & ~size_t(15) : 0) + slow;
}

void launch_rdf_cells(const PairArgs& a, const CellGrid& g, const int32_t* start, const double* spos,
                      const int32_t* stype, int64_t n, int waves_per_cell, int nblocks,
                      uint32_t* partial, hipStream_t s) {
  Stencil st{};
  int k = 0;
  st.off[k][0] = st.off[k][1] = st.off[k][2] = 0;
  ++k;
  const int zr = (a.ndim == 3) ? 1 : 0;
  for (int dz = -zr; dz <= zr; ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        const bool fwd = (dz > 0) || (dz == 0 && dy > 0) || (dz == 0 && dy == 0 && dx > 0);
        if (!fwd) continue;
        st.off[k][0] = dx;
        st.off[k][1] = dy;
        st.off[k][2] = dz;
        ++k;
      }
  st.n = k;
  const int64_t ncell = static_cast<int64_t>(g.nc[0]) * g.nc[1] * g.nc[2];
  bool shift = true;
  for (int k = 0; k < a.ndim; ++k) shift = shift && a.cellshift[k];
  // tasks: (home cell, chunk share); with reach 2, (x-row of cells, part)
  const bool row_tasks = shift && g.reach == 2;
  const int64_t ntasks = (row_tasks ? ncell / g.nc[0] : ncell) * waves_per_cell;
  const bool qq = a.q != nullptr;
  // FP32 prefilter bound (image-shift axes): |FP32 r2 - FP64 r2| < delta for
  // every pair with r2 <= side2 + delta.  Per axis, with U the largest
  // |coordinate relative to the home cell's corner| (2 cells; with reach 2,
  // 3 cells, and 5 along x, where a home chunk spans up to 3 cells and meets
  // 2 more on either side; plus L/4 + 1 cell on a non-periodic axis, whose
  // atoms may lie in [-L/4, 5L/4)) and
  // L the axis length: corner + shift and x - (corner + shift) in FP64
  // (< 2^-51 (2L + U) together), each coordinate rounded to FP32 (< 2^-24 U),
  // the FP32 difference (< 2^-24 (rc + 2e)); so e < 2^-23 (2U + rc) +
  // 2^-51 (2L + U), and sum_k (2 rc e_k + e_k^2) bounds the change of r2;
  // 2^-21 side2 covers the FP32 squares and sums of r2 itself (and the
  // FP32 rounding of the limit and of the cut).
  double delta = 0.0;
  float lim_pf = 0.f;
  if (shift) {
    const double rc = std::sqrt(a.side2) * (1.0 + std::ldexp(1.0, -20));
    double err = 0.0;
    for (int k = 0; k < a.ndim; ++k) {
      const double L = g.periodic[k] ? a.lmi[k] : a.lmi[k] / 2.5;
      const double cs = 1.0 / g.inv_h[k];
      const double reach_cells = row_tasks ? (k == 0 ? 5.0 : 3.0) : 2.0;
      const double U = reach_cells * cs + (g.periodic[k] ? 0.0 : 0.25 * L + cs);
      const double e = std::ldexp(2.0 * U + rc, -23) + std::ldexp(2.0 * L + U, -51);
      err += 2.0 * rc * e + e * e;
    }
    delta = err + std::ldexp(a.side2, -21);
    // the FP32 rounding of side2 + 2 delta stays above side2 + delta
    lim_pf = static_cast<float>(a.side2 + 2.0 * delta);
  }
  // Bin edges of the prefilter in FP32 registers: t_k = fl(fl(fl(k + off)^2)
  // * dr2), and an entry of FP32 r2 in [fl(t_k + m), fl(t_{k+1} - m)) is
  // binned to k without the FP64 rule.  Proven here for every k by
  // evaluating the same FP32 operations (IEEE, no contraction, as the
  // kernel does): fl(t_k + m) >= thr[k] + delta for k = 1..lsmax+1 and
  // fl(t_{k+1} - m) <= thr[k+1] - delta for k = 0..lsmax; with |FP32 r2 -
  // FP64 r2| < delta, such an entry's FP64 r2 lies in [thr[k], thr[k+1]).
  // m starts at the largest |t_k - thr[k]| plus 2 delta and doubles until
  // the check passes; without host edges, or after ten doublings, m = inf
  // (every entry takes the FP64 rule).
  float pf_margin = INFINITY;
  const float pf_off = a.floor_bins ? -1.0f : -0.5f;
  const float pf_dr2 = static_cast<float>(a.deltar * a.deltar);
  if (shift && a.thr_host) {
    const double* th = a.thr_host;
    auto edge = [&](int k) {
      const float x = static_cast<float>(k) + pf_off;
      const float x2 = x * x;
      return x2 * pf_dr2;
    };
    const double tol = delta * (1.0 + std::ldexp(1.0, -30));  // covers FP64 rounding of thr + delta
    double dev = 0.0;
    for (int k = 1; k <= a.lsmax + 1; ++k) dev = std::max(dev, std::fabs(static_cast<double>(edge(k)) - th[k]));
    float m = static_cast<float>((dev + 2.0 * delta) * (1.0 + std::ldexp(1.0, -20)));
    for (int tries = 0; tries <= 10; ++tries, m *= 2.0f) {
      bool ok = m > 0.0f;
      for (int k = 1; ok && k <= a.lsmax + 1; ++k) {
        const float lo = edge(k) + m;
        ok = static_cast<double>(lo) >= th[k] + tol;
      }
      for (int k = 0; ok && k <= a.lsmax; ++k) {
        const float hi = edge(k + 1) - m;
        ok = static_cast<double>(hi) <= th[k + 1] - tol;
      }
      if (ok) {
        pf_margin = m;
        break;
      }
    }
  }
  // One launch per histogram tile of rows_per_pass pair-type rows.
  const int rows = a.nrows > 0 ? std::min(a.nrows, a.nit) : a.nit;
  for (int row0 = 0; row0 < a.nit; row0 += rows) {
    PairArgs t = a;
    t.side2_pf = lim_pf;
    t.pf_delta = delta;
    t.pf_dr2 = pf_dr2;
    t.pf_off = pf_off;
    t.pf_est = a.floor_bins ? 1.0f : 0.5f;
    t.pf_margin = pf_margin;
    t.row0 = row0;
    t.nrows = std::min(rows, a.nit - row0);
    const size_t lds = cells_lds_bytes(t.nrows, t.lsmax, qq, shift);
#define TRJ_CELLS(S, D, Q, R)                                                                                     \
  do {                                                                                                            \
    if (lds > 65536)                                                                                              \
      (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_rdf_cells<S, D, Q, R>),                         \
                                hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));              \
    hipLaunchKernelGGL((k_rdf_cells<S, D, Q, R>), dim3(nblocks), dim3(64 * kCellWaves), lds, s, t, g, st, start,    \
                       spos, stype, n, waves_per_cell, ntasks, partial);                                      \
  } while (0)
    if (qq) {
      if (a.ndim == 3) {
        if (row_tasks) TRJ_CELLS(true, true, true, true);
        else if (shift) TRJ_CELLS(true, true, true, false);
        else TRJ_CELLS(false, true, true, false);
      } else {
        if (row_tasks) TRJ_CELLS(true, false, true, true);
        else if (shift) TRJ_CELLS(true, false, true, false);
        else TRJ_CELLS(false, false, true, false);
      }
    } else {
      if (a.ndim == 3) {
        if (row_tasks) TRJ_CELLS(true, true, false, true);
        else if (shift) TRJ_CELLS(true, true, false, false);
        else TRJ_CELLS(false, true, false, false);
      } else {
        if (row_tasks) TRJ_CELLS(true, false, false, true);
        else if (shift) TRJ_CELLS(true, false, false, false);
        else TRJ_CELLS(false, false, false, false);
      }
    }
#undef TRJ_CELLS
  }
}

void launch_rdf_allpairs(const PairArgs& a, const double* pos, const int32_t* itype, int64_t n,
                         int nblocks, uint32_t* partial, hipStream_t s, int64_t stride, const int32_t* range) {
  const int rows = a.nrows > 0 ? std::min(a.nrows, a.nit) : a.nit;
  for (int row0 = 0; row0 < a.nit; row0 += rows) {
    PairArgs t = a;
    t.row0 = row0;
    t.nrows = std::min(rows, a.nit - row0);
    const size_t lds = pair_lds_bytes(t.nrows, t.lsmax, t.q != nullptr);
    if (lds > 65536)
      (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_rdf_allpairs),
                                hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
    hipLaunchKernelGGL(k_rdf_allpairs, dim3(nblocks), dim3(256), lds, s, t, pos, itype, n, stride > 0 ? stride : n,
                       range, partial);
  }
}

void launch_slice_keys(const double* pos, int64_t n, const int32_t* itype, int nsp, const SliceSet& ss,
                       int32_t* key, unsigned long long* last, hipStream_t s) {
  hipLaunchKernelGGL(k_slice_keys, dim3(grid1d(n, 256)), dim3(256), 0, s, pos, n, itype, nsp, ss, key, last);
}

void launch_slice_counts(const int32_t* key, int64_t n, const unsigned long long* last, int nsp, int32_t* count,
                         int32_t* countp, const double* q, double* qsum, int32_t* qcnt, hipStream_t s) {
  hipLaunchKernelGGL(k_slice_counts, dim3(grid1d(n, 256)), dim3(256), 0, s, key, n, last, nsp, count, countp, q,
                     qsum, qcnt);
}

void launch_cell_count_slices(const double* pos, const int32_t* key, int64_t n, int nsp, const CellGrid& g,
                              int nslice, int32_t* cellid, int32_t* count, hipStream_t s) {
  hipLaunchKernelGGL(k_cell_count_slices, dim3(grid1d(n, 256)), dim3(256), 0, s, pos, key, n, nsp, g, nslice,
                     cellid, count);
}

void launch_fqt(const FqtArgs& a, hipStream_t s) {
  if (a.nvec <= 0) return;
  hipLaunchKernelGGL(k_fqt_partial, dim3(a.nvec, kFqtChunks, a.nact + 1), dim3(256), 0, s, a);
  hipLaunchKernelGGL(k_fqt_vectors, dim3((a.nvec + 63) / 64), dim3(64), 0, s, a);
  if (a.nact > 0) hipLaunchKernelGGL(k_fqt_bins, dim3((a.nqw * a.nact + 63) / 64), dim3(64), 0, s, a);
}

void launch_fqt_collect(int b, int nqw, int ndyn, double* fkt, double* fskt, double* tot, double* tots, hipStream_t s) {
  hipLaunchKernelGGL(k_fqt_collect, dim3((nqw * ndyn + 255) / 256), dim3(256), 0, s, b, nqw, ndyn, fkt, fskt, tot, tots);
}

size_t s2n_lds_bytes(int lsmax) { return sizeof(uint32_t) * (((lsmax + 1 + 3) & ~3) + 256); }

void launch_s2n(const S2nArgs& a, hipStream_t s) {
  if (a.np <= 0) return;
  hipLaunchKernelGGL(k_s2n, dim3(a.np), dim3(256), s2n_lds_bytes(a.lsmax), s, a);
}

void launch_xy_finish(int nslice, int nsp, int lsmax, int64_t* frame, const int32_t* count, const int32_t* countp,
                      int64_t* counts, double* hxy, unsigned long long* gq, const double* qp, const double* qsum,
                      const int32_t* qcnt, double* gqacc, hipStream_t s) {
  const int64_t total = static_cast<int64_t>(nslice) * (nsp * (nsp + 1) / 2) * lsmax;
  hipLaunchKernelGGL(k_xy_finish, dim3(grid1d(total, 256)), dim3(256), 0, s, nslice, nsp, lsmax, frame, count,
                     countp, counts, hxy, gq, qp, qsum, qcnt, gqacc);
}

void launch_qq_params(const double* q, int64_t n, double* qp, uint32_t* flags, hipStream_t s) {
  hipLaunchKernelGGL(k_qq_params, dim3(1), dim3(1024), 0, s, q, n, qp, flags);
}

void launch_qq_finish(unsigned long long* gq, int lsmax, const double* qp, double* acc,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_qq_finish, dim3(grid1d(lsmax, 256)), dim3(256), 0, s, gq, lsmax, qp, acc);
}

void launch_rdf_merge(const uint32_t* partial, int nslices, int nbins, int64_t* counts, hipStream_t s,
                      const QqWeights* qw, int lsmax, unsigned long long* gq) {
  const QqWeights w = qw ? *qw : QqWeights{};
  hipLaunchKernelGGL(k_rdf_merge, dim3(grid1d(nbins, 256), kMergeGroups), dim3(256), 0, s, partial, nslices,
                     nbins, counts, w, lsmax > 0 ? lsmax : 1, qw ? gq : nullptr);
}

// Grid of the contraction kernels: 8 * ceil(ntiles * ksplit / 8) workgroups
// (contract_work_index).
static dim3 contract_grid(const ContractArgs& a) {
  const int64_t nwork = static_cast<int64_t>(a.ntiles) * a.ksplit;
  return dim3(static_cast<unsigned>(8 * ((nwork + 7) / 8)));
}

void launch_sq_contract_f64(const ContractArgs& a, hipStream_t s) {
  if (a.ntiles <= 0) return;
  if (a.anc)
    hipLaunchKernelGGL(k_sq_contract_f64<true>, contract_grid(a), dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL(k_sq_contract_f64<false>, contract_grid(a), dim3(256), 0, s, a);
}

static AnchorAxes anchor_axes(const int rows[3], const int base[3]) {
  AnchorAxes ax;
  for (int k = 0; k < 3; ++k) {
    ax.rows[k] = rows[k];
    ax.base[k] = base[k];
    ax.row0[k] = k == 0 ? 0 : ax.row0[k - 1] + rows[k - 1];
  }
  return ax;
}

void launch_sq_anchors32(const double* u, int64_t n, const int rows[3], const int base[3], float2* anc,
                         hipStream_t s, const int32_t* live_off, const int32_t* live_nch) {
  const int nrow = rows[0] + rows[1] + rows[2];
  if (n <= 0 || nrow <= 0) return;
  hipLaunchKernelGGL(k_sq_anchors32, dim3(static_cast<unsigned>((n + 255) / 256), 3), dim3(256), 0, s, u, n,
                     anchor_axes(rows, base), nrow, anc, live_off, live_nch);
}

void launch_sq_anchors(const double* u, int64_t n, const int rows[3], const int base[3], double2* anc,
                       hipStream_t s, const int32_t* live_off, const int32_t* live_nch) {
  const int nrow = rows[0] + rows[1] + rows[2];
  if (n <= 0 || nrow <= 0) return;
  hipLaunchKernelGGL(k_sq_anchors, dim3(static_cast<unsigned>((n + 255) / 256), 3), dim3(256), 0, s, u, n,
                     anchor_axes(rows, base), anc, live_off, live_nch);
}

void launch_sq_contract_lp(const ContractArgs& a, int precision, hipStream_t s) {
  if (a.ntiles <= 0) return;
  const dim3 grid = contract_grid(a), block(256);
  if (precision == TRJ_PREC_FP32)
    hipLaunchKernelGGL(k_sq_contract_f32, grid, block, 0, s, a);
  else if (precision == TRJ_PREC_FP16X3)
    if (a.anc32)
      hipLaunchKernelGGL((k_sq_contract_lp16w<TRJ_PREC_FP16X3, true>), grid, dim3(512), 0, s, a);
    else
      hipLaunchKernelGGL((k_sq_contract_lp16w<TRJ_PREC_FP16X3, false>), grid, dim3(512), 0, s, a);
  else
    if (a.anc32)
      hipLaunchKernelGGL((k_sq_contract_lp16w<TRJ_PREC_BF16X3, true>), grid, dim3(512), 0, s, a);
    else
      hipLaunchKernelGGL((k_sq_contract_lp16w<TRJ_PREC_BF16X3, false>), grid, dim3(512), 0, s, a);
}

void launch_sq_shells(const void* rho, bool rho_f32, const int16_t* tile_shell, int ntiles, int tile_vec,
                      int ksplit, int nsp, int nslice, int nch, const double* bsc, int nqmax, const int32_t* norm,
                      const double* qsp, double* sqf, double* sqfp, double* sqfq, hipStream_t s) {
  if (ntiles <= 0) return;
  Bsc b{}, q{};
  for (int i = 0; i < nsp; ++i) b.b[i] = bsc[i];
  if (qsp)
    for (int i = 0; i < nsp; ++i) q.b[i] = qsp[i];
  const int nacc = nslice * (1 + nsp) + ((nch > nslice * nsp || qsp) ? nslice : 0);
  const size_t lds = static_cast<size_t>(nacc) * (nqmax + 1) * sizeof(double);
  const int use_lds = lds <= 48 * 1024;
  const int64_t total = static_cast<int64_t>(ntiles) * tile_vec;
  // Splits over four wavefronts when they are many (one thread per vector is
  // then bound by its chain of loads; with few, the barriers cost more) and
  // the channel sums fit in LDS beside the accumulators.
  const size_t lds4 = (use_lds ? ((lds + 15) & ~size_t(15)) : 0) + sizeof(double) * 2 * kWave * (8 + nch);
  if (ksplit >= 64 && lds4 <= 64 * 1024) {
    int64_t g4 = (total + kWave - 1) / kWave;
    if (g4 > 2048) g4 = 2048;
    if (rho_f32)
      hipLaunchKernelGGL(k_sq_shells4<float2>, dim3(static_cast<unsigned>(g4)), dim3(256), lds4, s,
                         static_cast<const float2*>(rho), tile_shell, ntiles, tile_vec, ksplit, nsp, nslice, nch, b,
                         nqmax, norm, q, qsp ? 1 : 0, sqf, sqfp, sqfq, use_lds);
    else
      hipLaunchKernelGGL(k_sq_shells4<double2>, dim3(static_cast<unsigned>(g4)), dim3(256), lds4, s,
                         static_cast<const double2*>(rho), tile_shell, ntiles, tile_vec, ksplit, nsp, nslice, nch, b,
                         nqmax, norm, q, qsp ? 1 : 0, sqf, sqfp, sqfq, use_lds);
    return;
  }
  int64_t g = (total + 255) / 256;
  if (g > 1024) g = 1024;
  if (rho_f32)
    hipLaunchKernelGGL(k_sq_shells<float2>, dim3(static_cast<unsigned>(g)), dim3(256), use_lds ? lds : 0, s,
                       static_cast<const float2*>(rho), tile_shell, ntiles, tile_vec, ksplit, nsp, nslice, nch, b,
                       nqmax, norm, q, qsp ? 1 : 0, sqf, sqfp, sqfq, use_lds);
  else
    hipLaunchKernelGGL(k_sq_shells<double2>, dim3(static_cast<unsigned>(g)), dim3(256), use_lds ? lds : 0, s,
                       static_cast<const double2*>(rho), tile_shell, ntiles, tile_vec, ksplit, nsp, nslice, nch, b,
                       nqmax, norm, q, qsp ? 1 : 0, sqf, sqfp, sqfq, use_lds);
}

size_t sq_list_scratch_bytes(int nvec, int nch) {
  return sizeof(double) * 2 * static_cast<size_t>(nvec) * kListBlocks * nch;
}

void launch_sq_list(const double* u, const double* w, const int32_t* sp_off, int64_t stride, int nsp, int nslice,
                    int nch, int ndim, const int32_t* kv, const int32_t* shell, int nvec, const double* bsc,
                    int nqmax, bool norm, const double* qsp, double* part, double* sqf, double* sqfp, double* sqfq,
                    hipStream_t s) {
  if (nvec <= 0) return;
  Bsc b{}, q{};
  for (int i = 0; i < nsp; ++i) b.b[i] = bsc[i];
  if (qsp)
    for (int i = 0; i < nsp; ++i) q.b[i] = qsp[i];
  hipLaunchKernelGGL(k_sq_list, dim3(nvec, kListBlocks), dim3(256), 0, s, u, w, sp_off, stride, nch, ndim, kv,
                     part);
  hipLaunchKernelGGL(k_sq_list_finish, dim3((nvec + 63) / 64), dim3(64), 0, s, part, shell, nvec, nsp, nslice,
                     nch, b, nqmax, norm ? sp_off : nullptr, q, qsp ? 1 : 0, sqf, sqfp, sqfq);
}

int64_t launch_lds_atomics(int nblocks, int iters, uint32_t* out, hipStream_t s) {
  hipLaunchKernelGGL(k_lds_atomics, dim3(nblocks), dim3(256), 0, s, iters, out);
  return static_cast<int64_t>(nblocks) * 256 * 16 * iters;
}

double launch_mfma_rate(int precision, int nblocks, int iters, float* out, unsigned long long* clk, hipStream_t s) {
  if (precision == TRJ_PREC_FP64) {
    hipLaunchKernelGGL(k_mfma_rate<TRJ_PREC_FP64>, dim3(nblocks), dim3(256), 0, s, iters, out, clk);
    return static_cast<double>(nblocks) * 4 * iters * 8 * (16.0 * 16 * 4 * 2);
  }
  if (precision == TRJ_PREC_FP16X3)
    hipLaunchKernelGGL(k_mfma_rate<TRJ_PREC_FP16X3>, dim3(nblocks), dim3(256), 0, s, iters, out, clk);
  else
    hipLaunchKernelGGL(k_mfma_rate<TRJ_PREC_BF16X3>, dim3(nblocks), dim3(256), 0, s, iters, out, clk);
  return static_cast<double>(nblocks) * 4 * iters * 8 * (16.0 * 16 * 32 * 2);
}

// Launches k_coissue<KV, mode> on `ncu` workgroups of 512 threads with
// `lds_bytes` of dynamic LDS; returns the MFMAs per SIMD (0 on bad arguments).
double launch_coissue(int kv, int mode, int ncu, int trips, size_t lds_bytes, float* out, unsigned long long* clk,
                      hipStream_t s) {
  typedef void (*Kern)(int, float*, unsigned long long*);
  static const Kern table[4][4] = {
      {k_coissue<1, 0>, k_coissue<1, 1>, k_coissue<1, 2>, k_coissue<1, 3>},
      {k_coissue<2, 0>, k_coissue<2, 1>, k_coissue<2, 2>, k_coissue<2, 3>},
      {k_coissue<3, 0>, k_coissue<3, 1>, k_coissue<3, 2>, k_coissue<3, 3>},
      {k_coissue<4, 0>, k_coissue<4, 1>, k_coissue<4, 2>, k_coissue<4, 3>}};
  if (kv < 1 || kv > 4 || mode < 0 || mode > 3 || trips < 2) return 0.0;
  const Kern k = table[kv - 1][mode];
  if (hipFuncSetAttribute(reinterpret_cast<const void*>(k), hipFuncAttributeMaxDynamicSharedMemorySize,
                          static_cast<int>(lds_bytes)) != hipSuccess)
    return 0.0;
  hipLaunchKernelGGL(k, dim3(ncu), dim3(512), lds_bytes, s, trips, out, clk);
  return 32.0 * trips;
}

void launch_batch_group(const BatchDev& d, int nblocks, size_t pair_lds, hipStream_t s) {
  hipLaunchKernelGGL(bk_zero, dim3(nblocks), dim3(256), 0, s, d);
  hipLaunchKernelGGL(bk_ingest, dim3(nblocks), dim3(256), 0, s, d);
  hipLaunchKernelGGL(bk_scan_reduce, dim3(nblocks), dim3(256), 0, s, d);
  hipLaunchKernelGGL(bk_scan_blocks, dim3(1), dim3(1024), 0, s, d, nblocks);
  hipLaunchKernelGGL(bk_scan_apply, dim3(nblocks), dim3(256), 0, s, d);
  hipLaunchKernelGGL(bk_scatter, dim3(nblocks), dim3(256), 0, s, d);
  hipLaunchKernelGGL(bk_pairs, dim3(nblocks), dim3(256), kRingBytes + pair_lds, s, d);
}

// ------------------------------------------------------- batched S(Q)
// The frame that holds count block b (frames in block order).
__device__ inline int bsq_frame_of(const BatchSqFrame* __restrict__ fr, int nframes, int b) {
  int lo = 0, hi = nframes - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (fr[mid].blk0 <= b) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Species of the count block's atoms: per-wavefront ballot counts in cnt,
// the lane's own species (-1: past the frame or outside 1..nsp) returned and
// its ballot mask in *mine.
__device__ inline int bsq_block_species(const BatchSqDev& d, const BatchSqFrame& F, int lb,
                                        int (*cnt)[kNspMax], unsigned long long* mine) {
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  const int li = lb * 256 + static_cast<int>(threadIdx.x);
  int s = -1;
  if (li < F.n) {
    const int t = d.itype_in[F.atom_off + li];
    s = (t >= 1 && t <= F.nsp) ? t - 1 : -1;
  }
  *mine = 0;
  for (int sp = 0; sp < F.nsp; ++sp) {
    const unsigned long long m = __ballot(s == sp);
    if (lane == 0) cnt[w][sp] = __popcll(m);
    if (s == sp) *mine = m;
  }
  return s;
}

// Block counts per species: blk[moff + s * nblk + lb].
__global__ __launch_bounds__(256) void bsq_count(BatchSqDev d) {
  __shared__ int cnt[4][kNspMax];
  const int nframes = d.hdr->nframes, nblocks = d.hdr->nblocks;
  for (int b = blockIdx.x; b < nblocks; b += gridDim.x) {
    const BatchSqFrame& F = d.fr[bsq_frame_of(d.fr, nframes, b)];
    const int lb = b - F.blk0;
    unsigned long long mine;
    (void)bsq_block_species(d, F, lb, cnt, &mine);
    __syncthreads();
    for (int sp = threadIdx.x; sp < F.nsp; sp += 256)
      d.blk[F.moff + sp * F.nblk + lb] = cnt[0][sp] + cnt[1][sp] + cnt[2][sp] + cnt[3][sp];
    __syncthreads();
  }
}

// Channel offsets: channel chan0 + s of frame f starts at the scanned count
// of its first block; sp_off[nch] is the number of entries.
__global__ __launch_bounds__(256) void bsq_offsets(BatchSqDev d) {
  const int nframes = d.hdr->nframes;
  for (int f = threadIdx.x; f < nframes; f += 256) {
    const BatchSqFrame& F = d.fr[f];
    for (int s = 0; s < F.nsp; ++s) d.sp_off[F.chan0 + s] = d.blkoff[F.moff + s * F.nblk];
  }
  if (threadIdx.x == 0) {
    const int last = d.hdr->ncounts - 1;
    d.sp_off[d.hdr->nch] = d.blkoff[last] + d.blk[last];
  }
}

// Each atom's turns into its sorted slot: positions converted to FP64 and
// wrapped into [0, L) on every axis of its frame (as k_ingest wraps the
// periodic axes of an S(Q) frame), then hi, lo, cos, sin per axis as
// k_species_scatter writes them with L the frame's box.
__global__ __launch_bounds__(256) void bsq_scatter(BatchSqDev d) {
  __shared__ int cnt[4][kNspMax];
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  const int nframes = d.hdr->nframes, nblocks = d.hdr->nblocks;
  for (int b = blockIdx.x; b < nblocks; b += gridDim.x) {
    const BatchSqFrame& F = d.fr[bsq_frame_of(d.fr, nframes, b)];
    const int lb = b - F.blk0;
    const int li = lb * 256 + static_cast<int>(threadIdx.x);
    unsigned long long mine;
    const int s = bsq_block_species(d, F, lb, cnt, &mine);
    __syncthreads();
    if (s >= 0) {
      int off = d.blkoff[F.moff + s * F.nblk + lb];
      for (int ww = 0; ww < w; ++ww) off += cnt[ww][s];
      off += __popcll(mine & ((1ull << lane) - 1ull));
      const int64_t ai = F.atom_off + li;
      uint32_t bad = 0;
      for (int k = 0; k < 3; ++k) {
        double x = 0.0, L = 1.0;
        if (k < d.ndim) {
          const int64_t idx = ai * d.ndim + k;
          x = d.dtype == TRJ_F32 ? static_cast<double>(static_cast<const float*>(d.pos_in)[idx])
                                 : static_cast<const double*>(d.pos_in)[idx];
          L = F.box[k];
          if (!isfinite(x)) bad |= kFlagRange;
          else x = wrap_periodic(x, L);
        }
        const double hi = x / L;
        const double lo = fma(-hi, L, x) / L;
        double sn, cs;
        sincospi(2.0 * frac_turns(1, hi, lo), &sn, &cs);
        d.u[k * d.stride + off] = hi;
        d.u[(3 + k) * d.stride + off] = lo;
        d.u[(6 + k) * d.stride + off] = cs;
        d.u[(9 + k) * d.stride + off] = sn;
      }
      if (bad) atomicOr(d.flags, bad);
    } else if (li < F.n) {
      atomicOr(d.flags, kFlagType);
    }
    __syncthreads();
  }
}

// The frame whose shell-kernel units hold unit u (frames in unit order).
__device__ inline int bsq_frame_of_unit(const BatchSqFrame* __restrict__ fr, int nframes, int u) {
  int lo = 0, hi = nframes - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (fr[mid].unit0 <= u) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Shells of every frame: a unit is vpu consecutive tile vectors of one frame
// (vpu set by the host so that every unit reads about the same number of
// split partials); for each vector the frame's splits are summed per
// channel, and |sum_s b_s rho_s|^2, |rho_s|^2 and a count of 1 go to the
// vector's shell nint(|q| / dq) of the frame's box, the FP64 rule of the
// half-space plan (shells 1 .. nqmax, the origin left out).  Sums gather in
// LDS per unit, then go to the outputs with FP64 / integer atomics.
__global__ __launch_bounds__(256) void bsq_shells(BatchSqDev d) {
  extern __shared__ __align__(16) unsigned char smem[];
  double* acc = reinterpret_cast<double*>(smem);
  const int nframes = d.hdr->nframes, nunits = d.hdr->nunits;
  const int tvec = d.tx * d.ty * d.tz;
  const int64_t total = static_cast<int64_t>(d.ntiles) * tvec;
  const double2* r2 = reinterpret_cast<const double2*>(d.rho);
  const float2* r2f = reinterpret_cast<const float2*>(d.rho);
  for (int unit = blockIdx.x; unit < nunits; unit += gridDim.x) {
    const BatchSqFrame& F = d.fr[bsq_frame_of_unit(d.fr, nframes, unit)];
    const int part = unit - F.unit0;
    const int64_t chunk = F.vpu;
    const int ns = F.nqmax + 1, nsp = F.nsp;
    const int nacc = (2 + nsp) * ns;  // sqf, sqfp [nsp], counts
    const bool lds = nacc <= d.shell_lds;
    if (lds) {
      for (int k = threadIdx.x; k < nacc; k += 256) acc[k] = 0.0;
      __syncthreads();
    }
    auto add = [&](int row, int sh, double v) {
      if (lds) {
        atomicAdd(&acc[row * ns + sh], v);
      } else if (row == 0) {
        atomicAdd(&d.sqf[F.out_off + sh], v);
      } else if (row <= nsp) {
        atomicAdd(&d.sqfp[F.outp_off + static_cast<int64_t>(row - 1) * ns + sh], v);
      } else {
        atomicAdd(&d.nq[F.out_off + sh], static_cast<unsigned long long>(v));
      }
    };
    const int64_t e1 = (part + 1) * chunk < total ? (part + 1) * chunk : total;
    for (int64_t e = part * chunk + threadIdx.x; e < e1; e += 256) {
      const int tile = static_cast<int>(e / tvec), v = static_cast<int>(e % tvec);
      const int a = d.tiles[4 * tile] + v / (d.ty * d.tz);
      const int bq = d.tiles[4 * tile + 1] + (v / d.tz) % d.ty;
      const int c = d.tiles[4 * tile + 2] + v % d.tz;
      if (a > d.kmax[0] || bq > d.kmax[1] || c > d.kmax[2] || (a == 0 && bq == 0 && c == 0)) continue;
      const double qa = a * F.fk[0], qb = bq * F.fk[1], qc = c * F.fk[2];
      const double sh = round(sqrt(qa * qa + qb * qb + qc * qc) / F.dq);
      if (!(sh >= 1.0 && sh <= static_cast<double>(F.nqmax))) continue;
      const int ish = static_cast<int>(sh);
      double tr = 0.0, ti = 0.0;
      for (int s = 0; s < nsp; ++s) {
        CSum sr, si;
        const int64_t first = static_cast<int64_t>(F.plane0 + s) * total + e;
        const int64_t step = static_cast<int64_t>(nsp) * total;
        if (d.rho_f32) {
          const float2* src = r2f + first;
#pragma unroll 8
          for (int k = 0; k < F.ks; ++k) {
            const float2 x = src[k * step];
            sr.add(static_cast<double>(x.x));
            si.add(static_cast<double>(x.y));
          }
        } else {
          const double2* src = r2 + first;
#pragma unroll 8
          for (int k = 0; k < F.ks; ++k) {
            const double2 x = src[k * step];
            sr.add(x.x);
            si.add(x.y);
          }
        }
        const double rr = sr.value(), ri = si.value();
        add(1 + s, ish, rr * rr + ri * ri);
        tr += F.bsc[s] * rr;
        ti += F.bsc[s] * ri;
      }
      add(0, ish, tr * tr + ti * ti);
      add(1 + nsp, ish, 1.0);
    }
    if (lds) {
      __syncthreads();
      for (int k = threadIdx.x; k < nacc; k += 256) {
        const double x = acc[k];
        if (x == 0.0) continue;
        const int row = k / ns, sh = k % ns;
        if (row == 0)
          atomicAdd(&d.sqf[F.out_off + sh], x);
        else if (row <= nsp)
          atomicAdd(&d.sqfp[F.outp_off + static_cast<int64_t>(row - 1) * ns + sh], x);
        else
          atomicAdd(&d.nq[F.out_off + sh], static_cast<unsigned long long>(x));
      }
      __syncthreads();
    }
  }
}

void launch_batch_sq_group(const BatchSqDev& d, int precision, hipStream_t s) {
  hipLaunchKernelGGL(bsq_count, dim3(d.nblocks), dim3(256), 0, s, d);
  launch_exclusive_scan_dev(d.blk, d.blkoff, &d.hdr->ncounts, s);
  hipLaunchKernelGGL(bsq_offsets, dim3(1), dim3(256), 0, s, d);
  hipLaunchKernelGGL(bsq_scatter, dim3(d.nblocks), dim3(256), 0, s, d);
  ContractArgs ca{};
  ca.u = d.u;
  ca.w = nullptr;
  ca.sp_off = d.sp_off;
  ca.tiles = d.tiles;
  ca.ntiles = d.ntiles;
  ca.ksplit = d.cap_items;
  ca.nsp = kNspMax;
  ca.ndim = d.ndim;
  ca.nch = 0;
  ca.n = d.stride;
  ca.rho = d.rho;
  ca.items = d.items;
  ca.item_plane = d.item_plane;
  ca.nitems = &d.hdr->nitems;
  if (d.anc) {
    // The group's entries are sp_off[0 .. nch): anchors for those only.
    if (precision == TRJ_PREC_FP64) {
      launch_sq_anchors(d.u, d.stride, d.anc_rows, d.anc_base, static_cast<double2*>(d.anc), s, d.sp_off,
                        &d.hdr->nch);
      ca.anc = static_cast<const double2*>(d.anc);
    } else {
      launch_sq_anchors32(d.u, d.stride, d.anc_rows, d.anc_base, static_cast<float2*>(d.anc), s, d.sp_off,
                          &d.hdr->nch);
      ca.anc32 = static_cast<const float2*>(d.anc);
      ca.anc_nrows = d.anc_rows[0] + d.anc_rows[1] + d.anc_rows[2];
    }
    ca.anc_row[0] = 0;
    ca.anc_row[1] = d.anc_rows[0];
    ca.anc_row[2] = d.anc_rows[0] + d.anc_rows[1];
    for (int k = 0; k < 3; ++k) ca.anc_base[k] = d.anc_base[k];
  }
  if (precision == TRJ_PREC_FP64)
    launch_sq_contract_f64(ca, s);
  else
    launch_sq_contract_lp(ca, precision, s);
  hipLaunchKernelGGL(bsq_shells, dim3(d.nblocks), dim3(256), static_cast<size_t>(d.shell_lds) * sizeof(double), s,
                     d);
}

}  // namespace trj
