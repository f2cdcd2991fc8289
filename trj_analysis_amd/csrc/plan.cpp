// plan.cpp -- host-side planning of the trj_hip library: RDF bin edges and
// minimum-image switch points (exact in FP64), the S(Q) q-set, its shells and
// normalisation counts, and the tiling of the dense q-box for the contraction.
//
// Compiled with -ffp-contract=off: the FP32 arithmetic below must round once
// per operation, as the reference's FP32 sq_init does (myprec = sp,
// src/core/precision.f90:22).
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <limits>

#include "trj_internal.h"

namespace trj {

namespace {

// Fortran NINT: round half away from zero.
inline long nint_d(double x) { return std::lround(x); }
inline int nint_f(float x) { return static_cast<int>(std::lroundf(x)); }

// Ordered bit pattern of a non-negative double (monotone in the value).
inline uint64_t dbits(double v) {
  uint64_t u;
  std::memcpy(&u, &v, 8);
  return u;
}
inline double dval(uint64_t u) {
  double v;
  std::memcpy(&v, &u, 8);
  return v;
}

// Smallest non-negative double v with pred(v) true, pred monotone (false..true).
template <class F>
double first_true(F pred, double hi) {
  uint64_t lo = 0, h = dbits(hi);  // pred(hi) must be true
  while (lo < h) {
    uint64_t mid = lo + (h - lo) / 2;
    if (pred(dval(mid)))
      h = mid;
    else
      lo = mid + 1;
  }
  return dval(lo);
}

}  // namespace

// rdf.cuf:55-58, with rcrdf = min(rcrdf, side/2) applied first
// (netcdf.f90:610-612 runs before RDF_init on the first frame).
int rdf_lsmax(int ndim, const double box[3], double rcrdf, double deltar, double* rc_eff) {
  double side = box[0];
  for (int k = 1; k < ndim; ++k) side = std::fmin(side, box[k]);
  double rc = std::fmin(rcrdf, side / 2.0);
  if (rc_eff) *rc_eff = rc;
  if (rc > 0.0) return static_cast<int>(nint_d(rc / deltar)) + 1;
  return static_cast<int>(nint_d(side / 2.0 / deltar)) + 1;
}

// The bin rule k = nint(sqrt(r2)/deltar) (rdf.cuf:461-462) is a monotone step
// function of r2 because sqrt, division by deltar > 0 and nint are each
// monotone under round-to-nearest.  Its steps are located exactly here, so the
// kernels bin by comparing r2 with these edges and never evaluate sqrt or a
// division, yet agree bit for bit with the FP64 rule.
std::vector<double> rdf_bin_thresholds(double deltar, int lsmax) {
  std::vector<double> thr(static_cast<size_t>(lsmax) + 2, 0.0);
  for (int m = 1; m <= lsmax + 1; ++m) {
    double hi = (m + 1.0) * deltar;
    hi = hi * hi * 4.0 + 1.0;
    thr[m] = first_true([&](double r2) { return nint_d(std::sqrt(r2) / deltar) >= m; }, hi);
  }
  return thr;
}

// Edges of rdfxy's rule k = int(r/deltar) + 1 (rdf.cuf:718), FP64 here:
// thr[m] = smallest r^2 with int(sqrt(r^2)/deltar) + 1 >= m.
std::vector<double> rdf_bin_thresholds_floor(double deltar, int lsmax) {
  std::vector<double> thr(static_cast<size_t>(lsmax) + 2, 0.0);
  for (int m = 2; m <= lsmax + 1; ++m) {
    double hi = (m + 1.0) * deltar;
    hi = hi * hi * 4.0 + 1.0;
    thr[m] = first_true([&](double r2) { return static_cast<int64_t>(std::sqrt(r2) / deltar) + 1 >= m; }, hi);
  }
  return thr;
}

// d - L*nint(d/L) (rdf.cuf:455-457) switches image where nint(d/L) reaches 1;
// by symmetry of division and of nint, at -t on the negative side.
double minimage_threshold(double L) {
  return first_true([&](double d) { return nint_d(d / L) >= 1; }, 2.0 * L);
}

FrameGeom frame_geometry(int ndim, const double box[3], const int periodic[3], double* rc_sticky) {
  FrameGeom g{};
  double side = box[0];
  for (int k = 1; k < ndim; ++k) side = std::fmin(side, box[k]);
  for (int k = 0; k < 3; ++k) {
    g.box[k] = box[k];
    g.lmi[k] = (k < ndim && !periodic[k]) ? 2.5 * box[k] : box[k];
    g.half[k] = (k < ndim) ? minimage_threshold(g.lmi[k]) : std::numeric_limits<double>::infinity();
  }
  double half = side / 2.0;
  if (half < *rc_sticky) *rc_sticky = half;
  g.side2 = (*rc_sticky > 0.0) ? (*rc_sticky) * (*rc_sticky) : half * half;
  double v = 1.0;
  for (int k = 0; k < ndim; ++k) v *= box[k];
  g.volume = v;
  return g;
}

// ---------------------------------------------------------------------------
// S(Q)

namespace {

// FP32 pi of the reference (common.cuf:43).
const float kPiF = 3.1415926535898f;

// sq.cuf:176-178 / 184-185: l <- l - m + mod(l, m) when mod(l, m) /= 0.
inline int lattice_adjust(int l, int m) {
  int r = l % m;
  return r != 0 ? l - m + r : l;
}

// sq.cuf:537-555: shell of octant vector k in the sqfnD set or -1.
inline int ref_octant_shell(const SqPlan& P, int ndim, int kx, int ky, int kz) {
  float a = static_cast<float>(kx) * P.fk[0];
  float b = static_cast<float>(ky) * P.fk[1];
  float q2 = a * a + b * b;
  if (ndim == 3) {
    float c = static_cast<float>(kz) * P.fk[2];
    q2 = q2 + c * c;
  }
  if (!(q2 > 0.0f) || !(q2 <= P.qmin2)) return -1;
  float q = std::sqrt(q2);
  int ind = static_cast<int>(q / P.dq + 0.501f);
  return ind <= P.nqmax ? ind : -1;
}

}  // namespace

void anchor_layout(const int32_t* org, int stride, size_t ntiles, const TileShape& tk, int rows[3], int base[3]) {
  // Segment origins of a tile: x kx0 + 8 s (s < tx / 8), y ky0 + 8 s
  // (s < ty / 8), z kz0 + 8 s (s < tz / 8).
  const int seg[3] = {tk.tx / 8, tk.ty / 8, tk.tz / 8};
  for (int k = 0; k < 3; ++k) {
    rows[k] = 0;
    base[k] = 0;
  }
  if (ntiles == 0) return;
  int lo[3], hi[3];
  for (int k = 0; k < 3; ++k) {
    lo[k] = INT32_MAX;
    hi[k] = INT32_MIN;
  }
  for (size_t t = 0; t < ntiles; ++t)
    for (int k = 0; k < 3; ++k) {
      const int o = org[t * stride + k];
      lo[k] = std::min(lo[k], o);
      hi[k] = std::max(hi[k], o + 8 * (seg[k] - 1));
    }
  for (int k = 0; k < 3; ++k) {
    base[k] = lo[k];
    rows[k] = (hi[k] - lo[k]) / 8 + 1;
  }
}

int sq_plan(const trj_sq_params& p_in, const double box[3], SqPlan* out, std::string* err) {
  SqPlan& P = *out;
  P = SqPlan();
  P.p = p_in;
  trj_sq_params& p = P.p;
  if (p.ndim != 2 && p.ndim != 3) {
    *err = "sq: ndim must be 2 or 3";
    return TRJ_EINVAL;
  }
  if (p.nsp < 1 || p.nsp > kNspMax) {
    *err = "sq: nsp must be in 1..10";
    return TRJ_EINVAL;
  }
  for (int k = 0; k < p.ndim; ++k)
    if (!(box[k] > 0.0)) {
      *err = "sq: box lengths must be positive";
      return TRJ_EINVAL;
    }
  if (!(p.qmax > 0.0)) {
    *err = "sq: qmax must be positive";
    return TRJ_EINVAL;
  }
  if (!(p.qmin > 0.0)) p.qmin = p.qmax;  // the reference has no default (input.f90:56)
  const int ndim = p.ndim;
  for (int k = 0; k < 3; ++k) P.box[k] = (k < ndim) ? box[k] : 1.0;

  if (p.qset == TRJ_QSET_REF) {
    // sq.cuf:156-200 in FP32.
    float side[3] = {static_cast<float>(P.box[0]), static_cast<float>(P.box[1]),
                     static_cast<float>(P.box[2])};
    for (int k = 0; k < 3; ++k) P.fk[k] = (2.0f * kPiF) / side[k];
    P.dq = std::fmin(P.fk[0], P.fk[1]);
    if (ndim == 3) P.dq = std::fmin(P.dq, P.fk[2]);
    const float nqf = static_cast<float>(p.qmax) / P.dq;
    if (!(nqf < 32767.0f)) {
      *err = "sq: qmax/dq above 32767 shells is not supported";
      return TRJ_ELIMIT;
    }
    P.nqmax = nint_f(nqf);
    float qminf = static_cast<float>(p.qmin);
    P.l[0] = nint_f(qminf / P.fk[0]);
    P.l[1] = nint_f(qminf / P.fk[1]);
    P.l[2] = (ndim == 3) ? nint_f(qminf / P.fk[2]) : 0;
    // sq.cuf:165-166 read qedge before it is set; treated as no-ops.
    const int m = (ndim == 3) ? 4 : 16;
    for (int k = 0; k < ndim; ++k) P.l[k] = lattice_adjust(P.l[k], m);
    for (int k = 0; k < 3; ++k) P.kext[k] = (k < ndim) ? m * (P.l[k] / m + 1) : 1;
    P.nqmin = std::min(P.l[0], P.l[1]);
    if (ndim == 3) P.nqmin = std::min(P.nqmin, P.l[2]);
    float qm = P.fk[0] * static_cast<float>(P.l[0]);
    qm = std::fmin(qm, P.fk[1] * static_cast<float>(P.l[1]));
    if (ndim == 3) qm = std::fmin(qm, P.fk[2] * static_cast<float>(P.l[2]));
    P.qminp = qm;
    if (P.nqmin == 0) {  // sq.cuf:195-200
      P.nqmin = 1;
      P.l[0] = P.l[1] = P.l[2] = 1;
    }
    P.qmin2 = P.qminp * P.qminp;

    // Host shell counts (sq.cuf:303-343).
    P.nq.assign(static_cast<size_t>(P.nqmax) + 1, 0);
    const int fill = (ndim == 3) ? 3 : 2;
    if (P.nqmin == 1) {
      for (int s = 1; s <= P.nqmax; ++s) P.nq[s] = fill;
    } else {
      int indm = 0;
      const int ix = (ndim == 3) ? P.l[0] - 1 : P.l[0];  // 2D loops include lmax
      const int iy = (ndim == 3) ? P.l[1] - 1 : P.l[1];
      const int iz = (ndim == 3) ? P.l[2] - 1 : 0;
      for (int i = 0; i <= ix; ++i)
        for (int j = 0; j <= iy; ++j)
          for (int k = 0; k <= iz; ++k) {
            if (i + j + k == 0) continue;
            float a = static_cast<float>(i) * P.fk[0];
            float b = static_cast<float>(j) * P.fk[1];
            float s2 = a * a + b * b;
            if (ndim == 3) {
              float c = static_cast<float>(k) * P.fk[2];
              s2 = s2 + c * c;
            }
            float rr = std::sqrt(s2);
            if (rr < P.qminp) {
              int ind = nint_f(rr / P.dq);
              indm = std::max(indm, ind);
              if (ind <= P.nqmax) P.nq[ind] += 1;
            }
          }
      for (int s = indm + 1; s <= P.nqmax; ++s) P.nq[s] = fill;
    }

    // Dense octant set (sqfnD), only launched when nqmin > 1 (sq.cuf:73).
    P.nq_summed.assign(static_cast<size_t>(P.nqmax) + 1, 0);
    for (int k = 0; k < 3; ++k) {
      P.kmin[k] = 0;
      P.kmax[k] = -1;
    }
    if (P.nqmin > 1) {
      for (int a = 0; a < P.kext[0]; ++a)
        for (int b = 0; b < P.kext[1]; ++b)
          for (int c = 0; c < P.kext[2]; ++c)
            if (ref_octant_shell(P, ndim, a, b, c) >= 0) {
              P.kmax[0] = std::max(P.kmax[0], a);
              P.kmax[1] = std::max(P.kmax[1], b);
              P.kmax[2] = std::max(P.kmax[2], c);
            }
    }
    // Axis directions (sqfact / sqfact2); when nqmin <= 1 the reference's
    // idir = 0 form applies from k = 1 (sq.cuf:110-113).
    int k0 = (P.nqmin > 1) ? P.nqmin + 1 : 1;
    for (int k = k0; k <= P.nqmax; ++k) {
      const int dirs = (ndim == 3) ? 3 : 2;
      for (int d = 0; d < dirs; ++d) {
        int v[3] = {0, 0, 0};
        v[d] = k;
        P.axis_kv.insert(P.axis_kv.end(), v, v + 3);
        P.axis_shell.push_back(k);
        P.nq_summed[k] += 1;
      }
    }
  } else if (p.qset == TRJ_QSET_HALFSPACE) {
    for (int k = 0; k < ndim; ++k)
      if (p.grid[k] < 2 || (p.grid[k] % 2) != 0) {
        *err = "sq: half-space grid extents must be even and >= 2";
        return TRJ_EINVAL;
      }
    double fk64[3];
    for (int k = 0; k < 3; ++k) {
      fk64[k] = 2.0 * M_PI / P.box[k];
      P.fk[k] = static_cast<float>(fk64[k]);
    }
    double dq64 = std::fmin(fk64[0], fk64[1]);
    if (ndim == 3) dq64 = std::fmin(dq64, fk64[2]);
    P.dq = static_cast<float>(dq64);
    P.kmin[0] = -p.grid[0] / 2;
    P.kmax[0] = p.grid[0] / 2 - 1;
    P.kmin[1] = -p.grid[1] / 2;
    P.kmax[1] = p.grid[1] / 2 - 1;
    P.kmin[2] = 0;
    P.kmax[2] = (ndim == 3) ? p.grid[2] / 2 - 1 : 0;
    int maxshell = 0;
    for (int a = P.kmin[0]; a <= P.kmax[0]; ++a)
      for (int b = P.kmin[1]; b <= P.kmax[1]; ++b)
        for (int c = P.kmin[2]; c <= P.kmax[2]; ++c) {
          double qa = a * fk64[0], qb = b * fk64[1], qc = c * fk64[2];
          maxshell = std::max(maxshell, static_cast<int>(nint_d(std::sqrt(qa * qa + qb * qb + qc * qc) / dq64)));
        }
    const double qcap = std::floor(p.qmax / dq64 + 0.5);  // nint without int overflow
    P.nqmax = (qcap < static_cast<double>(maxshell)) ? static_cast<int>(qcap) : maxshell;
    P.nqmin = 0;
    P.qminp = 0.0f;
    P.qmin2 = 0.0f;
    P.nq_summed.assign(static_cast<size_t>(P.nqmax) + 1, 0);
  } else {
    *err = "sq: unknown q-set";
    return TRJ_EINVAL;
  }

  // Tile the dense box; keep tiles with at least one summed vector.
  const bool ref = (p.qset == TRJ_QSET_REF);
  double fk64[3] = {2.0 * M_PI / P.box[0], 2.0 * M_PI / P.box[1], 2.0 * M_PI / P.box[2]};
  double dq64 = P.dq;
  if (!ref) {
    dq64 = std::fmin(fk64[0], fk64[1]);
    if (ndim == 3) dq64 = std::fmin(dq64, fk64[2]);
  }
  if (p.precision < TRJ_PREC_FP64 || p.precision > TRJ_PREC_BF16X3) {
    *err = "sq: unknown precision";
    return TRJ_EINVAL;
  }
  P.tk = tile_shape(p.precision);
  const int TKX = P.tk.tx, TKY = P.tk.ty, TKZ = P.tk.tz;
  if (P.kmax[0] >= P.kmin[0] && P.kmax[1] >= P.kmin[1] && P.kmax[2] >= P.kmin[2]) {
    std::vector<int16_t> sh(P.tk.vec());
    for (int x0 = P.kmin[0]; x0 <= P.kmax[0]; x0 += TKX)
      for (int y0 = P.kmin[1]; y0 <= P.kmax[1]; y0 += TKY)
        for (int z0 = P.kmin[2]; z0 <= P.kmax[2]; z0 += TKZ) {
          int count = 0;
          uint32_t blocks = 0;
          for (int i = 0; i < TKX; ++i)
            for (int j = 0; j < TKY; ++j)
              for (int k = 0; k < TKZ; ++k) {
                int a = x0 + i, b = y0 + j, c = z0 + k;
                int s = -1;
                if (a <= P.kmax[0] && b <= P.kmax[1] && c <= P.kmax[2]) {
                  if (ref) {
                    s = ref_octant_shell(P, ndim, a, b, c);
                  } else if (a != 0 || b != 0 || c != 0) {
                    double qa = a * fk64[0], qb = b * fk64[1], qc = c * fk64[2];
                    long ss = nint_d(std::sqrt(qa * qa + qb * qb + qc * qc) / dq64);
                    s = (ss >= 1 && ss <= P.nqmax) ? static_cast<int>(ss) : -1;
                  }
                }
                // Vector order within a tile: kz fastest, then ky, then kx.
                sh[(i * TKY + j) * TKZ + k] = static_cast<int16_t>(s);
                if (s >= 0) {
                  ++count;
                  P.nq_summed[s] += 1;
                  const int bit = ((i * TKY + j) / 16) * (TKZ / 16) + k / 16;
                  if (bit < 32) blocks |= 1u << bit;  // used by the FP64 tiles (16 blocks)
                }
              }
          if (count > 0) {
            P.tiles.push_back(SqTile{x0, y0, z0, static_cast<int32_t>(blocks)});
            P.tile_shell.insert(P.tile_shell.end(), sh.begin(), sh.end());
          }
        }
  }
  if (p.precision != TRJ_PREC_FP32 && !P.tiles.empty()) {
    std::vector<int32_t> org;
    org.reserve(3 * P.tiles.size());
    for (const SqTile& t : P.tiles) {
      org.push_back(t.kx0);
      org.push_back(t.ky0);
      org.push_back(t.kz0);
    }
    anchor_layout(org.data(), 3, P.tiles.size(), P.tk, P.anc_rows, P.anc_base);
  }
  if (P.nqmax > 32767) {
    *err = "sq: nqmax above 32767 is not supported";
    return TRJ_ELIMIT;
  }
  if (!ref || (p.flags & TRJ_SQ_NQ_CONSISTENT)) P.nq = P.nq_summed;

  trj_sq_info& I = P.info;
  I.dq = P.dq;
  for (int k = 0; k < 3; ++k) I.fk[k] = P.fk[k];
  I.qmin_eff = P.qminp;
  I.nqmax = P.nqmax;
  I.nqmin = P.nqmin;
  for (int k = 0; k < 3; ++k) I.lmax[k] = P.l[k];
  int64_t ng = 0;
  for (size_t t = 0; t < P.tile_shell.size(); ++t) ng += (P.tile_shell[t] >= 0);
  I.nvec_grid = ng;
  I.nvec_axis = static_cast<int64_t>(P.axis_shell.size());
  for (int k = 0; k < 3; ++k) I.box_ext[k] = P.kmax[k] >= P.kmin[k] ? P.kmax[k] - P.kmin[k] + 1 : 0;
  return TRJ_OK;
}

}  // namespace trj

// ---------------------------------------------------------------------------
// Host-only entry points (no device needed).

extern "C" int trj_sq_plan(const trj_sq_params* p, const double box[3], trj_sq_info* info,
                           int64_t* nq) {
  if (!p || !box || !info) return TRJ_EINVAL;
  trj::SqPlan P;
  std::string err;
  int rc = trj::sq_plan(*p, box, &P, &err);
  if (rc != TRJ_OK) return rc;
  *info = P.info;
  if (nq)
    for (int s = 0; s <= P.nqmax; ++s) nq[s] = P.nq[s];
  return TRJ_OK;
}

extern "C" int trj_rdf_plan(const trj_rdf_params* p, const double box[3], int32_t* lsmax,
                            double* rc_eff) {
  if (!p || !box) return TRJ_EINVAL;
  if (p->ndim != 2 && p->ndim != 3) return TRJ_EINVAL;
  if (!(p->deltar > 0.0)) return TRJ_EINVAL;
  double rc = 0.0;
  int ls = trj::rdf_lsmax(p->ndim, box, p->rcrdf, p->deltar, &rc);
  if (lsmax) *lsmax = ls;
  if (rc_eff) *rc_eff = rc;
  return TRJ_OK;
}
