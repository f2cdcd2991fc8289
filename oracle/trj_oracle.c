/*
 * trj_oracle.c -- CPU restatement of the trj_analysis structural hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see trj_oracle.h).  Built with -ffp-contract=off
 * so that every FP64/FP32 operation below rounds exactly once, in the order
 * written.  Reference citations are file:line in the reference repository.
 */
#define _GNU_SOURCE /* sincos */
#include "trj_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* The reference's pi is the FP32 rounding of 3.1415926535898D0
 * (src/core/common.cuf:43, myprec = sp). */
static const float OR_PI_F = 3.1415926535898f;

long or_nint_d(double x) { return (long)round(x); }
int or_nint_f(float x) { return (int)roundf(x); }

/* rdf.cuf:741-753 */
int or_ind_ij(int i, int j, int nsp) {
  if (i <= j) return (i - 1) * (nsp - 1) - (i - 1) * (i - 2) / 2 + j;
  return (j - 1) * (nsp - 1) - (j - 1) * (j - 2) / 2 + i;
}

/* netcdf.f90:602-622 */
void or_geometry(int ndim, const double box[3], const int32_t periodic[3],
                 double* rcrdf_inout, double lmi[3], double* side2, double* volume) {
  double side = box[0];
  for (int k = 1; k < ndim; ++k)
    if (box[k] < side) side = box[k];
  for (int k = 0; k < 3; ++k) lmi[k] = box[k];
  for (int k = 0; k < ndim; ++k) lmi[k] = periodic[k] ? box[k] : 2.5 * box[k];
  double half = side / 2.0;
  if (half < *rcrdf_inout) *rcrdf_inout = half; /* rcrdf = min(rcrdf, side/2) */
  if (*rcrdf_inout > 0.0)
    *side2 = (*rcrdf_inout) * (*rcrdf_inout);
  else
    *side2 = half * half;
  double v = 1.0;
  for (int k = 0; k < ndim; ++k) v = v * box[k];
  *volume = v;
}

/* Wrap into [0,L): floor, then fold the two rounding edge cases. */
double or_wrap1(double x, double L) {
  double f = floor(x / L);
  double t = f * L;
  double w = x - t;
  if (w < 0.0) w = w + L;
  if (w >= L) w = w - L;
  return w;
}

/* rdf.cuf:55-58 with the cap of netcdf.f90:610-612 applied first (RDF_init runs
 * after select_ncdfinput on the first frame, trj_analysis.f90:187-193). */
int or_rdf_lsmax(int ndim, const double box[3], double rcrdf, double deltar) {
  double side = box[0];
  for (int k = 1; k < ndim; ++k)
    if (box[k] < side) side = box[k];
  double rc = rcrdf;
  if (side / 2.0 < rc) rc = side / 2.0;
  if (rc > 0.0) return (int)or_nint_d(rc / deltar) + 1;
  return (int)or_nint_d(side / 2.0 / deltar) + 1;
}

/* One pair, rdf.cuf:452-463 restated in FP64.  Returns the 1-based bin or 0. */
static inline int pair_bin(int ndim, double xi, double yi, double zi, double xj, double yj,
                           double zj, const double lmi[3], double side2, double deltar,
                           int lsmax) {
  double xd = xj - xi;
  double yd = yj - yi;
  xd = xd - lmi[0] * round(xd / lmi[0]);
  yd = yd - lmi[1] * round(yd / lmi[1]);
  double r2;
  if (ndim == 3) {
    double zd = zj - zi;
    zd = zd - lmi[2] * round(zd / lmi[2]);
    r2 = xd * xd + yd * yd + zd * zd;
  } else {
    r2 = xd * xd + yd * yd;
  }
  if (!(r2 < side2)) return 0;
  long k = or_nint_d(sqrt(r2) / deltar);
  if (k > 0 && k <= lsmax) return (int)k;
  return 0;
}

void or_rdf_brute(int64_t n, const double* x, const double* y, const double* z,
                  const int32_t* itype, int ndim, int nsp, const double lmi[3],
                  double side2, double deltar, int lsmax, int64_t* counts) {
  for (int64_t i = 0; i < n; ++i) {
    double zi = (ndim == 3) ? z[i] : 0.0;
    for (int64_t j = i + 1; j < n; ++j) {
      double zj = (ndim == 3) ? z[j] : 0.0;
      int k = pair_bin(ndim, x[i], y[i], zi, x[j], y[j], zj, lmi, side2, deltar, lsmax);
      if (k) {
        int ij = or_ind_ij(itype[i], itype[j], nsp);
        counts[(int64_t)(ij - 1) * lsmax + (k - 1)] += 1;
      }
    }
  }
}

/* Linked cells: cell side >= cutoff, per-axis neighbour sets deduplicated so
 * that each unordered pair is visited once, from its lower index. */
void or_rdf_cells(int64_t n, const double* x, const double* y, const double* z,
                  const int32_t* itype, int ndim, int nsp, const double box[3],
                  const int32_t periodic[3], const double lmi[3], double side2,
                  double deltar, int lsmax, int64_t* counts) {
  const double* pos[3] = {x, y, z};
  double rc = sqrt(side2);
  int nc[3] = {1, 1, 1};
  double h[3] = {1.0, 1.0, 1.0}, lo[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k < ndim; ++k) {
    if (!periodic[k]) {
      double mn = pos[k][0], mx = pos[k][0];
      for (int64_t i = 1; i < n; ++i) {
        if (pos[k][i] < mn) mn = pos[k][i];
        if (pos[k][i] > mx) mx = pos[k][i];
      }
      lo[k] = mn;
      double span = mx - mn;
      int c = (int)floor(span / (rc * 1.000001));
      nc[k] = c < 1 ? 1 : c;
      h[k] = span > 0.0 ? span / nc[k] * 1.000001 : 1.0;
    } else {
      int c = (int)floor(box[k] / (rc * 1.000001));
      nc[k] = c < 1 ? 1 : c;
      h[k] = box[k] / nc[k];
    }
  }
  int64_t ncell = (int64_t)nc[0] * nc[1] * nc[2];
  int64_t* head = (int64_t*)malloc(sizeof(int64_t) * ncell);
  int64_t* next = (int64_t*)malloc(sizeof(int64_t) * (n > 0 ? n : 1));
  int32_t* cell = (int32_t*)malloc(sizeof(int32_t) * 3 * (n > 0 ? n : 1));
  for (int64_t c = 0; c < ncell; ++c) head[c] = -1;
  for (int64_t i = n - 1; i >= 0; --i) {
    int cc[3] = {0, 0, 0};
    for (int k = 0; k < ndim; ++k) {
      long c = (long)floor((pos[k][i] - lo[k]) / h[k]);
      if (c < 0) c = 0;
      if (c >= nc[k]) c = nc[k] - 1;
      cc[k] = (int)c;
    }
    for (int k = 0; k < 3; ++k) cell[3 * i + k] = cc[k];
    int64_t id = ((int64_t)cc[2] * nc[1] + cc[1]) * nc[0] + cc[0];
    next[i] = head[id];
    head[id] = i;
  }
  int nit = nsp * (nsp + 1) / 2;
  int64_t hsz = (int64_t)nit * lsmax;
#pragma omp parallel
  {
    int64_t* loc = (int64_t*)calloc((size_t)hsz, sizeof(int64_t));
#pragma omp for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; ++i) {
      int set[3][3], nset[3];
      for (int k = 0; k < 3; ++k) {
        int c = cell[3 * i + k];
        nset[k] = 0;
        for (int o = -1; o <= 1; ++o) {
          int cand = c + o;
          if (k >= ndim) {
            if (o != 0) continue;
          } else if (periodic[k]) {
            cand = ((cand % nc[k]) + nc[k]) % nc[k];
          } else if (cand < 0 || cand >= nc[k]) {
            continue;
          }
          int dup = 0;
          for (int s = 0; s < nset[k]; ++s) dup |= (set[k][s] == cand);
          if (!dup) set[k][nset[k]++] = cand;
        }
      }
      double zi = (ndim == 3) ? z[i] : 0.0;
      for (int a = 0; a < nset[2]; ++a)
        for (int b = 0; b < nset[1]; ++b)
          for (int c = 0; c < nset[0]; ++c) {
            int64_t id = ((int64_t)set[2][a] * nc[1] + set[1][b]) * nc[0] + set[0][c];
            for (int64_t j = head[id]; j >= 0; j = next[j]) {
              if (j <= i) continue;
              double zj = (ndim == 3) ? z[j] : 0.0;
              int k = pair_bin(ndim, x[i], y[i], zi, x[j], y[j], zj, lmi, side2, deltar, lsmax);
              if (k) loc[(int64_t)(or_ind_ij(itype[i], itype[j], nsp) - 1) * lsmax + (k - 1)] += 1;
            }
          }
    }
#pragma omp critical
    for (int64_t q = 0; q < hsz; ++q) counts[q] += loc[q];
    free(loc);
  }
  free(head);
  free(next);
  free(cell);
}

/* Charge-weighted pair histogram (rdf.cuf:466-472 g_qq, per frame).  The
 * reference adds q_i*q_j in FP32 atomics; here every product is rounded once
 * to the fixed point of or_qq_scale and summed exactly in int64, so the sum
 * does not depend on the order of the pairs.  acc[k-1] += llrint(qi*qj*scale)
 * for the in-range pairs of bin k (same pair rule as or_rdf_brute). */
double or_qq_scale(int64_t n, const double* q) {
  double m = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    double v = q[i] * q[i];
    if (v > m) m = v;
  }
  int e = 0;
  if (m > 0.0) frexp(m, &e); /* m < 2^e */
  return ldexp(1.0, 30 - e);
}

void or_rdf_qq_brute(int64_t n, const double* x, const double* y, const double* z,
                     const double* q, int ndim, const double lmi[3], double side2,
                     double deltar, int lsmax, double scale, int64_t* acc) {
  for (int64_t i = 0; i < n; ++i) {
    double zi = (ndim == 3) ? z[i] : 0.0;
    for (int64_t j = i + 1; j < n; ++j) {
      double zj = (ndim == 3) ? z[j] : 0.0;
      int k = pair_bin(ndim, x[i], y[i], zi, x[j], y[j], zj, lmi, side2, deltar, lsmax);
      if (k) {
        double p = q[i] * q[j];
        acc[k - 1] += llrint(p * scale);
      }
    }
  }
}

/* Host step of RDFcomp (rdf.cuf:285-287): gqq += gqq_frame / avq2 with
 * avq2 = sum(q^2)/N; here the frame's fixed-point sums are converted back
 * ((double)fixed / scale, exact power of two) and divided by avq2 (one
 * rounding).  avq2 = 0 (no charges) adds nothing. */
void or_qq_accumulate(int lsmax, const int64_t* fixed, double scale, int64_t n, const double* q,
                      double* acc) {
  double s2 = 0.0;
  for (int64_t i = 0; i < n; ++i) s2 += q[i] * q[i];
  double avq2 = s2 / (double)n;
  if (!(avq2 > 0.0)) return;
  for (int k = 0; k < lsmax; ++k) acc[k] += ((double)fixed[k] / scale) / avq2;
}

/* log.f90:501-503: g_qq(r_i) = 2 V gqq(i) / (N^2 dV Nconf), i = 1..lsmax-2,
 * dV the spherical (3D) or circular (2D) shell, FP64. */
void or_rdf_gqq(int ndim, int lsmax, double deltar, double volume, int64_t natoms, int64_t nconf,
                const double* acc, double* out) {
  const double pi = 3.14159265358979323846;
  for (int i = 1; i <= lsmax - 2; ++i) {
    double ri = i * deltar, a = ri + deltar / 2.0, b = ri - deltar / 2.0;
    double dv = (ndim == 3) ? 4.0 * pi * (a * a * a - b * b * b) / 3.0 : pi * (a * a - b * b);
    out[i - 1] = 2.0 * volume * acc[i - 1] / ((double)natoms * (double)natoms * dv * (double)nconf);
  }
}

/* log.f90:484-518, evaluated in FP64. */
void or_rdf_gmix(int ndim, int nsp, int lsmax, double deltar, double volume,
                 const int64_t* ntype, int64_t nconf, const int64_t* counts, double* out) {
  const double pi = 3.14159265358979323846;
  int nit = nsp * (nsp + 1) / 2;
  for (int i = 1; i <= lsmax - 2; ++i) {
    double ri = i * deltar;
    double dv;
    if (ndim == 3) {
      double a = ri + deltar / 2, b = ri - deltar / 2;
      dv = 4.0 * pi * (a * a * a - b * b * b) / 3.0;
    } else {
      double a = ri + deltar / 2, b = ri - deltar / 2;
      dv = pi * (a * a - b * b);
    }
    double* row = out + (int64_t)(i - 1) * (1 + nit);
    row[0] = ri;
    int col = 1;
    for (int j = 1; j <= nsp; ++j)
      for (int l = j; l <= nsp; ++l) {
        int ij = or_ind_ij(j, l, nsp);
        double h = (double)counts[(int64_t)(ij - 1) * lsmax + (i - 1)];
        double fac = (double)(j / l + 1);
        row[col++] = fac * volume * h / (dv * (double)ntype[l - 1] * (double)ntype[j - 1] * (double)nconf);
      }
  }
}

/* rdf.cuf:313-342 then rdf.cuf:755-822. */
double or_first_minimum(int ndim, int nsp, int lsmax, double deltar, double volume,
                        int64_t natoms, const int64_t* counts) {
  const double pi = 3.14159265358979323846;
  int nit = nsp * (nsp + 1) / 2;
  double* g = (double*)calloc((size_t)lsmax + 1, sizeof(double));
  for (int i = 1; i <= lsmax; ++i) {
    double ri = i * deltar, a = ri + deltar / 2, b = ri - deltar / 2, dv;
    if (ndim == 3)
      dv = 4.0 * pi * (a * a * a - b * b * b) / 3.0;
    else
      dv = pi * (a * a - b * b);
    double s = 0.0;
    for (int ij = 1; ij <= nit; ++ij) s += (double)counts[(int64_t)(ij - 1) * lsmax + (i - 1)];
    g[i] = 2.0 * volume * s / (dv * (double)natoms * (double)natoms);
  }
  int nrdf = lsmax, imax = 1, imin;
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
  imin = imax + 1;
  for (int i = imax + 1; i <= nrdf - 1; ++i)
    if (g[i] < g[i - 1] && g[i] < g[i + 1]) {
      imin = i;
      break;
    }
  double rmin;
  if (imin > 0 && imin < nrdf) {
    double x1 = deltar * (imin - 1), x2 = deltar * imin, x3 = deltar * (imin + 1);
    double y1 = g[imin - 1], y2 = g[imin], y3 = g[imin + 1];
    double denom = y1 - 2.0 * y2 + y3;
    if (fabs(denom) > 1e-12)
      rmin = x2 + 0.5 * (y1 - y3) / denom * (x3 - x1) / 2.0;
    else
      rmin = x2;
  } else {
    rmin = deltar * imin;
  }
  free(g);
  return rmin;
}

/* sq.cuf:176-178 (3D, ndg=4) and 184-185 (2D, ndg2=16). */
static int adj(int l, int m) { return (l % m != 0) ? l - m + (l % m) : l; }

/* sq.cuf:156-200 in FP32. */
void or_sq_setup(int ndim, const double box[3], double qmax, double qmin, or_sqgrid* g) {
  float side[3] = {(float)box[0], (float)box[1], (float)box[2]};
  float twopi = 2.0f * OR_PI_F;
  float fk[3];
  for (int k = 0; k < 3; ++k) fk[k] = twopi / side[k];
  float qmaxf = (float)qmax, qminf = (float)qmin;
  float dq;
  if (ndim == 2) {
    dq = fk[0] < fk[1] ? fk[0] : fk[1];
  } else {
    dq = fk[0];
    if (fk[1] < dq) dq = fk[1];
    if (fk[2] < dq) dq = fk[2];
  }
  int nqmax = or_nint_f(qmaxf / dq);
  int l[3] = {or_nint_f(qminf / fk[0]), or_nint_f(qminf / fk[1]), 0};
  if (ndim == 3) l[2] = or_nint_f(qminf / fk[2]);
  int nqmin, kext[3];
  if (ndim == 3) {
    for (int k = 0; k < 3; ++k) l[k] = adj(l[k], 4);
    for (int k = 0; k < 3; ++k) kext[k] = 4 * (l[k] / 4 + 1);
    nqmin = l[0];
    if (l[1] < nqmin) nqmin = l[1];
    if (l[2] < nqmin) nqmin = l[2];
  } else {
    for (int k = 0; k < 2; ++k) l[k] = adj(l[k], 16);
    kext[0] = 16 * (l[0] / 16 + 1);
    kext[1] = 16 * (l[1] / 16 + 1);
    kext[2] = 1;
    nqmin = l[0] < l[1] ? l[0] : l[1];
  }
  float qminp;
  {
    float a = fk[0] * (float)l[0], b = fk[1] * (float)l[1];
    qminp = a < b ? a : b;
    if (ndim == 3) {
      float c = fk[2] * (float)l[2];
      if (c < qminp) qminp = c;
    }
  }
  if (nqmin == 0) {
    nqmin = 1;
    l[0] = l[1] = l[2] = 1;
  }
  for (int k = 0; k < 3; ++k) g->fk[k] = fk[k];
  g->dq = dq;
  g->nqmax = nqmax;
  for (int k = 0; k < 3; ++k) g->l[k] = l[k];
  g->nqmin = nqmin;
  g->qminp = qminp;
  g->qmin2 = qminp * qminp;
  for (int k = 0; k < 3; ++k) g->kext[k] = kext[k];
}

/* sq.cuf:303-343 */
void or_sq_nq(int ndim, const or_sqgrid* g, int64_t* nq) {
  for (int s = 0; s <= g->nqmax; ++s) nq[s] = 0;
  int fill = (ndim == 3) ? 3 : 2;
  if (g->nqmin == 1) {
    for (int s = 1; s <= g->nqmax; ++s) nq[s] = fill;
    return;
  }
  int indm = 0;
  if (ndim == 3) {
    for (int i = 0; i <= g->l[0] - 1; ++i)
      for (int j = 0; j <= g->l[1] - 1; ++j)
        for (int k = 0; k <= g->l[2] - 1; ++k) {
          if (i + j + k == 0) continue;
          float a = (float)i * g->fk[0], b = (float)j * g->fk[1], c = (float)k * g->fk[2];
          float rr = sqrtf(a * a + b * b + c * c);
          if (rr < g->qminp) {
            int ind = or_nint_f(rr / g->dq);
            if (ind > indm) indm = ind;
            if (ind <= g->nqmax) nq[ind] += 1;
          }
        }
  } else {
    for (int i = 0; i <= g->l[0]; ++i)
      for (int j = 0; j <= g->l[1]; ++j) {
        if (i + j == 0) continue;
        float a = (float)i * g->fk[0], b = (float)j * g->fk[1];
        float rr = sqrtf(a * a + b * b);
        if (rr < g->qminp) {
          int ind = or_nint_f(rr / g->dq);
          if (ind > indm) indm = ind;
          if (ind <= g->nqmax) nq[ind] += 1;
        }
      }
  }
  for (int s = indm + 1; s <= g->nqmax; ++s) nq[s] = fill;
}

/* sq.cuf:533-555,578 */
int or_sq_member(int ndim, const or_sqgrid* g, int kx, int ky, int kz) {
  float kf1 = (float)kx * g->fk[0];
  float kf2 = (float)ky * g->fk[1];
  float q2;
  if (ndim == 3) {
    float kf3 = (float)kz * g->fk[2];
    q2 = kf1 * kf1 + kf2 * kf2 + kf3 * kf3;
  } else {
    q2 = kf1 * kf1 + kf2 * kf2;
  }
  if (!(q2 > 0.0f && q2 <= g->qmin2)) return -1;
  float q = sqrtf(q2);
  int ind = (int)(q / g->dq + 0.501f);
  if (ind > g->nqmax) return -1;
  return ind;
}

/* Phase in turns, frac(n . u), long double. */
static inline long double turns(int nx, int ny, int nz, long double ux, long double uy,
                                long double uz) {
  long double t = (long double)nx * ux + (long double)ny * uy + (long double)nz * uz;
  return t - floorl(t);
}

void or_sq_rho(int64_t n, const double* x, const double* y, const double* z,
               const int32_t* itype, int nsp, const double box[3], int64_t nvec,
               const int32_t* kv, double* rho) {
  const long double twopi = 6.283185307179586476925286766559005768L;
#pragma omp parallel for schedule(dynamic, 4)
  for (int64_t v = 0; v < nvec; ++v) {
    long double re[16], im[16];
    for (int s = 0; s < nsp; ++s) re[s] = im[s] = 0.0L;
    int nx = kv[3 * v], ny = kv[3 * v + 1], nz = kv[3 * v + 2];
    for (int64_t j = 0; j < n; ++j) {
      long double ux = (long double)x[j] / (long double)box[0];
      long double uy = (long double)y[j] / (long double)box[1];
      long double uz = z ? (long double)z[j] / (long double)box[2] : 0.0L;
      long double ph = twopi * turns(nx, ny, nz, ux, uy, uz);
      int s = itype[j] - 1;
      re[s] += cosl(ph);
      im[s] += sinl(ph);
    }
    for (int s = 0; s < nsp; ++s) {
      rho[(v * nsp + s) * 2] = (double)re[s];
      rho[(v * nsp + s) * 2 + 1] = (double)im[s];
    }
  }
}

/* The reference sum set: sqfnD octant (sq.cuf:89, when nqmin > 1) and the
 * axis shells of sqfact/sqfact2 (sq.cuf:94-98, 112-120; the idir = 0 form is
 * used when nqmin <= 1).  Returns the count; *kv_out [3 per vector] and
 * *sh_out (shell per vector) are malloc'ed. */
static int64_t sum_set(int ndim, const or_sqgrid* g, int32_t** kv_out, int32_t** sh_out) {
  int64_t cap = 1024, nv = 0;
  int32_t* kv = (int32_t*)malloc(sizeof(int32_t) * 3 * cap);
  int32_t* sh = (int32_t*)malloc(sizeof(int32_t) * cap);
#define PUSH(a, b, c, s)                                              \
  do {                                                                \
    if (nv == cap) {                                                  \
      cap *= 2;                                                       \
      kv = (int32_t*)realloc(kv, sizeof(int32_t) * 3 * cap);          \
      sh = (int32_t*)realloc(sh, sizeof(int32_t) * cap);              \
    }                                                                 \
    kv[3 * nv] = (a);                                                 \
    kv[3 * nv + 1] = (b);                                             \
    kv[3 * nv + 2] = (c);                                             \
    sh[nv++] = (s);                                                   \
  } while (0)
  int kmin_axis;
  if (g->nqmin > 1) {
    int kz_ext = (ndim == 3) ? g->kext[2] : 1;
    for (int a = 0; a < g->kext[0]; ++a)
      for (int b = 0; b < g->kext[1]; ++b)
        for (int c = 0; c < kz_ext; ++c) {
          int s = or_sq_member(ndim, g, a, b, c);
          if (s >= 0) PUSH(a, b, c, s);
        }
    kmin_axis = g->nqmin + 1;
  } else {
    kmin_axis = 1;
  }
  for (int k = kmin_axis; k <= g->nqmax; ++k) {
    PUSH(k, 0, 0, k);
    PUSH(0, k, 0, k);
    if (ndim == 3) PUSH(0, 0, k, k);
  }
#undef PUSH
  *kv_out = kv;
  *sh_out = sh;
  return nv;
}

void or_sq_frame(int64_t n, const double* x, const double* y, const double* z,
                 const int32_t* itype, int ndim, int nsp, const double box[3],
                 const or_sqgrid* g, const double* bsc, double* sqf, double* sqfp) {
  int32_t *kv, *sh;
  int64_t nv = sum_set(ndim, g, &kv, &sh);
  double* rho = (double*)malloc(sizeof(double) * 2 * nsp * (nv > 0 ? nv : 1));
  or_sq_rho(n, x, y, ndim == 3 ? z : NULL, itype, nsp, box, nv, kv, rho);
  int ns = g->nqmax + 1;
  for (int64_t v = 0; v < nv; ++v) {
    double tr = 0.0, ti = 0.0;
    for (int s = 0; s < nsp; ++s) {
      double r = rho[(v * nsp + s) * 2], i = rho[(v * nsp + s) * 2 + 1];
      sqfp[(int64_t)s * ns + sh[v]] += r * r + i * i;
      tr += bsc[s] * r;
      ti += bsc[s] * i;
    }
    sqf[sh[v]] += tr * tr + ti * ti;
  }
  free(rho);
  free(kv);
  free(sh);
}

/* rho_s(q) of or_sq_rho with FP64 phases (the CPU baseline's timing port):
 * per vector, the turns n.(r/L) in double, reduced to [0,1), one sincos per
 * atom, species sums in double; OpenMP over the vectors. */
void or_sq_rho_f64(int64_t n, const double* x, const double* y, const double* z,
                   const int32_t* itype, int nsp, const double box[3], int64_t nvec,
                   const int32_t* kv, double* rho) {
  const double twopi = 6.283185307179586476925286766559;
  const double ix = 1.0 / box[0], iy = 1.0 / box[1], iz = z ? 1.0 / box[2] : 0.0;
#pragma omp parallel for schedule(dynamic, 8)
  for (int64_t v = 0; v < nvec; ++v) {
    double re[16], im[16];
    for (int s = 0; s < nsp; ++s) re[s] = im[s] = 0.0;
    const double nx = kv[3 * v], ny = kv[3 * v + 1], nz = kv[3 * v + 2];
    for (int64_t j = 0; j < n; ++j) {
      double t = nx * (x[j] * ix) + ny * (y[j] * iy) + (z ? nz * (z[j] * iz) : 0.0);
      t -= floor(t);
      double sn, cs;
      sincos(twopi * t, &sn, &cs);
      const int s = itype[j] - 1;
      re[s] += cs;
      im[s] += sn;
    }
    for (int s = 0; s < nsp; ++s) {
      rho[(v * nsp + s) * 2] = re[s];
      rho[(v * nsp + s) * 2 + 1] = im[s];
    }
  }
}

/* sq.cuf:523-591 and 372-454 as or_sq_frame, with the FP64 phases of
 * or_sq_rho_f64 (the CPU baseline's timing port). */
void or_sq_frame_f64(int64_t n, const double* x, const double* y, const double* z,
                     const int32_t* itype, int ndim, int nsp, const double box[3],
                     const or_sqgrid* g, const double* bsc, double* sqf, double* sqfp) {
  int32_t *kv, *sh;
  int64_t nv = sum_set(ndim, g, &kv, &sh);
  double* rho = (double*)malloc(sizeof(double) * 2 * nsp * (nv > 0 ? nv : 1));
  or_sq_rho_f64(n, x, y, ndim == 3 ? z : NULL, itype, nsp, box, nv, kv, rho);
  int ns = g->nqmax + 1;
  for (int64_t v = 0; v < nv; ++v) {
    double tr = 0.0, ti = 0.0;
    for (int s = 0; s < nsp; ++s) {
      double r = rho[(v * nsp + s) * 2], i = rho[(v * nsp + s) * 2 + 1];
      sqfp[(int64_t)s * ns + sh[v]] += r * r + i * i;
      tr += bsc[s] * r;
      ti += bsc[s] * i;
    }
    sqf[sh[v]] += tr * tr + ti * ti;
  }
  free(rho);
  free(kv);
  free(sh);
}

/* Charge-weighted sums of the same set (sq.cuf:424-432 sqfact, sq.cuf:563-582
 * sqfnD with ex_qc): sqfq[shell] += |sum_j q_j exp(i q.r_j)|^2, long double. */
void or_sq_frame_q(int64_t n, const double* x, const double* y, const double* z, const double* q,
                   int ndim, const double box[3], const or_sqgrid* g, double* sqfq) {
  const long double twopi = 6.283185307179586476925286766559005768L;
  int32_t *kv, *sh;
  int64_t nv = sum_set(ndim, g, &kv, &sh);
#pragma omp parallel for schedule(dynamic, 4)
  for (int64_t v = 0; v < nv; ++v) {
    long double re = 0.0L, im = 0.0L;
    int nx = kv[3 * v], ny = kv[3 * v + 1], nz = kv[3 * v + 2];
    for (int64_t j = 0; j < n; ++j) {
      long double ux = (long double)x[j] / (long double)box[0];
      long double uy = (long double)y[j] / (long double)box[1];
      long double uz = (ndim == 3) ? (long double)z[j] / (long double)box[2] : 0.0L;
      long double ph = twopi * turns(nx, ny, ndim == 3 ? nz : 0, ux, uy, uz);
      re += (long double)q[j] * cosl(ph);
      im += (long double)q[j] * sinl(ph);
    }
    double r = (double)re, i = (double)im;
    double p = r * r + i * i;
#pragma omp atomic
    sqfq[sh[v]] += p;
  }
  free(kv);
  free(sh);
}

/* log.f90:317-401 */
int or_sq_columns(int nsp, int nqmax, float dq, int64_t natoms, const int64_t* ntype,
                  const double* bsc, int64_t nconf, const double* sqf, const double* sqfp,
                  const int64_t* nq, double* cols, double* mix) {
  double baver = 0.0;
  for (int s = 0; s < nsp; ++s) baver += (double)ntype[s] * bsc[s];
  baver /= (double)natoms;
  int bsc_one = 1;
  for (int s = 0; s < nsp; ++s)
    if (fabs(bsc[s] - 1.0) > 1.0e-4) bsc_one = 0;
  int binary = (nsp == 2 && bsc_one);
  int ncols = binary ? 7 : 3;
  double x1 = 0, x2 = 0;
  if (binary) {
    x1 = (double)ntype[0] / (double)natoms;
    x2 = (double)ntype[1] / (double)natoms;
  }
  int ns = nqmax + 1;
  for (int i = 1; i <= nqmax; ++i) {
    double* row = cols + (int64_t)(i - 1) * ncols;
    double nqi = (double)nq[i];
    double snn = (sqf[i] / (baver * baver)) / ((double)natoms * (double)nconf * nqi);
    row[0] = (double)((float)i * dq);
    row[1] = snn;
    if (binary) {
      double s11 = x1 * sqfp[0 * ns + i] / ((double)ntype[0] * (double)nconf * nqi);
      double s22 = x2 * sqfp[1 * ns + i] / ((double)ntype[1] * (double)nconf * nqi);
      double s12 = 0.5 * (snn - s11 - s22);
      double scc = x2 * x2 * s11 + x1 * x1 * s22 - 2.0 * x1 * x2 * s12;
      row[2] = scc;
      row[3] = s11;
      row[4] = s22;
      row[5] = s12;
      row[6] = nqi;
    } else {
      row[2] = nqi;
    }
    double* mrow = mix + (int64_t)(i - 1) * (1 + nsp);
    mrow[0] = row[0];
    for (int s = 0; s < nsp; ++s)
      mrow[1 + s] = sqfp[(int64_t)s * ns + i] / ((double)ntype[s] * (double)nconf * nqi);
  }
  return ncols;
}

/* ------------------------------------------------------------------ slices */

/* Slice of a z coordinate (rdf.cuf:692-696, sq.cuf:640-643, common.cuf:225-226),
 * in FP32 as the reference: the first zslice nearest to z (minloc), and -1
 * unless |zslice - z| <= zgrid/2. */
int or_slice_of(float z, int nslice, const float* zslice, float zgrid) {
  int best = 0;
  float bd = fabsf(zslice[0] - z);
  for (int s = 1; s < nslice; ++s) {
    float d = fabsf(zslice[s] - z);
    if (d < bd) {
      bd = d;
      best = s;
    }
  }
  return (bd <= zgrid / 2.0f) ? best : -1;
}

/* Index of the last atom of the reference's species-sorted order (species
 * 1..nsp, file order within a species): the last atom of the highest species
 * present.  rdfxy's thread loop (rdf.cuf:688) stops before it. */
static int64_t last_in_species_order(int64_t n, const int32_t* itype) {
  int64_t last = -1;
  int32_t ts = 0;
  for (int64_t i = 0; i < n; ++i)
    if (itype[i] >= ts) {
      ts = itype[i];
      last = i;
    }
  return last;
}

/* rdfxy (rdf.cuf:672-736): pairs i<j of one slice, xy distance with the
 * nint image on x and y (lxy), r^2 < side2, bin k = int(r/deltar) + 1 in
 * FP64 here, 1 <= k <= lsmax.  counts[(sl*nit + ij-1)*lsmax + k-1]; count[sl]
 * and countp[sl*nsp + s-1] are the slice members without the last atom of
 * the species-sorted order (rdf.cuf:688, 699-700).  With q (may be NULL)
 * fixed[sl*lsmax + k-1] += llrint(q_i q_j scale). */
void or_rdf_xy_brute(int64_t n, const double* x, const double* y, const double* z,
                     const int32_t* itype, int nsp, const double lxy[2], double side2, double deltar,
                     int lsmax, int nslice, const float* zslice, float zgrid, int64_t* counts,
                     int64_t* count, int64_t* countp, const double* q, double scale, int64_t* fixed) {
  int nit = nsp * (nsp + 1) / 2;
  int* sl = (int*)malloc(sizeof(int) * (n > 0 ? n : 1));
  for (int64_t i = 0; i < n; ++i) sl[i] = or_slice_of((float)z[i], nslice, zslice, zgrid);
  int64_t last = last_in_species_order(n, itype);
  for (int64_t i = 0; i < n; ++i) {
    if (sl[i] < 0 || i == last) continue;
    count[sl[i]] += 1;
    countp[(int64_t)sl[i] * nsp + itype[i] - 1] += 1;
  }
  for (int64_t i = 0; i < n; ++i) {
    if (sl[i] < 0) continue;
    for (int64_t j = i + 1; j < n; ++j) {
      if (sl[j] != sl[i]) continue;
      double xd = x[j] - x[i], yd = y[j] - y[i];
      xd = xd - lxy[0] * round(xd / lxy[0]);
      yd = yd - lxy[1] * round(yd / lxy[1]);
      double r2 = xd * xd + yd * yd;
      if (!(r2 < side2)) continue;
      long k = (long)(sqrt(r2) / deltar) + 1;
      if (k < 1 || k > lsmax) continue;
      int ij = or_ind_ij(itype[i], itype[j], nsp);
      counts[((int64_t)sl[i] * nit + ij - 1) * lsmax + k - 1] += 1;
      if (q) fixed[(int64_t)sl[i] * lsmax + k - 1] += llrint(q[i] * q[j] * scale);
    }
  }
  free(sl);
}

/* rdf.cuf:284-296 per frame, in FP64: hxy[sl][j][l][k] += counts / (countp_j countp_l)
 * for j <= l with both counts positive (hxy: [nslice][nsp][nsp][lsmax], j <= l
 * filled). */
void or_rdf_xy_accumulate(int nsp, int lsmax, int nslice, const int64_t* counts, const int64_t* countp,
                          double* hxy) {
  int nit = nsp * (nsp + 1) / 2;
  for (int sl = 0; sl < nslice; ++sl) {
    int ij = 0;
    for (int j = 0; j < nsp; ++j)
      for (int l = j; l < nsp; ++l, ++ij) {
        double cj = (double)countp[(int64_t)sl * nsp + j], cl = (double)countp[(int64_t)sl * nsp + l];
        if (!(cj > 0.0 && cl > 0.0)) continue;
        for (int k = 0; k < lsmax; ++k)
          hxy[(((int64_t)sl * nsp + j) * nsp + l) * lsmax + k] +=
              (double)counts[((int64_t)sl * nit + ij) * lsmax + k] / (cj * cl);
      }
  }
}

/* log.f90:494-496, 511-514, 545-551: g_xy for pair (j, l), j <= l, and
 * r_i = i deltar, i = 1..lsmax-2: (j/l + 1) V (zgrid/Lz) hxy / (dV Nconf),
 * dV = pi ((r+dr/2)^2 - (r-dr/2)^2) zgrid; the Fortran j/l + 1 is 2 for
 * j = l and 1 otherwise.  out[(pair*(lsmax-2) + i-1)*nslice + sl], pairs in
 * ind_ij order. */
void or_rdf_xy_gmix(int nsp, int lsmax, int nslice, double deltar, const double box[3], double zgrid,
                    int64_t nconf, const double* hxy, double* out) {
  const double pi = 3.14159265358979323846;
  double v = box[0] * box[1] * box[2];
  int pair = 0;
  for (int j = 0; j < nsp; ++j)
    for (int l = j; l < nsp; ++l, ++pair)
      for (int i = 1; i <= lsmax - 2; ++i) {
        double ri = i * deltar, a = ri + deltar / 2.0, b = ri - deltar / 2.0;
        double dv = pi * (a * a - b * b) * zgrid;
        double f = (j == l) ? 2.0 : 1.0;
        for (int sl = 0; sl < nslice; ++sl)
          out[((int64_t)pair * (lsmax - 2) + i - 1) * nslice + sl] =
              f * v * (zgrid / box[2]) * hxy[(((int64_t)sl * nsp + j) * nsp + l) * lsmax + i - 1] /
              (dv * (double)nconf);
      }
}

/* The slice sum set: the 2D q-set (sq.cuf:89 sqf3xy when nqmin > 1, kx, ky
 * >= 0) and the x and y axis vectors k = nqmin+1 .. nqmax of sqf2xy
 * (sq.cuf:91-92, 114-116: also from nqmin+1 when nqmin <= 1, so shell 1
 * gets no vector when nqmin = 1). */
static int64_t sum_set_xy(const or_sqgrid* g, int32_t** kv_out, int32_t** sh_out) {
  int64_t cap = 1024, nv = 0;
  int32_t* kv = (int32_t*)malloc(sizeof(int32_t) * 2 * cap);
  int32_t* sh = (int32_t*)malloc(sizeof(int32_t) * cap);
#define PUSH2(a, b, s)                                                \
  do {                                                                \
    if (nv == cap) {                                                  \
      cap *= 2;                                                       \
      kv = (int32_t*)realloc(kv, sizeof(int32_t) * 2 * cap);          \
      sh = (int32_t*)realloc(sh, sizeof(int32_t) * cap);              \
    }                                                                 \
    kv[2 * nv] = (a);                                                 \
    kv[2 * nv + 1] = (b);                                             \
    sh[nv++] = (s);                                                   \
  } while (0)
  if (g->nqmin > 1)
    for (int a = 0; a < g->kext[0]; ++a)
      for (int b = 0; b < g->kext[1]; ++b) {
        int s = or_sq_member(2, g, a, b, 0);
        if (s >= 0) PUSH2(a, b, s);
      }
  /* nqmin can be negative after the 16-adjust; the reference's axis loop
   * then starts below 1, and only k >= 1 lands in a shell. */
  int kmin = g->nqmin + 1 > 1 ? g->nqmin + 1 : 1;
  for (int k = kmin; k <= g->nqmax; ++k) {
    PUSH2(k, 0, k);
    PUSH2(0, k, k);
  }
#undef PUSH2
  *kv_out = kv;
  *sh_out = sh;
  return nv;
}

/* sqf3xy / sqf2xy (sq.cuf:596-771) for one frame, long-double sums: per
 * slice sl and vector, sqf[sl][shell] += |sum_s bsc_s rho_s|^2 / count_sl,
 * sqfp[sl][s][shell] += |rho_s|^2 / countsp_sl,s and (q not NULL)
 * sqfq[sl][shell] += |sum q_j e^{iq.r_j}|^2 / count_sl, over the slice's
 * members (all atoms, unlike rdfxy).  Phases use x and y only; g is the 2D
 * grid (or_sq_setup with ndim = 2).  Arrays are [nslice][...][nqmax+1]. */
void or_sq_frame_xy(int64_t n, const double* x, const double* y, const double* z, const int32_t* itype,
                    const double* q, int nsp, const double box[3], const or_sqgrid* g, const double* bsc,
                    int nslice, const float* zslice, float zgrid, double* sqf, double* sqfp, double* sqfq) {
  const long double twopi = 6.283185307179586476925286766559005768L;
  int32_t *kv, *sh;
  int64_t nv = sum_set_xy(g, &kv, &sh);
  int* sl = (int*)malloc(sizeof(int) * (n > 0 ? n : 1));
  int64_t cnt[64], cntp[64 * 16];
  for (int s = 0; s < nslice; ++s) {
    cnt[s] = 0;
    for (int t = 0; t < nsp; ++t) cntp[s * nsp + t] = 0;
  }
  for (int64_t i = 0; i < n; ++i) {
    sl[i] = or_slice_of((float)z[i], nslice, zslice, zgrid);
    if (sl[i] >= 0) {
      cnt[sl[i]] += 1;
      cntp[sl[i] * nsp + itype[i] - 1] += 1;
    }
  }
  int ns = g->nqmax + 1;
#pragma omp parallel for schedule(dynamic, 4)
  for (int64_t v = 0; v < nv; ++v) {
    long double re[64 * 16], im[64 * 16], qre[64], qim[64];
    for (int s = 0; s < nslice * nsp; ++s) re[s] = im[s] = 0.0L;
    for (int s = 0; s < nslice; ++s) qre[s] = qim[s] = 0.0L;
    int nx = kv[2 * v], ny = kv[2 * v + 1];
    for (int64_t j = 0; j < n; ++j) {
      if (sl[j] < 0) continue;
      long double ux = (long double)x[j] / (long double)box[0];
      long double uy = (long double)y[j] / (long double)box[1];
      long double ph = twopi * turns(nx, ny, 0, ux, uy, 0.0L);
      long double c = cosl(ph), s = sinl(ph);
      int ch = sl[j] * nsp + itype[j] - 1;
      re[ch] += c;
      im[ch] += s;
      if (q) {
        qre[sl[j]] += (long double)q[j] * c;
        qim[sl[j]] += (long double)q[j] * s;
      }
    }
    for (int s = 0; s < nslice; ++s) {
      double tr = 0.0, ti = 0.0;
      for (int t = 0; t < nsp; ++t) {
        double r = (double)re[s * nsp + t], i = (double)im[s * nsp + t];
        if (cntp[s * nsp + t] > 0) {
          double p = (r * r + i * i) / (double)cntp[s * nsp + t];
#pragma omp atomic
          sqfp[((int64_t)s * nsp + t) * ns + sh[v]] += p;
        }
        tr += bsc[t] * r;
        ti += bsc[t] * i;
      }
      if (cnt[s] > 0) {
        double p = (tr * tr + ti * ti) / (double)cnt[s];
#pragma omp atomic
        sqf[(int64_t)s * ns + sh[v]] += p;
        if (q) {
          double qr = (double)qre[s], qi = (double)qim[s];
          double pq = (qr * qr + qi * qi) / (double)cnt[s];
#pragma omp atomic
          sqfq[(int64_t)s * ns + sh[v]] += pq;
        }
      }
    }
  }
  free(sl);
  free(kv);
  free(sh);
}

/* ------------------------------------------------------ number fluctuations */

/* s2D3_sh / s2D2_sh (rdf.cuf:489-615) for given unit points (pts[p*3 + k] in
 * (0,1], point = box * pts): per point the atoms with r^2 < side2 binned by
 * k = nint(r/deltar) in FP64 (the image rule of or_rdf_brute), then the
 * cumulative N_k = sum_{b<=k} h_b, k = 0..lsmax, added as gi[k] += N_k and
 * g2i[k] += N_k^2 (int64). */
void or_s2n(int64_t np, const double* pts, const double box[3], int64_t n, const double* x, const double* y,
            const double* z, int ndim, const double lmi[3], double side2, double deltar, int lsmax, int64_t* gi,
            int64_t* g2i) {
  int64_t* h = (int64_t*)malloc(sizeof(int64_t) * (lsmax + 1));
  for (int64_t p = 0; p < np; ++p) {
    for (int k = 0; k <= lsmax; ++k) h[k] = 0;
    double xp = box[0] * pts[3 * p], yp = box[1] * pts[3 * p + 1], zp = box[2] * pts[3 * p + 2];
    for (int64_t j = 0; j < n; ++j) {
      double dx = xp - x[j], dy = yp - y[j];
      dx = dx - lmi[0] * round(dx / lmi[0]);
      dy = dy - lmi[1] * round(dy / lmi[1]);
      double r2 = dx * dx + dy * dy;
      if (ndim == 3) {
        double dz = zp - z[j];
        dz = dz - lmi[2] * round(dz / lmi[2]);
        r2 += dz * dz;
      }
      if (!(r2 < side2)) continue;
      long k = or_nint_d(sqrt(r2) / deltar);
      if (k >= 0 && k <= lsmax) h[k] += 1;
    }
    int64_t nk = 0;
    for (int k = 0; k <= lsmax; ++k) {
      nk += h[k];
      gi[k] += nk;
      g2i[k] += nk * nk;
    }
  }
  free(h);
}

/* ------------------------------------------------------------------ F(Q,t) */

/* The S(Q,w) q-set (sq.cuf:234-282, FP32; the storing loop's k range):
 * vectors (i,j,k) >= 0, not all 0, i <= nint(qmm/fk1), ..., in the first l with
 * qw(l) - dq/2 < |q| < qw(l) + dq/2.  Writes up to cap vectors grouped by l
 * (kv[3*v], bin[v]); returns the count. */
int64_t or_fqt_vectors(int ndim, const double box[3], int nqw, const double* qw, int64_t cap, int32_t* kv,
                       int32_t* bin) {
  float fk[3];
  for (int k = 0; k < 3; ++k) fk[k] = (2.0f * OR_PI_F) / (float)(k < ndim ? box[k] : 1.0);
  float dq = fk[0] < fk[1] ? fk[0] : fk[1];
  if (ndim == 3 && fk[2] < dq) dq = fk[2];
  float qmm = 0.0f;
  for (int l = 0; l < nqw; ++l)
    if ((float)qw[l] > qmm) qmm = (float)qw[l];
  int imax = or_nint_f(qmm / fk[0]), jmax = or_nint_f(qmm / fk[1]), kmax = ndim == 3 ? or_nint_f(qmm / fk[2]) : 0;
  int64_t nv = 0;
  for (int l = 0; l < nqw; ++l)
    for (int i = 0; i <= imax; ++i)
      for (int j = 0; j <= jmax; ++j)
        for (int k = 0; k <= kmax; ++k) {
          if (i + j + k == 0) continue;
          float a = i * fk[0], b = j * fk[1], c = k * fk[2];
          float rr = sqrtf(a * a + b * b + c * c);
          int first = -1;
          for (int m = 0; m < nqw; ++m)
            if ((float)qw[m] - dq / 2 < rr && (float)qw[m] + dq / 2 > rr) {
              first = m;
              break;
            }
          if (first != l) continue;
          if (nv < cap) {
            kv[3 * nv] = i;
            kv[3 * nv + 1] = j;
            kv[3 * nv + 2] = k;
            bin[nv] = l;
          }
          ++nv;
        }
  return nv;
}

/* FqtnD over a trajectory with rtcorr's origins (dynamics.cuf:157-245,
 * 421-508), long-double phases: pos[f][j][3] for f = 0..nframes-1 (frame
 * f+1).  tot/tots [nqw][ndyn] get the sums of the origins that completed
 * (itlimit frames), *count_org their number. */
void or_fqt(int ndim, int nsp, const double* bsc, const double box[3], int nqw, const double* qw, int nbuffer,
            int jump, int ndyn, int itlimit, int64_t nframes, int64_t n, const double* pos, const int32_t* itype,
            double* tot, double* tots, int64_t* count_org) {
  const long double twopi = 6.283185307179586476925286766559005768L;
  int64_t nv = or_fqt_vectors(ndim, box, nqw, qw, 0, NULL, NULL);
  int32_t* kv = (int32_t*)malloc(sizeof(int32_t) * 3 * (nv > 0 ? nv : 1));
  int32_t* bin = (int32_t*)malloc(sizeof(int32_t) * (nv > 0 ? nv : 1));
  or_fqt_vectors(ndim, box, nqw, qw, nv, kv, bin);
  int* origin = (int*)malloc(sizeof(int) * nbuffer);
  double* buf = (double*)calloc((size_t)nbuffer * nqw * ndyn, sizeof(double));
  double* bufs = (double*)calloc((size_t)nbuffer * nqw * ndyn, sizeof(double));
  int64_t* orgframe = (int64_t*)malloc(sizeof(int64_t) * nbuffer);
  for (int b = 0; b < nbuffer; ++b) {
    origin[b] = b * jump + 1;
    orgframe[b] = -1;
  }
  *count_org = 0;
  for (int64_t f = 0; f < nframes; ++f) {
    const int iconf = (int)f + 1;
    for (int b = 0; b < nbuffer; ++b)
      if (iconf == origin[b]) orgframe[b] = f;
    for (int b = 0; b < nbuffer; ++b) {
      int ix = iconf - origin[b] + 1;
      if (iconf < origin[b] || ix > ndyn) continue;
      const double* r0 = pos + orgframe[b] * n * 3;
      const double* rt = pos + f * n * 3;
#pragma omp parallel for schedule(dynamic, 4)
      for (int64_t v = 0; v < nv; ++v) {
        long double c0 = 0, s0 = 0, ct = 0, st = 0, self = 0;
        for (int64_t j = 0; j < n; ++j) {
          long double t0 = 0, tt = 0;
          for (int k = 0; k < ndim; ++k) {
            t0 += (long double)kv[3 * v + k] * (long double)r0[3 * j + k] / (long double)box[k];
            tt += (long double)kv[3 * v + k] * (long double)rt[3 * j + k] / (long double)box[k];
          }
          long double w = bsc[itype[j] - 1];
          c0 += w * cosl(twopi * t0);
          s0 += w * sinl(twopi * t0);
          ct += w * cosl(twopi * tt);
          st += w * sinl(twopi * tt);
          self += cosl(twopi * (tt - t0));
        }
        double fv = (double)(c0 * ct + s0 * st), fs = (double)self;
        size_t idx = ((size_t)b * nqw + bin[v]) * ndyn + (ix - 1);
#pragma omp atomic
        buf[idx] += fv;
#pragma omp atomic
        bufs[idx] += fs;
      }
    }
    for (int b = 0; b < nbuffer; ++b)
      if (iconf - origin[b] == itlimit) {
        origin[b] = iconf + 1;
        for (int i = 0; i < nqw * ndyn; ++i) {
          tot[i] += buf[(size_t)b * nqw * ndyn + i];
          tots[i] += bufs[(size_t)b * nqw * ndyn + i];
          buf[(size_t)b * nqw * ndyn + i] = 0.0;
          bufs[(size_t)b * nqw * ndyn + i] = 0.0;
        }
        *count_org += 1;
      }
  }
  free(kv);
  free(bin);
  free(origin);
  free(buf);
  free(bufs);
  free(orgframe);
}
