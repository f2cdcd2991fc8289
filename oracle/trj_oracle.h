/*
 * trj_oracle.h -- CPU restatement of the trj_analysis structural hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links or calls
 * this code; only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg use it, and only as the checker.
 *
 * Each function cites the reference file:line it restates (paths relative to
 * the reference repository root).  Arithmetic precision is stated per function:
 *   - the RDF pair rule is restated in FP64 with no FMA contraction;
 *   - the q-grid / shell bookkeeping is restated in FP32, because the reference
 *     does it in FP32 (myprec = sp, src/core/precision.f90:22) and the shipped
 *     n(Q) columns depend on FP32 rounding on the qmin sphere;
 *   - rho(q) is restated with long-double phases and sums (the FP64 oracle).
 *
 * Pinning status (what checks each function independently of its own formula):
 *   or_ind_ij            closed form for nsp<=10, exhaustive (tests/test_oracle_rdf.py)
 *   or_geometry          closed-form cases (cap, 2.5L non-periodic, volume)
 *   or_wrap              invariants: result in [0,L), x - result is a multiple of L
 *   or_rdf_brute         two-atom closed form; simple-cubic lattice shell counts
 *                        (sums of three squares); pair-count totals; permutation
 *                        and translation invariance
 *   or_rdf_cells         equals or_rdf_brute on random frames (several boxes/cutoffs)
 *   or_rdf_gmix          closed form of log.f90:489-518 on hand-made histograms;
 *                        ideal-gas g -> 1 statistically
 *   or_first_minimum     hand-made curves with a known parabola vertex
 *   or_sq_setup/or_sq_nq golden n(Q) vectors in the reference's shipped outputs
 *                        (examples/3D/LJmix, 3D/salr, 3D/meam-dyn, 2D/salr)
 *   or_sq_member         exact integer lattice-point counts per shell on a cubic
 *                        and a 1:2:1 box, away from the FP32 edges
 *                        (tests/test_oracle_pins.py); brute force on tiny grids
 *   or_sq_rho            single atom |rho|^2 = 1; two atoms 2+2cos(q.dr);
 *                        perfect lattice gives N^2 on reciprocal-lattice q and
 *                        ~0 elsewhere; translation invariance
 *   or_sq_columns        perfect lattices: exact Bragg values of S_NN, S_jj and the
 *                        ordered-alloy S_cc = (N/4) nB3/nq; ideal gas: S_jj -> 1,
 *                        S_NN -> <b^2>/<b>^2, s11 -> x1, s22 -> x2, s12 -> 0,
 *                        S_cc -> x1 x2; identical frames with Nconf = k; species
 *                        relabelling (tests/test_oracle_pins.py)
 *   or_sq_rho_f64,       the CPU-baseline timing port: equal to or_sq_rho and
 *   or_sq_frame_f64      or_sq_frame to 1e-12
 *   or_qq_scale          powers of two bracketing max q^2 (hand cases)
 *   or_rdf_qq_brute      unit charges give 2^29 x the or_rdf_brute counts;
 *                        +-1 by species gives c11 + c22 - c12; two-atom closed
 *                        form; invariant under q -> 2q (scale drops by 4)
 *   or_qq_accumulate     hand-made sums (exact arithmetic cases)
 *   or_rdf_gqq           equals or_rdf_gmix's formula with N_j N_l -> N^2 on the
 *                        same histogram (closed form of log.f90:501-503)
 *   or_slice_of          hand cases: nearest centre, first of two equidistant
 *                        centres (minloc), the zgrid/2 boundary in FP32
 *   or_rdf_xy_brute      one slice spanning every atom equals the 2D rule on the
 *                        xy projection with floor binning (a two-atom closed form,
 *                        pair totals); slices partition pairs (sum over slices of
 *                        a split equals the pairs with both atoms in one slice);
 *                        the last-atom count rule on hand cases
 *   or_rdf_xy_accumulate hand-made counts
 *   or_rdf_xy_gmix       an ideal slab: g_xy -> (2i-1)/(2i) per floor bin i, the
 *                        value the centred cylindrical shell fixes
 *                        (tests/test_oracle_pins.py)
 *   or_sq_frame_xy       one slice with every atom equals or_sq_frame (ndim 2)
 *                        divided by N; a single atom per slice gives 1 per vector;
 *                        unit charges give the bsc = 1 sum
 *   or_s2n               a single atom at a known distance; a point on a site of a
 *                        simple-cubic lattice, whose N_k is an integer-vector count;
 *                        N_k at lsmax equals the atoms within the cutoff
 *   or_fqt_vectors       brute-force shell membership in numpy (FP32) on hand grids
 *   or_fqt               identical frames: F = sum |rho|^2 per lag and Fs = N per
 *                        vector; a uniform drift d per frame: Fs = N cos(q.d lag),
 *                        F = |rho(0)|^2 cos(q.d lag); origin counts worked by hand
 *   or_sq_frame_q        unit charges equal or_sq_frame's S with bsc = 1;
 *                        uniform charge c scales by c^2; single atom gives q^2
 * None is "parity unpinned".
 */
#ifndef TRJ_ORACLE_H
#define TRJ_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Fortran NINT (round half away from zero) on doubles and floats. */
long or_nint_d(double x);
int or_nint_f(float x);

/* rdf.cuf:741-753: 1-based upper-triangle index of species pair (i,j). */
int or_ind_ij(int i, int j, int nsp);

/* netcdf.f90:602-622 per-frame geometry.
 * box: cell lengths; periodic: flags; rcrdf_inout: the sticky cutoff, capped to
 * half the shortest side (netcdf.f90:610-612).  Outputs: lmi = minimum-image
 * lengths (sidelv: 2.5*L on non-periodic axes, :603-609), side2 (:613-617),
 * volume (:618-622). */
void or_geometry(int ndim, const double box[3], const int32_t periodic[3],
                 double* rcrdf_inout, double lmi[3], double* side2, double* volume);

/* Coordinate wrap into [0,L) on periodic axes.  The reference truncates toward
 * zero (netcdf.f90:681); this build floors, then folds the two rounding edge
 * cases back into [0,L).  Non-periodic axes are copied unchanged. */
double or_wrap1(double x, double L);

/* rdf.cuf:55-58: lsmax = nint(rcrdf/deltar)+1, or from half the shortest side
 * when rcrdf <= 0. */
int or_rdf_lsmax(int ndim, const double box[3], double rcrdf, double deltar);

/* rdf.cuf:440-475 (3D) and rdf.cuf:631-659 (2D): all unordered pairs i<j,
 * minimum image d - Lmi*nint(d/Lmi), r2 < side2, k = nint(sqrt(r2)/deltar),
 * kept if 1 <= k <= lsmax.  FP64, no FMA.  Positions must already be wrapped.
 * counts[(ij-1)*lsmax + (k-1)] += 1 (layout of histomix_d(lsmax,nit)). */
void or_rdf_brute(int64_t n, const double* x, const double* y, const double* z,
                  const int32_t* itype, int ndim, int nsp, const double lmi[3],
                  double side2, double deltar, int lsmax, int64_t* counts);

/* Same result as or_rdf_brute, through linked cells (OpenMP over atoms). */
void or_rdf_cells(int64_t n, const double* x, const double* y, const double* z,
                  const int32_t* itype, int ndim, int nsp, const double box[3],
                  const int32_t periodic[3], const double lmi[3], double side2,
                  double deltar, int lsmax, int64_t* counts);

/* Charge-weighted histogram g_qq (rdf.cuf:466-472, 285-287; log.f90:501-503).
 * or_qq_scale: fixed-point scale 2^(30-e) with max q^2 < 2^e (2^30 for no
 * charges).  or_rdf_qq_brute: acc[k-1] += llrint(q_i q_j scale) over the pairs
 * of or_rdf_brute's rule.  or_qq_accumulate: acc += (fixed/scale)/avq2,
 * avq2 = sum(q^2)/N (FP64, index order).  or_rdf_gqq: 2 V acc / (N^2 dV Nconf)
 * for i = 1..lsmax-2. */
double or_qq_scale(int64_t n, const double* q);
void or_rdf_qq_brute(int64_t n, const double* x, const double* y, const double* z,
                     const double* q, int ndim, const double lmi[3], double side2,
                     double deltar, int lsmax, double scale, int64_t* acc);
void or_qq_accumulate(int lsmax, const int64_t* fixed, double scale, int64_t n, const double* q,
                      double* acc);
void or_rdf_gqq(int ndim, int lsmax, double deltar, double volume, int64_t natoms, int64_t nconf,
                const double* acc, double* out);

/* log.f90:484-518: g_jl(r_i) for i=1..lsmax-2.  out has (lsmax-2) rows of
 * (1 + nit) doubles: r, then g_11, g_12, ..., g_1n, g_22, ... */
void or_rdf_gmix(int ndim, int nsp, int lsmax, double deltar, double volume,
                 const int64_t* ntype, int64_t nconf, const int64_t* counts,
                 double* out);

/* rdf.cuf:313-342 running total g (3D/2D bulk) and rdf.cuf:755-822 first
 * minimum with parabolic interpolation.  Returns rmin. */
double or_first_minimum(int ndim, int nsp, int lsmax, double deltar, double volume,
                        int64_t natoms, const int64_t* counts);

/* sq.cuf:156-200: q-grid setup in FP32. */
typedef struct {
  float fk[3];     /* 2*pi/L per axis (sq.cuf:189) */
  float dq;        /* sq.cuf:156-160 */
  int nqmax;       /* sq.cuf:161 */
  int l[3];        /* lmax per axis after the mod-4 (3D) / mod-16 (2D) adjust */
  int nqmin;       /* sq.cuf:180/187, reset to 1 when 0 (sq.cuf:195-200) */
  float qminp;     /* qmin after sq.cuf:191/193 */
  float qmin2;     /* qmin**2 (sq.cuf:344) */
  int kext[3];     /* thread-grid extent of sqfnD per axis (sq.cuf:175-179, 183-186) */
} or_sqgrid;
void or_sq_setup(int ndim, const double box[3], double qmax, double qmin, or_sqgrid* g);

/* sq.cuf:303-343: host shell counts nq(0:nqmax). */
void or_sq_nq(int ndim, const or_sqgrid* g, int64_t* nq);

/* sq.cuf:533-555,578: shell of octant vector (kx,ky,kz) in the sqfnD sum set,
 * or -1 when it is not summed. */
int or_sq_member(int ndim, const or_sqgrid* g, int kx, int ky, int kz);

/* Long-double rho_s(n) = sum_{j in species s} exp(i 2 pi (n.(r_j/L))), for nvec
 * integer vectors kv[3*v..] (x,y,z).  itype 1..nsp.  rho out: [v][s][re,im]. */
void or_sq_rho(int64_t n, const double* x, const double* y, const double* z,
               const int32_t* itype, int nsp, const double box[3], int64_t nvec,
               const int32_t* kv, double* rho);

/* Reference sum set (octant below qmin + axis shells above, sq.cuf:58-126,
 * 372-454, 523-591) for one frame, added into sqf[0..nqmax] and
 * sqfp[s*(nqmax+1) + shell].  bsc: scattering lengths per species. */
void or_sq_frame(int64_t n, const double* x, const double* y, const double* z,
                 const int32_t* itype, int ndim, int nsp, const double box[3],
                 const or_sqgrid* g, const double* bsc, double* sqf, double* sqfp);

/* or_sq_rho with FP64 phases (sincos in double), OpenMP over the vectors: the
 * CPU-baseline timing port, not a checker. */
void or_sq_rho_f64(int64_t n, const double* x, const double* y, const double* z,
                   const int32_t* itype, int nsp, const double box[3], int64_t nvec,
                   const int32_t* kv, double* rho);

/* or_sq_frame with FP64 phases: sincos of 2 pi frac(n.(r/L)) in double, OpenMP
 * over the vectors.  The CPU-baseline timing port of the reference's sqfnD /
 * sqfact sums (bench.py), not a checker. */
void or_sq_frame_f64(int64_t n, const double* x, const double* y, const double* z,
                     const int32_t* itype, int ndim, int nsp, const double box[3],
                     const or_sqgrid* g, const double* bsc, double* sqf, double* sqfp);

/* Charge-weighted sums over the same sum set as or_sq_frame (sq.cuf:424-432,
 * 563-582): sqfq[shell] += |sum_j q_j exp(i 2 pi n.(r_j/L))|^2, long double
 * phases and sums; positions already wrapped. */
void or_sq_frame_q(int64_t n, const double* x, const double* y, const double* z, const double* q,
                   int ndim, const double box[3], const or_sqgrid* g, double* sqfq);

/* log.f90:317-401: printed columns.  Rows i=1..nqmax.  ncols_out = 3 (Q, S_NN,
 * n) or 7 (Q, S_NN, S_cc, S_11, S_22, S_12, n) for a binary mixture with all
 * bsc == 1 (log.f90:331,348,380-389); mix has nsp columns S_jj. */
int or_sq_columns(int nsp, int nqmax, float dq, int64_t natoms, const int64_t* ntype,
                  const double* bsc, int64_t nconf, const double* sqf, const double* sqfp,
                  const int64_t* nq, double* cols, double* mix);

/* Confined slices (twoDstruc_3D).  or_slice_of: FP32 nearest slice, -1 when
 * outside zgrid/2 (rdf.cuf:692-696).  or_rdf_xy_brute: per-slice xy pair
 * histogram, bin int(r/dr)+1, counts without the species-sorted last atom
 * (rdf.cuf:672-736).  or_rdf_xy_accumulate: per-frame normalisation by the
 * slice's species counts (rdf.cuf:284-296).  or_rdf_xy_gmix: gxy tables
 * (log.f90:494-496, 511-514, 545-551).  or_sq_frame_xy: sqf3xy/sqf2xy
 * per-slice sums divided per vector by the slice counts (sq.cuf:596-771). */
int or_slice_of(float z, int nslice, const float* zslice, float zgrid);
void or_rdf_xy_brute(int64_t n, const double* x, const double* y, const double* z,
                     const int32_t* itype, int nsp, const double lxy[2], double side2, double deltar,
                     int lsmax, int nslice, const float* zslice, float zgrid, int64_t* counts,
                     int64_t* count, int64_t* countp, const double* q, double scale, int64_t* fixed);
void or_rdf_xy_accumulate(int nsp, int lsmax, int nslice, const int64_t* counts, const int64_t* countp,
                          double* hxy);
void or_rdf_xy_gmix(int nsp, int lsmax, int nslice, double deltar, const double box[3], double zgrid,
                    int64_t nconf, const double* hxy, double* out);
void or_sq_frame_xy(int64_t n, const double* x, const double* y, const double* z, const int32_t* itype,
                    const double* q, int nsp, const double box[3], const or_sqgrid* g, const double* bsc,
                    int nslice, const float* zslice, float zgrid, double* sqf, double* sqfp, double* sqfq);

/* s2D3_sh / s2D2_sh for given unit points: int64 sums of the cumulative
 * counts N_k and N_k^2 over points (rdf.cuf:489-615). */
void or_s2n(int64_t np, const double* pts, const double box[3], int64_t n, const double* x, const double* y,
            const double* z, int ndim, const double lmi[3], double side2, double deltar, int lsmax, int64_t* gi,
            int64_t* g2i);

/* F(Q,t): the S(Q,w) q-set and FqtnD over rtcorr's origins (sq.cuf:234-282,
 * dynamics.cuf:157-245, 421-508). */
int64_t or_fqt_vectors(int ndim, const double box[3], int nqw, const double* qw, int64_t cap, int32_t* kv,
                       int32_t* bin);
void or_fqt(int ndim, int nsp, const double* bsc, const double box[3], int nqw, const double* qw, int nbuffer,
            int jump, int ndyn, int itlimit, int64_t nframes, int64_t n, const double* pos, const int32_t* itype,
            double* tot, double* tots, int64_t* count_org);

#ifdef __cplusplus
}
#endif
#endif
