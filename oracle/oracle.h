/*
 * oracle.h -- CPU restatement of the OERM Monte Carlo radiative solver of
 * RadChan (src/montecarlo), used ONLY as the checker for the HIP library
 * (tests, smoke check and the CPU-baseline leg of the benchmark).
 *
 * This is test infrastructure, not product code: nothing in the shipped
 * library links or calls it.
 *
 * Each function names the reference lines it restates.  Where the reference
 * has a defect (see the "quirk" notes), the oracle follows the corrected
 * semantics that the GPU library also implements.
 *
 * Numerics policy (shared with the HIP library, checked bit-for-bit by the
 * GPU parity tests): FP32 arithmetic with no contraction (-ffp-contract=off),
 * explicit fmaf where a fused op is wanted, correctly rounded division and
 * sqrt, and project-written exp/expm1/sincos(2*pi*u) built only from those
 * operations, so that CPU and GPU produce identical bits.  The full-FP64
 * mode (precision 1, the library's floatpoint 1) does the same in double,
 * with its own project-written expm1/exp and sin/cos of 2 pi u.
 */
#ifndef RADCHAN_ORACLE_H
#define RADCHAN_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAXLEV 9

/* Solver parameters: the runtime form of src/param/param.h and
 * src/param/NarrowBand.h.  The layout matches rc_params of the HIP library
 * (include/radchan.h) field by field; the Python tests fill one ctypes
 * structure for both. */
typedef struct orc_params {
    int32_t imax, jmax, kmax;        /* param.h:10-12 (global sizes)          */
    int32_t p_col;                   /* param.h:9  (z slabs; 1 = whole domain) */
    float   Lx, Ly, Lz;              /* param.h:13-15                          */
    int32_t nVar;                    /* param.h:18                             */
    double  fact_mesh;               /* param.h:16 (a double literal there)    */
    double  chanHeight;              /* param.h:20                             */
    int32_t photmax;                 /* param.h:19                             */
    int32_t grid_num;                /* param.h:24                             */
    int32_t srt, grey, calcwalls;    /* param.h:57-61                          */
    int32_t grey_true;               /* 1: one band, Ib = sigma T^4 / pi       */
    int32_t maxi[ORC_MAXLEV], maxj[ORC_MAXLEV], maxk[ORC_MAXLEV];  /* param.h:27-29 */
    int32_t maxs[ORC_MAXLEV];        /* param.h:31-49                          */
    int32_t bd[6];                   /* W,E,S,N,B,T  param.h:95-100            */
    int32_t adiabw, adiabe;          /* param.h:82-83                          */
    float   eps[6];                  /* param.h:68-73                          */
    float   Tw[6];                   /* param.h:75-80                          */
    float   toll, tollS;             /* param.h:86-87                          */
    int32_t nT, nQ, nB;              /* NarrowBand.h:10-12 / 35-37             */
    float   abscoeff, Tmins, wvmax;  /* NarrowBand.h:38-40 (grey block)        */
    uint64_t seed;                   /* Philox key                             */
    int32_t sampler;                 /* 1: Sobol fluid directions, 0: Philox   */
    int32_t fp64;                    /* the library's floatpoint switch; the   */
                                     /* oracle takes the precision per call    */
} orc_params;

/* Spectral tables as read by readT (read_tables.cpp:21-95).  Arrays are in
 * the ORIGINAL band order; kq is [nB][nT][nQ], prob [nT][nB], probg
 * [nT][nB][nQ]; wq holds the nQ quadrature weights (NarrowBand.h:21-27). */
typedef struct orc_tables {
    int32_t nB, nT, nQ, pad;
    float *Tnb, *kP;
    float *wvl, *wvr, *wvc, *kavg;
    float *kq, *prob, *probg;
    float *wq;
} orc_tables;

/* One traced ray (for per-ray parity with the GPU trace entry point). */
typedef struct orc_ray {
    int32_t nb, g;             /* band in sorted (rank) order, g-point        */
    int32_t ic, jc, kc, level; /* indices and level when the ray ended        */
    int32_t steps, status;     /* executed march steps; how it ended          */
    uint32_t path_hash;        /* FNV-1a over every visited (ic,jc,kc)        */
    float sx, sy, sz;          /* emission direction                          */
    float tra;                 /* transmissivity when the ray ended           */
    float de;                  /* FP32 tally of this ray                      */
    double de64;               /* FP64-energy tally on the same FP32 geometry */
    double sumabs;             /* sum over terms of |absorbed| + |emitted|    */
                               /* parts (FP64): the scale of the tally before */
                               /* the reciprocity difference cancels          */
} orc_ray;

/* ray end status */
#define ORC_ABSORBED 0   /* transmissivity fell under toll, rest dumped  */
#define ORC_WALL     1   /* black wall (bd 2 or 4)                        */
#define ORC_LOST     2   /* lost-ray wall (bd 5)                          */
#define ORC_CAPPED   3   /* last level's maxs reached; energy dropped     */
#define ORC_STOPPED  4   /* tra <= tol without a dump (tra == tol edge)   */

typedef struct orc_info {
    float Tmax, kappamax, Ibmax, Ibmax_w, Ibmax_e;
    int32_t tm, tstar;
    int32_t n_emit;            /* emitting fluid cells (imax*jmax*kmax/p_col) */
    double kappamax64, Ibmax64;
} orc_info;

typedef struct orc_ctx orc_ctx;

/* --- building blocks ---------------------------------------------------- */
void     orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
float    orc_u01(uint32_t x);
void     orc_sincos2pi(float u, float *s, float *c);
void     orc_expm1_exp(float x, float *em1, float *e);
/* the full-FP64 path's project math: sin/cos of 2 pi u (u given in single
 * precision) and expm1/exp for x <= 0 */
void     orc_sincos2pi_d(float u, double *s, double *c);
void     orc_expm1_exp_d(double x, double *em1, double *e);
float    orc_planck(float T, float nu);
int      orc_upper_bound(const float *cdf, int n, float R);
void     orc_sort_idx(const float *kavg, int nB, int32_t *idx);
void     orc_sobol2(uint32_t h, uint32_t x[2]);
void     orc_sobol_uniforms(uint64_t seed, uint32_t h, uint32_t v, uint32_t cell, uint32_t tag, float u[2]);
void     orc_iso_dir(float u0, float u1, float s[3]);
void     orc_lamb_dir(float u0, float u1, int bound, float s[3]);
int      orc_grid(const orc_params *p, int level, float *x, float *xu,
                  float *y, float *yu, float *z, float *zu);
float    orc_tmax(const orc_params *p, const float *T);

/* --- tables ------------------------------------------------------------- */
int      orc_read_tables(const char *dir, double chanHeight, orc_tables *t);
int      orc_grey_tables(const orc_params *p, float Tmax, orc_tables *t);

/* --- solver ------------------------------------------------------------- */
orc_ctx *orc_create(const orc_params *p, const orc_tables *t, const float *T,
                    int zstart, char *err, int errlen);
void     orc_destroy(orc_ctx *c);
void     orc_set_snapshot(orc_ctx *c, int s);
/* test-only: plant mistake `id` (see PLANT_* in oracle.cpp; 0 = none) */
void     orc_plant(int id);
/* the linear texture fetch of the temperature volume at texel coordinates */
float    orc_tex3(const orc_ctx *c, float tx, float ty, float tz);
/* temperatures read on level g: cells [km][jm][im], adiabatic x walls
 * [2][km][jm], bd=4 z walls [2][jm][im] (either of the last two may be NULL) */
int      orc_level_T(const orc_ctx *c, int g, float *Tc, float *Tw, float *Tz);
void     orc_get_info(const orc_ctx *c, orc_info *info);
/* band sequence of one (cell, run): photmax (rank, g) pairs in ray order */
void     orc_band_sequence(const orc_ctx *c, int surface, int bound, int i,
                           int j, int k, int v, int32_t *nb, int32_t *g);
void     orc_trace_fluid(const orc_ctx *c, int i, int j, int k, int v, int h, orc_ray *r);
void     orc_trace_surface(const orc_ctx *c, int bound, int j, int k, int v, int h, orc_ray *r);
/* Per (cell, run) tallies for a list of cells given as (i,j,k) triples
 * (k global).  precision 0: FP32 geometry, outputs de32 (FP32 sum in ray
 * order), de32d (FP64 sum of the per-ray FP32 tallies, may be NULL) and
 * de64 (FP64 energy on the same geometry); precision 1: full FP64 (de32 and
 * de32d receive the FP64 result).  steps receives the march steps per cell.
 * Runs on nthreads OpenMP threads (0 = runtime default). */
int      orc_solve_fluid(const orc_ctx *c, int precision, const int32_t *cells,
                         int ncells, float *de32, double *de32d, double *de64,
                         int64_t *steps, int nthreads);
/* orc_solve_fluid with every ray's tally added to its band group: group
 * of a ray = group_of_rank[sorted band rank]; out receives, per cell, run
 * and group, the FP64 sum of the per-ray tallies ([cell][run][group]) --
 * the FP32 tallies (precision 0) or the full FP64 ones (precision 1) */
int      orc_solve_fluid_groups(const orc_ctx *c, int precision, const int32_t *cells,
                                int ncells, const int32_t *group_of_rank, int ngroups,
                                double *out, int nthreads);
/* walls: cells are (bound, j, k) triples */
int      orc_solve_surface(const orc_ctx *c, int precision, const int32_t *cells,
                           int ncells, float *de32, double *de32d, double *de64,
                           int64_t *steps, int nthreads);

/* Correlated-k distributions of narrow bands from line-by-line spectra
 * (nbcorrk/correlatedK.cpp:144-225 and the k_avg sum of calc_prob.cpp:52-56),
 * one band of one temperature per spectrum s: points off[s] .. off[s+1]-1
 * of (wvn [cm^-1], kvn), band edges wvl[s], wvr[s].  The k range of the
 * band is cut into n_pwrk intervals of equal width in k^pwr; each pair of
 * neighbouring points adds its wavenumber width (over the band width) to
 * the interval of its mean k; the cumulative g is summed down from 1 at
 * the top interval; k at the quadrature points gq is interpolated
 * linearly in g between interval centres.  Outputs per spectrum:
 * kq[s][nq], kavg[s] (sum of interval centre k times its fraction),
 * kmin[s], kmax[s].  nthreads OpenMP threads (0 = runtime default). */
int      orc_ckdist(int nspec, const int64_t *off, const double *wvn, const double *kvn,
                    const double *wvl, const double *wvr, int n_pwrk, double pwr, int nq,
                    const double *gq, double *kq, double *kavg, double *kmin, double *kmax,
                    int nthreads);
/* interp3D (src/montecarlo/spline.cpp:36-177) of one coarse level: the
 * field T0 ((imax+2)(jmax+2)(kmax+2), idx_T layout) at the coarse cell
 * centres by natural cubic splines along x, then y, then z; the x walls
 * from the splines of the face means.  x0/y0/z0 and xc/yc/zc are the fine
 * and coarse cell centres (1-based, Gridn::mk_grid's x, y, z).  Outputs:
 * Tc [km][jm][im], Tw [2][km][jm] (may be NULL). */
int      orc_interp3d(int imax, int jmax, int kmax, const float *x0, const float *y0,
                      const float *z0, const float *T0, int im, int jm, int km,
                      const float *xc, const float *yc, const float *zc, float *Tc, float *Tw);
/* spline / splint (spline.cpp:180-226) on one line, 0-based: y2[n] of the
 * natural spline through (x, y) and its values ys[m] at xs */
int      orc_spline(const float *x, const float *y, int n, float *y2, const float *xs, int m,
                    float *ys);
/* coarse-level temperatures of a context: 0 = texture mean (default),
 * 1 = interp3D for cells and adiabatic x walls */
int      orc_set_restriction(orc_ctx *c, int mode);
/* the estimator's band/g pdf: 0 = the reference's nearest-row CDFs and
 * tabulated kP (default), 1 = exact: pdf and normalisation from the tables
 * interpolated at Tmax (narrow band only) */
int      orc_set_estimator(orc_ctx *c, int mode);
/* The generator's table step, calcProbT (nbcorrk/calc_prob.cpp:17-106),
 * after orc_ckdist's distributions: spectra row-major (s = i*nB + j, band j
 * at temperature T[i]).  Outputs: kq[nT][nB][nq], kavg[nT][nB], kP[nT] (the
 * correlated-k Planck mean), cuprob[nT][nB] (band CDF), cuprob_g[nT][nB][nq]
 * (g CDF per band).  Sums in the reference's order, per bin. */
int      orc_ck_tables(int nB, int nT, const int64_t *off, const double *wvn, const double *kvn,
                       const double *wvl, const double *wvr, const double *T, int n_pwrk,
                       double pwr, int nq, const double *gq, const double *wq, double *kq,
                       double *kavg, double *kP, double *cuprob, double *cuprob_g, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
