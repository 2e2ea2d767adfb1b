/*
 * radchan.h -- C ABI of the MI355X radiative-transfer library
 * (narrow-band correlated-k OERM Monte Carlo, a drop-in for RadChan's
 * src/montecarlo path on one GPU).
 *
 * Two groups of entry points:
 *
 *  1. The reference's own entry points, same symbols and by-reference
 *     signatures, called from Fortran through implicit F77 interfaces
 *     (src/dns/channel.f:242,251,255 -> gfortran mangles MC_GPU/GET_RESULTS
 *     to mc_gpu_/get_results_).  Parameters come from rc_set_params (defaults:
 *     the values of src/param/param.h and src/param/NarrowBand.h); tables are
 *     read from "<table dir>/" (default "properties", as read_tables.cpp:26,45,68).
 *
 *  2. Extensions (no counterpart in the reference): runtime parameters, error
 *     query, explicit tables, a device-resident session API for repeated
 *     solves and batches of temperature snapshots, and a per-ray trace used by
 *     the parity tests.
 *
 * All functions are host functions.  Integer returns: 0 on success, a
 * negative code on failure with a message available from rc_last_error().
 * The two reference entry points return void; on failure they print the
 * message to stderr and leave the outputs untouched (the reference aborts
 * the process, param.h:102-124), and rc_last_error() reports it.
 */
#ifndef RADCHAN_H
#define RADCHAN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RC_MAXLEV 9

/* Runtime form of src/param/param.h + src/param/NarrowBand.h. */
typedef struct rc_params {
    int32_t imax, jmax, kmax;        /* param.h:10-12, global sizes            */
    int32_t p_col;                   /* param.h:9, z slabs (nodes); 1 = whole  */
    float   Lx, Ly, Lz;              /* param.h:13-15                          */
    int32_t nVar;                    /* param.h:18                             */
    double  fact_mesh;               /* param.h:16                             */
    double  chanHeight;              /* param.h:20                             */
    int32_t photmax;                 /* param.h:19                             */
    int32_t grid_num;                /* param.h:24                             */
    int32_t srt, grey, calcwalls;    /* param.h:57-61                          */
    int32_t grey_true;               /* grey only: 1 band, Ib = sigma T^4/pi   */
    int32_t maxi[RC_MAXLEV], maxj[RC_MAXLEV], maxk[RC_MAXLEV]; /* param.h:27-29 */
    int32_t maxs[RC_MAXLEV];         /* param.h:31-49                          */
    int32_t bd[6];                   /* W,E,S,N,B,T faces, param.h:95-100      */
    int32_t adiabw, adiabe;          /* param.h:82-83                          */
    float   eps[6];                  /* param.h:68-73                          */
    float   Tw[6];                   /* param.h:75-80                          */
    float   toll, tollS;             /* param.h:86-87                          */
    int32_t nT, nQ, nB;              /* NarrowBand.h:10-12 (grey :35-37)       */
    float   abscoeff, Tmins, wvmax;  /* NarrowBand.h:38-40                     */
    uint64_t seed;                   /* Philox key (no reference counterpart)  */
    int32_t sampler;                 /* fluid directions, param.h:62-63 random: */
                                     /* 1 Sobol points (digitally shifted per   */
                                     /* cell and run), 0 Philox                 */
    int32_t fp64;                    /* param.h:64-65 floatpoint: 1 double      */
                                     /* precision (the full-FP64 kernels, plain */
                                     /* variant only), 0 single precision       */
} rc_params;

/* Spectral tables in the reference's original band order, as read by readT
 * (src/montecarlo/read_tables.cpp:21-95).  Caller-owned arrays:
 * Tnb[nT], kP[nT] (scaled by chanHeight), wvl/wvr/wvc/kavg[nB],
 * kq[nB][nT][nQ] (scaled), prob[nT][nB], probg[nT][nB][nQ], wq[nQ]. */
typedef struct rc_tables {
    int32_t nB, nT, nQ, pad;
    float *Tnb, *kP;
    float *wvl, *wvr, *wvc, *kavg;
    float *kq, *prob, *probg;
    float *wq;
} rc_tables;

/* Per-ray record of rc_trace (layout shared with the CPU oracle). */
typedef struct rc_ray {
    int32_t nb, g;              /* band (sorted rank order) and g-point       */
    int32_t ic, jc, kc, level;  /* cell and grid level where the ray ended     */
    int32_t steps, status;      /* march steps; end status (RC_RAY_*)          */
    uint32_t path_hash;         /* FNV-1a over every visited (ic, jc, kc)      */
    float sx, sy, sz;           /* emission direction                          */
    float tra;                  /* transmissivity at the end                   */
    float de;                   /* FP32 tally of the ray                       */
    double de64;                /* unused by the GPU (0)                       */
    double sumabs;              /* unused by the GPU (0)                       */
} rc_ray;

#define RC_RAY_ABSORBED 0
#define RC_RAY_WALL     1
#define RC_RAY_LOST     2
#define RC_RAY_CAPPED   3
#define RC_RAY_STOPPED  4

/* Solver variants */
#define RC_VARIANT_PLAIN   0   /* thread per emitting cell, FP32 per-cell sums  */
#define RC_VARIANT_BATCHED 1   /* persistent wavefront work queue, FP64 tallies */
#define RC_VARIANT_PLAIN_BANDS 2 /* the plain variant launched once per sorted  */
                                 /* band and run (srt=1 narrow band only): the  */
                                 /* same sums as RC_VARIANT_PLAIN               */

typedef struct rc_stats {
    double   kernel_ms;         /* device time of the last run (HIP events)   */
    double   march_ms;          /* device time of the march kernels alone     */
                                /* (band sort and march, not finalize); the   */
                                /* whole graph when replayed as a hipGraph    */
    int64_t  rays;              /* rays marched in the last run               */
    int64_t  steps;             /* march steps executed in the last run       */
    int64_t  capped;            /* rays that hit the last level's step cap    */
    int32_t  launches;          /* kernel launches in the last run            */
    int32_t  variant;
    /* batched variant: lane slots of the bulk march (64 per sub-step of a */
    /* wavefront, live ray or not) and the steps of live rays among them;  */
    /* their ratio is the active-lane fraction of the bulk march, and      */
    /* steps - bulk_steps are the steps of the drain                       */
    int64_t  bulk_slots;
    int64_t  bulk_steps;
} rc_stats;

/* ---- 1. reference entry points --------------------------------------- */

/* src/montecarlo/mc_gpu_noint.cu:660  extern "C" void mc_gpu_(myfloatF *Tfort,
 * int *zstart, int *rank).  Tfort: (imax+2)(jmax+2)(kmax+2) temperatures in
 * Kelvin, i fastest (idx_T, param.h:132).  Copies the field, launches the
 * solve asynchronously and returns. */
void mc_gpu_(float *Tfort, int *zstart, int *rank);

/* src/montecarlo/mc_gpu_noint.cu:876  extern "C" void get_results_(myfloatF
 * resfort[], myfloatF varfort[], int *rank).  Waits for the solve started by
 * mc_gpu_ and scatters div(q_r) and sqrt(variance) to the Fortran layout
 * idx_F (param.h:138, k local to the slab), size (imax+2)(jmax+2)(kmax/p_col+2).
 * Planes i=0 and i=imax+1 receive the wall fluxes when calcwalls==1 (else 0). */
void get_results_(float *resfort, float *varfort, int *rank);

/* ---- 2. extensions ----------------------------------------------------- */

void        rc_default_params(rc_params *p);
/* fill p from the values of the reference's compile-time macros, passed by
 * position (param.h/NarrowBand.h macro names equal the field names, so a
 * file that includes those headers cannot name the fields); used by
 * radchan_param_shim.h */
int         rc_params_compiled(rc_params *p, int imax_, int jmax_, int kmax_, int p_col_, float Lx_, float Ly_,
                               float Lz_, double fact_mesh_, int nVar_, int photmax_, double chanHeight_,
                               int grid_num_, const int *maxi_, const int *maxj_, const int *maxk_,
                               const int *maxs_, int srt_, int grey_, int calcwalls_, const float *eps6,
                               const float *Tw6, int adiabw_, int adiabe_, float toll_, float tollS_,
                               const int *bd6, int nT_, int nQ_, int nB_, float abscoeff_, float Tmins_,
                               float wvmax_);
int         rc_set_params(const rc_params *p);       /* used by mc_gpu_        */
int         rc_get_params(rc_params *p);
int         rc_set_table_dir(const char *dir);       /* default "properties"   */
int         rc_set_variant(int variant);             /* used by mc_gpu_        */
/* options of the mc_gpu_ session (no counterpart in the reference):
 * rc_set_estimator and rc_set_restriction modes, both 0 by default */
int         rc_set_dropin_modes(int estimator, int restriction);
const char *rc_last_error(void);
const char *rc_version(void);

/* tables: reader and writer of the reference text formats, grey synthesis */
int rc_read_tables(const char *dir, double chanHeight, rc_tables *t);
int rc_write_tables(const char *dir, const rc_tables *t, double chanHeight);
int rc_grey_tables(const rc_params *p, float Tmax, rc_tables *t);
/* sort_idx (mc_gpu_noint.cu:1278): idx[rank] = band, ascending kavg, ties by index */
int rc_sort_idx(const float *kavg, int nB, int32_t *idx);
/* correlated-k distributions of line-by-line spectra (correlatedKT,
 * nbcorrk/correlatedK.cpp:144-225) and their band means k_avg (calcProbT,
 * nbcorrk/calc_prob.cpp:52-56), on the GPU.  Spectrum s holds the points
 * off[s]..off[s+1]-1 of wvn/kvn (1..8192 points, wavenumbers ascending) and
 * the band edges wvl[s] < wvr[s]; n_pwrk bins of k^pwr (20000 and 0.1 in the
 * generator, param.h:25 and correlatedK.cpp:147); gq[nq] nondecreasing in
 * [0, 1].  Outputs: kq[nspec][nq], kavg, kmin, kmax per spectrum (host
 * arrays; the call is synchronous).  kernel_ms (may be NULL) receives the
 * kernel's device time. */
int rc_ckdist(int nspec, const int64_t *off, const double *wvn, const double *kvn, const double *wvl,
              const double *wvr, int n_pwrk, double pwr, int nq, const double *gq, double *kq,
              double *kavg, double *kmin, double *kmax, double *kernel_ms);
/* the generator's table step, calcProbT (nbcorrk/calc_prob.cpp:17-106),
 * on rc_ckdist's output: kq[nT][nB][nQ] and kavg[nT][nB] (FP64, row-major
 * by temperature T[nT], increasing), band edges wvl/wvr[nB], weights
 * wq[nQ].  Fills t (sizes nB, nT, nQ; caller-owned arrays) as
 * rc_read_tables leaves them: FP32, k values times chanHeight, kavg of the
 * last row, prob = band CDF, probg = g CDF, kP = the correlated-k Planck
 * mean (also in place of the line-by-line one the reference writes there). */
int rc_ck_tables(int nB, int nT, int nQ, const double *T, const double *wvl, const double *wvr,
                 const double *kq, const double *kavg, const double *wq, double chanHeight,
                 rc_tables *t);
/* The generator's front end on a line-by-line spectrum in file order (wv
 * in cm^-1 increasing, kv in cm^-1), in place of absWideBand
 * (nbcorrk/read_files.cpp:22-73) and the narrow-band split of its main
 * (nbcorrk/ckdist_hitran_part.cpp:55-114): wide bands where kv > toll, with
 * weak stretches of up to wvtoll (counted dwv per weak point) kept inside,
 * each split into (int)(width / deltawv[min(i/3, 2)]) (>= 1) narrow bands by
 * point count.  Narrow band b is points first[b] .. first[b]+count[b]-1 of
 * the input, with edges wv_left/wv_right[b] and centre wvc[b] (those three
 * may be NULL); at most max_bands are written.  Returns the number of
 * narrow bands (< 0 on error); nwide (may be NULL) gets the wide ones.
 * The reference's constants: toll 1e-6, wvtoll 25000, dwv 0.01,
 * deltawv {10, 10, 10} (nbcorrk/param.h:18-23). */
int rc_lbl_bands(int64_t n, const double *wv, const double *kv, double toll, double wvtoll, double dwv,
                 const double *deltawv, int max_bands, int64_t *first, int64_t *count, double *wv_left,
                 double *wv_right, double *wvc, int32_t *nwide);
/* readFillT (nbcorrk/read_files.cpp:76-114) for one temperature T [K]: the
 * spectrum (wv, kv, n points in file order) fills kvn[ntot], the narrow
 * bands' absorption coefficients in m^-1 (kv times 100), at the points whose
 * wavenumber equals the next one awaited in wvn[ntot] (the bands' points in
 * order, e.g. the reference spectrum's wv over the bands' ranges); *kPr gets
 * the line-by-line Planck mean in m^-1 accumulated over the points read
 * until the last band is full.  Returns the points filled (ntot when every
 * awaited wavenumber was found; < 0 on error); kvn past them is untouched. */
int64_t rc_lbl_fill(int64_t n, const double *wv, const double *kv, double T, double dwv, int64_t ntot,
                    const double *wvn, double *kvn, double *kPr);
/* interp3D (src/montecarlo/spline.cpp:36-177; not called by the
 * reference, mc_gpu_noint.cu:694): the field T ((imax+2)(jmax+2)(kmax+2)
 * host floats) restricted to grid level `level` (1..grid_num-1) by natural
 * cubic splines along x, y, z; Tc [km][jm][im] the cell temperatures, Tw
 * (may be NULL) [2][km][jm] the west/east x-face temperatures (the spline
 * of the face means; interp3D's coarse ghost is 2 Tw - Tc).  Uses the GPU;
 * synchronous. */
int rc_spline_restrict(const rc_params *p, int level, const float *T, float *Tc, float *Tw);
/* No counterpart in the reference: rc_spline_restrict for callers that
 * cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  T, Tc and Tw (may be NULL) are
 * DEVICE arrays of the sizes above; asynchronous on `stream` (a
 * hipStream_t; NULL = the null stream); calls no hipMalloc/hipFree.  The
 * grid and the spline's working space live in one fixed device array of
 * RC_SPLINE_NOALLOC_FLOATS floats (64 MiB, held by every GPU the library
 * runs on until the process ends); a field that needs more (about 2.7 times
 * the fine interior at level 1, e.g. 192x128x192 fits, 256x128x256 does
 * not) is refused with an error.  Calls that overlap on one GPU share that
 * array and race: order them on one stream. */
#define RC_SPLINE_NOALLOC_FLOATS (16u << 20)
int rc_spline_restrict_noalloc(const rc_params *p, int level, const float *T, float *Tc, float *Tw,
                               void *stream);
/* rc_ckdist on device arrays (spectra already in HBM, e.g. written there by
 * another kernel), asynchronous on `stream` (a hipStream_t; NULL = the null
 * stream), nothing allocated.  max_points bounds the points of every spectrum (1..8192; it
 * picks the kernel instance); a spectrum with more, or with none, gets NaN
 * outputs.  gq must be nondecreasing in [0, 1] (not checked here). */
int rc_ckdist_device(int nspec, const int64_t *off, const double *wvn, const double *kvn,
                     const double *wvl, const double *wvr, int n_pwrk, double pwr, int nq,
                     const double *gq, double *kq, double *kavg, double *kmin, double *kmax,
                     int max_points, void *stream);
/* Gridn::mk_grid (definitions.h:68-136) of one level */
int rc_grid(const rc_params *p, int level, float *x, float *xu, float *y,
            float *yu, float *z, float *zu);

/* session: device-resident solver state */
typedef struct rc_ctx rc_ctx;
rc_ctx *rc_create(const rc_params *p);
void    rc_destroy(rc_ctx *c);
/* tables for a narrow-band run (ignored for grey runs, which synthesise
 * their tables from each field's Tmax) */
int rc_load_tables(rc_ctx *c, const rc_tables *t);
/* nsnap temperature snapshots of the same shape, host arrays of
 * (imax+2)(jmax+2)(kmax+2) floats each; copied to the device */
int rc_set_fields(rc_ctx *c, int nsnap, const float *const *T, int zstart);
/* launch the solve of every loaded snapshot (asynchronous) */
int rc_run(rc_ctx *c, int variant);
/* wait for the last run; fills st (may be NULL) */
int rc_sync(rc_ctx *c, rc_stats *st);
/* per emitting cell of snapshot s (order: i fastest, then j, then local k):
 * mean over nVar and the reference's sqrt-variance (mc_gpu_noint.cu:234-239) */
int rc_results(rc_ctx *c, int s, float *mean, float *sdev);
/* per (emitting cell, run) FP64 tallies of snapshot s, [cell][nVar] */
int rc_tallies(rc_ctx *c, int s, double *de);
/* wall results of snapshot s: [bound][k][j] means and sdev, and tallies
 * [bound][k][j][nVar]; bound 0 = west (i=0), 1 = east (i=imax+1) */
int rc_surface_results(rc_ctx *c, int s, float *mean, float *sdev, double *de);
/* trace single rays of snapshot 0: fluid rays given as (i,j,k,v,h) with k
 * global, surface rays as (bound,j,k,v,h); one record per ray */
int rc_trace(rc_ctx *c, int surface, const int32_t *rays, int n, rc_ray *out);
/* scalar set-up values of snapshot s: Tmax, kappamax, Ibmax, Ibmax_w,
 * Ibmax_e; ints tm, tstar */
int rc_info(rc_ctx *c, int s, float *f5, int32_t *i2);
/* schedule of the batched solver: persistent workgroups (one wavefront
 * each) per compute unit, the number of idle lanes that triggers a
 * refill, and march steps per loop iteration; 0 selects the default.
 * Chosen by the autotuner. */
int rc_set_schedule(rc_ctx *c, int wg_per_cu, int threshold, int steps_per_iter);
int rc_get_schedule(rc_ctx *c, int32_t *wg_per_cu, int32_t *threshold, int32_t *steps_per_iter);
/* refills of the batched solver (no counterpart in the reference): 0 =
 * once per loop iteration (rays that end on an earlier step wait for it),
 * 1 = also between the iteration's march steps, -1 = the default (1 for a
 * single grid, 0 for several levels).  Defaults of the schedule with it:
 * (32, 4, 6) single grid; (20, 4, maxs[0]) several levels, the finest
 * level's step cap (clamped to 1..16), so that rays picked up together
 * hand off to the next level together.  Chosen by the autotuner;
 * rc_get_refill stores the mode set (-1: default) in *mode. */
int rc_set_refill(rc_ctx *c, int mode);
int rc_get_refill(rc_ctx *c, int32_t *mode);
/* storage of the k tables on the device: 0 = FP32 (default), 1 = BF16 k
 * pairs with a power-of-two scale per band (Ib stays FP32); rebuilds the
 * loaded tables */
int rc_set_table_precision(rc_ctx *c, int bf16);
/* record the batched solve (queue reset, tally reset, persistent kernel,
 * finalize) once as a hipGraph and replay it on later rc_run calls; the
 * graph is dropped when fields or the schedule change */
int rc_set_graph(rc_ctx *c, int on);
/* temperature of the coarse grid levels: 0 = the texture mean of the fine
 * field at the mapped cell centre, as the reference's kernels read it
 * (mc_gpu_noint.cu:334-338; default); 1 = interp3D's cubic-spline
 * restriction (spline.cpp:36-177, see rc_spline_restrict) for the cells and
 * the adiabatic x walls (bd=4 z walls keep the texture mean).  Applies to
 * fields loaded by later rc_set_fields calls. */
int rc_set_restriction(rc_ctx *c, int mode);
/* band and g pdf of the estimator: 0 = the reference's (CDFs of the table
 * row nearest Tmax, kappa_max = the tabulated Planck mean at Tmax,
 * mc_gpu_noint.cu:1110-1131, read_tables.cpp:35-38; default); 1 = exact:
 * pdf and normalisation from the tables interpolated at Tmax itself (p(b,g)
 * proportional to K_bg(Tmax) B_b(Tmax) w_g dnu_b, kappa_max = pi Z /
 * (sigma Tmax^4); walls: p(b) from B_b(Tmax) dnu_b and the banded emission
 * eps pi sum B dnu), unbiased at any Tmax and for any table normalisation.
 * Narrow-band runs only (grey runs are unaffected); applies to fields loaded
 * by later rc_set_fields calls. */
int rc_set_estimator(rc_ctx *c, int mode);
/* summation of the batched variant's tallies (no counterpart in the
 * reference): 0 = each ray's FP32 tally added to its unit's FP64 total as
 * the ray ends (LDS and FP64 atomics; the order, and so the last bits,
 * vary from run to run; default); 1 = deterministic: every ray's tally
 * stored at its own entry (4 bytes per ray in HBM: photmax nVar per
 * emitting cell), then summed per (cell, run) in a fixed order with
 * compensated FP64 arithmetic -- the same bits every run and for every
 * schedule.  photmax <= 65535.  The plain variant is always deterministic;
 * groups use mode 0. */
int rc_set_reduction(rc_ctx *c, int mode);
/* time the batched solve of the loaded snapshots over schedules (steps per
 * iteration 1..4 and, on several levels, the finest level's step cap, then
 * workgroups per CU, then refill threshold, each at the best of the
 * previous; then refills between the steps with steps per
 * iteration 2..8 at threshold 4 and, if they win, threshold 8) and keep
 * the fastest, refill mode included (rc_get_refill) */
int rc_autotune(rc_ctx *c, int32_t *best_wg, int32_t *best_th, int32_t *best_spi, float *best_ms);
/* groups (no counterpart in the reference): the loaded snapshots of
 * several sessions -- grids of different sizes -- solved by one batched
 * launch.  The sessions must agree in every parameter but the grid (nVar,
 * photmax, srt, table sizes, tolerances, boundary codes, domain lengths,
 * seed, table precision, calcwalls).  With wall fluxes, the walls of every
 * snapshot are a second batched launch.  Each snapshot draws the Philox
 * streams it draws in its own session, and the results land in the
 * sessions: after rc_group_sync, rc_results, rc_tallies and
 * rc_surface_results of each session return them. */
typedef struct rc_group rc_group;
rc_group *rc_group_create(rc_ctx *const *sessions, int n);
void      rc_group_destroy(rc_group *g);
int       rc_group_set_schedule(rc_group *g, int wg_per_cu, int threshold, int steps_per_iter);
int       rc_group_set_refill(rc_group *g, int mode);
/* launch (asynchronous): one batched kernel over every snapshot (and one
 * over their walls) */
int       rc_group_run(rc_group *g);
/* wait for the last group run; fills st (may be NULL) */
int       rc_group_sync(rc_group *g, rc_stats *st);
int rc_device_synchronize(void);

#ifdef __cplusplus
}
#endif
#endif
