/*
 * rc_kernels.h -- device-side descriptors shared by the kernels and the
 * session code, and the kernel launchers.
 */
#ifndef RC_KERNELS_H
#define RC_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/radchan.h"

namespace rc {

/* Philox stream tags (4th counter word); the snapshot index is added as
 * 16 * snapshot so every snapshot draws its own streams */
enum { TAG_FLUID = 0, TAG_WEST = 1, TAG_EAST = 2, TAG_NOISE = 3 };

/* solve-wide constants, passed by value */
struct DevCommon {
    int imax, jmax, kmax, kloc, zstart;
    int nVar, photmax, srt, nB, nQ, nT, grid_num;
    int ncell, nsurf;              /* emitting fluid cells, wall cells per wall */
    float Lx, Ly, Lz, toll, tollS;
    int bd[6], adiab[2];
    uint32_t k0, k1;               /* Philox key */
    /* deterministic reduction (batched variant): every ray's FP32 tally at
     * ray_de[unit photmax + h], summed per unit afterwards in a fixed order
     * with compensation; nullptr: FP64 atomics, the default */
    float *ray_de;
    int sampler;                   /* fluid directions: 1 Sobol points shifted per (cell, run), 0 Philox */
    int refill;                    /* batched: 1 also refill idle lanes between an iteration's march steps, */
                                   /* 0 only once per iteration, < 0 the default (batched_schedule)      */
    int maxs0;                     /* step cap of the finest level (the default schedule's steps)        */
    /* batched variant, srt above 256 keys: the radix sort's scratch, 2
     * rows of photmax (rounded up to 64) keys per workgroup of the launch
     * (batched_srt_keys) */
    uint16_t *srt_keys;
};

/* the counter word h of a (cell, run)'s Sobol shift: no ray has it */
constexpr uint32_t SOBOL_SHIFT_H = 0xFFFFFFFFu;

/* Table rows: nT temperatures, one pad entry (the clamped top row needs
 * no branch), then two entries per face (W,E,S,N,B,T) holding the Ib of a
 * black wall of fixed temperature with k = 0, so that every black wall's
 * emission is a table lookup at a position: row stride tab_stride(nT),
 * fixed wall of face f at position nT + 1 + 2 f */
__host__ __device__ inline int tab_stride(int nT) { return nT + 13; }
__host__ __device__ inline float fixed_wall_pos(int nT, int face) { return (float)(nT + 1 + 2 * face); }

/* one grid level of one snapshot.  P holds table positions (temperature
 * mapped to the table's row coordinate and clamped, the row t and weight
 * a being (int)p and p - t): the cells [km][jm][im], then the black walls
 * -- x faces [2][km][jm] (adiabatic: the texture T at the wall; fixed T:
 * the face's fixed position), y faces [2][km][im] (fixed), z faces
 * [2][jm][im] (bd 4: the adjacent-cell texture T; fixed) -- see
 * wall_index */
struct DevLevel {
    int im, jm, km, maxs;
    float dy, dz;                  /* uniform y/z spacing: face j at j*dy */
    int ncell, xoff;               /* ncell = im jm km; xu = xu_all + xoff */
    const float *xu;               /* im+1 x faces                        */
    const float *P;                /* table positions, cells then walls   */
    int poff, pad1;                /* P = P_all + poff                    */
    /* coarsening to the next level (mc_gpu_noint.cu:215-223): the index
     * ratios along x, y, z and, per ratio r >= 2, a multiplier m with
     * (i + r - 1) / r == umulhi(i + r - 1, m) for every index of the
     * level (0: divide) */
    int cr[3];
    uint32_t cmul[3];
};

/* entries of a level's position array */
__host__ __device__ inline size_t level_pos_count(int im, int jm, int km) {
    return (size_t)im * jm * km + 2 * ((size_t)jm * km + (size_t)im * km + (size_t)im * jm);
}

/* one temperature snapshot (device memory) */
struct DevSnap {
    float Tmax, Ibmax, Ibmax_w, Ibmax_e;
    float Tnb0, invdT;
    int tstar, snap;
    const float2 *kb;              /* [(rank*nQ+g)*S+t] = {k, Ib}, S = tab_stride(nT) */
    const uint32_t *kp16;          /* BF16 mode: [(rank*nQ+g)*S+t] = k_t | k_t+1 << 16 */
    const float4 *bp;              /* BF16 mode: [rank*S+t] = {Ib_t, Ib_t+1, scale, 0}: */
                                   /* Ib pair and the band's power-of-two k scale      */
    int kbf16, pad3;               /* 1: k tables stored as BF16                      */
    const float *cdf_b;            /* [nB] fluid band CDF, row t*         */
    const float *cdf_g;            /* [nB][nQ] g CDFs, row t*             */
    const float *cdf_bs;           /* [nB] wall band CDF, row t*          */
    const float *cdf_gs;           /* [nQ] cumulative quadrature weights  */
    const int *rank_of;            /* [nB] sorted rank of each band       */
    const int *guide_b;            /* [gM+1] first index with cdf_b > m/gM */
    const int *guide_bs;           /* [gM+1] same for cdf_bs               */
    int gM, pad2;
    const float *T0;               /* ghost-padded input field            */
    const float *xu_all;           /* x faces of every level, level g at lev[g].xoff */
    const float *P_all;            /* table positions of every level, g at lev[g].poff */
    int nxu, npos;                 /* their lengths                       */
    const float *x0, *y0, *z0;     /* level-0 cell centres                */
    /* the full-FP64 path (floatpoint 1): the temperatures behind P_all (same
     * layout), the weights in double and the table's temperature map */
    const float *T_all;
    double Ibmax64, Ibmax_w64, Ibmax_e64, Tnb0d, invdTd;
    /* the snapshot's grid (snapshots of one batched launch may differ in
     * it): global sizes, first z plane and counts of emitting cells and
     * wall cells per wall; ubase / sbase: the launch's emitting (wall)
     * cells before this snapshot, so its (cell, run) tallies start at
     * ubase nVar (sbase nVar) */
    int imax, jmax, kmax, zstart;
    int ncell, nsurf, ubase, sbase;
    DevLevel lev[RC_MAXLEV];
};

/* table positions of level g (cells and black walls, layout of DevLevel.P)
 * from the ghost-padded field T0; the table's first temperature, inverse
 * spacing and row count map temperatures to positions */
void launch_level_P(hipStream_t st, const float *T0, const rc_params &p, int g, float Tnb0, float invdT, int nT,
                    float *P, float *Tout = nullptr);
/* cell and x-wall temperatures of level g (Tl [km][jm][im], Tw [2][km][jm]
 * or NULL) converted to positions in P (the spline restriction's output) */
void launch_pos_from_T(hipStream_t st, const rc_params &p, int g, const float *Tl, const float *Tw, float Tnb0,
                       float invdT, int nT, float *P, float *Tout = nullptr);
/* cubic-spline restriction of level g (interp3D): cells Tl [km][jm][im] and,
 * if Tw, x-wall faces [2][km][jm]; scratch of spline_scratch_floats floats */
size_t spline_scratch_floats(int imax, int jmax, int kmax, int im, int jm, int km);
void launch_spline_level(hipStream_t st, const float *T0, const float *x0, const float *y0, const float *z0,
                         const float *xc, const float *yc, const float *zc, int imax, int jmax, int kmax, int im,
                         int jm, int km, float *Tl, float *Tw, float *scratch);
void launch_sort(hipStream_t st, const DevCommon &cm, const DevSnap *snaps, int s, int v, int surface, int nunit,
                 const int32_t *ulist, uint16_t *keys);
void launch_fluid_plain(hipStream_t st, const DevCommon &cm, const DevSnap *snaps, int s, int v, const uint16_t *keys,
                        double *de, unsigned long long *stats);
/* the plain variant in one launch per sorted band (srt keys required):
 * FP32 cell sums in acc (ncell floats), stored to de[cell][v] */
void launch_fluid_bands(hipStream_t st, const DevCommon &cm, const DevSnap *snaps, int s, int v, const uint16_t *keys,
                        float *acc, double *de, unsigned long long *stats);
void launch_surface_plain(hipStream_t st, const DevCommon &cm, const DevSnap *snaps, int s, int v,
                          const uint16_t *keys, double *de, unsigned long long *stats);
/* the full-FP64 variant (floatpoint 1) in the plain launch shape: FP64
 * geometry, energy and cell sums; the snapshot's T_all must be set */
void launch_fluid_f64(hipStream_t st, const DevCommon &cm, const DevSnap *snaps, int s, int v, const uint16_t *keys,
                      double *de, unsigned long long *stats);
void launch_surface_f64(hipStream_t st, const DevCommon &cm, const DevSnap *snaps, int s, int v,
                        const uint16_t *keys, double *de, unsigned long long *stats);
void launch_trace(hipStream_t st, const DevCommon &cm, const DevSnap *snaps, int surface, const int32_t *rays,
                  const int32_t *unit, int n, const uint16_t *keys, rc_ray *out);
void launch_finalize(hipStream_t st, const double *de, int n, int nVar, float *mean, float *sdev);
/* de[u] = the compensated FP64 sum of ray[u photmax + h], h = 0..photmax-1,
 * in a fixed order (one wavefront per unit) */
void launch_reduce_rays(hipStream_t st, const float *ray, size_t nunit, int photmax, double *de);
/* workgroups of a batched launch over nunit_cells emitting (wall) cells,
 * and the 16-bit keys of radix-sort scratch it needs (0: none) */
int batched_grid(const DevCommon &cm, long nunit_cells, int wg_per_cu);
/* the batched schedule with its defaults filled in (values <= 0, refill
 * < 0: the defaults) */
void batched_schedule(const DevCommon &cm, int &wg_per_cu, int &threshold, int &spi, int &refill);
size_t batched_srt_keys(const DevCommon &cm, long nunit_cells, int wg_per_cu);
void launch_zero(hipStream_t st, double *p, size_t n);
void launch_zero_u64(hipStream_t st, unsigned long long *p, int n);
/* input check and Tmax of a temperature field on the device (out: 2 words) */
void launch_field_scan(hipStream_t st, const float *T, int imax, int jmax, int kmax, uint32_t *out);

}  // namespace rc

#endif
