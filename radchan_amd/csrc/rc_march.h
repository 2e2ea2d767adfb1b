/*
 * rc_march.h -- the per-ray device code shared by the plain and batched
 * solvers: band draws, ray set-up and the multigrid DDA march.
 *
 * Reference: src/montecarlo/mc_gpu_noint.cu kernel_fluid :115-224,
 * march_ray :245-658, find_band :1103-1169, kernel_surface :1381-1462.
 */
#ifndef RC_MARCH_H
#define RC_MARCH_H

#include <hip/hip_runtime.h>

#include "rc_device.h"
#include "rc_internal.h"
#include "rc_kernels.h"

namespace rc {

using namespace rcd;

/* ------------------------------------------------------------------------ */
/* band draws and the srt enumeration                                        */

/* (rank, g) key of one band draw: band from the Planck-weighted band CDF of
 * the row nearest Tmax, g from that band's g CDF (find_band :1103-1169) */
__device__ __forceinline__ uint32_t band_key(const DevCommon &cm, const DevSnap &sn, bool surface, float u2, float u3) {
    const int nb = surface ? upper_bound_guided(sn.cdf_bs, cm.nB, sn.guide_bs, sn.gM, u2)
                           : upper_bound_guided(sn.cdf_b, cm.nB, sn.guide_b, sn.gM, u2);
    const float *gc = surface ? sn.cdf_gs : sn.cdf_g + (size_t)nb * cm.nQ;
    const int g = (cm.nQ <= 32) ? upper_bound_short(gc, cm.nQ, u3) : upper_bound(gc, cm.nQ, u3);
    return (uint32_t)(gld(sn.rank_of, nb) * cm.nQ + g);
}

/* ------------------------------------------------------------------------ */
/* the march                                                                 */

struct Ray {
    int ic, jc, kc, f0, f1, f2;
    float xp, yp, zp, sx, sy, sz, isx, isy, isz, tra;
};

struct En {
    uint32_t row;   /* table row: sorted band rank * nQ + g */
    int rank;
    float W, Bi, invBi;
};

/* storage of the k tables: read from the snapshot at run time, or fixed at
 * compile time (FP32 {k, Ib} pairs / BF16 k pairs) */
enum { TAB_RT = -1, TAB_F32 = 0, TAB_BF16 = 1 };

/* k and Ib of table row `row` (band rank * nQ + g) at temperature T:
 * clamp-to-edge linear interpolation in T (the reference's linear-filtered
 * textures, memory_copy_noint.cu:241-247, mc_gpu_noint.cu:784-791).  FP32
 * tables hold {k, Ib} pairs; BF16 tables hold k_t and k_t+1 as a BF16 pair
 * times a power-of-two band scale, and per band the Ib pairs in FP32 beside
 * the scale (one 16-byte word). */
/* the snapshot fields the march reads, held as wavefront-uniform values by
 * the batched kernel (the march functions accept it or a DevSnap) */
struct SnapView {
    float Tnb0, invdT;
    int kbf16;
    const float2 *kb;
    const uint32_t *kp16;
    const float4 *bp;
    const float *bfix;
};

template <int TM = TAB_RT, class S>
__device__ __forceinline__ void kb_at(const S &sn, uint32_t row, int rank, int nT, float T, float &K, float &B) {
    int t; float a;
    tpos(T, sn.Tnb0, sn.invdT, nT, t, a);
    if (TM == TAB_RT ? sn.kbf16 != 0 : TM == TAB_BF16) {
        const uint32_t w = gld32(sn.kp16, row * (uint32_t)nT + (uint32_t)t);
        const float4 b = gld32(sn.bp, (uint32_t)rank * (uint32_t)nT + (uint32_t)t);
        const float k0 = __uint_as_float(w << 16) * b.z;
        const float k1 = __uint_as_float(w & 0xFFFF0000u) * b.z;
        K = lerp(k0, k1, a);
        B = lerp(b.x, b.y, a);
    } else {
        const float4 q = gld32_pair(sn.kb, row * (uint32_t)(nT + 1) + (uint32_t)t);
        K = lerp(q.x, q.z, a);
        B = lerp(q.y, q.w, a);
    }
}

__device__ __forceinline__ void set_dir(Ray &r, float sx, float sy, float sz) {
    r.sx = sx; r.sy = sy; r.sz = sz;
    r.f0 = sx < 0.f; r.f1 = sy < 0.f; r.f2 = sz < 0.f;
    r.isx = 1.0f / sx; r.isy = 1.0f / sy; r.isz = 1.0f / sz;
}

/* wall emission tally W tra (Bw - Bi)/Bi, then the ray ends */
__device__ __forceinline__ void wall_hit(const En &e, Ray &r, float Bw, float &de) {
    de -= (e.W * r.tra) * ((Bw - e.Bi) * e.invBi);
    r.tra = 0.f;
}

/* boundary condition of the face the ray just crossed (the one index that
 * left 1..n): periodic, black (reference wall, adiabatic midpoint or fixed
 * temperature), specular, black at the adjacent-cell temperature (z), lost
 * (z) -- mc_gpu_noint.cu:375-654 */
template <int TM = TAB_RT, class S>
__device__ __forceinline__ void boundary(const DevCommon &cm, const S &sn, const DevLevel &L, Ray &r, const En &e,
                                      float &de, int &status) {
    const int nT = cm.nT;
    int face;
    if (r.ic == 0) face = FW;
    else if (r.ic == L.im + 1) face = FE;
    else if (r.jc == 0) face = FS;
    else if (r.jc == L.jm + 1) face = FN;
    else if (r.kc == 0) face = FB;
    else face = FT;
    const int bd = cm.bd[face];
    if (bd == 1) {
        switch (face) {
            case FW: r.ic = L.im; r.xp = cm.Lx; break;
            case FE: r.ic = 1; r.xp = 0.f; break;
            case FS: r.jc = L.jm; r.yp = cm.Ly; break;
            case FN: r.jc = 1; r.yp = 0.f; break;
            case FB: r.kc = L.km; r.zp = cm.Lz; break;
            default: r.kc = 1; r.zp = 0.f; break;
        }
    } else if (bd == 2) {
        float Bw;
        if ((face == FW && cm.adiab[0]) || (face == FE && cm.adiab[1])) {
            const float Tw = gld(L.Tw, (size_t)(face == FE) * L.jm * L.km + (size_t)(r.jc - 1) + (size_t)L.jm * (r.kc - 1));
            float Kw;
            kb_at<TM>(sn, e.row, e.rank, nT, Tw, Kw, Bw);
        } else {
            Bw = gld(sn.bfix, (size_t)face * cm.nB + e.rank);
        }
        wall_hit(e, r, Bw, de);
        status = RC_RAY_WALL;
    } else if (bd == 3) {
        switch (face) {
            case FW: r.xp = 0.f; r.sx = -r.sx; r.isx = -r.isx; r.ic = 1; r.f0 = r.sx < 0.f; break;
            case FE: r.xp = cm.Lx; r.sx = -r.sx; r.isx = -r.isx; r.ic = L.im; r.f0 = r.sx < 0.f; break;
            case FS: r.yp = 0.f; r.sy = -r.sy; r.isy = -r.isy; r.jc = 1; r.f1 = r.sy < 0.f; break;
            case FN: r.yp = cm.Ly; r.sy = -r.sy; r.isy = -r.isy; r.jc = L.jm; r.f1 = r.sy < 0.f; break;
            case FB: r.zp = 0.f; r.sz = -r.sz; r.isz = -r.isz; r.kc = 1; r.f2 = r.sz < 0.f; break;
            default: r.zp = cm.Lz; r.sz = -r.sz; r.isz = -r.isz; r.kc = L.km; r.f2 = r.sz < 0.f; break;
        }
    } else if (bd == 4) {
        const float Tw = gld(L.Tz, (size_t)(face == FT) * L.im * L.jm + (size_t)(r.ic - 1) + (size_t)L.im * (r.jc - 1));
        float Kw, Bw;
        kb_at<TM>(sn, e.row, e.rank, nT, Tw, Kw, Bw);
        wall_hit(e, r, Bw, de);
        status = RC_RAY_WALL;
    } else {
        r.tra = 0.f;
        status = RC_RAY_LOST;
    }
}

/* one march step (one iteration of march_ray's loop, mc_gpu_noint.cu:252-654):
 * distances to the three candidate faces, position update, attenuation and
 * reciprocity tally in the current cell, transmissivity cut-off, then one
 * index move.  Written with selects, no branch; returns whether an index
 * left the grid (the boundary condition is then due). */
template <int TM = TAB_RT, class S>
__device__ __forceinline__ bool march_core(const DevCommon &cm, const S &sn, const DevLevel &L, Ray &r,
                                           const En &e, float tol, float &de, int &status) {
    const int nT = cm.nT;
    /* both indices are in range for a ray inside the grid; the bounds only
     * matter to a lane whose state is stale (the batched kernel steps idle
     * lanes too and discards what they compute) */
    const float fx = gld32(L.xu, min((uint32_t)(r.ic - r.f0), (uint32_t)L.im));
    const float fy = (float)(r.jc - r.f1) * L.dy;
    const float fz = (float)(r.kc - r.f2) * L.dz;
    const float dsx = (fx - r.xp) * r.isx;
    const float dsy = (fy - r.yp) * r.isy;
    const float dsz = (fz - r.zp) * r.isz;
    const uint32_t cell = min((uint32_t)(r.ic - 1) + (uint32_t)L.im * ((uint32_t)(r.jc - 1) + (uint32_t)L.jm * (uint32_t)(r.kc - 1)),
                              (uint32_t)L.ncell - 1u);
    const float Tc = gld32(L.T, cell);
    float ds = dsx;
    ds = (ds > dsy) ? dsy : ds;
    ds = (ds > dsz) ? dsz : ds;
    r.xp = (ds == dsx) ? fx : fmaf(ds, r.sx, r.xp);
    r.yp = (ds == dsy) ? fy : fmaf(ds, r.sy, r.yp);
    r.zp = (ds == dsz) ? fz : fmaf(ds, r.sz, r.zp);
    float K, B;
    kb_at<TM>(sn, e.row, e.rank, nT, Tc, K, B);
    float em1, ex;
    expm1_exp(-(K * ds), em1, ex);
    const float alpha = -em1;
    const float fac = (B - e.Bi) * e.invBi;
    de -= ((e.W * r.tra) * alpha) * fac;
    const float tra = r.tra * ex;
    const bool dump = tra < tol;
    const float dd = (e.W * tra) * fac;
    de = dump ? de - dd : de;
    r.tra = dump ? 0.f : tra;
    status = dump ? RC_RAY_ABSORBED : status;
    /* one index moves along the crossed axis */
    const bool xa = dsx < dsy && dsx < dsz;
    const bool ya = !(dsx < dsy) && dsy < dsz;
    const bool za = !xa && !ya;
    r.ic += xa ? (r.f0 ? -1 : 1) : 0;
    r.jc += ya ? (r.f1 ? -1 : 1) : 0;
    r.kc += za ? (r.f2 ? -1 : 1) : 0;
    return (uint32_t)(r.ic - 1) >= (uint32_t)L.im || (uint32_t)(r.jc - 1) >= (uint32_t)L.jm ||
           (uint32_t)(r.kc - 1) >= (uint32_t)L.km;
}

/* march_core, then the boundary condition of the face crossed, if any */
template <int TM = TAB_RT, class S>
__device__ __forceinline__ void march_step(const DevCommon &cm, const S &sn, const DevLevel &L, Ray &r,
                                           const En &e, float tol, float &de, int &status) {
    if (march_core<TM>(cm, sn, L, r, e, tol, de, status)) boundary<TM>(cm, sn, L, r, e, de, status);
}

/* one grid level of the march (march_ray, mc_gpu_noint.cu:245-658);
 * returns the steps taken on this level */
template <bool TR>
__device__ __forceinline__ int march_level(const DevCommon &cm, const DevSnap &sn, const DevLevel &L, Ray &r,
                                           const En &e, float tol, float &de, int &steps, uint32_t &hash,
                                           int &status) {
    int counter = 0;
    while (r.tra > tol && counter < L.maxs) {
        if (TR) hash = fnv(fnv(fnv(hash, (uint32_t)r.ic), (uint32_t)r.jc), (uint32_t)r.kc);
        march_step(cm, sn, L, r, e, tol, de, status);
        steps += 1;
        counter += 1;
    }
    return counter;
}

/* all grid levels with index coarsening (mc_gpu_noint.cu:210-224) */
template <bool TR>
__device__ __forceinline__ void march_all(const DevCommon &cm, const DevSnap &sn, Ray &r, const En &e, float tol,
                                          float &de, int &steps, uint32_t &hash, int &status, int &level) {
    status = RC_RAY_STOPPED;
    level = 0;
    for (int g = 0; g < cm.grid_num; ++g) {
        const DevLevel &L = sn.lev[g];
        level = g;
        const int n = march_level<TR>(cm, sn, L, r, e, tol, de, steps, hash, status);
        if (!(r.tra > tol)) return;
        if (g == cm.grid_num - 1) {
            if (n >= L.maxs) status = RC_RAY_CAPPED;
            return;
        }
        const DevLevel &N = sn.lev[g + 1];
        const int ri = L.im / N.im, rj = L.jm / N.jm, rk = L.km / N.km;
        r.ic = (r.ic + ri - 1) / ri;
        r.jc = (r.jc + rj - 1) / rj;
        r.kc = (r.kc + rk - 1) / rk;
    }
}

/* per-ray energy constants: W = Ibmax B(Ti)/B(Tmax) K(Ti)/K(Tmax) for the
 * fluid (:181-186), W = Ibmax_s B(Tw)/B(Tmax) for walls (:1419-1424) */
template <int TM = TAB_RT>
__device__ __forceinline__ void energy(const DevCommon &cm, const DevSnap &sn, uint32_t key, float Ti, bool surface,
                                       int bound, En &e) {
    const int rank = (int)(key / (uint32_t)cm.nQ);
    e.rank = rank;
    e.row = key;
    float Ki, Bi, Km, Bm;
    kb_at<TM>(sn, e.row, rank, cm.nT, Ti, Ki, Bi);
    kb_at<TM>(sn, e.row, rank, cm.nT, sn.Tmax, Km, Bm);
    e.Bi = Bi;
    e.invBi = 1.0f / Bi;
    if (!surface) e.W = sn.Ibmax * (((Bi / Bm) * Ki) / Km);
    else e.W = (bound ? sn.Ibmax_e : sn.Ibmax_w) * (Bi / Bm);
}

/* emitting fluid cell n -> (i, j, k global) */
__device__ __forceinline__ void cell_ijk(const DevCommon &cm, int n, int &i, int &j, int &k) {
    i = n % cm.imax + 1;
    j = (n / cm.imax) % cm.jmax + 1;
    k = n / (cm.imax * cm.jmax) + cm.zstart;
}

/* isotropic emission direction of a fluid ray from its uniforms (a, b),
 * with exact zeros moved off the axes (kernel_fluid :131-160) */
__device__ __forceinline__ void fluid_dir(float a, float b, float &sx, float &sy, float &sz) {
    iso_dir(a, b, sx, sy, sz);
    if (sx == 0.f) sx = 1e-10f;
    if (sy == 0.f) sy = 1e-10f;
    if (sz == 0.f) sz = 1e-10f;
}

/* Lambertian emission direction from wall `bound` (kernel_surface
 * :1398-1417) */
__device__ __forceinline__ void surface_dir(float a, float b, int bound, float &sx, float &sy, float &sz) {
    lamb_dir(a, b, bound, sx, sy, sz);
    if (fabsf(sx) <= 1e-10) sx = (float)(1 - 2 * bound) * 1e-10f;
    if (fabsf(sy) <= 1e-10) sy = (float)((sy > 0.f) - (sy <= 0.f)) * 1e-10f;
    if (fabsf(sz) <= 1e-10) sz = (float)((sz > 0.f) - (sz <= 0.f)) * 1e-10f;
}

/* on-face index adjustment of a wall ray placed on the wall cell face
 * (:1429-1441): an index moves when the ray starts on a face it leaves */
__device__ __forceinline__ void surface_adjust(const DevSnap &sn, Ray &r) {
    const DevLevel &L0 = sn.lev[0];
    if (r.f0 && r.xp == gld(L0.xu, r.ic - 1)) r.ic -= 1;
    if (r.f1 && r.yp == (float)(r.jc - 1) * L0.dy) r.jc -= 1;
    if (r.f2 && r.zp == (float)(r.kc - 1) * L0.dz) r.kc -= 1;
    if (!r.f0 && r.xp == gld(L0.xu, r.ic)) r.ic += 1;
    if (!r.f1 && r.yp == (float)r.jc * L0.dy) r.jc += 1;
    if (!r.f2 && r.zp == (float)r.kc * L0.dz) r.kc += 1;
}

/* set-up of one fluid ray from the centre of cell (i,j,k) (kernel_fluid
 * :115-186): direction, band, per-ray energy constants */
__device__ __forceinline__ uint32_t fluid_setup(const DevCommon &cm, const DevSnap &sn, int i, int j, int k,
                                                uint32_t cid, int v, int h, uint32_t key, bool have_key, float Ti,
                                                Ray &r, En &e) {
    const U4 u = draw4(cm.k0, cm.k1, (uint32_t)h, (uint32_t)v, cid, TAG_FLUID + 16u * (uint32_t)sn.snap);
    if (!have_key) key = band_key(cm, sn, false, u.c, u.d);
    float sx, sy, sz;
    fluid_dir(u.a, u.b, sx, sy, sz);
    set_dir(r, sx, sy, sz);
    r.ic = i; r.jc = j; r.kc = k;
    r.xp = gld(sn.x0, i); r.yp = gld(sn.y0, j); r.zp = gld(sn.z0, k);
    r.tra = 1.0f;
    energy(cm, sn, key, Ti, false, 0, e);
    return key;
}

/* set-up of one ray from the centre of wall cell (j,k) (kernel_surface
 * :1381-1441) */
__device__ __forceinline__ uint32_t surface_setup(const DevCommon &cm, const DevSnap &sn, int bound, int j, int k,
                                                  uint32_t cid, int v, int h, uint32_t key, bool have_key, float Tw,
                                                  Ray &r, En &e) {
    const U4 u = draw4(cm.k0, cm.k1, (uint32_t)h, (uint32_t)v, cid,
                       (bound ? TAG_EAST : TAG_WEST) + 16u * (uint32_t)sn.snap);
    if (!have_key) key = band_key(cm, sn, true, u.c, u.d);
    float sx, sy, sz;
    surface_dir(u.a, u.b, bound, sx, sy, sz);
    set_dir(r, sx, sy, sz);
    r.ic = bound ? cm.imax : 1; r.jc = j; r.kc = k;
    r.xp = bound ? cm.Lx : 0.f; r.yp = gld(sn.y0, j); r.zp = gld(sn.z0, k);
    r.tra = 1.0f;
    energy(cm, sn, key, Tw, true, bound, e);
    surface_adjust(sn, r);
    return key;
}

template <bool TR>
__device__ __forceinline__ void fill_record(const DevCommon &cm, const Ray &r, const En &e, uint32_t key, int st,
                                            int status, int level, uint32_t hash, float de, rc_ray *rec) {
    if (TR) {
        rec->nb = e.rank; rec->g = (int)(key % (uint32_t)cm.nQ);
        rec->ic = r.ic; rec->jc = r.jc; rec->kc = r.kc; rec->level = level;
        rec->steps = st; rec->status = status; rec->path_hash = hash;
        rec->tra = r.tra; rec->de = de;
        rec->de64 = 0.0; rec->sumabs = 0.0;
    }
}

/* one fluid ray from the centre of cell (i,j,k) (kernel_fluid :115-224) */
template <bool TR>
__device__ __forceinline__ float fluid_ray(const DevCommon &cm, const DevSnap &sn, int i, int j, int k, uint32_t cid,
                                           int v, int h, uint32_t key, bool have_key, float Ti, int &steps,
                                           rc_ray *rec) {
    Ray r;
    En e;
    key = fluid_setup(cm, sn, i, j, k, cid, v, h, key, have_key, Ti, r, e);
    if (TR) { rec->sx = r.sx; rec->sy = r.sy; rec->sz = r.sz; }
    float de = 0.f;
    int st = 0, status, level;
    uint32_t hash = 2166136261u;
    march_all<TR>(cm, sn, r, e, cm.toll, de, st, hash, status, level);
    steps += st;
    fill_record<TR>(cm, r, e, key, st, status, level, hash, de, rec);
    return de;
}

/* one ray from the centre of wall cell (j,k) (kernel_surface :1381-1462) */
template <bool TR>
__device__ __forceinline__ float surface_ray(const DevCommon &cm, const DevSnap &sn, int bound, int j, int k,
                                             uint32_t cid, int v, int h, uint32_t key, bool have_key, float Tw,
                                             int &steps, rc_ray *rec) {
    Ray r;
    En e;
    key = surface_setup(cm, sn, bound, j, k, cid, v, h, key, have_key, Tw, r, e);
    if (TR) { rec->sx = r.sx; rec->sy = r.sy; rec->sz = r.sz; }
    float de = 0.f;
    int st = 0, status, level;
    uint32_t hash = 2166136261u;
    march_all<TR>(cm, sn, r, e, cm.tollS, de, st, hash, status, level);
    steps += st;
    fill_record<TR>(cm, r, e, key, st, status, level, hash, de, rec);
    return de;
}

__device__ __forceinline__ float cell_T(const DevCommon &cm, const DevSnap &sn, int i, int j, int k) {
    return gld(sn.T0, (size_t)i + (size_t)(cm.imax + 2) * ((size_t)j + (size_t)(cm.jmax + 2) * (size_t)k));
}
__device__ __forceinline__ float wall_T(const DevCommon &cm, const DevSnap &sn, int bound, int j, int k) {
    const size_t row = (size_t)(cm.imax + 2) * ((size_t)j + (size_t)(cm.jmax + 2) * (size_t)k);
    /* texel coordinate 1.0 (imax+1.0): halfway between ghost and first cell */
    return bound ? lerp(gld(sn.T0, row + cm.imax), gld(sn.T0, row + cm.imax + 1), 0.5f)
                 : lerp(gld(sn.T0, row), gld(sn.T0, row + 1), 0.5f);
}

}  // namespace rc

#endif
