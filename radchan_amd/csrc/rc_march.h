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
 * the scale (one 16-byte word).  Rows are tab_stride(nT) entries long. */
/* the snapshot fields the march reads, held as wavefront-uniform values by
 * the batched kernel (the march functions accept it or a DevSnap) */
struct SnapView {
    float Tnb0, invdT;
    int kbf16;
    const float2 *kb;
    const uint32_t *kp16;
    const float4 *bp;
    const float *xu_all, *P_all;
    int nxu, npos;
};

/* k and Ib of table row `row` at row t with weight a */
template <int TM = TAB_RT, class S>
__device__ __forceinline__ void kb_ta(const S &sn, uint32_t row, int rank, int nT, int t, float a, float &K, float &B) {
    const uint32_t st = (uint32_t)tab_stride(nT);
    if (TM == TAB_RT ? sn.kbf16 != 0 : TM == TAB_BF16) {
        const uint32_t w = gld32(sn.kp16, row * st + (uint32_t)t);
        const float4 b = gld32(sn.bp, (uint32_t)rank * st + (uint32_t)t);
        const float k0 = __uint_as_float(w << 16) * b.z;
        const float k1 = __uint_as_float(w & 0xFFFF0000u) * b.z;
        K = lerp(k0, k1, a);
        B = lerp(b.x, b.y, a);
    } else {
        const float4 q = gld32_pair(sn.kb, row * st + (uint32_t)t);
        K = lerp(q.x, q.z, a);
        B = lerp(q.y, q.w, a);
    }
}

/* at temperature T (the emission of a ray) */
template <int TM = TAB_RT, class S>
__device__ __forceinline__ void kb_at(const S &sn, uint32_t row, int rank, int nT, float T, float &K, float &B) {
    int t; float a;
    tpos(T, sn.Tnb0, sn.invdT, nT, t, a);
    kb_ta<TM>(sn, row, rank, nT, t, a, K, B);
}

/* at a table position (a cell or wall entry of a level's position array) */
template <int TM = TAB_RT, class S>
__device__ __forceinline__ void kb_pos(const S &sn, uint32_t row, int rank, int nT, float pos, float &K, float &B) {
    int t; float a;
    ppos(pos, t, a);
    kb_ta<TM>(sn, row, rank, nT, t, a, K, B);
}

/* kb_pos in two halves, for a fixed storage mode: the loads of the two
 * table rows t, t+1 (issued at once), and the interpolation (when the
 * values are needed); finish(fetch(...)) computes exactly kb_pos */
struct KbRaw {
    float4 q;        /* FP32: {k_t, Ib_t, k_t+1, Ib_t+1}; BF16: {Ib_t, Ib_t+1, scale, 0} */
    uint32_t w;      /* BF16: k_t | k_t+1 << 16 */
    float a;         /* interpolation weight */
};

template <int TM, class S>
__device__ __forceinline__ void kb_load(const S &sn, uint32_t row, int rank, int nT, int t, KbRaw &k) {
    static_assert(TM == TAB_F32 || TM == TAB_BF16, "fixed storage mode");
    const uint32_t st = (uint32_t)tab_stride(nT);
    if (TM == TAB_BF16) {
        k.w = gld32(sn.kp16, row * st + (uint32_t)t);
        k.q = gld32(sn.bp, (uint32_t)rank * st + (uint32_t)t);
    } else {
        k.w = 0u;
        k.q = gld32_pair(sn.kb, row * st + (uint32_t)t);
    }
}

template <int TM, class S>
__device__ __forceinline__ KbRaw kb_fetch(const S &sn, uint32_t row, int rank, int nT, float pos) {
    KbRaw k;
    int t;
    ppos(pos, t, k.a);
    kb_load<TM>(sn, row, rank, nT, t, k);
    return k;
}

template <int TM>
__device__ __forceinline__ void kb_finish(const KbRaw &k, float &K, float &B) {
    if (TM == TAB_BF16) {
        const float k0 = __uint_as_float(k.w << 16) * k.q.z;
        const float k1 = __uint_as_float(k.w & 0xFFFF0000u) * k.q.z;
        K = lerp(k0, k1, k.a);
        B = lerp(k.q.x, k.q.y, k.a);
    } else {
        K = lerp(k.q.x, k.q.z, k.a);
        B = lerp(k.q.y, k.q.w, k.a);
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

/* outcome of the boundary condition of a face */
enum { RB_ON = 0, RB_WALL = 1, RB_LOST = 2 };

/* the entry of the black wall the ray left the grid through (face `face`,
 * the other two indices in range) in the level's position array */
template <class G>
__device__ __forceinline__ uint32_t wall_index(const G &L, const Ray &r, int face) {
    const uint32_t im = (uint32_t)L.im, jm = (uint32_t)L.jm, km = (uint32_t)L.km;
    const uint32_t i = (uint32_t)(r.ic - 1), j = (uint32_t)(r.jc - 1), k = (uint32_t)(r.kc - 1);
    const uint32_t side = (uint32_t)(face & 1);
    const uint32_t nx = __umul24(jm, km), ny = __umul24(im, km), nz = __umul24(im, jm);
    const uint32_t x0 = __umul24(im, nx), y0 = x0 + 2u * nx, z0 = y0 + 2u * ny;
    return face < FS ? x0 + side * nx + j + jm * k : (face < FB ? y0 + side * ny + i + im * k : z0 + side * nz + i + im * j);
}

/* the face a ray that left the grid crossed (exactly one index left 1..n
 * in the step) */
template <class G>
__device__ __forceinline__ int exit_face(const G &L, const Ray &r) {
    const bool xo = (uint32_t)(r.ic - 1) >= (uint32_t)L.im;
    const bool yo = !xo && (uint32_t)(r.jc - 1) >= (uint32_t)L.jm;
    const bool lo = xo ? r.ic == 0 : (yo ? r.jc == 0 : r.kc == 0);
    return (xo ? FW : (yo ? FS : FB)) + (lo ? 0 : 1);
}

/* the condition of face f, from the six face codes packed 3 bits each (a
 * wavefront-uniform word, no per-lane memory access) */
__device__ __forceinline__ int face_bd(const DevCommon &cm, int face) {
    const uint32_t pack = (uint32_t)cm.bd[0] | ((uint32_t)cm.bd[1] << 3) | ((uint32_t)cm.bd[2] << 6) |
                          ((uint32_t)cm.bd[3] << 9) | ((uint32_t)cm.bd[4] << 12) | ((uint32_t)cm.bd[5] << 15);
    return (int)((pack >> (3 * face)) & 7u);
}

/* periodic (bd 1): to the opposite end (low face -> last cell, coordinate
 * at the high end); specular (bd 3): back into the edge cell, coordinate on
 * the face, direction component reversed */
template <class G>
__device__ __forceinline__ void face_move(const DevCommon &cm, const G &L, Ray &r, int face, bool per) {
    const bool xo = face < FS, yo = face >= FS && face < FB, zo = face >= FB;
    const bool lo = (face & 1) == 0;
    const bool hi_end = per == lo;
    r.ic = xo ? (hi_end ? L.im : 1) : r.ic;
    r.xp = xo ? (hi_end ? cm.Lx : 0.f) : r.xp;
    r.jc = yo ? (hi_end ? L.jm : 1) : r.jc;
    r.yp = yo ? (hi_end ? cm.Ly : 0.f) : r.yp;
    r.kc = zo ? (hi_end ? L.km : 1) : r.kc;
    r.zp = zo ? (hi_end ? cm.Lz : 0.f) : r.zp;
    const bool fx = !per && xo, fy = !per && yo, fz = !per && zo;
    r.sx = fx ? -r.sx : r.sx;
    r.isx = fx ? -r.isx : r.isx;
    r.sy = fy ? -r.sy : r.sy;
    r.isy = fy ? -r.isy : r.isy;
    r.sz = fz ? -r.sz : r.sz;
    r.isz = fz ? -r.isz : r.isz;
    r.f0 = r.sx < 0.f;
    r.f1 = r.sy < 0.f;
    r.f2 = r.sz < 0.f;
}

/* boundary condition of the face the ray just crossed (the one index that
 * left 1..n), without its tally: periodic and specular faces move the ray
 * on (RB_ON); a black face -- reference wall, adiabatic midpoint or fixed
 * temperature -- or a bd 4 z face ends it with the wall's Ib in Bw, the
 * table at the wall's position (RB_WALL); a bd 5 face loses it (RB_LOST)
 * -- mc_gpu_noint.cu:375-654 */
template <int TM = TAB_RT, class S>
__device__ __forceinline__ int boundary_move(const DevCommon &cm, const S &sn, const DevLevel &L, Ray &r,
                                             const En &e, float &Bw) {
    const int face = exit_face(L, r);
    const int bd = face_bd(cm, face);
    if (bd == 1 || bd == 3) {
        face_move(cm, L, r, face, bd == 1);
        return RB_ON;
    }
    if (bd == 2 || bd == 4) {
        float Kw;
        kb_pos<TM>(sn, e.row, e.rank, cm.nT, gld(L.P, wall_index(L, r, face)), Kw, Bw);
        return RB_WALL;
    }
    return RB_LOST;
}

/* the boundary condition with its tally: the wall emission tally ends the
 * ray at a wall; a lost ray ends with nothing */
template <int TM = TAB_RT, class S>
__device__ __forceinline__ void boundary(const DevCommon &cm, const S &sn, const DevLevel &L, Ray &r, const En &e,
                                      float &de, int &status) {
    float Bw = 0.f;
    const int b = boundary_move<TM>(cm, sn, L, r, e, Bw);
    if (b == RB_WALL) {
        wall_hit(e, r, Bw, de);
        status = RC_RAY_WALL;
    } else if (b == RB_LOST) {
        r.tra = 0.f;
        status = RC_RAY_LOST;
    }
}

/* The batched march carries the ray's linear cell index from step to step
 * (a step moves it by +-1, +-im or +-im jm) instead of forming it from the
 * three indices: the per-ray steps along y and z, set at the ray's start,
 * after a reflection or wrap and at a level change. */
struct CellWalk {
    uint32_t cell;
    int dcy, dcz;
};

template <class G>
__device__ __forceinline__ void cell_walk_setup(const G &L, const Ray &r, CellWalk &w) {
    w.cell = (uint32_t)(r.ic - 1) + __umul24((uint32_t)L.im, (uint32_t)(r.jc - 1) +
                                                              __umul24((uint32_t)L.jm, (uint32_t)(r.kc - 1)));
    const int imjm = (int)__umul24((uint32_t)L.im, (uint32_t)L.jm);
    w.dcy = r.f1 ? -L.im : L.im;
    w.dcz = r.f2 ? -imjm : imjm;
}

/* The boundary condition of the batched march, without memory access:
 * periodic and specular faces move the ray on; at a black or bd 4 face the
 * ray keeps its transmissivity and gets the index of the wall's entry in
 * the level's position array (wall = 1), and the next march step takes the
 * wall's emission through the step's own table lookup (march_core_bulk);
 * a bd 5 face loses the ray (tra = 0, no tally).  The moves are those of
 * face_move, and they move the ray's cell walk with it; the wall entries
 * are wall_index's. */
template <class G>
__device__ __forceinline__ void boundary_defer(const DevCommon &cm, const G &L, Ray &r, bool left, int &wall,
                                               uint32_t &widx, CellWalk &cw) {
    /* one block per axis, each taken only when a lane left through one of
     * its two faces; the face codes are wavefront-uniform */
    const uint32_t im = (uint32_t)L.im, jm = (uint32_t)L.jm, km = (uint32_t)L.km;
    const uint32_t imjm = __umul24(im, jm), ncell = __umul24(im, __umul24(jm, km));
    const bool xo = left && (uint32_t)(r.ic - 1) >= im;
    const bool yo = left && !xo && (uint32_t)(r.jc - 1) >= jm;
    const bool zo = left && !xo && !yo;
    if (xo) {
        const bool lo = r.ic == 0;
        const int bd = lo ? cm.bd[FW] : cm.bd[FE];
        if (bd == 1) {                     /* periodic: to the opposite end */
            r.ic = lo ? L.im : 1;
            r.xp = lo ? cm.Lx : 0.f;
            cw.cell += lo ? im : (uint32_t)-(int)im;
        } else if (bd == 3) {              /* specular: back into the edge cell, sx reversed */
            r.ic = lo ? 1 : L.im;
            r.xp = lo ? 0.f : cm.Lx;
            r.sx = -r.sx;
            r.isx = -r.isx;
            r.f0 = r.sx < 0.f;
            cw.cell += lo ? 1u : (uint32_t)-1;
        } else {                           /* black */
            wall = 1;
            widx = ncell + (lo ? 0u : __umul24(jm, km)) + (uint32_t)(r.jc - 1) +
                   __umul24(jm, (uint32_t)(r.kc - 1));
        }
    }
    if (yo) {
        const bool lo = r.jc == 0;
        const int bd = lo ? cm.bd[FS] : cm.bd[FN];
        if (bd == 1) {
            r.jc = lo ? L.jm : 1;
            r.yp = lo ? cm.Ly : 0.f;
            const uint32_t w = __umul24(jm, im);
            cw.cell += lo ? w : (uint32_t)-(int)w;
        } else if (bd == 3) {
            r.jc = lo ? 1 : L.jm;
            r.yp = lo ? 0.f : cm.Ly;
            r.sy = -r.sy;
            r.isy = -r.isy;
            r.f1 = r.sy < 0.f;
            cw.cell += lo ? im : (uint32_t)-(int)im;
            cw.dcy = -cw.dcy;
        } else {
            wall = 1;
            widx = ncell + 2u * __umul24(jm, km) + (lo ? 0u : __umul24(im, km)) +
                   (uint32_t)(r.ic - 1) + __umul24(im, (uint32_t)(r.kc - 1));
        }
    }
    if (zo) {
        const bool lo = r.kc == 0;
        const int bd = lo ? cm.bd[FB] : cm.bd[FT];
        if (bd == 1) {
            r.kc = lo ? L.km : 1;
            r.zp = lo ? cm.Lz : 0.f;
            const uint32_t w = ncell;
            cw.cell += lo ? w : (uint32_t)-(int)w;
        } else if (bd == 3) {
            r.kc = lo ? 1 : L.km;
            r.zp = lo ? 0.f : cm.Lz;
            r.sz = -r.sz;
            r.isz = -r.isz;
            r.f2 = r.sz < 0.f;
            cw.cell += lo ? imjm : (uint32_t)-(int)imjm;
            cw.dcz = -cw.dcz;
        } else if (bd == 2 || bd == 4) {
            wall = 1;
            widx = ncell + 2u * (__umul24(jm, km) + __umul24(im, km)) + (lo ? 0u : imjm) +
                   (uint32_t)(r.ic - 1) + __umul24(im, (uint32_t)(r.jc - 1));
        } else {                           /* bd 5: lost, no tally */
            r.tra = 0.f;
        }
    }
}

/* The geometry of one march step (one iteration of march_ray's loop,
 * mc_gpu_noint.cu:252-654), given the x face the ray heads for
 * (fx = xu[ic - f0]): distances to the three candidate faces, the segment
 * length ds, the position update and the index move along the crossed
 * axis (xa: x, ya: y, else z).  Nothing here depends on the temperature.
 * Returns whether an index left the grid (the boundary condition is then
 * due). */
template <class G>
__device__ __forceinline__ bool geo_move(const G &L, Ray &r, float fx, float &ds, bool &xa, bool &ya) {
    const float fy = (float)(r.jc - r.f1) * L.dy;
    const float fz = (float)(r.kc - r.f2) * L.dz;
    const float dsx = (fx - r.xp) * r.isx;
    const float dsy = (fy - r.yp) * r.isy;
    const float dsz = (fz - r.zp) * r.isz;
    ds = dsx;
    ds = (ds > dsy) ? dsy : ds;
    ds = (ds > dsz) ? dsz : ds;
    r.xp = (ds == dsx) ? fx : fmaf(ds, r.sx, r.xp);
    r.yp = (ds == dsy) ? fy : fmaf(ds, r.sy, r.yp);
    r.zp = (ds == dsz) ? fz : fmaf(ds, r.sz, r.zp);
    xa = dsx < dsy && dsx < dsz;
    ya = !(dsx < dsy) && dsy < dsz;
    const bool za = !xa && !ya;
    r.ic += xa ? (r.f0 ? -1 : 1) : 0;
    r.jc += ya ? (r.f1 ? -1 : 1) : 0;
    r.kc += za ? (r.f2 ? -1 : 1) : 0;
    return (uint32_t)(r.ic - 1) >= (uint32_t)L.im || (uint32_t)(r.jc - 1) >= (uint32_t)L.jm ||
           (uint32_t)(r.kc - 1) >= (uint32_t)L.km;
}

/* the linear index of the ray's cell on its level (clamped: a lane whose
 * state is stale reads in bounds) */
__device__ __forceinline__ uint32_t cell_index(const DevLevel &L, const Ray &r) {
    /* 24-bit products (full rate): the factors are below 2^24 for every
     * grid whose jm km and im are, and the product is the cell index */
    return min((uint32_t)(r.ic - 1) + __umul24((uint32_t)L.im, (uint32_t)(r.jc - 1) +
                                                                  __umul24((uint32_t)L.jm, (uint32_t)(r.kc - 1))),
               (uint32_t)L.ncell - 1u);
}

/* geo_move, with `cell` the cell the segment crosses */
__device__ __forceinline__ bool geo_core_fx(const DevLevel &L, Ray &r, float fx, uint32_t &cell, float &ds) {
    cell = cell_index(L, r);
    bool xa, ya;
    return geo_move(L, r, fx, ds, xa, ya);
}

/* geo_core_fx with the x face looked up: in global memory, or with XL in
 * an LDS copy */
template <bool XL = false>
__device__ __forceinline__ bool geo_core(const DevLevel &L, Ray &r, uint32_t &cell, float &ds,
                                         const float *xul = nullptr) {
    /* both indices are in range for a ray inside the grid; the bounds only
     * matter to a lane whose state is stale (the batched kernel steps idle
     * lanes too and discards what they compute) */
    const uint32_t ix = min((uint32_t)(r.ic - r.f0), (uint32_t)L.im);
    const float fx = XL ? xul[ix] : gld32(L.xu, ix);
    return geo_core_fx(L, r, fx, cell, ds);
}

/* The absorption of one step with optical depth -x (x = -K ds): attenuation
 * and reciprocity tally at the cell's Ib B, transmissivity cut-off (the
 * dump tally when it falls below tol). */
__device__ __forceinline__ void phys_core_x(float x, float B, const En &e, float tol, float &tra, float &de,
                                            int &status) {
    float em1, ex;
    expm1_exp_le88(x, em1, ex);
    const float alpha = -em1;
    const float fac = (B - e.Bi) * e.invBi;
    de -= ((e.W * tra) * alpha) * fac;
    const float tn = tra * ex;
    const bool dump = tn < tol;
    const float dd = (e.W * tn) * fac;
    de = dump ? de - dd : de;
    tra = dump ? 0.f : tn;
    status = dump ? RC_RAY_ABSORBED : status;
}

__device__ __forceinline__ void phys_core(float K, float B, float ds, const En &e, float tol, float &tra,
                                          float &de, int &status) {
    phys_core_x(-(K * ds), B, e, tol, tra, de, status);
}

/* one march step: geometry, then the absorption at the crossed cell's
 * table position (table interpolation, FP32).  Select-based, no branch;
 * returns whether an index left the grid. */
template <int TM = TAB_RT, class S>
__device__ __forceinline__ bool march_core(const DevCommon &cm, const S &sn, const DevLevel &L, Ray &r,
                                           const En &e, float tol, float &de, int &status) {
    uint32_t cell;
    float ds;
    const bool out = geo_core(L, r, cell, ds);
    float K, B;
    kb_pos<TM>(sn, e.row, e.rank, cm.nT, gld32(L.P, cell), K, B);
    phys_core(K, B, ds, e, tol, r.tra, de, status);
    return out;
}

/* A level of the batched march held per lane (the lanes of a multigrid
 * wavefront march rays on different levels): sizes, step cap, spacings and
 * the offsets of the level's x faces and table positions in the
 * snapshot's arrays of every level */
struct LaneLevel {
    int im, jm, km, maxs;
    float dy, dz;
    uint32_t xoff, poff;
};

__device__ __forceinline__ LaneLevel lane_level(const DevLevel &L) {
    LaneLevel l;
    l.im = L.im; l.jm = L.jm; l.km = L.km; l.maxs = L.maxs;
    l.dy = L.dy; l.dz = L.dz;
    l.xoff = (uint32_t)L.xoff; l.poff = (uint32_t)L.poff;
    return l;
}

/* the index ic on the next level, (ic + r - 1) / r (mc_gpu_noint.cu:215-223),
 * by the level's multiplier */
__device__ __forceinline__ int coarsen(int ic, int r, uint32_t m) {
    const uint32_t n = (uint32_t)(ic + r - 1);
    if (r == 1) return ic;
    return (int)(m ? __umulhi(n, m) : n / (uint32_t)r);
}

/* One step of the batched march.  A ray's table row is fixed, and along
 * its path the row t changes only every few cells, so the ray's last table
 * pair stays in registers and is loaded again only when t changes (kt: the
 * cached t, -1 for none; the loads of the other lanes are masked off).
 * The x face is loaded before the table position, so that the wait for the
 * position covers it and the geometry runs while a table pair is in
 * flight.  A ray waiting at a black wall (wall = 1, set by boundary_defer)
 * takes the wall's emission instead: the table at the wall's position,
 * with the whole transmissivity absorbed (optical depth infinite: alpha =
 * 1, the transmissivity then 0 and dumped), which is wall_hit's tally to
 * the bit; its geometry is discarded.  The cell index is carried in `cw`.
 * The level's x faces and positions are read at xu[xoff + i] and
 * P[poff + i], the indices clamped to xmax and pmax (a lane whose state is
 * stale reads in bounds).  Returns whether an index left the grid (never
 * for a wall step).  Same values as march_core and boundary, so the same
 * tallies. */
template <int TM, class S, class G>
__device__ __forceinline__ bool march_core_bulk(const DevCommon &cm, const S &sn, const G &L, const float *xu,
                                                uint32_t xoff, uint32_t xmax, const float *P, uint32_t poff,
                                                uint32_t pmax, Ray &r, const En &e, float tol, float &de, int &kt,
                                                KbRaw &kc, bool wall, uint32_t widx, CellWalk &cw) {
    const uint32_t ix = min(xoff + (uint32_t)(r.ic - r.f0), xmax);
    const float fx = gld32(xu, ix);
    asm volatile("" ::: "memory");
    const float pos = gld32(P, min(poff + (wall ? widx : cw.cell), pmax));
    int t;
    ppos(pos, t, kc.a);
    if (t != kt || wall) {
        kb_load<TM>(sn, e.row, e.rank, cm.nT, t, kc);
        kt = wall ? -1 : t;
    }
    const int stx = r.f0 ? -1 : 1;
    float ds;
    bool xa, ya;
    const bool out = geo_move(L, r, fx, ds, xa, ya);
    cw.cell += xa ? (uint32_t)stx : (uint32_t)(ya ? cw.dcy : cw.dcz);
    float K, B;
    kb_finish<TM>(kc, K, B);
    int status = 0;
    phys_core_x(wall ? -INFINITY : -(K * ds), B, e, tol, r.tra, de, status);
    return out && !wall;
}

/* A ray marched on one grid level to its end, or until its step counter
 * reaches `limit` (at most the level's step cap L.maxs), with the geometry
 * two steps ahead of the absorption: the temperature of step i+2 and the
 * table entries of step i+1 are loaded while step i is absorbed.  The
 * arithmetic of every step is that of march_step, so the tally,
 * transmissivity and step count are bit-identical; only a dumped ray's
 * final position differs (the geometry ran ahead of it).  A ray stopped by
 * the limit has taken exactly its steps up to the limit, so the march can
 * resume from it.  Expects a live ray below the limit.  With XL the
 * level's x faces come from an LDS copy (xul), so that the geometry never
 * waits on the vector-memory counter behind the prefetched loads. */
template <int TM = TAB_RT, bool XL = false, class S>
__device__ __forceinline__ void march_ahead(const DevCommon &cm, const S &sn, const DevLevel &L, Ray &r,
                                            const En &e, float tol, float &de, int &counter, int &nsteps,
                                            const float *xul, int limit) {
    const int nT = cm.nT;
    int status = 0;
    /* step i */
    uint32_t c0;
    float d0, w0;
    int b0 = RB_ON;
    if (geo_core<XL>(L, r, c0, d0, xul)) b0 = boundary_move<TM>(cm, sn, L, r, e, w0);
    KbRaw k0 = kb_fetch<TM>(sn, e.row, e.rank, nT, gld32(L.P, c0));
    /* step i+1 */
    /* The loads are issued on every path (cell 0 for a step not taken,
     * whose values are then never used), so that the number of loads in
     * flight does not depend on the path and the wait for the oldest one
     * does not have to drain the newer ones. */
    bool h1 = b0 == RB_ON && counter + 1 < limit;
    uint32_t c1 = 0u;
    float d1, w1;
    int b1 = RB_ON;
    if (h1)
        if (geo_core<XL>(L, r, c1, d1, xul)) b1 = boundary_move<TM>(cm, sn, L, r, e, w1);
    float p1 = gld32(L.P, c1);
    for (;;) {
        /* step i+2: geometry and temperature */
        const bool h2 = h1 && b1 == RB_ON && counter + 2 < limit;
        uint32_t c2 = 0u;
        float d2, w2;
        int b2 = RB_ON;
        if (h2)
            if (geo_core<XL>(L, r, c2, d2, xul)) b2 = boundary_move<TM>(cm, sn, L, r, e, w2);
        const float p2 = gld32(L.P, c2);
        /* step i+1: table entries */
        const KbRaw k1 = kb_fetch<TM>(sn, e.row, e.rank, nT, p1);
        /* step i: absorption, then the end of the ray, if any */
        float K0, B0;
        kb_finish<TM>(k0, K0, B0);
        phys_core(K0, B0, d0, e, tol, r.tra, de, status);
        counter += 1;
        nsteps += 1;
        if (b0 == RB_WALL) {
            wall_hit(e, r, w0, de);
            break;
        }
        if (b0 == RB_LOST) {
            r.tra = 0.f;
            break;
        }
        if (!(r.tra > tol) || !h1) break;
        k0 = k1; d0 = d1; w0 = w1; b0 = b1;
        h1 = h2; d1 = d2; w1 = w2; b1 = b2; p1 = p2;
    }
}

/* One ray marched on one grid level by a whole wavefront, every lane
 * holding the same ray (as per-lane values, so that the geometry stays on
 * the vector ALU): up to 64 steps of geometry run in every lane (each step
 * needs the position the one before left), lane s keeping the cell and
 * length of step s; the temperatures, table entries and exponentials of
 * those steps are then computed one step per lane, all at once; last, the
 * transmissivity is carried through the steps in order (lane s receives
 * the value before its step), the step that ends the ray found by a
 * ballot, and the tally terms, computed per lane, summed in step order.
 * Every product and sum is the one march_step makes, on the same values
 * and in the same order, so the tally, transmissivity and step count are
 * bit-identical to it.  Returns at the ray's end or at the level's step
 * cap; expects a live ray under the cap.  `xul`: the level's x faces in
 * LDS. */
template <int TM, class S>
__device__ __forceinline__ void march_coop(const DevCommon &cm, const S &sn, const DevLevel &L, Ray &r,
                                           const En &e, float tol, float &de, int &counter, int &nsteps,
                                           const float *xul, int lane) {
    /* Per-ray constants of the geometry (recomputed after a reflection):
     * index steps along each axis and the matching steps of the linear
     * cell index; the cell index itself is carried from step to step.  The
     * x face the ray heads for (fxc) and the next one along x (fxn, index
     * ixn) are read one step ahead of their use. */
    int dix, diy, diz, dcy, dcz;
    uint32_t cur, ixn;
    float fxc, fxn;
    auto setup = [&]() {
        dix = r.f0 ? -1 : 1;
        diy = r.f1 ? -1 : 1;
        diz = r.f2 ? -1 : 1;
        dcy = diy * L.im;
        dcz = diz * L.im * L.jm;
        cur = min((uint32_t)(r.ic - 1) + (uint32_t)L.im * ((uint32_t)(r.jc - 1) + (uint32_t)L.jm * (uint32_t)(r.kc - 1)),
                  (uint32_t)L.ncell - 1u);
        const uint32_t ix = min((uint32_t)(r.ic - r.f0), (uint32_t)L.im);
        ixn = min(ix + (uint32_t)dix, (uint32_t)L.im);
        fxc = xul[ix];
        fxn = xul[ixn];
    };
    setup();
    for (;;) {
        /* geometry of up to 64 steps (the arithmetic of geo_core_fx); only
         * the block's last step can meet a black or open face (the block
         * stops there) */
        uint32_t mcell = 0u;
        float mds = 0.f, wl = 0.f;
        int bl = RB_ON;
        const int nmax = min(64, L.maxs - __builtin_amdgcn_readfirstlane(counter));
        int n = 0;
        for (;;) {
            const float fy = (float)(r.jc - r.f1) * L.dy;
            const float fz = (float)(r.kc - r.f2) * L.dz;
            const float dsx = (fxc - r.xp) * r.isx;
            const float dsy = (fy - r.yp) * r.isy;
            const float dsz = (fz - r.zp) * r.isz;
            float ds = dsx;
            ds = (ds > dsy) ? dsy : ds;
            ds = (ds > dsz) ? dsz : ds;
            r.xp = (ds == dsx) ? fxc : fmaf(ds, r.sx, r.xp);
            r.yp = (ds == dsy) ? fy : fmaf(ds, r.sy, r.yp);
            r.zp = (ds == dsz) ? fz : fmaf(ds, r.sz, r.zp);
            const bool xa = dsx < dsy && dsx < dsz;
            const bool ya = !(dsx < dsy) && dsy < dsz;
            r.ic += xa ? dix : 0;
            r.jc += ya ? diy : 0;
            r.kc += (!xa && !ya) ? diz : 0;
            mcell = lane == n ? cur : mcell;
            mds = lane == n ? ds : mds;
            cur += xa ? (uint32_t)dix : (uint32_t)(ya ? dcy : dcz);
            fxc = xa ? fxn : fxc;
            ixn = xa ? min(ixn + (uint32_t)dix, (uint32_t)L.im) : ixn;
            fxn = xul[ixn];
            n += 1;
            const bool out = (uint32_t)(r.ic - 1) >= (uint32_t)L.im || (uint32_t)(r.jc - 1) >= (uint32_t)L.jm ||
                             (uint32_t)(r.kc - 1) >= (uint32_t)L.km;
            if (__builtin_amdgcn_readfirstlane((int)out)) {
                bl = boundary_move<TM>(cm, sn, L, r, e, wl);
                setup();
            }
            if (__builtin_amdgcn_readfirstlane(bl) != RB_ON || n >= nmax) break;
        }
        /* the absorption terms of the n steps, one per lane (the lanes
         * from n up work on cell 0 and are not read) */
        float K, B;
        kb_finish<TM>(kb_fetch<TM>(sn, e.row, e.rank, cm.nT, gld32(L.P, mcell)), K, B);
        float em1, ex;
        expm1_exp(-(K * mds), em1, ex);
        const float alpha = -em1;
        const float fac = (B - e.Bi) * e.invBi;
        /* transmissivity before each step: trl in lane s */
        float trl = 0.f, t = r.tra;
        for (int s = 0; s < n; ++s) {
            trl = lane == s ? t : trl;
            t = t * __int_as_float(__builtin_amdgcn_readlane(__float_as_int(ex), s));
        }
        /* the step that ends the ray: the transmissivity falls to tol or
         * below (dumped below it), or the block's last step met a black or
         * open face */
        const float tn = trl * ex;
        const bool ends = lane < n && (!(tn > tol) || (lane == n - 1 && bl != RB_ON));
        const unsigned long long em = __ballot(ends);
        const int m = em ? (int)__builtin_ctzll(em) + 1 : n;
        /* the tally terms of the steps taken, in order */
        const float term = ((e.W * trl) * alpha) * fac;
        for (int s = 0; s < m; ++s) de -= __int_as_float(__builtin_amdgcn_readlane(__float_as_int(term), s));
        counter += m;
        nsteps += m;
        if (em) {
            const float te = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(tn), m - 1));
            if (te < tol) {
                de -= (e.W * te) * __int_as_float(__builtin_amdgcn_readlane(__float_as_int(fac), m - 1));
                r.tra = 0.f;
            } else {
                r.tra = te;
            }
            if (m == n && bl == RB_WALL) wall_hit(e, r, wl, de);
            else if (m == n && bl == RB_LOST) r.tra = 0.f;
            return;
        }
        r.tra = t;
        if (counter >= L.maxs) return;
    }
}

/* march_core, then the boundary condition of the face crossed, if any */
template <int TM = TAB_RT, class S>
__device__ __forceinline__ void march_step(const DevCommon &cm, const S &sn, const DevLevel &L, Ray &r,
                                           const En &e, float tol, float &de, int &status) {
    if (march_core<TM>(cm, sn, L, r, e, tol, de, status)) boundary<TM>(cm, sn, L, r, e, de, status);
}

/* The boundary condition of the face crossed, as a branch over the face:
 * the same outcome and tally as boundary(), in the form a thread that
 * marches its own ray (plain variant, traced rays) runs fastest -- its
 * lanes cross faces at steps of their own, so the face code is read only
 * where one is crossed. */
__device__ __forceinline__ void boundary_thread(const DevCommon &cm, const DevSnap &sn, const DevLevel &L, Ray &r,
                                                const En &e, float &de, int &status) {
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
    } else if (bd == 3) {
        switch (face) {
            case FW: r.xp = 0.f; r.sx = -r.sx; r.isx = -r.isx; r.ic = 1; r.f0 = r.sx < 0.f; break;
            case FE: r.xp = cm.Lx; r.sx = -r.sx; r.isx = -r.isx; r.ic = L.im; r.f0 = r.sx < 0.f; break;
            case FS: r.yp = 0.f; r.sy = -r.sy; r.isy = -r.isy; r.jc = 1; r.f1 = r.sy < 0.f; break;
            case FN: r.yp = cm.Ly; r.sy = -r.sy; r.isy = -r.isy; r.jc = L.jm; r.f1 = r.sy < 0.f; break;
            case FB: r.zp = 0.f; r.sz = -r.sz; r.isz = -r.isz; r.kc = 1; r.f2 = r.sz < 0.f; break;
            default: r.zp = cm.Lz; r.sz = -r.sz; r.isz = -r.isz; r.kc = L.km; r.f2 = r.sz < 0.f; break;
        }
    } else if (bd == 2 || bd == 4) {
        float Bw, Kw;
        kb_pos(sn, e.row, e.rank, cm.nT, gld(L.P, wall_index(L, r, face)), Kw, Bw);
        wall_hit(e, r, Bw, de);
        status = RC_RAY_WALL;
    } else {
        r.tra = 0.f;
        status = RC_RAY_LOST;
    }
}

/* one march step of a thread that marches its own ray: march_core's
 * arithmetic with the table-position load issued as soon as the cell index
 * is known, then boundary_thread */
__device__ __forceinline__ void march_step_thread(const DevCommon &cm, const DevSnap &sn, const DevLevel &L, Ray &r,
                                                  const En &e, float tol, float &de, int &status) {
    const float fx = gld32(L.xu, min((uint32_t)(r.ic - r.f0), (uint32_t)L.im));
    const float fy = (float)(r.jc - r.f1) * L.dy;
    const float fz = (float)(r.kc - r.f2) * L.dz;
    const float dsx = (fx - r.xp) * r.isx;
    const float dsy = (fy - r.yp) * r.isy;
    const float dsz = (fz - r.zp) * r.isz;
    const uint32_t cell = min((uint32_t)(r.ic - 1) + __umul24((uint32_t)L.im, (uint32_t)(r.jc - 1) +
                                                                 __umul24((uint32_t)L.jm, (uint32_t)(r.kc - 1))),
                              (uint32_t)L.ncell - 1u);
    const float pc = gld32(L.P, cell);
    float ds = dsx;
    ds = (ds > dsy) ? dsy : ds;
    ds = (ds > dsz) ? dsz : ds;
    r.xp = (ds == dsx) ? fx : fmaf(ds, r.sx, r.xp);
    r.yp = (ds == dsy) ? fy : fmaf(ds, r.sy, r.yp);
    r.zp = (ds == dsz) ? fz : fmaf(ds, r.sz, r.zp);
    float K, B;
    kb_pos(sn, e.row, e.rank, cm.nT, pc, K, B);
    phys_core(K, B, ds, e, tol, r.tra, de, status);
    const bool xa = dsx < dsy && dsx < dsz;
    const bool ya = !(dsx < dsy) && dsy < dsz;
    const bool za = !xa && !ya;
    r.ic += xa ? (r.f0 ? -1 : 1) : 0;
    r.jc += ya ? (r.f1 ? -1 : 1) : 0;
    r.kc += za ? (r.f2 ? -1 : 1) : 0;
    if ((uint32_t)(r.ic - 1) >= (uint32_t)L.im || (uint32_t)(r.jc - 1) >= (uint32_t)L.jm ||
        (uint32_t)(r.kc - 1) >= (uint32_t)L.km)
        boundary_thread(cm, sn, L, r, e, de, status);
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
        march_step_thread(cm, sn, L, r, e, tol, de, status);
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
template <class D>
__device__ __forceinline__ void cell_ijk(const D &cm, int n, int &i, int &j, int &k) {
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
    const uint32_t tag = TAG_FLUID + 16u * (uint32_t)sn.snap;
    /* the ray's four uniforms: band and g (unless the sorted key is given)
     * and the direction (unless the Sobol sampler gives it) */
    U4 u{0.f, 0.f, 0.f, 0.f};
    if (!have_key || !cm.sampler) u = draw4(cm.k0, cm.k1, (uint32_t)h, (uint32_t)v, cid, tag);
    if (!have_key) key = band_key(cm, sn, false, u.c, u.d);
    if (cm.sampler) {
        /* the reference's random == 1 (emiss_angSobol, mc_gpu_noint.cu:1085-1092) */
        const uint4 sh = draw4_bits(cm.k0, cm.k1, SOBOL_SHIFT_H, (uint32_t)v, cid, tag);
        sobol_uniforms((uint32_t)h, sh.x, sh.y, u.a, u.b);
    }
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

/* grid sizes from D: the solve-wide DevCommon, or the snapshot's own */
template <class D>
__device__ __forceinline__ float cell_T(const D &cm, const DevSnap &sn, int i, int j, int k) {
    return gld(sn.T0, (size_t)i + (size_t)(cm.imax + 2) * ((size_t)j + (size_t)(cm.jmax + 2) * (size_t)k));
}
template <class D>
__device__ __forceinline__ float wall_T(const D &cm, const DevSnap &sn, int bound, int j, int k) {
    const size_t row = (size_t)(cm.imax + 2) * ((size_t)j + (size_t)(cm.jmax + 2) * (size_t)k);
    /* texel coordinate 1.0 (imax+1.0): halfway between ghost and first cell */
    return bound ? lerp(gld(sn.T0, row + cm.imax), gld(sn.T0, row + cm.imax + 1), 0.5f)
                 : lerp(gld(sn.T0, row), gld(sn.T0, row + 1), 0.5f);
}

}  // namespace rc

#endif
