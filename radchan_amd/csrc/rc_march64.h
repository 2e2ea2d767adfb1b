/*
 * rc_march64.h -- the full-FP64 march (param.h floatpoint 1, where the
 * reference compiles myfloat as double: param.h:64-65,158-166): ray
 * geometry, transmissivity, tallies and the per-ray constants in double;
 * the temperature field, the k/Ib tables and the band draws stay single
 * precision, as the reference's textures and float input are.
 *
 * One thread marches its rays in sequence (the reference's launch shape,
 * mc_gpu_noint.cu:55-243, :1316-1482); the march is march_ray
 * (:245-658) with its boundary conditions and the multigrid hand-off
 * (:210-224).  Every operation is IEEE double (+, -, *, /, sqrt, fma) or
 * the project-written expm1/exp and sin/cos of 2 pi u below, so the CPU
 * oracle's full-FP64 mode reproduces each ray bit for bit.
 */
#ifndef RC_MARCH64_H
#define RC_MARCH64_H

#include "rc_march.h"

namespace rc {

__device__ __forceinline__ double bits2d(uint64_t b) { return __longlong_as_double((long long)b); }

/* expm1(x) and exp(x) for x <= 0: x = n ln2 + r (Cody-Waite, two fma
 * steps), |r| <= ln2/2, expm1(r) by its Taylor series to r^13/13! (Horner),
 * scaled by 2^n, in two factors below 2^-1000 */
__device__ __forceinline__ void expm1_exp_d(double x, double &em1, double &e) {
    if (!(x < 0.0)) { em1 = x; e = 1.0; return; }
    if (x < -745.5) { em1 = -1.0; e = 0.0; return; }
    const double n = floor(x * 1.4426950408889634 + 0.5);
    double r = fma(-n, 6.93147180369123816490e-01, x);
    r = fma(-n, 1.90821492927058770002e-10, r);
    double p = fma(r, 1.0 / 6227020800.0, 1.0 / 479001600.0);
    p = fma(r, p, 1.0 / 39916800.0);
    p = fma(r, p, 1.0 / 3628800.0);
    p = fma(r, p, 1.0 / 362880.0);
    p = fma(r, p, 1.0 / 40320.0);
    p = fma(r, p, 1.0 / 5040.0);
    p = fma(r, p, 1.0 / 720.0);
    p = fma(r, p, 1.0 / 120.0);
    p = fma(r, p, 1.0 / 24.0);
    p = fma(r, p, 1.0 / 6.0);
    p = fma(r, p, 0.5);
    p = fma(r * r, p, r);
    const int ni = (int)n;
    if (ni < -1000) {
        const double s1 = bits2d((uint64_t)(ni + 600 + 1023) << 52);
        e = fma(s1, p, s1) * bits2d((uint64_t)(1023 - 600) << 52);
        em1 = -1.0;
        return;
    }
    const double s1 = bits2d((uint64_t)(ni + 1023) << 52);
    e = fma(s1, p, s1);
    em1 = (ni == 0) ? p : fma(s1, p, s1 - 1.0);
}

/* sin and cos of 2 pi u (u in [0,1], single precision): exact quadrant
 * split of 4u, Taylor polynomials to degree 17 / 18 on |y| <= pi/4 */
__device__ __forceinline__ void sincos2pi_d(float u, double &s, double &c) {
    const double x = 4.0 * (double)u;
    const double m = floor(x + 0.5);
    const double f = x - m;
    const double y = f * 1.5707963267948966;
    const double y2 = y * y;
    double ps = fma(y2, 1.0 / 355687428096000.0, -1.0 / 1307674368000.0);
    ps = fma(y2, ps, 1.0 / 6227020800.0);
    ps = fma(y2, ps, -1.0 / 39916800.0);
    ps = fma(y2, ps, 1.0 / 362880.0);
    ps = fma(y2, ps, -1.0 / 5040.0);
    ps = fma(y2, ps, 1.0 / 120.0);
    ps = fma(y2, ps, -1.0 / 6.0);
    const double sp = fma(y * y2, ps, y);
    double pc = fma(y2, 1.0 / 6402373705728000.0, -1.0 / 20922789888000.0);
    pc = fma(y2, pc, 1.0 / 87178291200.0);
    pc = fma(y2, pc, -1.0 / 479001600.0);
    pc = fma(y2, pc, 1.0 / 3628800.0);
    pc = fma(y2, pc, -1.0 / 40320.0);
    pc = fma(y2, pc, 1.0 / 720.0);
    pc = fma(y2, pc, -1.0 / 24.0);
    pc = fma(y2, pc, 0.5);
    const double cp = fma(-y2, pc, 1.0);
    switch ((int)m & 3) {
        case 0: s = sp;  c = cp;  break;
        case 1: s = cp;  c = -sp; break;
        case 2: s = -sp; c = -cp; break;
        default: s = -cp; c = sp; break;
    }
}

struct Ray64 {
    int ic, jc, kc, f0, f1, f2;
    double xp, yp, zp, sx, sy, sz, isx, isy, isz, tra;
};

/* per-ray constants: the ray's table row (k and Ib pairs of its band and
 * g), weight W = Ibmax ratioI and B(Ti) */
struct En64 {
    const float2 *row;
    double W, Bi;
};

/* K and B of the row at temperature T: clamp-to-edge linear interpolation
 * in double (the texture semantics of memory_copy_noint.cu:241-247) */
__device__ __forceinline__ void kb64(const DevSnap &sn, const float2 *row, int nT, double T, double &K, double &B) {
    const double pos = (T - sn.Tnb0d) * sn.invdTd;
    int t;
    double a;
    if (!(pos > 0.0)) { t = 0; a = 0.0; }
    else if (pos >= (double)(nT - 1)) { t = nT - 1; a = 0.0; }
    else { t = (int)pos; a = pos - (double)t; }
    const float2 v0 = row[t], v1 = row[t + 1];
    K = (double)v0.x + a * ((double)v1.x - (double)v0.x);
    B = (double)v0.y + a * ((double)v1.y - (double)v0.y);
}

__device__ __forceinline__ void set_dir64(Ray64 &r, double sx, double sy, double sz) {
    r.sx = sx; r.sy = sy; r.sz = sz;
    r.f0 = sx < 0.0; r.f1 = sy < 0.0; r.f2 = sz < 0.0;
    r.isx = 1.0 / sx; r.isy = 1.0 / sy; r.isz = 1.0 / sz;
}

/* one step of march_ray in double on level L (one iteration of the loop
 * at mc_gpu_noint.cu:252-654).  T holds the level's cell temperatures then
 * its black walls' (the layout of DevLevel.P).  A ray that ends here is
 * left with tra = 0 (absorbed below tl, at a black wall, or lost). */
__device__ __forceinline__ void step64(const DevCommon &cm, const DevSnap &sn, const DevLevel &L, const float *T,
                                       Ray64 &r, const En64 &e, double tl, double &de, int &steps, int &status) {
    {
        const double fx = (double)L.xu[r.ic - r.f0];
        const double fy = (double)((float)(r.jc - r.f1) * L.dy);
        const double fz = (double)((float)(r.kc - r.f2) * L.dz);
        const double dsx = (fx - r.xp) * r.isx;
        const double dsy = (fy - r.yp) * r.isy;
        const double dsz = (fz - r.zp) * r.isz;
        const float Tc = T[(size_t)(r.ic - 1) + (size_t)L.im * ((size_t)(r.jc - 1) + (size_t)L.jm * (r.kc - 1))];
        double ds = dsx;
        if (ds > dsy) ds = dsy;
        if (ds > dsz) ds = dsz;
        r.xp = (ds == dsx) ? fx : fma(ds, r.sx, r.xp);
        r.yp = (ds == dsy) ? fy : fma(ds, r.sy, r.yp);
        r.zp = (ds == dsz) ? fz : fma(ds, r.sz, r.zp);
        /* attenuation and the reciprocity tally (:358-370) */
        double K, B;
        kb64(sn, e.row, cm.nT, (double)Tc, K, B);
        const double x = K * ds;
        const double fac = (B - e.Bi) / e.Bi;
        double em, ex;
        expm1_exp_d(-x, em, ex);
        const double wta = e.W * r.tra * (-em);
        de -= wta * fac;
        r.tra = r.tra * ex;
        if (r.tra < tl) {
            de -= e.W * r.tra * fac;
            r.tra = 0.0;
            status = RC_RAY_ABSORBED;
        }
        steps += 1;
        /* one index along the crossed axis, with the tie rules */
        int axis;
        if (dsx < dsy) axis = (dsx < dsz) ? 0 : 2;
        else axis = (dsy < dsz) ? 1 : 2;
        int face = -1;
        if (axis == 0) {
            if (r.f0) { r.ic -= 1; if (r.ic == 0) face = FW; }
            else { r.ic += 1; if (r.ic == L.im + 1) face = FE; }
        } else if (axis == 1) {
            if (r.f1) { r.jc -= 1; if (r.jc == 0) face = FS; }
            else { r.jc += 1; if (r.jc == L.jm + 1) face = FN; }
        } else {
            if (r.f2) { r.kc -= 1; if (r.kc == 0) face = FB; }
            else { r.kc += 1; if (r.kc == L.km + 1) face = FT; }
        }
        if (face >= 0) {
            const int bd = cm.bd[face];
            if (bd == 1) {                                    /* periodic */
                switch (face) {
                    case FW: r.ic = L.im; r.xp = (double)cm.Lx; break;
                    case FE: r.ic = 1; r.xp = 0.0; break;
                    case FS: r.jc = L.jm; r.yp = (double)cm.Ly; break;
                    case FN: r.jc = 1; r.yp = 0.0; break;
                    case FB: r.kc = L.km; r.zp = (double)cm.Lz; break;
                    default: r.kc = 1; r.zp = 0.0; break;
                }
            } else if (bd == 3) {                             /* specular */
                switch (face) {
                    case FW: r.xp = 0.0; r.sx = -r.sx; r.isx = -r.isx; r.ic = 1; r.f0 = r.sx < 0.0; break;
                    case FE: r.xp = (double)cm.Lx; r.sx = -r.sx; r.isx = -r.isx; r.ic = L.im; r.f0 = r.sx < 0.0; break;
                    case FS: r.yp = 0.0; r.sy = -r.sy; r.isy = -r.isy; r.jc = 1; r.f1 = r.sy < 0.0; break;
                    case FN: r.yp = (double)cm.Ly; r.sy = -r.sy; r.isy = -r.isy; r.jc = L.jm; r.f1 = r.sy < 0.0; break;
                    case FB: r.zp = 0.0; r.sz = -r.sz; r.isz = -r.isz; r.kc = 1; r.f2 = r.sz < 0.0; break;
                    default: r.zp = (double)cm.Lz; r.sz = -r.sz; r.isz = -r.isz; r.kc = L.km; r.f2 = r.sz < 0.0; break;
                }
            } else if (bd == 2 || bd == 4) {
                /* black wall: a fixed temperature's Ib (the row's entry of
                 * the face), or the table at the wall's temperature
                 * (adiabatic x walls, bd 4 z walls) */
                const bool adiab = (face == FW && cm.adiab[0]) || (face == FE && cm.adiab[1]) || bd == 4;
                double Bw;
                if (adiab) {
                    const Ray rr{r.ic, r.jc, r.kc, r.f0, r.f1, r.f2};
                    double Kw;
                    kb64(sn, e.row, cm.nT, (double)T[wall_index(L, rr, face)], Kw, Bw);
                } else {
                    Bw = (double)e.row[cm.nT + 1 + 2 * face].y;
                }
                de -= e.W * r.tra * ((Bw - e.Bi) / e.Bi);
                r.tra = 0.0;
                status = RC_RAY_WALL;
            } else {                                          /* 5: lost */
                r.tra = 0.0;
                status = RC_RAY_LOST;
            }
        }
    }
}

/* one grid level of march_ray in double; returns the steps taken on it */
__device__ inline int march_level64(const DevCommon &cm, const DevSnap &sn, const DevLevel &L, const float *T,
                                    Ray64 &r, const En64 &e, float tol, double &de, int &steps, int &status) {
    const double tl = (double)tol;
    int counter = 0;
    while (r.tra > tl && counter < L.maxs) {
        step64(cm, sn, L, T, r, e, tl, de, steps, status);
        counter += 1;
    }
    return counter;
}

/* the indices on level g + 1 of a ray leaving level g (:215-223) */
__device__ __forceinline__ void coarsen64(const DevSnap &sn, int g, Ray64 &r) {
    const DevLevel &L = sn.lev[g], &M = sn.lev[g + 1];
    const int ri = L.im / M.im, rj = L.jm / M.jm, rk = L.km / M.km;
    r.ic = (r.ic + ri - 1) / ri;
    r.jc = (r.jc + rj - 1) / rj;
    r.kc = (r.kc + rk - 1) / rk;
}

/* the multigrid loop (:210-224): a ray still alive after a level's step
 * cap moves on to the next level with coarsened indices */
__device__ inline void march_all64(const DevCommon &cm, const DevSnap &sn, Ray64 &r, const En64 &e, float tol,
                                   double &de, int &steps, int &status) {
    status = RC_RAY_STOPPED;
    for (int g = 0; g < cm.grid_num; ++g) {
        const DevLevel &L = sn.lev[g];
        const int n = march_level64(cm, sn, L, sn.T_all + L.poff, r, e, tol, de, steps, status);
        if (!(r.tra > (double)tol)) return;
        if (g == cm.grid_num - 1) {
            if (n >= L.maxs) status = RC_RAY_CAPPED;
            return;
        }
        coarsen64(sn, g, r);
    }
}

/* per-ray constants in double (:154-186, :1419-1424) */
__device__ __forceinline__ void energy64(const DevCommon &cm, const DevSnap &sn, uint32_t key, float Ti, bool surface,
                                         int bound, En64 &e) {
    e.row = sn.kb + (size_t)key * tab_stride(cm.nT);
    double Ki, Bi, Km, Bm;
    kb64(sn, e.row, cm.nT, (double)Ti, Ki, Bi);
    kb64(sn, e.row, cm.nT, (double)sn.Tmax, Km, Bm);
    e.Bi = Bi;
    e.W = surface ? (bound ? sn.Ibmax_e64 : sn.Ibmax_w64) * (Bi / Bm) : sn.Ibmax64 * (Bi / Bm) * (Ki / Km);
}

/* a fluid ray's start at the centre of cell (i,j,k): direction, position
 * and per-ray constants (kernel_fluid :115-186) */
__device__ __forceinline__ void fluid_setup64(const DevCommon &cm, const DevSnap &sn, int i, int j, int k,
                                              uint32_t cid, int v, int h, uint32_t key, bool have_key, float Ti,
                                              Ray64 &r, En64 &e) {
    const uint32_t tag = TAG_FLUID + 16u * (uint32_t)sn.snap;
    U4 u{0.f, 0.f, 0.f, 0.f};
    if (!have_key || !cm.sampler) u = draw4(cm.k0, cm.k1, (uint32_t)h, (uint32_t)v, cid, tag);
    if (!have_key) key = band_key(cm, sn, false, u.c, u.d);
    if (cm.sampler) {
        const uint4 sh = draw4_bits(cm.k0, cm.k1, SOBOL_SHIFT_H, (uint32_t)v, cid, tag);
        sobol_uniforms((uint32_t)h, sh.x, sh.y, u.a, u.b);
    }
    const double ct = 1.0 - 2.0 * (double)u.b;
    const double st = 2.0 * sqrt((double)u.b * (1.0 - (double)u.b));
    double sn2, cs2;
    sincos2pi_d(u.a, sn2, cs2);
    double sx = ct, sy = st * cs2, sz = st * sn2;
    if (sx == 0.0) sx = (double)1e-10f;
    if (sy == 0.0) sy = (double)1e-10f;
    if (sz == 0.0) sz = (double)1e-10f;
    set_dir64(r, sx, sy, sz);
    r.ic = i; r.jc = j; r.kc = k;
    r.xp = (double)sn.x0[i]; r.yp = (double)sn.y0[j]; r.zp = (double)sn.z0[k];
    r.tra = 1.0;
    energy64(cm, sn, key, Ti, false, 0, e);
}

/* one fluid ray from the centre of cell (i,j,k) (kernel_fluid :115-224) */
__device__ inline double fluid_ray64(const DevCommon &cm, const DevSnap &sn, int i, int j, int k, uint32_t cid, int v,
                                     int h, uint32_t key, bool have_key, float Ti, int &steps) {
    Ray64 r;
    En64 e;
    fluid_setup64(cm, sn, i, j, k, cid, v, h, key, have_key, Ti, r, e);
    double de = 0.0;
    int status;
    march_all64(cm, sn, r, e, cm.toll, de, steps, status);
    return de;
}

/* one ray from the centre of wall cell (j,k) (kernel_surface :1381-1462) */
__device__ inline double surface_ray64(const DevCommon &cm, const DevSnap &sn, int bound, int j, int k, uint32_t cid,
                                       int v, int h, uint32_t key, bool have_key, float Tw, int &steps) {
    const U4 u = draw4(cm.k0, cm.k1, (uint32_t)h, (uint32_t)v, cid,
                       (bound ? TAG_EAST : TAG_WEST) + 16u * (uint32_t)sn.snap);
    if (!have_key) key = band_key(cm, sn, true, u.c, u.d);
    const double a = sqrt(1.0 - (double)u.b), b = sqrt((double)u.b);
    double sn2, cs2;
    sincos2pi_d(u.a, sn2, cs2);
    const double st = bound ? a : b;
    double sx = bound ? -b : a, sy = st * cs2, sz = st * sn2;
    if (fabs(sx) <= 1e-10) sx = (double)((float)(1 - 2 * bound) * 1e-10f);
    if (fabs(sy) <= 1e-10) sy = (double)((float)((sy > 0.0) - (sy <= 0.0)) * 1e-10f);
    if (fabs(sz) <= 1e-10) sz = (double)((float)((sz > 0.0) - (sz <= 0.0)) * 1e-10f);
    Ray64 r;
    set_dir64(r, sx, sy, sz);
    r.ic = bound ? cm.imax : 1; r.jc = j; r.kc = k;
    r.xp = bound ? (double)cm.Lx : 0.0; r.yp = (double)sn.y0[j]; r.zp = (double)sn.z0[k];
    r.tra = 1.0;
    En64 e;
    energy64(cm, sn, key, Tw, true, bound, e);
    /* on-face index adjustment (:1429-1441) */
    const DevLevel &L0 = sn.lev[0];
    if (r.f0 && r.xp == (double)L0.xu[r.ic - 1]) r.ic -= 1;
    if (r.f1 && r.yp == (double)((float)(r.jc - 1) * L0.dy)) r.jc -= 1;
    if (r.f2 && r.zp == (double)((float)(r.kc - 1) * L0.dz)) r.kc -= 1;
    if (!r.f0 && r.xp == (double)L0.xu[r.ic]) r.ic += 1;
    if (!r.f1 && r.yp == (double)((float)r.jc * L0.dy)) r.jc += 1;
    if (!r.f2 && r.zp == (double)((float)r.kc * L0.dz)) r.kc += 1;
    double de = 0.0;
    int status;
    march_all64(cm, sn, r, e, cm.tollS, de, steps, status);
    return de;
}

}  // namespace rc

#endif
