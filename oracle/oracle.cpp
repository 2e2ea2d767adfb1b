/*
 * oracle.cpp -- CPU restatement of RadChan's narrow-band correlated-k OERM
 * Monte Carlo solver (src/montecarlo).  TEST INFRASTRUCTURE ONLY: the checker
 * for the HIP library and the timed CPU baseline; never part of the product.
 *
 * Reference files restated (paths relative to the RadChan tree):
 *   src/montecarlo/mc_gpu_noint.cu    kernel_fluid :55-243, march_ray :245-658,
 *                                     mc_gpu_ :660-874, find_band :1103-1169,
 *                                     find_bandS :1170-1233, wave_find :1234-1267,
 *                                     sort_idx :1278-1294, kernel_surface :1316-1482
 *   src/montecarlo/memory_copy_noint.cu  I_blackC :288-297, texture setup :131-286
 *   src/montecarlo/read_tables.cpp    readT :21-95 / :98-175, readS :188-220
 *   src/montecarlo/definitions.h      Gridn::mk_grid :68-136
 *   src/param/param.h, src/param/NarrowBand.h
 *
 * Deliberate, documented departures from the reference (the GPU library
 * follows the same ones):
 *  - randomness: Philox4x32-10 (Salmon et al., SC'11) keyed by
 *    (seed; ray, run, cell, stream) instead of cuRAND XORWOW/Sobol;
 *  - texture filtering: explicit FP32 lerp with clamp-to-edge, full-precision
 *    weights (CUDA uses 8-bit weights);
 *  - face distances use a per-ray reciprocal of the direction cosines;
 *  - 1-exp(-x) is computed as -expm1(-x) (no cancellation for thin cells);
 *  - the reciprocity factor is (B-Bi)/Bi, exactly zero when B == Bi;
 *  - srt band enumeration is drawn per (cell, run) instead of once per GPU
 *    thread (mc_gpu_noint.cu:89-91);
 *  - specular walls reflect correctly on all faces (the reference zeroes or
 *    overwrites components on y/z, :541-547,565-571,592-598,630-636);
 *  - the surface band special case uses the surface CDF (reference :1189 uses
 *    the fluid one) and the surface g CDF is the cumulative wq (the reference
 *    constant probq, :34-36, has a transposed digit in its second entry);
 *  - variance uses the reference normalisation 1/((n-1/2)(n-1)) (:238) but is
 *    left to the caller; the oracle returns per-run tallies.
 */
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double kStefan = 5.670373e-08;                 /* param.h:85 */
constexpr double kPi = 3.14159265358979323846;           /* param.h:88 */
constexpr double kC1 = 3.741771790075259e-16;            /* memory_copy_noint.cu:292 */
constexpr double kC2 = 0.014387741858429;                /* memory_copy_noint.cu:293 */

/* Planted mistakes (orc_plant): test-only switches that make the oracle
 * commit one plausible error, so that the tests can show their closed-form
 * checks reject it.  0 = none (always, outside those tests). */
enum {
    PLANT_NONE = 0,
    PLANT_GREY_ROW = 1,        /* grey band CDF of row i built at Tnb[i+1]          */
    PLANT_GREY_WVL = 2,        /* grey band CDF with Ib at the band's left edge     */
    PLANT_WALL_ROW = 3,        /* wall band CDF (readS) at row t*+1                 */
    PLANT_WALL_G_UNIFORM = 4,  /* wall g CDF uniform instead of cumulative wq       */
    PLANT_TEX_NO_SHIFT = 5,    /* texture fetch without the half-texel shift        */
    PLANT_TMAP_NO_GHOST = 6,   /* coarse-cell texel coordinate without the ghost +1 */
    PLANT_XWALL_GHOST = 7,     /* adiabatic x-wall T at the ghost texel centre      */
};
int g_plant = PLANT_NONE;

enum { FW = 0, FE = 1, FS = 2, FN = 3, FB = 4, FT = 5 };
enum { TAG_FLUID = 0, TAG_WEST = 1, TAG_EAST = 2 };

/* ---------------- Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) -------- */
inline void philox_round(uint32_t c[4], uint32_t k0, uint32_t k1) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    const uint32_t n0 = hi1 ^ c[1] ^ k0;
    const uint32_t n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
}
inline void philox10(uint32_t c[4], uint32_t k0, uint32_t k1) {
    for (int r = 0; r < 10; ++r) {
        philox_round(c, k0, k1);
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}
/* uniform on (0,1], a multiple of 2^-24 (same range convention as curand_uniform) */
inline float u01f(uint32_t x) { return (float)((x >> 8) + 1u) * 0x1p-24f; }

inline void draw4(uint64_t seed, uint32_t h, uint32_t v, uint32_t cell, uint32_t tag, float u[4]) {
    uint32_t c[4] = {h, v, cell, tag};
    philox10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    for (int q = 0; q < 4; ++q) u[q] = u01f(c[q]);
}

/* Sobol directions (the reference's random == 1, emiss_angSobol,
 * mc_gpu_noint.cu:1085-1092, with cuRAND's Sobol32 generators of
 * :84-88 and :813-817): point h of the first two dimensions of the Sobol
 * sequence -- dimension 1 the van der Corput sequence (direction numbers
 * 2^(31-k)), dimension 2 Joe and Kuo's first primitive polynomial x + 1
 * with m_1 = 1 (v_1 = 2^31, v_k = v_(k-1) xor v_(k-1) >> 1) -- digitally
 * shifted (xor) by two Philox words of counter (2^32 - 1, v, cell, tag),
 * one shift per (cell, run), so that the runs are independent estimates
 * (the reference starts every thread's sequence at its thread index,
 * quirk A9). */
inline void sobol2_point(uint32_t h, uint32_t *x0, uint32_t *x1) {
    uint32_t r = 0u, y = 0u, v = 0x80000000u;
    for (int k = 0; k < 32; ++k)
        if (h >> k & 1u) r |= 0x80000000u >> k;
    for (uint32_t g = h; g; g >>= 1) {
        if (g & 1u) y ^= v;
        v ^= v >> 1;
    }
    *x0 = r;
    *x1 = y;
}

inline void sobol_dir_uniforms(uint64_t seed, uint32_t h, uint32_t v, uint32_t cell, uint32_t tag, float *a,
                               float *b) {
    uint32_t c[4] = {0xFFFFFFFFu, v, cell, tag};
    philox10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    uint32_t x0, x1;
    sobol2_point(h, &x0, &x1);
    *a = u01f(x0 ^ c[0]);
    *b = u01f(x1 ^ c[1]);
}

/* ---------------- project math: identical op sequence on GPU -------------- */
inline float bits2f(uint32_t b) { float f; std::memcpy(&f, &b, 4); return f; }

/* sin and cos of 2*pi*u for u in (0,1]: quadrant split on 4u (exact), octant
 * fold, Taylor polynomials of degree 9/10 on [0, pi/4]. */
inline void sincos2pi_f(float u, float *s, float *c) {
    const float x = u * 4.0f;
    int q = (int)x;
    float f = x - (float)q;
    const bool fold = f > 0.5f;
    if (fold) f = 1.0f - f;
    const float y = f * 1.57079637f;
    const float y2 = y * y;
    float ps = fmaf(y2, (float)(1.0 / 362880.0), (float)(-1.0 / 5040.0));
    ps = fmaf(y2, ps, (float)(1.0 / 120.0));
    ps = fmaf(y2, ps, (float)(-1.0 / 6.0));
    const float sp = fmaf(y * y2, ps, y);
    float pc = fmaf(y2, (float)(-1.0 / 3628800.0), (float)(1.0 / 40320.0));
    pc = fmaf(y2, pc, (float)(-1.0 / 720.0));
    pc = fmaf(y2, pc, (float)(1.0 / 24.0));
    pc = fmaf(y2, pc, -0.5f);
    const float cp = fmaf(y2, pc, 1.0f);
    const float a = fold ? cp : sp;   /* sin of the quadrant-local angle */
    const float b = fold ? sp : cp;   /* cos of the quadrant-local angle */
    switch (q & 3) {
        case 0: *s = a;  *c = b;  break;
        case 1: *s = b;  *c = -a; break;
        case 2: *s = -a; *c = -b; break;
        default: *s = -b; *c = a; break;
    }
}

/* expm1(x) and exp(x): x = n ln2 + r (Cody-Waite with fma), |r| <= ln2/2,
 * expm1(r) by its Taylor series to r^9/9!. */
inline void expm1_exp_f(float x, float *em1, float *e) {
    if (x < -87.0f) { *em1 = -1.0f; *e = 0.0f; return; }
    if (x > 88.0f) x = 88.0f;
    const float n = rintf(x * 1.44269502f);
    float r = fmaf(-n, 0.693145752f, x);
    r = fmaf(-n, 1.42860677e-6f, r);
    float p = fmaf(r, (float)(1.0 / 362880.0), (float)(1.0 / 40320.0));
    p = fmaf(r, p, (float)(1.0 / 5040.0));
    p = fmaf(r, p, (float)(1.0 / 720.0));
    p = fmaf(r, p, (float)(1.0 / 120.0));
    p = fmaf(r, p, (float)(1.0 / 24.0));
    p = fmaf(r, p, (float)(1.0 / 6.0));
    p = fmaf(r, p, 0.5f);
    p = fmaf(r * r, p, r);
    const int ni = (int)n;
    const float sc = bits2f((uint32_t)(ni + 127) << 23);
    *e = fmaf(sc, p, sc);
    *em1 = (ni == 0) ? p : fmaf(sc, p, sc - 1.0f);
}

/* ---------------- FP64 project math (the full-FP64 path, floatpoint 1) ----
 * The same operation sequence compiles on the GPU, so the FP64 paths agree
 * bit for bit; accuracy against libm is a test (a few ulp). */
inline double bits2d(uint64_t b) { double d; std::memcpy(&d, &b, 8); return d; }

/* expm1(x) and exp(x) for x <= 0: x = n ln2 + r (Cody-Waite, two fma
 * steps), |r| <= ln2/2, expm1(r) by its Taylor series to r^13/13! (Horner
 * with fma), then scaled by 2^n (two factors below the normal range). */
inline void expm1_exp_d(double x, double *em1, double *e) {
    if (!(x < 0.0)) { *em1 = x; *e = 1.0; return; }          /* x = 0 (or -0) */
    if (x < -745.5) { *em1 = -1.0; *e = 0.0; return; }
    const double n = std::floor(x * 1.4426950408889634 + 0.5);
    double r = std::fma(-n, 6.93147180369123816490e-01, x);
    r = std::fma(-n, 1.90821492927058770002e-10, r);
    double p = std::fma(r, 1.0 / 6227020800.0, 1.0 / 479001600.0);
    p = std::fma(r, p, 1.0 / 39916800.0);
    p = std::fma(r, p, 1.0 / 3628800.0);
    p = std::fma(r, p, 1.0 / 362880.0);
    p = std::fma(r, p, 1.0 / 40320.0);
    p = std::fma(r, p, 1.0 / 5040.0);
    p = std::fma(r, p, 1.0 / 720.0);
    p = std::fma(r, p, 1.0 / 120.0);
    p = std::fma(r, p, 1.0 / 24.0);
    p = std::fma(r, p, 1.0 / 6.0);
    p = std::fma(r, p, 0.5);
    p = std::fma(r * r, p, r);                                /* expm1(r) */
    const int ni = (int)n;                                    /* -1075 .. 0 */
    if (ni < -1000) {                                         /* exp(x) < 2^-1000: scaled twice */
        const double s1 = bits2d((uint64_t)(ni + 600 + 1023) << 52);
        *e = std::fma(s1, p, s1) * bits2d((uint64_t)(1023 - 600) << 52);
        *em1 = -1.0;
        return;
    }
    const double s1 = bits2d((uint64_t)(ni + 1023) << 52);
    *e = std::fma(s1, p, s1);
    *em1 = (ni == 0) ? p : std::fma(s1, p, s1 - 1.0);
}

/* sin and cos of 2 pi u for u in [0,1] given in single precision: 4u and
 * its quadrant split are exact; the quadrant-local angle f pi/2, |f| <= 1/2,
 * by Taylor polynomials to degree 17 / 18. */
inline void sincos2pi_d(float u, double *s, double *c) {
    const double x = 4.0 * (double)u;
    const double m = std::floor(x + 0.5);
    const double f = x - m;                                   /* exact, in [-1/2, 1/2] */
    const double y = f * 1.5707963267948966;
    const double y2 = y * y;
    double ps = std::fma(y2, 1.0 / 355687428096000.0, -1.0 / 1307674368000.0);
    ps = std::fma(y2, ps, 1.0 / 6227020800.0);
    ps = std::fma(y2, ps, -1.0 / 39916800.0);
    ps = std::fma(y2, ps, 1.0 / 362880.0);
    ps = std::fma(y2, ps, -1.0 / 5040.0);
    ps = std::fma(y2, ps, 1.0 / 120.0);
    ps = std::fma(y2, ps, -1.0 / 6.0);
    const double sp = std::fma(y * y2, ps, y);
    double pc = std::fma(y2, 1.0 / 6402373705728000.0, -1.0 / 20922789888000.0);
    pc = std::fma(y2, pc, 1.0 / 87178291200.0);
    pc = std::fma(y2, pc, -1.0 / 479001600.0);
    pc = std::fma(y2, pc, 1.0 / 3628800.0);
    pc = std::fma(y2, pc, -1.0 / 40320.0);
    pc = std::fma(y2, pc, 1.0 / 720.0);
    pc = std::fma(y2, pc, -1.0 / 24.0);
    pc = std::fma(y2, pc, 0.5);
    const double cp = std::fma(-y2, pc, 1.0);
    switch ((int)m & 3) {
        case 0: *s = sp;  *c = cp;  break;
        case 1: *s = cp;  *c = -sp; break;
        case 2: *s = -sp; *c = -cp; break;
        default: *s = -cp; *c = sp; break;
    }
}

/* Planck spectral intensity per unit wavenumber (nu in cm^-1), computed in
 * FP64 and rounded once; I_blackC, memory_copy_noint.cu:288-297. */
inline float planck_f(float T, float nu) {
    const double n100 = 100.0 * (double)nu;
    return (float)(1.0 / kPi * kC1 * n100 * n100 * n100 / std::expm1(kC2 * n100 / (double)T));
}
inline double planck_d(double T, double nu) {
    const double n100 = 100.0 * nu;
    return 1.0 / kPi * kC1 * n100 * n100 * n100 / std::expm1(kC2 * n100 / T);
}

/* first index i with cdf[i] > R, R == 1 -> n-1, past the end -> n-1
 * (the bisection of find_band, mc_gpu_noint.cu:1116-1130, with its special
 * cases; the clamp replaces an out-of-range read). */
inline int upper_bound_f(const float *cdf, int n, float R) {
    int lo = 0, hi = n;               /* answer in [lo, hi] */
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (R < cdf[mid]) hi = mid; else lo = mid + 1;
    }
    if (R == 1.0f) lo = n - 1;
    if (lo > n - 1) lo = n - 1;
    return lo;
}

inline uint32_t fnv(uint32_t h, uint32_t w) { return (h ^ w) * 16777619u; }

inline float lerpf_(float a, float b, float w) { return fmaf(w, b - a, a); }

/* table position of a temperature: clamp-to-edge, as the reference textures
 * (memory_copy_noint.cu:241-247, mc_gpu_noint.cu:784-791) */
inline void tpos_f(float T, float Tnb0, float invdT, int nT, int *t, float *a) {
    const float pos = (T - Tnb0) * invdT;
    if (!(pos > 0.0f)) { *t = 0; *a = 0.0f; return; }
    if (pos >= (float)(nT - 1)) { *t = nT - 1; *a = 0.0f; return; }
    const int ti = (int)pos;
    *t = ti; *a = pos - (float)ti;
}
inline void tpos_d(double T, double Tnb0, double invdT, int nT, int *t, double *a) {
    const double pos = (T - Tnb0) * invdT;
    if (!(pos > 0.0)) { *t = 0; *a = 0.0; return; }
    if (pos >= (double)(nT - 1)) { *t = nT - 1; *a = 0.0; return; }
    const int ti = (int)pos;
    *t = ti; *a = pos - (double)ti;
}

struct Level {
    int im, jm, km, maxs;
    int rx, ry, rz;                     /* imax/im etc. (int division, :334-336) */
    std::vector<float> x, xu, y, yu, z, zu;
    std::vector<float> Ts, Tws;         /* cubic-spline restriction (empty: texture mean) */
};

}  // namespace

struct orc_ctx {
    orc_params p;
    int zstart, kloc;
    int nB, nT, nQ;
    std::vector<float> T;               /* (imax+2)(jmax+2)(kmax+2) */
    std::vector<Level> lev;
    float Tmax, kappamax, Ibmax, Ibmax_s[2];
    double kappamax64, Ibmax64, Ibmax_s64[2];
    int tm, tstar;
    uint32_t tagoff;                    /* 16 * snapshot index              */
    float Tnb0, invdT;
    double Tnb0d, invdTd;
    std::vector<float> Tnb;
    std::vector<int32_t> idx, rank_of;  /* idx[rank] = band, rank_of[band] = rank */
    std::vector<float> ktab;            /* [rank][g][t], nT+1 entries (edge pad) */
    std::vector<float> btab;            /* [rank][t],    nT+1 entries            */
    std::vector<float> cdf_b;           /* fluid band CDF row t* (orig order)    */
    std::vector<float> cdf_g;           /* [orig band][nQ] at row t*             */
    std::vector<float> cdf_bs;          /* surface band CDF row t*               */
    std::vector<float> cdf_gs;          /* [nQ] cumulative wq                    */
    std::vector<float> bfix[6];         /* fixed-T wall Ib per rank              */
    std::vector<float> wq, wvl, wvr;    /* quadrature weights, band edges        */
    bool grey_true;
};

namespace {

inline float Tat(const orc_ctx &c, int i, int j, int k) {
    const orc_params &p = c.p;
    return c.T[(size_t)i + (size_t)(p.imax + 2) * ((size_t)j + (size_t)(p.jmax + 2) * (size_t)k)];
}

/* linear texture fetch of the temperature volume at unnormalised coordinates
 * (cudaFilterModeLinear + cudaAddressModeClamp, memory_copy_noint.cu:159-165):
 * texel i has its centre at i+0.5. */
float tex3_T(const orc_ctx &c, float tx, float ty, float tz) {
    const orc_params &p = c.p;
    const int n[3] = {p.imax + 2, p.jmax + 2, p.kmax + 2};
    const float crd[3] = {tx, ty, tz};
    int i0[3], i1[3];
    float w[3];
    for (int d = 0; d < 3; ++d) {
        const float b = (g_plant == PLANT_TEX_NO_SHIFT) ? crd[d] : crd[d] - 0.5f;
        const float fl = std::floor(b);
        w[d] = b - fl;
        int lo = (int)fl, hi = lo + 1;
        lo = std::min(std::max(lo, 0), n[d] - 1);
        hi = std::min(std::max(hi, 0), n[d] - 1);
        i0[d] = lo; i1[d] = hi;
    }
    const float c00 = lerpf_(Tat(c, i0[0], i0[1], i0[2]), Tat(c, i1[0], i0[1], i0[2]), w[0]);
    const float c10 = lerpf_(Tat(c, i0[0], i1[1], i0[2]), Tat(c, i1[0], i1[1], i0[2]), w[0]);
    const float c01 = lerpf_(Tat(c, i0[0], i0[1], i1[2]), Tat(c, i1[0], i0[1], i1[2]), w[0]);
    const float c11 = lerpf_(Tat(c, i0[0], i1[1], i1[2]), Tat(c, i1[0], i1[1], i1[2]), w[0]);
    const float c0 = lerpf_(c00, c10, w[1]);
    const float c1 = lerpf_(c01, c11, w[1]);
    return lerpf_(c0, c1, w[2]);
}

/* mapped cell-centre texel coordinate of a coarse index (:334-336) */
inline float tmap(int ic, int r) {
    return (g_plant == PLANT_TMAP_NO_GHOST) ? ((float)ic - 0.5f) * (float)r : ((float)ic - 0.5f) * (float)r + 1.0f;
}

/* texel x coordinate of the adiabatic wall temperature: halfway between the
 * ghost plane and the first (last) cell, x = 1.0 (imax + 1.0)
 * (mc_gpu_noint.cu:391-399, 422-430) */
inline float xwall_tex(const orc_params &p, int bound) {
    if (g_plant == PLANT_XWALL_GHOST) return bound ? (float)p.imax + 1.5f : 0.5f;
    return bound ? (float)p.imax + 1.0f : 1.0f;
}

/* Gridn::mk_grid, definitions.h:68-136 */
void build_level(const orc_params &p, int g, Level &L) {
    const int im = p.maxi[g], jm = p.maxj[g], km = p.maxk[g], im0 = p.maxi[0];
    L.im = im; L.jm = jm; L.km = km; L.maxs = p.maxs[g];
    L.rx = p.imax / im; L.ry = p.jmax / jm; L.rz = p.kmax / km;
    L.x.assign(im + 2, 0.f); L.xu.assign(im + 1, 0.f);
    L.y.assign(jm + 2, 0.f); L.yu.assign(jm + 1, 0.f);
    L.z.assign(km + 2, 0.f); L.zu.assign(km + 1, 0.f);
    const float dy = p.Ly * 1.0f / (float)jm;
    const float dz = p.Lz * 1.0f / (float)km;
    for (int i = 0; i < jm + 1; ++i) L.yu[i] = (float)i * dy;
    for (int i = 0; i < km + 1; ++i) L.zu[i] = (float)i * dz;
    L.y[0] = -dy / 2.0f;
    L.z[0] = -dz / 2.0f;
    for (int i = 1; i < jm + 1; ++i) L.y[i] = (L.yu[i - 1] + L.yu[i]) / 2.0f;
    for (int i = 1; i < km + 1; ++i) L.z[i] = (L.zu[i - 1] + L.zu[i]) / 2.0f;
    L.y[jm + 1] = L.y[jm] + dy;
    L.z[km + 1] = L.z[km] + dz;
    /* parabolic cell-size law dx(r) = 0.5 - f (r-1/2)^2, normalised and
     * mirrored (definitions.h:108-123; the DNS grid mk_grid.f:13-24) */
    std::vector<float> xu0(im0 + 1, 0.f);
    xu0[0] = 0.f;
    for (int i = 1; i < im0 / 2 + 1; ++i) {
        const float r = (1.f * (float)i) / (float)im0;
        const float dx = (float)(0.5 - p.fact_mesh * ((double)r - 0.5) * ((double)r - 0.5));
        xu0[i] = xu0[i - 1] + dx;
    }
    const float xnorm = xu0[im0 / 2];
    for (int i = 1; i < im0 / 2 + 1; ++i)
        xu0[i] = (float)((double)p.Lx / 2. * (double)xu0[i] / (double)xnorm);
    for (int i = im0; i > im0 / 2; --i) xu0[i] = p.Lx - xu0[im0 - i];
    L.xu[0] = 0.f;
    for (int i = 1; i < im + 1; ++i) L.xu[i] = xu0[(size_t)i * im0 / im];
    for (int i = 1; i < im + 1; ++i) L.x[i] = (float)(0.5 * (double)(L.xu[i] + L.xu[i - 1]));
    L.x[0] = -L.x[1];
    L.x[im + 1] = L.xu[im] + (L.xu[im] - L.x[im]);
}

/* ------------------------------------------------------------------------- */
/* Ray marching, templated on the geometry type G (float: modes a/b; double:
 * full FP64, mode c).  In float mode a double "shadow" carries the FP64
 * energy along the FP32 geometry and decisions. */

template <class G> struct Ray {
    int ic, jc, kc;
    G xp, yp, zp;
    G sx, sy, sz, isx, isy, isz;
    int f0, f1, f2;
    G tra;
    double tra64;
};

struct Energy {           /* per-ray constants */
    int rank, g;
    float W, Bi, invBi;   /* FP32 path  */
    double W64, Bi64;     /* FP64 path  */
    const float *kt;      /* k(t) of (rank,g), nT+1 entries */
    const float *bt;      /* Ib(t) of rank,     nT+1 entries */
};

struct Tally {
    float de;
    double de64, sumabs;
    int steps;
    uint32_t hash;
};

template <class G> inline G face_x(const Level &L, int i) { return (G)L.xu[i]; }

/* K and B of the current band at temperature T */
inline void kb_f(const orc_ctx &c, const Energy &e, float T, float *K, float *B) {
    int t; float a;
    tpos_f(T, c.Tnb0, c.invdT, c.nT, &t, &a);
    *K = lerpf_(e.kt[t], e.kt[t + 1], a);
    *B = lerpf_(e.bt[t], e.bt[t + 1], a);
}
inline void kb_d(const orc_ctx &c, const Energy &e, double T, double *K, double *B) {
    int t; double a;
    tpos_d(T, c.Tnb0d, c.invdTd, c.nT, &t, &a);
    *K = (double)e.kt[t] + a * ((double)e.kt[t + 1] - (double)e.kt[t]);
    *B = (double)e.bt[t] + a * ((double)e.bt[t + 1] - (double)e.bt[t]);
}

/* Wall / dump tally with transmissivity tra at emission Bw: De -= W tra (Bw-Bi)/Bi */
template <class G> inline void wall_tally(const orc_ctx &c, const Energy &e, Ray<G> &r, float Tw_, Tally &t, bool fixed, float Bfix) {
    if constexpr (std::is_same<G, float>::value) {
        float Bw;
        if (fixed) Bw = Bfix; else { float K; kb_f(c, e, Tw_, &K, &Bw); }
        const float term = (e.W * (float)r.tra) * ((Bw - e.Bi) * e.invBi);
        t.de -= term;
        double Bw64;
        if (fixed) Bw64 = (double)Bfix; else { double K64; kb_d(c, e, (double)Tw_, &K64, &Bw64); }
        const double term64 = e.W64 * r.tra64 * ((Bw64 - e.Bi64) / e.Bi64);
        t.de64 -= term64;
        t.sumabs += std::fabs(e.W64 * r.tra64) * (std::fabs(Bw64) / e.Bi64 + 1.0);
    } else {
        double Bw;
        if (fixed) Bw = (double)Bfix; else { double K; kb_d(c, e, (double)Tw_, &K, &Bw); }
        const double term = e.W64 * (double)r.tra * ((Bw - e.Bi64) / e.Bi64);
        t.de64 -= term;
        t.sumabs += std::fabs(e.W64 * (double)r.tra) * (std::fabs(Bw) / e.Bi64 + 1.0);
    }
}

/* one grid level of march_ray (mc_gpu_noint.cu:245-658); returns the number
 * of steps taken on this level */
template <class G>
int march_level(const orc_ctx &c, int lev, Ray<G> &r, const Energy &e, float tol_loop,
                float tol_dump, Tally &t, int *status) {
    const orc_params &p = c.p;
    const Level &L = c.lev[lev];
    constexpr bool isf = std::is_same<G, float>::value;
    int counter = 0;
    while (r.tra > (G)tol_loop && counter < L.maxs) {
        /* distances to the three candidate faces (:329-331) */
        const G fx = face_x<G>(L, r.ic - r.f0);
        const G fy = (G)L.yu[r.jc - r.f1];
        const G fz = (G)L.zu[r.kc - r.f2];
        const G dsx = (fx - r.xp) * r.isx;
        const G dsy = (fy - r.yp) * r.isy;
        const G dsz = (fz - r.zp) * r.isz;
        /* temperature at the mapped cell centre (:334-338) */
        const float Tc = L.Ts.empty()
            ? tex3_T(c, tmap(r.ic, L.rx), tmap(r.jc, L.ry), tmap(r.kc, L.rz))
            : L.Ts[(size_t)(r.ic - 1) + (size_t)L.im * ((size_t)(r.jc - 1) + (size_t)L.jm * (r.kc - 1))];
        /* ds = min, position snap to the crossed face (:346-356) */
        G ds = dsx;
        if (ds > dsy) ds = dsy;
        if (ds > dsz) ds = dsz;
        if constexpr (isf) {
            if (ds == dsx) r.xp = fx; else r.xp = (G)fmaf((float)ds, (float)r.sx, (float)r.xp);
            if (ds == dsy) r.yp = fy; else r.yp = (G)fmaf((float)ds, (float)r.sy, (float)r.yp);
            if (ds == dsz) r.zp = fz; else r.zp = (G)fmaf((float)ds, (float)r.sz, (float)r.zp);
        } else {
            if (ds == dsx) r.xp = fx; else r.xp = std::fma(ds, r.sx, r.xp);
            if (ds == dsy) r.yp = fy; else r.yp = std::fma(ds, r.sy, r.yp);
            if (ds == dsz) r.zp = fz; else r.zp = std::fma(ds, r.sz, r.zp);
        }
        /* attenuation and reciprocity tally (:358-370) */
        if constexpr (isf) {
            float K, B;
            kb_f(c, e, Tc, &K, &B);
            float em1, ex;
            expm1_exp_f(-(K * (float)ds), &em1, &ex);
            const float alpha = -em1;
            const float fac = (B - e.Bi) * e.invBi;
            t.de -= ((e.W * (float)r.tra) * alpha) * fac;
            double K64, B64;
            kb_d(c, e, (double)Tc, &K64, &B64);
            const double x64 = K64 * (double)ds;
            const double fac64 = (B64 - e.Bi64) / e.Bi64;
            double em64, ex64;
            expm1_exp_d(-x64, &em64, &ex64);
            const double wta64 = e.W64 * r.tra64 * (-em64);
            const double term64 = wta64 * fac64;
            t.de64 -= term64;
            t.sumabs += std::fabs(wta64) * (std::fabs(B64) / e.Bi64 + 1.0);
            r.tra = (G)((float)r.tra * ex);
            r.tra64 *= ex64;
            if ((float)r.tra < tol_dump) {
                t.de -= (e.W * (float)r.tra) * fac;
                const double d64 = e.W64 * r.tra64 * fac64;
                t.de64 -= d64;
                t.sumabs += std::fabs(e.W64 * r.tra64) * (std::fabs(B64) / e.Bi64 + 1.0);
                r.tra = 0; r.tra64 = 0.0;
                *status = ORC_ABSORBED;
            }
        } else {
            double K, B;
            kb_d(c, e, (double)Tc, &K, &B);
            const double x = K * (double)ds;
            const double fac = (B - e.Bi64) / e.Bi64;
            double em, ex;
            expm1_exp_d(-x, &em, &ex);
            const double wta = e.W64 * (double)r.tra * (-em);
            const double term = wta * fac;
            t.de64 -= term;
            t.sumabs += std::fabs(wta) * (std::fabs(B) / e.Bi64 + 1.0);
            r.tra = (G)((double)r.tra * ex);
            if ((double)r.tra < (double)tol_dump) {
                const double d = e.W64 * (double)r.tra * fac;
                t.de64 -= d;
                t.sumabs += std::fabs(e.W64 * (double)r.tra) * (std::fabs(B) / e.Bi64 + 1.0);
                r.tra = 0;
                *status = ORC_ABSORBED;
            }
        }
        t.hash = fnv(fnv(fnv(t.hash, (uint32_t)r.ic), (uint32_t)r.jc), (uint32_t)r.kc);
        t.steps += 1;
        /* advance one index along the crossed axis, with boundary conditions
         * (:375-654).  Ties: x only if strictly closest; y if dsx>=dsy and
         * dsy<dsz; z otherwise. */
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
            const int bd = p.bd[face];
            if (bd == 1) {                                    /* periodic */
                switch (face) {
                    case FW: r.ic = L.im; r.xp = (G)p.Lx; break;
                    case FE: r.ic = 1; r.xp = 0; break;
                    case FS: r.jc = L.jm; r.yp = (G)p.Ly; break;
                    case FN: r.jc = 1; r.yp = 0; break;
                    case FB: r.kc = L.km; r.zp = (G)p.Lz; break;
                    default: r.kc = 1; r.zp = 0; break;
                }
            } else if (bd == 2) {                             /* black */
                bool fixed = true;
                float Tw_ = 0.f;
                const size_t wq = (size_t)(r.jc - 1) + (size_t)L.jm * (r.kc - 1);
                if (face == FW && p.adiabw == 1) {
                    fixed = false;
                    Tw_ = L.Tws.empty() ? tex3_T(c, xwall_tex(p, 0), tmap(r.jc, L.ry), tmap(r.kc, L.rz)) : L.Tws[wq];
                }
                if (face == FE && p.adiabe == 1) {
                    fixed = false;
                    Tw_ = L.Tws.empty() ? tex3_T(c, xwall_tex(p, 1), tmap(r.jc, L.ry), tmap(r.kc, L.rz))
                                        : L.Tws[(size_t)L.jm * L.km + wq];
                }
                wall_tally<G>(c, e, r, Tw_, t, fixed, fixed ? c.bfix[face][e.rank] : 0.f);
                r.tra = 0; r.tra64 = 0.0;
                *status = ORC_WALL;
            } else if (bd == 3) {                             /* specular */
                switch (face) {
                    case FW: r.xp = 0; r.sx = -r.sx; r.isx = -r.isx; r.ic = 1; r.f0 = r.sx < 0; break;
                    case FE: r.xp = (G)p.Lx; r.sx = -r.sx; r.isx = -r.isx; r.ic = L.im; r.f0 = r.sx < 0; break;
                    case FS: r.yp = 0; r.sy = -r.sy; r.isy = -r.isy; r.jc = 1; r.f1 = r.sy < 0; break;
                    case FN: r.yp = (G)p.Ly; r.sy = -r.sy; r.isy = -r.isy; r.jc = L.jm; r.f1 = r.sy < 0; break;
                    case FB: r.zp = 0; r.sz = -r.sz; r.isz = -r.isz; r.kc = 1; r.f2 = r.sz < 0; break;
                    default: r.zp = (G)p.Lz; r.sz = -r.sz; r.isz = -r.isz; r.kc = L.km; r.f2 = r.sz < 0; break;
                }
            } else if (bd == 4) {                             /* black at adjacent-cell T (z) */
                const float Tw_ = (face == FB)
                    ? tex3_T(c, tmap(r.ic, L.rx), tmap(r.jc, L.ry), 1.5f)
                    : tex3_T(c, tmap(r.ic, L.rx), tmap(r.jc, L.ry), (float)p.kmax + 0.5f);
                wall_tally<G>(c, e, r, Tw_, t, false, 0.f);
                r.tra = 0; r.tra64 = 0.0;
                *status = ORC_WALL;
            } else {                                          /* 5: lost ray */
                r.tra = 0; r.tra64 = 0.0;
                *status = ORC_LOST;
            }
        }
        counter += 1;
    }
    return counter;
}

/* multigrid loop with index coarsening (mc_gpu_noint.cu:210-224).  A ray
 * that has ended stays where it ended (the reference keeps coarsening the
 * indices of a dead ray, which changes nothing else). */
template <class G>
void march_all(const orc_ctx &c, Ray<G> &r, const Energy &e, float tol, Tally &t, int *status, int *level) {
    const int ng = c.p.grid_num;
    *status = ORC_STOPPED;
    *level = 0;
    for (int g = 0; g < ng; ++g) {
        *level = g;
        const int n = march_level<G>(c, g, r, e, tol, tol, t, status);
        if (!(r.tra > (G)tol)) return;
        if (g == ng - 1) {
            if (n >= c.lev[g].maxs) *status = ORC_CAPPED;   /* energy dropped (:252) */
            return;
        }
        const Level &A = c.lev[g], &B = c.lev[g + 1];
        const int ri = A.im / B.im, rj = A.jm / B.jm, rk = A.km / B.km;
        r.ic = (r.ic + ri - 1) / ri;
        r.jc = (r.jc + rj - 1) / rj;
        r.kc = (r.kc + rk - 1) / rk;
    }
}

/* ------------------------------------------------------------------------- */
/* band draws and the srt enumeration (kernel_find :1023-1049, wave_find
 * :1234-1267): histogram of the photmax draws of one (cell, run), enumerated
 * in ascending (rank, g). */
inline void draw_band(const orc_ctx &c, bool surface, float u2, float u3, int *rank, int *g) {
    const int nb = surface ? upper_bound_f(c.cdf_bs.data(), c.nB, u2)
                           : upper_bound_f(c.cdf_b.data(), c.nB, u2);
    const float *gc = surface ? c.cdf_gs.data() : &c.cdf_g[(size_t)nb * c.nQ];
    *g = upper_bound_f(gc, c.nQ, u3);
    *rank = c.rank_of[nb];
}

void band_list(const orc_ctx &c, bool surface, uint32_t cell, uint32_t tag, int v,
               std::vector<int32_t> &nb, std::vector<int32_t> &gq) {
    const int P = c.p.photmax;
    nb.resize(P); gq.resize(P);
    if (c.p.srt != 1) {
        for (int h = 0; h < P; ++h) {
            float u[4];
            draw4(c.p.seed, (uint32_t)h, (uint32_t)v, cell, tag + c.tagoff, u);
            draw_band(c, surface, u[2], u[3], &nb[h], &gq[h]);
        }
        return;
    }
    std::vector<int32_t> cnt((size_t)c.nB * c.nQ, 0);
    for (int h = 0; h < P; ++h) {
        float u[4];
        int r, g;
        draw4(c.p.seed, (uint32_t)h, (uint32_t)v, cell, tag + c.tagoff, u);
        draw_band(c, surface, u[2], u[3], &r, &g);
        cnt[(size_t)r * c.nQ + g] += 1;
    }
    int h = 0;
    for (int r = 0; r < c.nB; ++r)
        for (int g = 0; g < c.nQ; ++g)
            for (int n = 0; n < cnt[(size_t)r * c.nQ + g]; ++n) { nb[h] = r; gq[h] = g; ++h; }
}

inline uint32_t fluid_cell_id(const orc_params &p, int i, int j, int k) {
    return (uint32_t)((i - 1) + p.imax * ((j - 1) + p.jmax * (k - 1)));
}
inline uint32_t surf_cell_id(const orc_params &p, int j, int k) {
    return (uint32_t)((j - 1) + p.jmax * (k - 1));
}

template <class G> inline void set_dir(Ray<G> &r, G sx, G sy, G sz) {
    r.sx = sx; r.sy = sy; r.sz = sz;
    r.f0 = sx < 0; r.f1 = sy < 0; r.f2 = sz < 0;
    r.isx = (G)1 / sx; r.isy = (G)1 / sy; r.isz = (G)1 / sz;
}

void energy_setup(const orc_ctx &c, int rank, int g, float Ti, bool surface, int bound, Energy &e) {
    e.rank = rank; e.g = g;
    e.kt = &c.ktab[((size_t)rank * c.nQ + g) * (c.nT + 1)];
    e.bt = &c.btab[(size_t)rank * (c.nT + 1)];
    float Ki, Bi, Km, Bm;
    kb_f(c, e, Ti, &Ki, &Bi);
    kb_f(c, e, c.Tmax, &Km, &Bm);
    double Ki64, Bi64, Km64, Bm64;
    kb_d(c, e, (double)Ti, &Ki64, &Bi64);
    kb_d(c, e, (double)c.Tmax, &Km64, &Bm64);
    e.Bi = Bi; e.invBi = 1.0f / Bi; e.Bi64 = Bi64;
    if (!surface) {
        /* ratioI = B(Ti)/B(Tmax) * K(Ti)/K(Tmax)  (:181-186) */
        const float ratioI = ((Bi / Bm) * Ki) / Km;
        e.W = c.Ibmax * ratioI;
        e.W64 = c.Ibmax64 * (Bi64 / Bm64) * (Ki64 / Km64);
    } else {
        /* surface: ratioI = B(Tw)/B(Tmax) (:1419-1424) */
        e.W = c.Ibmax_s[bound] * (Bi / Bm);
        e.W64 = c.Ibmax_s64[bound] * (Bi64 / Bm64);
    }
}

/* fluid ray (kernel_fluid per-ray body, :115-224) */
template <class G>
void fluid_ray(const orc_ctx &c, int i, int j, int k, int v, int h, int rank, int g,
               orc_ray *out, Tally &t) {
    const orc_params &p = c.p;
    const Level &L0 = c.lev[0];
    float u[4];
    draw4(p.seed, (uint32_t)h, (uint32_t)v, fluid_cell_id(p, i, j, k), TAG_FLUID + c.tagoff, u);
    if (p.sampler == 1)
        sobol_dir_uniforms(p.seed, (uint32_t)h, (uint32_t)v, fluid_cell_id(p, i, j, k), TAG_FLUID + c.tagoff, &u[0],
                           &u[1]);
    Ray<G> r;
    r.xp = (G)L0.x[i]; r.yp = (G)L0.y[j]; r.zp = (G)L0.z[k];
    r.ic = i; r.jc = j; r.kc = k;
    G sx, sy, sz;
    if constexpr (std::is_same<G, float>::value) {
        float s[3];
        orc_iso_dir(u[0], u[1], s);
        sx = (G)s[0]; sy = (G)s[1]; sz = (G)s[2];
    } else {
        const double ct = 1.0 - 2.0 * (double)u[1];
        const double st = 2.0 * std::sqrt((double)u[1] * (1.0 - (double)u[1]));
        double sn, cs;
        sincos2pi_d(u[0], &sn, &cs);
        sx = (G)ct; sy = (G)(st * cs); sz = (G)(st * sn);
    }
    /* zero components replaced (:132-137) */
    if (sx == 0) sx = (G)1e-10f;
    if (sy == 0) sy = (G)1e-10f;
    if (sz == 0) sz = (G)1e-10f;
    set_dir<G>(r, sx, sy, sz);
    r.tra = 1; r.tra64 = 1.0;
    Energy e;
    energy_setup(c, rank, g, Tat(c, i, j, k), false, 0, e);
    int status = ORC_STOPPED, level = 0;
    march_all<G>(c, r, e, p.toll, t, &status, &level);
    if (out) {
        out->nb = rank; out->g = g;
        out->ic = r.ic; out->jc = r.jc; out->kc = r.kc; out->level = level;
        out->steps = t.steps; out->status = status; out->path_hash = t.hash;
        out->sx = (float)sx; out->sy = (float)sy; out->sz = (float)sz;
        out->tra = (float)r.tra;
    }
}

/* surface ray (kernel_surface per-ray body, :1381-1462) */
template <class G>
void surface_ray(const orc_ctx &c, int bound, int j, int k, int v, int h, int rank, int g,
                 orc_ray *out, Tally &t) {
    const orc_params &p = c.p;
    const Level &L0 = c.lev[0];
    float u[4];
    draw4(p.seed, (uint32_t)h, (uint32_t)v, surf_cell_id(p, j, k), (bound ? TAG_EAST : TAG_WEST) + c.tagoff, u);
    Ray<G> r;
    const int i = bound ? p.imax : 1;
    r.xp = bound ? (G)p.Lx : (G)0;
    r.yp = (G)L0.y[j]; r.zp = (G)L0.z[k];
    r.ic = i; r.jc = j; r.kc = k;
    G sx, sy, sz;
    if constexpr (std::is_same<G, float>::value) {
        float s[3];
        orc_lamb_dir(u[0], u[1], bound, s);
        sx = (G)s[0]; sy = (G)s[1]; sz = (G)s[2];
    } else {
        const double a = std::sqrt(1.0 - (double)u[1]), b = std::sqrt((double)u[1]);
        double sn, cs;
        sincos2pi_d(u[0], &sn, &cs);
        const double st = bound ? a : b;
        sx = bound ? (G)(-b) : (G)a;
        sy = (G)(st * cs); sz = (G)(st * sn);
    }
    /* near-zero components (:1395-1400) */
    if (std::fabs((double)sx) <= 1e-10) sx = (G)((float)(1 - 2 * bound) * 1e-10f);
    if (std::fabs((double)sy) <= 1e-10) sy = (G)((float)((sy > 0) - (sy <= 0)) * 1e-10f);
    if (std::fabs((double)sz) <= 1e-10) sz = (G)((float)((sz > 0) - (sz <= 0)) * 1e-10f);
    set_dir<G>(r, sx, sy, sz);
    r.tra = 1; r.tra64 = 1.0;
    const float Tw_ = tex3_T(c, xwall_tex(p, bound), (float)j + 0.5f, (float)k + 0.5f);
    Energy e;
    energy_setup(c, rank, g, Tw_, true, bound, e);
    /* on-face index adjustment (:1429-1441) */
    if (r.f0 && r.xp == (G)L0.xu[r.ic - 1]) r.ic -= 1;
    if (r.f1 && r.yp == (G)L0.yu[r.jc - 1]) r.jc -= 1;
    if (r.f2 && r.zp == (G)L0.zu[r.kc - 1]) r.kc -= 1;
    if (!r.f0 && r.xp == (G)L0.xu[r.ic]) r.ic += 1;
    if (!r.f1 && r.yp == (G)L0.yu[r.jc]) r.jc += 1;
    if (!r.f2 && r.zp == (G)L0.zu[r.kc]) r.kc += 1;
    int status = ORC_STOPPED, level = 0;
    march_all<G>(c, r, e, p.tollS, t, &status, &level);
    if (out) {
        out->nb = rank; out->g = g;
        out->ic = r.ic; out->jc = r.jc; out->kc = r.kc; out->level = level;
        out->steps = t.steps; out->status = status; out->path_hash = t.hash;
        out->sx = (float)sx; out->sy = (float)sy; out->sz = (float)sz;
        out->tra = (float)r.tra;
    }
}

inline Tally tally0() { Tally t; t.de = 0.f; t.de64 = 0.0; t.sumabs = 0.0; t.steps = 0; t.hash = 2166136261u; return t; }

void set_err(char *err, int errlen, const std::string &m) {
    if (err && errlen > 0) { std::snprintf(err, (size_t)errlen, "%s", m.c_str()); }
}

}  // namespace

/* ======================== exported building blocks ======================= */
extern "C" {

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    philox10(c, key[0], key[1]);
    for (int q = 0; q < 4; ++q) out[q] = c[q];
}
float orc_u01(uint32_t x) { return u01f(x); }
void orc_sincos2pi(float u, float *s, float *c) { sincos2pi_f(u, s, c); }
void orc_expm1_exp(float x, float *em1, float *e) { expm1_exp_f(x, em1, e); }
void orc_sincos2pi_d(float u, double *s, double *c) { sincos2pi_d(u, s, c); }
void orc_expm1_exp_d(double x, double *em1, double *e) { expm1_exp_d(x, em1, e); }
float orc_planck(float T, float nu) { return planck_f(T, nu); }
int orc_upper_bound(const float *cdf, int n, float R) { return upper_bound_f(cdf, n, R); }

/* sort_idx (mc_gpu_noint.cu:1278-1294): ascending kavg; the reference's
 * std::sort is unstable, ties are ordered here by band index. */
void orc_sort_idx(const float *kavg, int nB, int32_t *idx) {
    std::vector<int32_t> v(nB);
    for (int i = 0; i < nB; ++i) v[i] = i;
    std::stable_sort(v.begin(), v.end(), [&](int32_t a, int32_t b) { return kavg[a] < kavg[b]; });
    for (int i = 0; i < nB; ++i) idx[i] = v[i];
}

/* isotropic direction (emiss_ang, mc_gpu_noint.cu:1077-1084):
 * phi = 2 pi u0, cos(theta) = 1 - 2 u1, s = (cos t, sin t cos p, sin t sin p);
 * sin(theta) = 2 sqrt(u1 (1-u1)) exactly from u1. */
/* point h of the two-dimensional Sobol sequence (32-bit fractions) */
void orc_sobol2(uint32_t h, uint32_t x[2]) { sobol2_point(h, &x[0], &x[1]); }

/* the direction uniforms (phi, theta) of fluid ray h of run v of cell id
 * `cell` under the Sobol sampler */
void orc_sobol_uniforms(uint64_t seed, uint32_t h, uint32_t v, uint32_t cell, uint32_t tag, float u[2]) {
    sobol_dir_uniforms(seed, h, v, cell, tag, &u[0], &u[1]);
}

void orc_iso_dir(float u0, float u1, float s[3]) {
    const float ct = 1.0f - 2.0f * u1;
    const float st = 2.0f * std::sqrt(u1 * (1.0f - u1));
    float sp, cp;
    sincos2pi_f(u0, &sp, &cp);
    s[0] = ct; s[1] = st * cp; s[2] = st * sp;
}

/* Lambertian wall direction (emiss_angS, mc_gpu_noint.cu:1093-1102):
 * theta = acos(1-2u1)/2 (+pi/2 on the east wall), i.e. sin^2 theta = u1. */
void orc_lamb_dir(float u0, float u1, int bound, float s[3]) {
    const float a = std::sqrt(1.0f - u1);   /* cos(theta0) */
    const float b = std::sqrt(u1);          /* sin(theta0) */
    float sp, cp;
    sincos2pi_f(u0, &sp, &cp);
    if (bound == 0) { s[0] = a;  s[1] = b * cp; s[2] = b * sp; }
    else            { s[0] = -b; s[1] = a * cp; s[2] = a * sp; }
}

int orc_grid(const orc_params *p, int level, float *x, float *xu, float *y, float *yu, float *z, float *zu) {
    if (level < 0 || level >= p->grid_num) return -1;
    Level L;
    build_level(*p, level, L);
    std::memcpy(x, L.x.data(), L.x.size() * 4);
    std::memcpy(xu, L.xu.data(), L.xu.size() * 4);
    std::memcpy(y, L.y.data(), L.y.size() * 4);
    std::memcpy(yu, L.yu.data(), L.yu.size() * 4);
    std::memcpy(z, L.z.data(), L.z.size() * 4);
    std::memcpy(zu, L.zu.data(), L.zu.size() * 4);
    return 0;
}

/* Tmax over interior cells and the x-wall midpoints (mc_gpu_noint.cu:696-710) */
float orc_tmax(const orc_params *p, const float *T) {
    float Tmax = 0.f;
    const size_t sx = 1, sy = (size_t)(p->imax + 2), sz = sy * (size_t)(p->jmax + 2);
    for (int k = 1; k < p->kmax + 1; ++k)
        for (int j = 1; j < p->jmax + 1; ++j) {
            const size_t b = (size_t)j * sy + (size_t)k * sz;
            for (int i = 1; i < p->imax + 1; ++i) Tmax = std::max(Tmax, T[b + (size_t)i * sx]);
            Tmax = std::max(Tmax, (T[b] + T[b + 1]) * 0.5f);
            Tmax = std::max(Tmax, (T[b + (size_t)p->imax] + T[b + (size_t)p->imax + 1]) * 0.5f);
        }
    return Tmax;
}

/* readT, narrow-band form (read_tables.cpp:21-95): planck-mean.txt,
 * NarrBand%d.txt, prob_new2.txt; k scaled by chanHeight (:30,56); kavg is
 * the value on the LAST temperature row (:48-52). */
int orc_read_tables(const char *dir, double chanHeight, orc_tables *t) {
    const int nB = t->nB, nT = t->nT, nQ = t->nQ;
    std::string base(dir);
    char path[4096];
    std::snprintf(path, sizeof path, "%s/planck-mean.txt", base.c_str());
    FILE *fp = std::fopen(path, "r");
    if (!fp) return -1;
    for (int i = 0; i < nT; ++i) {
        float d1, d2;
        if (std::fscanf(fp, "%f %e %e %e\n", &t->Tnb[i], &t->kP[i], &d1, &d2) != 4) { std::fclose(fp); return -2; }
        t->kP[i] *= (float)chanHeight;
    }
    std::fclose(fp);
    for (int b = 0; b < nB; ++b) {
        std::snprintf(path, sizeof path, "%s/NarrBand%d.txt", base.c_str(), b);
        fp = std::fopen(path, "r");
        if (!fp) return -3;
        if (std::fscanf(fp, "%e\t%e\t%e\n", &t->wvl[b], &t->wvr[b], &t->wvc[b]) != 3) { std::fclose(fp); return -4; }
        for (int j = 0; j < nT; ++j) {
            float d0, d1;
            if (std::fscanf(fp, " %f\t", &d0) != 1 || std::fscanf(fp, " %f\t", &d1) != 1 ||
                std::fscanf(fp, " %f\t", &t->kavg[b]) != 1) { std::fclose(fp); return -4; }
            for (int g = 0; g < nQ; ++g) {
                float &kq = t->kq[((size_t)b * nT + j) * nQ + g];
                if (std::fscanf(fp, " %e\t", &kq) != 1) { std::fclose(fp); return -4; }
                kq *= (float)chanHeight;
            }
            if (std::fscanf(fp, "\n") < 0) { /* end of file is fine */ }
        }
        std::fclose(fp);
    }
    std::snprintf(path, sizeof path, "%s/prob_new2.txt", base.c_str());
    fp = std::fopen(path, "r");
    if (!fp) return -5;
    for (int i = 0; i < nT; ++i)
        for (int j = 0; j < nB; ++j) {
            float d0; int d3;
            if (std::fscanf(fp, "%f \t %d \t %e \t", &d0, &d3, &t->prob[(size_t)i * nB + j]) != 3) { std::fclose(fp); return -6; }
            for (int g = 0; g < nQ; ++g)
                if (std::fscanf(fp, "%e\t", &t->probg[((size_t)i * nB + j) * nQ + g]) != 1) { std::fclose(fp); return -6; }
            if (std::fscanf(fp, "\n") < 0) { }
        }
    std::fclose(fp);
    return 0;
}

/* readT, grey form (read_tables.cpp:98-175): nT temperatures from 0.9 Tmins
 * to 1.1 Tmax, nB bands over [0, wvmax], k = abscoeff (unscaled), band CDF
 * from Ib * dnu; probg is set to 1 (the reference's write there is out of
 * bounds, :169-170).  grey_true: one band with Ib = sigma T^4 / pi. */
int orc_grey_tables(const orc_params *p, float Tmax, orc_tables *t) {
    const int nB = t->nB, nT = t->nT, nQ = t->nQ;
    if (nQ != 1) return -1;
    for (int i = 0; i < nT; ++i) {
        t->kP[i] = p->abscoeff;
        t->Tnb[i] = (float)((double)p->Tmins * 0.9 +
                            ((double)Tmax * 1.1 - (double)p->Tmins * 0.9) * (double)(float)i / (double)(float)nT);
    }
    if (p->grey_true) {
        if (nB != 1) return -2;
        t->wvl[0] = 0.f; t->wvr[0] = p->wvmax; t->wvc[0] = p->wvmax / 2.f; t->kavg[0] = p->abscoeff;
        for (int j = 0; j < nT; ++j) {
            t->kq[j] = p->abscoeff;
            t->prob[j] = 1.0f;
            t->probg[j] = 1.0f;
        }
        return 0;
    }
    const float dw = p->wvmax / (float)nB;
    std::vector<float> wvl(nB), wvr(nB);
    wvl[0] = 0.f;
    for (int i = 1; i < nB; ++i) wvl[i] = wvl[i - 1] + dw;
    wvr[0] = dw;
    for (int i = 1; i < nB; ++i) wvr[i] = wvr[i - 1] + dw;
    for (int i = 0; i < nB; ++i) {
        t->wvc[i] = (float)((double)(wvl[i] + wvr[i]) / 2.);
        t->wvl[i] = wvl[i]; t->wvr[i] = wvr[i];
        t->kavg[i] = p->abscoeff;
        for (int j = 0; j < nT; ++j) t->kq[(size_t)i * nT + j] = p->abscoeff;
    }
    for (int i = 0; i < nT; ++i) {
        float *row = &t->prob[(size_t)i * nB];
        const float Tr = t->Tnb[(g_plant == PLANT_GREY_ROW) ? std::min(i + 1, nT - 1) : i];
        const float *nu = (g_plant == PLANT_GREY_WVL) ? wvl.data() : t->wvc;
        row[0] = planck_f(Tr, nu[0]) * (wvr[0] - wvl[0]) * 100.f;
        for (int j = 1; j < nB; ++j) row[j] = row[j - 1] + planck_f(Tr, nu[j]) * (wvr[j] - wvl[j]) * 100.f;
        const float last = row[nB - 1];
        for (int j = 0; j < nB; ++j) { row[j] = row[j] / last; t->probg[(size_t)i * nB + j] = 1.0f; }
    }
    return 0;
}

/* mc_gpu_ setup (mc_gpu_noint.cu:660-830) without the device parts */
orc_ctx *orc_create(const orc_params *pp, const orc_tables *t, const float *T, int zstart, char *err, int errlen) {
    orc_ctx *c = new orc_ctx();
    c->p = *pp;
    c->tagoff = 0;
    orc_params &p = c->p;
    if (p.grey) p.srt = 0;                                  /* NarrowBand.h:220-221 */
    c->nB = t->nB; c->nT = t->nT; c->nQ = t->nQ;
    c->grey_true = p.grey && p.grey_true;
    if (p.grid_num < 1 || p.grid_num > ORC_MAXLEV || p.p_col < 1 || p.kmax % p.p_col) {
        set_err(err, errlen, "bad grid_num/p_col"); delete c; return nullptr;
    }
    c->zstart = zstart; c->kloc = p.kmax / p.p_col;
    for (int f = 0; f < 4; ++f)
        if (p.bd[f] < 1 || p.bd[f] > 3) { set_err(err, errlen, "x/y faces take bd 1..3"); delete c; return nullptr; }
    for (int f = 4; f < 6; ++f)
        if (p.bd[f] < 1 || p.bd[f] > 5) { set_err(err, errlen, "z faces take bd 1..5"); delete c; return nullptr; }
    const size_t nT3 = (size_t)(p.imax + 2) * (p.jmax + 2) * (p.kmax + 2);
    c->T.assign(T, T + nT3);
    c->lev.resize(p.grid_num);
    for (int g = 0; g < p.grid_num; ++g) build_level(p, g, c->lev[g]);

    const int nB = c->nB, nT = c->nT, nQ = c->nQ;
    c->Tnb.assign(t->Tnb, t->Tnb + nT);
    c->Tmax = orc_tmax(&p, T);
    const float dT = c->Tnb[1] - c->Tnb[0];
    c->Tnb0 = c->Tnb[0]; c->invdT = 1.0f / dT;
    c->Tnb0d = (double)c->Tnb[0]; c->invdTd = 1.0 / (double)dT;
    c->tm = (int)((c->Tmax - c->Tnb[0]) / dT);
    if (c->tm < 0 || c->tm + 1 >= nT) { set_err(err, errlen, "Tmax outside the temperature table"); delete c; return nullptr; }
    const int tm = c->tm;
    /* kappamax: lerp of kP at Tmax (read_tables.cpp:35-38; grey :108) */
    if (p.grey) { c->kappamax = p.abscoeff; c->kappamax64 = p.abscoeff; }
    else {
        c->kappamax = (t->kP[tm + 1] - t->kP[tm]) / (c->Tnb[tm + 1] - c->Tnb[tm]) * (c->Tmax - c->Tnb[tm]) + t->kP[tm];
        c->kappamax64 = ((double)t->kP[tm + 1] - t->kP[tm]) / ((double)c->Tnb[tm + 1] - c->Tnb[tm]) *
                        ((double)c->Tmax - c->Tnb[tm]) + (double)t->kP[tm];
    }
    const double T4 = (double)c->Tmax * c->Tmax * c->Tmax * c->Tmax;
    c->Ibmax64 = 4.0 * c->kappamax64 * T4 * kStefan / p.photmax;          /* :74 */
    c->Ibmax = (float)(4.0 * (double)c->kappamax * T4 * kStefan / p.photmax);
    c->Ibmax_s64[0] = (double)p.eps[FW] * kStefan * T4 / p.photmax;      /* :1362,1366 */
    c->Ibmax_s64[1] = (double)p.eps[FE] * kStefan * T4 / p.photmax;
    c->Ibmax_s[0] = (float)c->Ibmax_s64[0];
    c->Ibmax_s[1] = (float)c->Ibmax_s64[1];
    /* nearest tabulated row to Tmax for the band pdf (:1110-1131) */
    c->tstar = ((c->Tmax - c->Tnb[tm]) < (c->Tnb[tm + 1] - c->Tmax)) ? tm : tm + 1;
    const int ts = c->tstar;

    c->idx.resize(nB); c->rank_of.resize(nB);
    if (p.srt == 1) orc_sort_idx(t->kavg, nB, c->idx.data());
    else for (int b = 0; b < nB; ++b) c->idx[b] = b;
    for (int r = 0; r < nB; ++r) c->rank_of[c->idx[r]] = r;

    /* k and Ib tables in rank order, edge-padded (narrowband_copy
     * memory_copy_noint.cu:193-208; Ib mc_gpu_noint.cu:770-780) */
    c->ktab.assign((size_t)nB * nQ * (nT + 1), 0.f);
    c->btab.assign((size_t)nB * (nT + 1), 0.f);
    for (int r = 0; r < nB; ++r) {
        const int b = c->idx[r];
        for (int g = 0; g < nQ; ++g) {
            float *kt = &c->ktab[((size_t)r * nQ + g) * (nT + 1)];
            for (int tt = 0; tt < nT; ++tt) kt[tt] = t->kq[((size_t)b * nT + tt) * nQ + g];
            kt[nT] = kt[nT - 1];
        }
        float *bt = &c->btab[(size_t)r * (nT + 1)];
        for (int tt = 0; tt < nT; ++tt)
            bt[tt] = c->grey_true ? (float)(kStefan * std::pow((double)c->Tnb[tt], 4) / kPi)
                                  : planck_f(c->Tnb[tt], t->wvc[b]);
        bt[nT] = bt[nT - 1];
    }
    c->cdf_b.assign(t->prob + (size_t)ts * nB, t->prob + (size_t)(ts + 1) * nB);
    c->cdf_g.assign(t->probg + (size_t)ts * nB * nQ, t->probg + (size_t)(ts + 1) * nB * nQ);
    /* surface CDFs: readS (read_tables.cpp:188-220) at row t*, cumulative wq */
    c->cdf_bs.assign(nB, 0.f);
    {
        const float Tr = c->Tnb[(g_plant == PLANT_WALL_ROW) ? std::min(ts + 1, nT - 1) : ts];
        float acc = 0.f;
        for (int b = 0; b < nB; ++b) {
            const float ib = c->grey_true ? 1.0f : planck_f(Tr, t->wvc[b]);
            acc = (b == 0) ? ib * (t->wvr[0] - t->wvl[0]) * 100.f : acc + ib * (t->wvr[b] - t->wvl[b]) * 100.f;
            c->cdf_bs[b] = acc;
        }
        for (int b = 0; b < nB; ++b) c->cdf_bs[b] = c->cdf_bs[b] / acc;
    }
    c->wq.assign(t->wq, t->wq + nQ);
    c->wvl.assign(t->wvl, t->wvl + nB);
    c->wvr.assign(t->wvr, t->wvr + nB);
    c->cdf_gs.assign(nQ, 0.f);
    {
        float acc = 0.f;
        for (int g = 0; g < nQ; ++g) { acc = (g == 0) ? t->wq[0] : acc + t->wq[g]; c->cdf_gs[g] = acc; }
        for (int g = 0; g < nQ; ++g) c->cdf_gs[g] = c->cdf_gs[g] / acc;
        if (g_plant == PLANT_WALL_G_UNIFORM)
            for (int g = 0; g < nQ; ++g) c->cdf_gs[g] = (float)(g + 1) / (float)nQ;
    }
    /* fixed-temperature wall emission per rank (black_copy, memory_copy_noint.cu:27-52) */
    for (int f = 0; f < 6; ++f) {
        c->bfix[f].assign(nB, 0.f);
        for (int r = 0; r < nB; ++r) {
            const double Tw = p.Tw[f];
            c->bfix[f][r] = c->grey_true ? (float)(kStefan * Tw * Tw * Tw * Tw / kPi)
                                         : planck_f(p.Tw[f], t->wvc[c->idx[r]]);
        }
    }
    return c;
}

void orc_destroy(orc_ctx *c) { delete c; }

/* streams of snapshot s of a batch draw with stream tag + 16 s */
void orc_set_snapshot(orc_ctx *c, int s) { c->tagoff = 16u * (uint32_t)s; }

void orc_plant(int id) { g_plant = id; }

float orc_tex3(const orc_ctx *c, float tx, float ty, float tz) { return tex3_T(*c, tx, ty, tz); }

/* the temperatures the march reads on level g (texture mean; the spline
 * restriction's values when it is on): cells Tc [km][jm][im], adiabatic x
 * walls Tw [2][km][jm], bd=4 z walls Tz [2][jm][im] (either may be NULL) */
int orc_level_T(const orc_ctx *c, int g, float *Tc, float *Tw, float *Tz) {
    if (!c || g < 0 || g >= c->p.grid_num) return -1;
    const orc_params &p = c->p;
    const Level &L = c->lev[g];
    for (int k = 1; k <= L.km; ++k)
        for (int j = 1; j <= L.jm; ++j)
            for (int i = 1; i <= L.im; ++i) {
                const size_t n = (size_t)(i - 1) + (size_t)L.im * ((size_t)(j - 1) + (size_t)L.jm * (k - 1));
                if (Tc) Tc[n] = L.Ts.empty() ? tex3_T(*c, tmap(i, L.rx), tmap(j, L.ry), tmap(k, L.rz)) : L.Ts[n];
            }
    for (int b = 0; b < 2 && Tw; ++b)
        for (int k = 1; k <= L.km; ++k)
            for (int j = 1; j <= L.jm; ++j) {
                const size_t n = (size_t)(j - 1) + (size_t)L.jm * (k - 1);
                Tw[(size_t)b * L.jm * L.km + n] = L.Tws.empty()
                    ? tex3_T(*c, xwall_tex(p, b), tmap(j, L.ry), tmap(k, L.rz)) : L.Tws[(size_t)b * L.jm * L.km + n];
            }
    for (int b = 0; b < 2 && Tz; ++b)
        for (int j = 1; j <= L.jm; ++j)
            for (int i = 1; i <= L.im; ++i)
                Tz[(size_t)b * L.im * L.jm + (size_t)(i - 1) + (size_t)L.im * (j - 1)] =
                    tex3_T(*c, tmap(i, L.rx), tmap(j, L.ry), b ? (float)p.kmax + 0.5f : 1.5f);
    return 0;
}

void orc_get_info(const orc_ctx *c, orc_info *info) {
    info->Tmax = c->Tmax; info->kappamax = c->kappamax; info->Ibmax = c->Ibmax;
    info->Ibmax_w = c->Ibmax_s[0]; info->Ibmax_e = c->Ibmax_s[1];
    info->tm = c->tm; info->tstar = c->tstar;
    info->n_emit = c->p.imax * c->p.jmax * c->kloc;
    info->kappamax64 = c->kappamax64; info->Ibmax64 = c->Ibmax64;
}

void orc_band_sequence(const orc_ctx *c, int surface, int bound, int i, int j, int k, int v, int32_t *nb, int32_t *g) {
    std::vector<int32_t> a, b;
    if (surface) band_list(*c, true, surf_cell_id(c->p, j, k), bound ? TAG_EAST : TAG_WEST, v, a, b);
    else band_list(*c, false, fluid_cell_id(c->p, i, j, k), TAG_FLUID, v, a, b);
    std::memcpy(nb, a.data(), a.size() * 4);
    std::memcpy(g, b.data(), b.size() * 4);
}

void orc_trace_fluid(const orc_ctx *c, int i, int j, int k, int v, int h, orc_ray *r) {
    std::vector<int32_t> nb, g;
    band_list(*c, false, fluid_cell_id(c->p, i, j, k), TAG_FLUID, v, nb, g);
    Tally t = tally0();
    fluid_ray<float>(*c, i, j, k, v, h, nb[h], g[h], r, t);
    r->de = t.de; r->de64 = t.de64; r->sumabs = t.sumabs;
}

void orc_trace_surface(const orc_ctx *c, int bound, int j, int k, int v, int h, orc_ray *r) {
    std::vector<int32_t> nb, g;
    band_list(*c, true, surf_cell_id(c->p, j, k), bound ? TAG_EAST : TAG_WEST, v, nb, g);
    Tally t = tally0();
    surface_ray<float>(*c, bound, j, k, v, h, nb[h], g[h], r, t);
    r->de = t.de; r->de64 = t.de64; r->sumabs = t.sumabs;
}

/* kernel_fluid cell loop (:92-241) over a cell list; per (cell, run) tallies */
int orc_solve_fluid(const orc_ctx *c, int precision, const int32_t *cells, int ncells,
                    float *de32, double *de32d, double *de64, int64_t *steps, int nthreads) {
    const int nVar = c->p.nVar, P = c->p.photmax;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
    for (int n = 0; n < ncells; ++n) {
        const int i = cells[3 * n], j = cells[3 * n + 1], k = cells[3 * n + 2];
        std::vector<int32_t> nb, g;
        int64_t st = 0;
        for (int v = 0; v < nVar; ++v) {
            band_list(*c, false, fluid_cell_id(c->p, i, j, k), TAG_FLUID, v, nb, g);
            float s32 = 0.f;
            double s32d = 0.0, s64 = 0.0;
            for (int h = 0; h < P; ++h) {
                Tally t = tally0();
                if (precision == 0) fluid_ray<float>(*c, i, j, k, v, h, nb[h], g[h], nullptr, t);
                else fluid_ray<double>(*c, i, j, k, v, h, nb[h], g[h], nullptr, t);
                s32 += (precision == 0) ? t.de : (float)t.de64;
                s32d += (precision == 0) ? (double)t.de : t.de64;
                s64 += t.de64;
                st += t.steps;
            }
            de32[(size_t)n * nVar + v] = s32;
            if (de32d) de32d[(size_t)n * nVar + v] = s32d;
            de64[(size_t)n * nVar + v] = s64;
        }
        if (steps) steps[n] = st;
    }
    return 0;
}

/* orc_solve_fluid split by band group (test infrastructure: the precision
 * study of the mixed-precision solver against full FP64) */
int orc_solve_fluid_groups(const orc_ctx *c, int precision, const int32_t *cells, int ncells,
                           const int32_t *group_of_rank, int ngroups, double *out, int nthreads) {
    const int nVar = c->p.nVar, P = c->p.photmax;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
    for (int n = 0; n < ncells; ++n) {
        const int i = cells[3 * n], j = cells[3 * n + 1], k = cells[3 * n + 2];
        std::vector<int32_t> nb, g;
        for (int v = 0; v < nVar; ++v) {
            band_list(*c, false, fluid_cell_id(c->p, i, j, k), TAG_FLUID, v, nb, g);
            double *o = out + ((size_t)n * nVar + v) * ngroups;
            for (int q = 0; q < ngroups; ++q) o[q] = 0.0;
            for (int h = 0; h < P; ++h) {
                Tally t = tally0();
                if (precision == 0) fluid_ray<float>(*c, i, j, k, v, h, nb[h], g[h], nullptr, t);
                else fluid_ray<double>(*c, i, j, k, v, h, nb[h], g[h], nullptr, t);
                o[group_of_rank[nb[h]]] += (precision == 0) ? (double)t.de : t.de64;
            }
        }
    }
    return 0;
}

int orc_solve_surface(const orc_ctx *c, int precision, const int32_t *cells, int ncells,
                      float *de32, double *de32d, double *de64, int64_t *steps, int nthreads) {
    const int nVar = c->p.nVar, P = c->p.photmax;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
    for (int n = 0; n < ncells; ++n) {
        const int bound = cells[3 * n], j = cells[3 * n + 1], k = cells[3 * n + 2];
        std::vector<int32_t> nb, g;
        int64_t st = 0;
        for (int v = 0; v < nVar; ++v) {
            band_list(*c, true, surf_cell_id(c->p, j, k), bound ? TAG_EAST : TAG_WEST, v, nb, g);
            float s32 = 0.f;
            double s32d = 0.0, s64 = 0.0;
            for (int h = 0; h < P; ++h) {
                Tally t = tally0();
                if (precision == 0) surface_ray<float>(*c, bound, j, k, v, h, nb[h], g[h], nullptr, t);
                else surface_ray<double>(*c, bound, j, k, v, h, nb[h], g[h], nullptr, t);
                s32 += (precision == 0) ? t.de : (float)t.de64;
                s32d += (precision == 0) ? (double)t.de : t.de64;
                s64 += t.de64;
                st += t.steps;
            }
            de32[(size_t)n * nVar + v] = s32;
            if (de32d) de32d[(size_t)n * nVar + v] = s32d;
            de64[(size_t)n * nVar + v] = s64;
        }
        if (steps) steps[n] = st;
    }
    return 0;
}

/* correlatedKT (nbcorrk/correlatedK.cpp:144-225) of one spectrum, points
 * b .. e-1: the k^pwr histogram ff, the shifted interval values kk (both
 * n_pwrk + 1 long, left for the caller), the quadrature values kq[nq] and
 * kavg as calcProbT sums it (calc_prob.cpp:52-56).
 * Deviations: bins are clamped to [0, n_pwrk] and a band of constant k
 * (zero bin width) puts every point in bin 1. */
static void ck_one(int64_t b, int64_t e, const double *wvn, const double *kvn, double wl, double wr, int n_pwrk,
                   double pwr, int nq, const double *gq, std::vector<double> &kk, std::vector<double> &ff,
                   double *kq, double *kavg, double *kmin, double *kmax) {
    double lo = 1e10, hi = -1000.0;
    for (int64_t h = b; h < e; ++h) {
        lo = std::min(lo, kvn[h]);
        hi = std::max(hi, kvn[h]);
    }
    *kmin = lo;
    *kmax = hi;
    const double pmin = std::pow(lo, pwr), pmax = std::pow(hi, pwr);
    const double step = (pmax - pmin) / (n_pwrk - 1);
    ff.assign(n_pwrk + 1, 0.0);
    kk.resize(n_pwrk + 1);
    std::vector<double> gg(n_pwrk + 1, 0.0);
    for (int h = 0; h < n_pwrk + 1; ++h) kk[h] = std::pow(h * step + pmin, 1.0 / pwr);
    const double width = wr - wl;
    for (int64_t h = b + 1; h < e; ++h) {
        const double ki = (kvn[h] + kvn[h - 1]) / 2;
        int bin = 1;
        if (step > 0.0) {
            const double x = (std::pow(ki, pwr) - pmin) / step + 1;
            bin = std::max(0, (int)x);
            bin = std::min(bin, n_pwrk);
        }
        ff[bin] = ff[bin] + (wvn[h] - wvn[h - 1]) / width;
    }
    const int ng = n_pwrk - 1;
    for (int h = 0; h < n_pwrk; ++h) kk[h] = 0.5 * (kk[h] + kk[h + 1]);
    gg[ng] = 1.0;
    for (int h = ng; h > 0; --h) gg[h - 1] = gg[h] - ff[h];
    int t = 0;
    for (int r = 0; r < nq; ++r) {
        while (t + 1 < ng && !(gg[t + 1] >= gq[r])) ++t;
        double v = kk[t] + (kk[t + 1] - kk[t]) * (gq[r] - gg[t]) / (gg[t + 1] - gg[t]);
        if (gg[t + 1] == gg[t]) v = kk[t + 1];
        kq[r] = v;
    }
    double a = 0.0;
    for (int h = 1; h < n_pwrk; ++h) a += kk[h] * ff[h];
    *kavg = a;
}

int orc_ckdist(int nspec, const int64_t *off, const double *wvn, const double *kvn, const double *wvl,
               const double *wvr, int n_pwrk, double pwr, int nq, const double *gq, double *kq, double *kavg,
               double *kmin, double *kmax, int nthreads) {
    if (nspec < 0 || n_pwrk < 2 || nq < 1 || !(pwr > 0.0)) return -1;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
    for (int s = 0; s < nspec; ++s) {
        std::vector<double> kk, ff;
        ck_one(off[s], off[s + 1], wvn, kvn, wvl[s], wvr[s], n_pwrk, pwr, nq, gq, kk, ff, kq + (size_t)s * nq,
               kavg + s, kmin + s, kmax + s);
    }
    return 0;
}

/* spline (src/montecarlo/spline.cpp:180-207), natural ends (yp1 = ypn =
 * 1e30): 1-based x[1..n], y[1..n] -> y2[1..n]; float data, the double
 * constants of the source kept (those products round to float when
 * assigned) */
static void nr_spline(const float *x, const float *y, int n, float *y2) {
    std::vector<float> u(n + 1);
    y2[1] = u[1] = 0.0f;
    for (int i = 2; i <= n - 1; i++) {
        const float sig = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1]);
        const float p = sig * y2[i - 1] + 2.0;
        y2[i] = (sig - 1.0) / p;
        u[i] = (y[i + 1] - y[i]) / (x[i + 1] - x[i]) - (y[i] - y[i - 1]) / (x[i] - x[i - 1]);
        u[i] = (6.0 * u[i] / (x[i + 1] - x[i - 1]) - sig * u[i - 1]) / p;
    }
    const float qn = 0.0f, un = 0.0f;
    y2[n] = (un - qn * u[n - 1]) / (qn * y2[n - 1] + 1.0);
    for (int k = n - 1; k >= 1; k--) y2[k] = y2[k] * y2[k + 1] + u[k];
}

/* splint (spline.cpp:209-226) */
static float nr_splint(const float *xa, const float *ya, const float *y2a, int n, float x) {
    int klo = 1, khi = n;
    while (khi - klo > 1) {
        const int k = (khi + klo) >> 1;
        if (xa[k] > x) khi = k; else klo = k;
    }
    const float h = xa[khi] - xa[klo];
    const float a = (xa[khi] - x) / h, b = (x - xa[klo]) / h;
    return a * ya[klo] + b * ya[khi] + ((a * a * a - a) * y2a[klo] + (b * b * b - b) * y2a[khi]) * (h * h) / 6.0;
}

/* interp3D (spline.cpp:36-177) for one coarse level: splines along x over
 * every fine (j, k), then y over every (coarse i, fine k), then z over
 * every coarse (i, j); the x walls: face means (ghost + first)/2 splined
 * along y then z.  Tc [km][jm][im]; Tw [2][km][jm] (the coarse face mean,
 * since interp3D sets the ghost to 2 face - interior) or NULL.  The source
 * also fills coarse column im+1 before overwriting it with the ghost; that
 * column is not formed here. */
int orc_interp3d(int imax, int jmax, int kmax, const float *x0, const float *y0, const float *z0, const float *T0,
                 int im, int jm, int km, const float *xc, const float *yc, const float *zc, float *Tc, float *Tw) {
    if (imax < 2 || jmax < 2 || kmax < 2 || im < 1 || jm < 1 || km < 1) return -1;
    auto T0at = [&](int i, int j, int k) {
        return T0[(size_t)i + (size_t)(imax + 2) * ((size_t)j + (size_t)(jmax + 2) * (size_t)k)];
    };
    std::vector<float> tx((size_t)(im + 1) * (jmax + 1) * (kmax + 1)), txy((size_t)(im + 1) * (jm + 1) * (kmax + 1));
    auto TX = [&](int i, int j, int k) -> float & { return tx[(size_t)i + (size_t)(im + 1) * ((size_t)j + (size_t)(jmax + 1) * k)]; };
    auto TXY = [&](int i, int j, int k) -> float & { return txy[(size_t)i + (size_t)(im + 1) * ((size_t)j + (size_t)(jm + 1) * k)]; };
    const int nmax = std::max(imax, std::max(jmax, kmax)) + 2;
    std::vector<float> tmp(nmax), tmp2(nmax), tmp3(nmax), tmp4(nmax);
    for (int j = 1; j <= jmax; j++)
        for (int k = 1; k <= kmax; k++) {
            for (int i = 1; i <= imax; i++) tmp[i] = T0at(i, j, k);
            nr_spline(x0, tmp.data(), imax, tmp2.data());
            for (int i = 1; i <= im; i++) TX(i, j, k) = nr_splint(x0, tmp.data(), tmp2.data(), imax, xc[i]);
        }
    for (int i = 1; i <= im; i++)
        for (int k = 1; k <= kmax; k++) {
            for (int j = 1; j <= jmax; j++) tmp[j] = TX(i, j, k);
            nr_spline(y0, tmp.data(), jmax, tmp2.data());
            for (int j = 1; j <= jm; j++) TXY(i, j, k) = nr_splint(y0, tmp.data(), tmp2.data(), jmax, yc[j]);
        }
    for (int i = 1; i <= im; i++)
        for (int j = 1; j <= jm; j++) {
            for (int k = 1; k <= kmax; k++) tmp[k] = TXY(i, j, k);
            nr_spline(z0, tmp.data(), kmax, tmp2.data());
            for (int k = 1; k <= km; k++)
                Tc[(size_t)(i - 1) + (size_t)im * ((size_t)(j - 1) + (size_t)jm * (k - 1))] =
                    nr_splint(z0, tmp.data(), tmp2.data(), kmax, zc[k]);
        }
    if (!Tw) return 0;
    std::vector<float> by((size_t)2 * (jm + 1) * (kmax + 1));
    auto BY = [&](int side, int j, int k) -> float & { return by[(size_t)side + 2 * ((size_t)j + (size_t)(jm + 1) * k)]; };
    for (int k = 1; k <= kmax; k++) {
        for (int j = 1; j <= jmax; j++) {
            tmp[j] = 0.5 * (T0at(0, j, k) + T0at(1, j, k));
            tmp3[j] = 0.5 * (T0at(imax + 1, j, k) + T0at(imax, j, k));
        }
        nr_spline(y0, tmp.data(), jmax, tmp2.data());
        nr_spline(y0, tmp3.data(), jmax, tmp4.data());
        for (int j = 1; j <= jm; j++) {
            BY(0, j, k) = nr_splint(y0, tmp.data(), tmp2.data(), jmax, yc[j]);
            BY(1, j, k) = nr_splint(y0, tmp3.data(), tmp4.data(), jmax, yc[j]);
        }
    }
    for (int j = 1; j <= jm; j++) {
        for (int k = 1; k <= kmax; k++) {
            tmp[k] = BY(0, j, k);
            tmp3[k] = BY(1, j, k);
        }
        nr_spline(z0, tmp.data(), kmax, tmp2.data());
        nr_spline(z0, tmp3.data(), kmax, tmp4.data());
        for (int k = 1; k <= km; k++) {
            Tw[(size_t)(j - 1) + (size_t)jm * (k - 1)] = nr_splint(z0, tmp.data(), tmp2.data(), kmax, zc[k]);
            Tw[(size_t)jm * km + (size_t)(j - 1) + (size_t)jm * (k - 1)] =
                nr_splint(z0, tmp3.data(), tmp4.data(), kmax, zc[k]);
        }
    }
    return 0;
}

/* the estimator's exact mode: p(b, g) = q_bg / Z with q_bg = K_bg(Tmax)
 * B_b(Tmax) w_g dnu_b (K, B interpolated linearly between the rows around
 * Tmax), kappa_max = pi Z / (sigma Tmax^4); walls: p(b) = B_b dnu_b / Zs,
 * Ibmax_s = eps pi Zs / photmax.  With these, the expected tally of a cell
 * is its emission 4 pi sum K B w dnu (fluid) or eps pi sum B dnu (walls)
 * for any Tmax; the reference's nearest-row pdf (mc_gpu_noint.cu:1110-1131)
 * and tabulated kP (read_tables.cpp:35-38) are exact only at tabulated
 * temperatures of self-consistent tables.  FP64 sums in band order. */
int orc_set_estimator(orc_ctx *c, int mode) {
    if (!c || (mode != 0 && mode != 1)) return -1;
    const orc_params &p = c->p;
    if (mode == 0 || p.grey) return 0;
    const int nB = c->nB, nQ = c->nQ, nT = c->nT, tm = c->tm;
    const double f = ((double)c->Tmax - (double)c->Tnb[tm]) / ((double)c->Tnb[tm + 1] - (double)c->Tnb[tm]);
    std::vector<double> qsum(nB), q((size_t)nB * nQ), qs(nB);
    double Z = 0.0, Zs = 0.0;
    for (int b = 0; b < nB; ++b) {
        const int r = c->rank_of[b];
        const float *bt = &c->btab[(size_t)r * (nT + 1)];
        const double Bm = (double)bt[tm] + f * ((double)bt[tm + 1] - (double)bt[tm]);
        const double dnu = ((double)c->wvr[b] - (double)c->wvl[b]) * 100.0;
        double acc = 0.0;
        for (int g = 0; g < nQ; ++g) {
            const float *kt = &c->ktab[((size_t)r * nQ + g) * (nT + 1)];
            const double v = ((double)kt[tm] + f * ((double)kt[tm + 1] - (double)kt[tm])) * Bm * (double)c->wq[g] * dnu;
            q[(size_t)b * nQ + g] = v;
            acc += v;
        }
        qsum[b] = acc;
        Z += acc;
        qs[b] = Bm * dnu;
        Zs += qs[b];
    }
    double cum = 0.0, cums = 0.0;
    for (int b = 0; b < nB; ++b) {
        cum += qsum[b];
        cums += qs[b];
        c->cdf_b[b] = (float)(cum / Z);
        c->cdf_bs[b] = (float)(cums / Zs);
        double cg = 0.0;
        for (int g = 0; g < nQ; ++g) {
            cg += q[(size_t)b * nQ + g];
            c->cdf_g[(size_t)b * nQ + g] = qsum[b] > 0.0 ? (float)(cg / qsum[b]) : (float)(g + 1) / (float)nQ;
        }
    }
    c->cdf_b[nB - 1] = 1.0f;
    c->cdf_bs[nB - 1] = 1.0f;
    const double T4 = (double)c->Tmax * c->Tmax * c->Tmax * c->Tmax;
    c->kappamax64 = kPi * Z / (kStefan * T4);
    c->kappamax = (float)c->kappamax64;
    c->Ibmax64 = 4.0 * kPi * Z / p.photmax;
    c->Ibmax = (float)c->Ibmax64;
    c->Ibmax_s64[0] = (double)p.eps[FW] * kPi * Zs / p.photmax;
    c->Ibmax_s64[1] = (double)p.eps[FE] * kPi * Zs / p.photmax;
    c->Ibmax_s[0] = (float)c->Ibmax_s64[0];
    c->Ibmax_s[1] = (float)c->Ibmax_s64[1];
    return 0;
}

/* the spline routines on one line (0-based arrays in and out): y2[n] of
 * the natural spline through (x, y), and its values ys[m] at xs */
int orc_spline(const float *x, const float *y, int n, float *y2, const float *xs, int m, float *ys) {
    if (n < 2 || m < 0) return -1;
    std::vector<float> x1(n + 1), y1(n + 1), d(n + 1);
    for (int i = 0; i < n; ++i) { x1[i + 1] = x[i]; y1[i + 1] = y[i]; }
    nr_spline(x1.data(), y1.data(), n, d.data());
    for (int i = 0; i < n; ++i) y2[i] = d[i + 1];
    for (int q = 0; q < m; ++q) ys[q] = nr_splint(x1.data(), y1.data(), d.data(), n, xs[q]);
    return 0;
}

/* coarse levels of this field by interp3D (mode 1) or by the texture mean
 * (mode 0, the default) */
int orc_set_restriction(orc_ctx *c, int mode) {
    if (!c || (mode != 0 && mode != 1)) return -1;
    const orc_params &p = c->p;
    for (size_t g = 1; g < c->lev.size(); ++g) {
        Level &L = c->lev[g];
        if (mode == 0) { L.Ts.clear(); L.Tws.clear(); continue; }
        const Level &F = c->lev[0];
        L.Ts.assign((size_t)L.im * L.jm * L.km, 0.f);
        L.Tws.assign((size_t)2 * L.jm * L.km, 0.f);
        if (int e = orc_interp3d(p.imax, p.jmax, p.kmax, F.x.data(), F.y.data(), F.z.data(), c->T.data(), L.im, L.jm,
                                 L.km, L.x.data(), L.y.data(), L.z.data(), L.Ts.data(), L.Tws.data()))
            return e;
    }
    return 0;
}

/* I_black of the generator (nbcorrk/ckdist_hitran_part.cpp:229-242): the
 * exponential in single precision, as there */
static double ck_iblack(double T, double nu) {
    const double C1 = 3.741771790075259e-16, C2 = 0.014387741858429;
    return 1.0 / 3.14159265358979323846 * C1 * std::pow((nu * 100), 3) / (expf((float)(C2 * nu * 100 / T)) - 1);
}

/* calcProbT (nbcorrk/calc_prob.cpp:17-106) after correlatedKT, on spectra
 * ordered row-major (spectrum i*nB + j is band j at temperature T[i]):
 * per row the correlated-k Planck mean kP (its per-bin sum over every band),
 * the band CDF cuprob (per-bin sums over kP) and the g CDF cuprob_g of
 * every band (kq Ib dnu 100 wq, cumulated and divided by the last); plus
 * kq and kavg of every spectrum.  Sums run in the reference's order. */
int orc_ck_tables(int nB, int nT, const int64_t *off, const double *wvn, const double *kvn, const double *wvl,
                  const double *wvr, const double *T, int n_pwrk, double pwr, int nq, const double *gq,
                  const double *wq, double *kq, double *kavg, double *kP, double *cuprob, double *cuprob_g,
                  int nthreads) {
    if (nB < 1 || nT < 1 || n_pwrk < 2 || nq < 1 || !(pwr > 0.0)) return -1;
    const double pi = 3.14159265358979323846, stefan = 5.670373e-08;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
    for (int i = 0; i < nT; ++i) {
        std::vector<double> kk, ff;
        double lo, hi;
        double kPi = 0.0;
        for (int j = 0; j < nB; ++j) {
            const int s = i * nB + j;
            ck_one(off[s], off[s + 1], wvn, kvn, wvl[s], wvr[s], n_pwrk, pwr, nq, gq, kk, ff,
                   kq + (size_t)s * nq, kavg + s, &lo, &hi);
            const double wvc = (wvl[s] + wvr[s]) / 2;
            for (int h = 1; h < n_pwrk; ++h)
                kPi += kk[h] * ck_iblack(T[i], wvc) * (ff[h]) * (wvr[s] - wvl[s]) * 100 * pi /
                       (stefan * std::pow(T[i], 4));
        }
        kP[i] = kPi;
        for (int j = 0; j < nB; ++j) {
            const int s = i * nB + j;
            double dummy;
            ck_one(off[s], off[s + 1], wvn, kvn, wvl[s], wvr[s], n_pwrk, pwr, nq, gq, kk, ff,
                   kq + (size_t)s * nq, &dummy, &lo, &hi);
            const double wvc = (wvl[s] + wvr[s]) / 2;
            double prob = 0.0;
            for (int h = 1; h < n_pwrk; ++h)
                prob += kk[h] * ck_iblack(T[i], wvc) * (ff[h]) * (wvr[s] - wvl[s]) * 100 * pi /
                        (kPi * stefan * std::pow(T[i], 4));
            cuprob[s] = j > 0 ? cuprob[s - 1] + prob : prob;
            double *cg = cuprob_g + (size_t)s * nq;
            for (int g = 0; g < nq; ++g) {
                const double p2 = kq[(size_t)s * nq + g] * ck_iblack(T[i], wvc) * (wvr[s] - wvl[s]) * 100 * wq[g];
                cg[g] = g == 0 ? p2 : cg[g - 1] + p2;
            }
            for (int g = 0; g < nq; ++g) cg[g] /= cg[nq - 1];
        }
    }
    return 0;
}

}  // extern "C"
