/*
 * rc_device.h -- device arithmetic of the radiative solver (gfx950).
 *
 * The operation sequence of every function here is part of the library's
 * numerics contract: FP32 with no implicit contraction (built with
 * -ffp-contract=off), explicit fmaf, IEEE division and sqrt, and no
 * hardware transcendental instructions, so that a CPU restatement with the
 * same sequence reproduces every ray bit for bit.
 */
#ifndef RC_DEVICE_H
#define RC_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rcd {

/* load through a pointer known to address global memory (pointers read from
 * descriptors in memory are otherwise generic and compile to FLAT loads) */
template <class T> __device__ __forceinline__ T gld(const T *p, size_t i) {
    return ((const __attribute__((address_space(1))) T *)p)[i];
}
/* the same for arrays under 4 GiB, with the byte offset formed in 32 bits:
 * with a wavefront-uniform base the load takes a scalar base register and
 * a 32-bit vector offset */
typedef const __attribute__((address_space(1))) char rc_gchar;
template <class T> __device__ __forceinline__ T gld32(const T *p, uint32_t i) {
    return *(const __attribute__((address_space(1))) T *)((rc_gchar *)p + i * (uint32_t)sizeof(T));
}
typedef float rc_v2f __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float2 gld(const float2 *p, size_t i) {
    const rc_v2f v = ((const __attribute__((address_space(1))) rc_v2f *)p)[i];
    return make_float2(v.x, v.y);
}
__device__ __forceinline__ float2 gld32(const float2 *p, uint32_t i) {
    const rc_v2f v = *(const __attribute__((address_space(1))) rc_v2f *)((rc_gchar *)p + i * 8u);
    return make_float2(v.x, v.y);
}
/* two consecutive float2 (8-byte aligned) in one 16-byte load */
typedef float rc_v4f_a8 __attribute__((ext_vector_type(4), aligned(8)));
__device__ __forceinline__ float4 gld32_pair(const float2 *p, uint32_t i) {
    const rc_v4f_a8 v = *(const __attribute__((address_space(1))) rc_v4f_a8 *)((rc_gchar *)p + i * 8u);
    return make_float4(v.x, v.y, v.z, v.w);
}
typedef float rc_v4f __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 gld(const float4 *p, size_t i) {
    const rc_v4f v = ((const __attribute__((address_space(1))) rc_v4f *)p)[i];
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ float4 gld32(const float4 *p, uint32_t i) {
    const rc_v4f v = *(const __attribute__((address_space(1))) rc_v4f *)((rc_gchar *)p + i * 16u);
    return make_float4(v.x, v.y, v.z, v.w);
}

/* the value of lane (lane ^ m) of a wavefront: quad permutes for m = 1, 2,
 * LDS-crossbar swizzles within 32 lanes for m = 4, 8, 16, a permute for 32
 * (m is a constant once the caller's loops are unrolled) */
__device__ __forceinline__ uint32_t lane_xor(uint32_t v, int m) {
    switch (m) {
        case 1: return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false);
        case 2: return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, false);
        case 4: return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, 0x101F);
        case 8: return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, 0x201F);
        case 16: return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, 0x401F);
        default: return (uint32_t)__shfl_xor((int)v, m);
    }
}

/* ---- Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers:
 * as easy as 1, 2, 3", SC'11): 10 rounds of two 32x32->64 multiplies with
 * a Weyl-sequence key schedule. ---------------------------------------- */
__device__ __forceinline__ void philox10(uint32_t &c0, uint32_t &c1, uint32_t &c2, uint32_t &c3,
                                         uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        /* each 32x32 -> 64 product in one instruction (high and low word) */
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t lo0 = (uint32_t)p0, hi0 = (uint32_t)(p0 >> 32);
        const uint32_t lo1 = (uint32_t)p1, hi1 = (uint32_t)(p1 >> 32);
        const uint32_t n0 = hi1 ^ c1 ^ k0;
        const uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

/* uniform on (0,1], multiple of 2^-24 */
__device__ __forceinline__ float u01(uint32_t x) { return (float)((x >> 8) + 1u) * 0x1p-24f; }

struct U4 { float a, b, c, d; };

/* the four uniforms of ray h of run v of cell `cell` in stream `tag` */
__device__ __forceinline__ U4 draw4(uint32_t k0, uint32_t k1, uint32_t h, uint32_t v, uint32_t cell, uint32_t tag) {
    uint32_t c0 = h, c1 = v, c2 = cell, c3 = tag;
    philox10(c0, c1, c2, c3, k0, k1);
    U4 u;
    u.a = u01(c0); u.b = u01(c1); u.c = u01(c2); u.d = u01(c3);
    return u;
}

/* the four raw 32-bit Philox outputs of counter (h, v, cell, tag) */
__device__ __forceinline__ uint4 draw4_bits(uint32_t k0, uint32_t k1, uint32_t h, uint32_t v, uint32_t cell,
                                            uint32_t tag) {
    uint32_t c0 = h, c1 = v, c2 = cell, c3 = tag;
    philox10(c0, c1, c2, c3, k0, k1);
    return make_uint4(c0, c1, c2, c3);
}

/* Point h of the first two dimensions of the Sobol sequence, as 32-bit
 * fractions: dimension 1 has the direction numbers v_k = 2^(31-k) (the van
 * der Corput sequence, the bit reversal of h), dimension 2 those of the
 * primitive polynomial x + 1 with m_1 = 1 (Joe and Kuo's first entry):
 * v_1 = 2^31, v_k = v_(k-1) xor (v_(k-1) >> 1). */
__device__ __forceinline__ void sobol2(uint32_t h, uint32_t &x0, uint32_t &x1) {
    x0 = __builtin_bitreverse32(h);
    /* x1 = xor of v_k over the set bits k of h.  v_k (v_0 = 2^31, v_k =
     * v_(k-1) ^ v_(k-1) >> 1) holds bit 31 - j exactly when the binomial
     * C(k, j) is odd, i.e. when j's bits are a subset of k's (Lucas), so
     * bit 31 - j of x1 is the xor of h's bits at the supersets k of j: a
     * superset-sum over the five index bits, five shift-mask-xor steps,
     * then reversed -- branch-free, the same value as the recurrence. */
    uint32_t f = h;
    f ^= (f >> 1) & 0x55555555u;
    f ^= (f >> 2) & 0x33333333u;
    f ^= (f >> 4) & 0x0F0F0F0Fu;
    f ^= (f >> 8) & 0x00FF00FFu;
    f ^= (f >> 16) & 0x0000FFFFu;
    x1 = __builtin_bitreverse32(f);
}

/* the direction uniforms of ray h under the Sobol sampler: the Sobol
 * point h digitally shifted (xor) by its cell and run's shift words */
__device__ __forceinline__ void sobol_uniforms(uint32_t h, uint32_t s0, uint32_t s1, float &a, float &b) {
    uint32_t x0, x1;
    sobol2(h, x0, x1);
    a = u01(x0 ^ s0);
    b = u01(x1 ^ s1);
}

/* sin and cos of 2 pi u, u in (0,1]: exact quadrant split of 4u, octant
 * fold, Taylor polynomials (degree 9 / 10) on [0, pi/4]. */
__device__ __forceinline__ void sincos2pi(float u, float &s, float &c) {
    const float x = u * 4.0f;
    const int q = (int)x;
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
    const float a = fold ? cp : sp;
    const float b = fold ? sp : cp;
    switch (q & 3) {
        case 0: s = a;  c = b;  break;
        case 1: s = b;  c = -a; break;
        case 2: s = -a; c = -b; break;
        default: s = -b; c = a; break;
    }
}

/* expm1(x) and exp(x): x = n ln2 + r with |r| <= ln2/2 (Cody-Waite, fma),
 * expm1(r) from its Taylor series through r^9/9!, then rescaled by 2^n. */
__device__ __forceinline__ void expm1_exp(float x, float &em1, float &e) {
    const bool under = x < -87.0f;
    float xc = under ? -87.0f : x;
    xc = (xc > 88.0f) ? 88.0f : xc;
    const float n = rintf(xc * 1.44269502f);
    float r = fmaf(-n, 0.693145752f, xc);
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
    const float sc = __int_as_float((ni + 127) << 23);
    const float ev = fmaf(sc, p, sc);
    const float mv = (ni == 0) ? p : fmaf(sc, p, sc - 1.0f);
    e = under ? 0.0f : ev;
    em1 = under ? -1.0f : mv;
}

/* expm1_exp for x <= 88 (the march's exponents: minus an optical depth,
 * which is never below -0 by more than a rounding): the same values,
 * without the clamp at the top of the range */
__device__ __forceinline__ void expm1_exp_le88(float x, float &em1, float &e) {
    const bool under = x < -87.0f;
    const float xc = under ? -87.0f : x;
    const float n = rintf(xc * 1.44269502f);
    float r = fmaf(-n, 0.693145752f, xc);
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
    const float sc = __int_as_float((ni + 127) << 23);
    const float ev = fmaf(sc, p, sc);
    const float mv = (ni == 0) ? p : fmaf(sc, p, sc - 1.0f);
    e = under ? 0.0f : ev;
    em1 = under ? -1.0f : mv;
}

/* first index with cdf[i] > R; R == 1 and past-the-end map to n-1 */
__device__ __forceinline__ int upper_bound(const float *__restrict__ cdf, int n, float R) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (R < gld(cdf, mid)) hi = mid; else lo = mid + 1;
    }
    if (R == 1.0f) lo = n - 1;
    if (lo > n - 1) lo = n - 1;
    return lo;
}

/* upper_bound with a guide table g[m] = first index with cdf > m/M
 * (Chen and Asau, 1974): the answer lies in [g[m-1], g[m+2]] for
 * m = floor(R M), one slot of slack either side against rounding of R M.
 * Returns exactly what upper_bound returns. */
__device__ __forceinline__ int upper_bound_guided(const float *__restrict__ cdf, int n, const int *__restrict__ g,
                                                  int M, float R) {
    int m = (int)(R * (float)M);
    int lo = gld(g, max(m - 1, 0)), hi = gld(g, min(m + 2, M));
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (R < gld(cdf, mid)) hi = mid; else lo = mid + 1;
    }
    if (R == 1.0f) lo = n - 1;
    if (lo > n - 1) lo = n - 1;
    return lo;
}

/* upper_bound over a short non-decreasing CDF: count of entries <= R, all
 * loads issued together; a row of a multiple of 4 entries (16-byte aligned,
 * as every g CDF row is) is read as 16-byte words */
__device__ __forceinline__ int upper_bound_short(const float *__restrict__ cdf, int n, float R) {
    int c = 0;
    if ((n & 3) == 0) {
        const float4 *c4 = reinterpret_cast<const float4 *>(cdf);
        for (int q = 0; q < (n >> 2); ++q) {
            const float4 v = gld(c4, q);
            c += ((v.x <= R) ? 1 : 0) + ((v.y <= R) ? 1 : 0) + ((v.z <= R) ? 1 : 0) + ((v.w <= R) ? 1 : 0);
        }
    } else {
        for (int q = 0; q < n; ++q) c += (gld(cdf, q) <= R) ? 1 : 0;
    }
    if (R == 1.0f) c = n - 1;
    if (c > n - 1) c = n - 1;
    return c;
}

__device__ __forceinline__ float lerp(float a, float b, float w) { return fmaf(w, b - a, a); }

/* row t and weight of temperature T in a table starting at Tnb0 with
 * spacing 1/invdT, clamped to the first and last rows */
/* the table position of temperature T: (T - Tnb0) invdT clamped to
 * [0, nT - 1] (0 for NaN), so that ppos() of it gives exactly the t and a
 * of tpos(T) */
__device__ __forceinline__ float table_pos(float T, float Tnb0, float invdT, int nT) {
    const float pos = (T - Tnb0) * invdT;
    const bool lo = !(pos > 0.0f);
    const bool hi = pos >= (float)(nT - 1);
    return lo ? 0.0f : (hi ? (float)(nT - 1) : pos);
}

/* row t and interpolation weight a of a table position p >= 0 */
__device__ __forceinline__ void ppos(float p, int &t, float &a) {
    t = (int)p;
    a = p - (float)t;
}

__device__ __forceinline__ void tpos(float T, float Tnb0, float invdT, int nT, int &t, float &a) {
    const float pos = (T - Tnb0) * invdT;
    const bool lo = !(pos > 0.0f);
    const bool hi = pos >= (float)(nT - 1);
    const int ti = (int)pos;
    t = lo ? 0 : (hi ? nT - 1 : ti);
    a = (lo || hi) ? 0.0f : pos - (float)ti;
}

/* isotropic direction from (u0, u1): phi = 2 pi u0, cos theta = 1 - 2 u1 */
__device__ __forceinline__ void iso_dir(float u0, float u1, float &sx, float &sy, float &sz) {
    const float ct = 1.0f - 2.0f * u1;
    const float st = 2.0f * sqrtf(u1 * (1.0f - u1));
    float sp, cp;
    sincos2pi(u0, sp, cp);
    sx = ct; sy = st * cp; sz = st * sp;
}

/* Lambertian direction into the domain from the west (bound 0) or east
 * (bound 1) wall: sin^2 theta = u1 */
__device__ __forceinline__ void lamb_dir(float u0, float u1, int bound, float &sx, float &sy, float &sz) {
    const float a = sqrtf(1.0f - u1);
    const float b = sqrtf(u1);
    float sp, cp;
    sincos2pi(u0, sp, cp);
    if (bound == 0) { sx = a;  sy = b * cp; sz = b * sp; }
    else            { sx = -b; sy = a * cp; sz = a * sp; }
}

__device__ __forceinline__ uint32_t fnv(uint32_t h, uint32_t w) { return (h ^ w) * 16777619u; }

}  // namespace rcd

#endif
