/*
 * rc_solver.hip -- MI355X (gfx950) narrow-band correlated-k OERM Monte Carlo
 * solver for the radiative source term div(q_r) of a channel flow, a drop-in
 * for RadChan's src/montecarlo path on one GPU.
 *
 * Reference path (RadChan): src/montecarlo/mc_gpu_noint.cu
 *   kernel_fluid :55-243, march_ray :245-658, mc_gpu_ :660-874,
 *   get_results_ :876-1021, find_band :1103-1169, wave_find :1234-1267,
 *   kernel_surface :1316-1482; tables src/montecarlo/memory_copy_noint.cu.
 *
 * Device data layout (all structure-of-arrays in HBM):
 *   per snapshot and grid level: the table positions of the cells
 *   [km][jm][im] (the temperature the reference samples with a texture at
 *   the mapped cell centre, mapped to the table's row coordinate) and of
 *   the black walls (DevLevel); x faces per level;
 *   tables in sorted band ("rank") order: kb[(rank*nQ+g)*S+t] = {k, Ib}
 *   pairs, S = tab_stride(nT), one 16-byte load per march step, the last
 *   row repeated so that the clamped top row needs no branch, then the
 *   fixed-temperature walls' entries.
 *
 * Variants:
 *   plain   -- the reference's launch shape: one thread per emitting cell,
 *              photmax rays in sequence, FP32 per-cell sums, a sort pre-pass
 *              per run (nVar launches of each kernel);
 *   batched -- see rc_batched.hip.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "rc_device.h"
#include "rc_internal.h"
#include "rc_kernels.h"
#include "rc_march.h"
#include "rc_radix.h"
#include "rc_march64.h"

using namespace rcd;

namespace rc {

/* ------------------------------------------------------------------------ */
/* set-up kernels                                                            */

/* linear texture fetch of the ghost-padded temperature volume at
 * unnormalised texel coordinates with clamp-to-edge addressing (the
 * reference's cudaFilterModeLinear + cudaAddressModeClamp texture,
 * memory_copy_noint.cu:131-168): texel i is centred at i + 0.5 */
__device__ float tex3_T(const float *__restrict__ T, int nx, int ny, int nz, float tx, float ty, float tz) {
    const int n[3] = {nx, ny, nz};
    const float crd[3] = {tx, ty, tz};
    int i0[3], i1[3];
    float w[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const float b = crd[d] - 0.5f;
        const float fl = floorf(b);
        w[d] = b - fl;
        int lo = (int)fl, hi = lo + 1;
        lo = min(max(lo, 0), n[d] - 1);
        hi = min(max(hi, 0), n[d] - 1);
        i0[d] = lo; i1[d] = hi;
    }
    auto at = [&](int i, int j, int k) { return T[(size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * (size_t)k)]; };
    const float c00 = lerp(at(i0[0], i0[1], i0[2]), at(i1[0], i0[1], i0[2]), w[0]);
    const float c10 = lerp(at(i0[0], i1[1], i0[2]), at(i1[0], i1[1], i0[2]), w[0]);
    const float c01 = lerp(at(i0[0], i0[1], i1[2]), at(i1[0], i0[1], i1[2]), w[0]);
    const float c11 = lerp(at(i0[0], i1[1], i1[2]), at(i1[0], i1[1], i1[2]), w[0]);
    const float c0 = lerp(c00, c10, w[1]);
    const float c1 = lerp(c01, c11, w[1]);
    return lerp(c0, c1, w[2]);
}

/* mapped texel coordinate of coarse index ic at ratio r (:334-336) */
__device__ __forceinline__ float tmap(int ic, int r) { return ((float)ic - 0.5f) * (float)r + 1.0f; }

/* how the table position of a black wall face is found: 0 not a black
 * wall, 1 the field's texture at the wall (adiabatic x walls, bd 4 z
 * walls), 2 the face's fixed temperature */
struct FaceModes {
    int m[6];
};

FaceModes face_modes(const rc_params &p) {
    FaceModes f{};
    for (int s = 0; s < 2; ++s) {
        const int adiab = s ? p.adiabe : p.adiabw;
        f.m[FW + s] = p.bd[FW + s] == 2 ? (adiab == 1 ? 1 : 2) : 0;
        f.m[FS + s] = p.bd[FS + s] == 2 ? 2 : 0;
        f.m[FB + s] = p.bd[FB + s] == 4 ? 1 : (p.bd[FB + s] == 2 ? 2 : 0);
    }
    return f;
}

/* per-level table positions (layout of DevLevel.P): cells at the texture
 * sample of the mapped cell centre (:334-338), adiabatic x walls at texel
 * x = 1.0 / imax + 1.0 (:391-399, :422-430), bd 4 z walls at the adjacent
 * cell's centre plane (:466-475, :504-513), fixed-temperature walls at
 * their face's fixed position */
__global__ void __launch_bounds__(256) k_level_P(const float *__restrict__ T0, int imax, int jmax, int kmax, int im,
                                                 int jm, int km, FaceModes fm, float Tnb0, float invdT, int nT,
                                                 float *__restrict__ P, float *__restrict__ Tout) {
    const int rx = imax / im, ry = jmax / jm, rz = kmax / km;
    const int nx = imax + 2, ny = jmax + 2, nz = kmax + 2;
    const size_t ncell = (size_t)im * jm * km;
    const size_t nfx = (size_t)jm * km, nfy = (size_t)im * km, nfz = (size_t)im * jm;
    const size_t total = level_pos_count(im, jm, km);
    for (size_t id = blockIdx.x * (size_t)blockDim.x + threadIdx.x; id < total; id += (size_t)gridDim.x * blockDim.x) {
        if (id < ncell) {
            const int i = (int)(id % im) + 1, j = (int)((id / im) % jm) + 1, k = (int)(id / ((size_t)im * jm)) + 1;
            const float Tc = tex3_T(T0, nx, ny, nz, tmap(i, rx), tmap(j, ry), tmap(k, rz));
            P[id] = table_pos(Tc, Tnb0, invdT, nT);
            if (Tout) Tout[id] = Tc;
            continue;
        }
        size_t q = id - ncell;
        float v = 0.f, tw = 0.f;
        if (q < 2 * nfx) {
            const int b = (int)(q / nfx);
            const size_t r = q % nfx;
            const int j = (int)(r % jm) + 1, k = (int)(r / jm) + 1;
            const int mode = fm.m[FW + b];
            if (mode == 1) {
                const float tx = b ? (float)imax + 1.0f : 1.0f;
                tw = tex3_T(T0, nx, ny, nz, tx, tmap(j, ry), tmap(k, rz));
                v = table_pos(tw, Tnb0, invdT, nT);
            } else if (mode == 2) {
                v = fixed_wall_pos(nT, FW + b);
            }
        } else if ((q -= 2 * nfx) < 2 * nfy) {
            const int b = (int)(q / nfy);
            if (fm.m[FS + b] == 2) v = fixed_wall_pos(nT, FS + b);
        } else {
            q -= 2 * nfy;
            const int b = (int)(q / nfz);
            const size_t r = q % nfz;
            const int i = (int)(r % im) + 1, j = (int)(r / im) + 1;
            const int mode = fm.m[FB + b];
            if (mode == 1) {
                const float tz = b ? (float)kmax + 0.5f : 1.5f;
                tw = tex3_T(T0, nx, ny, nz, tmap(i, rx), tmap(j, ry), tz);
                v = table_pos(tw, Tnb0, invdT, nT);
            } else if (mode == 2) {
                v = fixed_wall_pos(nT, FB + b);
            }
        }
        P[id] = v;
        if (Tout) Tout[id] = tw;
    }
}

/* cell (and adiabatic x-wall) temperatures of a level, e.g. the spline
 * restriction's, to table positions in P (and, for the FP64 path, kept in
 * Tout beside them) */
__global__ void __launch_bounds__(256) k_pos_from_T(const float *__restrict__ Tl, const float *__restrict__ Tw,
                                                    int im, int jm, int km, FaceModes fm, float Tnb0, float invdT,
                                                    int nT, float *__restrict__ P, float *__restrict__ Tout) {
    const size_t ncell = (size_t)im * jm * km, nfx = (size_t)jm * km;
    const size_t total = ncell + (Tw ? 2 * nfx : 0);
    for (size_t id = blockIdx.x * (size_t)blockDim.x + threadIdx.x; id < total; id += (size_t)gridDim.x * blockDim.x) {
        if (id < ncell) {
            P[id] = table_pos(Tl[id], Tnb0, invdT, nT);
            if (Tout) Tout[id] = Tl[id];
        } else {
            const size_t q = id - ncell;
            if (fm.m[FW + (int)(q / nfx)] == 1) {
                P[id] = table_pos(Tw[q], Tnb0, invdT, nT);
                if (Tout) Tout[id] = Tw[q];
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* the srt enumeration                                                       */

/* the radix sort's LDS: photmax keys (rounded up to 64) twice as 16 bits,
 * then two 256-bucket histograms; within 64 KiB up to 15 872 keys */
__host__ __device__ __forceinline__ bool sort_radix(int P) { return P > 256 && ((P + 63) & ~63) * 4 + 2048 <= 65536; }

/* the Philox stream of sort unit u: cell id, run and tag (explicit
 * (cell id, run, tag) triples, else fluid cells or the two walls' cells of
 * the slab) */
__device__ __forceinline__ void unit_stream(const DevCommon &cm, int u, int surface, const int32_t *ulist,
                                            uint32_t &cid, int &vv, uint32_t &tag) {
    if (ulist) {
        cid = (uint32_t)ulist[3 * u]; vv = ulist[3 * u + 1]; tag = (uint32_t)ulist[3 * u + 2];
    } else if (!surface) {
        cid = (uint32_t)u + (uint32_t)cm.imax * cm.jmax * (cm.zstart - 1); tag = TAG_FLUID;
    } else {
        const int b = u / cm.nsurf;
        cid = (uint32_t)(u % cm.nsurf) + (uint32_t)cm.jmax * (cm.zstart - 1); tag = b ? TAG_EAST : TAG_WEST;
    }
}

/* One workgroup (one wavefront) per (cell, run): draws photmax keys and
 * sorts them in LDS, by a bitonic network up to 256 keys (and above the
 * radix sort's LDS, 15 872 keys), else by the two-pass LSD radix sort the
 * batched kernel uses (8-bit digits; O(photmax) like the reference's
 * histogram enumeration, where the network's stages grow as log^2).  The
 * sorted sequence is the one the reference's histogram enumeration
 * produces (kernel_find :1023-1049 + wave_find :1234-1267): ascending
 * sorted-band rank, then g; equal keys are identical draws, so any correct
 * sort gives the same sequence. */
__global__ void __launch_bounds__(64) k_sort_bands(DevCommon cm, const DevSnap *__restrict__ snaps, int s, int v,
                                                   int surface, int nunit, const int32_t *__restrict__ ulist,
                                                   uint16_t *__restrict__ keys) {
    extern __shared__ uint32_t lds[];
    const DevSnap &sn = snaps[s];
    const int P = cm.photmax;
    const int lane = threadIdx.x;
    if (sort_radix(P)) {
        /* keys in draw order at ka, after the low-byte pass at kb, the two
         * digit histograms after them */
        const int Pr = (P + 63) & ~63;
        uint16_t *const ka = reinterpret_cast<uint16_t *>(lds);
        uint16_t *const kb = ka + Pr;
        uint32_t *const hist = lds + Pr;
        for (int u = blockIdx.x; u < nunit; u += gridDim.x) {
            uint32_t cid, tag;
            int vv = v;
            unit_stream(cm, u, surface, ulist, cid, vv, tag);
            const uint32_t tagx = tag + 16u * (uint32_t)sn.snap;
            for (int w = lane; w < 512; w += 64) hist[w] = 0u;
            __syncthreads();
            for (int h = lane; h < P; h += 64) {
                const U4 r = draw4(cm.k0, cm.k1, (uint32_t)h, (uint32_t)vv, cid, tagx);
                const uint32_t key = band_key(cm, sn, tag != TAG_FLUID, r.c, r.d);
                ka[h] = (uint16_t)key;
                atomicAdd(&hist[key & 255u], 1u);
                atomicAdd(&hist[256u + (key >> 8)], 1u);
            }
            __syncthreads();
            radix_offsets(hist, lane);
            radix_pass(ka, kb, hist, P, 0, lane);
            radix_pass(kb, ka, hist + 256, P, 8, lane);
            for (int h = lane; h < P; h += 64) keys[(size_t)u * P + h] = ka[h];
            __syncthreads();
        }
        return;
    }
    int n2 = 1;
    while (n2 < P) n2 <<= 1;
    for (int u = blockIdx.x; u < nunit; u += gridDim.x) {
        uint32_t cid, tag;
        int vv = v;
        unit_stream(cm, u, surface, ulist, cid, vv, tag);
        const uint32_t tagx = tag + 16u * (uint32_t)sn.snap;
        for (int h = threadIdx.x; h < n2; h += blockDim.x) {
            uint32_t key = 0xFFFFFFFFu;
            if (h < P) {
                const U4 r = draw4(cm.k0, cm.k1, (uint32_t)h, (uint32_t)vv, cid, tagx);
                key = band_key(cm, sn, tag != TAG_FLUID, r.c, r.d);
            }
            lds[h] = key;
        }
        __syncthreads();
        for (int k = 2; k <= n2; k <<= 1)
            for (int j = k >> 1; j > 0; j >>= 1) {
                for (int i = threadIdx.x; i < n2; i += blockDim.x) {
                    const int ixj = i ^ j;
                    if (ixj > i) {
                        const uint32_t a = lds[i], b = lds[ixj];
                        const bool up = (i & k) == 0;
                        if ((a > b) == up) { lds[i] = b; lds[ixj] = a; }
                    }
                }
                __syncthreads();
            }
        for (int h = threadIdx.x; h < P; h += blockDim.x) keys[(size_t)u * P + h] = (uint16_t)lds[h];
        __syncthreads();
    }
}

/* ------------------------------------------------------------------------ */
/* plain variant: the reference's launch shape                               */

__global__ void __launch_bounds__(128) k_fluid_plain(DevCommon cm, const DevSnap *__restrict__ snaps, int s, int v,
                                                     const uint16_t *__restrict__ keys, double *__restrict__ de,
                                                     unsigned long long *__restrict__ stats) {
    const DevSnap &sn = snaps[s];
    int64_t steps = 0;
    int64_t rays = 0;
    for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < cm.ncell; n += gridDim.x * blockDim.x) {
        int i, j, k;
        cell_ijk(cm, n, i, j, k);
        const uint32_t cid = (uint32_t)n + (uint32_t)cm.imax * cm.jmax * (cm.zstart - 1);
        const float Ti = cell_T(cm, sn, i, j, k);
        float acc = 0.f;
        int st = 0;
        for (int h = 0; h < cm.photmax; ++h) {
            const uint32_t key = keys ? (uint32_t)keys[(size_t)n * cm.photmax + h] : 0u;
            acc += fluid_ray<false>(cm, sn, i, j, k, cid, v, h, key, keys != nullptr, Ti, st, nullptr);
        }
        steps += st;
        rays += cm.photmax;
        de[(size_t)n * cm.nVar + v] = (double)acc;
    }
    /* wavefront-reduced counters */
    for (int o = 32; o > 0; o >>= 1) {
        steps += __shfl_xor(steps, o);
        rays += __shfl_xor(rays, o);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&stats[0], (unsigned long long)rays);
        atomicAdd(&stats[1], (unsigned long long)steps);
    }
}

/* plain variant, one launch per band: the rays of sorted band `rank` of
 * every emitting cell (a contiguous range of the cell's sorted keys), added
 * to the cell's FP32 sum acc[n]; launched for every band in rank order, it
 * adds each cell's rays in the order k_fluid_plain does, so the sums are
 * the same to the bit */
__global__ void __launch_bounds__(128) k_fluid_band(DevCommon cm, const DevSnap *__restrict__ snaps, int s, int v,
                                                    int rank, const uint16_t *__restrict__ keys,
                                                    float *__restrict__ acc, unsigned long long *__restrict__ stats) {
    const DevSnap &sn = snaps[s];
    const int P = cm.photmax;
    const uint32_t klo = (uint32_t)rank * (uint32_t)cm.nQ, khi = klo + (uint32_t)cm.nQ;
    int64_t steps = 0, rays = 0;
    for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < cm.ncell; n += gridDim.x * blockDim.x) {
        const uint16_t *kk = keys + (size_t)n * P;
        /* first ray with key >= klo, and with key >= khi */
        int a = 0, b = P;
        while (a < b) { const int m = (a + b) >> 1; if ((uint32_t)kk[m] < klo) a = m + 1; else b = m; }
        int c = a, d = P;
        while (c < d) { const int m = (c + d) >> 1; if ((uint32_t)kk[m] < khi) c = m + 1; else d = m; }
        if (a == c) continue;
        int i, j, k;
        cell_ijk(cm, n, i, j, k);
        const uint32_t cid = (uint32_t)n + (uint32_t)cm.imax * cm.jmax * (cm.zstart - 1);
        const float Ti = cell_T(cm, sn, i, j, k);
        float sum = acc[n];
        int st = 0;
        for (int h = a; h < c; ++h) sum += fluid_ray<false>(cm, sn, i, j, k, cid, v, h, (uint32_t)kk[h], true, Ti, st, nullptr);
        acc[n] = sum;
        steps += st;
        rays += c - a;
    }
    for (int o = 32; o > 0; o >>= 1) {
        steps += __shfl_xor(steps, o);
        rays += __shfl_xor(rays, o);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&stats[0], (unsigned long long)rays);
        atomicAdd(&stats[1], (unsigned long long)steps);
    }
}

__global__ void k_acc_zero(float *__restrict__ acc, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) acc[i] = 0.f;
}

__global__ void k_acc_store(const float *__restrict__ acc, int n, int nVar, int v, double *__restrict__ de) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        de[(size_t)i * nVar + v] = (double)acc[i];
}

__global__ void __launch_bounds__(128) k_surface_plain(DevCommon cm, const DevSnap *__restrict__ snaps, int s, int v,
                                                       const uint16_t *__restrict__ keys, double *__restrict__ de,
                                                       unsigned long long *__restrict__ stats) {
    const DevSnap &sn = snaps[s];
    int64_t steps = 0, rays = 0;
    const int nsurf = cm.nsurf;
    for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < 2 * nsurf; n += gridDim.x * blockDim.x) {
        const int bound = n / nsurf, m = n % nsurf;
        if ((bound == 0 && !cm.adiab[0]) || (bound == 1 && !cm.adiab[1])) continue;
        const int j = m % cm.jmax + 1, k = m / cm.jmax + cm.zstart;
        const uint32_t cid = (uint32_t)m + (uint32_t)cm.jmax * (cm.zstart - 1);
        const float Tw = wall_T(cm, sn, bound, j, k);
        float acc = 0.f;
        int st = 0;
        for (int h = 0; h < cm.photmax; ++h) {
            const uint32_t key = keys ? (uint32_t)keys[(size_t)n * cm.photmax + h] : 0u;
            acc += surface_ray<false>(cm, sn, bound, j, k, cid, v, h, key, keys != nullptr, Tw, st, nullptr);
        }
        steps += st;
        rays += cm.photmax;
        de[(size_t)n * cm.nVar + v] = (double)acc;
    }
    for (int o = 32; o > 0; o >>= 1) {
        steps += __shfl_xor(steps, o);
        rays += __shfl_xor(rays, o);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&stats[0], (unsigned long long)rays);
        atomicAdd(&stats[1], (unsigned long long)steps);
    }
}

/* ------------------------------------------------------------------------ */
/* full-FP64 variant (floatpoint 1): the reference's launch shape, a thread  */
/* per emitting (wall) cell, rays in sequence, FP64 per-cell sums.  Both     */
/* kernels are held to 128 VGPRs (4 wavefronts per SIMD, a few spills): 7 %  */
/* faster at 192x128x192 than with the 151 registers and 3 wavefronts they   */
/* take unbounded.                                                          */
__global__ void __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(4, 8)))
k_fluid_f64(DevCommon cm, const DevSnap *__restrict__ snaps, int s, int v, const uint16_t *__restrict__ keys,
            double *__restrict__ de, unsigned long long *__restrict__ stats) {
    const DevSnap &sn = snaps[s];
    int64_t steps = 0, rays = 0;
    for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < cm.ncell; n += gridDim.x * blockDim.x) {
        int i, j, k;
        cell_ijk(cm, n, i, j, k);
        const uint32_t cid = (uint32_t)n + (uint32_t)cm.imax * cm.jmax * (cm.zstart - 1);
        const float Ti = cell_T(cm, sn, i, j, k);
        double acc = 0.0;
        int st = 0;
        for (int h = 0; h < cm.photmax; ++h) {
            const uint32_t key = keys ? (uint32_t)keys[(size_t)n * cm.photmax + h] : 0u;
            acc += fluid_ray64(cm, sn, i, j, k, cid, v, h, key, keys != nullptr, Ti, st);
        }
        steps += st;
        rays += cm.photmax;
        de[(size_t)n * cm.nVar + v] = acc;
    }
    for (int o = 32; o > 0; o >>= 1) {
        steps += __shfl_xor(steps, o);
        rays += __shfl_xor(rays, o);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&stats[0], (unsigned long long)rays);
        atomicAdd(&stats[1], (unsigned long long)steps);
    }
}

__global__ void __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(4, 8)))
k_surface_f64(DevCommon cm, const DevSnap *__restrict__ snaps, int s, int v, const uint16_t *__restrict__ keys,
              double *__restrict__ de, unsigned long long *__restrict__ stats) {
    const DevSnap &sn = snaps[s];
    int64_t steps = 0, rays = 0;
    for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < 2 * cm.nsurf; n += gridDim.x * blockDim.x) {
        const int bound = n / cm.nsurf, w = n % cm.nsurf;
        if (!cm.adiab[bound]) continue;
        const int j = w % cm.jmax + 1, k = w / cm.jmax + cm.zstart;
        const uint32_t cid = (uint32_t)(w + cm.jmax * (cm.zstart - 1));
        const float Tw = wall_T(cm, sn, bound, j, k);
        double acc = 0.0;
        int st = 0;
        for (int h = 0; h < cm.photmax; ++h) {
            const uint32_t key = keys ? (uint32_t)keys[(size_t)n * cm.photmax + h] : 0u;
            acc += surface_ray64(cm, sn, bound, j, k, cid, v, h, key, keys != nullptr, Tw, st);
        }
        steps += st;
        rays += cm.photmax;
        de[(size_t)n * cm.nVar + v] = acc;
    }
    for (int o = 32; o > 0; o >>= 1) {
        steps += __shfl_xor(steps, o);
        rays += __shfl_xor(rays, o);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&stats[0], (unsigned long long)rays);
        atomicAdd(&stats[1], (unsigned long long)steps);
    }
}

/* traced rays: rays[5*q] = (i or bound, j, k, v, h); unit[q] indexes keys */
__global__ void __launch_bounds__(64) k_trace(DevCommon cm, const DevSnap *__restrict__ snaps, int surface,
                                              const int32_t *__restrict__ rays, const int32_t *__restrict__ unit,
                                              int n, const uint16_t *__restrict__ keys, rc_ray *__restrict__ out) {
    const DevSnap &sn = snaps[0];
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int a = rays[5 * q], j = rays[5 * q + 1], k = rays[5 * q + 2], v = rays[5 * q + 3], h = rays[5 * q + 4];
    const bool hk = keys != nullptr;
    const uint32_t key = hk ? (uint32_t)keys[(size_t)unit[q] * cm.photmax + h] : 0u;
    int st = 0;
    if (!surface) {
        const uint32_t cid = (uint32_t)((a - 1) + cm.imax * ((j - 1) + cm.jmax * (k - 1)));
        fluid_ray<true>(cm, sn, a, j, k, cid, v, h, key, hk, cell_T(cm, sn, a, j, k), st, &out[q]);
    } else {
        const uint32_t cid = (uint32_t)((j - 1) + cm.jmax * (k - 1));
        surface_ray<true>(cm, sn, a, j, k, cid, v, h, key, hk, wall_T(cm, sn, a, j, k), st, &out[q]);
    }
}

/* mean over runs and the reference's variance 1/((n-1/2)(n-1)) sum (x-m)^2
 * (mc_gpu_noint.cu:234-239), reported as its square root (:984); 0 for a
 * single run */
__global__ void k_finalize(const double *__restrict__ de, int n, int nVar, float *__restrict__ mean,
                           float *__restrict__ sdev) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
        double m = 0.0;
        for (int v = 0; v < nVar; ++v) m += de[(size_t)c * nVar + v];
        m /= nVar;
        double var = 0.0;
        if (nVar > 1) {
            for (int v = 0; v < nVar; ++v) {
                const double d = de[(size_t)c * nVar + v] - m;
                var += d * d;
            }
            var = var / (nVar - 0.5) / (nVar - 1);
        }
        mean[c] = (float)m;
        sdev[c] = (float)sqrt(var);
    }
}

/* ------------------------------------------------------------------------ */
/* launchers used by the session code                                        */

void launch_level_P(hipStream_t st, const float *T0, const rc_params &p, int g, float Tnb0, float invdT, int nT,
                    float *P, float *Tout) {
    const size_t n = level_pos_count(p.maxi[g], p.maxj[g], p.maxk[g]);
    const int blocks = (int)std::min<size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_level_P, dim3(blocks), dim3(256), 0, st, T0, p.imax, p.jmax, p.kmax, p.maxi[g], p.maxj[g],
                       p.maxk[g], face_modes(p), Tnb0, invdT, nT, P, Tout);
}

void launch_pos_from_T(hipStream_t st, const rc_params &p, int g, const float *Tl, const float *Tw, float Tnb0,
                       float invdT, int nT, float *P, float *Tout) {
    const size_t n = (size_t)p.maxi[g] * p.maxj[g] * p.maxk[g] + (size_t)2 * p.maxj[g] * p.maxk[g];
    const int blocks = (int)std::min<size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_pos_from_T, dim3(blocks), dim3(256), 0, st, Tl, Tw, p.maxi[g], p.maxj[g], p.maxk[g],
                       face_modes(p), Tnb0, invdT, nT, P, Tout);
}

void launch_sort(hipStream_t st, const DevCommon &cm, const DevSnap *snaps, int s, int v, int surface, int nunit,
                 const int32_t *ulist, uint16_t *keys) {
    int n2 = 1;
    while (n2 < cm.photmax) n2 <<= 1;
    const size_t lds = sort_radix(cm.photmax) ? (size_t)((cm.photmax + 63) & ~63) * 4 + 2048 : (size_t)n2 * 4;
    const int blocks = std::min(nunit, 1 << 20);
    hipLaunchKernelGGL(k_sort_bands, dim3(blocks), dim3(64), lds, st, cm, snaps, s, v, surface, nunit,
                       ulist, keys);
}

void launch_fluid_plain(hipStream_t st, const DevCommon &cm, const DevSnap *snaps, int s, int v, const uint16_t *keys,
                        double *de, unsigned long long *stats) {
    const int blocks = std::max(1, std::min((cm.ncell + 127) / 128, 1 << 20));
    hipLaunchKernelGGL(k_fluid_plain, dim3(blocks), dim3(128), 0, st, cm, snaps, s, v, keys, de, stats);
}

void launch_fluid_bands(hipStream_t st, const DevCommon &cm, const DevSnap *snaps, int s, int v, const uint16_t *keys,
                        float *acc, double *de, unsigned long long *stats) {
    const int blocks = std::max(1, std::min((cm.ncell + 127) / 128, 1 << 20));
    const int zb = std::max(1, std::min((cm.ncell + 255) / 256, 65536));
    hipLaunchKernelGGL(k_acc_zero, dim3(zb), dim3(256), 0, st, acc, cm.ncell);
    for (int rank = 0; rank < cm.nB; ++rank)
        hipLaunchKernelGGL(k_fluid_band, dim3(blocks), dim3(128), 0, st, cm, snaps, s, v, rank, keys, acc, stats);
    hipLaunchKernelGGL(k_acc_store, dim3(zb), dim3(256), 0, st, acc, cm.ncell, cm.nVar, v, de);
}

void launch_surface_plain(hipStream_t st, const DevCommon &cm, const DevSnap *snaps, int s, int v,
                          const uint16_t *keys, double *de, unsigned long long *stats) {
    const int blocks = std::max(1, std::min((2 * cm.nsurf + 127) / 128, 1 << 20));
    hipLaunchKernelGGL(k_surface_plain, dim3(blocks), dim3(128), 0, st, cm, snaps, s, v, keys, de, stats);
}

/* Input check and Tmax of a temperature field (imax+2)(jmax+2)(kmax+2) on
 * the device: out[0] |= 1 when an entry (ghosts included) is not a
 * positive finite number; out[1] = the largest interior temperature and
 * x-wall midpoint (T_0 + T_1)/2, (T_imax + T_imax+1)/2 of the interior rows
 * (mc_gpu_noint.cu:696-710), as float bits: positive floats order as their
 * bits.  out zero on entry. */
__global__ void __launch_bounds__(256) k_field_scan(const float *__restrict__ T, int imax, int jmax, int kmax,
                                                    uint32_t *__restrict__ out) {
    const uint32_t nx = (uint32_t)imax + 2u, ny = (uint32_t)jmax + 2u;
    const size_t n = (size_t)nx * ny * (size_t)(kmax + 2);
    uint32_t bad = 0u, m = 0u;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const float t = T[q];
        bad |= (t > 0.f && t <= 3.40282347e38f) ? 0u : 1u;
        const uint32_t row = (uint32_t)(q / nx), i = (uint32_t)(q - (size_t)row * nx);
        const uint32_t j = row % ny, k = row / ny;
        if (j >= 1u && j <= (uint32_t)jmax && k >= 1u && k <= (uint32_t)kmax && i >= 1u && i <= (uint32_t)imax) {
            float c = t;
            if (i == 1u) c = fmaxf(c, (T[q - 1] + t) * 0.5f);
            if (i == (uint32_t)imax) c = fmaxf(c, (t + T[q + 1]) * 0.5f);
            m = max(m, __float_as_uint(c));
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        bad |= (uint32_t)__shfl_xor((int)bad, o);
        m = max(m, (uint32_t)__shfl_xor((int)m, o));
    }
    if ((threadIdx.x & 63) == 0) {
        if (bad) atomicOr(&out[0], 1u);
        atomicMax(&out[1], m);
    }
}

__global__ void k_zero_u32x2(uint32_t *p) {
    if (threadIdx.x < 2) p[threadIdx.x] = 0u;
}

void launch_field_scan(hipStream_t st, const float *T, int imax, int jmax, int kmax, uint32_t *out) {
    const size_t n = (size_t)(imax + 2) * (jmax + 2) * (kmax + 2);
    const int blocks = (int)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 8192));
    hipLaunchKernelGGL(k_zero_u32x2, dim3(1), dim3(64), 0, st, out);
    hipLaunchKernelGGL(k_field_scan, dim3(blocks), dim3(256), 0, st, T, imax, jmax, kmax, out);
}

void launch_fluid_f64(hipStream_t st, const DevCommon &cm, const DevSnap *snaps, int s, int v, const uint16_t *keys,
                      double *de, unsigned long long *stats) {
    const int blocks = std::max(1, std::min((cm.ncell + 127) / 128, 1 << 20));
    hipLaunchKernelGGL(k_fluid_f64, dim3(blocks), dim3(128), 0, st, cm, snaps, s, v, keys, de, stats);
}

void launch_surface_f64(hipStream_t st, const DevCommon &cm, const DevSnap *snaps, int s, int v,
                        const uint16_t *keys, double *de, unsigned long long *stats) {
    const int blocks = std::max(1, std::min((2 * cm.nsurf + 127) / 128, 1 << 20));
    hipLaunchKernelGGL(k_surface_f64, dim3(blocks), dim3(128), 0, st, cm, snaps, s, v, keys, de, stats);
}

void launch_trace(hipStream_t st, const DevCommon &cm, const DevSnap *snaps, int surface, const int32_t *rays,
                  const int32_t *unit, int n, const uint16_t *keys, rc_ray *out) {
    hipLaunchKernelGGL(k_trace, dim3((n + 63) / 64), dim3(64), 0, st, cm, snaps, surface, rays, unit, n, keys, out);
}

void launch_finalize(hipStream_t st, const double *de, int n, int nVar, float *mean, float *sdev) {
    const int blocks = std::max(1, std::min((n + 255) / 256, 65536));
    hipLaunchKernelGGL(k_finalize, dim3(blocks), dim3(256), 0, st, de, n, nVar, mean, sdev);
}

}  // namespace rc
