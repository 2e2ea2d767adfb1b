/*
 * rc_batched.hip -- the batched solver: persistent wavefronts pulling work
 * units from a device queue.
 *
 * A unit is one emitting cell (or wall cell) of one snapshot with all of its
 * nVar runs.  Per run, the wavefront draws the photmax (band, g) samples,
 * sorts them in LDS (the srt enumeration of mc_gpu_noint.cu:1234-1267), and
 * its 64 lanes march consecutive rays of the sorted list, so that lanes work
 * on neighbouring bands of similar optical thickness.  Per-ray tallies are
 * FP32 (the same arithmetic as the plain path) and are summed in FP64,
 * reduced across the wavefront with cross-lane shuffles.  Units from all
 * snapshots share one queue, so snapshots of different sizes and optical
 * thickness balance across the chip.
 */
#include <hip/hip_runtime.h>

#include <algorithm>

#include "rc_device.h"
#include "rc_internal.h"
#include "rc_kernels.h"
#include "rc_march.h"

using namespace rcd;

namespace rc {

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}

/* bitonic sort of n2 (power of two) keys in LDS by one 64-lane workgroup */
__device__ __forceinline__ void lds_sort(uint32_t *lds, int n2) {
    for (int k = 2; k <= n2; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < n2; i += 64) {
                const int ixj = i ^ j;
                if (ixj > i) {
                    const uint32_t a = lds[i], b = lds[ixj];
                    const bool up = (i & k) == 0;
                    if ((a > b) == up) { lds[i] = b; lds[ixj] = a; }
                }
            }
            __syncthreads();
        }
}

template <bool SURF>
__global__ void __launch_bounds__(64) k_batched(DevCommon cm, const DevSnap *__restrict__ snaps, int nsnap,
                                                double *__restrict__ de, unsigned long long *__restrict__ stats,
                                                int *__restrict__ queue) {
    extern __shared__ uint32_t lds[];
    __shared__ int s_unit;
    const int P = cm.photmax;
    int n2 = 1;
    while (n2 < P) n2 <<= 1;
    const int per_snap = SURF ? 2 * cm.nsurf : cm.ncell;
    const int total = per_snap * nsnap;
    int64_t steps = 0, rays = 0;
    for (;;) {
        if (threadIdx.x == 0) s_unit = atomicAdd(queue, 1);
        __syncthreads();
        const int u = s_unit;
        __syncthreads();
        if (u >= total) break;
        const int s = u / per_snap, n = u % per_snap;
        const DevSnap &sn = snaps[s];
        int i = 0, j, k, bound = 0;
        uint32_t cid, tag;
        float Ti;
        if (!SURF) {
            cell_ijk(cm, n, i, j, k);
            cid = (uint32_t)n + (uint32_t)cm.imax * cm.jmax * (cm.zstart - 1);
            tag = TAG_FLUID;
            Ti = cell_T(cm, sn, i, j, k);
        } else {
            bound = n / cm.nsurf;
            const int m = n % cm.nsurf;
            if ((bound == 0 && !cm.adiab[0]) || (bound == 1 && !cm.adiab[1])) continue;
            j = m % cm.jmax + 1;
            k = m / cm.jmax + cm.zstart;
            cid = (uint32_t)m + (uint32_t)cm.jmax * (cm.zstart - 1);
            tag = bound ? TAG_EAST : TAG_WEST;
            Ti = wall_T(cm, sn, bound, j, k);
        }
        for (int v = 0; v < cm.nVar; ++v) {
            if (cm.srt) {
                const uint32_t tagx = tag + 16u * (uint32_t)sn.snap;
                for (int h = threadIdx.x; h < n2; h += 64) {
                    uint32_t key = 0xFFFFFFFFu;
                    if (h < P) {
                        const U4 r = draw4(cm.k0, cm.k1, (uint32_t)h, (uint32_t)v, cid, tagx);
                        key = band_key(cm, sn, SURF, r.c, r.d);
                    }
                    lds[h] = key;
                }
                __syncthreads();
                lds_sort(lds, n2);
            }
            double acc = 0.0;
            int st = 0;
            for (int h = threadIdx.x; h < P; h += 64) {
                const uint32_t key = cm.srt ? lds[h] : 0u;
                float d;
                if (!SURF) d = fluid_ray<false>(cm, sn, i, j, k, cid, v, h, key, cm.srt != 0, Ti, st, nullptr);
                else d = surface_ray<false>(cm, sn, bound, j, k, cid, v, h, key, cm.srt != 0, Ti, st, nullptr);
                acc += (double)d;
                rays += 1;
            }
            steps += st;
            acc = wave_sum(acc);
            if (threadIdx.x == 0) de[((size_t)s * per_snap + n) * cm.nVar + v] = acc;
            __syncthreads();
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        steps += __shfl_xor(steps, o);
        rays += __shfl_xor(rays, o);
    }
    if (threadIdx.x == 0) {
        atomicAdd(&stats[0], (unsigned long long)rays);
        atomicAdd(&stats[1], (unsigned long long)steps);
    }
}

__global__ void k_zero_int(int *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

void launch_batched(hipStream_t st, const DevCommon &cm, const DevSnap *snaps, int nsnap, int surface, double *de,
                    unsigned long long *stats, int *queue, int wg_per_cu, int rays_per_lane) {
    (void)rays_per_lane;
    int dev = 0, ncu = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    if (ncu <= 0) ncu = 256;
    if (wg_per_cu <= 0) wg_per_cu = 16;
    int n2 = 1;
    while (n2 < cm.photmax) n2 <<= 1;
    const size_t lds = cm.srt ? (size_t)n2 * 4 : 4;
    const int per_snap = surface ? 2 * cm.nsurf : cm.ncell;
    const int total = per_snap * nsnap;
    const int grid = std::max(1, std::min(ncu * wg_per_cu, total));
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, st, queue, 1);
    if (!surface)
        hipLaunchKernelGGL(k_batched<false>, dim3(grid), dim3(64), lds, st, cm, snaps, nsnap, de, stats, queue);
    else
        hipLaunchKernelGGL(k_batched<true>, dim3(grid), dim3(64), lds, st, cm, snaps, nsnap, de, stats, queue);
}

}  // namespace rc
