/*
 * rc_batched.hip -- the batched solver: persistent wavefronts with ray
 * regeneration, fed by a device work queue.
 *
 * Work unit: one emitting cell (or wall cell) of one snapshot and one of its
 * nVar runs, i.e. photmax rays.  A wavefront (one 64-lane workgroup) takes
 * units from a global atomic queue.  For each unit it draws the photmax
 * (band, g) samples and bitonic-sorts them in LDS (the srt enumeration of
 * mc_gpu_noint.cu:1234-1267), double-buffered: the next unit's list is
 * prepared while lanes still march rays of the current one.
 *
 * The march loop advances every active lane by one step per iteration
 * (march_step, the loop body of march_ray) and refills idle lanes with the
 * next rays of the current unit once enough lanes are idle, so a long ray
 * does not hold 63 idle lanes.  Rays keep their (sorted) order of dispatch,
 * so neighbouring lanes march neighbouring bands.  Each finished ray adds its
 * FP32 tally to the (cell, run) FP64 total with a device-scope atomic add.
 * Every per-ray result is bit-identical to the plain path and the CPU
 * oracle; only the summation order of the FP64 totals differs.
 *
 * Units of all snapshots share the queue; a wavefront drains its lanes
 * before crossing to another snapshot.
 */
#include <hip/hip_runtime.h>

#include <algorithm>

#include "rc_device.h"
#include "rc_internal.h"
#include "rc_kernels.h"
#include "rc_march.h"

using namespace rcd;

namespace rc {

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

/* per-unit constants, identical in every lane */
struct Unit {
    int s, n, v;          /* snapshot, cell (or wall cell) index, run     */
    int i, j, k, bound;   /* emission cell                               */
    uint32_t cid, tag;    /* Philox cell id and stream tag               */
    float Ti;             /* emitting temperature                        */
    int skip;             /* wall unit of a non-adiabatic wall           */
};

template <bool SURF>
__device__ __forceinline__ Unit decode(const DevCommon &cm, const DevSnap *__restrict__ snaps, int u, int per_snap) {
    Unit U;
    U.v = u % cm.nVar;
    const int cs = u / cm.nVar;
    U.s = cs / per_snap;
    U.n = cs % per_snap;
    const DevSnap &sn = snaps[U.s];
    U.skip = 0;
    if (!SURF) {
        cell_ijk(cm, U.n, U.i, U.j, U.k);
        U.bound = 0;
        U.cid = (uint32_t)U.n + (uint32_t)cm.imax * cm.jmax * (cm.zstart - 1);
        U.tag = TAG_FLUID;
        U.Ti = cell_T(cm, sn, U.i, U.j, U.k);
    } else {
        U.bound = U.n / cm.nsurf;
        const int m = U.n % cm.nsurf;
        U.i = 0;
        U.j = m % cm.jmax + 1;
        U.k = m / cm.jmax + cm.zstart;
        U.cid = (uint32_t)m + (uint32_t)cm.jmax * (cm.zstart - 1);
        U.tag = U.bound ? TAG_EAST : TAG_WEST;
        U.skip = (U.bound == 0 && !cm.adiab[0]) || (U.bound == 1 && !cm.adiab[1]);
        U.Ti = U.skip ? 0.f : wall_T(cm, sn, U.bound, U.j, U.k);
    }
    return U;
}

/* values identical in every lane, moved to scalar registers */
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ uint32_t uni(uint32_t x) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)x); }
__device__ __forceinline__ float uni(float x) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x))); }
__device__ __forceinline__ Unit uni(const Unit &a) {
    Unit b;
    b.s = uni(a.s); b.n = uni(a.n); b.v = uni(a.v);
    b.i = uni(a.i); b.j = uni(a.j); b.k = uni(a.k); b.bound = uni(a.bound);
    b.cid = uni(a.cid); b.tag = uni(a.tag); b.Ti = uni(a.Ti); b.skip = uni(a.skip);
    return b;
}

/* LDS ray queue: rays generated in bulk (all 64 lanes) wait here until an
 * idle lane picks them up; structure of arrays, QCAP slots per wavefront */
constexpr int QCAP = 96;
enum { Q_IC, Q_JC, Q_KC, Q_XP, Q_YP, Q_ZP, Q_SX, Q_SY, Q_SZ, Q_ISX, Q_ISY, Q_ISZ, Q_W, Q_BI, Q_IBI, Q_KEY, Q_TIDX,
       Q_WORDS };

__device__ __forceinline__ void q_put(uint32_t *q, int slot, const Ray &r, const En &e, uint32_t key, int tidx) {
    q[Q_IC * QCAP + slot] = (uint32_t)r.ic;
    q[Q_JC * QCAP + slot] = (uint32_t)r.jc;
    q[Q_KC * QCAP + slot] = (uint32_t)r.kc;
    q[Q_XP * QCAP + slot] = __float_as_uint(r.xp);
    q[Q_YP * QCAP + slot] = __float_as_uint(r.yp);
    q[Q_ZP * QCAP + slot] = __float_as_uint(r.zp);
    q[Q_SX * QCAP + slot] = __float_as_uint(r.sx);
    q[Q_SY * QCAP + slot] = __float_as_uint(r.sy);
    q[Q_SZ * QCAP + slot] = __float_as_uint(r.sz);
    q[Q_ISX * QCAP + slot] = __float_as_uint(r.isx);
    q[Q_ISY * QCAP + slot] = __float_as_uint(r.isy);
    q[Q_ISZ * QCAP + slot] = __float_as_uint(r.isz);
    q[Q_W * QCAP + slot] = __float_as_uint(e.W);
    q[Q_BI * QCAP + slot] = __float_as_uint(e.Bi);
    q[Q_IBI * QCAP + slot] = __float_as_uint(e.invBi);
    q[Q_KEY * QCAP + slot] = key;
    q[Q_TIDX * QCAP + slot] = (uint32_t)tidx;
}

__device__ __forceinline__ void q_get(const uint32_t *q, int slot, const DevCommon &cm, const DevSnap &sn, Ray &r,
                                      En &e, int &tidx) {
    r.ic = (int)q[Q_IC * QCAP + slot];
    r.jc = (int)q[Q_JC * QCAP + slot];
    r.kc = (int)q[Q_KC * QCAP + slot];
    r.xp = __uint_as_float(q[Q_XP * QCAP + slot]);
    r.yp = __uint_as_float(q[Q_YP * QCAP + slot]);
    r.zp = __uint_as_float(q[Q_ZP * QCAP + slot]);
    r.sx = __uint_as_float(q[Q_SX * QCAP + slot]);
    r.sy = __uint_as_float(q[Q_SY * QCAP + slot]);
    r.sz = __uint_as_float(q[Q_SZ * QCAP + slot]);
    r.isx = __uint_as_float(q[Q_ISX * QCAP + slot]);
    r.isy = __uint_as_float(q[Q_ISY * QCAP + slot]);
    r.isz = __uint_as_float(q[Q_ISZ * QCAP + slot]);
    r.f0 = r.sx < 0.f; r.f1 = r.sy < 0.f; r.f2 = r.sz < 0.f;
    r.tra = 1.0f;
    e.W = __uint_as_float(q[Q_W * QCAP + slot]);
    e.Bi = __uint_as_float(q[Q_BI * QCAP + slot]);
    e.invBi = __uint_as_float(q[Q_IBI * QCAP + slot]);
    const uint32_t key = q[Q_KEY * QCAP + slot];
    e.row = key;
    e.rank = (int)(key / (uint32_t)cm.nQ);
    tidx = (int)q[Q_TIDX * QCAP + slot];
}

template <bool SURF, bool MULTI>
__global__ void __launch_bounds__(64) k_batched(DevCommon cm, const DevSnap *__restrict__ snaps, int nsnap,
                                                double *__restrict__ de, unsigned long long *__restrict__ stats,
                                                int *__restrict__ queue, int threshold) {
    extern __shared__ uint32_t lds[];             /* ray queue, then 2 x n2 sorted keys */
    uint32_t *rq = lds;
    uint32_t *keybuf = lds + Q_WORDS * QCAP;
    const int P = cm.photmax;
    int n2 = 1;
    while (n2 < P) n2 <<= 1;
    const int lane = threadIdx.x;
    const int per_snap = SURF ? 2 * cm.nsurf : cm.ncell;
    const int total = per_snap * nsnap * cm.nVar;
    const float tol = SURF ? cm.tollS : cm.toll;

    /* wavefront-uniform state: dispatch of units and the ray queue */
    int cur_h = P, nxt = -1, nxt_s = -1, qdone = 0, bcur = 0, snap = -1;
    int qhead = 0, qcount = 0;
    Unit U{};

    /* lane state */
    bool active = false;
    Ray r;
    En e;
    float dacc = 0.f;
    int tidx = 0, lev = 0, counter = 0;
    int nsteps = 0, nrays = 0;

    for (;;) {
        /* (a) prepare the next unit's sorted band list */
        if (nxt < 0 && !qdone) {
            int u = 0;
            if (lane == 0) u = atomicAdd(queue, 1);
            u = uni(u);
            if (u >= total) {
                qdone = 1;
            } else {
                nxt = u;
                const Unit N = uni(decode<SURF>(cm, snaps, u, per_snap));
                nxt_s = N.s;
                if (cm.srt && !N.skip) {
                    const DevSnap &sn = snaps[N.s];
                    uint32_t *kb = keybuf + (bcur ^ 1) * n2;
                    const uint32_t tagx = N.tag + 16u * (uint32_t)sn.snap;
                    for (int h = lane; h < n2; h += 64) {
                        uint32_t key = 0xFFFFFFFFu;
                        if (h < P) {
                            const U4 q = draw4(cm.k0, cm.k1, (uint32_t)h, (uint32_t)N.v, N.cid, tagx);
                            key = band_key(cm, sn, SURF, q.c, q.d);
                        }
                        kb[h] = key;
                    }
                    __syncthreads();
                    lds_sort(kb, n2);
                }
            }
        }
        const unsigned long long live = __ballot(active);
        /* (b) move to the prepared unit when the current one is dispatched
         * (a new snapshot only once every lane and the queue are empty) */
        if (cur_h >= P && nxt >= 0 && (nxt_s == snap || (live == 0ull && qcount == 0))) {
            U = uni(decode<SURF>(cm, snaps, nxt, per_snap));
            nxt = -1; cur_h = U.skip ? P : 0; bcur ^= 1; snap = U.s;
            continue;
        }
        /* (c) bulk generation: every lane sets up one ray of the current unit */
        const int left = P - cur_h;
        if (left > 0 && qcount <= QCAP - 64) {
            const int ng = min(64, left);
            if (lane < ng) {
                const DevSnap &sn = snaps[U.s];
                const int h = cur_h + lane;
                const uint32_t kin = cm.srt ? keybuf[bcur * n2 + h] : 0u;
                Ray g;
                En ge;
                const uint32_t key = !SURF ? fluid_setup(cm, sn, U.i, U.j, U.k, U.cid, U.v, h, kin, cm.srt != 0, U.Ti, g, ge)
                                           : surface_setup(cm, sn, U.bound, U.j, U.k, U.cid, U.v, h, kin, cm.srt != 0, U.Ti, g, ge);
                q_put(rq, (qhead + qcount + lane) % QCAP, g, ge, key, (U.s * per_snap + U.n) * cm.nVar + U.v);
            }
            __syncthreads();
            qcount += ng;
            cur_h += ng;
        }
        /* (d) idle lanes take queued rays */
        const unsigned long long idle = ~live;
        const int nidle = __popcll(idle);
        if (qcount > 0 && nidle >= min(threshold, qcount)) {
            const int rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
            const int npop = min(nidle, qcount);
            if (!active && rank < npop) {
                q_get(rq, (qhead + rank) % QCAP, cm, snaps[snap], r, e, tidx);
                active = true;
                dacc = 0.f;
                lev = 0;
                counter = 0;
            }
            qhead = (qhead + npop) % QCAP;
            qcount -= npop;
        }
        if (__ballot(active) == 0ull) {
            if (qdone && nxt < 0 && cur_h >= P && qcount == 0) break;
            continue;
        }
        /* (e) one march step of every active lane */
        if (active) {
            const DevSnap &sn = snaps[snap];
            const DevLevel &L = MULTI ? sn.lev[lev] : sn.lev[0];
            if (r.tra > tol && counter < L.maxs) {
                int status;
                march_step(cm, sn, L, r, e, tol, dacc, status);
                counter += 1;
                nsteps += 1;
            } else if (!MULTI || !(r.tra > tol) || lev == cm.grid_num - 1) {
                atomicAdd(&de[tidx], (double)dacc);
                nrays += 1;
                active = false;
            } else {
                /* coarsen the indices for the next grid level (:215-223) */
                const DevLevel &Nl = sn.lev[lev + 1];
                const int ri = L.im / Nl.im, rj = L.jm / Nl.jm, rk = L.km / Nl.km;
                r.ic = (r.ic + ri - 1) / ri;
                r.jc = (r.jc + rj - 1) / rj;
                r.kc = (r.kc + rk - 1) / rk;
                lev += 1;
                counter = 0;
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        nsteps += __shfl_xor(nsteps, o);
        nrays += __shfl_xor(nrays, o);
    }
    if (lane == 0) {
        atomicAdd(&stats[0], (unsigned long long)(uint32_t)nrays);
        atomicAdd(&stats[1], (unsigned long long)(uint32_t)nsteps);
    }
}

__global__ void k_zero_int(int *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void k_zero_f64(double *p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0.0;
}

__global__ void k_zero_u64(unsigned long long *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

void launch_zero_u64(hipStream_t st, unsigned long long *p, int n) {
    hipLaunchKernelGGL(k_zero_u64, dim3((n + 63) / 64), dim3(64), 0, st, p, n);
}

void launch_zero(hipStream_t st, double *p, size_t n) {
    const int blocks = (int)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 8192));
    hipLaunchKernelGGL(k_zero_f64, dim3(blocks), dim3(256), 0, st, p, n);
}

void launch_batched(hipStream_t st, const DevCommon &cm, const DevSnap *snaps, int nsnap, int surface, double *de,
                    unsigned long long *stats, int *queue, int wg_per_cu, int threshold) {
    static int ncu = 0;       /* compute units of the (single) device */
    if (ncu <= 0) {
        int dev = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
        if (ncu <= 0) ncu = 256;
    }
    if (wg_per_cu <= 0) wg_per_cu = 32;
    if (threshold <= 0) threshold = 8;
    int n2 = 1;
    while (n2 < cm.photmax) n2 <<= 1;
    const size_t lds = (size_t)Q_WORDS * QCAP * 4 + (cm.srt ? (size_t)2 * n2 * 4 : 0);
    const int per_snap = surface ? 2 * cm.nsurf : cm.ncell;
    const size_t nde = (size_t)per_snap * nsnap * cm.nVar;
    const long total = (long)nde;
    const int grid = (int)std::max<long>(1, std::min<long>((long)ncu * wg_per_cu, total));
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, st, queue, 1);
    launch_zero(st, de, nde);
    const bool multi = cm.grid_num > 1;
    if (!surface) {
        if (multi) hipLaunchKernelGGL((k_batched<false, true>), dim3(grid), dim3(64), lds, st, cm, snaps, nsnap, de, stats, queue, threshold);
        else hipLaunchKernelGGL((k_batched<false, false>), dim3(grid), dim3(64), lds, st, cm, snaps, nsnap, de, stats, queue, threshold);
    } else {
        if (multi) hipLaunchKernelGGL((k_batched<true, true>), dim3(grid), dim3(64), lds, st, cm, snaps, nsnap, de, stats, queue, threshold);
        else hipLaunchKernelGGL((k_batched<true, false>), dim3(grid), dim3(64), lds, st, cm, snaps, nsnap, de, stats, queue, threshold);
    }
}

}  // namespace rc
