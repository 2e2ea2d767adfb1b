/*
 * rc_batched.hip -- the batched solver: persistent wavefronts with ray
 * regeneration, fed by a device work queue.
 *
 * Work unit: one emitting cell (or wall cell) of one snapshot and one of its
 * nVar runs, i.e. photmax rays.  A wavefront (one 64-lane workgroup) takes
 * units from a global atomic queue.  For each unit it draws the photmax
 * (band, g) samples and bitonic-sorts them in LDS (the srt enumeration of
 * mc_gpu_noint.cu:1234-1267) once the previous unit's rays are all
 * generated, while lanes still march them.
 *
 * The march loop advances every active lane by one step per iteration
 * (march_step, the loop body of march_ray) and refills idle lanes with the
 * next rays of the current unit once enough lanes are idle, so a long ray
 * does not hold 63 idle lanes.  Rays keep their (sorted) order of dispatch,
 * so neighbouring lanes march neighbouring bands.  Each finished ray adds its
 * FP32 tally to its unit's FP64 total in LDS; the unit's last ray stores the
 * total.  Every per-ray result is bit-identical to the plain path and the
 * CPU oracle; only the summation order of the FP64 totals differs.
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

/* bitonic sort of 256 keys held in registers, element r*64 + lane in x[r]:
 * stages whose partner is in another lane exchange through a lane shuffle,
 * stages whose partner is in another register of the same lane use min/max
 * directly; no LDS traffic and no barrier */
__device__ __forceinline__ void wave_sort256(uint32_t (&x)[4], int lane) {
    /* the lane index as an opaque value, so that the per-stage lane masks
     * are formed here instead of being hoisted out of the caller's loop and
     * kept live across it */
    asm volatile("" : "+v"(lane));
#pragma unroll
    for (int k = 2; k <= 256; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            if (j >= 64) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int p = r ^ (j >> 6);
                    if (p > r) {
                        const bool up = ((r * 64) & k) == 0;
                        const uint32_t lo = min(x[r], x[p]), hi = max(x[r], x[p]);
                        x[r] = up ? lo : hi;
                        x[p] = up ? hi : lo;
                    }
                }
            } else {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const uint32_t y = (uint32_t)__shfl_xor((int)x[r], j);
                    /* all ones where this lane keeps the larger key */
                    const uint32_t big = 0u - ((uint32_t)(((r * 64 + lane) & k) != 0) ^ (uint32_t)((lane & j) != 0));
                    const uint32_t mn = min(x[r], y), mx = max(x[r], y);
                    x[r] = (mn & ~big) | (mx & big);
                }
            }
        }
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
template <typename T>
__device__ __forceinline__ T *uni(T *p) {
    const uint64_t v = reinterpret_cast<uint64_t>(p);
    const uint64_t lo = uni((uint32_t)v), hi = uni((uint32_t)(v >> 32));
    return reinterpret_cast<T *>((hi << 32) | lo);
}

/* the march's snapshot fields and level, read with scalar loads */
__device__ __forceinline__ SnapView snap_view(const DevSnap *snp) {
    SnapView v;
    v.Tnb0 = uni(snp->Tnb0);
    v.invdT = uni(snp->invdT);
    v.kbf16 = uni(snp->kbf16);
    v.kb = uni(snp->kb);
    v.kp16 = uni(snp->kp16);
    v.bp = uni(snp->bp);
    v.kscale = uni(snp->kscale);
    v.bfix = uni(snp->bfix);
    return v;
}

__device__ __forceinline__ DevLevel level_view(const DevLevel *lp) {
    DevLevel L;
    L.im = uni(lp->im); L.jm = uni(lp->jm); L.km = uni(lp->km); L.maxs = uni(lp->maxs);
    L.dy = uni(lp->dy); L.dz = uni(lp->dz);
    L.pad0 = 0; L.pad1 = 0;
    L.xu = uni(lp->xu); L.T = uni(lp->T); L.Tw = uni(lp->Tw); L.Tz = uni(lp->Tz);
    return L;
}

__device__ __forceinline__ Unit uni(const Unit &a) {
    Unit b;
    b.s = uni(a.s); b.n = uni(a.n); b.v = uni(a.v);
    b.i = uni(a.i); b.j = uni(a.j); b.k = uni(a.k); b.bound = uni(a.bound);
    b.cid = uni(a.cid); b.tag = uni(a.tag); b.Ti = uni(a.Ti); b.skip = uni(a.skip);
    return b;
}

/* LDS ray queue: rays generated in bulk (all 64 lanes) wait here until an
 * idle lane picks them up.  Structure of arrays, QCAP (power of two) slots
 * per wavefront; a queued ray holds its direction, energy constants, table
 * row and the unit slot it belongs to.  The start cell and position come
 * from the unit slot when a lane picks the ray up. */
constexpr int QCAP = 128;
enum { Q_SX, Q_SY, Q_SZ, Q_ISX, Q_ISY, Q_ISZ, Q_W, Q_BI, Q_IBI, Q_KEY, Q_SLOT, Q_WORDS };

/* Unit slots: the units a wavefront has in flight (rays still queued or
 * marching).  Each holds the FP64 tally of its rays, the number of rays not
 * yet finished, the tally index and the start cell and position of its
 * rays.  The ray that finishes last stores the tally, so every (cell, run)
 * total is written once, by the one wavefront that owns the unit, with a
 * plain store.  A slot is reused only once its count is zero. */
constexpr int NSLOT = 32;
constexpr int SLOT_WORDS = 10;     /* acc (2 words), cnt, tid, ic, jc, kc, xp, yp, zp */

struct Slots {
    double *acc;
    int *cnt, *tid, *ic, *jc, *kc;
    float *xp, *yp, *zp;
};

__device__ __forceinline__ Slots slot_view(uint32_t *lds) {
    Slots s;
    s.acc = reinterpret_cast<double *>(lds);
    s.cnt = reinterpret_cast<int *>(lds + 2 * NSLOT);
    s.tid = s.cnt + NSLOT;
    s.ic = s.tid + NSLOT;
    s.jc = s.ic + NSLOT;
    s.kc = s.jc + NSLOT;
    s.xp = reinterpret_cast<float *>(s.kc + NSLOT);
    s.yp = s.xp + NSLOT;
    s.zp = s.yp + NSLOT;
    return s;
}

__device__ __forceinline__ void q_put(uint32_t *q, int qs, float sx, float sy, float sz, const En &e, uint32_t key,
                                      int slot) {
    q[Q_SX * QCAP + qs] = __float_as_uint(sx);
    q[Q_SY * QCAP + qs] = __float_as_uint(sy);
    q[Q_SZ * QCAP + qs] = __float_as_uint(sz);
    q[Q_ISX * QCAP + qs] = __float_as_uint(1.0f / sx);
    q[Q_ISY * QCAP + qs] = __float_as_uint(1.0f / sy);
    q[Q_ISZ * QCAP + qs] = __float_as_uint(1.0f / sz);
    q[Q_W * QCAP + qs] = __float_as_uint(e.W);
    q[Q_BI * QCAP + qs] = __float_as_uint(e.Bi);
    q[Q_IBI * QCAP + qs] = __float_as_uint(e.invBi);
    q[Q_KEY * QCAP + qs] = key;
    q[Q_SLOT * QCAP + qs] = (uint32_t)slot;
}

template <bool SURF>
__device__ __forceinline__ void q_get(const uint32_t *q, int qs, const Slots &sl, const DevCommon &cm,
                                      const DevSnap &sn, Ray &r, En &e, int &slot) {
    r.sx = __uint_as_float(q[Q_SX * QCAP + qs]);
    r.sy = __uint_as_float(q[Q_SY * QCAP + qs]);
    r.sz = __uint_as_float(q[Q_SZ * QCAP + qs]);
    r.isx = __uint_as_float(q[Q_ISX * QCAP + qs]);
    r.isy = __uint_as_float(q[Q_ISY * QCAP + qs]);
    r.isz = __uint_as_float(q[Q_ISZ * QCAP + qs]);
    r.f0 = r.sx < 0.f; r.f1 = r.sy < 0.f; r.f2 = r.sz < 0.f;
    r.tra = 1.0f;
    e.W = __uint_as_float(q[Q_W * QCAP + qs]);
    e.Bi = __uint_as_float(q[Q_BI * QCAP + qs]);
    e.invBi = __uint_as_float(q[Q_IBI * QCAP + qs]);
    const uint32_t key = q[Q_KEY * QCAP + qs];
    e.row = key;
    e.rank = (int)(key / (uint32_t)cm.nQ);
    slot = (int)q[Q_SLOT * QCAP + qs];
    r.ic = sl.ic[slot]; r.jc = sl.jc[slot]; r.kc = sl.kc[slot];
    r.xp = sl.xp[slot]; r.yp = sl.yp[slot]; r.zp = sl.zp[slot];
    if (SURF) surface_adjust(sn, r);
}

/* LDS words of one wavefront: unit slots, ray queue, (multigrid) the grid
 * levels of the current snapshot, then (srt) the unit's sorted keys and the
 * direction uniforms of its rays */
constexpr int LEV_WORDS = (int)(sizeof(DevLevel) / 4);
__host__ __device__ __forceinline__ int batched_lds_words(int n2, int srt, int multi) {
    return SLOT_WORDS * NSLOT + Q_WORDS * QCAP + (multi ? LEV_WORDS * RC_MAXLEV : 0) + (srt ? 3 * n2 : 0);
}

template <bool SURF, bool MULTI, int TM>
__global__ void __launch_bounds__(64) k_batched(DevCommon cm, const DevSnap *__restrict__ snaps, int nsnap,
                                                double *__restrict__ de, unsigned long long *__restrict__ stats,
                                                int *__restrict__ queue, int threshold) {
    extern __shared__ uint32_t lds[];
    const Slots sl = slot_view(lds);
    uint32_t *rq = lds + SLOT_WORDS * NSLOT;
    DevLevel *llev = reinterpret_cast<DevLevel *>(rq + Q_WORDS * QCAP);   /* 8-byte aligned */
    const int P = cm.photmax;
    int n2 = 1;
    while (n2 < P) n2 <<= 1;
    uint32_t *keys = rq + Q_WORDS * QCAP + (MULTI ? LEV_WORDS * RC_MAXLEV : 0);
    float *dira = reinterpret_cast<float *>(keys + n2);
    float *dirb = dira + n2;
    const int lane = threadIdx.x;
    const int per_snap = SURF ? 2 * cm.nsurf : cm.ncell;
    const int total = per_snap * nsnap * cm.nVar;
    const float tol = SURF ? cm.tollS : cm.toll;
    if (lane < NSLOT) sl.cnt[lane] = 0;
    __syncthreads();

    /* wavefront-uniform state: dispatch of units, unit slots, ray queue */
    int cur_h = P, nxt = -1, nxt_s = -1, qdone = 0, snap = -1, seq = 0, cslot = 0;
    int qhead = 0, qcount = 0;
    Unit U{};

    /* lane state */
    bool active = false;
    Ray r;
    En e;
    float dacc = 0.f;
    int rslot = 0, lev = 0, counter = 0;
    int nsteps = 0, nrays = 0;

    /* one back-edge: every iteration either starts a unit or generates,
     * refills and steps */
    for (;;) {
        /* (a) once the current unit is fully generated, take the next unit
         * and (srt) draw and sort its band keys, keeping the direction
         * uniforms of every ray */
        if (nxt < 0 && !qdone && cur_h >= P) {
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
                    const uint32_t tagx = N.tag + 16u * (uint32_t)sn.snap;
                    if (n2 <= 256) {
                        /* up to 256 keys: sorted in registers */
                        uint32_t x[4];
#pragma unroll
                        for (int q4 = 0; q4 < 4; ++q4) {
                            const int h = q4 * 64 + lane;
                            uint32_t key = 0xFFFFFFFFu;
                            if (h < P) {
                                const U4 q = draw4(cm.k0, cm.k1, (uint32_t)h, (uint32_t)N.v, N.cid, tagx);
                                key = band_key(cm, sn, SURF, q.c, q.d);
                                dira[h] = q.a;
                                dirb[h] = q.b;
                            }
                            x[q4] = key;
                        }
                        wave_sort256(x, lane);
#pragma unroll
                        for (int q4 = 0; q4 < 4; ++q4)
                            if (q4 * 64 + lane < n2) keys[q4 * 64 + lane] = x[q4];
                        __syncthreads();
                    } else {
                        for (int h = lane; h < n2; h += 64) {
                            uint32_t key = 0xFFFFFFFFu;
                            if (h < P) {
                                const U4 q = draw4(cm.k0, cm.k1, (uint32_t)h, (uint32_t)N.v, N.cid, tagx);
                                key = band_key(cm, sn, SURF, q.c, q.d);
                                dira[h] = q.a;
                                dirb[h] = q.b;
                            }
                            keys[h] = key;
                        }
                        __syncthreads();
                        lds_sort(keys, n2);
                    }
                }
            }
        }
        const unsigned long long live = __ballot(active);
        if (cur_h >= P && nxt >= 0 && (nxt_s == snap || (live == 0ull && qcount == 0)) &&
            uni(sl.cnt[seq & (NSLOT - 1)]) == 0) {
            /* (b) start the prepared unit once its slot is free (a new
             * snapshot only once every lane and the queue are empty) */
            U = uni(decode<SURF>(cm, snaps, nxt, per_snap));
            nxt = -1;
            if (MULTI && U.s != snap) {
                const uint32_t *src = reinterpret_cast<const uint32_t *>(snaps[U.s].lev);
                for (int w = lane; w < LEV_WORDS * cm.grid_num; w += 64)
                    reinterpret_cast<uint32_t *>(llev)[w] = gld(src, w);
            }
            snap = U.s;
            const int tidx = (U.s * per_snap + U.n) * cm.nVar + U.v;
            if (U.skip) {
                if (lane == 0) de[tidx] = 0.0;
            } else {
                cslot = seq & (NSLOT - 1);
                seq += 1;
                cur_h = 0;
                if (lane == 0) {
                    const DevSnap &sn = snaps[U.s];
                    sl.acc[cslot] = 0.0;
                    sl.cnt[cslot] = P;
                    sl.tid[cslot] = tidx;
                    sl.ic[cslot] = SURF ? (U.bound ? cm.imax : 1) : U.i;
                    sl.jc[cslot] = U.j;
                    sl.kc[cslot] = U.k;
                    sl.xp[cslot] = SURF ? (U.bound ? cm.Lx : 0.f) : gld(sn.x0, U.i);
                    sl.yp[cslot] = gld(sn.y0, U.j);
                    sl.zp[cslot] = gld(sn.z0, U.k);
                }
            }
            __syncthreads();
        } else {
            /* (c) bulk generation: every lane sets up one ray of the current unit */
            const int left = P - cur_h;
            if (left > 0 && qcount <= QCAP - 64) {
                const int ng = min(64, left);
                if (lane < ng) {
                    const DevSnap &sn = snaps[U.s];
                    const int h = cur_h + lane;
                    float a, b;
                    uint32_t key;
                    if (cm.srt) {
                        key = keys[h];
                        a = dira[h];
                        b = dirb[h];
                    } else {
                        const U4 q = draw4(cm.k0, cm.k1, (uint32_t)h, (uint32_t)U.v, U.cid, U.tag + 16u * (uint32_t)sn.snap);
                        key = band_key(cm, sn, SURF, q.c, q.d);
                        a = q.a;
                        b = q.b;
                    }
                    float sx, sy, sz;
                    if (SURF) surface_dir(a, b, U.bound, sx, sy, sz);
                    else fluid_dir(a, b, sx, sy, sz);
                    En ge;
                    energy<TM>(cm, sn, key, U.Ti, SURF, U.bound, ge);
                    q_put(rq, (qhead + qcount + lane) & (QCAP - 1), sx, sy, sz, ge, key, cslot);
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
                    q_get<SURF>(rq, (qhead + rank) & (QCAP - 1), sl, cm, snaps[snap], r, e, rslot);
                    active = true;
                    dacc = 0.f;
                    lev = 0;
                    counter = 0;
                }
                qhead = (qhead + npop) & (QCAP - 1);
                qcount -= npop;
            }
            if (__ballot(active) == 0ull) {
                if (qdone && nxt < 0 && cur_h >= P && qcount == 0) break;
            } else {
                /* (e) one march step of every active lane; the snapshot's
                 * fields (and the level, single grid) are wavefront-uniform
                 * values, a lane's level comes from LDS (multigrid) */
                const SnapView sv = snap_view(snaps + snap);
                DevLevel l0;
                if (!MULTI) l0 = level_view(&snaps[snap].lev[0]);
                const DevLevel L = MULTI ? llev[lev] : l0;
                const bool alive = r.tra > tol;
                const bool go = active && alive && counter < L.maxs;
                const bool fin = active && !go && (!MULTI || !alive || lev == cm.grid_num - 1);
                const bool crs = MULTI && active && !go && !fin;
                if (go) {
                    int status;
                    march_step<TM>(cm, sv, L, r, e, tol, dacc, status);
                    counter += 1;
                    nsteps += 1;
                }
                if (fin) {
                    /* the ray ends: add its tally to its unit; the last ray
                     * of the unit stores the unit's total */
                    atomicAdd(&sl.acc[rslot], (double)dacc);
                    if (atomicSub(&sl.cnt[rslot], 1) == 1) de[sl.tid[rslot]] = sl.acc[rslot];
                    nrays += 1;
                    active = false;
                }
                if (crs) {
                    /* coarsen the indices for the next grid level (:215-223) */
                    const DevLevel Nl = llev[lev + 1];
                    const int ri = L.im / Nl.im, rj = L.jm / Nl.jm, rk = L.km / Nl.km;
                    r.ic = (r.ic + ri - 1) / ri;
                    r.jc = (r.jc + rj - 1) / rj;
                    r.kc = (r.kc + rk - 1) / rk;
                    lev += 1;
                    counter = 0;
                }
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

void launch_batched(hipStream_t st, const DevCommon &cm, const DevSnap *snaps, int nsnap, int surface, int snaps_bf16,
                    double *de, unsigned long long *stats, int *queue, int wg_per_cu, int threshold) {
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
    const size_t lds = (size_t)batched_lds_words(n2, cm.srt, cm.grid_num > 1) * 4;
    const int per_snap = surface ? 2 * cm.nsurf : cm.ncell;
    const size_t nde = (size_t)per_snap * nsnap * cm.nVar;
    const long total = (long)nde;
    const int grid = (int)std::max<long>(1, std::min<long>((long)ncu * wg_per_cu, total));
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, st, queue, 1);
    const bool multi = cm.grid_num > 1;
    const bool bf = snaps_bf16;
#define RC_LAUNCH(S, M, T) hipLaunchKernelGGL((k_batched<S, M, T>), dim3(grid), dim3(64), lds, st, cm, snaps, nsnap, de, \
                                           stats, queue, threshold)
    if (!surface) {
        if (multi) { if (bf) RC_LAUNCH(false, true, TAB_BF16); else RC_LAUNCH(false, true, TAB_F32); }
        else { if (bf) RC_LAUNCH(false, false, TAB_BF16); else RC_LAUNCH(false, false, TAB_F32); }
    } else {
        if (multi) { if (bf) RC_LAUNCH(true, true, TAB_BF16); else RC_LAUNCH(true, true, TAB_F32); }
        else { if (bf) RC_LAUNCH(true, false, TAB_BF16); else RC_LAUNCH(true, false, TAB_F32); }
    }
#undef RC_LAUNCH
}

}  // namespace rc
