/*
 * rc_batched.hip -- the batched solver: persistent wavefronts with ray
 * regeneration, fed by a device work queue.
 *
 * Work unit: one emitting cell (or wall cell) of one snapshot and one of its
 * nVar runs, i.e. photmax rays.  A wavefront (one 64-lane workgroup) takes
 * units in chunks of nVar from per-XCD ranges of a global atomic queue.  For
 * each unit it draws the photmax (band, g) samples and sorts them (the srt
 * enumeration of mc_gpu_noint.cu:1234-1267): in registers up to 256 keys,
 * above that by two 8-bit radix passes through the wavefront's own HBM
 * scratch; the previous unit's rays keep marching from the LDS ray queue
 * meanwhile.
 *
 * The main loop generates rays in bulk (all 64 lanes set up one ray each),
 * refills idle lanes from the queue once enough of them are idle (at the
 * top of the iteration, and in refill mode 1 also between its steps), and
 * advances every lane by a few steps (march_core_bulk, the loop body of
 * march_ray; a ray that meets a black wall takes the wall's emission on
 * its next step, through that step's own table lookup).  Rays keep their
 * (sorted) order of dispatch, so neighbouring lanes march neighbouring
 * bands.  Each finished ray adds its FP32 tally to
 * its unit's FP64 total in an LDS slot; a slot's total goes to HBM when the
 * slot is reused.  When no unit is left the wavefront drains its last rays
 * (look-ahead march, then a whole-wavefront march of the longest ones).
 * Every per-ray result is bit-identical to the plain path and the CPU
 * oracle; only the summation order of the FP64 totals differs.
 *
 * Units of all snapshots (of one session, or of a group of sessions with
 * different grids) share the queue; a wavefront drains its lanes before
 * crossing to another snapshot.
 */
#include <hip/hip_runtime.h>

#include <algorithm>

#include "rc_device.h"
#include "rc_internal.h"
#include "rc_kernels.h"
#include "rc_march.h"
#include "rc_radix.h"

using namespace rcd;

namespace rc {

/* bitonic sort of 64 R keys held in registers, element R lane + r in x[r]:
 * compare-exchanges at distances below R stay inside a lane; the stages at
 * distance R..32 R exchange with lane (lane ^ distance/R) (21 of them).  No
 * LDS traffic and no barrier. */
template <int R>
__device__ __forceinline__ void wave_sort(uint32_t (&x)[R], int lane) {
    /* the lane index as an opaque value, so that the per-stage lane masks
     * are formed here instead of being hoisted out of the caller's loop and
     * kept live across it */
    asm volatile("" : "+v"(lane));
#pragma unroll
    for (int k = 2; k <= 64 * R; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            if (j < R) {
                const bool upl = ((lane * R) & k) == 0;
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int p = r ^ j;
                    if (p > r) {
                        const bool up = (k < R) ? ((r & k) == 0) : upl;
                        const uint32_t lo = min(x[r], x[p]), hi = max(x[r], x[p]);
                        x[r] = up ? lo : hi;
                        x[p] = up ? hi : lo;
                    }
                }
            } else {
                const bool keep_min = (((lane * R) & j) == 0) == (((lane * R) & k) == 0);
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const uint32_t y = lane_xor(x[r], j / R);
                    x[r] = keep_min ? min(x[r], y) : max(x[r], y);
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
    uint32_t sh0, sh1;    /* Sobol sampler: the (cell, run)'s shift words */
};

/* unit u of the launch: (cell, run) of the snapshot whose range of
 * emitting (wall) cells holds cell u / nVar; every grid size from the
 * snapshot */
template <bool SURF>
__device__ __forceinline__ Unit decode(const DevCommon &cm, const DevSnap *__restrict__ snaps, int nsnap, int u) {
    Unit U;
    U.v = u % cm.nVar;
    const int cs = u / cm.nVar;
    int s = 0;
    while (s + 1 < nsnap && cs >= (SURF ? snaps[s + 1].sbase : snaps[s + 1].ubase)) ++s;
    const DevSnap &sn = snaps[s];
    U.s = s;
    U.n = cs - (SURF ? sn.sbase : sn.ubase);
    U.skip = 0;
    U.sh0 = U.sh1 = 0u;
    if (!SURF) {
        cell_ijk(sn, U.n, U.i, U.j, U.k);
        U.bound = 0;
        U.cid = (uint32_t)U.n + (uint32_t)sn.imax * sn.jmax * (sn.zstart - 1);
        U.tag = TAG_FLUID;
        U.Ti = cell_T(sn, sn, U.i, U.j, U.k);
        if (cm.sampler) {
            const uint4 sh = draw4_bits(cm.k0, cm.k1, SOBOL_SHIFT_H, (uint32_t)U.v, U.cid, TAG_FLUID + 16u * (uint32_t)sn.snap);
            U.sh0 = sh.x;
            U.sh1 = sh.y;
        }
    } else {
        U.bound = U.n / sn.nsurf;
        const int m = U.n % sn.nsurf;
        U.i = 0;
        U.j = m % sn.jmax + 1;
        U.k = m / sn.jmax + sn.zstart;
        U.cid = (uint32_t)m + (uint32_t)sn.jmax * (sn.zstart - 1);
        U.tag = U.bound ? TAG_EAST : TAG_WEST;
        U.skip = (U.bound == 0 && !cm.adiab[0]) || (U.bound == 1 && !cm.adiab[1]);
        U.Ti = U.skip ? 0.f : wall_T(sn, sn, U.bound, U.j, U.k);
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
    v.xu_all = uni(snp->xu_all);
    v.P_all = uni(snp->P_all);
    v.nxu = uni(snp->nxu);
    v.npos = uni(snp->npos);
    return v;
}

__device__ __forceinline__ DevLevel level_view(const DevLevel *lp) {
    DevLevel L;
    L.im = uni(lp->im); L.jm = uni(lp->jm); L.km = uni(lp->km); L.maxs = uni(lp->maxs);
    L.dy = uni(lp->dy); L.dz = uni(lp->dz);
    L.ncell = uni(lp->ncell); L.xoff = uni(lp->xoff); L.poff = uni(lp->poff); L.pad1 = 0;
    L.xu = uni(lp->xu); L.P = uni(lp->P);
    for (int d = 0; d < 3; ++d) { L.cr[d] = uni(lp->cr[d]); L.cmul[d] = uni(lp->cmul[d]); }
    return L;
}

__device__ __forceinline__ Unit uni(const Unit &a) {
    Unit b;
    b.s = uni(a.s); b.n = uni(a.n); b.v = uni(a.v);
    b.i = uni(a.i); b.j = uni(a.j); b.k = uni(a.k); b.bound = uni(a.bound);
    b.cid = uni(a.cid); b.tag = uni(a.tag); b.Ti = uni(a.Ti); b.skip = uni(a.skip);
    b.sh0 = uni(a.sh0); b.sh1 = uni(a.sh1);
    return b;
}

/* device unit queue: counters of the QPARTS ranges (fluid at queue[0..],
 * walls at queue[16..]) */
constexpr int QPARTS = 8;

/* LDS ray queue: rays generated in bulk (all 64 lanes) wait here until an
 * idle lane picks them up.  QCAP records of 12 words per
 * wavefront, written and read as three 16-byte words: direction,
 * reciprocals, W, Bi, 1/Bi, table row and the unit slot the ray belongs to.
 * The start cell and position come from the unit slot at pick-up. */
constexpr int QCAP = 96;

/* ring index: i < 2 QCAP */
__device__ __forceinline__ int qwrap(int i) { return i >= QCAP ? i - QCAP : i; }
constexpr int QREC = 12;

/* Unit slots: the units a wavefront has in flight.  The wavefront numbers
 * the units it starts (0, 1, 2, ...); unit n lives in slot n mod NSLOT
 * until unit n + NSLOT starts.  A tally record {FP64 tally, rays finished,
 * tally index} and a placement record {start cell, start position} per
 * slot.  A finishing ray adds its tally and its count to its unit's slot
 * with LDS atomics that return nothing; whether the slot is still its
 * unit's follows from the unit numbers alone.  When a slot is taken over
 * (and for the last units, at the end) the previous unit's total goes to
 * HBM: a plain store when all its rays have finished -- the one wavefront
 * that owns a unit writes its total once -- or else (a long ray still in
 * flight) an FP64 atomic add of the partial total, after which the unit's
 * remaining rays add theirs to HBM the same way. */
constexpr int NSLOT = 16;
constexpr int SLOT_WORDS = 12;     /* tally record 4 words + placement record 8 words */

struct SlotTally {
    double acc;
    int done, tid;
};
struct SlotPlace {
    int ic, jc, kc;
    float xp, yp, zp;
    int pad0, pad1;
};
struct Slots {
    SlotTally *tal;
    SlotPlace *pl;
};

__device__ __forceinline__ Slots slot_view(uint32_t *lds) {
    Slots s;
    s.tal = reinterpret_cast<SlotTally *>(lds);
    s.pl = reinterpret_cast<SlotPlace *>(lds + 4 * NSLOT);
    return s;
}

__device__ __forceinline__ void q_put(uint32_t *q, int qs, float sx, float sy, float sz, const En &e, uint32_t key,
                                      int unit, int tid, int h) {
    uint4 *rec = reinterpret_cast<uint4 *>(q + qs * QREC);
    rec[0] = make_uint4(__float_as_uint(sx), __float_as_uint(sy), __float_as_uint(sz), __float_as_uint(1.0f / sx));
    rec[1] = make_uint4(__float_as_uint(1.0f / sy), __float_as_uint(1.0f / sz), __float_as_uint(e.W),
                        __float_as_uint(e.Bi));
    rec[2] = make_uint4(__float_as_uint(e.invBi), key | ((uint32_t)e.rank << 16),
                        ((uint32_t)unit & 0xFFFFu) | ((uint32_t)h << 16), (uint32_t)tid);
}

template <bool SURF>
__device__ __forceinline__ void q_get(const uint32_t *q, int qs, const Slots &sl, const DevCommon &cm,
                                      const DevSnap &sn, Ray &r, En &e, int &unit, int &tid, int &h) {
    const uint4 *rec = reinterpret_cast<const uint4 *>(q + qs * QREC);
    const uint4 a = rec[0], b = rec[1], c = rec[2];
    r.sx = __uint_as_float(a.x);
    r.sy = __uint_as_float(a.y);
    r.sz = __uint_as_float(a.z);
    r.isx = __uint_as_float(a.w);
    r.isy = __uint_as_float(b.x);
    r.isz = __uint_as_float(b.y);
    r.f0 = r.sx < 0.f; r.f1 = r.sy < 0.f; r.f2 = r.sz < 0.f;
    r.tra = 1.0f;
    e.W = __uint_as_float(b.z);
    e.Bi = __uint_as_float(b.w);
    e.invBi = __uint_as_float(c.x);
    e.row = c.y & 0xFFFFu;
    e.rank = (int)(c.y >> 16);
    unit = (int)(c.z & 0xFFFFu);
    h = (int)(c.z >> 16);
    tid = (int)c.w;
    const uint4 *pr = reinterpret_cast<const uint4 *>(&sl.pl[unit & (NSLOT - 1)]);
    const uint4 p0 = pr[0], p1 = pr[1];
    r.ic = (int)p0.x; r.jc = (int)p0.y; r.kc = (int)p0.z;
    r.xp = __uint_as_float(p0.w); r.yp = __uint_as_float(p1.x); r.zp = __uint_as_float(p1.y);
    if (SURF) surface_adjust(sn, r);
}

/* the distance from unit number u (16 bits, as the ray records hold it)
 * to the wavefront's unit count seq */
__device__ __forceinline__ int unit_age(int seq, int u) { return (int)((uint32_t)(seq - u) & 0xFFFFu); }

/* a ray's end: its tally and count go to its unit's slot, or to HBM once
 * the slot has been taken over (seq: units the wavefront has started); in
 * the deterministic reduction, the tally goes to the ray's own entry */
__device__ __forceinline__ void finish_ray(const DevCommon &cm, const Slots &sl, int seq, int runit, int rtid, int h,
                                           float dacc, double *__restrict__ de) {
    if (cm.ray_de) {
        cm.ray_de[(size_t)rtid * (size_t)cm.photmax + (size_t)h] = dacc;
    } else if (unit_age(seq, runit) <= NSLOT) {
        SlotTally &t = sl.tal[runit & (NSLOT - 1)];
        atomicAdd(&t.acc, (double)dacc);
        atomicAdd(&t.done, 1);
    } else {
        atomicAdd(&de[rtid], (double)dacc);
    }
}

/* LDS words of one wavefront: unit slots, ray queue, (multigrid) the grid
 * levels of the current snapshot, then (srt) up to 256 keys the unit's
 * sorted keys (16-bit: keys are below nB nQ <= 65535; 256 of them, the
 * register sort's width) preceded by the direction uniforms of its rays,
 * above 256 keys the radix sort's two digit histograms (512 words) */
constexpr int LEV_WORDS = (int)(sizeof(DevLevel) / 4);

/* drain of the last rays: at most COOP_RAYS live rays are marched one at a
 * time by the whole wavefront; above that, in blocks of DRAIN_BLOCK steps
 * per lane */
constexpr int COOP_RAYS = 4;
constexpr int DRAIN_BLOCK = 64;

/* lane src's value in every lane, as a per-lane value (a crossbar
 * permute), so that the arithmetic on it stays on the vector ALU */
__device__ __forceinline__ int bcast(int x, int src) { return __shfl(x, src); }
__device__ __forceinline__ uint32_t bcast(uint32_t x, int src) { return (uint32_t)__shfl((int)x, src); }
__device__ __forceinline__ float bcast(float x, int src) { return __shfl(x, src); }
__device__ __forceinline__ Ray bcast_ray(const Ray &a, int src) {
    Ray b;
    b.ic = bcast(a.ic, src); b.jc = bcast(a.jc, src); b.kc = bcast(a.kc, src);
    b.f0 = bcast(a.f0, src); b.f1 = bcast(a.f1, src); b.f2 = bcast(a.f2, src);
    b.xp = bcast(a.xp, src); b.yp = bcast(a.yp, src); b.zp = bcast(a.zp, src);
    b.sx = bcast(a.sx, src); b.sy = bcast(a.sy, src); b.sz = bcast(a.sz, src);
    b.isx = bcast(a.isx, src); b.isy = bcast(a.isy, src); b.isz = bcast(a.isz, src);
    b.tra = bcast(a.tra, src);
    return b;
}
__device__ __forceinline__ En bcast_en(const En &a, int src) {
    En b;
    b.row = bcast(a.row, src);
    b.rank = bcast(a.rank, src);
    b.W = bcast(a.W, src);
    b.Bi = bcast(a.Bi, src);
    b.invBi = bcast(a.invBi, src);
    return b;
}
__host__ __device__ __forceinline__ int batched_lds_words(int n2, int srt, int multi) {
    return SLOT_WORDS * NSLOT + QREC * QCAP + (multi ? LEV_WORDS * RC_MAXLEV : 0) +
           (srt ? (n2 <= 256 ? 2 * n2 + 128 : 512) : 0);
}

template <bool SURF, bool MULTI, int TM, bool RADIX>
__device__ __forceinline__ void batched_body(const DevCommon &cm, const DevSnap *__restrict__ snaps, int nsnap,
                                             double *__restrict__ de, unsigned long long *__restrict__ stats,
                                             int *__restrict__ queue, int threshold, int spi) {
    extern __shared__ uint32_t lds[];
    const Slots sl = slot_view(lds);
    uint32_t *rq = lds + SLOT_WORDS * NSLOT;
    DevLevel *llev = reinterpret_cast<DevLevel *>(rq + QREC * QCAP);   /* 16-byte aligned */
    const int P = cm.photmax;
    int n2 = 1;
    while (n2 < P) n2 <<= 1;
    /* srt: up to 256 keys the direction uniforms of the unit's rays are
     * kept from the key draws and the sorted keys sit in LDS; above that
     * the uniforms are drawn again at the rays' generation, and the keys
     * are sorted in the wavefront's HBM scratch (raw and sorted keys in
     * kr, the first radix pass in ka) */
    const int nkeep = RADIX ? 0 : n2;
    float *dira = reinterpret_cast<float *>(rq + QREC * QCAP + (MULTI ? LEV_WORDS * RC_MAXLEV : 0));
    float *dirb = dira + nkeep;
    uint16_t *keys = reinterpret_cast<uint16_t *>(dirb + nkeep);
    const int Pr = (P + 63) & ~63;
    uint16_t *const kr = RADIX ? cm.srt_keys + (size_t)blockIdx.x * 2 * Pr : nullptr;
    uint16_t *const ka = kr ? kr + Pr : nullptr;
    const int lane = threadIdx.x;
    const DevSnap &slast = snaps[nsnap - 1];
    const int total = (SURF ? slast.sbase + 2 * slast.nsurf : slast.ubase + slast.ncell) * cm.nVar;
    const float tol = SURF ? cm.tollS : cm.toll;
    /* the drain march reads the x faces of every level from LDS */
    int nxf = 0;
    for (int s = 0; s < nsnap; ++s) {
        int n = 0;
        for (int g = 0; g < (MULTI ? cm.grid_num : 1); ++g) n += snaps[s].lev[g].im + 1;
        nxf = max(nxf, n);
    }
    const bool xfit = uni(nxf) <= QREC * QCAP;

    __syncthreads();

    /* wavefront-uniform state: dispatch of units, unit slots, ray queue */
    int cur_h = P, nxt = -1, nxt_s = -1, qdone = 0, snap = -1, seq = 0, cunit = 0, ctid = 0;
    int qhead = 0, qcount = 0;
    const int chunk = min(max(cm.nVar, 1), 16);
    const int nch = (total + chunk - 1) / chunk;
    int cnext = 0, cleft = 0;
    int part = (int)(blockIdx.x % QPARTS), ptries = 0;   /* workgroups are dealt round-robin over the XCDs */
    Unit U{};

    /* lane state; an idle lane keeps in-range values, since the march step
     * runs in every lane */
    bool active = false;
    Ray r;
    r.ic = r.jc = r.kc = 1;
    r.f0 = r.f1 = r.f2 = 0;
    r.xp = r.yp = r.zp = 0.f;
    r.sx = r.sy = r.sz = 1.f;
    r.isx = r.isy = r.isz = 1.f;
    r.tra = 0.f;
    En e;
    e.row = 0u;
    e.rank = 0;
    e.W = 0.f;
    e.Bi = 1.f;
    e.invBi = 1.f;
    float dacc = 0.f;
    int runit = 0, rtid = 0, rh = 0, lev = 0, counter = 0;
    int kt = -1;                  /* the lane's cached table pair (t, values) */
    KbRaw kc;
    kc.q = make_float4(0.f, 0.f, 0.f, 0.f);
    kc.w = 0u;
    kc.a = 0.f;
    int wall = 0;                 /* the ray waits at a black wall, entry widx of the level's positions */
    uint32_t widx = 0u;
    LaneLevel ll{};               /* multigrid: the level of the lane's ray */
    CellWalk cw;                  /* the ray's linear cell index and its steps along y and z */
    cw.cell = 0u;
    cw.dcy = cw.dcz = 0;
    int nsteps = 0, nrays = 0;
    int bsub = 0;                 /* bulk march sub-steps of the wavefront (every lane steps) */
    int64_t bsteps = 0;           /* steps of live rays in the bulk march (wavefront total) */

    for (;;) {
        /* (a) once the current unit is fully generated, take the next unit
         * and (srt) draw and sort its band keys, keeping the direction
         * uniforms of every ray */
        if (nxt < 0 && !qdone && cur_h >= P) {
            /* units are taken in chunks of nVar (one cell's runs, so that
             * one wavefront writes the cell's consecutive tallies).  The
             * chunks are split into QPARTS contiguous ranges (z slabs of
             * the field), one per XCD, each with its own counter, so that
             * an XCD's L2 holds its own slab; a range that runs out sends
             * the wavefront to the next one. */
            while (cleft == 0 && part >= 0) {
                const int pb = (int)((long)nch * part / QPARTS), pe = (int)((long)nch * (part + 1) / QPARTS);
                int c0 = 0;
                if (lane == 0) c0 = atomicAdd(&queue[part], 1);
                c0 = uni(c0);
                if (pb + c0 < pe) {
                    cnext = (pb + c0) * chunk;
                    cleft = min(chunk, total - cnext);
                } else {
                    part = (part + 1) % QPARTS;
                    if (++ptries >= QPARTS) part = -1;
                }
            }
            const int u = cleft > 0 ? cnext : total;
            cnext += 1;
            cleft -= (cleft > 0) ? 1 : 0;
            if (u >= total) {
                qdone = 1;
            } else {
                nxt = u;
                const Unit N = uni(decode<SURF>(cm, snaps, nsnap, u));
                nxt_s = N.s;
                if (cm.srt && !N.skip) {
                    const DevSnap &sn = snaps[N.s];
                    const uint32_t tagx = N.tag + 16u * (uint32_t)sn.snap;
                    if constexpr (!RADIX) {
                        /* up to 256 keys: drawn one round of 64 at a time
                         * up to photmax, padded to 128 or 256, then sorted
                         * in registers, 2 or 4 keys per lane */
                        const int nr = n2 <= 128 ? 128 : 256;
#pragma unroll 1
                        for (int h = lane; h < ((P + 63) & ~63); h += 64) {
                            uint32_t key = 0xFFFFFFFFu;
                            if (h < P) {
                                const U4 q = draw4(cm.k0, cm.k1, (uint32_t)h, (uint32_t)N.v, N.cid, tagx);
                                key = band_key(cm, sn, SURF, q.c, q.d);
                                dira[h] = q.a;
                                dirb[h] = q.b;
                            }
                            keys[h] = (uint16_t)key;
                        }
                        for (int h = ((P + 63) & ~63) + lane; h < nr; h += 64) keys[h] = (uint16_t)0xFFFFu;
                        __syncthreads();
                        if (nr == 128) {
                            /* keys 2 lane, 2 lane + 1: one 4-byte LDS word */
                            uint32_t *kw = reinterpret_cast<uint32_t *>(keys) + lane;
                            const uint32_t w = *kw;
                            uint32_t x[2] = {w & 0xFFFFu, w >> 16};
                            wave_sort<2>(x, lane);
                            *kw = x[0] | (x[1] << 16);
                        } else {
                            /* keys 4 lane .. 4 lane + 3: one 8-byte LDS word */
                            uint2 *kw = reinterpret_cast<uint2 *>(keys) + lane;
                            const uint2 w = *kw;
                            uint32_t x[4] = {w.x & 0xFFFFu, w.x >> 16, w.y & 0xFFFFu, w.y >> 16};
                            wave_sort<4>(x, lane);
                            *kw = make_uint2(x[0] | (x[1] << 16), x[2] | (x[3] << 16));
                        }
                        __syncthreads();
                    } else {
                        /* above 256 keys: drawn into HBM with both digit
                         * histograms counted in LDS, then two radix passes
                         * (low byte into ka, high byte back into kr) */
                        uint32_t *hist = reinterpret_cast<uint32_t *>(keys);
                        for (int w = lane; w < 512; w += 64) hist[w] = 0u;
                        __syncthreads();
#pragma unroll 1
                        for (int h = lane; h < P; h += 64) {
                            const U4 q = draw4(cm.k0, cm.k1, (uint32_t)h, (uint32_t)N.v, N.cid, tagx);
                            const uint32_t key = band_key(cm, sn, SURF, q.c, q.d);
                            kr[h] = (uint16_t)key;
                            atomicAdd(&hist[key & 255u], 1u);
                            atomicAdd(&hist[256u + (key >> 8)], 1u);
                        }
                        __syncthreads();
                        radix_offsets(hist, lane);
                        radix_pass(kr, ka, hist, P, 0, lane);
                        radix_pass(ka, kr, hist + 256, P, 8, lane);
                    }
                }
            }
        }
        const unsigned long long live = __ballot(active);
        if (cur_h >= P && nxt >= 0 && (nxt_s == snap || (live == 0ull && qcount == 0)) &&
            (qcount == 0 || unit_age(seq, uni((int)(rq[qhead * QREC + 10] & 0xFFFFu))) < NSLOT)) {
            /* (b) start the prepared unit (a new snapshot only once every
             * lane and the queue are empty; never over the slot of a ray
             * still queued) */
            U = uni(decode<SURF>(cm, snaps, nsnap, nxt));
            nxt = -1;
            if (MULTI && U.s != snap) {
                const uint32_t *src = reinterpret_cast<const uint32_t *>(snaps[U.s].lev);
                for (int w = lane; w < LEV_WORDS * cm.grid_num; w += 64)
                    reinterpret_cast<uint32_t *>(llev)[w] = gld(src, w);
            }
            snap = U.s;
            const int tidx = ((SURF ? snaps[U.s].sbase : snaps[U.s].ubase) + U.n) * cm.nVar + U.v;
            if (!U.skip) {
                const int cslot = seq & (NSLOT - 1);
                cunit = seq;
                ctid = tidx;
                cur_h = 0;
                if (lane == 0) {
                    const DevSnap &sn = snaps[U.s];
                    SlotTally &t = sl.tal[cslot];
                    if (seq >= NSLOT && !cm.ray_de) {
                        if (t.done == P) de[t.tid] = t.acc;
                        else atomicAdd(&de[t.tid], t.acc);
                    }
                    t.acc = 0.0;
                    t.done = 0;
                    t.tid = tidx;
                    SlotPlace &pl = sl.pl[cslot];
                    pl.ic = SURF ? (U.bound ? sn.imax : 1) : U.i;
                    pl.jc = U.j;
                    pl.kc = U.k;
                    pl.xp = SURF ? (U.bound ? cm.Lx : 0.f) : gld(sn.x0, U.i);
                    pl.yp = gld(sn.y0, U.j);
                    pl.zp = gld(sn.z0, U.k);
                }
                seq += 1;
            }
            __syncthreads();
        }
        /* one back-edge: a unit may start and generate in the same iteration;
         * an iteration with no live lane steps idle lanes only */
        {
            /* (c) bulk generation: every lane sets up one ray of the current unit */
            const int left = P - cur_h;
            if (left > 0 && qcount <= QCAP - 64) {
                const int ng = min(64, left);
                if (lane < ng) {
                    const DevSnap &sn = snaps[U.s];
                    const int h = cur_h + lane;
                    float a, b;
                    uint32_t key;
                    if (!RADIX && cm.srt) {
                        key = (uint32_t)keys[h];
                        a = dira[h];
                        b = dirb[h];
                    } else if (RADIX) {
                        key = (uint32_t)kr[h];
                        /* the direction uniforms drawn again, unless the
                         * Sobol sampler replaces them below */
                        a = b = 0.f;
                        if (SURF || !cm.sampler) {
                            const U4 q = draw4(cm.k0, cm.k1, (uint32_t)h, (uint32_t)U.v, U.cid,
                                               U.tag + 16u * (uint32_t)sn.snap);
                            a = q.a;
                            b = q.b;
                        }
                    } else {
                        const U4 q = draw4(cm.k0, cm.k1, (uint32_t)h, (uint32_t)U.v, U.cid, U.tag + 16u * (uint32_t)sn.snap);
                        key = band_key(cm, sn, SURF, q.c, q.d);
                        a = q.a;
                        b = q.b;
                    }
                    if (!SURF && cm.sampler) sobol_uniforms((uint32_t)h, U.sh0, U.sh1, a, b);
                    float sx, sy, sz;
                    if (SURF) surface_dir(a, b, U.bound, sx, sy, sz);
                    else fluid_dir(a, b, sx, sy, sz);
                    En ge;
                    energy<TM>(cm, sn, key, U.Ti, SURF, U.bound, ge);
                    q_put(rq, qwrap(qhead + qcount + lane), sx, sy, sz, ge, key, cunit, ctid, h);
                }
                __syncthreads();
                qcount += ng;
                cur_h += ng;
            }
            /* (d) idle lanes take queued rays, once `threshold` lanes
             * are idle (or the queue holds fewer rays) */
            auto take_rays = [&](unsigned long long idle) {
                const int nidle = __popcll(idle);
                if (qcount > 0 && nidle >= min(threshold, qcount)) {
                    const int rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32),
                                                               __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
                    const int npop = min(nidle, qcount);
                    if (!active && rank < npop) {
                        q_get<SURF>(rq, qwrap(qhead + rank), sl, cm, snaps[snap], r, e, runit, rtid, rh);
                        active = true;
                        dacc = 0.f;
                        lev = 0;
                        counter = 0;
                        kt = -1;
                        wall = 0;
                        if (MULTI) {
                            ll = lane_level(llev[0]);
                            cell_walk_setup(ll, r, cw);
                        } else {
                            cell_walk_setup(level_view(&snaps[snap].lev[0]), r, cw);
                        }
                    }
                    qhead = qwrap(qhead + npop);
                    qcount -= npop;
                }
            };
            take_rays(~live);
            const bool drained = qdone && nxt < 0 && cur_h >= P && qcount == 0;
            if (drained && (__ballot(active) == 0ull || xfit)) break;   /* to the drain (f) */
            if (__ballot(active) != 0ull) {
                /* (e) march steps of every active lane; the snapshot's
                 * fields (and the level, single grid) are wavefront-uniform
                 * values, a lane's level is held in its registers
                 * (multigrid: ll, set at pick-up and at each coarsening) */
                const SnapView sv = snap_view(snaps + snap);
                DevLevel l0;
                if (!MULTI) l0 = level_view(&snaps[snap].lev[0]);
                const float *xb = MULTI ? sv.xu_all : l0.xu;
                const float *pb = MULTI ? sv.P_all : l0.P;
                const uint32_t xmax = MULTI ? (uint32_t)(sv.nxu - 1) : (uint32_t)l0.im;
                const uint32_t pmax = MULTI ? (uint32_t)(sv.npos - 1)
                                            : (uint32_t)(level_pos_count(l0.im, l0.jm, l0.km) - 1);
                /* spi march steps.  Every active lane holds a live ray under
                 * its level's step cap: a ray is finished (or coarsened)
                 * right after the step that ends it.  Idle lanes step too, on
                 * stale state that the next pick-up overwrites; the step
                 * bounds its loads. */
                bsub += spi;
                auto substep = [&](const auto &L, uint32_t xo, uint32_t po) {
                    /* a wall step is no march step of the reference (its
                     * tally belongs to the step that crossed the face): it
                     * neither counts nor meets the step cap */
                    const bool ws = wall != 0;
                    const bool out = march_core_bulk<TM>(cm, sv, L, xb, xo, xmax, pb, po, pmax, r, e, tol, dacc, kt,
                                                         kc, ws, widx, cw);
                    boundary_defer(cm, L, r, active && out, wall, widx, cw);
                    counter += 1;           /* a wall step ends the ray: its count is never read */
                    bsteps += __popcll(__ballot(active && !ws));
                    const bool alive = r.tra > tol;
                    const bool done = active && (!alive || (counter >= L.maxs && wall == 0));
                    const bool fin = done && (!MULTI || !alive || lev == cm.grid_num - 1);
                    const bool crs = MULTI && done && !fin;
                    if (fin) {
                        finish_ray(cm, sl, seq, runit, rtid, rh, dacc, de);
                        nrays += 1;
                        active = false;
                        wall = 0;
                    }
                    if (crs) {
                        /* coarsen the indices for the next grid level (:215-223) */
                        const DevLevel &Lc = llev[lev];
                        r.ic = coarsen(r.ic, Lc.cr[0], Lc.cmul[0]);
                        r.jc = coarsen(r.jc, Lc.cr[1], Lc.cmul[1]);
                        r.kc = coarsen(r.kc, Lc.cr[2], Lc.cmul[2]);
                        ll = lane_level(llev[lev + 1]);
                        cell_walk_setup(ll, r, cw);
                        lev += 1;
                        counter = 0;
                    }
                };
#pragma unroll 1
                for (int q = 0; q < spi; ++q) {
                    /* refill between the steps too: a lane whose ray ended
                     * idles for less of the iteration */
                    if (cm.refill && q > 0) take_rays(~__ballot(active));
                    if (MULTI) substep(ll, ll.xoff, ll.poff);
                    else substep(l0, 0u, 0u);
                }
            }
        }
    }
    /* (f) no ray left to start: the rays still live here are
     * the longest ones.  Multigrid: every live lane marches
     * its ray to the end with the look-ahead march (the step
     * caps of the levels keep the rays short).  Single grid:
     * while more than COOP_RAYS lanes are live, every live lane
     * marches its ray DRAIN_BLOCK steps with the look-ahead
     * march; below that, the wavefront marches the remaining
     * rays one at a time, all lanes on one ray.  Then the
     * wavefront is done. */

    if (active && wall) {
        /* a ray still waiting at its wall: the wall's emission, and the ray is done */
        const SnapView sv = snap_view(snaps + snap);
        const DevLevel *lp = MULTI ? &llev[lev] : &snaps[snap].lev[0];
        float Kw, Bw;
        kb_pos<TM>(sv, e.row, e.rank, cm.nT, gld(lp->P, widx), Kw, Bw);
        wall_hit(e, r, Bw, dacc);
        finish_ray(cm, sl, seq, runit, rtid, rh, dacc, de);
        nrays += 1;
        active = false;
        wall = 0;
    }
    if (__ballot(active) != 0ull) {
        const SnapView sv = snap_view(snaps + snap);
        DevLevel l0;
        if (!MULTI) l0 = level_view(&snaps[snap].lev[0]);
        /* the x faces of every level to LDS (the ray queue's
         * space, empty now), level g from xoff(g) */
        float *xul = reinterpret_cast<float *>(rq);
        int off = 0;
        for (int g = 0; g < (MULTI ? cm.grid_num : 1); ++g) {
            const DevLevel Lg = MULTI ? llev[g] : l0;
            for (int w = lane; w <= Lg.im; w += 64) xul[off + w] = gld(Lg.xu, w);
            off += Lg.im + 1;
        }
        __syncthreads();
        if (MULTI) {
            if (active) {
                /* the lane's ray may be on a coarser level already */
                int xoff = 0;
                for (int g = 0; g < lev; ++g) xoff += llev[g].im + 1;
                for (;;) {
                    const DevLevel L = llev[lev];
                    march_ahead<TM, true>(cm, sv, L, r, e, tol, dacc, counter, nsteps, xul + xoff, L.maxs);
                    if (!(r.tra > tol) || lev == cm.grid_num - 1) break;
                    /* coarsen the indices for the next grid level */
                    r.ic = coarsen(r.ic, L.cr[0], L.cmul[0]);
                    r.jc = coarsen(r.jc, L.cr[1], L.cmul[1]);
                    r.kc = coarsen(r.kc, L.cr[2], L.cmul[2]);
                    xoff += L.im + 1;
                    lev += 1;
                    counter = 0;
                }
                finish_ray(cm, sl, seq, runit, rtid, rh, dacc, de);
                nrays += 1;
                active = false;
            }
        } else {
            for (;;) {
                const unsigned long long act = __ballot(active);
                if (act == 0ull) break;
                if (__popcll(act) > COOP_RAYS) {
                    if (active) {
                        march_ahead<TM, true>(cm, sv, l0, r, e, tol, dacc, counter, nsteps, xul,
                                              min(l0.maxs, counter + DRAIN_BLOCK));
                        if (!(r.tra > tol) || counter >= l0.maxs) {
                            finish_ray(cm, sl, seq, runit, rtid, rh, dacc, de);
                            nrays += 1;
                            active = false;
                        }
                    }
                } else {
                    /* the lowest live lane's ray, by every lane */
                    const int src = __builtin_ctzll(act);
                    Ray cr = bcast_ray(r, src);
                    const En ce = bcast_en(e, src);
                    float cd = bcast(dacc, src);
                    int cc = bcast(counter, src), cs = 0;
                    march_coop<TM>(cm, sv, l0, cr, ce, tol, cd, cc, cs, xul, lane);
                    if (lane == src) {
                        finish_ray(cm, sl, seq, runit, rtid, rh, cd, de);
                        nsteps += cs;
                        nrays += 1;
                        active = false;
                    }
                }
            }
        }
    }
    /* the last units' totals (every ray has finished) */
    __syncthreads();
    if (lane < NSLOT && lane < seq && !cm.ray_de) {
        const SlotTally &t = sl.tal[lane];
        if (t.done == P) de[t.tid] = t.acc;
        else atomicAdd(&de[t.tid], t.acc);
    }
    /* the drain's per-lane step counts stay 32-bit (a lane marches far
     * fewer than 2^32 steps: 1e13 steps over 16 k wavefronts is 1e7 per
     * lane); the wavefront's sums are 64-bit */
    unsigned long long ws = (uint32_t)nsteps, wr = (uint32_t)nrays;
    for (int o = 32; o > 0; o >>= 1) {
        ws += __shfl_xor(ws, o);
        wr += __shfl_xor(wr, o);
    }
    if (lane == 0) {
        atomicAdd(&stats[0], wr);
        atomicAdd(&stats[1], ws + (unsigned long long)bsteps);
        atomicAdd(&stats[3], 64ull * (uint32_t)bsub);
        atomicAdd(&stats[4], (unsigned long long)bsteps);
    }
}

/* single grid (registers held to 5 wavefronts per SIMD: the march is
 * issue-bound, and a fifth wavefront outweighs the few values the
 * compiler then keeps in scratch).  RADIX: srt above 256 keys, sorted by
 * the radix passes; otherwise (no srt, or up to 256 keys) the register
 * sort, with no radix code in the kernel. */
template <bool SURF, int TM, bool RADIX>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5, 8)))
k_batched(DevCommon cm, const DevSnap *__restrict__ snaps, int nsnap, double *__restrict__ de,
          unsigned long long *__restrict__ stats, int *__restrict__ queue, int threshold, int spi) {
    batched_body<SURF, false, TM, RADIX>(cm, snaps, nsnap, de, stats, queue, threshold, spi);
}

/* multigrid (registers held to 5 wavefronts per SIMD) */
template <bool SURF, int TM, bool RADIX>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5, 8)))
k_batched_mg(DevCommon cm, const DevSnap *__restrict__ snaps, int nsnap, double *__restrict__ de,
             unsigned long long *__restrict__ stats, int *__restrict__ queue, int threshold, int spi) {
    batched_body<SURF, true, TM, RADIX>(cm, snaps, nsnap, de, stats, queue, threshold, spi);
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

/* one wavefront per unit: every lane sums its rays h = lane, lane + 64, ...
 * in order (Neumaier), then the lanes' (sum, error) pairs are combined by
 * a butterfly of TwoSums; the same operations every run, whatever the
 * order the rays ended in */
__global__ void __launch_bounds__(256) k_reduce_rays(const float *__restrict__ ray, size_t nunit, int P,
                                                     double *__restrict__ de) {
    const int lane = threadIdx.x & 63;
    const size_t w0 = (blockIdx.x * (size_t)blockDim.x + threadIdx.x) >> 6;
    const size_t nw = ((size_t)gridDim.x * blockDim.x) >> 6;
    for (size_t u = w0; u < nunit; u += nw) {
        const float *x = ray + u * (size_t)P;
        double s = 0.0, c = 0.0;
        for (int h = lane; h < P; h += 64) {
            const double v = (double)x[h];
            const double t = s + v;
            c += fabs(s) >= fabs(v) ? (s - t) + v : (v - t) + s;
            s = t;
        }
        for (int o = 1; o < 64; o <<= 1) {
            const double s2 = __shfl_xor(s, o), c2 = __shfl_xor(c, o);
            const double t = s + s2;
            const double bp = t - s;
            const double e = (s - (t - bp)) + (s2 - bp);
            c = c + c2 + e;
            s = t;
        }
        if (lane == 0) de[u] = s + c;
    }
}

void launch_reduce_rays(hipStream_t st, const float *ray, size_t nunit, int photmax, double *de) {
    const int blocks = (int)std::max<size_t>(1, std::min<size_t>((nunit + 3) / 4, 65536));
    hipLaunchKernelGGL(k_reduce_rays, dim3(blocks), dim3(256), 0, st, ray, nunit, photmax, de);
}

void launch_zero_u64(hipStream_t st, unsigned long long *p, int n) {
    hipLaunchKernelGGL(k_zero_u64, dim3((n + 63) / 64), dim3(64), 0, st, p, n);
}

void launch_zero(hipStream_t st, double *p, size_t n) {
    const int blocks = (int)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 8192));
    hipLaunchKernelGGL(k_zero_f64, dim3(blocks), dim3(256), 0, st, p, n);
}

/* workgroups per CU by default: 32 on a single grid, 20 on several levels
 * (what the autotuner picks on 256x128x256 and 192x128x192) */
static int default_wg_per_cu(const DevCommon &cm, int wg_per_cu) {
    return wg_per_cu > 0 ? wg_per_cu : (cm.grid_num > 1 ? 20 : 32);
}

/* Defaults of the schedule (a refill mode < 0, a threshold or steps per
 * iteration <= 0).  Single grid: refills at 4 idle lanes, also between
 * the steps, 6 steps per iteration: within 1 % of the autotuner's pick at
 * 256x128x256 (8 steps) and of the best at 64^3, where rays of 6 steps end
 * faster than the one round of 64 new rays per iteration replaces them at
 * 8 steps (0.60 of the lanes active, 10 % slower).
 * Several levels: refills at 4 idle lanes once per iteration and as
 * many steps per iteration as the finest level's step cap, so that the
 * rays a wavefront picks up together reach that cap, and hand off to the
 * next level, on the same step: one pass through the hand-off serves all
 * of them (param.h's maxs 5 on its 406x576x2304 share: 6.27 s against
 * 6.60 s with 4 steps and refills between them). */
void batched_schedule(const DevCommon &cm, int &wg_per_cu, int &threshold, int &spi, int &refill) {
    const bool mg = cm.grid_num > 1;
    if (refill < 0) refill = mg ? 0 : 1;
    wg_per_cu = default_wg_per_cu(cm, wg_per_cu);
    if (threshold <= 0) threshold = 4;
    if (spi <= 0) spi = !mg ? (refill ? 6 : 3) : std::max(1, std::min(16, cm.maxs0));
}

int batched_grid(const DevCommon &cm, long nunit_cells, int wg_per_cu) {
    static int ncu = 0;       /* compute units of the (single) device */
    if (ncu <= 0) {
        int dev = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
        if (ncu <= 0) ncu = 256;
    }
    const long total = nunit_cells * (long)cm.nVar;
    return (int)std::max<long>(1, std::min<long>((long)ncu * default_wg_per_cu(cm, wg_per_cu), total));
}

size_t batched_srt_keys(const DevCommon &cm, long nunit_cells, int wg_per_cu) {
    if (!cm.srt || cm.photmax <= 256) return 0;
    return (size_t)batched_grid(cm, nunit_cells, wg_per_cu) * 2 * (size_t)((cm.photmax + 63) & ~63);
}

void launch_batched(hipStream_t st, const DevCommon &cm0, const DevSnap *snaps, int nsnap, int nunit_cells,
                    int surface, int snaps_bf16, double *de, unsigned long long *stats, int *queue, int wg_per_cu,
                    int threshold, int spi) {
    DevCommon cm = cm0;
    batched_schedule(cm, wg_per_cu, threshold, spi, cm.refill);
    int n2 = 1;
    while (n2 < cm.photmax) n2 <<= 1;
    const size_t lds = (size_t)batched_lds_words(n2, cm.srt, cm.grid_num > 1) * 4;
    const size_t nde = (size_t)nunit_cells * cm.nVar;
    const int grid = batched_grid(cm, nunit_cells, wg_per_cu);
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, st, queue, QPARTS);
    launch_zero(st, de, nde);
    const bool multi = cm.grid_num > 1;
    const bool bf = snaps_bf16;
    const bool radix = cm.srt && n2 > 256;
    if (radix && !cm.srt_keys) return;     /* the callers size it (batched_srt_keys); never launched without */
#define RC_LAUNCH(K, S, T, R) hipLaunchKernelGGL((K<S, T, R>), dim3(grid), dim3(64), lds, st, cm, snaps, nsnap, de, \
                                              stats, queue, threshold, spi)
#define RC_LAUNCH2(K, S, T) do { if (radix) RC_LAUNCH(K, S, T, true); else RC_LAUNCH(K, S, T, false); } while (0)
    if (!surface) {
        if (multi) { if (bf) RC_LAUNCH2(k_batched_mg, false, TAB_BF16); else RC_LAUNCH2(k_batched_mg, false, TAB_F32); }
        else { if (bf) RC_LAUNCH2(k_batched, false, TAB_BF16); else RC_LAUNCH2(k_batched, false, TAB_F32); }
    } else {
        if (multi) { if (bf) RC_LAUNCH2(k_batched_mg, true, TAB_BF16); else RC_LAUNCH2(k_batched_mg, true, TAB_F32); }
        else { if (bf) RC_LAUNCH2(k_batched, true, TAB_BF16); else RC_LAUNCH2(k_batched, true, TAB_F32); }
    }
#undef RC_LAUNCH2
#undef RC_LAUNCH
}

}  // namespace rc
