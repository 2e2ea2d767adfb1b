/*
 * rc_capi.cpp -- session state and the C ABI of the radiative solver.
 *
 * mc_gpu_ / get_results_ follow src/montecarlo/mc_gpu_noint.cu:660-874 and
 * :876-1021: copy the temperature field, launch asynchronously, later
 * synchronise and scatter mean div(q_r) and sqrt(variance) to the Fortran
 * layout.  The rc_* extensions expose the same machinery as a device-
 * resident session (tables and grids uploaded once, several snapshots).
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "rc_internal.h"
#include "rc_kernels.h"

namespace rc {
void launch_batched(hipStream_t st, const DevCommon &cm, const DevSnap *snaps, int nsnap, int nunit_cells, int surface,
                    int snaps_bf16, double *de, unsigned long long *stats, int *queue, int wg_per_cu, int threshold,
                    int spi);
}

using namespace rc;

namespace {

constexpr int kGuideM = 2048;   /* guide-table resolution for the band CDFs */

/* a roctx range over the enclosing scope (set-up, launch and harvest
 * phases show up on a rocprofv3 --marker-trace timeline; no cost without
 * a profiler attached) */
struct Range {
    explicit Range(const char *what) { roctxRangePushA(what); }
    ~Range() { roctxRangePop(); }
    Range(const Range &) = delete;
    Range &operator=(const Range &) = delete;
};

#define RC_HIP(call)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess) {                                                              \
            set_error(std::string("HIP error: ") + hipGetErrorString(e_) + " at " #call);   \
            return -10;                                                                      \
        }                                                                                    \
    } while (0)

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr; bytes = 0;
    }
    int ensure(size_t n) {
        if (n <= bytes && p) return 0;
        release();
        if (n == 0) n = 16;
        if (hipMalloc(&p, n) != hipSuccess) {
            p = nullptr;
            set_error("device allocation of " + std::to_string(n) + " bytes failed");
            return -11;
        }
        bytes = n;
        return 0;
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

template <class T> int upload(DevBuf &b, const std::vector<T> &v, hipStream_t st) {
    if (int e = b.ensure(v.size() * sizeof(T))) return e;
    RC_HIP(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
    return 0;
}

/* device tables in sorted band order for one temperature table */
struct TableSet {
    float Tnb0 = 0.f, invdT = 0.f;
    int bf16 = 0;
    std::vector<float> Tnb;
    DevBuf kb, cdf_gs, rank_of, kp16, bp;
};

/* round-to-nearest-even FP32 -> BF16 (finite inputs) */
inline uint16_t to_bf16(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

struct SnapState {
    float Tmax = 0.f, kappamax = 0.f, Ibmax = 0.f, Ibmax_w = 0.f, Ibmax_e = 0.f;
    double kappamax64 = 0.0, Ibmax64 = 0.0, Ibmax_w64 = 0.0, Ibmax_e64 = 0.0;   /* the FP64 path's */
    int tm = 0, tstar = 0;
    DevBuf T0, cdf_b, cdf_g, cdf_bs, guide_b, guide_bs;
    DevBuf Pall;                          /* table positions of every level (DevLevel.P) */
    DevBuf Tall;                          /* FP64 path: the temperatures behind Pall */
    std::unique_ptr<TableSet> grey;     /* grey runs: tables follow Tmax */
    Tables greyhost;
};

}  // namespace

struct rc_ctx {
    rc_params p{};
    int kloc = 0, ncell = 0, nsurf = 0, zstart = 1;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, evm0 = nullptr, evm1 = nullptr;
    Tables tab;
    bool have_tables = false;
    std::vector<int32_t> idx, rank_of;
    TableSet nbset;
    std::vector<Grid> grids;
    DevBuf xuall, x0, y0, z0;               /* x faces of every level, level g at xoff[g] */
    int xoff[RC_MAXLEV] = {}, nxu = 0;
    size_t poff[RC_MAXLEV] = {}, npos = 0;  /* level g's table positions at poff[g] */
    DevBuf cx[RC_MAXLEV], cy[RC_MAXLEV], cz[RC_MAXLEV];   /* cell centres per level */
    DevBuf spl;                             /* spline restriction scratch */
    int restriction = 0;                    /* coarse-level T: 0 texture mean, 1 cubic spline */
    int estimator = 0;                      /* band pdf: 0 nearest row (reference), 1 exact at Tmax */
    std::vector<std::unique_ptr<SnapState>> snaps;
    std::vector<DevSnap> hsnap;             /* host copy of the descriptors */
    DevBuf dsnap;
    DevBuf de, ds, mean, sdev, smean, ssdev;
    DevBuf keys, stats, queue;
    DevBuf scan;                            /* a field's check flag and Tmax bits (rc_set_fields) */
    DevBuf bacc;                            /* per-band plain variant: FP32 cell sums */
    int reduction = 0;                      /* batched tallies: 0 FP64 atomics, 1 deterministic */
    DevBuf rayf, rayw;                      /* deterministic reduction: per-ray tallies */
    DevBuf srtk;                            /* batched, srt above 256 keys: radix-sort scratch */
    DevBuf tr_rays, tr_unit, tr_ulist, tr_out;
    rc_stats last{};
    bool launched = false;
    int wg_per_cu = 0, threshold = 0, spi = 0;   /* batched schedule; 0 = default */
    int refill = -1;                        /* refills between march steps: 1 on, 0 off, -1 default */
    bool use_graph = false;                 /* replay the batched run as a hipGraph */
    int table_bf16 = 0;                     /* k tables stored as BF16 */
    int graph_launches = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    void drop_graph() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        graph_exec = nullptr; graph = nullptr;
    }
};

namespace {

int build_tableset(rc_ctx *c, const Tables &t, TableSet &ts, std::vector<int32_t> &idx, std::vector<int32_t> &rank_of) {
    const rc_params &p = c->p;
    const int nB = t.nB, nT = t.nT, nQ = t.nQ;
    idx.resize(nB); rank_of.resize(nB);
    if (p.srt == 1 && !p.grey) sort_idx(t.kavg.data(), nB, idx.data());
    else for (int b = 0; b < nB; ++b) idx[b] = b;
    for (int r = 0; r < nB; ++r) rank_of[idx[r]] = r;
    ts.Tnb = t.Tnb;
    const float dT = t.Tnb[1] - t.Tnb[0];
    ts.Tnb0 = t.Tnb[0];
    ts.invdT = 1.0f / dT;
    const bool gt = p.grey && p.grey_true;
    /* Ib of the fixed-temperature walls per face and band (black_copy,
     * memory_copy_noint.cu:27-87): the rows' fixed-wall entries */
    std::vector<float> bfix((size_t)6 * nB);
    for (int f = 0; f < 6; ++f)
        for (int r = 0; r < nB; ++r) {
            const double Tw = p.Tw[f];
            bfix[(size_t)f * nB + r] = gt ? (float)(kStefan * Tw * Tw * Tw * Tw / kPi) : planck(p.Tw[f], t.wvc[idx[r]]);
        }
    const int S = tab_stride(nT);
    std::vector<float2> kb((size_t)nB * nQ * S);
    std::vector<float> brow(nT + 1);
    for (int r = 0; r < nB; ++r) {
        const int b = idx[r];
        for (int q = 0; q < nT; ++q)
            brow[q] = gt ? (float)(kStefan * std::pow((double)t.Tnb[q], 4) / kPi) : planck(t.Tnb[q], t.wvc[b]);
        brow[nT] = brow[nT - 1];
        for (int g = 0; g < nQ; ++g) {
            float2 *row = &kb[((size_t)r * nQ + g) * S];
            for (int q = 0; q < nT; ++q) row[q] = make_float2(t.kq[((size_t)b * nT + q) * nQ + g], brow[q]);
            row[nT] = row[nT - 1];
            for (int f = 0; f < 6; ++f)
                row[nT + 1 + 2 * f] = row[nT + 2 + 2 * f] = make_float2(0.f, bfix[(size_t)f * nB + r]);
        }
    }
    ts.bf16 = c->table_bf16;
    if (ts.bf16) {
        /* k as BF16 pairs with a power-of-two scale per band: 2^floor(log2 max k);
         * the fixed-wall entries hold k = 0 and the wall's Ib pair */
        std::vector<uint32_t> kp((size_t)nB * nQ * S, 0u);
        std::vector<float4> bp((size_t)nB * S, make_float4(0.f, 0.f, 0.f, 0.f));
        for (int r = 0; r < nB; ++r) {
            float mx = 0.f;
            for (int g = 0; g < nQ; ++g)
                for (int q = 0; q < nT; ++q) mx = std::max(mx, kb[((size_t)r * nQ + g) * S + q].x);
            int ex = 0;
            (void)std::frexp(mx > 0.f ? mx : 1.f, &ex);
            const float sc = std::ldexp(1.0f, ex - 1), inv = std::ldexp(1.0f, 1 - ex);
            for (int g = 0; g < nQ; ++g) {
                const float2 *row = &kb[((size_t)r * nQ + g) * S];
                for (int q = 0; q < nT; ++q)
                    kp[((size_t)r * nQ + g) * S + q] =
                        (uint32_t)to_bf16(row[q].x * inv) | ((uint32_t)to_bf16(row[q + 1].x * inv) << 16);
            }
            const float2 *row0 = &kb[(size_t)r * nQ * S];
            for (int q = 0; q < nT; ++q) bp[(size_t)r * S + q] = make_float4(row0[q].y, row0[q + 1].y, sc, 0.f);
            for (int f = 0; f < 6; ++f) {
                const float w = bfix[(size_t)f * nB + r];
                bp[(size_t)r * S + nT + 1 + 2 * f] = bp[(size_t)r * S + nT + 2 + 2 * f] = make_float4(w, w, sc, 0.f);
            }
        }
        if (int e = upload(ts.kp16, kp, c->stream)) return e;
        if (int e = upload(ts.bp, bp, c->stream)) return e;
    }
    std::vector<float> cgs(nQ);
    {
        std::vector<float> wq = t.wq;
        if ((int)wq.size() != nQ) wq = reference_wq(nQ);
        if ((int)wq.size() != nQ) { set_error("no quadrature weights for nQ"); return -1; }
        float acc = 0.f;
        for (int g = 0; g < nQ; ++g) { acc = (g == 0) ? wq[0] : acc + wq[g]; cgs[g] = acc; }
        for (int g = 0; g < nQ; ++g) cgs[g] = cgs[g] / acc;
    }
    if (int e = upload(ts.kb, kb, c->stream)) return e;
    if (int e = upload(ts.cdf_gs, cgs, c->stream)) return e;
    if (int e = upload(ts.rank_of, rank_of, c->stream)) return e;
    RC_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

/* Tmax, kappamax, Ibmax and the band-pdf row of one field (mc_gpu_noint.cu
 * :696-710, read_tables.cpp:35-38, mc_gpu_noint.cu:74, :1110-1131) */
int snapshot_scalars(const rc_params &p, const Tables &t, SnapState &s) {
    const float dT = t.Tnb[1] - t.Tnb[0];
    s.tm = (int)((s.Tmax - t.Tnb[0]) / dT);
    if (!(s.Tmax >= t.Tnb[0]) || s.tm < 0 || s.tm + 1 >= t.nT) {
        set_error("Tmax " + std::to_string(s.Tmax) + " K lies outside the temperature table [" +
                  std::to_string(t.Tnb[0]) + ", " + std::to_string(t.Tnb[t.nT - 1]) + "]");
        return -5;
    }
    const int tm = s.tm;
    if (p.grey) { s.kappamax = p.abscoeff; s.kappamax64 = p.abscoeff; }
    else {
        s.kappamax = (t.kP[tm + 1] - t.kP[tm]) / (t.Tnb[tm + 1] - t.Tnb[tm]) * (s.Tmax - t.Tnb[tm]) + t.kP[tm];
        s.kappamax64 = ((double)t.kP[tm + 1] - t.kP[tm]) / ((double)t.Tnb[tm + 1] - t.Tnb[tm]) *
                       ((double)s.Tmax - t.Tnb[tm]) + (double)t.kP[tm];
    }
    const double T4 = (double)s.Tmax * s.Tmax * s.Tmax * s.Tmax;
    s.Ibmax = (float)(4.0 * (double)s.kappamax * T4 * kStefan / p.photmax);
    s.Ibmax64 = 4.0 * s.kappamax64 * T4 * kStefan / p.photmax;
    s.Ibmax_w64 = (double)p.eps[FW] * kStefan * T4 / p.photmax;
    s.Ibmax_e64 = (double)p.eps[FE] * kStefan * T4 / p.photmax;
    s.Ibmax_w = (float)s.Ibmax_w64;
    s.Ibmax_e = (float)s.Ibmax_e64;
    s.tstar = ((s.Tmax - t.Tnb[tm]) < (t.Tnb[tm + 1] - s.Tmax)) ? tm : tm + 1;
    return 0;
}

/* the estimator's exact mode: the band/g pdf and its normalisation taken
 * from the tables interpolated at Tmax itself, instead of the nearest row's
 * CDFs and the tabulated Planck mean (the reference's choice, whose pdf and
 * weights disagree away from tabulated temperatures, and whose kappa_max
 * need not match the CDF normalisation).  With the emission of (band b,
 * g) at Tmax q_bg = K_bg(Tmax) B_b(Tmax) w_g dnu_b (linear interpolation in
 * T, as the weights use) and Z = sum q_bg: p(b, g) = q_bg / Z and
 * kappa_max = pi Z / (sigma Tmax^4), so that photmax rays carry the cell's
 * emission 4 pi sum K B w dnu exactly.  Walls: p(b) from B_b(Tmax) dnu_b and
 * Ibmax_s = eps pi Zs / photmax, the banded emission rather than eps sigma
 * T^4.  Sums in FP64 in band order, rounded once. */
int snapshot_exact(const rc_params &p, const Tables &t, const std::vector<float> &kq, SnapState &s,
                   std::vector<float> &cb, std::vector<float> &cg, std::vector<float> &cs) {
    const int nB = t.nB, nQ = t.nQ, nT = t.nT, tm = s.tm;
    const double f = ((double)s.Tmax - (double)t.Tnb[tm]) / ((double)t.Tnb[tm + 1] - (double)t.Tnb[tm]);
    std::vector<double> qb(nB), qbg((size_t)nB * nQ), bs(nB);
    double Z = 0.0, Zs = 0.0;
    for (int b = 0; b < nB; ++b) {
        const double b0 = planck(t.Tnb[tm], t.wvc[b]), b1 = planck(t.Tnb[tm + 1], t.wvc[b]);
        const double B = b0 + f * (b1 - b0);
        const double dnu = ((double)t.wvr[b] - (double)t.wvl[b]) * 100.0;
        double sb = 0.0;
        for (int g = 0; g < nQ; ++g) {
            const double k0 = kq[((size_t)b * nT + tm) * nQ + g], k1 = kq[((size_t)b * nT + tm + 1) * nQ + g];
            const double q = (k0 + f * (k1 - k0)) * B * (double)t.wq[g] * dnu;
            qbg[(size_t)b * nQ + g] = q;
            sb += q;
        }
        qb[b] = sb;
        Z += sb;
        bs[b] = B * dnu;
        Zs += bs[b];
    }
    if (!(Z > 0.0) || !(Zs > 0.0) || !std::isfinite(Z) || !std::isfinite(Zs)) {
        set_error("exact estimator: the tables give no (or no finite) emission at Tmax");
        return -6;
    }
    double run = 0.0, runs = 0.0;
    for (int b = 0; b < nB; ++b) {
        run += qb[b];
        runs += bs[b];
        cb[b] = (float)(run / Z);
        cs[b] = (float)(runs / Zs);
        double rg = 0.0;
        for (int g = 0; g < nQ; ++g) {
            rg += qbg[(size_t)b * nQ + g];
            cg[(size_t)b * nQ + g] = qb[b] > 0.0 ? (float)(rg / qb[b]) : (float)(g + 1) / (float)nQ;
        }
    }
    cb[nB - 1] = 1.0f;
    cs[nB - 1] = 1.0f;
    const double T4 = (double)s.Tmax * s.Tmax * s.Tmax * s.Tmax;
    s.kappamax64 = kPi * Z / (kStefan * T4);
    s.kappamax = (float)s.kappamax64;
    s.Ibmax64 = 4.0 * kPi * Z / p.photmax;
    s.Ibmax = (float)s.Ibmax64;
    s.Ibmax_w64 = (double)p.eps[FW] * kPi * Zs / p.photmax;
    s.Ibmax_e64 = (double)p.eps[FE] * kPi * Zs / p.photmax;
    s.Ibmax_w = (float)s.Ibmax_w64;
    s.Ibmax_e = (float)s.Ibmax_e64;
    return 0;
}

/* the k-distribution values as the march reads them: with BF16 tables the
 * values build_tableset stores (per band a power-of-two scale 2^(e-1) of
 * the band's largest k, the scaled k rounded to BF16 nearest-even), so that
 * the exact estimator's emission uses the k of the march */
std::vector<float> kq_as_marched(const rc_ctx *c, const Tables &t) {
    if (!c->table_bf16) return t.kq;
    const int nB = t.nB, nT = t.nT, nQ = t.nQ;
    std::vector<float> out(t.kq.size());
    for (int b = 0; b < nB; ++b) {
        float mx = 0.f;
        for (size_t i = (size_t)b * nT * nQ; i < (size_t)(b + 1) * nT * nQ; ++i) mx = std::max(mx, t.kq[i]);
        int ex = 0;
        (void)std::frexp(mx > 0.f ? mx : 1.f, &ex);
        const float sc = std::ldexp(1.0f, ex - 1), inv = std::ldexp(1.0f, 1 - ex);
        for (size_t i = (size_t)b * nT * nQ; i < (size_t)(b + 1) * nT * nQ; ++i) {
            const uint32_t u = (uint32_t)to_bf16(t.kq[i] * inv) << 16;
            float v;
            std::memcpy(&v, &u, 4);
            out[i] = v * sc;
        }
    }
    return out;
}

int snapshot_rows(rc_ctx *c, const Tables &t, SnapState &s) {
    const rc_params &p = c->p;
    const int nB = t.nB, nQ = t.nQ, ts = s.tstar;
    std::vector<float> cb(t.prob.begin() + (size_t)ts * nB, t.prob.begin() + (size_t)(ts + 1) * nB);
    std::vector<float> cg(t.probg.begin() + (size_t)ts * nB * nQ, t.probg.begin() + (size_t)(ts + 1) * nB * nQ);
    /* wall band CDF: readS (read_tables.cpp:188-220) on row t* */
    std::vector<float> cs(nB);
    const bool gt = p.grey && p.grey_true;
    const float Tr = t.Tnb[ts];
    float acc = 0.f;
    for (int b = 0; b < nB; ++b) {
        const float ib = gt ? 1.0f : planck(Tr, t.wvc[b]);
        acc = (b == 0) ? ib * (t.wvr[0] - t.wvl[0]) * 100.f : acc + ib * (t.wvr[b] - t.wvl[b]) * 100.f;
        cs[b] = acc;
    }
    for (int b = 0; b < nB; ++b) cs[b] = cs[b] / acc;
    if (c->estimator == 1 && !p.grey)
        if (int e = snapshot_exact(p, t, kq_as_marched(c, t), s, cb, cg, cs)) return e;
    /* guide tables: g[m] = first index with cdf > m/M, m = 0..M */
    auto guide = [&](const std::vector<float> &cdf) {
        const int M = kGuideM;
        std::vector<int32_t> g(M + 1);
        for (int q = 0; q <= M; ++q) {
            const float thr = (float)q / (float)M;
            g[q] = (int32_t)(std::upper_bound(cdf.begin(), cdf.end(), thr) - cdf.begin());
        }
        return g;
    };
    if (int e = upload(s.guide_b, guide(cb), c->stream)) return e;
    if (int e = upload(s.guide_bs, guide(cs), c->stream)) return e;
    if (int e = upload(s.cdf_b, cb, c->stream)) return e;
    if (int e = upload(s.cdf_g, cg, c->stream)) return e;
    if (int e = upload(s.cdf_bs, cs, c->stream)) return e;
    return 0;
}

DevCommon make_common(const rc_ctx *c) {
    const rc_params &p = c->p;
    DevCommon cm{};
    cm.imax = p.imax; cm.jmax = p.jmax; cm.kmax = p.kmax; cm.kloc = c->kloc; cm.zstart = c->zstart;
    cm.nVar = p.nVar; cm.photmax = p.photmax; cm.srt = (p.srt == 1 && !p.grey) ? 1 : 0;
    cm.nB = p.nB; cm.nQ = p.nQ; cm.nT = p.nT; cm.grid_num = p.grid_num;
    cm.ncell = c->ncell; cm.nsurf = c->nsurf;
    cm.Lx = p.Lx; cm.Ly = p.Ly; cm.Lz = p.Lz; cm.toll = p.toll; cm.tollS = p.tollS;
    for (int f = 0; f < 6; ++f) cm.bd[f] = p.bd[f];
    cm.adiab[0] = p.adiabw == 1; cm.adiab[1] = p.adiabe == 1;
    cm.k0 = (uint32_t)p.seed; cm.k1 = (uint32_t)(p.seed >> 32);
    cm.ray_de = nullptr;
    cm.sampler = p.sampler == 1 ? 1 : 0;
    cm.refill = c->refill;
    cm.maxs0 = p.maxs[0];
    return cm;
}

/* push the DevSnap descriptors */
int push_snaps(rc_ctx *c) {
    const rc_params &p = c->p;
    std::vector<DevSnap> v(c->snaps.size());
    for (size_t s = 0; s < c->snaps.size(); ++s) {
        SnapState &S = *c->snaps[s];
        const TableSet &ts = S.grey ? *S.grey : c->nbset;
        DevSnap d{};
        d.Tmax = S.Tmax; d.Ibmax = S.Ibmax; d.Ibmax_w = S.Ibmax_w; d.Ibmax_e = S.Ibmax_e;
        d.Tnb0 = ts.Tnb0; d.invdT = ts.invdT; d.tstar = S.tstar; d.snap = (int)s;
        d.kb = ts.kb.as<float2>();
        d.kbf16 = ts.bf16;
        d.kp16 = ts.kp16.as<uint32_t>(); d.bp = ts.bp.as<float4>();
        d.cdf_b = S.cdf_b.as<float>(); d.cdf_g = S.cdf_g.as<float>(); d.cdf_bs = S.cdf_bs.as<float>();
        d.cdf_gs = ts.cdf_gs.as<float>(); d.rank_of = ts.rank_of.as<int>();
        d.guide_b = S.guide_b.as<int>(); d.guide_bs = S.guide_bs.as<int>(); d.gM = kGuideM;
        d.T0 = S.T0.as<float>();
        d.xu_all = c->xuall.as<float>();
        d.P_all = S.Pall.as<float>();
        d.T_all = S.Tall.as<float>();
        d.Ibmax64 = S.Ibmax64; d.Ibmax_w64 = S.Ibmax_w64; d.Ibmax_e64 = S.Ibmax_e64;
        d.Tnb0d = (double)ts.Tnb[0];
        d.invdTd = 1.0 / (double)(ts.Tnb[1] - ts.Tnb[0]);
        d.nxu = c->nxu;
        d.npos = (int)c->npos;
        d.x0 = c->x0.as<float>(); d.y0 = c->y0.as<float>(); d.z0 = c->z0.as<float>();
        d.imax = p.imax; d.jmax = p.jmax; d.kmax = p.kmax; d.zstart = c->zstart;
        d.ncell = c->ncell; d.nsurf = c->nsurf;
        d.ubase = (int)s * c->ncell; d.sbase = (int)s * 2 * c->nsurf;
        for (int g = 0; g < p.grid_num; ++g) {
            DevLevel &L = d.lev[g];
            L.im = p.maxi[g]; L.jm = p.maxj[g]; L.km = p.maxk[g]; L.maxs = p.maxs[g];
            L.ncell = L.im * L.jm * L.km;
            L.dy = p.Ly * 1.0f / (float)L.jm;
            L.dz = p.Lz * 1.0f / (float)L.km;
            L.xoff = c->xoff[g];
            L.xu = c->xuall.as<float>() + c->xoff[g];
            L.poff = (int)c->poff[g];
            L.P = S.Pall.as<float>() + c->poff[g];
            /* coarsening to level g + 1: ratios and their multipliers
             * m = ceil(2^32 / r), exact for every index n of the level
             * when n (m r - 2^32) < 2^32 (else the kernel divides) */
            const int dims[3] = {p.maxi[g], p.maxj[g], p.maxk[g]};
            for (int a = 0; a < 3; ++a) {
                const int next = g + 1 < p.grid_num ? (a == 0 ? p.maxi[g + 1] : a == 1 ? p.maxj[g + 1] : p.maxk[g + 1])
                                                    : dims[a];
                const int r = dims[a] / next;
                L.cr[a] = r;
                L.cmul[a] = 0u;
                if (r >= 2) {
                    const uint64_t m = (((uint64_t)1 << 32) + (uint64_t)r - 1) / (uint64_t)r;
                    const uint64_t e = m * (uint64_t)r - ((uint64_t)1 << 32);
                    if ((uint64_t)(dims[a] + r - 1) * e < ((uint64_t)1 << 32)) L.cmul[a] = (uint32_t)m;
                }
            }
        }
        v[s] = d;
    }
    c->hsnap = v;
    return upload(c->dsnap, c->hsnap, c->stream);
}

}  // namespace

/* ========================================================================= */
extern "C" {

void rc_default_params(rc_params *p) { default_params(*p); }
const char *rc_last_error(void) { return last_error(); }
const char *rc_version(void) { return "radchan-mi355x 0.1 (gfx950)"; }

int rc_read_tables(const char *dir, double chanHeight, rc_tables *t) {
    Tables tt;
    tt.resize(t->nB, t->nT, t->nQ);
    if (int e = read_tables(dir, chanHeight, tt)) return e;
    std::memcpy(t->Tnb, tt.Tnb.data(), tt.Tnb.size() * 4);
    std::memcpy(t->kP, tt.kP.data(), tt.kP.size() * 4);
    std::memcpy(t->wvl, tt.wvl.data(), tt.wvl.size() * 4);
    std::memcpy(t->wvr, tt.wvr.data(), tt.wvr.size() * 4);
    std::memcpy(t->wvc, tt.wvc.data(), tt.wvc.size() * 4);
    std::memcpy(t->kavg, tt.kavg.data(), tt.kavg.size() * 4);
    std::memcpy(t->kq, tt.kq.data(), tt.kq.size() * 4);
    std::memcpy(t->prob, tt.prob.data(), tt.prob.size() * 4);
    std::memcpy(t->probg, tt.probg.data(), tt.probg.size() * 4);
    return 0;
}

static void copy_in(Tables &tt, const rc_tables *t) {
    tt.resize(t->nB, t->nT, t->nQ);
    std::memcpy(tt.Tnb.data(), t->Tnb, tt.Tnb.size() * 4);
    std::memcpy(tt.kP.data(), t->kP, tt.kP.size() * 4);
    std::memcpy(tt.wvl.data(), t->wvl, tt.wvl.size() * 4);
    std::memcpy(tt.wvr.data(), t->wvr, tt.wvr.size() * 4);
    std::memcpy(tt.wvc.data(), t->wvc, tt.wvc.size() * 4);
    std::memcpy(tt.kavg.data(), t->kavg, tt.kavg.size() * 4);
    std::memcpy(tt.kq.data(), t->kq, tt.kq.size() * 4);
    std::memcpy(tt.prob.data(), t->prob, tt.prob.size() * 4);
    std::memcpy(tt.probg.data(), t->probg, tt.probg.size() * 4);
    if (t->wq) std::memcpy(tt.wq.data(), t->wq, tt.wq.size() * 4);
    else tt.wq = reference_wq(t->nQ);
}

int rc_write_tables(const char *dir, const rc_tables *t, double chanHeight) {
    Tables tt;
    copy_in(tt, t);
    return write_tables(dir, tt, chanHeight);
}

int rc_ck_tables(int nB, int nT, int nQ, const double *T, const double *wvl, const double *wvr, const double *kq,
                 const double *kavg, const double *wq, double chanHeight, rc_tables *t) {
    if (!t || !T || !wvl || !wvr || !kq || !kavg || !wq || nB < 1 || nT < 1 || nQ < 1 || nQ > 1024 ||
        t->nB != nB || t->nT != nT || t->nQ != nQ) {
        set_error("rc_ck_tables: null array, sizes out of range or tables of other sizes");
        return -1;
    }
    Tables tt;
    if (int e = ck_tables(nB, nT, nQ, T, wvl, wvr, kq, kavg, wq, chanHeight, tt)) return e;
    std::memcpy(t->Tnb, tt.Tnb.data(), tt.Tnb.size() * 4);
    std::memcpy(t->kP, tt.kP.data(), tt.kP.size() * 4);
    std::memcpy(t->wvl, tt.wvl.data(), tt.wvl.size() * 4);
    std::memcpy(t->wvr, tt.wvr.data(), tt.wvr.size() * 4);
    std::memcpy(t->wvc, tt.wvc.data(), tt.wvc.size() * 4);
    std::memcpy(t->kavg, tt.kavg.data(), tt.kavg.size() * 4);
    std::memcpy(t->kq, tt.kq.data(), tt.kq.size() * 4);
    std::memcpy(t->prob, tt.prob.data(), tt.prob.size() * 4);
    std::memcpy(t->probg, tt.probg.data(), tt.probg.size() * 4);
    if (t->wq) std::memcpy(t->wq, tt.wq.data(), tt.wq.size() * 4);
    return 0;
}

int rc_lbl_bands(int64_t n, const double *wv, const double *kv, double toll, double wvtoll, double dwv,
                 const double *deltawv, int max_bands, int64_t *first, int64_t *count, double *wv_left,
                 double *wv_right, double *wvc, int32_t *nwide) {
    if (n < 1 || !wv || !kv || !deltawv || max_bands < 0 || (max_bands > 0 && (!first || !count))) {
        set_error("rc_lbl_bands: empty spectrum, null array or negative max_bands");
        return -1;
    }
    for (int d = 0; d < 3; ++d)
        if (!(deltawv[d] > 0.0)) { set_error("rc_lbl_bands: deltawv must be positive"); return -1; }
    std::vector<int64_t> f, c;
    std::vector<double> l, r, m;
    int nw = 0;
    const int nb = lbl_bands(n, wv, kv, toll, wvtoll, dwv, deltawv, f, c, l, r, m, nw);
    if (nb < 0) return nb;
    if (nwide) *nwide = nw;
    const int k = std::min(nb, max_bands);
    for (int b = 0; b < k; ++b) {
        first[b] = f[b];
        count[b] = c[b];
        if (wv_left) wv_left[b] = l[b];
        if (wv_right) wv_right[b] = r[b];
        if (wvc) wvc[b] = m[b];
    }
    return nb;
}

int64_t rc_lbl_fill(int64_t n, const double *wv, const double *kv, double T, double dwv, int64_t ntot,
                    const double *wvn, double *kvn, double *kPr) {
    if (n < 1 || ntot < 1 || !wv || !kv || !wvn || !kvn || !kPr || !(T > 0.0)) {
        set_error("rc_lbl_fill: empty spectrum or bands, null array or T <= 0");
        return -1;
    }
    return lbl_fill(n, wv, kv, T, dwv, ntot, wvn, kvn, *kPr);
}

int rc_grey_tables(const rc_params *p, float Tmax, rc_tables *t) {
    Tables tt;
    tt.resize(t->nB, t->nT, t->nQ);
    if (int e = grey_tables(*p, Tmax, tt)) return e;
    std::memcpy(t->Tnb, tt.Tnb.data(), tt.Tnb.size() * 4);
    std::memcpy(t->kP, tt.kP.data(), tt.kP.size() * 4);
    std::memcpy(t->wvl, tt.wvl.data(), tt.wvl.size() * 4);
    std::memcpy(t->wvr, tt.wvr.data(), tt.wvr.size() * 4);
    std::memcpy(t->wvc, tt.wvc.data(), tt.wvc.size() * 4);
    std::memcpy(t->kavg, tt.kavg.data(), tt.kavg.size() * 4);
    std::memcpy(t->kq, tt.kq.data(), tt.kq.size() * 4);
    std::memcpy(t->prob, tt.prob.data(), tt.prob.size() * 4);
    std::memcpy(t->probg, tt.probg.data(), tt.probg.size() * 4);
    if (t->wq) t->wq[0] = 1.0f;
    return 0;
}

int rc_sort_idx(const float *kavg, int nB, int32_t *idx) {
    if (nB < 1) { set_error("nB must be positive"); return -1; }
    sort_idx(kavg, nB, idx);
    return 0;
}

int rc_grid(const rc_params *p, int level, float *x, float *xu, float *y, float *yu, float *z, float *zu) {
    if (level < 0 || level >= p->grid_num) { set_error("level out of range"); return -1; }
    Grid g;
    build_grid(*p, level, g);
    std::memcpy(x, g.x.data(), g.x.size() * 4);
    std::memcpy(xu, g.xu.data(), g.xu.size() * 4);
    std::memcpy(y, g.y.data(), g.y.size() * 4);
    std::memcpy(yu, g.yu.data(), g.yu.size() * 4);
    std::memcpy(z, g.z.data(), g.z.size() * 4);
    std::memcpy(zu, g.zu.data(), g.zu.size() * 4);
    return 0;
}

int rc_device_synchronize(void) {
    RC_HIP(hipDeviceSynchronize());
    return 0;
}

rc_ctx *rc_create(const rc_params *pp) {
    const Range range_("radchan: set-up (rc_create)");
    rc_params p = *pp;
    if (p.grey) p.srt = 0;                              /* NarrowBand.h:220-221 */
    if (check_params(p)) return nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        set_error("no HIP device available");
        return nullptr;
    }
    auto c = std::make_unique<rc_ctx>();
    c->p = p;
    c->kloc = p.kmax / p.p_col;
    c->ncell = p.imax * p.jmax * c->kloc;
    c->nsurf = p.jmax * c->kloc;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipEventCreate(&c->evm0) != hipSuccess || hipEventCreate(&c->evm1) != hipSuccess) {
        set_error("cannot create HIP stream/events");
        return nullptr;
    }
    c->grids.resize(p.grid_num);
    std::vector<float> xuall;
    for (int g = 0; g < p.grid_num; ++g) {
        build_grid(p, g, c->grids[g]);
        c->xoff[g] = (int)xuall.size();
        xuall.insert(xuall.end(), c->grids[g].xu.begin(), c->grids[g].xu.end());
        c->poff[g] = c->npos;
        c->npos += level_pos_count(p.maxi[g], p.maxj[g], p.maxk[g]);
        if (upload(c->cx[g], c->grids[g].x, c->stream) || upload(c->cy[g], c->grids[g].y, c->stream) ||
            upload(c->cz[g], c->grids[g].z, c->stream))
            return nullptr;
    }
    c->nxu = (int)xuall.size();
    if (c->npos >= ((size_t)1 << 31)) { set_error("grid too large: table positions of all levels beyond 2^31"); return nullptr; }
    if (upload(c->xuall, xuall, c->stream)) return nullptr;
    if (upload(c->x0, c->grids[0].x, c->stream) || upload(c->y0, c->grids[0].y, c->stream) ||
        upload(c->z0, c->grids[0].z, c->stream))
        return nullptr;
    if (c->stats.ensure(8 * sizeof(unsigned long long)) || c->queue.ensure(64 * sizeof(int))) return nullptr;
    if (hipStreamSynchronize(c->stream) != hipSuccess) { set_error("stream synchronize failed"); return nullptr; }
    return c.release();
}

void rc_destroy(rc_ctx *c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->evm0) (void)hipEventDestroy(c->evm0);
    if (c->evm1) (void)hipEventDestroy(c->evm1);
    c->drop_graph();
    hipStream_t st = c->stream;
    delete c;
    if (st) (void)hipStreamDestroy(st);
}

int rc_load_tables(rc_ctx *c, const rc_tables *t) {
    if (!c || !t) { set_error("null argument"); return -1; }
    if (c->p.grey) return 0;
    if (t->nB != c->p.nB || t->nT != c->p.nT || t->nQ != c->p.nQ) {
        set_error("table sizes differ from the parameters (nB, nT, nQ)");
        return -2;
    }
    copy_in(c->tab, t);
    for (int q = 1; q < c->tab.nT; ++q)
        if (!(c->tab.Tnb[q] > c->tab.Tnb[q - 1])) { set_error("temperature table is not increasing"); return -3; }
    if (int e = build_tableset(c, c->tab, c->nbset, c->idx, c->rank_of)) return e;
    c->have_tables = true;
    return 0;
}

int rc_set_fields(rc_ctx *c, int nsnap, const float *const *T, int zstart) {
    const Range range_("radchan: fields (rc_set_fields)");
    if (!c || nsnap < 1 || !T) { set_error("bad arguments to rc_set_fields"); return -1; }
    const rc_params &p = c->p;
    if (zstart < 1 || zstart + c->kloc - 1 > p.kmax) { set_error("zstart outside 1..kmax-kmax/p_col+1"); return -2; }
    if (!p.grey && !c->have_tables) { set_error("narrow-band run without tables (rc_load_tables)"); return -3; }
    /* (snapshot, cell, run) units and their tallies are numbered in 32 bits */
    if ((int64_t)nsnap * std::max<int64_t>(c->ncell, 2 * (int64_t)c->nsurf) * p.nVar >= ((int64_t)1 << 31)) {
        set_error("too many (snapshot, cell, run) units for one launch (2^31): fewer snapshots or runs");
        return -5;
    }
    if (c->launched) { (void)hipStreamSynchronize(c->stream); c->launched = false; }
    c->drop_graph();
    c->zstart = zstart;
    const size_t nT0 = (size_t)(p.imax + 2) * (p.jmax + 2) * (p.kmax + 2);
    c->snaps.resize(nsnap);
    for (int s = 0; s < nsnap; ++s) {
        if (!c->snaps[s]) c->snaps[s] = std::make_unique<SnapState>();
        SnapState &S = *c->snaps[s];
        /* the field to the device, checked and its Tmax found there */
        if (int e = S.T0.ensure(nT0 * 4)) return e;
        if (int e = c->scan.ensure(2 * sizeof(uint32_t))) return e;
        RC_HIP(hipMemcpyAsync(S.T0.p, T[s], nT0 * 4, hipMemcpyHostToDevice, c->stream));
        launch_field_scan(c->stream, S.T0.as<float>(), p.imax, p.jmax, p.kmax, c->scan.as<uint32_t>());
        uint32_t scan[2] = {0u, 0u};
        RC_HIP(hipMemcpyAsync(scan, c->scan.p, sizeof scan, hipMemcpyDeviceToHost, c->stream));
        RC_HIP(hipStreamSynchronize(c->stream));
        if (scan[0]) {
            set_error("temperature field " + std::to_string(s) + " holds a non-positive or non-finite value");
            return -4;
        }
        std::memcpy(&S.Tmax, &scan[1], 4);
        const Tables *tb = &c->tab;
        if (p.grey) {
            S.greyhost.resize(p.nB, p.nT, 1);
            if (int e = grey_tables(p, S.Tmax, S.greyhost)) return e;
            if (!S.grey) S.grey = std::make_unique<TableSet>();
            std::vector<int32_t> idx, ro;
            if (int e = build_tableset(c, S.greyhost, *S.grey, idx, ro)) return e;
            tb = &S.greyhost;
        }
        if (int e = snapshot_scalars(p, *tb, S)) return e;
        if (int e = snapshot_rows(c, *tb, S)) return e;
        const TableSet &tset = S.grey ? *S.grey : c->nbset;
        if (int e = S.Pall.ensure(c->npos * 4)) return e;
        const bool f64 = p.fp64 == 1;
        if (f64)
            if (int e = S.Tall.ensure(c->npos * 4)) return e;
        for (int g = 0; g < p.grid_num; ++g) {
            const size_t nl = (size_t)p.maxi[g] * p.maxj[g] * p.maxk[g];
            float *Pg = S.Pall.as<float>() + c->poff[g];
            float *Tg = f64 ? S.Tall.as<float>() + c->poff[g] : nullptr;
            launch_level_P(c->stream, S.T0.as<float>(), p, g, tset.Tnb0, tset.invdT, p.nT, Pg, Tg);
            if (c->restriction == 1 && g > 0) {
                /* cells and adiabatic x-wall faces by cubic splines (then to
                 * positions); bd=4 z walls keep the texture mean */
                const bool wx = (p.bd[FW] == 2 && p.adiabw == 1) || (p.bd[FE] == 2 && p.adiabe == 1);
                const size_t nw = (size_t)2 * p.maxj[g] * p.maxk[g];
                const size_t ns = spline_scratch_floats(p.imax, p.jmax, p.kmax, p.maxi[g], p.maxj[g], p.maxk[g]);
                if (int e = c->spl.ensure((ns + nl + nw) * 4)) return e;
                float *tl = c->spl.as<float>() + ns, *tw = tl + nl;
                launch_spline_level(c->stream, S.T0.as<float>(), c->cx[0].as<float>(), c->cy[0].as<float>(),
                                    c->cz[0].as<float>(), c->cx[g].as<float>(), c->cy[g].as<float>(),
                                    c->cz[g].as<float>(), p.imax, p.jmax, p.kmax, p.maxi[g], p.maxj[g], p.maxk[g],
                                    tl, wx ? tw : nullptr, c->spl.as<float>());
                launch_pos_from_T(c->stream, p, g, tl, wx ? tw : nullptr, tset.Tnb0, tset.invdT, p.nT, Pg, Tg);
            }
        }
        RC_HIP(hipGetLastError());
    }
    if (int e = push_snaps(c)) return e;
    const size_t nde = (size_t)nsnap * c->ncell * p.nVar;
    const size_t nds = (size_t)nsnap * 2 * c->nsurf * p.nVar;
    if (int e = c->de.ensure(nde * 8)) return e;
    if (int e = c->ds.ensure(nds * 8)) return e;
    if (int e = c->mean.ensure((size_t)nsnap * c->ncell * 4)) return e;
    if (int e = c->sdev.ensure((size_t)nsnap * c->ncell * 4)) return e;
    if (int e = c->smean.ensure((size_t)nsnap * 2 * c->nsurf * 4)) return e;
    if (int e = c->ssdev.ensure((size_t)nsnap * 2 * c->nsurf * 4)) return e;
    RC_HIP(hipMemsetAsync(c->ds.p, 0, nds * 8, c->stream));
    RC_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

/* the device work of one solve (no host synchronisation), so that it can
 * also be recorded into a graph; returns the launch count.  With `timed`,
 * the march events bracket the march kernels alone (never inside a graph). */
static int enqueue_solve(rc_ctx *c, int variant, int *launches, bool timed) {
    const rc_params &p = c->p;
    const DevCommon cm = make_common(c);
    const int nsnap = (int)c->snaps.size();
    const DevSnap *ds = c->dsnap.as<DevSnap>();
    const bool walls = p.calcwalls == 1 && (p.adiabw == 1 || p.adiabe == 1);
    int n = 0;
    launch_zero_u64(c->stream, c->stats.as<unsigned long long>(), 8);
    ++n;
    if (timed) RC_HIP(hipEventRecord(c->evm0, c->stream));
    if (variant == RC_VARIANT_PLAIN || variant == RC_VARIANT_PLAIN_BANDS) {
        for (int s = 0; s < nsnap; ++s) {
            double *de = c->de.as<double>() + (size_t)s * c->ncell * p.nVar;
            for (int v = 0; v < p.nVar; ++v) {
                if (cm.srt) { launch_sort(c->stream, cm, ds, s, v, 0, c->ncell, nullptr, c->keys.as<uint16_t>()); ++n; }
                if (variant == RC_VARIANT_PLAIN_BANDS) {
                    launch_fluid_bands(c->stream, cm, ds, s, v, c->keys.as<uint16_t>(), c->bacc.as<float>(), de,
                                       c->stats.as<unsigned long long>());
                    n += p.nB + 2;
                } else if (p.fp64 == 1) {
                    launch_fluid_f64(c->stream, cm, ds, s, v, cm.srt ? c->keys.as<uint16_t>() : nullptr, de,
                                     c->stats.as<unsigned long long>());
                    ++n;
                } else {
                    launch_fluid_plain(c->stream, cm, ds, s, v, cm.srt ? c->keys.as<uint16_t>() : nullptr, de,
                                       c->stats.as<unsigned long long>());
                    ++n;
                }
            }
            if (walls) {
                double *dw = c->ds.as<double>() + (size_t)s * 2 * c->nsurf * p.nVar;
                for (int v = 0; v < p.nVar; ++v) {
                    if (cm.srt) { launch_sort(c->stream, cm, ds, s, v, 1, 2 * c->nsurf, nullptr, c->keys.as<uint16_t>()); ++n; }
                    if (p.fp64 == 1)
                        launch_surface_f64(c->stream, cm, ds, s, v, cm.srt ? c->keys.as<uint16_t>() : nullptr, dw,
                                           c->stats.as<unsigned long long>());
                    else
                        launch_surface_plain(c->stream, cm, ds, s, v, cm.srt ? c->keys.as<uint16_t>() : nullptr,
                                             dw, c->stats.as<unsigned long long>());
                    ++n;
                }
            }
        }
    } else {
        DevCommon cf = cm, cw = cm;
        cf.srt_keys = cw.srt_keys = c->srtk.as<uint16_t>();
        if (c->reduction == 1) {
            cf.ray_de = c->rayf.as<float>();
            cw.ray_de = c->rayw.as<float>();
        }
        launch_batched(c->stream, cf, ds, nsnap, nsnap * c->ncell, 0, c->table_bf16, c->de.as<double>(),
                       c->stats.as<unsigned long long>(), c->queue.as<int>(), c->wg_per_cu, c->threshold, c->spi);
        n += 3;
        if (c->reduction == 1) {
            launch_reduce_rays(c->stream, cf.ray_de, (size_t)nsnap * c->ncell * p.nVar, p.photmax, c->de.as<double>());
            ++n;
        }
        if (walls) {
            launch_batched(c->stream, cw, ds, nsnap, nsnap * 2 * c->nsurf, 1, c->table_bf16, c->ds.as<double>(),
                           c->stats.as<unsigned long long>(), c->queue.as<int>() + 16, c->wg_per_cu, c->threshold,
                           c->spi);
            n += 3;
            if (c->reduction == 1) {
                launch_reduce_rays(c->stream, cw.ray_de, (size_t)nsnap * 2 * c->nsurf * p.nVar, p.photmax,
                                   c->ds.as<double>());
                ++n;
            }
        }
    }
    if (timed) RC_HIP(hipEventRecord(c->evm1, c->stream));
    for (int s = 0; s < nsnap; ++s) {
        launch_finalize(c->stream, c->de.as<double>() + (size_t)s * c->ncell * p.nVar, c->ncell, p.nVar,
                        c->mean.as<float>() + (size_t)s * c->ncell, c->sdev.as<float>() + (size_t)s * c->ncell);
        ++n;
        if (walls) {
            launch_finalize(c->stream, c->ds.as<double>() + (size_t)s * 2 * c->nsurf * p.nVar, 2 * c->nsurf, p.nVar,
                            c->smean.as<float>() + (size_t)s * 2 * c->nsurf,
                            c->ssdev.as<float>() + (size_t)s * 2 * c->nsurf);
            ++n;
        }
    }
    *launches = n;
    RC_HIP(hipGetLastError());
    return 0;
}

int rc_run(rc_ctx *c, int variant) {
    const Range range_("radchan: launch (rc_run)");
    if (!c || c->snaps.empty()) { set_error("rc_run before rc_set_fields"); return -1; }
    if (variant != RC_VARIANT_PLAIN && variant != RC_VARIANT_BATCHED && variant != RC_VARIANT_PLAIN_BANDS) {
        set_error("unknown variant");
        return -2;
    }
    const rc_params &p = c->p;
    if (p.fp64 == 1 && variant != RC_VARIANT_PLAIN) {
        set_error("floatpoint 1 (double precision) runs in the plain variant (RC_VARIANT_PLAIN)");
        return -2;
    }
    if (variant == RC_VARIANT_PLAIN_BANDS && !(p.srt == 1 && !p.grey)) {
        set_error("the per-band plain variant needs sorted band keys (srt=1, narrow band)");
        return -2;
    }
    if ((variant == RC_VARIANT_PLAIN || variant == RC_VARIANT_PLAIN_BANDS) && p.srt == 1 && !p.grey) {
        const size_t nk = (size_t)std::max(c->ncell, 2 * c->nsurf) * p.photmax;
        if (int e = c->keys.ensure(nk * 2)) return e;
    }
    if (variant == RC_VARIANT_PLAIN_BANDS)
        if (int e = c->bacc.ensure((size_t)c->ncell * 4)) return e;
    if (variant == RC_VARIANT_BATCHED) {
        /* radix-sort scratch for the larger of the fluid and wall launches;
         * a new buffer invalidates a captured graph */
        const DevCommon cm = make_common(c);
        const long ns = (long)c->snaps.size();
        const size_t nk = std::max(batched_srt_keys(cm, ns * c->ncell, c->wg_per_cu),
                                   p.calcwalls == 1 ? batched_srt_keys(cm, ns * 2 * c->nsurf, c->wg_per_cu) : 0);
        if (nk) {
            void *old = c->srtk.p;
            if (int e = c->srtk.ensure(nk * 2)) return e;
            if (c->srtk.p != old) c->drop_graph();
        }
    }
    if (variant == RC_VARIANT_BATCHED && c->reduction == 1) {
        const size_t ns = c->snaps.size();
        if (int e = c->rayf.ensure(ns * c->ncell * p.nVar * (size_t)p.photmax * 4)) return e;
        if (p.calcwalls == 1)
            if (int e = c->rayw.ensure(ns * 2 * c->nsurf * p.nVar * (size_t)p.photmax * 4)) return e;
    }
    int launches = 0;
    const bool graph = c->use_graph && variant == RC_VARIANT_BATCHED;
    if (graph && !c->graph_exec) {
        RC_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        const int e = enqueue_solve(c, variant, &launches, false);
        hipGraph_t g = nullptr;
        const hipError_t ce = hipStreamEndCapture(c->stream, &g);
        if (e) { if (g) (void)hipGraphDestroy(g); return e; }
        if (ce != hipSuccess) { set_error(std::string("graph capture failed: ") + hipGetErrorString(ce)); return -12; }
        c->graph = g;
        RC_HIP(hipGraphInstantiate(&c->graph_exec, c->graph, nullptr, nullptr, 0));
        c->graph_launches = launches;
    }
    RC_HIP(hipEventRecord(c->ev0, c->stream));
    if (graph) {
        /* a replayed graph is timed as a whole */
        RC_HIP(hipEventRecord(c->evm0, c->stream));
        RC_HIP(hipGraphLaunch(c->graph_exec, c->stream));
        RC_HIP(hipEventRecord(c->evm1, c->stream));
        launches = c->graph_launches;
    } else if (int e = enqueue_solve(c, variant, &launches, true)) {
        return e;
    }
    RC_HIP(hipEventRecord(c->ev1, c->stream));
    c->last = rc_stats{};
    c->last.launches = launches;
    c->last.variant = variant;
    c->launched = true;
    return 0;
}

int rc_set_table_precision(rc_ctx *c, int bf16) {
    if (!c || (bf16 != 0 && bf16 != 1)) { set_error("table precision: 0 (FP32) or 1 (BF16)"); return -1; }
    if (c->launched) (void)rc_sync(c, nullptr);
    c->table_bf16 = bf16;
    c->drop_graph();
    if (c->have_tables)
        if (int e = build_tableset(c, c->tab, c->nbset, c->idx, c->rank_of)) return e;
    for (auto &sp : c->snaps)
        if (sp && sp->grey) {
            std::vector<int32_t> idx, ro;
            if (int e = build_tableset(c, sp->greyhost, *sp->grey, idx, ro)) return e;
        }
    /* the exact estimator's emission follows the k the march reads */
    if (c->estimator == 1 && !c->p.grey)
        for (auto &sp : c->snaps)
            if (sp)
                if (int e = snapshot_rows(c, c->tab, *sp)) return e;
    if (!c->snaps.empty()) return push_snaps(c);
    return 0;
}

int rc_set_restriction(rc_ctx *c, int mode) {
    if (!c || (mode != 0 && mode != 1)) { set_error("restriction: 0 (texture mean) or 1 (cubic spline)"); return -1; }
    if (mode == 1 && (c->p.imax < 2 || c->p.jmax < 2 || c->p.kmax < 2)) {
        set_error("cubic-spline restriction needs imax, jmax, kmax >= 2");
        return -2;
    }
    c->restriction = mode;
    return 0;
}

int rc_set_reduction(rc_ctx *c, int mode) {
    if (!c || (mode != 0 && mode != 1)) { set_error("reduction: 0 (FP64 atomics) or 1 (deterministic)"); return -1; }
    if (mode == 1 && c->p.photmax > 65535) { set_error("deterministic reduction: photmax above 65535"); return -2; }
    if (c->launched) (void)rc_sync(c, nullptr);
    c->reduction = mode;
    c->drop_graph();
    return 0;
}

int rc_set_estimator(rc_ctx *c, int mode) {
    if (!c || (mode != 0 && mode != 1)) { set_error("estimator: 0 (reference) or 1 (exact at Tmax)"); return -1; }
    c->estimator = mode;
    return 0;
}

int rc_spline_restrict(const rc_params *p, int level, const float *T, float *Tc, float *Tw) {
    if (!p || !T || !Tc) { set_error("rc_spline_restrict: null argument"); return -1; }
    if (int e = check_params(*p)) return e;
    if (level < 1 || level >= p->grid_num) { set_error("rc_spline_restrict: level in 1..grid_num-1"); return -2; }
    if (p->imax < 2 || p->jmax < 2 || p->kmax < 2) { set_error("cubic-spline restriction needs imax, jmax, kmax >= 2"); return -2; }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) { set_error("no HIP device available"); return -3; }
    Grid g0, gc;
    build_grid(*p, 0, g0);
    build_grid(*p, level, gc);
    const int im = p->maxi[level], jm = p->maxj[level], km = p->maxk[level];
    const size_t nT0 = (size_t)(p->imax + 2) * (p->jmax + 2) * (p->kmax + 2);
    const size_t nc = (size_t)im * jm * km, nw = (size_t)2 * jm * km;
    hipStream_t st = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) { set_error("cannot create HIP stream"); return -3; }
    int rc = 0;
    {
        DevBuf dT, dc, dw, x0, y0, z0, xc, yc, zc, scr;
        const size_t ns = spline_scratch_floats(p->imax, p->jmax, p->kmax, im, jm, km);
        if ((rc = dT.ensure(nT0 * 4)) || (rc = dc.ensure(nc * 4)) || (rc = dw.ensure(nw * 4)) ||
            (rc = scr.ensure(ns * 4)) || (rc = upload(x0, g0.x, st)) || (rc = upload(y0, g0.y, st)) ||
            (rc = upload(z0, g0.z, st)) || (rc = upload(xc, gc.x, st)) || (rc = upload(yc, gc.y, st)) ||
            (rc = upload(zc, gc.z, st))) {
            (void)hipStreamSynchronize(st);
            (void)hipStreamDestroy(st);
            return rc;
        }
        hipError_t e = hipMemcpyAsync(dT.p, T, nT0 * 4, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) {
            launch_spline_level(st, dT.as<float>(), x0.as<float>(), y0.as<float>(), z0.as<float>(), xc.as<float>(),
                                yc.as<float>(), zc.as<float>(), p->imax, p->jmax, p->kmax, im, jm, km, dc.as<float>(),
                                Tw ? dw.as<float>() : nullptr, scr.as<float>());
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipMemcpyAsync(Tc, dc.p, nc * 4, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess && Tw) e = hipMemcpyAsync(Tw, dw.p, nw * 4, hipMemcpyDeviceToHost, st);
        const hipError_t es = hipStreamSynchronize(st);
        if (e == hipSuccess) e = es;
        if (e != hipSuccess) { set_error(std::string("rc_spline_restrict: ") + hipGetErrorString(e)); rc = -4; }
    }
    (void)hipStreamDestroy(st);
    return rc;
}

int rc_set_graph(rc_ctx *c, int on) {
    if (!c) { set_error("null context"); return -1; }
    if (c->launched) (void)rc_sync(c, nullptr);
    c->use_graph = on != 0;
    c->drop_graph();
    return 0;
}

/* coordinate search over schedules on the loaded snapshots (batched
 * variant, device-event timing of the march): steps per iteration at the
 * default occupancy, then workgroups per CU, then the refill threshold,
 * refilling once per iteration; then refills between the steps too */
int rc_autotune(rc_ctx *c, int32_t *best_wg, int32_t *best_th, int32_t *best_spi, float *best_ms) {
    const Range range_("radchan: autotune");
    if (!c || c->snaps.empty()) { set_error("rc_autotune before rc_set_fields"); return -1; }
    /* steps per iteration 1..4, and on several levels the finest level's
     * step cap (rays picked up together hand off together) */
    std::vector<int> spis = {1, 2, 3, 4};
    if (c->p.grid_num > 1 && c->p.maxs[0] > 4 && c->p.maxs[0] <= 16) spis.push_back(c->p.maxs[0]);
    static const int wgs[] = {8, 12, 16, 20, 24, 32, 40};
    static const int ths[] = {4, 8, 16, 32};
    static const int rspis[] = {2, 4, 6, 8};
    static const int rths[] = {4, 8};
    const int keep_wg = c->wg_per_cu, keep_th = c->threshold, keep_spi = c->spi, keep_rf = c->refill;
    float best = 1e30f;
    int bw = 32, bt = 8, bs = 1, br = 0;
    auto trial = [&](int wg, int th, int sp, int rf) -> int {
        c->wg_per_cu = wg; c->threshold = th; c->spi = sp; c->refill = rf;
        c->drop_graph();
        if (int e = rc_run(c, RC_VARIANT_BATCHED)) return e;
        rc_stats st;
        if (int e = rc_sync(c, &st)) return e;
        if ((float)st.march_ms < best) { best = (float)st.march_ms; bw = wg; bt = th; bs = sp; br = rf; }
        return 0;
    };
    /* refills once per iteration: steps per iteration, workgroups per CU,
     * threshold; then refills between the steps too: steps per iteration
     * at threshold 4 and the best workgroup count, then threshold 8 */
    int e = 0;
    for (int sp : spis) if (!e) e = trial(bw, bt, sp, 0);
    const int s1 = bs;
    for (int wg : wgs) if (!e) e = trial(wg, bt, s1, 0);
    const int w1 = bw;
    for (int th : ths) if (!e) e = trial(w1, th, s1, 0);
    const int w2 = bw;
    for (int sp : rspis) if (!e) e = trial(w2, rths[0], sp, 1);
    if (!e && br == 1) {
        const int s3 = bs;
        for (int th : rths) if (!e && th != rths[0]) e = trial(w2, th, s3, 1);
    }
    if (e) {
        c->wg_per_cu = keep_wg; c->threshold = keep_th; c->spi = keep_spi; c->refill = keep_rf;
        c->drop_graph();
        return e;
    }
    c->wg_per_cu = bw; c->threshold = bt; c->spi = bs; c->refill = br;
    c->drop_graph();
    if (best_wg) *best_wg = bw;
    if (best_th) *best_th = bt;
    if (best_spi) *best_spi = bs;
    if (best_ms) *best_ms = best;
    return 0;
}

int rc_sync(rc_ctx *c, rc_stats *st) {
    const Range range_("radchan: harvest (rc_sync)");
    if (!c) { set_error("null context"); return -1; }
    if (c->launched) {
        RC_HIP(hipStreamSynchronize(c->stream));
        float ms = 0.f, mms = 0.f;
        RC_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        RC_HIP(hipEventElapsedTime(&mms, c->evm0, c->evm1));
        unsigned long long h[8];
        RC_HIP(hipMemcpy(h, c->stats.p, sizeof h, hipMemcpyDeviceToHost));
        c->last.kernel_ms = ms;
        c->last.march_ms = mms;
        c->last.rays = (int64_t)h[0];
        c->last.steps = (int64_t)h[1];
        c->last.capped = (int64_t)h[2];
        c->last.bulk_slots = (int64_t)h[3];
        c->last.bulk_steps = (int64_t)h[4];
        c->launched = false;
    }
    if (st) *st = c->last;
    return 0;
}

int rc_results(rc_ctx *c, int s, float *mean, float *sdev) {
    if (!c || s < 0 || s >= (int)c->snaps.size()) { set_error("bad snapshot index"); return -1; }
    if (int e = rc_sync(c, nullptr)) return e;
    if (mean) RC_HIP(hipMemcpy(mean, c->mean.as<float>() + (size_t)s * c->ncell, (size_t)c->ncell * 4, hipMemcpyDeviceToHost));
    if (sdev) RC_HIP(hipMemcpy(sdev, c->sdev.as<float>() + (size_t)s * c->ncell, (size_t)c->ncell * 4, hipMemcpyDeviceToHost));
    return 0;
}

int rc_tallies(rc_ctx *c, int s, double *de) {
    if (!c || s < 0 || s >= (int)c->snaps.size()) { set_error("bad snapshot index"); return -1; }
    if (int e = rc_sync(c, nullptr)) return e;
    const size_t n = (size_t)c->ncell * c->p.nVar;
    RC_HIP(hipMemcpy(de, c->de.as<double>() + (size_t)s * n, n * 8, hipMemcpyDeviceToHost));
    return 0;
}

int rc_surface_results(rc_ctx *c, int s, float *mean, float *sdev, double *de) {
    if (!c || s < 0 || s >= (int)c->snaps.size()) { set_error("bad snapshot index"); return -1; }
    if (int e = rc_sync(c, nullptr)) return e;
    const size_t n = (size_t)2 * c->nsurf;
    if (mean) RC_HIP(hipMemcpy(mean, c->smean.as<float>() + (size_t)s * n, n * 4, hipMemcpyDeviceToHost));
    if (sdev) RC_HIP(hipMemcpy(sdev, c->ssdev.as<float>() + (size_t)s * n, n * 4, hipMemcpyDeviceToHost));
    if (de) RC_HIP(hipMemcpy(de, c->ds.as<double>() + (size_t)s * n * c->p.nVar, n * c->p.nVar * 8, hipMemcpyDeviceToHost));
    return 0;
}

int rc_set_schedule(rc_ctx *c, int wg_per_cu, int threshold, int steps_per_iter) {
    if (!c || wg_per_cu < 0 || wg_per_cu > 64 || threshold < 0 || threshold > 64 || steps_per_iter < 0 ||
        steps_per_iter > 16) {
        set_error("schedule: wg_per_cu in 0..64, threshold in 0..64, steps_per_iter in 0..16 (0 = default)");
        return -1;
    }
    c->wg_per_cu = wg_per_cu;
    c->threshold = threshold;
    c->spi = steps_per_iter;
    c->drop_graph();
    return 0;
}

int rc_set_refill(rc_ctx *c, int mode) {
    if (!c || mode < -1 || mode > 1) { set_error("rc_set_refill: mode -1 (default), 0 or 1"); return -1; }
    c->refill = mode;
    c->drop_graph();
    return 0;
}

int rc_get_refill(rc_ctx *c, int32_t *mode) {
    if (!c || !mode) { set_error("rc_get_refill: null context or output"); return -1; }
    *mode = c->refill;
    return 0;
}

int rc_get_schedule(rc_ctx *c, int32_t *wg_per_cu, int32_t *threshold, int32_t *steps_per_iter) {
    if (!c) { set_error("null context"); return -1; }
    *wg_per_cu = c->wg_per_cu;
    *threshold = c->threshold;
    if (steps_per_iter) *steps_per_iter = c->spi;
    return 0;
}

int rc_info(rc_ctx *c, int s, float *f5, int32_t *i2) {
    if (!c || s < 0 || s >= (int)c->snaps.size()) { set_error("bad snapshot index"); return -1; }
    const SnapState &S = *c->snaps[s];
    if (f5) { f5[0] = S.Tmax; f5[1] = S.kappamax; f5[2] = S.Ibmax; f5[3] = S.Ibmax_w; f5[4] = S.Ibmax_e; }
    if (i2) { i2[0] = S.tm; i2[1] = S.tstar; }
    return 0;
}

int rc_trace(rc_ctx *c, int surface, const int32_t *rays, int n, rc_ray *out) {
    if (!c || c->snaps.empty() || n < 1) { set_error("rc_trace needs a loaded field and rays"); return -1; }
    if (int e = rc_sync(c, nullptr)) return e;
    const rc_params &p = c->p;
    const DevCommon cm = make_common(c);
    for (int q = 0; q < n; ++q) {
        const int a = rays[5 * q], j = rays[5 * q + 1], k = rays[5 * q + 2], v = rays[5 * q + 3], h = rays[5 * q + 4];
        const bool ok = (surface ? (a == 0 || a == 1) : (a >= 1 && a <= p.imax)) && j >= 1 && j <= p.jmax &&
                        k >= 1 && k <= p.kmax && v >= 0 && v < p.nVar && h >= 0 && h < p.photmax;
        if (!ok) { set_error("traced ray out of range"); return -2; }
    }
    /* units (cell id, run, tag) whose sorted band lists are needed */
    std::vector<int32_t> ulist, unit(n);
    std::map<std::tuple<int, int, int>, int> seen;
    for (int q = 0; q < n; ++q) {
        const int a = rays[5 * q], j = rays[5 * q + 1], k = rays[5 * q + 2], v = rays[5 * q + 3];
        const int cid = surface ? (j - 1) + p.jmax * (k - 1) : (a - 1) + p.imax * ((j - 1) + p.jmax * (k - 1));
        const int tag = surface ? (a ? TAG_EAST : TAG_WEST) : TAG_FLUID;
        auto key = std::make_tuple(cid, v, tag);
        auto it = seen.find(key);
        if (it == seen.end()) {
            const int u = (int)seen.size();
            seen[key] = u;
            ulist.push_back(cid); ulist.push_back(v); ulist.push_back(tag);
            unit[q] = u;
        } else unit[q] = it->second;
    }
    std::vector<int32_t> rv(rays, rays + 5 * (size_t)n);
    if (int e = upload(c->tr_rays, rv, c->stream)) return e;
    if (int e = upload(c->tr_unit, unit, c->stream)) return e;
    if (int e = upload(c->tr_ulist, ulist, c->stream)) return e;
    const int nunit = (int)seen.size();
    uint16_t *keys = nullptr;
    if (cm.srt) {
        if (int e = c->keys.ensure((size_t)nunit * p.photmax * 2)) return e;
        keys = c->keys.as<uint16_t>();
        launch_sort(c->stream, cm, c->dsnap.as<DevSnap>(), 0, 0, surface, nunit, c->tr_ulist.as<int32_t>(), keys);
    }
    if (int e = c->tr_out.ensure((size_t)n * sizeof(rc_ray))) return e;
    launch_trace(c->stream, cm, c->dsnap.as<DevSnap>(), surface, c->tr_rays.as<int32_t>(), c->tr_unit.as<int32_t>(), n,
                 keys, c->tr_out.as<rc_ray>());
    RC_HIP(hipGetLastError());
    RC_HIP(hipMemcpyAsync(out, c->tr_out.p, (size_t)n * sizeof(rc_ray), hipMemcpyDeviceToHost, c->stream));
    RC_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

/* ------------------------------------------------------------------------- */
/* reference entry points                                                     */

/* ---- groups: snapshots of several sessions in one batched launch ------- */

struct rc_group {
    std::vector<rc_ctx *> cs;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, evm0 = nullptr, evm1 = nullptr;
    std::vector<DevSnap> h;
    DevBuf dsnap, de, ds, stats, queue, srtk;
    int wg_per_cu = 0, threshold = 0, spi = 0, refill = -1;
    rc_stats last{};
    bool launched = false;
};

/* the parameters one batched launch shares: everything in DevCommon that
 * the batched kernel reads (the grid sizes travel with each snapshot) */
static int group_compatible(const rc_ctx *a, const rc_ctx *b) {
    const DevCommon x = make_common(a), y = make_common(b);
    bool ok = x.nVar == y.nVar && x.photmax == y.photmax && x.srt == y.srt && x.nB == y.nB && x.nQ == y.nQ &&
              x.nT == y.nT && x.grid_num == y.grid_num && x.Lx == y.Lx && x.Ly == y.Ly && x.Lz == y.Lz &&
              x.toll == y.toll && x.tollS == y.tollS && x.k0 == y.k0 && x.k1 == y.k1 && x.sampler == y.sampler &&
              x.adiab[0] == y.adiab[0] && x.adiab[1] == y.adiab[1] && a->p.grey == b->p.grey &&
              a->table_bf16 == b->table_bf16 && a->p.calcwalls == b->p.calcwalls;
    for (int f = 0; f < 6; ++f) ok = ok && x.bd[f] == y.bd[f];
    return ok ? 0 : -1;
}

rc_group *rc_group_create(rc_ctx *const *cs, int n) {
    if (!cs || n < 1) { set_error("rc_group_create: no sessions"); return nullptr; }
    for (int i = 0; i < n; ++i) {
        if (!cs[i]) { set_error("rc_group_create: null session"); return nullptr; }
        if (cs[i]->p.fp64 == 1) {
            set_error("rc_group_create: a group runs the batched kernel, and floatpoint 1 (double precision) runs "
                      "in the plain variant");
            return nullptr;
        }
        if (group_compatible(cs[0], cs[i])) {
            set_error("rc_group_create: the sessions differ in a parameter other than the grid (nVar, photmax, "
                      "srt, table sizes, tolerances, boundary codes, domain lengths, seed or table precision)");
            return nullptr;
        }
    }
    auto *g = new rc_group();
    g->cs.assign(cs, cs + n);
    if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&g->ev0) != hipSuccess || hipEventCreate(&g->ev1) != hipSuccess ||
        hipEventCreate(&g->evm0) != hipSuccess || hipEventCreate(&g->evm1) != hipSuccess) {
        set_error("rc_group_create: stream or event creation failed");
        rc_group_destroy(g);
        return nullptr;
    }
    return g;
}

void rc_group_destroy(rc_group *g) {
    if (!g) return;
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    for (hipEvent_t e : {g->ev0, g->ev1, g->evm0, g->evm1})
        if (e) (void)hipEventDestroy(e);
    if (g->stream) (void)hipStreamDestroy(g->stream);
    delete g;
}

int rc_group_set_schedule(rc_group *g, int wg_per_cu, int threshold, int steps_per_iter) {
    if (!g) { set_error("null group"); return -1; }
    if (wg_per_cu < 0 || wg_per_cu > 64 || threshold < 0 || threshold > 64 || steps_per_iter < 0 || steps_per_iter > 16) {
        set_error("schedule out of range (wg_per_cu 0..64, threshold 0..64, steps_per_iter 0..16)");
        return -2;
    }
    g->wg_per_cu = wg_per_cu; g->threshold = threshold; g->spi = steps_per_iter;
    return 0;
}

int rc_group_set_refill(rc_group *g, int mode) {
    if (!g || mode < -1 || mode > 1) { set_error("rc_group_set_refill: mode -1 (default), 0 or 1"); return -1; }
    g->refill = mode;
    return 0;
}

int rc_group_run(rc_group *g) {
    const Range range_("radchan: group launch");
    if (!g) { set_error("null group"); return -1; }
    /* the sessions' own queued work on the buffers the group writes */
    for (rc_ctx *c : g->cs) {
        if (c->snaps.empty()) { set_error("rc_group_run: a session has no fields"); return -2; }
        if (group_compatible(g->cs[0], c)) { set_error("rc_group_run: sessions no longer compatible"); return -4; }
        RC_HIP(hipStreamSynchronize(c->stream));
    }
    rc_ctx *c0 = g->cs[0];
    const int nVar = c0->p.nVar;
    /* wall fluxes: the adiabatic walls of every snapshot in a second launch */
    const bool walls = c0->p.calcwalls == 1 && (c0->p.adiabw == 1 || c0->p.adiabe == 1);
    g->h.clear();
    int64_t units = 0, wunits = 0;
    for (rc_ctx *c : g->cs) {
        units += (int64_t)c->hsnap.size() * c->ncell * nVar;
        wunits += (int64_t)c->hsnap.size() * 2 * c->nsurf * nVar;
    }
    if (units >= ((int64_t)1 << 31) || (walls && wunits >= ((int64_t)1 << 31))) {
        set_error("rc_group_run: too many (snapshot, cell, run) units for one launch (2^31)");
        return -5;
    }
    int cells = 0, wcells = 0;
    for (rc_ctx *c : g->cs)
        for (const DevSnap &d0 : c->hsnap) {
            DevSnap d = d0;
            d.ubase = cells;
            d.sbase = wcells;
            cells += c->ncell;
            wcells += 2 * c->nsurf;
            g->h.push_back(d);
        }
    if (int e = upload(g->dsnap, g->h, g->stream)) return e;
    if (int e = g->de.ensure((size_t)cells * nVar * 8)) return e;
    if (walls) { if (int e = g->ds.ensure((size_t)wcells * nVar * 8)) return e; }
    if (int e = g->stats.ensure(8 * sizeof(unsigned long long))) return e;
    if (int e = g->queue.ensure(32 * sizeof(int))) return e;
    DevCommon cm = make_common(c0);
    cm.refill = g->refill;
    const size_t nk = std::max(batched_srt_keys(cm, cells, g->wg_per_cu),
                               walls ? batched_srt_keys(cm, wcells, g->wg_per_cu) : 0);
    if (nk)
        if (int e = g->srtk.ensure(nk * 2)) return e;
    cm.srt_keys = g->srtk.as<uint16_t>();
    int n = 0;
    RC_HIP(hipEventRecord(g->ev0, g->stream));
    launch_zero_u64(g->stream, g->stats.as<unsigned long long>(), 8);
    ++n;
    RC_HIP(hipEventRecord(g->evm0, g->stream));
    launch_batched(g->stream, cm, g->dsnap.as<DevSnap>(), (int)g->h.size(), cells, 0, c0->table_bf16,
                   g->de.as<double>(), g->stats.as<unsigned long long>(), g->queue.as<int>(), g->wg_per_cu,
                   g->threshold, g->spi);
    n += 3;
    if (walls) {
        launch_batched(g->stream, cm, g->dsnap.as<DevSnap>(), (int)g->h.size(), wcells, 1, c0->table_bf16,
                       g->ds.as<double>(), g->stats.as<unsigned long long>(), g->queue.as<int>() + 16,
                       g->wg_per_cu, g->threshold, g->spi);
        n += 3;
    }
    RC_HIP(hipEventRecord(g->evm1, g->stream));
    /* each session's tallies, means and deviations into its own buffers */
    size_t off = 0, woff = 0;
    for (rc_ctx *c : g->cs) {
        const size_t nde = (size_t)c->ncell * nVar, nds = (size_t)2 * c->nsurf * nVar;
        for (size_t s = 0; s < c->snaps.size(); ++s) {
            RC_HIP(hipMemcpyAsync(c->de.as<double>() + s * nde, g->de.as<double>() + off, nde * 8,
                                  hipMemcpyDeviceToDevice, g->stream));
            launch_finalize(g->stream, c->de.as<double>() + s * nde, c->ncell, nVar,
                            c->mean.as<float>() + s * c->ncell, c->sdev.as<float>() + s * c->ncell);
            ++n;
            off += nde;
            if (walls) {
                RC_HIP(hipMemcpyAsync(c->ds.as<double>() + s * nds, g->ds.as<double>() + woff, nds * 8,
                                      hipMemcpyDeviceToDevice, g->stream));
                launch_finalize(g->stream, c->ds.as<double>() + s * nds, 2 * c->nsurf, nVar,
                                c->smean.as<float>() + s * 2 * c->nsurf, c->ssdev.as<float>() + s * 2 * c->nsurf);
                ++n;
                woff += nds;
            }
        }
    }
    RC_HIP(hipEventRecord(g->ev1, g->stream));
    RC_HIP(hipGetLastError());
    g->last = rc_stats{};
    g->last.launches = n;
    g->last.variant = RC_VARIANT_BATCHED;
    g->launched = true;
    return 0;
}

int rc_group_sync(rc_group *g, rc_stats *st) {
    const Range range_("radchan: group harvest");
    if (!g) { set_error("null group"); return -1; }
    if (g->launched) {
        RC_HIP(hipStreamSynchronize(g->stream));
        float ms = 0.f, mms = 0.f;
        RC_HIP(hipEventElapsedTime(&ms, g->ev0, g->ev1));
        RC_HIP(hipEventElapsedTime(&mms, g->evm0, g->evm1));
        unsigned long long h[8];
        RC_HIP(hipMemcpy(h, g->stats.p, sizeof h, hipMemcpyDeviceToHost));
        g->last.kernel_ms = ms;
        g->last.march_ms = mms;
        g->last.rays = (int64_t)h[0];
        g->last.steps = (int64_t)h[1];
        g->last.capped = (int64_t)h[2];
        g->last.bulk_slots = (int64_t)h[3];
        g->last.bulk_steps = (int64_t)h[4];
        g->launched = false;
    }
    if (st) *st = g->last;
    return 0;
}

static std::mutex g_mu;
static bool g_params_set = false;
static rc_params g_params;
static std::string g_dir = "properties";
static int g_variant = RC_VARIANT_BATCHED;
static int g_estimator = 0, g_restriction = 0;   /* options of the drop-in session */
static rc_ctx *g_ctx = nullptr;
static bool g_pending = false;
static std::string g_tab_dir;

static rc_params &gparams() {
    if (!g_params_set) { default_params(g_params); g_params_set = true; }
    return g_params;
}

int rc_set_params(const rc_params *p) {
    std::lock_guard<std::mutex> lk(g_mu);
    rc_params q = *p;
    if (q.grey) q.srt = 0;
    if (check_params(q)) return -1;
    gparams() = q;
    g_params_set = true;
    if (g_ctx && !g_pending) { rc_destroy(g_ctx); g_ctx = nullptr; g_tab_dir.clear(); }
    return 0;
}

int rc_params_compiled(rc_params *p, int imax_, int jmax_, int kmax_, int p_col_, float Lx_, float Ly_, float Lz_,
                       double fact_mesh_, int nVar_, int photmax_, double chanHeight_, int grid_num_,
                       const int *maxi_, const int *maxj_, const int *maxk_, const int *maxs_, int srt_, int grey_,
                       int calcwalls_, const float *eps6, const float *Tw6, int adiabw_, int adiabe_, float toll_,
                       float tollS_, const int *bd6, int nT_, int nQ_, int nB_, float abscoeff_, float Tmins_,
                       float wvmax_) {
    if (!p || grid_num_ < 1 || grid_num_ > RC_MAXLEV) { set_error("rc_params_compiled: bad grid_num"); return -1; }
    default_params(*p);
    p->imax = imax_; p->jmax = jmax_; p->kmax = kmax_; p->p_col = p_col_;
    p->Lx = Lx_; p->Ly = Ly_; p->Lz = Lz_;
    p->fact_mesh = fact_mesh_; p->nVar = nVar_; p->photmax = photmax_; p->chanHeight = chanHeight_;
    p->grid_num = grid_num_;
    for (int g = 0; g < RC_MAXLEV; ++g) {
        const bool on = g < grid_num_;
        p->maxi[g] = on ? maxi_[g] : 1; p->maxj[g] = on ? maxj_[g] : 1;
        p->maxk[g] = on ? maxk_[g] : 1; p->maxs[g] = on ? maxs_[g] : 1;
    }
    p->srt = srt_; p->grey = grey_; p->calcwalls = calcwalls_; p->grey_true = 0;
    for (int f = 0; f < 6; ++f) { p->eps[f] = eps6[f]; p->Tw[f] = Tw6[f]; p->bd[f] = bd6[f]; }
    p->adiabw = adiabw_; p->adiabe = adiabe_;
    p->toll = toll_; p->tollS = tollS_;
    p->nT = nT_; p->nQ = nQ_; p->nB = nB_;
    p->abscoeff = abscoeff_; p->Tmins = Tmins_; p->wvmax = wvmax_;
    return 0;
}

int rc_get_params(rc_params *p) {
    std::lock_guard<std::mutex> lk(g_mu);
    *p = gparams();
    return 0;
}

int rc_set_table_dir(const char *dir) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_dir = dir ? dir : "properties";
    g_tab_dir.clear();
    return 0;
}

int rc_set_variant(int variant) {
    if (variant != RC_VARIANT_PLAIN && variant != RC_VARIANT_BATCHED && variant != RC_VARIANT_PLAIN_BANDS) {
        set_error("unknown variant");
        return -1;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    g_variant = variant;
    return 0;
}

int rc_set_dropin_modes(int estimator, int restriction) {
    if ((estimator != 0 && estimator != 1) || (restriction != 0 && restriction != 1)) {
        set_error("drop-in modes: estimator and restriction are 0 or 1");
        return -1;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    g_estimator = estimator;
    g_restriction = restriction;
    return 0;
}

static void report(const char *where) { std::fprintf(stderr, "radchan %s: %s\n", where, last_error()); }

void mc_gpu_(float *Tfort, int *zstart, int *rank) {
    const Range range_("radchan: mc_gpu_");
    (void)rank;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!Tfort || !zstart) { set_error("null argument"); report("mc_gpu_"); return; }
    if (g_pending && g_ctx) {
        /* a launch was never harvested: finish it and drop its results */
        (void)rc_sync(g_ctx, nullptr);
        g_pending = false;
        std::fprintf(stderr, "radchan mc_gpu_: previous solve was not harvested by get_results_; discarded\n");
    }
    const rc_params &p = gparams();
    if (g_ctx && std::memcmp(&g_ctx->p, &p, sizeof p) != 0) { rc_destroy(g_ctx); g_ctx = nullptr; g_tab_dir.clear(); }
    if (!g_ctx) {
        g_ctx = rc_create(&p);
        if (!g_ctx) { report("mc_gpu_"); return; }
    }
    if (!p.grey && g_tab_dir != g_dir) {
        Tables t;
        t.resize(p.nB, p.nT, p.nQ);
        if (read_tables(g_dir, p.chanHeight, t)) { report("mc_gpu_"); return; }
        t.wq = reference_wq(p.nQ);
        rc_tables v = t.view();
        if (rc_load_tables(g_ctx, &v)) { report("mc_gpu_"); return; }
        g_tab_dir = g_dir;
    }
    if (rc_set_estimator(g_ctx, g_estimator) || rc_set_restriction(g_ctx, g_restriction)) { report("mc_gpu_"); return; }
    const float *T = Tfort;
    if (rc_set_fields(g_ctx, 1, &T, *zstart)) { report("mc_gpu_"); return; }
    /* floatpoint 1: the full-FP64 kernels, which run in the plain launch shape */
    if (rc_run(g_ctx, p.fp64 == 1 ? RC_VARIANT_PLAIN : g_variant)) { report("mc_gpu_"); return; }
    g_pending = true;
}

void get_results_(float *resfort, float *varfort, int *rank) {
    const Range range_("radchan: get_results_");
    (void)rank;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_pending || !g_ctx) { set_error("get_results_ called without a preceding mc_gpu_"); report("get_results_"); return; }
    if (!resfort || !varfort) { set_error("null argument"); report("get_results_"); return; }
    rc_ctx *c = g_ctx;
    const rc_params &p = c->p;
    std::vector<float> m((size_t)c->ncell), sd((size_t)c->ncell);
    if (rc_results(c, 0, m.data(), sd.data())) { report("get_results_"); return; }
    g_pending = false;
    const size_t nx = (size_t)p.imax + 2, ny = (size_t)p.jmax + 2;
    auto F = [&](int i, int j, int kl) { return (size_t)i + nx * ((size_t)j + ny * (size_t)kl); };
    for (int kl = 1; kl <= c->kloc; ++kl)
        for (int j = 1; j <= p.jmax; ++j)
            for (int i = 1; i <= p.imax; ++i) {
                const size_t n = (size_t)(i - 1) + (size_t)p.imax * ((size_t)(j - 1) + (size_t)p.jmax * (kl - 1));
                resfort[F(i, j, kl)] = m[n];
                varfort[F(i, j, kl)] = sd[n];
            }
    if (p.adiabw == 1 || p.adiabe == 1) {
        std::vector<float> wm((size_t)2 * c->nsurf, 0.f), ws((size_t)2 * c->nsurf, 0.f);
        if (p.calcwalls == 1 && rc_surface_results(c, 0, wm.data(), ws.data(), nullptr)) { report("get_results_"); return; }
        for (int kl = 1; kl <= c->kloc; ++kl)
            for (int j = 1; j <= p.jmax; ++j) {
                const size_t n = (size_t)(j - 1) + (size_t)p.jmax * (kl - 1);
                resfort[F(0, j, kl)] = wm[n];
                varfort[F(0, j, kl)] = ws[n];
                resfort[F(p.imax + 1, j, kl)] = wm[c->nsurf + n];
                varfort[F(p.imax + 1, j, kl)] = ws[c->nsurf + n];
            }
    }
}

}  // extern "C"
