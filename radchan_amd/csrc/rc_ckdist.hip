/*
 * rc_ckdist.hip -- correlated-k distributions of line-by-line spectra
 * (nbcorrk/correlatedK.cpp:144-225, correlatedKT) and the band means k_avg
 * of calcProbT (nbcorrk/calc_prob.cpp:52-56), on the GPU.
 *
 * Per spectrum (one narrow band at one temperature, points (wvn, kvn)):
 *   kmin, kmax over the points; p = k^pwr spaced evenly in n_pwrk steps,
 *   kk[h] = (h*step + pmin)^(1/pwr), then shifted half a bin upwards;
 *   ff[bin] += dwv_h / (wv_right - wv_left) for the pair average
 *   ki = (k_h + k_{h-1}) / 2, bin = max(0, int((ki^pwr - pmin)/step + 1));
 *   gg[ng] = 1, gg[h-1] = gg[h] - ff[h] for h = ng..1 (ng = n_pwrk - 1);
 *   kq[r] interpolates kk linearly in gg at each quadrature point gq[r];
 *   k_avg = sum_{h=1}^{ng} kk[h] ff[h].
 *
 * The reference walks all n_pwrk + 1 bins per spectrum.  A narrow band holds
 * far fewer points (about 1000 at dwv = 0.01 cm^-1 and 10 cm^-1 bands) than
 * bins (20 000), so the histogram is sparse: this kernel sorts the point
 * pairs by bin and works on the occupied bins only.  gg is a step function
 * that changes at occupied bins, so the quadrature walk reduces to a
 * binary search over them, and kk is evaluated only where it is used.
 *
 * One workgroup of 256 threads per spectrum, everything in LDS:
 *   (1) kmin/kmax: block min/max (exact in any order);
 *   (2) per pair h = 1..n-1 the key bin << 13 | h, sorted (so that pairs of
 *       one bin come in ascending h) by a bitonic network held in registers:
 *       only 3 of its stages cross wavefronts through LDS;
 *   (3) run starts (occupied bins) compacted by a block scan;
 *   (4) per occupied bin, ff summed in ascending h by one thread from the
 *       pair weights staged in LDS -- the reference's order, so ff is bit
 *       for bit the reference's;
 *   (5) the gg chain (one lane, descending bins) and the k_avg sum (one lane
 *       of another wave, ascending bins): the reference's sequences of
 *       operations with the zero terms of empty bins left out (x - 0 and
 *       x + 0 are exact), so both are bit for bit given the same kk;
 *   (6) one thread per quadrature point: binary searches for the first
 *       occupied bin whose gg reaches gq[r].
 * The bin index uses the device's FP64 pow, which may differ from the host
 * libm's in the last bit, and kk = (h step + pmin)^(1/pwr) is formed by
 * squaring when 1/pwr is an integer (10 for pwr = 0.1), within 5 ulp of
 * pow: results agree with the CPU restatement to FP64 rounding (and a pair
 * exactly on a bin edge could land one bin over).
 *
 * LDS per spectrum, np2 = 256 R >= n - 1 (R = 1, 2, .., 32 keys per thread,
 * one kernel instance each, chosen by the launch's longest spectrum): keys u32[np2], run
 * starts u32[np2], ff f64[np2], run bins u16[np2] = 18 B per slot; the
 * products kk*ff (f64[np2]) overlay keys and run starts once those are dead.
 * Up to 8192 points per spectrum (13-bit pair index), n_pwrk <= 65535.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "rc_device.h"
#include "rc_internal.h"

using namespace rcd;

namespace rc {
namespace {

constexpr int CK_THREADS = 256;
constexpr int CK_WAVES = CK_THREADS / 64;
constexpr int CK_NMAX = 8192;        /* points per spectrum (13-bit pair index) */
constexpr int CK_IDX_BITS = 13;

struct CkArgs {
    const int64_t *off;
    const double *wvn, *kvn, *wvl, *wvr, *gq;
    double *kq, *kavg, *kmin, *kmax;
    double pwr, ipwr;
    int nq, n_pwrk, ie;
};

/* bitonic sort of 256 R keys held in registers, element R tid + r in x[r]:
 * stages at distance below R stay inside a thread, distances R..32 R
 * exchange with lane (lane ^ distance/R) of the same wavefront, and only
 * the 64 R and 128 R distances (3 of the log2(256R)(log2(256R)+1)/2
 * stages) go through LDS, with a barrier on each side. */
template <int R>
__device__ __forceinline__ void block_sort(uint32_t (&x)[R], uint32_t *K, int tid) {
    int base = tid * R;
    /* opaque, so the per-stage masks are formed here rather than hoisted */
    asm volatile("" : "+v"(base));
#pragma unroll
    for (int k = 2; k <= 256 * R; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            if (j < R) {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int p = r ^ j;
                    if (p > r) {
                        const bool up = ((base + r) & k) == 0;
                        const uint32_t lo = min(x[r], x[p]), hi = max(x[r], x[p]);
                        x[r] = up ? lo : hi;
                        x[p] = up ? hi : lo;
                    }
                }
            } else if (j < 64 * R) {
                const bool keep_min = ((base & j) == 0) == ((base & k) == 0);
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const uint32_t y = lane_xor(x[r], j / R);
                    x[r] = keep_min ? min(x[r], y) : max(x[r], y);
                }
            } else {
                __syncthreads();
#pragma unroll
                for (int r = 0; r < R; ++r) K[base + r] = x[r];
                __syncthreads();
                const bool keep_min = ((base & j) == 0) == ((base & k) == 0);
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const uint32_t y = K[(base + r) ^ j];
                    x[r] = keep_min ? min(x[r], y) : max(x[r], y);
                }
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; ++r) K[base + r] = x[r];
    __syncthreads();
}

/* x^(1/pwr): by squaring when 1/pwr is a small integer (ie > 0; 10 for the
 * generator's pwr = 0.1: four products, within 5 ulp of the rounded power),
 * else pow */
__device__ __forceinline__ double kpow(double x, double ipwr, int ie) {
    if (ie > 0) {
        double r = 1.0, b = x;
        for (int e = ie; e; e >>= 1) {
            if (e & 1) r = r * b;
            if (e > 1) b = b * b;
        }
        return r;
    }
    return pow(x, ipwr);
}

__device__ __forceinline__ double kk_at(int h, double step, double pmin, double ipwr, int ie) {
    /* kk[h] after the half-bin shift: 0.5 * (K(h) + K(h+1)), K(h) = (h*step + pmin)^(1/pwr) */
    const double a = kpow((double)h * step + pmin, ipwr, ie);
    const double b = kpow((double)(h + 1) * step + pmin, ipwr, ie);
    return 0.5 * (a + b);
}

/* the same, out of line: the quadrature phase calls it four times, and
 * inlined there it raises the kernel's register peak (and lowers its
 * occupancy) */
__device__ __noinline__ double kk_at_q(int h, double step, double pmin, double ipwr, int ie) {
    return kk_at(h, step, pmin, ipwr, ie);
}

/* gg(j) = G_lo of the first occupied bin above j, 1 above the last */
__device__ __forceinline__ double gg_at(int j, const uint16_t *RB, const double *G, int r0, int r1) {
    int lo = r0, hi = r1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((int)RB[mid] > j) hi = mid; else lo = mid + 1;
    }
    return lo < r1 ? G[lo] : 1.0;
}

template <int R>
__global__ __launch_bounds__(CK_THREADS) void k_ckdist(CkArgs a) {
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ double red_lo[CK_WAVES], red_hi[CK_WAVES];
    __shared__ int scan[CK_WAVES];
    __shared__ int s_m, s_r0, s_r1;

    const int s = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    constexpr int np2 = CK_THREADS * R;
    uint32_t *K = (uint32_t *)lds;                      /* [np2] sort keys          */
    uint32_t *RS = K + np2;                             /* [np2] run starts         */
    double *P = (double *)lds;                          /* [np2] kk*ff (over K, RS) */
    double *F = (double *)(lds + 8 * (size_t)np2);      /* [np2] ff, then gg below  */
    uint16_t *RB = (uint16_t *)(lds + 16 * (size_t)np2); /* [np2] run bins          */

    const int64_t base = a.off[s];
    const int64_t n64 = a.off[s + 1] - base;
    if (n64 < 1 || n64 - 1 > np2) {
        /* more pairs than this instance's sort slots (or none): the caller
         * of the device entry point under-stated the longest spectrum */
        for (int q = tid; q < a.nq; q += CK_THREADS) a.kq[(size_t)s * a.nq + q] = __builtin_nan("");
        if (tid == 0) { a.kavg[s] = __builtin_nan(""); a.kmin[s] = __builtin_nan(""); a.kmax[s] = __builtin_nan(""); }
        return;
    }
    const int n = (int)n64;
    const double *wvs = a.wvn + base, *kvs = a.kvn + base;

    /* (1) kmin, kmax (the reference's initial values 1e10 and -1000) */
    double lo = 1e10, hi = -1000.0;
    for (int h = tid; h < n; h += CK_THREADS) {
        const double k = kvs[h];
        lo = k < lo ? k : lo;
        hi = hi < k ? k : hi;
    }
    for (int o = 32; o > 0; o >>= 1) {
        const double l2 = __shfl_xor(lo, o), h2 = __shfl_xor(hi, o);
        lo = l2 < lo ? l2 : lo;
        hi = hi < h2 ? h2 : hi;
    }
    if (lane == 0) { red_lo[wv] = lo; red_hi[wv] = hi; }
    __syncthreads();
    lo = red_lo[0];
    hi = red_hi[0];
    for (int w = 1; w < CK_WAVES; ++w) {
        lo = red_lo[w] < lo ? red_lo[w] : lo;
        hi = hi < red_hi[w] ? red_hi[w] : hi;
    }
    const int n_pwrk = a.n_pwrk, ng = n_pwrk - 1;
    /* k^pwr of kmin (even lanes) and kmax (odd lanes): one pow per wave */
    const double pv = pow((lane & 1) ? hi : lo, a.pwr);
    const double pmin = __shfl(pv, 0), pmax = __shfl(pv, 1);
    const double step = (pmax - pmin) / (double)(n_pwrk - 1);
    const double width = a.wvr[s] - a.wvl[s];
    const int npair = n > 1 ? n - 1 : 0;

    /* (2) keys: bin << 13 | h for the pair (h-1, h), pair i = R tid + r in
     * x[r]; padding sorts last */
    uint32_t x[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = tid * R + r;
        uint32_t key = 0xFFFFFFFFu;
        if (i < npair) {
            const int h = i + 1;
            const double ki = (kvs[h] + kvs[h - 1]) / 2;
            int bin = 1;                                 /* a flat spectrum (step 0) */
            if (step > 0.0) {
                const double xb = (pow(ki, a.pwr) - pmin) / step + 1;
                bin = max(0, (int)xb);
                bin = min(bin, n_pwrk);
            }
            key = ((uint32_t)bin << CK_IDX_BITS) | (uint32_t)h;
            F[i] = (wvs[h] - wvs[h - 1]) / width;        /* the pair's weight, by pair */
        }
        x[r] = key;
    }
    block_sort<R>(x, K, tid);

    /* (3) run starts, compacted in order: each thread scans its R keys */
    constexpr int chunk = R;
    const int c0 = tid * chunk;
    int cnt = 0;
    for (int i = c0; i < c0 + chunk; ++i)
        cnt += (i < npair && (i == 0 || (K[i] >> CK_IDX_BITS) != (K[i - 1] >> CK_IDX_BITS))) ? 1 : 0;
    int inc = cnt;
    for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(inc, o);
        if (lane >= o) inc += v;
    }
    if (lane == 63) scan[wv] = inc;
    __syncthreads();
    int pos = inc - cnt;
    for (int w = 0; w < wv; ++w) pos += scan[w];
    if (tid == CK_THREADS - 1) s_m = pos + cnt;
    for (int i = c0; i < c0 + chunk; ++i) {
        if (i < npair && (i == 0 || (K[i] >> CK_IDX_BITS) != (K[i - 1] >> CK_IDX_BITS))) {
            RS[pos] = (uint32_t)i;
            RB[pos] = (uint16_t)(K[i] >> CK_IDX_BITS);
            ++pos;
        }
    }
    __syncthreads();
    const int m = s_m;

    /* (4) ff per occupied bin, summed in ascending h from the pair weights
     * (held in F by pair); the sums replace them after a barrier */
    double fr[R];
#pragma unroll
    for (int u = 0; u < R; ++u) {
        const int r = tid + u * CK_THREADS;
        double f = 0.0;
        if (r < m) {
            const int b = (int)RS[r], e = r + 1 < m ? (int)RS[r + 1] : npair;
            for (int q = b; q < e; ++q) {
                const int h = (int)(K[q] & ((1u << CK_IDX_BITS) - 1));
                f = f + F[h - 1];
            }
        }
        fr[u] = f;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < R; ++u) {
        const int r = tid + u * CK_THREADS;
        if (r < m) F[r] = fr[u];
    }
    __syncthreads();
    /* kk*ff of the bins the k_avg sum covers (1..ng), over the dead keys */
    for (int r = tid; r < m; r += CK_THREADS) {
        const int bin = RB[r];
        P[r] = (bin >= 1 && bin <= ng) ? kk_at(bin, step, pmin, a.ipwr, a.ie) * F[r] : 0.0;
    }
    if (tid == 0) {
        int r0 = (m > 0 && RB[0] == 0) ? 1 : 0;
        int r1 = (m > r0 && (int)RB[m - 1] > ng) ? m - 1 : m;
        s_r0 = r0;
        s_r1 = r1;
    }
    __syncthreads();
    const int r0 = s_r0, r1 = s_r1;

    /* (5) gg chain from the top (wave 0) and the k_avg sum (wave 1) */
    if (tid == 0) {
        double g = 1.0;
        int r = r1 - 1;
        for (; r - 7 >= r0; r -= 8) {
            double f[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) f[u] = F[r - u];
#pragma unroll
            for (int u = 0; u < 8; ++u) { g = g - f[u]; F[r - u] = g; }
        }
        for (; r >= r0; --r) { g = g - F[r]; F[r] = g; }
    } else if (tid == 64) {
        double acc = 0.0;
        int r = r0;
        for (; r + 7 < r1; r += 8) {
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = P[r + u];
#pragma unroll
            for (int u = 0; u < 8; ++u) acc = acc + v[u];
        }
        for (; r < r1; ++r) acc = acc + P[r];
        a.kavg[s] = acc;
        a.kmin[s] = lo;
        a.kmax[s] = hi;
    }
    __syncthreads();

    /* (6) quadrature: t + 1 = the first j >= 1 with gg(j) >= gq (at most ng) */
    for (int q = tid; q < a.nq; q += CK_THREADS) {
        const double g = a.gq[q];
        int jstar;
        if (gg_at(1, RB, F, r0, r1) >= g) {
            jstar = 1;
        } else {
            int lo2 = r0;                                /* first occupied bin >= 2 */
            int hi2 = r1;
            while (lo2 < hi2) {
                const int mid = (lo2 + hi2) >> 1;
                if ((int)RB[mid] >= 2) hi2 = mid; else lo2 = mid + 1;
            }
            hi2 = r1;                                    /* first whose upper gg >= g */
            while (lo2 < hi2) {
                const int mid = (lo2 + hi2) >> 1;
                const double up = mid + 1 < r1 ? F[mid + 1] : 1.0;
                if (up >= g) hi2 = mid; else lo2 = mid + 1;
            }
            jstar = lo2 < r1 ? (int)RB[lo2] : ng;
        }
        const int t = jstar - 1;
        const double g0 = gg_at(t, RB, F, r0, r1), g1 = gg_at(t + 1, RB, F, r0, r1);
        const double k0 = kk_at_q(t, step, pmin, a.ipwr, a.ie), k1 = kk_at_q(t + 1, step, pmin, a.ipwr, a.ie);
        double v = k0 + (k1 - k0) * (g - g0) / (g1 - g0);
        if (g1 == g0) v = k1;
        a.kq[(size_t)s * a.nq + q] = v;
    }
}

struct DevBuf {
    void *p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
};

}  // namespace
}  // namespace rc

using namespace rc;

namespace rc {
namespace {
/* sort slots per spectrum for spectra of up to nmax points (a power of two,
 * at least one per thread) */
int ck_slots(int nmax) {
    int np2 = CK_THREADS;
    while (np2 < nmax - 1) np2 <<= 1;
    return np2;
}

hipError_t ck_launch(const CkArgs &a, int nspec, int np2, hipStream_t st) {
    const size_t lds = (size_t)np2 * 18;
    switch (np2 / CK_THREADS) {
        case 1: hipLaunchKernelGGL(k_ckdist<1>, dim3(nspec), dim3(CK_THREADS), lds, st, a); break;
        case 2: hipLaunchKernelGGL(k_ckdist<2>, dim3(nspec), dim3(CK_THREADS), lds, st, a); break;
        case 4: hipLaunchKernelGGL(k_ckdist<4>, dim3(nspec), dim3(CK_THREADS), lds, st, a); break;
        case 8: hipLaunchKernelGGL(k_ckdist<8>, dim3(nspec), dim3(CK_THREADS), lds, st, a); break;
        case 16: hipLaunchKernelGGL(k_ckdist<16>, dim3(nspec), dim3(CK_THREADS), lds, st, a); break;
        default: hipLaunchKernelGGL(k_ckdist<32>, dim3(nspec), dim3(CK_THREADS), lds, st, a); break;
    }
    return hipGetLastError();
}

int ck_lds_check(int np2, int nmax) {
    int dev = 0, lds_max = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) {
        set_error("rc_ckdist: no HIP device");
        return -6;
    }
    const size_t lds = (size_t)np2 * 18;
    if (lds + 256 > (size_t)lds_max) {
        set_error("rc_ckdist: spectra of " + std::to_string(nmax) + " points need " + std::to_string(lds) +
                  " B of LDS; the device allows " + std::to_string(lds_max));
        return -6;
    }
    return 0;
}
}  // namespace
}  // namespace rc

extern "C" int rc_ckdist_device(int nspec, const int64_t *off, const double *wvn, const double *kvn,
                                const double *wvl, const double *wvr, int n_pwrk, double pwr, int nq,
                                const double *gq, double *kq, double *kavg, double *kmin, double *kmax,
                                int max_points, void *stream) {
    if (nspec < 1 || !off || !wvn || !kvn || !wvl || !wvr || !gq || !kq || !kavg || !kmin || !kmax) {
        set_error("rc_ckdist_device: null array or no spectra");
        return -1;
    }
    if (n_pwrk < 2 || n_pwrk > 65535 || !(pwr > 0.0) || !std::isfinite(pwr)) {
        set_error("rc_ckdist_device: n_pwrk in 2..65535 and pwr > 0");
        return -2;
    }
    if (nq < 1 || nq > 4096) { set_error("rc_ckdist_device: nq in 1..4096"); return -3; }
    if (max_points < 1 || max_points > CK_NMAX) {
        set_error("rc_ckdist_device: max_points in 1.." + std::to_string(CK_NMAX));
        return -4;
    }
    const int np2 = ck_slots(max_points);
    if (int e = ck_lds_check(np2, max_points)) return e;
    CkArgs a;
    a.off = off; a.wvn = wvn; a.kvn = kvn; a.wvl = wvl; a.wvr = wvr; a.gq = gq;
    a.kq = kq; a.kavg = kavg; a.kmin = kmin; a.kmax = kmax;
    a.pwr = pwr;
    a.ipwr = 1.0 / pwr;
    a.ie = (a.ipwr == std::rint(a.ipwr) && a.ipwr >= 1.0 && a.ipwr <= 64.0) ? (int)a.ipwr : 0;
    a.nq = nq;
    a.n_pwrk = n_pwrk;
    const hipError_t e = ck_launch(a, nspec, np2, (hipStream_t)stream);
    if (e != hipSuccess) { set_error(std::string("rc_ckdist_device: ") + hipGetErrorString(e)); return -7; }
    return 0;
}

extern "C" int rc_ckdist(int nspec, const int64_t *off, const double *wvn, const double *kvn, const double *wvl,
                         const double *wvr, int n_pwrk, double pwr, int nq, const double *gq, double *kq,
                         double *kavg, double *kmin, double *kmax, double *kernel_ms) {
    if (nspec < 1 || !off || !wvn || !kvn || !wvl || !wvr || !gq || !kq || !kavg || !kmin || !kmax) {
        set_error("rc_ckdist: null array or no spectra");
        return -1;
    }
    if (n_pwrk < 2 || n_pwrk > 65535 || !(pwr > 0.0) || !std::isfinite(pwr)) {
        set_error("rc_ckdist: n_pwrk in 2..65535 and pwr > 0");
        return -2;
    }
    if (nq < 1 || nq > 4096) { set_error("rc_ckdist: nq in 1..4096"); return -3; }
    for (int q = 0; q < nq; ++q) {
        if (!(gq[q] >= 0.0 && gq[q] <= 1.0) || (q > 0 && gq[q] < gq[q - 1])) {
            set_error("rc_ckdist: gq must be nondecreasing in [0, 1]");
            return -3;
        }
    }
    if (off[0] < 0) { set_error("rc_ckdist: negative offset"); return -4; }
    int nmax = 0;
    for (int s = 0; s < nspec; ++s) {
        const int64_t n = off[s + 1] - off[s];
        if (n < 1 || n > CK_NMAX) {
            set_error("rc_ckdist: spectrum " + std::to_string(s) + " has " + std::to_string(n) +
                      " points (1.." + std::to_string(CK_NMAX) + " per spectrum)");
            return -4;
        }
        if (!(wvr[s] > wvl[s])) {
            set_error("rc_ckdist: spectrum " + std::to_string(s) + " has wv_right <= wv_left");
            return -5;
        }
        nmax = std::max(nmax, (int)n);
    }
    const int np2 = ck_slots(nmax);
    if (int e = ck_lds_check(np2, nmax)) return e;

    const int64_t p0 = off[0], np = off[nspec] - p0;
    std::vector<int64_t> hoff(off, off + nspec + 1);
    for (auto &o : hoff) o -= p0;
    DevBuf d_off, d_wvn, d_kvn, d_wvl, d_wvr, d_gq, d_kq, d_kavg, d_kmin, d_kmax;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    auto fail = [&](const char *what, hipError_t e) {
        set_error(std::string("rc_ckdist: ") + what + ": " + hipGetErrorString(e));
        rc = -7;
    };
    hipError_t e;
    if ((e = hipStreamCreate(&st)) != hipSuccess) { fail("stream", e); return rc; }
#define CK_TRY(call)                                   \
    if ((e = (call)) != hipSuccess) { fail(#call, e); break; }
    const size_t nsb = (size_t)nspec * sizeof(double);
    do {
        CK_TRY(hipEventCreate(&e0));
        CK_TRY(hipEventCreate(&e1));
        CK_TRY(hipMalloc(&d_off.p, (nspec + 1) * sizeof(int64_t)));
        CK_TRY(hipMalloc(&d_wvn.p, np * sizeof(double)));
        CK_TRY(hipMalloc(&d_kvn.p, np * sizeof(double)));
        CK_TRY(hipMalloc(&d_wvl.p, nsb));
        CK_TRY(hipMalloc(&d_wvr.p, nsb));
        CK_TRY(hipMalloc(&d_gq.p, nq * sizeof(double)));
        CK_TRY(hipMalloc(&d_kq.p, nsb * nq));
        CK_TRY(hipMalloc(&d_kavg.p, nsb));
        CK_TRY(hipMalloc(&d_kmin.p, nsb));
        CK_TRY(hipMalloc(&d_kmax.p, nsb));
        CK_TRY(hipMemcpyAsync(d_off.p, hoff.data(), (nspec + 1) * sizeof(int64_t), hipMemcpyHostToDevice, st));
        CK_TRY(hipMemcpyAsync(d_wvn.p, wvn + p0, np * sizeof(double), hipMemcpyHostToDevice, st));
        CK_TRY(hipMemcpyAsync(d_kvn.p, kvn + p0, np * sizeof(double), hipMemcpyHostToDevice, st));
        CK_TRY(hipMemcpyAsync(d_wvl.p, wvl, nsb, hipMemcpyHostToDevice, st));
        CK_TRY(hipMemcpyAsync(d_wvr.p, wvr, nsb, hipMemcpyHostToDevice, st));
        CK_TRY(hipMemcpyAsync(d_gq.p, gq, nq * sizeof(double), hipMemcpyHostToDevice, st));
        CkArgs a;
        a.off = (const int64_t *)d_off.p;
        a.wvn = (const double *)d_wvn.p;
        a.kvn = (const double *)d_kvn.p;
        a.wvl = (const double *)d_wvl.p;
        a.wvr = (const double *)d_wvr.p;
        a.gq = (const double *)d_gq.p;
        a.kq = (double *)d_kq.p;
        a.kavg = (double *)d_kavg.p;
        a.kmin = (double *)d_kmin.p;
        a.kmax = (double *)d_kmax.p;
        a.pwr = pwr;
        a.ipwr = 1.0 / pwr;
        a.ie = (a.ipwr == std::rint(a.ipwr) && a.ipwr >= 1.0 && a.ipwr <= 64.0) ? (int)a.ipwr : 0;
        a.nq = nq;
        a.n_pwrk = n_pwrk;
        CK_TRY(hipEventRecord(e0, st));
        CK_TRY(ck_launch(a, nspec, np2, st));
        CK_TRY(hipEventRecord(e1, st));
        CK_TRY(hipMemcpyAsync(kq, d_kq.p, nsb * nq, hipMemcpyDeviceToHost, st));
        CK_TRY(hipMemcpyAsync(kavg, d_kavg.p, nsb, hipMemcpyDeviceToHost, st));
        CK_TRY(hipMemcpyAsync(kmin, d_kmin.p, nsb, hipMemcpyDeviceToHost, st));
        CK_TRY(hipMemcpyAsync(kmax, d_kmax.p, nsb, hipMemcpyDeviceToHost, st));
        CK_TRY(hipStreamSynchronize(st));
        if (kernel_ms) {
            float ms = 0.0f;
            CK_TRY(hipEventElapsedTime(&ms, e0, e1));
            *kernel_ms = ms;
        }
    } while (0);
#undef CK_TRY
    (void)hipStreamSynchronize(st);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(st);
    return rc;
}
