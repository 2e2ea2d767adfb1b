/* LSD radix sort of 16-bit keys by one wavefront (two 8-bit passes), shared
 * by the batched kernel (keys in HBM scratch) and the plain variant's band
 * sort (keys in LDS). */
#ifndef RC_RADIX_H
#define RC_RADIX_H

#include <hip/hip_runtime.h>

#include <cstdint>

namespace rc {

/* the bits of m below this lane */
__device__ __forceinline__ uint32_t lanes_below(unsigned long long m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

/* The two digit histograms (hist[0..255]: bits 0-7 of the keys,
 * hist[256..511]: bits 8-15) turned into exclusive prefix sums, each lane
 * holding 4 consecutive buckets of a histogram. */
__device__ __forceinline__ void radix_offsets(uint32_t *hist, int lane) {
#pragma unroll 1
    for (int half = 0; half < 2; ++half) {
        uint32_t *hh = hist + 256 * half + 4 * lane;
        const uint32_t c0 = hh[0], c1 = hh[1], c2 = hh[2], c3 = hh[3];
        const uint32_t s = c0 + c1 + c2 + c3;
        uint32_t inc = s;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t y = (uint32_t)__shfl_up((int)inc, d);
            if (lane >= d) inc += y;
        }
        const uint32_t ex = inc - s;
        hh[0] = ex;
        hh[1] = ex + c0;
        hh[2] = ex + c0 + c1;
        hh[3] = ex + c0 + c1 + c2;
    }
    __syncthreads();
}

/* One stable counting pass of an LSD radix sort over n 16-bit keys: the
 * digit (key >> sh) & 255 of key h goes to dst[off[digit] + (keys of that
 * digit before h)], off being the digit's running offset in LDS.  Keys are
 * taken 64 at a time in order; the lanes holding the same digit are found
 * by eight ballots, so the pass keeps the order of equal digits.  src and
 * dst are the wavefront's own scratch, in HBM or LDS.  One wavefront per
 * workgroup. */
__device__ __forceinline__ void radix_pass(const uint16_t *src, uint16_t *dst, uint32_t *off, int n, int sh,
                                           int lane) {
#pragma unroll 1
    for (int base = 0; base < n; base += 64) {
        const bool ok = base + lane < n;
        const uint32_t key = ok ? (uint32_t)src[base + lane] : 0u;
        {
            const uint32_t dig = (key >> sh) & 255u;
            unsigned long long m = __ballot(ok);
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                const unsigned long long bb = __ballot((dig >> b) & 1u);
                m &= ((dig >> b) & 1u) ? bb : ~bb;
            }
            if (ok) {
                const uint32_t o = off[dig];
                const uint32_t below = lanes_below(m);
                dst[o + below] = (uint16_t)key;
                if (below == 0) off[dig] = o + (uint32_t)__popcll(m);
            }
        }
    }
    __syncthreads();
}

}  // namespace rc

#endif
