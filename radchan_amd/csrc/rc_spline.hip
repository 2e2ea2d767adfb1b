/*
 * rc_spline.hip -- restriction of the temperature field to a coarse grid
 * level by cubic splines (interp3D, src/montecarlo/spline.cpp:36-177), an
 * alternative to the texture average the reference's kernels use for the
 * coarse levels (mc_gpu_noint.cu:334-338; interp3D is not called there,
 * :694).
 *
 * interp3D interpolates the fine field to the coarse cell centres with
 * natural cubic splines along x (every fine (j, k) line), then y (every
 * coarse i, fine k), then z (every coarse (i, j)); the x walls take the
 * spline (along y, then z) of the face means (T_ghost + T_first)/2, and
 * the coarse ghost is 2 face - interior, so the coarse face mean is that
 * spline.  The arithmetic is the reference's spline/splint
 * (spline.cpp:180-226, Numerical Recipes): float data with the double
 * constants there, so that some products are formed in double and rounded
 * to float on assignment -- kept as written, the CPU restatement and this
 * kernel agree bit for bit.
 *
 * One thread per line in every pass; the second derivatives y2 and the
 * sweep's u of line L live in scratch at [i][L] (consecutive lines in
 * consecutive addresses).  Outputs: Tl [km][jm][im] cell temperatures and,
 * when asked, Tw [2][km][jm] the west/east face temperatures.  This runs
 * once per snapshot and level, before the march.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <mutex>

#include "rc_internal.h"
#include "rc_kernels.h"

namespace rc {
namespace {

/* second derivatives of the natural cubic spline through (x[i], y(i)),
 * i = 1..n (spline.cpp:180-207 with yp1 = ypn = 1e30); y2 and u strided */
template <class YF>
__device__ void nr_spline(const float *x, YF y, int n, float *y2, float *u, size_t st) {
    y2[st] = 0.0f;
    u[st] = 0.0f;
    for (int i = 2; i <= n - 1; ++i) {
        const float sig = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1]);
        const float p = (float)((double)(sig * y2[(i - 1) * st]) + 2.0);
        y2[i * st] = (float)(((double)sig - 1.0) / (double)p);
        const float ui = (y(i + 1) - y(i)) / (x[i + 1] - x[i]) - (y(i) - y(i - 1)) / (x[i] - x[i - 1]);
        u[i * st] = (float)((6.0 * (double)ui / (double)(x[i + 1] - x[i - 1]) - (double)(sig * u[(i - 1) * st])) /
                            (double)p);
    }
    const float qn = 0.0f, un = 0.0f;
    y2[n * st] = (float)((double)(un - qn * u[(n - 1) * st]) / ((double)(qn * y2[(n - 1) * st]) + 1.0));
    for (int k = n - 1; k >= 1; --k) y2[k * st] = y2[k * st] * y2[(k + 1) * st] + u[k * st];
}

/* the spline's value at x (spline.cpp:209-226) */
template <class YF>
__device__ float nr_splint(const float *xa, YF ya, const float *y2a, size_t st, int n, float x) {
    int klo = 1, khi = n;
    while (khi - klo > 1) {
        const int k = (khi + klo) >> 1;
        if (xa[k] > x) khi = k; else klo = k;
    }
    const float h = xa[khi] - xa[klo];
    const float a = (xa[khi] - x) / h, b = (x - xa[klo]) / h;
    const float t1 = a * ya(klo) + b * ya(khi);
    const float t2 = ((a * a * a - a) * y2a[klo * st] + (b * b * b - b) * y2a[khi * st]) * (h * h);
    return (float)((double)t1 + (double)t2 / 6.0);
}

struct SplArgs {
    const float *T0, *x0, *y0, *z0, *xc, *yc, *zc;
    int imax, jmax, kmax, im, jm, km;
    float *tx;      /* [kmax][jmax][im] after the x pass  */
    float *txy;     /* [kmax][jm][im]   after the y pass  */
    float *by;      /* [2][kmax][jm]    face means along y */
    float *wk;      /* 2 (n+1) lines: y2 then u            */
    float *Tl, *Tw;
};

__device__ __forceinline__ float T0at(const SplArgs &a, int i, int j, int k) {
    return a.T0[(size_t)i + (size_t)(a.imax + 2) * ((size_t)j + (size_t)(a.jmax + 2) * (size_t)k)];
}

/* x pass: fine line (j, k) -> coarse i = 1..im */
__global__ void __launch_bounds__(256) k_spline_x(SplArgs a) {
    const size_t nl = (size_t)a.jmax * a.kmax;
    const size_t L = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (L >= nl) return;
    const int j = (int)(L % a.jmax) + 1, k = (int)(L / a.jmax) + 1;
    float *y2 = a.wk + L, *u = a.wk + (size_t)(a.imax + 1) * nl + L;
    auto y = [&](int i) { return T0at(a, i, j, k); };
    nr_spline(a.x0, y, a.imax, y2, u, nl);
    float *out = a.tx + ((size_t)(k - 1) * a.jmax + (j - 1)) * a.im;
    for (int i = 1; i <= a.im; ++i) out[i - 1] = nr_splint(a.x0, y, y2, nl, a.imax, a.xc[i]);
}

/* y pass: line (coarse i, fine k) -> coarse j = 1..jm */
__global__ void __launch_bounds__(256) k_spline_y(SplArgs a) {
    const size_t nl = (size_t)a.im * a.kmax;
    const size_t L = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (L >= nl) return;
    const int i = (int)(L % a.im) + 1, k = (int)(L / a.im) + 1;
    float *y2 = a.wk + L, *u = a.wk + (size_t)(a.jmax + 1) * nl + L;
    const float *src = a.tx + (size_t)(k - 1) * a.jmax * a.im + (i - 1);
    auto y = [&](int j) { return src[(size_t)(j - 1) * a.im]; };
    nr_spline(a.y0, y, a.jmax, y2, u, nl);
    float *out = a.txy + (size_t)(k - 1) * a.jm * a.im + (i - 1);
    for (int j = 1; j <= a.jm; ++j) out[(size_t)(j - 1) * a.im] = nr_splint(a.y0, y, y2, nl, a.jmax, a.yc[j]);
}

/* z pass: line (coarse i, coarse j) -> coarse k = 1..km */
__global__ void __launch_bounds__(256) k_spline_z(SplArgs a) {
    const size_t nl = (size_t)a.im * a.jm;
    const size_t L = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (L >= nl) return;
    const int i = (int)(L % a.im) + 1, j = (int)(L / a.im) + 1;
    float *y2 = a.wk + L, *u = a.wk + (size_t)(a.kmax + 1) * nl + L;
    const float *src = a.txy + (size_t)(j - 1) * a.im + (i - 1);
    const size_t sk = (size_t)a.jm * a.im;
    auto y = [&](int k) { return src[(size_t)(k - 1) * sk]; };
    nr_spline(a.z0, y, a.kmax, y2, u, nl);
    float *out = a.Tl + (size_t)(j - 1) * a.im + (i - 1);
    for (int k = 1; k <= a.km; ++k) out[(size_t)(k - 1) * sk] = nr_splint(a.z0, y, y2, nl, a.kmax, a.zc[k]);
}

/* x walls, along y: line (side, fine k) of face means -> coarse j */
__global__ void __launch_bounds__(256) k_spline_wall_y(SplArgs a) {
    const size_t nl = (size_t)2 * a.kmax;
    const size_t L = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (L >= nl) return;
    const int side = (int)(L / a.kmax), k = (int)(L % a.kmax) + 1;
    const int ig = side ? a.imax + 1 : 0, ii = side ? a.imax : 1;
    float *y2 = a.wk + L, *u = a.wk + (size_t)(a.jmax + 1) * nl + L;
    auto y = [&](int j) { return (float)(0.5 * (double)(T0at(a, ig, j, k) + T0at(a, ii, j, k))); };
    nr_spline(a.y0, y, a.jmax, y2, u, nl);
    float *out = a.by + ((size_t)side * a.kmax + (k - 1)) * a.jm;
    for (int j = 1; j <= a.jm; ++j) out[j - 1] = nr_splint(a.y0, y, y2, nl, a.jmax, a.yc[j]);
}

/* x walls, along z: line (side, coarse j) -> coarse k; Tw [2][km][jm] */
__global__ void __launch_bounds__(256) k_spline_wall_z(SplArgs a) {
    const size_t nl = (size_t)2 * a.jm;
    const size_t L = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (L >= nl) return;
    const int side = (int)(L / a.jm), j = (int)(L % a.jm) + 1;
    float *y2 = a.wk + L, *u = a.wk + (size_t)(a.kmax + 1) * nl + L;
    const float *src = a.by + (size_t)side * a.kmax * a.jm + (j - 1);
    auto y = [&](int k) { return src[(size_t)(k - 1) * a.jm]; };
    nr_spline(a.z0, y, a.kmax, y2, u, nl);
    float *out = a.Tw + (size_t)side * a.km * a.jm + (j - 1);
    for (int k = 1; k <= a.km; ++k) out[(size_t)(k - 1) * a.jm] = nr_splint(a.z0, y, y2, nl, a.kmax, a.zc[k]);
}

inline size_t blocks_for(size_t n) { return (n + 255) / 256; }

}  // namespace

size_t spline_scratch_floats(int imax, int jmax, int kmax, int im, int jm, int km) {
    const size_t tx = (size_t)im * jmax * kmax, txy = (size_t)im * jm * kmax, by = (size_t)2 * kmax * jm;
    size_t wk = 2 * (size_t)(imax + 1) * jmax * kmax;
    wk = std::max(wk, 2 * (size_t)(jmax + 1) * im * kmax);
    wk = std::max(wk, 2 * (size_t)(kmax + 1) * im * jm);
    wk = std::max(wk, 2 * (size_t)(jmax + 1) * 2 * kmax);
    wk = std::max(wk, 2 * (size_t)(kmax + 1) * 2 * jm);
    return tx + txy + by + wk;
}

void launch_spline_level(hipStream_t st, const float *T0, const float *x0, const float *y0, const float *z0,
                         const float *xc, const float *yc, const float *zc, int imax, int jmax, int kmax, int im,
                         int jm, int km, float *Tl, float *Tw, float *scratch) {
    SplArgs a;
    a.T0 = T0; a.x0 = x0; a.y0 = y0; a.z0 = z0; a.xc = xc; a.yc = yc; a.zc = zc;
    a.imax = imax; a.jmax = jmax; a.kmax = kmax; a.im = im; a.jm = jm; a.km = km;
    a.tx = scratch;
    a.txy = a.tx + (size_t)im * jmax * kmax;
    a.by = a.txy + (size_t)im * jm * kmax;
    a.wk = a.by + (size_t)2 * kmax * jm;
    a.Tl = Tl; a.Tw = Tw;
    hipLaunchKernelGGL(k_spline_x, dim3(blocks_for((size_t)jmax * kmax)), dim3(256), 0, st, a);
    hipLaunchKernelGGL(k_spline_y, dim3(blocks_for((size_t)im * kmax)), dim3(256), 0, st, a);
    hipLaunchKernelGGL(k_spline_z, dim3(blocks_for((size_t)im * jm)), dim3(256), 0, st, a);
    if (Tw) {
        hipLaunchKernelGGL(k_spline_wall_y, dim3(blocks_for((size_t)2 * kmax)), dim3(256), 0, st, a);
        hipLaunchKernelGGL(k_spline_wall_z, dim3(blocks_for((size_t)2 * jm)), dim3(256), 0, st, a);
    }
}

/* ---- the allocation-free entry point --------------------------------

rc_spline_restrict_noalloc works in `scratch`, a file-scope device array
of RC_SPLINE_NOALLOC_FLOATS floats that every GPU the library runs on holds
until the process ends.  The host never passes the array's name to a kernel
(in host code it is a stand-in, not a device address): the device address
is looked up once per device with hipGetSymbolAddress, with that device
current, and kept in scratch_ptr[device]. */

namespace {

constexpr size_t kNoallocFloats = RC_SPLINE_NOALLOC_FLOATS;
static __device__ __align__(256) float scratch[kNoallocFloats];

constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
float *scratch_ptr[kMaxDevices] = {};

/* the device address of this device's scratch (nullptr and a message on failure) */
float *device_scratch(int dev) {
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("rc_spline_restrict_noalloc: device ordinal beyond 63");
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(scratch_mu);
    if (!scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error(std::string("rc_spline_restrict_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
            return nullptr;
        }
        scratch_ptr[dev] = static_cast<float *>(p);
    }
    return scratch_ptr[dev];
}

/* working floats of the grid step: f[im0+1], then the centres x0, y0, z0
 * of level 0 and xc, yc, zc of the coarse level (n+2 each), rounded up to
 * 64 floats so the spline scratch after it stays 256-byte aligned */
size_t grid_floats(int im0, int jm0, int km0, int im, int jm, int km) {
    const size_t n = (size_t)(im0 + 1) + (im0 + 2) + (jm0 + 2) + (km0 + 2) + (im + 2) + (jm + 2) + (km + 2);
    return (n + 63) / 64 * 64;
}

struct GridArgs {
    int im0, jm0, km0, im, jm, km;
    float Lx, Ly, Lz;
    double fact_mesh;
    float *f, *x0, *y0, *z0, *xc, *yc, *zc;
};

/* centres of n uniform cells of length L with mirrored ghosts (the y and
 * z directions of Gridn::mk_grid, definitions.h:85-105) */
__device__ void uniform_centres(float L, int n, float *c) {
    const float d = L * 1.0f / (float)n;
    c[0] = -d / 2.0f;
    for (int j = 1; j <= n; ++j) c[j] = ((float)(j - 1) * d + (float)j * d) / 2.0f;
    c[n + 1] = c[n] + d;
}

/* x centres of a level with im cells from the fine faces f[0..im0]
 * (definitions.h:125-134) */
__device__ void x_centres(const float *f, int im0, int im, float *x) {
    float prev = 0.f;
    for (int i = 1; i <= im; ++i) {
        const float xu = f[(size_t)i * im0 / im];
        x[i] = (float)(0.5 * (double)(xu + prev));
        prev = xu;
    }
    x[0] = -x[1];
    x[im + 1] = prev + (prev - x[im]);
}

/* Gridn::mk_grid (definitions.h:68-136) of level 0 and of the coarse
 * level, in one thread: the operations of the host's build_grid in its
 * order, so the centres are equal to its bit for bit */
__global__ void k_level_grid(GridArgs a) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    float *f = a.f;
    const int half = a.im0 / 2;
    f[0] = 0.f;
    for (int i = 1; i <= half; ++i) {
        const double r = (double)((float)i / (float)a.im0) - 0.5;
        f[i] = f[i - 1] + (float)(0.5 - a.fact_mesh * r * r);
    }
    const double norm = (double)f[half];
    const double scale = (double)a.Lx / 2.;
    for (int i = 1; i <= half; ++i) f[i] = (float)(scale * (double)f[i] / norm);
    for (int i = a.im0; i > half; --i) f[i] = a.Lx - f[a.im0 - i];
    x_centres(f, a.im0, a.im0, a.x0);
    x_centres(f, a.im0, a.im, a.xc);
    uniform_centres(a.Ly, a.jm0, a.y0);
    uniform_centres(a.Lz, a.km0, a.z0);
    uniform_centres(a.Ly, a.jm, a.yc);
    uniform_centres(a.Lz, a.km, a.zc);
}

}  // namespace

}  // namespace rc

/* allocation-free rc_spline_restrict on device arrays (declared in
 * include/radchan.h): grid and spline work in this device's scratch,
 * asynchronous on `stream` */
extern "C" int rc_spline_restrict_noalloc(const rc_params *p, int level, const float *T, float *Tc, float *Tw,
                                          void *stream) {
    using namespace rc;
    if (!p || !T || !Tc) { set_error("rc_spline_restrict_noalloc: null argument"); return -1; }
    if (int e = check_params(*p)) return e;
    if (level < 1 || level >= p->grid_num) { set_error("rc_spline_restrict_noalloc: level in 1..grid_num-1"); return -2; }
    if (p->imax < 2 || p->jmax < 2 || p->kmax < 2) { set_error("cubic-spline restriction needs imax, jmax, kmax >= 2"); return -2; }
    const int im = p->maxi[level], jm = p->maxj[level], km = p->maxk[level];
    const size_t ng = grid_floats(p->imax, p->jmax, p->kmax, im, jm, km);
    const size_t need = ng + spline_scratch_floats(p->imax, p->jmax, p->kmax, im, jm, km);
    if (need > kNoallocFloats) {
        set_error("rc_spline_restrict_noalloc: the field needs " + std::to_string(need) +
                  " floats of scratch, beyond its fixed " + std::to_string(kNoallocFloats) +
                  " (use rc_spline_restrict)");
        return -2;
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) { set_error("no HIP device available"); return -3; }
    float *w = device_scratch(dev);
    if (!w) return -3;
    GridArgs g;
    g.im0 = p->imax; g.jm0 = p->jmax; g.km0 = p->kmax; g.im = im; g.jm = jm; g.km = km;
    g.Lx = p->Lx; g.Ly = p->Ly; g.Lz = p->Lz; g.fact_mesh = p->fact_mesh;
    g.f = w;
    g.x0 = g.f + (p->imax + 1);
    g.y0 = g.x0 + (p->imax + 2);
    g.z0 = g.y0 + (p->jmax + 2);
    g.xc = g.z0 + (p->kmax + 2);
    g.yc = g.xc + (im + 2);
    g.zc = g.yc + (jm + 2);
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_level_grid, dim3(1), dim3(1), 0, st, g);
    launch_spline_level(st, T, g.x0, g.y0, g.z0, g.xc, g.yc, g.zc, p->imax, p->jmax, p->kmax, im, jm, km, Tc, Tw,
                        w + ng);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) { set_error(std::string("rc_spline_restrict_noalloc: ") + hipGetErrorString(e)); return -4; }
    return 0;
}
