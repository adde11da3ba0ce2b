// Camera post-process on the device: camera::tone_map for its global modes
// (rt/camera.cc:1124-1461), camera::gamma_correct (rt/camera.cc:1466-1501)
// and the down-sampling clip to bytes of the image writers
// (rt/camera.h:289-351). Images are camera::image layouts: 0-255 floats,
// pixel (x, y) at 3 * (x + y * x_res).
//
// Tone mapping is three passes over the image (HBM-bound, 12 bytes read and
// written per pixel and pass):
//   k_tone_yxy      rgb -> (x, Y, y) in place (with the desaturation), and
//                   per block: the sum of log10 Y (double), max Y, count
//   k_tone_scalars  one block: the log-average luminance, the temporal
//                   adaptation, the mode's scale factors
//   k_tone_apply    the mode's luminance mapping and (x, Y, y) -> rgb; with
//                   the contrast stretch it first leaves the mapped Y and
//                   per-block min / max, and k_tone_stretch finishes.
// The block partials are combined in block order, so results repeat bit for
// bit; the reference sums log10 Y in float in pixel order, which the double
// sum here approximates more closely than that float sum does.
#include <algorithm>
#include <cmath>
#include <fstream>
#include <string>
#include <type_traits>
#include <vector>

#include "rt_internal.hpp"

namespace rt {

namespace {

constexpr int kPostThreads = 256;
constexpr int kPostMaxBlocks = 2048;
constexpr float kMaxDistF = 3.402823466e+38f;

// tone_mapping_mode_t (rt/camera.h:43-45).
enum ToneMode : int {
    kNone = 0, kContrast = 1, kHistogram = 2, kHumanHistogram = 3, kNonLinear = 4, kExposure = 5, kAvgLum = 6, kMaxLum = 7, kBilateral = 8, kFerwerda = 9
};

struct TonePartial {
    double log_sum;
    float max_y;
    int32_t count;
};

struct ToneScalars {
    float sf;         // luminance scale
    float max_y;      // global_non_linear: 1 / (sf * maxY)^2; global_exposure: sf * maxY
    float mp, ms, k;  // global_ferwerda
    float yi;         // adapted log-average luminance
    float level;      // adaptation level after this frame
    float offset, scale;  // contrast stretch
    int32_t stretch_on;
};

template <class T>
__device__ __forceinline__ T block_reduce(T v, T (*op)(T, T), T *lds) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = op(v, __shfl_down(v, off, 64));
    if (lane == 0) lds[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        T r = lds[0];
        for (int w = 1; w < kPostThreads / 64; ++w) r = op(r, lds[w]);
        lds[0] = r;
    }
    __syncthreads();
    const T r = lds[0];
    __syncthreads();
    return r;
}

__device__ double add_d(double a, double b) { return a + b; }
__device__ float max_f(float a, float b) { return fmaxf(a, b); }
__device__ float min_f(float a, float b) { return fminf(a, b); }
__device__ int add_i(int a, int b) { return a + b; }

struct ToneArgs {
    float *img;
    int64_t n;
    int mode;
    float key, xw, yw;
    int desaturate, stretch, adapt;
    float time_step, level_in;
};

// The histogram modes' mapping (rt/camera.cc:1843-2062): the branch the
// reference takes and the cumulative distribution it maps through.
constexpr int kHistBins = 100;
struct HistState {
    int branch;  // 0 unchanged, 1 linear g * sf + bd_min, 2 g * sf, 3 the distribution
    float sf, bw_min, bw_max, bd_min, bd_max;
    float cdf[kHistBins + 1];
};

// Per pixel: rt/camera.cc:1137-1195.
__global__ void __launch_bounds__(kPostThreads) k_tone_yxy(ToneArgs a, TonePartial *__restrict__ part) {
    __shared__ double lds_d[kPostThreads / 64];
    __shared__ float lds_f[kPostThreads / 64];
    __shared__ int lds_i[kPostThreads / 64];
    const float xb[3] = {0.4124f, 0.3576f, 0.1805f};
    const float yb[3] = {0.2126f, 0.7151f, 0.0721f};
    const float zb[3] = {0.0193f, 0.1192f, 0.9505f};
    const float vb[3] = {0.01477f, 0.497f, 0.631f};
    double log_sum = 0.0;
    float max_y = 0.0f;
    int count = 0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kPostThreads + threadIdx.x; i < a.n;
         i += static_cast<int64_t>(gridDim.x) * kPostThreads) {
        float *p = a.img + 3 * i;
        const float r = p[0] / 255.0f, g = p[1] / 255.0f, b = p[2] / 255.0f;
        const float X = 683.0f * ((xb[0] * r) + (xb[1] * g) + (xb[2] * b));
        float Y = 683.0f * ((yb[0] * r) + (yb[1] * g) + (yb[2] * b));
        const float Z = 683.0f * ((zb[0] * r) + (zb[1] * g) + (zb[2] * b));
        if (Y < 3.18e-7f) {
            Y = 3.18e-7f;
        } else {
            ++count;
            max_y = fmaxf(max_y, Y);
            log_sum += static_cast<double>(log10f(Y));
        }
        const float W = Y + X + Z;
        float x = X / W;
        float y = Y / W;
        if (a.desaturate) {
            float s;
            // The reference's unqualified log10 of a float is the double one.
            const float log10_Y = static_cast<float>(log10(static_cast<double>(Y)));
            if (log10_Y < -2.0f) {
                s = 0.0f;
            } else if (log10_Y < 0.6f) {
                const float sl = (log10_Y + 2.0f) / 2.6f;
                const float sl2 = sl * sl;
                s = (3.0f * sl2) - (2.0f * sl * sl2);
            } else {
                s = 1.0f;
            }
            const float V = 1700.0f * ((vb[0] * r) + (vb[1] * g) + (vb[2] * b));
            x = ((1.0f - s) * a.xw) + (s * (x + a.xw - (1.0f / 3.0f)));
            y = ((1.0f - s) * a.yw) + (s * (y + a.yw - (1.0f / 3.0f)));
            Y = (0.4468f * (1.0f - s) * V) + (s * Y);
        }
        p[0] = x;
        p[1] = Y;
        p[2] = y;
    }
    log_sum = block_reduce<double>(log_sum, add_d, lds_d);
    max_y = block_reduce<float>(max_y, max_f, lds_f);
    count = block_reduce<int>(count, add_i, lds_i);
    if (threadIdx.x == 0) part[blockIdx.x] = TonePartial{log_sum, max_y, count};
}

// One block: rt/camera.cc:1196-1414 (everything that is not per pixel).
// Thread t folds partials t, t + 256, ... and the block reduces the 256
// results in a fixed tree, so the order never changes.
__global__ void __launch_bounds__(kPostThreads) k_tone_scalars(ToneArgs a, const TonePartial *__restrict__ part,
                                                               int nparts, ToneScalars *__restrict__ out) {
    __shared__ double lds_d[kPostThreads / 64];
    __shared__ float lds_f[kPostThreads / 64];
    __shared__ int lds_i[kPostThreads / 64];
    double log_sum = 0.0;
    float maxY = 0.0f;
    int numpix = 0;
    for (int b = threadIdx.x; b < nparts; b += kPostThreads) {
        const TonePartial q = part[b];
        log_sum += q.log_sum;
        maxY = fmaxf(maxY, q.max_y);
        numpix += q.count;
    }
    log_sum = block_reduce<double>(log_sum, add_d, lds_d);
    maxY = block_reduce<float>(maxY, max_f, lds_f);
    numpix = block_reduce<int>(numpix, add_i, lds_i);
    if (threadIdx.x != 0) return;
    float Yi = static_cast<float>(log_sum);
    if (numpix) Yi /= static_cast<float>(numpix);
    Yi = powf(10.0f, Yi);
    ToneScalars t{};
    if (a.adapt && a.time_step > 0.0f) {
        float k;
        const float log_la = logf(Yi);
        if (log_la < -2.0f) {
            k = 1.0f;
        } else if (log_la < 1.0f) {
            k = 1.0f - ((log_la + 2.0f) * (1.0f / 3.0f));
        } else {
            k = 0.0f;
        }
        const float lvl = a.level_in, ts = a.time_step;
        float level;
        if (lvl > Yi) {
            const float a_cone = Yi + ((lvl - Yi) * expf(-(ts / 0.2f)));
            const float a_rod = Yi + ((lvl - Yi) * expf(-(ts / 6.0f)));
            level = ((1.0f - k) * a_cone) + (k * a_rod);
        } else {
            const float r = (Yi / lvl);
            const float c1 = lvl * (powf(r, 0.9f) - 1.0f);
            const float c2 = Yi * (1.0f - powf(r, -0.1f));
            const float a_cone = Yi - (c1 * expf(-(ts / 16.0e-3f))) - (c2 * expf(-(ts / 200.0e-3f)));
            const float a_rod = Yi - (c1 * expf(-(ts / 40.0e-3f))) - (c2 * expf(-(ts / 6.0f)));
            level = ((1.0f - k) * a_cone) + (k * a_rod);
        }
        t.level = level;
        Yi = level;
    } else {
        t.level = Yi;
    }
    t.yi = Yi;
    t.sf = 1.0f;
    switch (a.mode) {
        case kContrast: {
            const float Yw = 300.0f;
            const float num = 1.219f + powf(Yw * 0.5f, 0.4f);
            const float den = 1.219f + powf(Yi, 0.4f);
            t.sf = (1.0f / Yw) * powf((num / den), 2.5f);
            break;
        }
        case kNonLinear: {
            t.sf = a.key / Yi;
            float m = maxY * t.sf;
            t.max_y = 1.0f / (m * m);
            break;
        }
        case kExposure:
            t.sf = a.key / Yi;
            t.max_y = maxY * t.sf;
            break;
        case kAvgLum:
            // The reference averages rgb only for global_exposure
            // (rt/camera.cc:1141-1145), so here the average is 0 and the
            // scale stays 1 (rt/camera.cc:1331-1342).
            t.sf = 1.0f;
            break;
        case kMaxLum:
            t.sf = a.key / maxY;
            break;
        case kBilateral:  // the bilateral kernels have written the mapped Y
        case kHistogram:  // k_hist_build decides the mapping
        case kHumanHistogram:
            t.sf = 1.0f;
            break;
        case kFerwerda: {
            const float lda = 800.0f;
            const float log_lda = logf(lda * 0.5f);
            const float log_la = logf(Yi);
            float tpc;
            if (log_la <= -2.6f) {
                tpc = -0.72f;
            } else if (log_la <= 1.9f) {
                tpc = powf(((0.249f * log_la) + 0.65f), 2.7f) - 0.72f;
            } else {
                tpc = log_la - 1.255f;
            }
            t.mp = (log_lda - 1.255f) / tpc;
            float tsc;
            if (log_la <= -3.94f) {
                tsc = -2.86f;
            } else if (log_la < -1.44f) {
                tsc = powf(((0.405f * log_la) + 1.6f), 2.18f) - 2.86f;
            } else {
                tsc = log_la - 0.395f;
            }
            t.ms = (log_lda - 1.255f) / tsc;
            if (log_la < -2.0f) {
                t.k = 1.0f;
            } else if (log_la < 1.0f) {
                t.k = 1.0f - ((log_la + 2.0f) * (1.0f / 3.0f));
            } else {
                t.k = 0.0f;
            }
            break;
        }
        default:
            t.sf = 1.0f / 683.0f;
            break;
    }
    *out = t;
}

__device__ __forceinline__ float map_histogram(const HistState &h, float g) {
    if (h.branch == 1) return g * h.sf + h.bd_min;
    if (h.branch == 2) return g * h.sf;
    if (h.branch == 3 && g > 0.0f) {
        const float bw = logf(g);
        float bf = (bw - h.bw_min) / (h.bw_max - h.bw_min) * kHistBins;
        bf = fminf(fmaxf(bf, 0.0f), static_cast<float>(kHistBins));
        const int b0 = static_cast<int>(bf);
        const int b1 = min((b0 + 1), kHistBins);
        const float s = bf - b0;
        const float pbw = (1 - s) * h.cdf[b0] + s * h.cdf[b1];
        const float bde = h.bd_min + (h.bd_max - h.bd_min) * pbw;
        g = expf(bde);
        g /= 300.0f;
    }
    return g;
}

__device__ __forceinline__ float map_luminance(int mode, const ToneScalars &t, float g) {
    switch (mode) {
        case kNonLinear:
            g *= t.sf;
            if (t.max_y) g *= (1.0f + g * t.max_y) / (1.0f + g);
            return g;
        case kExposure:
            g *= t.sf;
            if (t.max_y) g = 1.0f - expf(-g);
            return g;
        case kFerwerda: {
            const float lda = 800.0f;
            const float ldp = t.mp * g;
            const float lds = t.ms * g;
            return (1.0f / lda) * (ldp + (t.k * lds));
        }
        default:
            return g * t.sf;
    }
}

// camera::convert_Yxy_to_rgb (rt/camera.cc:2080-2099).
__device__ __forceinline__ void yxy_to_rgb(float *p, float x, float Y, float y) {
    const float X = x * Y / y;
    const float Z = (1.0f - x - y) * Y / y;
    p[0] = (826.353f * X) + (-391.986f * Y) + (-127.143f * Z);
    p[1] = (-247.0695f * X) + (478.329f * Y) + (10.5825f * Z);
    p[2] = (14.2035f * X) + (-52.02f * Y) + (269.535f * Z);
}

// Mapped luminance, then rgb; with the stretch, the mapped Y and per-block
// min / max only.
__global__ void __launch_bounds__(kPostThreads) k_tone_apply(ToneArgs a, const ToneScalars *__restrict__ ts,
                                                             const HistState *__restrict__ hist,
                                                             float2 *__restrict__ minmax) {
    __shared__ float lds_f[kPostThreads / 64];
    __shared__ HistState hs;
    const ToneScalars t = *ts;
    const bool histogram = a.mode == kHistogram || a.mode == kHumanHistogram;
    if (histogram) {
        if (threadIdx.x == 0) hs = *hist;
        __syncthreads();
    }
    float lo = kMaxDistF, hi = -kMaxDistF;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kPostThreads + threadIdx.x; i < a.n;
         i += static_cast<int64_t>(gridDim.x) * kPostThreads) {
        float *p = a.img + 3 * i;
        const float g = histogram ? map_histogram(hs, p[1]) : map_luminance(a.mode, t, p[1]);
        if (a.stretch) {
            p[1] = g;
            lo = fminf(lo, g);
            hi = fmaxf(hi, g);
        } else {
            yxy_to_rgb(p, p[0], g, p[2]);
        }
    }
    if (!a.stretch) return;
    lo = block_reduce<float>(lo, min_f, lds_f);
    hi = block_reduce<float>(hi, max_f, lds_f);
    if (threadIdx.x == 0) minmax[blockIdx.x] = make_float2(lo, hi);
}

// rt/camera.cc:1417-1436.
__global__ void __launch_bounds__(kPostThreads) k_stretch_scalars(const float2 *__restrict__ minmax, int nparts,
                                                                  ToneScalars *__restrict__ ts) {
    __shared__ float lds_f[kPostThreads / 64];
    float minY2 = kMaxDistF, maxY2 = -kMaxDistF;
    for (int b = threadIdx.x; b < nparts; b += kPostThreads) {
        const float2 q = minmax[b];
        minY2 = fminf(minY2, q.x);
        maxY2 = fmaxf(maxY2, q.y);
    }
    minY2 = block_reduce<float>(minY2, min_f, lds_f);
    maxY2 = block_reduce<float>(maxY2, max_f, lds_f);
    if (threadIdx.x != 0) return;
    ToneScalars t = *ts;
    t.stretch_on = ((minY2 > 0.0f) || (maxY2 < 0.0f)) && (minY2 < maxY2);
    t.offset = fminf(-minY2, 0.0f);
    t.scale = 1.0f / (fminf(1.0f, maxY2) - fmaxf(0.0f, minY2));
    *ts = t;
}

__global__ void __launch_bounds__(kPostThreads) k_tone_stretch(float *__restrict__ img, int64_t n,
                                                               const ToneScalars *__restrict__ ts) {
    const ToneScalars t = *ts;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kPostThreads + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * kPostThreads) {
        float *p = img + 3 * i;
        float g = p[1];
        if (t.stretch_on) g = (g + t.offset) * t.scale;
        yxy_to_rgb(p, p[0], g, p[2]);
    }
}

// camera::gamma_correct (rt/camera.cc:1466-1501): a gamma within 0.001 of
// 2.2 takes the sRGB curve, any other the plain power (the reference's
// branch order).
__global__ void __launch_bounds__(kPostThreads) k_gamma(float *__restrict__ img, int64_t n, float gamma) {
    const bool plain = fabsf(gamma - 2.2f) > 0.001f;
    const bool srgb = !plain && fabsf(gamma - 1.0f) > 0.001f;
    const float inv = 1.0f / gamma;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kPostThreads + threadIdx.x; i < 3 * n;
         i += static_cast<int64_t>(gridDim.x) * kPostThreads) {
        float v = img[i] / 255.0f;
        v = fminf(fmaxf(v, 0.0f), 1.0f);
        if (plain) {
            v = powf(v, inv);
        } else if (srgb) {
            v = (v <= 0.0031308f) ? (12.92f * v) : (1.055f * powf(v, (1.0f / 2.4f)) - 0.055f);
        }
        img[i] = v * 255.0f;
    }
}

// clip_image_to_rgb / clip_image_to_bgr (rt/camera.h:289-351): one output
// pixel per thread, its xs x ys samples summed column by column.
__global__ void __launch_bounds__(kPostThreads) k_clip(const float *__restrict__ img, int x_res, int xs, int ys,
                                                       int ox, int64_t nout, int bgr, uint8_t *__restrict__ out) {
    const int samples = xs * ys;
    const float inv = 1.0f / static_cast<float>(xs * ys);
    for (int64_t o = static_cast<int64_t>(blockIdx.x) * kPostThreads + threadIdx.x; o < nout;
         o += static_cast<int64_t>(gridDim.x) * kPostThreads) {
        const int x = static_cast<int>(o % ox), y = static_cast<int>(o / ox);
        const int64_t addr = static_cast<int64_t>(x) * xs + static_cast<int64_t>(y) * ys * x_res;
        float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
        for (int ax = 0; ax < xs; ++ax)
            for (int ay = 0; ay < ys; ++ay) {
                const float *p = img + 3 * (addr + static_cast<int64_t>(ay) * x_res + ax);
                a0 += p[0];
                a1 += p[1];
                a2 += p[2];
            }
        uint8_t *q = out + 3 * o;
        if (bgr) {
            q[0] = static_cast<uint8_t>(fminf(255.0f, fmaxf(0.0f, (a2 * inv))));
            q[1] = static_cast<uint8_t>(fminf(255.0f, fmaxf(0.0f, (a1 * inv))));
            q[2] = static_cast<uint8_t>(fminf(255.0f, fmaxf(0.0f, (a0 * inv))));
        } else {
            q[0] = static_cast<uint8_t>(fminf(255.0f, fmaxf(0.0f, (a0 / static_cast<float>(samples)))));
            q[1] = static_cast<uint8_t>(fminf(255.0f, fmaxf(0.0f, (a1 / static_cast<float>(samples)))));
            q[2] = static_cast<uint8_t>(fminf(255.0f, fmaxf(0.0f, (a2 / static_cast<float>(samples)))));
        }
    }
}

// ---- global_bilateral_filter: camera::bilateral_filter_tone_map
// (rt/camera.cc:1632-1775), the bilateral grid of Paris and Durand: log10 Y
// splatted into a grid coarser by s_sa in x and y and by 0.4 in log
// intensity, blurred by a circular Gaussian of one cell in each axis, sliced
// back per pixel by trilinear interpolation (the base layer), and the base
// compressed to a contrast of 300 while the detail (log Y - base) is kept.
// The reference convolves with single-precision FFTs; a circular convolution
// is that same sum, evaluated here as three separable passes in double. Every
// sum has a fixed order (the splat reduces per-lane partials in lane order),
// so frames repeat bit for bit.

struct BilGrid {
    int W, H, D;
    int pxy, pz;
    float s_sa, r_sa;
    float in_min;
    int64_t N;  // cells per channel
};

// log10 Y of the bright pixels (the reference's unqualified log10 on a float
// is the double one), 0 for the others (never written in the reference; a
// fresh allocation of a frame reads zero), and per-block min / max of the
// bright ones.
__global__ void __launch_bounds__(kPostThreads) k_bil_log(const float *__restrict__ img, int64_t n,
                                                          float *__restrict__ logi, float2 *__restrict__ part) {
    __shared__ float lds_f[kPostThreads / 64];
    float lo = kMaxDistF, hi = -kMaxDistF;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kPostThreads + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * kPostThreads) {
        const float g = img[3 * i + 1];
        float l = 0.0f;
        if (g > 3.18e-7f) {
            l = static_cast<float>(log10(static_cast<double>(g)));
            lo = fminf(lo, l);
            hi = fmaxf(hi, l);
        }
        logi[i] = l;
    }
    lo = block_reduce<float>(lo, min_f, lds_f);
    hi = block_reduce<float>(hi, max_f, lds_f);
    if (threadIdx.x == 0) part[blockIdx.x] = make_float2(lo, hi);
}

__global__ void __launch_bounds__(kPostThreads) k_minmax_fold(const float2 *__restrict__ part, int nparts,
                                                              float2 *__restrict__ out) {
    __shared__ float lds_f[kPostThreads / 64];
    float lo = kMaxDistF, hi = -kMaxDistF;
    for (int b = threadIdx.x; b < nparts; b += kPostThreads) {
        lo = fminf(lo, part[b].x);
        hi = fmaxf(hi, part[b].y);
    }
    lo = block_reduce<float>(lo, min_f, lds_f);
    hi = block_reduce<float>(hi, max_f, lds_f);
    if (threadIdx.x == 0) *out = make_float2(lo, hi);
}

__global__ void k_fill_zero(float *__restrict__ p, int64_t n) {
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        p[i] = 0.0f;
}

__device__ __forceinline__ int bil_cell_x(float x, float s_sa, int pxy) {
    return static_cast<int>(x / s_sa + 0.5f) + pxy;
}

__device__ __forceinline__ int bil_cell_z(float l, const BilGrid &g) {
    const int z = static_cast<int>((l - g.in_min) / g.r_sa + 0.5f) + g.pz;
    return min(max(z, 0), g.D - 1);
}

constexpr int kBilZChunk = 32;

// One wave per (x, y) cell: the lanes take the pixels of a window around the
// cell in a fixed order, keep those that round to the cell, and add count
// and log intensity into per-lane LDS columns, kBilZChunk depth cells at a
// time; lane z then sums row z over the lanes in lane order.
__global__ void __launch_bounds__(64) k_bil_splat(BilGrid g, const float *__restrict__ logi, int x_res, int y_res,
                                                  int cells_x, float *__restrict__ grid) {
    __shared__ float cnt[kBilZChunk][64];
    __shared__ float sum[kBilZChunk][64];
    const int lane = threadIdx.x;
    const int cx = g.pxy + static_cast<int>(blockIdx.x % cells_x), cy = g.pxy + static_cast<int>(blockIdx.x / cells_x);
    // Pixels whose rounded cell is (cx, cy) lie within s_sa / 2 of its centre.
    const float fx = static_cast<float>(cx - g.pxy) * g.s_sa, fy = static_cast<float>(cy - g.pxy) * g.s_sa;
    const int x0 = max(0, static_cast<int>(floorf(fx - 0.5f * g.s_sa)) - 1);
    const int x1 = min(x_res - 1, static_cast<int>(ceilf(fx + 0.5f * g.s_sa)) + 1);
    const int y0 = max(0, static_cast<int>(floorf(fy - 0.5f * g.s_sa)) - 1);
    const int y1 = min(y_res - 1, static_cast<int>(ceilf(fy + 0.5f * g.s_sa)) + 1);
    const int ww = x1 - x0 + 1, npix = ww * (y1 - y0 + 1);
    for (int z0 = 0; z0 < g.D; z0 += kBilZChunk) {
        for (int z = 0; z < kBilZChunk; ++z) {
            cnt[z][lane] = 0.0f;
            sum[z][lane] = 0.0f;
        }
        for (int k = lane; k < npix; k += 64) {
            const int x = x0 + k % ww, y = y0 + k / ww;
            if (bil_cell_x(static_cast<float>(x), g.s_sa, g.pxy) != cx ||
                bil_cell_x(static_cast<float>(y), g.s_sa, g.pxy) != cy)
                continue;
            const float l = logi[x + static_cast<int64_t>(x_res) * y];
            const int z = bil_cell_z(l, g) - z0;
            if (z < 0 || z >= kBilZChunk) continue;
            cnt[z][lane] += 1.0f;
            sum[z][lane] += l;
        }
        __syncthreads();
        if (lane < kBilZChunk && z0 + lane < g.D) {
            float c = 0.0f, s = 0.0f;
            for (int j = 0; j < 64; ++j) {
                c += cnt[lane][j];
                s += sum[lane][j];
            }
            const int64_t cell = cx + static_cast<int64_t>(g.W) * (cy + static_cast<int64_t>(g.H) * (z0 + lane));
            grid[cell] = c;
            grid[cell + g.N] = s;
        }
        __syncthreads();
    }
}

// One separable pass of the circular Gaussian along axis (0 x, 1 y, 2 z) of
// both channels: out[o] = sum_j in[o - j mod n] f[j].
__global__ void __launch_bounds__(kPostThreads) k_bil_conv(BilGrid g, int axis, const float *__restrict__ f,
                                                           const float *__restrict__ in, float *__restrict__ out) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kPostThreads + threadIdx.x;
    if (i >= 2 * g.N) return;
    const int64_t c = i / g.N, r = i - c * g.N;
    const int x = static_cast<int>(r % g.W), y = static_cast<int>((r / g.W) % g.H), z = static_cast<int>(r / (static_cast<int64_t>(g.W) * g.H));
    const int n = axis == 0 ? g.W : (axis == 1 ? g.H : g.D);
    const int o = axis == 0 ? x : (axis == 1 ? y : z);
    const int64_t stride = axis == 0 ? 1 : (axis == 1 ? g.W : static_cast<int64_t>(g.W) * g.H);
    const float *src = in + i - static_cast<int64_t>(o) * stride;
    double acc = 0.0;
    for (int j = 0; j < n; ++j) {
        int s = o - j;
        s += (s < 0) ? n : 0;
        acc += static_cast<double>(src[s * stride]) * static_cast<double>(f[j]);
    }
    out[i] = static_cast<float>(acc);
}

// Base layer per pixel: the trilinear slice of the blurred grid, IW / W, and
// per-block min / max (rt/camera.cc:1739-1773).
__global__ void __launch_bounds__(kPostThreads) k_bil_slice(BilGrid g, const float *__restrict__ grid,
                                                            const float *__restrict__ logi, int x_res, int64_t n,
                                                            float *__restrict__ base, float2 *__restrict__ part) {
    __shared__ float lds_f[kPostThreads / 64];
    float lo = kMaxDistF, hi = -kMaxDistF;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kPostThreads + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * kPostThreads) {
        const int x = static_cast<int>(i % x_res), y = static_cast<int>(i / x_res);
        const float z = logi[i] - g.in_min;
        const float xa = static_cast<float>(x) / g.s_sa + g.pxy;
        const int xi = max(0, min(static_cast<int>(xa), g.W - 1)), xxi = max(0, min(xi + 1, g.W - 1));
        const float ya = static_cast<float>(y) / g.s_sa + g.pxy;
        const int yi = max(0, min(static_cast<int>(ya), g.H - 1)), yyi = max(0, min(yi + 1, g.H - 1));
        const float za = z / g.r_sa + g.pz;
        const int zi = max(0, min(static_cast<int>(za), g.D - 1)), zzi = max(0, min(zi + 1, g.D - 1));
        const float ax = xa - xi, ay = ya - yi, az = za - zi;
        float v[2];
        for (int c = 0; c < 2; ++c) {
            const float *q = grid + c * g.N;
            auto at = [&](int a, int b, int d) { return q[a + static_cast<int64_t>(g.W) * (b + static_cast<int64_t>(g.H) * d)]; };
            v[c] = (1.0f - ax) * (1.0f - ay) * (1.0f - az) * at(xi, yi, zi) +
                   ax * (1.0f - ay) * (1.0f - az) * at(xxi, yi, zi) +
                   (1.0f - ax) * ay * (1.0f - az) * at(xi, yyi, zi) +
                   ax * ay * (1.0f - az) * at(xxi, yyi, zi) +
                   (1.0f - ax) * (1.0f - ay) * az * at(xi, yi, zzi) +
                   ax * (1.0f - ay) * az * at(xxi, yi, zzi) +
                   (1.0f - ax) * ay * az * at(xi, yyi, zzi) + ax * ay * az * at(xxi, yyi, zzi);
        }
        const float b = v[1] / v[0];
        base[i] = b;
        lo = fminf(lo, b);
        hi = fmaxf(hi, b);
    }
    lo = block_reduce<float>(lo, min_f, lds_f);
    hi = block_reduce<float>(hi, max_f, lds_f);
    if (threadIdx.x == 0) part[blockIdx.x] = make_float2(lo, hi);
}

// Y = 10^(base * gamma + (log Y - base)) / 255 with gamma = log10(300) /
// (max base - min base) (rt/camera.cc:1777-1784).
__global__ void __launch_bounds__(kPostThreads) k_bil_apply(float *__restrict__ img, int64_t n,
                                                            const float *__restrict__ logi,
                                                            const float *__restrict__ base,
                                                            const float2 *__restrict__ range, float log10_contrast) {
    const float2 r = *range;
    const float gamma = log10_contrast / (r.y - r.x);
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kPostThreads + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * kPostThreads) {
        const float b = base[i];
        const float v = powf(10.0f, (b * gamma + (logi[i] - b)));
        img[3 * i + 1] = v * (1.0f / 255.0f);
    }
}

// ---- local_histogram / local_human_histogram (rt/camera.cc:1843-2062).

// One wave per sample of the one-sample-per-degree image: the block of
// pixels the reference averages, divided by its float extent. The block
// bounds follow the reference's running sums x1 += inc (clamped to
// x_res - 1), replayed per wave; lane r sums row y0 + r (and every 64th row
// after it) left to right, and the rows are combined in a fixed tree, so
// the sample repeats bit for bit (the reference sums row after row into one
// float).
__global__ void __launch_bounds__(64) k_hist_downscale(const float *__restrict__ img, int x_res, int y_res,
                                                       int sxr, int syr, float sxi, float syi,
                                                       float *__restrict__ sl) {
    const int k = static_cast<int>(blockIdx.x);
    const int i = k % sxr, j = k / sxr, lane = threadIdx.x;
    float x0 = 0.0f, x1 = sxi, y0 = 0.0f, y1 = syi;
    for (int q = 0; q < i; ++q) {
        x0 = x1;
        x1 = fminf((x_res - 1.0f), (x1 + sxi));
    }
    for (int q = 0; q < j; ++q) {
        y0 = y1;
        y1 = fminf((y_res - 1.0f), (y1 + syi));
    }
    // With one sample across (sxr or syr 1) the reference's first bound is
    // x_res or y_res, one past the image: the rows and columns stop inside.
    const int ye = min(static_cast<int>(y1), y_res - 1), xe = min(static_cast<int>(x1), x_res - 1);
    float acc = 0.0f;
    for (int y = static_cast<int>(y0) + lane; y <= ye; y += 64)
        for (int x = static_cast<int>(x0); x <= xe; ++x) acc += img[3 * (static_cast<int64_t>(y) * x_res + x) + 1];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if (lane == 0) sl[k] = acc / (((x1 - x0) + 1) * ((y1 - y0) + 1));
}

// camera::just_noticable_difference (rt/camera.cc:1778-1806).
__device__ float jnd(float La) {
    const float lLa = static_cast<float>(log10(static_cast<double>(La)));
    float lLt;
    if (lLa < -3.94f) {
        lLt = -2.86f;
    } else if (lLa < -1.44f) {
        lLt = powf(0.405f * lLa + 1.6f, 2.18f) - 2.86f;
    } else if (lLa < -0.0184f) {
        lLt = lLa - 0.395f;
    } else if (lLa < 1.9f) {
        lLt = powf(0.249f * lLa + 0.65f, 2.7f) - 0.72f;
    } else {
        lLt = lLa - 1.255f;
    }
    return powf(10.0f, lLt);
}

// One block: range of the samples, the histogram (integer LDS atomics, so
// exact), then the reference's ceiling iteration: each bin's ceiling and
// trim on its own thread, the trimmed total and the cumulative distribution
// on thread 0 (sums of integer counts, exact in float in any order).
__global__ void __launch_bounds__(kPostThreads) k_hist_build(const float *__restrict__ sl, int ns, int human,
                                                             const ToneScalars *__restrict__ ts,
                                                             HistState *__restrict__ out) {
    __shared__ float lds_f[kPostThreads / 64];
    __shared__ int lds_i[kPostThreads / 64];
    __shared__ int his[kHistBins + 1];
    __shared__ float cdf[kHistBins + 1];
    __shared__ int trim, total;
    float lo = kMaxDistF, hi = 0.0f;
    int cnt = 0;
    for (int k = threadIdx.x; k < kHistBins + 1; k += kPostThreads) {
        his[k] = 0;
        cdf[k] = 0.0f;
    }
    for (int k = threadIdx.x; k < ns; k += kPostThreads) {
        const float v = sl[k];
        if (v > 0.0f) {
            lo = fminf(lo, v);
            hi = fmaxf(hi, v);
            ++cnt;
        }
    }
    lo = block_reduce<float>(lo, min_f, lds_f);
    hi = block_reduce<float>(hi, max_f, lds_f);
    cnt = block_reduce<int>(cnt, add_i, lds_i);
    const float ld_max = 300.0f, ld_min = 1.0f;
    const float bd_max = logf(ld_max), bd_min = logf(ld_min);
    const float lw_min = fmaxf(lo, 1.0e-4f);
    const float bw_min = logf(lw_min);
    float bw_max = logf(hi);
    if (bw_max == bw_min) bw_max += 0.0001f * fabsf(bw_max);
    HistState h{};
    h.bw_min = bw_min;
    h.bw_max = bw_max;
    h.bd_min = bd_min;
    h.bd_max = bd_max;
    if (lo <= 0.0f) {  // black image: the reference leaves Y alone
        if (threadIdx.x == 0) *out = h;
        return;
    }
    if ((bd_max - bd_min) > (bw_max - bw_min)) {
        h.branch = 1;
        h.sf = (bd_max - bd_min) / (bw_max - bw_min);
        if (threadIdx.x == 0) *out = h;
        return;
    }
    for (int k = threadIdx.x; k < ns; k += kPostThreads) {
        const float v = sl[k];
        if (v > 0.0f) {
            int b = static_cast<int>((logf(v) - bw_min) / (bw_max - bw_min) * kHistBins) + 1;
            b = min(max(b, 0), kHistBins);
            atomicAdd(&his[b], 1);
        }
    }
    __syncthreads();
    // Thread 0: the bins' total and the cumulative distribution over div
    // samples (the first divides by every positive sample, as the reference
    // does, bin 0 included; later ones by the bins' total).
    auto distribution = [&](int div) {
        int t = 0;
        for (int i = 1; i <= kHistBins; ++i) t += his[i];
        if (div < 0) div = t;
        float sum = 0.0f;
        for (int i = 1; i <= kHistBins; ++i) {
            sum += his[i];
            cdf[i] = sum / static_cast<float>(div);
        }
        total = t;
    };
    if (threadIdx.x == 0) distribution(cnt);
    __syncthreads();
    const int pix0 = cnt;  // the first tolerance uses the count of positive samples
    const float tolerance = 0.025f * static_cast<float>(pix0);
    const float bin_step_size = (bw_max - bw_min) / kHistBins;
    const float lw_step = 1.0f / static_cast<float>(kHistBins) * (bw_max - bw_min);
    const float lw_avg = ts->level;
    int pix_gt0 = pix0;
    for (;;) {
        if (pix_gt0 < tolerance) {
            const float num = 1.219f + powf(ld_max * 0.5f, 0.4f);
            const float den = 1.219f + powf(lw_avg, 0.4f);
            h.branch = 2;
            h.sf = (1.0f / ld_max) * powf((num / den), 2.5f);
            if (threadIdx.x == 0) *out = h;
            return;
        }
        if (threadIdx.x == 0) trim = 0;
        __syncthreads();
        const int i = threadIdx.x + 1;
        if (i <= kHistBins) {
            float ceiling = (pix_gt0 * bin_step_size) / (bd_max - bd_min);
            if (human) {
                const float bde = bd_min + (bd_max - bd_min) * cdf[i];
                const float ld = expf(bde);
                const float bw = i * lw_step + bw_min;
                const float lw = expf(bw);
                ceiling *= (jnd(ld) * lw) / (jnd(lw) * ld);
            }
            if (his[i] > ceiling) {
                atomicAdd(&trim, his[i] - static_cast<int>(ceiling));
                his[i] = static_cast<int>(ceiling);
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) distribution(-1);
        __syncthreads();
        pix_gt0 = total;
        if (!(trim > tolerance)) break;
        __syncthreads();  // trim is reset only after every thread has read it
    }
    if (threadIdx.x == 0) {
        h.branch = 3;
        for (int i = 0; i <= kHistBins; ++i) h.cdf[i] = cdf[i];
        *out = h;
    }
}

int post_blocks(int64_t n) {
    return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(kPostMaxBlocks, (n + kPostThreads - 1) / kPostThreads)));
}

// The bilateral tone map of the Y channel of an (x, Y, y) image on the
// device (the arguments global_bilateral_filter passes: r_s = r_sa = 0.4,
// s_s = s_sa = 0.02 min(x_res, y_res)). The grid's depth depends on the
// image's range of log intensity, so that range is read back once.
int32_t bilateral(rt_ctx *ctx, float *img, int x_res, int y_res, int blocks) {
    hipStream_t s = ctx->stream;
    BuildArena &ar = ctx->arena;
    const int64_t n = static_cast<int64_t>(x_res) * y_res;
    void *v = nullptr;
    auto take = [&](auto **p, int64_t count) -> hipError_t {
        const hipError_t e = ar.take(&v, static_cast<size_t>(std::max<int64_t>(1, count)) * sizeof(**p));
        *p = static_cast<std::remove_reference_t<decltype(*p)>>(v);
        return e;
    };
    float *logi = nullptr, *base = nullptr;
    float2 *part = nullptr, *range = nullptr;
    RT_HIP_CHECK(ctx, take(&logi, n));
    RT_HIP_CHECK(ctx, take(&base, n));
    RT_HIP_CHECK(ctx, take(&part, blocks));
    RT_HIP_CHECK(ctx, take(&range, 2));
    hipLaunchKernelGGL(k_bil_log, dim3(blocks), dim3(kPostThreads), 0, s, img, n, logi, part);
    hipLaunchKernelGGL(k_minmax_fold, dim3(1), dim3(kPostThreads), 0, s, part, blocks, range);
    float2 mm;
    RT_HIP_CHECK(ctx, hipMemcpyAsync(&mm, range, sizeof(mm), hipMemcpyDeviceToHost, s));
    RT_HIP_CHECK(ctx, hipStreamSynchronize(s));
    BilGrid g{};
    const float s_s = 0.02f * static_cast<float>(std::min(x_res, y_res));
    const float r_s = 0.4f;
    g.r_sa = 0.4f;
    g.s_sa = s_s;
    const float sigma_r = r_s / g.r_sa, sigma_s = s_s / g.s_sa;
    g.pxy = static_cast<int>(2.0f * sigma_s) + 1;
    g.pz = static_cast<int>(2.0f * sigma_r) + 1;
    g.in_min = mm.x;
    const float delta = mm.y - mm.x;  // a frame with no bright pixel: -inf, an empty grid depth
    if (!(delta >= 0.0f) || delta / g.r_sa > 4096.0f)
        return fail(ctx, RT_ERR_INVALID, "rt_image_tone_map: the bilateral filter needs a frame with a finite range "
                                         "of bright pixels");
    // Grid extents from the float quotients, truncated as the reference's
    // int conversions do, but sized in double first: a frame of extreme
    // aspect ratio gives a quotient beyond any int (and a grid of gigabytes).
    const double wd = std::trunc(static_cast<double>(static_cast<float>(x_res - 1) / g.s_sa)) + 1 + (g.pxy << 1);
    const double hd = std::trunc(static_cast<double>(static_cast<float>(y_res - 1) / g.s_sa)) + 1 + (g.pxy << 1);
    const double dd = std::trunc(static_cast<double>(delta / g.r_sa)) + 1 + (g.pz << 1);
    if (wd * hd * dd > static_cast<double>(int64_t(1) << 27))
        return fail(ctx, RT_ERR_INVALID, "rt_image_tone_map: the bilateral grid would exceed 2^27 cells");
    g.W = static_cast<int>(wd);
    g.H = static_cast<int>(hd);
    g.D = static_cast<int>(dd);
    g.N = static_cast<int64_t>(g.W) * g.H * g.D;
    const int cells_x = static_cast<int>(static_cast<float>(x_res - 1) / g.s_sa + 0.5f) + 1;
    const int cells_y = static_cast<int>(static_cast<float>(y_res - 1) / g.s_sa + 0.5f) + 1;
    // Axis factors exp(-X^2 / (2 sigma^2)), X the offset wrapped as the
    // reference builds its kernel (rt/camera.cc:1668-1690).
    std::vector<float> f(g.W + g.H + g.D);
    auto axis = [&](float *out, int len, float sigma) {
        for (int i = 0; i < len; ++i) {
            const double X = i - ((i > (len >> 1)) ? len : 0);
            out[i] = static_cast<float>(std::exp(-0.5 * X * X / (double(sigma) * sigma)));
        }
    };
    axis(f.data(), g.W, sigma_s);
    axis(f.data() + g.W, g.H, sigma_s);
    axis(f.data() + g.W + g.H, g.D, sigma_r);
    float *d_f = nullptr, *grid = nullptr, *tmp = nullptr;
    RT_HIP_CHECK(ctx, take(&d_f, static_cast<int64_t>(f.size())));
    RT_HIP_CHECK(ctx, take(&grid, 2 * g.N));
    RT_HIP_CHECK(ctx, take(&tmp, 2 * g.N));
    RT_HIP_CHECK(ctx, hipMemcpyAsync(d_f, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice, s));
    const unsigned cb = static_cast<unsigned>((2 * g.N + kPostThreads - 1) / kPostThreads);
    hipLaunchKernelGGL(k_fill_zero, dim3(std::min(cb, 2048u)), dim3(kPostThreads), 0, s, grid, 2 * g.N);
    hipLaunchKernelGGL(k_bil_splat, dim3(static_cast<unsigned>(cells_x * cells_y)), dim3(64), 0, s, g, logi, x_res,
                       y_res, cells_x, grid);
    hipLaunchKernelGGL(k_bil_conv, dim3(cb), dim3(kPostThreads), 0, s, g, 0, d_f, grid, tmp);
    hipLaunchKernelGGL(k_bil_conv, dim3(cb), dim3(kPostThreads), 0, s, g, 1, d_f + g.W, tmp, grid);
    hipLaunchKernelGGL(k_bil_conv, dim3(cb), dim3(kPostThreads), 0, s, g, 2, d_f + g.W + g.H, grid, tmp);
    hipLaunchKernelGGL(k_bil_slice, dim3(blocks), dim3(kPostThreads), 0, s, g, tmp, logi, x_res, n, base, part);
    hipLaunchKernelGGL(k_minmax_fold, dim3(1), dim3(kPostThreads), 0, s, part, blocks, range);
    hipLaunchKernelGGL(k_bil_apply, dim3(blocks), dim3(kPostThreads), 0, s, img, n, logi, base, range,
                       std::log10(300.0f));
    RT_HIP_CHECK(ctx, hipGetLastError());
    return RT_OK;
}

}  // namespace

int32_t image_tone_map(rt_ctx *ctx, float *rgb, int32_t x_res, int32_t y_res, rt_tone_params *tp, uint32_t flags,
                       float *ms) {
    hipStream_t s = ctx->stream;
    const int m = tp->mode;
    if ((m == 2 || m == 3) && !(tp->x_angle * (360.0f / (2.0f * 3.1415926535897932384626433f)) >= 1.0f &&
                                tp->y_angle * (360.0f / (2.0f * 3.1415926535897932384626433f)) >= 1.0f &&
                                tp->x_angle < 6.3f && tp->y_angle < 6.3f))
        return fail(ctx, RT_ERR_INVALID,
                    "rt_image_tone_map: the histogram modes need the camera's field of view (x_angle, y_angle) of "
                    "at least one degree");
    if (m < 0 || m > 9) return fail(ctx, RT_ERR_INVALID, "rt_image_tone_map: unknown mode " + std::to_string(m));
    if (tp->glare && !(tp->x_angle > 0.0f && x_res <= 4096 && y_res <= 4096))
        return fail(ctx, RT_ERR_INVALID,
                    "rt_image_tone_map: glare needs the camera's field of view (x_angle > 0) and a frame of at most "
                    "4096 x 4096 pixels");
    const int64_t n = static_cast<int64_t>(x_res) * y_res;
    BuildArena &ar = ctx->arena;
    RT_HIP_CHECK(ctx, ar.begin());
    float *img = rgb;
    const bool dev = (flags & RT_DEVICE_PTRS) != 0;
    void *v = nullptr;
    if (!dev) {
        RT_HIP_CHECK(ctx, ar.take(&v, static_cast<size_t>(3 * n) * sizeof(float)));
        img = static_cast<float *>(v);
        RT_HIP_CHECK(ctx, hipMemcpyAsync(img, rgb, 3 * n * sizeof(float), hipMemcpyHostToDevice, s));
    }
    const int blocks = post_blocks(n);
    TonePartial *part = nullptr;
    ToneScalars *scal = nullptr;
    float2 *minmax = nullptr;
    RT_HIP_CHECK(ctx, ar.take(&v, blocks * sizeof(TonePartial)));
    part = static_cast<TonePartial *>(v);
    RT_HIP_CHECK(ctx, ar.take(&v, sizeof(ToneScalars)));
    scal = static_cast<ToneScalars *>(v);
    RT_HIP_CHECK(ctx, ar.take(&v, blocks * sizeof(float2)));
    minmax = static_cast<float2 *>(v);
    const ToneArgs a{img, n, m, tp->key, tp->xw, tp->yw, tp->desaturate, tp->stretch, tp->adapt, tp->time_step,
                     tp->adaptation_level};
    if (ms) RT_HIP_CHECK(ctx, hipEventRecord(ctx->ev[0], s));
    hipLaunchKernelGGL(k_tone_yxy, dim3(blocks), dim3(kPostThreads), 0, s, a, part);
    hipLaunchKernelGGL(k_tone_scalars, dim3(1), dim3(kPostThreads), 0, s, a, part, blocks, scal);
    HistState *hist = nullptr;
    if (m == kBilateral) {
        if (int32_t rc = bilateral(ctx, img, x_res, y_res, blocks)) return rc;
    } else if (m == kHistogram || m == kHumanHistogram) {
        // One sample per degree of view (rt/camera.cc:1846-1853).
        const float PI = 3.1415926535897932384626433f;
        const int sxr = static_cast<int>(tp->x_angle * (360.0f / (2.0f * PI)));
        const int syr = static_cast<int>(tp->y_angle * (360.0f / (2.0f * PI)));
        const float sxi = static_cast<float>(x_res) / static_cast<float>(sxr);
        const float syi = static_cast<float>(y_res) / static_cast<float>(syr);
        float *sl = nullptr;
        RT_HIP_CHECK(ctx, ar.take(&v, static_cast<size_t>(sxr) * syr * sizeof(float)));
        sl = static_cast<float *>(v);
        RT_HIP_CHECK(ctx, ar.take(&v, sizeof(HistState)));
        hist = static_cast<HistState *>(v);
        hipLaunchKernelGGL(k_hist_downscale, dim3(static_cast<unsigned>(sxr * syr)), dim3(64), 0, s, img, x_res,
                           y_res, sxr, syr, sxi, syi, sl);
        hipLaunchKernelGGL(k_hist_build, dim3(1), dim3(kPostThreads), 0, s, sl, sxr * syr, m == kHumanHistogram ? 1 : 0,
                           scal, hist);
    }
    hipLaunchKernelGGL(k_tone_apply, dim3(blocks), dim3(kPostThreads), 0, s, a, scal, hist, minmax);
    if (tp->stretch) {
        hipLaunchKernelGGL(k_stretch_scalars, dim3(1), dim3(kPostThreads), 0, s, minmax, blocks, scal);
        hipLaunchKernelGGL(k_tone_stretch, dim3(blocks), dim3(kPostThreads), 0, s, img, n, scal);
    }
    RT_HIP_CHECK(ctx, hipGetLastError());
    ToneScalars t{};
    RT_HIP_CHECK(ctx, hipMemcpyAsync(&t, scal, sizeof(t), hipMemcpyDeviceToHost, s));
    if (tp->glare) {
        // rt/camera.cc:1455-1458: the glare filter at the adapted luminance,
        // whose light level picks the filter.
        RT_HIP_CHECK(ctx, hipStreamSynchronize(s));
        if (int32_t rc = apply_glare(ctx, img, x_res, y_res, t.yi, tp->x_angle)) return rc;
    }
    if (ms) RT_HIP_CHECK(ctx, hipEventRecord(ctx->ev[1], s));
    if (!dev) RT_HIP_CHECK(ctx, hipMemcpyAsync(rgb, img, 3 * n * sizeof(float), hipMemcpyDeviceToHost, s));
    RT_HIP_CHECK(ctx, hipStreamSynchronize(s));
    tp->adaptation_level = t.level;
    tp->log_avg_luminance = t.yi;
    if (ms) RT_HIP_CHECK(ctx, hipEventElapsedTime(ms, ctx->ev[0], ctx->ev[1]));
    return RT_OK;
}

int32_t image_gamma(rt_ctx *ctx, float *rgb, int64_t npix, float gamma, uint32_t flags, float *ms) {
    hipStream_t s = ctx->stream;
    float *img = rgb;
    const bool dev = (flags & RT_DEVICE_PTRS) != 0;
    BuildArena &ar = ctx->arena;
    RT_HIP_CHECK(ctx, ar.begin());
    if (!dev) {
        void *v = nullptr;
        RT_HIP_CHECK(ctx, ar.take(&v, static_cast<size_t>(3 * npix) * sizeof(float)));
        img = static_cast<float *>(v);
        RT_HIP_CHECK(ctx, hipMemcpyAsync(img, rgb, 3 * npix * sizeof(float), hipMemcpyHostToDevice, s));
    }
    if (ms) RT_HIP_CHECK(ctx, hipEventRecord(ctx->ev[0], s));
    hipLaunchKernelGGL(k_gamma, dim3(post_blocks(3 * npix)), dim3(kPostThreads), 0, s, img, npix, gamma);
    RT_HIP_CHECK(ctx, hipGetLastError());
    if (ms) RT_HIP_CHECK(ctx, hipEventRecord(ctx->ev[1], s));
    if (!dev) RT_HIP_CHECK(ctx, hipMemcpyAsync(rgb, img, 3 * npix * sizeof(float), hipMemcpyDeviceToHost, s));
    RT_HIP_CHECK(ctx, hipStreamSynchronize(s));
    if (ms) RT_HIP_CHECK(ctx, hipEventElapsedTime(ms, ctx->ev[0], ctx->ev[1]));
    return RT_OK;
}

int32_t image_clip(rt_ctx *ctx, const float *rgb, int32_t x_res, int32_t y_res, int32_t xs, int32_t ys, int32_t bgr,
                   uint8_t *out, uint32_t flags, float *ms) {
    hipStream_t s = ctx->stream;
    const int ox = x_res / xs, oy = y_res / ys;
    const int64_t nin = static_cast<int64_t>(x_res) * y_res, nout = static_cast<int64_t>(ox) * oy;
    const bool dev = (flags & RT_DEVICE_PTRS) != 0;
    const float *img = rgb;
    uint8_t *o = out;
    BuildArena &ar = ctx->arena;
    RT_HIP_CHECK(ctx, ar.begin());
    if (!dev) {
        void *v = nullptr;
        RT_HIP_CHECK(ctx, ar.take(&v, static_cast<size_t>(3 * nin) * sizeof(float)));
        RT_HIP_CHECK(ctx, hipMemcpyAsync(v, rgb, 3 * nin * sizeof(float), hipMemcpyHostToDevice, s));
        img = static_cast<const float *>(v);
        RT_HIP_CHECK(ctx, ar.take(&v, static_cast<size_t>(3 * nout)));
        o = static_cast<uint8_t *>(v);
    }
    if (ms) RT_HIP_CHECK(ctx, hipEventRecord(ctx->ev[0], s));
    hipLaunchKernelGGL(k_clip, dim3(post_blocks(nout)), dim3(kPostThreads), 0, s, img, x_res, xs, ys, ox, nout, bgr, o);
    RT_HIP_CHECK(ctx, hipGetLastError());
    if (ms) RT_HIP_CHECK(ctx, hipEventRecord(ctx->ev[1], s));
    if (!dev) RT_HIP_CHECK(ctx, hipMemcpyAsync(out, o, 3 * nout, hipMemcpyDeviceToHost, s));
    RT_HIP_CHECK(ctx, hipStreamSynchronize(s));
    if (ms) RT_HIP_CHECK(ctx, hipEventElapsedTime(ms, ctx->ev[0], ctx->ev[1]));
    return RT_OK;
}

namespace {

// CRC-32 of PNG chunks (ISO 3309 polynomial 0xEDB88320, reflected).
uint32_t crc32_update(uint32_t crc, const uint8_t *p, size_t n) {
    static uint32_t table[256];
    static bool ready = false;
    if (!ready) {
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t c = i;
            for (int k = 0; k < 8; ++k) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            table[i] = c;
        }
        ready = true;
    }
    for (size_t i = 0; i < n; ++i) crc = table[(crc ^ p[i]) & 0xFFu] ^ (crc >> 8);
    return crc;
}

void put_be32(std::vector<uint8_t> &v, uint32_t x) {
    for (int k = 3; k >= 0; --k) v.push_back(static_cast<uint8_t>(x >> (8 * k)));
}

void png_chunk(std::vector<uint8_t> &file, const char *type, const std::vector<uint8_t> &data) {
    put_be32(file, static_cast<uint32_t>(data.size()));
    const size_t start = file.size();
    file.insert(file.end(), type, type + 4);
    file.insert(file.end(), data.begin(), data.end());
    put_be32(file, crc32_update(0xFFFFFFFFu, file.data() + start, file.size() - start) ^ 0xFFFFFFFFu);
}

}  // namespace

// write_png_file (rt/camera.cc:2158-2210): 8-bit RGB, the image's last row
// first (row y of camera::image is y rows up from the bottom). The zlib
// stream holds stored (uncompressed) deflate blocks: any PNG reader decodes
// the same pixels as from libpng's compressed stream.
int32_t write_png(const char *path, const uint8_t *rgb, int32_t x_res, int32_t y_res) {
    const size_t row = static_cast<size_t>(x_res) * 3 + 1;
    std::vector<uint8_t> raw;
    raw.reserve(row * y_res);
    for (int y = y_res - 1; y >= 0; --y) {
        raw.push_back(0);  // filter: none
        const uint8_t *r = rgb + static_cast<size_t>(y) * x_res * 3;
        raw.insert(raw.end(), r, r + static_cast<size_t>(x_res) * 3);
    }
    std::vector<uint8_t> z = {0x78, 0x01};
    uint32_t s1 = 1, s2 = 0;  // Adler-32
    for (const uint8_t b : raw) {
        s1 = (s1 + b) % 65521u;
        s2 = (s2 + s1) % 65521u;
    }
    for (size_t off = 0; off < raw.size() || off == 0;) {
        const size_t len = std::min<size_t>(65535, raw.size() - off);
        const bool last = off + len == raw.size();
        z.push_back(last ? 1 : 0);
        z.push_back(static_cast<uint8_t>(len & 0xFF));
        z.push_back(static_cast<uint8_t>(len >> 8));
        z.push_back(static_cast<uint8_t>(~len & 0xFF));
        z.push_back(static_cast<uint8_t>((~len >> 8) & 0xFF));
        z.insert(z.end(), raw.begin() + off, raw.begin() + off + len);
        off += len;
        if (last) break;
    }
    put_be32(z, (s2 << 16) | s1);
    std::vector<uint8_t> file = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    std::vector<uint8_t> ihdr;
    put_be32(ihdr, static_cast<uint32_t>(x_res));
    put_be32(ihdr, static_cast<uint32_t>(y_res));
    ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});  // 8-bit, RGB, deflate, no filter method, no interlace
    png_chunk(file, "IHDR", ihdr);
    png_chunk(file, "IDAT", z);
    png_chunk(file, "IEND", {});
    std::ofstream f(path, std::ios::binary);
    if (!f) return RT_ERR_INVALID;
    f.write(reinterpret_cast<const char *>(file.data()), static_cast<std::streamsize>(file.size()));
    return f ? RT_OK : RT_ERR_INVALID;
}

// camera::write_tga_file (rt/camera.cc:2102-2155): 18-byte header of an
// uncompressed 24-bit image with a bottom-left origin, the bgr bytes in
// image order, eight zero bytes.
int32_t write_tga(const char *path, const uint8_t *bgr, int32_t x_res, int32_t y_res) {
    uint8_t h[18] = {0, 0, 2, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 24, 0};
    h[12] = static_cast<uint8_t>(x_res & 0xFF);
    h[13] = static_cast<uint8_t>((x_res >> 8) & 0xFF);
    h[14] = static_cast<uint8_t>(y_res & 0xFF);
    h[15] = static_cast<uint8_t>((y_res >> 8) & 0xFF);
    std::ofstream f(path, std::ios::binary);
    if (!f) return RT_ERR_INVALID;
    f.write(reinterpret_cast<const char *>(h), sizeof(h));
    f.write(reinterpret_cast<const char *>(bgr), static_cast<std::streamsize>(x_res) * y_res * 3);
    const char footer[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    f.write(footer, sizeof(footer));
    return f ? RT_OK : RT_ERR_INVALID;
}

}  // namespace rt
