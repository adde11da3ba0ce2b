// The glare filter of camera::tone_map (rt/camera.cc:1455-1458) on the GPU.
//
// Filter (rt/camera.cc:145-462), built once per resolution, field of view
// and light level and kept on the context:
//   * the reference's random generator (common/common.cc:30-105: MT19937
//     seeded 5489, whose refill stops at index 622) gives the intensity and
//     angle of 2 x 200 flare lines; their cos, sin and tan are taken on the
//     host with the float library calls the reference makes;
//   * k_glare_lines: one thread per pixel walks the 200 lines of an image in
//     order (so each pixel sums its line values in the reference's order) and
//     decides coverage with the reference's row and column bounds;
//   * the bloom's level is the corona's float sum in the reference's order
//     (line, row from the centre outwards, column): k_glare_line_rows counts
//     and then writes each (line, row)'s values at offsets from a scan, and
//     one wave adds them up serially (k_glare_seq_sum). A float running sum
//     of ~10^5 values rounds with a bias of its own (about 1% at 1080p), which
//     this keeps;
//   * k_glare_psf: each of the three images (corona, halo, flat bloom at the
//     corona's mean) times the trapezoidal integral of its spreading function
//     over the pixel (the reference's sample points, samps = 3), summed;
//   * k_glare_scale: scaled so the filter totals 1e-4 per pixel, by the
//     filter's float sum in pixel order (k_glare_seq_sum again).
// Convolution (perform_glare_filter, rt/camera.cc:1504-1630): the pixels with
// a channel above 255, circularly convolved on the x_res x y_res grid with
// the filter shifted to put its centre on pixel 0, divided by 255 and added.
// The reference uses fftw3f; here the circular convolution is exact through
// a linear one: both operands zero-padded to Mx x My >= (2 x_res - 1) x
// (2 y_res - 1) (powers of two), transformed by radix-2 FFTs in LDS (one
// workgroup per row, bit-reversed load, in-place butterflies) with tiled
// transposes between the two axes, multiplied, transformed back, and the
// four quadrants folded onto the image. The filter's spectrum is cached with
// the filter.
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "rt_internal.hpp"

namespace rt {

struct GlareFilter {
    int x_res = 0, y_res = 0, level = -1;
    float x_angle = 0.0f;
    float *gf = nullptr;                 // x_res * y_res rgb
    float2 *spec[3] = {nullptr, nullptr, nullptr};  // per channel, transposed layout (Mx rows of My)
};

struct GlareCache {
    int mx = 0, my = 0;                  // padded sizes (powers of two)
    float2 *tw_x = nullptr, *tw_y = nullptr;  // exp(-2 pi i m / M), m < M / 2
    float2 *a = nullptr, *b = nullptr;   // work buffers, mx * my each
    uint8_t *mask = nullptr;             // bright pixels of the image, mx * my bytes
    GlareFilter f[3];                    // one per light level
};

namespace {

constexpr float kPi = 3.1415926535897932384626433f;
constexpr int kLines = 200;
constexpr int kThreads = 256;
constexpr int kMaxFft = 8192;            // 64 KiB of LDS per row transform

// common/common.cc:30-105 (see the file comment).
struct Mt {
    uint32_t mt[624];
    int index = 625;
    uint32_t raw() {
        if (index >= 624) {
            if (index == 625) {
                mt[0] = 5489u;
                for (int i = 1; i < 624; ++i)
                    mt[i] = 1812433253u * (mt[i - 1] ^ (mt[i - 1] >> 30)) + static_cast<uint32_t>(i);
            }
            for (int i = 0; i < 227; ++i) {
                const uint32_t y = (0x80000000u & mt[i]) + (0x7FFFFFFFu & mt[i + 1]);
                mt[i] = mt[i + 397] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
            }
            for (int i = 227; i < 623; ++i) {
                const uint32_t y = (0x80000000u & mt[i]) + (0x7FFFFFFFu & mt[i + 1]);
                mt[i] = mt[i - 227] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
            }
            index = 0;
        }
        uint32_t y = mt[index++];
        y ^= (y >> 11);
        y ^= ((y << 7) & 0x9d2c5680u);
        y ^= ((y << 15) & 0xefc60000u);
        y ^= (y >> 18);
        return y;
    }
    float next() { return static_cast<float>(raw()) * (1.0f / 4294967296.0f); }
};

// One flare line (lineRad = 1): intensity, direction, tan and sin of its
// angle, the sin-based half width dif, the row step and first row.
struct Line {
    float intens, c, s, t, dif;
    int dj, j0;
};

// The columns [k0, k1) that line ln visits in row j (rt/camera.cc:162-181).
__device__ __forceinline__ void row_bounds(const Line &ln, int j, float xrf, float yrf, int xr, int &k0, int &k1) {
    const float jminf = (j + 0.0f - yrf / 2.0f);
    const float jmaxf = (j + 1.0f - yrf / 2.0f);
    float iminf = jminf / ln.t;
    float imaxf = jmaxf / ln.t;
    if (iminf > imaxf) {
        const float tmp = iminf;
        iminf = imaxf;
        imaxf = tmp;
    }
    const int imin = static_cast<int>(iminf - ln.dif + 0.5f + xrf / 2.0f);
    const int imax = static_cast<int>(imaxf + ln.dif - 0.5f + xrf / 2.0f);
    k0 = max(imin, 0);
    k1 = min(imax + 1, xr);
}

// The value line ln adds to pixel (k, j) of a visited column, or -1 where the
// pixel is outside the line's radius (rt/camera.cc:183-210).
__device__ __forceinline__ float line_value(const Line &ln, int k, int j, float xrf, float yrf) {
    const float p0 = k - xrf / 2.0f + 0.5f;
    const float p1 = j - yrf / 2.0f + 0.5f;
    const float dp = (p0 * ln.c) + (p1 * ln.s);
    float dist;
    if (dp > 0.0f) {
        const float v = (p0 - dp * ln.c);  // the reference repeats the x component
        dist = sqrtf((v * v) + (v * v));
    } else {
        dist = sqrtf((p0 * p0) + (p1 * p1));
    }
    if (dist >= 1.0f) return -1.0f;
    const float filt = 1.0f / 1.0f - 1.0f / (1.0f * 1.0f) * dist;
    const float len = sqrtf((p0 * p0) + (p1 * p1));
    const float flo = fmaxf((len * 2.0f * kPi) / kLines, 1.0f / kLines);
    return filt * flo * ln.intens;
}

__global__ void __launch_bounds__(kThreads) k_glare_lines(const Line *__restrict__ lines, int xr, int yr,
                                                          float *__restrict__ out) {
    // blockIdx.y: image (0 corona, 1 halo); out is two planes of x_res * y_res.
    const int64_t n = static_cast<int64_t>(xr) * yr;
    const int64_t pix = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    const Line *L = lines + blockIdx.y * kLines;
    const float xrf = static_cast<float>(static_cast<unsigned>(xr)), yrf = static_cast<float>(static_cast<unsigned>(yr));
    if (pix >= n) return;
    const int k = static_cast<int>(pix % xr), j = static_cast<int>(pix / xr);
    float acc = 0.0f;
    for (int l = 0; l < kLines; ++l) {
        const Line ln = L[l];
        // Rows from j0 stepping by dj while inside the image.
        if (ln.j0 < 0 || ln.j0 >= yr) continue;
        if (ln.dj > 0 ? (j < ln.j0) : (j > ln.j0)) continue;
        int k0, k1;
        row_bounds(ln, j, xrf, yrf, xr, k0, k1);
        if (k < k0 || k >= k1) continue;
        const float val = line_value(ln, k, j, xrf, yrf);
        if (val >= 0.0f) acc += val;
    }
    out[blockIdx.y * n + pix] = acc;
}

// One thread per (line, row step r) of the corona's lines: the number of
// values the row gives (WRITE = false) or the values themselves, in column
// order, at vals + off[t] (WRITE = true). Row r of a line is j0 + r dj.
template <bool WRITE>
__global__ void __launch_bounds__(kThreads) k_glare_line_rows(const Line *__restrict__ lines, int xr, int yr,
                                                              int *__restrict__ cnt, const int64_t *__restrict__ off,
                                                              float *__restrict__ vals) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (t >= static_cast<int64_t>(kLines) * yr) return;
    const Line ln = lines[t / yr];
    const int j = ln.j0 + static_cast<int>(t % yr) * ln.dj;
    int c = 0;
    if (ln.j0 >= 0 && ln.j0 < yr && j >= 0 && j < yr) {
        const float xrf = static_cast<float>(static_cast<unsigned>(xr)), yrf = static_cast<float>(static_cast<unsigned>(yr));
        int k0, k1;
        row_bounds(ln, j, xrf, yrf, xr, k0, k1);
        for (int k = k0; k < k1; ++k) {
            const float val = line_value(ln, k, j, xrf, yrf);
            if (val < 0.0f) continue;
            if (WRITE) vals[off[t] + c] = val;
            ++c;
        }
    }
    if (!WRITE) cnt[t] = c;
}

// Exclusive scan of m counts into off[0..m], off[m] the total: one workgroup,
// each thread a contiguous segment.
__global__ void __launch_bounds__(1024) k_glare_scan(const int *__restrict__ cnt, int64_t m,
                                                     int64_t *__restrict__ off) {
    __shared__ int64_t tot[1024];
    const int64_t seg = (m + 1023) / 1024;
    const int64_t b = threadIdx.x * seg, e = min(m, b + seg);
    int64_t s = 0;
    for (int64_t i = b; i < e; ++i) s += cnt[i];
    tot[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t run = 0;
        for (int i = 0; i < 1024; ++i) {
            const int64_t v = tot[i];
            tot[i] = run;
            run += v;
        }
        off[m] = run;
    }
    __syncthreads();
    int64_t run = tot[threadIdx.x];
    for (int64_t i = b; i < e; ++i) {
        off[i] = run;
        run += cnt[i];
    }
}

// The float running sum of v[i * NCH + c], i = 0 .. n-1, for each channel
// c < NCH, in order, as one loop adds them. The adds are one dependent chain
// per channel, so wave 0 does them all; the block's four waves stream the
// next chunk of kSeqChunk entries into the other half of an LDS double
// buffer meanwhile, which hides the HBM latency behind the adds.
constexpr int kSeqChunk = 2048;

template <int NCH>
__global__ void __launch_bounds__(256) k_glare_seq_sum(const float *__restrict__ v, int64_t n,
                                                       float *__restrict__ out) {
    __shared__ float buf[2][kSeqChunk * NCH];
    constexpr int kPer = kSeqChunk * NCH / 256;  // floats each thread moves per chunk
    const int64_t chunks = (n + kSeqChunk - 1) / kSeqChunk;
    const int64_t total = n * NCH;
    float s[NCH];
    for (int c = 0; c < NCH; ++c) s[c] = 0.0f;
    float r[kPer];
    // Chunk 0 into buffer 0.
    for (int k = 0; k < kPer; ++k) {
        const int64_t e = static_cast<int64_t>(k) * 256 + threadIdx.x;
        buf[0][k * 256 + threadIdx.x] = (e < total) ? v[e] : 0.0f;
    }
    __syncthreads();
    for (int64_t ch = 0; ch < chunks; ++ch) {
        const int cur = static_cast<int>(ch & 1);
        const bool more = ch + 1 < chunks;
        if (more) {
            const int64_t base = (ch + 1) * kSeqChunk * NCH;
#pragma unroll
            for (int k = 0; k < kPer; ++k) {
                const int64_t e = base + static_cast<int64_t>(k) * 256 + threadIdx.x;
                r[k] = (e < total) ? v[e] : 0.0f;
            }
        }
        if (threadIdx.x < 64) {
            const int cnt = static_cast<int>(min(static_cast<int64_t>(kSeqChunk), n - ch * kSeqChunk));
            const float *b = buf[cur];
            int j = 0;
            for (; j + 8 <= cnt; j += 8) {
                float x[8 * NCH];
#pragma unroll
                for (int k = 0; k < 8 * NCH; ++k) x[k] = b[j * NCH + k];
#pragma unroll
                for (int k = 0; k < 8; ++k)
#pragma unroll
                    for (int c = 0; c < NCH; ++c) s[c] += x[k * NCH + c];
            }
            for (; j < cnt; ++j)
                for (int c = 0; c < NCH; ++c) s[c] += b[j * NCH + c];
        }
        if (more)
#pragma unroll
            for (int k = 0; k < kPer; ++k) buf[cur ^ 1][k * 256 + threadIdx.x] = r[k];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int c = 0; c < NCH; ++c) out[c] = s[c];
}

// Spencer's glare components (rt/camera.cc:12-35).
__device__ __forceinline__ float f0(float t) {
    const float tmp = t * (1.0f / 0.02f);
    return 2.61e6f * expf(-(tmp * tmp));
}
__device__ __forceinline__ float f1(float t) {
    const float tmp = t + 0.02f;
    return 20.91f / (tmp * tmp * tmp);
}
__device__ __forceinline__ float f2(float t) {
    const float tmp = t + 0.02f;
    return 72.37f / (tmp * tmp);
}
__device__ __forceinline__ float f3(float t, float l) {
    const float tmp = t - 3.0f * l * (1.0f / 568.0f);
    return 436.9f * (568.0f / l) * expf(-19.75f * (tmp * tmp));
}

struct Rgb3 {
    float r, g, b;
};
__device__ __forceinline__ Rgb3 add3(Rgb3 a, Rgb3 b) { return Rgb3{a.r + b.r, a.g + b.g, a.b + b.b}; }

// The spreading functions (rt/camera.cc:37-141); kind = 3 * level + component.
__device__ Rgb3 psf(int kind, float x, float y, float p) {
    const float dist = sqrtf((x * x) + (y * y));
    const float deg = dist * p;
    float a;
    switch (kind) {
        case 0: a = (0.478f * f1(deg)) + (0.207f * f2(deg)); return Rgb3{a, a, a};
        case 1: return Rgb3{0.033f * f3(deg, 611.0f), 0.033f * f3(deg, 549.0f), 0.033f * f3(deg, 467.0f)};
        case 2: a = 0.282f * f0(deg); return Rgb3{a, a, a};
        case 3: a = (0.478f * f1(deg)) + (0.138f * f2(deg)); return Rgb3{a, a, a};
        case 4: return Rgb3{0.016f * f3(deg, 611.0f), 0.016f * f3(deg, 549.0f), 0.016f * f3(deg, 467.0f)};
        case 5: a = 0.368f * f0(deg); return Rgb3{a, a, a};
        case 6: a = (0.478f * f1(deg)) + (0.138f * f2(deg)); return Rgb3{a, a, a};
        case 7: return Rgb3{0.0f, 0.0f, 0.0f};
        default: a = 0.384f * f0(deg); return Rgb3{a, a, a};
    }
}

// rt/camera.cc:224-322 for one pixel: the trapezoidal integral of the
// spreading function with the reference's sample points.
__device__ Rgb3 psf_integral(int kind, int i, int j, float xrf, float yrf, float pixdeg) {
    const float w = static_cast<float>(i) - xrf / 2.0f, x = w + 1.0f;
    const float y = static_cast<float>(j) - yrf / 2.0f, z = y + 1.0f;
    Rgb3 sum1 = psf(kind, w, y, pixdeg);
    sum1 = add3(sum1, psf(kind, x, y, pixdeg));
    sum1 = add3(sum1, psf(kind, w, z, pixdeg));
    sum1 = add3(sum1, psf(kind, x, z, pixdeg));
    const int n = 2;  // samps - 1
    const float h = (x - w) / n;
    const float kk = (z - y) / n;
    // One interior abscissa each way: (x - w) / (n * 1 + w), (z - y) / (n * 1 + y).
    const float x_k = (x - w) / (static_cast<float>(n * 1) + w);
    const float y_k = (z - y) / (static_cast<float>(n * 1) + y);
    Rgb3 sum2{0.0f, 0.0f, 0.0f};
    sum2 = add3(sum2, add3(add3(add3(psf(kind, x_k, y, pixdeg), psf(kind, x_k, z, pixdeg)), psf(kind, w, y_k, pixdeg)),
                           psf(kind, x, y_k, pixdeg)));
    sum1 = add3(sum1, Rgb3{2.0f * sum2.r, 2.0f * sum2.g, 2.0f * sum2.b});
    Rgb3 sum3{0.0f, 0.0f, 0.0f};
    sum3 = add3(sum3, psf(kind, x_k, y_k, pixdeg));
    sum1 = add3(sum1, Rgb3{4.0f * sum3.r, 4.0f * sum3.g, 4.0f * sum3.b});
    const float s = 0.25f * h * kk;
    return Rgb3{sum1.r * s, sum1.g * s, sum1.b * s};
}

// Spread and sum the three images into the filter (rgb), with per-block
// partial sums of each channel.
__global__ void __launch_bounds__(kThreads) k_glare_psf(const float *__restrict__ lines_img,
                                                        const float *__restrict__ corona_sum, int xr, int yr,
                                                        float pixdeg, int level, float *__restrict__ gf) {
    const int64_t n = static_cast<int64_t>(xr) * yr;
    const int64_t pix = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    const float xrf = static_cast<float>(static_cast<unsigned>(xr)), yrf = static_cast<float>(static_cast<unsigned>(yr));
    if (pix >= n) return;
    {
        const int i = static_cast<int>(pix % xr), j = static_cast<int>(pix / xr);
        const float bloom = corona_sum[0] / static_cast<float>(static_cast<unsigned>(xr) * static_cast<unsigned>(yr));
        const float v0 = lines_img[pix], v1 = lines_img[n + pix];
        const Rgb3 s0 = psf_integral(3 * level + 0, i, j, xrf, yrf, pixdeg);
        const Rgb3 s1 = psf_integral(3 * level + 1, i, j, xrf, yrf, pixdeg);
        const Rgb3 s2 = psf_integral(3 * level + 2, i, j, xrf, yrf, pixdeg);
        Rgb3 g{v0 * s0.r, v0 * s0.g, v0 * s0.b};
        g = add3(g, Rgb3{v1 * s1.r, v1 * s1.g, v1 * s1.b});
        g = add3(g, Rgb3{bloom * s2.r, bloom * s2.g, bloom * s2.b});
        gf[3 * pix + 0] = g.r;
        gf[3 * pix + 1] = g.g;
        gf[3 * pix + 2] = g.b;
    }
}

// Scale so the filter totals 1e-4 per pixel per channel (rt/camera.cc:406-435).
__global__ void __launch_bounds__(kThreads) k_glare_scale(float *__restrict__ gf, int64_t n, const float *__restrict__ sum,
                                                          float scat) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= 3 * n) return;
    const float s = sum[i % 3];
    const float mul = (s > 0.0f) ? scat / s : 1.0f;
    gf[i] = gf[i] * mul;
}

// The pixels that take part in the convolution: a channel above 255
// (rt/camera.cc:1554-1568), marked before any channel gets its glare.
__global__ void __launch_bounds__(kThreads) k_glare_mask(const float *__restrict__ img, int64_t n,
                                                         uint8_t *__restrict__ mask) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= n) return;
    const float *p = img + 3 * i;
    mask[i] = ((p[0] > 255.0f) || (p[1] > 255.0f) || (p[2] > 255.0f)) ? 1 : 0;
}

// Zero-padded complex operands of the convolution, My rows of Mx.
// FILTER = false: channel c of the image at the marked pixels. FILTER = true:
// channel c of the filter moved so that its centre lands on pixel 0:
// glare_in[i] = gf[i + off] for i < off, gf[i - off] after, in the image's
// linear pixel order (rt/camera.cc:1571-1587).
template <bool FILTER>
__global__ void __launch_bounds__(kThreads) k_glare_pad(const float *__restrict__ src, const uint8_t *__restrict__ mask,
                                                        int xr, int yr, int c, int64_t off, int mx, int my,
                                                        float2 *__restrict__ dst) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (e >= static_cast<int64_t>(mx) * my) return;
    const int x = static_cast<int>(e % mx), y = static_cast<int>(e / mx);
    float v = 0.0f;
    if (x < xr && y < yr) {
        const int64_t i = static_cast<int64_t>(y) * xr + x;
        if (FILTER) {
            v = src[3 * ((i < off) ? (i + off) : (i - off)) + c];
        } else if (mask[i]) {
            v = src[3 * i + c];
        }
    }
    dst[e] = make_float2(v, 0.0f);
}

// In-place radix-2 FFT of every row of length M (a power of two, <= kMaxFft):
// bit-reversed load into LDS, log2(M) butterfly stages, store.
// INV: the conjugate twiddles (unnormalised inverse).
template <bool INV>
__global__ void __launch_bounds__(512) k_fft_rows(float2 *__restrict__ data, int M, int logM,
                                                  const float2 *__restrict__ tw) {
    __shared__ float2 x[kMaxFft];
    float2 *row = data + static_cast<int64_t>(blockIdx.x) * M;
    for (int i = threadIdx.x; i < M; i += blockDim.x) x[__brev(static_cast<unsigned>(i)) >> (32 - logM)] = row[i];
    __syncthreads();
    for (int half = 1; half < M; half <<= 1) {
        const int step = M / (2 * half);
        for (int b = threadIdx.x; b < M / 2; b += blockDim.x) {
            const int pos = b & (half - 1);
            const int i0 = ((b - pos) << 1) + pos, i1 = i0 + half;
            float2 w = tw[pos * step];
            if (INV) w.y = -w.y;
            const float2 u = x[i0], v = x[i1];
            const float2 t = make_float2(v.x * w.x - v.y * w.y, v.x * w.y + v.y * w.x);
            x[i0] = make_float2(u.x + t.x, u.y + t.y);
            x[i1] = make_float2(u.x - t.x, u.y - t.y);
        }
        __syncthreads();
    }
    for (int i = threadIdx.x; i < M; i += blockDim.x) row[i] = x[i];
}

// out (cols x rows) = transpose of in (rows x cols), 32 x 32 tiles.
__global__ void __launch_bounds__(256) k_transpose(const float2 *__restrict__ in, float2 *__restrict__ out, int rows,
                                                   int cols) {
    __shared__ float2 tile[32][33];
    const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
    for (int k = threadIdx.y; k < 32; k += 8) {
        const int r = by + k, c = bx + threadIdx.x;
        if (r < rows && c < cols) tile[k][threadIdx.x] = in[static_cast<int64_t>(r) * cols + c];
    }
    __syncthreads();
    for (int k = threadIdx.y; k < 32; k += 8) {
        const int r = bx + k, c = by + threadIdx.x;  // out row = in column
        if (r < cols && c < rows) out[static_cast<int64_t>(r) * rows + c] = tile[threadIdx.x][k];
    }
}

__global__ void __launch_bounds__(kThreads) k_cmul(float2 *__restrict__ a, const float2 *__restrict__ b, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= n) return;
    const float2 u = a[i], v = b[i];
    a[i] = make_float2(u.x * v.x - u.y * v.y, u.x * v.y + u.y * v.x);
}

// Fold the linear convolution (My rows of Mx, times Mx * My from the
// unnormalised transform pair) onto the circular one and add it to channel c
// as the reference does: n times the convolution, times 1 / (n * 255).
__global__ void __launch_bounds__(kThreads) k_glare_fold_add(const float2 *__restrict__ a, int mx, int my, int xr,
                                                             int yr, int c, float inv_m, float npix_f,
                                                             float pixel_divisor, float *__restrict__ img) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= static_cast<int64_t>(xr) * yr) return;
    const int x = static_cast<int>(i % xr), y = static_cast<int>(i / xr);
    float s = a[static_cast<int64_t>(y) * mx + x].x;
    if (x + xr < mx) s += a[static_cast<int64_t>(y) * mx + x + xr].x;
    if (y + yr < my) {
        s += a[static_cast<int64_t>(y + yr) * mx + x].x;
        if (x + xr < mx) s += a[static_cast<int64_t>(y + yr) * mx + x + xr].x;
    }
    const float conv = s * inv_m;
    img[3 * i + c] += (conv * npix_f) * pixel_divisor;
}

int pow2_at_least(int v) {
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

int log2i(int p) {
    int l = 0;
    while ((1 << l) < p) ++l;
    return l;
}

void free_filter(GlareFilter &f) {
    if (f.gf) (void)hipFree(f.gf);
    for (auto &sp : f.spec)
        if (sp) (void)hipFree(sp);
    f = GlareFilter{};
}

// 2D forward transform of a (My rows of Mx) into b (transposed: Mx rows of My).
void forward_2d(GlareCache &g, float2 *a, float2 *b, hipStream_t s) {
    hipLaunchKernelGGL(k_fft_rows<false>, dim3(g.my), dim3(512), 0, s, a, g.mx, log2i(g.mx), g.tw_x);
    hipLaunchKernelGGL(k_transpose, dim3((g.mx + 31) / 32, (g.my + 31) / 32), dim3(32, 8), 0, s, a, b, g.my, g.mx);
    hipLaunchKernelGGL(k_fft_rows<false>, dim3(g.mx), dim3(512), 0, s, b, g.my, log2i(g.my), g.tw_y);
}

}  // namespace

void release_glare(rt_ctx *ctx) {
    GlareCache *g = ctx->glare;
    if (!g) return;
    for (auto &f : g->f) free_filter(f);
    if (g->tw_x) (void)hipFree(g->tw_x);
    if (g->tw_y) (void)hipFree(g->tw_y);
    if (g->a) (void)hipFree(g->a);
    if (g->b) (void)hipFree(g->b);
    if (g->mask) (void)hipFree(g->mask);
    delete g;
    ctx->glare = nullptr;
}

// Light level of an adapted luminance (rt/camera.cc:1508-1527).
int glare_level(float yi) {
    const float l = std::log10(yi);
    return (l < -2.0f) ? 0 : ((l < 0.6f) ? 1 : 2);
}

// The filter of one light level for x_res x y_res and field of view x_angle,
// built (and its spectrum taken) when missing.
int32_t glare_filter(rt_ctx *ctx, int xr, int yr, float x_angle, int level, const float **gf_out) {
    if (xr <= 0 || yr <= 0 || xr > kMaxFft / 2 || yr > kMaxFft / 2)
        return fail(ctx, RT_ERR_INVALID, "glare: frames up to 4096 x 4096 pixels");
    if (!(x_angle > 0.0f) || level < 0 || level > 2)
        return fail(ctx, RT_ERR_INVALID, "glare: needs the camera's field of view (x_angle > 0) and a level 0..2");
    hipStream_t s = ctx->stream;
    if (!ctx->glare) ctx->glare = new GlareCache();
    GlareCache &g = *ctx->glare;
    // At least 2 points per axis (a one-point transform has no stages).
    const int mx = pow2_at_least(std::max(2, 2 * xr - 1)), my = pow2_at_least(std::max(2, 2 * yr - 1));
    if (g.mx != mx || g.my != my) {
        for (auto &f : g.f) free_filter(f);
        if (g.tw_x) (void)hipFree(g.tw_x);
        if (g.tw_y) (void)hipFree(g.tw_y);
        if (g.a) (void)hipFree(g.a);
        if (g.b) (void)hipFree(g.b);
        if (g.mask) (void)hipFree(g.mask);
        g.tw_x = g.tw_y = g.a = g.b = nullptr;
        g.mask = nullptr;
        g.mx = g.my = 0;
        const size_t cells = static_cast<size_t>(mx) * my;
        RT_HIP_CHECK(ctx, hipMalloc(&g.tw_x, (mx / 2 + 1) * sizeof(float2)));
        RT_HIP_CHECK(ctx, hipMalloc(&g.tw_y, (my / 2 + 1) * sizeof(float2)));
        RT_HIP_CHECK(ctx, hipMalloc(&g.a, cells * sizeof(float2)));
        RT_HIP_CHECK(ctx, hipMalloc(&g.b, cells * sizeof(float2)));
        RT_HIP_CHECK(ctx, hipMalloc(&g.mask, cells));
        for (int t = 0; t < 2; ++t) {
            const int M = t ? my : mx;
            std::vector<float2> tw(M / 2 + 1);
            for (int m = 0; m <= M / 2; ++m) {
                const double a = -2.0 * 3.14159265358979323846 * m / M;
                tw[m] = make_float2(static_cast<float>(std::cos(a)), static_cast<float>(std::sin(a)));
            }
            RT_HIP_CHECK(ctx, hipMemcpy(t ? g.tw_y : g.tw_x, tw.data(), tw.size() * sizeof(float2),
                                        hipMemcpyHostToDevice));
        }
        g.mx = mx;
        g.my = my;
    }
    GlareFilter &f = g.f[level];
    if (f.gf && (f.x_res != xr || f.y_res != yr || f.x_angle != x_angle)) free_filter(f);
    if (!f.gf) {
        const int64_t n = static_cast<int64_t>(xr) * yr;
        const int blocks = static_cast<int>((n + kThreads - 1) / kThreads);
        // Line parameters: the reference's draws and float library calls.
        std::vector<Line> lines(2 * kLines);
        Mt rng;
        const float yrf = static_cast<float>(static_cast<unsigned>(yr));
        for (int img = 0; img < 2; ++img) {
            for (int i = 0; i < kLines; ++i) {
                Line &ln = lines[img * kLines + i];
                ln.intens = rng.next();
                const float lineRad = 1.0f;
                const float theta = ((i + rng.next()) * (2.0f * kPi)) / kLines;
                ln.c = std::cos(theta);
                ln.s = std::sin(theta);
                ln.t = std::tan(theta);
                ln.dif = std::fabs(lineRad / std::sin(theta));
                ln.dj = (theta <= kPi) ? 1 : -1;
                ln.j0 = static_cast<int>(yrf / 2.0f - (lineRad - 0.5f) * ln.dj);
            }
        }
        Line *d_lines = nullptr;
        float *d_img = nullptr, *d_vals = nullptr, *d_sums = nullptr;
        int *d_cnt = nullptr;
        int64_t *d_off = nullptr;
        RT_HIP_CHECK(ctx, hipMalloc(&f.gf, 3 * n * sizeof(float)));
        f.x_res = xr;
        f.y_res = yr;
        f.x_angle = x_angle;
        f.level = level;
        auto release = [&]() {
            for (void *p : {static_cast<void *>(d_lines), static_cast<void *>(d_img), static_cast<void *>(d_vals),
                            static_cast<void *>(d_sums), static_cast<void *>(d_cnt), static_cast<void *>(d_off)})
                if (p) (void)hipFree(p);
        };
        auto bail = [&](hipError_t e) -> int32_t {
            release();
            free_filter(f);
            return fail(ctx, e == hipErrorOutOfMemory ? RT_ERR_OOM : RT_ERR_HIP,
                        std::string("glare filter: ") + hipGetErrorString(e));
        };
        // The corona's (line, row) slots: 200 rows of y_res.
        const int64_t m = static_cast<int64_t>(kLines) * yr;
        const dim3 slot_grid(static_cast<unsigned>((m + kThreads - 1) / kThreads));
        hipError_t e = hipMalloc(&d_lines, lines.size() * sizeof(Line));
        if (e == hipSuccess) e = hipMalloc(&d_img, 2 * n * sizeof(float));
        if (e == hipSuccess) e = hipMalloc(&d_sums, 8 * sizeof(float));  // [0] corona sum, [4..6] filter sums
        if (e == hipSuccess) e = hipMalloc(&d_cnt, m * sizeof(int));
        if (e == hipSuccess) e = hipMalloc(&d_off, (m + 1) * sizeof(int64_t));
        if (e == hipSuccess) e = hipMemcpyAsync(d_lines, lines.data(), lines.size() * sizeof(Line), hipMemcpyHostToDevice, s);
        if (e != hipSuccess) return bail(e);
        hipLaunchKernelGGL(k_glare_lines, dim3(blocks, 2), dim3(kThreads), 0, s, d_lines, xr, yr, d_img);
        hipLaunchKernelGGL(k_glare_line_rows<false>, slot_grid, dim3(kThreads), 0, s, d_lines, xr, yr, d_cnt,
                           static_cast<const int64_t *>(nullptr), static_cast<float *>(nullptr));
        hipLaunchKernelGGL(k_glare_scan, dim3(1), dim3(1024), 0, s, d_cnt, m, d_off);
        int64_t nvals = 0;
        e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(&nvals, d_off + m, sizeof(int64_t), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e == hipSuccess) e = hipMalloc(&d_vals, std::max<int64_t>(nvals, 1) * sizeof(float));
        if (e != hipSuccess) return bail(e);
        hipLaunchKernelGGL(k_glare_line_rows<true>, slot_grid, dim3(kThreads), 0, s, d_lines, xr, yr, d_cnt, d_off,
                           d_vals);
        hipLaunchKernelGGL(k_glare_seq_sum<1>, dim3(1), dim3(256), 0, s, d_vals, nvals, d_sums);
        const float pixdeg = (std::atan(x_angle / 2.0f) * 2.0f) / static_cast<float>(static_cast<unsigned>(xr));
        hipLaunchKernelGGL(k_glare_psf, dim3(blocks), dim3(kThreads), 0, s, d_img, d_sums, xr, yr, pixdeg, level,
                           f.gf);
        hipLaunchKernelGGL(k_glare_seq_sum<3>, dim3(1), dim3(256), 0, s, f.gf, n, d_sums + 4);
        const float scat = 0.0001f * static_cast<float>(static_cast<unsigned>(xr)) *
                           static_cast<float>(static_cast<unsigned>(yr));
        hipLaunchKernelGGL(k_glare_scale, dim3(static_cast<unsigned>((3 * n + kThreads - 1) / kThreads)),
                           dim3(kThreads), 0, s, f.gf, n, d_sums + 4, scat);
        // The filter's spectrum per channel (transposed layout).
        const int64_t off = (xr >> 1) + static_cast<int64_t>(xr) * ((yr - 1) >> 1);
        const size_t cells = static_cast<size_t>(mx) * my;
        const dim3 pad_grid(static_cast<unsigned>((cells + kThreads - 1) / kThreads));
        for (int c = 0; c < 3 && e == hipSuccess; ++c) {
            e = hipMalloc(&f.spec[c], cells * sizeof(float2));
            if (e != hipSuccess) break;
            hipLaunchKernelGGL(k_glare_pad<true>, pad_grid, dim3(kThreads), 0, s, f.gf, nullptr, xr, yr, c, off, mx,
                               my, g.a);
            forward_2d(g, g.a, f.spec[c], s);
        }
        if (e == hipSuccess) e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return bail(e);
        release();
    }
    if (gf_out) *gf_out = f.gf;
    return RT_OK;
}

// perform_glare_filter on a tone-mapped device image at adapted luminance yi.
int32_t apply_glare(rt_ctx *ctx, float *img, int xr, int yr, float yi, float x_angle) {
    const int level = glare_level(yi);
    if (int32_t rc = glare_filter(ctx, xr, yr, x_angle, level, nullptr)) return rc;
    GlareCache &g = *ctx->glare;
    GlareFilter &f = g.f[level];
    hipStream_t s = ctx->stream;
    const int64_t n = static_cast<int64_t>(xr) * yr;
    const size_t cells = static_cast<size_t>(g.mx) * g.my;
    const dim3 pad_grid(static_cast<unsigned>((cells + kThreads - 1) / kThreads));
    const float inv_m = 1.0f / static_cast<float>(cells);
    const float pixel_divisor = 1.0f / static_cast<float>(n * 255);
    const dim3 pix_grid(static_cast<unsigned>((n + kThreads - 1) / kThreads));
    hipLaunchKernelGGL(k_glare_mask, pix_grid, dim3(kThreads), 0, s, img, n, g.mask);
    for (int c = 0; c < 3; ++c) {
        hipLaunchKernelGGL(k_glare_pad<false>, pad_grid, dim3(kThreads), 0, s, img, g.mask, xr, yr, c, int64_t(0),
                           g.mx, g.my, g.a);
        forward_2d(g, g.a, g.b, s);
        hipLaunchKernelGGL(k_cmul, pad_grid, dim3(kThreads), 0, s, g.b, f.spec[c], static_cast<int64_t>(cells));
        hipLaunchKernelGGL(k_fft_rows<true>, dim3(g.mx), dim3(512), 0, s, g.b, g.my, log2i(g.my), g.tw_y);
        hipLaunchKernelGGL(k_transpose, dim3((g.my + 31) / 32, (g.mx + 31) / 32), dim3(32, 8), 0, s, g.b, g.a, g.mx,
                           g.my);
        hipLaunchKernelGGL(k_fft_rows<true>, dim3(g.my), dim3(512), 0, s, g.a, g.mx, log2i(g.mx), g.tw_x);
        hipLaunchKernelGGL(k_glare_fold_add, pix_grid, dim3(kThreads), 0, s, g.a, g.mx, g.my, xr, yr, c, inv_m,
                           static_cast<float>(n), pixel_divisor, img);
    }
    RT_HIP_CHECK(ctx, hipGetLastError());
    return RT_OK;
}

}  // namespace rt

extern "C" int32_t rt_glare_filter(rt_ctx *ctx, int32_t x_res, int32_t y_res, float x_angle, int32_t level,
                                   float *out, uint32_t flags) {
    if (!ctx || !out) return RT_ERR_INVALID;
    RT_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    const float *gf = nullptr;
    if (int32_t rc = rt::glare_filter(ctx, x_res, y_res, x_angle, level, &gf)) return rc;
    const size_t bytes = static_cast<size_t>(x_res) * y_res * 3 * sizeof(float);
    RT_HIP_CHECK(ctx, hipMemcpyAsync(out, gf, bytes, (flags & RT_DEVICE_PTRS) ? hipMemcpyDeviceToDevice
                                                                             : hipMemcpyDeviceToHost,
                                     ctx->stream));
    RT_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return RT_OK;
}
