// Oracle — TEST INFRASTRUCTURE ONLY (see vec.hpp for the rules).
//
// The glare filter of camera::tone_map (rt/camera.cc:1455-1458): the
// reference's random generator (common/common.cc:30-105), the flare lines
// (rt/camera.cc:145-220), the point-spreading functions and their per-pixel
// integration (rt/camera.cc:1-141, 224-322), the filter's composition and
// normalisation (rt/camera.cc:325-462) and perform_glare_filter
// (rt/camera.cc:1504-1630). Every loop runs in the reference's order and
// every float expression keeps its operand order, so float sums round as they
// do there. The reference convolves with single-precision FFTs (fftw3f); here
// the circular convolution is summed directly in double over the pixels that
// take part (those with a channel above 255).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "rt_oracle.h"

namespace {

const float kPi = 3.1415926535897932384626433f;

// common/common.cc:30-105: MT19937 seeded 5489 on first use, tempered to a
// float in [0, 1). Its refill loop stops at index 622, so mt[623] keeps the
// value of the previous round (after seeding: the initialisation value).
struct RefMt {
    uint32_t mt[624];
    int index = 625;
    void seed(uint32_t s) {
        mt[0] = s;
        for (int i = 1; i < 624; ++i) mt[i] = 1812433253u * (mt[i - 1] ^ (mt[i - 1] >> 30)) + static_cast<uint32_t>(i);
    }
    void refill() {
        for (int i = 0; i < 227; ++i) {
            const uint32_t y = (0x80000000u & mt[i]) + (0x7FFFFFFFu & mt[i + 1]);
            mt[i] = (y & 1u) ? (mt[i + 397] ^ (y >> 1) ^ 0x9908b0dfu) : (mt[i + 397] ^ (y >> 1));
        }
        for (int i = 227; i < 623; ++i) {
            const uint32_t y = (0x80000000u & mt[i]) + (0x7FFFFFFFu & mt[i + 1]);
            mt[i] = (y & 1u) ? (mt[i - 227] ^ (y >> 1) ^ 0x9908b0dfu) : (mt[i - 227] ^ (y >> 1));
        }
    }
    uint32_t raw() {
        if (index >= 624) {
            if (index == 625) seed(5489u);
            refill();
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

struct Rgb {
    float r = 0.0f, g = 0.0f, b = 0.0f;
};

// rt/camera.cc:145-220 (lineRad = 1): 200 lines from the image centre at
// stratified random angles; each row a line crosses gets the tent-filtered
// distance of its pixels to the line (the reference's perpendicular vector
// repeats its x component), times the line's circumference weight and
// random intensity. Returns the sum of every value written.
float flare_lines(RefMt &rng, int xr, int yr, Rgb *f) {
    float sum = 0.0f;
    const int nr_of_lines = 200;
    const float xrf = static_cast<float>(static_cast<unsigned>(xr)), yrf = static_cast<float>(static_cast<unsigned>(yr));
    for (int i = 0; i < nr_of_lines; i++) {
        const float lineIntens = rng.next();
        const float lineRad = 1.0;
        const float theta = ((i + rng.next()) * (2.0f * kPi)) / nr_of_lines;
        const float dir[2] = {std::cos(theta), std::sin(theta)};
        const int dj = (theta <= kPi) ? 1 : -1;
        int j = static_cast<int>(yrf / 2.0f - (lineRad - 0.5f) * dj);
        while ((j >= 0) && (j < yr)) {
            const float jminf = (j + 0.0f - yrf / 2.0f);
            const float jmaxf = (j + 1.0f - yrf / 2.0f);
            float iminf = jminf / std::tan(theta);
            float imaxf = jmaxf / std::tan(theta);
            if (iminf > imaxf) std::swap(iminf, imaxf);
            const float dif = std::fabs(lineRad / std::sin(theta));
            const int imin = static_cast<int>(iminf - dif + 0.5f + xrf / 2.0f);
            const int imax = static_cast<int>(imaxf + dif - 0.5f + xrf / 2.0f);
            for (int k = std::max(imin, 0); k < std::min((imax + 1), xr); k++) {
                float p[2];
                p[0] = k - xrf / 2.0f + 0.5f;
                p[1] = j - yrf / 2.0f + 0.5f;
                const float dp = (p[0] * dir[0]) + (p[1] * dir[1]);
                float dist;
                if (dp > 0.0f) {
                    const float vec[2] = {(p[0] - dp * dir[0]), (p[0] - dp * dir[0])};
                    dist = std::sqrt((vec[0] * vec[0]) + (vec[1] * vec[1]));
                } else {
                    dist = std::sqrt((p[0] * p[0]) + (p[1] * p[1]));
                }
                if (dist >= lineRad) continue;
                const float filt = 1.0f / lineRad - 1.0f / (lineRad * lineRad) * dist;
                const float len = std::sqrt((p[0] * p[0]) + (p[1] * p[1]));
                const float flo = std::max((len * 2.0f * kPi) / nr_of_lines, 1.0f / nr_of_lines);
                const float val = filt * flo * lineIntens;
                Rgb &o = f[k + (j * xr)];
                o.r += val;
                o.g += val;
                o.b += val;
                sum += val;
            }
            j += dj;
        }
    }
    return sum;
}

// Spencer's glare components (rt/camera.cc:12-35).
float f0(float t) {
    const float tmp = t * (1.0f / 0.02f);
    return 2.61e6f * std::exp(-(tmp * tmp));
}
float f1(float t) {
    const float tmp = t + 0.02f;
    return 20.91f / (tmp * tmp * tmp);
}
float f2(float t) {
    const float tmp = t + 0.02f;
    return 72.37f / (tmp * tmp);
}
float f3(float t, float l) {
    const float tmp = t - 3.0f * l * (1.0f / 568.0f);
    return 436.9f * (568.0f / l) * std::exp(-19.75f * (tmp * tmp));
}

// The nine spreading functions (rt/camera.cc:37-141); kind = 3 * level +
// component (level 0 scotopic, 1 mesopic, 2 photopic; component 0 corona,
// 1 halo, 2 bloom).
Rgb psf(int kind, float x, float y, float p) {
    const float dist = std::sqrt((x * x) + (y * y));
    const float deg = dist * p;
    auto grey = [](float a) { return Rgb{a, a, a}; };
    auto halo = [&](float s) { return Rgb{s * f3(deg, 611.0f), s * f3(deg, 549.0f), s * f3(deg, 467.0f)}; };
    switch (kind) {
        case 0: return grey((0.478f * f1(deg)) + (0.207f * f2(deg)));
        case 1: return halo(0.033f);
        case 2: return grey(0.282f * f0(deg));
        case 3: return grey((0.478f * f1(deg)) + (0.138f * f2(deg)));
        case 4: return halo(0.016f);
        case 5: return grey(0.368f * f0(deg));
        case 6: return grey((0.478f * f1(deg)) + (0.138f * f2(deg)));
        case 7: return Rgb{0.0f, 0.0f, 0.0f};
        default: return grey(0.384f * f0(deg));
    }
}

Rgb add(Rgb a, Rgb b) { return Rgb{a.r + b.r, a.g + b.g, a.b + b.b}; }

// rt/camera.cc:224-322: each pixel times a trapezoidal integral of the
// spreading function over it, with the reference's sample points (its
// interior abscissae are (x - w) / (n k + w), kept as written; samps = 3).
void apply_psf(int kind, int xr, int yr, float pixdeg, Rgb *f) {
    const float xrf = static_cast<float>(static_cast<unsigned>(xr)), yrf = static_cast<float>(static_cast<unsigned>(yr));
    for (int j = 0; j < yr; ++j) {
        for (int i = 0; i < xr; ++i) {
            const float w = static_cast<float>(i) - xrf / 2.0f, x = w + 1.0f;
            const float y = static_cast<float>(j) - yrf / 2.0f, z = y + 1.0f;
            Rgb sum1 = psf(kind, w, y, pixdeg);
            sum1 = add(sum1, psf(kind, x, y, pixdeg));
            sum1 = add(sum1, psf(kind, w, z, pixdeg));
            sum1 = add(sum1, psf(kind, x, z, pixdeg));
            const int samps = 3;
            const int n = samps - 1;
            const float h = (x - w) / n;
            const float kk = (z - y) / n;
            Rgb sum2;
            for (int k = 1; k < n; k++) {
                const float x_k = (x - w) / static_cast<float>(n * k + w);
                const float y_k = (z - y) / static_cast<float>(n * k + y);
                sum2 = add(sum2, add(add(add(psf(kind, x_k, y, pixdeg), psf(kind, x_k, z, pixdeg)),
                                         psf(kind, w, y_k, pixdeg)),
                                     psf(kind, x, y_k, pixdeg)));
            }
            sum1 = add(sum1, Rgb{2.0f * sum2.r, 2.0f * sum2.g, 2.0f * sum2.b});
            sum2 = Rgb{};
            for (int k = 1; k < n; ++k) {
                for (int l = 1; l < n; ++l) {
                    const float x_k = (x - w) / static_cast<float>(n * k + w);
                    const float y_l = (z - y) / static_cast<float>(n * l + y);
                    sum2 = add(sum2, psf(kind, x_k, y_l, pixdeg));
                }
            }
            sum1 = add(sum1, Rgb{4.0f * sum2.r, 4.0f * sum2.g, 4.0f * sum2.b});
            const float s = 0.25f * h * kk;
            sum1 = Rgb{sum1.r * s, sum1.g * s, sum1.b * s};
            Rgb &o = f[i + (j * xr)];
            o.r *= sum1.r;
            o.g *= sum1.g;
            o.b *= sum1.b;
        }
    }
}

// rt/camera.cc:325-462: corona and halo from two sets of flare lines, a flat
// bloom of their mean level, each spread by its function, summed and scaled
// so that the filter's total is 1e-4 per pixel.
void glare_filter(int xr, int yr, float pixdeg, int level, Rgb *out) {
    const int64_t npix = static_cast<int64_t>(xr) * yr;
    std::vector<Rgb> gf(3 * npix);
    RefMt rng;
    const float flare_sum = flare_lines(rng, xr, yr, gf.data());
    flare_lines(rng, xr, yr, gf.data() + npix);
    const float bloom_lum = flare_sum / static_cast<float>(static_cast<unsigned>(xr) * static_cast<unsigned>(yr));
    for (int64_t i = 0; i < npix; ++i) gf[2 * npix + i] = Rgb{bloom_lum, bloom_lum, bloom_lum};
    for (int c = 0; c < 3; ++c) apply_psf(3 * level + c, xr, yr, pixdeg, gf.data() + c * npix);
    Rgb sum;
    for (int64_t i = 0; i < npix; ++i) {
        gf[i] = add(gf[i], gf[npix + i]);
        gf[i] = add(gf[i], gf[2 * npix + i]);
        sum = add(sum, gf[i]);
    }
    const float amountScattered = 0.0001f;
    const float scat = amountScattered * static_cast<float>(static_cast<unsigned>(xr)) *
                       static_cast<float>(static_cast<unsigned>(yr));
    Rgb mul{1.0f, 1.0f, 1.0f};
    if (sum.r > 0.0f) mul.r = scat / sum.r;
    if (sum.g > 0.0f) mul.g = scat / sum.g;
    if (sum.b > 0.0f) mul.b = scat / sum.b;
    for (int64_t i = 0; i < npix; ++i) out[i] = Rgb{gf[i].r * mul.r, gf[i].g * mul.g, gf[i].b * mul.b};
}

int light_level(float yi) {
    const float l = std::log10(yi);
    return (l < -2.0f) ? 0 : ((l < 0.6f) ? 1 : 2);
}

}  // namespace

extern "C" {

void oracle_mt_stream(int32_t count, uint32_t *out) {
    RefMt rng;
    for (int i = 0; i < count; ++i) out[i] = rng.raw();
}

// x_angle = 2 w / t: pixdeg = (atan(w / t) * 2) / x_res as the reference
// computes it from -x_m / t.
void oracle_glare_filter(int32_t x_res, int32_t y_res, float x_angle, int32_t level, float *out) {
    const float pixdeg = (std::atan(x_angle / 2.0f) * 2.0f) / static_cast<float>(static_cast<unsigned>(x_res));
    glare_filter(x_res, y_res, pixdeg, level, reinterpret_cast<Rgb *>(out));
}

// perform_glare_filter (rt/camera.cc:1504-1630) on an image of 0-255 floats
// after the tone map: the pixels with a channel above 255 convolved
// (circularly, on the x_res x y_res grid) with the filter shifted so that its
// centre lands on pixel 0, divided by 255 and added to the image.
int32_t oracle_glare(float *img, int32_t x_res, int32_t y_res, float yi, float x_angle, int32_t *level_out) {
    if (!(x_angle > 0.0f) || x_res <= 0 || y_res <= 0) return -1;
    const int level = light_level(yi);
    if (level_out) *level_out = level;
    const int64_t n = static_cast<int64_t>(x_res) * y_res;
    std::vector<Rgb> gf(n);
    oracle_glare_filter(x_res, y_res, x_angle, level, reinterpret_cast<float *>(gf.data()));
    // glare_in[i] = gf[i + off] for i < off, gf[i - off] after
    // (rt/camera.cc:1573-1587).
    const int64_t off = (x_res >> 1) + static_cast<int64_t>(x_res) * ((y_res - 1) >> 1);
    std::vector<Rgb> gin(n);
    for (int64_t i = 0; i < n; ++i) gin[i] = (i < off) ? gf[i + off] : gf[i - off];
    std::vector<int64_t> bright;
    for (int64_t i = 0; i < n; ++i) {
        const float *p = img + 3 * i;
        if ((p[0] > 255.0f) || (p[1] > 255.0f) || (p[2] > 255.0f)) bright.push_back(i);
    }
    const float pixel_divisor = 1.0f / static_cast<float>(n * 255);
    std::vector<double> acc(3 * n, 0.0);
    for (const int64_t q : bright) {
        const int qx = static_cast<int>(q % x_res), qy = static_cast<int>(q / x_res);
        const float *src = img + 3 * q;
        for (int py = 0; py < y_res; ++py) {
            const int dy = (py - qy + y_res) % y_res;
            for (int px = 0; px < x_res; ++px) {
                const int dx = (px - qx + x_res) % x_res;
                const Rgb &g = gin[static_cast<int64_t>(dy) * x_res + dx];
                double *a = &acc[3 * (static_cast<int64_t>(py) * x_res + px)];
                a[0] += static_cast<double>(src[0]) * g.r;
                a[1] += static_cast<double>(src[1]) * g.g;
                a[2] += static_cast<double>(src[2]) * g.b;
            }
        }
    }
    // The reference's inverse transform is unnormalised (n times the
    // convolution), then scaled by 1 / (n * 255).
    for (int64_t i = 0; i < 3 * n; ++i)
        img[i] += static_cast<float>(acc[i] * static_cast<double>(n)) * pixel_divisor;
    return 0;
}

}  // extern "C"
