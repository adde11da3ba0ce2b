// Oracle — TEST INFRASTRUCTURE ONLY.
//
// Camera post-process as the reference runs it on camera::image (0-255
// floats, pixel (x, y) at x + y * x_res): tone_map (rt/camera.cc:1124-1461)
// for its global modes, gamma_correct (rt/camera.cc:1466-1501) and the
// down-sampling clip to bytes of the image writers (rt/camera.h:289-351).
// Every loop runs in the reference's order, so float sums round as they do
// there.
#include <algorithm>
#include <cmath>
#include <cstdint>

#include "rt_oracle.h"

namespace {

constexpr float kMaxDist = 3.402823466e+38f;  // common/common.h MAX_DIST

// camera::convert_Yxy_to_rgb (rt/camera.cc:2080-2099): pixel (x, Y, y) -> rgb.
void yxy_to_rgb(float *p) {
    const float x = p[0], y = p[2], Y = p[1];
    const float X = x * Y / y;
    const float Z = (1.0f - x - y) * Y / y;
    p[0] = (826.353f * X) + (-391.986f * Y) + (-127.143f * Z);
    p[1] = (-247.0695f * X) + (478.329f * Y) + (10.5825f * Z);
    p[2] = (14.2035f * X) + (-52.02f * Y) + (269.535f * Z);
}

}  // namespace

extern "C" int32_t oracle_tone_map(float *img, int32_t x_res, int32_t y_res, const oracle_tone_params *tp,
                                   float *adaptation_level, float *yi_out) {
    static const float xb[3] = {0.4124f, 0.3576f, 0.1805f};
    static const float yb[3] = {0.2126f, 0.7151f, 0.0721f};
    static const float zb[3] = {0.0193f, 0.1192f, 0.9505f};
    static const float vb[3] = {0.01477f, 0.497f, 0.631f};
    const int mode = tp->mode;
    if (mode == 2 || mode == 3 || mode == 8 || mode < 0 || mode > 9 || tp->glare) return -1;
    const int n = x_res * y_res;
    float Yi = 0.0f, maxY = 0.0f, rgbAvg = 0.0f;
    int numpix = 0;
    for (int i = 0; i < n; ++i) {
        float *p = img + 3 * static_cast<int64_t>(i);
        p[0] /= 255.0f;
        p[1] /= 255.0f;
        p[2] /= 255.0f;
        if (mode == 5) rgbAvg += (p[0] + p[1] + p[2]) * (1.0f / 3.0f);
        float X = 683.0f * ((xb[0] * p[0]) + (xb[1] * p[1]) + (xb[2] * p[2]));
        float Y = 683.0f * ((yb[0] * p[0]) + (yb[1] * p[1]) + (yb[2] * p[2]));
        float Z = 683.0f * ((zb[0] * p[0]) + (zb[1] * p[1]) + (zb[2] * p[2]));
        if (Y < 3.18e-7f) {
            Y = 3.18e-7f;
        } else {
            ++numpix;
            maxY = std::max(maxY, Y);
            Yi += std::log10(Y);
        }
        const float W = Y + X + Z;
        float x = X / W;
        float y = Y / W;
        if (tp->desaturate) {
            float s;
            // Unqualified log10 on a float in the reference: the double
            // ::log10, rounded to float on assignment.
            const float log10_Y = static_cast<float>(::log10(static_cast<double>(Y)));
            if (log10_Y < -2.0f) {
                s = 0.0f;
            } else if (log10_Y < 0.6f) {
                const float sl = (log10_Y + 2.0f) / 2.6f;
                const float sl2 = sl * sl;
                s = (3.0f * sl2) - (2.0f * sl * sl2);
            } else {
                s = 1.0f;
            }
            const float V = 1700.0f * ((vb[0] * p[0]) + (vb[1] * p[1]) + (vb[2] * p[2]));
            x = ((1.0f - s) * tp->xw) + (s * (x + tp->xw - (1.0f / 3.0f)));
            y = ((1.0f - s) * tp->yw) + (s * (y + tp->yw - (1.0f / 3.0f)));
            Y = (0.4468f * (1.0f - s) * V) + (s * Y);
        }
        p[0] = x;
        p[1] = Y;
        p[2] = y;
    }
    if (numpix) Yi /= numpix;
    rgbAvg /= (x_res * y_res);
    Yi = std::pow(10.0f, Yi);

    // Temporal adaptation (rt/camera.cc:1193-1252).
    if (tp->adapt && tp->time_step > 0.0f) {
        float k;
        const float log_la = std::log(Yi);
        if (log_la < -2.0f) {
            k = 1.0f;
        } else if (log_la < 1.0f) {
            k = 1.0f - ((log_la + 2.0f) * (1.0f / 3.0f));
        } else {
            k = 0.0f;
        }
        const float lvl = *adaptation_level, ts = tp->time_step;
        if (lvl > Yi) {
            const float a_cone = Yi + ((lvl - Yi) * std::exp(-(ts / 0.2f)));
            const float a_rod = Yi + ((lvl - Yi) * std::exp(-(ts / 6.0f)));
            *adaptation_level = ((1.0f - k) * a_cone) + (k * a_rod);
        } else {
            const float r = (Yi / lvl);
            const float c1 = lvl * (std::pow(r, 0.9f) - 1.0f);
            const float c2 = Yi * (1.0f - std::pow(r, -0.1f));
            const float a_cone = Yi - (c1 * std::exp(-(ts / 16.0e-3f))) - (c2 * std::exp(-(ts / 200.0e-3f)));
            const float a_rod = Yi - (c1 * std::exp(-(ts / 40.0e-3f))) - (c2 * std::exp(-(ts / 6.0f)));
            *adaptation_level = ((1.0f - k) * a_cone) + (k * a_rod);
        }
        Yi = *adaptation_level;
    } else {
        *adaptation_level = Yi;
    }
    if (yi_out) *yi_out = Yi;

    const float key = tp->key;
    if (mode == 1) {  // global_contrast (Ward)
        const float Yw = 300.0f;
        const float num = 1.219f + std::pow(Yw * 0.5f, 0.4f);
        const float den = 1.219f + std::pow(Yi, 0.4f);
        const float sf = (1.0f / Yw) * std::pow((num / den), 2.5f);
        for (int i = 0; i < n; ++i) img[3 * static_cast<int64_t>(i) + 1] *= sf;
    } else if (mode == 4) {  // global_non_linear
        const float sf = key / Yi;
        maxY *= sf;
        maxY = 1.0f / (maxY * maxY);
        for (int i = 0; i < n; ++i) {
            float &g = img[3 * static_cast<int64_t>(i) + 1];
            g *= sf;
            if (maxY) g *= (1.0f + g * maxY) / (1.0f + g);
        }
    } else if (mode == 5) {  // global_exposure
        const float sf = key / Yi;
        maxY *= sf;
        for (int i = 0; i < n; ++i) {
            float &g = img[3 * static_cast<int64_t>(i) + 1];
            g *= sf;
            if (maxY) g = 1.0f - std::exp(-g);
        }
    } else if (mode == 6) {  // global_avg_luminance
        float sf = 1.0f;
        if (rgbAvg) sf *= key * (0.5f / (rgbAvg * 683.0f));
        for (int i = 0; i < n; ++i) img[3 * static_cast<int64_t>(i) + 1] *= sf;
    } else if (mode == 7) {  // global_max_luminance
        const float sf = key / maxY;
        for (int i = 0; i < n; ++i) img[3 * static_cast<int64_t>(i) + 1] *= sf;
    } else if (mode == 9) {  // global_ferwerda
        const float lda = 800.0f;
        const float log_lda = std::log(lda * 0.5f);
        const float log_la = std::log(Yi);
        float tpc;
        if (log_la <= -2.6f) {
            tpc = -0.72f;
        } else if (log_la <= 1.9f) {
            tpc = std::pow(((0.249f * log_la) + 0.65f), 2.7f) - 0.72f;
        } else {
            tpc = log_la - 1.255f;
        }
        const float mp = (log_lda - 1.255f) / tpc;
        float tsc;
        if (log_la <= -3.94f) {
            tsc = -2.86f;
        } else if (log_la < -1.44f) {
            tsc = std::pow(((0.405f * log_la) + 1.6f), 2.18f) - 2.86f;
        } else {
            tsc = log_la - 0.395f;
        }
        const float ms = (log_lda - 1.255f) / tsc;
        float k;
        if (log_la < -2.0f) {
            k = 1.0f;
        } else if (log_la < 1.0f) {
            k = 1.0f - ((log_la + 2.0f) * (1.0f / 3.0f));
        } else {
            k = 0.0f;
        }
        for (int i = 0; i < n; ++i) {
            float &g = img[3 * static_cast<int64_t>(i) + 1];
            const float ldp = mp * g;
            const float lds = ms * g;
            g = (1.0f / lda) * (ldp + (k * lds));
        }
    } else {  // none
        const float sf = 1.0f / 683.0f;
        for (int i = 0; i < n; ++i) img[3 * static_cast<int64_t>(i) + 1] *= sf;
    }

    if (tp->stretch) {
        float maxY2 = -kMaxDist, minY2 = kMaxDist;
        for (int i = 0; i < n; ++i) {
            maxY2 = std::max(maxY2, img[3 * static_cast<int64_t>(i) + 1]);
            minY2 = std::min(minY2, img[3 * static_cast<int64_t>(i) + 1]);
        }
        if (((minY2 > 0.0f) || (maxY2 < 0.0f)) && (minY2 < maxY2)) {
            const float offset = std::min(-minY2, 0.0f);
            const float scale = 1.0f / (std::min(1.0f, maxY2) - std::max(0.0f, minY2));
            for (int i = 0; i < n; ++i) {
                float &g = img[3 * static_cast<int64_t>(i) + 1];
                g = (g + offset) * scale;
            }
        }
    }
    for (int i = 0; i < n; ++i) yxy_to_rgb(img + 3 * static_cast<int64_t>(i));
    return 0;
}

// camera::gamma_correct (rt/camera.cc:1466-1501), its branch order kept: a
// gamma within 0.001 of 2.2 takes the sRGB curve, any other the plain power.
extern "C" void oracle_gamma(float *img, int64_t npix, float gamma) {
    for (int64_t i = 0; i < npix; ++i) {
        float *p = img + 3 * i;
        float c[3] = {p[0] / 255.0f, p[1] / 255.0f, p[2] / 255.0f};
        for (float &v : c) v = std::min(std::max(v, 0.0f), 1.0f);
        if (std::fabs(gamma - 2.2f) > 0.001f) {
            for (float &v : c) v = std::pow(v, 1.0f / gamma);
        } else if (std::fabs(gamma - 1.0f) > 0.001f) {
            for (float &v : c) v = (v <= 0.0031308f) ? (12.92f * v) : (1.055f * std::pow(v, (1.0f / 2.4f)) - 0.055f);
        }
        p[0] = c[0] * 255.0f;
        p[1] = c[1] * 255.0f;
        p[2] = c[2] * 255.0f;
    }
}

// camera::clip_image_to_rgb / clip_image_to_bgr (rt/camera.h:289-351): box
// down-sampling by (xs, ys) samples, clamped to [0, 255] and truncated.
extern "C" void oracle_clip(const float *img, int32_t x_res, int32_t y_res, int32_t xs, int32_t ys, int32_t bgr,
                            uint8_t *out) {
    const int ox = x_res / xs, oy = y_res / ys;
    const int samples = xs * ys;
    const float inv = 1.0f / (xs * ys);
    for (int x = 0; x < ox; ++x) {
        for (int y = 0; y < oy; ++y) {
            float a[3] = {0.0f, 0.0f, 0.0f};
            const int addr = (x * xs) + (y * ys * x_res);
            for (int ax = 0; ax < xs; ++ax)
                for (int ay = 0; ay < ys; ++ay)
                    for (int k = 0; k < 3; ++k) a[k] += img[3 * static_cast<int64_t>(addr + (ay * x_res) + ax) + k];
            uint8_t *o = out + 3 * static_cast<int64_t>(x + y * ox);
            if (bgr) {
                for (int k = 0; k < 3; ++k)
                    o[k] = static_cast<uint8_t>(std::min(255.0f, std::max(0.0f, (a[2 - k] * inv))));
            } else {
                for (int k = 0; k < 3; ++k)
                    o[k] = static_cast<uint8_t>(std::min(255.0f, std::max(0.0f, (a[k] / samples))));
            }
        }
    }
}
