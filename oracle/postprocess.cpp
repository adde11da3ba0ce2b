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
#include <vector>

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

// The circular convolution the reference performs with FFTW
// (rt/camera.cc:1697-1734): both channels of a W x H x D grid (channel c at
// c * W * H * D, cell (x, y, z) at x + W * (y + H * z)) convolved with the
// periodic kernel exp(-0.5 * ((X^2 + Y^2) / ss^2 + Z^2 / sr^2)), X the offset
// wrapped to (-W/2, W/2] as the reference builds it. An r2c / c2r DFT pair
// with the 1 / (W H D) scale is this sum; it is evaluated directly here in
// double (the reference's single-precision FFT rounds differently).
void gauss_circular3d(float *grid, int W, int H, int D, float ss, float sr) {
    const int64_t N = static_cast<int64_t>(W) * H * D;
    std::vector<double> kx(W), ky(H), kz(D);
    auto wrap = [](int i, int n) { return static_cast<double>(i - ((i > (n >> 1)) ? n : 0)); };
    for (int i = 0; i < W; ++i) kx[i] = wrap(i, W) * wrap(i, W) / (double(ss) * ss);
    for (int i = 0; i < H; ++i) ky[i] = wrap(i, H) * wrap(i, H) / (double(ss) * ss);
    for (int i = 0; i < D; ++i) kz[i] = wrap(i, D) * wrap(i, D) / (double(sr) * sr);
    std::vector<double> in(N), out(N);
    for (int c = 0; c < 2; ++c) {
        float *g = grid + c * N;
        for (int64_t i = 0; i < N; ++i) in[i] = g[i];
        // Separable passes: the kernel is a product of its axis factors.
        auto pass = [&](int axis) {
            const int n = axis == 0 ? W : (axis == 1 ? H : D);
            const std::vector<double> &k = axis == 0 ? kx : (axis == 1 ? ky : kz);
            std::vector<double> f(n);
            for (int i = 0; i < n; ++i) f[i] = std::exp(-0.5 * k[i]);
            for (int z = 0; z < D; ++z)
                for (int y = 0; y < H; ++y)
                    for (int x = 0; x < W; ++x) {
                        const int o = axis == 0 ? x : (axis == 1 ? y : z);
                        double acc = 0.0;
                        for (int j = 0; j < n; ++j) {
                            const int src = (o - j + n) % n;
                            const int xs = axis == 0 ? src : x, ys = axis == 1 ? src : y, zs = axis == 2 ? src : z;
                            acc += in[xs + static_cast<int64_t>(W) * (ys + static_cast<int64_t>(H) * zs)] * f[j];
                        }
                        out[x + static_cast<int64_t>(W) * (y + static_cast<int64_t>(H) * z)] = acc;
                    }
            in.swap(out);
        };
        pass(0);
        pass(1);
        pass(2);
        for (int64_t i = 0; i < N; ++i) g[i] = static_cast<float>(in[i]);
    }
}

// camera::bilateral_filter_tone_map (rt/camera.cc:1632-1775) on the Y channel
// of an (x, Y, y) image, with the arguments global_bilateral_filter passes
// (rt/camera.cc:1354-1358). Pixels with Y <= 3.18e-7 have no log intensity in
// the reference (its array entry is never written; a fresh allocation of a
// frame's size reads zero), so they take 0; their grid depth index is
// clamped to the grid (the reference would index outside it when the
// smallest log intensity is positive).
void bilateral_tone_map(float *img, int x_res, int y_res) {
    const int n = x_res * y_res;
    const float s_s = 0.02f * static_cast<float>(std::min(x_res, y_res));
    const float r_s = 0.4f, r_sa = 0.4f, s_sa = s_s;
    std::vector<float> logi(n, 0.0f);
    float input_max = -kMaxDist, input_min = kMaxDist;
    for (int i = 0; i < n; ++i) {
        const float g = img[3 * static_cast<int64_t>(i) + 1];
        if (g > 3.18e-7f) {
            logi[i] = static_cast<float>(::log10(static_cast<double>(g)));
            input_max = std::max(input_max, logi[i]);
            input_min = std::min(input_min, logi[i]);
        }
    }
    const float input_delta = input_max - input_min;
    const float sigma_r = r_s / r_sa, sigma_s = s_s / s_sa;
    const int pxy = static_cast<int>(2.0f * sigma_s) + 1, pz = static_cast<int>(2.0f * sigma_r) + 1;
    const int W = static_cast<int>(static_cast<float>(x_res - 1) / s_sa) + 1 + (pxy << 1);
    const int H = static_cast<int>(static_cast<float>(y_res - 1) / s_sa) + 1 + (pxy << 1);
    const int D = static_cast<int>(input_delta / r_sa) + 1 + (pz << 1);
    const int64_t N = static_cast<int64_t>(W) * H * D;
    std::vector<float> grid(2 * N, 0.0f);
    for (int x = 0; x < x_res; ++x)
        for (int y = 0; y < y_res; ++y) {
            const float l = logi[x + x_res * y];
            const int sx = static_cast<int>(static_cast<float>(x) / s_sa + 0.5f) + pxy;
            const int sy = static_cast<int>(static_cast<float>(y) / s_sa + 0.5f) + pxy;
            const int sz = std::max(0, std::min(D - 1, static_cast<int>((l - input_min) / r_sa + 0.5f) + pz));
            const int64_t c = sx + static_cast<int64_t>(W) * (sy + static_cast<int64_t>(H) * sz);
            grid[c] += 1.0f;
            grid[c + N] += l;
        }
    gauss_circular3d(grid.data(), W, H, D, sigma_s, sigma_r);
    std::vector<float> result(n);
    float max_value = -kMaxDist, min_value = kMaxDist;
    for (int x = 0; x < x_res; ++x)
        for (int y = 0; y < y_res; ++y) {
            const float z = logi[x + x_res * y] - input_min;
            const float xa = static_cast<float>(x) / s_sa + pxy;
            const int xi = std::max(0, std::min(static_cast<int>(xa), W - 1)), xxi = std::max(0, std::min(xi + 1, W - 1));
            const float ya = static_cast<float>(y) / s_sa + pxy;
            const int yi = std::max(0, std::min(static_cast<int>(ya), H - 1)), yyi = std::max(0, std::min(yi + 1, H - 1));
            const float za = z / r_sa + pz;
            const int zi = std::max(0, std::min(static_cast<int>(za), D - 1)), zzi = std::max(0, std::min(zi + 1, D - 1));
            const float ax = xa - xi, ay = ya - yi, az = za - zi;
            auto at = [&](int c, int i, int j, int k) {
                return grid[c * N + i + static_cast<int64_t>(W) * (j + static_cast<int64_t>(H) * k)];
            };
            auto tri = [&](int c) {
                return (1.0f - ax) * (1.0f - ay) * (1.0f - az) * at(c, xi, yi, zi) +
                       ax * (1.0f - ay) * (1.0f - az) * at(c, xxi, yi, zi) +
                       (1.0f - ax) * ay * (1.0f - az) * at(c, xi, yyi, zi) +
                       ax * ay * (1.0f - az) * at(c, xxi, yyi, zi) +
                       (1.0f - ax) * (1.0f - ay) * az * at(c, xi, yi, zzi) +
                       ax * (1.0f - ay) * az * at(c, xxi, yi, zzi) +
                       (1.0f - ax) * ay * az * at(c, xi, yyi, zzi) + ax * ay * az * at(c, xxi, yyi, zzi);
            };
            const float r = tri(1) / tri(0);
            result[x + x_res * y] = r;
            max_value = std::max(max_value, r);
            min_value = std::min(min_value, r);
        }
    const float gamma = std::log10(300.0f) / (max_value - min_value);
    for (int i = 0; i < n; ++i) {
        const float v = std::pow(10.0f, (result[i] * gamma + (logi[i] - result[i])));
        img[3 * static_cast<int64_t>(i) + 1] = v * (1.0f / 255.0f);
    }
}

// camera::just_noticable_difference (rt/camera.cc:1778-1806); log10 of a
// float unqualified in the reference is the double one.
float jnd(float La) {
    const float lLa = static_cast<float>(::log10(static_cast<double>(La)));
    float lLt;
    if (lLa < -3.94f) {
        lLt = -2.86f;
    } else if (lLa < -1.44f) {
        lLt = std::pow(0.405f * lLa + 1.6f, 2.18f) - 2.86f;
    } else if (lLa < -0.0184f) {
        lLt = lLa - 0.395f;
    } else if (lLa < 1.9f) {
        lLt = std::pow(0.249f * lLa + 0.65f, 2.7f) - 0.72f;
    } else {
        lLt = lLa - 1.255f;
    }
    return std::pow(10.0f, lLt);
}

// camera::histogram_tone_map (rt/camera.cc:1843-2062), Ward Larson's
// histogram adjustment, on the Y channel of an (x, Y, y) image: the frame
// averaged to one sample per degree of view, a 100-bin histogram of log
// luminance whose bins are cut to a ceiling (linear, or scaled by the
// just-noticeable-difference ratio when human is set) until fewer than 2.5%
// of the samples are trimmed, and each pixel mapped through the cumulative
// distribution to log display luminance in [log 1, log 300].
void histogram_tone_map(float *img, int x_res, int y_res, bool human, float x_angle, float y_angle, float lw_avg) {
    auto G = [&](int i) -> float & { return img[3 * static_cast<int64_t>(i) + 1]; };
    const float PI = 3.1415926535897932384626433f;
    const int sxr = static_cast<int>(x_angle * (360.0f / (2.0f * PI)));
    const int syr = static_cast<int>(y_angle * (360.0f / (2.0f * PI)));
    std::vector<float> sl(static_cast<size_t>(std::max(0, sxr * syr)), 0.0f);
    const float sxi = static_cast<float>(x_res) / static_cast<float>(sxr);
    const float syi = static_cast<float>(y_res) / static_cast<float>(syr);
    float y0 = 0.0f, y1 = syi;
    for (int j = 0; j < syr; ++j) {
        float x0 = 0.0f, x1 = sxi;
        for (int i = 0; i < sxr; ++i) {
            float &acc = sl[j * sxr + i];
            // (a single sample across would reach one past the image: stop inside)
            const int ye = std::min(static_cast<int>(y1), y_res - 1), xe = std::min(static_cast<int>(x1), x_res - 1);
            for (int y = static_cast<int>(y0); y <= ye; ++y)
                for (int x = static_cast<int>(x0); x <= xe; ++x) acc += G(y * x_res + x);
            acc /= ((x1 - x0) + 1) * ((y1 - y0) + 1);
            x0 = x1;
            x1 = std::min((x_res - 1.0f), (x1 + sxi));
        }
        y0 = y1;
        y1 = std::min((y_res - 1.0f), (y1 + syi));
    }
    const int bins = 100;
    int pix_gt0 = 0;
    float lw_min = kMaxDist, lw_max = 0.0f;
    for (const float v : sl)
        if (v > 0.0f) {
            lw_min = std::min(lw_min, v);
            lw_max = std::max(lw_max, v);
            ++pix_gt0;
        }
    if (lw_min <= 0.0f) return;  // black image (never reached: lw_min starts at MAX_DIST)
    lw_min = std::max(lw_min, 1.0e-4f);
    const float bw_min = std::log(lw_min);
    float bw_max = std::log(lw_max);
    if (bw_max == bw_min) bw_max += 0.0001f * std::fabs(bw_max);
    const float ld_max = 300.0f, ld_min = 1.0f;
    const float bd_max = std::log(ld_max), bd_min = std::log(ld_min);
    const int n = x_res * y_res;
    if ((bd_max - bd_min) > (bw_max - bw_min)) {
        const float sf = (bd_max - bd_min) / (bw_max - bw_min);
        for (int i = 0; i < n; ++i) G(i) = G(i) * sf + bd_min;
        return;
    }
    int his[bins + 1] = {0};
    float cdf[bins + 1] = {0.0f};
    for (const float v : sl)
        if (v > 0.0f) {
            int b = static_cast<int>((std::log(v) - bw_min) / (bw_max - bw_min) * bins) + 1;
            b = std::min(std::max(b, 0), bins);
            his[b]++;
        }
    float sum = 0.0f;
    for (int i = 1; i <= bins; ++i) {
        sum += his[i];
        cdf[i] = sum / static_cast<float>(pix_gt0);
    }
    const float tolerance = 0.025f * static_cast<float>(pix_gt0);
    const float bin_step_size = (bw_max - bw_min) / bins;
    const float lw_step = 1.0f / static_cast<float>(bins) * (bw_max - bw_min);
    int trimmings;
    do {
        trimmings = 0;
        if (pix_gt0 < tolerance) {
            const float num = 1.219f + std::pow(ld_max * 0.5f, 0.4f);
            const float den = 1.219f + std::pow(lw_avg, 0.4f);
            const float sf = (1.0f / ld_max) * std::pow((num / den), 2.5f);
            for (int i = 0; i < n; ++i) G(i) *= sf;
            return;
        }
        for (int i = 1; i <= bins; ++i) {
            float ceiling = (pix_gt0 * bin_step_size) / (bd_max - bd_min);
            if (human) {
                const float bde = bd_min + (bd_max - bd_min) * cdf[i];
                const float ld = std::exp(bde);
                const float bw = i * lw_step + bw_min;
                const float lw = std::exp(bw);
                ceiling *= (jnd(ld) * lw) / (jnd(lw) * ld);
            }
            if (his[i] > ceiling) {
                trimmings += (his[i] - static_cast<int>(ceiling));
                his[i] = static_cast<int>(ceiling);
            }
        }
        pix_gt0 = 0;
        for (int i = 1; i <= bins; ++i) pix_gt0 += his[i];
        float s2 = 0.0f;
        for (int i = 1; i <= bins; ++i) {
            s2 += his[i];
            cdf[i] = s2 / static_cast<float>(pix_gt0);
        }
    } while (trimmings > tolerance);
    for (int i = 0; i < n; ++i) {
        float &g = G(i);
        if (g > 0.0f) {
            const float bw = std::log(g);
            float bf = (bw - bw_min) / (bw_max - bw_min) * bins;
            bf = std::min(std::max(bf, 0.0f), static_cast<float>(bins));
            const int b0 = static_cast<int>(bf);
            const int b1 = std::min((b0 + 1), bins);
            const float s = bf - b0;
            const float pbw = (1 - s) * cdf[b0] + s * cdf[b1];
            const float bde = bd_min + (bd_max - bd_min) * pbw;
            g = std::exp(bde);
            g /= ld_max;
        }
    }
}

}  // namespace

extern "C" void oracle_gauss_circular3d(float *grid, int32_t W, int32_t H, int32_t D, float ss, float sr) {
    gauss_circular3d(grid, W, H, D, ss, sr);
}

extern "C" int32_t oracle_tone_map(float *img, int32_t x_res, int32_t y_res, const oracle_tone_params *tp,
                                   float *adaptation_level, float *yi_out) {
    static const float xb[3] = {0.4124f, 0.3576f, 0.1805f};
    static const float yb[3] = {0.2126f, 0.7151f, 0.0721f};
    static const float zb[3] = {0.0193f, 0.1192f, 0.9505f};
    static const float vb[3] = {0.01477f, 0.497f, 0.631f};
    const int mode = tp->mode;
    if (mode < 0 || mode > 9) return -1;
    if (tp->glare && !(tp->x_angle > 0.0f)) return -3;
    if ((mode == 2 || mode == 3) && !(tp->x_angle * (360.0f / (2.0f * 3.1415926535897932384626433f)) >= 1.0f &&
                                      tp->y_angle * (360.0f / (2.0f * 3.1415926535897932384626433f)) >= 1.0f))
        return -2;
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
    } else if (mode == 2 || mode == 3) {  // local_histogram, local_human_histogram
        histogram_tone_map(img, x_res, y_res, mode == 3, tp->x_angle, tp->y_angle, *adaptation_level);
    } else if (mode == 8) {  // global_bilateral_filter
        bilateral_tone_map(img, x_res, y_res);
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
    // rt/camera.cc:1455-1458: the glare filter at the adapted luminance.
    if (tp->glare) oracle_glare(img, x_res, y_res, Yi, tp->x_angle, nullptr);
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
