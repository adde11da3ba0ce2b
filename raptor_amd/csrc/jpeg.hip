// JPEG writer: camera::write_jpeg_file (rt/camera.cc:2305-2366), which hands
// the clipped rgb bytes, last row first, to libjpeg with jpeg_set_defaults
// and jpeg_set_quality(q, force_baseline). Those defaults are the JPEG
// baseline of ITU-T T.81 as the IJG library sets it up: JFIF YCbCr, chroma
// subsampled 2 x 2 (4:2:0), the Annex K quantisation tables scaled by the IJG
// quality rule, the Annex K Huffman tables, no restart markers.
//
// The per-pixel and per-block work runs on the device, one workgroup per
// 16 x 16 minimum coded unit (MCU): the pixels (edge-replicated past the
// image, rows taken last-first), YCbCr, the chroma averaged over 2 x 2, the
// level shift, the 8 x 8 forward DCT of the four luminance and two chroma
// blocks as two passes through LDS, and the quantisation, written as
// zig-zag-ordered coefficients. The host then entropy-codes the MCUs in
// order (DC differences, AC run lengths, Huffman codes, 0xFF stuffing) and
// writes the markers. The DCT is the float transform of T.81 A.3.3; libjpeg's
// default integer DCT rounds a few coefficients differently, so files are
// alike in content, not byte for byte.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "rt_internal.hpp"

namespace rt {

namespace {

// T.81 Annex K.1 tables, in natural (row-major) order.
const int kLumaQ[64] = {16, 11, 10, 16, 24,  40,  51,  61,  12, 12, 14, 19, 26,  58,  60,  55,
                        14, 13, 16, 24, 40,  57,  69,  56,  14, 17, 22, 29, 51,  87,  80,  62,
                        18, 22, 37, 56, 68,  109, 103, 77,  24, 35, 55, 64, 81,  104, 113, 92,
                        49, 64, 78, 87, 103, 121, 120, 101, 72, 92, 95, 98, 112, 100, 103, 99};
const int kChromaQ[64] = {17, 18, 24, 47, 99, 99, 99, 99, 18, 21, 26, 66, 99, 99, 99, 99, 24, 26, 56, 99, 99, 99,
                          99, 99, 47, 66, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99,
                          99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99};
// Zig-zag scan: kZigzag[k] is the natural index of the k-th coefficient.
const int kZigzag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,  12, 19, 26, 33, 40, 48,
                         41, 34, 27, 20, 13, 6,  7,  14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23,
                         30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};
// Annex K.3 Huffman tables: code counts per length 1..16, then the symbols.
const uint8_t kDcLumaBits[16] = {0, 1, 5, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0};
const uint8_t kDcChromaBits[16] = {0, 3, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0};
const uint8_t kDcVals[12] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11};
const uint8_t kAcLumaBits[16] = {0, 2, 1, 3, 3, 2, 4, 3, 5, 5, 4, 4, 0, 0, 1, 0x7d};
const uint8_t kAcLumaVals[162] = {
    0x01, 0x02, 0x03, 0x00, 0x04, 0x11, 0x05, 0x12, 0x21, 0x31, 0x41, 0x06, 0x13, 0x51, 0x61, 0x07, 0x22, 0x71, 0x14,
    0x32, 0x81, 0x91, 0xa1, 0x08, 0x23, 0x42, 0xb1, 0xc1, 0x15, 0x52, 0xd1, 0xf0, 0x24, 0x33, 0x62, 0x72, 0x82, 0x09,
    0x0a, 0x16, 0x17, 0x18, 0x19, 0x1a, 0x25, 0x26, 0x27, 0x28, 0x29, 0x2a, 0x34, 0x35, 0x36, 0x37, 0x38, 0x39, 0x3a,
    0x43, 0x44, 0x45, 0x46, 0x47, 0x48, 0x49, 0x4a, 0x53, 0x54, 0x55, 0x56, 0x57, 0x58, 0x59, 0x5a, 0x63, 0x64, 0x65,
    0x66, 0x67, 0x68, 0x69, 0x6a, 0x73, 0x74, 0x75, 0x76, 0x77, 0x78, 0x79, 0x7a, 0x83, 0x84, 0x85, 0x86, 0x87, 0x88,
    0x89, 0x8a, 0x92, 0x93, 0x94, 0x95, 0x96, 0x97, 0x98, 0x99, 0x9a, 0xa2, 0xa3, 0xa4, 0xa5, 0xa6, 0xa7, 0xa8, 0xa9,
    0xaa, 0xb2, 0xb3, 0xb4, 0xb5, 0xb6, 0xb7, 0xb8, 0xb9, 0xba, 0xc2, 0xc3, 0xc4, 0xc5, 0xc6, 0xc7, 0xc8, 0xc9, 0xca,
    0xd2, 0xd3, 0xd4, 0xd5, 0xd6, 0xd7, 0xd8, 0xd9, 0xda, 0xe1, 0xe2, 0xe3, 0xe4, 0xe5, 0xe6, 0xe7, 0xe8, 0xe9, 0xea,
    0xf1, 0xf2, 0xf3, 0xf4, 0xf5, 0xf6, 0xf7, 0xf8, 0xf9, 0xfa};
const uint8_t kAcChromaBits[16] = {0, 2, 1, 2, 4, 4, 3, 4, 7, 5, 4, 4, 0, 1, 2, 0x77};
const uint8_t kAcChromaVals[162] = {
    0x00, 0x01, 0x02, 0x03, 0x11, 0x04, 0x05, 0x21, 0x31, 0x06, 0x12, 0x41, 0x51, 0x07, 0x61, 0x71, 0x13, 0x22, 0x32,
    0x81, 0x08, 0x14, 0x42, 0x91, 0xa1, 0xb1, 0xc1, 0x09, 0x23, 0x33, 0x52, 0xf0, 0x15, 0x62, 0x72, 0xd1, 0x0a, 0x16,
    0x24, 0x34, 0xe1, 0x25, 0xf1, 0x17, 0x18, 0x19, 0x1a, 0x26, 0x27, 0x28, 0x29, 0x2a, 0x35, 0x36, 0x37, 0x38, 0x39,
    0x3a, 0x43, 0x44, 0x45, 0x46, 0x47, 0x48, 0x49, 0x4a, 0x53, 0x54, 0x55, 0x56, 0x57, 0x58, 0x59, 0x5a, 0x63, 0x64,
    0x65, 0x66, 0x67, 0x68, 0x69, 0x6a, 0x73, 0x74, 0x75, 0x76, 0x77, 0x78, 0x79, 0x7a, 0x82, 0x83, 0x84, 0x85, 0x86,
    0x87, 0x88, 0x89, 0x8a, 0x92, 0x93, 0x94, 0x95, 0x96, 0x97, 0x98, 0x99, 0x9a, 0xa2, 0xa3, 0xa4, 0xa5, 0xa6, 0xa7,
    0xa8, 0xa9, 0xaa, 0xb2, 0xb3, 0xb4, 0xb5, 0xb6, 0xb7, 0xb8, 0xb9, 0xba, 0xc2, 0xc3, 0xc4, 0xc5, 0xc6, 0xc7, 0xc8,
    0xc9, 0xca, 0xd2, 0xd3, 0xd4, 0xd5, 0xd6, 0xd7, 0xd8, 0xd9, 0xda, 0xe2, 0xe3, 0xe4, 0xe5, 0xe6, 0xe7, 0xe8, 0xe9,
    0xea, 0xf2, 0xf3, 0xf4, 0xf5, 0xf6, 0xf7, 0xf8, 0xf9, 0xfa};

// jpeg_set_quality with force_baseline: the IJG quality-to-scale rule, each
// table entry (q * scale + 50) / 100 clamped to [1, 255].
void scaled_table(const int *base, int quality, int *out) {
    quality = std::max(1, std::min(100, quality));
    const int scale = quality < 50 ? 5000 / quality : 200 - quality * 2;
    for (int i = 0; i < 64; ++i) {
        long v = (static_cast<long>(base[i]) * scale + 50L) / 100L;
        out[i] = static_cast<int>(std::max(1L, std::min(255L, v)));
    }
}

struct JpegArgs {
    const uint8_t *rgb;  // x_res * y_res * 3, camera rows (row 0 is written last)
    int x_res, y_res;
    int mcu_x;          // MCUs per row
    float qy[64], qc[64];  // divisors, natural order
    int zig[64];
    float cosk[64];      // cosk[u * 8 + x] = C(u) / 2 * cos((2x + 1) u pi / 16)
    int16_t *coef;       // per MCU: 6 blocks x 64 zig-zag coefficients (Y0 Y1 Y2 Y3 Cb Cr)
};

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// One workgroup of 256 threads per MCU: thread t owns pixel (t % 16, t / 16)
// for the colour conversion, then one coefficient of one block per pass.
__global__ void __launch_bounds__(256) k_jpeg_mcu(JpegArgs a) {
    __shared__ float ycc[3][16][16];
    __shared__ float blk[6][8][8];
    __shared__ float tmp[6][8][8];
    const int t = threadIdx.x, px = t % 16, py = t / 16;
    const int mx = static_cast<int>(blockIdx.x) % a.mcu_x, my = static_cast<int>(blockIdx.x) / a.mcu_x;
    // File row r is camera row y_res - 1 - r; past the image the last
    // column / row repeats (as libjpeg pads an MCU).
    const int fx = clampi(mx * 16 + px, 0, a.x_res - 1);
    const int fr = clampi(my * 16 + py, 0, a.y_res - 1);
    const uint8_t *p = a.rgb + 3 * (static_cast<int64_t>(a.y_res - 1 - fr) * a.x_res + fx);
    const float r = p[0], g = p[1], b = p[2];
    ycc[0][py][px] = 0.299f * r + 0.587f * g + 0.114f * b - 128.0f;
    ycc[1][py][px] = -0.168736f * r - 0.331264f * g + 0.5f * b;  // + 128 - 128
    ycc[2][py][px] = 0.5f * r - 0.418688f * g - 0.081312f * b;
    __syncthreads();
    // Six 8 x 8 blocks: four luminance quadrants, then the 2 x 2 chroma means.
    for (int k = t; k < 6 * 64; k += 256) {
        const int bi = k / 64, e = k % 64, x = e % 8, y = e / 8;
        float v;
        if (bi < 4) {
            v = ycc[0][(bi / 2) * 8 + y][(bi % 2) * 8 + x];
        } else {
            const int c = bi - 3;
            v = 0.25f * ((ycc[c][2 * y][2 * x] + ycc[c][2 * y][2 * x + 1]) +
                         (ycc[c][2 * y + 1][2 * x] + ycc[c][2 * y + 1][2 * x + 1]));
        }
        blk[bi][y][x] = v;
    }
    __syncthreads();
    // Rows: tmp[y][u] = sum_x cosk[u][x] blk[y][x]; columns: F[v][u] = sum_y cosk[v][y] tmp[y][u].
    for (int k = t; k < 6 * 64; k += 256) {
        const int bi = k / 64, e = k % 64, u = e % 8, y = e / 8;
        float s = 0.0f;
#pragma unroll
        for (int x = 0; x < 8; ++x) s += a.cosk[u * 8 + x] * blk[bi][y][x];
        tmp[bi][y][u] = s;
    }
    __syncthreads();
    int16_t *out = a.coef + static_cast<int64_t>(blockIdx.x) * 6 * 64;
    for (int k = t; k < 6 * 64; k += 256) {
        const int bi = k / 64, zz = k % 64, nat = a.zig[zz], u = nat % 8, v = nat / 8;
        float s = 0.0f;
#pragma unroll
        for (int y = 0; y < 8; ++y) s += a.cosk[v * 8 + y] * tmp[bi][y][u];
        const float q = bi < 4 ? a.qy[nat] : a.qc[nat];
        out[bi * 64 + zz] = static_cast<int16_t>(rintf(s / q));
    }
}

struct HuffTable {
    uint16_t code[256];
    uint8_t size[256];
};

// Annex C: code lengths and canonical codes from the counts per length.
HuffTable make_table(const uint8_t *bits, const uint8_t *vals) {
    HuffTable h{};
    int k = 0;
    uint16_t code = 0;
    for (int len = 1; len <= 16; ++len) {
        for (int i = 0; i < bits[len - 1]; ++i, ++k) {
            h.code[vals[k]] = code++;
            h.size[vals[k]] = static_cast<uint8_t>(len);
        }
        code <<= 1;
    }
    return h;
}

struct BitWriter {
    std::vector<uint8_t> &out;
    uint32_t acc = 0;
    int n = 0;
    void put(uint32_t bits, int len) {
        for (int i = len - 1; i >= 0; --i) {
            acc = (acc << 1) | ((bits >> i) & 1u);
            if (++n == 8) {
                out.push_back(static_cast<uint8_t>(acc));
                if (acc == 0xFF) out.push_back(0x00);  // byte stuffing
                acc = 0;
                n = 0;
            }
        }
    }
    void flush() {  // pad with one bits
        while (n) put(1, 1);
    }
};

// The magnitude category (bit length) of v and its appended bits (F.1.2.1).
inline int category(int v) {
    int a = v < 0 ? -v : v, c = 0;
    while (a) {
        ++c;
        a >>= 1;
    }
    return c;
}
inline uint32_t extra_bits(int v, int c) { return static_cast<uint32_t>(v < 0 ? v + (1 << c) - 1 : v); }

void encode_block(BitWriter &w, const int16_t *z, int &pred, const HuffTable &dc, const HuffTable &ac) {
    const int diff = z[0] - pred;
    pred = z[0];
    const int c = category(diff);
    w.put(dc.code[c], dc.size[c]);
    if (c) w.put(extra_bits(diff, c), c);
    int run = 0;
    for (int k = 1; k < 64; ++k) {
        const int v = z[k];
        if (v == 0) {
            ++run;
            continue;
        }
        while (run > 15) {  // ZRL
            w.put(ac.code[0xF0], ac.size[0xF0]);
            run -= 16;
        }
        const int s = category(v);
        const int sym = (run << 4) | s;
        w.put(ac.code[sym], ac.size[sym]);
        w.put(extra_bits(v, s), s);
        run = 0;
    }
    if (run) w.put(ac.code[0x00], ac.size[0x00]);  // EOB
}

void put16(std::vector<uint8_t> &o, int v) {
    o.push_back(static_cast<uint8_t>(v >> 8));
    o.push_back(static_cast<uint8_t>(v & 0xFF));
}

void put_dht(std::vector<uint8_t> &o, int cls_id, const uint8_t *bits, const uint8_t *vals, int nvals) {
    o.push_back(0xFF);
    o.push_back(0xC4);
    put16(o, 2 + 1 + 16 + nvals);
    o.push_back(static_cast<uint8_t>(cls_id));
    for (int i = 0; i < 16; ++i) o.push_back(bits[i]);
    for (int i = 0; i < nvals; ++i) o.push_back(vals[i]);
}

}  // namespace

int32_t write_jpeg(rt_ctx *ctx, const char *path, const uint8_t *rgb, int32_t x_res, int32_t y_res, int32_t quality,
                   uint32_t flags, float *ms) {
    hipStream_t s = ctx->stream;
    const int mcu_x = (x_res + 15) / 16, mcu_y = (y_res + 15) / 16;
    const int64_t nmcu = static_cast<int64_t>(mcu_x) * mcu_y;
    const int64_t nbytes = static_cast<int64_t>(x_res) * y_res * 3;
    JpegArgs a{};
    a.x_res = x_res;
    a.y_res = y_res;
    a.mcu_x = mcu_x;
    int qy[64], qc[64];
    scaled_table(kLumaQ, quality, qy);
    scaled_table(kChromaQ, quality, qc);
    for (int i = 0; i < 64; ++i) {
        a.qy[i] = static_cast<float>(qy[i]);
        a.qc[i] = static_cast<float>(qc[i]);
        a.zig[i] = kZigzag[i];
    }
    for (int u = 0; u < 8; ++u)
        for (int x = 0; x < 8; ++x)
            a.cosk[u * 8 + x] = static_cast<float>((u == 0 ? std::sqrt(0.5) : 1.0) * 0.5 *
                                                   std::cos((2.0 * x + 1.0) * u * M_PI / 16.0));
    BuildArena &ar = ctx->arena;
    RT_HIP_CHECK(ctx, ar.begin());
    void *v = nullptr;
    if (flags & RT_DEVICE_PTRS) {
        a.rgb = rgb;
    } else {
        RT_HIP_CHECK(ctx, ar.take(&v, static_cast<size_t>(nbytes)));
        RT_HIP_CHECK(ctx, hipMemcpyAsync(v, rgb, nbytes, hipMemcpyHostToDevice, s));
        a.rgb = static_cast<const uint8_t *>(v);
    }
    RT_HIP_CHECK(ctx, ar.take(&v, static_cast<size_t>(nmcu) * 6 * 64 * sizeof(int16_t)));
    a.coef = static_cast<int16_t *>(v);
    if (ms) RT_HIP_CHECK(ctx, hipEventRecord(ctx->ev[0], s));
    hipLaunchKernelGGL(k_jpeg_mcu, dim3(static_cast<unsigned>(nmcu)), dim3(256), 0, s, a);
    RT_HIP_CHECK(ctx, hipGetLastError());
    if (ms) RT_HIP_CHECK(ctx, hipEventRecord(ctx->ev[1], s));
    std::vector<int16_t> coef(static_cast<size_t>(nmcu) * 6 * 64);
    RT_HIP_CHECK(ctx, hipMemcpyAsync(coef.data(), a.coef, coef.size() * sizeof(int16_t), hipMemcpyDeviceToHost, s));
    RT_HIP_CHECK(ctx, hipStreamSynchronize(s));
    if (ms) RT_HIP_CHECK(ctx, hipEventElapsedTime(ms, ctx->ev[0], ctx->ev[1]));

    std::vector<uint8_t> o;
    o.reserve(static_cast<size_t>(nbytes / 4 + 1024));
    const uint8_t jfif[] = {0xFF, 0xD8, 0xFF, 0xE0, 0x00, 0x10, 'J', 'F', 'I', 'F', 0x00, 0x01, 0x01,
                            0x00, 0x00, 0x01, 0x00, 0x01, 0x00, 0x00};
    o.insert(o.end(), jfif, jfif + sizeof(jfif));
    for (int t = 0; t < 2; ++t) {  // DQT, tables in zig-zag order
        o.push_back(0xFF);
        o.push_back(0xDB);
        put16(o, 67);
        o.push_back(static_cast<uint8_t>(t));
        for (int k = 0; k < 64; ++k) o.push_back(static_cast<uint8_t>((t == 0 ? qy : qc)[kZigzag[k]]));
    }
    o.push_back(0xFF);  // SOF0: 8-bit, three components, Y 2x2, Cb and Cr 1x1
    o.push_back(0xC0);
    put16(o, 17);
    o.push_back(8);
    put16(o, y_res);
    put16(o, x_res);
    o.push_back(3);
    const uint8_t comps[9] = {1, 0x22, 0, 2, 0x11, 1, 3, 0x11, 1};
    o.insert(o.end(), comps, comps + 9);
    put_dht(o, 0x00, kDcLumaBits, kDcVals, 12);
    put_dht(o, 0x10, kAcLumaBits, kAcLumaVals, 162);
    put_dht(o, 0x01, kDcChromaBits, kDcVals, 12);
    put_dht(o, 0x11, kAcChromaBits, kAcChromaVals, 162);
    const uint8_t sos[] = {0xFF, 0xDA, 0x00, 0x0C, 3, 1, 0x00, 2, 0x11, 3, 0x11, 0, 63, 0};
    o.insert(o.end(), sos, sos + sizeof(sos));
    const HuffTable dcl = make_table(kDcLumaBits, kDcVals), acl = make_table(kAcLumaBits, kAcLumaVals);
    const HuffTable dcc = make_table(kDcChromaBits, kDcVals), acc = make_table(kAcChromaBits, kAcChromaVals);
    BitWriter w{o};
    int py = 0, pcb = 0, pcr = 0;
    for (int64_t m = 0; m < nmcu; ++m) {
        const int16_t *z = coef.data() + m * 6 * 64;
        for (int b = 0; b < 4; ++b) encode_block(w, z + b * 64, py, dcl, acl);
        encode_block(w, z + 4 * 64, pcb, dcc, acc);
        encode_block(w, z + 5 * 64, pcr, dcc, acc);
    }
    w.flush();
    o.push_back(0xFF);
    o.push_back(0xD9);
    FILE *f = std::fopen(path, "wb");
    if (!f) return fail(ctx, RT_ERR_INVALID, std::string("rt_write_jpeg: cannot open ") + path);
    const size_t wr = std::fwrite(o.data(), 1, o.size(), f);
    std::fclose(f);
    if (wr != o.size()) return fail(ctx, RT_ERR_INVALID, std::string("rt_write_jpeg: short write to ") + path);
    return RT_OK;
}

}  // namespace rt

extern "C" int32_t rt_write_jpeg(rt_ctx *ctx, const char *path, const uint8_t *rgb, int32_t x_res, int32_t y_res,
                                 int32_t quality, uint32_t flags, float *ms) {
    if (!ctx || !path || !rgb || x_res <= 0 || y_res <= 0 || x_res > 65535 || y_res > 65535) return RT_ERR_INVALID;
    if (quality < 1 || quality > 100) return rt::fail(ctx, RT_ERR_INVALID, "rt_write_jpeg: quality must be 1..100");
    RT_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    return rt::write_jpeg(ctx, path, rgb, x_res, y_res, quality, flags, ms);
}
