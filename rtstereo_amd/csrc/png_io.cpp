// PNG files in and out of the stereo pipeline's host buffers, on zlib: the
// replacement for the reference driver's OpenCV imread / imwrite
// (seq/Main.cpp:72-73 gray read, :87-88 colour read, :275-282 disparity
// write) and its display scaling (seq/Main.cpp:335-360 scalecvMat).
//
// Decoding follows the PNG specification (ISO/IEC 15948): chunk CRCs are
// checked, the IDAT stream is inflated, every scanline is unfiltered (None,
// Sub, Up, Average, Paeth), Adam7 interlacing is undone, and every colour
// type and bit depth is converted to 8-bit samples as imread does: 16-bit
// samples keep their high byte, 1/2/4-bit gray levels are scaled to 0..255,
// palettes are expanded, alpha is dropped.  The result is B,G,R interleaved
// (imread's colour order) or gray.
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <zlib.h>

#include <math.h>

#include <string>
#include <vector>

#include "rtstereo.h"

namespace rts {
int fail(int code, const char *fmt, ...);
}

namespace {

const uint8_t kSig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};

uint32_t be32(const uint8_t *p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

void put32(std::vector<uint8_t> &v, uint32_t x) {
    v.push_back((uint8_t)(x >> 24));
    v.push_back((uint8_t)(x >> 16));
    v.push_back((uint8_t)(x >> 8));
    v.push_back((uint8_t)x);
}

bool read_file(const char *path, std::vector<uint8_t> &out) {
    FILE *f = fopen(path, "rb");
    if (!f) return false;
    uint8_t buf[1 << 16];
    size_t n;
    while ((n = fread(buf, 1, sizeof buf, f)) > 0) out.insert(out.end(), buf, buf + n);
    const bool ok = !ferror(f);
    fclose(f);
    return ok;
}

// Paeth predictor (PNG specification, filter type 4).
inline int paeth(int a, int b, int c) {
    const int p = a + b - c, pa = abs(p - a), pb = abs(p - b), pc = abs(p - c);
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

// Undoes the filter of one scanline in place; `prev` is the unfiltered
// previous line of the same pass (NULL for its first line); bpp is the
// filter's byte distance (bytes per pixel, at least 1).
bool unfilter(uint8_t type, uint8_t *cur, const uint8_t *prev, size_t len, int bpp) {
    switch (type) {
        case 0: return true;
        case 1:
            for (size_t i = bpp; i < len; i++) cur[i] = (uint8_t)(cur[i] + cur[i - bpp]);
            return true;
        case 2:
            if (prev)
                for (size_t i = 0; i < len; i++) cur[i] = (uint8_t)(cur[i] + prev[i]);
            return true;
        case 3:
            for (size_t i = 0; i < len; i++) {
                const int a = i >= (size_t)bpp ? cur[i - bpp] : 0, b = prev ? prev[i] : 0;
                cur[i] = (uint8_t)(cur[i] + ((a + b) >> 1));
            }
            return true;
        case 4:
            for (size_t i = 0; i < len; i++) {
                const int a = i >= (size_t)bpp ? cur[i - bpp] : 0, b = prev ? prev[i] : 0;
                const int c = (prev && i >= (size_t)bpp) ? prev[i - bpp] : 0;
                cur[i] = (uint8_t)(cur[i] + paeth(a, b, c));
            }
            return true;
    }
    return false;
}

struct Header {
    uint32_t w = 0, h = 0;
    int depth = 0, ctype = 0, interlace = 0;
    int samples() const { return ctype == 0 ? 1 : ctype == 2 ? 3 : ctype == 3 ? 1 : ctype == 4 ? 2 : 4; }
};

// Sample s (0-based within the row) of an unfiltered scanline, as an 8-bit
// value (palette index for colour type 3).
inline int sample8(const uint8_t *row, size_t s, int depth, int ctype) {
    if (depth == 8) return row[s];
    if (depth == 16) return row[2 * s];   // high byte
    const int per = 8 / depth;
    const int v = (row[s / per] >> (8 - depth * (int)(s % per + 1))) & ((1 << depth) - 1);
    return ctype == 3 ? v : v * 255 / ((1 << depth) - 1);
}

int decode(const std::vector<uint8_t> &file, int channels, uint8_t **pixels, int *W, int *H) {
    if (file.size() < 8 || memcmp(file.data(), kSig, 8) != 0)
        return rts::fail(RTS_ERR_INVALID, "not a PNG file (bad signature)");
    Header hd;
    std::vector<uint8_t> idat, palette;
    bool have_hdr = false, have_end = false;
    size_t pos = 8;
    while (pos + 12 <= file.size() && !have_end) {
        const uint32_t len = be32(&file[pos]);
        if (len > file.size() - pos - 12) return rts::fail(RTS_ERR_INVALID, "PNG chunk runs past the end of the file");
        const uint8_t *type = &file[pos + 4], *data = &file[pos + 8];
        const uint32_t crc = be32(data + len);
        if ((uint32_t)crc32(0L, type, len + 4) != crc) return rts::fail(RTS_ERR_INVALID, "PNG chunk CRC mismatch");
        if (!memcmp(type, "IHDR", 4)) {
            if (len != 13) return rts::fail(RTS_ERR_INVALID, "bad IHDR length");
            hd.w = be32(data);
            hd.h = be32(data + 4);
            hd.depth = data[8];
            hd.ctype = data[9];
            hd.interlace = data[12];
            if (data[10] != 0 || data[11] != 0) return rts::fail(RTS_ERR_INVALID, "unknown PNG compression or filter method");
            have_hdr = true;
        } else if (!memcmp(type, "PLTE", 4)) {
            palette.assign(data, data + len);
        } else if (!memcmp(type, "IDAT", 4)) {
            idat.insert(idat.end(), data, data + len);
        } else if (!memcmp(type, "IEND", 4)) {
            have_end = true;
        } else if (!(type[0] & 0x20)) {   // critical chunk we do not know
            return rts::fail(RTS_ERR_INVALID, "unsupported critical PNG chunk %.4s", (const char *)type);
        }
        pos += 12 + (size_t)len;
    }
    if (!have_hdr || !have_end || idat.empty()) return rts::fail(RTS_ERR_INVALID, "incomplete PNG file");
    const int ct = hd.ctype, dp = hd.depth;
    const bool ok_depth = (ct == 0 && (dp == 1 || dp == 2 || dp == 4 || dp == 8 || dp == 16)) ||
                          (ct == 3 && (dp == 1 || dp == 2 || dp == 4 || dp == 8)) ||
                          ((ct == 2 || ct == 4 || ct == 6) && (dp == 8 || dp == 16));
    if (!ok_depth) return rts::fail(RTS_ERR_INVALID, "invalid PNG colour type %d / bit depth %d", ct, dp);
    if (hd.w == 0 || hd.h == 0 || hd.w > (1u << 16) || hd.h > (1u << 16) || (uint64_t)hd.w * hd.h > (1ull << 30))
        return rts::fail(RTS_ERR_INVALID, "PNG size %ux%u not supported", hd.w, hd.h);
    if (hd.interlace > 1) return rts::fail(RTS_ERR_INVALID, "unknown PNG interlace method");
    if (ct == 3 && (palette.empty() || palette.size() % 3)) return rts::fail(RTS_ERR_INVALID, "missing PNG palette");
    if (channels != 1 && channels != 3) return rts::fail(RTS_ERR_INVALID, "channels must be 1 or 3");

    const int ns = hd.samples();
    const int bits_px = ns * dp;
    const int bpp = (bits_px + 7) / 8;
    // passes: Adam7 (start x, start y, step x, step y) or the whole image
    static const int A7[7][4] = {{0, 0, 8, 8}, {4, 0, 8, 8}, {0, 4, 4, 8}, {2, 0, 4, 4},
                                 {0, 2, 2, 4}, {1, 0, 2, 2}, {0, 1, 1, 2}};
    const int npass = hd.interlace ? 7 : 1;
    size_t raw_size = 0;
    for (int p = 0; p < npass; p++) {
        const uint32_t sx = hd.interlace ? A7[p][0] : 0, sy = hd.interlace ? A7[p][1] : 0;
        const uint32_t dx = hd.interlace ? A7[p][2] : 1, dy = hd.interlace ? A7[p][3] : 1;
        const uint32_t pw = hd.w > sx ? (hd.w - sx + dx - 1) / dx : 0, ph = hd.h > sy ? (hd.h - sy + dy - 1) / dy : 0;
        if (pw && ph) raw_size += (size_t)ph * (1 + ((size_t)pw * bits_px + 7) / 8);
    }
    std::vector<uint8_t> raw(raw_size);
    z_stream zs;
    memset(&zs, 0, sizeof zs);
    if (inflateInit(&zs) != Z_OK) return rts::fail(RTS_ERR_INVALID, "zlib inflateInit failed");
    zs.next_in = idat.data();
    zs.avail_in = (uInt)idat.size();
    zs.next_out = raw.data();
    zs.avail_out = (uInt)raw.size();
    const int zr = inflate(&zs, Z_FINISH);
    const size_t got = raw.size() - zs.avail_out;
    inflateEnd(&zs);
    if ((zr != Z_STREAM_END && zr != Z_OK && zr != Z_BUF_ERROR) || got != raw.size())
        return rts::fail(RTS_ERR_INVALID, "PNG image data is corrupt or truncated");

    const size_t npx = (size_t)hd.w * hd.h;
    uint8_t *out = (uint8_t *)malloc(npx * channels);
    if (!out) return rts::fail(RTS_ERR_NOMEM, "out of host memory for a %ux%u image", hd.w, hd.h);
    size_t off = 0;
    for (int p = 0; p < npass; p++) {
        const uint32_t sx = hd.interlace ? A7[p][0] : 0, sy = hd.interlace ? A7[p][1] : 0;
        const uint32_t dx = hd.interlace ? A7[p][2] : 1, dy = hd.interlace ? A7[p][3] : 1;
        const uint32_t pw = hd.w > sx ? (hd.w - sx + dx - 1) / dx : 0, ph = hd.h > sy ? (hd.h - sy + dy - 1) / dy : 0;
        if (!pw || !ph) continue;
        const size_t stride = ((size_t)pw * bits_px + 7) / 8;
        const uint8_t *prev = nullptr;
        for (uint32_t r = 0; r < ph; r++) {
            uint8_t *line = &raw[off];
            if (!unfilter(line[0], line + 1, prev, stride, bpp)) {
                free(out);
                return rts::fail(RTS_ERR_INVALID, "unknown PNG filter type %d", line[0]);
            }
            const uint8_t *row = line + 1;
            const uint32_t y = sy + r * dy;
            for (uint32_t c = 0; c < pw; c++) {
                const uint32_t x = sx + c * dx;
                int R, G, B;
                if (ct == 3) {
                    const int idx = sample8(row, c, dp, ct);
                    if ((size_t)idx * 3 + 2 >= palette.size()) {
                        free(out);
                        return rts::fail(RTS_ERR_INVALID, "PNG palette index %d out of range", idx);
                    }
                    R = palette[3 * idx];
                    G = palette[3 * idx + 1];
                    B = palette[3 * idx + 2];
                } else if (ct == 0 || ct == 4) {
                    R = G = B = sample8(row, (size_t)c * ns, dp, ct);
                } else {
                    R = sample8(row, (size_t)c * ns, dp, ct);
                    G = sample8(row, (size_t)c * ns + 1, dp, ct);
                    B = sample8(row, (size_t)c * ns + 2, dp, ct);
                }
                uint8_t *o = out + ((size_t)y * hd.w + x) * channels;
                if (channels == 3) {
                    o[0] = (uint8_t)B;
                    o[1] = (uint8_t)G;
                    o[2] = (uint8_t)R;
                } else if (ct == 0 || ct == 4) {
                    o[0] = (uint8_t)R;   // gray images keep their level
                } else {
                    o[0] = (uint8_t)((29u * B + 150u * G + 77u * R + 128u) >> 8);
                }
            }
            prev = row;
            off += 1 + stride;
        }
    }
    *pixels = out;
    *W = (int)hd.w;
    *H = (int)hd.h;
    return RTS_OK;
}

void chunk(std::vector<uint8_t> &f, const char *type, const uint8_t *data, size_t len) {
    put32(f, (uint32_t)len);
    const size_t start = f.size();
    f.insert(f.end(), type, type + 4);
    if (len) f.insert(f.end(), data, data + len);
    put32(f, (uint32_t)crc32(0L, &f[start], (uInt)(len + 4)));
}

}  // namespace

extern "C" {

rts_status rts_png_read(const char *path, int channels, uint8_t **pixels, int *W, int *H) {
    if (!path || !pixels || !W || !H) return rts::fail(RTS_ERR_INVALID, "null argument");
    *pixels = nullptr;
    std::vector<uint8_t> file;
    if (!read_file(path, file)) return rts::fail(RTS_ERR_INVALID, "cannot read %s", path);
    return decode(file, channels, pixels, W, H);
}

rts_status rts_png_decode(const uint8_t *data, size_t bytes, int channels, uint8_t **pixels, int *W, int *H) {
    if (!data || !pixels || !W || !H) return rts::fail(RTS_ERR_INVALID, "null argument");
    *pixels = nullptr;
    std::vector<uint8_t> file(data, data + bytes);
    return decode(file, channels, pixels, W, H);
}

void rts_png_free(uint8_t *pixels) { free(pixels); }

rts_status rts_png_write(const char *path, const uint8_t *pixels, int W, int H, int channels) {
    if (!path || !pixels) return rts::fail(RTS_ERR_INVALID, "null argument");
    if (W < 1 || H < 1 || W > (1 << 16) || H > (1 << 16))
        return rts::fail(RTS_ERR_INVALID, "image size %dx%d not supported", W, H);
    if (channels != 1 && channels != 3) return rts::fail(RTS_ERR_INVALID, "channels must be 1 or 3");
    const size_t stride = (size_t)W * channels;
    // per row, the filter (None, Sub, Up, Paeth) with the smallest sum of
    // absolute signed bytes, the usual compression heuristic
    std::vector<uint8_t> raw((stride + 1) * H), cand(stride), best(stride), rgb(stride), prev(stride, 0);
    for (int y = 0; y < H; y++) {
        const uint8_t *src = pixels + (size_t)y * stride;
        if (channels == 3)
            for (int x = 0; x < W; x++) {   // B,G,R -> R,G,B
                rgb[3 * x] = src[3 * x + 2];
                rgb[3 * x + 1] = src[3 * x + 1];
                rgb[3 * x + 2] = src[3 * x];
            }
        else
            memcpy(rgb.data(), src, stride);
        long best_cost = -1;
        uint8_t best_type = 0;
        for (uint8_t t : {0, 1, 2, 4}) {
            long cost = 0;
            for (size_t i = 0; i < stride; i++) {
                const int a = i >= (size_t)channels ? rgb[i - channels] : 0, b = y ? prev[i] : 0;
                const int c = (y && i >= (size_t)channels) ? prev[i - channels] : 0;
                const int pred = t == 0 ? 0 : t == 1 ? a : t == 2 ? b : paeth(a, b, c);
                cand[i] = (uint8_t)(rgb[i] - pred);
                cost += abs((int)(int8_t)cand[i]);
            }
            if (best_cost < 0 || cost < best_cost) {
                best_cost = cost;
                best_type = t;
                best.swap(cand);
            }
        }
        raw[(stride + 1) * y] = best_type;
        memcpy(&raw[(stride + 1) * y + 1], best.data(), stride);
        prev = rgb;
    }
    uLongf zlen = compressBound((uLong)raw.size());
    std::vector<uint8_t> z(zlen);
    if (compress2(z.data(), &zlen, raw.data(), (uLong)raw.size(), 6) != Z_OK)
        return rts::fail(RTS_ERR_INVALID, "zlib compression failed");
    std::vector<uint8_t> f(kSig, kSig + 8);
    uint8_t ihdr[13];
    const uint32_t w = (uint32_t)W, h = (uint32_t)H;
    const uint8_t hdr[13] = {(uint8_t)(w >> 24), (uint8_t)(w >> 16), (uint8_t)(w >> 8), (uint8_t)w,
                             (uint8_t)(h >> 24), (uint8_t)(h >> 16), (uint8_t)(h >> 8), (uint8_t)h,
                             8, (uint8_t)(channels == 3 ? 2 : 0), 0, 0, 0};
    memcpy(ihdr, hdr, 13);
    chunk(f, "IHDR", ihdr, 13);
    chunk(f, "IDAT", z.data(), zlen);
    chunk(f, "IEND", nullptr, 0);
    FILE *fp = fopen(path, "wb");
    if (!fp) return rts::fail(RTS_ERR_INVALID, "cannot create %s", path);
    const bool ok = fwrite(f.data(), 1, f.size(), fp) == f.size();
    if (fclose(fp) != 0 || !ok) return rts::fail(RTS_ERR_INVALID, "write to %s failed", path);
    return RTS_OK;
}

rts_status rts_disp_scale(uint8_t *disp, int W, int H, float coeff) {
    if (!disp || W < 1 || H < 1) return rts::fail(RTS_ERR_INVALID, "bad disparity map");
    const size_t n = (size_t)W * H;
    if (!(coeff > 0.0f)) {   // automatic: the largest disparity becomes 255
        uint8_t mx = 0;
        for (size_t i = 0; i < n; i++) mx = disp[i] > mx ? disp[i] : mx;
        if (mx == 0) return RTS_OK;
        coeff = (float)255 / (float)mx;
    }
    // the product in float, rounded half up through a double, as the
    // reference's round macro does with its double 0.5
    for (size_t i = 0; i < n; i++) {
        const float prod = (float)disp[i] * coeff;
        const double r = floor((double)prod + 0.5);
        disp[i] = (uint8_t)(r > 255.0 ? 255.0 : r);
    }
    return RTS_OK;
}

}  // extern "C"
