// image_io.cpp — image files for textures and renders, with the conventions of
// cpu::Image::open / save (cpu/image.cpp:6-72):
//   open: rows flipped so that row 0 is the bottom; 8-bit channels scaled by
//         1/255; RGB linearised with pow(., 2.2); alpha kept as stored; float
//         (HDR) RGB tonemapped (Drago et al. 2003), scaled by 4 and clipped to
//         [0, 1] before the pow; `grayscale` reads one channel (colour converted
//         with the ITU-R BT.601 luma weights) into the alpha image.
//   save: pow(., 1/2.2), x255, rounded and saturated to 8 bits, top row first.
// Formats read: PNG (1/2/4/8/16-bit grey, grey+alpha, RGB, RGBA, palette with
// tRNS; non-interlaced), PNM (P2/P3/P5/P6), PFM (Pf/PF), Radiance RGBE (.hdr).
// Written: PNG (8-bit RGB / grey), PPM / PGM.  No JPEG: such a file fails with
// a message naming the file.
#include <zlib.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>

#include "tinypt.h"

namespace tinypt {
namespace {

std::vector<uint8_t> read_file(const std::string &path) {
    std::ifstream f(path, std::ios::binary);
    TINYPT_CHECK_EX(f.good(), std::runtime_error, "cannot open image " + path);
    return std::vector<uint8_t>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

std::string lower_ext(const std::string &path) {
    auto dot = path.find_last_of('.');
    std::string e = dot == std::string::npos ? "" : path.substr(dot + 1);
    for (auto &c : e) c = (char)std::tolower((unsigned char)c);
    return e;
}

// Decoded file: top row first, `c` interleaved channels in RGB(A) order, either
// 8-bit (`u8`), 16-bit (`u16`) or float (`f`) samples.
struct Raw {
    int w = 0, h = 0, c = 0;
    std::vector<uint8_t> u8;
    std::vector<uint16_t> u16;
    std::vector<float> f;
    bool is_float() const { return !f.empty(); }
    float sample(size_t i) const { // 8/16-bit normalised to [0, 1]; float as stored
        if (!u8.empty()) return u8[i] / 255.f;
        if (!u16.empty()) return u16[i] / 65535.f;
        return f[i];
    }
};

// ---------------------------------------------------------------- PNG
uint32_t be32(const uint8_t *p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

uint8_t paeth(int a, int b, int c) {
    int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    return (uint8_t)(pa <= pb && pa <= pc ? a : (pb <= pc ? b : c));
}

Raw decode_png(const std::vector<uint8_t> &d, const std::string &path) {
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    TINYPT_CHECK_EX(d.size() > 8 && std::memcmp(d.data(), sig, 8) == 0, std::runtime_error, "not a PNG file: " + path);
    int w = 0, h = 0, depth = 0, ctype = 0, interlace = 0;
    std::vector<uint8_t> idat, plte, trns;
    size_t p = 8;
    while (p + 12 <= d.size()) {
        uint32_t len = be32(&d[p]);
        TINYPT_CHECK_EX(p + 12 + (size_t)len <= d.size(), std::runtime_error, "truncated PNG chunk in " + path);
        const uint8_t *type = &d[p + 4], *data = &d[p + 8];
        uint32_t crc = (uint32_t)crc32(0L, type, len + 4);
        TINYPT_CHECK_EX(crc == be32(data + len), std::runtime_error, "PNG chunk CRC mismatch in " + path);
        if (!std::memcmp(type, "IHDR", 4)) {
            TINYPT_CHECK_EX(len == 13, std::runtime_error, "bad PNG header in " + path);
            w = (int)be32(data);
            h = (int)be32(data + 4);
            depth = data[8];
            ctype = data[9];
            interlace = data[12];
        } else if (!std::memcmp(type, "PLTE", 4)) {
            plte.assign(data, data + len);
        } else if (!std::memcmp(type, "tRNS", 4)) {
            trns.assign(data, data + len);
        } else if (!std::memcmp(type, "IDAT", 4)) {
            idat.insert(idat.end(), data, data + len);
        } else if (!std::memcmp(type, "IEND", 4)) {
            break;
        }
        p += 12 + len;
    }
    TINYPT_CHECK_EX(w > 0 && h > 0, std::runtime_error, "PNG without image header: " + path);
    TINYPT_CHECK_EX(interlace == 0, std::runtime_error, "interlaced PNG is not supported: " + path);
    int chans = ctype == 0 ? 1 : ctype == 2 ? 3 : ctype == 3 ? 1 : ctype == 4 ? 2 : ctype == 6 ? 4 : 0;
    TINYPT_CHECK_EX(chans > 0, std::runtime_error, "unsupported PNG colour type in " + path);
    TINYPT_CHECK_EX(depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16, std::runtime_error,
                    "unsupported PNG bit depth in " + path);
    const size_t bits = (size_t)chans * depth, row = ((size_t)w * bits + 7) / 8, bpp = std::max<size_t>(1, bits / 8);
    std::vector<uint8_t> raw((row + 1) * (size_t)h);
    uLongf out_len = (uLongf)raw.size();
    TINYPT_CHECK_EX(uncompress(raw.data(), &out_len, idat.data(), (uLong)idat.size()) == Z_OK && out_len == raw.size(),
                    std::runtime_error, "corrupt PNG image data in " + path);
    // undo the per-row filters (PNG spec, section 9)
    std::vector<uint8_t> img(row * (size_t)h);
    for (int y = 0; y < h; y++) {
        const uint8_t *src = &raw[(size_t)y * (row + 1)];
        uint8_t *cur = &img[(size_t)y * row];
        const uint8_t *prev = y ? &img[(size_t)(y - 1) * row] : nullptr;
        int ft = src[0];
        for (size_t i = 0; i < row; i++) {
            int a = i >= bpp ? cur[i - bpp] : 0, b = prev ? prev[i] : 0, c = (prev && i >= bpp) ? prev[i - bpp] : 0;
            int x = src[1 + i];
            switch (ft) {
            case 0: break;
            case 1: x += a; break;
            case 2: x += b; break;
            case 3: x += (a + b) >> 1; break;
            case 4: x += paeth(a, b, c); break;
            default: TINYPT_THROW("bad PNG filter type in " + path);
            }
            cur[i] = (uint8_t)x;
        }
    }
    Raw r;
    r.w = w;
    r.h = h;
    auto sample = [&](int y, size_t k) -> int { // k-th sample of row y at the file's bit depth
        const uint8_t *rp = &img[(size_t)y * row];
        if (depth == 8) return rp[k];
        if (depth == 16) return rp[2 * k] << 8 | rp[2 * k + 1];
        size_t bit = k * depth;
        return (rp[bit >> 3] >> (8 - depth - (bit & 7))) & ((1 << depth) - 1);
    };
    if (ctype == 3) { // palette -> RGB, or RGBA with tRNS
        TINYPT_CHECK_EX(!plte.empty() && plte.size() % 3 == 0, std::runtime_error, "PNG palette missing in " + path);
        r.c = trns.empty() ? 3 : 4;
        r.u8.resize((size_t)w * h * r.c);
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) {
                size_t idx = (size_t)sample(y, (size_t)x);
                TINYPT_CHECK_EX(3 * idx + 2 < plte.size(), std::runtime_error, "PNG palette index out of range in " + path);
                uint8_t *o = &r.u8[((size_t)y * w + x) * r.c];
                for (int c = 0; c < 3; c++) o[c] = plte[3 * idx + c];
                if (r.c == 4) o[3] = idx < trns.size() ? trns[idx] : 255;
            }
        return r;
    }
    r.c = chans;
    const size_t n = (size_t)w * h * chans;
    if (depth == 16) {
        r.u16.resize(n);
        for (int y = 0; y < h; y++)
            for (size_t k = 0; k < (size_t)w * chans; k++) r.u16[(size_t)y * w * chans + k] = (uint16_t)sample(y, k);
    } else {
        r.u8.resize(n);
        const int scale = depth == 8 ? 1 : 255 / ((1 << depth) - 1); // low-depth grey expands to 8 bits
        for (int y = 0; y < h; y++)
            for (size_t k = 0; k < (size_t)w * chans; k++) r.u8[(size_t)y * w * chans + k] = (uint8_t)(sample(y, k) * scale);
    }
    return r;
}

// ---------------------------------------------------------------- PNM / PFM / Radiance
struct Tokens { // whitespace-separated header fields, '#' comments skipped (PNM)
    const std::vector<uint8_t> &d;
    size_t p = 0;
    std::string next() {
        for (;;) {
            while (p < d.size() && std::isspace(d[p])) p++;
            if (p < d.size() && d[p] == '#') {
                while (p < d.size() && d[p] != '\n') p++;
                continue;
            }
            break;
        }
        std::string s;
        while (p < d.size() && !std::isspace(d[p])) s += (char)d[p++];
        return s;
    }
};

Raw decode_pnm(const std::vector<uint8_t> &d, const std::string &path) {
    Tokens t{d};
    std::string magic = t.next();
    TINYPT_CHECK_EX(magic == "P2" || magic == "P3" || magic == "P5" || magic == "P6", std::runtime_error,
                    "unsupported PNM type in " + path);
    Raw r;
    r.w = std::stoi(t.next());
    r.h = std::stoi(t.next());
    int maxval = std::stoi(t.next());
    TINYPT_CHECK_EX(r.w > 0 && r.h > 0 && maxval > 0 && maxval < 65536, std::runtime_error, "bad PNM header in " + path);
    r.c = (magic == "P3" || magic == "P6") ? 3 : 1;
    const size_t n = (size_t)r.w * r.h * r.c;
    std::vector<int> v(n);
    if (magic == "P2" || magic == "P3") {
        for (size_t i = 0; i < n; i++) v[i] = std::stoi(t.next());
    } else {
        size_t p = t.p + 1; // one whitespace byte after maxval
        size_t bytes = maxval < 256 ? 1 : 2;
        TINYPT_CHECK_EX(p + n * bytes <= d.size(), std::runtime_error, "truncated PNM data in " + path);
        for (size_t i = 0; i < n; i++) v[i] = bytes == 1 ? d[p + i] : d[p + 2 * i] << 8 | d[p + 2 * i + 1];
    }
    if (maxval == 255) {
        r.u8.assign(v.begin(), v.end());
    } else { // other ranges rescaled to 16 bits
        r.u16.resize(n);
        for (size_t i = 0; i < n; i++) r.u16[i] = (uint16_t)std::lround((double)v[i] * 65535.0 / maxval);
    }
    return r;
}

Raw decode_pfm(const std::vector<uint8_t> &d, const std::string &path) {
    Tokens t{d};
    std::string magic = t.next();
    TINYPT_CHECK_EX(magic == "PF" || magic == "Pf", std::runtime_error, "not a PFM file: " + path);
    Raw r;
    r.c = magic == "PF" ? 3 : 1;
    r.w = std::stoi(t.next());
    r.h = std::stoi(t.next());
    double scale = std::stod(t.next());
    size_t p = t.p + 1;
    const size_t n = (size_t)r.w * r.h * r.c;
    TINYPT_CHECK_EX(r.w > 0 && r.h > 0 && p + 4 * n <= d.size(), std::runtime_error, "bad PFM data in " + path);
    const bool little = scale < 0;
    r.f.resize(n);
    for (int y = 0; y < r.h; y++) // PFM stores the bottom row first
        for (size_t k = 0; k < (size_t)r.w * r.c; k++) {
            const uint8_t *q = &d[p + 4 * ((size_t)(r.h - 1 - y) * r.w * r.c + k)];
            uint32_t u = little ? (uint32_t)q[0] | (uint32_t)q[1] << 8 | (uint32_t)q[2] << 16 | (uint32_t)q[3] << 24
                                : be32(q);
            float f;
            std::memcpy(&f, &u, 4);
            r.f[(size_t)y * r.w * r.c + k] = f;
        }
    return r;
}

// Radiance RGBE (Ward, "Real Pixels", Graphics Gems II): value = m * 2^(e - 136);
// scanlines flat or new-style run-length encoded per component.
Raw decode_hdr(const std::vector<uint8_t> &d, const std::string &path) {
    size_t p = 0;
    auto line = [&]() {
        std::string s;
        while (p < d.size() && d[p] != '\n') s += (char)d[p++];
        if (p < d.size()) p++;
        return s;
    };
    std::string first = line();
    TINYPT_CHECK_EX(first.rfind("#?", 0) == 0, std::runtime_error, "not a Radiance HDR file: " + path);
    for (std::string s = line(); !s.empty(); s = line())
        if (s.rfind("FORMAT=", 0) == 0)
            TINYPT_CHECK_EX(s == "FORMAT=32-bit_rle_rgbe", std::runtime_error, "unsupported HDR format in " + path);
    std::istringstream res(line());
    std::string ya, xa;
    Raw r;
    res >> ya >> r.h >> xa >> r.w;
    TINYPT_CHECK_EX(ya == "-Y" && xa == "+X" && r.w > 0 && r.h > 0, std::runtime_error,
                    "unsupported HDR orientation in " + path);
    r.c = 3;
    r.f.resize((size_t)r.w * r.h * 3);
    std::vector<uint8_t> sl((size_t)r.w * 4);
    for (int y = 0; y < r.h; y++) {
        TINYPT_CHECK_EX(p + 4 <= d.size(), std::runtime_error, "truncated HDR data in " + path);
        if (r.w >= 8 && r.w < 32768 && d[p] == 2 && d[p + 1] == 2 && ((d[p + 2] << 8) | d[p + 3]) == r.w) {
            p += 4;
            for (int c = 0; c < 4; c++)
                for (int x = 0; x < r.w;) {
                    TINYPT_CHECK_EX(p < d.size(), std::runtime_error, "truncated HDR data in " + path);
                    int cnt = d[p++];
                    if (cnt > 128) {
                        cnt -= 128;
                        TINYPT_CHECK_EX(x + cnt <= r.w && p < d.size(), std::runtime_error, "bad HDR run in " + path);
                        for (int k = 0; k < cnt; k++) sl[(size_t)(x + k) * 4 + c] = d[p];
                        p++;
                    } else {
                        TINYPT_CHECK_EX(cnt > 0 && x + cnt <= r.w && p + cnt <= d.size(), std::runtime_error,
                                        "bad HDR run in " + path);
                        for (int k = 0; k < cnt; k++) sl[(size_t)(x + k) * 4 + c] = d[p++];
                    }
                    x += cnt;
                }
        } else {
            TINYPT_CHECK_EX(p + sl.size() <= d.size(), std::runtime_error, "truncated HDR data in " + path);
            std::memcpy(sl.data(), &d[p], sl.size());
            p += sl.size();
        }
        for (int x = 0; x < r.w; x++) {
            const uint8_t *q = &sl[(size_t)x * 4];
            float s = q[3] ? std::ldexp(1.f, (int)q[3] - 136) : 0.f;
            for (int c = 0; c < 3; c++) r.f[((size_t)y * r.w + x) * 3 + c] = q[c] * s;
        }
    }
    return r;
}

Raw decode(const std::string &path) {
    std::vector<uint8_t> d = read_file(path);
    std::string e = lower_ext(path);
    if (d.size() >= 8 && d[0] == 137 && d[1] == 'P' && d[2] == 'N' && d[3] == 'G') return decode_png(d, path);
    if (d.size() >= 2 && d[0] == 'P' && (d[1] == 'F' || d[1] == 'f')) return decode_pfm(d, path);
    if (d.size() >= 2 && d[0] == 'P' && d[1] >= '2' && d[1] <= '6') return decode_pnm(d, path);
    if (d.size() >= 2 && d[0] == '#' && d[1] == '?') return decode_hdr(d, path);
    TINYPT_THROW("cannot decode image " + path + " (supported: PNG, PNM, PFM, Radiance HDR)");
}

float luma(float r, float g, float b) { return 0.299f * r + 0.587f * g + 0.114f * b; }

} // namespace

// Affine map of all samples onto [0, 1] (min -> 0, max -> 1), then ^(1 / gamma).
void normalize(std::vector<float> &v, float gamma) {
    if (v.empty()) return;
    float lo = *std::min_element(v.begin(), v.end()), hi = *std::max_element(v.begin(), v.end());
    float span = hi > lo ? hi - lo : 1.f;
    for (auto &x : v) {
        x = std::max((x - lo) / span, 0.f);
        if (gamma != 1.f) x = std::pow(x, 1.f / gamma);
    }
}

// Drago, Myszkowski, Annen, Chiba, "Adaptive Logarithmic Mapping For Displaying
// High Contrast Scenes" (Eurographics 2003), arranged as OpenCV's TonemapDrago
// is described (bias 0.85, saturation 1): the image is first mapped onto
// [0, 1]; its luminance L, divided by the log-average, becomes
// log(1 + L) / log(2 + 8 (L / L_max)^(log 0.85 / log 0.5)); each colour is
// scaled by mapped / L; the result is mapped onto [0, 1] again and raised to
// 1 / gamma.  OpenCV is not available here, so the match with its output is
// not checked.
void tonemap_drago(std::vector<float> &rgb, size_t n, float gamma) {
    const float bias = 0.85f;
    normalize(rgb, 1.f);
    std::vector<float> lum(n);
    double log_sum = 0;
    for (size_t i = 0; i < n; i++) {
        lum[i] = luma(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]);
        log_sum += std::log(std::max(lum[i], 1e-30f));
    }
    const float mean = (float)std::exp(log_sum / (double)std::max<size_t>(n, 1));
    float lmax = 0;
    for (size_t i = 0; i < n; i++) lmax = std::max(lmax, lum[i] /= mean);
    TINYPT_CHECK_EX(lmax > 0, std::runtime_error, "HDR image is black");
    const float ex = std::log(bias) / std::log(0.5f);
    for (size_t i = 0; i < n; i++) {
        float L = lum[i];
        float mapped = std::log(L + 1.f) / std::log(2.f + 8.f * std::pow(L / lmax, ex));
        float k = L > 0 ? mapped / L : 0.f;
        for (int c = 0; c < 3; c++) rgb[3 * i + c] *= k;
    }
    normalize(rgb, gamma);
}

RGBAImage Image::open(const std::string &path, bool grayscale) {
    Raw r = decode(path);
    const size_t n = (size_t)r.w * r.h;
    RGBAImage out;
    auto flip_row = [&](int y) { return r.h - 1 - y; }; // file rows are top first; ours bottom first
    if (grayscale) {
        std::vector<float> a(n);
        for (int y = 0; y < r.h; y++)
            for (int x = 0; x < r.w; x++) {
                size_t i = ((size_t)y * r.w + x) * r.c;
                float v;
                if (r.c >= 3) {
                    if (!r.u8.empty()) { // OpenCV's fixed-point BT.601 conversion of 8-bit colour
                        int g = (r.u8[i] * 4899 + r.u8[i + 1] * 9617 + r.u8[i + 2] * 1868 + 8192) >> 14;
                        v = g / 255.f;
                    } else {
                        v = luma(r.sample(i), r.sample(i + 1), r.sample(i + 2));
                    }
                } else {
                    v = r.sample(i);
                }
                a[(size_t)flip_row(y) * r.w + x] = v;
            }
        out.alpha = std::make_shared<Image>(Image::from_f32(r.w, r.h, 1, a.data()));
        return out;
    }
    const int cc = r.c >= 3 ? 3 : 1;
    std::vector<float> rgb(n * 3);
    std::vector<float> alpha;
    if (r.c == 2 || r.c == 4) alpha.resize(n);
    for (int y = 0; y < r.h; y++)
        for (int x = 0; x < r.w; x++) {
            size_t i = ((size_t)y * r.w + x) * r.c, o = (size_t)flip_row(y) * r.w + x;
            for (int c = 0; c < 3; c++) rgb[3 * o + c] = r.sample(i + (cc == 3 ? c : 0));
            if (!alpha.empty()) alpha[o] = r.sample(i + r.c - 1);
        }
    if (r.is_float()) {
        tonemap_drago(rgb, n, SRGB_GAMMA);
        for (auto &v : rgb) v = std::min(std::max(v * 4.f, 0.f), 1.f); // LDR_SCALE (image.cpp:55-56)
    }
    for (auto &v : rgb) v = std::pow(v, SRGB_GAMMA);
    out.rgb = std::make_shared<Image>(Image::from_f32(r.w, r.h, 3, rgb.data()));
    if (!alpha.empty()) out.alpha = std::make_shared<Image>(Image::from_f32(r.w, r.h, 1, alpha.data()));
    return out;
}

namespace {
void put32(std::vector<uint8_t> &o, uint32_t v) {
    for (int s = 24; s >= 0; s -= 8) o.push_back((uint8_t)(v >> s));
}
void png_chunk(std::vector<uint8_t> &o, const char *type, const std::vector<uint8_t> &data) {
    put32(o, (uint32_t)data.size());
    size_t start = o.size();
    o.insert(o.end(), type, type + 4);
    o.insert(o.end(), data.begin(), data.end());
    put32(o, (uint32_t)crc32(0L, &o[start], (uInt)(o.size() - start)));
}
} // namespace

void Image::save(const std::string &path) const {
    TINYPT_CHECK_EX(!empty() && (channels() == 3 || channels() == 1), std::runtime_error,
                    "invalid image channels: " + std::to_string(channels()));
    const int w = width(), h = height(), c = channels();
    std::vector<uint8_t> px((size_t)w * h * c);
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++)
            for (int k = 0; k < c; k++) {
                float v = at(x, h - 1 - y, k); // top row first
                if (c == 3) v = std::pow(std::max(v, 0.f), 1.f / SRGB_GAMMA);
                px[((size_t)y * w + x) * c + k] = (uint8_t)std::min(std::max(std::lround(v * 255.f), 0L), 255L);
            }
    std::string e = lower_ext(path);
    std::vector<uint8_t> file;
    if (e == "png") {
        static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
        file.assign(sig, sig + 8);
        std::vector<uint8_t> ihdr;
        put32(ihdr, (uint32_t)w);
        put32(ihdr, (uint32_t)h);
        ihdr.insert(ihdr.end(), {8, (uint8_t)(c == 3 ? 2 : 0), 0, 0, 0});
        png_chunk(file, "IHDR", ihdr);
        std::vector<uint8_t> raw;
        raw.reserve((size_t)h * (w * c + 1));
        for (int y = 0; y < h; y++) {
            raw.push_back(0); // filter: none
            raw.insert(raw.end(), &px[(size_t)y * w * c], &px[(size_t)(y + 1) * w * c]);
        }
        uLongf len = compressBound((uLong)raw.size());
        std::vector<uint8_t> z(len);
        TINYPT_CHECK_EX(compress2(z.data(), &len, raw.data(), (uLong)raw.size(), 6) == Z_OK, std::runtime_error,
                        "PNG compression failed");
        z.resize(len);
        png_chunk(file, "IDAT", z);
        png_chunk(file, "IEND", {});
    } else if (e == "ppm" || e == "pgm" || e == "pnm") {
        std::string hdr = std::string(c == 3 ? "P6" : "P5") + "\n" + std::to_string(w) + " " + std::to_string(h) + "\n255\n";
        file.assign(hdr.begin(), hdr.end());
        file.insert(file.end(), px.begin(), px.end());
    } else {
        TINYPT_THROW("cannot save image to " + path + " (supported: .png, .ppm, .pgm)");
    }
    std::ofstream f(path, std::ios::binary);
    TINYPT_CHECK_EX(f.good(), std::runtime_error, "cannot save image to " + path);
    f.write((const char *)file.data(), (std::streamsize)file.size());
    TINYPT_CHECK_EX(f.good(), std::runtime_error, "cannot save image to " + path);
}

} // namespace tinypt
