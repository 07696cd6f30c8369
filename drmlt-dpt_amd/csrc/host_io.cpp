// Output files of a render, on the host: the OpenEXR film the reference
// writes with OpenImageIO (WriteImage, src/image.cpp:45-58: three HALF
// channels), the statistics text (Stats::output,
// src/integrators/delayed/delayed.cpp:311-324) and the partial-image time
// log (delayed.cpp:620-632).
//
// The EXR writer follows the published OpenEXR file layout (single-part
// scanline file, version 2): magic and version, the header attributes
// (channels, compression, dataWindow, displayWindow, lineOrder,
// pixelAspectRatio, screenWindowCenter, screenWindowWidth), the chunk offset
// table and the chunks. Channels are stored in name order (B, G, R), each
// scanline channel after channel. ZIP chunks hold 16 scanlines: the bytes are
// split into even- and odd-indexed halves, delta-coded (each byte minus its
// predecessor, plus 128) and deflated with zlib; a chunk that does not
// shrink is stored raw, as the format allows.
#include "host_io.h"

#include <zlib.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

namespace drmlt {

// IEEE 754 binary32 -> binary16, round to nearest even; overflow to
// infinity, NaN kept quiet with its sign.
uint16_t float_to_half(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t ax = x & 0x7fffffffu;
    if (ax > 0x7f800000u) return uint16_t(sign | 0x7e00u);  // NaN
    if (ax >= 0x477ff000u) return uint16_t(sign | 0x7c00u); // >= 65520 rounds to infinity (and infinity)
    if (ax >= 0x38800000u) {                                 // normal half (>= 2^-14)
        const uint32_t mant = ax & 0x7fffffu;
        const uint32_t e = (ax >> 23) - 127 + 15;
        uint32_t h = (e << 10) | (mant >> 13);
        const uint32_t rem = mant & 0x1fffu;
        if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;  // a carry moves into the exponent, as it should
        return uint16_t(sign | h);
    }
    if (ax < 0x33000000u) return uint16_t(sign);  // below 2^-25: rounds to zero
    // subnormal half: units of 2^-24; value = m 2^(e - 150) = (m >> (126 - e)) 2^-24
    const int e = int(ax >> 23);
    const uint32_t mant = (ax & 0x7fffffu) | 0x800000u;
    const int shift = 126 - e;  // 14 .. 24
    uint32_t h = mant >> shift;
    const uint32_t rem = mant & ((1u << shift) - 1u), half = 1u << (shift - 1);
    if (rem > half || (rem == half && (h & 1u))) h++;
    return uint16_t(sign | h);
}

namespace {
void put_bytes(std::string &s, const void *p, size_t n) { s.append(static_cast<const char *>(p), n); }
template <class T> void put(std::string &s, T v) { put_bytes(s, &v, sizeof(T)); }  // little-endian host
void attr(std::string &s, const char *name, const char *type, const std::string &value) {
    s.append(name);
    s.push_back('\0');
    s.append(type);
    s.push_back('\0');
    put<int32_t>(s, int32_t(value.size()));
    s.append(value);
}
}  // namespace

int write_exr(const char *path, const float *rgb, int w, int h, int compression, std::string &err) {
    if (!path || !rgb || w <= 0 || h <= 0) { err = "bad image arguments"; return -1; }
    if (compression != kExrNone && compression != kExrZip) { err = "unsupported EXR compression (0 none, 3 zip)"; return -1; }
    const int lines = compression == kExrZip ? 16 : 1;
    const int nChunks = (h + lines - 1) / lines;

    std::string hdr;
    put<uint32_t>(hdr, 20000630u);  // magic 0x762f3101
    put<uint32_t>(hdr, 2u);         // version 2, single-part scanline
    std::string ch;
    for (const char *name : {"B", "G", "R"}) {
        ch.append(name);
        ch.push_back('\0');
        put<int32_t>(ch, 1);  // HALF
        put<uint8_t>(ch, 0);  // pLinear
        put<uint8_t>(ch, 0);
        put<uint8_t>(ch, 0);
        put<uint8_t>(ch, 0);
        put<int32_t>(ch, 1);  // xSampling
        put<int32_t>(ch, 1);  // ySampling
    }
    ch.push_back('\0');
    attr(hdr, "channels", "chlist", ch);
    attr(hdr, "compression", "compression", std::string(1, char(compression)));
    std::string box;
    put<int32_t>(box, 0);
    put<int32_t>(box, 0);
    put<int32_t>(box, w - 1);
    put<int32_t>(box, h - 1);
    attr(hdr, "dataWindow", "box2i", box);
    attr(hdr, "displayWindow", "box2i", box);
    attr(hdr, "lineOrder", "lineOrder", std::string(1, '\0'));  // increasing y
    std::string f1;
    put<float>(f1, 1.0f);
    attr(hdr, "pixelAspectRatio", "float", f1);
    std::string v2;
    put<float>(v2, 0.0f);
    put<float>(v2, 0.0f);
    attr(hdr, "screenWindowCenter", "v2f", v2);
    attr(hdr, "screenWindowWidth", "float", f1);
    hdr.push_back('\0');

    std::vector<std::string> chunks(nChunks);
    std::vector<uint16_t> raw;
    std::vector<unsigned char> tmp, packed;
    for (int k = 0; k < nChunks; k++) {
        const int y0 = k * lines, y1 = std::min(h, y0 + lines);
        raw.clear();
        for (int y = y0; y < y1; y++)
            for (int c = 2; c >= 0; c--)  // B, G, R
                for (int x = 0; x < w; x++) raw.push_back(float_to_half(rgb[(size_t(y) * w + x) * 3 + c]));
        const unsigned char *data = reinterpret_cast<const unsigned char *>(raw.data());
        size_t n = raw.size() * 2;
        std::string &out = chunks[k];
        put<int32_t>(out, y0);
        if (compression == kExrZip) {
            tmp.resize(n);
            size_t a = 0, b = (n + 1) / 2;
            for (size_t i = 0; i < n; i++) tmp[(i & 1) ? b++ : a++] = data[i];
            for (size_t i = n - 1; i >= 1; i--) tmp[i] = (unsigned char)(int(tmp[i]) - int(tmp[i - 1]) + 128);
            uLongf cap = compressBound(uLong(n));
            packed.resize(cap);
            if (compress(packed.data(), &cap, tmp.data(), uLong(n)) != Z_OK) { err = "zlib compression failed"; return -1; }
            if (size_t(cap) < n) {
                put<int32_t>(out, int32_t(cap));
                put_bytes(out, packed.data(), cap);
                continue;
            }
        }
        put<int32_t>(out, int32_t(n));
        put_bytes(out, data, n);
    }

    std::string file = hdr;
    uint64_t off = hdr.size() + 8ull * nChunks;
    for (int k = 0; k < nChunks; k++) {
        put<uint64_t>(file, off);
        off += chunks[k].size();
    }
    for (int k = 0; k < nChunks; k++) file += chunks[k];
    FILE *fp = std::fopen(path, "wb");
    if (!fp) { err = std::string("cannot open ") + path; return -1; }
    const size_t wr = std::fwrite(file.data(), 1, file.size(), fp);
    const int cl = std::fclose(fp);
    if (wr != file.size() || cl != 0) { err = std::string("write failed: ") + path; return -1; }
    return 0;
}

// Stats::output with the reference's Float (float) ratios and default
// stream formatting. Format 1 gives the two second-stage lines the
// reference's swapped labels (delayed.cpp:320-321: the "isotropic" line
// prints the H2MC ratio and count and vice versa); format 0 names each
// line's own component. Format 2 is the Stats::output of the mcmc and h2mc
// integrators (mcmc.cpp:76-86, h2mc.cpp:140-150): large steps, small steps
// (isotropic first-stage moves plus H2MC moves), and lens steps (mcmc's
// SmallStep never makes one, mutation.cpp:126-165: 0 of 0 there, printed as
// the reference's 0 / 0).
std::string stats_text(const uint64_t s[10], int format) {
    const float rLarge = float(s[1]) / float(s[0]), r1 = float(s[3]) / float(s[2]);
    const float rIso = float(s[5]) / float(s[4]), rH2 = float(s[7]) / float(s[6]);
    std::ostringstream o;
    o << "\n";
    o << "Rendering Statistics \n";
    if (format == 2) {
        // divided at run time, as the reference's is (x86-64: 0 / 0 is the negative NaN)
        volatile float lensAcc = float(s[9]), lensAll = float(s[8]);
        const float rLens = lensAcc / lensAll;
        o << "    Large step mutations acceptance ratio       : " << rLarge << " (" << s[0] << ")\n";
        const float rSmall = float(s[3] + s[7]) / float(s[2] + s[6]);
        o << "    Small step mutations acceptance ratio       : " << rSmall << " (" << (s[2] + s[6]) << ")\n";
        o << "    Lens mutations acceptance ratio    :                    " << rLens << " (" << s[8] << ")\n";
        o << "\n";
        return o.str();
    }
    const bool referenceLabels = format == 1;
    o << "    Large step mutations acceptance ratio                        : " << rLarge << " (" << s[0] << ")\n";
    o << "    First stage mutations acceptance ratio                       : " << r1 << " (" << s[2] << ")\n";
    if (referenceLabels) {
        o << "    Second stage mutations acceptance ratio, isotropic component : " << rH2 << " (" << s[6] << ")\n";
        o << "    Second stage mutations acceptance ratio, H2MC component      : " << rIso << " (" << s[4] << ")\n";
    } else {
        o << "    Second stage mutations acceptance ratio, isotropic component : " << rIso << " (" << s[4] << ")\n";
        o << "    Second stage mutations acceptance ratio, H2MC component      : " << rH2 << " (" << s[6] << ")\n";
    }
    o << "\n";
    return o.str();
}

std::string float_text(float v) {
    std::ostringstream o;
    o << v;
    return o.str();
}

}  // namespace drmlt
