// Reading the reference's serialized subpaths (serialized.h).
#include "serialized.h"

#include <cstring>

namespace drmlt {

bool TriKeyMap::Key::operator==(const Key &o) const { return std::memcmp(w, o.w, sizeof(w)) == 0; }

size_t TriKeyMap::KeyHash::operator()(const Key &k) const {
    // FNV-1a over the 36 bytes
    uint64_t h = 1469598103934665603ull;
    const unsigned char *b = reinterpret_cast<const unsigned char *>(k.w);
    for (size_t i = 0; i < sizeof(k.w); i++) h = (h ^ b[i]) * 1099511628211ull;
    return size_t(h);
}

void TriKeyMap::build(int nt, const float *p0e1e2) {
    map_.clear();
    map_.reserve(size_t(nt) * 2);
    for (int i = 0; i < nt; i++) {
        Key k;
        std::memcpy(k.w, p0e1e2 + 9 * size_t(i), sizeof(k.w));
        map_.emplace(k, i);  // an exact duplicate triangle keeps its first index
    }
}

int TriKeyMap::find(const float *shape) const {
    Key k;
    std::memcpy(k.w, shape + 2, sizeof(k.w));  // after type and isMoving: p0, p1 - p0, p2 - p0 at time 0
    const auto it = map_.find(k);
    return it == map_.end() ? -1 : it->second;
}

namespace {

// A sequential reader over one state's vertParams, bounded by its length.
struct Reader {
    const float *v;
    int64_t len, pos = 0;
    bool ok = true;
    const float *take(int64_t n) {
        if (pos + n > len) {
            ok = false;
            return nullptr;
        }
        const float *p = v + pos;
        pos += n;
        return p;
    }
};

}  // namespace

int decode_subpath(const TriKeyMap &tm, const SceneFacts &sf, int c, int l, const float *lens, const float *primary, const float *vert,
                   int64_t vertLen, drmlt_path_rec &R, std::string &err) {
    std::memset(&R, 0, sizeof(R));
    R.c = c;
    R.l = l;
    R.light = -1;
    R.lprim = -1;
    for (int k = 0; k < 16; k++) {
        R.cprim[k] = -1; R.lgprim[k] = -1;
        R.crr[k] = 1.0f; R.lrr[k] = 1.0f;
    }
    R.screen[0] = lens[0];
    R.screen[1] = lens[1];
    const int P = c + l - 1, D = 2 * (P > 2 ? P : 2);
    for (int k = 0; k < D; k++) R.u[k] = primary[1 + k];  // primary[0] is the time
    Reader rd{vert, vertLen};
    auto tri = [&](const float *shape, int &out) -> bool {
        if (!shape) return false;
        if (shape[0] != 0.0f) { err = "shape record of a kind other than a triangle mesh"; return false; }
        out = tm.find(shape);
        if (out < 0) { err = "a serialized triangle is not in the context's scene"; return false; }
        return true;
    };
    auto bsdf = [&](int prim) -> bool {
        const float *b = rd.take(kSerBsdf);
        if (!b) return false;
        if (int(b[0]) != sf.triMatType[prim]) { err = "a serialized BSDF differs in type from the scene's material"; return false; }
        return true;
    };
    auto light = [&](int &lightOut, int &primOut) -> bool {
        const float *L = rd.take(kSerLight);
        if (!L) return false;
        if (L[0] != 1.0f) { err = "only area lights are supported (light record of another type)"; return false; }
        if (!tri(L + 1, primOut)) return false;
        lightOut = sf.triLight[primOut];
        if (lightOut < 0) { err = "a serialized light triangle is not emissive in the context's scene"; return false; }
        return true;
    };
    err.clear();
    if (!rd.take(3)) goto short_buffer;  // lens vertex position
    if (l > 1) {
        if (!rd.take(1)) goto short_buffer;  // light pick probability
        if (!light(R.light, R.lprim)) goto fail;
        for (int d = 0; d < l - 1; d++) {
            if (!tri(rd.take(kSerShape), R.lgprim[d])) goto fail;
            const float *dv = rd.take(2);  // bsdfDiscrete, useAbsoluteParam
            if (!dv) goto short_buffer;
            R.ldisc[d] = dv[0];
            if (!bsdf(R.lgprim[d])) goto fail;
            if (d == l - 2 && c == 1) return 0;
            if (d == l - 2) break;
            const float *rr = rd.take(1);
            if (!rr) goto short_buffer;
            R.lrr[d] = rr[0];
        }
    }
    for (int d = 0; d < c - 1; d++) {
        if (!tri(rd.take(kSerShape), R.cprim[d])) goto fail;
        if (d == c - 2) {
            if (l == 0) {
                int li = -1, lp = -1;
                if (!light(li, lp)) goto fail;
                if (lp != R.cprim[d]) { err = "the light hit by the last camera vertex is another triangle"; goto fail; }
                if (!rd.take(1)) goto short_buffer;
            } else if (l == 1) {
                if (!light(R.light, R.lprim)) goto fail;
                if (!bsdf(R.cprim[d])) goto fail;
                if (!rd.take(1)) goto short_buffer;
            } else {
                if (!bsdf(R.cprim[d])) goto fail;
            }
            return 0;
        }
        const float *dv = rd.take(2);
        if (!dv) goto short_buffer;
        R.cdisc[d] = dv[0];
        if (!bsdf(R.cprim[d])) goto fail;
        const float *rr = rd.take(1);
        if (!rr) goto short_buffer;
        R.crr[d] = rr[0];
    }
    return 0;
short_buffer:
    err = "vertParams shorter than the serialized subpath";
fail:
    if (err.empty()) err = "vertParams shorter than the serialized subpath";
    return -1;
}

bool material_from_serialized(const float *b, bool twoSided, float *m) {
    std::memset(m, 0, 16 * sizeof(float));
    const int type = int(b[0]);
    if (float(type) != b[0] || type < 0 || type > 2) return false;
    m[0] = float(type);
    for (int k = 0; k < 3; k++) m[1 + k] = b[1 + k];  // Kd (Lambertian, Phong) or Ks (dielectric)
    if (type == 1) {
        for (int k = 0; k < 3; k++) m[4 + k] = b[4 + k];  // Ks
        m[7] = b[7];   // exponent
        m[11] = b[8];  // KsWeight
    } else if (type == 2) {
        for (int k = 0; k < 3; k++) m[4 + k] = b[4 + k];  // Kt
        m[9] = b[7];   // eta
        m[10] = b[8];  // invEta
        m[8] = b[9];   // alpha
    }
    m[12] = twoSided ? 1.0f : 0.0f;
    return true;
}

}  // namespace drmlt
