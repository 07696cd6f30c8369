// Device math for the DRMLT path kernels (gfx950): a 3-vector templated on the
// scalar type, the scalar overloads the path code calls, a second-order
// forward-mode jet for gradient/Hessian rows, and the Philox4x32-10 counter RNG.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define DEV __device__ __forceinline__
#define HD __host__ __device__ __forceinline__

namespace drmlt {

// ---------------------------------------------------------------- scalars
DEV float s_sqrt(float x) { return sqrtf(x); }
DEV float s_fabs(float x) { return fabsf(x); }
DEV float s_sin(float x) { return sinf(x); }
DEV float s_cos(float x) { return cosf(x); }
DEV float s_exp(float x) { return expf(x); }
DEV float s_log(float x) { return logf(x); }
DEV float s_pow(float x, float e) { return powf(x, e); }
DEV float s_max(float a, float b) { return a > b ? a : b; }  // fmax on non-NaN inputs
DEV double s_sqrt(double x) { return sqrt(x); }
DEV double s_fabs(double x) { return fabs(x); }
DEV double s_sin(double x) { return sin(x); }
DEV double s_cos(double x) { return cos(x); }
DEV double s_exp(double x) { return exp(x); }
DEV double s_log(double x) { return log(x); }
DEV double s_pow(double x, double e) { return pow(x, e); }
DEV double s_max(double a, double b) { return a > b ? a : b; }
DEV float val(float x) { return x; }
DEV double val(double x) { return x; }

// ---------------------------------------------------------------- jets
// Jet<S,K>: value v, tangent t = d/du_i, gradient block g[j] = d/du_j for the
// K columns j of one chunk, and h[j] = d2/(du_i du_j). One evaluation of the
// path with u_i seeded in t and u_j (j in the chunk) seeded in g yields the
// gradient block and K entries of Hessian row i (forward-over-forward).
template <class S, int K> struct Jet {
    S v, t;
    S g[K], h[K];
    DEV Jet() {}
    DEV Jet(S c) : v(c), t(0) {
#pragma unroll
        for (int j = 0; j < K; j++) { g[j] = 0; h[j] = 0; }
    }
};
template <class S, int K> DEV S val(const Jet<S, K> &a) { return a.v; }

template <class S, int K> DEV Jet<S, K> operator+(const Jet<S, K> &a, const Jet<S, K> &b) {
    Jet<S, K> r; r.v = a.v + b.v; r.t = a.t + b.t;
#pragma unroll
    for (int j = 0; j < K; j++) { r.g[j] = a.g[j] + b.g[j]; r.h[j] = a.h[j] + b.h[j]; }
    return r;
}
template <class S, int K> DEV Jet<S, K> operator-(const Jet<S, K> &a, const Jet<S, K> &b) {
    Jet<S, K> r; r.v = a.v - b.v; r.t = a.t - b.t;
#pragma unroll
    for (int j = 0; j < K; j++) { r.g[j] = a.g[j] - b.g[j]; r.h[j] = a.h[j] - b.h[j]; }
    return r;
}
template <class S, int K> DEV Jet<S, K> operator-(const Jet<S, K> &a) {
    Jet<S, K> r; r.v = -a.v; r.t = -a.t;
#pragma unroll
    for (int j = 0; j < K; j++) { r.g[j] = -a.g[j]; r.h[j] = -a.h[j]; }
    return r;
}
template <class S, int K> DEV Jet<S, K> operator*(const Jet<S, K> &a, const Jet<S, K> &b) {
    Jet<S, K> r; r.v = a.v * b.v; r.t = a.t * b.v + a.v * b.t;
#pragma unroll
    for (int j = 0; j < K; j++) {
        r.g[j] = a.g[j] * b.v + a.v * b.g[j];
        r.h[j] = a.h[j] * b.v + a.g[j] * b.t + a.t * b.g[j] + a.v * b.h[j];
    }
    return r;
}
template <class S, int K> DEV Jet<S, K> scale(const Jet<S, K> &a, S s) {
    Jet<S, K> r; r.v = a.v * s; r.t = a.t * s;
#pragma unroll
    for (int j = 0; j < K; j++) { r.g[j] = a.g[j] * s; r.h[j] = a.h[j] * s; }
    return r;
}
// f(a) given f(a.v), f'(a.v), f''(a.v)
template <class S, int K> DEV Jet<S, K> chain(const Jet<S, K> &a, S f0, S f1, S f2) {
    Jet<S, K> r; r.v = f0; r.t = f1 * a.t;
#pragma unroll
    for (int j = 0; j < K; j++) { r.g[j] = f1 * a.g[j]; r.h[j] = f2 * a.t * a.g[j] + f1 * a.h[j]; }
    return r;
}
template <class S, int K> DEV Jet<S, K> operator*(const Jet<S, K> &a, S s) { return scale(a, s); }
template <class S, int K> DEV Jet<S, K> operator*(S s, const Jet<S, K> &a) { return scale(a, s); }
template <class S, int K> DEV Jet<S, K> operator+(const Jet<S, K> &a, S s) { Jet<S, K> r = a; r.v += s; return r; }
template <class S, int K> DEV Jet<S, K> operator+(S s, const Jet<S, K> &a) { Jet<S, K> r = a; r.v += s; return r; }
template <class S, int K> DEV Jet<S, K> operator-(const Jet<S, K> &a, S s) { Jet<S, K> r = a; r.v -= s; return r; }
template <class S, int K> DEV Jet<S, K> operator-(S s, const Jet<S, K> &a) { Jet<S, K> r = -a; r.v += s; return r; }
template <class S, int K> DEV Jet<S, K> recip(const Jet<S, K> &a) {
    S i = S(1) / a.v;
    return chain(a, i, -i * i, S(2) * i * i * i);
}
template <class S, int K> DEV Jet<S, K> operator/(const Jet<S, K> &a, const Jet<S, K> &b) { return a * recip(b); }
template <class S, int K> DEV Jet<S, K> operator/(const Jet<S, K> &a, S s) { return scale(a, S(1) / s); }
template <class S, int K> DEV Jet<S, K> operator/(S s, const Jet<S, K> &a) { return scale(recip(a), s); }
template <class S, int K> DEV Jet<S, K> &operator+=(Jet<S, K> &a, const Jet<S, K> &b) { a = a + b; return a; }
template <class S, int K> DEV Jet<S, K> &operator*=(Jet<S, K> &a, const Jet<S, K> &b) { a = a * b; return a; }
template <class S, int K> DEV Jet<S, K> &operator*=(Jet<S, K> &a, S s) { a = scale(a, s); return a; }
template <class S, int K> DEV Jet<S, K> &operator-=(Jet<S, K> &a, const Jet<S, K> &b) { a = a - b; return a; }

template <class S, int K> DEV Jet<S, K> s_sqrt(const Jet<S, K> &a) {
    S r = s_sqrt(a.v);
    S d1 = S(0.5) / r;
    return chain(a, r, d1, -d1 / (S(2) * a.v));
}
// |x| takes the x >= 0 branch at 0, as chad does (chad.h:1241-1249)
template <class S, int K> DEV Jet<S, K> s_fabs(const Jet<S, K> &a) { return a.v >= S(0) ? a : -a; }
template <class S, int K> DEV Jet<S, K> s_sin(const Jet<S, K> &a) { S s = s_sin(a.v), c = s_cos(a.v); return chain(a, s, c, -s); }
template <class S, int K> DEV Jet<S, K> s_cos(const Jet<S, K> &a) { S s = s_sin(a.v), c = s_cos(a.v); return chain(a, c, -s, -c); }
template <class S, int K> DEV Jet<S, K> s_exp(const Jet<S, K> &a) { S e = s_exp(a.v); return chain(a, e, e, e); }
template <class S, int K> DEV Jet<S, K> s_log(const Jet<S, K> &a) { S i = S(1) / a.v; return chain(a, s_log(a.v), i, -i * i); }
template <class S, int K> DEV Jet<S, K> s_pow(const Jet<S, K> &a, S e) {
    S p = s_pow(a.v, e);
    S d1 = a.v != S(0) ? e * p / a.v : S(0);
    S d2 = a.v != S(0) ? (e - S(1)) * d1 / a.v : S(0);
    return chain(a, p, d1, d2);
}
template <class S, int K> DEV Jet<S, K> s_max(const Jet<S, K> &a, S b) { return a.v > b ? a : Jet<S, K>(b); }
template <class S, int K> DEV bool operator<(const Jet<S, K> &a, S b) { return a.v < b; }
template <class S, int K> DEV bool operator>(const Jet<S, K> &a, S b) { return a.v > b; }
template <class S, int K> DEV bool operator<=(const Jet<S, K> &a, S b) { return a.v <= b; }
template <class S, int K> DEV bool operator>=(const Jet<S, K> &a, S b) { return a.v >= b; }
template <class S, int K> DEV bool operator<(const Jet<S, K> &a, const Jet<S, K> &b) { return a.v < b.v; }
template <class S, int K> DEV bool operator>(const Jet<S, K> &a, const Jet<S, K> &b) { return a.v > b.v; }
template <class S, int K> DEV bool operator==(const Jet<S, K> &a, S b) { return a.v == b; }
template <class S, int K> DEV bool operator!=(const Jet<S, K> &a, S b) { return a.v != b; }

// scalar type traits: base precision of T
template <class T> struct Base { typedef T type; };
template <class S, int K> struct Base<Jet<S, K>> { typedef S type; };

// ---------------------------------------------------------------- vectors
template <class T> struct V3 {
    T x, y, z;
    DEV V3() {}
    DEV V3(T a, T b, T c) : x(a), y(b), z(c) {}
};
template <class T> DEV V3<T> operator+(const V3<T> &a, const V3<T> &b) { return V3<T>(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class T> DEV V3<T> operator-(const V3<T> &a, const V3<T> &b) { return V3<T>(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class T> DEV V3<T> operator-(const V3<T> &a) { return V3<T>(-a.x, -a.y, -a.z); }
template <class T> DEV V3<T> operator*(const V3<T> &a, const T &s) { return V3<T>(a.x * s, a.y * s, a.z * s); }
template <class T> DEV V3<T> operator*(const T &s, const V3<T> &a) { return V3<T>(a.x * s, a.y * s, a.z * s); }
template <class T> DEV V3<T> mul(const V3<T> &a, const V3<T> &b) { return V3<T>(a.x * b.x, a.y * b.y, a.z * b.z); }
template <class T> DEV T dot(const V3<T> &a, const V3<T> &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> DEV V3<T> cross(const V3<T> &a, const V3<T> &b) {
    return V3<T>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <class T> DEV T length2(const V3<T> &a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
template <class T> DEV V3<T> normalize(const V3<T> &a) { T inv = T(1) / s_sqrt(length2(a)); return a * inv; }
template <class T> DEV bool is_zero(const V3<T> &a) { return val(a.x) == 0 && val(a.y) == 0 && val(a.z) == 0; }
template <class T> DEV T luminance(const V3<T> &v) { return v.x * T(0.212671f) + v.y * T(0.715160f) + v.z * T(0.072169f); }
template <class T> DEV V3<T> vconst(float a, float b, float c) { return V3<T>(T(a), T(b), T(c)); }
template <class T> DEV V3<T> lift(const V3<float> &a) { return V3<T>(T(a.x), T(a.y), T(a.z)); }
DEV V3<float> lower(const V3<float> &a) { return a; }
DEV V3<float> lower(const V3<double> &a) { return V3<float>(float(a.x), float(a.y), float(a.z)); }
template <class S, int K> DEV V3<float> lower(const V3<Jet<S, K>> &a) { return V3<float>(float(a.x.v), float(a.y.v), float(a.z.v)); }

// ---------------------------------------------------------------- Philox4x32-10
// Salmon et al., "Parallel random numbers: as easy as 1, 2, 3" (SC'11).
// Key = 64-bit seed, counter = (block, iteration, chain, stream tag): every
// draw of every stage of every chain is addressable without stored RNG state.
struct Philox {
    uint32_t k0, k1, c0, c1, c2, c3;
    uint32_t buf[4];
    int have;
    DEV Philox(uint64_t seed, uint32_t iter, uint32_t chain, uint32_t tag)
        : k0(uint32_t(seed)), k1(uint32_t(seed >> 32)), c0(0), c1(iter), c2(chain), c3(tag), have(0) {}
    DEV void block() {
        uint32_t x0 = c0, x1 = c1, x2 = c2, x3 = c3, a = k0, b = k1;
#pragma unroll
        for (int r = 0; r < 10; r++) {
            uint32_t hi0 = __umulhi(0xD2511F53u, x0), lo0 = 0xD2511F53u * x0;
            uint32_t hi1 = __umulhi(0xCD9E8D57u, x2), lo1 = 0xCD9E8D57u * x2;
            uint32_t n0 = hi1 ^ x1 ^ a, n2 = hi0 ^ x3 ^ b;
            x0 = n0; x1 = lo1; x2 = n2; x3 = lo0;
            a += 0x9E3779B9u; b += 0xBB67AE85u;
        }
        buf[0] = x0; buf[1] = x1; buf[2] = x2; buf[3] = x3;
        have = 4;
        c0++;
    }
    DEV uint32_t next() {
        if (have == 0) block();
        return buf[--have];
    }
    DEV float uni() { return float(next() >> 8) * (1.0f / 16777216.0f); }
    DEV static double box_muller(uint32_t w1, uint32_t w2) {
        double u1 = (double(w1 >> 8) + 1.0) * (1.0 / 16777216.0);
        double u2 = double(w2 >> 8) * (1.0 / 16777216.0);
        return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    }
    DEV double normal() {
        const uint32_t w1 = next();
        return box_muller(w1, next());
    }
    // The value the j-th normal() call returns on this stream when it is at a
    // block boundary (have == 0): normal() takes two words, next() hands out
    // words 3, 2, 1, 0 of each block.
    DEV double normal_at(int j) const {
        Philox t = *this;
        t.c0 = c0 + uint32_t(j / 2);
        t.block();
        return (j & 1) ? box_muller(t.buf[1], t.buf[0]) : box_muller(t.buf[3], t.buf[2]);
    }
    // Advances a stream at a block boundary past `count` normal() calls.
    DEV void skip_normals(int count) {
        const int words = 2 * count;
        c0 += uint32_t(words / 4);
        have = 0;
        if (words % 4) {
            block();
            have -= words % 4;
        }
    }
};

constexpr float kPI = 3.14159265358979323846f;
constexpr float kINVPI = 1.0f / kPI;
constexpr float kTWOPI = 2.0f * kPI;
constexpr float kINVTWOPI = 1.0f / kTWOPI;
constexpr float kCosEps = 1e-4f;      // commondef.h:62
constexpr float kIsectEps = 1e-3f;    // commondef.h:56
constexpr float kShadowEps = 1e-3f;   // commondef.h:57

}  // namespace drmlt
