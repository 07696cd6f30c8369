// ORACLE (test infrastructure only). CPU restatement of the reference's scene,
// camera, triangle, BSDF and area-light code, templated on the scalar type so
// that the same text runs in float (as the reference ships) and in double (the
// parity reference). Nothing in the product links this file; it is compiled
// into oracle/liborc.so and loaded by tests/, bench.py's CPU baseline leg and
// the smoke check only.
//
// Each function cites the reference file:line it restates.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <limits>

namespace orc {

template <class F> struct V3 {
    F x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(F a, F b, F c) : x(a), y(b), z(c) {}
    F operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    V3 operator+(const V3 &o) const { return V3(x + o.x, y + o.y, z + o.z); }
    V3 operator-(const V3 &o) const { return V3(x - o.x, y - o.y, z - o.z); }
    V3 operator-() const { return V3(-x, -y, -z); }
    V3 operator*(F s) const { return V3(x * s, y * s, z * s); }
    V3 operator/(F s) const { return V3(x / s, y / s, z / s); }
    V3 mul(const V3 &o) const { return V3(x * o.x, y * o.y, z * o.z); }
    bool isZero() const { return x == F(0) && y == F(0) && z == F(0); }
    F maxCoeff() const { return std::max(x, std::max(y, z)); }
    F sum() const { return x + y + z; }
};
template <class F> inline V3<F> operator*(F s, const V3<F> &v) { return v * s; }
template <class F> inline F Dot(const V3<F> &a, const V3<F> &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class F> inline V3<F> Cross(const V3<F> &a, const V3<F> &b) {
    return V3<F>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <class F> inline F sq(F x) { return x * x; }
template <class F> inline F LengthSquared(const V3<F> &v) { return sq(v.x) + sq(v.y) + sq(v.z); }
template <class F> inline F Length(const V3<F> &v) { return std::sqrt(LengthSquared(v)); }
template <class F> inline V3<F> Normalize(const V3<F> &v) { F inv = F(1) / Length(v); return v * inv; }
template <class F> inline F DistanceSquared(const V3<F> &a, const V3<F> &b) { return LengthSquared(a - b); }
// utils.h:212-215
template <class F> inline F Luminance(const V3<F> &v) {
    return v.x * F(0.212671) + v.y * F(0.715160) + v.z * F(0.072169);
}
// utils.h:383-386
template <class F> inline F Modulo(F a, F b) { F r = std::fmod(a, b); return (r < F(0)) ? r + b : r; }
// utils.h:178-180
template <class F> inline V3<F> Reflect(const V3<F> &wi, const V3<F> &n) { return (F(2) * Dot(wi, n)) * n - wi; }
// utils.h:182-190
template <class F> inline V3<F> Refract(const V3<F> &wi, const V3<F> &n, F cosThetaT, F eta, F invEta) {
    F e = cosThetaT < F(0) ? invEta : eta;
    return n * (Dot(wi, n) * e + cosThetaT) - wi * e;
}
// utils.h:221-232 (Duff et al. orthonormal basis with the n.z ~ -1 special case)
template <class F> inline void CoordinateSystem(const V3<F> &n, V3<F> &b1, V3<F> &b2) {
    if (n.z < F(-1.0 + 1e-6)) {
        b1 = V3<F>(0, -1, 0);
        b2 = V3<F>(-1, 0, 0);
        return;
    }
    const F a = F(1) / (F(1) + n.z);
    const F b = -n.x * n.y * a;
    b1 = V3<F>(F(1) - sq(n.x) * a, b, -n.x);
    b2 = V3<F>(b, F(1) - sq(n.y) * a, -n.y);
}

template <class F> struct C {
    static constexpr F PI = F(3.14159265358979323846);
    static constexpr F INVPI = F(1) / PI;
    static constexpr F TWOPI = F(2) * PI;
    static constexpr F INVTWOPI = F(1) / TWOPI;
    static constexpr F CosEps = F(1e-4);  // commondef.h:62
};
// commondef.h:55-59: ray epsilons of the shipped (FP32) build. The oracle
// keeps the FP32 epsilons in both precisions so that it restates the shipped
// build's geometry; the -DDOUBLE_PRECISION epsilons are a separate build.
constexpr double kIsectEps = 1e-3;
constexpr double kShadowEps = 1e-3;

// sampling.h:6-15
template <class F> inline V3<F> SampleSphere(F u0, F u1, F &jac) {
    const F th = C<F>::TWOPI * u0, ph = C<F>::PI * u1;
    const F sp = std::sin(ph), cp = std::cos(ph);
    jac = std::fabs(sp);
    return V3<F>(sp * std::cos(th), sp * std::sin(th), cp);
}
// sampling.h:23-33
template <class F> inline F patan2(F y, F x) {
    if (y == F(0) && x == F(0)) return F(0);
    F r = std::atan2(y, x);
    if (r < F(0)) r += F(2.0 * M_PI);
    return r;
}
// sampling.h:35-41
template <class F> inline void ToSphericalCoord(const V3<F> &d, F &u0, F &u1, F &jac) {
    u0 = patan2(d.y, d.x) * C<F>::INVTWOPI;
    F ph = std::acos(d.z);
    jac = std::fabs(std::sin(ph));
    u1 = ph * C<F>::INVPI;
}
// sampling.h:99-106 (primal: ADEpsilon == 0)
template <class F> inline V3<F> SampleCosHemisphere(F u0, F u1) {
    F phi = C<F>::TWOPI * u0;
    F tmp = std::sqrt(std::max(F(1) - u1, F(0)));
    return V3<F>(std::cos(phi) * tmp, std::sin(phi) * tmp, std::sqrt(std::max(u1, F(0))));
}

// ---------------------------------------------------------------- materials
// Material record layout (16 floats), shared with the product's scene format:
// [0] type (0 Lambertian, 1 Phong, 2 RoughDielectric)
// [1..3] Kd (Lambertian, Phong) | Ks (RoughDielectric)
// [4..6] Ks (Phong) | Kt (RoughDielectric)
// [7] Phong exponent  [8] RoughDielectric alpha  [9] eta=intIOR/extIOR
// [10] 1/eta  [11] Phong KsWeight  [12] twoSided
enum { MAT_LAMBERT = 0, MAT_PHONG = 1, MAT_ROUGHDIEL = 2 };

template <class F> struct Material {
    int type;
    V3<F> a, b;
    F expo, alpha, eta, invEta, ksWeight;
    bool twoSided;
};

template <class F> struct BsdfOut {
    V3<F> contrib;
    F cosWo = 0, pdf = 0, revPdf = 0;
};

// Planted mistakes, for the tests that show each BSDF check fails on the
// mistake it is meant to catch (orc_plant; 0, the default, plants nothing):
// 1 Phong lobe normalised by (n + 1) / 2pi instead of (n + 2) / 2pi,
// 2 Phong lobe pdf normalised by (n + 2) / 2pi instead of (n + 1) / 2pi,
// 3 rough dielectric reflection without the Fresnel factor,
// 4 rough dielectric without the shadowing-masking term G,
// 5 rough dielectric transmission without eta'^2 in the numerator,
// 6 rough dielectric radiance transmission without the 1 / eta'^2 scaling,
// 7 rough dielectric transmission weighted by F instead of 1 - F.
inline int g_plant = 0;

// lambertian.cpp:15-38
template <class F> inline void LambertEval(const Material<F> &m, const V3<F> &wi, const V3<F> &n, const V3<F> &wo, BsdfOut<F> &o) {
    F cosWi = Dot(n, wi);
    V3<F> n_ = n;
    if (m.twoSided && cosWi < F(0)) { cosWi = -cosWi; n_ = -n_; }
    o.cosWo = Dot(n_, wo);
    o.contrib = V3<F>();
    if (cosWi < C<F>::CosEps || o.cosWo < C<F>::CosEps) return;
    F fwd = o.cosWo * C<F>::INVPI, rev = cosWi * C<F>::INVPI;
    o.contrib = fwd * m.a;
    o.pdf = fwd;
    o.revPdf = rev;
}
// lambertian.cpp:42-92
template <class F> inline bool LambertSample(const Material<F> &m, const V3<F> &wi, const V3<F> &n, F r0, F r1, V3<F> &wo, BsdfOut<F> &o) {
    F cosWi = Dot(wi, n);
    V3<F> n_ = n;
    if (std::fabs(cosWi) < C<F>::CosEps) return false;
    if (cosWi < F(0)) {
        if (m.twoSided) { cosWi = -cosWi; n_ = -n_; } else return false;
    }
    V3<F> b0, b1;
    CoordinateSystem(n_, b0, b1);
    V3<F> r = SampleCosHemisphere(r0, r1);
    wo = r.x * b0 + r.y * b1 + r.z * n_;
    o.cosWo = r.z;
    o.pdf = r.z * C<F>::INVPI;
    if (o.cosWo < C<F>::CosEps) return false;
    o.revPdf = cosWi * C<F>::INVPI;
    o.contrib = m.a;
    return true;
}

// phong.cpp:21-60
template <class F> inline void PhongEval(const Material<F> &m, const V3<F> &wi, const V3<F> &n, const V3<F> &wo, BsdfOut<F> &o) {
    o.contrib = V3<F>(); o.pdf = 0; o.revPdf = 0;
    F cosWi = Dot(n, wi);
    V3<F> n_ = n;
    if (m.twoSided && cosWi < F(0)) { cosWi = -cosWi; n_ = -n_; }
    o.cosWo = Dot(n_, wo);
    if (cosWi <= C<F>::CosEps || o.cosWo <= C<F>::CosEps) return;
    if (m.ksWeight > F(0)) {
        const F alpha = std::max(Dot(Reflect(wi, n_), wo), F(0));
        const F w = std::pow(alpha, m.expo) * C<F>::INVTWOPI;
        if (w > F(1e-10)) {
            o.contrib = m.b * ((m.expo + F(g_plant == 1 ? 1 : 2)) * w);
            o.pdf = m.ksWeight * (m.expo + F(g_plant == 2 ? 2 : 1)) * w;
            o.revPdf = o.pdf;
        }
    }
    if (m.ksWeight < F(1)) {
        o.pdf += (F(1) - m.ksWeight) * o.cosWo * C<F>::INVPI;
        o.revPdf += (F(1) - m.ksWeight) * cosWi * C<F>::INVPI;
        o.contrib = o.contrib + m.a * C<F>::INVPI;
    }
    o.contrib = o.contrib * o.cosWo;
    if (o.contrib.maxCoeff() < F(1e-10)) o.contrib = V3<F>();
}
// phong.cpp:62-146
template <class F> inline bool PhongSample(const Material<F> &m, const V3<F> &wi, const V3<F> &n, F r0, F r1, F uDisc, V3<F> &wo, BsdfOut<F> &o) {
    F cosWi = Dot(wi, n);
    if (std::fabs(cosWi) < C<F>::CosEps) return false;
    V3<F> n_ = n;
    if (cosWi < F(0)) {
        if (m.twoSided) { cosWi = -cosWi; n_ = -n_; } else return false;
    }
    const V3<F> R = Reflect(wi, n_);
    F g; V3<F> ax;
    if (uDisc > m.ksWeight) { g = F(1); ax = n_; } else { g = m.expo; ax = R; }
    const F power = F(1) / (g + F(1));
    const F cosA = std::pow(r1, power);
    const F sinA = std::sqrt(F(1) - sq(cosA));
    const F phi = C<F>::TWOPI * r0;
    V3<F> b0, b1;
    CoordinateSystem(ax, b0, b1);
    wo = (sinA * std::cos(phi)) * b0 + (sinA * std::sin(phi)) * b1 + cosA * ax;
    o.cosWo = Dot(n_, wo);
    if (o.cosWo < C<F>::CosEps) return false;
    o.contrib = V3<F>(); o.pdf = 0;
    if (m.ksWeight > F(0)) {
        const F alpha = std::max(Dot(R, wo), F(0));
        const F w = std::pow(alpha, m.expo) * C<F>::INVTWOPI;
        if (w > F(1e-10)) {
            o.contrib = m.b * ((m.expo + F(2)) * w);
            o.pdf = m.ksWeight * (m.expo + F(1)) * w;
        }
        o.revPdf = o.pdf;
    }
    if (m.ksWeight < F(1)) {
        o.contrib = o.contrib + m.a * C<F>::INVPI;
        o.pdf += (F(1) - m.ksWeight) * o.cosWo * C<F>::INVPI;
        o.revPdf += (F(1) - m.ksWeight) * cosWi * C<F>::INVPI;
    }
    o.contrib = o.contrib * o.cosWo;
    if (o.pdf < F(1e-10)) return false;
    o.contrib = o.contrib * (F(1) / o.pdf);
    return true;
}

// microfacet.h:6-18
template <class F> inline F BeckmannD(const V3<F> &h, F a) {
    F c2 = sq(h.z);
    F e = (sq(h.x) / sq(a) + sq(h.y) / sq(a)) / c2;
    return std::exp(-e) / (C<F>::PI * a * a * sq(c2));
}
// microfacet.h:20-37
template <class F> inline F BeckmannG1(F alpha, F cosT) {
    F tanT = std::sqrt(std::fabs(F(1) - sq(cosT))) / cosT;
    if (tanT <= F(0)) return F(1);
    F a = F(1) / (alpha * tanT);
    if (a >= F(1.6)) return F(1);
    F a2 = a * a;
    return (F(3.535) * a + F(2.181) * a2) / (F(1) + F(2.276) * a + F(2.577) * a2);
}
// microfacet.h:70-91
template <class F> inline F Fresnel(F cosI_, F &cosT_, F eta, F invEta) {
    F scale = (cosI_ > F(0)) ? invEta : eta;
    F ct2 = F(1) - (F(1) - sq(cosI_)) * sq(scale);
    if (ct2 <= F(0)) { cosT_ = F(0); return F(1); }
    F ci = std::fabs(cosI_), ct = std::sqrt(ct2);
    F Rs = (ci - eta * ct) / (ci + eta * ct);
    F Rp = (eta * ci - ct) / (eta * ci + ct);
    cosT_ = (cosI_ > F(0)) ? -ct : ct;
    return F(0.5) * (sq(Rs) + sq(Rp));
}
// microfacet.h:162-182
template <class F> inline V3<F> SampleMicronormal(F r0, F r1, F alpha, F &pdfW) {
    F phi = C<F>::TWOPI * r1;
    F a2 = sq(alpha);
    F tan2 = a2 * (-std::log(std::max(F(1) - r0, F(1e-6))));
    F cm = F(1) / std::sqrt(F(1) + tan2);
    F cm2 = sq(cm);
    pdfW = (F(1) - r0) / (C<F>::PI * a2 * cm * cm2);
    F sm = std::sqrt(std::max(F(1) - cm2, F(0)));
    return V3<F>(sm * std::cos(phi), sm * std::sin(phi), cm);
}
// roughdielectric.cpp:23-121
template <class F> inline void DielEval(const Material<F> &m, bool adjoint, const V3<F> &wi, const V3<F> &n, const V3<F> &wo, BsdfOut<F> &o) {
    F cosWi = Dot(wi, n);
    o.contrib = V3<F>(); o.cosWo = 0; o.pdf = 0; o.revPdf = 0;
    if (std::fabs(cosWi) < C<F>::CosEps) return;
    o.cosWo = Dot(wo, n);
    F cosWo = o.cosWo;
    if (std::fabs(cosWo) < C<F>::CosEps) return;
    bool reflect = cosWi * cosWo > F(0);
    F eta_ = cosWi > F(0) ? m.eta : m.invEta;
    F revEta_ = cosWo > F(0) ? m.eta : m.invEta;
    V3<F> H = reflect ? Normalize(wi + wo) : Normalize(wi + wo * eta_);
    if (Dot(H, n) < F(0)) H = -H;
    F cHi = Dot(wi, H), cHo = Dot(wo, H);
    if (std::fabs(cHi) < C<F>::CosEps || std::fabs(cHo) < C<F>::CosEps) return;
    if (cHi * cosWi <= F(0)) return;
    if (cHo * cosWo <= F(0)) return;
    V3<F> b0, b1;
    CoordinateSystem(n, b0, b1);
    V3<F> lh(Dot(b0, H), Dot(b1, H), Dot(n, H));
    F alp = m.alpha;
    F D = BeckmannD(lh, alp);
    if (D <= F(0)) return;
    F revCHi = cHo, revCHo = cHi;
    F dummy;
    F Fr = Fresnel(cHi, dummy, m.eta, m.invEta);
    F aCi = std::fabs(cosWi), aCo = std::fabs(cosWo);
    F G = g_plant == 4 ? F(1) : BeckmannG1(alp, aCi) * BeckmannG1(alp, aCo);
    F sa = alp * (F(1.2) - F(0.2) * std::sqrt(aCi));
    F prob = lh.z * BeckmannD(lh, sa);
    if (prob < F(1e-20)) { o.contrib = V3<F>(); return; }
    F rsa = alp * (F(1.2) - F(0.2) * std::sqrt(aCo));
    F revProb = lh.z * BeckmannD(lh, rsa);
    if (reflect) {
        F s = std::fabs((g_plant == 3 ? F(1) : Fr) * D * G / (F(4) * cosWi));
        o.contrib = m.a * s;
        o.pdf = std::fabs(prob * Fr / (F(4) * cHo));
        o.revPdf = std::fabs(revProb * Fr / (F(4) * revCHo));
    } else {
        F sd = cHi + eta_ * cHo;
        F rsd = revCHi + revEta_ * revCHo;
        F factor = (adjoint || g_plant == 6) ? F(1) : sq(F(1) / eta_);
        F tw = g_plant == 7 ? Fr : F(1) - Fr;
        F e2 = g_plant == 5 ? F(1) : sq(eta_);
        F s = std::fabs(factor * (tw * D * G * e2 * cHi * cHo) / (cosWi * sq(sd)));
        o.contrib = m.b * s;
        o.pdf = std::fabs(prob * (F(1) - Fr) * (sq(eta_) * cHo) / sq(sd));
        o.revPdf = std::fabs(revProb * (F(1) - Fr) * (sq(revEta_) * revCHo) / sq(rsd));
    }
}
// roughdielectric.cpp:146-253
template <class F> inline bool DielSample(const Material<F> &m, bool adjoint, const V3<F> &wi, const V3<F> &n, F r0, F r1, F uDisc, V3<F> &wo, BsdfOut<F> &o) {
    F cosWi = Dot(wi, n);
    if (std::fabs(cosWi) < C<F>::CosEps) return false;
    F alp = m.alpha;
    F sa = alp * (F(1.2) - F(0.2) * std::sqrt(std::fabs(cosWi)));
    F mPdf;
    V3<F> lh = SampleMicronormal(r0, r1, sa, mPdf);
    o.pdf = mPdf;
    V3<F> b0, b1;
    CoordinateSystem(n, b0, b1);
    V3<F> H = lh.x * b0 + lh.y * b1 + lh.z * n;
    F cHi = Dot(wi, H);
    if (std::fabs(cHi) < C<F>::CosEps) return false;
    F cosT = 0;
    F Fr = Fresnel(cHi, cosT, m.eta, m.invEta);
    bool reflect = uDisc <= Fr;
    V3<F> refl;
    F cHo;
    if (reflect) {
        wo = Reflect(wi, H);
        if (Fr <= F(0) || Dot(n, wo) * Dot(n, wi) <= F(0)) return false;
        refl = m.a;
        cHo = Dot(wo, H);
        o.pdf = std::fabs(o.pdf * Fr / (F(4) * cHo));
        F revCHo = cHi;
        F rev_dwh = F(1) / (F(4) * revCHo);
        o.cosWo = Dot(wo, n);
        if (std::fabs(o.cosWo) < C<F>::CosEps) return false;
        F rsa = alp * (F(1.2) - F(0.2) * std::sqrt(std::fabs(o.cosWo)));
        F revD = BeckmannD(lh, rsa);
        o.revPdf = std::fabs(Fr * revD * lh.z * rev_dwh);
    } else {
        wo = Refract(wi, H, cosT, m.eta, m.invEta);
        if (Fr >= F(1) || cosT == F(0) || Dot(n, wo) * Dot(n, wi) >= F(0)) return false;
        F eta_ = cosWi > F(0) ? m.eta : m.invEta;
        F factor = adjoint ? F(1) : sq(F(1) / eta_);
        refl = m.b * factor;
        cHo = Dot(wo, H);
        F sd = cHi + eta_ * cHo;
        F dwh = (sq(eta_) * cHo) / sq(sd);
        o.pdf = std::fabs(o.pdf * (F(1) - Fr) * std::fabs(dwh));
        o.cosWo = Dot(wo, n);
        if (std::fabs(o.cosWo) < C<F>::CosEps) return false;
        F revEta_ = o.cosWo > F(0) ? m.eta : m.invEta;
        F rsd = cHo + revEta_ * cHi;
        F rdwh = (sq(revEta_) * cHi) / sq(rsd);
        F rsa = alp * (F(1.2) - F(0.2) * std::sqrt(std::fabs(o.cosWo)));
        F revD = BeckmannD(lh, rsa);
        o.revPdf = std::fabs((F(1) - Fr) * revD * lh.z * rdwh);
    }
    if (std::fabs(cHo) < C<F>::CosEps) return false;
    if (o.pdf < F(1e-20)) return false;
    if (cHi * cosWi <= F(0)) return false;
    if (cHo * o.cosWo <= F(0)) return false;
    F aCi = std::fabs(cosWi), aCo = std::fabs(o.cosWo);
    F D = BeckmannD(lh, alp);
    F G = BeckmannG1(alp, aCi) * BeckmannG1(alp, aCo);
    o.contrib = refl * std::fabs((D * G * cHi) / (mPdf * aCi));
    return true;
}

template <class F> inline void BsdfEval(const Material<F> &m, bool adjoint, const V3<F> &wi, const V3<F> &n, const V3<F> &wo, BsdfOut<F> &o) {
    if (m.type == MAT_LAMBERT) LambertEval(m, wi, n, wo, o);
    else if (m.type == MAT_PHONG) PhongEval(m, wi, n, wo, o);
    else DielEval(m, adjoint, wi, n, wo, o);
}
template <class F> inline bool BsdfSample(const Material<F> &m, bool adjoint, const V3<F> &wi, const V3<F> &n, F r0, F r1, F uDisc, V3<F> &wo, BsdfOut<F> &o) {
    if (m.type == MAT_LAMBERT) return LambertSample(m, wi, n, r0, r1, wo, o);
    if (m.type == MAT_PHONG) return PhongSample(m, wi, n, r0, r1, uDisc, wo, o);
    return DielSample(m, adjoint, wi, n, r0, r1, uDisc, wo, o);
}
// lambertian.h:39-41; phong.cpp:148-155; roughdielectric.h:61-63
template <class F> inline F Roughness(const Material<F> &m, F uDisc) {
    if (m.type == MAT_LAMBERT) return F(1);
    if (m.type == MAT_PHONG) {
        if (uDisc < m.ksWeight) return std::sqrt(F(2) / (F(2) + F(4) * m.expo));
        return F(1);
    }
    return m.alpha;
}

}  // namespace orc
