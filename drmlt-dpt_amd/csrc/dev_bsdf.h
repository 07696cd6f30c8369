// Device BSDFs, templated on the scalar type (float for the primal, jets for
// derivatives): Lambertian (lambertian.cpp:15-92), normalised modified Phong
// (phong.cpp:21-166) and the rough dielectric of Walter et al. 2007 with the
// Beckmann distribution and the rational shadowing approximation
// (roughdielectric.cpp:23-253, microfacet.h:6-182). The material record is 16
// floats: type, Kd|Ks, Ks|Kt, exponent, alpha, eta, 1/eta, KsWeight, twoSided.
#pragma once
#include "dev_math.h"

namespace drmlt {

enum { MAT_LAMBERT = 0, MAT_PHONG = 1, MAT_ROUGHDIEL = 2 };

struct Mat {
    int type;
    float a[3], b[3];
    float expo, alpha, eta, invEta, ksWeight;
    bool twoSided;
};
DEV Mat load_mat(const float *m) {
    Mat r;
    r.type = int(m[0]);
    r.a[0] = m[1]; r.a[1] = m[2]; r.a[2] = m[3];
    r.b[0] = m[4]; r.b[1] = m[5]; r.b[2] = m[6];
    r.expo = m[7]; r.alpha = m[8]; r.eta = m[9]; r.invEta = m[10]; r.ksWeight = m[11];
    r.twoSided = m[12] != 0.0f;
    return r;
}

template <class T> struct BOut {
    V3<T> contrib;
    T cosWo, pdf, revPdf;
};

// utils.h:221-232
template <class T> DEV void coord_system(const V3<T> &n, V3<T> &b1, V3<T> &b2) {
    if (val(n.z) < float(-1.0 + 1e-6)) {
        b1 = vconst<T>(0, -1, 0);
        b2 = vconst<T>(-1, 0, 0);
        return;
    }
    const T a = T(1) / (T(1) + n.z);
    const T b = -n.x * n.y * a;
    b1 = V3<T>(T(1) - n.x * n.x * a, b, -n.x);
    b2 = V3<T>(b, T(1) - n.y * n.y * a, -n.y);
}
template <class T> DEV V3<T> reflect(const V3<T> &wi, const V3<T> &n) { return (T(2) * dot(wi, n)) * n - wi; }
// sampling.h:99-106
template <class T> DEV V3<T> sample_cos_hemisphere(const T &u0, const T &u1) {
    T phi = T(kTWOPI) * u0;
    T tmp = s_sqrt(s_max(T(1) - u1, typename Base<T>::type(0)));
    return V3<T>(s_cos(phi) * tmp, s_sin(phi) * tmp, s_sqrt(s_max(u1, typename Base<T>::type(0))));
}
// sampling.h:6-15
template <class T> DEV V3<T> sample_sphere(const T &u0, const T &u1, T &jac) {
    const T th = T(kTWOPI) * u0, ph = T(kPI) * u1;
    const T sp = s_sin(ph), cp = s_cos(ph);
    jac = s_fabs(sp);
    return V3<T>(sp * s_cos(th), sp * s_sin(th), cp);
}
// sampling.h:23-41 (float only)
DEV void to_spherical(const V3<float> &d, float &u0, float &u1, float &jac) {
    float th = (d.y == 0.0f && d.x == 0.0f) ? 0.0f : atan2f(d.y, d.x);
    if (th < 0.0f) th += float(2.0 * 3.14159265358979323846);
    u0 = th * kINVTWOPI;
    float ph = acosf(d.z);
    jac = fabsf(sinf(ph));
    u1 = ph * kINVPI;
}

template <class T> DEV V3<T> mat_a(const Mat &m) { return vconst<T>(m.a[0], m.a[1], m.a[2]); }
template <class T> DEV V3<T> mat_b(const Mat &m) { return vconst<T>(m.b[0], m.b[1], m.b[2]); }

template <class T> DEV void lambert_eval(const Mat &m, const V3<T> &wi, const V3<T> &n, const V3<T> &wo, BOut<T> &o) {
    T cosWi = dot(n, wi);
    V3<T> n_ = n;
    if (m.twoSided && val(cosWi) < 0) { cosWi = -cosWi; n_ = -n_; }
    o.cosWo = dot(n_, wo);
    o.contrib = vconst<T>(0, 0, 0);
    if (val(cosWi) < kCosEps || val(o.cosWo) < kCosEps) return;
    T fwd = o.cosWo * T(kINVPI), rev = cosWi * T(kINVPI);
    o.contrib = mat_a<T>(m) * fwd;
    o.pdf = fwd;
    o.revPdf = rev;
}
template <class T> DEV bool lambert_sample(const Mat &m, const V3<T> &wi, const V3<T> &n, const T &r0, const T &r1, V3<T> &wo, BOut<T> &o) {
    T cosWi = dot(wi, n);
    V3<T> n_ = n;
    if (s_fabs(val(cosWi)) < kCosEps) return false;
    if (val(cosWi) < 0) {
        if (m.twoSided) { cosWi = -cosWi; n_ = -n_; } else return false;
    }
    V3<T> b0, b1;
    coord_system(n_, b0, b1);
    V3<T> r = sample_cos_hemisphere(r0, r1);
    wo = r.x * b0 + r.y * b1 + r.z * n_;
    o.cosWo = r.z;
    o.pdf = r.z * T(kINVPI);
    if (val(o.cosWo) < kCosEps) return false;
    o.revPdf = cosWi * T(kINVPI);
    o.contrib = mat_a<T>(m);
    return true;
}

template <class T> DEV void phong_eval(const Mat &m, const V3<T> &wi, const V3<T> &n, const V3<T> &wo, BOut<T> &o) {
    o.contrib = vconst<T>(0, 0, 0); o.pdf = T(0); o.revPdf = T(0);
    T cosWi = dot(n, wi);
    V3<T> n_ = n;
    if (m.twoSided && val(cosWi) < 0) { cosWi = -cosWi; n_ = -n_; }
    o.cosWo = dot(n_, wo);
    if (val(cosWi) <= kCosEps || val(o.cosWo) <= kCosEps) return;
    if (m.ksWeight > 0.0f) {
        const T alpha = s_max(dot(reflect(wi, n_), wo), typename Base<T>::type(0));
        const T w = s_pow(alpha, typename Base<T>::type(m.expo)) * T(kINVTWOPI);
        if (val(w) > 1e-10f) {
            o.contrib = mat_b<T>(m) * ((T(m.expo) + T(2)) * w);
            o.pdf = T(m.ksWeight) * (T(m.expo) + T(1)) * w;
            o.revPdf = o.pdf;
        }
    }
    if (m.ksWeight < 1.0f) {
        o.pdf = o.pdf + T(1.0f - m.ksWeight) * o.cosWo * T(kINVPI);
        o.revPdf = o.revPdf + T(1.0f - m.ksWeight) * cosWi * T(kINVPI);
        o.contrib = o.contrib + mat_a<T>(m) * T(kINVPI);
    }
    o.contrib = o.contrib * o.cosWo;
    float mx = fmaxf(float(val(o.contrib.x)), fmaxf(float(val(o.contrib.y)), float(val(o.contrib.z))));
    if (mx < 1e-10f) o.contrib = vconst<T>(0, 0, 0);
}
template <class T> DEV bool phong_sample(const Mat &m, const V3<T> &wi, const V3<T> &n, const T &r0, const T &r1, float uDisc, V3<T> &wo,
                                         BOut<T> &o) {
    T cosWi = dot(wi, n);
    if (s_fabs(val(cosWi)) < kCosEps) return false;
    V3<T> n_ = n;
    if (val(cosWi) < 0) {
        if (m.twoSided) { cosWi = -cosWi; n_ = -n_; } else return false;
    }
    const V3<T> R = reflect(wi, n_);
    float g;
    V3<T> ax;
    if (uDisc > m.ksWeight) { g = 1.0f; ax = n_; } else { g = m.expo; ax = R; }
    const float power = 1.0f / (g + 1.0f);
    const T cosA = s_pow(r1, typename Base<T>::type(power));
    const T sinA = s_sqrt(T(1) - cosA * cosA);
    const T phi = T(kTWOPI) * r0;
    V3<T> b0, b1;
    coord_system(ax, b0, b1);
    wo = (sinA * s_cos(phi)) * b0 + (sinA * s_sin(phi)) * b1 + cosA * ax;
    o.cosWo = dot(n_, wo);
    if (val(o.cosWo) < kCosEps) return false;
    o.contrib = vconst<T>(0, 0, 0); o.pdf = T(0); o.revPdf = T(0);
    if (m.ksWeight > 0.0f) {
        const T alpha = s_max(dot(R, wo), typename Base<T>::type(0));
        const T w = s_pow(alpha, typename Base<T>::type(m.expo)) * T(kINVTWOPI);
        if (val(w) > 1e-10f) {
            o.contrib = mat_b<T>(m) * ((T(m.expo) + T(2)) * w);
            o.pdf = T(m.ksWeight) * (T(m.expo) + T(1)) * w;
        }
        o.revPdf = o.pdf;
    }
    if (m.ksWeight < 1.0f) {
        o.contrib = o.contrib + mat_a<T>(m) * T(kINVPI);
        o.pdf = o.pdf + T(1.0f - m.ksWeight) * o.cosWo * T(kINVPI);
        o.revPdf = o.revPdf + T(1.0f - m.ksWeight) * cosWi * T(kINVPI);
    }
    o.contrib = o.contrib * o.cosWo;
    if (val(o.pdf) < 1e-10f) return false;
    o.contrib = o.contrib * (T(1) / o.pdf);
    return true;
}

// microfacet.h:6-18
template <class T> DEV T beckmann_d(const V3<T> &h, const T &a) {
    T c2 = h.z * h.z;
    T e = (h.x * h.x / (a * a) + h.y * h.y / (a * a)) / c2;
    return s_exp(-e) / (T(kPI) * a * a * c2 * c2);
}
// microfacet.h:20-37
template <class T> DEV T beckmann_g1(float alpha, const T &cosT) {
    T tanT = s_sqrt(s_fabs(T(1) - cosT * cosT)) / cosT;
    if (val(tanT) <= 0) return T(1);
    T a = T(1) / (T(alpha) * tanT);
    if (val(a) >= 1.6f) return T(1);
    T a2 = a * a;
    return (T(3.535f) * a + T(2.181f) * a2) / (T(1) + T(2.276f) * a + T(2.577f) * a2);
}
// microfacet.h:70-91
template <class T> DEV T fresnel(const T &cosI_, T &cosT_, float eta, float invEta) {
    float scale = val(cosI_) > 0 ? invEta : eta;
    T ct2 = T(1) - (T(1) - cosI_ * cosI_) * T(scale * scale);
    if (val(ct2) <= 0) { cosT_ = T(0); return T(1); }
    T ci = s_fabs(cosI_), ct = s_sqrt(ct2);
    T Rs = (ci - T(eta) * ct) / (ci + T(eta) * ct);
    T Rp = (T(eta) * ci - ct) / (T(eta) * ci + ct);
    cosT_ = val(cosI_) > 0 ? -ct : ct;
    return T(0.5f) * (Rs * Rs + Rp * Rp);
}
// microfacet.h:162-182
template <class T> DEV V3<T> sample_micronormal(const T &r0, const T &r1, const T &alpha, T &pdfW) {
    T phi = T(kTWOPI) * r1;
    T a2 = alpha * alpha;
    T lg = val(T(1) - r0) > 1e-6f ? s_log(T(1) - r0) : T(s_log(typename Base<T>::type(1e-6f)));
    T tan2 = a2 * (-lg);
    T cm = T(1) / s_sqrt(T(1) + tan2);
    T cm2 = cm * cm;
    pdfW = (T(1) - r0) / (T(kPI) * a2 * cm * cm2);
    T sm = s_sqrt(s_max(T(1) - cm2, typename Base<T>::type(0)));
    return V3<T>(sm * s_cos(phi), sm * s_sin(phi), cm);
}
// roughdielectric.cpp:23-121
template <class T> DEV void diel_eval(const Mat &m, bool adjoint, const V3<T> &wi, const V3<T> &n, const V3<T> &wo, BOut<T> &o) {
    T cosWi = dot(wi, n);
    o.contrib = vconst<T>(0, 0, 0); o.cosWo = T(0); o.pdf = T(0); o.revPdf = T(0);
    if (s_fabs(val(cosWi)) < kCosEps) return;
    o.cosWo = dot(wo, n);
    const T cosWo = o.cosWo;
    if (s_fabs(val(cosWo)) < kCosEps) return;
    const bool refl = val(cosWi) * val(cosWo) > 0;
    const float eta_ = val(cosWi) > 0 ? m.eta : m.invEta;
    const float revEta_ = val(cosWo) > 0 ? m.eta : m.invEta;
    V3<T> H = refl ? normalize(wi + wo) : normalize(wi + wo * T(eta_));
    if (val(dot(H, n)) < 0) H = -H;
    const T cHi = dot(wi, H), cHo = dot(wo, H);
    if (s_fabs(val(cHi)) < kCosEps || s_fabs(val(cHo)) < kCosEps) return;
    if (val(cHi) * val(cosWi) <= 0) return;
    if (val(cHo) * val(cosWo) <= 0) return;
    V3<T> b0, b1;
    coord_system(n, b0, b1);
    V3<T> lh(dot(b0, H), dot(b1, H), dot(n, H));
    const float alp = m.alpha;
    T D = beckmann_d(lh, T(alp));
    if (val(D) <= 0) return;
    const T revCHi = cHo, revCHo = cHi;
    T dummy;
    T Fr = fresnel(cHi, dummy, m.eta, m.invEta);
    T aCi = s_fabs(cosWi), aCo = s_fabs(cosWo);
    T G = beckmann_g1(alp, aCi) * beckmann_g1(alp, aCo);
    T sa = T(alp) * (T(1.2f) - T(0.2f) * s_sqrt(aCi));
    T prob = lh.z * beckmann_d(lh, sa);
    if (val(prob) < 1e-20f) { o.contrib = vconst<T>(0, 0, 0); return; }
    T rsa = T(alp) * (T(1.2f) - T(0.2f) * s_sqrt(aCo));
    T revProb = lh.z * beckmann_d(lh, rsa);
    if (refl) {
        T s = s_fabs(Fr * D * G / (T(4) * cosWi));
        o.contrib = mat_a<T>(m) * s;
        o.pdf = s_fabs(prob * Fr / (T(4) * cHo));
        o.revPdf = s_fabs(revProb * Fr / (T(4) * revCHo));
    } else {
        T sd = cHi + T(eta_) * cHo;
        T rsd = revCHi + T(revEta_) * revCHo;
        float factor = adjoint ? 1.0f : (1.0f / eta_) * (1.0f / eta_);
        T s = s_fabs(T(factor) * ((T(1) - Fr) * D * G * T(eta_ * eta_) * cHi * cHo) / (cosWi * sd * sd));
        o.contrib = mat_b<T>(m) * s;
        o.pdf = s_fabs(prob * (T(1) - Fr) * (T(eta_ * eta_) * cHo) / (sd * sd));
        o.revPdf = s_fabs(revProb * (T(1) - Fr) * (T(revEta_ * revEta_) * revCHo) / (rsd * rsd));
    }
}
// roughdielectric.cpp:146-253
template <class T> DEV bool diel_sample(const Mat &m, bool adjoint, const V3<T> &wi, const V3<T> &n, const T &r0, const T &r1, float uDisc,
                                        V3<T> &wo, BOut<T> &o) {
    T cosWi = dot(wi, n);
    if (s_fabs(val(cosWi)) < kCosEps) return false;
    const float alp = m.alpha;
    T sa = T(alp) * (T(1.2f) - T(0.2f) * s_sqrt(s_fabs(cosWi)));
    T mPdf;
    V3<T> lh = sample_micronormal(r0, r1, sa, mPdf);
    o.pdf = mPdf;
    V3<T> b0, b1;
    coord_system(n, b0, b1);
    V3<T> H = lh.x * b0 + lh.y * b1 + lh.z * n;
    T cHi = dot(wi, H);
    if (s_fabs(val(cHi)) < kCosEps) return false;
    T cosT = T(0);
    T Fr = fresnel(cHi, cosT, m.eta, m.invEta);
    const bool refl = uDisc <= float(val(Fr));
    V3<T> rf;
    T cHo;
    if (refl) {
        wo = reflect(wi, H);
        if (val(Fr) <= 0 || val(dot(n, wo)) * val(dot(n, wi)) <= 0) return false;
        rf = mat_a<T>(m);
        cHo = dot(wo, H);
        o.pdf = s_fabs(o.pdf * Fr / (T(4) * cHo));
        T revCHo = cHi;
        T rev_dwh = T(1) / (T(4) * revCHo);
        o.cosWo = dot(wo, n);
        if (s_fabs(val(o.cosWo)) < kCosEps) return false;
        T rsa = T(alp) * (T(1.2f) - T(0.2f) * s_sqrt(s_fabs(o.cosWo)));
        T revD = beckmann_d(lh, rsa);
        o.revPdf = s_fabs(Fr * revD * lh.z * rev_dwh);
    } else {
        const float e = val(cosT) < 0 ? m.invEta : m.eta;
        wo = H * (dot(wi, H) * T(e) + cosT) - wi * T(e);  // Refract(wi, H, ...), utils.h:182-190
        if (val(Fr) >= 1 || val(cosT) == 0 || val(dot(n, wo)) * val(dot(n, wi)) >= 0) return false;
        const float eta_ = val(cosWi) > 0 ? m.eta : m.invEta;
        const float factor = adjoint ? 1.0f : (1.0f / eta_) * (1.0f / eta_);
        rf = mat_b<T>(m) * T(factor);
        cHo = dot(wo, H);
        T sd = cHi + T(eta_) * cHo;
        T dwh = (T(eta_ * eta_) * cHo) / (sd * sd);
        o.pdf = s_fabs(o.pdf * (T(1) - Fr) * s_fabs(dwh));
        o.cosWo = dot(wo, n);
        if (s_fabs(val(o.cosWo)) < kCosEps) return false;
        const float revEta_ = val(o.cosWo) > 0 ? m.eta : m.invEta;
        T rsd = cHo + T(revEta_) * cHi;
        T rdwh = (T(revEta_ * revEta_) * cHi) / (rsd * rsd);
        T rsa = T(alp) * (T(1.2f) - T(0.2f) * s_sqrt(s_fabs(o.cosWo)));
        T revD = beckmann_d(lh, rsa);
        o.revPdf = s_fabs((T(1) - Fr) * revD * lh.z * rdwh);
    }
    if (s_fabs(val(cHo)) < kCosEps) return false;
    if (val(o.pdf) < 1e-20f) return false;
    if (val(cHi) * val(cosWi) <= 0) return false;
    if (val(cHo) * val(o.cosWo) <= 0) return false;
    T aCi = s_fabs(cosWi), aCo = s_fabs(o.cosWo);
    T D = beckmann_d(lh, T(alp));
    T G = beckmann_g1(alp, aCi) * beckmann_g1(alp, aCo);
    o.contrib = rf * s_fabs((D * G * cHi) / (mPdf * aCi));
    return true;
}

template <class T> DEV void bsdf_eval(const Mat &m, bool adjoint, const V3<T> &wi, const V3<T> &n, const V3<T> &wo, BOut<T> &o) {
    if (m.type == MAT_LAMBERT) lambert_eval(m, wi, n, wo, o);
    else if (m.type == MAT_PHONG) phong_eval(m, wi, n, wo, o);
    else diel_eval(m, adjoint, wi, n, wo, o);
}
template <class T> DEV bool bsdf_sample(const Mat &m, bool adjoint, const V3<T> &wi, const V3<T> &n, const T &r0, const T &r1, float uDisc,
                                        V3<T> &wo, BOut<T> &o) {
    if (m.type == MAT_LAMBERT) return lambert_sample(m, wi, n, r0, r1, wo, o);
    if (m.type == MAT_PHONG) return phong_sample(m, wi, n, r0, r1, uDisc, wo, o);
    return diel_sample(m, adjoint, wi, n, r0, r1, uDisc, wo, o);
}
// lambertian.h:39-41, phong.cpp:148-155, roughdielectric.h:61-63
DEV float roughness(const Mat &m, float uDisc) {
    if (m.type == MAT_LAMBERT) return 1.0f;
    if (m.type == MAT_PHONG) return uDisc < m.ksWeight ? sqrtf(2.0f / (2.0f + 4.0f * m.expo)) : 1.0f;
    return m.alpha;
}

}  // namespace drmlt
