// Forward-mode dual number with three tangents: instantiating the templated stylization
// nodes with it yields each node's 3x3 Jacobian with respect to its RGB input, which the
// adjoint kernels chain in reverse.
#pragma once
#include "style.hpp"

namespace nlpf {

struct Dual3 {
    float v, d[3];
    NLPF_HD Dual3() : v(0.0f), d{0.0f, 0.0f, 0.0f} {}
    NLPF_HD Dual3(float x) : v(x), d{0.0f, 0.0f, 0.0f} {}
    NLPF_HD Dual3(double x) : v((float)x), d{0.0f, 0.0f, 0.0f} {}
    NLPF_HD Dual3(int x) : v((float)x), d{0.0f, 0.0f, 0.0f} {}
    NLPF_HD Dual3(unsigned x) : v((float)x), d{0.0f, 0.0f, 0.0f} {}
    NLPF_HD static Dual3 var(float x, int i) {
        Dual3 r(x);
        r.d[i] = 1.0f;
        return r;
    }
};

NLPF_HD Dual3 operator+(Dual3 a, Dual3 b) {
    Dual3 r(a.v + b.v);
    for (int i = 0; i < 3; ++i) r.d[i] = a.d[i] + b.d[i];
    return r;
}
NLPF_HD Dual3 operator-(Dual3 a, Dual3 b) {
    Dual3 r(a.v - b.v);
    for (int i = 0; i < 3; ++i) r.d[i] = a.d[i] - b.d[i];
    return r;
}
NLPF_HD Dual3 operator-(Dual3 a) {
    Dual3 r(-a.v);
    for (int i = 0; i < 3; ++i) r.d[i] = -a.d[i];
    return r;
}
NLPF_HD Dual3 operator*(Dual3 a, Dual3 b) {
    Dual3 r(a.v * b.v);
    for (int i = 0; i < 3; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
    return r;
}
NLPF_HD Dual3 operator/(Dual3 a, Dual3 b) {
    const float inv = 1.0f / b.v;
    Dual3 r(a.v * inv);
    for (int i = 0; i < 3; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * inv;
    return r;
}
NLPF_HD bool operator<(Dual3 a, Dual3 b) { return a.v < b.v; }
NLPF_HD bool operator<=(Dual3 a, Dual3 b) { return a.v <= b.v; }
NLPF_HD bool operator>(Dual3 a, Dual3 b) { return a.v > b.v; }
NLPF_HD bool operator>=(Dual3 a, Dual3 b) { return a.v >= b.v; }
NLPF_HD bool operator==(Dual3 a, Dual3 b) { return a.v == b.v; }
NLPF_HD bool operator!=(Dual3 a, Dual3 b) { return a.v != b.v; }

// piecewise-constant functions have zero derivative almost everywhere
NLPF_HD Dual3 m_floor(Dual3 a) { return Dual3(floorf(a.v)); }
NLPF_HD Dual3 m_round(Dual3 a) { return Dual3(roundf(a.v)); }
NLPF_HD Dual3 m_fmod(Dual3 a, Dual3 b) {
    Dual3 r = a;
    const float q = truncf(a.v / b.v);
    r.v = fmodf(a.v, b.v);
    for (int i = 0; i < 3; ++i) r.d[i] = a.d[i] - q * b.d[i];
    return r;
}
NLPF_HD Dual3 m_sqrt(Dual3 a) {
    Dual3 r(sqrtf(a.v));
    const float k = r.v > 0.0f ? 0.5f / r.v : 0.0f;
    for (int i = 0; i < 3; ++i) r.d[i] = a.d[i] * k;
    return r;
}
NLPF_HD Dual3 m_pow(Dual3 a, Dual3 b) {
    Dual3 r(powf(a.v, b.v));
    const float da = a.v != 0.0f ? b.v * powf(a.v, b.v - 1.0f) : 0.0f;
    const float db = a.v > 0.0f ? r.v * logf(a.v) : 0.0f;
    for (int i = 0; i < 3; ++i) r.d[i] = da * a.d[i] + db * b.d[i];
    return r;
}

}  // namespace nlpf
