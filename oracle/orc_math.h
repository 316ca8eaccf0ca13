// CPU oracle -- TEST INFRASTRUCTURE ONLY.
//
// Scalar/vector arithmetic of the oracle.  Everything here exists so that the
// oracle evaluates the reference shader's expressions in a fixed, documented
// order (no FMA contraction: this file is compiled with -ffp-contract=off) and
// with transcendental functions whose float results do not depend on the
// platform libm.
//
// Float precision ("fp32 oracle"): sin/cos/acos/atan2 follow the published
// single-precision Cephes algorithms (S. L. Moshier, Cephes Math Library,
// sinf.c / asinf.c / atanf.c: Cody-Waite reduction by pi/4 with a three-part
// constant, then the minimax polynomials of that library).  They are accurate
// to ~1-2 ulp and are pinned against libm in tests/test_oracle_math.py.
//
// Double precision ("fp64 convergence oracle"): std:: libm.
//
// Vector ops follow HLSL/Slang semantics; dot() sums left to right
// ((x*x + y*y) + z*z); lerp(a, b, t) = a + t*(b - a).
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>

namespace orc {

// ---- dual numbers (fp64 tangent-linear mode) ------------------------------
// Dual{v, d}: value and derivative along one parameter direction.  Arithmetic
// applies the chain rule; comparisons and casts use the value only, so a
// PT<Dual> run takes the decisions of the fp64 run and returns the exact
// derivative of the estimator with its random numbers fixed.
struct Dual {
    double v = 0.0, d = 0.0;
    constexpr Dual() = default;
    constexpr Dual(double x) : v(x), d(0.0) {}
    constexpr Dual(double x, double dx) : v(x), d(dx) {}
    explicit operator double() const { return v; }
    explicit operator float() const { return (float)v; }
    explicit operator int() const { return (int)v; }
    explicit operator uint32_t() const { return (uint32_t)v; }
};
inline Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
inline Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
inline Dual operator-(Dual a) { return {-a.v, -a.d}; }
inline Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
inline Dual operator/(Dual a, Dual b) { return {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
inline Dual& operator*=(Dual& a, Dual b) { a = a * b; return a; }
inline Dual& operator+=(Dual& a, Dual b) { a = a + b; return a; }
inline bool operator<(Dual a, Dual b) { return a.v < b.v; }
inline bool operator>(Dual a, Dual b) { return a.v > b.v; }
inline bool operator<=(Dual a, Dual b) { return a.v <= b.v; }
inline bool operator>=(Dual a, Dual b) { return a.v >= b.v; }
inline bool operator==(Dual a, Dual b) { return a.v == b.v; }
inline bool operator!=(Dual a, Dual b) { return a.v != b.v; }
inline Dual osqrt(Dual a) {
    const double s = std::sqrt(a.v);
    return {s, s > 0.0 ? a.d / (2.0 * s) : 0.0};
}
inline Dual oabs(Dual a) { return a.v < 0.0 ? -a : a; }
inline Dual ofloor(Dual a) { return {std::floor(a.v), 0.0}; }
inline float osqrt(float a) { return std::sqrt(a); }
inline double osqrt(double a) { return std::sqrt(a); }
inline float oabs(float a) { return std::fabs(a); }
inline double oabs(double a) { return std::fabs(a); }
inline float ofloor(float a) { return std::floor(a); }
inline double ofloor(double a) { return std::floor(a); }
inline double value_of(float a) { return a; }
inline double value_of(double a) { return a; }
inline double value_of(Dual a) { return a.v; }

template <class R> struct V3 {
    R x, y, z;
};
template <class R> inline V3<R> v3(R x, R y, R z) { return V3<R>{x, y, z}; }
template <class R> inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> inline V3<R> operator-(V3<R> a) { return {-a.x, -a.y, -a.z}; }
template <class R> inline V3<R> operator*(V3<R> a, V3<R> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class R> inline V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> inline V3<R> operator*(R s, V3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class R> inline V3<R> operator/(V3<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }
template <class R> inline V3<R> operator/(V3<R> a, V3<R> b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
template <class R> inline R dot(V3<R> a, V3<R> b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
template <class R> inline V3<R> cross(V3<R> a, V3<R> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class R> inline R length(V3<R> a) { return osqrt(dot(a, a)); }
// normalize(v) = v * (1 / sqrt(dot(v, v))): one correctly rounded sqrt, one
// correctly rounded division and three multiplies.
template <class R> inline V3<R> normalize(V3<R> a) {
    R inv = R(1) / osqrt(dot(a, a));
    return a * inv;
}
template <class R> inline V3<R> reflect(V3<R> i, V3<R> n) {
    R d = dot(n, i);
    return i - n * (R(2) * d);
}
// HLSL refract(): zero vector on total internal reflection.
template <class R> inline V3<R> refract(V3<R> i, V3<R> n, R eta) {
    R ni = dot(n, i);
    R k = R(1) - (eta * eta) * (R(1) - ni * ni);
    if (k < R(0)) return {R(0), R(0), R(0)};
    return i * eta - n * (eta * ni + osqrt(k));
}
template <class R> inline R maxc(V3<R> a) { return std::max(a.x, std::max(a.y, a.z)); }
template <class R> inline bool any_pos(V3<R> a) { return a.x > R(0) || a.y > R(0) || a.z > R(0); }
template <class R> inline R clampv(R x, R lo, R hi) { return std::min(std::max(x, lo), hi); }
template <class R> inline V3<R> lerp3(V3<R> a, V3<R> b, R t) { return a + (b - a) * t; }
// The shutter-time lerp of motion blur, in the defining form of GLSL's mix,
// a (1 - t) + b t: at t = 1 it returns b exactly, which the reference relies
// on for bit-identical images with motion blur off (PathTrace.slang:2048-2053).
template <class R> inline V3<R> mix3(V3<R> a, V3<R> b, R t) {
    const R u = R(1) - t;
    return {a.x * u + b.x * t, a.y * u + b.y * t, a.z * u + b.z * t};
}
// quatRotate (PathTrace.slang:2028-2031): v + 2 * cross(u, cross(u, v) + w * v).
template <class R> inline V3<R> quat_rotate(const R q[4], V3<R> v) {
    V3<R> u{q[0], q[1], q[2]};
    V3<R> c = cross(u, cross(u, v) + v * q[3]);
    return v + c * R(2);
}

// ---- transcendental functions -------------------------------------------

// Cephes logf (logf.c): x = m 2^e, m in [sqrt(1/2), sqrt(2)), the degree-9
// minimax polynomial of log(1 + f), ln 2 in two parts; and pow(x, y) for
// x >= 0 as exp(y log x) from it.
inline float cephes_logf(float xin) {
    if (!(xin > 0.0f)) return -INFINITY;
    if (std::isinf(xin)) return xin;
    uint32_t bits;
    std::memcpy(&bits, &xin, 4);
    int e = 0;
    if (bits < 0x00800000u) {
        xin = xin * 8388608.0f;
        std::memcpy(&bits, &xin, 4);
        e = -23;
    }
    e += (int)((bits >> 23) & 0xFFu) - 126;
    const uint32_t mb = (bits & 0x007FFFFFu) | 0x3F000000u;
    float x;
    std::memcpy(&x, &mb, 4);
    if (x < 0.707106781186547524f) {
        e -= 1;
        x = (x + x) - 1.0f;
    } else {
        x = x - 1.0f;
    }
    const float z = x * x;
    float y = ((((((((7.0376836292e-2f * x - 1.1514610310e-1f) * x + 1.1676998740e-1f) * x - 1.2420140846e-1f) * x +
                   1.4249322787e-1f) * x - 1.6668057665e-1f) * x + 2.0000714765e-1f) * x - 2.4999993993e-1f) * x +
               3.3333331174e-1f) * x * z;
    const float fe = (float)e;
    y = y + -2.12194440e-4f * fe;
    y = y + -0.5f * z;
    float r = x + y;
    r = r + 0.693359375f * fe;
    return r;
}

struct MathF {
    // Cephes sinf/cosf: reduce |x| by multiples of pi/4 (DP1+DP2+DP3), select
    // octant, evaluate the sine or cosine minimax polynomial.
    static void sincos(float xin, float* s_out, float* c_out) {
        const float DP1 = 0.78515625f;
        const float DP2 = 2.4187564849853515625e-4f;
        const float DP3 = 3.77489497744594108e-8f;
        const float FOPI = 1.27323954473516f;
        float x = xin;
        int sign_s = 1;
        if (x < 0.0f) { x = -x; sign_s = -1; }
        int sign_c = 1;
        unsigned j = (unsigned)(x * FOPI);
        float y = (float)j;
        if (j & 1u) { j += 1u; y += 1.0f; }
        j &= 7u;
        if (j > 3u) { j -= 4u; sign_s = -sign_s; sign_c = -sign_c; }
        if (j > 1u) sign_c = -sign_c;
        float r = ((x - y * DP1) - y * DP2) - y * DP3;
        float z = r * r;
        float ps = ((-1.9515295891e-4f * z + 8.3321608736e-3f) * z - 1.6666654611e-1f) * z * r + r;
        float pc = ((2.443315711809948e-5f * z - 1.388731625493765e-3f) * z + 4.166664568298827e-2f) * z * z
                   - 0.5f * z + 1.0f;
        float s, c;
        if (j == 1u || j == 2u) { s = pc; c = ps; } else { s = ps; c = pc; }
        *s_out = sign_s < 0 ? -s : s;
        *c_out = sign_c < 0 ? -c : c;
    }
    static float sin(float x) { float s, c; sincos(x, &s, &c); return s; }
    static float cos(float x) { float s, c; sincos(x, &s, &c); return c; }
    // Cephes asinf kernel on [0, 0.5] after the half-angle identity.
    static float asin_kernel(float z, float x) {
        return ((((4.2163199048e-2f * z + 2.4181311049e-2f) * z + 4.5470025998e-2f) * z + 7.4953002686e-2f) * z
                + 1.6666752422e-1f) * z * x + x;
    }
    static float asin(float xin) {
        float a = xin < 0.0f ? -xin : xin;
        float r;
        if (a > 0.5f) {
            float z = 0.5f * (1.0f - a);
            float x = std::sqrt(z);
            r = asin_kernel(z, x);
            r = 1.5707963267948966f - (r + r);
        } else {
            r = asin_kernel(a * a, a);
        }
        return xin < 0.0f ? -r : r;
    }
    static float acos(float x) {
        if (x < -0.5f) return 3.14159265358979f - 2.0f * asin(std::sqrt(0.5f * (1.0f + x)));
        if (x > 0.5f) return 2.0f * asin(std::sqrt(0.5f * (1.0f - x)));
        return 1.5707963267948966f - asin(x);
    }
    static float atan(float xin) {
        float x = xin < 0.0f ? -xin : xin;
        float y;
        if (x > 2.414213562373095f) { y = 1.5707963267948966f; x = -1.0f / x; }
        else if (x > 0.4142135623730950f) { y = 0.7853981633974483f; x = (x - 1.0f) / (x + 1.0f); }
        else { y = 0.0f; }
        float z = x * x;
        y = y + ((((8.05374449538e-2f * z - 1.38776856032e-1f) * z + 1.99777106478e-1f) * z - 3.33329491539e-1f) * z * x + x);
        return xin < 0.0f ? -y : y;
    }
    static float atan2(float y, float x) {
        int code = 0;
        if (x < 0.0f) code = 2;
        if (y < 0.0f) code |= 1;
        if (x == 0.0f) {
            if (code & 1) return -1.5707963267948966f;
            if (y == 0.0f) return 0.0f;
            return 1.5707963267948966f;
        }
        if (y == 0.0f) return (code & 2) ? 3.14159265358979f : 0.0f;
        float w = 0.0f;
        if (code == 2) w = 3.14159265358979f;
        else if (code == 3) w = -3.14159265358979f;
        return w + atan(y / x);
    }
    // Cephes expf: n = floor(x log2(e) + 1/2), x - n ln 2 in two parts
    // (C1 + C2), the degree-6 minimax polynomial, then the exact scaling by
    // 2^n.  Below -87 (where e^x nears the smallest normal float) the result
    // is 0, so no denormal is ever formed; above 88.72 it is +inf.
    static float exp(float xin) {
        if (xin > 88.72283905206835f) return INFINITY;
        if (xin < -87.0f) return 0.0f;
        const float LOG2EF = 1.44269504088896341f;
        const float C1 = 0.693359375f;
        const float C2 = -2.12194440e-4f;
        float z = std::floor(LOG2EF * xin + 0.5f);
        float x = xin - z * C1;
        x = x - z * C2;
        const int n = (int)z;
        const float xx = x * x;
        float p = (((((1.9875691500e-4f * x + 1.3981999507e-3f) * x + 8.3334519073e-3f) * x + 4.1665795894e-2f) * x
                    + 1.6666665459e-1f) * x + 5.0000001201e-1f);
        p = (p * xx + x) + 1.0f;
        uint32_t bits = (uint32_t)(n + 127) << 23;
        float scale;
        std::memcpy(&scale, &bits, 4);
        return p * scale;
    }
    static float log(float x) { return cephes_logf(x); }
    // log2 as log(x) * log2(e) (the ray-cone LOD's log2)
    static float log2(float x) { return cephes_logf(x) * 1.44269504088896341f; }
    static float pow(float x, float y) {
        if (x == 0.0f) return y > 0.0f ? 0.0f : (y == 0.0f ? 1.0f : INFINITY);
        return exp(y * log(x));
    }
};

struct MathD {
    static double sin(double x) { return std::sin(x); }
    static double cos(double x) { return std::cos(x); }
    static double acos(double x) { return std::acos(x); }
    static double atan2(double y, double x) { return std::atan2(y, x); }
    static double exp(double x) { return std::exp(x); }
    static double asin(double x) { return std::asin(x); }
    static double log(double x) { return std::log(x); }
    static double log2(double x) { return std::log2(x); }
    static double pow(double x, double y) { return std::pow(x, y); }
};

struct MathDual {
    static Dual sin(Dual x) { return {std::sin(x.v), std::cos(x.v) * x.d}; }
    static Dual cos(Dual x) { return {std::cos(x.v), -std::sin(x.v) * x.d}; }
    static Dual acos(Dual x) {
        const double s = std::sqrt(std::max(0.0, 1.0 - x.v * x.v));
        return {std::acos(x.v), s > 0.0 ? -x.d / s : 0.0};
    }
    static Dual atan2(Dual y, Dual x) {
        const double r2 = x.v * x.v + y.v * y.v;
        return {std::atan2(y.v, x.v), r2 > 0.0 ? (x.v * y.d - y.v * x.d) / r2 : 0.0};
    }
    static Dual exp(Dual x) {
        const double e = std::exp(x.v);
        return {e, e * x.d};
    }
    static Dual asin(Dual x) {
        const double s = std::sqrt(std::max(0.0, 1.0 - x.v * x.v));
        return {std::asin(x.v), s > 0.0 ? x.d / s : 0.0};
    }
    static Dual log(Dual x) { return {std::log(x.v), x.v > 0.0 ? x.d / x.v : 0.0}; }
    static Dual log2(Dual x) { return {std::log2(x.v), x.v > 0.0 ? x.d / (x.v * 0.6931471805599453) : 0.0}; }
    static Dual pow(Dual x, Dual y) {
        const double p = std::pow(x.v, y.v);
        const double dx = x.v > 0.0 ? y.v * std::pow(x.v, y.v - 1.0) * x.d : 0.0;
        const double dy = x.v > 0.0 ? p * std::log(x.v) * y.d : 0.0;
        return {p, dx + dy};
    }
};

template <class R> struct MathOf;
template <> struct MathOf<float> { using M = MathF; };
template <> struct MathOf<double> { using M = MathD; };
template <> struct MathOf<Dual> { using M = MathDual; };

// PCG hash + 24-bit uniform (PathTraceMath.slang:16-24).
inline uint32_t pcg_hash(uint32_t x) {
    x = x * 747796405u + 2891336453u;
    x = ((x >> ((x >> 28u) + 4u)) ^ x) * 277803737u;
    return (x >> 22u) ^ x;
}
template <class R> inline R randf(uint32_t& seed) {
    seed = pcg_hash(seed);
    return R(seed & 0x00FFFFFFu) / R(16777216.0);
}

inline float u2f(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
inline uint32_t f2u(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }

}  // namespace orc
