// Device-side ray algebra of the trace loop, templated on the bulk
// arithmetic type R (double for the FP64 path, float for the FP32 path).
// Frame transforms between global, stage and element frames are always
// double (positions reach kilometres and the 1e-5 m self-intersection nudge
// of coretrace/raytrace.cpp:729-731 is below the float ulp there); the torus
// quartic is always double; the Newton start plane is found in double.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

#include "st_types.h"

namespace st {

#define STD __device__ __forceinline__

static constexpr double kDeg = 0.017453292519943295;  // coretrace/types.h:64

// -------------------------------------------------------------- math shims
template <typename R> struct M;
template <> struct M<double> {
    static STD double sqrt_(double x) { return ::sqrt(x); }
    static STD double sin_(double x) { return ::sin(x); }
    static STD double cos_(double x) { return ::cos(x); }
    static STD double atan2_(double y, double x) { return ::atan2(y, x); }
    static STD double asin_(double x) { return ::asin(x); }
    static STD double acos_(double x) { return ::acos(x); }
    static STD double log_(double x) { return ::log(x); }
    static STD double pow_(double x, double y) { return ::pow(x, y); }
    static STD double fabs_(double x) { return ::fabs(x); }
};
template <> struct M<float> {
    static STD float sqrt_(float x) { return ::sqrtf(x); }
    static STD float sin_(float x) { return ::sinf(x); }
    static STD float cos_(float x) { return ::cosf(x); }
    static STD float atan2_(float y, float x) { return ::atan2f(y, x); }
    static STD float asin_(float x) { return ::asinf(x); }
    static STD float acos_(float x) { return ::acosf(x); }
    static STD float log_(float x) { return ::__logf(x); }
    static STD float pow_(float x, float y) { return ::powf(x, y); }
    static STD float fabs_(float x) { return ::fabsf(x); }
};

// ------------------------------------------------------------------ RNG ---
// Philox4x32-10 (Salmon et al., SC'11).  Draw k of the stream (seed, stage,
// id) is word k%4 of the block with counter (id_lo, id_hi, stage, k/4) and
// key (seed, 0x53544F52), u = w 2^-32.  Both precisions draw the same double
// uniforms and take every accept/reject decision in double, so the FP32 path
// samples exactly the errors the FP64 path samples.
STD void philox10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        // full 64-bit products: one v_mad_u64_u32 each instead of a
        // separate high and low multiply
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        // three-input XOR in one v_bitop3_b32 (truth table 0x96)
        const uint32_t n0 = __builtin_amdgcn_bitop3_b32((uint32_t)(p1 >> 32), c[1], k0, 0x96);
        const uint32_t n2 = __builtin_amdgcn_bitop3_b32((uint32_t)(p0 >> 32), c[3], k1, 0x96);
        c[0] = n0; c[1] = (uint32_t)p1; c[2] = n2; c[3] = (uint32_t)p0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

STD uint32_t pick4(const uint32_t b[4], uint32_t i) {
    return i == 0 ? b[0] : (i == 1 ? b[1] : (i == 2 ? b[2] : b[3]));
}

STD double word_uniform(uint32_t w) { return (double)w * 2.3283064365386963e-10; }

struct Rng {
    uint32_t seed, stage, id_lo, id_hi;
    uint32_t k;
    uint32_t blk[4];
    bool have_saved;
    double saved;

    STD void begin(uint32_t seed_, uint32_t stage_, uint64_t id) {
        seed = seed_; stage = stage_; id_lo = (uint32_t)id; id_hi = (uint32_t)(id >> 32);
        k = 0; have_saved = false; saved = 0.0;
    }
    STD void fill(uint32_t block) {
        blk[0] = id_lo; blk[1] = id_hi; blk[2] = stage; blk[3] = block;
        philox10(blk, seed, 0x53544F52u);
    }
    STD uint32_t next_word() {
        if ((k & 3u) == 0u) fill(k >> 2);
        uint32_t w = pick4(blk, k & 3u);
        ++k;
        return w;
    }
    // Continue the stream at draw index kk (the draws before it were consumed elsewhere).
    STD void seek(uint32_t kk) {
        k = kk;
        if (k & 3u) fill(k >> 2);
    }
    STD double uniform() { return word_uniform(next_word()); }
    // Marsaglia polar with one cached deviate, the algorithm of libstdc++'s
    // normal_distribution (bits/random.tcc:1802-1835): y*m first, x*m cached.
    // The rejection test is double in both modes; FAST takes the radial
    // factor from float log1p/sqrt (relative error ~1e-7 of a deviate).
    template <bool FAST>
    STD double normal(double mean, double sd) {
        double ret;
        if (have_saved) {
            have_saved = false;
            ret = saved;
        } else {
            double x, y, r2;
            do {
                x = 2.0 * uniform() - 1.0;
                y = 2.0 * uniform() - 1.0;
                r2 = x * x + y * y;
            } while (r2 > 1.0 || r2 == 0.0);
            double mult;
            if constexpr (FAST) {
                // log(r2) to float relative accuracy over (0, 1]: log1p of
                // r2 - 1 (exact in double) near 1, log of r2 below 1/2
                const float lg = (r2 < 0.5) ? ::logf((float)r2) : ::log1pf((float)(r2 - 1.0));
                mult = (double)::sqrtf(-2.0f * lg / (float)r2);
            } else {
                mult = ::sqrt(-2 * ::log(r2) / r2);
            }
            saved = x * mult;
            have_saved = true;
            ret = y * mult;
        }
        return mean + ret * sd;
    }
};

// ------------------------------------------------------------ transforms --
STD void mat_vec(const double *Mx, const double v[3], double o[3]) {
    o[0] = Mx[0] * v[0] + Mx[1] * v[1] + Mx[2] * v[2];
    o[1] = Mx[3] * v[0] + Mx[4] * v[1] + Mx[5] * v[2];
    o[2] = Mx[6] * v[0] + Mx[7] * v[1] + Mx[8] * v[2];
}
STD void mat_t_vec(const double *Mx, const double v[3], double o[3]) {
    o[0] = Mx[0] * v[0] + Mx[3] * v[1] + Mx[6] * v[2];
    o[1] = Mx[1] * v[0] + Mx[4] * v[1] + Mx[7] * v[2];
    o[2] = Mx[2] * v[0] + Mx[5] * v[1] + Mx[8] * v[2];
}
// coretrace/mathproc.cpp:115-173 with R the reference->local rotation
STD void to_local(const double p[3], const double c[3], const double org[3], const double *Rm, double pl[3],
                  double cl[3]) {
    double d[3] = {p[0] - org[0], p[1] - org[1], p[2] - org[2]};
    double cc[3] = {c[0], c[1], c[2]};
    mat_vec(Rm, d, pl);
    mat_vec(Rm, cc, cl);
}
STD void to_ref(const double pl[3], const double cl[3], const double org[3], const double *Rm, double p[3],
                double c[3]) {
    double t[3], cc[3] = {cl[0], cl[1], cl[2]};
    mat_t_vec(Rm, pl, t);
    mat_t_vec(Rm, cc, c);
    p[0] = t[0] + org[0];
    p[1] = t[1] + org[1];
    p[2] = t[2] + org[2];
}
template <typename R>
STD R dot3(const R a[3], const R b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// ----------------------------------------------------- aperture tests -----
// Final aperture test (coretrace/determineelementintersectionnew.cpp:104-623)
// and, with plane = true, the aperture-plane variant used to choose the
// Newton start (coretrace/intersect.cpp:272-583; its rectangle test never
// fails, :402,408).
template <typename R>
STD bool in_tri(R x1, R y1, R x2, R y2, R x3, R y3, R xt, R yt) {
    R a = (x1 - xt) * (y2 - yt) - (x2 - xt) * (y1 - yt);
    R b = (x2 - xt) * (y3 - yt) - (x3 - xt) * (y2 - yt);
    R c = (x3 - xt) * (y1 - yt) - (x1 - xt) * (y3 - yt);
    bool sa = a >= 0, sb = b >= 0, sc = c >= 0;
    return sa == sb && sb == sc;
}

template <typename R>
STD bool aperture_contains(const DevElem &e, R x, R y, bool plane) {
    const R A = (R)e.p[0], B = (R)e.p[1], C = (R)e.p[2];
    R r = M<R>::sqrt_(x * x + y * y);
    const R c30 = (R)0.86602540378443871, s30 = (R)0.49999999999999994;  // cos, sin of 30 deg
    const R sl60 = (R)1.7320508075688767;
    switch (e.ap) {
    case AP_CIRCLE:
        return !(r > A / R(2));
    case AP_HEX: {
        R Ro = A / R(2);
        if (r > Ro) return false;
        R Ri = Ro * c30;
        if (r <= Ri) return true;
        R XL = M<R>::sqrt_(Ro * Ro - Ri * Ri);
        if (x <= Ro && x > XL) {
            R Y1 = sl60 * (x - Ro);
            return y >= Y1 && y <= -Y1;
        }
        if (x <= XL && x >= -XL) return y >= -Ri && y <= Ri;
        if (x < -XL && x >= -Ro) {
            R Y3 = sl60 * (x + Ro);
            return y >= -Y3 && y <= Y3;
        }
        return false;
    }
    case AP_TRI: {
        R Ro = A / R(2);
        if (r > Ro) return false;
        R Ri = Ro * s30;
        if (r <= Ri) return true;
        if (x <= Ro && x > R(0)) {
            R Y1 = -sl60 * (x - Ri / c30);
            return y <= Y1 && y >= -Ri;
        }
        if (x >= -Ro && x <= R(0)) {
            R Y3 = sl60 * (x + Ri / c30);
            return y >= -Ri && y <= Y3;
        }
        return false;
    }
    case AP_RECT:
        if (plane) return true;
        if (x > A / R(2) || x < -A / R(2)) return false;
        if (y > B / R(2) || y < -B / R(2)) return false;
        return true;
    case AP_ANNULUS: {
        bool radial = plane || !(e.p[0] == 0.0 && e.p[1] == 0.0);
        if (radial && (r < A || r > B)) return false;
        R half = C * (R)kDeg / R(2);
        if (x >= R(0)) {
            R a = M<R>::asin_(y / r);
            return !(a > half || a < -half);
        }
        if (x < R(0)) {
            if (y >= R(0) && (M<R>::acos_(y / r) + (R)(M_PI / 2.0)) > half) return false;
            if (y < R(0) && (-M<R>::acos_(-y / r) - (R)(M_PI / 2.0)) < -half) return false;
            return true;
        }
        return false;
    }
    case AP_LINE:
        if (!(e.p[0] == 0.0 && e.p[1] == 0.0)) {
            if (x < A || x > B) return false;
        }
        return !(y < -C / R(2) || y > C / R(2));
    case AP_ITRI:
        return in_tri<R>((R)e.p[0], (R)e.p[1], (R)e.p[2], (R)e.p[3], (R)e.p[4], (R)e.p[5], x, y);
    case AP_QUAD:
        return in_tri<R>((R)e.p[0], (R)e.p[1], (R)e.p[2], (R)e.p[3], (R)e.p[4], (R)e.p[5], x, y) ||
               in_tri<R>((R)e.p[0], (R)e.p[1], (R)e.p[4], (R)e.p[5], (R)e.p[6], (R)e.p[7], x, y);
    default:
        return false;
    }
}

// ------------------------------------------------------- closed forms -----
template <typename R>
STD R sqr_grad(R K, R d, R a2) { return (((((R(2) * K) * d) / a2) * R(2)) * K) * d / a2; }

// Unit normal of a closed-form quadric at X
// (coretrace/quadricsurfaceclosedform.cpp:294-306; same constants as below).
template <typename R>
STD void quadric_normal(const DevElem &e, const R X[3], R DF[3]) {
    R Zc = 0, Kx = 1, Ky = 1, Kz = 1, a2 = 1, b2 = 1, c2 = 1;
    switch (e.surf) {
    case SURF_SPHERE: Ky = (e.stype == 7) ? R(0) : R(1); Zc = R(1) / (R)e.vcx; break;
    case SURF_PARABOLA: {
        a2 = R(2) / (R)e.vcx;
        b2 = e.vcy > 0.0 ? R(2) / (R)e.vcy : R(1e10);
        Ky = (e.stype == 7) ? R(0) : R(1);
        R sm = M<R>::sqrt_(sqr_grad(Kx, X[0], a2) + sqr_grad(Ky, X[1], b2) + R(1));
        DF[0] = -(R(2) * Kx * X[0] / a2) / sm;
        DF[1] = -(R(2) * Ky * X[1] / b2) / sm;
        DF[2] = R(1) / sm;
        return;
    }
    case SURF_CONIC: Kz = (R)e.kappa; Zc = R(1) / (R)e.kappa / (R)e.vcx; break;
    default: Ky = 0; Zc = R(1) / (R)e.curv_rev; break;  // SURF_CYLINDER
    }
    R sm = M<R>::sqrt_(sqr_grad(Kx, X[0], a2) + sqr_grad(Ky, X[1], b2) + sqr_grad(Kz, X[2] - Zc, c2));
    DF[0] = -(R(2) * Kx * X[0] / a2) / sm;
    DF[1] = -(R(2) * Ky * X[1] / b2) / sm;
    DF[2] = -(R(2) * Kz * (X[2] - Zc) / c2) / sm;
}

// Ray-quadric intersection (coretrace/quadricsurfaceclosedform.cpp:58-308).
// Returns false on a miss (ErrorFlag 1).
template <typename R>
STD bool quadric(const DevElem &e, const R P[3], const R D[3], R X[3], R DF[3], R *path) {
    R Xc = 0, Yc = 0, Zc = 0, Kx = 0, Ky = 0, Kz = 0, a2 = 1, b2 = 1, c2 = 1, A = 0, B = 0, C = 0;
    switch (e.surf) {
    case SURF_SPHERE: {
        Kx = 1; Ky = (e.stype == 7) ? R(0) : R(1); Kz = 1;
        R r = R(1) / (R)e.vcx;
        Zc = r;
        R xd = P[0], yd = P[1], zd = P[2] - Zc;
        A = D[0] * D[0] * Kx / a2 + D[1] * D[1] * Ky / b2 + D[2] * D[2] * Kz / c2;
        B = R(2) * (Kx * xd * D[0] / a2 + Ky * yd * D[1] / b2 + Kz * zd * D[2] / c2);
        C = Kx * xd * xd / a2 + Ky * yd * yd / b2 + Kz * zd * zd / c2 - r * r;
        break;
    }
    case SURF_PARABOLA: {
        a2 = R(2) / (R)e.vcx;
        b2 = e.vcy > 0.0 ? R(2) / (R)e.vcy : R(1e10);
        c2 = 1;
        Kx = 1; Ky = (e.stype == 7) ? R(0) : R(1); Kz = 0;
        R xd = P[0], yd = P[1], zd = P[2];
        A = D[0] * D[0] * Kx / a2 + D[1] * D[1] * Ky / b2;
        B = R(2) * (Kx * xd * D[0] / a2 + Ky * yd * D[1] / b2) - D[2] / c2;
        C = Kx * xd * xd / a2 + Ky * yd * yd / b2 - zd / c2;
        break;
    }
    case SURF_CONIC: {
        Kx = 1; Ky = 1; Kz = (R)e.kappa;
        Zc = R(1) / (R)e.kappa / (R)e.vcx;
        R xd = P[0], yd = P[1], zd = P[2] - Zc;
        A = D[0] * D[0] * Kx / a2 + D[1] * D[1] * Ky / b2 + D[2] * D[2] * Kz / c2;
        B = R(2) * (Kx * xd * D[0] / a2 + Ky * yd * D[1] / b2 + Kz * zd * D[2] / c2);
        C = Kx * xd * xd / a2 + Ky * yd * yd / b2 + Kz * zd * zd / c2 - R(1) / (R)e.kappa / ((R)e.vcx * (R)e.vcx);
        break;
    }
    default: {  // SURF_CYLINDER
        Kx = 1; Ky = 0; Kz = 1;
        R r = R(1) / (R)e.curv_rev;
        Zc = r;
        R xd = P[0], yd = P[1], zd = P[2] - Zc;
        A = D[0] * D[0] * Kx / a2 + D[1] * D[1] * Ky / b2 + D[2] * D[2] * Kz / c2;
        B = R(2) * (Kx * xd * D[0] / a2 + Ky * yd * D[1] / b2 + Kz * zd * D[2] / c2);
        C = Kx * xd * xd / a2 + Ky * yd * yd / a2 + Kz * zd * zd / c2 - r * r;
        break;
    }
    }
    if constexpr (sizeof(R) == 4) {
        // FP32: (z - zc)^2 - zc^2 cancels catastrophically when the centre
        // is far from the vertex (heliostat spheres have R ~ 1 km); use the
        // expanded forms, which are algebraically equal.
        if (e.surf == SURF_SPHERE || e.surf == SURF_CYLINDER) {
            R r2x = Zc + Zc;
            C = Kx * P[0] * P[0] + Ky * P[1] * P[1] + P[2] * (P[2] - r2x);
        } else if (e.surf == SURF_CONIC) {
            C = P[0] * P[0] + P[1] * P[1] + Kz * P[2] * P[2] - R(2) * P[2] / (R)e.vcx;
        }
    }
    R t;
    if (M<R>::fabs_(A) < R(1e-12)) {
        t = -C / B;
        if (!(t > R(0))) return false;
    } else {
        R disc = (B * B) - R(4) * A * C;
        if (!((B * B) > R(4) * A * C)) return false;
        R sq = M<R>::sqrt_(disc);
        R t1, t2;
        if constexpr (sizeof(R) == 4) {
            // roots without cancellation: t1 t2 = C / A
            R q = R(-0.5) * (B + (B >= R(0) ? sq : -sq));
            if (B >= R(0)) { t2 = q / A; t1 = C / q; }
            else { t1 = q / A; t2 = C / q; }
        } else {
            t1 = (-B + sq) / (R(2) * A);
            t2 = (-B - sq) / (R(2) * A);
        }
        if (t2 > R(0)) {
            t = t2;
            R z2 = P[2] + t2 * D[2];
            if (z2 > (R)e.zap) t = t1;
            bool len_check = e.surf == SURF_CYLINDER || (e.surf == SURF_SPHERE && e.stype == 7);
            if (len_check) {
                R y = P[1] + t * D[1];
                R half = (R)e.p[2] / R(2);
                if (y < -half || y > half) t = t1;
            }
        } else if (t2 == R(0)) {
            t = t1;
        } else if (t2 < R(0) && t1 > R(0)) {
            t = t1;
        } else if (t1 <= R(0)) {
            return false;
        } else {
            return false;  // NaN roots
        }
    }
    for (int i = 0; i < 3; i++) X[i] = P[i] + t * D[i];
    *path = t;
    quadric_normal<R>(e, X, DF);
    return true;
}

// Torus normal at a hit point (coretrace/torusclosedform.cpp:176-187).
STD void torus_normal(const DevElem &e, const double X[3], double DF[3]) {
    const double Rs = e.cross_r, Ra = e.ann_r;
    const double rxy = ::sqrt(X[0] * X[0] + X[1] * X[1]);
    DF[0] = 2.0 * X[0] * (Ra - rxy) / rxy;
    DF[1] = 2.0 * X[1] * (Ra - rxy) / rxy;
    DF[2] = -(2.0 * (X[2] - Rs));
}

// Torus by the Ferrari quartic with a bisected resolvent cubic and the
// reference's root bookkeeping (coretrace/torusclosedform.cpp:57-188,
// root432.cpp:53-284, mathproc.cpp:333-348).  Always double.
struct QuarticState {
    double a30, a31, a32, a20, a21;
    double r[5];
    double im;
    STD double f(double x) const { return x * x * x + a32 * x * x + a31 * x + a30; }
    STD void root2() {
        double w = a21 * a21 - 4 * a20;
        double q1 = -a21 / 2.0;
        if (w < 0) {
            im = ::sqrt(-w) / 2.0;
            r[1] = q1; r[2] = q1;
        } else if (w == 0) {
            r[1] = q1; r[2] = q1; im = 0.0;
        } else if (w > 0) {
            double q2 = ::sqrt(w) / 2.0;
            im = 0.0;
            r[1] = q1 + q2; r[2] = q1 - q2;
        }
    }
    STD double root3_one() {
        if (a30 == 0) return 0.0;
        double am = ::fabs(a30);
        double t = ::fabs(a31); if (am < t) am = t;
        t = ::fabs(a32); if (am < t) am = t;
        t = 0.0;  // a[3][3] is zero
        double aa, b;
        if (a30 > 0) { aa = -am - 1; b = 0; } else { aa = 0; b = am + 1; }
        double xa = aa, xb = b;
        if (f(aa) == 0) return aa;
        if (f(b) == 0) return b;
        double xc;
        for (int it = 0; it < 2000; ++it) {
            xc = (xa + xb) / 2.0;
            double te = f(xc);
            if (te == 0) break;
            if (xb - xa < 0.000001) break;
            if (f(xa) * te > 0) xa = xc; else xb = xc;
        }
        return xc;
    }
};

// The torus (x^2 + y^2 + (z - Rs)^2 + Ra^2 - Rs^2)^2 = 4 Ra^2 (x^2 + y^2) of
// the quartic below lies in the shell |z - Rs| <= |Rs|, rho <= |Ra| + |Rs|
// and, for a ring torus, rho >= |Ra| - |Rs| (rho^2 = x^2 + y^2).  A line
// that stays a margin outside the shell has no real root; its quartic solve
// is skipped.  The margin, 1e-3 (|Ra| + |Rs|), keeps near-tangent lines,
// whose root pairs rounding could turn real or complex, on the solve.
STD bool torus_shell(const DevElem &e, const double P[3], const double D[3]) {
    const double Rs = e.cross_r, Ra = e.ann_r;
    const double m = 1e-3 * (::fabs(Ra) + ::fabs(Rs));
    const double z0 = Rs - ::fabs(Rs) - m, z1 = Rs + ::fabs(Rs) + m;
    double ta = -1e300, tb = 1e300;  // the line's parameters inside the slab z0 <= z <= z1
    if (D[2] != 0.0) {
        const double u = (z0 - P[2]) / D[2], v = (z1 - P[2]) / D[2];
        ta = ::fmin(u, v);
        tb = ::fmax(u, v);
    } else if (P[2] < z0 || P[2] > z1) {
        return false;
    }
    // rho^2(t) = a t^2 + 2 b t + c on [ta, tb]: least at the vertex (clamped), most at an end
    const double a = D[0] * D[0] + D[1] * D[1], b = P[0] * D[0] + P[1] * D[1], c = P[0] * P[0] + P[1] * P[1];
    const double tv = ::fmin(::fmax(a > 0.0 ? -b / a : ta, ta), tb);
    const double hi = ::fabs(Ra) + ::fabs(Rs) + m;
    if ((a * tv + 2.0 * b) * tv + c > hi * hi) return false;
    const double lo = ::fabs(Ra) - ::fabs(Rs) - m;
    if (lo > 0.0 && ta > -1e300 && tb < 1e300) {
        const double ra = (a * ta + 2.0 * b) * ta + c, rb = (a * tb + 2.0 * b) * tb + c;
        if (::fmax(ra, rb) < lo * lo) return false;
    }
    return true;
}

STD bool torus(const DevElem &e, const double P[3], const double D[3], double X[3], double DF[3], double *path) {
    const double Rs = e.cross_r, Ra = e.ann_r;
    const double Xo = P[0], Yo = P[1], Zo = P[2];
    const double Ep = D[0], Et = D[1], Rh = D[2];
    double a44 = ::pow(Ep, 4) + 2.0 * Ep * Ep * (Et * Et + Rh * Rh) + ::pow(Et, 4) + 2.0 * Et * Et * Rh * Rh + ::pow(Rh, 4);
    double a43 = 4.0 * (Ep * Ep + Et * Et + Rh * Rh) * (Ep * Xo + Et * Yo + Rh * Zo - Rh * Rs);
    double a42 = Xo * Xo * (6.0 * Ep * Ep + 2.0 * Et * Et + 2.0 * Rh * Rh) + 8.0 * Ep * Xo * (Et * Yo + Rh * Zo - Rh * Rs) +
                 2.0 * Yo * Yo * (Ep * Ep + 3.0 * Et * Et + Rh * Rh) + 8.0 * Et * Rh * Yo * (Zo - Rs) +
                 (Ep * Ep + Et * Et + 3.0 * Rh * Rh) * (2.0 * Zo * Zo - 4.0 * Rs * Zo) -
                 2.0 * Ra * Ra * (Ep * Ep + Et * Et - Rh * Rh) + 4.0 * Rh * Rh * Rs * Rs;
    double a41 = 4.0 * (Xo * Xo * (Ep * Xo + Et * Yo + Rh * Zo - Rh * Rs) + Yo * Yo * (Ep * Xo + Et * Yo + Rh * Zo - Rh * Rs) +
                        Zo * Zo * (Ep * Xo + Et * Yo + Rh * Zo - 3.0 * Rh * Rs) - 2.0 * Ep * Rs * Xo * Zo - Ep * Ra * Ra * Xo -
                        2.0 * Et * Rs * Yo * Zo - Et * Ra * Ra * Yo + Rh * Ra * Ra * (Zo - Rs) + 2.0 * Rh * Rs * Rs * Zo);
    double a40 = ::pow(Xo, 4) + 2.0 * Xo * Xo * (Yo * Yo + Zo * Zo - 2.0 * Rs * Zo - Ra * Ra) + ::pow(Yo, 4) +
                 2.0 * Yo * Yo * (Zo * Zo - 2.0 * Rs * Zo - Ra * Ra) + ::pow(Zo, 4) - 4.0 * Rs * Zo * Zo * Zo +
                 2.0 * Ra * Ra * Zo * Zo + 4.0 * Rs * Rs * Zo * Zo - 4.0 * Ra * Ra * Rs * Zo + ::pow(Ra, 4);
    // normalisation (root432.cpp Root_4)
    a40 = a40 / a44; a41 = a41 / a44; a42 = a42 / a44; a43 = a43 / a44;
    QuarticState Q;
    for (int i = 0; i < 5; i++) Q.r[i] = 0.0;
    Q.im = 0.0;
    double aa = a43, b = a42, c = a41, d = a40;
    double q = b - (3.0 * aa * aa / 8.0);
    double rr = c - (aa * b / 2) + (aa * aa * aa / 8.0);
    double s = d - (aa * c / 4) + (aa * aa * b / 16.0) - (3 * aa * aa * aa * aa / 256.0);
    Q.a32 = q / 2.0;
    Q.a31 = (q * q - 4 * s) / 16.0;
    Q.a30 = -(rr * rr / 64.0);
    Q.a20 = 0; Q.a21 = 0;
    if ((rr != 0) || (Q.a31 >= 0)) {
        Q.r[3] = Q.root3_one();
    } else {
        Q.a21 = Q.a32; Q.a20 = Q.a31;
        Q.root2();
        Q.r[3] = Q.r[1];
    }
    double k = ::sqrt(Q.r[3]);
    if (k == 0) {
        rr = ::sqrt(q * q - 4 * s);
    } else {
        q = q + (4 * Q.r[3]);
        rr = rr / (2 * k);
    }
    double l = (q - rr) / 2.0, m = (q + rr) / 2.0;
    Q.a21 = 2 * k; Q.a20 = l;
    Q.root2();
    double rv[5];
    rv[0] = 0.0;
    rv[3] = Q.r[1] - (a43 / 4);
    rv[4] = Q.r[2] - (a43 / 4.0);
    double ii = Q.im;
    Q.a21 = -2 * k; Q.a20 = m;
    Q.root2();
    rv[2] = Q.r[2] - (a43 / 4.0);
    rv[1] = Q.r[1] - (a43 / 4.0);
    double im1 = Q.im, im2 = ii;
    double L = 0.0;
    if (im1 == 0.0 && im2 == 0.0) {
        // insertion sort of slots 0..3 where a new minimum lands in slot 1
        for (int j = 1; j < 4; j++) {
            double a = rv[j];
            int i = j - 1;
            bool stopped = false;
            for (; i >= 0; i--) {
                if (rv[i] <= a) { stopped = true; break; }
                rv[i + 1] = rv[i];
            }
            if (!stopped) i = 0;
            rv[i + 1] = a;
        }
        if (rv[4] <= 0.0) return false;
        L = rv[4];
        if (!(rv[3] <= 0.0)) {
            L = rv[3];
            if (!(rv[2] <= 0.0)) {
                L = rv[2];
                if (!(rv[1] <= 0.0)) L = rv[1];
            }
        }
    } else if (im1 == 0.0 && im2 != 0.0) {
        L = (rv[1] <= rv[2]) ? rv[1] : rv[2];
    } else if (im1 != 0.0 && im2 == 0.0) {
        L = (rv[3] <= rv[4]) ? rv[3] : rv[4];
    }
    *path = L;
    if (L == 0.0) return false;
    double x = Xo + L * Ep, y = Yo + L * Et, z = Zo + L * Rh;
    double rxy = ::sqrt(x * x + y * y);
    X[0] = x; X[1] = y; X[2] = z;
    DF[0] = 2.0 * x * (Ra - rxy) / rxy;
    DF[1] = 2.0 * y * (Ra - rxy) / rxy;
    DF[2] = -(2.0 * (z - Rs));
    return true;
}

// --------------------------------------------- Newton-Raphson surfaces ----
// Zernike monomials z = sum B_ij x^j y^(i-j) and slopes (coretrace/vshot.cpp:59-139).
template <typename R>
STD void zernike(const DevElem &e, const double *aux, R x, R y, R *z, R *dzdx, R *dzdy, bool slopes) {
    const int n = e.fit_order;
    const int w = e.aux_n;  // row stride (order + 1)
    const double *Bm = aux + e.aux;
    R xs = x, ys = y;
    R zz = 0, dx = 0, dy = 0;
    R px[16], py[16];
    px[0] = 1; py[0] = 1;
    for (int i = 1; i <= n; i++) { px[i] = px[i - 1] * x; py[i] = py[i - 1] * y; }
    for (int i = 0; i <= n; i++)
        for (int j = 0; j <= i; j++) zz = zz + (R)Bm[i * w + j] * px[j] * py[i - j];
    *z = zz;
    if (!slopes) return;
    if (xs == R(0)) xs = xs + R(1e-10);
    if (ys == R(0)) ys = ys + R(1e-10);
    px[0] = 1; py[0] = 1;
    for (int i = 1; i <= n; i++) { px[i] = px[i - 1] * xs; py[i] = py[i - 1] * ys; }
    for (int i = 1; i <= n; i++)
        for (int j = 0; j <= i; j++) {
            R b = (R)Bm[i * w + j];
            if (j > 0) dx = dx + R(j) * b * px[j - 1] * py[i - j];
            if (i - j > 0) dy = dy + R(i - j) * b * px[j] * py[i - j - 1];
        }
    *dzdx = dx;
    *dzdy = dy;
}

// The same monomial surface and its slopes in Horner form, for the float
// search of FP32 runs: z = sum_j x^j P_j(y), P_j(y) = sum_k B[j+k][j] y^k,
// each P_j and its derivative by Horner in y, then Horner in x for z and
// dz/dx (d/dx of Z x + p is DX x + Z) and dz/dy.  No power tables: the loop
// state is six scalars.  FP64 runs keep the reference's term order
// (zernike above), and every FP32 hit is finished in double by that.
template <typename R>
STD void zernike_horner(const DevElem &e, const double *aux, R x, R y, R *z, R *dzdx, R *dzdy) {
    const int n = e.fit_order;
    const int w = e.aux_n;
    const double *Bm = aux + e.aux;
    R Z = 0, DX = 0, DY = 0;
    for (int j = n; j >= 0; --j) {
        R p = 0, dp = 0;
        for (int k = n - j; k >= 0; --k) {
            dp = dp * y + p;
            p = p * y + (R)Bm[(j + k) * w + j];
        }
        DX = DX * x + Z;
        Z = Z * x + p;
        DY = DY * x + dp;
    }
    *z = Z;
    *dzdx = DX;
    *dzdy = DY;
}

// Numerical-Recipes splint (coretrace/mathproc.cpp:244-276)
template <typename R>
STD bool splint(const double *xa, const double *ya, const double *y2a, int n, R x, R *y, R *dydx) {
    int klo = 0, khi = n - 1;
    while (khi - klo > 1) {
        int k = (khi + klo) / 2;
        if ((R)xa[k] > x) khi = k; else klo = k;
    }
    R xk = (R)xa[khi], xl = (R)xa[klo];
    R h = xk - xl;
    if (h == R(0)) return false;
    R a = (xk - x) / h, b = (x - xl) / h;
    R yl = (R)ya[klo], yk = (R)ya[khi], y2l = (R)y2a[klo], y2k = (R)y2a[khi];
    *y = a * yl + b * yk + ((a * a * a - a) * y2l + (b * b * b - b) * y2k) * (h * h) / R(6);
    *dydx = (yk - yl) / (xk - xl) - (R(3) * a * a - R(1)) * (xk - xl) * y2l / R(6) +
            (R(3) * b * b - R(1)) * (xk - xl) * y2k / R(6);
    return true;
}

// Polynomial surface of revolution (coretrace/mathproc.cpp:218-241)
template <typename R>
STD void polyrev(const DevElem &e, const double *aux, R x, R y, R *z, R *dzdx, R *dzdy) {
    const double *c = aux + e.aux;
    R r = M<R>::sqrt_(x * x + y * y);
    R zz = 0, dr = 0, rp = 1, rpm1 = 0;  // rp = r^i, rpm1 = r^(i-1)
    for (int i = 0; i <= e.fit_order; i++) {
        if (i >= 1) dr = dr + R(i) * (R)c[i] * rpm1;
        zz = zz + (R)c[i] * rp;
        rpm1 = rp;
        rp = rp * r;
    }
    *z = zz;
    *dzdx = dr * (x / r);
    *dzdy = dr * (y / r);
}

// Measured slopes (angles, rad) at (x, y): modified Shepard weights 1/r^2 of
// the VSHOT points within the region of interest, the point's own slopes on
// an exact hit, zero with none in reach (coretrace/vshot.cpp:290-367).
STD void vshot_slopes(const DevElem &e, const double *aux, double x, double y, double *zx, double *zy) {
    const double *v = aux + e.aux2;
    double s = 0.0, wx = 0.0, wy = 0.0;
    for (int i = 0; i < e.aux2_n; i++, v += 4) {
        const double xx = v[0] - x, yy = v[1] - y;
        if (xx == 0.0 && yy == 0.0) {
            *zx = v[2];
            *zy = v[3];
            return;
        }
        const double r2 = xx * xx + yy * yy;
        if (r2 < e.roi2) {
            s = s + 1.0 / r2;
            wx = wx + v[2] / r2;
            wy = wy + v[3] / r2;
        }
    }
    if (s == 0.0) {
        *zx = 0.0;
        *zy = 0.0;
        return;
    }
    *zx = wx / s;
    *zy = wy / s;
}

// Finite-element surface at (x, y) (coretrace/vshot.cpp:166-238): the data
// nodes of the hash-tree zone holding the point and of its neighbours
// (coretrace/treemesh.cpp:209-245, 511-537, 757-800), the seven nearest of
// them, a kriging model over those with the power variogram
// alpha r^1.999, alpha fitted by least squares over the node pairs
// (coretrace/vshot.cpp:43-163), solved by LU with implicit partial pivoting,
// then the model's value at the point and its forward differences one step
// away in x and in y (the y step taken from the point itself).  Returns
// false where the tree has no zone, or fewer than three nodes, for the
// point (the reference throws there) or the kriging matrix has a zero row.
//
// aux + e.aux holds the flattened tree: {x0, x1, y0, y1, address bits,
// step, zones, nodes}, five doubles per zone {child 0, child 1, terminal,
// offset of its candidate list, candidates}, x y z per node, then the
// candidate lists (node indices: the zone's own nodes, then its
// neighbours').  Kept out of line: its arrays would otherwise weigh on the
// registers of every kernel that can meet such a surface.
__device__ __attribute__((noinline)) bool fe_interp(const DevElem &e, const double *aux, double x, double y,
                                                    double *zr, double *dzdx, double *dzdy) {
    const double *hd = aux + e.aux;
    const int nc = (int)hd[4], nz = (int)hd[6];
    const double step = hd[5];
    const double *zones = hd + 8, *nodes = zones + 5 * nz;
    // the zone: halve x and y alternately, down an address of nc digits,
    // stopping at a terminal zone; an 'x' split has both children equal
    double x0 = hd[0], x1 = hd[1], y0 = hd[2], y1 = hd[3];
    int zn = 0;
    for (int i = 0; i < nc; i++) {
        const double *z = zones + 5 * zn;
        if (z[2] != 0.0) break;
        bool up;
        if ((i & 1) == 0) {
            const double c = (x0 + x1) * 0.5;
            up = x > c;
            if (up) x0 = c; else x1 = c;
        } else {
            const double c = (y0 + y1) * 0.5;
            up = y > c;
            if (up) y0 = c; else y1 = c;
        }
        const int next = (int)(up ? z[1] : z[0]);
        if (next < 0) return false;
        zn = next;
    }
    const double *cand = aux + (int64_t)zones[5 * zn + 3];
    const int nn = (int)zones[5 * zn + 4];
    if (nn < 3) return false;
    // the seven nearest candidates, by distance, ties in list order
    double rk[7];
    int ik[7];
    int m = 0;
    for (int c = 0; c < nn; c++) {
        const int i = (int)cand[c];
        const double dx = x - nodes[3 * i], dy = y - nodes[3 * i + 1];
        const double r = ::sqrt(dx * dx + dy * dy);
        if (m == 7 && !(r < rk[6])) continue;
        int k = m < 7 ? m++ : 6;
        for (; k > 0 && rk[k - 1] > r; k--) {
            rk[k] = rk[k - 1];
            ik[k] = ik[k - 1];
        }
        rk[k] = r;
        ik[k] = i;
    }
    double px[7], py[7], pz[8];
    for (int k = 0; k < m; k++) {
        px[k] = nodes[3 * ik[k]];
        py[k] = nodes[3 * ik[k] + 1];
        pz[k] = nodes[3 * ik[k] + 2];
    }
    const double bet = 1.999;
    double num = 0.0, den = 0.0;
    for (int i = 0; i < m; i++)
        for (int j = i + 1; j < m; j++) {
            double rb = 0.0;
            rb += (px[i] - px[j]) * (px[i] - px[j]);
            rb += (py[i] - py[j]) * (py[i] - py[j]);
            rb = ::pow(rb, 0.5 * bet);
            num += rb * (0.5 * ((pz[i] - pz[j]) * (pz[i] - pz[j])) - 0.0);
            den += rb * rb;
        }
    const double alph = num / den;
    const int n = m + 1;
    double lu[8][8];
    for (int i = 0; i < m; i++) {
        for (int j = i; j < m; j++) {
            double d = 0.0;
            d += (px[i] - px[j]) * (px[i] - px[j]);
            d += (py[i] - py[j]) * (py[i] - py[j]);
            lu[i][j] = lu[j][i] = 0.0 + alph * ::pow(::sqrt(d), bet);
        }
        lu[i][m] = lu[m][i] = 1.0;
    }
    lu[m][m] = pz[m] = 0.0;
    int indx[8];
    double vv[8];
    for (int i = 0; i < n; i++) {
        double big = 0.0;
        for (int j = 0; j < n; j++) big = ::fabs(lu[i][j]) > big ? ::fabs(lu[i][j]) : big;
        if (big == 0.0) return false;
        vv[i] = 1.0 / big;
    }
    // the pivot row carries over from the previous column when none in this
    // one compares greater than zero (a zero or NaN column, as duplicate
    // nodes give), as in the reference (coretrace/interpolate.cpp:5-50)
    int imax = 0;
    for (int k = 0; k < n; k++) {
        double big = 0.0;
        for (int i = k; i < n; i++) {
            const double t = vv[i] * ::fabs(lu[i][k]);
            if (t > big) { big = t; imax = i; }
        }
        if (k != imax) {
            for (int j = 0; j < n; j++) {
                const double t = lu[imax][j];
                lu[imax][j] = lu[k][j];
                lu[k][j] = t;
            }
            vv[imax] = vv[k];
        }
        indx[k] = imax;
        if (lu[k][k] == 0.0) lu[k][k] = 1.0e-40;
        for (int i = k + 1; i < n; i++) {
            const double t = lu[i][k] /= lu[k][k];
            for (int j = k + 1; j < n; j++) lu[i][j] -= t * lu[k][j];
        }
    }
    double w[8];
    for (int i = 0; i < n; i++) w[i] = pz[i];
    int ii = 0;
    for (int i = 0; i < n; i++) {
        const int ip = indx[i];
        double sum = w[ip];
        w[ip] = w[i];
        if (ii != 0)
            for (int j = ii - 1; j < i; j++) sum -= lu[i][j] * w[j];
        else if (sum != 0.0)
            ii = i + 1;
        w[i] = sum;
    }
    for (int i = n - 1; i >= 0; i--) {
        double sum = w[i];
        for (int j = i + 1; j < n; j++) sum -= lu[i][j] * w[j];
        w[i] = sum / lu[i][i];
    }
    const double qx[3] = {x, x + step, x + step - step}, qy[3] = {y, y, y + step};
    double v[3];
    for (int q = 0; q < 3; q++) {
        double s = 0.0;
        for (int i = 0; i < m; i++) {
            double d = 0.0;
            d += (qx[q] - px[i]) * (qx[q] - px[i]);
            d += (qy[q] - py[i]) * (qy[q] - py[i]);
            s += w[i] * (0.0 + alph * ::pow(::sqrt(d), bet));
        }
        s += w[m] * 1.0;
        v[q] = s;
    }
    *zr = v[0];
    *dzdx = (v[0] - v[1]) / step;
    *dzdy = (v[0] - v[2]) / step;
    return true;
}

// F and grad F for Newton (coretrace/surface.cpp:57-413).  X[2] is replaced
// by the surface z for the tabulated surfaces, as in the reference.  G = 2 is the
// kernel variant for stages holding a finite-element surface: only there is
// its interpolation compiled in (its scratch arrays would otherwise cost the
// other generic kernels registers).
template <typename R, int G>
STD int surface_eval(const DevElem &e, const double *aux, R X[3], R *F, R DF[3]) {
    R x = X[0], y = X[1], z = X[2];
    switch (e.surf) {
    case SURF_FLAT:
        *F = z;
        DF[0] = 0; DF[1] = 0; DF[2] = 1;
        return 0;
    case SURF_GENERAL: {
        R rho2 = (e.stype == 1) ? x * x + y * y : x * x;
        if (e.cone_tan != 0.0) {
            R sr = M<R>::sqrt_(rho2);
            R ct = (R)e.cone_tan;
            *F = z - sr / ct;
            DF[0] = -x / (sr * ct);
            DF[1] = -y / (sr * ct);
            DF[2] = 1;
            return 0;
        }
        R cx = (R)e.vcx, cy = (R)e.vcy, k = (R)e.kappa;
        R arg = R(1) - k * (cx * cx * x * x + cy * cy * y * y);
        int err = (arg < R(0)) ? 1 : 0;
        R term = M<R>::sqrt_(arg);
        *F = z - (cx * x * x + cy * y * y) / (R(1) + term);
        DF[0] = -x * (cx / term);
        DF[1] = -y * (cy / term);
        DF[2] = 1;
        return err;
    }
    case SURF_ZERNIKE: {
        R zz, dx, dy;
        if constexpr (sizeof(R) == 4) zernike_horner<R>(e, aux, x, y, &zz, &dx, &dy);
        else zernike<R>(e, aux, x, y, &zz, &dx, &dy, true);
        X[2] = zz;
        *F = z - zz;
        DF[0] = -dx; DF[1] = -dy; DF[2] = 1;
        return 0;
    }
    case SURF_POLY: {
        R yv = (e.ap == AP_LINE) ? R(0) : y;
        R zz, dx, dy;
        polyrev<R>(e, aux, x, yv, &zz, &dx, &dy);
        X[2] = zz;
        *F = z - zz;
        DF[0] = -dx; DF[1] = -dy; DF[2] = 1;
        return 0;
    }
    case SURF_VSHOT: {
        // z from the monomial fit, the normal from the interpolated measured
        // slopes (coretrace/surface.cpp:244-298)
        if (x * x + y * y == R(0)) {
            *F = z;
            DF[0] = 0; DF[1] = 0; DF[2] = 1;
            return 0;
        }
        R zz, dx, dy;
        zernike<R>(e, aux, x, y, &zz, &dx, &dy, false);
        double zx, zy;
        vshot_slopes(e, aux, (double)x, (double)y, &zx, &zy);
        X[2] = zz;
        *F = z - zz;
        DF[0] = (R)-::tan(zx); DF[1] = (R)-::tan(zy); DF[2] = 1;
        return 0;
    }
    case SURF_FE: {
        // kriging over the finite-element nodes (coretrace/surface.cpp:218-242);
        // 4 where the interpolation fails
        if (x * x + y * y == R(0)) {
            *F = z;
            DF[0] = 0; DF[1] = 0; DF[2] = 1;
            return 0;
        }
        if constexpr (G != 2) {
            return 4;  // not in this variant (the host never sends it here)
        } else {
            double zz, dx, dy;
            if (!fe_interp(e, aux, (double)x, (double)y, &zz, &dx, &dy)) return 4;
            X[2] = (R)zz;
            *F = z - (R)zz;
            DF[0] = (R)dx; DF[1] = (R)dy; DF[2] = 1;
            return 0;
        }
    }
    default: {  // SURF_SPLINE
        R rho = M<R>::sqrt_(x * x + y * y);
        R drdx = x / rho, drdy = y / rho;
        if (e.ap == AP_LINE) { rho = x; drdx = 1; drdy = 0; }
        const int n = e.aux_n;
        const double *sx = aux + e.aux;
        R zz, dz;
        if (!splint<R>(sx, sx + n, sx + 2 * n, n, rho, &zz, &dz)) return 3;
        X[2] = zz;
        *F = z - zz;
        DF[0] = -(dz * drdx); DF[1] = -(dz * drdy); DF[2] = 1;
        return 0;
    }
    }
}

// z of the surface at (x, y) for the start-plane choice
// (coretrace/surfacezatxypair.cpp:56-280), double.  HORNER evaluates the
// monomial surface in Horner form (the FP32 path: the start plane only
// picks where Newton starts, and a 1-ulp difference of z cannot change that
// choice for a ray nudged 1e-5 m off its last hit).
template <bool HORNER = false>
STD bool surface_z_at_xy(const DevElem &e, const double *aux, double x, double y, double *F) {
    if constexpr (HORNER) {
        if (e.surf == SURF_ZERNIKE) {
            double z, dx, dy;
            zernike_horner<double>(e, aux, x, y, &z, &dx, &dy);
            *F = z;
            return true;
        }
    }
    switch (e.surf) {
    case SURF_GENERAL: {
        double rho2 = (e.stype == 1) ? x * x + y * y : x * x;
        if (e.cone_tan != 0.0) { *F = ::sqrt(rho2) / e.cone_tan; return true; }
        double q = e.kappa * (e.vcx * e.vcx * x * x + e.vcy * e.vcy * y * y);
        if (q > 1.0) { *F = 0.0; return true; }
        *F = (e.vcx * x * x + e.vcy * y * y) / (1.0 + ::sqrt(1.0 - q));
        return true;
    }
    case SURF_ZERNIKE: {
        double z, dx, dy;
        zernike<double>(e, aux, x, y, &z, &dx, &dy, false);
        *F = z;
        return true;
    }
    case SURF_VSHOT: {  // coretrace/surfacezatxypair.cpp:195-207
        double z = 0.0, dx, dy;
        if (x * x + y * y != 0.0) zernike<double>(e, aux, x, y, &z, &dx, &dy, false);
        *F = z;
        return true;
    }
    case SURF_POLY: {
        double yv = (e.ap == AP_LINE) ? 0.0 : y;
        double z, dx, dy;
        polyrev<double>(e, aux, x, yv, &z, &dx, &dy);
        *F = z;
        return true;
    }
    default:
        return false;
    }
}

// coretrace/newzstartforcubicsplinesurf.cpp:57-142
STD void spline_cyl(double CR, const double P[3], const double D[3], bool isA, double *zs, int *ef) {
    *ef = 0;
    if (isA) {
        double A = D[0] * D[0] + D[1] * D[1];
        double B = 2.0 * (P[0] * D[0] + P[1] * D[1]);
        double C = P[0] * P[0] + P[1] * P[1] - CR * CR;
        if (B * B > 4.0 * A * C) {
            double t1 = (-B + ::sqrt(B * B - 4.0 * A * C)) / (2.0 * A);
            double t2 = (-B - ::sqrt(B * B - 4.0 * A * C)) / (2.0 * A);
            if (t2 > 0) { *zs = P[2] + t2 * D[2]; return; }
            if (t2 == 0) { *zs = P[2] + t1 * D[2]; return; }
            if (t2 < 0 && t1 > 0) { *zs = P[2] + t1 * D[2]; return; }
            *ef = 1;
        } else {
            *ef = 2;
        }
    } else {
        double t1 = (CR - P[0]) / D[0];
        if (t1 >= 0.0) *zs = P[2] + t1 * D[2]; else *ef = 1;
    }
}

// Newton start plane for cubic-spline surfaces (coretrace/intersect.cpp:637-946).
STD bool spline_start(const DevElem &e, const double *aux, const double X[3], const double D[3], double *ZStart) {
    const int n = e.aux_n;
    const double *sx = aux + e.aux, *sy = sx + n, *sy2 = sx + 2 * n;
    const double OuterR = sx[n - 1], InnerR = sx[0];
    const bool isA = e.ap == AP_ANNULUS, isL = e.ap == AP_LINE;
    const double ZA = sy[n - 1];
    double S00 = -X[2] / (D[2] + 0.00000000001);
    double X10 = X[0] + D[0] * S00, Y10 = X[1] + D[1] * S00;
    double R10 = ::sqrt(X10 * X10 + Y10 * Y10);
    double S0A = (ZA - X[2]) / (D[2] + 0.00000000001);
    double X1A = X[0] + D[0] * S0A, Y1A = X[1] + D[1] * S0A;
    double R1A = ::sqrt(X1A * X1A + Y1A * Y1A);
    double zcs = 0;
    int ef = 0;
#define ST_CYL(CR, FROM)                                   \
    do {                                                   \
        spline_cyl((CR), (FROM), D, isA, &zcs, &ef);       \
        if (ef == 0) *ZStart = zcs;                        \
        return true;                                       \
    } while (0)
    const double PA[3] = {X1A, Y1A, ZA}, PZ[3] = {X10, Y10, 0.0};
    if (X[2] >= ZA && D[2] < 0.0) {
        if ((isA && R1A > OuterR) || (isL && X1A > OuterR)) ST_CYL(OuterR * 0.999999, PA);
        if ((isA && R1A <= OuterR) || (isL && X1A <= OuterR)) {
            if ((isA && R10 >= OuterR) || (isL && X10 >= OuterR)) ST_CYL(OuterR * 0.999999, PA);
            *ZStart = 0.0;
            return true;
        }
    }
    if (X[2] <= 0.0 && D[2] > 0.0) {
        if ((isA && R10 >= OuterR) || (isL && X10 >= OuterR)) ST_CYL(OuterR * 0.999999, PZ);
        if ((isA && R10 < OuterR && R10 > InnerR) || (isL && X10 < OuterR && X10 > InnerR)) {
            *ZStart = 0.0;
            return true;
        }
        if ((isA && R10 <= InnerR) || (isL && X10 <= InnerR)) ST_CYL(InnerR / 0.999999, PZ);
    }
    if (X[2] < ZA && X[2] > 0.0 && D[2] < 0.0) {
        double R1 = ::sqrt(X[0] * X[0] + X[1] * X[1]);
        if ((isA && R1 >= OuterR) || (isL && X[0] >= OuterR)) ST_CYL(OuterR * 0.999999, X);
        if ((isA && R1 < OuterR && R1 > InnerR) || (isL && X[0] < OuterR && X[0] > InnerR)) {
            double Z1, dz;
            if (!splint<double>(sx, sy, sy2, n, isA ? R1 : X[0], &Z1, &dz)) return false;
            if (Z1 < X[2]) {
                if ((isA && R10 >= OuterR) || (isL && X10 >= OuterR)) ST_CYL(OuterR * 0.999999, X);
                *ZStart = 0.0;
                return true;
            }
            if (Z1 >= X[2]) *ZStart = X[2];
            return true;
        }
        if ((isA && R1 <= InnerR) || (isL && X[0] <= InnerR)) ST_CYL(InnerR / 0.999999, X);
    }
    if (X[2] < ZA && X[2] > 0.0 && D[2] > 0.0) {
        double R1 = ::sqrt(X[0] * X[0] + X[1] * X[1]);
        if ((isA && R1 >= OuterR) || (isL && X[0] >= OuterR)) ST_CYL(OuterR * 0.999999, X);
        if ((isA && R1 < OuterR && R1 > InnerR) || (isL && X[0] < OuterR && X[0] > InnerR)) {
            double Z1, dz;
            if (!splint<double>(sx, sy, sy2, n, isA ? R1 : X[0], &Z1, &dz)) return false;
            if (Z1 < X[2]) {
                if ((isA && R1A >= OuterR) || (isL && X1A >= OuterR)) ST_CYL(OuterR * 0.999999, X);
                return true;
            }
            if (Z1 >= X[2]) {
                if ((isA && R10 < OuterR) || (isL && X10 < OuterR)) {
                    *ZStart = 0.0;
                    return true;
                }
                if ((isA && R10 >= OuterR) || (isL && X10 >= OuterR)) ST_CYL(OuterR * 0.999999, PZ);
            }
        }
        if ((isA && R1 <= InnerR) || (isL && X[0] <= InnerR)) ST_CYL(OuterR * 0.999999, X);
    }
#undef ST_CYL
    return true;
}

// ------------------------------------------------------ element hit -------
template <typename R>
struct ElemHit {
    double path;      // path length from the nudged element-frame origin
    R X[3];           // hit point, element frame
    R DF[3];          // surface normal (flipped on back-side hits)
    int back;
};

// Ray vs one element: Intersect + DetermineElementIntersectionNew
// (coretrace/intersect.cpp:140-1011, determineelementintersectionnew.cpp:58-633).
// P, D: the nudged ray in the element frame (double).  Returns true on an
// intercept (before the caller's closest-hit / ZAperture rule); on a miss
// where a finite-element interpolation failed, h.back is set to 2.  G = false
// is the variant for stages whose elements are all closed-form quadrics or
// flats: the torus and Newton code is left out of the kernel.
template <typename R, int G>
STD bool element_hit(const DevElem &e, const double *aux, const double P[3], const double D[3], ElemHit<R> &h) {
    const int s = e.surf;
    R X[3], DF[3];
    double path;
    bool ok;
    if (s == SURF_FLAT) {
        // The reference runs its Newton iteration on F = z from the start
        // plane z = 0 (coretrace/intersect.cpp:589-592, 945-1011), which
        // stops at the first iterate: that iterate in closed form.
        const double S0 = (0.0 - P[2] == 0.0) ? 0.0 : (0.0 - P[2]) / (D[2] + 0.00000000001);
        X[0] = (R)(P[0] + D[0] * S0);
        X[1] = (R)(P[1] + D[1] * S0);
        X[2] = R(0);
        DF[0] = R(0); DF[1] = R(0); DF[2] = R(1);
        path = S0 + 0.0;
        ok = !(path < 0.0);
    } else if (!G && s > SURF_CYLINDER) {
        return false;  // not in this variant (the host never sends it here)
    } else if (s == SURF_TORUS) {
        // The reference quartic solve (root432) is ill-conditioned for an
        // origin kilometres away (a sun ray starts 10 km back): there a
        // 1-ulp change of the origin flips the outcome of a sizeable share
        // of rays.  For such origins the solve starts where the ray enters
        // the sphere of radius 2r+1 about the element's hit-bounding sphere;
        // every real intersection lies inside, so no hit is skipped.  Nearer
        // origins run the reference arithmetic unchanged.
        double t0 = 0.0;
        {
            double d[3] = {P[0] - e.bsphere_c[0], P[1] - e.bsphere_c[1], P[2] - e.bsphere_c[2]};
            double tc = -(d[0] * D[0] + d[1] * D[1] + d[2] * D[2]);
            double dist2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            double rr = 2.0 * e.bsphere_r + 1.0;
            if (dist2 > rr * rr && tc > rr) t0 = tc - rr;
        }
        const double Ps[3] = {P[0] + t0 * D[0], P[1] + t0 * D[1], P[2] + t0 * D[2]};
        double Xd[3] = {0.0, 0.0, 0.0}, DFd[3] = {0.0, 0.0, 0.0}, L = 0.0;
        ok = torus_shell(e, t0 != 0.0 ? Ps : P, D) && torus(e, t0 != 0.0 ? Ps : P, D, Xd, DFd, &L);
        path = t0 + L;
        if (ok && path < 0.0) ok = false;  // coretrace/determineelementintersectionnew.cpp:86-100
        for (int i = 0; i < 3; i++) { X[i] = (R)Xd[i]; DF[i] = (R)DFd[i]; }
    } else if (s <= SURF_CYLINDER) {
        // closed-form quadrics; a float ray starting far away is first moved
        // along itself to the element's bounding sphere
        double t0 = 0.0;
        if (sizeof(R) == 4) {
            double d[3] = {P[0] - e.bsphere_c[0], P[1] - e.bsphere_c[1], P[2] - e.bsphere_c[2]};
            double tc = -(d[0] * D[0] + d[1] * D[1] + d[2] * D[2]);
            double dist2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            double rr = 2.0 * e.bsphere_r + 1.0;
            if (dist2 > rr * rr && tc > rr) t0 = tc - rr;
        }
        R Pr[3] = {(R)(P[0] + t0 * D[0]), (R)(P[1] + t0 * D[1]), (R)(P[2] + t0 * D[2])};
        R Dr[3] = {(R)D[0], (R)D[1], (R)D[2]};
        R t;
        ok = quadric<R>(e, Pr, Dr, X, DF, &t);
        path = t0 + (double)t;
        if (ok && !(path >= 0.0)) ok = false;
    } else {
        // Newton-Raphson: start plane in double, iterations in R
        double S0Ap = (e.zap - P[2] == 0.0) ? 0.0 : (e.zap - P[2]) / (D[2] + 0.00000000001);
        bool za_in = aperture_contains<double>(e, P[0] + D[0] * S0Ap, P[1] + D[1] * S0Ap, true);
        double ZStart = 0.0;
        if (s == SURF_SPLINE) {
            if (!spline_start(e, aux, P, D, &ZStart)) return false;
        } else if (s != SURF_FLAT) {
            double FXY = 0.0;
            surface_z_at_xy<sizeof(R) == 4>(e, aux, P[0], P[1], &FXY);
            if (P[2] <= 0.0 && D[2] > 0.0) ZStart = 0.0;
            else if (P[2] <= FXY && D[2] > 0.0) ZStart = 0.0;
            else if (P[2] <= FXY && D[2] < 0.0 && P[2] > e.zap && za_in) ZStart = 0.0;
            else if (P[2] <= FXY && D[2] < 0.0) ZStart = e.zap;
            else if (P[2] > FXY && D[2] < 0.0) ZStart = 0.0;
            else if (P[2] > FXY && D[2] > 0.0) ZStart = e.zap;
        }
        double S0 = (ZStart - P[2] == 0.0) ? 0.0 : (ZStart - P[2]) / (D[2] + 0.00000000001);
        const R X1 = (R)(P[0] + D[0] * S0), Y1 = (R)(P[1] + D[1] * S0), Z1 = (R)ZStart;
        const R Dr[3] = {(R)D[0], (R)D[1], (R)D[2]};
        R lower = R(-1e10), upper = R(1e10);
        if (s == SURF_SPLINE) {
            const double *sx = aux + e.aux;
            R a = ((R)sx[0] - X1) / (Dr[0] + R(0.00000000001));
            R b = ((R)sx[e.aux_n - 1] - X1) / (Dr[0] + R(0.00000000001));
            lower = a < b ? a : b;
            upper = a < b ? b : a;
        }
        R SJ1 = 0, SJ = 0;
        ok = false;
        const R eps = R(0.000001);
        for (int it = 0; it < 20; ++it) {
            SJ = SJ1;
            X[0] = X1 + Dr[0] * SJ;
            X[1] = Y1 + Dr[1] * SJ;
            X[2] = Z1 + Dr[2] * SJ;
            R F;
            const int code = surface_eval<R, G>(e, aux, X, &F, DF);
            if (code != 0) {
                if (code == 4) h.back = 2;  // finite-element interpolation failed
                break;
            }
            R dfd = dot3<R>(DF, Dr);
            if (M<R>::fabs_(F) <= eps * M<R>::fabs_(dfd)) { ok = true; break; }
            SJ1 = SJ - F / dfd;
            if (s == SURF_SPLINE) {
                if ((F < 0 && Dr[2] > 0) || (F > 0 && Dr[2] < 0)) lower = SJ > lower ? SJ : lower;
                else if ((F < 0 && Dr[2] < 0) || (F > 0 && Dr[2] > 0)) upper = SJ < upper ? SJ : upper;
                if (SJ1 < lower || SJ1 > upper) SJ1 = R(0.5) * (lower + upper);
                if (upper <= lower) break;
            }
        }
        path = S0 + (double)SJ;
        if (ok && path < 0.0) ok = false;
    }
    if (!ok) return false;
    if (!aperture_contains<R>(e, X[0], X[1], false)) return false;
    R Dr[3] = {(R)D[0], (R)D[1], (R)D[2]};
    h.back = (dot3<R>(Dr, DF) < R(0)) ? 0 : 1;
    if (h.back) { DF[0] = -DF[0]; DF[1] = -DF[1]; DF[2] = -DF[2]; }
    for (int i = 0; i < 3; i++) { h.X[i] = X[i]; h.DF[i] = DF[i]; }
    h.path = path;
    return true;
}

// Double-precision finish of a hit the FP32 search found: the point is
// moved onto the surface along the ray (path from the float solve, then one
// Newton step on the surface function, which squares the float error away)
// and the normal is evaluated there, flipped towards the incoming ray as
// element_hit flips it.  A float hit point is off by ~1e-6 m, which tilts a
// heliostat normal (curvature radius ~1 km) by ~1e-9 rad; reflected towards
// a receiver hundreds of metres away at grazing incidence that grows to
// 1e-4 m.  P, D: the nudged ray in the element frame.
// Returns false where a finite-element interpolation failed.
template <int G>
STD bool hit_finish_d(const DevElem &e, const double *aux, const double P[3], const double D[3], double path,
                      double X[3], double DF[3], int *back) {
    bool ok = true;
    const int s = e.surf;
    if (s == SURF_FLAT || (!G && s > SURF_CYLINDER)) {
        // the reference's start plane z = 0, taken as the hit
        const double S0 = (0.0 - P[2] == 0.0) ? 0.0 : (0.0 - P[2]) / (D[2] + 0.00000000001);
        X[0] = P[0] + D[0] * S0; X[1] = P[1] + D[1] * S0; X[2] = 0.0;
        DF[0] = 0.0; DF[1] = 0.0; DF[2] = 1.0;
    } else if (s == SURF_TORUS) {
        for (int i = 0; i < 3; i++) X[i] = P[i] + path * D[i];
        torus_normal(e, X, DF);
    } else if (s <= SURF_CYLINDER) {
        for (int i = 0; i < 3; i++) X[i] = P[i] + path * D[i];
        // F and grad F of the implicit quadric (cancellation-free forms)
        double F, Gr[3];
        const double x = X[0], y = X[1], z = X[2];
        if (s == SURF_PARABOLA) {
            const double a2 = 2.0 / e.vcx, b2 = e.vcy > 0.0 ? 2.0 / e.vcy : 1e10;
            const double ky = (e.stype == 7) ? 0.0 : 1.0;
            F = x * x / a2 + ky * y * y / b2 - z;
            Gr[0] = 2.0 * x / a2; Gr[1] = 2.0 * ky * y / b2; Gr[2] = -1.0;
        } else if (s == SURF_CONIC) {
            F = x * x + y * y + e.kappa * z * z - 2.0 * z / e.vcx;
            Gr[0] = 2.0 * x; Gr[1] = 2.0 * y; Gr[2] = 2.0 * e.kappa * z - 2.0 / e.vcx;
        } else {
            const double zc = (s == SURF_SPHERE) ? 1.0 / e.vcx : 1.0 / e.curv_rev;
            const double ky = (s == SURF_CYLINDER || e.stype == 7) ? 0.0 : 1.0;
            F = x * x + ky * y * y + z * (z - 2.0 * zc);
            Gr[0] = 2.0 * x; Gr[1] = 2.0 * ky * y; Gr[2] = 2.0 * (z - zc);
        }
        const double gd = Gr[0] * D[0] + Gr[1] * D[1] + Gr[2] * D[2];
        if (gd != 0.0) {
            const double t = path - F / gd;
            for (int i = 0; i < 3; i++) X[i] = P[i] + t * D[i];
        }
        quadric_normal<double>(e, X, DF);
    } else {
        // Newton surfaces: one more iteration in double from the float
        // iterate, then the surface z and slopes at the final point
        double Xc[3], F;
        for (int i = 0; i < 3; i++) Xc[i] = P[i] + path * D[i];
        ok = surface_eval<double, G>(e, aux, Xc, &F, DF) != 4;
        const double dfd = DF[0] * D[0] + DF[1] * D[1] + DF[2] * D[2];
        const double t = (ok && dfd != 0.0) ? path - F / dfd : path;
        for (int i = 0; i < 3; i++) X[i] = P[i] + t * D[i];
        ok = ok && surface_eval<double, G>(e, aux, X, &F, DF) != 4;
        if (!ok) { DF[0] = 0.0; DF[1] = 0.0; DF[2] = 1.0; }
    }
    *back = (dot3<double>(D, DF) < 0.0) ? 0 : 1;
    if (*back) { DF[0] = -DF[0]; DF[1] = -DF[1]; DF[2] = -DF[2]; }
    return ok;
}

// ------------------------------------------------------------ optics ------
// Everything after the closest hit runs on double directions in both
// precisions: a direction error of 1e-7 rad already moves a receiver hit by
// 0.1 mm over a kilometre of flight, and on CDNA4 double add/multiply/FMA
// issue at the non-packed float rate.  FAST (the FP32 path) differs only in
// how it gets there: ray frames are built algebraically instead of through
// atan2/sin/cos, and the sines and cosines of the small perturbation angles,
// whose relative error is what matters, are float.

// 1/sqrt(x) in double: the hardware estimate refined by two Newton steps.
STD double rsqrt_d(double x) {
    double y = __builtin_amdgcn_rsq(x);
    y = y * (1.5 - 0.5 * x * y * y);
    y = y * (1.5 - 0.5 * x * y * y);
    return y;
}

// Rotation taking the local z axis onto direction c, stored local ->
// reference (coretrace/errors.cpp:98-121, surfacenormalerrors.cpp:87-106:
// Euler angles alpha = atan2(cx, cz), beta = atan2(cy, sqrt(cx^2 + cz^2)),
// gamma = 0).  The algebraic form uses cos(alpha) = cz/h, sin(alpha) = cx/h,
// cos(beta) = h/n, sin(beta) = cy/n (h = sqrt(cx^2 + cz^2), n = |c|), and
// keeps the reference's special cases for cz == 0 (alpha fixed at pi/2, or
// alpha = 0 and beta = pi/2 when cx == 0 too).
template <bool FAST>
STD void ray_frame(const double c[3], double RT[9]) {
    double ca, cb, sa, sb;
    if constexpr (FAST) {
        if (c[2] == 0.0) {
            if (c[0] == 0.0) { ca = 1.0; sa = 0.0; cb = 0.0; sb = 1.0; }
            else {
                const double h = ::fabs(c[0]), in = rsqrt_d(h * h + c[1] * c[1]);
                ca = 0.0; sa = 1.0; cb = h * in; sb = c[1] * in;
            }
        } else {
            const double h2 = c[0] * c[0] + c[2] * c[2];
            const double ih = rsqrt_d(h2), in = rsqrt_d(h2 + c[1] * c[1]);
            ca = c[2] * ih; sa = c[0] * ih; cb = h2 * ih * in; sb = c[1] * in;
        }
    } else {
        double e0, e1;
        if (c[2] == 0.0) {
            if (c[0] == 0.0) { e0 = 0.0; e1 = M_PI / 2.0; }
            else { e0 = M_PI / 2.0; e1 = ::atan2(c[1], ::sqrt(c[0] * c[0] + c[2] * c[2])); }
        } else {
            e0 = ::atan2(c[0], c[2]);
            e1 = ::atan2(c[1], ::sqrt(c[0] * c[0] + c[2] * c[2]));
        }
        ::sincos(e0, &sa, &ca);
        ::sincos(e1, &sb, &cb);
    }
    RT[0] = ca;       RT[1] = -sa * sb; RT[2] = sa * cb;
    RT[3] = 0;        RT[4] = cb;       RT[5] = sb;
    RT[6] = -sa;      RT[7] = -ca * sb; RT[8] = ca * cb;
}

// Direction at polar angle sqrt(theta2) and azimuth 2 pi u about the frame's
// z axis.  FAST: sin(theta)/theta and (1 - cos theta)/theta^2 from their
// Taylor series in theta^2 (truncation below 1e-9 for theta <= pi/2; larger
// angles take sinf/cosf), cos(theta) = 1 - theta^2 (1 - cos)/theta^2 formed
// in double, azimuth from sinpi/cospi of 2u.
template <bool FAST>
STD void cone_dir(double theta2, double u, const double RT[9], double out[3]) {
    double l0, l1, l2;
    if constexpr (FAST) {
        const float t2 = (float)theta2;
        const float th = ::sqrtf(t2);
        float st, vc;  // sin(theta), 1 - cos(theta)
        if (t2 <= 2.5f) {
            const float sinc =
                1.0f - t2 / 6.0f * (1.0f - t2 / 20.0f * (1.0f - t2 / 42.0f * (1.0f - t2 / 72.0f *
                (1.0f - t2 / 110.0f * (1.0f - t2 / 156.0f)))));
            const float vers =
                0.5f * (1.0f - t2 / 12.0f * (1.0f - t2 / 30.0f * (1.0f - t2 / 56.0f * (1.0f - t2 / 90.0f *
                (1.0f - t2 / 132.0f * (1.0f - t2 / 182.0f))))));
            st = th * sinc;
            vc = t2 * vers;
        } else {
            st = ::sinf(th);
            vc = 1.0f - ::cosf(th);
        }
        float sp, cp;
        ::sincospif((float)(2.0 * u), &sp, &cp);
        l0 = (double)(st * cp);
        l1 = (double)(st * sp);
        l2 = 1.0 - (double)vc;
    } else {
        const double phi = u * 2.0 * 3.1415926535897932385;
        const double th = ::sqrt(theta2);
        double st, ct, sp, cp;
        ::sincos(th, &st, &ct);
        ::sincos(phi, &sp, &cp);
        l0 = st * cp; l1 = st * sp; l2 = ct;
    }
    out[0] = RT[0] * l0 + RT[1] * l1 + RT[2] * l2;
    out[1] = RT[3] * l0 + RT[4] * l1 + RT[5] * l2;
    out[2] = RT[6] * l0 + RT[7] * l1 + RT[8] * l2;
}

// First index i with table[i] >= theta, capped at n-1 (the linear walk of
// coretrace/errors.cpp:167-169 on an ascending table).
STD int table_index(const double *ang, int n, double theta, bool ascending) {
    if (ascending) {
        int lo = 0, hi = n - 1;
        while (lo < hi) {
            int mid = (lo + hi) >> 1;
            if (ang[mid] < theta) lo = mid + 1; else hi = mid;
        }
        return lo;
    }
    int i = 0;
    while (i < n - 1 && ang[i] < theta) i++;
    return i;
}

// The 'd' sunshape rejection loop of coretrace/errors.cpp:160-185, split
// in two.  A try proposes (tx, ty) uniform over the square of half-side
// max_angle and accepts with probability profile(theta)/max_intensity, and
// only inside the circle of radius max_angle.  sunshape_d_screen rejects on
// that circle and on a precomputed upper envelope of the profile per
// polar-angle bin (padded by more than the float error of the bin index),
// so it only ever rejects what the exact test rejects; sunshape_d_exact is
// the reference's table walk and test.  theta^2 (t2) is in mrad^2.
// The draws enter unscaled: 2 ma (w 2^-32) is the exact product w (2 ma
// 2^-32), since scaling by 2^-32 is exact, so tx and ty round as the
// reference's 2 ma u - ma does; the envelope holds its bounds times 2^32, so
// u2 > bound compares as w2 > bound 2^32, exactly.
STD bool sunshape_d_screen(uint32_t w0, uint32_t w1, uint32_t w2, const DevSun &sun, const double *aux,
                           double *t2out) {
    const double ma = sun.max_angle;
    const double sc = 2.0 * ma * 2.3283064365386963e-10;
    const double tx = (double)w0 * sc - ma;
    const double ty = (double)w1 * sc - ma;
    const double t2 = tx * tx + ty * ty;
    *t2out = t2;
    if (t2 > ma * ma) return false;
    if (sun.env_n > 0) {
        int b = (int)(::sqrtf((float)t2) * (float)sun.env_scale);
        b = b < 0 ? 0 : (b >= sun.env_n ? sun.env_n - 1 : b);
        if ((double)w2 > aux[sun.env + b]) return false;
    }
    return true;
}

STD bool sunshape_d_exact(double t2, double u2, const DevSun &sun, const double *aux) {
    const int n = sun.tab_n;
    const double *ang = aux + sun.tab, *inten = ang + n;
    const double th = ::sqrt(t2);
    const int i = table_index(ang, n, th, sun.tab_ascending != 0);
    double stest;
    if (i == 0) stest = inten[0];
    else stest = inten[i - 1] + (inten[i] - inten[i - 1]) * (th - ang[i - 1]) / (ang[i] - ang[i - 1]);
    if (u2 > stest / sun.max_intensity) return false;
    if (t2 > sun.max_angle * sun.max_angle) return false;
    return true;
}

STD bool sunshape_d_try(uint32_t w0, uint32_t w1, uint32_t w2, const DevSun &sun, const double *aux,
                        double *t2out) {
    return sunshape_d_screen(w0, w1, w2, sun, aux, t2out) && sunshape_d_exact(*t2out, word_uniform(w2), sun, aux);
}

// Group q of the 'd' loop for the stream (seed, stage 0, sun_index) whose
// loop starts at draw k0 (try j uses draws k0+3j .. k0+3j+2): tries
// 4q - shift .. 4q - shift + 3, where shift = 3 k0 mod 4 puts the group's
// twelve draws on three whole Philox blocks (tries before the first are
// skipped).  Returns the first accepted try of the group as 0..3, or -1;
// *theta2 is its (mrad^2) sample.  The screens run first and branch-free; the
// exact test then runs only on the tries that survive them, in order.
STD uint32_t sunshape_d_shift(uint32_t k0) { return (3u * k0) & 3u; }

STD int sunshape_d_group(uint32_t seed, uint64_t sun_index, uint32_t k0, uint32_t q, const DevSun &sun,
                         const double *aux, double *theta2) {
    const int64_t j0 = 4 * (int64_t)q - (int64_t)sunshape_d_shift(k0);
    const uint32_t blk0 = (uint32_t)((k0 + 3 * j0) >> 2);  // k0 + 3 j0 is a multiple of 4 and >= 0 for q >= 1
    uint32_t w[12];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        uint32_t b[4] = {(uint32_t)sun_index, (uint32_t)(sun_index >> 32), 0u, blk0 + (uint32_t)c};
        philox10(b, seed, 0x53544F52u);
#pragma unroll
        for (int i = 0; i < 4; ++i) w[4 * c + i] = b[i];
    }
    double t2s[4];
    uint32_t cand = 0;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        if (j0 + t < 0) continue;
        if (sunshape_d_screen(w[3 * t], w[3 * t + 1], w[3 * t + 2], sun, aux, &t2s[t])) cand |= 1u << t;
    }
    while (cand != 0u) {
        const int t = __ffs(cand) - 1;
        cand &= cand - 1u;
        const double t2 = t == 0 ? t2s[0] : (t == 1 ? t2s[1] : (t == 2 ? t2s[2] : t2s[3]));
        const double u2 = word_uniform(t == 0 ? w[2] : (t == 1 ? w[5] : (t == 2 ? w[8] : w[11])));
        if (sunshape_d_exact(t2, u2, sun, aux)) {
            *theta2 = t2;
            return t;
        }
    }
    return -1;
}

// The 'd' sunshape's polar angle by inversion, one 53-bit uniform per ray.
// The reference's rejection loop (coretrace/errors.cpp:160-185) proposes
// (tx, ty) uniform on the square of half-side max_angle and keeps it with
// probability I(theta)/I_max inside the circle, I the table's linear
// interpolation: the kept theta has density proportional to I(theta) theta
// on [0, max_angle].  On each table segment that density is a quadratic in
// x = theta - lo, (a + b x)(lo + x), and its integral from lo the cubic
// G(x) = x (p0 + x (p1 + x p2)).  The host lays the segments out in aux, 8
// doubles each: lo, width w, a, b, p0, p1, p2, and C, the mass of the
// segments before it; the total mass follows the last segment.  The draw
// picks the segment holding u * total by bisection over C and solves G(x) = r
// by Newton steps kept inside a shrinking bracket, until a step is within a
// few ulps of theta (about 3 steps).  Only
// +, -, *, / are used, so an FP64 build without contraction gives the same
// bits as the host.  Returns theta^2 in mrad^2.
STD double sunshape_icdf_solve(const double *seg, double r) {
    const double w = seg[1], p0 = seg[4], p1 = seg[5], p2 = seg[6];
    const double a = seg[2], b = seg[3], lo = seg[0];
    const double m = w * (p0 + w * (p1 + w * p2));
    if (!(r < m)) r = m;
    if (!(r > 0.0)) r = 0.0;
    double xl = 0.0, xh = w;
    double x = m > 0.0 ? w * (r / m) : 0.0;
    const double tol = 1e-15 * (lo + w);  // a few ulps of theta
    for (int it = 0; it < 60; ++it) {
        const double f = x * (p0 + x * (p1 + x * p2)) - r;
        if (f > 0.0) xh = x;
        else xl = x;
        const double d = (a + b * x) * (lo + x);
        double xn;
        if (d > 0.0) {
            xn = x - f / d;
            if (::fabs(xn - x) <= tol) {  // converged: take the last step
                x = xn < 0.0 ? 0.0 : (xn > w ? w : xn);
                break;
            }
        } else {
            xn = 0.5 * (xl + xh);
        }
        if (!(xn > xl && xn < xh)) xn = 0.5 * (xl + xh);
        if (!(xh > xl)) break;
        x = xn;
    }
    return lo + x;
}

STD double sunshape_icdf(uint32_t w0, uint32_t w1, const DevSun &sun, const double *aux) {
    const double *tab = aux + sun.icdf;
    const int n = sun.icdf_n;
    const double total = tab[kIcdfStride * n];
    // 53-bit uniform in [0, 1): the high word then the top 21 bits of the next
    const double u = ((double)w0 * 2097152.0 + (double)(w1 >> 11)) * 1.1102230246251565e-16;
    const double m = u * total;
    int lo = 0, hi = n - 1;  // the last segment whose mass-before is <= m
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tab[kIcdfStride * mid + 7] <= m) lo = mid;
        else hi = mid - 1;
    }
    const double *seg = tab + kIcdfStride * lo;
    const double th = sunshape_icdf_solve(seg, m - seg[7]);
    return th * th;
}

// One angular perturbation of a direction: the sunshape error
// (coretrace/errors.cpp:60-226, source 1), the slope error of the surface
// normal (coretrace/surfacenormalerrors.cpp:58-167; slope = true, only 'g'
// and 'p' perturb) or the specularity error (errors.cpp, source 2).  The
// polar angle is drawn by the distribution `dist` with width `delop` (rad),
// or taken from `pre` (has_pre) when the 'd' rejection loop already ran
// elsewhere;
// then the azimuth draw and the rotation about `base`.  With resample (the
// specular step of a reflection) a direction into the surface DF is drawn
// again, up to 50000 times (errors.cpp:206-216).
template <bool FAST>
STD void perturb(Rng &rng, const double base[3], int dist, double delop, bool slope, const DevSun &sun,
                 const double *aux, bool has_pre, double pre, bool resample, const double DF[3], double out[3]) {
    double RT[9];
    ray_frame<FAST>(base, RT);
    unsigned maxcall = 0;
    for (;;) {
        double theta2 = 0.0;
        if (has_pre) {
            theta2 = pre;
        } else if (dist == 'g' || dist == 'G') {
            const double tx = rng.normal<FAST>(0.0, delop);
            const double ty = rng.normal<FAST>(0.0, delop);
            theta2 = tx * tx + ty * ty;
        } else if (dist == 'p' || dist == 'P') {
            double tx, ty;
            do {
                tx = 2.0 * delop * rng.uniform() - delop;
                ty = 2.0 * delop * rng.uniform() - delop;
                theta2 = tx * tx + ty * ty;
            } while (theta2 > delop * delop);
        } else if (!slope && (dist == 'd' || dist == 'D')) {
            if (sun.icdf_n > 0) {
                const uint32_t w0 = rng.next_word(), w1 = rng.next_word();
                theta2 = sunshape_icdf(w0, w1, sun, aux);
            } else {
                for (;;) {
                    const uint32_t w0 = rng.next_word(), w1 = rng.next_word(), w2 = rng.next_word();
                    if (sunshape_d_try(w0, w1, w2, sun, aux, &theta2)) break;
                }
            }
            theta2 = theta2 / 1000000.0;
        } else if (!slope && (dist == 'f' || dist == 'F')) {
            const double u = rng.uniform();
            double a;
            if constexpr (FAST) a = (double)::asinf(::sqrtf((float)u));
            else a = ::asin(::sqrt(u));
            theta2 = a * a;
        }
        cone_dir<FAST>(theta2, rng.uniform(), RT, out);
        if (resample && dot3<double>(out, DF) < 0.0 && maxcall++ < 50000u) continue;
        break;
    }
}

// Reflection / refraction (coretrace/interaction.cpp:58-201); cout is left
// untouched when the refraction Newton loop does not converge.
STD void interact(Rng &rng, const double c[3], const double DF[3], int type, const DevFace &f, double cout[3]) {
    if (type == 1) {
        const double n1 = f.n_in, n2 = f.n_out;
        const double mu = n1 / n2;
        const double D2 = dot3<double>(DF, DF);
        const double B = (mu * mu - 1.0) / D2;
        const double A = mu * dot3<double>(c, DF) / D2;
        if (B > A * A) {
            const double a = dot3<double>(c, DF) / dot3<double>(DF, DF);
            for (int i = 0; i < 3; i++) cout[i] = c[i] - 2.0 * a * DF[i];
            return;
        }
        const double nn = ::sqrt(dot3<double>(DF, DF));
        const double u[3] = {-DF[0] / nn, -DF[1] / nn, -DF[2] / nn};
        const double ia = ::acos(dot3<double>(c, u));
        const double si = ::sin(ia), ci = ::cos(ia);
        const double st = n1 * si / n2;
        const double ct = ::sqrt(1.0 - st * st);
        const double rs = (n1 * ci - n2 * ct) / (n1 * ci + n2 * ct);
        const double rp = (n1 * ct - n2 * ci) / (n1 * ct + n2 * ci);
        const double rave = (rp * rp + rs * rs) / 2.0;
        if (rave < rng.uniform()) {
            double g = -B / (2.0 * A), g1 = 0;
            for (int i = 1; i < 10; i++) {
                g1 = (g * g - B) / (2.0 * (g + A));
                if (::fabs(g - g1) < 0.000005) {
                    for (int j = 0; j < 3; j++) cout[j] = mu * c[j] + g1 * DF[j];
                    return;
                }
                g = g1;
            }
            return;
        }
        const double a = dot3<double>(c, DF) / dot3<double>(DF, DF);
        for (int i = 0; i < 3; i++) cout[i] = c[i] - 2.0 * a * DF[i];
        return;
    }
    const double a = dot3<double>(c, DF) / dot3<double>(DF, DF);
    for (int i = 0; i < 3; i++) cout[i] = c[i] - 2.0 * a * DF[i];
}

// Optical efficiency (coretrace/raytrace.cpp:885-946).
STD double test_value(int interaction, const DevFace &f, const double *aux, const double c[3], const double DF[3]) {
    int tab, n;
    if (interaction == 1) {
        if (f.trans_n <= 0) return f.transmissivity;
        tab = f.trans_tab; n = f.trans_n;
    } else {
        if (f.refl_n <= 0) return f.reflectivity;
        tab = f.refl_tab; n = f.refl_n;
    }
    const double *ang = aux + tab, *val = ang + n;
    const double nn = ::sqrt(dot3<double>(DF, DF));
    const double u[3] = {-DF[0] / nn, -DF[1] / nn, -DF[2] / nn};
    const double ia = ::acos(dot3<double>(c, u)) * 1000.0;
    if (ia >= ang[n - 1]) return val[n - 1];
    int m = 0;
    while (ang[m] < ia) m++;
    if (m == 0) return val[0];
    return (val[m] + val[m - 1]) / 2.0;
}

}  // namespace st
