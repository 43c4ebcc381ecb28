// Operation counting for the oracle.  TEST AND MEASUREMENT INFRASTRUCTURE
// ONLY: st_oracle_count.cpp compiles the oracle with every `double` replaced
// by `Counted`, a double that counts the arithmetic done on it, so a run of
// the oracle on a workload gives the algorithmic floating-point work of the
// trace loop per ray, by phase.
//
// Counting rule: add, subtract, multiply, divide and square root count 1
// each (a fused multiply-add would count 2: the oracle has none, it is built
// without contraction); comparisons, negation, fabs, floor and ceil count 0;
// transcendentals (sin, cos, tan, atan, atan2, asin, acos, exp, log, pow)
// count 1 each in their own column.
//
// Phases: the trace loop names the phase its work belongs to (ST_OPS_*
// hooks in st_oracle.cpp, no-ops in the normal build):
//   0 stage-1 work of sun rays that hit nothing (sun-ray generation, stage
//     transform), element tests excluded
//   1 the same work for sun rays that hit
//   2 the rest of stage 1 for hitting rays (records, optics, errors, the
//     multi-hit loop), element tests excluded
//   3 tabulated-sunshape polar-angle draws (within stage 1's optics)
//   4 + k  stage k+1 (k >= 1), element tests excluded
//   12 + k element tests (ray-element intersections) of stage k+1, k = 0..2
//   15 scratch (a sun ray's work until it is known to hit or miss)
// and element tests are counted per stage (st_opcount_tests).
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <type_traits>

namespace opcount {

constexpr int kCols = 5;  // add/sub, mul, div, sqrt, transcendental
constexpr int kPhases = 16;
constexpr int kScratch = 15;

struct State {
    uint64_t n[kPhases][kCols];
    uint64_t tests[3];
    int phase;
    int saved;
};
inline State &state() {
    static State s = {{{0}}, {0}, 2, 2};
    return s;
}
inline void bump(int col) { state().n[state().phase][col]++; }

inline void set_phase(int p) { state().phase = p; }
// Move the scratch phase's counts to phase `to` and continue in `next`.
inline void settle(int to, int next) {
    State &s = state();
    for (int c = 0; c < kCols; c++) {
        s.n[to][c] += s.n[kScratch][c];
        s.n[kScratch][c] = 0;
    }
    s.phase = next;
}
inline void push(int p) {
    state().saved = state().phase;
    state().phase = p;
}
inline void pop() { state().phase = state().saved; }
// An element test of stage si (0-based) begins; its work goes to phase
// 12 + min(si, 2) until pop().
inline void push_test(int si) {
    const int k = si < 2 ? si : 2;
    state().tests[k]++;
    push(12 + k);
}

struct Counted {
    double v;
    Counted() = default;
    constexpr Counted(double x) : v(x) {}
    template <typename T, typename std::enable_if<std::is_integral<T>::value, int>::type = 0>
    constexpr Counted(T x) : v((double)x) {}
    template <typename T, typename std::enable_if<std::is_arithmetic<T>::value, int>::type = 0>
    explicit constexpr operator T() const { return (T)v; }

    Counted &operator+=(Counted o) { bump(0); v += o.v; return *this; }
    Counted &operator-=(Counted o) { bump(0); v -= o.v; return *this; }
    Counted &operator*=(Counted o) { bump(1); v *= o.v; return *this; }
    Counted &operator/=(Counted o) { bump(2); v /= o.v; return *this; }
    Counted operator-() const { return Counted(-v); }
    Counted operator+() const { return *this; }
    Counted &operator++() { bump(0); v += 1.0; return *this; }
};
static_assert(sizeof(Counted) == sizeof(double) && std::is_trivially_copyable<Counted>::value,
              "Counted must pass where a double does");

inline Counted operator+(Counted a, Counted b) { bump(0); return Counted(a.v + b.v); }
inline Counted operator-(Counted a, Counted b) { bump(0); return Counted(a.v - b.v); }
inline Counted operator*(Counted a, Counted b) { bump(1); return Counted(a.v * b.v); }
inline Counted operator/(Counted a, Counted b) { bump(2); return Counted(a.v / b.v); }
inline bool operator<(Counted a, Counted b) { return a.v < b.v; }
inline bool operator>(Counted a, Counted b) { return a.v > b.v; }
inline bool operator<=(Counted a, Counted b) { return a.v <= b.v; }
inline bool operator>=(Counted a, Counted b) { return a.v >= b.v; }
inline bool operator==(Counted a, Counted b) { return a.v == b.v; }
inline bool operator!=(Counted a, Counted b) { return a.v != b.v; }

#define ST_COUNT_T1(name, col)                                    \
    inline Counted name(Counted a) {                              \
        if (col >= 0) bump(col);                                  \
        return Counted(std::name(a.v));                           \
    }
ST_COUNT_T1(sqrt, 3)
ST_COUNT_T1(sin, 4)
ST_COUNT_T1(cos, 4)
ST_COUNT_T1(tan, 4)
ST_COUNT_T1(atan, 4)
ST_COUNT_T1(asin, 4)
ST_COUNT_T1(acos, 4)
ST_COUNT_T1(exp, 4)
ST_COUNT_T1(log, 4)
ST_COUNT_T1(fabs, -1)
ST_COUNT_T1(floor, -1)
ST_COUNT_T1(ceil, -1)
#undef ST_COUNT_T1
inline Counted atan2(Counted a, Counted b) { bump(4); return Counted(std::atan2(a.v, b.v)); }
inline Counted pow(Counted a, Counted b) { bump(4); return Counted(std::pow(a.v, b.v)); }
inline bool isnan(Counted a) { return std::isnan(a.v); }
inline bool isfinite(Counted a) { return std::isfinite(a.v); }
inline Counted abs(Counted a) { return Counted(std::fabs(a.v)); }
inline Counted fmin(Counted a, Counted b) { return Counted(std::fmin(a.v, b.v)); }
inline Counted fmax(Counted a, Counted b) { return Counted(std::fmax(a.v, b.v)); }

}  // namespace opcount

// The oracle calls these unqualified or through std::.
namespace std {
using opcount::acos;
using opcount::asin;
using opcount::atan;
using opcount::atan2;
using opcount::ceil;
using opcount::cos;
using opcount::exp;
using opcount::fabs;
using opcount::floor;
using opcount::isfinite;
using opcount::isnan;
using opcount::log;
using opcount::pow;
using opcount::sin;
using opcount::sqrt;
using opcount::tan;
inline opcount::Counted max(opcount::Counted a, double b) { return a.v < b ? opcount::Counted(b) : a; }
inline opcount::Counted max(double a, opcount::Counted b) { return a < b.v ? b : opcount::Counted(a); }
inline opcount::Counted min(opcount::Counted a, double b) { return b < a.v ? opcount::Counted(b) : a; }
inline opcount::Counted min(double a, opcount::Counted b) { return b.v < a ? b : opcount::Counted(a); }
}  // namespace std
using opcount::acos;
using opcount::asin;
using opcount::atan;
using opcount::atan2;
using opcount::ceil;
using opcount::cos;
using opcount::exp;
using opcount::fabs;
using opcount::floor;
using opcount::log;
using opcount::pow;
using opcount::sin;
using opcount::sqrt;
using opcount::tan;

using opcount::fmax;
using opcount::fmin;

#define ST_HAVE_LONG_DOUBLE_T
typedef long double st_long_double;

#define ST_OPS_PHASE(p) opcount::set_phase(p)
#define ST_OPS_SETTLE(to, next) opcount::settle(to, next)
#define ST_OPS_PUSH(p) opcount::push(p)
#define ST_OPS_POP() opcount::pop()
#define ST_OPS_TEST(si) opcount::push_test(si)
