// Random streams for the CPU oracle.  TEST INFRASTRUCTURE ONLY: this file is
// compiled into oracle/_build/st_oracle.so, which only tests/, the smoke
// check and the CPU-baseline leg of bench.py load.  Nothing shipped links it.
//
// Two sources of uniforms are restated here:
//
//  * RefStream -- the reference's single sequential stream.  The reference's
//    MTRand (coretrace/mtrand.h:136-139,298-305,327-343) is a thin wrapper
//    over libstdc++ <random> from GCC 11.4.0 (__GLIBCXX__ 20230528):
//      - std::default_random_engine == std::minstd_rand0
//        (a = 16807, m = 2^31-1; bits/random.h:1555), seeded with
//        x = seed mod m, or 1 when that is 0 (bits/random.tcc:117-124);
//      - uniform_real_distribution<double>(0,1) -> generate_canonical<double,53>
//        with k = 2 engine calls: sum = (g1-1) + (g2-1)*r, divided by
//        RN_double(r*r) where r = 2^31-2; a result >= 1 becomes
//        nextafter(1,0) (bits/random.tcc:3348-3380);
//      - normal_distribution<double>: Marsaglia polar, one cached deviate,
//        returns y*m first and caches x*m (bits/random.tcc:1802-1835).
//    The restatement is checked draw-for-draw against the std:: types
//    themselves (tests/test_oracle_rng.py).
//
//  * CounterStream -- the counter-based per-ray stream the GPU kernels use:
//    Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
//    easy as 1, 2, 3", SC'11).  Uniform k of stream (stage, id) is word k%4
//    of the Philox block with counter (id_lo, id_hi, stage, k/4) and key
//    (seed, 0x53544F52), scaled by 2^-32.  The id of every stage's stream
//    is the ray's 0-based sun-ray index.  Normals use the same polar method
//    as above with the same caching, so both streams feed identical
//    algorithms.  Checked against the Random123 known-answer vectors.
#pragma once
#include <cmath>
#include <cstdint>

// long double, named so the operation-counting build (st_count.h) can keep it
#ifndef ST_HAVE_LONG_DOUBLE_T
#define ST_HAVE_LONG_DOUBLE_T
typedef long double st_long_double;
#endif

namespace oracle {

// ---------------------------------------------------------------- Philox ---
struct Philox4 {
    uint32_t v[4];
};

inline Philox4 philox4x32_10(const uint32_t ctr[4], const uint32_t key[2]) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round) {
        uint64_t p0 = (uint64_t)M0 * c0;
        uint64_t p1 = (uint64_t)M1 * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += W0; k1 += W1;
    }
    Philox4 r;
    r.v[0] = c0; r.v[1] = c1; r.v[2] = c2; r.v[3] = c3;
    return r;
}

static const uint32_t kCounterKeyHi = 0x53544F52u;  // second key word

// Uniform k of the per-ray stream (seed, stage, id): word k%4 of block k/4.
inline double counter_uniform(uint32_t seed, uint32_t stage, uint64_t id, uint64_t k) {
    uint32_t ctr[4] = {(uint32_t)id, (uint32_t)(id >> 32), stage, (uint32_t)(k >> 2)};
    uint32_t key[2] = {seed, kCounterKeyHi};
    Philox4 b = philox4x32_10(ctr, key);
    return (double)b.v[k & 3] * 2.3283064365386963e-10;  // 2^-32
}

// ------------------------------------------------------------ RefStream ---
class RefStream {
public:
    explicit RefStream(uint32_t seed) {
        uint64_t s = (uint64_t)seed % 2147483647ull;
        x_ = (s == 0) ? 1ull : s;
        have_saved_ = false;
        saved_ = 0.0;
    }
    uint64_t engine() {
        x_ = (x_ * 16807ull) % 2147483647ull;
        return x_;
    }
    double uniform() {
        const st_long_double r = (st_long_double)2147483646.0;  // max - min + 1
        double sum = 0.0;
        double tmp = 1.0;
        for (int k = 0; k < 2; ++k) {
            sum += (double)(engine() - 1ull) * tmp;
            tmp = (double)((st_long_double)tmp * r);
        }
        double ret = sum / tmp;
        if (ret >= 1.0) ret = std::nextafter(1.0, 0.0);
        return ret;
    }
    double normal(double mean, double stddev) {
        double ret;
        if (have_saved_) {
            have_saved_ = false;
            ret = saved_;
        } else {
            double x, y, r2;
            do {
                x = 2.0 * uniform() - 1.0;
                y = 2.0 * uniform() - 1.0;
                r2 = x * x + y * y;
            } while (r2 > 1.0 || r2 == 0.0);
            const double mult = std::sqrt(-2 * std::log(r2) / r2);
            saved_ = x * mult;
            have_saved_ = true;
            ret = y * mult;
        }
        ret = ret * 1.0 + 0.0;  // the distribution's own (0,1) parameters
        return mean + ret * stddev;
    }

private:
    uint64_t x_;
    bool have_saved_;
    double saved_;
};

// -------------------------------------------------------- CounterStream ---
class CounterStream {
public:
    CounterStream() : seed_(0), stage_(0), id_(0), k_(0), have_saved_(false), saved_(0.0) {}
    void begin(uint32_t seed, uint32_t stage, uint64_t id) {
        seed_ = seed; stage_ = stage; id_ = id; k_ = 0;
        have_saved_ = false; saved_ = 0.0;
    }
    double uniform() { return counter_uniform(seed_, stage_, id_, k_++); }
    uint64_t draws() const { return k_; }
    double normal(double mean, double stddev) {
        double ret;
        if (have_saved_) {
            have_saved_ = false;
            ret = saved_;
        } else {
            double x, y, r2;
            do {
                x = 2.0 * uniform() - 1.0;
                y = 2.0 * uniform() - 1.0;
                r2 = x * x + y * y;
            } while (r2 > 1.0 || r2 == 0.0);
            const double mult = std::sqrt(-2 * std::log(r2) / r2);
            saved_ = x * mult;
            have_saved_ = true;
            ret = y * mult;
        }
        return mean + ret * stddev;
    }

private:
    uint32_t seed_, stage_;
    uint64_t id_, k_;
    bool have_saved_;
    double saved_;
};

// One object the trace loop draws from; mode 0 = reference stream,
// mode 1 = per-ray counter streams.
class Rng {
public:
    Rng(int mode, uint32_t seed) : mode_(mode), seed_(seed), ref_(seed) {}
    void begin_ray(uint32_t stage, uint64_t id) {
        if (mode_ == 1) ctr_.begin(seed_, stage, id);
    }
    double operator()() { return mode_ == 1 ? ctr_.uniform() : ref_.uniform(); }
    double normal(double mean, double stddev) {
        return mode_ == 1 ? ctr_.normal(mean, stddev) : ref_.normal(mean, stddev);
    }
    int mode() const { return mode_; }

private:
    int mode_;
    uint32_t seed_;
    RefStream ref_;
    CounterStream ctr_;
};

}  // namespace oracle
