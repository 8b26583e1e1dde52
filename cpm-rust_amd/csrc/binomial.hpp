// k ~ Binomial(n, p) as the EdgeList proposal draws it once per MCS
// (crates/cpm-core/src/monte_carlo.rs:336-366, via rand_distr 0.4.3's
// Binomial::sample, binomial.rs; rand_distr is a dependency that the
// reference does not vendor).  The algorithm is the published one:
//
//  * p > 1/2 is mirrored (sample with 1 - p, return n - k);
//  * n * p < 10 (and n < 2^31): BINV, inversion by the recurrence
//    f(x + 1) = f(x) ((n + 1) s / (x + 1) - s), s = p / q, started from
//    q^n (binary powering, as f64::powi) with one gen::<f64>() draw;
//  * otherwise BTPE (Kachitvichyanukul & Schmeiser, Commun. ACM 31(2),
//    1988): a triangle, two parallelograms and two exponential tails
//    under the scaled mass function; each round draws u ~ U[0, p4) and
//    v ~ U[0, 1) through rand 0.8's UniformFloat (52 random bits as
//    1.x - 1, times the scale, plus the low end); acceptance by the
//    recursive ratio within 20 of the mode, else by squeezing and the
//    final test with Stirling corrections (GSL's signs).
//
// The logarithms are the libm's; a decision can differ from the
// reference's only when two compared values agree to the last place.
#pragma once

#include <cmath>
#include <cstdint>

#include "core.hpp"

namespace cpm {

// rand 0.8 UniformFloat::new: the largest scale s <= high - low such that
// low + s * (1 - 2^-52) < high
CPM_HD double uniform_scale(double low, double high) {
    const double max_rand = 1.0 - 0x1p-52;
    double scale = high - low;
    while (scale * max_rand + low >= high) scale = nextafter(scale, 0.0);
    return scale;
}

CPM_HD double uniform_sample(Xoshiro& rng, double low, double scale) {
    const uint64_t bits = (xo_next(rng) >> 12) | 0x3FF0000000000000ull;
    double v12;
    __builtin_memcpy(&v12, &bits, 8);
    return (v12 - 1.0) * scale + low;
}

// x^n for n >= 0 by binary powering, multiplication order as f64::powi
CPM_HD double powi_nonneg(double a, uint32_t b) {
    double r = 1.0;
    for (;;) {
        if (b & 1u) r *= a;
        b >>= 1;
        if (b == 0) break;
        a *= a;
    }
    return r;
}

// Rust's `x as i64`: truncation toward zero, saturating, NaN -> 0 (a
// sample of v = 0 gives log(v) = -inf in the tails)
CPM_HD int64_t rust_f64_to_i64(double x) {
    if (x != x) return 0;
    if (x <= -0x1p63) return INT64_MIN;
    if (x >= 0x1p63) return INT64_MAX;
    return (int64_t)x;
}

CPM_HD double btpe_stirling(double a) {
    const double a2 = a * a;
    return (13860.0 - (462.0 - (132.0 - (99.0 - 140.0 / a2) / a2) / a2) / a2) / a / 166320.0;
}

// Returns false only if a loop ran away (BINV past n, or BTPE over
// kMaxRounds rejections: neither happens with probability above ~1e-16);
// *k is then n or the last candidate, and the caller reports an error.
constexpr uint32_t kBinomialMaxRounds = 100000;

CPM_HD bool binomial_sample(Xoshiro& rng, uint64_t n_u, double p_in, uint64_t* k_out) {
    if (p_in == 0.0) {
        *k_out = 0;
        return true;
    }
    if (p_in == 1.0) {
        *k_out = n_u;
        return true;
    }
    const double p = p_in <= 0.5 ? p_in : 1.0 - p_in;
    const double q = 1.0 - p;
    uint64_t result = 0;
    bool ok = true;
    if ((double)n_u * p < 10.0 && n_u <= 0x7fffffffull) {
        const double s = p / q;
        const double a = (double)(n_u + 1) * s;
        double r = powi_nonneg(q, (uint32_t)n_u);
        double u = gen_f64(rng);
        uint64_t x = 0;
        while (u > r) {
            u -= r;
            x += 1;
            if (x > n_u) {  // only when rounding leaves u above the summed masses
                x = n_u;
                ok = false;
                break;
            }
            r *= a / (double)x - s;
        }
        result = x;
    } else {
        const double n = (double)n_u;
        const double np = n * p;
        const double npq = np * q;
        const double f_m = np + p;
        const int64_t m = rust_f64_to_i64(f_m);
        const double p1 = floor(2.195 * sqrt(npq) - 4.6 * q) + 0.5;
        const double x_m = (double)m + 0.5;
        const double x_l = x_m - p1;
        const double x_r = x_m + p1;
        const double c = 0.134 + 20.5 / (15.3 + (double)m);
        const double p2 = p1 * (1.0 + 2.0 * c);
        const double al = (f_m - x_l) / (f_m - x_l * p);
        const double lambda_l = al * (1.0 + 0.5 * al);
        const double ar = (x_r - f_m) / (x_r * q);
        const double lambda_r = ar * (1.0 + 0.5 * ar);
        const double p3 = p2 + c / lambda_l;
        const double p4 = p3 + c / lambda_r;
        const double scale_u = uniform_scale(0.0, p4);
        const double scale_v = uniform_scale(0.0, 1.0);
        int64_t y = 0;
        uint32_t round = 0;
        for (;; ++round) {
            if (round >= kBinomialMaxRounds) {
                ok = false;
                break;
            }
            const double u = uniform_sample(rng, 0.0, scale_u);
            double v = uniform_sample(rng, 0.0, scale_v);
            if (!(u > p1)) {  // triangle
                y = rust_f64_to_i64(x_m - p1 * v + u);
                break;
            }
            if (!(u > p2)) {  // parallelograms
                const double x = x_l + (u - p1) / c;
                v = v * c + 1.0 - fabs(x - x_m) / p1;
                if (v > 1.0) continue;
                y = rust_f64_to_i64(x);
            } else if (!(u > p3)) {  // left tail
                y = rust_f64_to_i64(x_l + log(v) / lambda_l);
                if (y < 0) continue;
                v *= (u - p2) * lambda_l;
            } else {  // right tail
                y = rust_f64_to_i64(x_r - log(v) / lambda_r);
                if (y > 0 && (uint64_t)y > n_u) continue;
                v *= (u - p3) * lambda_r;
            }
            const int64_t k = y > m ? y - m : m - y;
            if (!(k > 20 && (double)k < 0.5 * npq - 1.0)) {
                // f(y) / f(m) by the recurrence from the mode
                const double s = p / q;
                const double a = s * (n + 1.0);
                double f = 1.0;
                if (m < y) {
                    for (int64_t i = m + 1; i <= y; ++i) f *= a / (double)i - s;
                } else if (m > y) {
                    for (int64_t i = y + 1; i <= m; ++i) f /= a / (double)i - s;
                }
                if (v > f) continue;
                break;
            }
            const double kf = (double)k;
            const double rho = (kf / npq) * ((kf * (kf / 3.0 + 0.625) + 1.0 / 6.0) / npq + 0.5);
            const double t = -0.5 * kf * kf / npq;
            const double alpha = log(v);
            if (alpha < t - rho) break;
            if (alpha > t + rho) continue;
            const double x1 = (double)(y + 1);
            const double f1 = (double)(m + 1);
            const double z = (double)(rust_f64_to_i64(n) + 1 - m);
            const double w = (double)(rust_f64_to_i64(n) - y + 1);
            if (alpha > x_m * log(f1 / x1) + (n - (double)m + 0.5) * log(z / w) +
                            (double)(y - m) * log(w * p / (x1 * q)) + btpe_stirling(f1) +
                            btpe_stirling(z) - btpe_stirling(x1) - btpe_stirling(w))
                continue;
            break;
        }
        result = y < 0 ? 0 : (uint64_t)y;
    }
    *k_out = p != p_in ? n_u - result : result;
    return ok;
}

}  // namespace cpm
