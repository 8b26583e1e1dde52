// The Metropolis test at T = 1 (crates/cpm-core/src/monte_carlo.rs:147-158):
// an uphill move (dH > 0) is accepted iff u < exp(-dH), u = gen::<f64>().
//
// The reference evaluates exp with the platform libm (glibc), which is
// correctly rounded in nearly all cases.  A device libm exp can differ
// from it in the last place, and a decision changes only when u falls
// between the two values.  metropolis_accept therefore decides with the
// libm value unless u lies within 2^-40 (relative) of it, and there with
// the correctly rounded exp(-dH) computed by exp_dd below; draws that take
// that path are reported as near-ties.
//
// exp_dd: double-double exp for x <= 0 (results >= 2^-1022 are correctly
// rounded except when exp(x) lies within ~2^-95 relative of a midpoint
// between doubles).  Method: Cody-Waite reduction x = k ln2 + r with ln2
// split into three parts (the first with 42 significant bits, so k * L1 is
// exact for |k| < 2^11), r scaled by 2^-10, exp of the scaled argument by
// its Taylor series (terms to s^3 in double-double, the rest in double),
// ten double-double squarings, and the final scaling by 2^k.  The split of
// ln2 and 1/6 in double-double are the leading bits of those numbers.
#pragma once

#include <cmath>

#include "core.hpp"

namespace cpm {

struct DD {
    double hi, lo;
};

// Knuth's two-sum: s + e == a + b exactly
CPM_HD DD dd_two_sum(double a, double b) {
    const double s = a + b;
    const double bb = s - a;
    const double e = (a - (s - bb)) + (b - bb);
    return {s, e};
}
// |a| >= |b|
CPM_HD DD dd_fast_two_sum(double a, double b) {
    const double s = a + b;
    return {s, b - (s - a)};
}
// p + e == a * b exactly (fused multiply-add for the error term)
CPM_HD DD dd_two_prod(double a, double b) {
    const double p = a * b;
    return {p, fma(a, b, -p)};
}
CPM_HD DD dd_add(DD a, DD b) {
    const DD s = dd_two_sum(a.hi, b.hi);
    const DD t = dd_two_sum(a.lo, b.lo);
    DD u = dd_fast_two_sum(s.hi, s.lo + t.hi);
    return dd_fast_two_sum(u.hi, u.lo + t.lo);
}
CPM_HD DD dd_mul(DD a, DD b) {
    DD p = dd_two_prod(a.hi, b.hi);
    const double cross = a.hi * b.lo + a.lo * b.hi;
    return dd_fast_two_sum(p.hi, p.lo + cross);
}

// exp(x) for -745 < x <= 0 as a double-double scaled by 2^k: returns the
// unscaled value y (1/2 < y < 2) and k.
CPM_HD DD exp_dd_unscaled(double x, int* k_out) {
    const double kf = rint(x * 0x1.71547652b82fep+0);  // nearest integer to x / ln2
    const double L1 = 0x1.62e42fefa38p-1;              // ln2 = L1 + L2 + L3 + O(2^-155)
    const double L2 = 0x1.ef35793c7673p-45;
    const double L3 = 0x1.f97b57a079a19p-103;
    const double r1 = x - kf * L1;  // exact: kf * L1 fits 53 bits, x and kf * L1 are close
    const DD t = dd_two_prod(kf, L2);
    DD r = dd_two_sum(r1, -t.hi);
    r = dd_fast_two_sum(r.hi, r.lo - t.lo - kf * L3);
    const DD s = {r.hi * 0x1p-10, r.lo * 0x1p-10};  // |s| < 3.4e-4
    // exp(s) - 1 = s + s^2/2 + s^3/6 + s^4 (1/24 + s/120 + ... + s^4/8!)
    const DD s2 = dd_mul(s, s);
    const DD s3 = dd_mul(s2, s);
    const DD sixth = {0x1.5555555555555p-3, 0x1.5555555555555p-57};
    const double h = s.hi;
    const double tail =
        (s2.hi * s2.hi) *
        (1.0 / 24.0 + h * (1.0 / 120.0 + h * (1.0 / 720.0 + h * (1.0 / 5040.0 + h / 40320.0))));
    DD em1 = dd_add(s, DD{s2.hi * 0.5, s2.lo * 0.5});
    em1 = dd_add(em1, dd_mul(s3, sixth));
    em1 = dd_add(em1, DD{tail, 0.0});
    DD y = dd_add(DD{1.0, 0.0}, em1);
    for (int i = 0; i < 10; ++i) y = dd_mul(y, y);
    *k_out = (int)kf;
    return y;
}

// exp(x) rounded to nearest for -708.39 < x <= 0 (see above).
CPM_HD double exp_cr(double x) {
    int k = 0;
    const DD y = exp_dd_unscaled(x, &k);
    return ldexp(y.hi + y.lo, k);
}

// exp(x) for a double-double argument |x| < 709 (either sign), as a
// double-double scaled by 2^k: the reduction above with the low part of x
// folded into the reduced argument.
CPM_HD DD exp_dd2_unscaled(DD x, int* k_out) {
    const double kf = rint(x.hi * 0x1.71547652b82fep+0);
    const double L1 = 0x1.62e42fefa38p-1;
    const double L2 = 0x1.ef35793c7673p-45;
    const double L3 = 0x1.f97b57a079a19p-103;
    const double r1 = x.hi - kf * L1;
    const DD t = dd_two_prod(kf, L2);
    DD r = dd_two_sum(r1, -t.hi);
    r = dd_two_sum(r.hi, ((r.lo - t.lo) - kf * L3) + x.lo);
    const DD s = {r.hi * 0x1p-10, r.lo * 0x1p-10};
    const DD s2 = dd_mul(s, s);
    const DD s3 = dd_mul(s2, s);
    const DD sixth = {0x1.5555555555555p-3, 0x1.5555555555555p-57};
    const double h = s.hi;
    const double tail =
        (s2.hi * s2.hi) *
        (1.0 / 24.0 + h * (1.0 / 120.0 + h * (1.0 / 720.0 + h * (1.0 / 5040.0 + h / 40320.0))));
    DD em1 = dd_add(s, DD{s2.hi * 0.5, s2.lo * 0.5});
    em1 = dd_add(em1, dd_mul(s3, sixth));
    em1 = dd_add(em1, DD{tail, 0.0});
    DD y = dd_add(DD{1.0, 0.0}, em1);
    for (int i = 0; i < 10; ++i) y = dd_mul(y, y);
    *k_out = (int)kf;
    return y;
}

// exp(x) rounded to nearest for any x whose result is a normal double or
// overflows (x > 709.79: +inf; x < -708.39: the libm value; NaN: NaN).
CPM_HD double exp_cr_any(double x) {
    if (!(x > -708.39)) return exp(x);
    if (x > 709.79) return INFINITY;
    int k = 0;
    const DD y = exp_dd2_unscaled(DD{x, 0.0}, &k);
    // y * 2^k in two steps: 2^k alone overflows for k = 1024
    return ldexp(ldexp(y.hi + y.lo, k - k / 2), k / 2);
}

// The Act term's geometric-mean root (crates/cpm-core/src/energy.rs:443:
// product.powf(1.0 / count as f64)) for a product of non-negative integer
// activities: x^(1/n rounded to double), rounded to nearest.  x = 0 gives
// 0 and n = 1 gives x, as pow does; otherwise ln x comes from the libm log
// refined by one Newton step against the double-double exp (error ~2^-104),
// is multiplied by y = RN(1/n) in double-double, and exponentiated in
// double-double, so the result is correctly rounded except within ~2^-95
// of a midpoint.  glibc's pow is not correctly rounded on a small fraction
// of arguments (about 1e-3 of random activity products; none of the
// powers m^n with m <= 228, n <= 27); there the two can differ in the last
// place.
CPM_HD double act_root(double x, uint32_t n) {
    if (n == 1 || x == 0.0) return x;
    const double y = 1.0 / (double)n;
    const double l0 = log(x);
    int k = 0;
    const DD e = exp_dd2_unscaled(DD{l0, 0.0}, &k);
    const double eh = ldexp(e.hi, k), el = ldexp(e.lo, k);
    const double d = ((x - eh) - el) / eh;  // x = exp(l0) (1 + d), |d| ~ 2^-52
    const DD l = dd_two_sum(l0, d - 0.5 * d * d);
    DD t = dd_two_prod(y, l.hi);
    t = dd_fast_two_sum(t.hi, t.lo + y * l.lo);
    int k2 = 0;
    const DD r = exp_dd2_unscaled(t, &k2);
    return ldexp(r.hi + r.lo, k2);
}

// The Metropolis-Hastings test (crates/cpm-core/src/monte_carlo.rs:168-196):
// u < min((n_A / n_B) exp(-dH), 1), with the near-tie rule of
// metropolis_accept (the ratio recomputed from the correctly rounded exp
// when u lies within 2^-40 of the libm one).  n_A > 0.
CPM_HD bool mh_accept(double u, uint32_t n_a, uint32_t n_b, double dh, bool* near_tie) {
    const double q = (double)n_a / (double)n_b;
    const double p = fmin(q * exp(-dh), 1.0);
    if (!(fabs(u - p) <= p * 0x1p-40)) return u < p;
    *near_tie = true;
    return u < fmin(q * exp_cr_any(-dh), 1.0);
}

// u < exp(-dh) for dh > 0, decided as a correctly rounded exp would decide
// it; *near_tie is set when the libm value was too close to u to trust.
CPM_HD bool metropolis_accept(double u, double dh, bool* near_tie) {
    const double p = exp(-dh);
    if (!(fabs(u - p) <= p * 0x1p-40)) return u < p;
    *near_tie = true;
    return u < exp_cr(-dh);
}

}  // namespace cpm
