// Jump-ahead for Xoshiro256++ streams: the characteristic polynomial of
// the generator's linear state transition, found by Berlekamp-Massey on
// one output bit of the state sequence, and x^J mod p(x) for a jump of J
// draws.  Applying the polynomial to a state (sum of q_i * T^i s) yields
// the state J draws ahead, so one replica's serial stream can be generated
// as independent segments in parallel.  Host-only; the device applies the
// polynomials in the fill kernel.
#include "rng_jump.hpp"

#include <cstring>
#include <stdexcept>
#include <vector>

namespace cpm {

namespace {

// GF(2) polynomials as little-endian bit vectors (bit i = coefficient of x^i).
struct Poly {
    std::vector<uint64_t> w;
    explicit Poly(size_t bits = 0) : w((bits + 64) / 64, 0) {}
    bool get(size_t i) const { return i / 64 < w.size() && ((w[i / 64] >> (i % 64)) & 1); }
    void flip(size_t i) {
        if (i / 64 >= w.size()) w.resize(i / 64 + 1, 0);
        w[i / 64] ^= 1ull << (i % 64);
    }
    int degree() const {
        for (size_t k = w.size(); k-- > 0;)
            if (w[k]) return (int)(k * 64 + 63 - __builtin_clzll(w[k]));
        return -1;
    }
};

// Berlekamp-Massey over GF(2): connection polynomial C(x) of the shortest
// LFSR producing s (C(x) = 1 + c1 x + ... + cL x^L).
Poly berlekamp_massey(const std::vector<uint8_t>& s, int* L_out) {
    const size_t n = s.size();
    Poly C(n + 1), B(n + 1);
    C.flip(0);
    B.flip(0);
    int L = 0, m = 1;
    for (size_t i = 0; i < n; ++i) {
        int d = s[i];
        for (int j = 1; j <= L; ++j) d ^= C.get(j) & s[i - j];
        if (!d) {
            ++m;
            continue;
        }
        Poly T = C;
        for (size_t j = 0; j + m <= n; ++j)
            if (B.get(j)) C.flip(j + m);
        if (2 * L <= (int)i) {
            L = (int)i + 1 - L;
            B = T;
            m = 1;
        } else {
            ++m;
        }
    }
    *L_out = L;
    return C;
}

// (a * b) mod p for degree(p) = 256, a,b < x^256.
void mulmod(const uint64_t a[4], const uint64_t b[4], const uint64_t p[5], uint64_t out[4]) {
    uint64_t prod[8] = {0};
    for (int i = 0; i < 256; ++i)
        if ((a[i / 64] >> (i % 64)) & 1)
            for (int k = 0; k < 4; ++k) {
                int sh = i % 64, wo = i / 64;
                prod[k + wo] ^= b[k] << sh;
                if (sh) prod[k + wo + 1] ^= b[k] >> (64 - sh);
            }
    for (int i = 511; i >= 256; --i)
        if ((prod[i / 64] >> (i % 64)) & 1) {
            int shift = i - 256;  // subtract x^shift * p
            for (int k = 0; k < 5; ++k) {
                int sh = shift % 64, wo = shift / 64;
                if (k + wo < 8) prod[k + wo] ^= p[k] << sh;
                if (sh && k + wo + 1 < 8) prod[k + wo + 1] ^= p[k] >> (64 - sh);
            }
        }
    std::memcpy(out, prod, 32);
}

const uint64_t* char_poly() {
    static uint64_t p[5];
    static bool ready = false;
    if (ready) return p;
    Xoshiro x = xo_seed(0x5eed);
    std::vector<uint8_t> bits(1024);
    for (auto& b : bits) {
        b = (uint8_t)(x.s[0] & 1);
        xo_next(x);
    }
    int L = 0;
    Poly C = berlekamp_massey(bits, &L);
    if (L != 256) throw std::runtime_error("xoshiro256 characteristic polynomial search failed");
    // p(x) = x^L C(1/x): coefficient of x^(L-j) is c_j
    std::memset(p, 0, sizeof p);
    for (int j = 0; j <= L; ++j)
        if (C.get(j)) p[(L - j) / 64] |= 1ull << ((L - j) % 64);
    ready = true;
    return p;
}

}  // namespace

JumpPoly xo_jump_poly(uint64_t distance) {
    const uint64_t* p = char_poly();
    uint64_t result[4] = {1, 0, 0, 0};  // x^0
    uint64_t base[4] = {2, 0, 0, 0};    // x^1
    for (uint64_t e = distance; e; e >>= 1) {
        if (e & 1) mulmod(result, base, p, result);
        mulmod(base, base, p, base);
    }
    JumpPoly q;
    std::memcpy(q.w, result, 32);
    return q;
}

void xo_apply_jump(Xoshiro& s, const JumpPoly& q) {
    Xoshiro acc{{0, 0, 0, 0}};
    Xoshiro cur = s;
    for (int i = 0; i < 256; ++i) {
        if ((q.w[i / 64] >> (i % 64)) & 1)
            for (int k = 0; k < 4; ++k) acc.s[k] ^= cur.s[k];
        xo_next(cur);
    }
    s = acc;
}

void xo_jump_selfcheck() {
    for (uint64_t dist : {0ull, 1ull, 255ull, 256ull, 1000ull, 4097ull}) {
        Xoshiro a = xo_seed(dist * 7 + 3), b = a;
        for (uint64_t i = 0; i < dist; ++i) xo_next(a);
        xo_apply_jump(b, xo_jump_poly(dist));
        if (std::memcmp(a.s, b.s, 32) != 0)
            throw std::runtime_error("xoshiro jump-ahead self-check failed");
    }
}

}  // namespace cpm
