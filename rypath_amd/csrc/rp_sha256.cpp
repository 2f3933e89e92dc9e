// SHA-256 (FIPS 180-4).  The round constants are the first 32 fractional
// bits of the cube roots of the first 64 primes and the initial hash the
// first 32 fractional bits of the square roots of the first 8 primes
// (FIPS 180-4 sections 4.2.2 and 5.3.3); they are derived here from that
// definition instead of being tabulated.
#include "rp_sha256.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace rp {
namespace {

struct Constants {
    uint32_t k[64];
    uint32_t h0[8];
    Constants() {
        int primes[64], n = 0;
        for (int c = 2; n < 64; ++c) {
            bool prime = true;
            for (int d = 2; d * d <= c; ++d)
                if (c % d == 0) {
                    prime = false;
                    break;
                }
            if (prime) primes[n++] = c;
        }
        // 32 fractional bits: integer parts are below 8, so long double's
        // 64-bit significand holds them exactly; frac_bits(x) rounds down.
        auto frac_bits = [](long double x) {
            long double f = x - std::floor(x);
            return uint32_t(std::floor(std::ldexp(f, 32)));
        };
        for (int i = 0; i < 64; ++i) k[i] = frac_bits(std::cbrt((long double)primes[i]));
        for (int i = 0; i < 8; ++i) h0[i] = frac_bits(std::sqrt((long double)primes[i]));
    }
};
const Constants &constants() {
    static const Constants c;
    return c;
}
inline uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

} // namespace

Sha256::Sha256() {
    std::memcpy(h_, constants().h0, sizeof h_);
}

void Sha256::block(const uint8_t *p) {
    const uint32_t *k = constants().k;
    uint32_t w[64];
    for (int i = 0; i < 16; ++i)
        w[i] = uint32_t(p[4 * i]) << 24 | uint32_t(p[4 * i + 1]) << 16 | uint32_t(p[4 * i + 2]) << 8 | p[4 * i + 3];
    for (int i = 16; i < 64; ++i) {
        uint32_t s0 = rotr(w[i - 15], 7) ^ rotr(w[i - 15], 18) ^ (w[i - 15] >> 3);
        uint32_t s1 = rotr(w[i - 2], 17) ^ rotr(w[i - 2], 19) ^ (w[i - 2] >> 10);
        w[i] = w[i - 16] + s0 + w[i - 7] + s1;
    }
    uint32_t a = h_[0], b = h_[1], c = h_[2], d = h_[3], e = h_[4], f = h_[5], g = h_[6], h = h_[7];
    for (int i = 0; i < 64; ++i) {
        uint32_t S1 = rotr(e, 6) ^ rotr(e, 11) ^ rotr(e, 25);
        uint32_t ch = (e & f) ^ (~e & g);
        uint32_t t1 = h + S1 + ch + k[i] + w[i];
        uint32_t S0 = rotr(a, 2) ^ rotr(a, 13) ^ rotr(a, 22);
        uint32_t maj = (a & b) ^ (a & c) ^ (b & c);
        uint32_t t2 = S0 + maj;
        h = g;
        g = f;
        f = e;
        e = d + t1;
        d = c;
        c = b;
        b = a;
        a = t1 + t2;
    }
    h_[0] += a, h_[1] += b, h_[2] += c, h_[3] += d, h_[4] += e, h_[5] += f, h_[6] += g, h_[7] += h;
}

void Sha256::update(const void *data, size_t n) {
    const uint8_t *p = static_cast<const uint8_t *>(data);
    bytes_ += n;
    if (fill_) {
        size_t take = std::min(n, size_t(64) - fill_);
        std::memcpy(buf_ + fill_, p, take);
        fill_ += take, p += take, n -= take;
        if (fill_ == 64) {
            block(buf_);
            fill_ = 0;
        }
    }
    while (n >= 64) {
        block(p);
        p += 64, n -= 64;
    }
    if (n) {
        std::memcpy(buf_, p, n);
        fill_ = n;
    }
}

std::array<uint8_t, 32> Sha256::finish() {
    const uint64_t bits = bytes_ * 8;
    const uint8_t one = 0x80, zero = 0;
    update(&one, 1);
    while (fill_ != 56) update(&zero, 1);
    uint8_t len[8];
    for (int i = 0; i < 8; ++i) len[i] = uint8_t(bits >> (56 - 8 * i));
    update(len, 8);
    std::array<uint8_t, 32> out{};
    for (int i = 0; i < 8; ++i)
        for (int j = 0; j < 4; ++j) out[size_t(4 * i + j)] = uint8_t(h_[i] >> (24 - 8 * j));
    return out;
}

} // namespace rp

extern "C" void rp_sha256(const void *data, size_t size, uint8_t digest[32]) {
    rp::Sha256 h;
    h.update(data, size);
    auto d = h.finish();
    std::memcpy(digest, d.data(), 32);
}
