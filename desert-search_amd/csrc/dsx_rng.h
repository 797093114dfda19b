// Seed derivation and Minecraft's positional RNG, host + device (gfx950).
//
// Integer-exact restatements of the reference's seed arithmetic:
//   splitmix64 finalizer          standalone.hpp:13-17 (gate.cuh:20-24)
//   xoroshiro128++ with MC seeding standalone.hpp:18-37 (gate.cuh:25-42)
//   temperature octave -10 phase  dsearch_desert.cpp:80-94 (gate.cuh:71-84)
// Double arithmetic here is plain IEEE add/mul/floor (no contraction: built with -ffp-contract=off).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../data/mc_noise.h"

#define DSX_HD __host__ __device__ __forceinline__

namespace dsx {

DSX_HD uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

DSX_HD uint64_t mix64(uint64_t x) {
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

struct Xoro {
    uint64_t a, b;
    DSX_HD static Xoro seeded(uint64_t seed) {
        seed ^= DSX_SEED_XOR;
        Xoro r;
        r.a = mix64(seed);
        r.b = mix64(seed + DSX_GOLDEN);
        return r;
    }
    DSX_HD static Xoro raw(uint64_t a, uint64_t b) {
        Xoro r;
        r.a = a;
        r.b = b;
        return r;
    }
    DSX_HD uint64_t next() {
        const uint64_t out = rotl64(a + b, 17) + a;
        const uint64_t t = a ^ b;
        a = rotl64(a, 49) ^ t ^ (t << 21);
        b = rotl64(t, 28);
        return out;
    }
    DSX_HD double uniform() { return double(next() >> 11) * 0x1p-53; }
    // Lemire bounded draw on the low 32 bits, rejecting the biased low range.
    DSX_HD uint32_t bounded(uint32_t n) {
        uint64_t m = uint64_t(uint32_t(next())) * n;
        if (uint32_t(m) < n) {
            const uint32_t lim = uint32_t(0u - n) % n;
            while (uint32_t(m) < lim) m = uint64_t(uint32_t(next())) * n;
        }
        return uint32_t(m >> 32);
    }
};

// Positional RNG of the temperature channel: world seed -> (a, b) -> fork with the channel hash.
DSX_HD Xoro temperature_fork(uint64_t seed) {
    Xoro w = Xoro::seeded(seed);
    const uint64_t a = w.next(), b = w.next();
    return Xoro::raw(a ^ DSX_NOISE[DSX_CH_T].fork_a, b ^ DSX_NOISE[DSX_CH_T].fork_b);
}

// Sum over the two noise sides of |fy - 1/2|, fy = fractional y offset of temperature octave -10.
DSX_HD double temperature_phase(uint64_t seed, double *fy) {
    Xoro c = temperature_fork(seed);
    double s = 0;
    for (int side = 0; side < 2; side++) {
        const uint64_t fa = c.next(), fb = c.next();
        Xoro r = Xoro::raw(fa ^ DSX_OCT_SALT[0][0], fb ^ DSX_OCT_SALT[0][1]);
        r.next();
        const double y = r.uniform() * 256.;
        const double f = y - floor(y);
        if (fy) fy[side] = f;
        s += fabs(f - .5);
    }
    return s;
}

// |fy - 1/2| summed over the two sides of octave j (the j-th octave of non-zero amplitude) of channel
// ch: the initialisation of Perlin::init (standalone.hpp:55-61, 123-139) up to the y offset.
DSX_HD double octave_phase(uint64_t seed, int ch, int j) {
    const dsx_noise_cfg &cfg = DSX_NOISE[ch];
    Xoro w = Xoro::seeded(seed);
    const uint64_t a = w.next(), b = w.next();
    Xoro r = Xoro::raw(a ^ cfg.fork_a, b ^ cfg.fork_b);
    double s = 0;
    for (int side = 0; side < 2; side++) {
        const uint64_t ra = r.next(), rb = r.next();
        int i = 0, cnt = -1;
        for (; i < cfg.len; i++)
            if (cfg.amp[i] != 0 && ++cnt == j) break;
        const uint64_t *salt = DSX_OCT_SALT[cfg.first + i + 10];
        Xoro q = Xoro::raw(ra ^ salt[0], rb ^ salt[1]);
        q.uniform();
        const double oy = q.uniform() * 256;
        s += fabs((oy - floor(oy)) - .5);
    }
    return s;
}

DSX_HD double fade(double t) { return t * t * t * (t * (t * 6. - 15.) + 10.); }

// Minecraft's 16 gradient vectors as (gx*x + gy*y) + gz*z with components in {0, +-1}
// (standalone.hpp:40-49). The components are decoded from bit masks so no local array is indexed.
// Component a (0 x, 1 y, 2 z) of gradient h: bit h of a +1 mask and of a -1 mask.
DSX_HD double grad_component(int a, unsigned h) {
    constexpr unsigned pos[3] = {0x1055u, 0x5503u, 0x2330u}, neg[3] = {0x40AAu, 0xAA0Cu, 0x8CC0u};
    return ((pos[a] >> h) & 1) ? 1.0 : (((neg[a] >> h) & 1) ? -1.0 : 0.0);
}
DSX_HD double grad(unsigned h, double x, double y, double z) {
    h &= 15;
    const double gx = grad_component(0, h), gy = grad_component(1, h), gz = grad_component(2, h);
    return gx * x + gy * y + gz * z;
}

}  // namespace dsx
