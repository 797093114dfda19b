"""The synthetic area-query workload (dsx_area_workload) against an independent Python restatement of
its definition: xoroshiro128++ with Minecraft's seeding (standalone.hpp:18-37), Lemire bounded
draws, 53-bit uniform doubles, splitmix64 finalizer seeds."""
import math

import numpy as np

from conftest import load_package

M64 = (1 << 64) - 1


def mix(x):
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & M64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & M64
    return x ^ (x >> 31)


def rotl(x, k):
    return ((x << k) | (x >> (64 - k))) & M64


class Rng:
    def __init__(self, seed):
        seed ^= 0x6A09E667F3BCC909
        self.a = mix(seed)
        self.b = mix((seed + 0x9E3779B97F4A7C15) & M64)

    def next(self):
        a, b = self.a, self.b
        out = (rotl((a + b) & M64, 17) + a) & M64
        t = a ^ b
        self.a = rotl(a, 49) ^ t ^ ((t << 21) & M64)
        self.b = rotl(t, 28)
        return out

    def uniform(self):
        return (self.next() >> 11) * 2.0 ** -53

    def bounded(self, n):
        m = (self.next() & 0xFFFFFFFF) * n
        if (m & 0xFFFFFFFF) < n:
            lim = ((1 << 32) - n) % n
            while (m & 0xFFFFFFFF) < lim:
                m = (self.next() & 0xFFFFFFFF) * n
        return m >> 32


def test_workload_matches_definition():
    dsx = load_package().dsx
    q, caps = dsx.area_workload(4096)
    r = Rng(0xD5E27C4)
    for j in range(4096):
        X = r.bounded(60000001) - 30000000
        Z = r.bounded(60000001) - 30000000
        cap = math.floor(math.pow(10.0, 3.0 + 4.0 * r.uniform()))
        half = math.floor((math.sqrt(cap) - 1.0) / 2.0)
        assert (int(q[j]["seed"]), int(q[j]["x"]), int(q[j]["z"]), int(q[j]["half"]), int(caps[j])) == \
            (mix(j), X, Z, half, cap), j
    assert caps.min() >= 1000 and caps.max() < 10 ** 7
    assert abs(q["x"]).max() <= 30000000 and abs(q["z"]).max() <= 30000000
    # log-uniform caps: about a quarter per decade
    dec = np.floor(np.log10(caps.astype(np.float64))).astype(int)
    assert all(abs((dec == d).mean() - 0.25) < 0.04 for d in (3, 4, 5, 6))
    # window edge (2 half + 1)^2 <= cap < (2 half + 3)^2
    h = q["half"].astype(np.int64)
    assert np.all((2 * h + 1) ** 2 <= caps) and np.all(caps < (2 * h + 3) ** 2)
