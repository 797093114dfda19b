// Minecraft 26.3 overworld climate-noise constants (data, shared by the HIP library and the oracle).
//
// These are fixed by the game and must be identical to the values the reference samples with
// (reference src/standalone.hpp:85-100):
//   * DSX_OCT_SALT[k]  positional-RNG salts (MD5 of "octave_<k-10>") for octaves -10 .. 0
//   * DSX_NOISE[c]     per channel c (T, H, C, E, Shift, W): the positional fork hash (two words),
//                      the first octave and the octave amplitudes.
#pragma once
#include <stdint.h>

typedef struct {
    uint64_t fork_a, fork_b;   // xored into the world's positional RNG state
    int first;                 // first octave (frequency 2^first)
    int len;                   // number of amplitude slots
    double amp[9];             // amplitude per octave slot (0 = octave absent)
} dsx_noise_cfg;

enum { DSX_CH_T = 0, DSX_CH_H = 1, DSX_CH_C = 2, DSX_CH_E = 3, DSX_CH_SHIFT = 4, DSX_CH_W = 5 };

static const uint64_t DSX_OCT_SALT[11][2] = {
    {0x36d326eed40efeb2ull, 0x5be9ce18223c636aull}, {0x082fe255f8be6631ull, 0x4e96119e22dedc81ull},
    {0x0ef68ec68504005eull, 0x48b6bf93a2789640ull}, {0xf11268128982754full, 0x257a1d670430b0aaull},
    {0xe51c98ce7d1de664ull, 0x5f9478a733040c45ull}, {0x6d7b49e7e429850aull, 0x2e3063c622a24777ull},
    {0xbd90d5377ba1b762ull, 0xc07317d419a7548dull}, {0x53d39c6752dac858ull, 0xbcd1c5a80ab65b3eull},
    {0xb4a24d7a84e7677bull, 0x023ff9668e89b5c4ull}, {0xdffa22b534c5f608ull, 0xb9b67517d3665ca9ull},
    {0xd50708086cef4d7cull, 0x6e1651ecc7f43309ull}};

static const dsx_noise_cfg DSX_NOISE[6] = {
    {0x5c7e6b29735f0d7full, 0xf7d86f1bbc734988ull, -10, 6, {1.5, 0, 1, 0, 0, 0}},
    {0x81bb4d22e8dc168eull, 0xf1c8b4bea16303cdull, -8, 6, {1, 1, 0, 0, 0, 0}},
    {0x83886c9d0ae3a662ull, 0xafa638a61b42e8adull, -9, 9, {1, 1, 2, 2, 2, 1, 1, 1, 1}},
    {0xd02491e6058f6fd8ull, 0x4792512c94c17a80ull, -9, 5, {1, 1, 0, 1, 1}},
    {0x080518cf6af25384ull, 0x3f3dfb40a54febd5ull, -3, 4, {1, 1, 1, 0}},
    {0xefc8ef4d36102b34ull, 0x1beeeb324a0f24eaull, -7, 6, {1, 2, 1, 0, 0, 0}}};

// World-seed whitening constant and golden-ratio increment of the seeded RNG.
#define DSX_SEED_XOR 0x6a09e667f3bcc909ull
#define DSX_GOLDEN 0x9e3779b97f4a7c15ull
