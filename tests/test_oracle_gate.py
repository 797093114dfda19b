"""Pins for the oracle's restatement of the GPU gate (gate.cuh:205-326), which has no CPU counterpart
in the reference: closed forms, identities with reference-pinned functions, brute force."""
import numpy as np
import pytest

T_W0 = 1.25 * (32.0 / 63.0) * 1.5
SEEDS = [0x9E3779B97F4A7C15, 3008578943537042204, 12345, (1 << 64) - 7]


def test_phase_pass_fraction_closed_form(oracle):
    """|fyA - 1/2| + |fyB - 1/2| <= s has probability 2 s^2 for uniform fy (s = 0.06 -> 0.72%)."""
    n = 1 << 20
    k = oracle.phase_list(0, n, 0.06)
    p = 2 * 0.06 ** 2
    sd = (n * p * (1 - p)) ** 0.5
    assert abs(len(k) - n * p) < 5 * sd


def test_phase_fy_matches_reference_seeds_csv(oracle):
    """fyA/fyB of the reference's seeds CSV (%.4f) come from the same phase computation."""
    import ctypes, os
    from conftest import GOLDEN
    rows = open(os.path.join(GOLDEN, "c1.seeds.csv")).read().splitlines()[:50]
    for r in rows:
        f = r.split(",")
        fy = (ctypes.c_double * 2)()
        oracle.or_phase(int(f[0]) % (1 << 64), fy)
        assert "%.4f" % fy[0] == f[1] and "%.4f" % fy[1] == f[2]


@pytest.mark.parametrize("seed", SEEDS)
def test_full_grid_is_reference_tabulation(oracle, seed):
    """The full gate's 512^2 grid equals lowfield.hpp tabulate_octave (pinned by the reference's
    maxta/maxtb outputs) run with octave weight float(T_W0): same corner terms, same sum order."""
    w = float(np.float32(T_W0))
    for side in (0, 1):
        V = oracle.gate_grid(seed, side, 512, 0)
        t = np.zeros((512, 512), np.float32)
        oracle.or_tabulate_side(seed, side, w, t.ctypes.data)
        assert np.array_equal(V.view(np.uint32), t.view(np.uint32))


@pytest.mark.parametrize("seed", SEEDS)
def test_legacy_grid_is_even_parity_of_full(oracle, seed):
    V = oracle.gate_grid(seed, 0, 512, 0)
    L = oracle.gate_grid(seed, 0, 256, 0)
    assert np.array_equal(L.view(np.uint32), V[::2, ::2].view(np.uint32))


@pytest.mark.parametrize("seed", SEEDS[:2])
def test_box_grid_is_mean_of_four_parities(oracle, seed):
    """Anti-aliased cell (i, j) = w/4 * sum over the 4 half-lattice parities of the octave sampled
    directly (improved noise at lattice (i + px/2, j + pz/2)). The merged 3x3 table only regroups the
    corner terms, so the two agree to float rounding (1e-6); a wrong shift or index fails by ~0.1."""
    w = float(np.float32(T_W0))
    V = oracle.gate_grid(seed, 1, 256, 1)
    rng = np.random.default_rng(seed % 1000)
    for i, j in rng.integers(0, 256, size=(200, 2)):
        d = sum(oracle.or_perlin_side(seed, 1, i + 0.5 * px, j + 0.5 * pz) for px in (0, 1) for pz in (0, 1))
        assert abs(V[j, i] - 0.25 * w * d) < 1e-6


def _blur_max_numpy(V, R, k):
    """Sequential float32 blur (sum from 0, tap order), as gate.cuh:296-313, then max * norm."""
    G = V.shape[0]
    H = np.zeros_like(V)
    for d in range(2 * R + 1):
        H = (H + np.float32(k[d]) * np.roll(V, -(d - R), axis=1)).astype(np.float32)
    O = np.zeros_like(V)
    for d in range(2 * R + 1):
        O = (O + np.float32(k[d]) * np.roll(H, -(d - R), axis=0)).astype(np.float32)
    ks = np.float32(0)
    for d in range(2 * R + 1):
        ks = np.float32(ks + k[d])
    norm = np.float32(np.float32(1) / np.float32(ks * ks))
    return np.float32(O.max() * norm)


@pytest.mark.parametrize("G,R,sig", [(256, 4, 1.0), (512, 7, 2.0)])
def test_blur_max_bruteforce(oracle, G, R, sig):
    k = oracle.gate_kern(R, sig)
    rng = np.random.default_rng(G)
    V = rng.standard_normal((G, G)).astype(np.float32) * np.float32(0.2)
    got = oracle.or_gate_blur_max(V.ctypes.data, G, R, k.ctypes.data)
    assert np.float32(got).view(np.uint32) == _blur_max_numpy(V, R, k).view(np.uint32)
    # periodic: a cyclic shift of the grid leaves the maximum bit-identical
    Vs = np.ascontiguousarray(np.roll(np.roll(V, 37, axis=0), -91, axis=1))
    assert np.float32(oracle.or_gate_blur_max(Vs.ctypes.data, G, R, k.ctypes.data)).view(np.uint32) == \
        np.float32(got).view(np.uint32)


def test_gate_kernel_weights(oracle):
    """expf(-0.5 d^2 / sig^2) in float: symmetric, centre 1, tails match exp in double to 1 ulp."""
    for R, sig in ((4, 1.0), (7, 2.0)):
        k = oracle.gate_kern(R, sig)
        assert k[R] == 1.0 and np.array_equal(k, k[::-1])
        ref = np.exp(-0.5 * (np.arange(-R, R + 1) / sig) ** 2)
        assert np.all(np.abs(k - ref) <= np.spacing(k.astype(np.float32)))


def test_gate_run_consistent_with_amplitudes(oracle):
    """or_gate_run's survivor set equals thresholding the per-seed amplitudes (4-parity pre-gate,
    then the full gate) over the phase survivors, brute force on a small range."""
    start, count = 10 ** 12, 1 << 16
    idx, amps, cnt = oracle.gate_run(start, count)
    ks = oracle.phase_list(start, count, 0.06)
    assert cnt[0] == len(ks)
    want = []
    for k in ks:
        seed = oracle.or_mix(start + int(k))
        if oracle.or_gate_amp(seed, 256, 4, 1, 1.0) >= np.float32(0.645):
            a = oracle.or_gate_amp(seed, 512, 7, 0, 2.0)
            if a >= np.float32(0.70):
                want.append((int(k), a))
    assert [int(i) for i in idx] == [k for k, _ in want]
    assert np.array_equal(amps, np.array([a for _, a in want], np.float32))
