"""PCG32 of the oracle against published known answers.

The reference's Sampler wraps pcg32.h (src/core/sampler.cpp:14-60); no KAT
exists in the reference's tests, so the stream is pinned against the output
of the public pcg-c-basic demo (pcg32-demo, seed 42, sequence 54)."""
import numpy as np

import oracle

# pcg32-demo, "Round 1, 32bit" line (pcg_random.org, pcg-c-basic).
PCG_DEMO_42_54 = [0xa15c02b7, 0x7b47f409, 0xba1d3330, 0x83d2f293, 0xbfa4784b, 0xcbed606e]


def test_pcg32_demo_kat():
    assert list(oracle.pcg32_u32(42, 54, 6)) == PCG_DEMO_42_54


def test_pcg32_float_mapping():
    u = oracle.pcg32_u32(7, 3, 1000)
    f = oracle.pcg32_f32(7, 3, 1000)
    expect = ((u >> 9).astype(np.float64)) / float(1 << 23)
    assert np.array_equal(f.astype(np.float64), expect)
    assert f.min() >= 0.0 and f.max() < 1.0


def test_pcg32_streams_differ_and_are_uniform():
    a = oracle.pcg32_f32(1, 1, 200000)
    b = oracle.pcg32_f32(1, 2, 200000)
    assert not np.array_equal(a[:16], b[:16])
    hist, _ = np.histogram(a, bins=20, range=(0, 1))
    chi2 = ((hist - 10000.0) ** 2 / 10000.0).sum()
    assert chi2 < 45.0   # 19 dof, p ~ 1e-3
