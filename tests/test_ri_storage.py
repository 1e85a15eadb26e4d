"""Compensated FP32 storage of the RI factors (VBE_PRECISION_FP32C_RI), checked on the CPU against
what the format's mathematics fixes: hi = fp32(b), k = round((b - hi) 2^(158 - e)) clamped to int8, e
the biased exponent of hi (at least 1); the decoded hi + k 2^(e - 158) is exact in FP64 and within
2^-31 |b| of b (2^-32 unless k is clamped, which happens when b lies within 2^-8 ulp(hi) of a rounding
tie of fp32). The device encoder and decoder follow these formulas; their results are checked by
the GPU parity tests (tests/test_gpu_parity.py, test_gpu_precision.py)."""
import numpy as np


def encode(b):
    hi = b.astype(np.float32)
    e = ((hi.view(np.uint32) >> 23) & 0xFF).astype(np.int64)
    e = np.maximum(e, 1)
    r = (b - hi.astype(np.float64)) * np.ldexp(1.0, 158 - e)
    k = np.clip(np.rint(r), -128, 127).astype(np.int8)
    return hi, k


def decode(hi, k):
    e = ((hi.view(np.uint32) >> 23) & 0xFF).astype(np.int64)
    e = np.maximum(e, 1)
    return hi.astype(np.float64) + k.astype(np.float64) * np.ldexp(1.0, e - 158)


def test_decoded_factors_within_two_to_minus_32():
    rng = np.random.default_rng(7)
    # the synthetic generator's scale (B ~ N(0, 1/n_bf)) and a wide range of magnitudes
    b = np.concatenate([rng.normal(scale=np.sqrt(1 / 120), size=200000),
                        rng.normal(size=50000) * np.exp(rng.uniform(-40, 40, size=50000))])
    hi, k = encode(b)
    d = decode(hi, k)
    rel = np.abs(d - b) / np.abs(b)
    assert rel.max() <= 2.0 ** -31
    clamped = np.abs(k.astype(np.int64)) == 127
    assert rel[~clamped].max() <= 2.0 ** -32
    # plain FP32 storage is 2^8 times coarser
    assert (np.abs(hi.astype(np.float64) - b) / np.abs(b)).max() > 2.0 ** -26


def test_decode_is_exact_in_fp64():
    rng = np.random.default_rng(8)
    b = rng.normal(size=10000)
    hi, k = encode(b)
    # hi has 24 significant bits and k 2^(e-158) sits 8 bits below hi's last bit: the sum has at
    # most 33 significant bits, so the FP64 addition is exact, in either order
    d1 = hi.astype(np.float64) + k.astype(np.float64) * np.ldexp(1.0, np.maximum(((hi.view(np.uint32) >> 23) & 0xFF).astype(np.int64), 1) - 158)
    d2 = k.astype(np.float64) * np.ldexp(1.0, np.maximum(((hi.view(np.uint32) >> 23) & 0xFF).astype(np.int64), 1) - 158) + hi.astype(np.float64)
    assert np.array_equal(d1, d2)
    assert np.array_equal((d1 - hi.astype(np.float64)).astype(np.float64) * np.ldexp(1.0, 158 - np.maximum(((hi.view(np.uint32) >> 23) & 0xFF).astype(np.int64), 1)), k.astype(np.float64))


def test_zero_and_edge_values():
    b = np.array([0.0, 1.0, -1.0, 2.0 ** -126, 3.0e38, 1.0 - 2.0 ** -30, np.nextafter(1.0, 2.0)])
    hi, k = encode(b)
    d = decode(hi, k)
    assert d[0] == 0.0 and d[1] == 1.0 and d[2] == -1.0
    nz = b != 0.0
    assert np.all(np.abs(d[nz] - b[nz]) <= 2.0 ** -31 * np.abs(b[nz]))
