"""The FP32 uniforms of the device streams (Stream::u32 and Stream::v32 in
csrc/tm_device.h) are formed from bits and one fused operation.  Their values are
those of the plain arithmetic definitions, for every word:

  u32(w) = ((w >> 9) + 1/2) 2^-23, built as float(0x3F800000 | w >> 9) - (1 - 2^-24)
           (the bit pattern v_alignbit_b32(0x7F, w, 9) gives);
  v32(w) = (float(~w) + 1/2) 2^-32 with float rounding, built as
           fma(float(~w), 2^-32, 2^-33) (one rounding of the exact value).

Checked here in numpy on every edge class of w and on 2^20 random words, against
the definitions evaluated step by step in float32.  No GPU needed."""
import numpy as np


def words():
    rng = np.random.default_rng(5)
    edges = [0, 1, 2, 511, 512, 513, (1 << 23) - 1, 1 << 23, (1 << 24) - 1, 1 << 24, (1 << 24) + 1,
             (1 << 31) - 1, 1 << 31, (1 << 32) - 512, (1 << 32) - 2, (1 << 32) - 1]
    near = [e + d for e in (1 << 23, 1 << 24, 1 << 25, 1 << 31, (1 << 32) - (1 << 8)) for d in range(-600, 600)]
    w = np.concatenate([np.array(edges + near, np.int64) % (1 << 32),
                        rng.integers(0, 1 << 32, 1 << 20, dtype=np.int64)])
    return w.astype(np.uint32)


def test_u32_from_bits_equals_the_definition():
    w = words()
    f32 = np.float32
    want = ((w >> 9).astype(f32) + f32(0.5)) * f32(2.0 ** -23)
    one_plus = ((w >> 9) | np.uint32(0x3F800000)).view(f32)
    got = one_plus - f32(1.0 - 2.0 ** -24)
    assert got.dtype == f32 and np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert got.min() > 0 and got.max() < 1
    # a plausible slip, subtracting 1 (the half step dropped), is caught
    assert not np.array_equal(one_plus - f32(1.0), want)


def test_v32_single_rounding_equals_the_definition():
    w = words()
    f32 = np.float32
    x = (~w).astype(f32)                   # rounded to 24 significant bits
    want = (x + f32(0.5)) * f32(2.0 ** -32)
    exact = x.astype(np.float64) * 2.0 ** -32 + 2.0 ** -33   # exact in FP64 (at most 33 bits)
    got = exact.astype(f32)                # the fused multiply-add's one rounding
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert got.min() > 0 and got.max() <= 1
