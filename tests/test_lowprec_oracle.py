"""CPU checks of the low-precision code restatement against the published formats."""
import numpy as np

import lowprec


def test_bf16_encode_known_values():
    y = np.array([1.0, -2.5, 0.0, 3.0e-39, 1.00390625, 1.01171875], np.float32)
    c = lowprec.bf16_encode(y)
    assert c[0] == 0x3F80 and c[1] == 0xC020 and c[2] == 0
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: rounds to even (1.0)
    assert c[4] == 0x3F80
    # 1 + 3*2^-8 is a tie between 1+2^-7 and 1+2^-6: rounds to even (1 + 2^-6)
    assert c[5] == 0x3F82
    assert np.array_equal(lowprec.bf16_decode(c)[:3], y[:3])


def test_bf16_power_of_two_scale_is_lossless_on_small_integers():
    x = np.arange(0, 257, dtype=np.float32)[None, :]
    codes, s, deq = lowprec.quantize(x, "bf16")
    assert s[0] == 512.0
    assert np.array_equal(deq, x)


def test_e4m3_known_codes():
    # OCP E4M3: 1.0 = 0x38, 448 = 0x7E, smallest subnormal 2^-9 = 0x01, 2^-6 = 0x08
    y = np.array([1.0, 448.0, 2.0 ** -9, 2.0 ** -6, -1.5, 0.0], np.float32)
    c = lowprec.e4m3_encode(y)
    assert list(c) == [0x38, 0x7E, 0x01, 0x08, 0xBC, 0x00]
    assert np.array_equal(lowprec.e4m3_decode(c), y)


def test_e4m3_round_trip_all_codes_and_error_bound():
    codes = np.array([c for c in range(256) if (c & 0x7F) != 0x7F], np.uint8)  # skip NaNs
    vals = lowprec.e4m3_decode(codes)
    assert np.array_equal(lowprec.e4m3_encode(vals)[vals != 0], codes[vals != 0])
    rng = np.random.default_rng(0)
    y = rng.uniform(-448, 448, 10000).astype(np.float32)
    d = lowprec.e4m3_decode(lowprec.e4m3_encode(y))
    big = np.abs(y) >= 2.0 ** -6
    assert np.all(np.abs(d - y)[big] <= np.abs(y)[big] * 2.0 ** -4 + 1e-12)
