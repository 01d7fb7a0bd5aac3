"""Low-precision row codes in numpy -- TEST INFRASTRUCTURE ONLY.

The precision form has no counterpart in the reference; its quantizers are
pinned here by the published formats:
  bfloat16: the upper 16 bits of an IEEE float32, rounded to nearest even;
  FP8 E4M3 (OCP "e4m3fn"): 1 sign, 4 exponent (bias 7), 3 mantissa bits,
            subnormals at exponent 0 (step 2^-9), no infinities, max 448.
Per-row scales: BF16 uses the power of two 2^e with max|x| = m * 2^e, m in
[0.5, 1); FP8 uses max|x| / 448; an all-zero row has scale 1.
"""
from __future__ import annotations

import numpy as np


def bf16_scales(x: np.ndarray) -> np.ndarray:
    amax = np.abs(x).max(axis=1).astype(np.float32)
    _, e = np.frexp(amax)
    s = np.ldexp(np.float32(1.0), e).astype(np.float32)
    return np.where(amax > 0, s, np.float32(1.0)).astype(np.float32)


def bf16_encode(y: np.ndarray) -> np.ndarray:
    u = np.ascontiguousarray(y, dtype=np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def bf16_decode(c: np.ndarray) -> np.ndarray:
    return (c.astype(np.uint32) << 16).view(np.float32)


def fp8_scales(x: np.ndarray) -> np.ndarray:
    amax = np.abs(x).max(axis=1).astype(np.float32)
    return np.where(amax > 0, amax / np.float32(448.0), np.float32(1.0)).astype(np.float32)


def _rne(v: np.ndarray) -> np.ndarray:
    return np.round(v)  # numpy rounds halves to even


def e4m3_encode(y: np.ndarray) -> np.ndarray:
    """float32 values in [-448, 448] -> OCP E4M3 codes, round to nearest even."""
    y = np.asarray(y, dtype=np.float64)
    sign = (np.signbit(y)).astype(np.uint8)
    a = np.minimum(np.abs(y), 448.0)
    code = np.zeros(a.shape, dtype=np.int64)
    sub = a < 2.0 ** -6
    m_sub = _rne(a / 2.0 ** -9).astype(np.int64)  # 0..8; 8 is the smallest normal
    code = np.where(sub, m_sub, code)
    with np.errstate(divide="ignore"):
        e = np.floor(np.log2(np.where(sub, 1.0, a))).astype(np.int64)
    frac = np.where(sub, 0.0, a / np.ldexp(1.0, e) - 1.0)
    m = _rne(frac * 8).astype(np.int64)
    e = np.where(m == 8, e + 1, e)
    m = np.where(m == 8, 0, m)
    norm = ((e + 7) << 3) | m
    code = np.where(sub, code, norm)
    return (code.astype(np.uint8) | (sign << 7)).astype(np.uint8)


def e4m3_decode(c: np.ndarray) -> np.ndarray:
    c = np.asarray(c, dtype=np.uint8).astype(np.int64)
    s = np.where(c >> 7, -1.0, 1.0)
    e = (c >> 3) & 15
    m = c & 7
    v = np.where(e == 0, m * 2.0 ** -9, (1.0 + m / 8.0) * np.ldexp(1.0, e - 7))
    return (s * v).astype(np.float32)


def quantize(x: np.ndarray, storage: str):
    """(codes, scales, dequantized float32 rows) as the GPU quantizer defines them."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    if storage == "bf16":
        s = bf16_scales(x)
        codes = bf16_encode(x / s[:, None])
        return codes, s, bf16_decode(codes) * s[:, None]
    s = fp8_scales(x)
    y = np.clip(x / s[:, None], -448.0, 448.0).astype(np.float32)
    codes = e4m3_encode(y)
    return codes, s, e4m3_decode(codes) * s[:, None]
