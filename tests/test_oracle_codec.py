"""Pins the oracle's 5-bit codec (sseqs/xbit.cu:8-62) by hand vectors,
the closed-form bit layout and round trips."""
import numpy as np
import pytest


def _bits_msb_first(syms):
    """Reference layout restated as a bit string: symbol q at bits 5q..5q+4."""
    bits = "".join(format(s & 31, "05b") for s in syms)
    bits += "0" * (-len(bits) % 8)
    return np.array([int(bits[i:i + 8], 2) for i in range(0, len(bits), 8)], dtype=np.uint8)


def test_hand_vector(orc):
    # '@' = 0, 'A' = 1, ..., '_' = 31: "@ABCDEFG" -> 00000 00001 00010 ... 00111
    s = b"@ABCDEFG"
    packed = orc.xbit_pack(s)
    assert packed.tolist() == [0b00000000, 0b01000100, 0b00110010, 0b00010100, 0b11000111]
    assert orc.xbit_unpack(packed, 8).tobytes() == s


def test_all_symbols_window(orc):
    s = bytes(range(64, 96))  # every 5-bit value once
    packed = orc.xbit_pack(s)
    assert packed.tolist() == _bits_msb_first([c - 64 for c in s]).tolist()
    assert orc.xbit_unpack(packed, 32).tobytes() == s


@pytest.mark.parametrize("n", [0, 1, 7, 8, 9, 15, 16, 17, 1000, 4099])
def test_roundtrip_any_length(orc, n):
    rng = np.random.default_rng(n)
    s = (rng.integers(0, 32, n) + 64).astype(np.uint8)
    packed = orc.xbit_pack(s)
    assert packed.size == (5 * n + 7) // 8
    assert packed.tolist() == _bits_msb_first((s - 64).tolist()).tolist()
    assert np.array_equal(orc.xbit_unpack(packed, n), s)


def test_unpack_offset(orc):
    rng = np.random.default_rng(3)
    s = (rng.integers(0, 32, 800) + 64).astype(np.uint8)
    packed = orc.xbit_pack(s)
    # byte offset lo = 5*g starts at symbol 8*g (sseqs/xbit.cu:69, lo = j*bs)
    for g in (0, 1, 7, 50):
        out = orc.xbit_unpack(packed, 800 - 8 * g, 5 * g)
        assert np.array_equal(out, s[8 * g:])
