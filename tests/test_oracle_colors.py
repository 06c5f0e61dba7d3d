"""Pins the oracle's compressed-colour codec (variable_weight_size_colors.h:10-721,
color_utils.h:112-148) against an independent model written here:

* a leaf is a list of blocks; block k starts at local colour index idx_k of its 16384-colour
  macro block, carries colorBits and a bits-per-weight b_k; a new block starts at every
  macro boundary and whenever colorBits or b changes;
* weights form one MSB-first bit stream over the bytes of the stored (byte-swapped) words;
  colour i of block k reads b_k bits at macroWeight + blockWeight + (local - idx_k) * b_k.

Decoding every index through the oracle must return the encoded sequence, and the
oracle's stored arrays must match this model field by field.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import orc


def _p(a, t=C.c_uint32):
    return a.ctypes.data_as(C.POINTER(t))


def encode(bits, w, bpw):
    n = len(bits)
    bits = np.ascontiguousarray(bits, np.uint32)
    w = np.ascontiguousarray(w, np.uint8)
    bpw = np.ascontiguousarray(bpw, np.uint8)
    capB, capW, capM = n + 16, n + 16, 2 * (n // 16384 + 2)
    blocks = np.zeros(capB, np.uint64)
    weights = np.zeros(capW, np.uint32)
    macro = np.zeros(capM, np.uint64)
    sizes = np.zeros(3, np.uint64)
    assert orc.orc_encode_colors(_p(bits), _p(w, C.c_uint8), _p(bpw, C.c_uint8), n, _p(blocks, C.c_uint64),
                                 _p(weights), _p(macro, C.c_uint64), _p(sizes, C.c_uint64), capB, capW, capM) == 0
    return blocks[: sizes[0]], weights[: sizes[1]], macro[: sizes[2]]


def decode(blocks, weights, macro, offset, i):
    out = np.zeros(2, np.uint32)
    r = orc.orc_decode_color(_p(blocks, C.c_uint64), blocks.size, _p(weights), weights.size, _p(macro, C.c_uint64),
                             macro.size, offset, i, _p(out))
    return None if r else (int(out[0]), int(out[1]) & 0xFF, int(out[1]) >> 8)


def model(bits, w, bpw):
    """Independent model of the encoder output (blocks as tuples, bit stream as a str)."""
    blocks, stream, macros = [], [], []
    last = None
    for i in range(len(bits)):
        if i % 16384 == 0:
            macros.append((len(blocks), len(stream)))
            last = None
        key = (int(bits[i]), int(bpw[i]))
        if key != last:
            mw = macros[-1][1]
            blocks.append((key[0], key[1], len(stream) - mw, i % 16384))
            last = key
        if bpw[i]:
            stream.extend(format(int(w[i]), f"0{int(bpw[i])}b"))
    return blocks, "".join(stream), macros


def stored_stream(weights):
    return "".join(format(b, "08b") for b in weights.astype("<u4").tobytes())


def random_sequence(rng, n, runs=True):
    bits = np.zeros(n, np.uint32)
    bpw = np.zeros(n, np.uint8)
    i = 0
    while i < n:
        length = int(rng.integers(1, 40)) if runs else 1
        b = int(rng.integers(0, 5))
        bits[i:i + length] = rng.integers(0, 2**32)
        bpw[i:i + length] = b
        i += length
    w = np.array([rng.integers(0, 1 << b) if b else 0 for b in bpw], np.uint8)
    return bits, w, bpw


@pytest.mark.parametrize("n,seed", [(1, 0), (37, 1), (5000, 2), (16384, 3), (16385, 4), (40000, 5)])
def test_round_trip_and_model(n, seed):
    rng = np.random.default_rng(seed)
    bits, w, bpw = random_sequence(rng, n)
    blocks, weights, macro = encode(bits, w, bpw)
    mblocks, mstream, mmacros = model(bits, w, bpw)
    # block layout: colour bits, header fields
    assert len(blocks) == len(mblocks)
    for b, (cb, bp, woff, idx) in zip(blocks.tolist(), mblocks):
        hdr = b & 0xFFFFFFFF
        assert b >> 32 == cb
        assert hdr & 0x3FFF == idx
        if bp == 0:
            assert hdr >> 16 == 0xFFFF
        else:
            assert hdr >> 16 == woff and ((hdr >> 14) & 3) + 1 == bp
    assert macro.tolist() == [v for m in mmacros for v in m]
    # weights: the stored bytes read MSB-first are the model stream (zero padded)
    s = stored_stream(weights)
    assert s[: len(mstream)] == mstream and set(s[len(mstream):]) <= {"0"}
    # decode every index
    for i in range(n):
        cb, ww, bp = decode(blocks, weights, macro, ~np.uint64(0), i)
        assert (cb, bp) == (int(bits[i]), int(bpw[i]))
        assert ww == (int(w[i]) if bpw[i] else 0)


def test_shared_offset_and_validity():
    rng = np.random.default_rng(9)
    bits, w, bpw = random_sequence(rng, 20000)
    blocks, weights, macro = encode(bits, w, bpw)
    for off in (0, 7, 16384):
        for i in range(0, 20000 - off, 997):
            cb, ww, bp = decode(blocks, weights, macro, off, i)
            assert cb == bits[i + off]
    # is_valid_index: index < 16384 * macro blocks (2 * index / 16384 < macro.size())
    assert decode(blocks, weights, macro, ~np.uint64(0), 2 * 16384 - 1) is not None
    assert decode(blocks, weights, macro, ~np.uint64(0), 2 * 16384) is None
