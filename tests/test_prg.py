"""The counter-based sampler of seeded clients (onionpirv2_amd/host/prg.h),
CPU only. ChaCha20 is pinned by the known-answer block of RFC 8439 section
2.3.2 and by an independent restatement of the RFC written here; the
sampling rules on top of it (uniform residues, ternary secret, rounded
Gaussian by a cumulative table) by restating them from the raw blocks, and
the error distribution by its moments.
"""
import math
import struct

import numpy as np
import pytest


def _rotl(x, n):
    return ((x << n) | (x >> (32 - n))) & 0xFFFFFFFF


def _qr(s, a, b, c, d):
    s[a] = (s[a] + s[b]) & 0xFFFFFFFF; s[d] = _rotl(s[d] ^ s[a], 16)
    s[c] = (s[c] + s[d]) & 0xFFFFFFFF; s[b] = _rotl(s[b] ^ s[c], 12)
    s[a] = (s[a] + s[b]) & 0xFFFFFFFF; s[d] = _rotl(s[d] ^ s[a], 8)
    s[c] = (s[c] + s[d]) & 0xFFFFFFFF; s[b] = _rotl(s[b] ^ s[c], 7)


def chacha20(key, counter, nonce_words):
    """RFC 8439 section 2.3 block function; nonce given as three words."""
    st = [0x61707865, 0x3320646E, 0x79622D32, 0x6B206574] + list(struct.unpack("<8I", key)) + [counter] + list(nonce_words)
    w = st[:]
    for _ in range(10):
        _qr(w, 0, 4, 8, 12); _qr(w, 1, 5, 9, 13); _qr(w, 2, 6, 10, 14); _qr(w, 3, 7, 11, 15)
        _qr(w, 0, 5, 10, 15); _qr(w, 1, 6, 11, 12); _qr(w, 2, 7, 8, 13); _qr(w, 3, 4, 9, 14)
    return [(a + b) & 0xFFFFFFFF for a, b in zip(w, st)]


def test_chacha20_rfc8439_block():
    import onionpirv2_amd.host as H
    key = bytes(range(32))
    nonce = struct.unpack("<3I", bytes.fromhex("000000090000004a00000000"))
    want = bytes.fromhex("10f1e7e4d13b5915500fdd1fa32071c4c7d1f4c733c068030422aa9ac3d46c4e"
                         "d2826446079faa0914c2d705d98b02a2b5129cd1de164eb9cbd083e8a2503c4e")
    assert struct.pack("<16I", *chacha20(key, 1, nonce)) == want
    assert H.chacha20_block(key, 1, list(nonce)) == want


def test_chacha20_matches_restatement_on_random_inputs():
    import onionpirv2_amd.host as H
    rng = np.random.default_rng(1)
    for _ in range(20):
        key = rng.integers(0, 256, size=32, dtype=np.uint8).tobytes()
        ctr = int(rng.integers(0, 1 << 32))
        nonce = [int(x) for x in rng.integers(0, 1 << 32, size=3)]
        assert H.chacha20_block(key, ctr, nonce) == struct.pack("<16I", *chacha20(key, ctr, nonce))


def _draws128(key, domain, i, j, n):
    out = []
    for b in range((n + 3) // 4):
        w = chacha20(key, b, (domain, i, j))
        for m in range(4):
            lo = w[4 * m] | w[4 * m + 1] << 32
            hi = w[4 * m + 2] | w[4 * m + 3] << 32
            out.append(hi << 64 | lo)
    return out[:n]


def _draws64(key, domain, i, j, n):
    out = []
    for b in range((n + 7) // 8):
        w = chacha20(key, b, (domain, i, j))
        out += [w[2 * m] | w[2 * m + 1] << 32 for m in range(8)]
    return out[:n]


@pytest.mark.parametrize("q", [288135826036838401, 536813569, 1152921504606830593, 3])
def test_uniform_rule(q):
    import onionpirv2_amd.host as H
    key = bytes(range(100, 132))
    got = H.prg_uniform(key, 4, 17, 1, q, 37)
    assert [int(x) for x in got] == [d % q for d in _draws128(key, 4, 17, 1, 37)]


def test_ternary_rule():
    import onionpirv2_amd.host as H
    key = bytes(range(7, 39))
    got = H.prg_ternary(key, 50)
    assert [int(x) for x in got] == [[0, 1, -1][d % 3] for d in _draws64(key, 1, 0, 0, 50)]
    big = H.prg_ternary(key, 30000)
    counts = [int((big == v).sum()) for v in (-1, 0, 1)]
    assert all(abs(c - 10000) < 500 for c in counts)


def test_gaussian_table_and_rule():
    """Entry v of the table: floor(P(round(X) <= v - 24) 2^64) for X ~ N(0,
    sigma^2), i.e. Phi((v - 24 + 1/2) / sigma) (last entry saturated); a
    draw u maps to the first entry >= u."""
    import onionpirv2_amd.host as H
    sigma = 2.55
    cdt = H.gaussian_cdt(sigma)
    assert len(cdt) == 49 and cdt[-1] == (1 << 64) - 1
    for v in range(-24, 24):
        p = 0.5 * math.erfc(-(v + 0.5) / (sigma * math.sqrt(2)))
        assert abs(cdt[v + 24] - p * 2.0 ** 64) <= max(2.0 ** 12, p * 2.0 ** 64 * 1e-12), v
    assert all(a <= b for a, b in zip(cdt, cdt[1:]))
    key = bytes(range(50, 82))
    got = H.prg_gaussian(key, 5, 9, sigma, 40)
    want = []
    for u in _draws64(key, 5, 9, 0, 40):
        v = 0
        while u > cdt[v]:
            v += 1
        want.append(v - 24)
    assert [int(x) for x in got] == want
    big = H.prg_gaussian(key, 5, 10, sigma, 200000)
    assert abs(big.mean()) < 0.03 and abs(big.var() - (sigma ** 2 + 1 / 12)) < 0.15
    assert np.abs(big).max() < 24


def test_query_seed_rule():
    import onionpirv2_amd.host as H
    key = bytes(range(32))
    w = chacha20(key, 0, (8, 5, 1))
    assert H.query_seed(key, (1 << 32) + 5) == struct.pack("<8I", *w[:8])
