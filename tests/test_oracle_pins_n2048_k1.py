"""Pins for the oracle's single-prime path (CONFIG_N2048_K1, database_constants.h:61-75),
CPU only: N = 2048, one ~60-bit NTT prime, the single-modulus gadgets
(gsw.cpp:252-315, bv_keyswitch.cpp:20-40,296-367), the u64 first dimension
(matrix.cpp:24-151, server.cpp:408-527) and the direct centred rescale
(server.cpp:1034-1043).

The reference's tests are compiled for one config at a time and hold no
fixtures for this one, so each piece is pinned by what the mathematics
fixes: closed-form constants, brute-force NTT evaluation, exact balanced
digits (ties, boundaries), exact rescale against rationals, RGSW selection,
BV keyswitch = automorphism and end-to-end decryption. Parity unpinned, as for
the other schemes: the NTT evaluation order (HEXL's bit-reversed convention)
and the choice of minimal primitive root.
"""
from fractions import Fraction

import numpy as np
import pytest

from test_oracle_pins_k2 import _is_prime, _largest_prime_1_mod, _smallest_prime_above, brv


@pytest.fixture(scope="module")
def p1(O):
    return O.PirParams("CONFIG_N2048_K1")


def _centred(x, q):
    return x - q if x > q // 2 else x


def test_n2048_k1_constants_exact(O, p1):
    """pir.cpp:14-98 rules: q the largest 60-bit prime 1 mod 2N, t the
    smallest prime above 2^13 (PlainMod 14), small_q the largest 22-bit prime
    1 mod 2N; data / key gadget widths ceil(60 / l), BV width 60 / L_KS + 1
    (bv_keyswitch.cpp:73-76)."""
    assert p1.N == 2048 and len(p1.rns_mods) == 1
    q = p1.rns_mods[0]
    assert q == _largest_prime_1_mod(60, 2 * p1.N) and _is_prime(q) and q.bit_length() == 60
    assert p1.plain_mod == _smallest_prime_above(1 << 13) == 8209
    assert p1.small_q == _largest_prime_1_mod(22, 2 * p1.N) == 4169729
    assert (p1.l_ep, p1.base_log2, p1.l_key, p1.base_log2_key, p1.l_ks, p1.bv_base_log2) == (5, 12, 8, 8, 8, 8)


def test_n2048_k1_ntt_brute_force(O, p1):
    N, q = p1.N, p1.rns_mods[0]
    psi = O.ntt_root(N, q)
    assert pow(psi, N, q) == q - 1
    rng = np.random.default_rng(3)
    a = rng.integers(0, q, size=N, dtype=np.uint64)
    got = O.ntt_fwd(a, N, q)
    ai = [int(x) for x in a]
    for j in [0, 1, 1234, N - 1]:
        w = pow(psi, 2 * brv(j, 11) + 1, q)
        assert int(got[j]) == sum(x * pow(w, i, q) for i, x in enumerate(ai)) % q
    assert np.array_equal(O.ntt_inv(got, N, q), a)


def _balanced_reference(v, q, b, nd):
    """Balanced base-2^b digits of the centred v, LSB-first, each in
    [-2^(b-1), 2^(b-1)), computed by exact integer division with the low
    digit rounded toward the tie's negative side (sign extension)."""
    d, out, B = _centred(v, q), [], 1 << b
    for _ in range(nd):
        r = d % B
        if r >= B // 2:
            r -= B
        out.append(r)
        d = (d - r) // B
    return out


def test_n2048_k1_signed_digits_exact(O, p1):
    """signed_gadget_decompose (bv_keyswitch.cpp:20-40) at the 60-bit prime:
    digits in [-B/2, B/2) (a low digit of exactly B/2 becomes -B/2 with a
    carry), stored mod q, recomposing the centred value exactly."""
    q, b, nd = p1.rns_mods[0], p1.bv_base_log2, p1.l_ks
    B = 1 << b
    vals = [0, 1, q - 1, q // 2, q // 2 + 1, B // 2, q - B // 2, (B // 2) * (1 + B), q - (B // 2) * (1 + B + B * B),
            (3 * B + B // 2) * B, 123456789012345678]
    vals += [int(x) for x in O.mt19937_64_draws(77, 500) % np.uint64(q)]
    for v in vals:
        got = [int(x) for x in O.signed_gadget_decompose(v, b, q, nd)]
        signed = [x if x < q // 2 else x - q for x in got]
        assert signed == _balanced_reference(v, q, b, nd), v
        assert all(-B // 2 <= s < B // 2 for s in signed)
        assert sum(s * B ** i for i, s in enumerate(signed)) == _centred(v, q)
    # a tie in the low digit carries up: B/2 -> (-B/2, +1)
    got = [int(x) for x in O.signed_gadget_decompose(B // 2, b, q, nd)]
    assert got[0] == q - B // 2 and got[1] == 1


@pytest.mark.parametrize("which", ["data", "key"])
def test_n2048_k1_ext_prod_rows_exact(O, p1, which):
    """decomp_rlwe_single_mod (gsw.cpp:252-295): row poly*l + (l-1-i) holds
    balanced digit i of the centred coefficient."""
    l, b = (p1.l_ep, p1.base_log2) if which == "data" else (p1.l_key, p1.base_log2_key)
    q, N = p1.rns_mods[0], p1.N
    rng = np.random.default_rng(12)
    ct = rng.integers(0, q, size=(2, N), dtype=np.uint64)
    ct[0, :4] = [0, q - 1, q // 2, q // 2 + 1]
    ct[1, :2] = [1 << (b - 1), q - (1 << (b - 1))]
    rows = O.decompose_rows(p1, ct, l, b)
    for poly in range(2):
        for k in [0, 1, 2, 3, 999, N - 1]:
            got = [int(rows[poly * l + (l - 1 - i)][k]) for i in range(l)]
            signed = [x if x < q // 2 else x - q for x in got]
            assert signed == _balanced_reference(int(ct[poly, k]), q, b, l)


def test_n2048_k1_rescale_exact(O, p1):
    """utils.cpp:445-462 with in = q: centred round(v * q' / q) into [0, q')."""
    q, sq = p1.rns_mods[0], p1.small_q
    rng = np.random.default_rng(6)
    vals = [0, 1, q // 2 - 1, q // 2, q // 2 + 1, q - 1] + [int(x) for x in rng.integers(0, q, 300, dtype=np.uint64)]
    for a in vals:
        v = a if a < q // 2 else a - q
        r = Fraction(v * sq, q)
        nearest = (abs(r) + Fraction(1, 2)).__floor__() * (1 if v >= 0 else -1)
        assert O.rescale(a, q, sq) == nearest % sq


def test_n2048_k1_external_product_selects(O, p1):
    N, t = p1.N, p1.plain_mod
    msg = np.zeros(N, np.uint64); msg[:3] = [t // 2 + 1, 1234, t - 1]
    sk, ct = O.gen_sk_and_encrypt(p1, msg, 13)
    for m, expect in [(1, msg), (0, np.zeros(N, np.uint64))]:
        pt = np.zeros(N, np.uint64); pt[0] = m
        g = O.plain_to_gsw(p1, pt, sk, p1.l_ep, p1.base_log2, 14 + m)
        out = O.external_product(p1, g, ct, p1.l_ep, p1.base_log2)
        dec, budget = O.decrypt_and_budget(p1, sk, out, True)
        assert np.array_equal(dec, expect) and budget > 0


def test_n2048_k1_bv_keyswitch_is_automorphism(O, p1):
    N, t = p1.N, p1.plain_mod
    msg = np.zeros(N, np.uint64)
    msg[[0, 1, 2, 10, 2000]] = [1, 2, 3, t - 5, 77]
    sk, ct = O.gen_sk_and_encrypt(p1, msg, 23)
    for k in (N + 1, N // 8 + 1):
        key = O.gen_bv_ks_key(p1, sk, k, 24)
        out = O.bv_apply(p1, ct, k, key)
        assert np.array_equal(O.decrypt_rns(p1, sk, out, False), O.automorphism_coeff(msg, k, t))


def test_n2048_k1_ext_prod_mux_selects(O, p1):
    p = p1.with_layout(43, 5)
    N, t = p.N, p.plain_mod
    c = O.PirClient(p, 1, 61)
    sk = c.secret_key()
    rng = np.random.default_rng(62)
    a = rng.integers(0, t, size=N, dtype=np.uint64)
    b = rng.integers(0, t, size=N, dtype=np.uint64)
    ca, cb = c.encrypt_bfv(a, 63), c.encrypt_bfv(b, 64)
    srv = O.PirServer(p)
    for bit, want in [(1, b), (0, a)]:
        g = O.plain_to_gsw(p, np.eye(1, N, 0, np.uint64)[0] * np.uint64(bit), sk, p.l_ep, p.base_log2, 65 + bit)
        out = srv.ext_prod_mux(ca, cb, g)
        assert np.array_equal(O.decrypt_rns(p, sk, out, False), want)


def test_n2048_k1_mod_switch_round_trip(O, p1):
    p = p1.with_layout(43, 5)
    c = O.PirClient(p, 1, 33)
    msg = np.zeros(p.N, np.uint64)
    msg[:10] = (np.arange(10, dtype=np.uint64) * np.uint64(7919)) % np.uint64(p.plain_mod)
    ct = c.encrypt_bfv(msg, 34)
    ms = O.PirServer(p).mod_switch(ct)
    assert ms.shape == (2, p.N) and int(ms.max()) < p.small_q
    pt, budget = c.decrypt_mod_q(ms)
    assert np.array_equal(pt, msg) and budget > 0


@pytest.mark.parametrize("layout", [(43, 5, True), (700, 8, True)])
def test_n2048_k1_end_to_end_decrypts(O, p1, layout):
    p = p1.with_layout(*layout)
    c = O.PirClient(p, 9, 4321 + layout[0])
    srv = O.PirServer(p)
    idx = p.target_num_pt - 1
    srv.load_data(p.target_num_pt, "random", 19, [idx, 0])
    srv.set_client_keys(9, c.create_bv_galois_keys(), c.generate_gsw_from_key())
    for i in [idx, 0]:
        resp = srv.make_query(9, c.fast_generate_query(p, i))
        wire = srv.save_resp_to_stream(resp)
        assert len(wire) == (2 * p.N * 22 + 7) // 8
        pt, budget = c.decrypt_mod_q(c.load_resp_from_stream(wire))
        assert np.array_equal(pt, srv.direct_get_original_plaintext(i)) and budget >= 1
