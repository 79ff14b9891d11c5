"""Pins for the oracle's two-prime path (CONFIG_N4096_K2_MP), CPU only.

The reference's own tests are compiled for one config at a time and hold no
fixtures for this one, so each piece is pinned by what the mathematics fixes:
closed-form NTT evaluation, CRT recomposition of the multi-precision gadget
digits, RGSW selection, BV keyswitch = automorphism, the mod-switch round trip
and end-to-end decryption (tests/test_pir.cpp's contract).

Parity unpinned, as for K = 1: the NTT evaluation order (HEXL's bit-reversed
convention) and the choice of minimal primitive root; both only matter for
NTT-form data crossing the API.
"""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def p2(O):
    return O.PirParams("CONFIG_N4096_K2_MP")


def brv(x, bits):
    return int(format(x, f"0{bits}b")[::-1], 2)


def crt(r0, r1, q0, q1):
    return (r0 + q0 * (((r1 - r0) * pow(q0, -1, q1)) % q1)) % (q0 * q1)


def test_k2_params(O, p2):
    """database_constants.h:111-124; pir.cpp:14-98."""
    q0, q1 = p2.rns_mods
    assert p2.N == 4096 and q0 != q1
    for q in (q0, q1):
        assert O.is_prime(q) and q.bit_length() == 60 and (q - 1) % (2 * p2.N) == 0
    assert (p2.l_ep, p2.l_key, p2.l_ks) == (5, 8, 8)
    assert p2.small_q.bit_length() == 50 and O.is_prime(p2.small_q)
    assert p2.plain_mod.bit_length() == 40
    assert p2.ct_mod_width == 120
    assert p2.l_ep * p2.base_log2 >= 120 and p2.l_key * p2.base_log2_key >= 120
    assert p2.q0_inv_mod_q1 == pow(q0, -1, q1)


def test_k2_ntt_brute_force(O, p2):
    N = p2.N
    rng = np.random.default_rng(1)
    for q in p2.rns_mods:
        psi = O.ntt_root(N, q)
        assert pow(psi, N, q) == q - 1
        a = rng.integers(0, q, size=N, dtype=np.uint64)
        got = O.ntt_fwd(a, N, q)
        ai = [int(x) for x in a]
        for j in [0, 1, 1000, N - 1]:
            e = (2 * brv(j, 12) + 1)
            w = pow(psi, e, q)
            acc, wp = 0, 1
            for x in ai:
                acc = (acc + x * wp) % q
                wp = wp * w % q
            assert int(got[j]) == acc
        assert np.array_equal(O.ntt_inv(got, N, q), a)


def test_k2_ntt_kat_product(O, p2):
    N = p2.N
    for q in p2.rns_mods:
        a = np.zeros(N, np.uint64); a[:3] = [1, 2, 3]
        b = np.zeros(N, np.uint64); b[:2] = [4, 5]
        fa, fb = O.ntt_fwd(a, N, q), O.ntt_fwd(b, N, q)
        prod = np.array([(int(x) * int(y)) % q for x, y in zip(fa, fb)], dtype=np.uint64)
        assert list(O.ntt_inv(prod, N, q)[:5]) == [4, 13, 22, 15, 0]


@pytest.mark.parametrize("which", ["data", "key"])
def test_k2_mp_digits_recompose(O, p2, which):
    """gsw.cpp:22-85,184-250: unsigned base-2^b digits of the CRT value, row
    poly*l + (l-1-i) holds digit i, the same small integer in both limbs."""
    l, b = (p2.l_ep, p2.base_log2) if which == "data" else (p2.l_key, p2.base_log2_key)
    q0, q1 = p2.rns_mods
    N = p2.N
    rng = np.random.default_rng(9)
    ct = np.empty((2, 2 * N), np.uint64)
    for m, q in enumerate((q0, q1)):
        ct[:, m * N:(m + 1) * N] = rng.integers(0, q, size=(2, N), dtype=np.uint64)
    ct[0, 0], ct[0, N] = q0 - 1, q1 - 1            # X = Q - 1
    rows = O.decompose_rows(p2, ct, l, b)
    for poly in range(2):
        for k in [0, 1, 2049, N - 1]:
            X = crt(int(ct[poly, k]), int(ct[poly, N + k]), q0, q1)
            digits = [int(rows[poly * l + (l - 1 - i)][k]) for i in range(l)]
            assert all(d < (1 << b) for d in digits)
            assert [int(rows[poly * l + (l - 1 - i)][N + k]) for i in range(l)] == digits
            assert sum(d << (b * i) for i, d in enumerate(digits)) == X


def test_k2_external_product_selects(O, p2):
    N, t = p2.N, p2.plain_mod
    msg = np.zeros(N, np.uint64); msg[:3] = [t // 2 + 1, 12345, t - 1]
    sk, ct = O.gen_sk_and_encrypt(p2, msg, 11)
    for m, expect in [(1, msg), (0, np.zeros(N, np.uint64))]:
        pt = np.zeros(N, np.uint64); pt[0] = m
        g = O.plain_to_gsw(p2, pt, sk, p2.l_ep, p2.base_log2, 12 + m)
        out = O.external_product(p2, g, ct, p2.l_ep, p2.base_log2)  # NTT form
        # decrypt_rns is the RNS decryption (rlwe.cpp); decrypt_and_budget is single-modulus only
        assert np.array_equal(O.decrypt_rns(p2, sk, out, True), expect)


def test_k2_bv_keyswitch_is_automorphism(O, p2):
    """bv_keyswitch.cpp:373-481 (signed MP digits, tie (-B/2, B/2])."""
    N, t = p2.N, p2.plain_mod
    msg = np.zeros(N, np.uint64)
    msg[[0, 1, 2, 10, 4000]] = [1, 2, 3, t - 5, 77]
    sk, ct = O.gen_sk_and_encrypt(p2, msg, 21)
    for k in (N + 1, N // 4 + 1):
        key = O.gen_bv_ks_key(p2, sk, k, 22)
        out = O.bv_apply(p2, ct, k, key)
        assert np.array_equal(O.decrypt_rns(p2, sk, out, False), O.automorphism_coeff(msg, k, t))


def test_k2_ext_prod_mux_selects(O, p2):
    """tests/test_ext_prod_mux.cpp with two RNS limbs: RGSW(1) selects y,
    RGSW(0) keeps x."""
    p = p2.with_layout(43, 5)
    N, t = p.N, p.plain_mod
    c = O.PirClient(p, 1, 51)
    sk = c.secret_key()
    rng = np.random.default_rng(52)
    a = rng.integers(0, t, size=N, dtype=np.uint64)
    b = rng.integers(0, t, size=N, dtype=np.uint64)
    a[:3] = [5, 7, t - 1]
    b[:3] = [3, 9, 1]
    ca, cb = c.encrypt_bfv(a, 53), c.encrypt_bfv(b, 54)
    srv = O.PirServer(p)
    for bit, want in [(1, b), (0, a)]:
        g = O.plain_to_gsw(p, np.eye(1, N, 0, np.uint64)[0] * np.uint64(bit), sk, p.l_ep, p.base_log2, 55 + bit)
        out = srv.ext_prod_mux(ca, cb, g)
        assert np.array_equal(O.decrypt_rns(p, sk, out, False), want)


def test_k2_mod_switch_round_trip(O, p2):
    p = p2.with_layout(43, 5)
    c = O.PirClient(p, 1, 31)
    msg = np.zeros(p.N, np.uint64)
    msg[:10] = (np.arange(10, dtype=np.uint64) * np.uint64(987654321)) % np.uint64(p.plain_mod)
    ct = c.encrypt_bfv(msg, 32)
    ms = O.PirServer(p).mod_switch(ct)
    assert ms.shape == (2, p.N) and int(ms.max()) < p.small_q
    pt, budget = c.decrypt_mod_q(ms)
    assert np.array_equal(pt, msg) and budget > 0


def test_k2_level_mat_mat_u64(O, p2):
    """matrix.cpp:24-151 with 60-bit moduli: exact u128 sums, seeded inputs."""
    q = p2.rns_mods[0]
    m, n, lv = 3, 512, 2
    r = O.mt19937_64_draws(0xC0FFEE, m * n * lv + n * 2 * lv)
    A = r[: m * n * lv] % np.uint64(q)
    B = r[m * n * lv:] % np.uint64(q)
    got = O.level_mat_mat_u64(A, B, m, n, lv, q)
    for level in range(lv):
        for i in range(m):
            for c in range(2):
                want = sum(int(A[level * m * n + i * n + k]) * int(B[level * n * 2 + 2 * k + c]) for k in range(n)) % q
                assert int(got[level * m * 2 + 2 * i + c]) == want


@pytest.mark.parametrize("layout", [(43, 5, True), (700, 8, True)])
def test_k2_end_to_end_decrypts(O, p2, layout):
    p = p2.with_layout(*layout)
    c = O.PirClient(p, 9, 5678 + layout[0])
    srv = O.PirServer(p)
    idx = p.target_num_pt - 1
    srv.load_data(p.target_num_pt, "random", 17, [idx, 0])
    srv.set_client_keys(9, c.create_bv_galois_keys(), c.generate_gsw_from_key())
    for i in [idx, 0]:
        resp = srv.make_query(9, c.fast_generate_query(p, i))
        wire = srv.save_resp_to_stream(resp)
        assert len(wire) == (2 * p.N * 50 + 7) // 8
        pt, budget = c.decrypt_mod_q(c.load_resp_from_stream(wire))
        assert np.array_equal(pt, srv.direct_get_original_plaintext(i)) and budget >= 1


def _is_prime(n):
    """Deterministic Miller-Rabin for n < 2^64 (first 12 prime bases)."""
    if n < 2:
        return False
    bases = [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37]
    for b in bases:
        if n % b == 0:
            return n == b
    d, r = n - 1, 0
    while d % 2 == 0:
        d, r = d // 2, r + 1
    for a in bases:
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(r - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def _largest_prime_1_mod(bits, m, skip=()):
    x = ((1 << bits) - 1) // m * m + 1
    while x not in skip and not _is_prime(x) or x in skip:
        x -= m
    return x


def _smallest_prime_above(v):
    x = v + 1
    while not _is_prime(x):
        x += 1
    return x


def test_k2_constants_exact(p2):
    """The CONFIG_N4096_K2_MP values re-derived from the reference's prime and
    plaintext-modulus rules (database_constants.h:111-124, pir.cpp:14-98,
    utils.cpp:136-287): the two largest 60-bit primes 1 mod 2N, t the smallest
    prime above 2^39, small_q the largest 50-bit prime 1 mod 2N."""
    q0, q1 = p2.rns_mods
    assert (q0, q1) == (1152921504606830593, 1152921504606748673)
    assert q0 == _largest_prime_1_mod(60, 2 * p2.N) and q1 == _largest_prime_1_mod(60, 2 * p2.N, skip=(q0,))
    assert p2.plain_mod == 549755813911 == _smallest_prime_above(1 << 39)
    assert p2.small_q == 1125899906826241 == _largest_prime_1_mod(50, 2 * p2.N)
    assert p2.bv_base_log2 == 16 and p2.l_ks == 8


def _centred(x, Q):
    return x if x <= Q // 2 else x - Q


@pytest.mark.parametrize("b", [16, 8])
def test_k2_signed_mp_digits_balanced_and_exact(O, p2, b):
    """bv_keyswitch.cpp:42-64: the balanced digits of the centred CRT value lie
    in (-B/2, B/2] -- a tie goes to +B/2 -- and recompose it exactly over the
    integers. With enough digits this representation is unique, so range +
    recomposition pin every digit. Ties are forced at every digit position and
    at the centring boundary +-Q/2."""
    q0, q1 = p2.rns_mods
    Q = q0 * q1
    B, nd = 1 << b, -(-Q.bit_length() // b) + 1
    rng = np.random.default_rng(77)
    vals = [0, 1, Q - 1, Q // 2, Q // 2 + 1, Q // 2 - 1, B // 2, Q - B // 2, B // 2 - 1, Q - B // 2 - 1]
    for pos in range(nd - 1):
        for k in (0, 1, 5):
            vals.append((k * B + B // 2) * B ** pos)                    # digit pos is a tie
            vals.append(Q - (k * B + B // 2) * B ** pos)                # negative tie
            vals.append(((k * B + B // 2) * B + B // 2) * B ** pos % Q)  # two ties in a row
    vals += [int(x) % Q for x in rng.integers(0, 1 << 62, size=40, dtype=np.uint64) * (1 << 58)]
    vals += [int(rng.integers(0, 1 << 60)) * (1 << 60) % Q + int(rng.integers(0, 1 << 60)) for _ in range(40)]
    ties = 0
    for x in vals:
        x %= Q
        d = O.signed_gadget_decompose_mp(x & ((1 << 64) - 1), x >> 64, Q & ((1 << 64) - 1), Q >> 64, b, nd)
        assert all(-B // 2 < di <= B // 2 for di in d), (x, d)
        assert sum(di * B ** i for i, di in enumerate(d)) == _centred(x, Q), x
        ties += sum(di == B // 2 for di in d)
    assert ties >= 3 * (nd - 1)


def test_k2_drop_q1_round_half_up_exact(O, p2):
    """server.cpp:1053-1061: x = CRT(r0, r1) in [0, Q), out = round(x / q1)
    (nearest, half up) reduced into [0, q0). Checked with exact rationals at
    the rounding boundary (x mod q1 around q1/2) and at the top of [0, Q),
    where round(x / q1) = q0 wraps to 0."""
    from fractions import Fraction
    import math
    q0, q1 = p2.rns_mods
    Q = q0 * q1
    inv = pow(q0, -1, q1)
    assert p2.q0_inv_mod_q1 == inv
    xs = [0, 1, q1 - 1, q1, Q - 1, Q - q1 // 2, Q - q1 // 2 - 1, Q - q1 // 2 - 2, Q - q1]
    rng = np.random.default_rng(5)
    for k in [0, 1, 12345, q0 - 2, q0 - 1] + [int(v) for v in rng.integers(0, q0, size=10, dtype=np.uint64)]:
        for r in (q1 // 2 - 1, q1 // 2, q1 // 2 + 1, 0, q1 - 1):
            xs.append(k * q1 + r)
    wraps = 0
    for x in xs:
        assert 0 <= x < Q
        want = math.floor(Fraction(x, q1) + Fraction(1, 2)) % q0
        wraps += math.floor(Fraction(x, q1) + Fraction(1, 2)) == q0
        assert O.drop_q1(x % q0, x % q1, q0, q1, inv) == want, x
    assert wraps >= 2


def test_k2_mod_switch_is_drop_then_rescale(O, p2):
    """mod_switch_inplace's two-prime branch (server.cpp:1044-1075) on crafted
    coefficients: each output is rescale(drop_q1(...)), both pinned above and
    in the single-modulus pins."""
    p = p2.with_layout(1, 0)
    q0, q1 = p2.rns_mods
    N, sq = p.N, p.small_q
    rng = np.random.default_rng(6)
    ct = np.empty((2, 2 * N), np.uint64)
    ct[:, :N] = rng.integers(0, q0, size=(2, N), dtype=np.uint64)
    ct[:, N:] = rng.integers(0, q1, size=(2, N), dtype=np.uint64)
    Q = q0 * q1
    special = [Q - 1, Q // 2, Q // 2 + 1, 5 * q1 + q1 // 2, 5 * q1 + q1 // 2 + 1, 0]
    for i, x in enumerate(special):
        ct[0, i], ct[0, N + i] = x % q0, x % q1
    out = O.PirServer(p).mod_switch(ct)
    inv = pow(q0, -1, q1)
    for poly in range(2):
        for i in list(range(len(special))) + [100, N - 1]:
            d = O.drop_q1(int(ct[poly, i]), int(ct[poly, N + i]), q0, q1, inv)
            assert int(out[poly, i]) == O.rescale(d, q0, sq)


# ---------------------------------------------------------------------------
# CONFIG_N2048_K2_MP (database_constants.h:96-109): the same two-prime
# algorithms at N = 2048 with two 29-bit primes.
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def p2s(O):
    return O.PirParams("CONFIG_N2048_K2_MP")


def test_k2_n2048_constants_exact(O, p2s):
    """Re-derived from the reference's rules: the two largest 29-bit primes
    1 mod 2N (the composite scheme's q1, q2), t the smallest prime above
    2^9, small_q the largest 22-bit prime 1 mod 2N; gadget widths
    ceil(58 / l)."""
    assert p2s.N == 2048 and p2s.rns_mods == [536813569, 536752129]
    q0 = _largest_prime_1_mod(29, 2 * p2s.N)
    assert p2s.rns_mods == [q0, _largest_prime_1_mod(29, 2 * p2s.N, skip=(q0,))]
    assert p2s.plain_mod == 521 == _smallest_prime_above(1 << 9)
    assert p2s.small_q == 4169729 == _largest_prime_1_mod(22, 2 * p2s.N)
    assert (p2s.l_ep, p2s.base_log2, p2s.l_key, p2s.base_log2_key, p2s.l_ks, p2s.bv_base_log2) == (5, 12, 8, 8, 8, 8)


def test_k2_n2048_ntt_brute_force(O, p2s):
    N = p2s.N
    rng = np.random.default_rng(2)
    for q in p2s.rns_mods:
        psi = O.ntt_root(N, q)
        assert pow(psi, N, q) == q - 1
        a = rng.integers(0, q, size=N, dtype=np.uint64)
        got = O.ntt_fwd(a, N, q)
        ai = [int(x) for x in a]
        for j in [0, 1, 777, N - 1]:
            w = pow(psi, 2 * brv(j, 11) + 1, q)
            assert int(got[j]) == sum(x * pow(w, i, q) for i, x in enumerate(ai)) % q


def test_k2_n2048_digits_and_drop_exact(O, p2s):
    """The MP gadget digits and drop_q1 with this scheme's Q = q0 * q1 < 2^58."""
    from fractions import Fraction
    import math
    q0, q1 = p2s.rns_mods
    Q = q0 * q1
    b = p2s.bv_base_log2
    B, nd = 1 << b, p2s.l_ks
    for x in [0, 1, Q - 1, Q // 2, Q // 2 + 1, B // 2, Q - B // 2, (3 * B + B // 2) * B, Q - (B // 2) * B ** 3,
              123456789012345]:
        d = O.signed_gadget_decompose_mp(x, 0, Q, 0, b, nd)
        assert all(-B // 2 < di <= B // 2 for di in d)
        assert sum(di * B ** i for i, di in enumerate(d)) == _centred(x, Q)
    inv = pow(q0, -1, q1)
    assert p2s.q0_inv_mod_q1 == inv
    for x in [0, Q - 1, Q - q1 // 2, Q - q1 // 2 - 1, 77 * q1 + q1 // 2, 77 * q1 + q1 // 2 + 1]:
        want = math.floor(Fraction(x, q1) + Fraction(1, 2)) % q0
        assert O.drop_q1(x % q0, x % q1, q0, q1, inv) == want


@pytest.mark.parametrize("layout", [(43, 5, True), (700, 8, True)])
def test_k2_n2048_end_to_end_decrypts(O, p2s, layout):
    p = p2s.with_layout(*layout)
    c = O.PirClient(p, 9, 99 + layout[0])
    srv = O.PirServer(p)
    idx = p.target_num_pt - 1
    srv.load_data(p.target_num_pt, "random", 18, [idx, 0])
    srv.set_client_keys(9, c.create_bv_galois_keys(), c.generate_gsw_from_key())
    for i in [idx, 0]:
        resp = srv.make_query(9, c.fast_generate_query(p, i))
        wire = srv.save_resp_to_stream(resp)
        assert len(wire) == (2 * p.N * 22 + 7) // 8
        pt, budget = c.decrypt_mod_q(c.load_resp_from_stream(wire))
        assert np.array_equal(pt, srv.direct_get_original_plaintext(i)) and budget >= 1
