"""Pins for the CPU oracle (CPU only).

Each oracle function is checked against the reference's own known-answer
tests (src/tests/*.cpp) or against what the mathematics fixes, chosen so that
a plausible slip (dropped term, wrong sign or index, transposed operand)
fails one of them. The reference itself cannot be built here (Intel HEXL
1.2.6 is not vendored), so there is no compiled reference to compare with.

Unpinned by any reference test: the NTT evaluation ORDER. The oracle uses
HEXL's documented convention (root powers bit-reversed, output index j holds
a(w^(2*bitrev(j)+1))); the order only matters for NTT-form data crossing the
API (keys, database), never for coefficient-form responses.
"""
import numpy as np
import pytest

N = 2048


def brv(x, bits):
    return int(format(x, f"0{bits}b")[::-1], 2) if bits else 0


# ---------------------------------------------------------------------------
# Parameters (src/pir.cpp, tests/test_runtime_layout.cpp, tests/test_utils.cpp)
# ---------------------------------------------------------------------------
def test_runtime_layout_kats(O, comp_params):
    """tests/test_runtime_layout.cpp:3-40."""
    f = comp_params.with_layout(349526, 10, True)
    assert (f.target_num_pt, f.fst_dim_sz, f.num_dims, f.num_pt, f.expansion_height) == (349526, 512, 11, 349696, 10)
    assert comp_params.scheme_compatible(f)
    s = comp_params.with_layout(1, 0, True)
    assert (s.num_pt, s.num_dims, s.expansion_height) == (1, 1, 0)
    assert O.calculate_db_shape(43, 6, 5, False) == (26, 2)
    with pytest.raises(ValueError):
        comp_params.with_layout(0, 0, True)
    # the recorded physical size of the 1 GB row (outputs/merkle_baselines/*.json:62)
    assert f.num_pt * N * 8 == 5729419264


def test_utils_arith_kats(O):
    """tests/test_utils.cpp: Miller-Rabin and modular inverse KATs."""
    for p in [2, 3, 5, 7, 11, 13, 97, 1000003, 1152921504606748673, 1152921504606830593, 36028797018820609,
              (1 << 61) - 1]:
        assert O.is_prime(p)
    for c in [0, 1, 4, 6, 9, 15, 21, 25, 49, 1000000, 1152921504606748673 - 2, 1 << 61, (1 << 60) + 1, 3215031751]:
        assert not O.is_prime(c)
    for v, m in [(3, 7), (2, 1152921504606748673), (1234567, 1152921504606748673), (7, 36028797018820609),
                 (1 << 20, 1152921504606830593)]:
        inv = O.try_invert(v, m)
        assert inv is not None and v * inv % m == 1
    for v, m in [(0, 7), (6, 9), (10, 100)]:
        assert O.try_invert(v, m) is None


def is_prime_py(n):
    if n < 2:
        return False
    for p in range(2, int(n ** 0.5) + 1):
        if n % p == 0:
            return False
    return True


def test_composite_modulus_from_definitions(comp_params):
    """pir.cpp:14-63 restated from first principles: q1 > q2 are the two largest
    primes < 2^29 that are 1 mod 2N; w_i is the smallest primitive 2N-th root
    mod q_i; w_crt is their CRT; t is the smallest prime above 2^12; q' is the
    largest 22-bit prime that is 1 mod 2N."""
    p = comp_params
    step = 2 * N
    cands = [c for c in range((1 << 29) - 1, (1 << 29) - 200 * step, -1) if c % step == 1]
    primes = [c for c in cands if pow(3, c - 1, c) == 1 and is_prime_py(c)][:2]
    assert [p.q1, p.q2] == primes
    Q = p.q1 * p.q2
    assert p.rns_mods == [Q] and Q % (2 * N) == 1 and Q < 2**58
    for q, w in [(p.q1, p.w1), (p.q2, p.w2)]:
        assert pow(w, N, q) == q - 1          # primitive 2N-th root
        smaller = [x for x in range(2, w) if pow(x, N, q) == q - 1]
        assert smaller == []                   # and the minimal one
    assert p.w_crt % p.q1 == p.w1 and p.w_crt % p.q2 == p.w2 and p.w_crt < Q
    assert pow(p.w_crt, N, Q) == Q - 1
    assert p.plain_mod == 4099 and all(not is_prime_py(x) for x in range(4097, 4099))
    sq = p.small_q
    assert sq < 2**22 and sq % step == 1 and is_prime_py(sq)
    assert not any(is_prime_py(c) for c in range(sq + step, 2**22, step))
    assert (p.base_log2, p.base_log2_key, p.bv_base_log2) == (10, 6, 8)  # ceil(58/6), ceil(58/10), 58/8+1


# ---------------------------------------------------------------------------
# NTT (utils.cpp:84-90, tests/test_hexl_ntt.cpp)
# ---------------------------------------------------------------------------
def test_ntt_brute_force_evaluation(O, comp_params):
    """NTT(a)[j] = a(w^(2*bitrev(j)+1)) mod Q, checked at 24 indices by Horner."""
    Q, w = comp_params.rns_mods[0], comp_params.w_crt
    rng = np.random.default_rng(1)
    a = rng.integers(0, Q, size=N, dtype=np.uint64)
    A = O.ntt_fwd(a, N, Q)
    coeffs = [int(x) for x in a]
    for j in list(rng.integers(0, N, 22)) + [0, N - 1]:
        x = pow(w, 2 * brv(int(j), 11) + 1, Q)
        acc = 0
        for c in reversed(coeffs):
            acc = (acc * x + c) % Q
        assert int(A[j]) == acc
    assert np.array_equal(O.ntt_inv(A, N, Q), a)


def test_ntt_hexl_kats(O, comp_params):
    """tests/test_hexl_ntt.cpp:25-107: round trip, product, add, decrypt."""
    Q, t = comp_params.rns_mods[0], comp_params.plain_mod
    rng = np.random.default_rng(42)
    x = rng.integers(0, Q, size=N, dtype=np.uint64)
    assert np.array_equal(O.ntt_inv(O.ntt_fwd(x, N, Q), N, Q), x)
    a = np.zeros(N, np.uint64); a[:3] = [1, 2, 3]
    b = np.zeros(N, np.uint64); b[:2] = [4, 5]
    fa, fb = O.ntt_fwd(a, N, Q), O.ntt_fwd(b, N, Q)
    c = O.ntt_inv(np.array([int(u) * int(v) % Q for u, v in zip(fa, fb)], np.uint64), N, Q)
    assert list(c[:5]) == [4, 13, 22, 15, 0]
    # simulated BFV decryption with s = 1 - x + x^2, m = 7
    s = np.zeros(N, np.uint64); s[0], s[1], s[2] = 1, Q - 1, 1
    ap = rng.integers(0, Q, size=N, dtype=np.uint64)
    fs, fa = O.ntt_fwd(s, N, Q), O.ntt_fwd(ap, N, Q)
    as_ = O.ntt_inv(np.array([int(u) * int(v) % Q for u, v in zip(fa, fs)], np.uint64), N, Q)
    ct0 = np.array([(Q - int(v)) % Q for v in as_], np.uint64)
    ct0[0] = (int(ct0[0]) + (Q // t) * 7) % Q
    f1 = O.ntt_fwd(ap, N, Q)
    ph = O.ntt_inv(np.array([int(u) * int(v) % Q for u, v in zip(f1, fs)], np.uint64), N, Q)
    p0 = (int(ph[0]) + int(ct0[0])) % Q
    assert ((p0 * t + Q // 2) // Q) % t == 7


def negacyclic_mul(a, b, q):
    """Schoolbook product in Z_q[x]/(x^N+1) of sparse {index: value} polys."""
    out = {}
    for i, x in a.items():
        for j, y in b.items():
            k, v = i + j, x * y
            if k >= N:
                k, v = k - N, -v
            out[k] = (out.get(k, 0) + v) % q
    return out


def test_ntt_negacyclic_convolution(O, comp_params):
    Q = comp_params.rns_mods[0]
    rng = np.random.default_rng(3)
    da = {int(i): int(rng.integers(0, Q)) for i in rng.choice(N, 40, replace=False)}
    db = {int(i): int(rng.integers(0, Q)) for i in rng.choice(N, 40, replace=False)}
    da[N - 1] = 5; db[N - 3] = 7  # force wrap-around terms
    a = np.zeros(N, np.uint64); b = np.zeros(N, np.uint64)
    for k, v in da.items(): a[k] = v
    for k, v in db.items(): b[k] = v
    fa, fb = O.ntt_fwd(a, N, Q), O.ntt_fwd(b, N, Q)
    c = O.ntt_inv(np.array([int(u) * int(v) % Q for u, v in zip(fa, fb)], np.uint64), N, Q)
    want = np.zeros(N, dtype=object)
    for k, v in negacyclic_mul(da, db, Q).items():
        want[k] = v
    assert [int(x) for x in c] == list(want)


# ---------------------------------------------------------------------------
# Automorphism, shift, gadget, rescale (utils.cpp:58-72,166-186,206-222,445-462)
# ---------------------------------------------------------------------------
def test_automorphism_and_shift(O, comp_params):
    Q = comp_params.rns_mods[0]
    a = np.zeros(N, np.uint64); a[1] = 5; a[3] = 9; a[0] = 2
    for k in [3, 257, 2049]:
        s = O.automorphism_coeff(a, k, Q)
        want = np.zeros(N, dtype=object)
        for i in [0, 1, 3]:
            e = (i * k) % (2 * N)
            want[e % N] = (want[e % N] + (int(a[i]) if e < N else Q - int(a[i]))) % Q
        assert [int(x) for x in s] == list(want)
    rng = np.random.default_rng(4)
    b = rng.integers(0, Q, size=N, dtype=np.uint64)
    for k in [1, 4, 64]:
        got = O.negacyclic_shift(b, -k, Q)  # times X^-k (server.cpp:808)
        xk = np.zeros(N, np.uint64); xk[N - k] = Q - 1  # X^-k = -X^(N-k)
        fb, fx = O.ntt_fwd(b, N, Q), O.ntt_fwd(xk, N, Q)
        want = O.ntt_inv(np.array([int(u) * int(v) % Q for u, v in zip(fb, fx)], np.uint64), N, Q)
        assert np.array_equal(got, want)


def test_gsw_gadget_msb_first(O, comp_params):
    Q = comp_params.rns_mods[0]
    g = O.gsw_gadget(6, 10, [Q])
    assert g[0] == [pow(2, 10 * (5 - p), Q) for p in range(6)]


def test_rescale_exact(O, comp_params):
    """Centred round(a * out / in) reduced into [0, out) vs exact rationals."""
    from fractions import Fraction
    Q, sq = comp_params.rns_mods[0], comp_params.small_q
    rng = np.random.default_rng(5)
    vals = [0, 1, Q // 2 - 1, Q // 2, Q // 2 + 1, Q - 1] + [int(x) for x in rng.integers(0, Q, 300, dtype=np.uint64)]
    for a in vals:
        v = a if a < Q // 2 else a - Q          # centre (utils.cpp:451-452)
        r = Fraction(v * sq, Q)                  # never a half-integer: 2*v*sq is even, Q is odd
        nearest = (abs(r) + Fraction(1, 2)).__floor__() * (1 if v >= 0 else -1)
        assert O.rescale(a, Q, sq) == nearest % sq


# ---------------------------------------------------------------------------
# Barrett (tests/test_barrett.cpp)
# ---------------------------------------------------------------------------
def test_barrett_grid(O):
    rng = np.random.default_rng(0xBA77E77F)
    qs64 = [1, 2, 3, 17, O.generate_prime(29), O.generate_prime(32), O.generate_prime(40), O.generate_prime(50),
            O.generate_prime(60), (1 << 63) - 25, (1 << 63) + 29, (1 << 64) - 1]
    for q in qs64:
        for x in [0, 1, q - 1, q, (q + 1) % (1 << 64), (1 << 64) - 1] + [int(v) for v in rng.integers(0, 2**63, 300, dtype=np.uint64)]:
            assert O.barrett_u64(x, q) == x % q
    for qb in [8, 17, 32, 40, 50, 60, 61]:
        q = O.generate_prime(qb)
        for xb in [16, 32, 64, 96, 120, 127, 128]:
            for _ in range(60):
                x = int(rng.integers(0, 2**63)) << 65 | int(rng.integers(0, 2**63)) << 2 | int(rng.integers(0, 4))
                x &= (1 << xb) - 1
                assert O.barrett_u128(x & (2**64 - 1), x >> 64, q) == x % q
    q60 = O.generate_prime(60)
    for x in [0, 1, q60 - 1, q60, q60 + 1, q60 * q60, 2**128 - 1, 2**128 - 2, 2**127, 2**127 + 1]:
        assert O.barrett_u128(x & (2**64 - 1), x >> 64, q60) == x % q60


# ---------------------------------------------------------------------------
# First-dimension mat-mat (tests/test_fst_dim.cpp:8-93, seeded mt19937_64)
# ---------------------------------------------------------------------------
def test_fst_dim_seeded_kats(O, comp_params):
    q1 = comp_params.q1
    n = 512  # default layout's fst_dim_sz
    r = O.mt19937_64_draws(0x51A1A5, 3 * n)
    a = (r[:n] % np.uint64(q1)).astype(np.uint32)
    b = (r[n:] % np.uint64(q1)).astype(np.uint32)
    got = O.level_mat_mat_32(a, b, 1, n, 1, q1)
    ref0 = sum(int(x) * int(y) for x, y in zip(a, b[0::2])) % q1
    ref1 = sum(int(x) * int(y) for x, y in zip(a, b[1::2])) % q1
    assert list(got) == [ref0, ref1]
    m, n, lv = 4, 1024, 3
    Q = comp_params.rns_mods[0]
    r = O.mt19937_64_draws(0xC0FFEE, m * n * lv + n * 2 * lv)
    A = r[: m * n * lv] % np.uint64(Q)
    B = r[m * n * lv:] % np.uint64(Q)
    got = O.level_mat_mat_u64(A, B, m, n, lv, Q)
    for level in range(lv):
        for i in range(m):
            t0 = t1 = 0
            for k in range(n):
                av = int(A[level * m * n + i * n + k])
                t0 = (t0 + av * int(B[level * n * 2 + 2 * k])) % Q
                t1 = (t1 + av * int(B[level * n * 2 + 2 * k + 1])) % Q
            assert int(got[level * m * 2 + 2 * i]) == t0 and int(got[level * m * 2 + 2 * i + 1]) == t1


# ---------------------------------------------------------------------------
# Gadget decompositions (gsw.cpp:252-295, bv_keyswitch.cpp:20-40,
# tests/test_bv_keyswitch.cpp:10-57)
# ---------------------------------------------------------------------------
def test_signed_decompose_reconstructs(O, comp_params):
    Q = comp_params.rns_mods[0]
    b = (58 + 8 - 1) // 8
    r = O.mt19937_64_draws(42, 2000) % np.uint64(Q)
    for v in list(r) + [0, 1, Q // 2, Q // 2 + 1, Q - 1]:
        d = O.signed_gadget_decompose(int(v), b, Q, 8)
        assert sum(x * (1 << (b * i)) for i, x in enumerate(d)) % Q == int(v)
        mags = [x if x <= Q // 2 else Q - x for x in d]
        assert max(mags) <= (1 << b) // 2
        assert all(x < (1 << (b - 1)) or Q - x <= (1 << (b - 1)) for x in d)  # [-B/2, B/2)


@pytest.mark.parametrize("which", ["data", "key"])
def test_ext_prod_rows_reconstruct(O, comp_params, which):
    p = comp_params
    l, b = (p.l_ep, p.base_log2) if which == "data" else (p.l_key, p.base_log2_key)
    Q = p.rns_mods[0]
    rng = np.random.default_rng(8)
    ct = rng.integers(0, Q, size=(2, N), dtype=np.uint64)
    rows = O.decompose_rows(p, ct, l, b)
    for poly in range(2):
        for k in [0, 1, 777, N - 1]:
            v = sum(int(rows[poly * l + (l - 1 - i)][k]) * pow(2, b * i, Q) for i in range(l)) % Q
            assert v == int(ct[poly, k])


# ---------------------------------------------------------------------------
# Scheme invariants (tests/test_ext_prod.cpp, test_bv_keyswitch.cpp,
# test_mod_switch.cpp)
# ---------------------------------------------------------------------------
def test_external_product_selects(O, comp_params):
    p = comp_params
    t = p.plain_mod
    msg = np.zeros(N, np.uint64); msg[:3] = [t // 2 + 1, t // 2 + 2, t // 2 + 3]
    sk, ct = O.gen_sk_and_encrypt(p, msg, 11)
    for m, expect in [(1, msg), (0, np.zeros(N, np.uint64))]:
        pt = np.zeros(N, np.uint64); pt[0] = m
        g = O.plain_to_gsw(p, pt, sk, p.l_ep, p.base_log2, 12 + m)
        out = O.external_product(p, g, ct, p.l_ep, p.base_log2)
        dec, budget = O.decrypt_and_budget(p, sk, out, True)
        assert np.array_equal(dec, expect) and budget > 0


def test_ext_prod_mux_selects(O, comp_params):
    """tests/test_ext_prod_mux.cpp: mux(RGSW(1), a, b) decrypts to b (and,
    completing the identity x + b*(y - x), RGSW(0) keeps a)."""
    p = comp_params.with_layout(43, 5)
    t = p.plain_mod
    c = O.PirClient(p, 1, 41)
    sk = c.secret_key()
    rng = np.random.default_rng(42)
    a = rng.integers(0, t, size=N, dtype=np.uint64)
    b = rng.integers(0, t, size=N, dtype=np.uint64)
    a[:3] = [5, 7, t - 1]
    b[:3] = [3, 9, 1]
    ca, cb = c.encrypt_bfv(a, 43), c.encrypt_bfv(b, 44)
    srv = O.PirServer(p)
    for bit, want in [(1, b), (0, a)]:
        g = O.plain_to_gsw(p, np.eye(1, N, 0, np.uint64)[0] * np.uint64(bit), sk, p.l_ep, p.base_log2, 45 + bit)
        out = srv.ext_prod_mux(ca, cb, g)
        dec, budget = O.decrypt_and_budget(p, sk, out, False)
        assert np.array_equal(dec, want) and budget > 0


def test_bv_keyswitch_is_automorphism(O, comp_params):
    p = comp_params
    t = p.plain_mod
    msg = np.zeros(N, np.uint64)
    msg[[0, 1, 2, 10, 11, 12]] = [1, 2, 3, 10, 11, 12]
    sk, ct = O.gen_sk_and_encrypt(p, msg, 21)
    key = O.gen_bv_ks_key(p, sk, 257, 22)
    out = O.bv_apply(p, ct, 257, key)
    want = O.automorphism_coeff(msg, 257, t)
    assert np.array_equal(O.decrypt_rns(p, sk, out, False), want)


def test_mod_switch_round_trip(O, comp_params):
    """tests/test_mod_switch.cpp: encrypt -> mod_switch_inplace -> decrypt_mod_q."""
    p = comp_params.with_layout(43, 5)
    c = O.PirClient(p, 1, 31)
    msg = np.zeros(N, np.uint64)
    msg[:10] = np.arange(10) * 400 % p.plain_mod
    ct = c.encrypt_bfv(msg, 32)
    assert np.array_equal(O.decrypt_rns(p, c.secret_key(), ct, False), msg)
    ms = O.PirServer(p).mod_switch(ct)
    assert int(ms.max()) < p.small_q
    pt, budget = c.decrypt_mod_q(ms)
    assert np.array_equal(pt, msg) and budget > 0


# ---------------------------------------------------------------------------
# Expansion per-slot oracle, first dimension, end to end
# ---------------------------------------------------------------------------
def test_expansion_slots_decrypt_to_packed_constants(O, comp_params):
    """Leaf j of fast_expand_qry holds phase ~ round(Q/t)*[j == col] for
    first-dimension slots and ~ bit * B^(l-1-k) for selector rows."""
    p = comp_params.with_layout(700, 8)
    c = O.PirClient(p, 3, 77)
    srv = O.PirServer(p)
    srv.set_client_keys(3, c.create_bv_galois_keys(), c.generate_gsw_from_key())
    idx = 517
    bits = c.get_query_indices(p, idx)
    leaves = srv.fast_expand_qry(3, c.fast_generate_query(p, idx))
    Q, t, sk = p.rns_mods[0], p.plain_mod, c.secret_key()
    delta = (Q + t // 2) // t
    bound = 2**44
    for j in range(p.fst_dim_sz + p.l_ep * (p.num_dims - 1)):
        ph = O.phase(p, sk, leaves[j], False).astype(object)
        if j < p.fst_dim_sz:
            want = delta if j == bits[0] else 0
        else:
            sel, k = divmod(j - p.fst_dim_sz, p.l_ep)
            want = bits[sel + 1] * pow(2, p.base_log2 * (p.l_ep - 1 - k), Q)
        err = (int(ph[0]) - want + Q // 2) % Q - Q // 2
        assert abs(err) < bound, (j, err)
        assert max(abs(int(x)) for x in ph[1:]) < bound


def test_first_dim_selects_column(O, comp_params):
    p = comp_params.with_layout(43, 5)
    c = O.PirClient(p, 1, 5)
    srv = O.PirServer(p)
    srv.load_data(43, "affine", 0, list(range(p.num_pt)))
    srv.set_client_keys(1, c.create_bv_galois_keys(), c.generate_gsw_from_key())
    col = 9
    leaves = srv.fast_expand_qry(1, c.fast_generate_query(p, col))
    mid = srv.evaluate_first_dim(leaves[: p.fst_dim_sz])
    sk = c.secret_key()
    for r in range(p.other_dim_sz):
        got = O.decrypt_rns(p, sk, mid[r], False)
        assert np.array_equal(got, srv.direct_get_original_plaintext(r * p.fst_dim_sz + col))


@pytest.mark.parametrize("layout", [(1, 0, True), (43, 5, True), (3, 2, True), (700, 8, True)])
def test_end_to_end_decrypts(O, comp_params, layout):
    """tests/test_pir.cpp / test_shared_session.cpp: decrypt(response) = record."""
    p = comp_params.with_layout(*layout)
    c = O.PirClient(p, 9, 1234 + layout[0])
    srv = O.PirServer(p)
    q_idx = p.target_num_pt - 1
    srv.load_data(p.target_num_pt, "affine", 0, [q_idx, 0])
    srv.set_client_keys(9, c.create_bv_galois_keys(), c.generate_gsw_from_key())
    for idx in [q_idx, 0]:
        resp = srv.make_query(9, c.fast_generate_query(p, idx))
        wire = srv.save_resp_to_stream(resp)
        assert len(wire) == (2 * N * 22 + 7) // 8 == 11264
        pt, budget = c.decrypt_mod_q(c.load_resp_from_stream(wire))
        assert np.array_equal(pt, srv.direct_get_original_plaintext(idx)) and budget >= 1


def test_loader_contract(O, comp_params):
    """tests/test_server_loader.cpp: padding row recorded as zero, bad input rejected."""
    p = comp_params.with_layout(3, 2)
    srv = O.PirServer(p)
    srv.load_data(3, "affine", 0, [0, 2, 3])
    assert srv.direct_get_original_plaintext(0)[7] == (3 * 7 + 11) % p.plain_mod
    assert not srv.direct_get_original_plaintext(3).any()
    bad = np.zeros((1, N), np.uint64); bad[0, 17] = p.plain_mod
    with pytest.raises(ValueError):
        srv.load_data(1, "array", 0, [], bad)
    with pytest.raises(ValueError):
        srv.load_data(5, "affine", 0, [])


def test_codec_bit_layout(O, comp_params):
    p = comp_params.with_layout(1, 0)
    srv = O.PirServer(p)
    ct = np.zeros((2, N), np.uint64)
    ct[0, 0] = 0x2ABCDE  # 22 bits, LSB-first
    ct[0, 1] = 1
    ct[1, N - 1] = (1 << 22) - 1
    w = srv.save_resp_to_stream(ct)
    assert w[0] == 0xDE and w[1] == 0xBC and w[2] == 0x2A | (1 << 6)
    assert w[-1] == 0xFF and w[-2] == 0xFF and w[-3] >> 2 == 0x3F
