"""Host API (product C++) on the two-prime scheme CONFIG_N4096_K2_MP.

CPU: PirClient is bit-exact with the oracle (secret key, BV keys, RGSW(s),
packed queries) for K = 2.
GPU: PirServer on a K = 2 layout answers bit-exactly like the oracle on the
same database; gen_data_device fills the database in HBM with the seeded
formula the host uses for recorded plaintexts, and answers decrypt to them.
"""
import numpy as np
import pytest

CFG = "CONFIG_N4096_K2_MP"


def splitmix_pts(seed, first, count, N, t):
    """Coefficient c of plaintext i = splitmix64(seed + i*N + c) mod t (uint64 wrap-around)."""
    with np.errstate(over="ignore"):
        x = np.uint64(seed) + np.arange(first * N, (first + count) * N, dtype=np.uint64)
        x = x + np.uint64(0x9E3779B97F4A7C15)
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        x = x ^ (x >> np.uint64(31))
    return (x % np.uint64(t)).reshape(count, N)


def splitmix64_scalar(x):
    """Vigna's splitmix64 output function for state x (python integers)."""
    M = (1 << 64) - 1
    z = (x + 0x9E3779B97F4A7C15) & M
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
    return z ^ (z >> 31)


def test_splitmix_formula():
    # the published generator's first output from seed 0
    assert splitmix64_scalar(0) == 0xE220A8397B1DCDAF
    t = 549755813911
    x = splitmix_pts(77, 5, 2, 8, t)
    for i in range(2):
        for c in range(8):
            assert int(x[i, c]) == splitmix64_scalar(77 + (5 + i) * 8 + c) % t


@pytest.mark.parametrize("layout", [(700, 8, True), (43, 5, False)])
def test_host_client_k2_bit_exact_with_oracle(O, layout):
    import onionpirv2_amd.host as H
    hp = H.PirParams(CFG).with_layout(*layout)
    op = O.PirParams(CFG).with_layout(*layout)
    assert list(hp.rns_mods) == list(op.rns_mods) and hp.small_q == op.small_q and hp.plain_mod == op.plain_mod
    hc, oc = H.PirClient(hp, 4, 99), O.PirClient(op, 4, 99)
    assert np.array_equal(hc.secret_key(), oc.secret_key())
    assert np.array_equal(hc.create_bv_galois_keys(), oc.create_bv_galois_keys())
    assert np.array_equal(hc.generate_gsw_from_key(), oc.generate_gsw_from_key())
    for idx in (0, hp.num_pt - 1):
        assert np.array_equal(hc.fast_generate_query(hp, idx), oc.fast_generate_query(op, idx))


@pytest.mark.gpu
def test_host_server_k2_matches_oracle(O):
    import onionpirv2_amd.host as H
    print("[gpu] host K=2 server vs oracle", flush=True)
    layout = (700, 8, True)
    hp = H.PirParams(CFG).with_layout(*layout)
    op = O.PirParams(CFG).with_layout(*layout)
    pts = splitmix_pts(1234, 0, hp.target_num_pt, hp.N, hp.plain_mod)
    rec = [0, 333, hp.target_num_pt - 1]
    hs = H.PirServer(hp)
    hs.load_data_array(pts, rec)
    os_ = O.PirServer(op)
    os_.load_data(op.target_num_pt, "array", 0, rec, pts)
    hc, oc = H.PirClient(hp, 5, 17), O.PirClient(op, 5, 17)
    bv, gsw = hc.create_bv_galois_keys(), hc.generate_gsw_from_key()
    hs.set_client_bv_galois_key(5, bv)
    hs.set_client_gsw_key(5, gsw)
    os_.set_client_keys(5, oc.create_bv_galois_keys(), oc.generate_gsw_from_key())
    qs = []
    for idx in rec:
        q = hc.fast_generate_query(hp, idx)
        assert np.array_equal(q, oc.fast_generate_query(op, idx))
        r = hs.make_query(5, q.copy())
        assert np.array_equal(r, os_.make_query(5, q))
        pt, budget = hc.decrypt_mod_q(r)
        assert np.array_equal(pt, pts[idx]) and budget > 0
        qs.append(q)
    batch = hs.make_queries([5] * len(qs), np.stack(qs))
    for j, q in enumerate(qs):
        assert np.array_equal(batch[j], hs.make_query(5, q.copy()))


@pytest.mark.gpu
def test_gen_data_device_k2(O):
    import onionpirv2_amd.host as H
    print("[gpu] K=2 device-generated database", flush=True)
    hp = H.PirParams(CFG).with_layout(1024, 8, True)
    hs = H.PirServer(hp)
    rec = [7, 1023]
    hs.gen_data_device(rec, 4242)
    c = H.PirClient(hp, 6, 23)
    hs.set_client_bv_galois_key(6, c.create_bv_galois_keys())
    hs.set_client_gsw_key(6, c.generate_gsw_from_key())
    for idx in rec:
        want = splitmix_pts(4242, idx, 1, hp.N, hp.plain_mod)[0]
        assert np.array_equal(hs.direct_get_original_plaintext(idx), want)
        pt, budget = c.decrypt_mod_q(hs.make_query(6, c.fast_generate_query(hp, idx)))
        assert np.array_equal(pt, want) and budget > 0


@pytest.mark.gpu
def test_gen_data_device_k1():
    import onionpirv2_amd.host as H
    print("[gpu] K=1 device-generated database", flush=True)
    hp = H.PirParams().with_layout(4096, 9, True)
    hs = H.PirServer(hp)
    rec = [0, 4095]
    hs.gen_data_device(rec, 99)
    c = H.PirClient(hp, 8, 24)
    hs.set_client_bv_galois_key(8, c.create_bv_galois_keys())
    hs.set_client_gsw_key(8, c.generate_gsw_from_key())
    for idx in rec:
        want = splitmix_pts(99, idx, 1, hp.N, hp.plain_mod)[0]
        assert np.array_equal(hs.direct_get_original_plaintext(idx), want)
        pt, _ = c.decrypt_mod_q(hs.make_query(8, c.fast_generate_query(hp, idx)))
        assert np.array_equal(pt, want)
