"""Host API (product C++ PirParams/PirClient) against the oracle, on the CPU.

Same seed -> the same secret key, BV Galois keys, RGSW(s) and packed
queries, bit for bit; derived parameters and layouts equal the oracle's.
"""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def H():
    import onionpirv2_amd.host as host
    return host


LAYOUTS = [(43, 5, True), (4096, 9, True), (349526, 10, True), (1, 0, True), (43, 5, False), (174763, 9, True)]


def test_params_match_oracle(H, comp_params):
    p = H.PirParams()
    o = comp_params
    assert (p.q1, p.q2, p.ntt_root) == (o.q1, o.q2, o.w_crt)
    assert list(p.rns_mods) == list(o.rns_mods)
    assert (p.plain_mod, p.small_q) == (o.plain_mod, o.small_q)
    assert (p.base_log2, p.base_log2_key, p.bv_base_log2) == (o.base_log2, o.base_log2_key, o.bv_base_log2)
    for lay in LAYOUTS:
        a, b = p.with_layout(*lay), o.with_layout(*lay)
        assert (a.fst_dim_sz, a.num_dims, a.num_pt, a.other_dim_sz) == (b.fst_dim_sz, b.num_dims, b.num_pt, b.other_dim_sz)


@pytest.mark.parametrize("layout", [(43, 5, True), (4096, 9, True), (1, 0, True)])
def test_client_bit_exact_with_oracle(H, O, comp_params, layout):
    seed = 4242 + layout[0]
    hp = H.PirParams().with_layout(*layout)
    op = comp_params.with_layout(*layout)
    hc = H.PirClient(hp, 3, seed)
    oc = O.PirClient(op, 3, seed)
    assert np.array_equal(hc.secret_key(), oc.secret_key())
    assert np.array_equal(hc.create_bv_galois_keys(), oc.create_bv_galois_keys())
    assert np.array_equal(hc.generate_gsw_from_key(), oc.generate_gsw_from_key())
    for idx in [0, hp.num_pt - 1, hp.num_pt // 3]:
        assert hc.get_query_indices(hp, idx) == oc.get_query_indices(op, idx)
        assert np.array_equal(hc.fast_generate_query(hp, idx), oc.fast_generate_query(op, idx))


def test_decrypt_and_codec_match_oracle(H, O, comp_params):
    lay = (43, 5, True)
    hp, op = H.PirParams().with_layout(*lay), comp_params.with_layout(*lay)
    hc, oc = H.PirClient(hp, 1, 99), O.PirClient(op, 1, 99)
    srv = O.PirServer(op)
    srv.load_data(43, "affine", 0, [42])
    srv.set_client_keys(1, oc.create_bv_galois_keys(), oc.generate_gsw_from_key())
    hc.create_bv_galois_keys(); hc.generate_gsw_from_key()
    q = hc.fast_generate_query(hp, 42)
    resp = srv.make_query(1, q)
    wire = srv.save_resp_to_stream(resp)
    back = hc.load_resp_from_stream(wire)
    assert np.array_equal(back, resp)
    pt_h, b_h = hc.decrypt_mod_q(back)
    pt_o, b_o = oc.decrypt_mod_q(back)
    assert np.array_equal(pt_h, pt_o) and b_h == b_o
    assert np.array_equal(pt_h, srv.direct_get_original_plaintext(42))
    with pytest.raises(RuntimeError):
        hc.load_resp_from_stream(wire[:-3])


def test_query_index_out_of_range(H):
    hp = H.PirParams().with_layout(43, 5)
    c = H.PirClient(hp, 1, 1)
    with pytest.raises(IndexError):
        c.fast_generate_query(hp, hp.num_pt)


def test_layout_rejections(H):
    p = H.PirParams()
    with pytest.raises(ValueError):
        p.with_layout(0, 0)
    with pytest.raises(ValueError):
        p.with_layout(10, 12)
    with pytest.raises(ValueError):
        p.with_query_shape(600, 1, 9)
