"""Seed-compressed query and key wire formats (onionpirv2_amd/host/codec.h)
for seeded clients, CPU only: encode -> decode round trips exactly, the
sizes equal the reference's size model (src/includes/pir.h:267-291) where
it is exact, and an answer computed by the CPU oracle from the decoded
query and keys decrypts to the stored plaintext.
"""
import numpy as np
import pytest

CONFIGS = ["CONFIG_N2048_K1_COMP", "CONFIG_N2048_K1", "CONFIG_N2048_K2_MP", "CONFIG_N4096_K2_MP"]


def _seeded(H, cfg, layout=(43, 5, True), cid=3, tag=0):
    p = H.PirParams(cfg).with_layout(*layout)
    key = bytes((7 * i + tag) % 256 for i in range(32))
    return p, H.PirClient(p, cid, key)


@pytest.mark.parametrize("cfg", CONFIGS)
def test_query_round_trip_and_size(cfg):
    import onionpirv2_amd.host as H
    p, c = _seeded(H, cfg)
    assert c.seeded
    q = c.fast_generate_query(p, 40)
    wire = H.encode_query(p, c.last_query_seed(), q)
    assert len(wire) == H.query_wire_bytes(p) == p.get_BFV_size(True)
    assert np.array_equal(H.decode_query(p, wire), q)
    q2 = c.fast_generate_query(p, 40)  # a fresh query seed per query
    assert not np.array_equal(q2[1], q[1])
    with pytest.raises(ValueError):
        H.decode_query(p, wire[:-1])
    bad = bytearray(wire)
    bad[32:40] = b"\xff" * 8  # first value all ones: at or above the modulus
    with pytest.raises(ValueError):
        H.decode_query(p, bytes(bad))


@pytest.mark.parametrize("cfg", CONFIGS)
def test_keys_round_trip_and_size(cfg):
    import onionpirv2_amd.host as H
    p, c = _seeded(H, cfg, tag=1)
    bv, gsw = c.create_bv_galois_keys(), c.generate_gsw_from_key()
    wire = H.encode_keys(p, c.key_seed(), bv, gsw)
    h = p.expansion_height
    assert len(wire) == H.keys_wire_bytes(p, h)
    # one seed for the bundle, so 32 bytes per row less than the per-row-seed model
    model = p.get_gsw_key_size(True) + p.get_bv_galois_key_size(True)
    assert len(wire) == model - 32 * (h * p.l_ks + 2 * p.l_key) + 33
    d = H.decode_keys(p, wire)
    assert d["h"] == h and d["key"] == c.key_seed()
    bv2, gsw2 = H.expand_keys(p, wire)
    assert np.array_equal(bv2, bv) and np.array_equal(gsw2, gsw)
    with pytest.raises(ValueError):
        H.decode_keys(p, wire[:-5])


def test_headline_wire_sizes():
    """The 256 x 1024 layout (expansion height 9): 14880-byte query (paper: 15
    KB) and 1366049 bytes of keys (paper: 1.5 MB)."""
    import onionpirv2_amd.host as H
    p = H.PirParams().with_layout(262144, 9, True)
    assert H.query_wire_bytes(p) == 14880
    assert H.keys_wire_bytes(p, 9) == 33 + (9 * 8 + 20) * 14848 == 1366049


@pytest.mark.parametrize("cfg", CONFIGS)
def test_wire_answer_decrypts_on_the_oracle(O, cfg):
    import onionpirv2_amd.host as H
    p, c = _seeded(H, cfg, layout=(43, 5, True), cid=9, tag=2)
    op = O.PirParams(cfg).with_layout(43, 5, True)
    kw = H.encode_keys(p, c.key_seed(), c.create_bv_galois_keys(), c.generate_gsw_from_key())
    srv = O.PirServer(op)
    srv.load_data(op.target_num_pt, "random", 5, [17, 42])
    bv, gsw = H.expand_keys(p, kw)
    srv.set_client_keys(9, bv, gsw)
    for idx in (17, 42):
        q = c.fast_generate_query(p, idx)
        qw = H.encode_query(p, c.last_query_seed(), q)
        resp = srv.make_query(9, H.decode_query(p, qw))
        pt, budget = c.decrypt_mod_q(resp)
        assert np.array_equal(pt, srv.direct_get_original_plaintext(idx)) and budget > 0
