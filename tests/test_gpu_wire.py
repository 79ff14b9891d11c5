"""Seed-compressed keys ingested on the GPU (opir_client_set_keys_seeded:
the device regenerates the uniform halves from the key seed) answer exactly
like the fully expanded keys, for all three parameter sets, and the answers
to wire-decoded queries equal the oracle's, bit for bit.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CONFIGS = ["CONFIG_N2048_K1_COMP", "CONFIG_N2048_K1", "CONFIG_N2048_K2_MP", "CONFIG_N4096_K2_MP"]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("cfg", CONFIGS)
def test_wire_keys_answer_like_expanded_keys(O, cfg):
    import onionpirv2_amd.host as H
    say(f"[gpu] wire keys {cfg}")
    layout = (700, 8, True)
    p = H.PirParams(cfg).with_layout(*layout)
    op = O.PirParams(cfg).with_layout(*layout)
    key = bytes((11 * i + 5) % 256 for i in range(32))
    c = H.PirClient(p, 21, key)
    kw = H.encode_keys(p, c.key_seed(), c.create_bv_galois_keys(), c.generate_gsw_from_key())
    bv, gsw = H.expand_keys(p, kw)
    srv = H.PirServer(p)
    rec = [0, 333, p.target_num_pt - 1]
    srv.gen_data_device(rec, 99)
    srv.set_client_keys_wire(21, kw)           # uniform halves regenerated on the device
    srv.set_client_bv_galois_key(22, bv)       # the same keys, expanded on the host
    srv.set_client_gsw_key(22, gsw)
    ref = O.PirServer(op)
    ref.load_data(op.target_num_pt, "random", 7, rec)
    ref.set_client_keys(21, bv, gsw)
    for i in rec:
        q = c.fast_generate_query(p, i)
        qw = H.encode_query(p, c.last_query_seed(), q)
        r_wire = srv.make_query_wire(21, qw)
        assert np.array_equal(r_wire, srv.make_query(22, q.copy()))
        pt, budget = c.decrypt_mod_q(r_wire)
        assert np.array_equal(pt, srv.direct_get_original_plaintext(i)) and budget > 0
    # bit-exact against the oracle on a database both hold
    srv2 = H.PirServer(p)
    pts = O.mt19937_64_draws(7, op.target_num_pt * op.N) % np.uint64(op.plain_mod)
    srv2.load_data_array(pts.reshape(op.target_num_pt, op.N), rec)
    srv2.set_client_keys_wire(21, kw)
    q = c.fast_generate_query(p, 333)
    qw = H.encode_query(p, c.last_query_seed(), q)
    assert np.array_equal(srv2.make_query_wire(21, qw), ref.make_query(21, H.decode_query(p, qw)))


def test_seeded_ingest_through_the_c_abi(O, comp_params):
    """opir_client_set_keys_seeded with the b / c0 halves only equals
    opir_client_set_keys with the full keys (batch answers compared)."""
    import onionpirv2_amd.host as H
    from onionpirv2_amd.gpu import GpuContext, GpuDatabase
    from conftest import scheme_from_oracle
    say("[gpu] seeded ingest through the C ABI")
    p = comp_params.with_layout(43, 5)
    hp = H.PirParams().with_layout(43, 5)
    ctx = GpuContext(scheme_from_oracle(p))
    try:
        db = GpuDatabase(ctx, p.fst_dim_sz, p.other_dim_sz, p.num_dims, p.expansion_height)
        db.load_plain(np.arange(p.num_pt * p.N, dtype=np.uint64).reshape(p.num_pt, p.N) % np.uint64(p.plain_mod))
        c = H.PirClient(hp, 1, bytes(range(32)))
        bv, gsw = c.create_bv_galois_keys(), c.generate_gsw_from_key()
        ctx.set_client_keys(1, bv, gsw)
        ctx.set_client_keys_seeded(2, c.key_seed(), bv[:, :, 1, :], gsw[:, : p.N])
        qs = np.stack([c.fast_generate_query(hp, i) for i in (0, 7, 40)])
        assert np.array_equal(db.answer_batch([1, 1, 1], qs), db.answer_batch([2, 2, 2], qs))
    finally:
        ctx.close()
