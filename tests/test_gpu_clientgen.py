"""Client-side helpers on the GPU (opir_client_keygen_device,
opir_query_gen): keys and queries of seeded clients generated in HBM equal
the host PirClient(params, id, key)'s, value for value, and the answers to
device-generated queries under device-generated keys decrypt to the stored
plaintexts. Reference: src/client.cpp:20-225, src/bv_keyswitch.cpp:191-276.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CONFIGS = ["CONFIG_N2048_K1_COMP", "CONFIG_N2048_K2_MP"]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _scheme(O, cfg):
    p = O.PirParams(cfg)
    if cfg == "CONFIG_N2048_K1_COMP":
        from conftest import scheme_from_oracle
        return scheme_from_oracle(p)
    q1, q2 = p.rns_mods
    return dict(poly_degree=p.N, composite=0, q1=q1, q2=q2, ntt_root=O.ntt_root(p.N, q1), ntt_root2=O.ntt_root(p.N, q2),
                plain_mod=p.plain_mod, small_q=p.small_q, l_ep=p.l_ep, base_log2=p.base_log2, l_key=p.l_key,
                base_log2_key=p.base_log2_key, l_ks=p.l_ks, bv_base_log2=p.bv_base_log2)


def _keys(n, tag):
    return [bytes((31 * i + 7 * j + tag) % 256 for i in range(32)) for j in range(n)]


@pytest.mark.parametrize("cfg", CONFIGS)
def test_device_keygen_equals_host_seeded_client(O, cfg):
    import onionpirv2_amd.host as H
    from onionpirv2_amd.gpu import GpuContext
    say(f"[gpu] device keygen {cfg}")
    hp = H.PirParams(cfg).with_layout(4096, 9, True)
    ctx = GpuContext(_scheme(O, cfg))
    try:
        keys = _keys(3, 1)
        ids = [100, 101, 102]
        ctx.keygen_device(ids, keys, hp.expansion_height)
        for cid, k in zip(ids, keys):
            c = H.PirClient(hp, cid, k)
            bv, gsw = ctx.get_client_keys(cid, hp.expansion_height)
            assert np.array_equal(bv, c.create_bv_galois_keys()), cid
            assert np.array_equal(gsw, c.generate_gsw_from_key()), cid
    finally:
        ctx.close()


@pytest.mark.parametrize("cfg", CONFIGS)
def test_device_queries_equal_host_and_decrypt(O, cfg):
    import onionpirv2_amd.host as H
    from onionpirv2_amd.gpu import GpuContext, GpuDatabase, InvalidArgument, OutOfRange
    say(f"[gpu] device queries {cfg}")
    layout = (700, 8, True)
    hp = H.PirParams(cfg).with_layout(*layout)
    op = O.PirParams(cfg).with_layout(*layout)
    ctx = GpuContext(_scheme(O, cfg))
    try:
        db = GpuDatabase(ctx, op.fst_dim_sz, op.other_dim_sz, op.num_dims, op.expansion_height)
        pts = (O.mt19937_64_draws(3, op.num_pt * op.N) % np.uint64(op.plain_mod)).reshape(op.num_pt, op.N)
        db.load_plain(pts)
        keys = _keys(2, 9)
        ids = [7, 8]
        ctx.keygen_device(ids, keys, op.expansion_height)
        clients = [H.PirClient(hp, cid, k) for cid, k in zip(ids, keys)]
        idx = [0, 5, 699, 350, 128, 64]
        owner = [0, 1, 0, 1, 0, 0]
        counters, host_q = [], []
        for i, o in zip(idx, owner):
            counters.append(sum(1 for x in owner[: len(counters)] if x == o))
            host_q.append(clients[o].fast_generate_query(hp, i))
        dev_q = db.query_gen([keys[o] for o in owner], counters, idx)
        assert np.array_equal(dev_q, np.stack(host_q))
        resp = db.answer_batch([ids[o] for o in owner], dev_q)
        for r, i, o in zip(resp, idx, owner):
            pt, budget = clients[o].decrypt_mod_q(r)
            assert np.array_equal(pt, pts[i]) and budget > 0, i
        with pytest.raises(OutOfRange):
            db.query_gen([keys[0]], [0], [op.num_pt])
    finally:
        ctx.close()


def test_two_prime_n4096_context_refuses_device_generation(O):
    from onionpirv2_amd.gpu import GpuContext, InvalidArgument
    p = O.PirParams("CONFIG_N4096_K2_MP")
    q0, q1 = p.rns_mods
    sch = dict(poly_degree=p.N, composite=0, q1=q0, q2=q1, ntt_root=O.ntt_root(p.N, q0), ntt_root2=O.ntt_root(p.N, q1),
               plain_mod=p.plain_mod, small_q=p.small_q, l_ep=p.l_ep, base_log2=p.base_log2, l_key=p.l_key,
               base_log2_key=p.base_log2_key, l_ks=p.l_ks, bv_base_log2=p.bv_base_log2)
    ctx = GpuContext(sch)
    try:
        with pytest.raises(InvalidArgument):
            ctx.keygen_device([1], _keys(1, 0), 2)
    finally:
        ctx.close()
