"""Tree PIR on the GPU (opir_tree_*: unpack_tree_query, the alpha pyramid and
select_level of src/tree_query.cpp / src/tree_select.cpp on the answer
path's kernels) against the CPU oracle, bit-exact, on the tree shapes of
src/tests/test_tree_select.cpp; every level decrypts to D_l[p_l].
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _scheme(O, cfg):
    p = O.PirParams(cfg)
    if cfg == "CONFIG_N2048_K1_COMP":
        from conftest import scheme_from_oracle
        return p, scheme_from_oracle(p)
    q1, q2 = p.rns_mods
    return p, dict(poly_degree=p.N, composite=0, q1=q1, q2=q2, ntt_root=O.ntt_root(p.N, q1),
                   ntt_root2=O.ntt_root(p.N, q2), plain_mod=p.plain_mod, small_q=p.small_q, l_ep=p.l_ep,
                   base_log2=p.base_log2, l_key=p.l_key, base_log2_key=p.base_log2_key, l_ks=p.l_ks,
                   bv_base_log2=p.bv_base_log2)


@pytest.mark.parametrize("cfg,L,a", [("CONFIG_N2048_K1_COMP", 16, 3), ("CONFIG_N2048_K1_COMP", 13, 2),
                                     ("CONFIG_N2048_K2_MP", 14, 2)])
def test_tree_select_matches_oracle_and_decrypts(O, cfg, L, a):
    from onionpirv2_amd.gpu import GpuContext, GpuTree
    say(f"[gpu] tree PIR {cfg} L={L} a={a}")
    scheme, sch = _scheme(O, cfg)
    tp = O.TreeParams(L, a, scheme)
    qv = scheme.with_query_shape(tp.N0, tp.b + tp.r, tp.h_q)
    ctx = GpuContext(sch)
    try:
        tree = GpuTree(ctx, L, a)
        assert (tree.w, tree.h_q, tree.N0, tree.nsel) == (tp.w, tp.h_q, tp.N0, tp.b + tp.r)
        levels = [O.pack_tree_level(level, tp) for level in range(L + 1)]
        for level, d in enumerate(levels):
            tree.load_level(level, d)
        c = O.PirClient(qv, 5, 1000 + L)
        bv, gsw = c.create_bv_galois_keys(), c.generate_gsw_from_key()
        ctx.set_client_keys(5, bv, gsw)
        srv = O.PirServer(qv)
        srv.set_client_keys(5, bv, gsw)
        leaves = [0, tp.N - 1, tp.P + tp.B + (1 if tp.b >= 2 else 0), 0x2BADBEEF % tp.N]
        for leaf in leaves:
            q = O.make_tree_query(c, tp, qv, leaf)
            got = tree.select(5, q)
            want = O.tree_select_all(srv, tp, 5, q)
            for level in range(L + 1):
                assert np.array_equal(got[level], want[level]), (leaf, level)
                _, p, _ = O.tree_level_oracle(leaf, level, tp)
                assert np.array_equal(c.decrypt_ct(got[level]), levels[level][p]), (leaf, level)
        tree.close()
    finally:
        ctx.close()


@pytest.mark.parametrize("cfg", ["CONFIG_N2048_K1_COMP", "CONFIG_N2048_K2_MP"])
def test_host_tree_api_matches_oracle(O, cfg):
    """The host C++ mirror (make_tree_pir_params_for_scheme, make_tree_query,
    TreePirServer) produces the oracle's query and per-level answers."""
    import onionpirv2_amd.host as H
    say(f"[gpu] host tree API {cfg}")
    L, a = 14, 2
    hs, os_ = H.PirParams(cfg), O.PirParams(cfg)
    ht, tp = H.make_tree_pir_params_for_scheme(L, a, hs), O.TreeParams(L, a, os_)
    assert (ht.w, ht.h_q, ht.N0, ht.b, ht.P, ht.B) == (tp.w, tp.h_q, tp.N0, tp.b, tp.P, tp.B)
    assert H.build_level_plans(ht) == O.tree_level_plans(tp)
    hq, oq = H.tree_query_expansion_params(ht, hs), os_.with_query_shape(tp.N0, tp.b + tp.r, tp.h_q)
    hc, oc = H.PirClient(hq, 8, 4242), O.PirClient(oq, 8, 4242)
    bv, gsw = hc.create_bv_galois_keys(), hc.generate_gsw_from_key()
    assert np.array_equal(bv, oc.create_bv_galois_keys()) and np.array_equal(gsw, oc.generate_gsw_from_key())
    srv = H.TreePirServer(hs, ht)
    srv.preprocess_synthetic()
    srv.set_client_keys(8, bv, gsw)
    ref = O.PirServer(oq)
    ref.set_client_keys(8, bv, gsw)
    for leaf in (3, ht.N - 2):
        q = H.make_tree_query(hc, hs, ht, leaf)
        assert np.array_equal(q, O.make_tree_query(oc, tp, oq, leaf))
        got = srv.select_path(8, q)
        want = O.tree_select_all(ref, tp, 8, q)
        for level in range(L + 1):
            assert np.array_equal(got[level], want[level]), (leaf, level)
            _, p, _ = O.tree_level_oracle(leaf, level, tp)
            assert np.array_equal(oc.decrypt_ct(got[level]), H.pack_tree_level_synthetic(level, ht)[p])
