"""Pins for the oracle's tree PIR layer (oracle/tree.cpp), CPU only, after
the reference's own tests: the index identities of
src/tests/test_tree_index.cpp (exhausted on one tree, sampled on another),
the plan case boundaries, the packing identity and zero fill and the
MSB-first fold of src/tests/test_tree_select.cpp, and its encrypted gate:
every level of the path decrypts to the packed plaintext D_l[p_l].
"""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def scheme(O):
    return O.PirParams("CONFIG_N2048_K1_COMP")


def _check_leaf(O, tp, plans, leaf):
    alpha, beta_bits, gamma_bits = O.tree_client_index(leaf, tp)
    gamma, beta = leaf // tp.P, (leaf % tp.P) % tp.B
    assert sum(b << u for u, b in enumerate(beta_bits)) == beta
    assert sum(g << v for v, g in enumerate(gamma_bits)) == gamma
    assert gamma * tp.P + alpha * tp.B + beta == leaf
    for level in range(tp.L + 1):
        node, p, pos = O.tree_level_oracle(leaf, level, tp)
        R = plans[level]["R"]
        assert node == leaf >> (tp.L - level)
        assert node == pos * R + p and p < R
        if level >= tp.r:
            assert pos == gamma
            s = tp.L - level  # invariant 2: p_l from (alpha, beta)
            want = alpha * (1 << (tp.b - s)) + (beta >> s) if s <= tp.b else alpha >> (s - tp.b)
            assert p == want
        else:
            assert p == 0 and pos == gamma >> (tp.r - level)


@pytest.mark.parametrize("L,a,exhaust", [(13, 2, True), (16, 3, False), (12, 1, True)])
def test_tree_index_identities(O, scheme, L, a, exhaust):
    tp = O.TreeParams(L, a, scheme)
    assert (tp.N, tp.n, tp.r, tp.N0) == (1 << L, 2048, 11, 1 << a)
    assert tp.b == L - 11 - a and tp.P == tp.N // 2048 and tp.B == 1 << tp.b
    assert tp.w == tp.N0 + tp.ell * (tp.b + tp.r) and tp.W >= tp.w > tp.W // 2 and tp.W == 1 << tp.h_q
    plans = O.tree_level_plans(tp)
    assert len(plans) == L + 1
    for pl in plans:
        level = pl["level"]
        assert pl["R"] == (1 << (level - tp.r) if level >= tp.r else 1)
        if pl["R"] == 1:
            assert pl["case"] == 0 and pl["coarsen"] == tp.a and pl["beta_count"] == 0 and pl["beta_begin"] == tp.b
        elif pl["R"] < tp.N0:
            assert pl["case"] == 1 and pl["coarsen"] == tp.a - (level - tp.r) and pl["beta_count"] == 0
        else:
            assert pl["case"] == 2 and pl["beta_begin"] == tp.L - level
            assert pl["beta_count"] == level - tp.r - tp.a and pl["beta_begin"] + pl["beta_count"] == tp.b
    leaves = range(tp.N) if exhaust else [0, tp.N - 1] + [int(x) for x in np.random.default_rng(L).integers(0, tp.N, 300)]
    for leaf in leaves:
        _check_leaf(O, tp, plans, leaf)


def test_tree_packing_identity(O, scheme):
    tp = O.TreeParams(13, 2, scheme)
    for level in range(tp.L + 1):
        D = O.pack_tree_level(level, tp)
        R, nodes = D.shape[0], 1 << level
        j = np.arange(nodes)
        want = np.array([O.synthetic_tree_node(level, int(x), tp.t) for x in j], np.uint64)
        assert np.array_equal(D[j % R, j // R], want)
        filled = np.zeros(D.shape, bool)
        filled[j % R, j // R] = True
        assert not D[~filled].any()


def _plain_fold(values, target, bits, msb_first):
    values = list(values)
    for step in range(bits):
        bit = (target >> (bits - 1 - step if msb_first else step)) & 1
        half = len(values) // 2
        values = [values[j + half] if bit else values[j] for j in range(half)]
    return values[0]


def test_plain_fold_is_msb_first():
    for bits in (1, 2, 3):
        vals = [100 + i for i in range(1 << bits)]
        assert all(_plain_fold(vals, t, bits, True) == vals[t] for t in range(1 << bits))
        assert bits < 2 or any(_plain_fold(vals, t, bits, False) != vals[t] for t in range(1 << bits))


@pytest.mark.parametrize("cfg,L,a", [("CONFIG_N2048_K1_COMP", 16, 3), ("CONFIG_N2048_K1_COMP", 13, 2),
                                     ("CONFIG_N2048_K2_MP", 14, 2)])
def test_tree_select_decrypts_to_the_packed_plaintext(O, cfg, L, a):
    """tests/test_tree_select.cpp's encrypted gate on the oracle: boundary
    leaves and a beta-asymmetric one (gamma = alpha = beta = 1)."""
    scheme = O.PirParams(cfg)
    tp = O.TreeParams(L, a, scheme)
    qv = scheme.with_query_shape(tp.N0, tp.b + tp.r, tp.h_q)
    c = O.PirClient(qv, 3, 77 + L)
    srv = O.PirServer(qv)
    srv.set_client_keys(3, c.create_bv_galois_keys(), c.generate_gsw_from_key())
    levels = [O.pack_tree_level(level, tp) for level in range(tp.L + 1)]
    leaves = [0, tp.N - 1, tp.P + tp.B + (1 if tp.b >= 2 else 0)]
    for leaf in leaves:
        q = O.make_tree_query(c, tp, qv, leaf)
        out = O.tree_select_all(srv, tp, 3, q)
        assert len(out) == tp.L + 1
        for level, ct in enumerate(out):
            _, p, _ = O.tree_level_oracle(leaf, level, tp)
            assert np.array_equal(c.decrypt_ct(ct), levels[level][p]), (leaf, level)
