"""GPU parity of the RNS-limb schemes against the CPU oracle, bit-exact,
through the C ABI. Every test runs for all three:
  CONFIG_N4096_K2_MP: N = 4096, two ~60-bit RNS limbs (u64 kernels);
  CONFIG_N2048_K2_MP: N = 2048, two 29-bit RNS limbs (the 32-bit kernels of
    the composite scheme with the MP gadgets and the drop-then-rescale mod
    switch, database_constants.h:96-109);
  CONFIG_N2048_K1: N = 2048, one ~60-bit prime (the u64 kernels with the
    single-modulus balanced gadgets and the direct rescale,
    database_constants.h:61-75).

Layouts (with_layout, K = 2 gadgets l_ep = 5):
  (43, 5)        N0=16,  dims=3, Nrest=3   ragged deepest level (offset 1)
  (700, 8)       N0=128, dims=4, Nrest=6   ragged with offset 2
  (1024, 8)      N0=128, dims=4, Nrest=8   perfect tree
  (43, 5, False) N0=27,  dims=2, Nrest=2   non-power-of-two first dimension
  (1, 0)         singleton: no expansion, no selectors
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LAYOUTS = [(43, 5, True), (700, 8, True), (1024, 8, True), (43, 5, False), (1, 0, True)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


K2_CONFIGS = ["CONFIG_N4096_K2_MP", "CONFIG_N2048_K2_MP", "CONFIG_N2048_K1"]


@pytest.fixture(scope="module", params=K2_CONFIGS)
def k2name(request):
    return request.param


@pytest.fixture(scope="module")
def k2p(O, k2name):
    return O.PirParams(k2name)


def scheme_k2(O, p):
    q0, q1 = (p.rns_mods + [0])[:2]  # one prime: q2 = 0
    return dict(poly_degree=p.N, composite=0, q1=q0, q2=q1, ntt_root=O.ntt_root(p.N, q0),
                ntt_root2=O.ntt_root(p.N, q1) if q1 else 0,
                plain_mod=p.plain_mod, small_q=p.small_q, l_ep=p.l_ep, base_log2=p.base_log2, l_key=p.l_key,
                base_log2_key=p.base_log2_key, l_ks=p.l_ks, bv_base_log2=p.bv_base_log2)


@pytest.fixture(scope="module")
def ctx2(O, k2p, k2name):
    from onionpirv2_amd.gpu import GpuContext
    say(f"[gpu] creating RNS-limb context ({k2name})")
    ctx = GpuContext(scheme_k2(O, k2p))
    yield ctx
    ctx.close()


def rand_rns(rng, p, shape):
    """Random polynomials [..., K limbs, N], each limb reduced mod its prime."""
    out = np.empty(shape + (len(p.rns_mods), p.N), dtype=np.uint64)
    for m, q in enumerate(p.rns_mods):
        out[..., m, :] = rng.integers(0, q, size=shape + (p.N,), dtype=np.uint64)
    return out


class Setup:
    def __init__(self, O, base, layout, seed, ctx, cid=11):
        from onionpirv2_amd.gpu import GpuDatabase
        self.p = p = base.with_layout(*layout)
        self.cid = cid
        self.client = O.PirClient(p, cid, seed)
        self.bv = self.client.create_bv_galois_keys()
        self.gsw = self.client.generate_gsw_from_key()
        self.server = O.PirServer(p)
        logical = p.target_num_pt
        self.indices = sorted({0, logical - 1, (seed * 7919) % logical})
        self.server.load_data(logical, "random", seed, self.indices)
        self.server.set_client_keys(cid, self.bv, self.gsw)
        self.db = GpuDatabase(ctx, p.fst_dim_sz, p.other_dim_sz, p.num_dims, p.expansion_height)
        draws = O.mt19937_64_draws(seed, logical * p.N)
        self.db.load_plain((draws % np.uint64(p.plain_mod)).reshape(logical, p.N))
        ctx.set_client_keys(cid, self.bv, self.gsw)


@pytest.mark.parametrize("nq", [1, 3])
def test_k2_ntt_matches_oracle(O, k2p, ctx2, nq):
    say(f"[gpu] ntt nq={nq}")
    rng = np.random.default_rng(50 + nq)
    polys = rand_rns(rng, k2p, (nq,))
    exp = polys.copy()
    for i in range(nq):
        for m, q in enumerate(k2p.rns_mods):
            exp[i, m] = O.ntt_fwd(polys[i, m], k2p.N, q)
    got = ctx2.ntt_fwd(polys.reshape(nq, -1)).reshape(exp.shape)
    assert np.array_equal(got, exp)
    assert np.array_equal(ctx2.ntt_inv(got.reshape(nq, -1)).reshape(exp.shape), polys)


def test_k2_ntt_kat_product(k2p, ctx2):
    """tests/test_hexl_ntt.cpp:41-59 on every limb: (1+2x+3x^2)(4+5x)."""
    N, K = k2p.N, len(k2p.rns_mods)
    a = np.zeros((K, N), np.uint64); a[:, :3] = [1, 2, 3]
    b = np.zeros((K, N), np.uint64); b[:, :2] = [4, 5]
    fa = ctx2.ntt_fwd(a.reshape(1, -1)).reshape(K, N)
    fb = ctx2.ntt_fwd(b.reshape(1, -1)).reshape(K, N)
    prod = np.array([[(int(x) * int(y)) % q for x, y in zip(fa[m], fb[m])] for m, q in enumerate(k2p.rns_mods)],
                    dtype=np.uint64)
    c = ctx2.ntt_inv(prod.reshape(1, -1)).reshape(K, N)
    for m in range(K):
        assert list(c[m, :5]) == [4, 13, 22, 15, 0]


@pytest.mark.parametrize("which", ["key", "data"])
def test_k2_external_product_matches_oracle(O, k2p, ctx2, which):
    say(f"[gpu] external product ({which})")
    p = k2p
    l, b = (p.l_key, p.base_log2_key) if which == "key" else (p.l_ep, p.base_log2)
    sk, _ = O.gen_sk_and_encrypt(p, np.zeros(p.N, np.uint64), 5)
    one = np.zeros(p.N, np.uint64); one[0] = 1
    gsw = O.plain_to_gsw(p, one, sk, l, b, 9)
    ct = rand_rns(np.random.default_rng(7), p, (2,)).reshape(2, -1)
    exp = O.external_product(p, gsw, ct, l, b)
    assert np.array_equal(ctx2.ext_prod(gsw, l, b, ct), exp)


@pytest.mark.parametrize("sel", ["one", "random"])
def test_k2_ext_prod_mux_matches_oracle(O, k2p, ctx2, k2name, sel):
    """PirServer::ext_prod_mux (src/server.cpp:652-722) with two 60-bit limbs,
    through the C ABI and the host PirServer, bit-exact against the oracle."""
    import onionpirv2_amd.host as H
    say(f"[gpu] ext_prod_mux ({sel} selector)")
    p = k2p
    l, b = p.l_ep, p.base_log2
    rng = np.random.default_rng(23)
    if sel == "random":
        gsw = rand_rns(rng, p, (2 * l, 2)).reshape(2 * l, -1)
    else:
        sk, _ = O.gen_sk_and_encrypt(p, np.zeros(p.N, np.uint64), 5)
        one = np.zeros(p.N, np.uint64); one[0] = 1
        gsw = O.plain_to_gsw(p, one, sk, l, b, 9)
    x = rand_rns(rng, p, (2,))
    y = rand_rns(rng, p, (2,))
    q0, q1 = p.rns_mods[0], p.rns_mods[-1]
    x[0, 0, :3] = [0, q0 - 1, 7]
    y[0, 0, :3] = [q0 - 1, 0, 7]
    x[1, -1, :2] = [q1 - 1, 0]
    y[1, -1, :2] = [q1 - 1, q1 - 1]
    x, y = x.reshape(2, -1), y.reshape(2, -1)
    exp = O.PirServer(p.with_layout(43, 5)).ext_prod_mux(x.copy(), y.copy(), gsw)
    assert np.array_equal(ctx2.ext_prod_mux(gsw, l, b, x, y), exp)
    hp = H.PirParams(k2name).with_layout(43, 5)
    assert np.array_equal(H.PirServer(hp).ext_prod_mux(x, y, gsw), exp)


def test_k2_mod_switch_matches_oracle(O, k2p, ctx2):
    say("[gpu] mod switch")
    p = k2p
    ct = rand_rns(np.random.default_rng(3), p, (2,))
    for m, q in enumerate(p.rns_mods):
        ct[0, m, :5] = [0, 1, q // 2, q // 2 + 1, q - 1]
    ct = ct.reshape(2, -1)
    srv = O.PirServer(p.with_layout(1, 0))
    assert np.array_equal(ctx2.mod_switch(ct), srv.mod_switch(ct))


@pytest.mark.parametrize("layout", LAYOUTS, ids=lambda l: "x".join(map(str, l)))
def test_k2_stages_match_oracle(O, k2p, ctx2, layout):
    say(f"[gpu] stages layout={layout}")
    s = Setup(O, k2p, layout, seed=3000 + layout[0], ctx=ctx2)
    p = s.p
    query = s.client.fast_generate_query(p, s.indices[-1])
    say("  expand")
    exp_o = s.server.fast_expand_qry(s.cid, query)
    exp_g = s.db.expand(s.cid, query)
    bad = [j for j in range(len(exp_o)) if not np.array_equal(exp_g[j], exp_o[j])]
    assert not bad, f"leaves differing from the oracle: {bad}"
    sel_o = None
    if p.num_dims > 1:
        say("  complete")
        sel_o = s.server.complete_selectors(s.cid, exp_o)
        assert np.array_equal(s.db.complete(s.cid, exp_o), sel_o)
    say("  first dim")
    fq = exp_o[: p.fst_dim_sz]
    mid_o = s.server.evaluate_first_dim(fq)
    assert np.array_equal(s.db.first_dim(fq), mid_o)
    if p.num_dims > 1:
        say("  other dim")
        res_o = s.server.evaluate_other_dim(mid_o, sel_o)
        assert np.array_equal(s.db.other_dim(mid_o, sel_o), res_o)


@pytest.mark.parametrize("layout", LAYOUTS, ids=lambda l: "x".join(map(str, l)))
def test_k2_answer_bit_exact_and_decrypts(O, k2p, ctx2, layout):
    say(f"[gpu] answer layout={layout}")
    s = Setup(O, k2p, layout, seed=4000 + layout[0], ctx=ctx2)
    p = s.p
    for idx in s.indices:
        query = s.client.fast_generate_query(p, idx)
        resp_o = s.server.make_query(s.cid, query)
        resp_g = s.db.answer(s.cid, query)
        assert np.array_equal(resp_g, resp_o), f"index {idx}"
        wire = s.server.save_resp_to_stream(resp_g)
        assert len(wire) == (2 * p.N * int(p.small_q - 1).bit_length() + 7) // 8
        pt, budget = s.client.decrypt_mod_q(s.client.load_resp_from_stream(wire))
        assert np.array_equal(pt, s.server.direct_get_original_plaintext(idx)) and budget > 0


def test_k2_batch_equals_single(O, k2p, ctx2):
    say("[gpu] batch vs single")
    s = Setup(O, k2p, (1024, 8, True), seed=88, ctx=ctx2, cid=31)
    other = O.PirClient(s.p, 32, 89)
    ctx2.set_client_keys(32, other.create_bv_galois_keys(), other.generate_gsw_from_key())
    qs, ids = [], []
    for j, idx in enumerate([5, 1023, 600, 5]):
        c = s.client if j % 2 == 0 else other
        qs.append(c.fast_generate_query(s.p, idx))
        ids.append(c.client_id)
    batch = s.db.answer_batch(ids, np.stack(qs))
    for j in range(len(qs)):
        assert np.array_equal(batch[j], s.db.answer(ids[j], qs[j]))
    assert np.array_equal(batch[0], s.server.make_query(31, qs[0]))


def test_k2_graph_replays_match_eager(O, k2p, ctx2):
    """The two-prime answer captured as a HIP graph and replayed: every
    replay equals the eager answer, with eager calls in between and a
    changing client order in the batch."""
    say("[gpu] graph replays")
    s = Setup(O, k2p, (1024, 8, True), seed=91, ctx=ctx2, cid=35)
    other = O.PirClient(s.p, 36, 92)
    ctx2.set_client_keys(36, other.create_bv_galois_keys(), other.generate_gsw_from_key())
    qa = s.client.fast_generate_query(s.p, 7)
    qb = other.fast_generate_query(s.p, 1000)
    want_a, want_b = s.db.answer(35, qa), s.db.answer(36, qb)
    try:
        ctx2.set_graphs(True)
        for r in range(10):  # more calls than key-table slots
            assert np.array_equal(s.db.answer(35, qa), want_a), (r, "a")
            ids, qs, want = ([35, 36], [qa, qb], [want_a, want_b]) if r % 2 else ([36, 35], [qb, qa], [want_b, want_a])
            got = s.db.answer_batch(ids, np.stack(qs))
            assert np.array_equal(got, np.stack(want)), (r, "batch")
            if r % 4 == 3:
                ctx2.set_graphs(False)
                assert np.array_equal(s.db.answer(36, qb), want_b), (r, "eager")
                ctx2.set_graphs(True)
    finally:
        ctx2.set_graphs(False)
        ctx2.remove_client(36)


def test_k2_completion_ahead_of_scan_matches(O, k2p, ctx2):
    """The N = 4096 two-prime completion (one 1024-thread workgroup per CU)
    runs ahead of the scan on the answer stream. On 8 compute units, eager
    and replayed as graphs, a 4-query batch (60 completion workgroups) and a
    single query answer as on the unrestricted context."""
    say("[gpu] completion ahead of the scan (8-CU mask)")
    s = Setup(O, k2p, (1024, 8, True), seed=95, ctx=ctx2, cid=37)
    other = O.PirClient(s.p, 38, 96)
    ctx2.set_client_keys(38, other.create_bv_galois_keys(), other.generate_gsw_from_key())
    ids = [37, 38, 37, 38]
    qs = np.stack([(s.client if c == 37 else other).fast_generate_query(s.p, i)
                   for c, i in zip(ids, [3, 900, 511, 64])])
    want = s.db.answer_batch(ids, qs)
    assert np.array_equal(want[0], s.server.make_query(37, qs[0]))
    try:
        ctx2.set_cu_mask(range(8))
        assert np.array_equal(s.db.answer_batch(ids, qs), want)
        assert np.array_equal(s.db.answer(38, qs[1]), want[1])
        ctx2.set_graphs(True)
        for r in range(3):
            assert np.array_equal(s.db.answer_batch(ids, qs), want), r
    finally:
        ctx2.set_graphs(False)
        ctx2.set_cu_mask(None)
    try:
        assert np.array_equal(s.db.answer_batch(ids, qs), want)  # unrestricted again
    finally:
        ctx2.remove_client(38)


@pytest.mark.parametrize("layout", LAYOUTS[:4], ids=lambda l: "x".join(map(str, l)))
def test_k2_every_scan_shape_matches_oracle(O, k2p, ctx2, layout):
    """Every tile shape of the two-prime scan gives the oracle's first
    dimension (one query) and the same 9-query batch answers."""
    from test_gpu_parity import _each_scan_shape
    say(f"[gpu] scan shapes layout={layout}")
    s = Setup(O, k2p, layout, seed=4000 + layout[0], ctx=ctx2, cid=41)
    p = s.p
    query = s.client.fast_generate_query(p, s.indices[-1])
    fq = s.server.fast_expand_qry(s.cid, query)[: p.fst_dim_sz]
    mid_o = s.server.evaluate_first_dim(fq)
    mids = _each_scan_shape(ctx2, lambda: s.db.first_dim(fq))
    assert len(mids) >= 4
    for v, m in mids.items():
        assert np.array_equal(m, mid_o), f"scan shape {v}"
    idxs = [(s.indices[-1] + 13 * j) % p.target_num_pt for j in range(9)]
    qs = np.stack([s.client.fast_generate_query(p, i) for i in idxs])
    ids = [s.cid] * len(idxs)
    want = s.db.answer_batch(ids, qs)
    assert np.array_equal(want[0], s.server.make_query(s.cid, qs[0]))
    got = _each_scan_shape(ctx2, lambda: s.db.answer_batch(ids, qs))
    assert len(got) >= 4
    for v, b in got.items():
        assert np.array_equal(b, want), f"scan shape {v}"


@pytest.mark.parametrize("kexpr", ["N+1", "N//4+1", "3", "2*N-1"])
def test_k2_bv_keyswitch_matches_oracle(O, k2p, ctx2, kexpr):
    """bv_apply_galois_inplace_k2 (bv_keyswitch.cpp:373-481: signed MP digits,
    ties to +B/2) alone through opir_bv_ks, against the oracle, bit-exact."""
    p = k2p
    N = p.N
    k = eval(kexpr, {"N": N})
    say(f"[gpu] BV keyswitch k={k}")
    sk, _ = O.gen_sk_and_encrypt(p, np.zeros(N, np.uint64), 25)
    key = O.gen_bv_ks_key(p, sk, k, 26 + k)[0]
    ct = rand_rns(np.random.default_rng(k), p, (2,)).reshape(2, -1)
    want = O.bv_apply(p, ct, k, key[None])
    assert np.array_equal(ctx2.bv_ks(ct, k, key), want)
