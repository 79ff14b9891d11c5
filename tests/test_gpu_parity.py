"""GPU parity: every stage of the HIP answer path against the CPU oracle,
bit-exact, through the C ABI (onionpirv2_amd/gpu.py -> libonionpir_gpu.so).

Layouts are small enough for the oracle to finish in seconds; each covers a
shape feature of the reference's evaluate_other_dim (src/server.cpp:599-649):
  (43, 5)        N0=16,  dims=3, Nrest=3   ragged deepest level (offset 1)
  (700, 8)       N0=128, dims=4, Nrest=6   ragged with offset 2
  (4096, 9)      N0=256, dims=5, Nrest=16  perfect tree (the 4096-plaintext workload)
  (43, 5, False) N0=26,  dims=2, Nrest=2   non-power-of-two first dimension
  (1, 0)         singleton: no expansion, no selectors
"""
import sys

import numpy as np
import pytest

from conftest import scheme_from_oracle

pytestmark = pytest.mark.gpu

LAYOUTS = [(43, 5, True), (700, 8, True), (4096, 9, True), (43, 5, False), (1, 0, True)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def gpu_ctx(comp_params):
    from onionpirv2_amd.gpu import GpuContext
    say("[gpu] creating context")
    ctx = GpuContext(scheme_from_oracle(comp_params))
    yield ctx
    ctx.close()


class Setup:
    """Oracle server+client and a GPU database holding the same plaintexts."""

    def __init__(self, O, base, layout, seed, gpu_ctx, cid=11):
        self.p = base.with_layout(*layout)
        p = self.p
        self.cid = cid
        self.client = O.PirClient(p, cid, seed)
        self.bv = self.client.create_bv_galois_keys()
        self.gsw = self.client.generate_gsw_from_key()
        self.server = O.PirServer(p)
        logical = p.target_num_pt
        self.indices = sorted({0, logical - 1, (seed * 7919) % logical})
        self.server.load_data(logical, "random", seed, self.indices)
        self.server.set_client_keys(cid, self.bv, self.gsw)
        from onionpirv2_amd.gpu import GpuDatabase
        self.db = GpuDatabase(gpu_ctx, p.fst_dim_sz, p.other_dim_sz, p.num_dims, p.expansion_height)
        draws = O.mt19937_64_draws(seed, logical * p.N)
        pts = (draws % np.uint64(p.plain_mod)).astype(np.uint16).reshape(logical, p.N)
        self.db.load_plain(pts)
        gpu_ctx.set_client_keys(cid, self.bv, self.gsw)


@pytest.mark.parametrize("nq", [1, 7])
def test_ntt_fwd_inv_match_oracle(O, comp_params, gpu_ctx, nq):
    say(f"[gpu] ntt fwd/inv nq={nq}")
    Q = comp_params.rns_mods[0]
    rng = np.random.default_rng(100 + nq)
    polys = rng.integers(0, Q, size=(nq, comp_params.N), dtype=np.uint64)
    exp = O.ntt_fwd(polys.ravel(), comp_params.N, Q).reshape(nq, -1)
    got = gpu_ctx.ntt_fwd(polys)
    assert np.array_equal(got, exp)
    back = gpu_ctx.ntt_inv(got)
    assert np.array_equal(back, polys)
    assert np.array_equal(gpu_ctx.ntt_inv(polys), O.ntt_inv(polys.ravel(), comp_params.N, Q).reshape(nq, -1))


def test_ntt_kat_product(comp_params, gpu_ctx):
    """tests/test_hexl_ntt.cpp:41-59: (1+2x+3x^2)(4+5x) = 4+13x+22x^2+15x^3."""
    say("[gpu] ntt product KAT")
    N, Q = comp_params.N, comp_params.rns_mods[0]
    a = np.zeros(N, np.uint64); a[:3] = [1, 2, 3]
    b = np.zeros(N, np.uint64); b[:2] = [4, 5]
    fa, fb = gpu_ctx.ntt_fwd(a), gpu_ctx.ntt_fwd(b)
    prod = np.array([(int(x) * int(y)) % Q for x, y in zip(fa.ravel(), fb.ravel())], dtype=np.uint64)
    c = gpu_ctx.ntt_inv(prod).ravel()
    assert list(c[:5]) == [4, 13, 22, 15, 0]


@pytest.mark.parametrize("which", ["key", "data"])
def test_external_product_matches_oracle(O, comp_params, gpu_ctx, which):
    say(f"[gpu] external product ({which} gadget)")
    p = comp_params
    l, b = (p.l_key, p.base_log2_key) if which == "key" else (p.l_ep, p.base_log2)
    rng = np.random.default_rng(7)
    sk, _ = O.gen_sk_and_encrypt(p, np.zeros(p.N, np.uint64), 5)
    one = np.zeros(p.N, np.uint64); one[0] = 1
    gsw = O.plain_to_gsw(p, one, sk, l, b, 9)
    ct = rng.integers(0, p.rns_mods[0], size=(2, p.N), dtype=np.uint64)
    exp = O.external_product(p, gsw, ct, l, b)
    got = gpu_ctx.ext_prod(gsw, l, b, ct)
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("sel", ["zero", "one", "random"])
def test_ext_prod_mux_matches_oracle(O, comp_params, gpu_ctx, sel):
    """PirServer::ext_prod_mux (src/server.cpp:652-722) through the C ABI and
    through the host PirServer, against the oracle, bit-exact."""
    import onionpirv2_amd.host as H
    say(f"[gpu] ext_prod_mux ({sel} selector)")
    p = comp_params
    Q = p.rns_mods[0]
    l, b = p.l_ep, p.base_log2
    rng = np.random.default_rng(19)
    if sel == "random":
        gsw = rng.integers(0, Q, size=(2 * l, 2 * p.N), dtype=np.uint64)
    else:
        sk, _ = O.gen_sk_and_encrypt(p, np.zeros(p.N, np.uint64), 5)
        pt = np.zeros(p.N, np.uint64); pt[0] = 1 if sel == "one" else 0
        gsw = O.plain_to_gsw(p, pt, sk, l, b, 9)
    x = rng.integers(0, Q, size=(2, p.N), dtype=np.uint64)
    y = rng.integers(0, Q, size=(2, p.N), dtype=np.uint64)
    # subtraction and addition wrap-around edges
    x[0, :4] = [0, Q - 1, 5, Q - 1]
    y[0, :4] = [Q - 1, 0, 5, Q - 1]
    exp = O.PirServer(p.with_layout(1, 0)).ext_prod_mux(x.copy(), y.copy(), gsw)
    x0, y0 = x.copy(), y.copy()
    assert np.array_equal(gpu_ctx.ext_prod_mux(gsw, l, b, x, y), exp)
    assert np.array_equal(x, x0) and np.array_equal(y, y0)
    hp = H.PirParams("CONFIG_N2048_K1_COMP").with_layout(43, 5)
    assert np.array_equal(H.PirServer(hp).ext_prod_mux(x, y, gsw), exp)


def test_mod_switch_matches_oracle(O, comp_params, gpu_ctx):
    say("[gpu] mod switch")
    p = comp_params
    rng = np.random.default_rng(3)
    ct = rng.integers(0, p.rns_mods[0], size=(2, p.N), dtype=np.uint64)
    # include the centring boundary values
    Q = p.rns_mods[0]
    ct[0, :6] = [0, 1, Q // 2 - 1, Q // 2, Q // 2 + 1, Q - 1]
    srv = O.PirServer(p.with_layout(1, 0))
    assert np.array_equal(gpu_ctx.mod_switch(ct), srv.mod_switch(ct))


@pytest.mark.parametrize("layout", LAYOUTS, ids=lambda l: "x".join(map(str, l)))
def test_stages_match_oracle(O, comp_params, gpu_ctx, layout):
    say(f"[gpu] stages layout={layout}")
    s = Setup(O, comp_params, layout, seed=1000 + layout[0], gpu_ctx=gpu_ctx)
    p = s.p
    idx = s.indices[-1]
    query = s.client.fast_generate_query(p, idx)
    say("  expand")
    exp_o = s.server.fast_expand_qry(s.cid, query)
    exp_g = s.db.expand(s.cid, query)
    assert np.array_equal(exp_g, exp_o)
    if p.num_dims > 1:
        say("  complete")
        sel_o = s.server.complete_selectors(s.cid, exp_o)
        sel_g = s.db.complete(s.cid, exp_o)
        assert np.array_equal(sel_g, sel_o)
    else:
        sel_o = None
    say("  first dim")
    fq = exp_o[: p.fst_dim_sz]
    mid_o = s.server.evaluate_first_dim(fq)
    mid_g = s.db.first_dim(fq)
    assert np.array_equal(mid_g, mid_o)
    say("  other dim")
    if p.num_dims > 1:
        res_o = s.server.evaluate_other_dim(mid_o, sel_o)
        res_g = s.db.other_dim(mid_o, sel_o)
        assert np.array_equal(res_g, res_o)


@pytest.mark.parametrize("layout", LAYOUTS, ids=lambda l: "x".join(map(str, l)))
def test_answer_bit_exact_and_decrypts(O, comp_params, gpu_ctx, layout):
    say(f"[gpu] answer layout={layout}")
    s = Setup(O, comp_params, layout, seed=2000 + layout[0], gpu_ctx=gpu_ctx)
    p = s.p
    for idx in s.indices:
        query = s.client.fast_generate_query(p, idx)
        resp_o = s.server.make_query(s.cid, query)
        resp_g = s.db.answer(s.cid, query)
        assert np.array_equal(resp_g, resp_o), f"index {idx}"
        wire = s.server.save_resp_to_stream(resp_g)
        assert len(wire) == 11264
        pt, budget = s.client.decrypt_mod_q(s.client.load_resp_from_stream(wire))
        assert np.array_equal(pt, s.server.direct_get_original_plaintext(idx))
        assert budget > 0


def test_batch_equals_single(O, comp_params, gpu_ctx):
    """Each query's response is the same alone and inside a batch."""
    say("[gpu] batch vs single")
    s = Setup(O, comp_params, (4096, 9, True), seed=77, gpu_ctx=gpu_ctx, cid=21)
    other = O.PirClient(s.p, 22, 78)
    gpu_ctx.set_client_keys(22, other.create_bv_galois_keys(), other.generate_gsw_from_key())
    qs, ids = [], []
    for j, idx in enumerate([5, 4095, 777, 3000, 5]):
        c = s.client if j % 2 == 0 else other
        qs.append(c.fast_generate_query(s.p, idx))
        ids.append(c.client_id)
    batch = s.db.answer_batch(ids, np.stack(qs))
    for j in range(len(qs)):
        assert np.array_equal(batch[j], s.db.answer(ids[j], qs[j]))
    assert np.array_equal(batch[0], s.server.make_query(21, qs[0]))


def test_errors_map_to_reference_exceptions(comp_params, gpu_ctx):
    from onionpirv2_amd.gpu import GpuDatabase, InvalidArgument, OutOfRange
    say("[gpu] error paths")
    p = comp_params.with_layout(43, 5)
    db = GpuDatabase(gpu_ctx, p.fst_dim_sz, p.other_dim_sz, p.num_dims, p.expansion_height)
    q = np.zeros((2, p.N), np.uint64)
    with pytest.raises(OutOfRange):
        db.answer(987654, q)
    bad = np.full((2, p.N), p.rns_mods[0], np.uint64)
    with pytest.raises(InvalidArgument):
        db.answer(987654, bad)
    with pytest.raises(InvalidArgument):
        db.load_plain(np.zeros((p.num_pt + 1, p.N), np.uint16))


def _each_scan_shape(ctx, run):
    """run() under every first-dimension scan tile shape; shapes that cannot
    take the batch raise InvalidArgument. Returns {shape: result}."""
    from onionpirv2_amd.gpu import InvalidArgument
    out = {}
    try:
        for v in range(ctx.scan_variant_count()):
            ctx.set_scan_variant(v)
            try:
                out[v] = run()
            except InvalidArgument:
                continue
    finally:
        ctx.set_scan_variant(-1)
    return out


@pytest.mark.parametrize("layout", LAYOUTS[:4], ids=lambda l: "x".join(map(str, l)))
def test_every_scan_shape_matches_oracle(O, comp_params, gpu_ctx, layout):
    """Every tile shape of the first-dimension scan gives the oracle's
    evaluate_first_dim output (one query), and the same 9-query batch answers
    (18 query columns: one full and one partial column tile)."""
    say(f"[gpu] scan shapes layout={layout}")
    s = Setup(O, comp_params, layout, seed=3000 + layout[0], gpu_ctx=gpu_ctx)
    p = s.p
    query = s.client.fast_generate_query(p, s.indices[-1])
    fq = s.server.fast_expand_qry(s.cid, query)[: p.fst_dim_sz]
    mid_o = s.server.evaluate_first_dim(fq)
    mids = _each_scan_shape(gpu_ctx, lambda: s.db.first_dim(fq))
    assert len(mids) >= 4
    for v, m in mids.items():
        assert np.array_equal(m, mid_o), f"scan shape {v}"
    idxs = [(s.indices[-1] + 13 * j) % p.target_num_pt for j in range(9)]
    qs = np.stack([s.client.fast_generate_query(p, i) for i in idxs])
    ids = [s.cid] * len(idxs)
    want = s.db.answer_batch(ids, qs)
    assert np.array_equal(want[0], s.server.make_query(s.cid, qs[0]))
    got = _each_scan_shape(gpu_ctx, lambda: s.db.answer_batch(ids, qs))
    assert len(got) >= 4
    for v, b in got.items():
        assert np.array_equal(b, want), f"scan shape {v}"


def test_every_scan_shape_two_row_tiles_64_queries(O, comp_params, gpu_ctx):
    """256 x 32 layout: four full k blocks per level (N0 = 256), two row
    tiles, and a 64-query batch (128 query columns, eight full column tiles). Every shape gives the oracle's evaluate_first_dim output
    and the same batch answers; one answer is checked against the oracle."""
    say("[gpu] scan shapes, 256 x 32 layout, 64 queries")
    s = Setup(O, comp_params, (8192, 9, True), seed=4242, gpu_ctx=gpu_ctx, cid=42)
    p = s.p
    assert (p.fst_dim_sz, p.other_dim_sz) == (256, 32)
    query = s.client.fast_generate_query(p, s.indices[-1])
    fq = s.server.fast_expand_qry(s.cid, query)[: p.fst_dim_sz]
    mid_o = s.server.evaluate_first_dim(fq)
    mids = _each_scan_shape(gpu_ctx, lambda: s.db.first_dim(fq))
    assert len(mids) == gpu_ctx.scan_variant_count()
    for v, m in mids.items():
        assert np.array_equal(m, mid_o), f"scan shape {v}"
    idxs = [(29 * j + 5) % p.target_num_pt for j in range(64)]
    qs = np.stack([s.client.fast_generate_query(p, i) for i in idxs])
    ids = [s.cid] * len(idxs)
    want = s.db.answer_batch(ids, qs)
    assert np.array_equal(want[63], s.server.make_query(s.cid, qs[63]))
    got = _each_scan_shape(gpu_ctx, lambda: s.db.answer_batch(ids, qs))
    assert len(got) >= 6  # the shapes that stage 2 columns take one query only
    for v, b in got.items():
        assert np.array_equal(b, want), f"scan shape {v}"


def test_autotuned_scan_matches_default(O, comp_params, gpu_ctx):
    """Autotune mode times the applicable shapes once per batch-size class,
    keeps one, and answers exactly as the default shape does."""
    from onionpirv2_amd.gpu import SCAN_AUTOTUNE, SCAN_DEFAULT
    say("[gpu] autotuned scan")
    s = Setup(O, comp_params, (4096, 9, True), seed=5150, gpu_ctx=gpu_ctx, cid=51)
    p = s.p
    idxs = [(17 * j + 3) % p.target_num_pt for j in range(9)]
    qs = np.stack([s.client.fast_generate_query(p, i) for i in idxs])
    ids = [s.cid] * len(idxs)
    want1 = s.db.answer(s.cid, qs[0])
    want9 = s.db.answer_batch(ids, qs)
    assert s.db.tuned_scan_variant(1) == -1
    try:
        gpu_ctx.set_scan_variant(SCAN_AUTOTUNE)
        for _ in range(2):
            assert np.array_equal(s.db.answer(s.cid, qs[0]), want1)
            assert np.array_equal(s.db.answer_batch(ids, qs), want9)
    finally:
        gpu_ctx.set_scan_variant(SCAN_DEFAULT)
    n = gpu_ctx.scan_variant_count()
    assert 0 <= s.db.tuned_scan_variant(1) < n
    assert 0 <= s.db.tuned_scan_variant(9) < n
    assert np.array_equal(want1, s.server.make_query(s.cid, qs[0]))


def test_reserve_settles_autotuned_shapes(O, comp_params, gpu_ctx):
    """In autotune mode, opir_db_reserve(12) settles the scan shape of every
    batch-size class up to 12 queries before any answer; answers then equal
    the default shape's."""
    from onionpirv2_amd.gpu import SCAN_AUTOTUNE, SCAN_DEFAULT
    say("[gpu] reserve settles the autotuned scan shapes")
    s = Setup(O, comp_params, (4096, 9, True), seed=5252, gpu_ctx=gpu_ctx, cid=52)
    p = s.p
    idxs = [(19 * j + 7) % p.target_num_pt for j in range(12)]
    qs = np.stack([s.client.fast_generate_query(p, i) for i in idxs])
    ids = [s.cid] * len(idxs)
    want = {b: s.db.answer_batch(ids[:b], qs[:b]) for b in (1, 3, 12)}
    n = gpu_ctx.scan_variant_count()
    try:
        gpu_ctx.set_scan_variant(SCAN_AUTOTUNE)
        s.db.reserve(12)
        for b in (1, 2, 3, 4, 5, 8, 9, 12):
            assert 0 <= s.db.tuned_scan_variant(b) < n, b
        for b, w in want.items():
            assert np.array_equal(s.db.answer_batch(ids[:b], qs[:b]), w)
    finally:
        gpu_ctx.set_scan_variant(SCAN_DEFAULT)
    assert s.db.tuned_scan_variant(17) == -1  # 17..32 queries: a class above 12


def test_cu_mask_answers_match(O, comp_params, gpu_ctx):
    """Answers on streams restricted to 64 compute units, and after the
    restriction is lifted, equal the unrestricted answers."""
    from onionpirv2_amd.gpu import InvalidArgument
    say("[gpu] CU-masked context streams")
    s = Setup(O, comp_params, (700, 8, True), seed=6161, gpu_ctx=gpu_ctx, cid=61)
    p = s.p
    idxs = [(11 * j + 2) % p.target_num_pt for j in range(5)]
    qs = np.stack([s.client.fast_generate_query(p, i) for i in idxs])
    ids = [s.cid] * len(idxs)
    want = s.db.answer_batch(ids, qs)
    try:
        gpu_ctx.set_cu_mask(range(64))
        assert np.array_equal(s.db.answer_batch(ids, qs), want)
        assert np.array_equal(s.db.answer(s.cid, qs[1]), want[1])
        with pytest.raises(InvalidArgument):
            gpu_ctx.set_cu_mask([])  # selects no compute unit
    finally:
        gpu_ctx.set_cu_mask(None)
    assert np.array_equal(s.db.answer_batch(ids, qs), want)


def test_context_close_destroys_its_databases(comp_params):
    """Closing a context first destroys the databases still open on it
    (a database uses its context's stream), so a database that outlives its
    context, as one held by a failed test's traceback does, is never
    destroyed against a freed context."""
    from onionpirv2_amd.gpu import GpuContext, GpuDatabase
    say("[gpu] context close with a database still open")
    p = comp_params.with_layout(43, 5)
    ctx = GpuContext(scheme_from_oracle(comp_params))
    db = GpuDatabase(ctx, p.fst_dim_sz, p.other_dim_sz, p.num_dims, p.expansion_height)
    ctx.close()
    assert db.h is None and ctx.h is None
    db.close()


def test_graph_replays_match_eager(O, comp_params, gpu_ctx):
    """With graphs on, each answer's launch sequence is captured on first use
    (per batch size and key-table slot) and replayed after: every replay
    equals the eager answer, with eager calls in between, for single queries
    and for batches whose client order changes between replays (the key
    table is read at replay time)."""
    say("[gpu] graph replays")
    s = Setup(O, comp_params, (4096, 9, True), seed=6161, gpu_ctx=gpu_ctx, cid=61)
    p = s.p
    c2 = O.PirClient(p, 62, 6262)
    gpu_ctx.set_client_keys(62, c2.create_bv_galois_keys(), c2.generate_gsw_from_key())
    idxs = [(29 * j + 5) % p.target_num_pt for j in range(4)]
    qa = np.stack([s.client.fast_generate_query(p, i) for i in idxs])
    qb = np.stack([c2.fast_generate_query(p, i) for i in idxs])
    want_a = [s.db.answer(61, q) for q in qa]
    want_b = [s.db.answer(62, q) for q in qb]
    mix_ids = [61, 62, 61, 62]
    mix_q = np.stack([qa[0], qb[1], qa[2], qb[3]])
    want_mix = s.db.answer_batch(mix_ids, mix_q)
    assert np.array_equal(want_mix[1], want_b[1])
    try:
        gpu_ctx.set_graphs(True)
        for r in range(20):  # more calls than key-table slots: every slot's graph is replayed
            i = r % 4
            assert np.array_equal(s.db.answer(61, qa[i]), want_a[i]), (r, "a")
            assert np.array_equal(s.db.answer(62, qb[i]), want_b[i]), (r, "b")
            if r % 3 == 0:
                perm = [(k + r) % 4 for k in range(4)]
                got = s.db.answer_batch([mix_ids[k] for k in perm], mix_q[perm])
                assert np.array_equal(got, want_mix[perm]), (r, "mix")
            if r % 5 == 4:  # eager work between replays
                gpu_ctx.set_graphs(False)
                assert np.array_equal(s.db.answer(62, qb[i]), want_b[i]), (r, "eager")
                gpu_ctx.set_graphs(True)
    finally:
        gpu_ctx.set_graphs(False)
        gpu_ctx.remove_client(62)


def test_resident_key_bytes_track_clients(O, comp_params, gpu_ctx):
    """opir_ctx_resident_key_bytes (serving admission) counts each resident
    client's keys once, at its expansion height."""
    say("[gpu] resident key bytes")
    p = comp_params.with_layout(4096, 9, True)
    c = O.PirClient(p, 71, 7)
    bv, gsw = c.create_bv_galois_keys(), c.generate_gsw_from_key()
    before = gpu_ctx.resident_key_bytes()
    gpu_ctx.set_client_keys(9071, bv, gsw)
    gpu_ctx.set_client_keys(9072, bv, gsw)
    gpu_ctx.set_client_keys(9072, bv, gsw)  # replacing keeps one copy
    assert gpu_ctx.resident_key_bytes() - before == 2 * gpu_ctx.key_bytes(bv.shape[0])
    gpu_ctx.remove_client(9071)
    gpu_ctx.remove_client(9072)
    assert gpu_ctx.resident_key_bytes() == before


def test_loaders_reject_unreduced_values(comp_params, gpu_ctx):
    """server.cpp:129-137: a plaintext coefficient >= t is rejected with
    invalid_argument, by the Python wrapper before any narrowing and by the C
    ABI itself; a reference-layout residue >= its limb prime likewise."""
    import ctypes
    from onionpirv2_amd.gpu import GpuDatabase, InvalidArgument, OPIR_EINVAL
    say("[gpu] loader value checks")
    p = comp_params.with_layout(43, 5)
    db = GpuDatabase(gpu_ctx, p.fst_dim_sz, p.other_dim_sz, p.num_dims, p.expansion_height)
    pts = np.zeros((p.num_pt, p.N), np.uint16)
    pts[3, 7] = p.plain_mod
    with pytest.raises(InvalidArgument):
        db.load_plain(pts)
    with pytest.raises(InvalidArgument):
        db.load_plain(np.full((1, p.N), 1 << 16, np.uint32))  # would wrap to 0 as uint16
    rc = db.lib.opir_db_load_plain(db.h, ctypes.c_void_p(pts.ctypes.data), 0, p.num_pt)
    assert rc == OPIR_EINVAL and "mod t" in gpu_ctx.lib.opir_last_error(gpu_ctx.h).decode()
    n = p.N * p.num_pt
    lo = np.zeros(n, np.uint32)
    hi = np.zeros(n, np.uint32)
    lo[n - 1] = p.q1
    rc = db.lib.opir_db_load_ntt(db.h, ctypes.c_void_p(lo.ctypes.data), ctypes.c_void_p(hi.ctypes.data))
    assert rc == OPIR_EINVAL
    lo[n - 1] = 0
    hi[5] = p.q2
    with pytest.raises(InvalidArgument):
        db.load_ntt(lo, hi)
    db.close()


def test_mod_switch_small_composite_modulus_exact(O):
    """The final rescale on a composite modulus far below the reference's
    2^58 (two 20-bit limb primes): every output equals the exact centred
    rescale (utils.cpp:445-462, pinned against exact rationals)."""
    from onionpirv2_amd.gpu import GpuContext
    say("[gpu] mod switch, 40-bit composite modulus")
    N = 2048
    q1, q2 = O.generate_ntt_friendly_primes([20, 20], N)
    Q = q1 * q2
    root = O.crt_combine(O.minimal_primitive_root(2 * N, q1), q1, O.minimal_primitive_root(2 * N, q2), q2)
    sq = 12289
    sch = dict(poly_degree=N, composite=1, q1=q1, q2=q2, ntt_root=root, plain_mod=257, small_q=sq,
               l_ep=4, base_log2=10, l_key=4, base_log2_key=10, l_ks=4, bv_base_log2=10)
    ctx = GpuContext(sch)
    try:
        rng = np.random.default_rng(41)
        ct = rng.integers(0, Q, size=(2, N), dtype=np.uint64)
        ct[0, :7] = [0, 1, Q // 2 - 1, Q // 2, Q // 2 + 1, Q - 1, Q - 2]
        got = ctx.mod_switch(ct)
        want = np.array([[O.rescale(int(x), Q, sq) for x in row] for row in ct], np.uint64)
        assert np.array_equal(got, want)
    finally:
        ctx.close()


@pytest.mark.parametrize("k", [2049, 1025, 5, 3, 4095])
def test_bv_keyswitch_matches_oracle(O, comp_params, gpu_ctx, k):
    """bvks::bv_apply_galois_inplace (bv_keyswitch.cpp:296-367) alone through
    opir_bv_ks, against the oracle's bv_apply on a random ciphertext (every
    digit path, including the centring edges), bit-exact."""
    from onionpirv2_amd.gpu import InvalidArgument
    say(f"[gpu] BV keyswitch k={k}")
    p = comp_params
    Q = p.rns_mods[0]
    sk, _ = O.gen_sk_and_encrypt(p, np.zeros(p.N, np.uint64), 15)
    key = O.gen_bv_ks_key(p, sk, k, 16 + k)[0]
    rng = np.random.default_rng(k)
    ct = rng.integers(0, Q, size=(2, p.N), dtype=np.uint64)
    ct[1, :6] = [0, 1, Q // 2, Q // 2 + 1, Q - 1, Q // 2 - 1]
    want = O.bv_apply(p, ct, k, key[None])
    assert np.array_equal(gpu_ctx.bv_ks(ct, k, key), want)
    with pytest.raises(InvalidArgument):
        gpu_ctx.bv_ks(ct, 2, key)


def test_clone_and_view_answer_like_the_original(O, comp_params, gpu_ctx):
    """opir_ctx_clone + opir_db_view: the clone reads the parent's keys, the
    view shares the resident database; answers from two host threads, one on
    each, equal the original's; key setters on the clone are refused."""
    import threading
    from onionpirv2_amd.gpu import InvalidArgument
    say("[gpu] context clone and database view")
    s = Setup(O, comp_params, (700, 8, True), seed=7171, gpu_ctx=gpu_ctx, cid=71)
    p = s.p
    qs = np.stack([s.client.fast_generate_query(p, i) for i in (3, 99, 700 - 1, 350)])
    ids = [s.cid] * len(qs)
    want = s.db.answer_batch(ids, qs)
    clone = gpu_ctx.clone()
    view = s.db.view(clone)
    got = {}

    def run(name, db):
        got[name] = [db.answer_batch(ids, qs) for _ in range(5)]

    th = [threading.Thread(target=run, args=("orig", s.db)), threading.Thread(target=run, args=("view", view))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for name in ("orig", "view"):
        for b in got[name]:
            assert np.array_equal(b, want), name
    with pytest.raises(InvalidArgument):
        clone.set_client_keys(72, s.bv, s.gsw)
    view.close()
    clone.close()
    assert np.array_equal(s.db.answer(s.cid, qs[0]), want[0])
