"""Continuous-batching serving loop (onionpirv2_amd/host/serving.cpp).

CPU: the HBM-budget admission planner.
GPU: requests for two resident databases from several clients, submitted as
a burst and as Poisson arrivals, are answered in batches of whatever is
queued; every response equals the same query answered alone
(PirServer::make_query), bit for bit, and decrypts to the stored plaintext.
"""
import numpy as np
import pytest


def test_plan_batch_limits_budget_shares():
    import onionpirv2_amd.host as H
    # workspace of database d at batch B: 100 B + 10 d bytes
    ws = lambda d, B: 100 * B + 10 * d
    assert H.plan_batch_limits(2, ws, 50, 1050, 64) == [5, 4]
    assert H.plan_batch_limits(2, ws, 0, 0, 64) == [64, 64]          # no budget: max_batch
    assert H.plan_batch_limits(1, ws, 0, 10**9, 8) == [8]            # capped by max_batch
    # a workspace with steps (column tiles of 8 queries), like the scan fragments
    step = lambda d, B: 1000 * ((B + 7) // 8) + B
    assert H.plan_batch_limits(1, step, 0, 2000 + 16, 64) == [16]
    with pytest.raises(ValueError):
        H.plan_batch_limits(1, ws, 1000, 1050, 64)                   # keys leave no room for one query
    with pytest.raises(ValueError):
        H.plan_batch_limits(1, ws, 2000, 1000, 64)                   # keys alone exceed the budget
    with pytest.raises(ValueError):
        H.plan_batch_limits(0, ws, 0, 0, 64)


@pytest.fixture(scope="module")
def served():
    import onionpirv2_amd.host as H
    print("[gpu] serving: building two databases and 4 clients", flush=True)
    layouts = [(4096, 9), (700, 8)]
    params = [H.PirParams().with_layout(t, h, True) for t, h in layouts]
    servers = []
    for j, p in enumerate(params):
        s = H.PirServer(p)  # own GPU context and stream per database
        s.gen_data([5, 17, 300, 699], 900 + j)
        servers.append(s)
    clients = {}
    for c in range(4):
        cid = 500 + c
        cl = H.PirClient(params[0], cid, 77 + c)
        bv, gsw = cl.create_bv_galois_keys(), cl.generate_gsw_from_key()
        for s in servers:
            s.set_client_bv_galois_key(cid, bv)
            s.set_client_gsw_key(cid, gsw)
        clients[cid] = cl
    pool = []  # (db, client, index, query)
    for d, p in enumerate(params):
        for k, idx in enumerate([5, 17, 300, 699]):
            cid = 500 + (k + d) % 4
            pool.append((d, cid, idx, clients[cid].fast_generate_query(p, idx)))
    alone = [servers[d].make_query(cid, q.copy()) for d, cid, _, q in pool]
    yield H, servers, clients, pool, alone


@pytest.mark.gpu
def test_burst_batches_equal_alone(served):
    H, servers, clients, pool, alone = served
    print("[gpu] serving burst", flush=True)
    loop = H.ServingLoop(servers, max_batch=6)
    try:
        tickets = {}
        for rep in range(3):
            for k, (d, cid, _, q) in enumerate(pool):
                tickets[loop.submit(d, cid, q)] = k
        loop.wait_idle()
        res = loop.poll()
    finally:
        loop.stop()
    assert len(res) == len(tickets)
    assert max(r["batch"] for r in res) > 1, "the burst should have been batched"
    assert all(r["batch"] <= 6 for r in res)
    for r in res:
        assert r["error"] == ""
        assert np.array_equal(r["resp"], alone[tickets[r["ticket"]]])


@pytest.mark.gpu
def test_poisson_mix_equal_alone_and_decrypts(served):
    H, servers, clients, pool, alone = served
    print("[gpu] serving poisson", flush=True)
    from onionpirv2_amd.gpu import load_library  # noqa: F401  (the HIP library must be the one answering)
    # budget: keys of 4 clients on each context plus room for batches of a few queries
    loop = H.ServingLoop(servers, max_batch=8, hbm_budget=2 << 30, keyed_clients=4)
    try:
        assert all(1 <= b <= 8 for b in loop.batch_limits)
        dbs = [d for d, _, _, _ in pool]
        cids = [c for _, c, _, _ in pool]
        qs = np.stack([q for _, _, _, q in pool])
        run = loop.run_poisson(dbs, cids, qs, 2000.0, 0.4, 12345)
    finally:
        loop.stop()
    n = len(run["pool_index"])
    assert n > 100 and run["errors"] == 0
    for i in range(n):
        assert np.array_equal(run["resp"][i], alone[run["pool_index"][i]])
    for k in sorted(set(run["pool_index"].tolist()))[:4]:
        d, cid, idx, _ = pool[k]
        pt, budget = clients[cid].decrypt_mod_q(alone[k])
        assert np.array_equal(pt, servers[d].direct_get_original_plaintext(idx)) and budget > 0
    assert np.all(run["latency_ms"] >= run["queue_ms"])


@pytest.mark.gpu
def test_budget_too_small_is_rejected(served):
    H, servers, _, _, _ = served
    with pytest.raises(ValueError):
        H.ServingLoop(servers, max_batch=8, hbm_budget=1 << 20, keyed_clients=4)


@pytest.mark.gpu
def test_mixed_schemes_served_equal_alone(served):
    """An n=4096 two-prime database served beside the n=2048 ones: queries of
    both schemes in one Poisson mix, every response equal to the same query
    answered alone; admission counts the keys resident on each context."""
    H, servers, clients, pool, alone = served
    print("[gpu] serving mixed schemes", flush=True)
    p2 = H.PirParams("CONFIG_N4096_K2_MP").with_layout(1024, 8, True)
    s2 = H.PirServer(p2)
    s2.gen_data_device([3, 1000], 4242)
    c2 = H.PirClient(p2, 800, 31)
    s2.set_client_bv_galois_key(800, c2.create_bv_galois_keys())
    s2.set_client_gsw_key(800, c2.generate_gsw_from_key())
    q2 = [c2.fast_generate_query(p2, i) for i in (3, 1000)]
    a2 = [s2.make_query(800, q.copy()) for q in q2]
    loop = H.ServingLoop(servers + [s2], max_batch=8, hbm_budget=24 << 30)
    try:
        assert all(1 <= b <= 8 for b in loop.batch_limits)
        dbs = [d for d, _, _, _ in pool] + [2, 2]
        cids = [c for _, c, _, _ in pool] + [800, 800]
        qs = [q for _, _, _, q in pool] + q2
        run = loop.run_poisson(dbs, cids, qs, 600.0, 0.5, 777)
    finally:
        loop.stop()
    want = alone + a2
    n = len(run["pool_index"])
    assert n > 100 and run["errors"] == 0
    seen_k2 = 0
    for i in range(n):
        k = run["pool_index"][i]
        assert np.array_equal(run["resp"][i], want[k])
        seen_k2 += k >= len(pool)
    assert seen_k2 > 0
    pt, budget = c2.decrypt_mod_q(a2[1])
    assert np.array_equal(pt, s2.direct_get_original_plaintext(1000)) and budget > 0


@pytest.mark.gpu
def test_bad_request_fails_alone(served):
    """A request from an unkeyed client, batched with valid ones, fails on
    its own: every other request still equals its answer alone. A query with
    unreduced values is rejected at submit."""
    H, servers, clients, pool, alone = served
    print("[gpu] serving: one bad request in a burst", flush=True)
    loop = H.ServingLoop(servers, max_batch=8)
    try:
        tickets = {}
        bad = None
        for k, (d, cid, _, q) in enumerate(pool):
            tickets[loop.submit(d, cid, q)] = k
            if k == 3:
                bad = loop.submit(d, 987654, q)  # no keys for this client
        with pytest.raises(ValueError):
            q = pool[0][3].copy()
            q[0, 0] = np.uint64(1 << 62)
            loop.submit(0, pool[0][1], q)
        loop.wait_idle()
        res = loop.poll()
    finally:
        loop.stop()
    assert len(res) == len(tickets) + 1
    for r in res:
        if r["ticket"] == bad:
            assert "unknown client" in r["error"]
        else:
            assert r["error"] == "", r["error"]
            assert np.array_equal(r["resp"], alone[tickets[r["ticket"]]])
    assert max(r["batch"] for r in res) > 1


@pytest.mark.gpu
@pytest.mark.parametrize("lanes", [1, 3])
def test_lanes_overlap_batches_of_one_database(served, lanes):
    """With several answer lanes per database (context clones sharing the
    parent's keys, views sharing the resident database), consecutive batches
    of one database run at the same time on the device; every response still
    equals its answer alone."""
    H, servers, clients, pool, alone = served
    print(f"[gpu] serving with {lanes} lanes", flush=True)
    loop = H.ServingLoop(servers[:1], max_batch=2, lanes=lanes)
    try:
        tickets = {}
        own = [k for k, e in enumerate(pool) if e[0] == 0]
        for rep in range(6):
            for k in own:
                d, cid, _, q = pool[k]
                tickets[loop.submit(d, cid, q)] = k
        loop.wait_idle()
        res = loop.poll()
    finally:
        loop.stop()
    assert len(res) == len(tickets)
    for r in res:
        assert r["error"] == ""
        assert np.array_equal(r["resp"], alone[tickets[r["ticket"]]])
    spans = sorted({(r["t_start"], r["t_done"]) for r in res})
    overlapping = sum(1 for a, b in zip(spans, spans[1:]) if b[0] < a[1])
    if lanes == 1:
        assert overlapping == 0
    else:
        assert overlapping > 0, spans


def _ws(server, batch):
    """opir_db_workspace_bytes of a host PirServer's database."""
    import ctypes
    from onionpirv2_amd.gpu import load_library
    out = ctypes.c_size_t()
    assert load_library().opir_db_workspace_bytes(ctypes.c_void_p(server.db_handle()), batch, ctypes.byref(out)) == 0
    return out.value


def _resident_keys(server):
    import ctypes
    from onionpirv2_amd.gpu import load_library
    out = ctypes.c_size_t()
    assert load_library().opir_ctx_resident_key_bytes(ctypes.c_void_p(server.ctx_handle()), ctypes.byref(out)) == 0
    return out.value


@pytest.mark.gpu
def test_replanning_moves_budget_to_the_busy_database(served):
    """Admission re-planned as the load moves: with a budget whose equal
    share admits a few queries per database, an idle database trims to one
    query's workspace and a burst on the other grows past its share into the
    freed budget; a later burst on both reclaims the shares. Every response
    still equals its answer alone."""
    import time
    H, servers, clients, pool, alone = served
    print("[gpu] serving: re-planned admission", flush=True)
    keys = sum(_resident_keys(s) for s in servers)
    budget = keys + _ws(servers[0], 8) + _ws(servers[1], 1)
    loop = H.ServingLoop(servers, max_batch=8, hbm_budget=budget, lanes=1, shrink_after_s=0.1)
    try:
        share = loop.batch_limits
        assert share[0] < 8, share
        time.sleep(0.5)  # both lanes idle past shrink_after_s: one query's workspace each
        assert loop.reserved_batch(0) == 1 and loop.reserved_batch(1) == 1
        own = [k for k, e in enumerate(pool) if e[0] == 0]
        tickets = {}
        for rep in range(8):
            for k in own:
                d, cid, _, q = pool[k]
                tickets[loop.submit(d, cid, q)] = k
        loop.wait_idle()
        res = loop.poll()
        assert loop.peak_batch(0) > share[0], (loop.peak_batch(0), share)
        assert max(r["batch"] for r in res) > share[0]
        for r in res:
            assert r["error"] == ""
            assert np.array_equal(r["resp"], alone[tickets[r["ticket"]]])
        # both databases busy: the second reclaims its share from the first
        tickets = {}
        for rep in range(8):
            for k, (d, cid, _, q) in enumerate(pool):
                tickets[loop.submit(d, cid, q)] = k
        loop.wait_idle()
        res = loop.poll()
        for r in res:
            assert r["error"] == ""
            assert np.array_equal(r["resp"], alone[tickets[r["ticket"]]])
        assert max(r["batch"] for r in res if r["db"] == 1) > 1
        assert loop.grows > 0 and loop.trims > 0
        time.sleep(0.5)
        assert loop.reserved_batch(0) == 1 and loop.reserved_batch(1) == 1
    finally:
        loop.stop()
