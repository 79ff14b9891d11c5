"""The HIP answer path on the two largest workloads, at full size.

512 x 4096: CONFIG_N4096_K2_MP, with_layout({2^21, 10}) = 512 x 4096 plaintexts at
n = 4096 (128 GiB in NTT form, generated in HBM): recorded entries decrypt,
the first and last rows included, and a 32-query batch of distinct clients
equals the queries answered alone.
Serving: the serving loop over the three resident databases of the serving
workload (256 x 256 and 256 x 1024 at n = 2048, the 128 GiB one at n = 4096)
under Poisson arrivals: every served response equals the same query
answered alone.
Reference test: src/tests/test_pir.cpp:45-96.
"""
import sys
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def k2db():
    import onionpirv2_amd.host as H
    p = H.PirParams("CONFIG_N4096_K2_MP").with_layout(2 ** 21, 10, True)
    assert (p.fst_dim_sz, p.other_dim_sz, p.num_dims) == (512, 4096, 13)
    rec = [0, 511, 777777, p.target_num_pt - 512, p.target_num_pt - 1]
    t0 = time.time()
    say("[gpu] 512x4096: generating the 128 GiB database in HBM")
    srv = H.PirServer(p)
    srv.gen_data_device(rec, 31337)
    clients = []
    for j in range(32):
        c = H.PirClient(p, 2000 + j, 99 + j)
        srv.set_client_bv_galois_key(2000 + j, c.create_bv_galois_keys())
        srv.set_client_gsw_key(2000 + j, c.generate_gsw_from_key())
        clients.append(c)
    say(f"[gpu] 512x4096: database and 32 clients ready ({time.time() - t0:.1f} s)")
    yield p, srv, clients, rec
    del srv


def test_512x4096_recorded_entries_decrypt(k2db):
    p, srv, clients, rec = k2db
    for j, i in enumerate(rec):
        c = clients[j]
        say(f"[gpu] 512x4096: answer for plaintext {i}")
        r = srv.make_query(c.client_id, c.fast_generate_query(p, i))
        pt, budget = c.decrypt_mod_q(r)
        assert np.array_equal(pt, srv.direct_get_original_plaintext(i)) and budget > 0, i


def test_512x4096_batch_32_equals_single(k2db):
    p, srv, clients, rec = k2db
    rng = np.random.default_rng(3)
    idxs = [int(x) for x in rng.choice(p.target_num_pt, size=32, replace=False)]
    idxs[0], idxs[31] = rec[0], rec[-1]
    ids = [c.client_id for c in clients]
    qs = np.stack([clients[j].fast_generate_query(p, idxs[j]) for j in range(32)])
    say("[gpu] 512x4096: batch of 32")
    batch = srv.make_queries(ids, qs)
    say("[gpu] 512x4096: 32 single answers")
    for j in range(32):
        assert np.array_equal(batch[j], srv.make_query(ids[j], qs[j].copy())), j
    for j in (0, 31):
        pt, _ = clients[j].decrypt_mod_q(batch[j])
        assert np.array_equal(pt, srv.direct_get_original_plaintext(idxs[j]))


def test_three_databases_served_equal_alone(k2db):
    """Poisson arrivals over the three resident databases; 24 clients keyed
    on the n = 2048 contexts and 4 on the n = 4096 one; the pool follows the
    serving workload's mix (n = 4096 queries are a small share)."""
    import onionpirv2_amd.host as H
    p2, s2, c2s, rec2 = k2db
    layouts = [(65536, 9), (262144, 9)]
    servers, params, recs = [], [], []
    for j, (t, h) in enumerate(layouts):
        p = H.PirParams().with_layout(t, h, True)
        rec = [0, 4321, p.target_num_pt - 1]
        say(f"[gpu] serving: generating the {p.fst_dim_sz} x {p.other_dim_sz} database")
        s = H.PirServer(p)
        s.gen_data_device(rec, 600 + j)
        servers.append(s)
        params.append(p)
        recs.append(rec)
    clients = []
    for j in range(24):
        c = H.PirClient(params[0], 3000 + j, 555 + j)
        bv, gsw = c.create_bv_galois_keys(), c.generate_gsw_from_key()
        for s in servers:
            s.set_client_bv_galois_key(3000 + j, bv)
            s.set_client_gsw_key(3000 + j, gsw)
        clients.append(c)
    pool = []  # (db, client, index, query)
    for d in range(2):
        for k in range(24):
            idx = recs[d][k % 3] if k < 3 else (97 * k + 13 * d) % params[d].target_num_pt
            pool.append((d, 3000 + k, idx, clients[k].fast_generate_query(params[d], idx)))
    for k in range(2):
        c = c2s[k]
        pool.append((2, c.client_id, rec2[k], c.fast_generate_query(p2, rec2[k])))
    dbs = servers + [s2]
    say("[gpu] serving: answering the pool alone")
    alone = [dbs[d].make_query(cid, q.copy()) for d, cid, _, q in pool]
    loop = H.ServingLoop(dbs, max_batch=32, hbm_budget=96 << 30)
    try:
        say(f"[gpu] serving: Poisson arrivals, batch limits {list(loop.batch_limits)}")
        qs = [e[3] for e in pool]
        run = loop.run_poisson([e[0] for e in pool], [e[1] for e in pool], qs, 800.0, 0.6, 4711)
    finally:
        loop.stop()
    n = len(run["pool_index"])
    assert n > 200 and run["errors"] == 0
    per_db = [0, 0, 0]
    for i in range(n):
        k = run["pool_index"][i]
        assert np.array_equal(run["resp"][i], alone[k])
        per_db[pool[k][0]] += 1
    assert min(per_db) > 0, per_db
    for d in range(2):
        k = d * 24
        pt, budget = clients[0].decrypt_mod_q(alone[k])
        assert np.array_equal(pt, servers[d].direct_get_original_plaintext(pool[k][2])) and budget > 0
