"""The HIP answer path at the full size of the headline workload: the
256 x 1024 database at n = 2048 (CONFIG_N2048_K1_COMP, with_layout({262144, 9}),
4 GiB in NTT form), plain and batched, against the CPU oracle.

The oracle's database is filled with uniform NTT-domain residues (its own
load_data at this size runs minutes on one core) and handed to the GPU as is
through opir_db_load_ntt, the reference's composite [coeff][num_pt] layout
(src/server.cpp:162-170); the answer path's arithmetic does not depend on
whether the residues are transforms of small plaintexts. Decryption at this
size is checked with device-generated plaintexts (gen_data_device).
Reference test: src/tests/test_pir.cpp:45-96.
"""
import sys
import time

import numpy as np
import pytest

from conftest import scheme_from_oracle

pytestmark = pytest.mark.gpu

LAYOUT = (262144, 9, True)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def full(O):
    from onionpirv2_amd.gpu import GpuContext, GpuDatabase
    import onionpirv2_amd.host as H
    p = O.PirParams("CONFIG_N2048_K1_COMP").with_layout(*LAYOUT)
    assert (p.fst_dim_sz, p.other_dim_sz, p.num_dims) == (256, 1024, 11)
    t0 = time.time()
    say("[gpu] 256x1024: oracle database (uniform NTT residues, 4 GiB)")
    ref = O.PirServer(p)
    ref.fill_random_ntt(0x5EED)
    lo, hi = ref.db_lo(), ref.db_hi()
    say(f"[gpu] 256x1024: loading it through opir_db_load_ntt ({time.time() - t0:.1f} s)")
    ctx = GpuContext(scheme_from_oracle(p))
    db = GpuDatabase(ctx, p.fst_dim_sz, p.other_dim_sz, p.num_dims, p.expansion_height)
    db.load_ntt(lo, hi)
    del lo, hi
    hp = H.PirParams().with_layout(*LAYOUT)
    clients = []
    for j in range(128):
        c = H.PirClient(hp, 1000 + j, 4242 + j)
        bv, gsw = c.create_bv_galois_keys(), c.generate_gsw_from_key()
        ctx.set_client_keys(1000 + j, bv, gsw)
        clients.append((c, bv, gsw))
    say(f"[gpu] 256x1024: 128 clients keyed ({time.time() - t0:.1f} s)")
    yield p, hp, ref, ctx, db, clients
    db.close()
    ctx.close()


def _indices(n, num_pt, seed):
    rng = np.random.default_rng(seed)
    return [int(x) for x in rng.choice(num_pt, size=n, replace=False)]


def test_256x1024_plain_answer_bit_exact_vs_oracle(full):
    p, hp, ref, ctx, db, clients = full
    c, bv, gsw = clients[0]
    idx = p.target_num_pt - 1
    q = c.fast_generate_query(hp, idx)
    say("[gpu] 256x1024: plain answer, last plaintext")
    got = db.answer(1000, q)
    ref.set_client_keys(1000, bv, gsw)
    say("[gpu] 256x1024: oracle make_query")
    want = ref.make_query(1000, q)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("B", [64, 128])
def test_256x1024_batch_of_distinct_clients_equals_single(full, B):
    """The batched form: B queries of B distinct clients (their own
    key sets) in one fused pass; every response equals the query answered
    alone, and two equal the oracle's make_query."""
    p, hp, ref, ctx, db, clients = full
    idxs = _indices(B, p.target_num_pt, B)
    ids = [1000 + j for j in range(B)]
    qs = np.stack([clients[j][0].fast_generate_query(hp, idxs[j]) for j in range(B)])
    say(f"[gpu] 256x1024 batch: batch of {B}")
    batch = db.answer_batch(ids, qs)
    say(f"[gpu] 256x1024 batch: {B} single answers")
    for j in range(B):
        assert np.array_equal(batch[j], db.answer(ids[j], qs[j])), j
    for j in (1, B - 1):
        _, bv, gsw = clients[j]
        ref.set_client_keys(ids[j], bv, gsw)
        say(f"[gpu] 256x1024 batch: oracle make_query for client {ids[j]}")
        assert np.array_equal(batch[j], ref.make_query(ids[j], qs[j]))


def test_256x1024_device_generated_database_decrypts(full):
    """The same layout through the host PirServer with plaintexts generated
    on the device: recorded entries (first, last, one inside) decrypt."""
    import onionpirv2_amd.host as H
    p, hp, ref, ctx, db, clients = full
    rec = [0, 123457, hp.target_num_pt - 1]
    say("[gpu] 256x1024: gen_data_device on a second 4 GiB database")
    srv = H.PirServer(hp)
    srv.gen_data_device(rec, 777)
    c, bv, gsw = clients[5]
    srv.set_client_bv_galois_key(1005, bv)
    srv.set_client_gsw_key(1005, gsw)
    qs = np.stack([c.fast_generate_query(hp, i) for i in rec])
    resps = srv.make_queries([1005] * 3, qs)
    for i, r in zip(rec, resps):
        pt, budget = c.decrypt_mod_q(r)
        assert np.array_equal(pt, srv.direct_get_original_plaintext(i)) and budget > 0, i
    del srv
