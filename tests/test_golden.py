"""Committed golden answers (tests/golden/answer_*.npz, written by
tests/golden/make_golden.py with the CPU oracle).

CPU: the oracle and the product's C++ client still reproduce the fixtures
(queries from the client seed, responses and wire bytes from the seeded DB).
GPU: the product path -- host PirServer -> C ABI -> HIP kernels -- answers
each fixture query with the committed response, byte for byte.
"""
import glob
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "answer_*.npz")))


def load(path):
    d = np.load(path)  # plain arrays, allow_pickle stays False
    target, height, pow2, db_seed, cid, cseed, index = (int(x) for x in d["case"])
    cfg = str(d["config"]) if "config" in d.files else "CONFIG_N2048_K1_COMP"
    return dict(target=target, height=height, pow2=bool(pow2), db_seed=db_seed, cid=cid, cseed=cseed,
                index=index, query=d["query"], resp=d["resp"], wire=d["wire"].tobytes(), config=cfg)


def db_plaintexts(O, f, N, t):
    draws = O.mt19937_64_draws(f["db_seed"], f["target"] * N)
    return (draws % np.uint64(t)).astype(np.uint64)


def test_fixtures_present():
    assert len(FIXTURES) >= 6
    assert {load(f)["config"] for f in FIXTURES} == {"CONFIG_N2048_K1_COMP", "CONFIG_N2048_K1", "CONFIG_N2048_K2_MP",
                                                     "CONFIG_N4096_K2_MP"}


@pytest.mark.parametrize("path", FIXTURES, ids=os.path.basename)
def test_oracle_reproduces_fixture(O, path):
    f = load(path)
    p = O.PirParams(f["config"]).with_layout(f["target"], f["height"], f["pow2"])
    client = O.PirClient(p, f["cid"], f["cseed"])
    # the client's seeded stream runs keys first, then the query
    bv, gsw = client.create_bv_galois_keys(), client.generate_gsw_from_key()
    assert np.array_equal(client.fast_generate_query(p, f["index"]), f["query"])
    server = O.PirServer(p)
    server.load_data(f["target"], "random", f["db_seed"], [f["index"]])
    server.set_client_keys(f["cid"], bv, gsw)
    resp = server.make_query(f["cid"], f["query"].copy())
    assert np.array_equal(resp, f["resp"])
    assert server.save_resp_to_stream(resp) == f["wire"]


@pytest.mark.parametrize("path", FIXTURES, ids=os.path.basename)
def test_host_client_reproduces_fixture_query(path):
    import onionpirv2_amd.host as H
    f = load(path)
    hp = H.PirParams(f["config"]).with_layout(f["target"], f["height"], f["pow2"])
    c = H.PirClient(hp, f["cid"], f["cseed"])
    c.create_bv_galois_keys()
    c.generate_gsw_from_key()
    assert np.array_equal(c.fast_generate_query(hp, f["index"]), f["query"])


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIXTURES, ids=os.path.basename)
def test_gpu_answer_matches_fixture(O, path):
    import onionpirv2_amd.host as H
    f = load(path)
    print(f"[gpu] golden {os.path.basename(path)}", flush=True)
    hp = H.PirParams(f["config"]).with_layout(f["target"], f["height"], f["pow2"])
    server = H.PirServer(hp)
    server.load_data_array(db_plaintexts(O, f, hp.N, hp.plain_mod), [f["index"]])
    c = H.PirClient(hp, f["cid"], f["cseed"])
    server.set_client_bv_galois_key(f["cid"], c.create_bv_galois_keys())
    server.set_client_gsw_key(f["cid"], c.generate_gsw_from_key())
    resp = server.make_query(f["cid"], f["query"].copy())
    assert np.array_equal(resp, f["resp"])
    assert server.save_resp_to_stream(resp) == f["wire"]
    pt, budget = c.decrypt_mod_q(resp)
    assert np.array_equal(pt, server.direct_get_original_plaintext(f["index"])) and budget > 0
