"""Write the golden answer fixtures of tests/golden/ with the CPU oracle.

Each fixture is one seeded PIR exchange under one parameter set
(CONFIG_N2048_K1_COMP, CONFIG_N2048_K1, CONFIG_N2048_K2_MP or CONFIG_N4096_K2_MP):
  database   plaintext i, coefficient c = mt19937_64(db_seed) draw i*N + c mod t
             (the oracle's "random" loader, src/server.cpp:80-203 contract)
  client     PirClient(params, client_id, client_seed)
  query      fast_generate_query(params, index)      [2][K*N] u64 (mod Q, or per RNS limb)
  response   PirServer::make_query                   [2][N] u64 mod small_q
  wire       save_resp_to_stream(response)           ceil(2N log2(small_q) / 8) bytes

Run from the repository root after building the oracle (make -C oracle):
  python tests/golden/make_golden.py [fixture names to write; default all]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle", "_build"))

COMP = "CONFIG_N2048_K1_COMP"
# (target_num_pt, expansion_height, fst_dim_pow2, db_seed, client_id, client_seed, index[, config])
CASES = [
    (43, 5, True, 4301, 5, 91, 42),
    (700, 8, True, 7001, 6, 92, 3),
    (4096, 9, True, 40961, 7, 93, 2989),
    (43, 5, True, 4302, 8, 94, 40, "CONFIG_N2048_K2_MP"),
    (700, 8, True, 7002, 9, 95, 699, "CONFIG_N2048_K2_MP"),
    (43, 5, True, 4303, 10, 96, 17, "CONFIG_N4096_K2_MP"),
    (43, 5, True, 4304, 11, 97, 41, "CONFIG_N2048_K1"),
    (700, 8, True, 7004, 12, 98, 512, "CONFIG_N2048_K1"),
]


def config(case):
    return case[7] if len(case) > 7 else COMP


def name(case):
    cfg = "" if config(case) == COMP else config(case).lower().replace("config_", "") + "_"
    return f"answer_{cfg}{case[0]}x{case[1]}x{int(case[2])}.npz"


def make(O, case):
    target, height, pow2, db_seed, cid, cseed, index = case[:7]
    p = O.PirParams(config(case)).with_layout(target, height, pow2)
    client = O.PirClient(p, cid, cseed)
    server = O.PirServer(p)
    server.load_data(target, "random", db_seed, [index])
    server.set_client_keys(cid, client.create_bv_galois_keys(), client.generate_gsw_from_key())
    query = client.fast_generate_query(p, index)
    resp = server.make_query(cid, query)
    wire = np.frombuffer(server.save_resp_to_stream(resp), dtype=np.uint8)
    pt, _ = client.decrypt_mod_q(resp)
    assert np.array_equal(pt, server.direct_get_original_plaintext(index))
    return dict(case=np.array([target, height, int(pow2), db_seed, cid, cseed, index], dtype=np.uint64),
                config=np.array(config(case)),
                query=np.asarray(query, dtype=np.uint64), resp=np.asarray(resp, dtype=np.uint64), wire=wire)


def main():
    import onionpir_oracle as O
    only = set(sys.argv[1:])
    for case in CASES:
        if only and name(case) not in only:
            continue
        d = make(O, case)
        np.savez_compressed(os.path.join(HERE, name(case)), **d)
        print("wrote", name(case))


if __name__ == "__main__":
    main()
