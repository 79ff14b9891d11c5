#!/usr/bin/env python3
"""Benchmark of the MI355X OnionPIRv2 server answer path.

Workload (the reference's PirServer::make_query on CONFIG_N2048_K1_COMP):
  layout with_layout({262144, 9, true}) -> N0=256 x Nrest=1024, 11 dims,
  262144 plaintexts of 2048 coefficients < t (synthetic, seeded), 768 MiB of
  logical payload, 4 GiB NTT-form database resident in HBM.

A step is one answer pass over one batch: the plain form answers one query
per step (the headline `value`), the batched form answers --batch queries from
distinct clients per step. Queries/keys/database are resident in HBM before
timing; the timed region is bracketed by torch.cuda.synchronize().

  value     = DB GiB/s of the plain form = target_num_pt * 3072 B / s
              (the reference's throughput definition, src/tests/test_pir.cpp:92-93)
  roofline  = first-dimension scan kernel: algorithmic bytes / kernel time
  cpu_baseline = the CPU oracle (a plain single-thread restatement) running
              one full make_query of the same layout on the host cores.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# Each database context has two streams (answer path + selector completion);
# with HIP's default of 4 hardware queues, the streams of the serving leg's
# three databases would share queues and serialise behind each other. Set
# before the HIP runtime starts; an explicit setting in the environment wins.
os.environ.setdefault("GPU_MAX_HW_QUEUES", "16")

HBM_PEAK_GBS = 8000.0
INT8_MFMA_PEAK_TOPS = 5000.0  # dense int8 MFMA, MI355X


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--target", type=int, default=262144)
    ap.add_argument("--height", type=int, default=9)
    ap.add_argument("--no-cpu", action="store_true", help="skip the cpu_baseline leg")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--sweep", default="1,8,32,64,128",
                    help="batch sizes of the batch sweep (comma list; empty to skip)")
    ap.add_argument("--seed", type=int, default=20251142)
    ap.add_argument("--no-serve", action="store_true", help="skip the serving (continuous batching) leg")
    ap.add_argument("--serve-lambdas", default="10,50,200,1000,2000", help="Poisson arrival rates (1/s)")
    ap.add_argument("--serve-k2-cus", type=int, default=224,
                    help="serving: compute units the n=4096 database's context may use during the serving leg "
                         "(0 = all); the rest stay free for the n=2048 databases' short kernels")
    ap.add_argument("--serve-k2-entries", type=int, default=1,
                    help="0 leaves the n=4096 database out of the serving mix")
    ap.add_argument("--serve-seconds", type=float, default=2.0, help="arrival window per rate")
    ap.add_argument("--serve-clients", type=int, default=512,
                    help="seeded clients spread over the serving databases (the n=4096 one gets the k2 leg's)")
    ap.add_argument("--serve-replan", type=float, default=1.0,
                    help="serving: re-plan batch limits, trimming lanes idle this many seconds (< 0: fixed plan)")
    ap.add_argument("--serve-lanes", type=int, default=1,
                    help="concurrent answer lanes per database in the serving loop")
    ap.add_argument("--serve-budget-gib", type=float, default=96.0,
                    help="HBM budget for client keys and batch workspaces of the serving loop")
    ap.add_argument("--no-k2", action="store_true", help="skip the CONFIG_N4096_K2_MP leg")
    ap.add_argument("--tree-L", type=int, default=20, help="tree PIR leg: tree height (0: skip the leg)")
    ap.add_argument("--tree-a", type=int, default=4, help="tree PIR leg: first-dimension bits")
    ap.add_argument("--other-schemes", default="CONFIG_N2048_K1,CONFIG_N2048_K2_MP",
                    help="n=2048 parameter sets also measured on the headline layout (empty: none)")
    ap.add_argument("--no-graphs", action="store_true", help="skip the HIP-graph replay leg")
    ap.add_argument("--no-paper-row", action="store_true",
                    help="skip the with_layout({349526,10}) leg (the reference's 1 GB row, 5.3 GiB NTT form)")
    ap.add_argument("--k2-target", type=int, default=2**21, help="n=4096 layout: target plaintexts")
    ap.add_argument("--k2-height", type=int, default=10)
    ap.add_argument("--k2-batch", type=int, default=32)
    ap.add_argument("--k2-steps", type=int, default=5)
    ap.add_argument("--scan-shape", default="autotune",
                    help="first-dimension scan tile shape: autotune (timed per batch-size class), default "
                         "(measured table) or a shape index")
    ap.add_argument("--copy-probe-gib", type=float, default=0.0,
                    help="run a plain device copy of this many GiB first (calibrates rocprofv3 "
                         "FETCH_SIZE/WRITE_SIZE against known bytes)")
    return ap.parse_args()


def splitmix64(x):
    x = (x + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return x ^ (x >> 31)


def distinct_indices(n, num_pt, seed):
    out, s = [], seed
    while len(out) < n:
        s = splitmix64(s)
        v = s % num_pt
        if v not in out:
            out.append(v)
    return out


def scan_shape_arg(args):
    return {"autotune": -2, "default": -1}.get(args.scan_shape, None) if not args.scan_shape.isdigit() \
        else int(args.scan_shape)


def tuned_shape(G, db_h, b):
    import ctypes
    out = ctypes.c_int(-1)
    G.ctx_call_raw("opir_db_tuned_scan_variant", db_h, ctypes.c_size_t(b), ctypes.byref(out))
    return out.value


def batched_scan_leg(G, ctx_h, answer, B, params, log):
    """The first-dimension scan of a B-query batch on its own: HIP-event time
    of the scan launch (profiling slot 5), against both bounds it has at this
    size: HBM bytes (database + query columns + results) and the int8 MFMA
    work of the digit-plane products (16 x 16 x 64 tiles over the padded
    rows, columns and k)."""
    import ctypes
    G.ctx_call_raw("opir_ctx_set_profiling", ctx_h, 1)
    ms = []
    for _ in range(3):
        answer(B)
        out = (ctypes.c_float * 6)()
        assert G.ctx_call_raw("opir_last_stage_ms", ctx_h, out) == 0
        ms.append(out[5])
    G.ctx_call_raw("opir_ctx_set_profiling", ctx_h, 0)
    scan = sorted(ms)[1]
    N, N0, Nr = params.N, params.fst_dim_sz, params.other_dim_sz
    nbytes = params.num_pt * N * 8 + B * 2 * N0 * N * 8 + B * Nr * 2 * N * 8
    rows, cols, ks = -(-Nr // 16) * 16, -(-2 * B // 16) * 16, -(-N0 // 64) * 64
    mfma_ops = 2 * rows * cols * ks * 16 * 2 * N  # 4 x 4 digit planes, 2 limbs x N levels
    r = {"kernel_ms": round(scan, 4), "algorithmic_bytes": nbytes,
         "achieved_gbs": round(nbytes / scan / 1e6, 1), "frac_of_hbm_peak": round(nbytes / scan / 1e6 / HBM_PEAK_GBS, 4),
         "int8_mfma_ops": mfma_ops, "achieved_tops": round(mfma_ops / scan / 1e9, 1),
         "frac_of_int8_mfma_peak": round(mfma_ops / scan / 1e9 / INT8_MFMA_PEAK_TOPS, 4)}
    log(f"[batch] scan {scan:.3f} ms: {r['achieved_gbs']:.0f} GB/s ({r['frac_of_hbm_peak']:.2f} of HBM peak), "
        f"{r['achieved_tops']:.0f} int8 TOP/s ({r['frac_of_int8_mfma_peak']:.2f} of the dense MFMA peak)")
    return r


def pct(xs, p):
    xs = sorted(xs)
    k = (len(xs) - 1) * p / 100.0
    lo = int(k)
    hi = min(lo + 1, len(xs) - 1)
    return xs[lo] + (xs[hi] - xs[lo]) * (k - lo)


def committed_traffic(args):
    """HBM bytes per first-dimension scan launch from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE, calibrated on a 1 GiB copy), when the
    workload is the one it was measured on."""
    path = os.path.join(ROOT, "profiles", "k_scan_traffic_256x1024.md")
    if (args.target, args.height) != (262144, 9) or not os.path.exists(path):
        return None, None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(line.split(":")[1]), os.path.relpath(path, ROOT)
    return None, None


def graph_leg(args, G, ctx_h, answer, torch, B, d_r, server, queries, log):
    """The same plain and batched answers with the launch sequence captured
    as a HIP graph on first use and replayed (fixed buffers, so one graph per
    key-table slot); the eager numbers above are the headline."""
    import numpy as np
    assert G.ctx_call_raw("opir_ctx_set_graphs", ctx_h, 1) == 0
    out = {}
    try:
        for b, steps in ((1, args.steps), (B, max(3, args.steps // 4))):
            if b == B and args.no_batch:
                continue
            for _ in range(10):  # capture every key-table slot's graph
                answer(b)
            torch.cuda.synchronize()
            evs = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
            evs[0].record()
            for i in range(steps):
                answer(b)
                evs[i + 1].record()
            torch.cuda.synchronize()
            per = [evs[i].elapsed_time(evs[i + 1]) for i in range(steps)]
            ms = evs[0].elapsed_time(evs[-1]) / steps
            out[f"b{b}"] = {"ms_per_batch": round(ms, 4), "p50_ms": round(pct(per, 50), 4),
                            "p99_ms": round(pct(per, 99), 4), "steps": steps}
            log(f"[graphs] B={b}: {ms:.3f} ms/batch (captured launch sequence replayed)")
        rb = d_r.cpu().numpy().view(np.uint64)
        assert np.array_equal(rb[0], server.make_query(1000, queries[0].copy())), "graph replay != plain"
    finally:
        assert G.ctx_call_raw("opir_ctx_set_graphs", ctx_h, 0) == 0
    return out


def serving(args, H, params, server, log, k2=None):
    """Continuous batching over resident databases: the benchmark DB
    (256 x 1024, n=2048), a 256 x 256 one (1 GiB NTT form) and, unless
    --no-k2, the n=4096 two-prime DB (512 x 4096, 128 GiB NTT form).
    --serve-clients seeded clients are spread over the three: the n=4096 one
    gets the k2 leg's clients, the rest split evenly between the n=2048 ones,
    each keyed only on its database (keys generated on the device from the
    client seeds); every client has one query in the pool, and Poisson
    arrivals draw uniformly from the pool, so each database's share of the
    traffic is its share of the clients. Every response is checked against
    the same query answered alone."""
    import numpy as np
    p2 = H.PirParams().with_layout(65536, 9, True)
    s2 = H.PirServer(p2)
    s2.gen_data_device([], args.seed + 11)
    servers, plist = [server, s2], [params, p2]
    names = ["256x1024 n=2048 (4 GiB NTT form)", "256x256 n=2048 (1 GiB NTT form)"]
    nk2 = len(k2["queries"]) if (k2 is not None and args.serve_k2_entries > 0) else 0
    nc = args.serve_clients - nk2
    t0 = time.time()
    entries, ids, qs = [], [], []
    for d in range(2):
        cids = [5000 + 100000 * d + j for j in range(d, nc, 2)]
        keys = [(args.seed + 7919 * c).to_bytes(8, "little") * 4 for c in cids]
        servers[d].client_keygen_device(cids, keys)
        for c, k in zip(cids, keys):
            cl = H.PirClient(plist[d], c, k)
            idx = splitmix64(args.seed ^ c) % plist[d].target_num_pt
            entries.append(d)
            ids.append(c)
            qs.append(cl.fast_generate_query(plist[d], idx))
    k2_masked = None
    if nk2:
        if args.serve_k2_cus > 0:
            import ctypes
            from onionpirv2_amd import gpu as G
            words = [0] * ((args.serve_k2_cus + 31) // 32)
            for c in range(args.serve_k2_cus):
                words[c // 32] |= 1 << (c % 32)
            arr = (ctypes.c_uint32 * len(words))(*words)
            k2_masked = k2["server"].ctx_handle()
            assert G.ctx_call_raw("opir_ctx_set_cu_mask", k2_masked, arr, len(words)) == 0
        servers.append(k2["server"])
        names.append("512x4096 n=4096 K=2 (128 GiB NTT form)")
        for j in range(nk2):
            entries.append(2)
            ids.append(7000 + j)
            qs.append(k2["queries"][j])
    alone = [servers[d].make_query(cid, q.copy()) for d, cid, q in zip(entries, ids, qs)]
    log(f"[serve] {nc} n=2048 clients keyed on the device ({time.time() - t0:.1f}s) and {nk2} n=4096 clients; "
        f"{len(entries)}-query pool over {len(servers)} databases answered alone")
    # the n=4096 database's batches are bounded by its own measured batch size
    caps = [64] * (len(servers) - (1 if nk2 else 0)) + ([args.k2_batch] if nk2 else [])
    loop = H.ServingLoop(servers, max_batch=64, hbm_budget=int(args.serve_budget_gib * 2**30), keyed_clients=0,
                         lanes=args.serve_lanes, shrink_after_s=args.serve_replan, max_batch_db=caps)
    rows = []
    try:
        for lam in [float(x) for x in args.serve_lambdas.split(",") if x]:
            # queries of different schemes differ in length: the loop takes a list
            run = loop.run_poisson(entries, ids, qs, lam, args.serve_seconds, args.seed + int(lam))
            n = len(run["pool_index"])
            exact = sum(np.array_equal(run["resp"][i], alone[run["pool_index"][i]]) for i in range(n))
            lat = run["latency_ms"].tolist()
            per_db = {}
            for d in range(len(servers)):
                ld = [lat[i] for i in range(n) if entries[run["pool_index"][i]] == d]
                if ld:
                    per_db[names[d]] = {"requests": len(ld), "p50_ms": round(pct(ld, 50), 3),
                                        "p99_ms": round(pct(ld, 99), 3)}
            rows.append({"lambda_qps": lam, "requests": n, "achieved_qps": round(run["achieved_qps"], 1),
                         "p50_ms": round(pct(lat, 50), 3), "p99_ms": round(pct(lat, 99), 3),
                         "mean_batch": round(float(run["batch"].mean()), 2), "bit_exact_vs_alone": f"{exact}/{n}",
                         "errors": run["errors"], "per_database": per_db})
            log(f"[serve] lambda {lam:.0f}/s: {n} requests, {run['achieved_qps']:.0f} q/s, p50 {pct(lat, 50):.2f} ms "
                f"p99 {pct(lat, 99):.2f} ms, mean batch {run['batch'].mean():.1f}, exact {exact}/{n}, "
                f"per db {[(k.split()[0], v['requests'], v['p50_ms'], v['p99_ms']) for k, v in per_db.items()]}")
            assert exact == n and run["errors"] == 0, "served response differs from the query answered alone"
        limits = list(loop.batch_limits)
        replan = {"peak_batch": [loop.peak_batch(d) for d in range(len(servers))], "grows": loop.grows,
                  "trims": loop.trims} if args.serve_replan >= 0 else None
    finally:
        loop.stop()
        if k2_masked is not None:  # the n=4096 leg runs unrestricted
            assert G.ctx_call_raw("opir_ctx_set_cu_mask", k2_masked, None, 0) == 0
    return {"databases": names, "clients": {names[d]: entries.count(d) for d in range(len(servers))},
            "client_keys": "seeded clients; n=2048 keys generated on the device, n=4096 keys on the host",
            "lanes_per_database": args.serve_lanes,
            "replan_idle_s": args.serve_replan if args.serve_replan >= 0 else None,
            "n4096_cus": args.serve_k2_cus if k2_masked else None,
            "hbm_budget_gib": args.serve_budget_gib, "batch_limits": limits, "replanned": replan, "rates": rows}


def paper_row(args, H, G, torch, log):
    """The reference's 1 GB row: with_layout({349526, 10}) = 512 x 683
    plaintexts (1.0 GiB logical, 5.3 GiB NTT form, ragged MUX tree), data
    generated in HBM from a seed; plain path and one batch of 64."""
    import ctypes
    import numpy as np
    from concurrent.futures import ThreadPoolExecutor
    p = H.PirParams().with_layout(349526, 10, True)
    B = 64
    t0 = time.time()
    server = H.PirServer(p)
    idxs = distinct_indices(B, p.num_pt, args.seed + 5)
    idxs = [i % p.target_num_pt for i in idxs]
    server.gen_data_device(idxs[:2], args.seed + 5)

    def mk(j):
        c = H.PirClient(p, 9000 + j, args.seed + 17 * j + 3)
        return c, c.create_bv_galois_keys(), c.generate_gsw_from_key(), c.fast_generate_query(p, idxs[j])

    with ThreadPoolExecutor(8) as ex:
        cl = list(ex.map(mk, range(B)))
    for j, (c, bv, gsw, _) in enumerate(cl):
        server.set_client_bv_galois_key(9000 + j, bv)
        server.set_client_gsw_key(9000 + j, gsw)
    queries = np.stack([x[3] for x in cl])
    log(f"[1gb] layout N0={p.fst_dim_sz} Nrest={p.other_dim_sz} dims={p.num_dims}, {B} clients, "
        f"setup {time.time() - t0:.1f}s")
    for j in range(2):
        pt, budget = cl[j][0].decrypt_mod_q(server.make_query(9000 + j, queries[j].copy()))
        assert np.array_equal(pt, server.direct_get_original_plaintext(idxs[j])), "wrong 1 GB-row PIR answer"
    db_h, ctx_h = server.db_handle(), server.ctx_handle()
    stream = torch.cuda.current_stream()
    assert G.ctx_call_raw("opir_ctx_set_stream", ctx_h, ctypes.c_void_p(stream.cuda_stream)) == 0
    assert G.ctx_call_raw("opir_ctx_set_scan_variant", ctx_h, scan_shape_arg(args)) == 0
    dev = torch.device("cuda:0")
    d_q = torch.from_numpy(queries.view(np.int64)).to(dev)
    d_r = torch.empty_like(d_q)
    ids = np.arange(9000, 9000 + B, dtype=np.uint64)

    def answer(b, off=0):
        rc = G.answer_batch_device_raw(db_h, ids[off:off + b], d_q[off:off + b].data_ptr(), d_r[off:off + b].data_ptr())
        if rc:
            raise RuntimeError(G.last_error_raw(ctx_h))

    def timed(b, steps):
        for i in range(2):
            answer(b)
        torch.cuda.synchronize()
        evs = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
        evs[0].record()
        for i in range(steps):
            answer(b, (i * b) % B if b < B else 0)
            evs[i + 1].record()
        torch.cuda.synchronize()
        per = [evs[i].elapsed_time(evs[i + 1]) for i in range(steps)]
        return sum(per) / steps, per

    ms_plain, per = timed(1, args.steps)
    ms_b, per_b = timed(B, 4)
    answer(B)
    torch.cuda.synchronize()
    rb = d_r.cpu().numpy().view(np.uint64)
    assert np.array_equal(rb[B - 1], server.make_query(9000 + B - 1, queries[B - 1].copy())), "1 GB row batch != plain"
    gib = p.target_num_pt * p.pt_size / 2**30
    log(f"[1gb] plain {ms_plain:.3f} ms -> {gib / (ms_plain * 1e-3):.0f} DB GiB/s; batch {B}: {ms_b:.2f} ms -> "
        f"{B * 1e3 / ms_b:.0f} q/s ({B * ms_plain / ms_b:.1f}x plain)")
    res = {"workload": "OnionPIRv2 make_query, CONFIG_N2048_K1_COMP, with_layout({349526,10}) = 512 x 683 "
                       "plaintexts (1.0 GiB logical, 5.3 GiB NTT form), synthetic seeded data generated in HBM",
           "fst_dim_sz": p.fst_dim_sz, "other_dim_sz": p.other_dim_sz, "num_dims": p.num_dims,
           "db_gib_per_s_plain": round(gib / (ms_plain * 1e-3), 2), "ms_per_query_plain": round(ms_plain, 4),
           "queries_per_s_plain": round(1e3 / ms_plain, 2),
           "latency_ms_plain": {"p50": round(pct(per, 50), 4), "p99": round(pct(per, 99), 4)},
           "batched": {"batch": B, "ms_per_batch": round(ms_b, 3), "queries_per_s": round(B * 1e3 / ms_b, 2),
                       "speedup_vs_plain_qps": round(B * ms_plain / ms_b, 3)},
           "scan_shape_plain": tuned_shape(G, db_h, 1), "scan_shape_batched": tuned_shape(G, db_h, B)}
    del d_q, d_r, server
    torch.cuda.synchronize()
    return res


def k2_setup(args, H, log):
    """The n=4096 two-prime database (with_layout({2^21, 10}) = 512 x 4096
    plaintexts, 128 GiB in NTT form, generated in HBM from a seed) and
    k2_batch keyed clients with one query each."""
    import numpy as np
    from concurrent.futures import ThreadPoolExecutor
    p = H.PirParams("CONFIG_N4096_K2_MP").with_layout(args.k2_target, args.k2_height, True)
    B = args.k2_batch
    log(f"[k2] layout N0={p.fst_dim_sz} Nrest={p.other_dim_sz} dims={p.num_dims} num_pt={p.num_pt} "
        f"({p.num_pt * p.N * 2 * 8 / 2**30:.0f} GiB NTT form)")
    t0 = time.time()
    server = H.PirServer(p)
    idxs = distinct_indices(max(B, 2), p.num_pt, args.seed + 3)
    server.gen_data_device(idxs[:2], args.seed + 3)
    log(f"[k2] database generated on the device in {time.time() - t0:.1f}s")

    def mk(j):
        c = H.PirClient(p, 7000 + j, args.seed + 31 * j + 1)
        bv, gsw = c.create_bv_galois_keys(), c.generate_gsw_from_key()
        return c, bv, gsw, c.fast_generate_query(p, idxs[j])

    t0 = time.time()
    with ThreadPoolExecutor(8) as ex:
        cl = list(ex.map(mk, range(max(B, 2))))
    for j, (c, bv, gsw, _) in enumerate(cl):
        server.set_client_bv_galois_key(7000 + j, bv)
        server.set_client_gsw_key(7000 + j, gsw)
    queries = np.stack([x[3] for x in cl])
    log(f"[k2] {len(cl)} clients keyed, queries built in {time.time() - t0:.1f}s")
    return {"params": p, "server": server, "clients": [x[0] for x in cl], "queries": queries, "idxs": idxs}


def k2_leg(args, H, G, torch, log, k2):
    """CONFIG_N4096_K2_MP (two ~60-bit RNS limbs, N = 4096) on the layout
    with_layout({2^21, 10}) = 512 x 4096 plaintexts, 128 GiB in NTT form:
    plain single-query path and one fused batch."""
    import ctypes
    import numpy as np
    p, server, queries, idxs = k2["params"], k2["server"], k2["queries"], k2["idxs"]
    cl = [(c,) for c in k2["clients"]]
    N, B = p.N, args.k2_batch
    for j in range(2):
        r = server.make_query(7000 + j, queries[j].copy())
        pt, budget = cl[j][0].decrypt_mod_q(r)
        assert np.array_equal(pt, server.direct_get_original_plaintext(idxs[j])), "wrong n=4096 PIR answer"
        log(f"[k2] query {j} decrypts to plaintext {idxs[j]}, noise budget {budget}")
    db_h, ctx_h = server.db_handle(), server.ctx_handle()
    stream = torch.cuda.current_stream()
    assert G.ctx_call_raw("opir_ctx_set_stream", ctx_h, ctypes.c_void_p(stream.cuda_stream)) == 0
    assert G.ctx_call_raw("opir_ctx_set_scan_variant", ctx_h, scan_shape_arg(args)) == 0
    dev = torch.device("cuda:0")
    d_q = torch.from_numpy(queries.view(np.int64)).to(dev)
    d_r = torch.empty((B, 2, N), dtype=torch.int64, device=dev)
    ids = np.arange(7000, 7000 + B, dtype=np.uint64)

    def answer(b, off=0):
        rc = G.answer_batch_device_raw(db_h, ids[off:off + b], d_q[off:off + b].data_ptr(), d_r[off:off + b].data_ptr())
        if rc:
            raise RuntimeError(G.last_error_raw(ctx_h))

    def timed(b, steps):
        answer(b)
        torch.cuda.synchronize()
        evs = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
        evs[0].record()
        for i in range(steps):
            answer(b, (i * b) % B if b < B else 0)
            evs[i + 1].record()
        torch.cuda.synchronize()
        per = [evs[i].elapsed_time(evs[i + 1]) for i in range(steps)]
        return sum(per) / steps, per

    ms_plain, per = timed(1, args.k2_steps)
    G.ctx_call_raw("opir_ctx_set_profiling", ctx_h, 1)
    answer(1)
    out = (ctypes.c_float * 6)()
    assert G.ctx_call_raw("opir_last_stage_ms", ctx_h, out) == 0
    G.ctx_call_raw("opir_ctx_set_profiling", ctx_h, 0)
    names = ["Expand", "Convert", "First dim", "Other dim", "Modulus switching"]
    stages = {k: round(v, 3) for k, v in zip(names, out[:5])}
    scan_ms = out[5]
    scan_bytes = p.num_pt * N * 2 * 8 + p.fst_dim_sz * 2 * 2 * N * 8 + p.other_dim_sz * 2 * 2 * N * 8
    gib = p.target_num_pt * p.pt_size / 2**30
    log(f"[k2] plain {ms_plain:.2f} ms/query (p50 {pct(per, 50):.2f} p99 {pct(per, 99):.2f}), stages {stages}, "
        f"scan {scan_ms:.2f} ms -> {scan_bytes / scan_ms / 1e6:.0f} GB/s")
    res = {"workload": "OnionPIRv2 make_query, CONFIG_N4096_K2_MP, with_layout({2097152,10}) = 512 x 4096 "
                       "plaintexts (39 GiB logical, 128 GiB NTT form), synthetic seeded data generated in HBM",
           "fst_dim_sz": p.fst_dim_sz, "other_dim_sz": p.other_dim_sz, "num_dims": p.num_dims,
           "db_gib_per_s_plain": round(gib / (ms_plain * 1e-3), 2), "ms_per_query_plain": round(ms_plain, 3),
           "queries_per_s_plain": round(1e3 / ms_plain, 2),
           "latency_ms_plain": {"p50": round(pct(per, 50), 3), "p99": round(pct(per, 99), 3)},
           "stage_ms_plain": stages,
           "scan": {"kernel": "k_scan_direct / k_scan_pp / k_scan (first-dimension scan, int8 MFMA over 8 digit planes)", "kernel_ms": round(scan_ms, 3),
                    "algorithmic_bytes": scan_bytes, "achieved_gbs": round(scan_bytes / scan_ms / 1e6, 1),
                    "frac_of_hbm_peak": round(scan_bytes / scan_ms / 1e6 / HBM_PEAK_GBS, 4)}}
    if B > 1:
        ms_b, per_b = timed(B, 2)
        res["batched"] = {"batch": B, "ms_per_batch": round(ms_b, 2), "queries_per_s": round(B * 1e3 / ms_b, 2),
                          "speedup_vs_plain_qps": round(B * ms_plain / ms_b, 3)}
        log(f"[k2] batch {B}: {ms_b:.1f} ms -> {B * 1e3 / ms_b:.1f} q/s")
        answer(B)
        torch.cuda.synchronize()
        rb = d_r.cpu().numpy().view(np.uint64)
        for j in (0, B - 1):
            assert np.array_equal(rb[j], server.make_query(7000 + j, queries[j].copy())), "n=4096 batch != plain"
    del d_q, d_r
    torch.cuda.synchronize()
    return res


def scheme_leg(args, H, G, torch, log, cfg):
    """Another n=2048 parameter set (CONFIG_N2048_K1: one ~60-bit prime, u64
    kernels; CONFIG_N2048_K2_MP: two 29-bit RNS limbs, MP gadgets) on the
    headline layout with_layout({262144, 9}) = 256 x 1024 plaintexts, data
    generated in HBM from a seed: plain single-query path with its stage
    split and scan rate, and one fused batch of 64."""
    import ctypes
    import numpy as np
    from concurrent.futures import ThreadPoolExecutor
    p = H.PirParams(cfg).with_layout(args.target, args.height, True)
    B = 64
    K = len(p.rns_mods)
    t0 = time.time()
    server = H.PirServer(p)
    idxs = distinct_indices(B, p.num_pt, args.seed + 11)
    idxs = [i % p.target_num_pt for i in idxs]
    server.gen_data_device(idxs[:2], args.seed + 11)

    def mk(j):
        c = H.PirClient(p, 11000 + j, args.seed + 13 * j + 5)
        return c, c.create_bv_galois_keys(), c.generate_gsw_from_key(), c.fast_generate_query(p, idxs[j])

    with ThreadPoolExecutor(8) as ex:
        cl = list(ex.map(mk, range(B)))
    for j, (c, bv, gsw, _) in enumerate(cl):
        server.set_client_bv_galois_key(11000 + j, bv)
        server.set_client_gsw_key(11000 + j, gsw)
    queries = np.stack([x[3] for x in cl])
    tag = cfg.lower().replace("config_", "")
    log(f"[{tag}] layout N0={p.fst_dim_sz} Nrest={p.other_dim_sz} dims={p.num_dims}, {K} RNS limb(s), {B} clients, "
        f"setup {time.time() - t0:.1f}s")
    for j in range(2):
        pt, budget = cl[j][0].decrypt_mod_q(server.make_query(11000 + j, queries[j].copy()))
        assert np.array_equal(pt, server.direct_get_original_plaintext(idxs[j])), f"wrong {cfg} PIR answer"
    db_h, ctx_h = server.db_handle(), server.ctx_handle()
    stream = torch.cuda.current_stream()
    assert G.ctx_call_raw("opir_ctx_set_stream", ctx_h, ctypes.c_void_p(stream.cuda_stream)) == 0
    assert G.ctx_call_raw("opir_ctx_set_scan_variant", ctx_h, scan_shape_arg(args)) == 0
    dev = torch.device("cuda:0")
    d_q = torch.from_numpy(queries.view(np.int64)).to(dev)
    d_r = torch.empty((B, 2, p.N), dtype=torch.int64, device=dev)
    ids = np.arange(11000, 11000 + B, dtype=np.uint64)

    def answer(b, off=0):
        rc = G.answer_batch_device_raw(db_h, ids[off:off + b], d_q[off:off + b].data_ptr(), d_r[off:off + b].data_ptr())
        if rc:
            raise RuntimeError(G.last_error_raw(ctx_h))

    def timed(b, steps):
        for i in range(2):
            answer(b)
        torch.cuda.synchronize()
        evs = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
        evs[0].record()
        for i in range(steps):
            answer(b, (i * b) % B if b < B else 0)
            evs[i + 1].record()
        torch.cuda.synchronize()
        per = [evs[i].elapsed_time(evs[i + 1]) for i in range(steps)]
        return sum(per) / steps, per

    ms_plain, per = timed(1, args.steps)
    G.ctx_call_raw("opir_ctx_set_profiling", ctx_h, 1)
    acc = []
    for i in range(5):
        answer(1, i)
        out = (ctypes.c_float * 6)()
        assert G.ctx_call_raw("opir_last_stage_ms", ctx_h, out) == 0
        acc.append(list(out))
    G.ctx_call_raw("opir_ctx_set_profiling", ctx_h, 0)
    names = ["Expand", "Convert", "First dim", "Other dim", "Modulus switching"]
    stages = {k: round(sum(a[i] for a in acc) / len(acc), 4) for i, k in enumerate(names)}
    scan_ms = sum(a[5] for a in acc) / len(acc)
    # DB (K limbs of 8 B for the 60-bit prime, of 4 B for 29-bit limbs) + query columns + results
    wb = 8 if p.rns_mods[0] >= 1 << 32 else 4
    scan_bytes = p.num_pt * p.N * K * wb + p.fst_dim_sz * 2 * K * p.N * wb + p.other_dim_sz * 2 * K * p.N * wb
    ms_b, per_b = timed(B, 4)
    answer(B)
    torch.cuda.synchronize()
    rb = d_r.cpu().numpy().view(np.uint64)
    for j in (0, B - 1):
        assert np.array_equal(rb[j], server.make_query(11000 + j, queries[j].copy())), f"{cfg} batch != plain"
    gib = p.target_num_pt * p.pt_size / 2**30
    log(f"[{tag}] plain {ms_plain:.3f} ms -> {gib / (ms_plain * 1e-3):.0f} DB GiB/s, stages {stages}, scan "
        f"{scan_ms:.3f} ms -> {scan_bytes / scan_ms / 1e6:.0f} GB/s; batch {B}: {ms_b:.2f} ms -> "
        f"{B * 1e3 / ms_b:.0f} q/s ({B * ms_plain / ms_b:.1f}x plain)")
    res = {"workload": f"OnionPIRv2 make_query, {cfg}, with_layout({{{args.target},{args.height}}}) = "
                       f"{p.fst_dim_sz} x {p.other_dim_sz} plaintexts, synthetic seeded data generated in HBM",
           "rns_limbs": K, "plain_mod": p.plain_mod,
           "db_gib_per_s_plain": round(gib / (ms_plain * 1e-3), 2), "ms_per_query_plain": round(ms_plain, 4),
           "queries_per_s_plain": round(1e3 / ms_plain, 2),
           "latency_ms_plain": {"p50": round(pct(per, 50), 4), "p99": round(pct(per, 99), 4)},
           "stage_ms_plain": stages,
           "scan": {"kernel_ms": round(scan_ms, 4), "algorithmic_bytes": scan_bytes,
                    "achieved_gbs": round(scan_bytes / scan_ms / 1e6, 1),
                    "frac_of_hbm_peak": round(scan_bytes / scan_ms / 1e6 / HBM_PEAK_GBS, 4)},
           "batched": {"batch": B, "ms_per_batch": round(ms_b, 3), "queries_per_s": round(B * 1e3 / ms_b, 2),
                       "speedup_vs_plain_qps": round(B * ms_plain / ms_b, 3)}}
    del d_q, d_r, server
    torch.cuda.synchronize()
    return res


def tree_leg(args, H, log):
    """Tree PIR on the answer path's kernels (the reference's tree layer,
    src/tree_query.cpp:198-240, src/tree_select.cpp:121-238): a tree of 2^L
    leaves (CONFIG_N2048_K1_COMP, first dimension 2^a), every level's packed
    plaintexts resident on the device; one query selects one node on every
    level. Timed through the host TreePirServer (query upload and the L + 1
    level ciphertexts' download included); parity is in tests/test_gpu_tree.py."""
    import numpy as np
    hs = H.PirParams()
    L, a = args.tree_L, args.tree_a
    ht = H.make_tree_pir_params_for_scheme(L, a, hs)
    t0 = time.time()
    srv = H.TreePirServer(hs, ht)
    srv.preprocess_synthetic()
    hq = H.tree_query_expansion_params(ht, hs)
    c = H.PirClient(hq, 8, args.seed + 99)
    srv.set_client_keys(8, c.create_bv_galois_keys(), c.generate_gsw_from_key())
    leaves = [(args.seed * (j + 1) * 2654435761) % ht.N for j in range(8)]
    qs = [H.make_tree_query(c, hs, ht, leaf) for leaf in leaves]
    setup = time.time() - t0
    first = srv.select_path(8, qs[0])
    srv.select_path(8, qs[1])
    per = []
    for j in range(max(10, args.steps)):
        t1 = time.perf_counter()
        srv.select_path(8, qs[j % len(qs)])
        per.append((time.perf_counter() - t1) * 1e3)
    assert np.array_equal(srv.select_path(8, qs[0]), first), "tree select is not deterministic"
    npt = sum(1 << max(lv - (ht.n.bit_length() - 1), 0) for lv in range(L + 1))
    ms = sum(per) / len(per)
    log(f"[tree] L={L} a={a}: {1 << L} leaves, {npt} packed plaintexts over {L + 1} levels (setup {setup:.1f}s); "
        f"select {ms:.3f} ms per query (p50 {pct(per, 50):.3f} p99 {pct(per, 99):.3f})")
    return {"workload": f"tree PIR, CONFIG_N2048_K1_COMP, 2^{L} leaves, first dimension 2^{a}, synthetic node "
                        f"values; one query selects a node on each of the {L + 1} levels",
            "packed_plaintexts": npt, "packed_width": ht.w, "query_expansion_height": ht.h_q,
            "ms_per_query": round(ms, 4), "queries_per_s": round(1e3 / ms, 2),
            "latency_ms": {"p50": round(pct(per, 50), 4), "p99": round(pct(per, 99), 4)},
            "timing": "host TreePirServer::select_path, query upload and level downloads included"}


def host_cpu():
    """CPU model (as lscpu reports it, from /proc/cpuinfo) and the cores this
    process may use (nproc)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        nproc = len(os.sched_getaffinity(0))
    except AttributeError:
        nproc = os.cpu_count()
    return model, nproc


def cpu_baseline(args, log, warmup=3, trials=5):
    """The oracle's make_query on the same layout on the host: `warmup`
    untimed queries, then the median of `trials` timed ones (one thread)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle", "_build"))
    import numpy as np
    import onionpir_oracle as O
    p = O.PirParams("CONFIG_N2048_K1_COMP").with_layout(args.target, args.height, True)
    c = O.PirClient(p, 1, args.seed)
    bv, g = c.create_bv_galois_keys(), c.generate_gsw_from_key()
    s = O.PirServer(p)
    s.fill_random_ntt(args.seed)
    s.set_client_keys(1, bv, g)
    times, stages = [], []
    for k in range(warmup + trials):
        q = c.fast_generate_query(p, (12345 + 7919 * k) % p.num_pt)
        t0 = time.perf_counter()
        s.make_query(1, q)
        dt = time.perf_counter() - t0
        if k >= warmup:
            times.append(dt)
            stages.append(s.last_stage_ms())
        log(f"[cpu] oracle make_query {'warmup' if k < warmup else 'trial'} {dt:.2f}s")
    med = float(np.median(times))
    mid = stages[int(np.argsort(times)[len(times) // 2])]
    gib = p.target_num_pt * p.pt_size / 2**30
    model, nproc = host_cpu()
    del s
    return {"value": round(gib / med, 4), "unit": "DB GiB/s", "cores": 1, "kind": "port",
            "sample": f"{trials} timed full make_query calls (after {warmup} untimed) of the same layout "
                      f"({p.fst_dim_sz}x{p.other_dim_sz}, {p.num_dims} dims) by the single-thread CPU oracle, "
                      f"an unoptimized restatement of the reference's algorithm (plain C++, no HEXL/AVX-512), "
                      f"database filled with uniform NTT-domain residues (cost is value-independent); "
                      f"median {med:.2f} s",
            "trial_s": [round(t, 3) for t in times],
            "median_s": round(med, 3), "min_s": round(min(times), 3), "max_s": round(max(times), 3),
            "host_cpu": model, "nproc": nproc,
            "stage_ms": {k: round(v, 1) for k, v in mid.items()}}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 (across GPUs the "
                                   "path shards by replicating the database, one process per GPU)"}))
        sys.exit(2)
    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    import numpy as np
    import torch
    import onionpirv2_amd.host as H
    from onionpirv2_amd import gpu as G

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    params = H.PirParams().with_layout(args.target, args.height, True)
    N, B = params.N, args.batch
    sweep = [int(x) for x in args.sweep.split(",") if x] if not args.no_batch else []
    nclients = max([B] + sweep)
    log(f"[setup] layout N0={params.fst_dim_sz} Nrest={params.other_dim_sz} dims={params.num_dims} "
        f"num_pt={params.num_pt} h={params.expansion_height}")
    t0 = time.time()
    server = H.PirServer(params)
    idxs = distinct_indices(max(nclients, 8), params.num_pt, args.seed)
    server.gen_data(idxs[:8], args.seed)
    log(f"[setup] database generated and loaded in {time.time() - t0:.1f}s")
    t0 = time.time()
    clients = []
    for j in range(nclients):
        c = H.PirClient(params, 1000 + j, args.seed + 7 * j)
        server.set_client_bv_galois_key(1000 + j, c.create_bv_galois_keys())
        server.set_client_gsw_key(1000 + j, c.generate_gsw_from_key())
        clients.append(c)
    queries = np.stack([clients[j].fast_generate_query(params, idxs[j]) for j in range(nclients)])
    log(f"[setup] {nclients} clients keyed, queries built in {time.time() - t0:.1f}s")

    # PCIe-inclusive plain rate: the host-buffer entry point (query upload,
    # answer, response download, host conversions), for DESIGN notes only.
    for j in range(2):
        server.make_query(1000 + j, queries[j].copy())
    t0 = time.perf_counter()
    for j in range(10):
        server.make_query(1000 + j % 2, queries[j % 2].copy())
    host_ms = (time.perf_counter() - t0) * 100.0

    # correctness gate on the benchmarked database: decrypt two answers
    for j in range(2):
        r = server.make_query(1000 + j, queries[j].copy())
        pt, budget = clients[j].decrypt_mod_q(r)
        assert np.array_equal(pt, server.direct_get_original_plaintext(idxs[j])), "wrong PIR answer"
        log(f"[check] query {j} decrypts to plaintext {idxs[j]}, noise budget {budget}")

    db_h, ctx_h = server.db_handle(), server.ctx_handle()
    # One explicit stream (default priority) shared by torch's events and the
    # library: torch's default stream is the NULL stream, which the C ABI
    # would not adopt.
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    assert stream.cuda_stream, "expected a non-NULL HIP stream"
    assert G.ctx_call_raw("opir_ctx_set_stream", ctx_h, __import__("ctypes").c_void_p(stream.cuda_stream)) == 0
    assert G.ctx_call_raw("opir_ctx_set_scan_variant", ctx_h, scan_shape_arg(args)) == 0
    if args.copy_probe_gib > 0:
        n = int(args.copy_probe_gib * 2**30) // 4
        src = torch.ones(n, dtype=torch.int32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(2):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[probe] copied {n * 4} bytes twice")
        del src, dst
    d_q = torch.from_numpy(queries.view(np.int64)).to(dev)
    d_r = torch.empty_like(d_q)
    ids = np.arange(1000, 1000 + nclients, dtype=np.uint64)

    def answer(b, off=0):
        rc = G.answer_batch_device_raw(db_h, ids[off:off + b], d_q[off:off + b].data_ptr(), d_r[off:off + b].data_ptr())
        if rc != 0:
            raise RuntimeError(G.last_error_raw(ctx_h))

    def timed(b, steps, warmup):
        off = lambda i: (i * b) % nclients if b * 2 <= nclients else 0
        for i in range(warmup):
            answer(b, off(i))
        torch.cuda.synchronize()
        evs = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
        wall0 = time.perf_counter()
        evs[0].record()
        for i in range(steps):
            answer(b, off(i))
            evs[i + 1].record()
        torch.cuda.synchronize()
        wall = time.perf_counter() - wall0
        per = [evs[i].elapsed_time(evs[i + 1]) for i in range(steps)]
        return evs[0].elapsed_time(evs[-1]), per, wall

    gib_db = params.target_num_pt * params.pt_size / 2**30
    total_ms, per, wall = timed(1, args.steps, args.warmup)
    ms_plain = total_ms / args.steps
    log(f"[plain] {ms_plain:.3f} ms/query  p50 {pct(per, 50):.3f} p99 {pct(per, 99):.3f}  wall {wall:.2f}s")

    # profiled pass: reference stage breakdown + scan kernel time
    G.ctx_call_raw("opir_ctx_set_profiling", ctx_h, 1)
    stage_acc, scan_ms = {}, []
    import ctypes
    for i in range(max(3, min(args.steps, 10))):
        answer(1, i % B)
        out = (ctypes.c_float * 6)()
        assert G.ctx_call_raw("opir_last_stage_ms", ctx_h, out) == 0
        names = ["Expand", "Convert", "First dim", "Other dim", "Modulus switching"]
        for k, v in zip(names, out[:5]):
            stage_acc.setdefault(k, []).append(v)
        scan_ms.append(out[5])
    G.ctx_call_raw("opir_ctx_set_profiling", ctx_h, 0)
    stages = {k: round(sum(v) / len(v), 4) for k, v in stage_acc.items()}
    scan = sum(scan_ms) / len(scan_ms)
    N0, Nr = params.fst_dim_sz, params.other_dim_sz
    scan_bytes = params.num_pt * N * 8 + N * N0 * 2 * 8 + Nr * 2 * N * 8
    achieved = scan_bytes / (scan * 1e-3) / 1e9
    log(f"[plain] stages {stages} scan {scan:.3f} ms -> {achieved:.0f} GB/s")

    traffic, traffic_src = committed_traffic(args)
    result = {
        "metric": "DB GiB/s & queries/s, 1GB n=2048, batched-64 vs plain GPU; p50/p99 ms",
        "value": round(gib_db / (ms_plain * 1e-3), 3),
        "unit": "DB GiB/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_plain, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u32",
        "arithmetic": "exact modular: two 29-bit RNS limbs on 32-bit VALU; scan on int8 MFMA digit planes "
                      "with int32 accumulation; no floating point",
        "data": "synthetic (seeded DB coefficients < t, seeded client keys and queries)",
        "config": {"workload": "OnionPIRv2 make_query, CONFIG_N2048_K1_COMP, plain single-query path, "
                               "with_layout({262144,9}) = 256 x 1024 plaintexts (0.75 GiB logical, 4 GiB NTT-form)",
                   "fst_dim_sz": N0, "other_dim_sz": Nr, "num_dims": params.num_dims,
                   "target_num_pt": params.target_num_pt, "parallelism": "single GPU"},
        "queries_per_s_plain": round(1e3 / ms_plain, 2),
        "latency_ms_plain": {"p50": round(pct(per, 50), 4), "p99": round(pct(per, 99), 4)},
        "plain_ms_pcie_inclusive": round(host_ms, 4),
        "stage_ms_plain": stages,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": traffic_src,
                     "kernel": "k_scan_direct / k_scan_pp / k_scan (first-dimension scan, int8 MFMA over 4 digit planes)", "kernel_ms": round(scan, 4),
                     "algorithmic_bytes": scan_bytes, "scan_shape": tuned_shape(G, db_h, 1)},
        "scan_shape_mode": args.scan_shape,
    }
    if not args.no_batch:
        total_ms, per_b, wall = timed(B, max(3, args.steps // 4), 1)
        steps_b = max(3, args.steps // 4)
        ms_b = total_ms / steps_b
        result["batched"] = {"batch": B, "ms_per_batch": round(ms_b, 3),
                             "queries_per_s": round(B * 1e3 / ms_b, 2),
                             "db_gib_per_s": round(B * gib_db / (ms_b * 1e-3), 3),
                             "speedup_vs_plain_qps": round((B * 1e3 / ms_b) / (1e3 / ms_plain), 3),
                             "latency_ms": {"p50": round(pct(per_b, 50), 3), "p99": round(pct(per_b, 99), 3)},
                             "scan_shape": tuned_shape(G, db_h, B)}
        log(f"[batch] B={B}: {ms_b:.2f} ms/batch -> {B * 1e3 / ms_b:.1f} q/s")
        result["batched"]["scan"] = batched_scan_leg(G, ctx_h, answer, B, params, log)
        # batched answers must equal the plain ones
        answer(B)
        torch.cuda.synchronize()
        rb = d_r.cpu().numpy().view(np.uint64)
        for j in [0, B - 1]:
            assert np.array_equal(rb[j], server.make_query(1000 + j, queries[j].copy())), "batch != plain"
        # batch-size sweep: one fused pass per batch, distinct clients
        rows = []
        for b in sweep:
            steps_s = max(3, min(args.steps, 256 // b))
            total_ms, per_s, _ = timed(b, steps_s, 1)
            ms_s = total_ms / steps_s
            rows.append({"batch": b, "ms_per_batch": round(ms_s, 3), "queries_per_s": round(b * 1e3 / ms_s, 2),
                         "p50_ms": round(pct(per_s, 50), 3), "p99_ms": round(pct(per_s, 99), 3), "steps": steps_s,
                         "scan_shape": tuned_shape(G, db_h, b)})
            log(f"[sweep] B={b}: {ms_s:.2f} ms/batch -> {b * 1e3 / ms_s:.1f} q/s")
        if rows:
            result["batch_sweep"] = rows
    if not args.no_graphs:
        result["graphs"] = graph_leg(args, G, ctx_h, answer, torch, B, d_r, server, queries, log)
    if not args.no_paper_row:
        result["paper_1gb_row"] = paper_row(args, H, G, torch, log)
    if args.tree_L > 0:
        result["tree_pir"] = tree_leg(args, H, log)
    others = [c for c in args.other_schemes.split(",") if c]
    if others:
        result["other_schemes"] = {c: scheme_leg(args, H, G, torch, log, c) for c in others}
    k2 = k2_setup(args, H, log) if not args.no_k2 else None
    if not args.no_serve:
        torch.cuda.synchronize()
        result["serving"] = serving(args, H, params, server, log, k2)
    if k2 is not None:
        result["n4096_k2"] = k2_leg(args, H, G, torch, log, k2)
        del k2
    if not args.no_cpu:
        try:
            result["cpu_baseline"] = cpu_baseline(args, log)
        except Exception as e:  # the CPU leg is a reported baseline only
            result["cpu_baseline"] = {"value": None, "error": str(e)}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
