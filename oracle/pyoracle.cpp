// OnionPIRv2 CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle.h).
// pybind11 module `onionpir_oracle`: numpy-facing wrappers for the tests,
// smoke() and bench.py's cpu_baseline leg.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cstring>

#include "oracle.h"

namespace py = pybind11;
using namespace oracle;
using U64Arr = py::array_t<u64, py::array::c_style | py::array::forcecast>;
using U32Arr = py::array_t<uint32_t, py::array::c_style | py::array::forcecast>;

static U64Arr vec_to_arr(const std::vector<u64> &v) {
  U64Arr a((py::ssize_t)v.size());
  std::memcpy(a.mutable_data(), v.data(), v.size() * sizeof(u64));
  return a;
}
static std::vector<u64> arr_to_vec(const U64Arr &a) {
  return std::vector<u64>(a.data(), a.data() + a.size());
}

// ct <-> array [2, K*N]
static U64Arr ct_to_arr(const RlweCt &ct) {
  const size_t n = ct.c0.size();
  U64Arr a({(py::ssize_t)2, (py::ssize_t)n});
  std::memcpy(a.mutable_data(), ct.c0.data(), n * 8);
  std::memcpy(a.mutable_data() + n, ct.c1.data(), n * 8);
  return a;
}
static RlweCt arr_to_ct(const U64Arr &a, bool ntt_form = false) {
  if (a.ndim() != 2 || a.shape(0) != 2) throw std::invalid_argument("ciphertext array must be [2, K*N]");
  const size_t n = a.shape(1);
  RlweCt ct;
  ct.c0.assign(a.data(), a.data() + n);
  ct.c1.assign(a.data() + n, a.data() + 2 * n);
  ct.ntt_form = ntt_form;
  return ct;
}
static U64Arr cts_to_arr(const std::vector<RlweCt> &v, size_t KN) {
  U64Arr a({(py::ssize_t)v.size(), (py::ssize_t)2, (py::ssize_t)KN});
  for (size_t i = 0; i < v.size(); i++) {
    std::memcpy(a.mutable_data() + (2 * i) * KN, v[i].c0.data(), KN * 8);
    std::memcpy(a.mutable_data() + (2 * i + 1) * KN, v[i].c1.data(), KN * 8);
  }
  return a;
}
static std::vector<RlweCt> arr_to_cts(const U64Arr &a, bool ntt) {
  const size_t n = a.shape(0), KN = a.shape(2);
  std::vector<RlweCt> v(n);
  for (size_t i = 0; i < n; i++) {
    v[i].c0.assign(a.data() + 2 * i * KN, a.data() + (2 * i + 1) * KN);
    v[i].c1.assign(a.data() + (2 * i + 1) * KN, a.data() + (2 * i + 2) * KN);
    v[i].ntt_form = ntt;
  }
  return v;
}
static U64Arr gsw_to_arr(const GSWCt &g) {
  const size_t r = g.size(), w = r ? g[0].size() : 0;
  U64Arr a({(py::ssize_t)r, (py::ssize_t)w});
  for (size_t i = 0; i < r; i++) std::memcpy(a.mutable_data() + i * w, g[i].data(), w * 8);
  return a;
}
static GSWCt arr_to_gsw(const U64Arr &a) {
  GSWCt g(a.shape(0));
  const size_t w = a.shape(1);
  for (size_t i = 0; i < g.size(); i++) g[i].assign(a.data() + i * w, a.data() + (i + 1) * w);
  return g;
}
// BV keys <-> array [h, L_KS, 2 (a,b), K*N] + galois list
static U64Arr bv_to_arr(const BvGaloisKeys &k) {
  const size_t h = k.keys.size(), L = h ? k.keys[0].rows.size() : 0, KN = (h && L) ? k.keys[0].rows[0].a.size() : 0;
  U64Arr a({(py::ssize_t)h, (py::ssize_t)L, (py::ssize_t)2, (py::ssize_t)KN});
  u64 *p = a.mutable_data();
  for (size_t i = 0; i < h; i++)
    for (size_t j = 0; j < L; j++) {
      std::memcpy(p + ((i * L + j) * 2 + 0) * KN, k.keys[i].rows[j].a.data(), KN * 8);
      std::memcpy(p + ((i * L + j) * 2 + 1) * KN, k.keys[i].rows[j].b.data(), KN * 8);
    }
  return a;
}
static BvGaloisKeys arr_to_bv(const U64Arr &a, size_t N) {
  BvGaloisKeys k;
  const size_t h = a.shape(0), L = a.shape(1), KN = a.shape(3);
  for (size_t i = 0; i < h; i++) {
    BvKey key;
    key.galois_k = (uint32_t)((N >> i) + 1);
    key.rows.resize(L);
    for (size_t j = 0; j < L; j++) {
      key.rows[j].a.assign(a.data() + ((i * L + j) * 2) * KN, a.data() + ((i * L + j) * 2 + 1) * KN);
      key.rows[j].b.assign(a.data() + ((i * L + j) * 2 + 1) * KN, a.data() + ((i * L + j) * 2 + 2) * KN);
    }
    k.keys.push_back(std::move(key));
  }
  return k;
}

static PlaintextSource make_source(const std::string &kind, u64 seed, u64 t, size_t N, std::mt19937_64 *rng,
                                   const U64Arr *arr) {
  if (kind == "random") {
    // server.cpp:53-66 (seeded): coefficient = rng() % t, pt-major order.
    return [rng, t, N](size_t, std::vector<u64> &out) {
      out.resize(N);
      for (auto &c : out) c = (*rng)() % t;
    };
  }
  if (kind == "affine") {
    // tests/test_shared_session.cpp:72-79: (17*i + 3*c + 11) % t.
    return [t, N](size_t i, std::vector<u64> &out) {
      out.resize(N);
      for (size_t c = 0; c < N; c++) out[c] = (17 * i + 3 * c + 11) % t;
    };
  }
  if (kind == "array") {
    return [arr, N](size_t i, std::vector<u64> &out) {
      out.assign(arr->data() + i * N, arr->data() + (i + 1) * N);
    };
  }
  (void)seed;
  throw std::invalid_argument("unknown source kind " + kind);
}

PYBIND11_MODULE(onionpir_oracle, m) {
  m.doc() = "OnionPIRv2 CPU oracle (test infrastructure only)";

  py::class_<PirParams>(m, "PirParams")
      .def(py::init([](const std::string &name) { return PirParams(config_by_name(name)); }))
      .def("with_layout", &PirParams::with_layout, py::arg("target_num_pt"), py::arg("expansion_height"),
           py::arg("fst_dim_pow2") = true)
      .def("with_query_shape", &PirParams::with_query_shape)
      .def("scheme_compatible", &PirParams::scheme_compatible)
      .def_property_readonly("N", &PirParams::N)
      .def_property_readonly("K", &PirParams::K)
      .def_readonly("rns_mods", &PirParams::rns_mods)
      .def_readonly("composite", &PirParams::composite)
      .def_readonly("q1", &PirParams::q1)
      .def_readonly("q2", &PirParams::q2)
      .def_readonly("w1", &PirParams::w1)
      .def_readonly("w2", &PirParams::w2)
      .def_readonly("w_crt", &PirParams::w_crt)
      .def_readonly("q1_inv_mod_q2", &PirParams::q1_inv_mod_q2)
      .def_readonly("q0_inv_mod_q1", &PirParams::q0_inv_mod_q1)
      .def_readonly("plain_mod", &PirParams::plain_mod)
      .def_readonly("small_q", &PirParams::small_q)
      .def_readonly("base_log2", &PirParams::base_log2)
      .def_readonly("base_log2_key", &PirParams::base_log2_key)
      .def_readonly("target_num_pt", &PirParams::target_num_pt)
      .def_readonly("expansion_height", &PirParams::expansion_height)
      .def_readonly("num_pt", &PirParams::num_pt)
      .def_readonly("fst_dim_sz", &PirParams::fst_dim_sz)
      .def_readonly("num_dims", &PirParams::num_dims)
      .def_property_readonly("other_dim_sz", &PirParams::other_dim_sz)
      .def_property_readonly("l_ep", &PirParams::l_ep)
      .def_property_readonly("l_key", &PirParams::l_key)
      .def_property_readonly("l_ks", &PirParams::l_ks)
      .def_property_readonly("bv_base_log2", &PirParams::bv_base_log2)
      .def_property_readonly("pt_size", &PirParams::pt_size)
      .def_property_readonly("ct_mod_width", &PirParams::ct_mod_width);

  // number theory
  m.def("is_prime", &is_prime);
  m.def("generate_prime", &generate_prime);
  m.def("generate_ntt_friendly_primes", &generate_ntt_friendly_primes);
  m.def("minimal_primitive_root", &minimal_primitive_root);
  m.def("crt_combine", &crt_combine);
  m.def("calculate_db_shape", &calculate_db_shape);
  m.def("bit_reverse", &bit_reverse);
  m.def("rescale", &rescale);
  m.def("try_invert", [](u64 v, u64 mod) -> py::object {
    u64 r;
    if (!try_invert_uint_mod(v, mod, r)) return py::none();
    return py::int_(r);
  });
  m.def("barrett_u64", [](u64 x, u64 q) { return barrett_reduce_u64(x, barrett_u64_setup(q)); });
  m.def("barrett_u128", [](u64 lo, u64 hi, u64 q) {
    return barrett_reduce_u128(((u128)hi << 64) | lo, barrett_u128_setup(q));
  });
  m.def("ntt_root", [](size_t N, u64 q) { return ntt_table(N, q).root; });
  m.def("ntt_fwd", [](U64Arr a, size_t N, u64 q) {
    auto v = arr_to_vec(a);
    if (v.size() % N) throw std::invalid_argument("size");
    for (size_t i = 0; i < v.size(); i += N) ntt_fwd(v.data() + i, N, q);
    return vec_to_arr(v);
  });
  m.def("ntt_inv", [](U64Arr a, size_t N, u64 q) {
    auto v = arr_to_vec(a);
    for (size_t i = 0; i < v.size(); i += N) ntt_inv(v.data() + i, N, q);
    return vec_to_arr(v);
  });
  m.def("automorphism_coeff", [](U64Arr a, uint32_t k, u64 q) {
    std::vector<u64> out(a.size());
    automorphism_coeff(a.data(), a.size(), k, q, out.data());
    return vec_to_arr(out);
  });
  m.def("negacyclic_shift", [](U64Arr a, int64_t shift, u64 q) {
    std::vector<u64> out(a.size());
    negacyclic_shift(a.data(), a.size(), (size_t)shift, q, out.data());
    return vec_to_arr(out);
  });
  m.def("gsw_gadget", &gsw_gadget);
  m.def("signed_gadget_decompose", [](u64 v, size_t b, u64 q, size_t nd) {
    std::vector<u64> out(nd);
    signed_gadget_decompose(v, b, q, out.data(), nd);
    return out;
  });
  // two-prime gadget digits: val and Q as (lo, hi) u64 halves of u128 values
  m.def("signed_gadget_decompose_mp", [](u64 v_lo, u64 v_hi, u64 Q_lo, u64 Q_hi, size_t b, size_t nd) {
    std::vector<int64_t> out(nd);
    signed_gadget_decompose_mp(((u128)v_hi << 64) | v_lo, ((u128)Q_hi << 64) | Q_lo, b, out.data(), nd);
    return out;
  });
  m.def("drop_q1", &drop_q1);
  m.def("level_mat_mat_32", [](U32Arr A, U32Arr B, size_t m_, size_t n, size_t levels, u64 q) {
    if ((size_t)A.size() != m_ * n * levels || (size_t)B.size() != n * 2 * levels)
      throw std::invalid_argument("shape");
    std::vector<u64> out(m_ * 2 * levels);
    level_mat_mat_32(A.data(), B.data(), out.data(), m_, n, levels, q);
    return vec_to_arr(out);
  });
  m.def("level_mat_mat_u64", [](U64Arr A, U64Arr B, size_t m_, size_t n, size_t levels, u64 q) {
    if ((size_t)A.size() != m_ * n * levels || (size_t)B.size() != n * 2 * levels)
      throw std::invalid_argument("shape");
    std::vector<u128> out(m_ * 2 * levels);
    std::vector<u64> lq(levels, q);
    level_mat_mat_u64(A.data(), B.data(), out.data(), m_, n, levels, lq.data());
    std::vector<u64> o(out.size());
    for (size_t i = 0; i < o.size(); i++) o[i] = (u64)out[i];
    return vec_to_arr(o);
  });
  m.def("mt19937_64_draws", [](u64 seed, size_t n) {
    std::mt19937_64 r(seed);
    std::vector<u64> v(n);
    for (auto &x : v) x = r();
    return vec_to_arr(v);
  });

  // Secret-key level helpers for invariant tests.
  m.def("gen_sk_and_encrypt", [](const PirParams &p, U64Arr msg, u64 seed) {
    std::mt19937_64 rng(seed);
    RlweSk sk = gen_secret_key_rns(p.N(), p.rns_mods, rng);
    RlweCt ct;
    encrypt_bfv_rns(arr_to_vec(msg), sk, p.N(), p.rns_mods, p.plain_mod, p.sigma(), rng, ct);
    return py::make_tuple(vec_to_arr(sk.data), ct_to_arr(ct));
  });
  m.def("decrypt_rns", [](const PirParams &p, U64Arr sk, U64Arr ct, bool ntt) {
    RlweSk s{arr_to_vec(sk)};
    std::vector<u64> pt;
    decrypt_rns(arr_to_ct(ct, ntt), s, p.N(), p.rns_mods, p.plain_mod, p.q0_inv_mod_q1, pt);
    return vec_to_arr(pt);
  });
  m.def("phase", [](const PirParams &p, U64Arr sk, U64Arr ct, bool ntt) {
    // c0 + c1*s mod Q (K = 1), centred to (-Q/2, Q/2]; coefficient form.
    const size_t N = p.N();
    const u64 Q = p.rns_mods[0];
    RlweCt c = arr_to_ct(ct, ntt);
    std::vector<u64> c0(c.c0), c1(c.c1), ph(N);
    if (ntt) ntt_inv(c0.data(), N, Q);
    else ntt_fwd(c1.data(), N, Q);
    for (size_t i = 0; i < N; i++) ph[i] = mulmod(c1[i], sk.data()[i], Q);
    ntt_inv(ph.data(), N, Q);
    py::array_t<int64_t> out((py::ssize_t)N);
    for (size_t i = 0; i < N; i++) {
      const u64 v = (ph[i] + c0[i]) % Q;
      out.mutable_data()[i] = v > Q / 2 ? (int64_t)v - (int64_t)Q : (int64_t)v;
    }
    return out;
  });
  m.def("decrypt_and_budget", [](const PirParams &p, U64Arr sk, U64Arr ct, bool ntt) {
    RlweSk s{arr_to_vec(sk)};
    std::vector<u64> pt;
    int b = decrypt_and_budget(arr_to_ct(ct, ntt), s, p.N(), p.rns_mods[0], p.plain_mod, pt);
    return py::make_tuple(vec_to_arr(pt), b);
  });
  m.def("plain_to_gsw", [](const PirParams &p, U64Arr pt, U64Arr sk, size_t l, size_t b, u64 seed) {
    std::mt19937_64 rng(seed);
    GSWEval g(p, l, b);
    RlweSk s{arr_to_vec(sk)};
    return gsw_to_arr(g.plain_to_gsw(arr_to_vec(pt), s, rng));
  });
  m.def("external_product", [](const PirParams &p, U64Arr gsw, U64Arr ct, size_t l, size_t b) {
    GSWEval g(p, l, b);
    RlweCt res;
    g.external_product(arr_to_gsw(gsw), arr_to_ct(ct), res);
    return ct_to_arr(res);
  });
  m.def("decompose_rows", [](const PirParams &p, U64Arr ct, size_t l, size_t b) {
    GSWEval g(p, l, b);
    std::vector<std::vector<u64>> rows(2 * l, std::vector<u64>(p.N() * p.K()));
    if (p.K() == 1) g.decomp_single_mod(arr_to_ct(ct), rows);
    else g.decomp_mp(arr_to_ct(ct), rows);
    return gsw_to_arr(rows);
  });
  m.def("gen_bv_ks_key", [](const PirParams &p, U64Arr sk, uint32_t k, u64 seed) {
    std::mt19937_64 rng(seed);
    RlweSk s{arr_to_vec(sk)};
    BvGaloisKeys ks;
    ks.keys.push_back(gen_bv_ks_key(p, s, k, rng));
    return bv_to_arr(ks);
  });
  m.def("bv_apply", [](const PirParams &p, U64Arr ct, uint32_t k, U64Arr key) {
    RlweCt c = arr_to_ct(ct);
    BvGaloisKeys ks = arr_to_bv(key, p.N());
    ks.keys[0].galois_k = k;
    bv_apply_galois_inplace(c, k, ks.keys[0], p);
    return ct_to_arr(c);
  });

  // tree PIR
  py::class_<TreeParams>(m, "TreeParams")
      .def(py::init(&make_tree_params), py::arg("L"), py::arg("a"), py::arg("scheme"))
      .def_readonly("L", &TreeParams::L).def_readonly("N", &TreeParams::N).def_readonly("n", &TreeParams::n)
      .def_readonly("r", &TreeParams::r).def_readonly("a", &TreeParams::a).def_readonly("N0", &TreeParams::N0)
      .def_readonly("P", &TreeParams::P).def_readonly("B", &TreeParams::B).def_readonly("b", &TreeParams::b)
      .def_readonly("ell", &TreeParams::ell).def_readonly("w", &TreeParams::w).def_readonly("W", &TreeParams::W)
      .def_readonly("h_q", &TreeParams::h_q).def_readonly("t", &TreeParams::t);
  m.def("tree_client_index", [](size_t leaf, const TreeParams &p) {
    const TreeCoords c = tree_client_index(leaf, p);
    return py::make_tuple(c.alpha, std::vector<int>(c.beta.begin(), c.beta.end()),
                          std::vector<int>(c.gamma.begin(), c.gamma.end()));
  });
  m.def("tree_level_plans", [](const TreeParams &p) {
    py::list out;
    for (const TreePlan &pl : tree_level_plans(p)) {
      py::dict d;
      d["level"] = pl.level; d["R"] = pl.R; d["case"] = pl.select_case; d["coarsen"] = pl.coarsen;
      d["beta_begin"] = pl.beta_begin; d["beta_count"] = pl.beta_count;
      out.append(d);
    }
    return out;
  });
  m.def("synthetic_tree_node", &synthetic_tree_node);
  m.def("pack_tree_level", [](size_t level, const TreeParams &p) {
    const auto d = pack_tree_level(level, p);
    U64Arr a({(py::ssize_t)d.size(), (py::ssize_t)p.n});
    for (size_t i = 0; i < d.size(); i++) std::memcpy(a.mutable_data() + i * p.n, d[i].data(), p.n * 8);
    return a;
  });
  m.def("tree_level_oracle", [](size_t leaf, size_t level, const TreeParams &p) {
    const TreeLevelOracle o = tree_level_oracle(leaf, level, p);
    return py::make_tuple(o.node, o.plaintext, o.position);
  });
  m.def("make_tree_query", [](PirClient &c, const TreeParams &tp, const PirParams &qview, size_t leaf) {
    return ct_to_arr(make_tree_query(c, tp, qview, leaf));
  });
  m.def("tree_pt_ct_mul", [](U64Arr pt, U64Arr ct, const PirParams &p) {
    return ct_to_arr(tree_pt_ct_mul(arr_to_vec(pt), arr_to_ct(ct), p));
  });
  m.def("tree_select_all", [](PirServer &srv, const TreeParams &tp, size_t id, U64Arr q) {
    std::vector<std::vector<std::vector<u64>>> levels;
    for (size_t l = 0; l <= tp.L; l++) levels.push_back(pack_tree_level(l, tp));
    py::list out;
    for (const RlweCt &ct : tree_select_all(srv, tp, id, arr_to_ct(q), levels)) out.append(ct_to_arr(ct));
    return out;
  });

  py::class_<PirClient>(m, "PirClient")
      .def(py::init<const PirParams &, size_t, u64>(), py::arg("params"), py::arg("client_id"), py::arg("seed"))
      .def_readonly("client_id", &PirClient::client_id)
      .def("secret_key", [](const PirClient &c) { return vec_to_arr(c.sk.data); })
      .def("get_query_indices", &PirClient::get_query_indices)
      .def("fast_generate_query",
           [](PirClient &c, const PirParams &qp, size_t idx) { return ct_to_arr(c.fast_generate_query(qp, idx)); })
      .def("create_bv_galois_keys", [](PirClient &c) { return bv_to_arr(c.create_bv_galois_keys()); })
      .def("generate_gsw_from_key", [](PirClient &c) { return gsw_to_arr(c.generate_gsw_from_key()); })
      .def("decrypt_mod_q",
           [](const PirClient &c, U64Arr ct) {
             int b = 0;
             auto pt = c.decrypt_mod_q(arr_to_ct(ct), &b);
             return py::make_tuple(vec_to_arr(pt), b);
           })
      .def("decrypt_ct", [](const PirClient &c, U64Arr ct) { return vec_to_arr(c.decrypt_ct(arr_to_ct(ct))); })
      .def("encrypt_bfv",
           [](const PirClient &c, U64Arr msg, u64 seed) {
             // encrypt_bfv_rns (rlwe.cpp:278-308) under this client's key, own RNG
             std::mt19937_64 rng(seed);
             RlweCt ct;
             encrypt_bfv_rns(arr_to_vec(msg), c.sk, c.params.N(), c.params.rns_mods, c.params.plain_mod,
                             c.params.sigma(), rng, ct);
             return ct_to_arr(ct);
           })
      .def("load_resp_from_stream",
           [](const PirClient &c, py::bytes b) { return ct_to_arr(c.load_resp_from_stream(std::string(b))); });

  py::class_<PirServer>(m, "PirServer")
      .def(py::init<const PirParams &>())
      .def("load_data",
           [](PirServer &s, size_t logical, const std::string &kind, u64 seed, std::vector<size_t> record,
              py::object arr) {
             std::mt19937_64 rng(seed);
             U64Arr a;
             if (kind == "array") a = arr.cast<U64Arr>();
             s.load_data(logical, make_source(kind, seed, s.params.plain_mod, s.params.N(), &rng, &a), record);
           },
           py::arg("logical_num_pt"), py::arg("kind"), py::arg("seed") = 0, py::arg("record") = std::vector<size_t>{},
           py::arg("array") = py::none())
      .def("fill_random_ntt",
           [](PirServer &s, u64 seed) {
             // Uniform NTT-domain residues for timing only: the answer path's cost
             // does not depend on the database values.
             std::mt19937_64 r(seed);
             for (auto &x : s.db_lo) x = (uint32_t)(r() % s.params.q1);
             for (auto &x : s.db_hi) x = (uint32_t)(r() % s.params.q2);
           })
      .def("set_client_keys",
           [](PirServer &s, size_t id, U64Arr bv, U64Arr gsw) {
             s.set_client_keys(id, arr_to_bv(bv, s.params.N()), arr_to_gsw(gsw));
           })
      .def("ext_prod_mux",
           [](const PirServer &s, U64Arr x, U64Arr y, U64Arr sel) {
             RlweCt cx = arr_to_ct(x), cy = arr_to_ct(y), r;
             s.ext_prod_mux(cx, cy, arr_to_gsw(sel), r);
             return ct_to_arr(r);
           })
      .def("direct_get_original_plaintext", [](const PirServer &s, size_t i) {
        auto it = s.recorded.find(i);
        if (it == s.recorded.end()) throw std::invalid_argument("Plaintext index was not recorded during database loading");
        return vec_to_arr(it->second);
      })
      .def("db_lo", [](const PirServer &s) {
        U32Arr a((py::ssize_t)s.db_lo.size());
        std::memcpy(a.mutable_data(), s.db_lo.data(), s.db_lo.size() * 4);
        return a;
      })
      .def("db_hi", [](const PirServer &s) {
        U32Arr a((py::ssize_t)s.db_hi.size());
        std::memcpy(a.mutable_data(), s.db_hi.data(), s.db_hi.size() * 4);
        return a;
      })
      .def("fast_expand_qry",
           [](const PirServer &s, size_t id, U64Arr q) {
             return cts_to_arr(s.fast_expand_qry(id, arr_to_ct(q)), s.params.N() * s.params.K());
           })
      .def("complete_selectors",
           [](const PirServer &s, size_t id, U64Arr exp) {
             auto g = s.complete_selectors(id, arr_to_cts(exp, false));
             const size_t ns = g.size(), r = ns ? g[0].size() : 0, w = r ? g[0][0].size() : 0;
             U64Arr a({(py::ssize_t)ns, (py::ssize_t)r, (py::ssize_t)w});
             for (size_t i = 0; i < ns; i++)
               for (size_t j = 0; j < r; j++) std::memcpy(a.mutable_data() + (i * r + j) * w, g[i][j].data(), w * 8);
             return a;
           })
      .def("evaluate_first_dim",
           [](PirServer &s, U64Arr q) {
             auto v = arr_to_cts(q, false);
             return cts_to_arr(s.evaluate_first_dim(v), s.params.N() * s.params.K());
           })
      .def("evaluate_other_dim",
           [](const PirServer &s, U64Arr mid, U64Arr sel) {
             auto m_ = arr_to_cts(mid, false);
             std::vector<GSWCt> g(sel.shape(0));
             const size_t r = sel.shape(1), w = sel.shape(2);
             for (size_t i = 0; i < g.size(); i++) {
               g[i].resize(r);
               for (size_t j = 0; j < r; j++) g[i][j].assign(sel.data() + (i * r + j) * w, sel.data() + (i * r + j + 1) * w);
             }
             return ct_to_arr(s.evaluate_other_dim(m_, g));
           })
      .def("mod_switch",
           [](const PirServer &s, U64Arr ct) {
             RlweCt c = arr_to_ct(ct);
             s.mod_switch_inplace(c, s.params.small_q);
             return ct_to_arr(c);
           })
      .def("make_query", [](PirServer &s, size_t id, U64Arr q) { return ct_to_arr(s.make_query(id, arr_to_ct(q))); })
      .def("last_stage_ms",
           [](const PirServer &s) {
             py::dict d;
             d["Expand"] = s.last_times.expand_ms;
             d["Convert"] = s.last_times.convert_ms;
             d["First dim"] = s.last_times.first_dim_ms;
             d["Other dim"] = s.last_times.other_dim_ms;
             d["Modulus switching"] = s.last_times.mod_switch_ms;
             return d;
           })
      .def("save_resp_to_stream", [](const PirServer &s, U64Arr ct) {
        return py::bytes(s.save_resp_to_stream(arr_to_ct(ct)));
      });
}
