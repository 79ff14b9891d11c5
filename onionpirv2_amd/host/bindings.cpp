// pybind11 module `_opir_host`: Python view of the host C++ API
// (PirParams / PirClient / PirServer) with ciphertexts as numpy arrays
// [2, N] (c0, c1), RGSW as [2l, 2N] and BV keys as [h, l_ks, 2 (a, b), N].
#include <pybind11/functional.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cstring>

#include "codec.h"
#include "onionpir_host.h"
#include "serving.h"
#include "tree.h"

namespace py = pybind11;
using namespace onionpir;
using Arr = py::array_t<uint64_t, py::array::c_style | py::array::forcecast>;

static Arr ct_arr(const RlweCt &ct) {
  const size_t n = ct.c0.size();
  Arr a({(py::ssize_t)2, (py::ssize_t)n});
  std::memcpy(a.mutable_data(), ct.c0.data(), n * 8);
  std::memcpy(a.mutable_data() + n, ct.c1.data(), n * 8);
  return a;
}
static RlweCt arr_ct(const Arr &a) {
  if (a.ndim() != 2 || a.shape(0) != 2) throw std::invalid_argument("ciphertext array must be [2, N]");
  const size_t n = a.shape(1);
  RlweCt ct;
  ct.c0.assign(a.data(), a.data() + n);
  ct.c1.assign(a.data() + n, a.data() + 2 * n);
  return ct;
}
static Arr gsw_arr(const GSWCt &g) {
  const size_t r = g.size(), w = r ? g[0].size() : 0;
  Arr a({(py::ssize_t)r, (py::ssize_t)w});
  for (size_t i = 0; i < r; i++) std::memcpy(a.mutable_data() + i * w, g[i].data(), w * 8);
  return a;
}
static GSWCt arr_gsw(const Arr &a) {
  GSWCt g(a.shape(0));
  const size_t w = a.shape(1);
  for (size_t i = 0; i < g.size(); i++) g[i].assign(a.data() + i * w, a.data() + (i + 1) * w);
  return g;
}
static Arr bv_arr(const bvks::BvGaloisKeys &k) {
  const size_t h = k.keys.size(), L = h ? k.keys[0].cts.size() : 0, N = (h && L) ? k.keys[0].cts[0].a.size() : 0;
  Arr a({(py::ssize_t)h, (py::ssize_t)L, (py::ssize_t)2, (py::ssize_t)N});
  for (size_t i = 0; i < h; i++)
    for (size_t r = 0; r < L; r++) {
      std::memcpy(a.mutable_data() + ((i * L + r) * 2 + 0) * N, k.keys[i].cts[r].a.data(), N * 8);
      std::memcpy(a.mutable_data() + ((i * L + r) * 2 + 1) * N, k.keys[i].cts[r].b.data(), N * 8);
    }
  return a;
}
static bvks::BvGaloisKeys arr_bv(const Arr &a, size_t N) {
  bvks::BvGaloisKeys k;
  if (a.ndim() != 4) throw std::invalid_argument("BV keys must be [h, l_ks, 2, N]");
  const size_t h = a.shape(0), L = a.shape(1), n = a.shape(3);
  for (size_t i = 0; i < h; i++) {
    bvks::BvKeySwitchKey key;
    key.galois_k = (uint32_t)((N >> i) + 1);
    key.cts.resize(L);
    for (size_t r = 0; r < L; r++) {
      key.cts[r].a.assign(a.data() + ((i * L + r) * 2) * n, a.data() + ((i * L + r) * 2 + 1) * n);
      key.cts[r].b.assign(a.data() + ((i * L + r) * 2 + 1) * n, a.data() + ((i * L + r) * 2 + 2) * n);
    }
    k.keys.push_back(std::move(key));
  }
  return k;
}

static prg::Seed to_seed(const py::bytes &b) {
  const std::string s = b;
  if (s.size() != 32) throw std::invalid_argument("a seed is 32 bytes");
  prg::Seed r;
  std::memcpy(r.data(), s.data(), 32);
  return r;
}
static py::bytes seed_bytes(const prg::Seed &s) { return py::bytes(reinterpret_cast<const char *>(s.data()), 32); }

PYBIND11_MODULE(_opir_host, m) {
  m.doc() = "Host API of the MI355X OnionPIRv2 server (PirParams, PirClient, PirServer)";
  // tree PIR (tree.h)
  py::class_<TreePirParams>(m, "TreePirParams")
      .def_readonly("N", &TreePirParams::N).def_readonly("L", &TreePirParams::L).def_readonly("n", &TreePirParams::n)
      .def_readonly("r", &TreePirParams::r).def_readonly("N0", &TreePirParams::N0).def_readonly("a", &TreePirParams::a)
      .def_readonly("P", &TreePirParams::P).def_readonly("B", &TreePirParams::B).def_readonly("b", &TreePirParams::b)
      .def_readonly("w", &TreePirParams::w).def_readonly("W", &TreePirParams::W).def_readonly("h_q", &TreePirParams::h_q)
      .def_readonly("t", &TreePirParams::t);
  m.def("make_tree_pir_params_for_scheme", &make_tree_pir_params_for_scheme);
  m.def("tree_query_expansion_params", &tree_query_expansion_params);
  m.def("client_index", [](size_t leaf, const TreePirParams &p) {
    const ClientCoordinates c = client_index(leaf, p);
    return py::make_tuple(c.alpha, std::vector<int>(c.beta_bits_le.begin(), c.beta_bits_le.end()),
                          std::vector<int>(c.gamma_bits_le.begin(), c.gamma_bits_le.end()));
  });
  m.def("build_level_plans", [](const TreePirParams &p) {
    py::list out;
    for (const LevelPlan &pl : build_level_plans(p)) {
      py::dict d;
      d["level"] = pl.level; d["R"] = pl.R; d["case"] = (int)pl.select_case; d["coarsen"] = pl.coarsen_count;
      d["beta_begin"] = pl.beta_begin; d["beta_count"] = pl.beta_count;
      out.append(d);
    }
    return out;
  });
  m.def("synthetic_tree_node_value", &synthetic_tree_node_value);
  m.def("pack_tree_level_synthetic", [](size_t level, const TreePirParams &p) {
    const auto d = pack_tree_level(level, p, [&](size_t l, size_t i) { return synthetic_tree_node_value(l, i, p.t); });
    Arr a({(py::ssize_t)d.size(), (py::ssize_t)p.n});
    for (size_t i = 0; i < d.size(); i++) std::memcpy(a.mutable_data() + i * p.n, d[i].data(), p.n * 8);
    return a;
  });
  m.def("make_tree_query", [](PirClient &c, const PirParams &scheme, const TreePirParams &t, size_t leaf) {
    return ct_arr(make_tree_query(c, scheme, t, leaf));
  });
  py::class_<TreePirServer>(m, "TreePirServer")
      .def(py::init([](const PirParams &scheme, const TreePirParams &t) {
             return std::make_unique<TreePirServer>(scheme, t, nullptr);
           }))
      .def("preprocess_synthetic",
           [](TreePirServer &s) {
             const uint64_t t = s.tree().t;
             s.preprocess([t](size_t l, size_t i) { return synthetic_tree_node_value(l, i, t); });
           })
      .def("load_level",
           [](TreePirServer &s, size_t level, Arr pts) {
             std::vector<std::vector<uint64_t>> v(pts.shape(0));
             for (size_t p = 0; p < v.size(); p++) v[p].assign(pts.data() + p * pts.shape(1), pts.data() + (p + 1) * pts.shape(1));
             s.load_level(level, v);
           })
      .def("set_client_keys",
           [](TreePirServer &s, size_t id, Arr bv, Arr g) {
             s.gpu_scheme()->set_client_keys(id, arr_bv(bv, s.tree().n), arr_gsw(g));
           })
      .def("select_path", [](TreePirServer &s, size_t id, Arr q) {
        std::vector<RlweCt> r;
        {
          RlweCt ct = arr_ct(q);
          py::gil_scoped_release nogil;
          r = s.select_path(id, ct);
        }
        const size_t KN = r.empty() ? 0 : r[0].c0.size();
        Arr a({(py::ssize_t)r.size(), (py::ssize_t)2, (py::ssize_t)KN});
        for (size_t l = 0; l < r.size(); l++) {
          std::memcpy(a.mutable_data() + l * 2 * KN, r[l].c0.data(), KN * 8);
          std::memcpy(a.mutable_data() + l * 2 * KN + KN, r[l].c1.data(), KN * 8);
        }
        return a;
      });
  // counter-based sampler (prg.h)
  m.def("chacha20_block", [](py::bytes key, uint32_t counter, std::array<uint32_t, 3> nonce) {
    uint32_t w[16];
    prg::chacha20_block(to_seed(key), counter, nonce.data(), w);
    return py::bytes(reinterpret_cast<const char *>(w), 64);
  });
  m.def("prg_uniform", [](py::bytes key, uint32_t domain, uint32_t i, uint32_t j, uint64_t q, size_t n) {
    Arr a((py::ssize_t)n);
    prg::uniform(to_seed(key), domain, i, j, q, n, a.mutable_data());
    return a;
  });
  m.def("prg_ternary", [](py::bytes key, size_t n) {
    py::array_t<int8_t> a((py::ssize_t)n);
    prg::ternary(to_seed(key), n, a.mutable_data());
    return a;
  });
  m.def("gaussian_cdt", &prg::gaussian_cdt);
  m.def("prg_gaussian", [](py::bytes key, uint32_t domain, uint32_t i, double sigma, size_t n) {
    py::array_t<int64_t> a((py::ssize_t)n);
    prg::gaussian(to_seed(key), domain, i, prg::gaussian_cdt(sigma), n, a.mutable_data());
    return a;
  });
  m.def("query_seed", [](py::bytes key, uint64_t ctr) { return seed_bytes(prg::query_seed(to_seed(key), ctr)); });
  // seed-compressed wire codec (codec.h)
  m.def("query_wire_bytes", &codec::query_wire_bytes);
  m.def("keys_wire_bytes", &codec::keys_wire_bytes);
  m.def("encode_query", [](const PirParams &p, py::bytes qseed, Arr q) {
    return py::bytes(codec::encode_query(p, to_seed(qseed), arr_ct(q)));
  });
  m.def("decode_query", [](const PirParams &p, py::bytes w) { return ct_arr(codec::decode_query(p, std::string(w))); });
  m.def("encode_keys", [](const PirParams &p, py::bytes key, Arr bv, Arr gsw) {
    return py::bytes(codec::encode_keys(p, to_seed(key), arr_bv(bv, p.get_poly_degree()), arr_gsw(gsw)));
  });
  m.def("decode_keys", [](const PirParams &p, py::bytes w) {
    const codec::SeededKeys k = codec::decode_keys(p, std::string(w));
    const size_t KN = p.get_poly_degree() * p.K();
    Arr b({(py::ssize_t)k.h, (py::ssize_t)p.get_l_ks(), (py::ssize_t)KN});
    Arr g({(py::ssize_t)(2 * p.get_l_key()), (py::ssize_t)KN});
    if (!k.bv_b.empty()) std::memcpy(b.mutable_data(), k.bv_b.data(), k.bv_b.size() * 8);
    std::memcpy(g.mutable_data(), k.gsw_c0.data(), k.gsw_c0.size() * 8);
    py::dict d;
    d["key"] = seed_bytes(k.key);
    d["h"] = k.h;
    d["bv_b"] = b;
    d["gsw_c0"] = g;
    return d;
  });
  m.def("expand_keys", [](const PirParams &p, py::bytes w) {
    bvks::BvGaloisKeys bv;
    GSWCt g;
    codec::expand_keys(p, codec::decode_keys(p, std::string(w)), bv, g);
    return py::make_tuple(bv_arr(bv), gsw_arr(g));
  });
  m.def("pack_bits", [](Arr v, size_t width) {
    std::string out;
    codec::pack_bits(v.data(), (size_t)v.size(), width, out);
    return py::bytes(out);
  });
  m.def("unpack_bits", [](py::bytes b, size_t n, size_t width) {
    std::string s = b;
    size_t pos = 0;
    Arr a((py::ssize_t)n);
    codec::unpack_bits(s, pos, n, width, a.mutable_data());
    return py::make_tuple(a, pos);
  });
  py::register_exception_translator([](std::exception_ptr p) {
    try {
      if (p) std::rethrow_exception(p);
    } catch (const std::out_of_range &e) {
      PyErr_SetString(PyExc_IndexError, e.what());
    }
  });

  py::class_<PirParams>(m, "PirParams")
      .def(py::init([](const std::string &name) { return PirParams(SchemeConfig::by_name(name)); }),
           py::arg("config") = "CONFIG_N2048_K1_COMP")
      .def("with_layout",
           [](const PirParams &p, size_t t, size_t h, bool pow2) { return p.with_layout({t, h, pow2}); },
           py::arg("target_num_pt"), py::arg("expansion_height"), py::arg("fst_dim_pow2") = true)
      .def("with_query_shape", &PirParams::with_query_shape)
      .def("scheme_compatible", &PirParams::scheme_compatible)
      .def_property_readonly("N", &PirParams::get_poly_degree)
      .def_property_readonly("rns_mods", &PirParams::get_rns_mods)
      .def_property_readonly("q1", &PirParams::q1)
      .def_property_readonly("q2", &PirParams::q2)
      .def_property_readonly("ntt_root", &PirParams::ntt_root)
      .def_property_readonly("plain_mod", &PirParams::get_plain_mod)
      .def_property_readonly("small_q", &PirParams::get_small_q)
      .def_property_readonly("base_log2", &PirParams::get_base_log2)
      .def_property_readonly("base_log2_key", &PirParams::get_base_log2_key)
      .def_property_readonly("bv_base_log2", &PirParams::get_bv_base_log2)
      .def_property_readonly("l_ep", &PirParams::get_l)
      .def_property_readonly("l_key", &PirParams::get_l_key)
      .def_property_readonly("l_ks", &PirParams::get_l_ks)
      .def_property_readonly("num_pt", &PirParams::get_num_pt)
      .def_property_readonly("target_num_pt", &PirParams::get_target_num_pt)
      .def_property_readonly("fst_dim_sz", &PirParams::get_fst_dim_sz)
      .def_property_readonly("other_dim_sz", &PirParams::get_other_dim_sz)
      .def_property_readonly("num_dims", &PirParams::get_num_dims)
      .def_property_readonly("expansion_height", &PirParams::get_expan_height)
      .def_property_readonly("pt_size", &PirParams::get_pt_size)
      .def_property_readonly("db_size_mb", &PirParams::get_DBSize_MB)
      .def_property_readonly("physical_storage_mb", &PirParams::get_physical_storage_MB)
      .def("get_BFV_size", &PirParams::get_BFV_size, py::arg("use_seed") = true)
      .def("get_gsw_key_size", &PirParams::get_gsw_key_size, py::arg("use_seed") = true)
      .def("get_bv_galois_key_size", &PirParams::get_bv_galois_key_size, py::arg("use_seed") = true);

  py::class_<PirClient>(m, "PirClient")
      .def(py::init<const PirParams &, size_t, uint64_t>(), py::arg("params"), py::arg("client_id"), py::arg("seed"))
      .def(py::init([](const PirParams &p, size_t id, py::bytes key) {
             return std::make_unique<PirClient>(p, id, to_seed(key));
           }),
           py::arg("params"), py::arg("client_id"), py::arg("key"))
      .def_property_readonly("client_id", &PirClient::get_client_id)
      .def_property_readonly("seeded", &PirClient::seeded)
      .def("key_seed", [](const PirClient &c) { return seed_bytes(c.key_seed()); })
      .def("last_query_seed", [](const PirClient &c) { return seed_bytes(c.last_query_seed()); })
      .def("secret_key", [](const PirClient &c) {
        const auto &d = c.secret_key().data;
        Arr a((py::ssize_t)d.size());
        std::memcpy(a.mutable_data(), d.data(), d.size() * 8);
        return a;
      })
      .def("get_query_indices", &PirClient::get_query_indices)
      .def("fast_generate_query",
           [](PirClient &c, const PirParams &p, size_t i) {
             RlweCt q;
             {
               py::gil_scoped_release nogil;
               q = c.fast_generate_query(p, i);
             }
             return ct_arr(q);
           })
      .def("create_bv_galois_keys",
           [](PirClient &c) {
             bvks::BvGaloisKeys k;
             {
               py::gil_scoped_release nogil;
               k = c.create_bv_galois_keys();
             }
             return bv_arr(k);
           })
      .def("generate_gsw_from_key",
           [](PirClient &c) {
             GSWCt g;
             {
               py::gil_scoped_release nogil;
               g = c.generate_gsw_from_key();
             }
             return gsw_arr(g);
           })
      .def("load_resp_from_stream",
           [](const PirClient &c, py::bytes b) {
             std::stringstream ss(std::string(b), std::ios::in | std::ios::binary);
             return ct_arr(c.load_resp_from_stream(ss));
           })
      .def("decrypt_mod_q", [](const PirClient &c, Arr ct) {
        int budget = 0;
        RlwePt pt = c.decrypt_mod_q(arr_ct(ct), &budget);
        Arr a((py::ssize_t)pt.data.size());
        std::memcpy(a.mutable_data(), pt.data.data(), pt.data.size() * 8);
        return py::make_tuple(a, budget);
      });

  py::class_<GpuScheme, std::shared_ptr<GpuScheme>>(m, "GpuScheme")
      .def(py::init<const PirParams &, int>(), py::arg("params"), py::arg("device") = 0)
      .def("ctx_handle", [](const GpuScheme &s) { return (uintptr_t)s.ctx(); });

  py::class_<PirServer>(m, "PirServer")
      .def(py::init<const PirParams &, int>(), py::arg("params"), py::arg("device") = 0)
      .def(py::init<const PirParams &, std::shared_ptr<GpuScheme>>(), py::arg("params"), py::arg("scheme"))
      .def("gen_data", &PirServer::gen_data, py::arg("record_indices") = std::vector<size_t>{}, py::arg("seed") = 0,
           py::call_guard<py::gil_scoped_release>())
      .def("gen_data_device", &PirServer::gen_data_device, py::arg("record_indices") = std::vector<size_t>{},
           py::arg("seed") = 0, py::call_guard<py::gil_scoped_release>())
      .def("load_data",
           [](PirServer &s, size_t logical, py::function src, std::vector<size_t> rec) {
             const size_t N = s.get_params().get_poly_degree();
             s.load_data(logical, [&](size_t i, RlwePt &out) {
               Arr a = src(i).cast<Arr>();
               out.data.assign(a.data(), a.data() + a.size());
               (void)N;
             }, rec);
           },
           py::arg("logical_num_pt"), py::arg("source"), py::arg("record_indices") = std::vector<size_t>{})
      .def("load_data_array",
           [](PirServer &s, Arr pts, std::vector<size_t> rec) {
             const size_t N = s.get_params().get_poly_degree(), n = pts.size() / N;
             const uint64_t *d = pts.data();
             py::gil_scoped_release nogil;
             s.load_data(n, [&](size_t i, RlwePt &out) { out.data.assign(d + i * N, d + (i + 1) * N); }, rec);
           },
           py::arg("pts"), py::arg("record_indices") = std::vector<size_t>{})
      .def("set_client_bv_galois_key",
           [](PirServer &s, size_t id, Arr bv) { s.set_client_bv_galois_key(id, arr_bv(bv, s.get_params().get_poly_degree())); })
      .def("set_client_gsw_key", [](PirServer &s, size_t id, Arr g) { s.set_client_gsw_key(id, arr_gsw(g)); })
      .def("set_client_session_keys",
           [](PirServer &s, size_t id, Arr bv, Arr g) {
             auto k = std::make_shared<PirSessionKeys>();
             k->bv_galois_keys = arr_bv(bv, s.get_params().get_poly_degree());
             k->gsw_key = arr_gsw(g);
             s.set_client_session_keys(id, k);
           })
      .def("make_query",
           [](PirServer &s, size_t id, Arr q) {
             RlweCt ct = arr_ct(q);
             RlweCt r;
             {
               py::gil_scoped_release nogil;
               r = s.make_query(id, ct);
             }
             return ct_arr(r);
           })
      .def("client_keygen_device",
           [](PirServer &s, std::vector<size_t> ids, std::vector<py::bytes> keys) {
             std::vector<prg::Seed> k;
             for (auto &b : keys) k.push_back(to_seed(b));
             s.client_keygen_device(ids, k);
           })
      .def("set_client_keys_wire",
           [](PirServer &s, size_t id, py::bytes w) { s.set_client_keys_wire(id, std::string(w)); })
      .def("make_query_wire",
           [](PirServer &s, size_t id, py::bytes w) {
             const std::string wire = w;
             RlweCt r;
             {
               py::gil_scoped_release nogil;
               r = s.make_query_wire(id, wire);
             }
             return ct_arr(r);
           })
      .def("make_queries",
           [](PirServer &s, std::vector<size_t> ids, Arr qs) {
             std::vector<RlweCt> v;
             const size_t n = qs.shape(0), N = qs.shape(2);
             for (size_t i = 0; i < n; i++) {
               RlweCt ct;
               ct.c0.assign(qs.data() + i * 2 * N, qs.data() + i * 2 * N + N);
               ct.c1.assign(qs.data() + i * 2 * N + N, qs.data() + (i + 1) * 2 * N);
               v.push_back(std::move(ct));
             }
             std::vector<RlweCt> r;
             {
               py::gil_scoped_release nogil;
               r = s.make_queries(ids, v);
             }
             const size_t rn = n ? r[0].c0.size() : 0;  // responses hold one limb mod small_q
             Arr a({(py::ssize_t)n, (py::ssize_t)2, (py::ssize_t)rn});
             for (size_t i = 0; i < n; i++) {
               std::memcpy(a.mutable_data() + i * 2 * rn, r[i].c0.data(), rn * 8);
               std::memcpy(a.mutable_data() + i * 2 * rn + rn, r[i].c1.data(), rn * 8);
             }
             return a;
           })
      .def("ext_prod_mux",
           [](PirServer &s, Arr x, Arr y, Arr sel) {
             RlweCt cx = arr_ct(x), cy = arr_ct(y), r;
             GSWCt g = arr_gsw(sel);
             s.ext_prod_mux(cx, cy, g, r);
             return ct_arr(r);
           })
      .def("save_resp_to_stream",
           [](PirServer &s, Arr ct) {
             std::stringstream ss;
             s.save_resp_to_stream(arr_ct(ct), ss);
             return py::bytes(ss.str());
           })
      .def("direct_get_original_plaintext",
           [](const PirServer &s, size_t i) {
             RlwePt p = s.direct_get_original_plaintext(i);
             Arr a((py::ssize_t)p.data.size());
             std::memcpy(a.mutable_data(), p.data.data(), p.data.size() * 8);
             return a;
           })
      .def("db_handle", [](const PirServer &s) { return (uintptr_t)s.gpu_db(); })
      .def("ctx_handle", [](const PirServer &s) { return (uintptr_t)s.gpu_scheme()->ctx(); })
      .def("gpu_scheme", &PirServer::gpu_scheme)
      .def_property_readonly("params", &PirServer::get_params);

  m.def("plan_batch_limits", &plan_batch_limits, py::arg("ndb"), py::arg("ws"), py::arg("fixed"), py::arg("budget"),
        py::arg("max_batch"));

  py::class_<ServingLoop>(m, "ServingLoop")
      .def(py::init([](std::vector<PirServer *> dbs, size_t max_batch, size_t hbm_budget, size_t keyed_clients,
                       size_t lanes, double shrink_after_s, std::vector<size_t> max_batch_db) {
             ServeConfig c;
             c.max_batch = max_batch;
             c.max_batch_db = std::move(max_batch_db);
             c.hbm_budget = hbm_budget;
             c.keyed_clients = keyed_clients;
             c.lanes = lanes;
             c.shrink_after_s = shrink_after_s;
             return std::make_unique<ServingLoop>(std::move(dbs), c);
           }),
           py::arg("dbs"), py::arg("max_batch") = 64, py::arg("hbm_budget") = 0, py::arg("keyed_clients") = 0,
           py::arg("lanes") = 2, py::arg("shrink_after_s") = -1.0,
           py::arg("max_batch_db") = std::vector<size_t>{}, py::keep_alive<1, 2>())
      .def("submit",
           [](ServingLoop &l, size_t db, size_t cid, Arr q) {
             std::vector<u64> v(q.data(), q.data() + q.size());
             return l.submit(db, cid, std::move(v));
           })
      .def("poll",
           [](ServingLoop &l) {
             std::vector<ServeResult> r;
             l.poll(r);
             py::list out;
             for (auto &x : r) {
               py::dict d;
               d["ticket"] = x.ticket;
               d["db"] = x.db;
               d["client_id"] = x.client_id;
               d["t_submit"] = x.t_submit;
               d["t_start"] = x.t_start;
               d["t_done"] = x.t_done;
               d["batch"] = x.batch;
               d["error"] = x.error;
               const size_t n = x.resp.size() / 2;
               Arr a({(py::ssize_t)2, (py::ssize_t)n});
               if (n) std::memcpy(a.mutable_data(), x.resp.data(), 2 * n * 8);
               d["resp"] = a;
               out.append(d);
             }
             return out;
           })
      .def("wait_idle", &ServingLoop::wait_idle, py::call_guard<py::gil_scoped_release>())
      .def("stop", &ServingLoop::stop, py::call_guard<py::gil_scoped_release>())
      .def_property_readonly("batch_limits", &ServingLoop::batch_limits)
      .def("reserved_batch", &ServingLoop::reserved_batch, py::arg("db"), py::arg("lane") = 0)
      .def("peak_batch", &ServingLoop::peak_batch, py::arg("db"))
      .def_property_readonly("grows", &ServingLoop::grows)
      .def_property_readonly("trims", &ServingLoop::trims)
      .def("run_poisson",
           // qs: one query [2, K*N] per pool entry (a list, or a [P, 2, K*N] array when all
           // databases share a scheme). resp comes back as a list of [2, N] arrays.
           [](ServingLoop &l, std::vector<size_t> dbs, std::vector<size_t> cids, py::object qs, double lambda,
              double duration, u64 seed) {
             std::vector<Arr> qv;
             if (py::isinstance<py::array>(qs) && Arr::ensure(qs).ndim() == 3) {
               Arr a = Arr::ensure(qs);
               for (py::ssize_t i = 0; i < a.shape(0); i++) qv.push_back(Arr::ensure(qs[py::int_(i)]));
             } else {
               for (auto h : qs) qv.push_back(Arr::ensure(h));
             }
             if (qv.size() != dbs.size() || dbs.size() != cids.size())
               throw std::invalid_argument("pool: dbs, client ids and queries must have one entry each");
             std::vector<PoolEntry> pool(qv.size());
             for (size_t i = 0; i < qv.size(); i++) {
               Arr q = Arr::ensure(qv[i]);
               if (!q || q.ndim() != 2 || q.shape(0) != 2) throw std::invalid_argument("each query must be [2, K*N]");
               pool[i] = PoolEntry{dbs[i], cids[i], std::vector<u64>(q.data(), q.data() + q.size())};
             }
             LoadRun run;
             {
               py::gil_scoped_release nogil;
               run = run_poisson(l, pool, lambda, duration, seed);
             }
             const size_t n = run.results.size();
             py::array_t<int64_t> idx((py::ssize_t)n), batch((py::ssize_t)n);
             py::array_t<double> lat((py::ssize_t)n), wait((py::ssize_t)n);
             py::list resp;
             size_t errors = 0;
             for (size_t i = 0; i < n; i++) {
               const ServeResult &x = run.results[i];
               idx.mutable_data()[i] = (int64_t)run.pool_index[i];
               batch.mutable_data()[i] = (int64_t)x.batch;
               lat.mutable_data()[i] = (x.t_done - x.t_submit) * 1e3;
               wait.mutable_data()[i] = (x.t_start - x.t_submit) * 1e3;
               const size_t RN = x.resp.size() / 2;  // response values per polynomial
               Arr a({(py::ssize_t)2, (py::ssize_t)RN});
               if (RN) std::memcpy(a.mutable_data(), x.resp.data(), 2 * RN * 8);
               if (x.resp.empty() || !x.error.empty()) errors++;
               resp.append(a);
             }
             py::dict d;
             d["offered_qps"] = run.offered_qps;
             d["achieved_qps"] = run.achieved_qps;
             d["span_s"] = run.span_s;
             d["pool_index"] = idx;
             d["batch"] = batch;
             d["latency_ms"] = lat;
             d["queue_ms"] = wait;
             d["resp"] = resp;
             d["errors"] = errors;
             return d;
           });
}
