// PirServer of the host API: the reference's server-side interface
// (src/includes/server.h:15-172) with every homomorphic stage executed on the
// GPU through include/onionpir_gpu.h. Value semantics and exceptions follow
// src/server.cpp; the database and the client keys live in HBM.
#include <bit>
#include <cstring>
#include <stdexcept>
#include <unordered_set>

#include "onionpir_gpu.h"
#include "codec.h"
#include "onionpir_host.h"

namespace onionpir {

GpuScheme::GpuScheme(const PirParams &p, int device) : params_(p) {
  opir_scheme s{};
  s.poly_degree = (uint32_t)p.get_poly_degree();
  if (p.q1()) {  // CONFIG_N2048_K1_COMP: one logical limb mod q1*q2
    s.composite = 1;
    s.q1 = p.q1();
    s.q2 = p.q2();
    s.ntt_root = p.ntt_root();
  } else if (p.K() == 2) {  // CONFIG_N2048_K2_MP / CONFIG_N4096_K2_MP: two RNS limbs
    const size_t N = p.get_poly_degree();
    s.composite = 0;
    s.q1 = p.get_rns_mods()[0];
    s.q2 = p.get_rns_mods()[1];
    s.ntt_root = utils::minimal_primitive_root(2 * N, s.q1);
    s.ntt_root2 = utils::minimal_primitive_root(2 * N, s.q2);
  } else if (p.K() == 1) {  // CONFIG_N2048_K1: one prime, q2 = 0
    s.composite = 0;
    s.q1 = p.get_rns_mods()[0];
    s.q2 = 0;
    s.ntt_root = utils::minimal_primitive_root(2 * p.get_poly_degree(), s.q1);
  } else {
    throw std::invalid_argument("the GPU path runs CONFIG_N2048_K1_COMP, CONFIG_N2048_K1, CONFIG_N2048_K2_MP and "
                                "CONFIG_N4096_K2_MP");
  }
  s.plain_mod = p.get_plain_mod();
  s.small_q = p.get_small_q();
  s.l_ep = (uint32_t)p.get_l();
  s.base_log2 = (uint32_t)p.get_base_log2();
  s.l_key = (uint32_t)p.get_l_key();
  s.base_log2_key = (uint32_t)p.get_base_log2_key();
  s.l_ks = (uint32_t)p.get_l_ks();
  s.bv_base_log2 = (uint32_t)p.get_bv_base_log2();
  s.noise_std_dev = p.get_noise_std_dev();
  const int rc = opir_ctx_create(&s, device, &ctx_);
  if (rc != OPIR_OK) {
    const std::string msg = opir_last_error(nullptr);
    if (rc == OPIR_EINVAL) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
  }
}

GpuScheme::~GpuScheme() { opir_ctx_destroy(ctx_); }

void GpuScheme::check(int rc) const {
  if (rc == OPIR_OK) return;
  const std::string msg = opir_last_error(ctx_);
  if (rc == OPIR_EINVAL) throw std::invalid_argument(msg);
  if (rc == OPIR_ERANGE) throw std::out_of_range(msg);
  throw std::runtime_error(msg);
}

void GpuScheme::set_client_keys(size_t id, const bvks::BvGaloisKeys &bv, const GSWCt &gsw) {
  const size_t N = params_.get_poly_degree(), L = params_.get_l_ks(), KN = N * params_.K();
  std::vector<uint64_t> fb(bv.keys.size() * L * 2 * KN), fg(gsw.size() * 2 * KN);
  for (size_t i = 0; i < bv.keys.size(); i++) {
    const auto &k = bv.get((uint32_t)((N >> i) + 1));
    if (k.cts.size() != L) throw std::invalid_argument("BV key row count mismatch");
    for (size_t r = 0; r < L; r++) {
      if (k.cts[r].a.size() != KN || k.cts[r].b.size() != KN) throw std::invalid_argument("BV key row size mismatch");
      std::memcpy(&fb[((i * L + r) * 2 + 0) * KN], k.cts[r].a.data(), KN * 8);
      std::memcpy(&fb[((i * L + r) * 2 + 1) * KN], k.cts[r].b.data(), KN * 8);
    }
  }
  if (gsw.size() != 2 * params_.get_l_key()) throw std::invalid_argument("RGSW key row count mismatch");
  for (size_t r = 0; r < gsw.size(); r++) {
    if (gsw[r].size() != 2 * KN) throw std::invalid_argument("RGSW key row size mismatch");
    std::memcpy(&fg[r * 2 * KN], gsw[r].data(), 2 * KN * 8);
  }
  check(opir_client_set_keys(ctx_, id, fb.data(), bv.keys.size(), fg.data()));
}

PirServer::PirServer(const PirParams &params, int device)
    : PirServer(params, std::make_shared<GpuScheme>(params, device)) {}

PirServer::PirServer(const PirParams &params, std::shared_ptr<GpuScheme> scheme)
    : params_(params), scheme_(std::move(scheme)) {
  if (!params_.scheme_compatible(scheme_->params()))
    throw std::invalid_argument("PirServer layout is not scheme-compatible with its GPU scheme");
  scheme_->check(opir_db_create(scheme_->ctx(), params_.get_fst_dim_sz(), params_.get_other_dim_sz(),
                                params_.get_num_dims(), params_.get_expan_height(), &db_));
}

PirServer::~PirServer() { opir_db_destroy(db_); }

void PirServer::gen_data(const std::vector<size_t> &record, u64 seed) {
  // server.cpp:53-66 with a seeded generator: coefficient = rng() % t
  std::mt19937_64 rng(seed);
  const u64 t = params_.get_plain_mod();
  const size_t N = params_.get_poly_degree();
  load_data(params_.get_num_pt(), [&](size_t, RlwePt &out) {
    out.data.resize(N);
    for (auto &c : out.data) c = rng() % t;
  }, record);
}

void PirServer::load_data(size_t logical, const PlaintextSource &source, const std::vector<size_t> &record) {
  // server.cpp:80-203: validate, record, zero the rounded tail; the NTT and
  // the transpose into the scan layout run on the GPU.
  if (!source) throw std::invalid_argument("PirServer::load_data requires a source");
  const size_t num_pt = params_.get_num_pt(), N = params_.get_poly_degree();
  if (logical > num_pt) throw std::invalid_argument("PirServer::load_data logical plaintext count exceeds its layout");
  for (size_t r : record)
    if (r >= num_pt) throw std::invalid_argument("PirServer::load_data record index exceeds its layout");
  recorded_.clear();
  const std::unordered_set<size_t> rs(record.begin(), record.end());
  const u64 t = params_.get_plain_mod();
  const size_t chunk = 16384;
  std::vector<uint64_t> buf(chunk * N);
  RlwePt pt;
  for (size_t base = 0; base < num_pt; base += chunk) {
    const size_t n = std::min(chunk, num_pt - base);
    for (size_t j = 0; j < n; j++) {
      const size_t id = base + j;
      uint64_t *dst = &buf[j * N];
      if (id < logical) {
        source(id, pt);
        if (pt.data.size() != N) throw std::invalid_argument("Plaintext source returned the wrong coefficient count");
        for (size_t c = 0; c < N; c++) {
          if (pt.data[c] >= t) throw std::invalid_argument("Plaintext source coefficient exceeds the plaintext modulus");
          dst[c] = pt.data[c];
        }
      } else {
        std::memset(dst, 0, N * sizeof(uint64_t));
      }
      if (rs.count(id)) {
        RlwePt rec;
        rec.data.assign(dst, dst + N);
        recorded_[id] = std::move(rec);
      }
    }
    scheme_->check(opir_db_load_plain_u64(db_, buf.data(), base, n));
  }
}

static u64 splitmix64_h(u64 x) {
  x += 0x9E3779B97F4A7C15ULL;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}

void PirServer::gen_data_device(const std::vector<size_t> &record, u64 seed) {
  // server.cpp:53-66 with the plaintexts generated in HBM (opir_db_gen_random)
  const size_t num_pt = params_.get_num_pt(), N = params_.get_poly_degree();
  const u64 t = params_.get_plain_mod();
  for (size_t r : record)
    if (r >= num_pt) throw std::invalid_argument("PirServer::gen_data record index exceeds its layout");
  scheme_->check(opir_db_gen_random(db_, seed));
  recorded_.clear();
  for (size_t r : record) {
    RlwePt rec;
    rec.data.resize(N);
    for (size_t c = 0; c < N; c++) rec.data[c] = splitmix64_h(seed + r * N + c) % t;
    recorded_[r] = std::move(rec);
  }
}

void PirServer::push_keys(size_t id) {
  auto s = sessions_.find(id);
  if (s != sessions_.end()) {
    scheme_->set_client_keys(id, s->second->bv_galois_keys, s->second->gsw_key);
    return;
  }
  auto b = bv_keys_.find(id);
  auto g = gsw_keys_.find(id);
  if (b != bv_keys_.end() && g != gsw_keys_.end()) scheme_->set_client_keys(id, b->second, g->second);
}

void PirServer::set_client_bv_galois_key(size_t id, bvks::BvGaloisKeys bv) {
  wire_keyed_.erase(id);
  bv_keys_[id] = std::move(bv);
  push_keys(id);
}

void PirServer::set_client_gsw_key(size_t id, GSWCt gsw) {
  wire_keyed_.erase(id);
  gsw_keys_[id] = std::move(gsw);
  push_keys(id);
}

void PirServer::set_client_session_keys(size_t id, SharedPirSessionKeys keys) {  // server.cpp:865-875
  if (!keys) throw std::invalid_argument("PirServer session key bundle cannot be null");
  if (keys->bv_galois_keys.keys.size() < params_.get_expan_height())
    throw std::invalid_argument("PirServer session lacks keys for its expansion height");
  sessions_[id] = std::move(keys);
  push_keys(id);
}

SharedPirSessionKeys PirServer::client_session_keys(size_t id) const { return sessions_.at(id); }

RlwePt PirServer::direct_get_original_plaintext(size_t index) const {  // server.cpp:885-892
  auto it = recorded_.find(index);
  if (it == recorded_.end())
    throw std::invalid_argument("Plaintext index " + std::to_string(index) + " was not recorded during database loading");
  RlwePt out;
  out.data.assign(it->second.data.begin(), it->second.data.end());
  return out;
}

static std::vector<uint64_t> flatten(const RlweCt &ct, size_t N) {
  if (ct.c0.size() != N || ct.c1.size() != N) throw std::invalid_argument("ciphertext has the wrong size");
  std::vector<uint64_t> f(2 * N);
  std::memcpy(f.data(), ct.c0.data(), N * 8);
  std::memcpy(f.data() + N, ct.c1.data(), N * 8);
  return f;
}

static RlweCt unflatten(const uint64_t *f, size_t N, bool ntt) {
  RlweCt ct;
  ct.c0.assign(f, f + N);
  ct.c1.assign(f + N, f + 2 * N);
  ct.ntt_form = ntt;
  return ct;
}

static void require_keys(const std::map<size_t, bvks::BvGaloisKeys> &bv, const std::map<size_t, GSWCt> &g,
                         const std::map<size_t, SharedPirSessionKeys> &s, const std::set<size_t> &w, size_t id) {
  // server.cpp:739-743,841-845: std::map::at on an unknown client
  if (!s.count(id) && !w.count(id) && (!bv.count(id) || !g.count(id))) throw std::out_of_range("map::at");
}

void PirServer::set_client_keys_wire(size_t id, const std::string &wire) {
  const codec::SeededKeys k = codec::decode_keys(params_, wire);
  if (k.h < params_.get_expan_height()) throw std::invalid_argument("PirServer session lacks keys for its expansion height");
  scheme_->check(opir_client_set_keys_seeded(scheme_->ctx(), id, k.key.data(), k.bv_b.data(), k.h, k.gsw_c0.data()));
  bv_keys_.erase(id);
  gsw_keys_.erase(id);
  sessions_.erase(id);
  wire_keyed_.insert(id);
}

void PirServer::client_keygen_device(const std::vector<size_t> &ids, const std::vector<std::array<uint8_t, 32>> &keys) {
  if (ids.size() != keys.size()) throw std::invalid_argument("one key seed per client");
  std::vector<uint64_t> id64(ids.begin(), ids.end());
  std::vector<uint8_t> kb(32 * keys.size());
  for (size_t i = 0; i < keys.size(); i++) std::memcpy(&kb[32 * i], keys[i].data(), 32);
  scheme_->check(opir_client_keygen_device(scheme_->ctx(), ids.size(), id64.data(), kb.data(), params_.get_expan_height()));
  for (size_t id : ids) {
    bv_keys_.erase(id);
    gsw_keys_.erase(id);
    sessions_.erase(id);
    wire_keyed_.insert(id);
  }
}

RlweCt PirServer::make_query_wire(size_t id, const std::string &wire) {
  RlweCt q = codec::decode_query(params_, wire);
  return make_query(id, q);
}

RlweCt PirServer::make_query(size_t id, RlweCt &query) {  // server.cpp:895-947
  require_keys(bv_keys_, gsw_keys_, sessions_, wire_keyed_, id);
  const size_t N = params_.get_poly_degree();
  auto q = flatten(query, N * params_.K());
  std::vector<uint64_t> r(2 * N);
  scheme_->check(opir_answer(db_, id, q.data(), r.data()));
  return unflatten(r.data(), N, false);
}

std::vector<RlweCt> PirServer::make_queries(const std::vector<size_t> &ids, const std::vector<RlweCt> &queries) {
  if (ids.size() != queries.size()) throw std::invalid_argument("one client id per query");
  const size_t N = params_.get_poly_degree(), B = ids.size(), KN = N * params_.K();
  std::vector<uint64_t> q(B * 2 * KN), r(B * 2 * N), id64(ids.begin(), ids.end());
  for (size_t i = 0; i < B; i++) {
    require_keys(bv_keys_, gsw_keys_, sessions_, wire_keyed_, ids[i]);
    auto f = flatten(queries[i], KN);
    std::memcpy(&q[i * 2 * KN], f.data(), 2 * KN * 8);
  }
  scheme_->check(opir_answer_batch(db_, B, id64.data(), q.data(), r.data()));
  std::vector<RlweCt> out;
  for (size_t i = 0; i < B; i++) out.push_back(unflatten(&r[i * 2 * N], N, false));
  return out;
}

std::vector<RlweCt> PirServer::expand_query(size_t id, RlweCt &query) const {  // server.cpp:818-821
  require_keys(bv_keys_, gsw_keys_, sessions_, wire_keyed_, id);
  const size_t N = params_.get_poly_degree() * params_.K();  // values per polynomial
  const size_t u = params_.get_fst_dim_sz() + params_.get_l() * (params_.get_num_dims() - 1);
  auto q = flatten(query, N);
  std::vector<uint64_t> out(u * 2 * N);
  scheme_->check(opir_expand(db_, id, q.data(), out.data()));
  std::vector<RlweCt> v;
  for (size_t i = 0; i < u; i++) v.push_back(unflatten(&out[i * 2 * N], N, false));
  return v;
}

std::vector<GSWCt> PirServer::complete_selectors(size_t id, const std::vector<RlweCt> &expanded) {
  const size_t N = params_.get_poly_degree() * params_.K(), l = params_.get_l(), ns = params_.get_num_dims() - 1;
  std::vector<GSWCt> sel(ns);
  if (!ns) return sel;
  const size_t u = params_.get_fst_dim_sz() + l * ns;
  if (expanded.size() < u) throw std::invalid_argument("complete_selectors expects one expanded row group per selector");
  require_keys(bv_keys_, gsw_keys_, sessions_, wire_keyed_, id);
  std::vector<uint64_t> e(u * 2 * N), out(ns * 2 * l * 2 * N);
  for (size_t i = 0; i < u; i++) {
    auto f = flatten(expanded[i], N);
    std::memcpy(&e[i * 2 * N], f.data(), 2 * N * 8);
  }
  scheme_->check(opir_complete(db_, id, e.data(), out.data()));
  for (size_t a = 0; a < ns; a++)
    for (size_t r = 0; r < 2 * l; r++) {
      const uint64_t *p = &out[(a * 2 * l + r) * 2 * N];
      sel[a].emplace_back(p, p + 2 * N);
    }
  return sel;
}

std::vector<RlweCt> PirServer::evaluate_first_dim(std::vector<RlweCt> &sv) {
  const size_t N = params_.get_poly_degree() * params_.K(), N0 = params_.get_fst_dim_sz(), Nr = params_.get_other_dim_sz();
  if (sv.size() < N0) throw std::invalid_argument("evaluate_first_dim expects fst_dim_sz ciphertexts");
  std::vector<uint64_t> q(N0 * 2 * N), out(Nr * 2 * N);
  for (size_t i = 0; i < N0; i++) {
    auto f = flatten(sv[i], N);
    std::memcpy(&q[i * 2 * N], f.data(), 2 * N * 8);
  }
  scheme_->check(opir_first_dim(db_, q.data(), out.data()));
  std::vector<RlweCt> v;
  for (size_t i = 0; i < Nr; i++) v.push_back(unflatten(&out[i * 2 * N], N, false));
  return v;
}

RlweCt PirServer::evaluate_other_dim(std::vector<RlweCt> &mid, std::vector<GSWCt> &sel) {
  const size_t N = params_.get_poly_degree() * params_.K(), Nr = params_.get_other_dim_sz(), l = params_.get_l();
  if (mid.size() < Nr) throw std::invalid_argument("evaluate_other_dim expects other_dim_sz candidates");
  std::vector<uint64_t> m(Nr * 2 * N), s(sel.size() * 2 * l * 2 * N + 1), out(2 * N);
  for (size_t i = 0; i < Nr; i++) {
    auto f = flatten(mid[i], N);
    std::memcpy(&m[i * 2 * N], f.data(), 2 * N * 8);
  }
  for (size_t a = 0; a < sel.size(); a++)
    for (size_t r = 0; r < 2 * l; r++) std::memcpy(&s[(a * 2 * l + r) * 2 * N], sel[a][r].data(), 2 * N * 8);
  scheme_->check(opir_other_dim(db_, m.data(), s.data(), out.data()));
  return unflatten(out.data(), N, false);
}

void PirServer::ext_prod_mux(RlweCt &x, RlweCt &y, GSWCt &sel, RlweCt &result) {
  // server.cpp:652-722: result = x + INTT(sel [x] (y - x)), all on the GPU
  const size_t N = params_.get_poly_degree() * params_.K(), l = params_.get_l();
  if (sel.size() != 2 * l) throw std::invalid_argument("selection RGSW must hold 2*l rows");
  auto fx = flatten(x, N), fy = flatten(y, N);
  std::vector<uint64_t> g(2 * l * 2 * N), out(2 * N);
  for (size_t r = 0; r < 2 * l; r++) {
    if (sel[r].size() != 2 * N) throw std::invalid_argument("selection RGSW row has the wrong size");
    std::memcpy(&g[r * 2 * N], sel[r].data(), 2 * N * 8);
  }
  scheme_->check(
      opir_ext_prod_mux(scheme_->ctx(), g.data(), l, params_.get_base_log2(), fx.data(), fy.data(), out.data()));
  result = unflatten(out.data(), N, x.ntt_form);
}

size_t PirServer::save_resp_to_stream(const RlweCt &resp, std::stringstream &stream) {
  // server.cpp:950-1009: c0 then c1, ceil(log2 q') bits each, LSB-first
  const size_t N = params_.get_poly_degree();
  const size_t w = std::bit_width(params_.get_small_q() - 1);
  if (resp.c0.size() < N || resp.c1.size() < N) throw std::invalid_argument("response must hold N coefficients per poly");
  std::vector<u64> v(2 * N);
  std::copy(resp.c0.begin(), resp.c0.begin() + N, v.begin());
  std::copy(resp.c1.begin(), resp.c1.begin() + N, v.begin() + N);
  std::string out;
  const size_t bytes = codec::pack_bits(v.data(), 2 * N, w, out);
  stream.write(out.data(), (std::streamsize)out.size());
  return bytes;
}

}  // namespace onionpir
