// Host-side C++ API of the MI355X OnionPIRv2 server: the reference's
// PirParams / RlweCt / GSWCt / bvks keys / PirClient / PirServer
// (src/includes/pir.h, rlwe.h, gsw.h, bv_keyswitch.h, client.h, server.h)
// with the same names, argument meaning and exceptions. PirServer's work runs
// on the GPU through the C ABI of include/onionpir_gpu.h; PirClient runs on
// the CPU (it produces queries/keys and checks responses).
//
// Differences from the reference, all forced by making it testable and
// runtime-configurable:
//   - parameters are a runtime SchemeConfig instead of compile-time DBConsts;
//   - every random source takes an explicit seed (the reference seeds from
//     std::random_device), drawn with the same std:: distributions in the
//     same order, so a given seed reproduces the same keys and queries.
#pragma once
#include <array>
#include <cstddef>
#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <random>
#include <set>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

struct opir_ctx;
struct opir_db;

namespace onionpir {

using u64 = uint64_t;
using u32 = uint32_t;

// src/includes/database_constants.h:47-124 as a runtime value.
struct SchemeConfig {
  std::string name = "CONFIG_N2048_K1_COMP";
  size_t poly_degree = 2048;
  size_t l_ep = 6, l_key = 10, l_ks = 8;
  size_t tree_height = 10;
  size_t plain_mod_bits = 13;
  size_t small_q_width = 22;
  std::vector<size_t> rns_mod_bits = {58};
  bool composite_first_dim = true;
  std::vector<size_t> first_dim_rns_bits = {29, 29};
  double noise_std_dev = 2.55;
  size_t db_size_mb = 128;
  bool fst_dim_pow2 = true;
  static SchemeConfig by_name(const std::string &name);
};

// src/includes/pir.h:187-200
struct PirLayoutConfig {
  size_t target_num_pt;
  size_t expansion_height;
  bool fst_dim_pow2;
};

class PirParams {
 public:
  PirParams() : PirParams(SchemeConfig{}) {}
  explicit PirParams(const SchemeConfig &cfg);
  PirParams with_layout(const PirLayoutConfig &layout) const;
  PirParams with_query_shape(size_t fst_dim_sz, size_t num_selector_bits, size_t expansion_height) const;
  bool scheme_compatible(const PirParams &o) const;

  size_t get_ct_mod_width() const;
  size_t get_num_bits_per_coeff() const { return cfg_.plain_mod_bits - 1; }
  size_t get_pt_size() const { return get_num_bits_per_coeff() * cfg_.poly_degree / 8; }
  double get_DBSize_MB() const { return (double)num_pt_ * get_pt_size() / 1024 / 1024; }
  double get_physical_storage_MB() const { return (double)get_coeff_val_cnt() * num_pt_ * 8 / 1024 / 1024; }
  size_t get_num_pt() const { return num_pt_; }
  size_t get_target_num_pt() const { return target_num_pt_; }
  size_t get_num_dims() const { return num_dims_; }
  size_t get_l() const { return cfg_.l_ep; }
  size_t get_l_key() const { return cfg_.l_key; }
  size_t get_l_ks() const { return cfg_.l_ks; }
  u64 get_small_q() const { return small_q_; }
  size_t get_base_log2() const { return base_log2_; }
  size_t get_base_log2_key() const { return base_log2_key_; }
  size_t get_bv_base_log2() const { return get_ct_mod_width() / cfg_.l_ks + 1; }  // bv_keyswitch.cpp:73-76
  size_t get_fst_dim_sz() const { return fst_dim_sz_; }
  size_t get_other_dim_sz() const { return num_pt_ / fst_dim_sz_; }
  size_t K() const { return rns_mods_.size(); }
  size_t get_coeff_val_cnt() const { return cfg_.poly_degree * K(); }
  u64 get_plain_mod() const { return plain_mod_; }
  const std::vector<u64> &get_rns_mods() const { return rns_mods_; }
  size_t get_poly_degree() const { return cfg_.poly_degree; }
  size_t get_expan_height() const { return expansion_height_; }
  bool get_fst_dim_pow2() const { return fst_dim_pow2_; }
  size_t get_num_other_dims() const { return num_dims_ - 1; }
  double get_noise_std_dev() const { return cfg_.noise_std_dev; }
  size_t get_BFV_size(bool use_seed = true) const {  // pir.h:267-274 (size model)
    return use_seed ? 32 + get_ct_mod_width() * cfg_.poly_degree / 8 : get_ct_mod_width() * cfg_.poly_degree * 2 / 8;
  }
  size_t get_gsw_key_size(bool use_seed = true) const { return 2 * cfg_.l_key * get_BFV_size(use_seed); }
  size_t get_bv_galois_key_size(bool use_seed = true) const {
    const size_t pp = (cfg_.poly_degree * get_ct_mod_width() + 7) / 8;
    return use_seed ? expansion_height_ * cfg_.l_ks * (pp + 32) : expansion_height_ * cfg_.l_ks * 2 * pp;
  }
  const SchemeConfig &config() const { return cfg_; }
  // composite CRT limbs and the NTT root registered for Q (pir.cpp:14-47)
  u64 q1() const { return q1_; }
  u64 q2() const { return q2_; }
  u64 ntt_root() const { return w_crt_; }

 private:
  void apply_layout(const PirLayoutConfig &layout);
  SchemeConfig cfg_;
  std::vector<u64> rns_mods_;
  u64 q1_ = 0, q2_ = 0, w_crt_ = 0;
  u64 plain_mod_ = 0, small_q_ = 0;
  size_t base_log2_ = 0, base_log2_key_ = 0;
  size_t target_num_pt_ = 0, expansion_height_ = 0, num_pt_ = 0, fst_dim_sz_ = 0, num_dims_ = 0;
  bool fst_dim_pow2_ = true;
};

// src/includes/rlwe.h:17-49
struct RlweCt {
  std::vector<u64> c0, c1;
  bool ntt_form = false;
  u64 *data(size_t i) { return i == 0 ? c0.data() : c1.data(); }
  const u64 *data(size_t i) const { return i == 0 ? c0.data() : c1.data(); }
  size_t poly_size() const { return c0.size(); }
};
struct RlweSk { std::vector<u64> data; };
struct RlwePt { std::vector<u64> data; size_t coeff_count() const { return data.size(); } };
using GSWCt = std::vector<std::vector<u64>>;  // src/includes/gsw.h:17

namespace bvks {  // src/includes/bv_keyswitch.h:48-66
struct BvRlweCt { std::vector<u64> a, b; };
struct BvKeySwitchKey { uint32_t galois_k = 0; std::vector<BvRlweCt> cts; };
class BvGaloisKeys {
 public:
  std::vector<BvKeySwitchKey> keys;
  const BvKeySwitchKey &get(uint32_t galois_k) const;
};
}  // namespace bvks

// src/includes/pir_session.h:8-18
struct PirSessionKeys {
  bvks::BvGaloisKeys bv_galois_keys;
  GSWCt gsw_key;
};
using SharedPirSessionKeys = std::shared_ptr<const PirSessionKeys>;

// CPU polynomial arithmetic mod one modulus with the reference's NTT
// (negacyclic, bit-reversed evaluation order). The composite Q runs as two
// 29-bit limbs and is recombined by CRT.
class RingMod {
 public:
  RingMod(size_t N, u64 q, u64 root, u64 q1 = 0, u64 q2 = 0);
  void ntt_fwd(u64 *a) const;
  void ntt_inv(u64 *a) const;
  void mul(const u64 *a, const u64 *b, u64 *out) const;  // pointwise
  u64 q() const { return q_; }

 private:
  struct Limb {
    u32 q;
    std::vector<u32> w, ws, iw, iws;
    u32 ninv, ninvs;
  };
  void limb_fwd(const Limb &L, u32 *a) const;
  void limb_inv(const Limb &L, u32 *a) const;
  size_t N_;
  u64 q_;
  bool split_;
  std::vector<Limb> limbs_;
  u64 q1inv_mod_q2_ = 0;
  // direct path (q < 2^62 prime, not split)
  std::vector<u64> w64_, iw64_;
  u64 ninv64_ = 0;
};

using PlaintextSource = std::function<void(size_t logical_index, RlwePt &out)>;

class PirClient {
 public:
  PirClient(const PirParams &params, size_t client_id, u64 seed);
  // Seeded client: every random polynomial (secret, uniform halves, errors)
  // comes from the counter-based sampler of prg.h keyed by `key`, so a server
  // can regenerate the uniform halves of its keys and queries from seeds
  // (the wire codec of codec.h) and the device can generate the same keys
  // and queries (opir_client_keygen_device, opir_query_gen_device).
  PirClient(const PirParams &params, size_t client_id, const std::array<uint8_t, 32> &key);
  bool seeded() const { return prg_; }
  const std::array<uint8_t, 32> &key_seed() const { return key_; }
  // Seed of the last query a seeded client generated (its c1 is
  // uniform(seed, kQueryA, 0, limb)).
  const std::array<uint8_t, 32> &last_query_seed() const { return last_qseed_; }
  RlweCt fast_generate_query(size_t pt_idx);
  RlweCt fast_generate_query(const PirParams &query_params, size_t pt_idx);
  void add_gsw_to_query(const PirParams &qp, RlweCt &query, const std::vector<size_t> &idx);
  // fast_generate_query with explicit query indices: {first-dim column,
  // selector bits...} of layout qp (what get_query_indices yields for a
  // plaintext index; tree queries pass alpha, beta and gamma bits).
  RlweCt generate_query_for_indices(const PirParams &qp, const std::vector<size_t> &idx);
  RlweCt fresh_zero_ct();  // client.cpp:290-299
  bvks::BvGaloisKeys create_bv_galois_keys();
  GSWCt generate_gsw_from_key();
  SharedPirSessionKeys create_session_keys();
  size_t get_client_id() const { return client_id_; }
  RlweCt load_resp_from_stream(std::stringstream &stream) const;
  RlwePt decrypt_mod_q(const RlweCt &ct, int *budget = nullptr) const;
  std::vector<size_t> get_query_indices(const PirParams &qp, size_t pt_idx) const;
  const RlweSk &secret_key() const { return sk_; }

 private:
  void encrypt_zero(RlweCt &ct);
  void encrypt_zero_seeded(RlweCt &ct, const std::array<uint8_t, 32> &qseed);
  void shared_error(uint32_t domain, uint32_t i, std::vector<u64> &e) const;  // [K][N], same integer per limb
  size_t client_id_;
  PirParams params_;
  std::mt19937_64 rng_;
  u64 seed_;
  std::vector<u64> qs_;                           // RNS limb moduli (one logical Q for K = 1)
  std::vector<std::shared_ptr<RingMod>> rings_;   // one negacyclic NTT ring per limb
  RlweSk sk_;                                     // K*N values, NTT form per limb
  bool prg_ = false;                              // seeded client (counter-based sampler)
  std::array<uint8_t, 32> key_{}, last_qseed_{};
  uint64_t query_ctr_ = 0;
  std::vector<uint64_t> cdt_;                     // error table for the seeded sampler
};

// Owns one GPU context (scheme tables, device keys, stream); shareable by
// several PirServer objects of scheme-compatible layouts.
class GpuScheme {
 public:
  explicit GpuScheme(const PirParams &params, int device = 0);
  ~GpuScheme();
  GpuScheme(const GpuScheme &) = delete;
  GpuScheme &operator=(const GpuScheme &) = delete;
  opir_ctx *ctx() const { return ctx_; }
  const PirParams &params() const { return params_; }
  void check(int rc) const;  // throws the reference's exception type
  void set_client_keys(size_t client_id, const bvks::BvGaloisKeys &bv, const GSWCt &gsw);

 private:
  opir_ctx *ctx_ = nullptr;
  PirParams params_;
};

class PirServer {
 public:
  explicit PirServer(const PirParams &params, int device = 0);
  PirServer(const PirParams &params, std::shared_ptr<GpuScheme> scheme);
  ~PirServer();
  PirServer(const PirServer &) = delete;
  PirServer &operator=(const PirServer &) = delete;

  void gen_data(const std::vector<size_t> &record_indices = {}, u64 seed = 0);
  // gen_data with the plaintexts generated on the GPU (no host copy of the
  // database): coefficient c of plaintext i = splitmix64(seed + i*N + c) mod t.
  void gen_data_device(const std::vector<size_t> &record_indices = {}, u64 seed = 0);
  void load_data(size_t logical_num_pt, const PlaintextSource &source,
                 const std::vector<size_t> &record_indices = {});
  RlweCt make_query(size_t client_id, RlweCt &query);
  std::vector<RlweCt> make_queries(const std::vector<size_t> &client_ids, const std::vector<RlweCt> &queries);
  std::vector<RlweCt> expand_query(size_t client_id, RlweCt &query) const;
  std::vector<GSWCt> complete_selectors(size_t client_id, const std::vector<RlweCt> &expanded);
  std::vector<RlweCt> evaluate_first_dim(std::vector<RlweCt> &selection_vector);
  RlweCt evaluate_other_dim(std::vector<RlweCt> &intermediate_db, std::vector<GSWCt> &selectors);
  // Homomorphic select: result = x when selection_cipher encrypts 0, y when
  // it encrypts 1 (server.h:96-99). x and y stay unchanged; result may be x.
  void ext_prod_mux(RlweCt &x, RlweCt &y, GSWCt &selection_cipher, RlweCt &result);
  size_t save_resp_to_stream(const RlweCt &response, std::stringstream &resp_stream);
  void set_client_bv_galois_key(size_t client_id, bvks::BvGaloisKeys bv_keys);
  void set_client_gsw_key(size_t client_id, GSWCt gsw_key);
  void set_client_session_keys(size_t client_id, SharedPirSessionKeys keys);
  // Seed-compressed wire forms (codec.h): keys whose uniform halves the
  // device regenerates (they live on the device only), and a query whose c1
  // is regenerated from its seed.
  void set_client_keys_wire(size_t client_id, const std::string &wire);
  // Keys of seeded clients (PirClient(params, id, key)) generated on the GPU
  // (opir_client_keygen_device), for load generation; device-only like wire keys.
  void client_keygen_device(const std::vector<size_t> &client_ids, const std::vector<std::array<uint8_t, 32>> &keys);
  RlweCt make_query_wire(size_t client_id, const std::string &wire);
  SharedPirSessionKeys client_session_keys(size_t client_id) const;
  RlwePt direct_get_original_plaintext(size_t index) const;
  const PirParams &get_params() const { return params_; }
  opir_db *gpu_db() const { return db_; }
  std::shared_ptr<GpuScheme> gpu_scheme() const { return scheme_; }

 private:
  void push_keys(size_t client_id);
  PirParams params_;
  std::shared_ptr<GpuScheme> scheme_;
  opir_db *db_ = nullptr;
  std::map<size_t, bvks::BvGaloisKeys> bv_keys_;
  std::map<size_t, GSWCt> gsw_keys_;
  std::map<size_t, SharedPirSessionKeys> sessions_;
  std::set<size_t> wire_keyed_;  // clients keyed through set_client_keys_wire
  std::unordered_map<size_t, RlwePt> recorded_;
};

// Free helpers of the reference (src/utils.cpp)
namespace utils {
bool is_prime(u64 n);
u64 generate_prime(size_t bit_width);
std::vector<u64> generate_ntt_friendly_primes(const std::vector<size_t> &bit_widths, size_t N);
u64 minimal_primitive_root(u64 degree, u64 q);
std::pair<size_t, size_t> calculate_db_shape(size_t target_num_pt, size_t l, size_t h, bool fst_dim_pow2);
size_t bit_reverse(size_t x, size_t logn);
bool try_invert_uint_mod(u64 value, u64 modulus, u64 &result);
u64 crt_combine(u64 w1, u64 q1, u64 w2, u64 q2);
std::vector<std::vector<u64>> gsw_gadget(size_t l, u64 base_log2, const std::vector<u64> &rns_mods);
}  // namespace utils

}  // namespace onionpir
