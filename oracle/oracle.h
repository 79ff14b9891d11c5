// OnionPIRv2 CPU oracle -- TEST INFRASTRUCTURE ONLY.
//
// A plain, single-threaded C++ restatement of the reference's server answer
// path (PirServer::make_query, src/server.cpp:895-947) and the client pieces
// that produce its inputs and check its outputs. It exists to check the HIP
// library, never to be shipped or measured as the product: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
//
// Every function names the reference file:line it follows. Arithmetic is
// written the direct way (u128 products, `%` reductions, O(n log n) textbook
// NTT loops) so that it is easy to audit against the reference, not fast.
//
// Parity status of each part is stated in the header of oracle_pin tests
// (tests/test_oracle_*.py); the NTT evaluation ORDER (HEXL's bit-reversed
// output convention) is the one item that no reference test pins.
#pragma once
#include <cstddef>
#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <random>
#include <string>
#include <unordered_map>
#include <vector>

namespace oracle {

using u64 = uint64_t;
using u128 = unsigned __int128;
using i128 = __int128;

// ---------------------------------------------------------------------------
// Compile-time configs of the reference, as runtime values.
// src/includes/database_constants.h:47-124
// ---------------------------------------------------------------------------
struct SchemeConfig {
  std::string name;
  size_t poly_degree;
  size_t l_ep, l_key, l_ks;
  size_t tree_height;
  size_t plain_mod_bits;
  size_t small_q_width;
  std::vector<size_t> rns_mod_bits;       // DBConsts::RnsMods
  bool composite_first_dim;               // DBConsts::CompositeFirstDim
  std::vector<size_t> first_dim_rns_bits; // DBConsts::FirstDimRNSMods
  double sigma = 2.55;                    // DBConsts::NoiseStdDev
  size_t db_size_mb = 128;                // DBConsts::DB_SIZE_MB
  bool fst_dim_pow2 = true;               // DBConsts::FST_DIM_POW2
};
SchemeConfig config_by_name(const std::string &name);

// ---------------------------------------------------------------------------
// Number theory helpers. src/utils.cpp:92-287
// ---------------------------------------------------------------------------
bool is_prime(u64 n);                                         // utils.cpp:136-153
bool try_invert_uint_mod(u64 value, u64 modulus, u64 &out);   // utils.cpp:155-163
u64 generate_prime(size_t bit_width);                         // utils.cpp:240-251
std::vector<u64> generate_ntt_friendly_primes(const std::vector<size_t> &bws, size_t N); // utils.cpp:253-287
u64 crt_combine(u64 w1, u64 q1, u64 w2, u64 q2);              // utils.cpp:229-238
u64 minimal_primitive_root(u64 degree, u64 q);                // HEXL MinimalPrimitiveRoot (published def.)
u64 mulmod(u64 a, u64 b, u64 m);
u64 powmod(u64 b, u64 e, u64 m);
std::pair<size_t, size_t> calculate_db_shape(size_t target, size_t l, size_t h, bool pow2); // utils.cpp:376-401
size_t bit_reverse(size_t x, size_t logn);                    // utils.h:277-285
u64 rescale(u64 a, u64 inp_mod, u64 out_mod);                 // utils.cpp:445-462

// Barrett reducers. src/includes/utils.h:126-203
struct BarrettU64 { u64 q, mu; };
BarrettU64 barrett_u64_setup(u64 q);
u64 barrett_reduce_u64(u64 x, const BarrettU64 &b);
struct BarrettU128 { u64 q, mu_lo, mu_hi; };
BarrettU128 barrett_u128_setup(u64 q);
u64 barrett_reduce_u128(u128 x, const BarrettU128 &b);

// ---------------------------------------------------------------------------
// Negacyclic NTT with the HEXL convention (root powers stored bit-reversed,
// Cooley-Tukey forward to bit-reversed order, Gentleman-Sande inverse, scale
// by n^-1). Works for any odd q < 2^62 with a given primitive 2N-th root,
// which covers the composite Q = q1*q2 of CONFIG_N2048_K1_COMP.
// src/utils.cpp:13-56,84-90
// ---------------------------------------------------------------------------
struct NttTable {
  size_t N = 0;
  u64 q = 0, root = 0, inv_n = 0;
  std::vector<u64> psi_brv;      // psi_brv[bitrev(i)] = root^i
  std::vector<u64> psi_inv_brv;  // psi_inv_brv[bitrev(i)] = root^-i
};
const NttTable &ntt_table(size_t N, u64 q);   // cached per (N,q); root from registry or minimal root
void register_ntt_root(size_t N, u64 q, u64 root);  // utils.cpp:47-56
void ntt_fwd(u64 *a, size_t N, u64 q);
void ntt_inv(u64 *a, size_t N, u64 q);

void automorphism_coeff(const u64 *in, size_t N, uint32_t k, u64 q, u64 *out);  // utils.cpp:58-72
void automorphism_ntt(const u64 *in, size_t N, uint32_t k, u64 q, u64 *out);    // utils.cpp:74-82
void negacyclic_shift(const u64 *poly, size_t N, size_t shift, u64 q, u64 *out); // utils.cpp:166-186
std::vector<std::vector<u64>> gsw_gadget(size_t l, u64 base_log2, const std::vector<u64> &mods); // utils.cpp:206-222

// ---------------------------------------------------------------------------
// Runtime parameters. src/pir.cpp:14-186, src/includes/pir.h
// ---------------------------------------------------------------------------
struct PirParams {
  SchemeConfig cfg;
  std::vector<u64> rns_mods;
  std::vector<size_t> rns_mod_bits;
  // composite
  bool composite = false;
  u64 q1 = 0, q2 = 0, w1 = 0, w2 = 0, w_crt = 0, q1_inv_mod_q2 = 0;
  // K=2 tables
  u64 q0_inv_mod_q1 = 0;
  std::vector<u64> r64_mod_q;
  u64 plain_mod = 0, small_q = 0;
  size_t base_log2 = 0, base_log2_key = 0;
  // layout
  size_t target_num_pt = 0, expansion_height = 0, num_pt = 0, fst_dim_sz = 0, num_dims = 0;
  bool fst_dim_pow2 = true;

  explicit PirParams(const SchemeConfig &c);
  PirParams with_layout(size_t target, size_t h, bool pow2) const;
  PirParams with_query_shape(size_t fst_dim_sz, size_t num_selector_bits, size_t h) const;
  bool scheme_compatible(const PirParams &o) const;

  size_t N() const { return cfg.poly_degree; }
  size_t K() const { return rns_mods.size(); }
  size_t ct_mod_width() const;
  size_t pt_size() const { return (cfg.plain_mod_bits - 1) * N() / 8; }
  size_t other_dim_sz() const { return num_pt / fst_dim_sz; }
  size_t l_ep() const { return cfg.l_ep; }
  size_t l_key() const { return cfg.l_key; }
  size_t l_ks() const { return cfg.l_ks; }
  size_t bv_base_log2() const { return ct_mod_width() / cfg.l_ks + 1; }  // bv_keyswitch.cpp:73-76
  double sigma() const { return cfg.sigma; }

 private:
  void apply_layout(size_t target, size_t h, bool pow2);
};

// ---------------------------------------------------------------------------
// RLWE types and primitives. src/includes/rlwe.h, src/rlwe.cpp
// ---------------------------------------------------------------------------
struct RlweCt {
  std::vector<u64> c0, c1;  // K*N each, limb-major
  bool ntt_form = false;
  u64 *data(size_t i) { return i == 0 ? c0.data() : c1.data(); }
  const u64 *data(size_t i) const { return i == 0 ? c0.data() : c1.data(); }
};
struct RlweSk { std::vector<u64> data; };  // NTT form, limb-major
using GSWCt = std::vector<std::vector<u64>>;  // 2l rows x [c0 K*N][c1 K*N]

void sample_gaussian_shared(u64 *out, size_t N, const std::vector<u64> &qs, double sigma, std::mt19937_64 &rng);
void sample_uniform_poly(u64 *out, size_t N, u64 q, std::mt19937_64 &rng);  // utils.cpp:421-430
RlweSk gen_secret_key_rns(size_t N, const std::vector<u64> &qs, std::mt19937_64 &rng);  // rlwe.cpp:219-228
void encrypt_zero_rns(const RlweSk &sk, size_t N, const std::vector<u64> &qs, double sigma,
                      std::mt19937_64 &rng, RlweCt &ct, bool ntt_form);           // rlwe.cpp:230-276
void encrypt_bfv_rns(const std::vector<u64> &m, const RlweSk &sk, size_t N, const std::vector<u64> &qs,
                     u64 t, double sigma, std::mt19937_64 &rng, RlweCt &ct);      // rlwe.cpp:278-308
void decrypt_rns(const RlweCt &ct, const RlweSk &sk, size_t N, const std::vector<u64> &qs, u64 t,
                 u64 q0_inv_mod_q1, std::vector<u64> &pt);                         // rlwe.cpp:310-379
int decrypt_and_budget(const RlweCt &ct, const RlweSk &sk, size_t N, u64 q, u64 t, std::vector<u64> &pt); // rlwe.cpp:83-113

// ---------------------------------------------------------------------------
// RGSW. src/gsw.cpp
// ---------------------------------------------------------------------------
struct GSWEval {
  const PirParams *p;
  size_t l, base_log2;
  GSWEval(const PirParams &pp, size_t l_, size_t b) : p(&pp), l(l_), base_log2(b) {}
  void decomp_single_mod(const RlweCt &ct, std::vector<std::vector<u64>> &out) const;  // gsw.cpp:252-295
  void decomp_mp(const RlweCt &ct, std::vector<std::vector<u64>> &out) const;          // gsw.cpp:184-250
  void decomp_to_ntt(std::vector<std::vector<u64>> &rows) const;                       // gsw.cpp:297-315
  void external_product(const GSWCt &g, const RlweCt &bfv, RlweCt &res) const;          // gsw.cpp:119-182
  void gsw_ntt_forward(GSWCt &g) const;                                                // gsw.cpp:105-117
  void query_to_gsw(std::vector<RlweCt> query, const GSWCt &key, GSWCt &out) const;    // gsw.cpp:317-357
  GSWCt plain_to_gsw(const std::vector<u64> &pt, const RlweSk &sk, std::mt19937_64 &rng) const; // gsw.cpp:359-422
};

// ---------------------------------------------------------------------------
// BV keyswitch. src/bv_keyswitch.cpp
// ---------------------------------------------------------------------------
struct BvRow { std::vector<u64> a, b; };  // NTT form, K*N each
struct BvKey { uint32_t galois_k = 0; std::vector<BvRow> rows; };
struct BvGaloisKeys {
  std::vector<BvKey> keys;
  const BvKey &get(uint32_t k) const;  // bv_keyswitch.cpp:92-98
};
void signed_gadget_decompose(u64 val, size_t base_log2, u64 q, u64 *out, size_t nd);          // bv_keyswitch.cpp:20-40
void signed_gadget_decompose_mp(u128 val, u128 Q, size_t base_log2, int64_t *out, size_t nd); // bv_keyswitch.cpp:42-64
u64 drop_q1(u64 r0, u64 r1, u64 q0, u64 q1, u64 q0_inv_mod_q1);                     // server.cpp:1053-1061
BvKey gen_bv_ks_key(const PirParams &p, const RlweSk &sk, uint32_t galois_k, std::mt19937_64 &rng); // 222-259
BvGaloisKeys gen_bv_galois_keys(const PirParams &p, const RlweSk &sk, u64 seed);                     // 261-276
void bv_apply_galois_inplace(RlweCt &ct, uint32_t galois_k, const BvKey &key, const PirParams &p);  // 296-500

// ---------------------------------------------------------------------------
// First-dimension mat-mat. src/matrix.cpp
// ---------------------------------------------------------------------------
void level_mat_mat_32(const uint32_t *A, const uint32_t *B, u64 *out, size_t m, size_t n,
                      size_t levels, u64 q);                                      // matrix.cpp:315-350
void level_mat_mat_u64(const u64 *A, const u64 *B, u128 *out, size_t m, size_t n,
                       size_t levels, const u64 *level_qs);                       // matrix.cpp:36-151

// ---------------------------------------------------------------------------
// Client. src/client.cpp
// ---------------------------------------------------------------------------
struct PirClient {
  size_t client_id;
  PirParams params;
  std::mt19937_64 rng;
  u64 seed;
  RlweSk sk;
  PirClient(const PirParams &p, size_t id, u64 seed);  // client.cpp:20-23 (seeded)
  std::vector<size_t> get_query_indices(const PirParams &qp, size_t pt_idx) const;  // client.cpp:48-98
  RlweCt fast_generate_query(const PirParams &qp, size_t pt_idx);                   // client.cpp:107-166
  void add_gsw_to_query(const PirParams &qp, RlweCt &query, const std::vector<size_t> &idx); // 174-225
  BvGaloisKeys create_bv_galois_keys();          // client.h:38-40 (seeded)
  GSWCt generate_gsw_from_key();                 // client.cpp:25-38
  std::vector<u64> decrypt_mod_q(const RlweCt &ct, int *budget) const;  // client.cpp:364-410
  RlweCt load_resp_from_stream(const std::string &bytes) const;         // client.cpp:315-361
  std::vector<u64> decrypt_ct(const RlweCt &ct) const;                  // client.cpp:280-288
};

// ---------------------------------------------------------------------------
// Server. src/server.cpp
// ---------------------------------------------------------------------------
using PlaintextSource = std::function<void(size_t, std::vector<u64> &)>;

struct StageTimes {  // reference stage names, src/includes/logging.h:122-134
  double expand_ms = 0, convert_ms = 0, first_dim_ms = 0, other_dim_ms = 0, mod_switch_ms = 0;
};

struct PirServer {
  PirParams params;
  size_t num_pt;
  std::vector<uint32_t> db_lo, db_hi;   // composite: [coeff * num_pt + pt]
  std::vector<u64> db_u64;              // non-composite: [(limb*N + coeff) * num_pt + pt]
  std::map<size_t, BvGaloisKeys> bv_keys;
  std::map<size_t, GSWCt> gsw_keys;
  std::unordered_map<size_t, std::vector<u64>> recorded;
  StageTimes last_times;

  explicit PirServer(const PirParams &p);  // server.cpp:29-46
  void load_data(size_t logical_num_pt, const PlaintextSource &src, const std::vector<size_t> &record); // 80-203
  void set_client_keys(size_t id, BvGaloisKeys bv, GSWCt gsw);
  std::vector<RlweCt> fast_expand_qry(size_t id, const RlweCt &query) const;            // 731-816
  std::vector<GSWCt> complete_selectors(size_t id, const std::vector<RlweCt> &expanded) const; // 823-855
  std::vector<RlweCt> evaluate_first_dim(std::vector<RlweCt> &q);                       // 313-405
  RlweCt evaluate_other_dim(std::vector<RlweCt> &mid, std::vector<GSWCt> &sel) const;   // 599-649
  void ext_prod_mux(RlweCt &x, RlweCt &y, const GSWCt &sel, RlweCt &res) const;         // 652-722
  void mod_switch_inplace(RlweCt &ct, u64 small_q) const;                               // 1025-1075
  RlweCt make_query(size_t id, const RlweCt &query);                                    // 895-947
  std::string save_resp_to_stream(const RlweCt &resp) const;                            // 950-1009
};

size_t small_q_width(u64 small_q);  // ceil(log2(small_q)), server.cpp:963-964

// ---------------------------------------------------------------------------
// Tree PIR. src/tree_index.cpp, src/tree_query.cpp, src/tree_select.cpp
// ---------------------------------------------------------------------------
struct TreeParams {
  size_t L = 0, N = 0, n = 0, r = 0, a = 0, N0 = 0, P = 0, B = 0, b = 0, ell = 0, w = 0, W = 0, h_q = 0;
  u64 t = 0;
};
struct TreeCoords {
  size_t alpha = 0;
  std::vector<uint8_t> beta, gamma;  // little-endian bits
};
struct TreePlan {
  size_t level = 0, R = 0;
  int select_case = 0;  // 0 Single, 1 CoarsenedAlpha, 2 AlphaBeta
  size_t coarsen = 0, beta_begin = 0, beta_count = 0;
};
struct TreeLevelOracle {
  size_t node = 0, plaintext = 0, position = 0;
};
TreeParams make_tree_params(size_t L, size_t a, const PirParams &scheme);   // tree_query.cpp:44-61
TreeCoords tree_client_index(size_t leaf, const TreeParams &p);            // tree_index.cpp:150-170
std::vector<TreePlan> tree_level_plans(const TreeParams &p);               // tree_index.cpp:172-215
u64 synthetic_tree_node(size_t level, size_t index, u64 t);                // tree_select.cpp:41-50
std::vector<std::vector<u64>> pack_tree_level(size_t level, const TreeParams &p);  // tree_select.cpp:52-74
TreeLevelOracle tree_level_oracle(size_t leaf, size_t level, const TreeParams &p); // tree_index.cpp:217-241
RlweCt make_tree_query(PirClient &c, const TreeParams &tp, const PirParams &qview, size_t leaf); // tree_query.cpp:140-196
RlweCt tree_pt_ct_mul(const std::vector<u64> &pt, const RlweCt &ct, const PirParams &p);        // tree_select.cpp:86-119
// unpack_tree_query + build_alpha_pyramid + select_level for every level 0..L
std::vector<RlweCt> tree_select_all(PirServer &srv, const TreeParams &tp, size_t id, const RlweCt &query,
                                    const std::vector<std::vector<std::vector<u64>>> &levels);

}  // namespace oracle
