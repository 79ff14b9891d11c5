// Internal launch interface between the C-ABI layer and the HIP kernels.
// Ciphertext layout on device ("RNS ct"): [poly p][limb][N] u32, 4*N words;
// NTT-form polynomials are stored in HEXL evaluation order (bit-reversed).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "opir_device.h"
#include "opir_k2.h"

namespace opir {

constexpr size_t kCtWords = 4 * kN;  // one RNS ciphertext

struct GadgetCfg {
  int l_ep, b_ep;     // data selectors (external product in the MUX tree)
  int l_key, b_key;   // RGSW(s) completion key
  int l_ks, b_ks;     // BV keyswitch
};

struct ExpandArgs {
  const u32 *in;
  u32 *out;
  size_t in_stride, out_stride;  // words between queries
  const u32 *const *bv_keys;     // per query
  size_t key_level_words;        // words per level in a client's BV key
  int level;                     // k = 2^level, galois = N/k + 1
  int kinv;                      // galois^-1 mod 2N
  int l_ks, b_ks;
  int bv_only = 0;               // 1: out node 0 receives the keyswitched c' alone (no children)
};

struct CompleteArgs {
  const u32 *leaves;
  u32 *sel;
  size_t leaf_stride, sel_stride;
  const u32 *const *gsw_keys;
  int l_ep, l_key, b_key;
};

struct MuxArgs {
  const u32 *in;
  u32 *out;
  size_t stride;       // words between queries (in and out)
  const u32 *sel;      // this level's selector for query 0
  size_t sel_stride;   // words between queries
  int offset, half;    // offset >= 0: ragged level; else perfect level with `half`
  int l, b;
};

// Per-query pointers for batched launches (device arrays of length B).
struct BatchPtrs {
  const u32 *const *bv_keys;   // client BV keys: [h][l_ks][2 (b,a)][2 limbs][N]
  const u32 *const *gsw_keys;  // client RGSW(s): [2 l_key][2][2][N]
};

void launch_split_u64(const u64 *in, u32 *out, size_t npoly, const SchemeConst &S, hipStream_t st);
// B groups of npoly polynomials: in + b*in_stride (u64 words) -> out + b*out_stride (u32 words).
void launch_split_u64_batch(const u64 *in, size_t in_stride, u32 *out, size_t out_stride, size_t npoly, int B,
                            const SchemeConst &S, hipStream_t st);
void launch_join_u64(const u32 *in, u64 *out, size_t npoly, const SchemeConst &S, hipStream_t st);
void launch_ntt_u64(u64 *data, size_t npoly, bool inverse, const SchemeConst &S, hipStream_t st);
void launch_ntt_rns(const u32 *in, u32 *out, size_t npoly_rns, bool inverse, const SchemeConst &S,
                    hipStream_t st);

// Database geometry: pt = row * N0 + k; rows padded to 16 (nrt tiles), k to
// 64 (nkb blocks); nlv NTT levels per limb; P balanced int8 digit planes per
// residue (4 for 29-bit limbs, 8 for 60-bit limbs). Device layout: digit
// planes in MFMA A-fragment order, see opir_scan.hip.
struct DbGeom {
  size_t N0 = 0, Nrest = 0, nrt = 0, nkb = 0;
  int P = 4, nlv = kN;
  int nl = 2;  // RNS limbs (one scan per limb)
};
struct QfragGeom {
  int N0 = 0, nkb = 0, nct = 0, ncm = 0;  // column tiles of 16, stored columns per tile
  int P = 4, nlv = kN;
  int nl = 2;
};
DbGeom make_db_geom(size_t N0, size_t Nrest, int P = 4, int nlv = kN, int nl = 2);
size_t db_bytes(const DbGeom &g);
QfragGeom make_qfrag_geom(size_t N0, int B, int P = 4, int nlv = kN, int nl = 2);
size_t qfrag_bytes(const QfragGeom &q);

// Seeded synthetic plaintexts [count][n]: coefficient c of plaintext i is
// splitmix64(seed + i*n + c) mod t.
void launch_gen_plain_u16(uint16_t *out, size_t first_pt, size_t count, int n, u64 seed, u64 t, hipStream_t st);
void launch_gen_plain_u64(u64 *out, size_t first_pt, size_t count, int n, u64 seed, u64 t, hipStream_t st);

// DB preprocessing (NTT of plaintexts, then scatter into the scan layout).
void launch_db_ntt_plain(const uint16_t *pts, u32 *tmp, size_t count, const SchemeConst &S, hipStream_t st);
void launch_db_scatter(const u32 *tmp, int8_t *db, size_t first_pt, size_t count, const DbGeom &g, hipStream_t st);
void launch_db_from_ref(const u32 *ref, int8_t *db, const DbGeom &g, hipStream_t st);

// Ciphertext blocks of the scratch for split (narrow) tree levels.
constexpr size_t kMuxSplitMaxBlocks = 1024;

// Expansion level: nodes [0, n) of input buffer -> children 2r, 2r+1 of output buffer.
// part (kMuxSplitMaxBlocks cts, or null): scratch for spreading a narrow
// level's digits over several workgroups.
void launch_expand_level(const u32 *in, u32 *out, size_t in_stride, size_t out_stride, int nnodes, int B,
                         int level, const BatchPtrs &P, const GadgetCfg &G, const SchemeConst &S, hipStream_t st,
                         u32 *part);
// Stand-alone BV keyswitch (bv_keyswitch.cpp:296-367) of one coefficient-form
// RNS ct: out = (sigma_k(c0) + sum d_i b_i, sum d_i a_i), d_i the balanced
// digits of sigma_k(c1); key_ptr: device array holding one pointer to l_ks
// device key rows [b (-> c0), a (-> c1)][2 limbs][N]; kinv = k^-1 mod 2N.
void launch_bv_apply(const u32 *in, u32 *out, const u32 *const *key_ptr, int kinv, const GadgetCfg &G,
                     const SchemeConst &S, hipStream_t st);
// Forward NTT of ciphertexts (both polys): in coef RNS cts -> out NTT RNS cts.
void launch_ct_ntt(const u32 *in, u32 *out, size_t in_stride, size_t out_stride, int count, int B,
                   const SchemeConst &S, hipStream_t st);
// Selector completion: rows = nsel * l_ep expanded leaves (coef form) -> sel [nsel][2 l_ep][2][2][N].
// part (kMuxSplitMaxBlocks cts, or null): scratch for spreading the rows of
// a small batch over digit groups.
void launch_complete(const u32 *leaves, u32 *sel, size_t leaf_stride, size_t sel_stride, int nrows, int B,
                     const BatchPtrs &P, const GadgetCfg &G, const SchemeConst &S, hipStream_t st,
                     u32 *part = nullptr);
// First-dim scan in two launches: launch_qfrag turns the selectors qntt
// [B][N0][2][2][N] into MFMA B fragments (qfrag_bytes(make_qfrag_geom(N0, B))
// bytes), launch_scan multiplies the DB by them -> mid [B][Nrest][2][2][N]
// (NTT form). variant: index into the scan's tile-shape table (-1 = the
// default for B); returns -1 if that shape cannot take B queries.
void launch_qfrag(const u32 *qntt, int8_t *qfrag, size_t q_stride, int B, const DbGeom &g, hipStream_t st);
int launch_scan(const int8_t *db, const int8_t *qfrag, u32 *mid, size_t mid_stride, int B, const DbGeom &g,
                const SchemeConst &S, hipStream_t st, int variant);
// Number of scan tile shapes for P digit planes (4: K = 1, 8: K = 2), and
// whether shape `variant` can take B queries on geometry g.
int scan_variant_count(int P);
bool scan_variant_ok(int P, int variant, int B, const DbGeom &g);
// The same for the two-prime scheme (P = 8 planes, u64 residues):
// tmp [limb][count][nlv] u64 -> DB; qntt [B][N0][2][2][nlv] u64 -> qfrag;
// mid [B][Nrest][2][2][nlv] u64.
void launch_db_scatter64(const u64 *tmp, int8_t *db, size_t first_pt, size_t count, const DbGeom &g, hipStream_t st);
void launch_qfrag64(const u64 *qntt, int8_t *qfrag, size_t q_stride, int B, const DbGeom &g, hipStream_t st);
int launch_scan64(const int8_t *db, const int8_t *qfrag, u64 *mid, size_t mid_stride, int B, const DbGeom &g,
                  const k2::Scheme64 &S, hipStream_t st, int variant);
// MUX level. ragged: slot s < offset copies in[s]; s >= offset muxes in[off+2j], in[off+2j+1].
// perfect (offset < 0): slot s muxes in[s], in[s+half].
void launch_mux(const u32 *in, u32 *out, size_t stride, int nslots, int offset, int half, int B,
                const u32 *sel, size_t sel_stride, const GadgetCfg &G, const SchemeConst &S, hipStream_t st);
// The same level with each node's digits spread over 2G workgroups
// (partials in part, kMuxSplitMaxBlocks ciphertexts), then combined; for
// levels with few nodes. mux_split_groups returns G, or 0 when the level is
// too wide for the split (then launch_mux).
int mux_split_groups(int nslots, int offset, int B, const GadgetCfg &G);
void launch_mux_split(const u32 *in, u32 *out, size_t stride, int nslots, int offset, int half, int B,
                      const u32 *sel, size_t sel_stride, u32 *part, int Gs, const GadgetCfg &G, const SchemeConst &S,
                      hipStream_t st);
// INTT + CRT + centred rescale to small_q: res [B][ct] NTT form -> out [B][2][N] u64.
void launch_final(const u32 *res, u64 *out, size_t res_stride, int B, const SchemeConst &S, hipStream_t st);
// NTT-form RNS ct -> coefficient-form u64 mod Q ct (test entry points).
void launch_ct_to_coef_u64(const u32 *in, u64 *out, size_t count, const SchemeConst &S, hipStream_t st);
// Stand-alone external product (test entry point).
void launch_ext_prod(const u32 *gsw, int l, int b, const u32 *ct_coef, u32 *out_ntt, const SchemeConst &S,
                     hipStream_t st);
// out = a - b (sub) or a + b per residue, n values laid out as [..][K][limb_len]
// with modulus q[k] for limb k (K = 1 or 2; a, b already reduced, q < 2^63).
// out may alias a or b.
void launch_addsub_mod(const u64 *a, const u64 *b, u64 *out, size_t n, size_t limb_len, int K, u64 q0, u64 q1,
                       bool sub, hipStream_t st);

}  // namespace opir
