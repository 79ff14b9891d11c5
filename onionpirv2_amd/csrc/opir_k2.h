// Launch interface of the 64-bit-limb answer path. Two schemes run on it:
//   CONFIG_N4096_K2_MP: N = 4096, Q = q0 * q1 with two ~60-bit NTT primes,
//     the whole pipeline in two RNS limbs with the multi-precision gadgets
//     (src/server.cpp:205-262,363-527; src/gsw.cpp:22-85,184-250;
//     src/bv_keyswitch.cpp:373-481; src/server.cpp:1044-1075);
//   CONFIG_N2048_K1: N = 2048, one ~60-bit NTT prime q, the single-modulus
//     gadgets (src/gsw.cpp:252-315; src/bv_keyswitch.cpp:20-40,296-367) and
//     the direct centred rescale (src/server.cpp:1034-1043).
//
// Device ciphertext layout: [poly][limb][N] u64, 2*K*N values, each limb in
// [0, q_limb); NTT-form polynomials in HEXL evaluation order, as the
// reference stores them (c0 = [limb0 N][limb1 N], then c1).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace opir {
namespace k2 {

typedef uint64_t u64;
typedef uint32_t u32;

constexpr int kMaxN = 4096;                  // largest ring degree of this path
constexpr size_t kMaxCt = 4 * (size_t)kMaxN;  // u64 values of its largest ciphertext

struct Limb64 {
  u64 q, q2;              // q, 2q
  u64 r64, r64s;          // 2^64 mod q and its Shoup companion
  u64 one_s;              // floor(2^64 / q)
  u64 ninv, ninv_s;       // N^-1 mod q and Shoup companion
  const u64 *psi, *psi_s;     // forward twiddles, bit-reversed powers
  const u64 *ipsi, *ipsi_s;   // inverse twiddles
};

struct Scheme64 {
  int N = 4096, K = 2;    // ring degree and RNS limbs: (4096, 2) or (2048, 1)
  Limb64 limb[2];
  u64 q0inv, q0inv_s;     // K = 2: q0^-1 mod q1 (and Shoup companion mod q1)
  u64 Q_lo, Q_hi;         // K = 2: Q = q0 * q1
  u64 hQ_lo, hQ_hi;       // K = 2: Q >> 1
  u64 small_q;
  u64 rm[2];              // Barrett reciprocals floor(2^rk[i] / q_i), in [2^63, 2^64)
  int rk[2];
};

// u64 values of one ciphertext of the scheme.
inline size_t ct_values(const Scheme64 &S) { return 2 * (size_t)S.K * S.N; }

struct Gadget64 {
  int l_ep, b_ep;         // data selectors
  int l_key, b_key;       // RGSW(s)
  int l_ks, b_ks;         // BV keyswitch
};

struct KeyPtrs64 {
  const u64 *const *bv;   // per query: [h][l_ks][2 (b, a)][K limbs][N]
  const u64 *const *gsw;  // per query: [2 l_key][2 polys][K limbs][N]
};

// NTT of `count` polynomials of K limbs each ([count][K][N]), in place.
void launch_k2_ntt(u64 *data, size_t count, bool inverse, const Scheme64 &S, hipStream_t st);
// Scratch for split (narrow) tree levels: at most kSplitTargetBlocks64
// ciphertexts of partial products; narrow levels spread each node over
// several workgroups when `part` is given and the level fits.
constexpr size_t kSplitTargetBlocks64 = 512;
// Expansion level: nodes [0, n) of in -> children 2r, 2r+1 of out (coefficient form).
void launch_k2_expand_level(const u64 *in, u64 *out, size_t stride, int nnodes, int B, int level,
                            const KeyPtrs64 &P, size_t key_level_values, const Gadget64 &G, const Scheme64 &S,
                            hipStream_t st, u64 *part);
// Stand-alone BV keyswitch (bv_keyswitch.cpp:296-481) of one coefficient-form
// ct: out = (sigma_k(c0) + sum d_i b_i, sum d_i a_i); key_ptr: device array
// holding one pointer to l_ks device key rows [b (-> c0), a (-> c1)][limb][N];
// kinv = k^-1 mod 2N.
void launch_k2_bv_apply(const u64 *in, u64 *out, const u64 *const *key_ptr, int kinv, const Gadget64 &G,
                        const Scheme64 &S, hipStream_t st);
// Selector completion: nrows expanded rows (coefficient form) -> sel [nsel][2 l_ep][ct] (NTT form).
void launch_k2_complete(const u64 *rows, u64 *sel, size_t row_stride, size_t sel_stride, int nrows, int B,
                        const KeyPtrs64 &P, const Gadget64 &G, const Scheme64 &S, hipStream_t st);
// Forward NTT of ciphertexts: in (coef) -> out (NTT), count cts per query.
void launch_k2_ct_ntt(const u64 *in, u64 *out, size_t in_stride, size_t out_stride, int count, int B,
                      const Scheme64 &S, hipStream_t st);
// MUX level on NTT-form candidates (same slot rules as the composite path).
void launch_k2_mux(const u64 *in, u64 *out, size_t stride, int nslots, int offset, int half, int B, const u64 *sel,
                   size_t sel_stride, const Gadget64 &G, const Scheme64 &S, hipStream_t st, u64 *part);
// INTT, then the mod switch to small_q (K = 2: drop q1 with round half up,
// then the centred rescale; K = 1: the centred rescale): res [B][ct] -> out [B][2][N].
void launch_k2_final(const u64 *res, u64 *out, size_t res_stride, int B, const Scheme64 &S, hipStream_t st);
// Plaintext coefficients (< t) [count][N] -> NTT per limb into out [limb][first + i][N] (of num_pt).
void launch_k2_db_load(const u64 *pts, u64 *db, size_t num_pt, size_t first, size_t count, const Scheme64 &S,
                       hipStream_t st);
// External product alone (test entry): gsw [2l][ct] NTT, ct coef -> out NTT.
void launch_k2_ext_prod(const u64 *gsw, int l, int b, const u64 *ct, u64 *out, const Scheme64 &S, hipStream_t st);
// Mod switch alone (test entry): coefficient-form ct -> [2][N] mod small_q.
void launch_k2_mod_switch(const u64 *ct, u64 *out, const Scheme64 &S, hipStream_t st);

}  // namespace k2
}  // namespace opir
