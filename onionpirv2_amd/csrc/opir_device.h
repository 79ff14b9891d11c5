// Device-side arithmetic for the OnionPIRv2 answer path on CDNA4 (gfx950).
//
// Ring elements of the CONFIG_N2048_K1_COMP scheme live mod Q = q1*q2 with
// two 29-bit NTT primes. Z_Q is isomorphic to Z_q1 x Z_q2, so every ring
// operation here runs on two 32-bit RNS limbs; the CRT is applied only where
// the reference needs the integer value mod Q (gadget decomposition and the
// final mod switch). The NTT with root w_crt mod Q equals, limb by limb, the
// NTT with root w_crt mod q_i, so limb-wise transforms reproduce
// utils::ntt_fwd/ntt_inv (src/utils.cpp:84-90) exactly.
//
// Thread layout of one 2048-point transform (a "lane block" of 256 threads,
// 8 values per thread):
//   map0: thread t holds indices t + 256*c            (c = 0..7)
//   map3: thread t holds indices 8*t + c              (c = 0..7)
// The forward transform takes map0 to map3, the inverse map3 to map0.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace opir {

typedef uint32_t u32;
typedef uint64_t u64;
typedef int64_t i64;

constexpr int kN = 2048;        // ring degree of the composite scheme
constexpr int kLogN = 11;
constexpr int kLB = 256;        // threads per lane block
constexpr int kPer = 8;         // values per thread
constexpr int kLdsOne = kN + (kN / 32) * 4;  // padded LDS words of one polynomial
// NTT scratch per lane block: three exchange buffers used in rotation, so an
// exchange needs one barrier (write, barrier, read): a buffer is rewritten
// only three exchanges later, after two more barriers.
constexpr int kLdsPoly = 3 * kLdsOne;

// Constants of one 29-bit RNS limb.
struct LimbConst {
  u32 q, q2;          // q, 2q
  u32 r32, r32s;      // 2^32 mod q and its Shoup companion
  u32 one_s;          // floor(2^32 / q)
  u32 ninv, ninv_s;   // N^-1 mod q and Shoup companion
  const u32 *psi;     // forward twiddles, bit-reversed order: psi[brv(i)] = w^i
  const u32 *psi_s;
  const u32 *ipsi;    // inverse twiddles: ipsi[brv(i)] = w^-i
  const u32 *ipsi_s;
};

// Whole-scheme constants passed by value to every kernel.
struct SchemeConst {
  LimbConst limb[2];
  u64 Q;              // q1 * q2
  u64 half_Q;         // Q >> 1
  u32 q1_mod_q2;      // q1 mod q2 (q1 > q2 is allowed)
  u32 q1inv, q1inv_s; // q1^-1 mod q2 and Shoup companion (mod q2)
  u32 q2_mod_q1;
  u64 small_q;
  u64 q_recip;        // floor(2^k / Q), k = 63 + bit length of Q: Barrett reciprocal for the final rescale
  int q_shift;        // k - 64
  // 0: composite Q = q1 * q2 seen as one modulus (CONFIG_N2048_K1_COMP);
  // 1: the two primes are the ciphertext's own RNS limbs (CONFIG_N2048_K2_MP):
  //    MP gadgets (src/gsw.cpp:22-85,184-250, src/bv_keyswitch.cpp:42-64) and
  //    the drop-then-rescale mod switch (src/server.cpp:1044-1075)
  int mp;
};

__device__ __forceinline__ u32 mul_shoup(u32 a, u32 w, u32 ws, u32 q) {
  // a * w mod q, lazily in [0, 2q); valid for any a < 2^32, w < q.
  const u32 hi = __umulhi(a, ws);
  return a * w - hi * q;
}

__device__ __forceinline__ u32 csub(u32 x, u32 m) { return x >= m ? x - m : x; }

__device__ __forceinline__ u32 reduce_4q(u32 x, u32 q, u32 q2) { return csub(csub(x, q2), q); }

// Exact x mod q for any 64-bit x.
__device__ __forceinline__ u32 reduce_u64(u64 x, const LimbConst &L) {
  const u32 hi = (u32)(x >> 32), lo = (u32)x;
  const u32 a = mul_shoup(hi, L.r32, L.r32s, L.q);          // [0, 2q)
  const u32 b = lo - __umulhi(lo, L.one_s) * L.q;           // [0, 2q)
  return reduce_4q(a + b, L.q, L.q2);
}

__device__ __forceinline__ u32 add_mod(u32 a, u32 b, u32 q) { return csub(a + b, q); }
__device__ __forceinline__ u32 sub_mod(u32 a, u32 b, u32 q) { return a >= b ? a - b : a + q - b; }

// CRT: (lo mod q1, hi mod q2) -> value in [0, Q). Garner as in
// inter_to_cts_composite (src/server.cpp:548-553).
__device__ __forceinline__ u64 crt_compose(u32 lo, u32 hi, const SchemeConst &S) {
  const u32 q2 = S.limb[1].q;
  const u32 lo2 = csub(lo, q2);               // lo < q1 < 2*q2
  const u32 diff = sub_mod(hi, lo2, q2);
  const u32 k = csub(mul_shoup(diff, S.q1inv, S.q1inv_s, q2), q2);
  return (u64)lo + (u64)S.limb[0].q * k;
}

// Signed digit r in [-2^(b-1), 2^(b-1)) rendered mod q (q > 2^(b-1)).
__device__ __forceinline__ u32 digit_mod(i64 r, u32 q) { return r >= 0 ? (u32)r : (u32)(r + (i64)q); }

// Digit extraction step, LSB-first: r = sext_b(d - delta) + delta, the
// representative of d mod 2^b in [delta - 2^(b-1), delta + 2^(b-1)), then
// d <- (d - r) >> b. delta = 0: balanced digits in [-B/2, B/2)
// (src/gsw.cpp:285-293, src/bv_keyswitch.cpp:32-39); delta = 1: (-B/2, B/2]
// (signed_gadget_decompose_mp, src/bv_keyswitch.cpp:42-64); delta = B/2: the
// unsigned digits of d >= 0 (decomp_rlwe_mp, src/gsw.cpp:184-250).
__device__ __forceinline__ i64 next_digit(i64 &d, int b, i64 delta = 0) {
  const int sh = 64 - b;
  const i64 r = ((i64)((u64)(d - delta) << sh) >> sh) + delta;
  d = (d - r) >> b;
  return r;
}

// The integer a gadget decomposes, from both limbs of one coefficient:
// composite, or the MP BV gadget: the CRT value centred to (-Q/2, Q/2];
// the MP data and key gadgets: the CRT value in [0, Q) itself.
__device__ __forceinline__ i64 gadget_value(u32 lo, u32 hi, const SchemeConst &S, bool centred) {
  const u64 x = crt_compose(lo, hi, S);
  return centred ? (x > S.half_Q ? (i64)x - (i64)S.Q : (i64)x) : (i64)x;
}

// Centre a value in [0, Q) to (-Q/2, Q/2] the way the reference does
// (value > Q>>1 -> value - Q).
__device__ __forceinline__ i64 centre(u64 x, const SchemeConst &S) {
  return x > S.half_Q ? (i64)x - (i64)S.Q : (i64)x;
}

__device__ __forceinline__ int lds_idx(int i) { return i + ((i >> 5) << 2); }

// ---------------------------------------------------------------------------
// 2048-point negacyclic NTT on one lane block (256 threads, 8 values each).
// lds: kLdsPoly words private to this lane block. Values in [0, q) on entry
// and exit. Caller must ensure all threads of the lane block call it.
// bar() is the barrier for the lane block (whole workgroup barrier).
// ---------------------------------------------------------------------------
// Forward butterfly (X, Y) -> (X + Y*w, X - Y*w), lazily reduced: t = Y*w
// mod q by Shoup lies in [0, 2q) for any Y < 2^32; tn = -t mod 2^32 lets
// both outputs take one instruction each (x - tn, and a three-input add for
// x + 2q - t). The conditional subtractions applied to X first depend on
// the bound X has on entry (every value of a stage has the same bound):
//   kBfC: X < 4q, subtract 2q if X >= 2q             -> outputs < 4q
//   kBfH: X < 8q, subtract 4q if X >= 4q             -> outputs < 6q
//   kBfD: X < 8q, subtract 4q, then 2q, if at least  -> outputs < 4q
//   kBfN: none                                       -> outputs < bound + 2q
// With q < 2^29 every bound below stays under 8q < 2^32.
enum { kBfN, kBfC, kBfH, kBfD };
template <int kKind>
__device__ __forceinline__ void bf_fwd(u32 &X, u32 &Y, u32 w, u32 ws, const LimbConst &L) {
  u32 x = X;
  if constexpr (kKind == kBfD || kKind == kBfH) x = csub(x, 2 * L.q2);
  if constexpr (kKind == kBfD || kKind == kBfC) x = csub(x, L.q2);
  const u32 hi = __umulhi(Y, ws);
  const u32 tn = hi * L.q - Y * w;
  X = x - tn;
  Y = x + tn + L.q2;
}

__device__ __forceinline__ void bf_inv(u32 &X, u32 &Y, u32 w, u32 ws, const LimbConst &L) {
  const u32 x = X, y = Y;
  X = csub(x + y, L.q2);
  Y = mul_shoup(x - y + L.q2, w, ws, L.q);
}

// Twiddles of one radix-8 group b: w[b], w[2b..2b+1], w[4b..4b+3], loaded as
// one dword, one dwordx2 and one dwordx4 (tables are 256-byte aligned).
struct Tw8 {
  u32 w1, s1;
  uint2 w2, s2;
  uint4 w4, s4;
};
// Load at a uniform base plus a 32-bit byte offset, so that the address is
// formed by the load itself (scalar base + vector offset) instead of 64-bit
// vector adds.
template <typename T>
__device__ __forceinline__ T ld_off(const u32 *__restrict__ base, unsigned byte_off) {
  return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + byte_off);
}

__device__ __forceinline__ Tw8 load_tw8(const u32 *__restrict__ w, const u32 *__restrict__ ws, unsigned b) {
  Tw8 r;
  r.w1 = ld_off<u32>(w, b << 2);
  r.s1 = ld_off<u32>(ws, b << 2);
  r.w2 = ld_off<uint2>(w, b << 3);
  r.s2 = ld_off<uint2>(ws, b << 3);
  r.w4 = ld_off<uint4>(w, b << 4);
  r.s4 = ld_off<uint4>(ws, b << 4);
  return r;
}

// Forward radix-8 round on a column (indices base + c*T); twiddle group
// b = m0 + g (stage offsets m0, 2 m0, 4 m0); K1, K2, K3: the butterfly kind
// of each of its three stages.
template <int K1, int K2, int K3>
__device__ __forceinline__ void fwd_round8(u32 v[8], int m0, int g, const LimbConst &L) {
  const Tw8 T = load_tw8(L.psi, L.psi_s, (unsigned)(m0 + g));
#pragma unroll
  for (int c = 0; c < 4; c++) bf_fwd<K1>(v[c], v[c + 4], T.w1, T.s1, L);
  bf_fwd<K2>(v[0], v[2], T.w2.x, T.s2.x, L);
  bf_fwd<K2>(v[1], v[3], T.w2.x, T.s2.x, L);
  bf_fwd<K2>(v[4], v[6], T.w2.y, T.s2.y, L);
  bf_fwd<K2>(v[5], v[7], T.w2.y, T.s2.y, L);
  bf_fwd<K3>(v[0], v[1], T.w4.x, T.s4.x, L);
  bf_fwd<K3>(v[2], v[3], T.w4.y, T.s4.y, L);
  bf_fwd<K3>(v[4], v[5], T.w4.z, T.s4.z, L);
  bf_fwd<K3>(v[6], v[7], T.w4.w, T.s4.w, L);
}

__device__ __forceinline__ void inv_round8(u32 v[8], int m0, int g, const LimbConst &L) {
  const Tw8 T = load_tw8(L.ipsi, L.ipsi_s, (unsigned)(m0 + g));
  bf_inv(v[0], v[1], T.w4.x, T.s4.x, L);
  bf_inv(v[2], v[3], T.w4.y, T.s4.y, L);
  bf_inv(v[4], v[5], T.w4.z, T.s4.z, L);
  bf_inv(v[6], v[7], T.w4.w, T.s4.w, L);
  bf_inv(v[0], v[2], T.w2.x, T.s2.x, L);
  bf_inv(v[1], v[3], T.w2.x, T.s2.x, L);
  bf_inv(v[4], v[6], T.w2.y, T.s2.y, L);
  bf_inv(v[5], v[7], T.w2.y, T.s2.y, L);
#pragma unroll
  for (int c = 0; c < 4; c++) bf_inv(v[c], v[c + 4], T.w1, T.s1, L);
}

// Index of value c of thread t in radix-8 round r (r = 0, 1, 2).
__device__ __forceinline__ int r8_index(int r, int t, int c) {
  const int T = 256 >> (3 * r);
  const int g = t / T, u = t % T;
  return g * 8 * T + u + c * T;
}

// Padded LDS word of value c of thread t in mapping r, written as
// lds_base<r>(t) + c * kLdsStride<r>: the padding i + (i>>5)*4 adds a
// constant per c because every mapping's c-step is either a multiple of 32
// (r = 0, 1) or stays inside one 32-word row (r = 2, 3). The constant
// strides let every LDS access use an immediate offset.
//   r = 0: i = t + 256c                  -> (t + 4(t>>5)) + 288c
//   r = 1: i = 256(t>>5) + (t&31) + 32c  -> (288(t>>5) + (t&31)) + 36c
//   r = 2: i = 32(t>>2) + (t&3) + 4c     -> (36(t>>2) + (t&3)) + 4c
//   r = 3: i = 8t + c                    -> (8t + 4(t>>2)) + c
template <int R> __device__ __forceinline__ int lds_base(int t) {
  if constexpr (R == 0) return t + ((t >> 5) << 2);
  else if constexpr (R == 1) return 288 * (t >> 5) + (t & 31);
  else if constexpr (R == 2) return 36 * (t >> 2) + (t & 3);
  else return 8 * t + ((t >> 2) << 2);
}
template <int R> constexpr int kLdsStride = R == 0 ? 288 : R == 1 ? 36 : R == 2 ? 4 : 1;

template <int R> __device__ __forceinline__ void lds_put(u32 *lds, int t, const u32 v[8]) {
  u32 *p = lds + lds_base<R>(t);
  if constexpr (R == 3) {  // 8 consecutive words, 16-byte aligned
    reinterpret_cast<uint4 *>(p)[0] = make_uint4(v[0], v[1], v[2], v[3]);
    reinterpret_cast<uint4 *>(p)[1] = make_uint4(v[4], v[5], v[6], v[7]);
  } else {
#pragma unroll
    for (int c = 0; c < 8; c++) p[c * kLdsStride<R>] = v[c];
  }
}
template <int R> __device__ __forceinline__ void lds_get(const u32 *lds, int t, u32 v[8]) {
  const u32 *p = lds + lds_base<R>(t);
  if constexpr (R == 3) {
    const uint4 a = reinterpret_cast<const uint4 *>(p)[0], b = reinterpret_cast<const uint4 *>(p)[1];
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
  } else {
#pragma unroll
    for (int c = 0; c < 8; c++) v[c] = p[c * kLdsStride<R>];
  }
}

// One exchange through buffer `buf` of the lane block's rotation. Every
// transform uses buffers 0, 1, 2 in this order, so consecutive transforms on
// the same scratch keep the rotation.
template <int FROM, int TO, typename Bar>
__device__ __forceinline__ void lds_exchange(u32 v[8], u32 *lds, int buf, int t, Bar bar) {
  lds_put<FROM>(lds + buf * kLdsOne, t, v);
  bar();
  lds_get<TO>(lds + buf * kLdsOne, t, v);
}

// kReduce = false leaves outputs lazily in [0, 4q) (< 2^31), for callers that
// only multiply-accumulate them into 64-bit sums. Inputs in [0, q) (any
// values below 2q are fine as well). Value bounds (for inputs below q)
// through the 11 stages: 3q, 5q, 7q | 6q, 8q, 6q | 8q, 6q, 8q | 6q, 4q: six
// conditional subtractions per value where a reduction every stage takes 11.
template <bool kReduce = true, typename Bar>
__device__ __forceinline__ void ntt_fwd_2048(u32 v[8], u32 *lds, int t, const LimbConst &L, Bar bar) {
  fwd_round8<kBfN, kBfN, kBfN>(v, 1, 0, L);     // stages 0-2 (map0, g = 0)
  lds_exchange<0, 1>(v, lds, 0, t, bar);
  fwd_round8<kBfH, kBfN, kBfH>(v, 8, t / 32, L);   // stages 3-5
  lds_exchange<1, 2>(v, lds, 1, t, bar);
  fwd_round8<kBfN, kBfH, kBfN>(v, 64, t / 4, L);   // stages 6-8
  lds_exchange<2, 3>(v, lds, 2, t, bar);
  // stages 9-10 on 8 consecutive values: twiddles w[512 + 2t + h], w[1024 + 4t + h]
  {
    const unsigned u = (unsigned)t;
    const uint2 w2 = ld_off<uint2>(L.psi, (512 * 4) + (u << 3));
    const uint2 s2 = ld_off<uint2>(L.psi_s, (512 * 4) + (u << 3));
    const uint4 w4 = ld_off<uint4>(L.psi, (1024 * 4) + (u << 4));
    const uint4 s4 = ld_off<uint4>(L.psi_s, (1024 * 4) + (u << 4));
    bf_fwd<kBfH>(v[0], v[2], w2.x, s2.x, L);
    bf_fwd<kBfH>(v[1], v[3], w2.x, s2.x, L);
    bf_fwd<kBfH>(v[4], v[6], w2.y, s2.y, L);
    bf_fwd<kBfH>(v[5], v[7], w2.y, s2.y, L);
    bf_fwd<kBfD>(v[0], v[1], w4.x, s4.x, L);
    bf_fwd<kBfD>(v[2], v[3], w4.y, s4.y, L);
    bf_fwd<kBfD>(v[4], v[5], w4.z, s4.z, L);
    bf_fwd<kBfD>(v[6], v[7], w4.w, s4.w, L);
  }
#pragma unroll
  for (int c = 0; c < 8; c++) v[c] = kReduce ? reduce_4q(v[c], L.q, L.q2) : v[c];
}

template <typename Bar>
__device__ __forceinline__ void ntt_inv_2048(u32 v[8], u32 *lds, int t, const LimbConst &L, Bar bar) {
  {
    const unsigned u = (unsigned)t;
    const uint4 w4 = ld_off<uint4>(L.ipsi, (1024 * 4) + (u << 4));
    const uint4 s4 = ld_off<uint4>(L.ipsi_s, (1024 * 4) + (u << 4));
    const uint2 w2 = ld_off<uint2>(L.ipsi, (512 * 4) + (u << 3));
    const uint2 s2 = ld_off<uint2>(L.ipsi_s, (512 * 4) + (u << 3));
    bf_inv(v[0], v[1], w4.x, s4.x, L);
    bf_inv(v[2], v[3], w4.y, s4.y, L);
    bf_inv(v[4], v[5], w4.z, s4.z, L);
    bf_inv(v[6], v[7], w4.w, s4.w, L);
    bf_inv(v[0], v[2], w2.x, s2.x, L);
    bf_inv(v[1], v[3], w2.x, s2.x, L);
    bf_inv(v[4], v[6], w2.y, s2.y, L);
    bf_inv(v[5], v[7], w2.y, s2.y, L);
  }
  lds_exchange<3, 2>(v, lds, 0, t, bar);
  inv_round8(v, 64, t / 4, L);
  lds_exchange<2, 1>(v, lds, 1, t, bar);
  inv_round8(v, 8, t / 32, L);
  lds_exchange<1, 0>(v, lds, 2, t, bar);
  inv_round8(v, 1, 0, L);
#pragma unroll
  for (int c = 0; c < 8; c++) v[c] = csub(mul_shoup(v[c], L.ninv, L.ninv_s, L.q), L.q);
}

// map0 / map3 global index of value c of thread t.
__device__ __forceinline__ int map0(int t, int c) { return t + 256 * c; }
__device__ __forceinline__ int map3(int t, int c) { return 8 * t + c; }

__device__ __forceinline__ void load_map3(u32 v[8], const u32 *p, int t) {
  const uint4 a = reinterpret_cast<const uint4 *>(p)[2 * t];
  const uint4 b = reinterpret_cast<const uint4 *>(p)[2 * t + 1];
  v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
  v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
}
__device__ __forceinline__ void store_map3(u32 *p, const u32 v[8], int t) {
  reinterpret_cast<uint4 *>(p)[2 * t] = make_uint4(v[0], v[1], v[2], v[3]);
  reinterpret_cast<uint4 *>(p)[2 * t + 1] = make_uint4(v[4], v[5], v[6], v[7]);
}
__device__ __forceinline__ void load_map0(u32 v[8], const u32 *p, int t) {
#pragma unroll
  for (int c = 0; c < 8; c++) v[c] = p[t + 256 * c];
}
__device__ __forceinline__ void store_map0(u32 *p, const u32 v[8], int t) {
#pragma unroll
  for (int c = 0; c < 8; c++) p[t + 256 * c] = v[c];
}

}  // namespace opir
