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
constexpr int kLdsPoly = kN + (kN / 32) * 4;  // padded LDS words per polynomial

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

// Balanced digit extraction step (src/gsw.cpp:285-293, src/bv_keyswitch.cpp:32-39):
// r = sign-extended low b bits of d, d <- (d - r) >> b.
__device__ __forceinline__ i64 next_digit(i64 &d, int b) {
  const int sh = 64 - b;
  const i64 r = (i64)((u64)d << sh) >> sh;
  d = (d - r) >> b;
  return r;
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
__device__ __forceinline__ void bf_fwd(u32 &X, u32 &Y, u32 w, u32 ws, const LimbConst &L) {
  const u32 x = csub(X, L.q2);
  const u32 t = mul_shoup(Y, w, ws, L.q);
  X = x + t;
  Y = x - t + L.q2;
}

__device__ __forceinline__ void bf_inv(u32 &X, u32 &Y, u32 w, u32 ws, const LimbConst &L) {
  const u32 x = X, y = Y;
  X = csub(x + y, L.q2);
  Y = mul_shoup(x - y + L.q2, w, ws, L.q);
}

// Forward radix-8 round on a column (indices base + c*T); twiddle groups g.
__device__ __forceinline__ void fwd_round8(u32 v[8], int m0, int g, const LimbConst &L) {
  {
    const u32 w = L.psi[m0 + g], ws = L.psi_s[m0 + g];
#pragma unroll
    for (int c = 0; c < 4; c++) bf_fwd(v[c], v[c + 4], w, ws, L);
  }
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const int idx = 2 * m0 + 2 * g + h;
    const u32 w = L.psi[idx], ws = L.psi_s[idx];
    bf_fwd(v[4 * h + 0], v[4 * h + 2], w, ws, L);
    bf_fwd(v[4 * h + 1], v[4 * h + 3], w, ws, L);
  }
#pragma unroll
  for (int h = 0; h < 4; h++) {
    const int idx = 4 * m0 + 4 * g + h;
    bf_fwd(v[2 * h], v[2 * h + 1], L.psi[idx], L.psi_s[idx], L);
  }
}

__device__ __forceinline__ void inv_round8(u32 v[8], int m0, int g, const LimbConst &L) {
#pragma unroll
  for (int h = 0; h < 4; h++) {
    const int idx = 4 * m0 + 4 * g + h;
    bf_inv(v[2 * h], v[2 * h + 1], L.ipsi[idx], L.ipsi_s[idx], L);
  }
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const int idx = 2 * m0 + 2 * g + h;
    const u32 w = L.ipsi[idx], ws = L.ipsi_s[idx];
    bf_inv(v[4 * h + 0], v[4 * h + 2], w, ws, L);
    bf_inv(v[4 * h + 1], v[4 * h + 3], w, ws, L);
  }
  {
    const u32 w = L.ipsi[m0 + g], ws = L.ipsi_s[m0 + g];
#pragma unroll
    for (int c = 0; c < 4; c++) bf_inv(v[c], v[c + 4], w, ws, L);
  }
}

// Index of value c of thread t in radix-8 round r (r = 0, 1, 2).
__device__ __forceinline__ int r8_index(int r, int t, int c) {
  const int T = 256 >> (3 * r);
  const int g = t / T, u = t % T;
  return g * 8 * T + u + c * T;
}

template <typename Bar>
__device__ __forceinline__ void lds_exchange(u32 v[8], u32 *lds, int t, int from_r, int to_r, Bar bar) {
  // from_r/to_r in {0,1,2,3}; 3 = consecutive mapping.
#pragma unroll
  for (int c = 0; c < 8; c++) {
    const int i = from_r == 3 ? 8 * t + c : r8_index(from_r, t, c);
    lds[lds_idx(i)] = v[c];
  }
  bar();
#pragma unroll
  for (int c = 0; c < 8; c++) {
    const int i = to_r == 3 ? 8 * t + c : r8_index(to_r, t, c);
    v[c] = lds[lds_idx(i)];
  }
  bar();
}

template <typename Bar>
__device__ __forceinline__ void ntt_fwd_2048(u32 v[8], u32 *lds, int t, const LimbConst &L, Bar bar) {
  fwd_round8(v, 1, 0, L);                       // stages 0-2 (map0, g = 0)
  lds_exchange(v, lds, t, 0, 1, bar);
  fwd_round8(v, 8, t / 32, L);                  // stages 3-5
  lds_exchange(v, lds, t, 1, 2, bar);
  fwd_round8(v, 64, t / 4, L);                  // stages 6-8
  lds_exchange(v, lds, t, 2, 3, bar);
  // stages 9-10 on 8 consecutive values
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const int idx = 512 + 2 * t + h;
    const u32 w = L.psi[idx], ws = L.psi_s[idx];
    bf_fwd(v[4 * h + 0], v[4 * h + 2], w, ws, L);
    bf_fwd(v[4 * h + 1], v[4 * h + 3], w, ws, L);
  }
#pragma unroll
  for (int h = 0; h < 4; h++) {
    const int idx = 1024 + 4 * t + h;
    bf_fwd(v[2 * h], v[2 * h + 1], L.psi[idx], L.psi_s[idx], L);
  }
#pragma unroll
  for (int c = 0; c < 8; c++) v[c] = reduce_4q(v[c], L.q, L.q2);
}

template <typename Bar>
__device__ __forceinline__ void ntt_inv_2048(u32 v[8], u32 *lds, int t, const LimbConst &L, Bar bar) {
#pragma unroll
  for (int h = 0; h < 4; h++) {
    const int idx = 1024 + 4 * t + h;
    bf_inv(v[2 * h], v[2 * h + 1], L.ipsi[idx], L.ipsi_s[idx], L);
  }
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const int idx = 512 + 2 * t + h;
    const u32 w = L.ipsi[idx], ws = L.ipsi_s[idx];
    bf_inv(v[4 * h + 0], v[4 * h + 2], w, ws, L);
    bf_inv(v[4 * h + 1], v[4 * h + 3], w, ws, L);
  }
  lds_exchange(v, lds, t, 3, 2, bar);
  inv_round8(v, 64, t / 4, L);
  lds_exchange(v, lds, t, 2, 1, bar);
  inv_round8(v, 8, t / 32, L);
  lds_exchange(v, lds, t, 1, 0, bar);
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
