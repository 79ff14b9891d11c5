// The two-prime answer path (CONFIG_N4096_K2_MP) on CDNA4.
//
// Arithmetic: each limb is a ~60-bit NTT prime, so values are u64 and every
// product is a 64x64->128 product. Modular multiplication by a known
// constant uses Shoup's method with a 64-bit companion (floor(w*2^64/q));
// the multiply-accumulates against key rows (products of two variables) sum
// exact 128-bit products and reduce each sum once, as hi*(2^64 mod q) + lo
// with both halves by Shoup. NTT butterflies are Harvey-lazy in [0, 4q)
// (4q < 2^62).
//
// Thread layout: a 4096-point transform is a lane block of 512 threads with
// 8 values each, four radix-8 rounds and three LDS exchanges; value c of
// thread t sits at index g*8T + (t % T) + c*T with T = 512 >> 3r, g = t / T in
// round r (r = 3: 8t + c). Tree nodes are 1024-thread workgroups, one lane
// block per limb.
//
// Gadgets (src/gsw.cpp:22-85,184-250; src/bv_keyswitch.cpp:42-64,373-481):
// the two limbs are recombined by CRT into X in [0, Q), Q ~ 2^120.
//   data / key selectors: unsigned digits (X >> (b*i)) & (2^b - 1),
//     stored at row poly*l + (l-1-i);
//   BV keyswitch: signed digits in (-B/2, B/2] of X centred to (-Q/2, Q/2],
//     digit i = ((floor((D + H_i) / B^i) + h) mod B) - h with h = B/2 - 1,
//     H_i = h * (1 + B + ... + B^(i-1)) -- the closed form of the
//     reference's carry loop, so every digit is computed independently.
#include "opir_k2.h"

namespace opir {
namespace k2 {

typedef unsigned __int128 u128;
typedef __int128 i128;

namespace {

struct Barrier {
  __device__ void operator()() const { __syncthreads(); }
};

constexpr int kLds = kN + kN / 32;  // padded u64 words of one exchange buffer

__device__ __forceinline__ u64 csub(u64 x, u64 m) { return x >= m ? x - m : x; }
__device__ __forceinline__ u64 addmod(u64 a, u64 b, u64 q) { return csub(a + b, q); }
__device__ __forceinline__ u64 submod(u64 a, u64 b, u64 q) { return a >= b ? a - b : a + q - b; }

// a * w mod q lazily in [0, 2q), any a < 2^64, w < q.
__device__ __forceinline__ u64 shoup(u64 a, u64 w, u64 ws, u64 q) { return a * w - __umul64hi(a, ws) * q; }


// CRT (r0 mod q0, r1 mod q1) -> X in [0, Q) (Garner, server.cpp:529-553 form).
__device__ __forceinline__ u128 crt2(u64 r0, u64 r1, const Scheme64 &S) {
  const u64 q1 = S.limb[1].q;
  const u64 r0m = csub(r0, q1);  // r0 < q0 < 2 q1
  const u64 s = csub(shoup(submod(r1, r0m, q1), S.q0inv, S.q0inv_s, q1), q1);
  return (u128)S.limb[0].q * s + r0;
}

__device__ __forceinline__ u128 Q128(const Scheme64 &S) { return ((u128)S.Q_hi << 64) | S.Q_lo; }
__device__ __forceinline__ u128 hQ128(const Scheme64 &S) { return ((u128)S.hQ_hi << 64) | S.hQ_lo; }

// Signed residue r (|r| < 2^63) mod q.
__device__ __forceinline__ u64 sres(int64_t r, u64 q) { return r >= 0 ? (u64)r : q - (u64)(-r); }

// floor(num / d) for num < 2^121 by Barrett: m = floor(2^k / d) in
// [2^63, 2^64) (k = 63 + bit length of d). floor(num * m / 2^k), keeping only
// the high 64 bits of the low product, is the quotient or up to two less;
// two conditional corrections make it exact.
__device__ __forceinline__ u64 div_barrett(u128 num, u64 d, u64 m, int k) {
  const u64 nh = (u64)(num >> 64), nl = (u64)num;
  const u128 t = (u128)nh * m + __umul64hi(nl, m);  // about num * m / 2^64
  u64 quo = (u64)(t >> (k - 64));
  u128 r = num - (u128)quo * d;
#pragma unroll
  for (int i = 0; i < 2; i++)
    if (r >= d) {
      r -= d;
      quo++;
    }
  return quo;
}

// ---------------------------------------------------------------------------
// 4096-point NTT on a 512-thread lane block
// ---------------------------------------------------------------------------
__device__ __forceinline__ int pidx(int i) { return i + (i >> 5); }
__device__ __forceinline__ int r_index(int r, int t, int c) {
  const int T = 512 >> (3 * r);
  return (t / T) * 8 * T + (t % T) + c * T;
}
__device__ __forceinline__ int map0(int t, int c) { return t + 512 * c; }
__device__ __forceinline__ int map3(int t, int c) { return 8 * t + c; }

__device__ __forceinline__ void bf_fwd(u64 &X, u64 &Y, u64 w, u64 ws, const Limb64 &L) {
  const u64 x = csub(X, L.q2);
  const u64 t = shoup(Y, w, ws, L.q);
  X = x + t;
  Y = x - t + L.q2;
}
__device__ __forceinline__ void bf_inv(u64 &X, u64 &Y, u64 w, u64 ws, const Limb64 &L) {
  const u64 x = X, y = Y;
  X = csub(x + y, L.q2);
  Y = shoup(x - y + L.q2, w, ws, L.q);
}

__device__ __forceinline__ void fwd8(u64 v[8], unsigned b, const Limb64 &L) {
  {
    const u64 w = L.psi[b], ws = L.psi_s[b];
#pragma unroll
    for (int c = 0; c < 4; c++) bf_fwd(v[c], v[c + 4], w, ws, L);
  }
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const u64 w = L.psi[2 * b + h], ws = L.psi_s[2 * b + h];
    bf_fwd(v[4 * h + 0], v[4 * h + 2], w, ws, L);
    bf_fwd(v[4 * h + 1], v[4 * h + 3], w, ws, L);
  }
#pragma unroll
  for (int h = 0; h < 4; h++) bf_fwd(v[2 * h], v[2 * h + 1], L.psi[4 * b + h], L.psi_s[4 * b + h], L);
}

__device__ __forceinline__ void inv8(u64 v[8], unsigned b, const Limb64 &L) {
#pragma unroll
  for (int h = 0; h < 4; h++) bf_inv(v[2 * h], v[2 * h + 1], L.ipsi[4 * b + h], L.ipsi_s[4 * b + h], L);
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const u64 w = L.ipsi[2 * b + h], ws = L.ipsi_s[2 * b + h];
    bf_inv(v[4 * h + 0], v[4 * h + 2], w, ws, L);
    bf_inv(v[4 * h + 1], v[4 * h + 3], w, ws, L);
  }
  {
    const u64 w = L.ipsi[b], ws = L.ipsi_s[b];
#pragma unroll
    for (int c = 0; c < 4; c++) bf_inv(v[c], v[c + 4], w, ws, L);
  }
}

// LDS exchange between rounds. t is passed by reference and made opaque
// afterwards, so the next round's twiddle addresses (functions of t) cannot be
// computed and loaded before the exchange: hoisting all rounds' twiddles
// exceeds the 128-register budget of a 1024-thread workgroup.
__device__ __forceinline__ void exch(u64 v[8], u64 *lds, int &t, int from, int to) {
#pragma unroll
  for (int c = 0; c < 8; c++) lds[pidx(r_index(from, t, c))] = v[c];
  __syncthreads();
#pragma unroll
  for (int c = 0; c < 8; c++) v[c] = lds[pidx(r_index(to, t, c))];
  __syncthreads();
  asm volatile("" : "+v"(t));
}

// map0 [0, q) in -> map3 out, [0, q) when reduce, else [0, 4q).
__device__ __forceinline__ void ntt_fwd(u64 v[8], u64 *lds, int t, const Limb64 &L, bool reduce = true) {
  asm volatile("" : "+v"(t));
  fwd8(v, 1, L);
  exch(v, lds, t, 0, 1);
  fwd8(v, 8 + t / 64, L);
  exch(v, lds, t, 1, 2);
  fwd8(v, 64 + t / 8, L);
  exch(v, lds, t, 2, 3);
  fwd8(v, 512 + t, L);
  if (reduce) {
#pragma unroll
    for (int c = 0; c < 8; c++) v[c] = csub(csub(v[c], L.q2), L.q);
  }
}

// map3 [0, q) in -> map0 out in [0, q).
__device__ __forceinline__ void ntt_inv(u64 v[8], u64 *lds, int t, const Limb64 &L) {
  asm volatile("" : "+v"(t));
  inv8(v, 512 + t, L);
  exch(v, lds, t, 3, 2);
  inv8(v, 64 + t / 8, L);
  exch(v, lds, t, 2, 1);
  inv8(v, 8 + t / 64, L);
  exch(v, lds, t, 1, 0);
  inv8(v, 1, L);
#pragma unroll
  for (int c = 0; c < 8; c++) v[c] = csub(shoup(v[c], L.ninv, L.ninv_s, L.q), L.q);
}

__device__ __forceinline__ void load0(u64 v[8], const u64 *p, int t) {
#pragma unroll
  for (int c = 0; c < 8; c++) v[c] = p[map0(t, c)];
}
__device__ __forceinline__ void store0(u64 *p, const u64 v[8], int t) {
#pragma unroll
  for (int c = 0; c < 8; c++) p[map0(t, c)] = v[c];
}
__device__ __forceinline__ void load3(u64 v[8], const u64 *p, int t) {
  const ulonglong2 *q = reinterpret_cast<const ulonglong2 *>(p + 8 * t);
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const ulonglong2 x = q[c];
    v[2 * c] = x.x;
    v[2 * c + 1] = x.y;
  }
}
__device__ __forceinline__ void store3(u64 *p, const u64 v[8], int t) {
  ulonglong2 *q = reinterpret_cast<ulonglong2 *>(p + 8 * t);
#pragma unroll
  for (int c = 0; c < 4; c++) q[c] = make_ulonglong2(v[2 * c], v[2 * c + 1]);
}

// acc[c] += v[c] * key[8t + c] as exact 128-bit sums, two key values at a
// time; mac_finish reduces each sum once. With v < 4q and key < q < 2^60
// every product is below 2^122, so a sum of up to 64 products cannot wrap
// (the scheme validation bounds every digit count by that).
__device__ __forceinline__ void mac_row(u128 (&acc)[8], const u64 v[8], const u64 *key, int t, const Limb64 &L) {
  const ulonglong2 *kp = reinterpret_cast<const ulonglong2 *>(key + 8 * t);
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const ulonglong2 k = kp[c];
    acc[2 * c] += (u128)v[2 * c] * k.x;
    acc[2 * c + 1] += (u128)v[2 * c + 1] * k.y;
  }
}
// (hi * 2^64 + lo) mod q in [0, q): hi * (2^64 mod q) and lo * 1 by Shoup, each in [0, 2q).
__device__ __forceinline__ u64 reduce128(u128 x, const Limb64 &L) {
  const u64 hi = (u64)(x >> 64), lo = (u64)x;
  const u64 a = shoup(hi, L.r64, L.r64s, L.q);
  const u64 b = lo - __umul64hi(lo, L.one_s) * L.q;
  return csub(csub(a + b, L.q2), L.q);
}
__device__ __forceinline__ void mac_finish(const u128 (&acc)[8], u64 (&out)[8], const Limb64 &L) {
#pragma unroll
  for (int c = 0; c < 8; c++) out[c] = reduce128(acc[c], L);
}

// One polynomial's external-product contribution (gsw.cpp:119-182 with the
// MP decomposition): X[c] are the recombined coefficients (map0) of poly p;
// digits i < l of base 2^b go through the forward NTT of this lane block's
// limb and are multiplied into acc0/acc1 with rows p*l + (l-1-i) of g
// ([rows][2 outputs][2 limbs][N], NTT form).
__device__ __forceinline__ void ep_poly(const u128 *xs, int p, int l, int b, const u64 *__restrict__ g, int limb,
                                        int t_in, const Limb64 &L, u64 *lds, u128 (&acc0)[8], u128 (&acc1)[8],
                                        int i0 = 0, int i1 = -1) {
  const u64 mask = ((u64)1 << b) - 1;
  if (i1 < 0) i1 = l;
#pragma unroll 1
  for (int i = i0; i < i1; i++) {
    // opaque copy of t: keeps per-thread addresses inside the loop (register pressure)
    int t = t_in;
    asm volatile("" : "+v"(t));
    u64 v[8];
#pragma unroll
    for (int c = 0; c < 8; c++) v[c] = (u64)(xs[map0(t, c)] >> (b * i)) & mask;  // < 2^b < q
    ntt_fwd(v, lds, t, L, false);
    const u64 *row = g + (size_t)(p * l + (l - 1 - i)) * kCt;
    mac_row(acc0, v, row + (size_t)(0 * 2 + limb) * kN, t, L);
    mac_row(acc1, v, row + (size_t)(1 * 2 + limb) * kN, t, L);
  }
}

// Stage this lane block's limb of a coefficient-form polynomial (map0) and
// recombine both limbs of every coefficient into xs[j] (LDS).
__device__ __forceinline__ void stage_compose(const u64 v[8], u64 (*stage)[kLds], int limb, int t, const Scheme64 &S,
                                              u128 *xs) {
  store0(stage[limb], v, t);
  __syncthreads();
  for (int j = threadIdx.x; j < kN; j += 1024) xs[j] = crt2(stage[0][j], stage[1][j], S);
  __syncthreads();
}

// Centred rescale of a mod-q0 value to small_q (utils.cpp:445-462).
__device__ __forceinline__ u64 rescale_q0(u64 a, const Scheme64 &S) {
  const u64 q0 = S.limb[0].q, h = q0 / 2;
  const bool neg = a >= h;  // v >= inp/2 -> v - inp
  const u64 mag = neg ? q0 - a : a;
  const u64 quo = div_barrett((u128)mag * S.small_q + h, q0, S.rm[0], S.rk[0]);
  const u64 r = quo < S.small_q ? quo : quo % S.small_q;  // mag <= q0/2 keeps quo below small_q
  return neg ? (r ? S.small_q - r : 0) : r;
}

// Drop q1 with round half up (server.cpp:1044-1075), then rescale.
__device__ __forceinline__ u64 final_coeff(u64 r0, u64 r1, const Scheme64 &S) {
  const u64 q0 = S.limb[0].q, q1 = S.limb[1].q;
  const u128 x = crt2(r0, r1, S);
  u64 o = div_barrett(x + (q1 >> 1), q1, S.rm[1], S.rk[1]);
  if (o >= q0) o -= q0;
  return rescale_q0(o, S);
}

}  // namespace

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k2_ntt(u64 *data, int inverse, Scheme64 S) {
  __shared__ u64 lds[2][kLds];
  const int limb = __builtin_amdgcn_readfirstlane(threadIdx.x >> 9), t = threadIdx.x & 511;
  const Limb64 &L = S.limb[limb];
  u64 *p = data + ((size_t)blockIdx.x * 2 + limb) * kN;
  u64 v[8];
  if (!inverse) {
    load0(v, p, t);
    ntt_fwd(v, lds[limb], t, L);
    store3(p, v, t);
  } else {
    load3(v, p, t);
    ntt_inv(v, lds[limb], t, L);
    store0(p, v, t);
  }
}

struct ExpandArgs64 {
  const u64 *in;
  u64 *out;
  size_t stride;
  const u64 *const *bv;
  size_t key_level_values;
  int level, kinv, l_ks, b_ks;
  int bv_only = 0;  // 1: out node 0 receives the keyswitched c' alone (no children)
};

// Expansion level (server.cpp:786-810, bv_keyswitch.cpp:373-481).
__global__ __launch_bounds__(1024) void k2_expand_level(ExpandArgs64 A, Scheme64 S) {
  __shared__ u128 xs[kN];  // recombined coefficients of the polynomial being decomposed
  __shared__ u64 lds[2][kLds];
  const int limb = __builtin_amdgcn_readfirstlane(threadIdx.x >> 9), t = threadIdx.x & 511;
  const Limb64 &L = S.limb[limb];
  const int node = blockIdx.x, qi = blockIdx.y;
  const u64 *c = A.in + qi * A.stride + (size_t)node * kCt;
  const u64 *c1 = c + (size_t)(2 + limb) * kN, *c0 = c + (size_t)limb * kN;
  // sigma_k(c1) (coefficient form): out[j] = in[i0] or -in[i0 - N], i0 = j*kinv mod 2N
  u64 v[8];
#pragma unroll
  for (int cc = 0; cc < 8; cc++) {
    const int j = map0(t, cc);
    const int i0 = (int)(((unsigned)j * (unsigned)A.kinv) & (2 * kN - 1));
    const u64 x = c1[i0 < kN ? i0 : i0 - kN];
    v[cc] = (i0 < kN || !x) ? x : L.q - x;
  }
  // centred CRT value D of sigma(c1), as a two's-complement 128-bit integer
  stage_compose(v, lds, limb, t, S, xs);
  {
    const u128 Q = Q128(S), hQ = hQ128(S);
    for (int j = threadIdx.x; j < kN; j += 1024)
      if (xs[j] > hQ) xs[j] -= Q;  // wraps to the two's complement of X - Q
    __syncthreads();
  }
  u128 sb[8], sa[8];
#pragma unroll
  for (int cc = 0; cc < 8; cc++) sb[cc] = sa[cc] = 0;
  const int b = A.b_ks;
  const i128 h = ((i128)1 << (b - 1)) - 1;
  const u64 *key = A.bv[qi] + (size_t)A.level * A.key_level_values;
  i128 H = 0;
  const int t_in = t;
  for (int i = 0; i < A.l_ks; i++) {
    int t = t_in;
    asm volatile("" : "+v"(t));
    u64 d[8];
#pragma unroll
    for (int cc = 0; cc < 8; cc++) {
      const i128 di = ((i128)xs[map0(t, cc)] + H) >> (b * i);
      const int64_t r = (int64_t)(((u64)(di + h)) & (((u64)1 << b) - 1)) - (int64_t)h;
      d[cc] = sres(r, L.q);
    }
    H += h << (b * i);
    ntt_fwd(d, lds[limb], t, L, false);
    const u64 *row = key + (size_t)i * kCt;  // [b (-> c0), a (-> c1)][limb][N]
    mac_row(sb, d, row + (size_t)(0 * 2 + limb) * kN, t, L);
    mac_row(sa, d, row + (size_t)(1 * 2 + limb) * kN, t, L);
  }
  u64 accb[8], acca[8];
  mac_finish(sb, accb, L);
  mac_finish(sa, acca, L);
  ntt_inv(accb, lds[limb], t, L);  // delta_b, coefficient form (map0)
  ntt_inv(acca, lds[limb], t, L);  // delta_a
  // c' = (sigma(c0) + delta_b, delta_a); left = c + c', right = (c - c') * X^-k
  u64 *left = A.out + qi * A.stride + (size_t)(2 * node) * kCt;
  u64 *right = left + kCt;
  const int k = 1 << A.level;
  for (int p = 0; p < 2; p++) {
    const u64 *cp = p ? c1 : c0;
    u64 cv[8], dv[8];
    load0(cv, cp, t);
#pragma unroll
    for (int cc = 0; cc < 8; cc++) {
      u64 w = p ? acca[cc] : accb[cc];
      if (p == 0) {
        const int j = map0(t, cc);
        const int i0 = (int)(((unsigned)j * (unsigned)A.kinv) & (2 * kN - 1));
        const u64 x = c0[i0 < kN ? i0 : i0 - kN];
        w = addmod(w, (i0 < kN || !x) ? x : L.q - x, L.q);
      }
      left[(size_t)(p * 2 + limb) * kN + map0(t, cc)] = A.bv_only ? w : addmod(cv[cc], w, L.q);
      dv[cc] = submod(cv[cc], w, L.q);
    }
    if (A.bv_only) continue;  // uniform over the workgroup
    store0(lds[limb], dv, t);
    __syncthreads();
#pragma unroll
    for (int cc = 0; cc < 8; cc++) {
      const int j = map0(t, cc), s = j + k;
      const u64 x = lds[limb][s < kN ? s : s - kN];
      right[(size_t)(p * 2 + limb) * kN + j] = (s < kN || !x) ? x : L.q - x;
    }
    __syncthreads();
  }
}

struct CompleteArgs64 {
  const u64 *rows;
  u64 *sel;
  size_t row_stride, sel_stride;
  const u64 *const *gsw;
  int l_ep, l_key, b_key;
};

// Selector completion (server.cpp:823-855, gsw.cpp:317-357).
__global__ __launch_bounds__(1024) void k2_complete(CompleteArgs64 A, Scheme64 S) {
  __shared__ u128 xs[kN];  // recombined coefficients of the polynomial being decomposed
  __shared__ u64 lds[2][kLds];
  const int limb = __builtin_amdgcn_readfirstlane(threadIdx.x >> 9), t = threadIdx.x & 511;
  const Limb64 &L = S.limb[limb];
  const int g = blockIdx.x, qi = blockIdx.y;
  const int a = g / A.l_ep, r = g % A.l_ep;
  const u64 *row = A.rows + qi * A.row_stride + (size_t)g * kCt;
  u64 *selb = A.sel + qi * A.sel_stride + (size_t)a * (2 * A.l_ep) * kCt;
  u128 s0[8], s1[8];
#pragma unroll
  for (int c = 0; c < 8; c++) s0[c] = s1[c] = 0;
#pragma unroll 1
  for (int p = 0; p < 2; p++) {
    u64 v[8];
    load0(v, row + (size_t)(p * 2 + limb) * kN, t);
    stage_compose(v, lds, limb, t, S, xs);
    ntt_fwd(v, lds[limb], t, L);
    store3(selb + (size_t)r * kCt + (size_t)(p * 2 + limb) * kN, v, t);
    ep_poly(xs, p, A.l_key, A.b_key, A.gsw[qi], limb, t, L, lds[limb], s0, s1);
    __syncthreads();  // xs is rewritten for the next polynomial
  }
  u64 acc0[8], acc1[8];
  mac_finish(s0, acc0, L);
  mac_finish(s1, acc1, L);
  u64 *bot = selb + (size_t)(A.l_ep + r) * kCt;
  store3(bot + (size_t)(0 * 2 + limb) * kN, acc0, t);
  store3(bot + (size_t)(1 * 2 + limb) * kN, acc1, t);
}

__global__ __launch_bounds__(1024) void k2_ct_ntt(const u64 *in, u64 *out, size_t in_stride, size_t out_stride,
                                                  Scheme64 S) {
  __shared__ u64 lds[2][kLds];
  const int limb = __builtin_amdgcn_readfirstlane(threadIdx.x >> 9), t = threadIdx.x & 511;
  const Limb64 &L = S.limb[limb];
  const int ct = blockIdx.x >> 1, p = blockIdx.x & 1, qi = blockIdx.y;
  u64 v[8];
  load0(v, in + qi * in_stride + (size_t)ct * kCt + (size_t)(p * 2 + limb) * kN, t);
  ntt_fwd(v, lds[limb], t, L);
  store3(out + qi * out_stride + (size_t)ct * kCt + (size_t)(p * 2 + limb) * kN, v, t);
}

struct MuxArgs64 {
  const u64 *in;
  u64 *out;
  size_t stride;
  const u64 *sel;
  size_t sel_stride;
  int offset, half, l, b;
};

// MUX (server.cpp:599-722) on NTT-form candidates: res = x + EP(sel, INTT(y - x)).
__global__ __launch_bounds__(1024) void k2_mux(MuxArgs64 A, Scheme64 S) {
  __shared__ u128 xs[kN];  // recombined coefficients of the polynomial being decomposed
  __shared__ u64 lds[2][kLds];
  const int limb = __builtin_amdgcn_readfirstlane(threadIdx.x >> 9), t = threadIdx.x & 511;
  const Limb64 &L = S.limb[limb];
  const int s = blockIdx.x, qi = blockIdx.y;
  const u64 *in = A.in + qi * A.stride;
  u64 *dst = A.out + qi * A.stride + (size_t)s * kCt;
  int xi, yi;
  if (A.offset >= 0) {
    if (s < A.offset) {
      const u64 *src = in + (size_t)s * kCt;
      for (int e = threadIdx.x; e < (int)kCt; e += 1024) dst[e] = src[e];
      return;
    }
    xi = A.offset + 2 * (s - A.offset);
    yi = xi + 1;
  } else {
    xi = s;
    yi = s + A.half;
  }
  const u64 *x = in + (size_t)xi * kCt, *y = in + (size_t)yi * kCt;
  const u64 *g = A.sel + qi * A.sel_stride;
  u128 acc0[8], acc1[8];
#pragma unroll
  for (int c = 0; c < 8; c++) acc0[c] = acc1[c] = 0;
#pragma unroll 1
  for (int p = 0; p < 2; p++) {
    u64 v[8], w[8];
    load3(v, y + (size_t)(p * 2 + limb) * kN, t);
    load3(w, x + (size_t)(p * 2 + limb) * kN, t);
#pragma unroll
    for (int c = 0; c < 8; c++) v[c] = submod(v[c], w[c], L.q);
    ntt_inv(v, lds[limb], t, L);
    stage_compose(v, lds, limb, t, S, xs);
    ep_poly(xs, p, A.l, A.b, g, limb, t, L, lds[limb], acc0, acc1);
    __syncthreads();  // xs is rewritten for the next polynomial
  }
  u64 e0[8], e1[8];
  mac_finish(acc0, e0, L);
  mac_finish(acc1, e1, L);
  u64 x0[8], x1[8];
  load3(x0, x + (size_t)(0 * 2 + limb) * kN, t);
  load3(x1, x + (size_t)(1 * 2 + limb) * kN, t);
#pragma unroll
  for (int c = 0; c < 8; c++) {
    x0[c] = addmod(x0[c], e0[c], L.q);
    x1[c] = addmod(x1[c], e1[c], L.q);
  }
  store3(dst + (size_t)(0 * 2 + limb) * kN, x0, t);
  store3(dst + (size_t)(1 * 2 + limb) * kN, x1, t);
}

// INTT (when ntt_in), then the K = 2 mod switch per coefficient. grid (2 polys, B).
__global__ __launch_bounds__(1024) void k2_final(const u64 *res, u64 *out, size_t res_stride, int ntt_in,
                                                 Scheme64 S) {
  __shared__ u64 lds[2][kLds];
  const int limb = __builtin_amdgcn_readfirstlane(threadIdx.x >> 9), t = threadIdx.x & 511;
  const Limb64 &L = S.limb[limb];
  const int p = blockIdx.x, qi = blockIdx.y;
  const u64 *src = res + qi * res_stride + (size_t)(p * 2 + limb) * kN;
  u64 v[8];
  if (ntt_in) {
    load3(v, src, t);
    ntt_inv(v, lds[limb], t, L);
  } else {
    load0(v, src, t);
  }
  __syncthreads();
  store0(lds[limb], v, t);
  __syncthreads();
  u64 *o = out + ((size_t)qi * 2 + p) * kN;
  for (int j = threadIdx.x; j < kN; j += 1024) o[j] = final_coeff(lds[0][j], lds[1][j], S);
}

__global__ __launch_bounds__(1024) void k2_db_load(const u64 *pts, u64 *db, size_t num_pt, size_t first,
                                                   Scheme64 S) {
  __shared__ u64 lds[2][kLds];
  const int limb = __builtin_amdgcn_readfirstlane(threadIdx.x >> 9), t = threadIdx.x & 511;
  const Limb64 &L = S.limb[limb];
  const size_t i = blockIdx.x;
  u64 v[8];
  load0(v, pts + i * kN, t);  // < t < q: already reduced
  ntt_fwd(v, lds[limb], t, L);
  store3(db + ((size_t)limb * num_pt + first + i) * kN, v, t);
}

__global__ __launch_bounds__(1024) void k2_ext_prod(const u64 *gsw, int l, int b, const u64 *ct, u64 *out,
                                                    Scheme64 S) {
  __shared__ u128 xs[kN];  // recombined coefficients of the polynomial being decomposed
  __shared__ u64 lds[2][kLds];
  const int limb = __builtin_amdgcn_readfirstlane(threadIdx.x >> 9), t = threadIdx.x & 511;
  const Limb64 &L = S.limb[limb];
  u128 s0[8], s1[8];
#pragma unroll
  for (int c = 0; c < 8; c++) s0[c] = s1[c] = 0;
#pragma unroll 1
  for (int p = 0; p < 2; p++) {
    u64 v[8];
    load0(v, ct + (size_t)(p * 2 + limb) * kN, t);
    stage_compose(v, lds, limb, t, S, xs);
    ep_poly(xs, p, l, b, gsw, limb, t, L, lds[limb], s0, s1);
    __syncthreads();  // xs is rewritten for the next polynomial
  }
  u64 acc0[8], acc1[8];
  mac_finish(s0, acc0, L);
  mac_finish(s1, acc1, L);
  store3(out + (size_t)(0 * 2 + limb) * kN, acc0, t);
  store3(out + (size_t)(1 * 2 + limb) * kN, acc1, t);
}

// ---------------------------------------------------------------------------
// Split tree levels for levels with few nodes: a node's digits are spread
// over several workgroups, each writing its partial external product (NTT
// form, reduced mod q) as one ciphertext block of `part`; a combine kernel
// adds them up. Part block of (query qi, node or muxed slot m, group):
//   MUX:       qi * (nmux * 2G) + (m * 2 + p) * G + g
//   expansion: qi * (nnodes * G) + node * G + g   ([delta_b, delta_a][limb][N])
// ---------------------------------------------------------------------------
__device__ __forceinline__ void mux_operands64(const MuxArgs64 &A, int m, int &xi, int &yi) {
  if (A.offset >= 0) {
    xi = A.offset + 2 * m;
    yi = xi + 1;
  } else {
    xi = m;
    yi = m + A.half;
  }
}

__global__ __launch_bounds__(1024) void k2_mux_part(MuxArgs64 A, u64 *__restrict__ part, int G, Scheme64 S) {
  __shared__ u128 xs[kN];
  __shared__ u64 lds[2][kLds];
  const int limb = __builtin_amdgcn_readfirstlane(threadIdx.x >> 9), t = threadIdx.x & 511;
  const Limb64 &L = S.limb[limb];
  const int qi = blockIdx.y;
  const int g = blockIdx.x % G, p = (blockIdx.x / G) & 1, m = blockIdx.x / (2 * G);
  int xi, yi;
  mux_operands64(A, m, xi, yi);
  const u64 *in = A.in + qi * A.stride;
  const u64 *x = in + (size_t)xi * kCt, *y = in + (size_t)yi * kCt;
  u64 v[8], w[8];
  load3(v, y + (size_t)(p * 2 + limb) * kN, t);
  load3(w, x + (size_t)(p * 2 + limb) * kN, t);
#pragma unroll
  for (int c = 0; c < 8; c++) v[c] = submod(v[c], w[c], L.q);
  ntt_inv(v, lds[limb], t, L);
  stage_compose(v, lds, limb, t, S, xs);
  u128 s0[8], s1[8];
#pragma unroll
  for (int c = 0; c < 8; c++) s0[c] = s1[c] = 0;
  const int dg = (A.l + G - 1) / G, i0 = g * dg, i1 = min(A.l, i0 + dg);
  ep_poly(xs, p, A.l, A.b, A.sel + qi * A.sel_stride, limb, t, L, lds[limb], s0, s1, i0, i1);
  u64 acc0[8], acc1[8];
  mac_finish(s0, acc0, L);
  mac_finish(s1, acc1, L);
  u64 *dst = part + ((size_t)qi * gridDim.x + blockIdx.x) * kCt;
  store3(dst + (size_t)(0 * 2 + limb) * kN, acc0, t);
  store3(dst + (size_t)(1 * 2 + limb) * kN, acc1, t);
}

// out[s] = in[s] for ragged slots s < offset, else x + the node's 2G partials.
// 32 workgroups per slot, one 16-byte word per thread.
__global__ __launch_bounds__(256) void k2_mux_fin(MuxArgs64 A, const u64 *__restrict__ part, int G, int nmux,
                                                  Scheme64 S) {
  constexpr int kWords = (int)kCt / 2;  // 16-byte words per ciphertext
  const int s = blockIdx.x >> 5, qi = blockIdx.y;
  const int e = (blockIdx.x & 31) * 256 + threadIdx.x;
  const u64 *in = A.in + qi * A.stride;
  ulonglong2 *dst = reinterpret_cast<ulonglong2 *>(A.out + qi * A.stride + (size_t)s * kCt);
  if (A.offset >= 0 && s < A.offset) {
    dst[e] = reinterpret_cast<const ulonglong2 *>(in + (size_t)s * kCt)[e];
    return;
  }
  const int m = A.offset >= 0 ? s - A.offset : s;
  int xi, yi;
  mux_operands64(A, m, xi, yi);
  const ulonglong2 a = reinterpret_cast<const ulonglong2 *>(in + (size_t)xi * kCt)[e];
  const ulonglong2 *pp =
      reinterpret_cast<const ulonglong2 *>(part + ((size_t)qi * nmux * 2 * G + (size_t)m * 2 * G) * kCt);
  constexpr int kMaxParts = 16;
  ulonglong2 b[kMaxParts];
#pragma unroll
  for (int j = 0; j < kMaxParts; j++)
    if (j < 2 * G) b[j] = pp[(size_t)j * kWords + e];
  const Limb64 &L = S.limb[((2 * e) / kN) & 1];  // polynomial index = output * 2 + limb
  u64 s0 = a.x, s1 = a.y;
#pragma unroll
  for (int j = 0; j < kMaxParts; j++)
    if (j < 2 * G) {
      s0 = addmod(s0, b[j].x, L.q);
      s1 = addmod(s1, b[j].y, L.q);
    }
  dst[e] = make_ulonglong2(s0, s1);
}

__global__ __launch_bounds__(1024) void k2_expand_part(ExpandArgs64 A, u64 *__restrict__ part, int G, Scheme64 S) {
  __shared__ u128 xs[kN];
  __shared__ u64 lds[2][kLds];
  const int limb = __builtin_amdgcn_readfirstlane(threadIdx.x >> 9), t = threadIdx.x & 511;
  const Limb64 &L = S.limb[limb];
  const int node = blockIdx.x / G, g = blockIdx.x % G, qi = blockIdx.y;
  const u64 *c = A.in + qi * A.stride + (size_t)node * kCt;
  const u64 *c1 = c + (size_t)(2 + limb) * kN;
  u64 v[8];
#pragma unroll
  for (int cc = 0; cc < 8; cc++) {
    const int j = map0(t, cc);
    const int i0 = (int)(((unsigned)j * (unsigned)A.kinv) & (2 * kN - 1));
    const u64 x = c1[i0 < kN ? i0 : i0 - kN];
    v[cc] = (i0 < kN || !x) ? x : L.q - x;
  }
  stage_compose(v, lds, limb, t, S, xs);
  {
    const u128 Q = Q128(S), hQ = hQ128(S);
    for (int j = threadIdx.x; j < kN; j += 1024)
      if (xs[j] > hQ) xs[j] -= Q;
    __syncthreads();
  }
  u128 sb[8], sa[8];
#pragma unroll
  for (int cc = 0; cc < 8; cc++) sb[cc] = sa[cc] = 0;
  const int b = A.b_ks;
  const i128 h = ((i128)1 << (b - 1)) - 1;
  const u64 *key = A.bv[qi] + (size_t)A.level * A.key_level_values;
  const int dg = (A.l_ks + G - 1) / G, d0 = g * dg, d1 = min(A.l_ks, d0 + dg);
  i128 H = 0;
  for (int i = 0; i < d0; i++) H += h << (b * i);
  const int t_in = t;
  for (int i = d0; i < d1; i++) {
    int t = t_in;
    asm volatile("" : "+v"(t));
    u64 d[8];
#pragma unroll
    for (int cc = 0; cc < 8; cc++) {
      const i128 di = ((i128)xs[map0(t, cc)] + H) >> (b * i);
      const int64_t r = (int64_t)(((u64)(di + h)) & (((u64)1 << b) - 1)) - (int64_t)h;
      d[cc] = sres(r, L.q);
    }
    H += h << (b * i);
    ntt_fwd(d, lds[limb], t, L, false);
    const u64 *row = key + (size_t)i * kCt;
    mac_row(sb, d, row + (size_t)(0 * 2 + limb) * kN, t, L);
    mac_row(sa, d, row + (size_t)(1 * 2 + limb) * kN, t, L);
  }
  u64 accb[8], acca[8];
  mac_finish(sb, accb, L);
  mac_finish(sa, acca, L);
  u64 *dst = part + ((size_t)qi * gridDim.x + blockIdx.x) * kCt;
  store3(dst + (size_t)(0 * 2 + limb) * kN, accb, t);
  store3(dst + (size_t)(1 * 2 + limb) * kN, acca, t);
}

// Workgroup (node, h), lane block = limb: delta_h = INTT(sum of the G
// partials), then child polynomials h as in k2_expand_level.
__global__ __launch_bounds__(1024) void k2_expand_fin(ExpandArgs64 A, const u64 *__restrict__ part, int G, int nnodes,
                                                      Scheme64 S) {
  __shared__ u64 lds[2][kLds];
  const int limb = __builtin_amdgcn_readfirstlane(threadIdx.x >> 9), t = threadIdx.x & 511;
  const Limb64 &L = S.limb[limb];
  const int node = blockIdx.x >> 1, h = blockIdx.x & 1, qi = blockIdx.y;
  const u64 *c = A.in + qi * A.stride + (size_t)node * kCt;
  const u64 *pp = part + ((size_t)qi * nnodes * G + (size_t)node * G) * kCt + (size_t)(h * 2 + limb) * kN;
  u64 w[8];
  load3(w, pp, t);
  for (int j = 1; j < G; j++) {
    u64 u[8];
    load3(u, pp + (size_t)j * kCt, t);
#pragma unroll
    for (int cc = 0; cc < 8; cc++) w[cc] = addmod(w[cc], u[cc], L.q);
  }
  ntt_inv(w, lds[limb], t, L);  // delta_h, coefficient form (map0)
  u64 *left = A.out + qi * A.stride + (size_t)(2 * node) * kCt;
  u64 *right = left + kCt;
  const int k = 1 << A.level;
  const u64 *cp = c + (size_t)(h * 2 + limb) * kN;
  u64 cv[8], dv[8];
  load0(cv, cp, t);
#pragma unroll
  for (int cc = 0; cc < 8; cc++) {
    u64 ww = w[cc];
    if (h == 0) {  // c'_0 = sigma(c0) + delta_b
      const int j = map0(t, cc);
      const int i0 = (int)(((unsigned)j * (unsigned)A.kinv) & (2 * kN - 1));
      const u64 x = cp[i0 < kN ? i0 : i0 - kN];
      ww = addmod(ww, (i0 < kN || !x) ? x : L.q - x, L.q);
    }
    left[(size_t)(h * 2 + limb) * kN + map0(t, cc)] = addmod(cv[cc], ww, L.q);
    dv[cc] = submod(cv[cc], ww, L.q);
  }
  __syncthreads();  // the INTT's last reads of lds are done
  store0(lds[limb], dv, t);
  __syncthreads();
#pragma unroll
  for (int cc = 0; cc < 8; cc++) {
    const int j = map0(t, cc), s = j + k;
    const u64 x = lds[limb][s < kN ? s : s - kN];
    right[(size_t)(h * 2 + limb) * kN + j] = (s < kN || !x) ? x : L.q - x;
  }
}

// Digit groups of a split level: one digit per workgroup, at most 8 groups.
static int split_groups64(int l) {
  const int dg = (l + 7) / 8;
  return (l + dg - 1) / dg;
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
void launch_k2_ntt(u64 *data, size_t count, bool inverse, const Scheme64 &S, hipStream_t st) {
  if (count) k2_ntt<<<(unsigned)count, 1024, 0, st>>>(data, inverse ? 1 : 0, S);
}

void launch_k2_bv_apply(const u64 *in, u64 *out, const u64 *const *key_ptr, int kinv, const Gadget64 &G,
                        const Scheme64 &S, hipStream_t st) {
  ExpandArgs64 A;
  A.in = in; A.out = out; A.stride = 0; A.bv = key_ptr; A.key_level_values = 0;
  A.level = 0; A.l_ks = G.l_ks; A.b_ks = G.b_ks; A.kinv = kinv;
  A.bv_only = 1;
  k2_expand_level<<<dim3(1, 1), 1024, 0, st>>>(A, S);
}

void launch_k2_expand_level(const u64 *in, u64 *out, size_t stride, int nnodes, int B, int level,
                            const KeyPtrs64 &P, size_t key_level_values, const Gadget64 &G, const Scheme64 &S,
                            hipStream_t st, u64 *part) {
  if (nnodes <= 0 || B <= 0) return;
  ExpandArgs64 A;
  A.in = in; A.out = out; A.stride = stride; A.bv = P.bv; A.key_level_values = key_level_values;
  A.level = level; A.l_ks = G.l_ks; A.b_ks = G.b_ks;
  const unsigned gk = (unsigned)(kN >> level) + 1;
  unsigned inv = 1;  // gk^-1 mod 2N = gk^(N-1): the unit group mod 2N = 2^13 has order N
  for (int e = 0; e < kN - 1; e++) inv = (inv * gk) & (2 * kN - 1);
  A.kinv = (int)inv;
  const int Gs = split_groups64(G.l_ks);
  if (part && (size_t)nnodes * Gs * B <= kSplitTargetBlocks64) {
    k2_expand_part<<<dim3(nnodes * Gs, B), 1024, 0, st>>>(A, part, Gs, S);
    k2_expand_fin<<<dim3(nnodes * 2, B), 1024, 0, st>>>(A, part, Gs, nnodes, S);
    return;
  }
  k2_expand_level<<<dim3(nnodes, B), 1024, 0, st>>>(A, S);
}

void launch_k2_complete(const u64 *rows, u64 *sel, size_t row_stride, size_t sel_stride, int nrows, int B,
                        const KeyPtrs64 &P, const Gadget64 &G, const Scheme64 &S, hipStream_t st) {
  if (nrows <= 0 || B <= 0) return;
  CompleteArgs64 A;
  A.rows = rows; A.sel = sel; A.row_stride = row_stride; A.sel_stride = sel_stride; A.gsw = P.gsw;
  A.l_ep = G.l_ep; A.l_key = G.l_key; A.b_key = G.b_key;
  k2_complete<<<dim3(nrows, B), 1024, 0, st>>>(A, S);
}

void launch_k2_ct_ntt(const u64 *in, u64 *out, size_t in_stride, size_t out_stride, int count, int B,
                      const Scheme64 &S, hipStream_t st) {
  if (count <= 0 || B <= 0) return;
  k2_ct_ntt<<<dim3(2 * count, B), 1024, 0, st>>>(in, out, in_stride, out_stride, S);
}

void launch_k2_mux(const u64 *in, u64 *out, size_t stride, int nslots, int offset, int half, int B, const u64 *sel,
                   size_t sel_stride, const Gadget64 &G, const Scheme64 &S, hipStream_t st, u64 *part) {
  if (nslots <= 0 || B <= 0) return;
  MuxArgs64 A;
  A.in = in; A.out = out; A.stride = stride; A.sel = sel; A.sel_stride = sel_stride;
  A.offset = offset; A.half = half; A.l = G.l_ep; A.b = G.b_ep;
  const int nmux = offset >= 0 ? nslots - offset : nslots;
  const size_t pairs = (size_t)nmux * 2 * B;
  if (part && nmux > 0 && pairs <= kSplitTargetBlocks64) {
    int Gs = (int)std::min<size_t>(split_groups64(G.l_ep), kSplitTargetBlocks64 / pairs);
    const int dg = (G.l_ep + Gs - 1) / Gs;
    Gs = (G.l_ep + dg - 1) / dg;  // no empty groups
    k2_mux_part<<<dim3(nmux * 2 * Gs, B), 1024, 0, st>>>(A, part, Gs, S);
    k2_mux_fin<<<dim3(nslots * 32, B), 256, 0, st>>>(A, part, Gs, nmux, S);
    return;
  }
  k2_mux<<<dim3(nslots, B), 1024, 0, st>>>(A, S);
}

void launch_k2_final(const u64 *res, u64 *out, size_t res_stride, int B, const Scheme64 &S, hipStream_t st) {
  if (B > 0) k2_final<<<dim3(2, B), 1024, 0, st>>>(res, out, res_stride, 1, S);
}

void launch_k2_mod_switch(const u64 *ct, u64 *out, const Scheme64 &S, hipStream_t st) {
  k2_final<<<dim3(2, 1), 1024, 0, st>>>(ct, out, 0, 0, S);
}

void launch_k2_db_load(const u64 *pts, u64 *db, size_t num_pt, size_t first, size_t count, const Scheme64 &S,
                       hipStream_t st) {
  if (count) k2_db_load<<<(unsigned)count, 1024, 0, st>>>(pts, db, num_pt, first, S);
}

void launch_k2_ext_prod(const u64 *gsw, int l, int b, const u64 *ct, u64 *out, const Scheme64 &S, hipStream_t st) {
  k2_ext_prod<<<1, 1024, 0, st>>>(gsw, l, b, ct, out, S);
}

}  // namespace k2
}  // namespace opir
