// The 64-bit-limb answer path on CDNA4: CONFIG_N4096_K2_MP (N = 4096, two
// ~60-bit RNS primes) and CONFIG_N2048_K1 (N = 2048, one ~60-bit prime).
// Every kernel is a template over (N, K); the launchers pick the
// instantiation from the scheme.
//
// Arithmetic: each limb is a ~60-bit NTT prime, so values are u64 and every
// product is a 64x64->128 product. Modular multiplication by a known
// constant uses Shoup's method with a 64-bit companion (floor(w*2^64/q));
// the multiply-accumulates against key rows (products of two variables) sum
// exact 128-bit products and reduce each sum once, as hi*(2^64 mod q) + lo
// with both halves by Shoup. NTT butterflies are Harvey-lazy in [0, 4q)
// (4q < 2^62).
//
// Thread layout: an N-point transform is a lane block of N/8 threads with 8
// values each and three LDS exchanges. Value c of thread t sits at index
// g*8T + (t % T) + c*T with T = (N/8) >> 3r, g = t / T in rounds r = 0..2,
// and at 8t + c in round 3. Rounds 0-2 are radix-8; round 3 is radix-8 at
// N = 4096 and two radix-4 groups (the last two stages) at N = 2048. Tree
// nodes are workgroups of K lane blocks, one per limb.
//
// Gadgets:
//   K = 2 (src/gsw.cpp:22-85,184-250; src/bv_keyswitch.cpp:42-64,373-481):
//     the two limbs are recombined by CRT into X in [0, Q), Q ~ 2^120.
//     data / key selectors: unsigned digits (X >> (b*i)) & (2^b - 1),
//       stored at row poly*l + (l-1-i);
//     BV keyswitch: signed digits in (-B/2, B/2] of X centred to (-Q/2, Q/2],
//       digit i = ((floor((D + H_i) / B^i) + h) mod B) - h with h = B/2 - 1,
//       H_i = h * (1 + B + ... + B^(i-1)).
//   K = 1 (src/gsw.cpp:252-315; src/bv_keyswitch.cpp:20-40,296-367): every
//     gadget takes the value centred to (-q/2, q/2] as D and its balanced
//     digits in [-B/2, B/2), LSB-first: digit i = sext_b(floor((D + H_i) /
//     B^i)) with H_i = (B/2) * (1 + B + ... + B^(i-1)), stored mod q (data
//     and key digits at row poly*l + (l-1-i)).
//   Both closed forms restate the reference's serial carry loops, so every
//   digit is computed on its own.
#include <type_traits>

#include "opir_k2.h"

namespace opir {
namespace k2 {

typedef unsigned __int128 u128;
typedef __int128 i128;

template <int N_, int K_>
struct Cfg {
  static constexpr int N = N_, K = K_;
  static constexpr int LB = N / 8;               // threads per lane block (one limb)
  static constexpr int WG = K * LB;              // threads per tree-node workgroup
  static constexpr size_t Ct = 2 * (size_t)K * N;  // u64 values per ciphertext
  static constexpr int Lds = N + N / 32;         // padded u64 words of one exchange buffer
  // Waves per SIMD the tree kernels are compiled for. One prime: three, i.e.
  // three 256-thread workgroups per CU (168 registers, a few spilled): the
  // 64-query batch in 18.9 ms against 19.7-20.2 ms at two (its plain query
  // unchanged). Two primes: one 1024-thread workgroup per CU, bound by LDS.
  static constexpr int Waves = K == 1 ? 3 : 1;
};
typedef Cfg<4096, 2> CfgK2;
typedef Cfg<2048, 1> CfgK1;

namespace {

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
// N-point NTT on an N/8-thread lane block
// ---------------------------------------------------------------------------
__device__ __forceinline__ int pidx(int i) { return i + (i >> 5); }
template <int N>
__device__ __forceinline__ int r_index(int r, int t, int c) {
  if (r == 3) return 8 * t + c;
  const int T = (N / 8) >> (3 * r);
  return (t / T) * 8 * T + (t % T) + c * T;
}
template <int N>
__device__ __forceinline__ int map0(int t, int c) { return t + (N / 8) * c; }

// Forward butterfly (X, Y) -> (X + Y*w, X - Y*w) with t = Y*w mod q in
// [0, 2q) (Shoup, any Y < 2^64). Conditional subtractions on X first, by
// the bound X has on entry (every value of a stage has the same bound):
//   kFwN: none                                     -> outputs < bound + 2q
//   kFwH: X < 16q, subtract 8q                     -> outputs < 10q
//   kFwD: X < 16q, subtract 8q, then 4q, then 2q   -> outputs < 4q
// With q < 2^60 every bound stays under 16q < 2^64.
enum { kFwN, kFwH, kFwD };
template <int kKind>
__device__ __forceinline__ void bf_fwd(u64 &X, u64 &Y, u64 w, u64 ws, const Limb64 &L) {
  u64 x = X;
  if constexpr (kKind != kFwN) x = csub(x, 4 * L.q2);
  if constexpr (kKind == kFwD) x = csub(csub(x, 2 * L.q2), L.q2);
  const u64 t = shoup(Y, w, ws, L.q);
  X = x + t;
  Y = x - t + L.q2;
}
__device__ __forceinline__ void bf_inv(u64 &X, u64 &Y, u64 w, u64 ws, const Limb64 &L) {
  const u64 x = X, y = Y;
  X = csub(x + y, L.q2);
  Y = shoup(x - y + L.q2, w, ws, L.q);
}

// Radix-8 group b: stages with twiddles psi[b], psi[2b + h], psi[4b + h];
// K1, K2, K3: the butterfly kind of each stage.
template <int K1, int K2, int K3>
__device__ __forceinline__ void fwd8(u64 v[8], unsigned b, const Limb64 &L) {
  {
    const u64 w = L.psi[b], ws = L.psi_s[b];
#pragma unroll
    for (int c = 0; c < 4; c++) bf_fwd<K1>(v[c], v[c + 4], w, ws, L);
  }
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const u64 w = L.psi[2 * b + h], ws = L.psi_s[2 * b + h];
    bf_fwd<K2>(v[4 * h + 0], v[4 * h + 2], w, ws, L);
    bf_fwd<K2>(v[4 * h + 1], v[4 * h + 3], w, ws, L);
  }
#pragma unroll
  for (int h = 0; h < 4; h++) bf_fwd<K3>(v[2 * h], v[2 * h + 1], L.psi[4 * b + h], L.psi_s[4 * b + h], L);
}
// The last two stages of fwd8 only: two radix-4 groups (N = 2048's round 3).
template <int K2, int K3>
__device__ __forceinline__ void fwd4x2(u64 v[8], unsigned b, const Limb64 &L) {
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const u64 w = L.psi[2 * b + h], ws = L.psi_s[2 * b + h];
    bf_fwd<K2>(v[4 * h + 0], v[4 * h + 2], w, ws, L);
    bf_fwd<K2>(v[4 * h + 1], v[4 * h + 3], w, ws, L);
  }
#pragma unroll
  for (int h = 0; h < 4; h++) bf_fwd<K3>(v[2 * h], v[2 * h + 1], L.psi[4 * b + h], L.psi_s[4 * b + h], L);
}

__device__ __forceinline__ void inv4x2(u64 v[8], unsigned b, const Limb64 &L) {
#pragma unroll
  for (int h = 0; h < 4; h++) bf_inv(v[2 * h], v[2 * h + 1], L.ipsi[4 * b + h], L.ipsi_s[4 * b + h], L);
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const u64 w = L.ipsi[2 * b + h], ws = L.ipsi_s[2 * b + h];
    bf_inv(v[4 * h + 0], v[4 * h + 2], w, ws, L);
    bf_inv(v[4 * h + 1], v[4 * h + 3], w, ws, L);
  }
}
__device__ __forceinline__ void inv8(u64 v[8], unsigned b, const Limb64 &L) {
  inv4x2(v, b, L);
  {
    const u64 w = L.ipsi[b], ws = L.ipsi_s[b];
#pragma unroll
    for (int c = 0; c < 4; c++) bf_inv(v[c], v[c + 4], w, ws, L);
  }
}

// LDS exchanges between rounds. t is passed by reference and made opaque
// afterwards, so the next round's twiddle addresses (functions of t) cannot be
// computed and loaded before the exchange: hoisting all rounds' twiddles
// exceeds the 128-register budget of a 1024-thread workgroup.
//
// In maps 1, 2 and 3 the 64 threads of wave w (t in [64w, 64w + 64)) hold
// exactly indices [512w, 512w + 512), for N = 4096 and N = 2048 alike, so
// only the exchanges from and to map 0 move values between waves. The others
// go through the wave's own region of the buffer with no workgroup barrier:
// the hardware runs a wave's LDS instructions in order, and the wavefront
// fences only keep the compiler from reordering them.
__device__ __forceinline__ void wave_order() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
template <int N>
__device__ __forceinline__ void put(const u64 v[8], u64 *lds, int t, int r) {
#pragma unroll
  for (int c = 0; c < 8; c++) lds[pidx(r_index<N>(r, t, c))] = v[c];
}
template <int N>
__device__ __forceinline__ void get(u64 v[8], const u64 *lds, int t, int r) {
#pragma unroll
  for (int c = 0; c < 8; c++) v[c] = lds[pidx(r_index<N>(r, t, c))];
}
// Between maps 1, 2 and 3, inside each wave.
template <int N>
__device__ __forceinline__ void exch_wave(u64 v[8], u64 *lds, int &t, int from, int to) {
  wave_order();
  put<N>(v, lds, t, from);
  wave_order();
  get<N>(v, lds, t, to);
  asm volatile("" : "+v"(t));
}
// Map 0 -> map 1 (forward transforms): the first barrier lets every wave
// finish its reads of the buffer, whatever they were, before the scattered
// writes; the second publishes them. Afterwards each wave reads only its own
// region.
template <int N>
__device__ __forceinline__ void exch_from0(u64 v[8], u64 *lds, int &t) {
  __syncthreads();
  put<N>(v, lds, t, 0);
  __syncthreads();
  get<N>(v, lds, t, 1);
  asm volatile("" : "+v"(t));
}
// Map 1 -> map 0 (inverse transforms): each wave writes its own region (no
// other wave reads the buffer on entry), the barrier publishes them, and the
// last barrier lets every wave finish its scattered reads before the buffer
// is written again.
template <int N>
__device__ __forceinline__ void exch_to0(u64 v[8], u64 *lds, int &t) {
  wave_order();
  put<N>(v, lds, t, 1);
  __syncthreads();
  get<N>(v, lds, t, 0);
  __syncthreads();
  asm volatile("" : "+v"(t));
}

// map0 values below 2q in -> map3 out, [0, q) when reduce, else [0, 4q).
// Two workgroup barriers, both before the first round's values cross waves;
// on return no wave reads another wave's region of lds. Value bounds: six
// stages without a subtraction (below 14q), then 10q, 12q, 14q | 10q, 12q,
// 4q at N = 4096 (10q, 4q for the last two at N = 2048): five conditional
// subtractions per value where one per stage takes 11-12.
template <int N>
__device__ __forceinline__ void ntt_fwd(u64 v[8], u64 *lds, int t, const Limb64 &L, bool reduce = true) {
  constexpr int LB = N / 8;
  asm volatile("" : "+v"(t));
  fwd8<kFwN, kFwN, kFwN>(v, 1, L);
  exch_from0<N>(v, lds, t);
  fwd8<kFwN, kFwN, kFwN>(v, 8 + t / (LB / 8), L);
  exch_wave<N>(v, lds, t, 1, 2);
  fwd8<kFwH, kFwN, kFwN>(v, 64 + t / (LB / 64), L);
  exch_wave<N>(v, lds, t, 2, 3);
  if constexpr (N == 4096)
    fwd8<kFwH, kFwN, kFwD>(v, 512 + t, L);
  else
    fwd4x2<kFwH, kFwD>(v, 256 + t, L);
  if (reduce) {
#pragma unroll
    for (int c = 0; c < 8; c++) v[c] = csub(csub(v[c], L.q2), L.q);
  }
}

// map3 [0, q) in -> map0 out in [0, q). On entry no other wave may still
// read this wave's region of lds (every transform and every staging use of
// lds in these kernels leaves it so); two workgroup barriers, the last after
// the scattered reads, so lds is free for any writes on return.
template <int N>
__device__ __forceinline__ void ntt_inv(u64 v[8], u64 *lds, int t, const Limb64 &L) {
  constexpr int LB = N / 8;
  asm volatile("" : "+v"(t));
  if constexpr (N == 4096)
    inv8(v, 512 + t, L);
  else
    inv4x2(v, 256 + t, L);
  exch_wave<N>(v, lds, t, 3, 2);
  inv8(v, 64 + t / (LB / 64), L);
  exch_wave<N>(v, lds, t, 2, 1);
  inv8(v, 8 + t / (LB / 8), L);
  exch_to0<N>(v, lds, t);
  inv8(v, 1, L);
#pragma unroll
  for (int c = 0; c < 8; c++) v[c] = csub(shoup(v[c], L.ninv, L.ninv_s, L.q), L.q);
}

template <int N>
__device__ __forceinline__ void load0(u64 v[8], const u64 *p, int t) {
#pragma unroll
  for (int c = 0; c < 8; c++) v[c] = p[map0<N>(t, c)];
}
template <int N>
__device__ __forceinline__ void store0(u64 *p, const u64 v[8], int t) {
#pragma unroll
  for (int c = 0; c < 8; c++) p[map0<N>(t, c)] = v[c];
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
__device__ __forceinline__ void mac_row(u128 (&acc)[8], const u64 v[8], const u64 *key, int t) {
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
__device__ __forceinline__ void zero(u128 (&a)[8], u128 (&b)[8]) {
#pragma unroll
  for (int c = 0; c < 8; c++) a[c] = b[c] = 0;
}

// ---------------------------------------------------------------------------
// Gadget sources: the coefficients of one coefficient-form polynomial, as the
// decomposition needs them. K = 2: X = CRT of both limbs in [0, Q), staged in
// LDS (both lane blocks read every coefficient's two residues); K = 1: the
// limb value centred to (-q/2, q/2], in registers.
// ---------------------------------------------------------------------------
template <typename C, int K = C::K>
struct Coef;

template <typename C>
struct Coef<C, 2> {
  u128 *xs;  // [N] in LDS
  // v: this lane block's limb of the polynomial (map0). centre: map X to
  // the two's complement of X - Q when X > Q/2 (the BV gadget's view).
  __device__ __forceinline__ void put(const u64 v[8], u64 (*stage)[C::Lds], int limb, int t, const Scheme64 &S,
                                      bool centre) {
    store0<C::N>(stage[limb], v, t);
    __syncthreads();
    for (int j = threadIdx.x; j < C::N; j += C::WG) {
      u128 x = crt2(stage[0][j], stage[1][j], S);
      if (centre && x > hQ128(S)) x -= Q128(S);  // wraps to the two's complement of X - Q
      xs[j] = x;
    }
    __syncthreads();
  }
};

template <typename C>
struct Coef<C, 1> {
  int64_t D[8];  // centred values of this thread's coefficients (map0)
  __device__ __forceinline__ void put(const u64 v[8], u64 (*)[C::Lds], int, int, const Scheme64 &S, bool) {
    const u64 q = S.limb[0].q, hq = q >> 1;
#pragma unroll
    for (int c = 0; c < 8; c++) D[c] = v[c] > hq ? (int64_t)v[c] - (int64_t)q : (int64_t)v[c];
  }
};

// Balanced digit i (base 2^b) of centred D with the running offset H = H_i,
// as r + q, a representative mod q (K = 1 gadgets, gsw.cpp:285-293,
// bv_keyswitch.cpp:32-39).
__device__ __forceinline__ u64 bal_digit(int64_t D, int64_t H, int i, int b, u64 q) {
  const int64_t x = (D + H) >> (b * i);
  const int64_t r = (int64_t)((u64)x << (64 - b)) >> (64 - b);
  return (u64)(r + (int64_t)q);  // in (0, 2q): the forward transform takes inputs below 2q
}
// H_i = sum_{j=1..i} 2^(b*j - 1)
__device__ __forceinline__ int64_t bal_offset(int i, int b) {
  int64_t H = 0;
  for (int j = 1; j <= i; j++) H += (int64_t)1 << (b * j - 1);
  return H;
}

// One polynomial's external-product contribution (gsw.cpp:119-182): digits
// i in [i0, i1) of l (base 2^b) of X go through the forward NTT of this lane
// block's limb and are multiplied into acc0/acc1 with rows p*l + (l-1-i) of
// g ([rows][2 outputs][K limbs][N], NTT form).
template <typename C>
__device__ __forceinline__ void ep_poly(const Coef<C> &X, int p, int l, int b, const u64 *__restrict__ g, int limb,
                                        int t_in, const Limb64 &L, u64 *lds, u128 (&acc0)[8], u128 (&acc1)[8],
                                        int i0 = 0, int i1 = -1) {
  constexpr int N = C::N, K = C::K;
  if (i1 < 0) i1 = l;
  int64_t H = 0;
  if constexpr (K == 1) H = bal_offset(i0, b);
#pragma unroll 1
  for (int i = i0; i < i1; i++) {
    // opaque copy of t: keeps per-thread addresses inside the loop (register pressure)
    int t = t_in;
    asm volatile("" : "+v"(t));
    u64 v[8];
    if constexpr (K == 2) {
      const u64 mask = ((u64)1 << b) - 1;
#pragma unroll
      for (int c = 0; c < 8; c++) v[c] = (u64)(X.xs[map0<N>(t, c)] >> (b * i)) & mask;  // < 2^b < q
    } else {
      if (i > i0) H += (int64_t)1 << (b * i - 1);
#pragma unroll
      for (int c = 0; c < 8; c++) v[c] = bal_digit(X.D[c], H, i, b, L.q);
    }
    ntt_fwd<N>(v, lds, t, L, false);
    const u64 *row = g + (size_t)(p * l + (l - 1 - i)) * C::Ct;
    mac_row(acc0, v, row + (size_t)(0 * K + limb) * N, t);
    mac_row(acc1, v, row + (size_t)(1 * K + limb) * N, t);
  }
}

// BV keyswitch digits [d0, d1) of sigma(c1) (gadget source X, centred) times
// key rows i ([b (-> c0), a (-> c1)][K limbs][N]) into sb, sa.
template <typename C>
__device__ __forceinline__ void bv_mac(const Coef<C> &X, const u64 *key, int d0, int d1, int b, int limb, int t_in,
                                       const Limb64 &L, u64 *lds, u128 (&sb)[8], u128 (&sa)[8]) {
  constexpr int N = C::N, K = C::K;
  if constexpr (K == 2) {
    const i128 h = ((i128)1 << (b - 1)) - 1;
    i128 H = 0;
    for (int i = 0; i < d0; i++) H += h << (b * i);
    for (int i = d0; i < d1; i++) {
      int t = t_in;
      asm volatile("" : "+v"(t));
      u64 d[8];
#pragma unroll
      for (int cc = 0; cc < 8; cc++) {
        const i128 di = ((i128)X.xs[map0<N>(t, cc)] + H) >> (b * i);
        const int64_t r = (int64_t)(((u64)(di + h)) & (((u64)1 << b) - 1)) - (int64_t)h;
        d[cc] = (u64)(r + (int64_t)L.q);  // in (0, 2q), taken as is by the forward transform
      }
      H += h << (b * i);
      ntt_fwd<N>(d, lds, t, L, false);
      const u64 *row = key + (size_t)i * C::Ct;
      mac_row(sb, d, row + (size_t)(0 * K + limb) * N, t);
      mac_row(sa, d, row + (size_t)(1 * K + limb) * N, t);
    }
  } else {
    int64_t H = bal_offset(d0, b);
    for (int i = d0; i < d1; i++) {
      int t = t_in;
      asm volatile("" : "+v"(t));
      if (i > d0) H += (int64_t)1 << (b * i - 1);
      u64 d[8];
#pragma unroll
      for (int cc = 0; cc < 8; cc++) d[cc] = bal_digit(X.D[cc], H, i, b, L.q);
      ntt_fwd<N>(d, lds, t, L, false);
      const u64 *row = key + (size_t)i * C::Ct;
      mac_row(sb, d, row + (size_t)(0 * K + limb) * N, t);
      mac_row(sa, d, row + (size_t)(1 * K + limb) * N, t);
    }
  }
}

// sigma_k of coefficient j of a coefficient-form limb p (mod q): in[i0] or
// -in[i0 - N] with i0 = j * kinv mod 2N.
template <int N>
__device__ __forceinline__ u64 autom(const u64 *p, int j, int kinv, u64 q) {
  const int i0 = (int)(((unsigned)j * (unsigned)kinv) & (2 * N - 1));
  const u64 x = p[i0 < N ? i0 : i0 - N];
  return (i0 < N || !x) ? x : q - x;
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

template <typename C>
__device__ __forceinline__ int lane_limb() {
  if constexpr (C::K == 1) return 0;
  else return __builtin_amdgcn_readfirstlane(threadIdx.x / C::LB);
}

}  // namespace

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------
template <typename C>
__global__ __launch_bounds__(C::WG) __attribute__((amdgpu_waves_per_eu(C::Waves))) void k2_ntt(u64 *data, int inverse, Scheme64 S) {
  __shared__ u64 lds[C::K][C::Lds];
  const int limb = lane_limb<C>(), t = threadIdx.x % C::LB;
  const Limb64 &L = S.limb[limb];
  u64 *p = data + ((size_t)blockIdx.x * C::K + limb) * C::N;
  u64 v[8];
  if (!inverse) {
    load0<C::N>(v, p, t);
    ntt_fwd<C::N>(v, lds[limb], t, L);
    store3(p, v, t);
  } else {
    load3(v, p, t);
    ntt_inv<C::N>(v, lds[limb], t, L);
    store0<C::N>(p, v, t);
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

// Expansion level (server.cpp:786-810, bv_keyswitch.cpp:296-481).
template <typename C>
__global__ __launch_bounds__(C::WG) __attribute__((amdgpu_waves_per_eu(C::Waves))) void k2_expand_level(ExpandArgs64 A, Scheme64 S) {
  constexpr int N = C::N, K = C::K;
  __shared__ u128 xs[K == 2 ? N : 1];  // K = 2: recombined coefficients being decomposed
  __shared__ u64 lds[K][C::Lds];
  const int limb = lane_limb<C>(), t = threadIdx.x % C::LB;
  const Limb64 &L = S.limb[limb];
  const int node = blockIdx.x, qi = blockIdx.y;
  const u64 *c = A.in + qi * A.stride + (size_t)node * C::Ct;
  const u64 *c1 = c + (size_t)(K + limb) * N, *c0 = c + (size_t)limb * N;
  u64 v[8];
#pragma unroll
  for (int cc = 0; cc < 8; cc++) v[cc] = autom<N>(c1, map0<N>(t, cc), A.kinv, L.q);  // sigma_k(c1)
  Coef<C> X;
  if constexpr (K == 2) X.xs = xs;
  X.put(v, lds, limb, t, S, true);
  u128 sb[8], sa[8];
  zero(sb, sa);
  bv_mac<C>(X, A.bv[qi] + (size_t)A.level * A.key_level_values, 0, A.l_ks, A.b_ks, limb, t, L, lds[limb], sb, sa);
  u64 accb[8], acca[8];
  mac_finish(sb, accb, L);
  mac_finish(sa, acca, L);
  ntt_inv<N>(accb, lds[limb], t, L);  // delta_b, coefficient form (map0)
  ntt_inv<N>(acca, lds[limb], t, L);  // delta_a
  // c' = (sigma(c0) + delta_b, delta_a); left = c + c', right = (c - c') * X^-k
  u64 *left = A.out + qi * A.stride + (size_t)(2 * node) * C::Ct;
  u64 *right = left + C::Ct;
  const int k = 1 << A.level;
  for (int p = 0; p < 2; p++) {
    const u64 *cp = p ? c1 : c0;
    u64 cv[8], dv[8];
    load0<N>(cv, cp, t);
#pragma unroll
    for (int cc = 0; cc < 8; cc++) {
      u64 w = p ? acca[cc] : accb[cc];
      if (p == 0) w = addmod(w, autom<N>(c0, map0<N>(t, cc), A.kinv, L.q), L.q);
      left[(size_t)(p * K + limb) * N + map0<N>(t, cc)] = A.bv_only ? w : addmod(cv[cc], w, L.q);
      dv[cc] = submod(cv[cc], w, L.q);
    }
    if (A.bv_only) continue;  // uniform over the workgroup
    store0<N>(lds[limb], dv, t);
    __syncthreads();
#pragma unroll
    for (int cc = 0; cc < 8; cc++) {
      const int j = map0<N>(t, cc), s = j + k;
      const u64 x = lds[limb][s < N ? s : s - N];
      right[(size_t)(p * K + limb) * N + j] = (s < N || !x) ? x : L.q - x;
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
template <typename C>
__global__ __launch_bounds__(C::WG) __attribute__((amdgpu_waves_per_eu(C::Waves))) void k2_complete(CompleteArgs64 A, Scheme64 S) {
  constexpr int N = C::N, K = C::K;
  __shared__ u128 xs[K == 2 ? N : 1];
  __shared__ u64 lds[K][C::Lds];
  const int limb = lane_limb<C>(), t = threadIdx.x % C::LB;
  const Limb64 &L = S.limb[limb];
  const int g = blockIdx.x, qi = blockIdx.y;
  const int a = g / A.l_ep, r = g % A.l_ep;
  const u64 *row = A.rows + qi * A.row_stride + (size_t)g * C::Ct;
  u64 *selb = A.sel + qi * A.sel_stride + (size_t)a * (2 * A.l_ep) * C::Ct;
  u128 s0[8], s1[8];
  zero(s0, s1);
  Coef<C> X;
  if constexpr (K == 2) X.xs = xs;
#pragma unroll 1
  for (int p = 0; p < 2; p++) {
    u64 v[8];
    load0<N>(v, row + (size_t)(p * K + limb) * N, t);
    X.put(v, lds, limb, t, S, false);
    ntt_fwd<N>(v, lds[limb], t, L);
    store3(selb + (size_t)r * C::Ct + (size_t)(p * K + limb) * N, v, t);
    ep_poly<C>(X, p, A.l_key, A.b_key, A.gsw[qi], limb, t, L, lds[limb], s0, s1);
    if constexpr (K == 2) __syncthreads();  // xs is rewritten for the next polynomial
  }
  u64 acc0[8], acc1[8];
  mac_finish(s0, acc0, L);
  mac_finish(s1, acc1, L);
  u64 *bot = selb + (size_t)(A.l_ep + r) * C::Ct;
  store3(bot + (size_t)(0 * K + limb) * N, acc0, t);
  store3(bot + (size_t)(1 * K + limb) * N, acc1, t);
}

template <typename C>
__global__ __launch_bounds__(C::WG) __attribute__((amdgpu_waves_per_eu(C::Waves))) void k2_ct_ntt(const u64 *in, u64 *out, size_t in_stride, size_t out_stride,
                                                   Scheme64 S) {
  constexpr int N = C::N, K = C::K;
  __shared__ u64 lds[K][C::Lds];
  const int limb = lane_limb<C>(), t = threadIdx.x % C::LB;
  const Limb64 &L = S.limb[limb];
  const int ct = blockIdx.x >> 1, p = blockIdx.x & 1, qi = blockIdx.y;
  u64 v[8];
  load0<N>(v, in + qi * in_stride + (size_t)ct * C::Ct + (size_t)(p * K + limb) * N, t);
  ntt_fwd<N>(v, lds[limb], t, L);
  store3(out + qi * out_stride + (size_t)ct * C::Ct + (size_t)(p * K + limb) * N, v, t);
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
template <typename C>
__global__ __launch_bounds__(C::WG) __attribute__((amdgpu_waves_per_eu(C::Waves))) void k2_mux(MuxArgs64 A, Scheme64 S) {
  constexpr int N = C::N, K = C::K;
  __shared__ u128 xs[K == 2 ? N : 1];
  __shared__ u64 lds[K][C::Lds];
  const int limb = lane_limb<C>(), t = threadIdx.x % C::LB;
  const Limb64 &L = S.limb[limb];
  const int s = blockIdx.x, qi = blockIdx.y;
  const u64 *in = A.in + qi * A.stride;
  u64 *dst = A.out + qi * A.stride + (size_t)s * C::Ct;
  int xi, yi;
  if (A.offset >= 0) {
    if (s < A.offset) {
      const u64 *src = in + (size_t)s * C::Ct;
      for (int e = threadIdx.x; e < (int)C::Ct; e += C::WG) dst[e] = src[e];
      return;
    }
    xi = A.offset + 2 * (s - A.offset);
    yi = xi + 1;
  } else {
    xi = s;
    yi = s + A.half;
  }
  const u64 *x = in + (size_t)xi * C::Ct, *y = in + (size_t)yi * C::Ct;
  const u64 *g = A.sel + qi * A.sel_stride;
  u128 acc0[8], acc1[8];
  zero(acc0, acc1);
  Coef<C> X;
  if constexpr (K == 2) X.xs = xs;
#pragma unroll 1
  for (int p = 0; p < 2; p++) {
    u64 v[8], w[8];
    load3(v, y + (size_t)(p * K + limb) * N, t);
    load3(w, x + (size_t)(p * K + limb) * N, t);
#pragma unroll
    for (int c = 0; c < 8; c++) v[c] = submod(v[c], w[c], L.q);
    ntt_inv<N>(v, lds[limb], t, L);
    X.put(v, lds, limb, t, S, false);
    ep_poly<C>(X, p, A.l, A.b, g, limb, t, L, lds[limb], acc0, acc1);
    if constexpr (K == 2) __syncthreads();  // xs is rewritten for the next polynomial
  }
  u64 e0[8], e1[8];
  mac_finish(acc0, e0, L);
  mac_finish(acc1, e1, L);
  u64 x0[8], x1[8];
  load3(x0, x + (size_t)(0 * K + limb) * N, t);
  load3(x1, x + (size_t)(1 * K + limb) * N, t);
#pragma unroll
  for (int c = 0; c < 8; c++) {
    x0[c] = addmod(x0[c], e0[c], L.q);
    x1[c] = addmod(x1[c], e1[c], L.q);
  }
  store3(dst + (size_t)(0 * K + limb) * N, x0, t);
  store3(dst + (size_t)(1 * K + limb) * N, x1, t);
}

// INTT (when ntt_in), then the mod switch per coefficient. grid (2 polys, B).
template <typename C>
__global__ __launch_bounds__(C::WG) __attribute__((amdgpu_waves_per_eu(C::Waves))) void k2_final(const u64 *res, u64 *out, size_t res_stride, int ntt_in,
                                                  Scheme64 S) {
  constexpr int N = C::N, K = C::K;
  __shared__ u64 lds[K][C::Lds];
  const int limb = lane_limb<C>(), t = threadIdx.x % C::LB;
  const Limb64 &L = S.limb[limb];
  const int p = blockIdx.x, qi = blockIdx.y;
  const u64 *src = res + qi * res_stride + (size_t)(p * K + limb) * N;
  u64 v[8];
  if (ntt_in) {
    load3(v, src, t);
    ntt_inv<N>(v, lds[limb], t, L);
  } else {
    load0<N>(v, src, t);
  }
  __syncthreads();
  store0<N>(lds[limb], v, t);
  __syncthreads();
  u64 *o = out + ((size_t)qi * 2 + p) * N;
  for (int j = threadIdx.x; j < N; j += C::WG) {
    if constexpr (K == 2)
      o[j] = final_coeff(lds[0][j], lds[1][j], S);
    else
      o[j] = rescale_q0(lds[0][j], S);  // server.cpp:1034-1043
  }
}

template <typename C>
__global__ __launch_bounds__(C::WG) __attribute__((amdgpu_waves_per_eu(C::Waves))) void k2_db_load(const u64 *pts, u64 *db, size_t num_pt, size_t first,
                                                    Scheme64 S) {
  constexpr int N = C::N, K = C::K;
  __shared__ u64 lds[K][C::Lds];
  const int limb = lane_limb<C>(), t = threadIdx.x % C::LB;
  const Limb64 &L = S.limb[limb];
  const size_t i = blockIdx.x;
  u64 v[8];
  load0<N>(v, pts + i * N, t);  // < t < q: already reduced
  ntt_fwd<N>(v, lds[limb], t, L);
  store3(db + ((size_t)limb * num_pt + first + i) * N, v, t);
}

template <typename C>
__global__ __launch_bounds__(C::WG) __attribute__((amdgpu_waves_per_eu(C::Waves))) void k2_ext_prod(const u64 *gsw, int l, int b, const u64 *ct, u64 *out,
                                                     Scheme64 S) {
  constexpr int N = C::N, K = C::K;
  __shared__ u128 xs[K == 2 ? N : 1];
  __shared__ u64 lds[K][C::Lds];
  const int limb = lane_limb<C>(), t = threadIdx.x % C::LB;
  const Limb64 &L = S.limb[limb];
  u128 s0[8], s1[8];
  zero(s0, s1);
  Coef<C> X;
  if constexpr (K == 2) X.xs = xs;
#pragma unroll 1
  for (int p = 0; p < 2; p++) {
    u64 v[8];
    load0<N>(v, ct + (size_t)(p * K + limb) * N, t);
    X.put(v, lds, limb, t, S, false);
    ep_poly<C>(X, p, l, b, gsw, limb, t, L, lds[limb], s0, s1);
    if constexpr (K == 2) __syncthreads();  // xs is rewritten for the next polynomial
  }
  u64 acc0[8], acc1[8];
  mac_finish(s0, acc0, L);
  mac_finish(s1, acc1, L);
  store3(out + (size_t)(0 * K + limb) * N, acc0, t);
  store3(out + (size_t)(1 * K + limb) * N, acc1, t);
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

template <typename C>
__global__ __launch_bounds__(C::WG) __attribute__((amdgpu_waves_per_eu(C::Waves))) void k2_mux_part(MuxArgs64 A, u64 *__restrict__ part, int G, Scheme64 S) {
  constexpr int N = C::N, K = C::K;
  __shared__ u128 xs[K == 2 ? N : 1];
  __shared__ u64 lds[K][C::Lds];
  const int limb = lane_limb<C>(), t = threadIdx.x % C::LB;
  const Limb64 &L = S.limb[limb];
  const int qi = blockIdx.y;
  const int g = blockIdx.x % G, p = (blockIdx.x / G) & 1, m = blockIdx.x / (2 * G);
  int xi, yi;
  mux_operands64(A, m, xi, yi);
  const u64 *in = A.in + qi * A.stride;
  const u64 *x = in + (size_t)xi * C::Ct, *y = in + (size_t)yi * C::Ct;
  u64 v[8], w[8];
  load3(v, y + (size_t)(p * K + limb) * N, t);
  load3(w, x + (size_t)(p * K + limb) * N, t);
#pragma unroll
  for (int c = 0; c < 8; c++) v[c] = submod(v[c], w[c], L.q);
  ntt_inv<N>(v, lds[limb], t, L);
  Coef<C> X;
  if constexpr (K == 2) X.xs = xs;
  X.put(v, lds, limb, t, S, false);
  u128 s0[8], s1[8];
  zero(s0, s1);
  const int dg = (A.l + G - 1) / G, i0 = g * dg, i1 = min(A.l, i0 + dg);
  ep_poly<C>(X, p, A.l, A.b, A.sel + qi * A.sel_stride, limb, t, L, lds[limb], s0, s1, i0, i1);
  u64 acc0[8], acc1[8];
  mac_finish(s0, acc0, L);
  mac_finish(s1, acc1, L);
  u64 *dst = part + ((size_t)qi * gridDim.x + blockIdx.x) * C::Ct;
  store3(dst + (size_t)(0 * K + limb) * N, acc0, t);
  store3(dst + (size_t)(1 * K + limb) * N, acc1, t);
}

// out[s] = in[s] for ragged slots s < offset, else x + the node's 2G partials.
// kFinWgs workgroups per slot, one 16-byte word per thread.
template <typename C>
constexpr int kFinWgs = (int)(C::Ct / 2 / 256);
template <typename C>
__global__ __launch_bounds__(256) void k2_mux_fin(MuxArgs64 A, const u64 *__restrict__ part, int G, int nmux,
                                                  Scheme64 S) {
  constexpr int N = C::N, K = C::K;
  constexpr int kWords = (int)C::Ct / 2;  // 16-byte words per ciphertext
  const int s = blockIdx.x / kFinWgs<C>, qi = blockIdx.y;
  const int e = (blockIdx.x % kFinWgs<C>) * 256 + threadIdx.x;
  const u64 *in = A.in + qi * A.stride;
  ulonglong2 *dst = reinterpret_cast<ulonglong2 *>(A.out + qi * A.stride + (size_t)s * C::Ct);
  if (A.offset >= 0 && s < A.offset) {
    dst[e] = reinterpret_cast<const ulonglong2 *>(in + (size_t)s * C::Ct)[e];
    return;
  }
  const int m = A.offset >= 0 ? s - A.offset : s;
  int xi, yi;
  mux_operands64(A, m, xi, yi);
  const ulonglong2 a = reinterpret_cast<const ulonglong2 *>(in + (size_t)xi * C::Ct)[e];
  const ulonglong2 *pp =
      reinterpret_cast<const ulonglong2 *>(part + ((size_t)qi * nmux * 2 * G + (size_t)m * 2 * G) * C::Ct);
  constexpr int kMaxParts = 16;
  ulonglong2 b[kMaxParts];
#pragma unroll
  for (int j = 0; j < kMaxParts; j++)
    if (j < 2 * G) b[j] = pp[(size_t)j * kWords + e];
  const Limb64 &L = S.limb[((2 * e) / N) % K];  // polynomial index = output * K + limb
  u64 s0 = a.x, s1 = a.y;
#pragma unroll
  for (int j = 0; j < kMaxParts; j++)
    if (j < 2 * G) {
      s0 = addmod(s0, b[j].x, L.q);
      s1 = addmod(s1, b[j].y, L.q);
    }
  dst[e] = make_ulonglong2(s0, s1);
}

template <typename C>
__global__ __launch_bounds__(C::WG) __attribute__((amdgpu_waves_per_eu(C::Waves))) void k2_expand_part(ExpandArgs64 A, u64 *__restrict__ part, int G, Scheme64 S) {
  constexpr int N = C::N, K = C::K;
  __shared__ u128 xs[K == 2 ? N : 1];
  __shared__ u64 lds[K][C::Lds];
  const int limb = lane_limb<C>(), t = threadIdx.x % C::LB;
  const Limb64 &L = S.limb[limb];
  const int node = blockIdx.x / G, g = blockIdx.x % G, qi = blockIdx.y;
  const u64 *c = A.in + qi * A.stride + (size_t)node * C::Ct;
  const u64 *c1 = c + (size_t)(K + limb) * N;
  u64 v[8];
#pragma unroll
  for (int cc = 0; cc < 8; cc++) v[cc] = autom<N>(c1, map0<N>(t, cc), A.kinv, L.q);
  Coef<C> X;
  if constexpr (K == 2) X.xs = xs;
  X.put(v, lds, limb, t, S, true);
  u128 sb[8], sa[8];
  zero(sb, sa);
  const int dg = (A.l_ks + G - 1) / G, d0 = g * dg, d1 = min(A.l_ks, d0 + dg);
  bv_mac<C>(X, A.bv[qi] + (size_t)A.level * A.key_level_values, d0, d1, A.b_ks, limb, t, L, lds[limb], sb, sa);
  u64 accb[8], acca[8];
  mac_finish(sb, accb, L);
  mac_finish(sa, acca, L);
  u64 *dst = part + ((size_t)qi * gridDim.x + blockIdx.x) * C::Ct;
  store3(dst + (size_t)(0 * K + limb) * N, accb, t);
  store3(dst + (size_t)(1 * K + limb) * N, acca, t);
}

// Workgroup (node, h), lane block = limb: delta_h = INTT(sum of the G
// partials), then child polynomials h as in k2_expand_level.
template <typename C>
__global__ __launch_bounds__(C::WG) __attribute__((amdgpu_waves_per_eu(C::Waves))) void k2_expand_fin(ExpandArgs64 A, const u64 *__restrict__ part, int G,
                                                       int nnodes, Scheme64 S) {
  constexpr int N = C::N, K = C::K;
  __shared__ u64 lds[K][C::Lds];
  const int limb = lane_limb<C>(), t = threadIdx.x % C::LB;
  const Limb64 &L = S.limb[limb];
  const int node = blockIdx.x >> 1, h = blockIdx.x & 1, qi = blockIdx.y;
  const u64 *c = A.in + qi * A.stride + (size_t)node * C::Ct;
  const u64 *pp = part + ((size_t)qi * nnodes * G + (size_t)node * G) * C::Ct + (size_t)(h * K + limb) * N;
  u64 w[8];
  load3(w, pp, t);
  for (int j = 1; j < G; j++) {
    u64 u[8];
    load3(u, pp + (size_t)j * C::Ct, t);
#pragma unroll
    for (int cc = 0; cc < 8; cc++) w[cc] = addmod(w[cc], u[cc], L.q);
  }
  ntt_inv<N>(w, lds[limb], t, L);  // delta_h, coefficient form (map0)
  u64 *left = A.out + qi * A.stride + (size_t)(2 * node) * C::Ct;
  u64 *right = left + C::Ct;
  const int k = 1 << A.level;
  const u64 *cp = c + (size_t)(h * K + limb) * N;
  u64 cv[8], dv[8];
  load0<N>(cv, cp, t);
#pragma unroll
  for (int cc = 0; cc < 8; cc++) {
    u64 ww = w[cc];
    if (h == 0) ww = addmod(ww, autom<N>(cp, map0<N>(t, cc), A.kinv, L.q), L.q);  // c'_0 = sigma(c0) + delta_b
    left[(size_t)(h * K + limb) * N + map0<N>(t, cc)] = addmod(cv[cc], ww, L.q);
    dv[cc] = submod(cv[cc], ww, L.q);
  }
  __syncthreads();  // the INTT's last reads of lds are done
  store0<N>(lds[limb], dv, t);
  __syncthreads();
#pragma unroll
  for (int cc = 0; cc < 8; cc++) {
    const int j = map0<N>(t, cc), s = j + k;
    const u64 x = lds[limb][s < N ? s : s - N];
    right[(size_t)(h * K + limb) * N + j] = (s < N || !x) ? x : L.q - x;
  }
}

// Digit groups of a split level: one digit per workgroup, at most 8 groups.
static int split_groups64(int l) {
  const int dg = (l + 7) / 8;
  return (l + dg - 1) / dg;
}

// ---------------------------------------------------------------------------
// Launchers: the instantiation follows the scheme's (N, K).
// ---------------------------------------------------------------------------
template <typename F>
static void with_cfg(const Scheme64 &S, F &&f) {
  if (S.K == 2)
    f(CfgK2{});
  else
    f(CfgK1{});
}

void launch_k2_ntt(u64 *data, size_t count, bool inverse, const Scheme64 &S, hipStream_t st) {
  if (!count) return;
  with_cfg(S, [&](auto c) {
    using C = decltype(c);
    k2_ntt<C><<<(unsigned)count, C::WG, 0, st>>>(data, inverse ? 1 : 0, S);
  });
}

void launch_k2_bv_apply(const u64 *in, u64 *out, const u64 *const *key_ptr, int kinv, const Gadget64 &G,
                        const Scheme64 &S, hipStream_t st) {
  ExpandArgs64 A;
  A.in = in; A.out = out; A.stride = 0; A.bv = key_ptr; A.key_level_values = 0;
  A.level = 0; A.l_ks = G.l_ks; A.b_ks = G.b_ks; A.kinv = kinv;
  A.bv_only = 1;
  with_cfg(S, [&](auto c) {
    using C = decltype(c);
    k2_expand_level<C><<<dim3(1, 1), C::WG, 0, st>>>(A, S);
  });
}

void launch_k2_expand_level(const u64 *in, u64 *out, size_t stride, int nnodes, int B, int level,
                            const KeyPtrs64 &P, size_t key_level_values, const Gadget64 &G, const Scheme64 &S,
                            hipStream_t st, u64 *part) {
  if (nnodes <= 0 || B <= 0) return;
  ExpandArgs64 A;
  A.in = in; A.out = out; A.stride = stride; A.bv = P.bv; A.key_level_values = key_level_values;
  A.level = level; A.l_ks = G.l_ks; A.b_ks = G.b_ks;
  const unsigned gk = (unsigned)(S.N >> level) + 1;
  unsigned inv = 1;  // gk^-1 mod 2N = gk^(N-1): the unit group mod 2N (a power of two) has order N
  for (int e = 0; e < S.N - 1; e++) inv = (inv * gk) & (unsigned)(2 * S.N - 1);
  A.kinv = (int)inv;
  const int Gs = split_groups64(G.l_ks);
  with_cfg(S, [&](auto c) {
    using C = decltype(c);
    if (part && (size_t)nnodes * Gs * B <= kSplitTargetBlocks64) {
      k2_expand_part<C><<<dim3(nnodes * Gs, B), C::WG, 0, st>>>(A, part, Gs, S);
      k2_expand_fin<C><<<dim3(nnodes * 2, B), C::WG, 0, st>>>(A, part, Gs, nnodes, S);
      return;
    }
    k2_expand_level<C><<<dim3(nnodes, B), C::WG, 0, st>>>(A, S);
  });
}

void launch_k2_complete(const u64 *rows, u64 *sel, size_t row_stride, size_t sel_stride, int nrows, int B,
                        const KeyPtrs64 &P, const Gadget64 &G, const Scheme64 &S, hipStream_t st) {
  if (nrows <= 0 || B <= 0) return;
  CompleteArgs64 A;
  A.rows = rows; A.sel = sel; A.row_stride = row_stride; A.sel_stride = sel_stride; A.gsw = P.gsw;
  A.l_ep = G.l_ep; A.l_key = G.l_key; A.b_key = G.b_key;
  with_cfg(S, [&](auto c) {
    using C = decltype(c);
    k2_complete<C><<<dim3(nrows, B), C::WG, 0, st>>>(A, S);
  });
}

void launch_k2_ct_ntt(const u64 *in, u64 *out, size_t in_stride, size_t out_stride, int count, int B,
                      const Scheme64 &S, hipStream_t st) {
  if (count <= 0 || B <= 0) return;
  with_cfg(S, [&](auto c) {
    using C = decltype(c);
    k2_ct_ntt<C><<<dim3(2 * count, B), C::WG, 0, st>>>(in, out, in_stride, out_stride, S);
  });
}

void launch_k2_mux(const u64 *in, u64 *out, size_t stride, int nslots, int offset, int half, int B, const u64 *sel,
                   size_t sel_stride, const Gadget64 &G, const Scheme64 &S, hipStream_t st, u64 *part) {
  if (nslots <= 0 || B <= 0) return;
  MuxArgs64 A;
  A.in = in; A.out = out; A.stride = stride; A.sel = sel; A.sel_stride = sel_stride;
  A.offset = offset; A.half = half; A.l = G.l_ep; A.b = G.b_ep;
  const int nmux = offset >= 0 ? nslots - offset : nslots;
  const size_t pairs = (size_t)nmux * 2 * B;
  with_cfg(S, [&](auto c) {
    using C = decltype(c);
    if (part && nmux > 0 && pairs <= kSplitTargetBlocks64) {
      int Gs = (int)std::min<size_t>(split_groups64(G.l_ep), kSplitTargetBlocks64 / pairs);
      const int dg = (G.l_ep + Gs - 1) / Gs;
      Gs = (G.l_ep + dg - 1) / dg;  // no empty groups
      k2_mux_part<C><<<dim3(nmux * 2 * Gs, B), C::WG, 0, st>>>(A, part, Gs, S);
      k2_mux_fin<C><<<dim3(nslots * kFinWgs<C>, B), 256, 0, st>>>(A, part, Gs, nmux, S);
      return;
    }
    k2_mux<C><<<dim3(nslots, B), C::WG, 0, st>>>(A, S);
  });
}

void launch_k2_final(const u64 *res, u64 *out, size_t res_stride, int B, const Scheme64 &S, hipStream_t st) {
  if (B <= 0) return;
  with_cfg(S, [&](auto c) {
    using C = decltype(c);
    k2_final<C><<<dim3(2, B), C::WG, 0, st>>>(res, out, res_stride, 1, S);
  });
}

void launch_k2_mod_switch(const u64 *ct, u64 *out, const Scheme64 &S, hipStream_t st) {
  with_cfg(S, [&](auto c) {
    using C = decltype(c);
    k2_final<C><<<dim3(2, 1), C::WG, 0, st>>>(ct, out, 0, 0, S);
  });
}

void launch_k2_db_load(const u64 *pts, u64 *db, size_t num_pt, size_t first, size_t count, const Scheme64 &S,
                       hipStream_t st) {
  if (!count) return;
  with_cfg(S, [&](auto c) {
    using C = decltype(c);
    k2_db_load<C><<<(unsigned)count, C::WG, 0, st>>>(pts, db, num_pt, first, S);
  });
}

void launch_k2_ext_prod(const u64 *gsw, int l, int b, const u64 *ct, u64 *out, const Scheme64 &S, hipStream_t st) {
  with_cfg(S, [&](auto c) {
    using C = decltype(c);
    k2_ext_prod<C><<<1, C::WG, 0, st>>>(gsw, l, b, ct, out, S);
  });
}

}  // namespace k2
}  // namespace opir
