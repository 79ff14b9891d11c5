// Query expansion, selector completion and the MUX tree of the answer path
// (src/server.cpp:599-855, src/gsw.cpp:119-182,317-357,
// src/bv_keyswitch.cpp:296-367).
//
// Each tree node (BV keyswitch, completion row, MUX) is one 1024-thread
// workgroup of four 256-thread lane blocks. A lane block runs 2048-point
// transforms of one RNS limb; the four blocks split a node's gadget digits so
// that its sequential chain of transforms is half as long as with one block
// per limb. Partial digit sums of two blocks are combined through LDS.
//
// Digits: the reference extracts balanced digits LSB-first with
//   r = sign-extended low b bits of d,  d <- (d - r) >> b
// (gsw.cpp:285-293, bv_keyswitch.cpp:32-39). Unrolled, digit i of the centred
// value D is r_i = sext_b(floor((D + H_i) / 2^(b*i))) with
// H_i = 2^(b-1) * (1 + 2^b + ... + 2^(b(i-1))), so any block can compute any
// digit directly.
#include "opir_kernels.h"

namespace opir {

namespace {

struct Barrier {
  __device__ void operator()() const { __syncthreads(); }
};

// Digit groups of a split level: one digit per workgroup, at most
// kSplitMaxG groups (then ceil(l / kSplitMaxG) digits each).
constexpr int kSplitMaxG = 8;
constexpr size_t kSplitTargetBlocks = 512;
inline int split_groups(int l) {
  const int dg = (l + kSplitMaxG - 1) / kSplitMaxG;
  return (l + dg - 1) / dg;
}

// Digit i of the base-2^b expansion of D whose digits lie in
// [delta - 2^(b-1), delta + 2^(b-1)) (next_digit's rule), as a residue mod q:
// sext_b(floor((D + H_i) / 2^(b*i)) - delta) + delta with
// H_i = (2^(b-1) - delta) * sum_{j<i} 2^(b*j).
__device__ __forceinline__ i64 digit_offset(int i, int b, i64 delta) {
  i64 H = 0;
  for (int j = 0; j < i; j++) H += (((i64)1 << (b - 1)) - delta) << (b * j);
  return H;
}
// The digit r (|r| < q for every gadget base) as r + q in (0, 2q): the
// forward transform it goes into takes inputs below 2q.
__device__ __forceinline__ u32 digit_residue(i64 D, i64 H, int shift, int b, u32 q, int delta) {
  const int r = __builtin_amdgcn_sbfe((int)(u32)((D + H) >> shift) - delta, 0, b) + delta;
  return (u32)r + q;
}

// Per-gadget digit rule: the composite scheme centres and takes balanced
// digits (delta 0) everywhere; the MP scheme's BV gadget centres with ties to
// +B/2 (delta 1), its data and key gadgets take unsigned digits of [0, Q)
// (delta B/2).
template <bool MP>
__device__ __forceinline__ int ks_delta() { return MP ? 1 : 0; }
template <bool MP>
__device__ __forceinline__ int sel_delta(int b) { return MP ? 1 << (b - 1) : 0; }

__device__ __forceinline__ void compose_centre8(const u32 *lo, const u32 *hi, int t, const SchemeConst &S,
                                                i64 (&D)[8], bool centred = true) {
#pragma unroll
  for (int c = 0; c < 8; c++) {
    const int j = map0(t, c);
    D[c] = gadget_value(lo[j], hi[j], S, centred);
  }
}

// Digits [i0, i1) of one polynomial: NTT each digit (this block's limb) and
// accumulate digit x key row into two u64 accumulators per value.
// row(i) = row_base + (rev ? l-1-i : i); key rows are [2 outputs][2 limbs][N].
// kKeyFirst: issue the key row's loads before the digit's NTT (16 more live
// registers), so that their latency hides behind it.
template <bool kKeyFirst = false>
__device__ __forceinline__ void digit_mac(const i64 (&D)[8], int i0, int i1, int l, int b, bool rev, int row_base,
                                          const u32 *__restrict__ key, int limb, int t_in, const LimbConst &L,
                                          u32 *nttlds, u64 (&acc0)[8], u64 (&acc1)[8], int delta) {
  Barrier bar;
#pragma unroll 1
  for (int i = i0; i < i1; i++) {
    // An opaque copy of t keeps the per-thread twiddle and key addresses
    // inside the loop; hoisted, they do not fit in 128 registers.
    int t = t_in;
    asm volatile("" : "+v"(t));
    const i64 H = digit_offset(i, b, delta);
    // key row: [2 outputs][2 limbs][N] words; byte offsets fit in 32 bits
    const unsigned row = (unsigned)(row_base + (rev ? l - 1 - i : i));
    const unsigned off = row * (unsigned)(kCtWords * 4) + (unsigned)limb * (kN * 4) + ((unsigned)t << 5);
    uint4 a0, a1, b0, b1;
    if (kKeyFirst) {
      a0 = ld_off<uint4>(key, off), a1 = ld_off<uint4>(key, off + 16);
      b0 = ld_off<uint4>(key, off + 2 * kN * 4), b1 = ld_off<uint4>(key, off + 2 * kN * 4 + 16);
    }
    u32 v[8];
#pragma unroll
    for (int c = 0; c < 8; c++) v[c] = digit_residue(D[c], H, b * i, b, L.q, delta);
    ntt_fwd_2048<false>(v, nttlds, t, L, bar);  // [0, 4q): at most 10 products of < 2^60 per sum
    if (!kKeyFirst) {
      a0 = ld_off<uint4>(key, off), a1 = ld_off<uint4>(key, off + 16);
      b0 = ld_off<uint4>(key, off + 2 * kN * 4), b1 = ld_off<uint4>(key, off + 2 * kN * 4 + 16);
    }
    const u32 k0[8] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
    const u32 k1[8] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
#pragma unroll
    for (int c = 0; c < 8; c++) {
      acc0[c] += (u64)v[c] * k0[c];
      acc1[c] += (u64)v[c] * k1[c];
    }
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// Expansion level (server.cpp:786-810 with bv_keyswitch.cpp:296-367):
//   c' = (sigma(c0) + sum d_i*b_i, sum d_i*a_i), d_i digits of sigma(c1)
//   out[2r] = c + c',  out[2r+1] = (c - c') * X^-k
// Wide levels: one 512-thread workgroup per node (lane block = limb), all
// l_ks digits, then delta_b and delta_a in turn; two nodes share a CU.
// ---------------------------------------------------------------------------
template <bool MP>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) void k_expand_serial(ExpandArgs A,
                                                                                               SchemeConst S) {
  __shared__ __align__(16) u32 nlds[2][kLdsPoly];
  __shared__ u32 cs[2][kN];  // c1 per limb, later shift staging
  // sigma(c1) per limb, in lane block 0's NTT scratch (free until the digits)
  u32(*sig)[kN] = reinterpret_cast<u32(*)[kN]>(&nlds[0][0]);
  const int limb = __builtin_amdgcn_readfirstlane(threadIdx.x >> 8), t = threadIdx.x & 255;
  const LimbConst &L = S.limb[limb];
  const int node = blockIdx.x, qi = blockIdx.y;
  const u32 *c = A.in + qi * A.in_stride + (size_t)node * kCtWords;
  {
    u32 cv[8];
    load_map0(cv, c + (1 * 2 + limb) * kN, t);
    store_map0(cs[limb], cv, t);
  }
  __syncthreads();
#pragma unroll
  for (int cc = 0; cc < 8; cc++) {
    const int j = map0(t, cc);
    const int i0 = (int)(((unsigned)j * (unsigned)A.kinv) & (2 * kN - 1));
    const u32 x = cs[limb][i0 < kN ? i0 : i0 - kN];
    sig[limb][j] = (i0 < kN || !x) ? x : L.q - x;
  }
  __syncthreads();
  u64 accb[8], acca[8];
#pragma unroll
  for (int cc = 0; cc < 8; cc++) accb[cc] = acca[cc] = 0;
  {
    i64 D[8];
    compose_centre8(sig[0], sig[1], t, S, D);
    __syncthreads();  // sig is read; the digit transforms reuse its LDS
    const u32 *key = A.bv_keys[qi] + (size_t)A.level * A.key_level_words;
    digit_mac<true>(D, 0, A.l_ks, A.l_ks, A.b_ks, false, 0, key, limb, t, L, nlds[limb], accb, acca, ks_delta<MP>());
  }
  u32 *left = A.out + qi * A.out_stride + (size_t)(2 * node) * kCtWords;
  u32 *right = left + kCtWords;
  const int k = 1 << A.level;
#pragma unroll
  for (int h = 0; h < 2; h++) {
    u32 keep[8];
#pragma unroll
    for (int cc = 0; cc < 8; cc++) keep[cc] = reduce_u64(h ? acca[cc] : accb[cc], L);
    ntt_inv_2048(keep, nlds[limb], t, L, Barrier());  // delta_h, coefficient form
    const u32 *ch = c + (h * 2 + limb) * kN;
    u32 cv[8];
    load_map0(cv, ch, t);
    if (h == 0) {  // c'_0 = sigma(c0) + delta_b
#pragma unroll
      for (int cc = 0; cc < 8; cc++) {
        const int j = map0(t, cc);
        const int i0 = (int)(((unsigned)j * (unsigned)A.kinv) & (2 * kN - 1));
        const u32 x = ch[i0 < kN ? i0 : i0 - kN];
        keep[cc] = add_mod(keep[cc], (i0 < kN || !x) ? x : L.q - x, L.q);
      }
    }
    if (A.bv_only) {  // uniform over the workgroup: every thread skips the same barriers
      store_map0(left + (h * 2 + limb) * kN, keep, t);
      continue;
    }
    u32 l_out[8];
#pragma unroll
    for (int cc = 0; cc < 8; cc++) {
      l_out[cc] = add_mod(cv[cc], keep[cc], L.q);
      cv[cc] = sub_mod(cv[cc], keep[cc], L.q);
    }
    store_map0(left + (h * 2 + limb) * kN, l_out, t);
    // right = (c - c') * X^-k; the INTT's barriers ordered the previous reads of cs
    store_map0(cs[limb], cv, t);
    __syncthreads();
#pragma unroll
    for (int cc = 0; cc < 8; cc++) {
      const int j = map0(t, cc), s = j + k;
      const u32 x = cs[limb][s < kN ? s : s - kN];
      right[(h * 2 + limb) * kN + j] = (s < kN || !x) ? x : L.q - x;
    }
  }
}

// ---------------------------------------------------------------------------
// Split expansion level for levels with few nodes: the node's l_ks digits are
// spread over G workgroups of two lane blocks (both limbs), each writing its
// partial (delta_b, delta_a) sums (NTT form, reduced mod q) to part block
// qi * (nnodes * G) + node * G + g; k_expand_fin adds them up, runs the INTTs
// and forms both children as k_expand_level does.
// ---------------------------------------------------------------------------
template <bool MP>
__global__ __launch_bounds__(512) void k_expand_part(ExpandArgs A, u32 *__restrict__ part, int G, SchemeConst S) {
  __shared__ __align__(16) u32 nlds[2][kLdsPoly];
  __shared__ u32 c1[2][kN];  // c1 per limb
  // sigma(c1) per limb, in lane block 0's NTT scratch (free until the digits)
  u32(*sig)[kN] = reinterpret_cast<u32(*)[kN]>(&nlds[0][0]);
  const int limb = __builtin_amdgcn_readfirstlane(threadIdx.x >> 8), t = threadIdx.x & 255;
  const LimbConst &L = S.limb[limb];
  const int node = blockIdx.x / G, g = blockIdx.x % G, qi = blockIdx.y;
  const u32 *c = A.in + qi * A.in_stride + (size_t)node * kCtWords;
  {
    u32 cv[8];
    load_map0(cv, c + (1 * 2 + limb) * kN, t);
    store_map0(c1[limb], cv, t);
  }
  __syncthreads();
#pragma unroll
  for (int cc = 0; cc < 8; cc++) {
    const int j = map0(t, cc);
    const int i0 = (int)(((unsigned)j * (unsigned)A.kinv) & (2 * kN - 1));
    const u32 x = c1[limb][i0 < kN ? i0 : i0 - kN];
    sig[limb][j] = (i0 < kN || !x) ? x : L.q - x;
  }
  __syncthreads();
  i64 D[8];
  compose_centre8(sig[0], sig[1], t, S, D);
  __syncthreads();  // sig is read; the digit transforms reuse its LDS
  u64 accb[8], acca[8];
#pragma unroll
  for (int cc = 0; cc < 8; cc++) accb[cc] = acca[cc] = 0;
  const int dg = (A.l_ks + G - 1) / G, i0 = g * dg, i1 = min(A.l_ks, i0 + dg);
  const u32 *key = A.bv_keys[qi] + (size_t)A.level * A.key_level_words;
  digit_mac<true>(D, i0, i1, A.l_ks, A.b_ks, false, 0, key, limb, t, L, nlds[limb], accb, acca, ks_delta<MP>());
  u32 ob[8], oa[8];
#pragma unroll
  for (int cc = 0; cc < 8; cc++) {
    ob[cc] = reduce_u64(accb[cc], L);
    oa[cc] = reduce_u64(acca[cc], L);
  }
  u32 *dst = part + ((size_t)qi * gridDim.x + blockIdx.x) * kCtWords;
  store_map3(dst + (0 * 2 + limb) * kN, ob, t);
  store_map3(dst + (1 * 2 + limb) * kN, oa, t);
}

// One workgroup per (node, h, limb): delta_h = INTT(sum of the G partials),
// then child polynomials h of this limb as in k_expand_level.
__global__ __launch_bounds__(256) void k_expand_fin(ExpandArgs A, const u32 *__restrict__ part, int G, int nnodes,
                                                    SchemeConst S) {
  __shared__ __align__(16) u32 nlds[kLdsPoly];
  __shared__ u32 sh[kN];
  const int t = threadIdx.x;
  const int limb = blockIdx.x & 1, h = (blockIdx.x >> 1) & 1;
  const LimbConst &L = S.limb[limb];
  const int node = blockIdx.x >> 2, qi = blockIdx.y;
  const u32 *c = A.in + qi * A.in_stride + (size_t)node * kCtWords;
  const u32 *pp = part + ((size_t)qi * nnodes * G + (size_t)node * G) * kCtWords + (h * 2 + limb) * kN;
  u32 keep[8];
  {
    uint4 w[kSplitMaxG][2];  // all G partials in flight at once
#pragma unroll
    for (int j = 0; j < kSplitMaxG; j++)
      if (j < G) {
        w[j][0] = reinterpret_cast<const uint4 *>(pp + (size_t)j * kCtWords)[2 * t];
        w[j][1] = reinterpret_cast<const uint4 *>(pp + (size_t)j * kCtWords)[2 * t + 1];
      }
    u64 acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int j = 0; j < kSplitMaxG; j++)
      if (j < G) {
        acc[0] += w[j][0].x; acc[1] += w[j][0].y; acc[2] += w[j][0].z; acc[3] += w[j][0].w;
        acc[4] += w[j][1].x; acc[5] += w[j][1].y; acc[6] += w[j][1].z; acc[7] += w[j][1].w;
      }
#pragma unroll
    for (int cc = 0; cc < 8; cc++) keep[cc] = reduce_u64(acc[cc], L);
  }
  // c (poly h of this limb) and, for h = 0, sigma(c0), loaded ahead of the INTT
  const u32 *ch = c + (h * 2 + limb) * kN;
  u32 cv[8], sg[8];
  load_map0(cv, ch, t);
  if (h == 0) {
#pragma unroll
    for (int cc = 0; cc < 8; cc++) {
      const int j = map0(t, cc);
      const int i0 = (int)(((unsigned)j * (unsigned)A.kinv) & (2 * kN - 1));
      const u32 x = ch[i0 < kN ? i0 : i0 - kN];
      sg[cc] = (i0 < kN || !x) ? x : L.q - x;
    }
  }
  ntt_inv_2048(keep, nlds, t, L, Barrier());  // delta_b (h=0) or delta_a (h=1), coefficient form
  u32 *left = A.out + qi * A.out_stride + (size_t)(2 * node) * kCtWords;
  u32 *right = left + kCtWords;
  if (h == 0) {  // c'_0 = sigma(c0) + delta_b
#pragma unroll
    for (int cc = 0; cc < 8; cc++) keep[cc] = add_mod(keep[cc], sg[cc], L.q);
  }
  u32 l_out[8];
#pragma unroll
  for (int cc = 0; cc < 8; cc++) {
    l_out[cc] = add_mod(cv[cc], keep[cc], L.q);
    cv[cc] = sub_mod(cv[cc], keep[cc], L.q);
  }
  store_map0(left + (h * 2 + limb) * kN, l_out, t);
  // right = (c - c') * X^-k: out[j] = d[j+k] if j+k < N else -d[j+k-N]
  store_map0(sh, cv, t);
  __syncthreads();
  const int k = 1 << A.level;
#pragma unroll
  for (int cc = 0; cc < 8; cc++) {
    const int j = map0(t, cc), s = j + k;
    const u32 x = sh[s < kN ? s : s - kN];
    right[(h * 2 + limb) * kN + j] = (s < kN || !x) ? x : L.q - x;
  }
}

// ---------------------------------------------------------------------------
// Selector completion (server.cpp:823-855, gsw.cpp:317-357): top row r =
// NTT(row); bottom row = external_product(RGSW(s), row), key gadget.
// Lane block (limb, p): poly p of the row, all l_key digits of it.
// ---------------------------------------------------------------------------
template <bool MP>
__global__ __launch_bounds__(1024) void k_complete(CompleteArgs A, SchemeConst S) {
  __shared__ __align__(16) u32 nlds[4][kLdsPoly];
  __shared__ u32 sh[2][2][kN];
  const int lb = __builtin_amdgcn_readfirstlane(threadIdx.x >> 8), t = threadIdx.x & 255;  // lb is wave-uniform
  const int limb = lb & 1, p = lb >> 1;
  const LimbConst &L = S.limb[limb];
  const int g = blockIdx.x, qi = blockIdx.y;
  const int a = g / A.l_ep, r = g % A.l_ep;
  const u32 *row = A.leaves + qi * A.leaf_stride + (size_t)g * kCtWords;
  u32 *selb = A.sel + qi * A.sel_stride + (size_t)a * (2 * A.l_ep) * kCtWords;
  u32 v[8];
  load_map0(v, row + (p * 2 + limb) * kN, t);
  store_map0(sh[p][limb], v, t);
  ntt_fwd_2048(v, nlds[lb], t, L, Barrier());  // its barriers also publish sh
  store_map3(selb + (size_t)r * kCtWords + (p * 2 + limb) * kN, v, t);
  i64 D[8];
  compose_centre8(sh[p][0], sh[p][1], t, S, D, !MP);
  u64 acc0[8], acc1[8];
#pragma unroll
  for (int c = 0; c < 8; c++) acc0[c] = acc1[c] = 0;
  digit_mac<true>(D, 0, A.l_key, A.l_key, A.b_key, true, p * A.l_key, A.gsw_keys[qi], limb, t, L, nlds[lb], acc0, acc1,
                  sel_delta<MP>(A.b_key));
  u32 o0[8], o1[8];
#pragma unroll
  for (int c = 0; c < 8; c++) {
    o0[c] = reduce_u64(acc0[c], L);
    o1[c] = reduce_u64(acc1[c], L);
  }
  __syncthreads();
  if (p == 1) {
#pragma unroll
    for (int c = 0; c < 8; c++) {
      sh[0][limb][map3(t, c)] = o0[c];
      sh[1][limb][map3(t, c)] = o1[c];
    }
  }
  __syncthreads();
  if (p == 0) {
#pragma unroll
    for (int c = 0; c < 8; c++) {
      o0[c] = add_mod(o0[c], sh[0][limb][map3(t, c)], L.q);
      o1[c] = add_mod(o1[c], sh[1][limb][map3(t, c)], L.q);
    }
    u32 *bot = selb + (size_t)(A.l_ep + r) * kCtWords;
    store_map3(bot + (0 * 2 + limb) * kN, o0, t);
    store_map3(bot + (1 * 2 + limb) * kN, o1, t);
  }
}

// The same row with one 512-thread workgroup (lane block = limb) that handles
// poly p = 0 and then p = 1. The sum is reduced between the two halves: each
// half adds l_key products of < 2^60.
template <bool MP>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) void k_complete_serial(CompleteArgs A,
                                                                                                 SchemeConst S) {
  __shared__ __align__(16) u32 nlds[2][kLdsPoly];
  __shared__ u32 sh[2][kN];
  const int limb = __builtin_amdgcn_readfirstlane(threadIdx.x >> 8), t = threadIdx.x & 255;
  const LimbConst &L = S.limb[limb];
  const int g = blockIdx.x, qi = blockIdx.y;
  const int a = g / A.l_ep, r = g % A.l_ep;
  const u32 *row = A.leaves + qi * A.leaf_stride + (size_t)g * kCtWords;
  u32 *selb = A.sel + qi * A.sel_stride + (size_t)a * (2 * A.l_ep) * kCtWords;
  u64 acc0[8], acc1[8];
#pragma unroll
  for (int c = 0; c < 8; c++) acc0[c] = acc1[c] = 0;
#pragma unroll
  for (int p = 0; p < 2; p++) {
    u32 v[8];
    load_map0(v, row + (p * 2 + limb) * kN, t);
    store_map0(sh[limb], v, t);  // the previous half's digit transforms ordered its reads of sh
    ntt_fwd_2048(v, nlds[limb], t, L, Barrier());  // its barriers also publish sh
    store_map3(selb + (size_t)r * kCtWords + (p * 2 + limb) * kN, v, t);
    i64 D[8];
    compose_centre8(sh[0], sh[1], t, S, D, !MP);
    if (p == 1) {
#pragma unroll
      for (int c = 0; c < 8; c++) {
        acc0[c] = reduce_u64(acc0[c], L);
        acc1[c] = reduce_u64(acc1[c], L);
      }
    }
    digit_mac<true>(D, 0, A.l_key, A.l_key, A.b_key, true, p * A.l_key, A.gsw_keys[qi], limb, t, L, nlds[limb], acc0,
                    acc1, sel_delta<MP>(A.b_key));
  }
  u32 o0[8], o1[8];
#pragma unroll
  for (int c = 0; c < 8; c++) {
    o0[c] = reduce_u64(acc0[c], L);
    o1[c] = reduce_u64(acc1[c], L);
  }
  u32 *bot = selb + (size_t)(A.l_ep + r) * kCtWords;
  store_map3(bot + (0 * 2 + limb) * kN, o0, t);
  store_map3(bot + (1 * 2 + limb) * kN, o1, t);
}

// Split completion for small batches, whose completion runs ahead of the
// scan with its latency on the answer's path: workgroup (row, poly p, digit
// group j) of two lane blocks (limbs) runs digits [j*dg, (j+1)*dg) of poly
// p and writes its partial bottom row (both outputs, both limbs, reduced) to
// part block qi * gridDim.x + blockIdx.x; the last group, which has the
// fewest digits, also writes the top row NTT(poly p). k_complete_fin adds
// the 2G partials of a row. Chains of ceil(l_key / G) + 1 transforms instead
// of l_key + 1.
template <bool MP>
__global__ __launch_bounds__(512) void k_complete_part(CompleteArgs A, u32 *__restrict__ part, int G, SchemeConst S) {
  __shared__ __align__(16) u32 nlds[2][kLdsPoly];
  __shared__ u32 sh[2][kN];
  const int limb = __builtin_amdgcn_readfirstlane(threadIdx.x >> 8), t = threadIdx.x & 255;
  const LimbConst &L = S.limb[limb];
  const int j = blockIdx.x % G, p = (blockIdx.x / G) & 1, g = blockIdx.x / (2 * G), qi = blockIdx.y;
  const int a = g / A.l_ep, r = g % A.l_ep;
  const u32 *row = A.leaves + qi * A.leaf_stride + (size_t)g * kCtWords;
  u32 v[8];
  load_map0(v, row + (p * 2 + limb) * kN, t);
  store_map0(sh[limb], v, t);
  __syncthreads();
  i64 D[8];
  compose_centre8(sh[0], sh[1], t, S, D, !MP);
  u64 acc0[8], acc1[8];
#pragma unroll
  for (int c = 0; c < 8; c++) acc0[c] = acc1[c] = 0;
  const int dg = (A.l_key + G - 1) / G, i0 = j * dg, i1 = min(A.l_key, i0 + dg);
  digit_mac<true>(D, i0, i1, A.l_key, A.b_key, true, p * A.l_key, A.gsw_keys[qi], limb, t, L, nlds[limb], acc0, acc1,
                  sel_delta<MP>(A.b_key));
  if (j == G - 1) {  // top row: NTT of poly p, on the same scratch rotation
    ntt_fwd_2048(v, nlds[limb], t, L, Barrier());
    u32 *selb = A.sel + qi * A.sel_stride + (size_t)a * (2 * A.l_ep) * kCtWords;
    store_map3(selb + (size_t)r * kCtWords + (p * 2 + limb) * kN, v, t);
  }
  u32 o0[8], o1[8];
#pragma unroll
  for (int c = 0; c < 8; c++) {
    o0[c] = reduce_u64(acc0[c], L);
    o1[c] = reduce_u64(acc1[c], L);
  }
  u32 *dst = part + ((size_t)qi * gridDim.x + blockIdx.x) * kCtWords;
  store_map3(dst + (0 * 2 + limb) * kN, o0, t);
  store_map3(dst + (1 * 2 + limb) * kN, o1, t);
}

// Bottom row r of a split completion: the sum of its 2G partials (blocks
// [g * 2G, g * 2G + 2G) of the query). Eight workgroups per row, one 16-byte
// word per thread, all partials in flight at once (G <= kSplitMaxG).
__global__ __launch_bounds__(256) void k_complete_fin(CompleteArgs A, const u32 *__restrict__ part, int G, int nrows,
                                                      SchemeConst S) {
  const int g = blockIdx.x >> 3, qi = blockIdx.y;
  const int e = (blockIdx.x & 7) * 256 + threadIdx.x;  // 16-byte word of the ciphertext, < kCtWords / 4
  const int a = g / A.l_ep, r = g % A.l_ep;
  const uint4 *pp =
      reinterpret_cast<const uint4 *>(part + ((size_t)qi * nrows * 2 * G + (size_t)g * 2 * G) * kCtWords);
  uint4 b[2 * kSplitMaxG];
#pragma unroll
  for (int k = 0; k < 2 * kSplitMaxG; k++)
    if (k < 2 * G) b[k] = pp[(size_t)k * (kCtWords / 4) + e];
  u64 s0 = 0, s1 = 0, s2 = 0, s3 = 0;
#pragma unroll
  for (int k = 0; k < 2 * kSplitMaxG; k++)
    if (k < 2 * G) {
      s0 += b[k].x; s1 += b[k].y; s2 += b[k].z; s3 += b[k].w;
    }
  const LimbConst &L = S.limb[((4 * e) / kN) & 1];  // polynomial index = output * 2 + limb
  uint4 *dst = reinterpret_cast<uint4 *>(A.sel + qi * A.sel_stride + ((size_t)a * 2 * A.l_ep + A.l_ep + r) * kCtWords);
  dst[e] = make_uint4(reduce_u64(s0, L), reduce_u64(s1, L), reduce_u64(s2, L), reduce_u64(s3, L));
}

// ---------------------------------------------------------------------------
// MUX level (server.cpp:599-722) on NTT-form candidates:
//   res = x + EP(sel, INTT(y - x))   (EP output in NTT form)
// equal to the reference's coefficient-form x + INTT(EP(sel, y - x)).
// ---------------------------------------------------------------------------

__device__ __forceinline__ void mux_operands(const MuxArgs &A, int m, int &xi, int &yi) {
  if (A.offset >= 0) {  // ragged: muxed slot offset + m pairs in[offset + 2m], in[offset + 2m + 1]
    xi = A.offset + 2 * m;
    yi = xi + 1;
  } else {
    xi = m;
    yi = m + A.half;
  }
}

// Wide levels: one 512-thread workgroup per node (lane block = limb) that
// handles poly p = 0 and then p = 1 (INTT of poly p of y - x, then its l
// digits), accumulating both into one sum; two nodes share a CU and their
// barrier waits overlap.
template <bool MP>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) void k_mux_serial(MuxArgs A, SchemeConst S) {
  __shared__ __align__(16) u32 nlds[2][kLdsPoly];
  __shared__ u32 sh[2][kN];
  const int limb = __builtin_amdgcn_readfirstlane(threadIdx.x >> 8), t = threadIdx.x & 255;
  const LimbConst &L = S.limb[limb];
  const int s = blockIdx.x, qi = blockIdx.y;
  const u32 *in = A.in + qi * A.stride;
  u32 *dst = A.out + qi * A.stride + (size_t)s * kCtWords;
  if (A.offset >= 0 && s < A.offset) {  // unpaired candidate of the ragged level stays in place
    const u32 *src = in + (size_t)s * kCtWords;
    for (int e = threadIdx.x; e < (int)kCtWords / 4; e += 512)
      reinterpret_cast<uint4 *>(dst)[e] = reinterpret_cast<const uint4 *>(src)[e];
    return;
  }
  int xi, yi;
  mux_operands(A, A.offset >= 0 ? s - A.offset : s, xi, yi);
  const u32 *x = in + (size_t)xi * kCtWords, *y = in + (size_t)yi * kCtWords;
  u64 acc0[8], acc1[8];
#pragma unroll
  for (int c = 0; c < 8; c++) acc0[c] = acc1[c] = 0;
#pragma unroll
  for (int p = 0; p < 2; p++) {
    u32 v[8], w[8];
    load_map3(v, y + (p * 2 + limb) * kN, t);
    load_map3(w, x + (p * 2 + limb) * kN, t);
#pragma unroll
    for (int c = 0; c < 8; c++) v[c] = sub_mod(v[c], w[c], L.q);
    ntt_inv_2048(v, nlds[limb], t, L, Barrier());  // its first barrier also orders sh reuse
    store_map0(sh[limb], v, t);
    __syncthreads();
    i64 D[8];
    compose_centre8(sh[0], sh[1], t, S, D, !MP);
    digit_mac<true>(D, 0, A.l, A.l, A.b, true, p * A.l, A.sel + qi * A.sel_stride, limb, t, L, nlds[limb], acc0,
                    acc1, sel_delta<MP>(A.b));
  }
  u32 x0[8], x1[8];
  load_map3(x0, x + (0 * 2 + limb) * kN, t);
  load_map3(x1, x + (1 * 2 + limb) * kN, t);
#pragma unroll
  for (int c = 0; c < 8; c++) {
    x0[c] = add_mod(x0[c], reduce_u64(acc0[c], L), L.q);
    x1[c] = add_mod(x1[c], reduce_u64(acc1[c], L), L.q);
  }
  store_map3(dst + (0 * 2 + limb) * kN, x0, t);
  store_map3(dst + (1 * 2 + limb) * kN, x1, t);
}

// ---------------------------------------------------------------------------
// Split MUX for levels with few nodes (latency-bound: k_mux keeps one CU per
// node busy with a chain of six transforms per lane block). The node's work
// is spread over 2G workgroups of two lane blocks (both limbs of poly p):
// each runs INTT(y - x) of poly p and the digits [g*dg, (g+1)*dg) of it, and
// writes its partial external product (both outputs, both limbs, NTT form,
// reduced mod q) to part; k_mux_fin adds x and the 2G partials.
// part block of (query qi, muxed slot m, poly p, group g):
//   qi * (nmux * 2G) + (m * 2 + p) * G + g, one ciphertext each.
// ---------------------------------------------------------------------------

template <bool MP>
__global__ __launch_bounds__(512) void k_mux_part(MuxArgs A, u32 *__restrict__ part, int G, SchemeConst S) {
  __shared__ __align__(16) u32 nlds[2][kLdsPoly];
  __shared__ u32 sh[2][kN];
  const int limb = __builtin_amdgcn_readfirstlane(threadIdx.x >> 8), t = threadIdx.x & 255;
  const LimbConst &L = S.limb[limb];
  const int qi = blockIdx.y;
  const int g = blockIdx.x % G, p = (blockIdx.x / G) & 1, m = blockIdx.x / (2 * G);
  int xi, yi;
  mux_operands(A, m, xi, yi);
  const u32 *in = A.in + qi * A.stride;
  const u32 *x = in + (size_t)xi * kCtWords, *y = in + (size_t)yi * kCtWords;
  u32 v[8], w[8];
  load_map3(v, y + (p * 2 + limb) * kN, t);
  load_map3(w, x + (p * 2 + limb) * kN, t);
#pragma unroll
  for (int c = 0; c < 8; c++) v[c] = sub_mod(v[c], w[c], L.q);
  ntt_inv_2048(v, nlds[limb], t, L, Barrier());
  store_map0(sh[limb], v, t);
  __syncthreads();
  i64 D[8];
  compose_centre8(sh[0], sh[1], t, S, D, !MP);
  u64 acc0[8], acc1[8];
#pragma unroll
  for (int c = 0; c < 8; c++) acc0[c] = acc1[c] = 0;
  const int dg = (A.l + G - 1) / G, i0 = g * dg, i1 = min(A.l, i0 + dg);
  digit_mac<true>(D, i0, i1, A.l, A.b, true, p * A.l, A.sel + qi * A.sel_stride, limb, t, L, nlds[limb], acc0,
                  acc1, sel_delta<MP>(A.b));
  u32 o0[8], o1[8];
#pragma unroll
  for (int c = 0; c < 8; c++) {
    o0[c] = reduce_u64(acc0[c], L);
    o1[c] = reduce_u64(acc1[c], L);
  }
  u32 *dst = part + ((size_t)qi * gridDim.x + blockIdx.x) * kCtWords;
  store_map3(dst + (0 * 2 + limb) * kN, o0, t);
  store_map3(dst + (1 * 2 + limb) * kN, o1, t);
}

// out[s] = in[s] for ragged slots s < offset, else x + sum of the node's 2G
// partials. Eight workgroups per slot, one 16-byte word per thread, all
// 2G + 1 loads in flight at once (G <= kSplitMaxG).
__global__ __launch_bounds__(256) void k_mux_fin(MuxArgs A, const u32 *__restrict__ part, int G, int nmux,
                                                 SchemeConst S) {
  const int s = blockIdx.x >> 3, qi = blockIdx.y;
  const int e = (blockIdx.x & 7) * 256 + threadIdx.x;  // 16-byte word of the ciphertext, < kCtWords / 4
  const u32 *in = A.in + qi * A.stride;
  uint4 *dst = reinterpret_cast<uint4 *>(A.out + qi * A.stride + (size_t)s * kCtWords);
  if (A.offset >= 0 && s < A.offset) {
    dst[e] = reinterpret_cast<const uint4 *>(in + (size_t)s * kCtWords)[e];
    return;
  }
  const int m = A.offset >= 0 ? s - A.offset : s;
  int xi, yi;
  mux_operands(A, m, xi, yi);
  const uint4 a = reinterpret_cast<const uint4 *>(in + (size_t)xi * kCtWords)[e];
  const uint4 *pp = reinterpret_cast<const uint4 *>(part + ((size_t)qi * nmux * 2 * G + (size_t)m * 2 * G) * kCtWords);
  uint4 b[2 * kSplitMaxG];
#pragma unroll
  for (int j = 0; j < 2 * kSplitMaxG; j++)
    if (j < 2 * G) b[j] = pp[(size_t)j * (kCtWords / 4) + e];
  u64 s0 = a.x, s1 = a.y, s2 = a.z, s3 = a.w;
#pragma unroll
  for (int j = 0; j < 2 * kSplitMaxG; j++)
    if (j < 2 * G) {
      s0 += b[j].x; s1 += b[j].y; s2 += b[j].z; s3 += b[j].w;
    }
  const LimbConst &L = S.limb[((4 * e) / kN) & 1];  // polynomial index = output * 2 + limb
  dst[e] = make_uint4(reduce_u64(s0, L), reduce_u64(s1, L), reduce_u64(s2, L), reduce_u64(s3, L));
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
void launch_expand_level(const u32 *in, u32 *out, size_t in_stride, size_t out_stride, int nnodes, int B,
                         int level, const BatchPtrs &P, const GadgetCfg &G, const SchemeConst &S, hipStream_t st,
                         u32 *part) {
  ExpandArgs A;
  A.in = in; A.out = out; A.in_stride = in_stride; A.out_stride = out_stride;
  A.bv_keys = P.bv_keys;
  A.key_level_words = (size_t)G.l_ks * kCtWords;
  A.level = level;
  const unsigned gk = (unsigned)(kN >> level) + 1;
  unsigned inv = 1;  // gk^-1 mod 2N = gk^(N-1): the unit group mod 2N = 2^12 has order N
  for (int e = 0; e < kN - 1; e++) inv = (inv * gk) & (2 * kN - 1);
  A.kinv = (int)inv;
  A.l_ks = G.l_ks; A.b_ks = G.b_ks;
  const int Gs = split_groups(G.l_ks);
  if (part && (size_t)nnodes * Gs * B <= kMuxSplitMaxBlocks) {
    if (S.mp) k_expand_part<true><<<dim3(nnodes * Gs, B), 512, 0, st>>>(A, part, Gs, S);
    else k_expand_part<false><<<dim3(nnodes * Gs, B), 512, 0, st>>>(A, part, Gs, S);
    k_expand_fin<<<dim3(nnodes * 4, B), 256, 0, st>>>(A, part, Gs, nnodes, S);
    return;
  }
  dim3 grid(nnodes, B);
  if (S.mp) k_expand_serial<true><<<grid, 512, 0, st>>>(A, S);
  else k_expand_serial<false><<<grid, 512, 0, st>>>(A, S);
}

void launch_bv_apply(const u32 *in, u32 *out, const u32 *const *key_ptr, int kinv, const GadgetCfg &G,
                     const SchemeConst &S, hipStream_t st) {
  ExpandArgs A;
  A.in = in; A.out = out; A.in_stride = 0; A.out_stride = 0;
  A.bv_keys = key_ptr;
  A.key_level_words = 0;
  A.level = 0;
  A.kinv = kinv;
  A.l_ks = G.l_ks; A.b_ks = G.b_ks;
  A.bv_only = 1;
  if (S.mp) k_expand_serial<true><<<dim3(1, 1), 512, 0, st>>>(A, S);
  else k_expand_serial<false><<<dim3(1, 1), 512, 0, st>>>(A, S);
}

void launch_complete(const u32 *leaves, u32 *sel, size_t leaf_stride, size_t sel_stride, int nrows, int B,
                     const BatchPtrs &P, const GadgetCfg &G, const SchemeConst &S, hipStream_t st, u32 *part) {
  if (nrows <= 0) return;
  CompleteArgs A;
  A.leaves = leaves; A.sel = sel; A.leaf_stride = leaf_stride; A.sel_stride = sel_stride;
  A.gsw_keys = P.gsw_keys; A.l_ep = G.l_ep; A.l_key = G.l_key; A.b_key = G.b_key;
  // small batches: digit groups of about a third of l_key each
  const int Gs = std::min(kSplitMaxG, std::max(1, (G.l_key + 3) / 4));
  if (part && (size_t)nrows * B < 256 && (size_t)nrows * 2 * Gs * B <= kMuxSplitMaxBlocks) {
    if (S.mp) k_complete_part<true><<<dim3(nrows * 2 * Gs, B), 512, 0, st>>>(A, part, Gs, S);
    else k_complete_part<false><<<dim3(nrows * 2 * Gs, B), 512, 0, st>>>(A, part, Gs, S);
    k_complete_fin<<<dim3(nrows * 8, B), 256, 0, st>>>(A, part, Gs, nrows, S);
    return;
  }
  dim3 grid(nrows, B);
  if ((size_t)nrows * B >= 256)  // enough rows to fill the device twice over
    if (S.mp) k_complete_serial<true><<<grid, 512, 0, st>>>(A, S);
    else k_complete_serial<false><<<grid, 512, 0, st>>>(A, S);
  else
    if (S.mp) k_complete<true><<<grid, 1024, 0, st>>>(A, S);
    else k_complete<false><<<grid, 1024, 0, st>>>(A, S);
}

void launch_mux(const u32 *in, u32 *out, size_t stride, int nslots, int offset, int half, int B,
                const u32 *sel, size_t sel_stride, const GadgetCfg &G, const SchemeConst &S, hipStream_t st) {
  if (nslots <= 0) return;
  MuxArgs A;
  A.in = in; A.out = out; A.stride = stride; A.sel = sel; A.sel_stride = sel_stride;
  A.offset = offset; A.half = half; A.l = G.l_ep; A.b = G.b_ep;
  dim3 grid(nslots, B);
  if (S.mp) k_mux_serial<true><<<grid, 512, 0, st>>>(A, S);
  else k_mux_serial<false><<<grid, 512, 0, st>>>(A, S);
}

int mux_split_groups(int nslots, int offset, int B, const GadgetCfg &G) {
  const int nmux = offset >= 0 ? nslots - offset : nslots;
  if (nmux <= 0) return 0;
  // about kSplitTargetBlocks workgroups: one digit each for narrow levels,
  // several digits per workgroup for wider ones, no split for the widest
  const size_t pairs = (size_t)nmux * 2 * B;
  if (pairs > kSplitTargetBlocks) return 0;
  int Gs = (int)std::min<size_t>(split_groups(G.l_ep), kSplitTargetBlocks / pairs);
  const int dg = (G.l_ep + Gs - 1) / Gs;
  Gs = (G.l_ep + dg - 1) / dg;  // no empty groups
  return pairs * Gs <= kMuxSplitMaxBlocks ? Gs : 0;
}

void launch_mux_split(const u32 *in, u32 *out, size_t stride, int nslots, int offset, int half, int B,
                      const u32 *sel, size_t sel_stride, u32 *part, int Gs, const GadgetCfg &G, const SchemeConst &S,
                      hipStream_t st) {
  MuxArgs A;
  A.in = in; A.out = out; A.stride = stride; A.sel = sel; A.sel_stride = sel_stride;
  A.offset = offset; A.half = half; A.l = G.l_ep; A.b = G.b_ep;
  const int nmux = offset >= 0 ? nslots - offset : nslots;
  if (S.mp) k_mux_part<true><<<dim3(nmux * 2 * Gs, B), 512, 0, st>>>(A, part, Gs, S);
  else k_mux_part<false><<<dim3(nmux * 2 * Gs, B), 512, 0, st>>>(A, part, Gs, S);
  k_mux_fin<<<dim3(8 * nslots, B), 256, 0, st>>>(A, part, Gs, nmux, S);
}

}  // namespace opir
