// HIP kernels of the OnionPIRv2 answer path for CDNA4 (gfx950).
// Reference algorithm: src/server.cpp:731-947 (make_query and its stages),
// src/gsw.cpp:119-182,252-357 (external product, completion),
// src/bv_keyswitch.cpp:296-367 (BV keyswitch), src/matrix.cpp:232-350 (scan).
#include "opir_kernels.h"

namespace opir {

namespace {

struct Barrier {
  __device__ void operator()() const { __syncthreads(); }
};

// Gadget-decomposition MAC shared by the external product, the completion
// and the BV keyswitch: for every digit of every input polynomial, forward
// NTT the digit (limb of this lane block) and accumulate digit * key row
// into two u64 accumulators per value. Digits are produced LSB-first from
// the running centred value D, exactly like decomp_rlwe_single_mod
// (src/gsw.cpp:285-293) and signed_gadget_decompose (bv_keyswitch.cpp:32-39).
// GSW_ROWS: key row for digit i of poly p is p*l + (l-1-i) (MSB-first rows,
// gsw.cpp:286); otherwise row i (BV keys, bv_keyswitch.cpp:344-356).
// Key row layout: [2 outputs][2 limbs][N] u32, NTT form.
template <int NP, bool GSW_ROWS>
__device__ __forceinline__ void gadget_mac(i64 (&D)[NP][8], int l, int b, const u32 *__restrict__ key, int sb,
                                           int t, const LimbConst &L, u32 *nttlds, u64 (&acc0)[8],
                                           u64 (&acc1)[8], i64 delta) {
  Barrier bar;
#pragma unroll
  for (int p = 0; p < NP; p++) {
    for (int i = 0; i < l; i++) {
      u32 v[8];
#pragma unroll
      for (int c = 0; c < 8; c++) v[c] = digit_mod(next_digit(D[p][c], b, delta), L.q);
      ntt_fwd_2048(v, nttlds, t, L, bar);
      const int row = GSW_ROWS ? (p * l + (l - 1 - i)) : i;
      const u32 *r = key + (size_t)row * (4 * kN);
      u32 k0[8], k1[8];
      load_map3(k0, r + (0 * 2 + sb) * kN, t);
      load_map3(k1, r + (1 * 2 + sb) * kN, t);
#pragma unroll
      for (int c = 0; c < 8; c++) {
        acc0[c] += (u64)v[c] * k0[c];
        acc1[c] += (u64)v[c] * k1[c];
      }
    }
  }
}

// Compose both limbs of a coefficient-form polynomial held in LDS (natural
// order) at this thread's map0 indices and centre the result.
__device__ __forceinline__ void compose_centre(const u32 *lo, const u32 *hi, int t, const SchemeConst &S,
                                               i64 (&D)[8]) {
#pragma unroll
  for (int c = 0; c < 8; c++) {
    const int j = map0(t, c);
    D[c] = centre(crt_compose(lo[j], hi[j], S), S);
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// Elementwise helpers.
// ---------------------------------------------------------------------------
// Group blockIdx.y: npoly polynomials at in + y*in_stride (u64 words) ->
// out + y*out_stride (u32 words).
// Two-limb MP scheme: the u64 input already holds [npoly][2 limbs][N] residues.
__global__ void k_split_u64(const u64 *__restrict__ in, u32 *__restrict__ out, size_t npoly, size_t in_stride,
                            size_t out_stride, SchemeConst S) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npoly * kN) return;
  in += blockIdx.y * in_stride;
  out += blockIdx.y * out_stride;
  if (S.mp) {
    out[2 * i] = (u32)in[2 * i];
    out[2 * i + 1] = (u32)in[2 * i + 1];
    return;
  }
  const size_t p = i / kN, j = i % kN;
  const u64 x = in[i];
  out[(2 * p + 0) * kN + j] = reduce_u64(x, S.limb[0]);
  out[(2 * p + 1) * kN + j] = reduce_u64(x, S.limb[1]);
}

__global__ void k_join_u64(const u32 *__restrict__ in, u64 *__restrict__ out, size_t npoly, SchemeConst S) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npoly * kN) return;
  if (S.mp) {  // [npoly][2 limbs][N] residues out
    out[2 * i] = in[2 * i];
    out[2 * i + 1] = in[2 * i + 1];
    return;
  }
  const size_t p = i / kN, j = i % kN;
  out[i] = crt_compose(in[(2 * p) * kN + j], in[(2 * p + 1) * kN + j], S);
}

// Residue-wise a -/+ b, the EltwiseSubMod / EltwiseAddMod steps of the
// homomorphic MUX (src/server.cpp:657-683). Limb k of every [K][limb_len]
// group is reduced by q_k.
__global__ void k_addsub_mod(const u64 *a, const u64 *b, u64 *out, size_t n, size_t limb_len, int K, u64 q0,
                             u64 q1, int sub) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const u64 q = ((i / limb_len) % (size_t)K) ? q1 : q0;
  const u64 x = a[i], y = b[i];
  u64 r;
  if (sub) {
    r = x >= y ? x - y : x + (q - y);
  } else {
    r = x + y;
    if (r >= q) r -= q;
  }
  out[i] = r;
}

// In-place NTT of u64 (mod Q) polynomials; one workgroup (2 lane blocks,
// one per limb) per polynomial.
__global__ __launch_bounds__(512) void k_ntt_u64(u64 *data, int inverse, SchemeConst S) {
  __shared__ __align__(16) u32 nlds[2][kLdsPoly];
  __shared__ u32 res[2][kN];
  const int sb = __builtin_amdgcn_readfirstlane(threadIdx.x >> 8), t = threadIdx.x & 255;  // sb: wave-uniform
  const LimbConst &L = S.limb[sb];
  u64 *poly = data + (size_t)blockIdx.x * (S.mp ? 2 : 1) * kN;
  u32 v[8];
  Barrier bar;
  if (S.mp) {  // each lane block transforms its own limb's residues in place
    u64 *lp = poly + (size_t)sb * kN;
#pragma unroll
    for (int c = 0; c < 8; c++) v[c] = reduce_u64(lp[inverse ? map3(t, c) : map0(t, c)], L);
    if (!inverse) ntt_fwd_2048(v, nlds[sb], t, L, bar);
    else ntt_inv_2048(v, nlds[sb], t, L, bar);
#pragma unroll
    for (int c = 0; c < 8; c++) lp[inverse ? map0(t, c) : map3(t, c)] = v[c];
    return;
  }
  if (!inverse) {
#pragma unroll
    for (int c = 0; c < 8; c++) v[c] = reduce_u64(poly[map0(t, c)], L);
    ntt_fwd_2048(v, nlds[sb], t, L, bar);
#pragma unroll
    for (int c = 0; c < 8; c++) res[sb][map3(t, c)] = v[c];
  } else {
#pragma unroll
    for (int c = 0; c < 8; c++) v[c] = reduce_u64(poly[map3(t, c)], L);
    ntt_inv_2048(v, nlds[sb], t, L, bar);
#pragma unroll
    for (int c = 0; c < 8; c++) res[sb][map0(t, c)] = v[c];
  }
  __syncthreads();
  for (int j = threadIdx.x; j < kN; j += 512) poly[j] = crt_compose(res[0][j], res[1][j], S);
}

// RNS polys [npoly][2 limbs][N]: coef (map0) <-> NTT (map3).
__global__ __launch_bounds__(512) void k_ntt_rns(const u32 *in, u32 *out, int inverse, SchemeConst S) {
  __shared__ __align__(16) u32 nlds[2][kLdsPoly];
  const int sb = __builtin_amdgcn_readfirstlane(threadIdx.x >> 8), t = threadIdx.x & 255;  // sb: wave-uniform
  const LimbConst &L = S.limb[sb];
  const size_t off = ((size_t)blockIdx.x * 2 + sb) * kN;
  u32 v[8];
  Barrier bar;
  if (!inverse) {
    load_map0(v, in + off, t);
    ntt_fwd_2048(v, nlds[sb], t, L, bar);
    store_map3(out + off, v, t);
  } else {
    load_map3(v, in + off, t);
    ntt_inv_2048(v, nlds[sb], t, L, bar);
    store_map0(out + off, v, t);
  }
}

// ---------------------------------------------------------------------------
// DB preprocessing (src/server.cpp:80-203). Plaintext coefficients are < t,
// so NTT_Q(pt) mod q_i = NTT_{q_i}(pt): each limb is transformed directly.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(512) void k_db_ntt_plain(const uint16_t *__restrict__ pts, u32 *__restrict__ tmp,
                                                      SchemeConst S) {
  __shared__ __align__(16) u32 nlds[2][kLdsPoly];
  const int sb = __builtin_amdgcn_readfirstlane(threadIdx.x >> 8), t = threadIdx.x & 255;  // sb: wave-uniform
  const LimbConst &L = S.limb[sb];
  const uint16_t *p = pts + (size_t)blockIdx.x * kN;
  u32 v[8];
#pragma unroll
  for (int c = 0; c < 8; c++) v[c] = p[map0(t, c)];
  ntt_fwd_2048(v, nlds[sb], t, L, Barrier());
  store_map3(tmp + ((size_t)sb * gridDim.x + blockIdx.x) * kN, v, t);
}

static inline unsigned cdiv(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

// Seeded synthetic plaintexts, generated where they are stored: coefficient
// c of plaintext i is splitmix64(seed + i*N + c) mod t (the host recomputes
// recorded plaintexts with the same formula).
__device__ __forceinline__ u64 splitmix64_d(u64 x) {
  x += 0x9E3779B97F4A7C15ULL;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}

template <typename T>
__global__ void k_gen_plain(T *__restrict__ out, u64 first_pt, size_t count, int n, u64 seed, u64 t) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count * (size_t)n) return;
  out[i] = (T)(splitmix64_d(seed + first_pt * (u64)n + i) % t);
}

void launch_gen_plain_u16(uint16_t *out, size_t first_pt, size_t count, int n, u64 seed, u64 t, hipStream_t st) {
  if (count) k_gen_plain<uint16_t><<<cdiv(count * n, 256), 256, 0, st>>>(out, first_pt, count, n, seed, t);
}
void launch_gen_plain_u64(u64 *out, size_t first_pt, size_t count, int n, u64 seed, u64 t, hipStream_t st) {
  if (count) k_gen_plain<u64><<<cdiv(count * n, 256), 256, 0, st>>>(out, first_pt, count, n, seed, t);
}

// Forward NTT of both polys of `count` cts per query (evaluate_first_dim's
// query transform, server.cpp:331-337). grid (count*2, B).
__global__ __launch_bounds__(512) void k_ct_ntt(const u32 *in, u32 *out, size_t in_stride, size_t out_stride,
                                                SchemeConst S) {
  __shared__ __align__(16) u32 nlds[2][kLdsPoly];
  const int sb = __builtin_amdgcn_readfirstlane(threadIdx.x >> 8), t = threadIdx.x & 255;  // sb: wave-uniform
  const LimbConst &L = S.limb[sb];
  const size_t poly = blockIdx.x;  // ct*2 + p
  const size_t off = (poly * 2 + sb) * kN;
  u32 v[8];
  load_map0(v, in + blockIdx.y * in_stride + off, t);
  ntt_fwd_2048(v, nlds[sb], t, L, Barrier());
  store_map3(out + blockIdx.y * out_stride + off, v, t);
}

// ---------------------------------------------------------------------------
// Final stage: INTT, CRT compose, centred rescale Q -> small_q
// (server.cpp:1025-1043 with utils.cpp:445-462). grid (2 polys, B).
// ---------------------------------------------------------------------------
__device__ __forceinline__ u64 rescale_to_small(u64 x, const SchemeConst &S) {
  i64 v = (i64)x;
  if (v >= (i64)(S.Q / 2)) v -= (i64)S.Q;
  const bool neg = v < 0;
  const u64 a = neg ? (u64)(-v) : (u64)v;
  // floor((a * small_q + floor(Q/2)) / Q), numerator < Q * 2^32. Barrett
  // with m = floor(2^k / Q) in [2^63, 2^64), k = 63 + bit length of Q:
  // floor(num * m / 2^k), taking the high 64 bits of the low product only,
  // is the quotient or up to two less (num / 2^k < 2^-31); then correct.
  const unsigned __int128 num = (unsigned __int128)a * S.small_q + (S.Q >> 1);
  const u64 nh = (u64)(num >> 64), nl = (u64)num;
  const unsigned __int128 t = (unsigned __int128)nh * S.q_recip + __umul64hi(nl, S.q_recip);
  u64 quo = (u64)(t >> S.q_shift);
  unsigned __int128 r = num - (unsigned __int128)quo * S.Q;
#pragma unroll
  for (int k = 0; k < 2; k++)
    if (r >= S.Q) {
      r -= S.Q;
      quo++;
    }
  // a <= Q/2 gives quo <= (small_q + 1) / 2 < small_q, so the signed result
  // lies in (-small_q, small_q)
  return neg ? (quo ? S.small_q - quo : 0) : quo;
}

// Two-limb MP scheme (server.cpp:1044-1075): drop the second limb's prime
// with round half up, x = CRT(r0, r1) in [0, q0 q1) -> floor((x + floor(q1/2))
// / q1) reduced into [0, q0), then the centred rescale q0 -> small_q
// (utils.cpp:445-462). Both limbs are below 2^30, so every value fits u64.
__device__ __forceinline__ u64 drop_and_rescale(u32 r0, u32 r1, const SchemeConst &S) {
  const u64 q0 = S.limb[0].q, q1 = S.limb[1].q, sq = S.small_q;
  const u64 x = crt_compose(r0, r1, S);
  u64 o = (x + (q1 >> 1)) / q1;
  if (o >= q0) o -= q0;
  const u64 h = q0 / 2;
  const bool neg = o >= h;  // v >= inp/2 -> v - inp
  const u64 mag = neg ? q0 - o : o;
  const u64 quo = (mag * sq + h) / q0;  // mag <= q0/2 keeps quo below small_q
  return neg ? (quo ? sq - quo : 0) : quo;
}

__global__ __launch_bounds__(512) void k_final(const u32 *res, u64 *out, size_t res_stride, SchemeConst S) {
  __shared__ __align__(16) u32 nlds[2][kLdsPoly];
  __shared__ u32 sh[2][kN];
  const int sb = __builtin_amdgcn_readfirstlane(threadIdx.x >> 8), t = threadIdx.x & 255;  // sb: wave-uniform
  const LimbConst &L = S.limb[sb];
  const int p = blockIdx.x, qi = blockIdx.y;
  u32 v[8];
  load_map3(v, res + qi * res_stride + (p * 2 + sb) * kN, t);
  ntt_inv_2048(v, nlds[sb], t, L, Barrier());
  store_map0(sh[sb], v, t);
  __syncthreads();
  u64 *o = out + ((size_t)qi * 2 + p) * kN;
  if (S.mp)
    for (int j = threadIdx.x; j < kN; j += 512) o[j] = drop_and_rescale(sh[0][j], sh[1][j], S);
  else
    for (int j = threadIdx.x; j < kN; j += 512) o[j] = rescale_to_small(crt_compose(sh[0][j], sh[1][j], S), S);
}

// NTT-form RNS cts -> coefficient-form u64 mod Q. grid (count*2).
__global__ __launch_bounds__(512) void k_ct_to_coef_u64(const u32 *in, u64 *out, SchemeConst S) {
  __shared__ __align__(16) u32 nlds[2][kLdsPoly];
  __shared__ u32 sh[2][kN];
  const int sb = __builtin_amdgcn_readfirstlane(threadIdx.x >> 8), t = threadIdx.x & 255;  // sb: wave-uniform
  const LimbConst &L = S.limb[sb];
  const size_t poly = blockIdx.x;
  u32 v[8];
  load_map3(v, in + (poly * 2 + sb) * kN, t);
  ntt_inv_2048(v, nlds[sb], t, L, Barrier());
  store_map0(sh[sb], v, t);
  __syncthreads();
  if (S.mp) {  // [count][2 polys][2 limbs][N] residues
    for (int j = threadIdx.x; j < kN; j += 512) {
      out[(poly * 2 + 0) * kN + j] = sh[0][j];
      out[(poly * 2 + 1) * kN + j] = sh[1][j];
    }
    return;
  }
  for (int j = threadIdx.x; j < kN; j += 512) out[poly * kN + j] = crt_compose(sh[0][j], sh[1][j], S);
}

// Stand-alone external product (gsw.cpp:119-182): ct coef RNS -> out NTT RNS.
__global__ __launch_bounds__(512) void k_ext_prod(const u32 *gsw, int l, int b, const u32 *ct, u32 *out,
                                                  SchemeConst S) {
  __shared__ __align__(16) u32 nlds[2][kLdsPoly];
  const int sb = __builtin_amdgcn_readfirstlane(threadIdx.x >> 8), t = threadIdx.x & 255;  // sb: wave-uniform
  const LimbConst &L = S.limb[sb];
  i64 D[2][8];
#pragma unroll
  for (int p = 0; p < 2; p++)
#pragma unroll
    for (int c = 0; c < 8; c++) {
      const int j = map0(t, c);
      D[p][c] = gadget_value(ct[(p * 2 + 0) * kN + j], ct[(p * 2 + 1) * kN + j], S, !S.mp);
    }
  u64 acc0[8], acc1[8];
#pragma unroll
  for (int c = 0; c < 8; c++) acc0[c] = acc1[c] = 0;
  gadget_mac<2, true>(D, l, b, gsw, sb, t, L, nlds[sb], acc0, acc1, S.mp ? (i64)1 << (b - 1) : 0);
  u32 o0[8], o1[8];
#pragma unroll
  for (int c = 0; c < 8; c++) {
    o0[c] = reduce_u64(acc0[c], L);
    o1[c] = reduce_u64(acc1[c], L);
  }
  store_map3(out + (0 * 2 + sb) * kN, o0, t);
  store_map3(out + (1 * 2 + sb) * kN, o1, t);
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

void launch_split_u64(const u64 *in, u32 *out, size_t npoly, const SchemeConst &S, hipStream_t st) {
  k_split_u64<<<cdiv(npoly * kN, 256), 256, 0, st>>>(in, out, npoly, 0, 0, S);
}

void launch_split_u64_batch(const u64 *in, size_t in_stride, u32 *out, size_t out_stride, size_t npoly, int B,
                            const SchemeConst &S, hipStream_t st) {
  if (B <= 0) return;
  dim3 grid((unsigned)cdiv(npoly * kN, 256), (unsigned)B);
  k_split_u64<<<grid, 256, 0, st>>>(in, out, npoly, in_stride, out_stride, S);
}
void launch_join_u64(const u32 *in, u64 *out, size_t npoly, const SchemeConst &S, hipStream_t st) {
  k_join_u64<<<cdiv(npoly * kN, 256), 256, 0, st>>>(in, out, npoly, S);
}
void launch_addsub_mod(const u64 *a, const u64 *b, u64 *out, size_t n, size_t limb_len, int K, u64 q0, u64 q1,
                       bool sub, hipStream_t st) {
  if (n) k_addsub_mod<<<cdiv(n, 256), 256, 0, st>>>(a, b, out, n, limb_len, K, q0, q1, sub ? 1 : 0);
}
void launch_ntt_u64(u64 *data, size_t npoly, bool inverse, const SchemeConst &S, hipStream_t st) {
  if (npoly) k_ntt_u64<<<(unsigned)npoly, 512, 0, st>>>(data, inverse ? 1 : 0, S);
}
void launch_ntt_rns(const u32 *in, u32 *out, size_t npoly_rns, bool inverse, const SchemeConst &S,
                    hipStream_t st) {
  if (npoly_rns) k_ntt_rns<<<(unsigned)npoly_rns, 512, 0, st>>>(in, out, inverse ? 1 : 0, S);
}
void launch_db_ntt_plain(const uint16_t *pts, u32 *tmp, size_t count, const SchemeConst &S, hipStream_t st) {
  if (count) k_db_ntt_plain<<<(unsigned)count, 512, 0, st>>>(pts, tmp, S);
}
void launch_ct_ntt(const u32 *in, u32 *out, size_t in_stride, size_t out_stride, int count, int B,
                   const SchemeConst &S, hipStream_t st) {
  if (count <= 0) return;
  dim3 grid(count * 2, B);
  k_ct_ntt<<<grid, 512, 0, st>>>(in, out, in_stride, out_stride, S);
}

void launch_final(const u32 *res, u64 *out, size_t res_stride, int B, const SchemeConst &S, hipStream_t st) {
  dim3 grid(2, B);
  k_final<<<grid, 512, 0, st>>>(res, out, res_stride, S);
}

void launch_ct_to_coef_u64(const u32 *in, u64 *out, size_t count, const SchemeConst &S, hipStream_t st) {
  if (count) k_ct_to_coef_u64<<<(unsigned)(count * 2), 512, 0, st>>>(in, out, S);
}

void launch_ext_prod(const u32 *gsw, int l, int b, const u32 *ct_coef, u32 *out_ntt, const SchemeConst &S,
                     hipStream_t st) {
  k_ext_prod<<<1, 512, 0, st>>>(gsw, l, b, ct_coef, out_ntt, S);
}


}  // namespace opir
