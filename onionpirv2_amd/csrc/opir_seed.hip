// Seed-compressed client keys (opir_client_set_keys_seeded): the uniform
// halves of the BV key rows (a) and of the RGSW(s) rows (c1) are regenerated
// on the device from the client's 32-byte key seed, in NTT form, straight
// into the resident key layout; only the other halves cross PCIe. The
// sampling rules are those of the seeded host client (host/prg.h,
// csrc/opir_prg.h).
#include "opir_capi_internal.h"
#include "opir_prg.h"

using namespace capi;

namespace {

// (hi * 2^64 + lo) mod q for a 32-bit limb: hi mod q, times 2^64 mod q, plus lo.
__device__ __forceinline__ u32 mod128_u32(u64 hi, u64 lo, const LimbConst &L) {
  const u32 r64 = reduce_u64((u64)L.r32 * L.r32, L);
  return reduce_u64((u64)reduce_u64(hi, L) * r64 + reduce_u64(lo, L), L);
}

// The same for a 60-bit limb: hi * (2^64 mod q) and lo by Shoup, each in [0, 2q).
__device__ __forceinline__ u64 mod128_u64(u64 hi, u64 lo, const k2::Limb64 &L) {
  const u64 a = hi * L.r64 - __umul64hi(hi, L.r64s) * L.q;
  const u64 b = lo - __umul64hi(lo, L.one_s) * L.q;
  u64 s = a + b;  // < 4q < 2^62
  if (s >= L.q2) s -= L.q2;
  if (s >= L.q) s -= L.q;
  return s;
}

// Polynomial p (blockIdx.y) of a uniform slot set: stream (domain, i0 + p,
// limb) written to out + p * stride + limb * N in the device key layout.
// Each thread expands one ChaCha20 block into four coefficients.
// 32-bit limbs: composite (mp == 0) draws stream limb 0 once and stores its
// residues mod both primes; MP draws one stream per limb (blockIdx.z).
__global__ void k_seed_uniform_u32(u32 *out, size_t stride, prg::Key key, uint32_t domain, uint32_t i0, SchemeConst S) {
  const int p = blockIdx.y, limb = blockIdx.z;
  const uint32_t blk = blockIdx.x * blockDim.x + threadIdx.x;
  if (blk >= (uint32_t)(kN / 4)) return;
  uint32_t w[16];
  prg::chacha20_block(key, blk, domain, i0 + p, S.mp ? limb : 0, w);
  u32 *o = out + (size_t)p * stride;
#pragma unroll
  for (int m = 0; m < 4; m++) {
    u64 hi, lo;
    prg::draw128(w, m, hi, lo);
    const size_t c = 4 * (size_t)blk + m;
    if (S.mp) {
      o[(size_t)limb * kN + c] = mod128_u32(hi, lo, S.limb[limb]);
    } else {
      o[c] = mod128_u32(hi, lo, S.limb[0]);
      o[kN + c] = mod128_u32(hi, lo, S.limb[1]);
    }
  }
}

__global__ void k_seed_uniform_u64(u64 *out, size_t stride, prg::Key key, uint32_t domain, uint32_t i0,
                                   k2::Scheme64 S) {
  const int p = blockIdx.y, limb = blockIdx.z;
  const uint32_t blk = blockIdx.x * blockDim.x + threadIdx.x;
  if (blk >= (uint32_t)(S.N / 4)) return;
  uint32_t w[16];
  prg::chacha20_block(key, blk, domain, i0 + p, limb, w);
  u64 *o = out + (size_t)p * stride + (size_t)limb * S.N;
#pragma unroll
  for (int m = 0; m < 4; m++) {
    u64 hi, lo;
    prg::draw128(w, m, hi, lo);
    o[4 * (size_t)blk + m] = mod128_u64(hi, lo, S.limb[limb]);
  }
}

prg::Key key_words(const uint8_t *seed) {
  prg::Key k;
  for (int i = 0; i < 8; i++)
    k.w[i] = (uint32_t)seed[4 * i] | (uint32_t)seed[4 * i + 1] << 8 | (uint32_t)seed[4 * i + 2] << 16 |
             (uint32_t)seed[4 * i + 3] << 24;
  return k;
}

// Fill npoly uniform polynomials of stream (domain, i0 + p) at dev + p * stride
// (device key layout; stride and offset in elements of the key type).
int fill_uniform(opir_ctx *ctx, void *dev, size_t stride, size_t npoly, const prg::Key &key, uint32_t domain) {
  if (!npoly) return OPIR_OK;
  if (ctx->k2) {
    dim3 grid((unsigned)((ctx->N / 4 + 255) / 256), (unsigned)npoly, (unsigned)ctx->K);
    k_seed_uniform_u64<<<grid, 256, 0, ctx->stream>>>((u64 *)dev, stride, key, domain, 0, ctx->S64);
  } else {
    dim3 grid((unsigned)((kN / 4 + 255) / 256), (unsigned)npoly, ctx->S.mp ? 2 : 1);
    k_seed_uniform_u32<<<grid, 256, 0, ctx->stream>>>((u32 *)dev, stride, key, domain, 0, ctx->S);
  }
  LAUNCH_CHECK(ctx);
  return OPIR_OK;
}

}  // namespace

extern "C" int opir_client_set_keys_seeded(opir_ctx *ctx, uint64_t id, const uint8_t *seed, const uint64_t *bv_b,
                                           size_t h, const uint64_t *gsw_c0) {
  if (!ctx) return OPIR_EINVAL;
  if (!seed || (!bv_b && h) || !gsw_c0) return set_err(ctx, OPIR_EINVAL, "PirServer session key bundle cannot be null");
  const size_t KN = (size_t)ctx->K * ctx->N;
  const size_t L = ctx->k2 ? (size_t)ctx->G64.l_ks : (size_t)ctx->G.l_ks;
  const size_t lk = ctx->k2 ? (size_t)ctx->G64.l_key : (size_t)ctx->G.l_key;
  if (h > 12) return set_err(ctx, OPIR_EINVAL, "BV key count out of range");
  // the reference-layout bundle with zero uniform halves; the regular setter
  // validates and installs it, then the device writes the uniform halves
  std::vector<uint64_t> bv(h * L * 2 * KN, 0), gsw(2 * lk * 2 * KN, 0);
  for (size_t r = 0; r < h * L; r++) std::memcpy(&bv[(r * 2 + 1) * KN], bv_b + r * KN, KN * sizeof(uint64_t));
  for (size_t r = 0; r < 2 * lk; r++) std::memcpy(&gsw[(r * 2 + 0) * KN], gsw_c0 + r * KN, KN * sizeof(uint64_t));
  int rc = opir_client_set_keys(ctx, id, bv.data(), h, gsw.data());
  if (rc) return rc;
  const ClientKeys &k = ctx->clients()[id];
  const prg::Key key = key_words(seed);
  // device key rows: [b or c0 (-> c0), a or c1 (-> c1)][2 limbs][N]; the
  // uniform half is the second polynomial of each row
  const size_t row = ctx->k2 ? k2::ct_values(ctx->S64) : kCtWords, half = row / 2;
  if (ctx->k2) {
    if ((rc = fill_uniform(ctx, (u64 *)k.bv + half, row, h * L, key, prg::kBvA))) return rc;
    if ((rc = fill_uniform(ctx, (u64 *)k.gsw + half, row, 2 * lk, key, prg::kGswA))) return rc;
  } else {
    if ((rc = fill_uniform(ctx, (u32 *)k.bv + half, row, h * L, key, prg::kBvA))) return rc;
    if ((rc = fill_uniform(ctx, (u32 *)k.gsw + half, row, 2 * lk, key, prg::kGswA))) return rc;
  }
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return OPIR_OK;
}
