// Allocation-free homomorphic select (opir_ext_prod_mux_device_noalloc):
// the device-pointer form of opir_ext_prod_mux for callers that cannot call
// hipMalloc (code under HIP graph capture, tight serving loops). Its
// intermediates live in one file-scope device array per GPU, `scratch`,
// whose address each device resolves once (noalloc_init, from the context's
// creation); every call on that GPU shares it, so calls that overlap on one
// GPU race, and each call rewrites everything it reads from it.
#include "opir_capi_internal.h"

#include <mutex>

using namespace capi;

namespace {

constexpr size_t kScratchBytes = size_t{2} << 20;
constexpr size_t kMaxL1 = 16;  // 32-bit-limb schemes: gadget rows of the selector RGSW
constexpr size_t kMaxL2 = 32;  // 64-bit-limb schemes (the MAC sums bound it, as in scheme validation)

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_ptr[kMaxDevices] = {};

}  // namespace

namespace capi {

// Resolves this device's copy of `scratch` (the device is current).
int noalloc_init(opir_ctx *ctx) {
  if (ctx->device < 0 || ctx->device >= kMaxDevices)
    return set_err(nullptr, OPIR_EINVAL, "device ordinal out of range");
  std::lock_guard<std::mutex> g(g_mu);
  if (!g_ptr[ctx->device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_err(nullptr, OPIR_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_ptr[ctx->device] = p;
  }
  ctx->scratch = g_ptr[ctx->device];
  return OPIR_OK;
}

}  // namespace capi

extern "C" int opir_ext_prod_mux_device_noalloc(opir_ctx *ctx, const uint64_t *d_gsw, size_t l, size_t b,
                                                const uint64_t *d_x, const uint64_t *d_y, uint64_t *d_out) {
  if (!ctx || !d_gsw || !d_x || !d_y || !d_out) return OPIR_EINVAL;
  if (!ctx->scratch) return set_err(ctx, OPIR_ERUNTIME, "scratch buffer not resolved for this device");
  if (b == 0 || b > 32) return set_err(ctx, OPIR_EINVAL, "invalid gadget");
  unsigned char *s = static_cast<unsigned char *>(ctx->scratch);
  const hipStream_t st = ctx->stream;
  if (ctx->k2) {
    if (l == 0 || l > kMaxL2)
      return set_err(ctx, OPIR_EINVAL, "selector gadget of " + std::to_string(l) + " rows exceeds the " +
                                           std::to_string(kMaxL2) + "-row scratch limit");
    const size_t ct = k2::ct_values(ctx->S64);  // u64 values per ciphertext
    u64 *diff = reinterpret_cast<u64 *>(s);
    u64 *o = diff + ct;
    static_assert(2 * k2::kMaxCt * sizeof(u64) <= kScratchBytes, "scratch too small for the 64-bit-limb layouts");
    const u64 q0 = ctx->S64.limb[0].q, q1 = ctx->K == 2 ? ctx->S64.limb[1].q : q0;
    // y - x, external product (NTT form), INTT, + x (server.cpp:696-721)
    launch_addsub_mod(d_y, d_x, diff, ct, ctx->N, ctx->K, q0, q1, true, st);
    k2::launch_k2_ext_prod(d_gsw, (int)l, (int)b, diff, o, ctx->S64, st);
    k2::launch_k2_ntt(o, 2, true, ctx->S64, st);
    launch_addsub_mod(d_x, o, d_out, ct, ctx->N, ctx->K, q0, q1, false, st);
    LAUNCH_CHECK(ctx);
    return OPIR_OK;
  }
  if (l == 0 || l > kMaxL1)
    return set_err(ctx, OPIR_EINVAL, "selector gadget of " + std::to_string(l) + " rows exceeds the " +
                                         std::to_string(kMaxL1) + "-row scratch limit");
  // scratch: diff (u64 ct) | o64 (u64 ct) | c32 | o32 | g32 (2l RNS rows)
  const size_t ct64 = 2 * (size_t)ctx->K * kN;  // u64 values of an ABI ciphertext
  u64 *diff = reinterpret_cast<u64 *>(s);
  u64 *o64 = diff + ct64;
  u32 *c32 = reinterpret_cast<u32 *>(o64 + ct64);
  u32 *o32 = c32 + kCtWords;
  u32 *g32 = o32 + kCtWords;
  static_assert(2 * 4 * (size_t)kN * sizeof(u64) + (2 + 2 * kMaxL1) * kCtWords * sizeof(u32) <= kScratchBytes,
                "scratch too small for the 32-bit limb layouts");
  const u64 m0 = ctx->S.mp ? ctx->S.limb[0].q : ctx->S.Q, m1 = ctx->S.mp ? ctx->S.limb[1].q : ctx->S.Q;
  launch_addsub_mod(d_y, d_x, diff, ct64, kN, ctx->K, m0, m1, true, st);
  launch_split_u64(d_gsw, g32, 2 * l * 2, ctx->S, st);
  launch_split_u64(diff, c32, 2, ctx->S, st);
  launch_ext_prod(g32, (int)l, (int)b, c32, o32, ctx->S, st);
  launch_join_u64(o32, o64, 2, ctx->S, st);
  launch_ntt_u64(o64, 2, true, ctx->S, st);
  launch_addsub_mod(d_x, o64, d_out, ct64, kN, ctx->K, m0, m1, false, st);
  LAUNCH_CHECK(ctx);
  return OPIR_OK;
}
