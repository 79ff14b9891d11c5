// Client-side helpers on the GPU, for load generation: keys and queries of
// seeded clients (the counter-based sampler of host/prg.h, device copy in
// opir_prg.h) produced in HBM, equal value for value to what the host
// PirClient(params, id, key) produces. They follow the client side of the
// reference, src/client.cpp:20-225 (query packing), src/bv_keyswitch.cpp:
// 191-276 (BV Galois keys) and src/gsw.cpp:359-422 with src/client.cpp:25-38
// (RGSW(s)); the 32-bit-limb schemes only (CONFIG_N2048_K1_COMP,
// CONFIG_N2048_K2_MP).
//
// One 512-thread workgroup per key row or query: lane block = limb, each
// ChaCha20 block of a stream expanded by one thread into LDS first.
#include "opir_capi_internal.h"
#include "opir_prg.h"

using namespace capi;

namespace {

struct Barrier {
  __device__ void operator()() const { __syncthreads(); }
};

struct Cdt {
  u64 t[2 * prg::kCdtMax + 1];
};

// Per-client inputs of a generation launch.
struct GenArgs {
  const prg::Key *keys;   // per client (blockIdx.y)
  u32 *const *bv;         // per client: device BV key rows [h][l_ks][b, a][2 limbs][N]
  u32 *const *gsw;        // per client: device RGSW(s) rows [2 l_key][c0, c1][2 limbs][N]
  int l_ks, l_key;
  u32 bv_pow[16][2];      // 2^(b_ks * r) mod q_limb
  u32 gsw_g[32][2];       // gadget value of row r mod q_limb (2^(b_key * (l_key - 1 - r)))
  Cdt cdt;
};

__device__ __forceinline__ u32 mulmod(u32 a, u32 b, const LimbConst &L) { return reduce_u64((u64)a * b, L); }
__device__ __forceinline__ u32 neg_mod(u32 a, u32 q) { return a ? q - a : 0; }
__device__ __forceinline__ u32 signed_res(int v, u32 q) { return v < 0 ? q - (u32)(-v) : (u32)v; }

// The ternary secret {0, 1, -1} of stream (kSecret, 0, 0) into s8[N].
__device__ void stream_ternary(const prg::Key &key, int8_t *s8) {
  for (int b = threadIdx.x; b < kN / 8; b += blockDim.x) {
    uint32_t w[16];
    prg::chacha20_block(key, b, prg::kSecret, 0, 0, w);
#pragma unroll
    for (int m = 0; m < 8; m++) {
      const int r = (int)(prg::draw64(w, m) % 3);
      s8[8 * b + m] = (int8_t)(r == 2 ? -1 : r);
    }
  }
}

// Rounded Gaussian errors of stream (domain, i, 0) into e8[N] (|e| <= kCdtMax).
__device__ void stream_gauss(const prg::Key &key, uint32_t domain, uint32_t i, const Cdt &cdt, int8_t *e8) {
  for (int b = threadIdx.x; b < kN / 8; b += blockDim.x) {
    uint32_t w[16];
    prg::chacha20_block(key, b, domain, i, 0, w);
#pragma unroll
    for (int m = 0; m < 8; m++) {
      const u64 u = prg::draw64(w, m);
      int v = 0;
      while (u > cdt.t[v]) v++;
      e8[8 * b + m] = (int8_t)(v - prg::kCdtMax);
    }
  }
}

__device__ __forceinline__ u32 mod128(u64 hi, u64 lo, const LimbConst &L) {
  const u32 r64 = reduce_u64((u64)L.r32 * L.r32, L);
  return reduce_u64((u64)reduce_u64(hi, L) * r64 + reduce_u64(lo, L), L);
}

// Uniform residues of stream (domain, i, limb) into u[2][N]: composite draws
// stream limb 0 once for both residues, MP one stream per limb.
__device__ void stream_uniform(const prg::Key &key, uint32_t domain, uint32_t i, const SchemeConst &S, u32 (*u)[kN]) {
  const int nstreams = S.mp ? 2 : 1;
  for (int job = threadIdx.x; job < nstreams * (kN / 4); job += blockDim.x) {
    const int limb = job / (kN / 4), b = job % (kN / 4);
    uint32_t w[16];
    prg::chacha20_block(key, b, domain, i, limb, w);
#pragma unroll
    for (int m = 0; m < 4; m++) {
      u64 hi, lo;
      prg::draw128(w, m, hi, lo);
      if (S.mp) {
        u[limb][4 * b + m] = mod128(hi, lo, S.limb[limb]);
      } else {
        u[0][4 * b + m] = mod128(hi, lo, S.limb[0]);
        u[1][4 * b + m] = mod128(hi, lo, S.limb[1]);
      }
    }
  }
}

// NTT (map3 out) of a small signed polynomial held in LDS, this lane block's limb.
template <typename F>
__device__ __forceinline__ void ntt_small(F coef, u32 v[8], u32 *nlds, int t, const LimbConst &L) {
#pragma unroll
  for (int c = 0; c < 8; c++) v[c] = signed_res(coef(map0(t, c)), L.q);
  ntt_fwd_2048(v, nlds, t, L, Barrier());
}

// BV key row (level, r) of client blockIdx.y: b = sigma_g(s) 2^(b r) - a s + e,
// a uniform, NTT form (bv_keyswitch.cpp:222-276), g = N / 2^level + 1.
__global__ __launch_bounds__(512) void k_gen_bv(GenArgs A, SchemeConst S) {
  __shared__ __align__(16) u32 nlds[2][kLdsPoly];
  __shared__ u32 ua[2][kN];
  __shared__ int8_t s8[kN], e8[kN];
  const int limb = __builtin_amdgcn_readfirstlane(threadIdx.x >> 8), t = threadIdx.x & 255;
  const LimbConst &L = S.limb[limb];
  const int row = blockIdx.x, ci = blockIdx.y;
  const int level = row / A.l_ks, r = row % A.l_ks;
  const prg::Key key = A.keys[ci];
  stream_ternary(key, s8);
  stream_gauss(key, prg::kBvE, row, A.cdt, e8);
  stream_uniform(key, prg::kBvA, row, S, ua);
  __syncthreads();
  // sigma_g(s)[j] = s[j g^-1 mod 2N] with the negacyclic sign (utils.cpp:58-72)
  const unsigned g = (unsigned)(kN >> level) + 1;
  unsigned ginv = 1;
  for (int e = 0; e < kN - 1; e++) ginv = (ginv * g) & (2 * kN - 1);
  u32 sig[8], s[8], e[8];
  ntt_small([&](int j) {
    const int i0 = (int)(((unsigned)j * ginv) & (2 * kN - 1));
    return i0 < kN ? (int)s8[i0] : -(int)s8[i0 - kN];
  }, sig, nlds[limb], t, L);
  ntt_small([&](int j) { return (int)s8[j]; }, s, nlds[limb], t, L);
  ntt_small([&](int j) { return (int)e8[j]; }, e, nlds[limb], t, L);
  u32 *dst = A.bv[ci] + (size_t)row * kCtWords;
  const u32 br = A.bv_pow[r][limb];
#pragma unroll
  for (int c = 0; c < 8; c++) {
    const int k = map3(t, c);
    const u32 a = ua[limb][k];
    const u32 msg = mulmod(sig[c], br, L), as = mulmod(a, s[c], L);
    dst[(size_t)limb * kN + k] = add_mod(sub_mod(msg, as, L.q), e[c], L.q);  // b (-> c0)
    dst[(size_t)(2 + limb) * kN + k] = a;                                    // a (-> c1)
  }
}

// RGSW(s) row (half, r), row id = half * l_key + r: c1 uniform, c0 =
// -(c1 s + e) + g_r (half ? s^2 : s), NTT form (the seeded client's rule).
__global__ __launch_bounds__(512) void k_gen_gsw(GenArgs A, SchemeConst S) {
  __shared__ __align__(16) u32 nlds[2][kLdsPoly];
  __shared__ u32 ua[2][kN];
  __shared__ int8_t s8[kN], e8[kN];
  const int limb = __builtin_amdgcn_readfirstlane(threadIdx.x >> 8), t = threadIdx.x & 255;
  const LimbConst &L = S.limb[limb];
  const int row = blockIdx.x, ci = blockIdx.y;
  const int half = row / A.l_key, r = row % A.l_key;
  const prg::Key key = A.keys[ci];
  stream_ternary(key, s8);
  stream_gauss(key, prg::kGswE, row, A.cdt, e8);
  stream_uniform(key, prg::kGswA, row, S, ua);
  __syncthreads();
  u32 s[8], e[8];
  ntt_small([&](int j) { return (int)s8[j]; }, s, nlds[limb], t, L);
  ntt_small([&](int j) { return (int)e8[j]; }, e, nlds[limb], t, L);
  u32 *dst = A.gsw[ci] + (size_t)row * kCtWords;
  const u32 gr = A.gsw_g[r][limb];
#pragma unroll
  for (int c = 0; c < 8; c++) {
    const int k = map3(t, c);
    const u32 c1 = ua[limb][k];
    const u32 m = mulmod(gr, half ? mulmod(s[c], s[c], L) : s[c], L);
    const u32 negv = neg_mod(add_mod(mulmod(c1, s[c], L), e[c], L.q), L.q);
    dst[(size_t)limb * kN + k] = add_mod(negv, m, L.q);  // c0
    dst[(size_t)(2 + limb) * kN + k] = c1;               // c1
  }
}

// Plaintext terms of one query: slot and residue per limb to add to c0.
constexpr int kMaxTerms = 1 + 40 * 8;
struct QueryArgs {
  const prg::Key *keys;     // per query: the client's key
  const u64 *counters;      // per query: the client's query counter
  const u32 *terms;         // per query: kMaxTerms x (slot, residue limb 0, residue limb 1)
  const int *nterms;        // per query
  u64 *out;                 // [B][2][K*N] ABI layout
  Cdt cdt;
};

// One query of a seeded client (client.cpp:107-225 with the seeded
// encryption): seed = block 0 of (kQuerySeed, counter), c1 = uniform(seed,
// kQueryA), c0 = -(c1 s + e) + plaintext terms, coefficient form.
__global__ __launch_bounds__(512) void k_gen_query(QueryArgs A, SchemeConst S) {
  __shared__ __align__(16) u32 nlds[2][kLdsPoly];
  __shared__ u32 ua[2][kN], c0s[2][kN];
  __shared__ int8_t s8[kN], e8[kN];
  __shared__ prg::Key qkey;
  const int limb = __builtin_amdgcn_readfirstlane(threadIdx.x >> 8), t = threadIdx.x & 255;
  const LimbConst &L = S.limb[limb];
  const int qi = blockIdx.x;
  const prg::Key key = A.keys[qi];
  if (threadIdx.x == 0) {
    const u64 ctr = A.counters[qi];
    uint32_t w[16];
    prg::chacha20_block(key, 0, prg::kQuerySeed, (uint32_t)ctr, (uint32_t)(ctr >> 32), w);
    for (int i = 0; i < 8; i++) qkey.w[i] = w[i];
  }
  stream_ternary(key, s8);
  __syncthreads();
  const prg::Key qk = qkey;
  stream_gauss(qk, prg::kQueryE, 0, A.cdt, e8);
  stream_uniform(qk, prg::kQueryA, 0, S, ua);
  __syncthreads();
  u32 s[8], a[8];
  ntt_small([&](int j) { return (int)s8[j]; }, s, nlds[limb], t, L);
#pragma unroll
  for (int c = 0; c < 8; c++) a[c] = ua[limb][map0(t, c)];
  ntt_fwd_2048(a, nlds[limb], t, L, Barrier());
#pragma unroll
  for (int c = 0; c < 8; c++) a[c] = mulmod(a[c], s[c], L);
  ntt_inv_2048(a, nlds[limb], t, L, Barrier());  // c1 * s, coefficient form (map0)
#pragma unroll
  for (int c = 0; c < 8; c++) {
    const int j = map0(t, c);
    c0s[limb][j] = neg_mod(add_mod(a[c], signed_res(e8[j], L.q), L.q), L.q);
  }
  __syncthreads();
  if (threadIdx.x < 2) {  // the plaintext terms, one limb per thread, in order
    const u32 *tm = A.terms + (size_t)qi * kMaxTerms * 3;
    const u32 q = S.limb[threadIdx.x].q;
    for (int k = 0; k < A.nterms[qi]; k++) {
      const u32 slot = tm[3 * k];
      c0s[threadIdx.x][slot] = add_mod(c0s[threadIdx.x][slot], tm[3 * k + 1 + threadIdx.x], q);
    }
  }
  __syncthreads();
  const size_t KN = (size_t)(S.mp ? 2 : 1) * kN;
  u64 *o = A.out + (size_t)qi * 2 * KN;
  for (int j = threadIdx.x; j < kN; j += 512) {
    if (S.mp) {
      o[j] = c0s[0][j];
      o[kN + j] = c0s[1][j];
      o[KN + j] = ua[0][j];
      o[KN + kN + j] = ua[1][j];
    } else {
      o[j] = crt_compose(c0s[0][j], c0s[1][j], S);
      o[KN + j] = crt_compose(ua[0][j], ua[1][j], S);
    }
  }
}

prg::Key key_words(const uint8_t *seed) {
  prg::Key k;
  for (int i = 0; i < 8; i++)
    k.w[i] = (uint32_t)seed[4 * i] | (uint32_t)seed[4 * i + 1] << 8 | (uint32_t)seed[4 * i + 2] << 16 |
             (uint32_t)seed[4 * i + 3] << 24;
  return k;
}

// Rounded Gaussian table: entry v = floor(Phi((v - kCdtMax + 1/2) / sigma) 2^64),
// the last one saturated (the host sampler's rule, host/prg.cpp).
Cdt make_cdt(double sigma) {
  Cdt c;
  for (int v = -prg::kCdtMax; v <= prg::kCdtMax; v++) {
    const double p = 0.5 * std::erfc(-(v + 0.5) / (sigma * std::sqrt(2.0)));
    const double scaled = std::ldexp(p, 64);
    c.t[v + prg::kCdtMax] = scaled >= 18446744073709551615.0 ? ~u64{0} : (u64)scaled;
  }
  c.t[2 * prg::kCdtMax] = ~u64{0};
  return c;
}

u64 powmod2(u64 e, u64 m) { return powmod_h(2, e, m); }

// Query packing (client.cpp:48-225): plaintext terms of index pt_idx on db's
// layout -- Delta (2^h)^-1 at BitRev(column), and for every selector bit 1 the
// gadget powers times (2^h)^-1 at BitRev(N0 + (i-1) l + k) -- as residues per limb.
int query_terms(const opir_db *db, uint64_t pt_idx, u32 *terms, int &nt) {
  const opir_ctx *ctx = db->ctx;
  const size_t N0 = db->N0, other = db->Nrest, d = db->dims, h = db->h;
  if (pt_idx >= db->num_pt) return OPIR_ERANGE;
  const u64 t = ctx->sch.plain_mod, q1 = ctx->S.limb[0].q, q2 = ctx->S.limb[1].q;
  const size_t col = pt_idx % N0, row = pt_idx / N0;
  std::vector<size_t> idx{col};
  if (d > 1) {
    const size_t hs = d - 1, deep = 2 * other - (size_t{1} << hs), shallow = other - deep;
    size_t perfect = row;
    if (row < shallow) {
      idx.push_back(0);
    } else {
      idx.push_back((row - shallow) % 2);
      perfect = shallow + (row - shallow) / 2;
    }
    for (size_t k = hs - 1; k-- > 0;) idx.push_back((perfect >> k) & 1);
  }
  u64 inv_t = 0;
  if (!inv_h((u64{1} << h) % t, t, inv_t)) return OPIR_EINVAL;
  nt = 0;
  auto add = [&](size_t slot, u64 r1, u64 r2) {
    terms[3 * nt] = (u32)slot;
    terms[3 * nt + 1] = (u32)r1;
    terms[3 * nt + 2] = (u32)r2;
    nt++;
  };
  const int lg = (int)h;
  const size_t slot0 = brv(col, lg);
  const u128 Q = (u128)q1 * q2;
  if (!ctx->S.mp) {  // round(Q inv / t) mod Q
    const u64 scaled = (u64)((Q * inv_t + (t >> 1)) / t % Q);
    add(slot0, scaled % q1, scaled % q2);
  } else {  // Delta inv + round((Q - Delta t) inv / t), per limb
    const u128 Delta = Q / t;
    const u64 rem = (u64)(Q - Delta * t);
    const u64 rr = (u64)(((u128)rem * inv_t + (t >> 1)) / t);
    const u128 sm = Delta * inv_t + rr;
    add(slot0, (u64)(sm % q1), (u64)(sm % q2));
  }
  const size_t l = (size_t)ctx->G.l_ep, b = (size_t)ctx->G.b_ep;
  for (size_t i = 1; i < idx.size(); i++) {
    if (idx[i] != 1) continue;
    for (size_t k = 0; k < l; k++) {
      const size_t slot = brv(N0 + (i - 1) * l + k, lg);
      const u64 e = b * (l - 1 - k);
      if (!ctx->S.mp) {
        const u64 Qm = ctx->S.Q;
        u64 inv_q = 0;
        inv_h((u64{1} << h) % Qm, Qm, inv_q);
        const u64 v = mulmod_h(powmod2(e, Qm), inv_q, Qm);
        add(slot, v % q1, v % q2);
      } else {
        u64 i1 = 0, i2 = 0;
        inv_h((u64{1} << h) % q1, q1, i1);
        inv_h((u64{1} << h) % q2, q2, i2);
        add(slot, mulmod_h(powmod2(e, q1), i1, q1), mulmod_h(powmod2(e, q2), i2, q2));
      }
    }
  }
  return OPIR_OK;
}

int check_k1(opir_ctx *ctx) {
  if (ctx->k2) return set_err(ctx, OPIR_EINVAL, "device client generation runs the 32-bit-limb schemes (N = 2048)");
  return OPIR_OK;
}

int gen_queries(opir_db *db, size_t B, const uint8_t *keys, const uint64_t *counters, const uint64_t *indices,
                u64 *d_out) {
  opir_ctx *ctx = db->ctx;
  std::vector<prg::Key> k(B);
  std::vector<u32> terms(B * kMaxTerms * 3, 0);
  std::vector<int> nts(B);
  for (size_t i = 0; i < B; i++) {
    k[i] = key_words(keys + 32 * i);
    const int rc = query_terms(db, indices[i], &terms[i * kMaxTerms * 3], nts[i]);
    if (rc == OPIR_ERANGE) return set_err(ctx, OPIR_ERANGE, "PIR plaintext index exceeds the runtime layout");
    if (rc) return set_err(ctx, rc, "query packing: 2^h not invertible mod t");
    if (nts[i] > kMaxTerms) return set_err(ctx, OPIR_EINVAL, "too many query terms");
  }
  DevBuf dk, dc, dt, dn;
  DEV_ALLOC(ctx, dk, B * sizeof(prg::Key));
  DEV_ALLOC(ctx, dc, B * sizeof(u64));
  DEV_ALLOC(ctx, dt, terms.size() * sizeof(u32));
  DEV_ALLOC(ctx, dn, B * sizeof(int));
  HIP_TRY(ctx, hipMemcpy(dk.p, k.data(), B * sizeof(prg::Key), hipMemcpyHostToDevice));
  HIP_TRY(ctx, hipMemcpy(dc.p, counters, B * sizeof(u64), hipMemcpyHostToDevice));
  HIP_TRY(ctx, hipMemcpy(dt.p, terms.data(), terms.size() * sizeof(u32), hipMemcpyHostToDevice));
  HIP_TRY(ctx, hipMemcpy(dn.p, nts.data(), B * sizeof(int), hipMemcpyHostToDevice));
  QueryArgs A;
  A.keys = (const prg::Key *)dk.p;
  A.counters = (const u64 *)dc.p;
  A.terms = (const u32 *)dt.p;
  A.nterms = (const int *)dn.p;
  A.out = d_out;
  A.cdt = make_cdt(ctx->noise_sigma);
  k_gen_query<<<(unsigned)B, 512, 0, ctx->stream>>>(A, ctx->S);
  LAUNCH_CHECK(ctx);
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));  // the argument buffers are freed on return
  return OPIR_OK;
}

}  // namespace

extern "C" {

int opir_client_keygen_device(opir_ctx *ctx, size_t n, const uint64_t *ids, const uint8_t *keys, size_t h) {
  if (!ctx || (n && (!ids || !keys))) return OPIR_EINVAL;
  if (ctx->key_owner) return set_err(ctx, OPIR_EINVAL, "a cloned context reads its parent's client keys");
  if (int rc = check_k1(ctx)) return rc;
  if (h > 11) return set_err(ctx, OPIR_EINVAL, "BV key count out of range");
  if (!n) return OPIR_OK;
  const size_t L = ctx->G.l_ks, lk = ctx->G.l_key;
  if (L > 16 || lk > 32) return set_err(ctx, OPIR_EINVAL, "gadget too long for device key generation");
  GenArgs A;
  A.l_ks = (int)L;
  A.l_key = (int)lk;
  for (int limb = 0; limb < 2; limb++) {
    const u64 q = ctx->S.limb[limb].q;
    for (size_t r = 0; r < L; r++) A.bv_pow[r][limb] = (u32)powmod2((u64)ctx->G.b_ks * r, q);
    // composite: 2^e mod Q reduced mod q equals 2^e mod q; MP: per limb
    for (size_t r = 0; r < lk; r++) A.gsw_g[r][limb] = (u32)powmod2((u64)ctx->G.b_key * (lk - 1 - r), q);
  }
  A.cdt = make_cdt(ctx->noise_sigma);
  std::vector<ClientKeys> made(n);
  std::vector<prg::Key> kw(n);
  std::vector<u32 *> pb(n), pg(n);
  auto release = [&]() {
    for (auto &k : made) {
      if (k.bv) (void)hipFree(k.bv);
      if (k.gsw) (void)hipFree(k.gsw);
    }
  };
  for (size_t i = 0; i < n; i++) {
    kw[i] = key_words(keys + 32 * i);
    made[i].h = h;
    if ((h && hipMalloc(&made[i].bv, h * L * kCtWords * sizeof(u32)) != hipSuccess) ||
        hipMalloc(&made[i].gsw, 2 * lk * kCtWords * sizeof(u32)) != hipSuccess) {
      release();
      return set_err(ctx, OPIR_ENOMEM, "hipMalloc keys");
    }
    pb[i] = (u32 *)made[i].bv;
    pg[i] = (u32 *)made[i].gsw;
  }
  DevBuf dk, db, dg;
  if (hipMalloc(&dk.p, n * sizeof(prg::Key)) != hipSuccess || hipMalloc(&db.p, n * sizeof(u32 *)) != hipSuccess ||
      hipMalloc(&dg.p, n * sizeof(u32 *)) != hipSuccess) {
    release();
    return set_err(ctx, OPIR_ENOMEM, "hipMalloc generation tables");
  }
  if (hipMemcpy(dk.p, kw.data(), n * sizeof(prg::Key), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(db.p, pb.data(), n * sizeof(u32 *), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dg.p, pg.data(), n * sizeof(u32 *), hipMemcpyHostToDevice) != hipSuccess) {
    release();
    return set_err(ctx, OPIR_EHIP, "upload generation tables");
  }
  A.keys = (const prg::Key *)dk.p;
  A.bv = (u32 *const *)db.p;
  A.gsw = (u32 *const *)dg.p;
  if (h) k_gen_bv<<<dim3((unsigned)(h * L), (unsigned)n), 512, 0, ctx->stream>>>(A, ctx->S);
  k_gen_gsw<<<dim3((unsigned)(2 * lk), (unsigned)n), 512, 0, ctx->stream>>>(A, ctx->S);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    release();
    return set_err(ctx, OPIR_EHIP, std::string("device key generation: ") + hipGetErrorString(e));
  }
  for (size_t i = 0; i < n; i++) {
    auto it = ctx->clients().find(ids[i]);
    if (it != ctx->clients().end()) {
      (void)hipFree(it->second.bv);
      (void)hipFree(it->second.gsw);
    }
    ctx->clients()[ids[i]] = made[i];
  }
  return OPIR_OK;
}

int opir_query_gen_device(opir_db *db, size_t B, const uint8_t *keys, const uint64_t *counters,
                          const uint64_t *indices, uint64_t *d_queries) {
  if (!db || (B && (!keys || !counters || !indices || !d_queries))) return OPIR_EINVAL;
  if (int rc = check_k1(db->ctx)) return rc;
  if (!B) return OPIR_OK;
  return gen_queries(db, B, keys, counters, indices, d_queries);
}

int opir_query_gen(opir_db *db, size_t B, const uint8_t *keys, const uint64_t *counters, const uint64_t *indices,
                   uint64_t *queries) {
  if (!db || (B && (!keys || !counters || !indices || !queries))) return OPIR_EINVAL;
  opir_ctx *ctx = db->ctx;
  if (int rc = check_k1(ctx)) return rc;
  if (!B) return OPIR_OK;
  const size_t bytes = B * 2 * (size_t)ctx->K * kN * sizeof(u64);
  DevBuf o;
  DEV_ALLOC(ctx, o, bytes);
  if (int rc = gen_queries(db, B, keys, counters, indices, (u64 *)o.p)) return rc;
  HIP_TRY(ctx, hipMemcpy(queries, o.p, bytes, hipMemcpyDeviceToHost));
  return OPIR_OK;
}

int opir_client_get_keys(opir_ctx *ctx, uint64_t id, uint64_t *bv, size_t h, uint64_t *gsw) {
  if (!ctx || (!bv && h) || !gsw) return OPIR_EINVAL;
  auto it = ctx->clients().find(id);
  if (it == ctx->clients().end()) return set_err(ctx, OPIR_ERANGE, "unknown client id " + std::to_string(id));
  const ClientKeys &k = it->second;
  if (h != k.h) return set_err(ctx, OPIR_EINVAL, "BV key count differs from the resident keys");
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  const size_t N = ctx->N, KN = (size_t)ctx->K * N;
  const size_t L = ctx->k2 ? (size_t)ctx->G64.l_ks : (size_t)ctx->G.l_ks;
  const size_t lk = ctx->k2 ? (size_t)ctx->G64.l_key : (size_t)ctx->G.l_key;
  // device rows [first, second][limbs][N]; BV ABI rows are [a, b] (device b, a), RGSW [c0, c1]
  auto export_rows = [&](const void *dev, size_t nrows, uint64_t *out, bool swap) -> int {
    if (!nrows) return OPIR_OK;
    if (ctx->k2) {
      const size_t ct = k2::ct_values(ctx->S64);
      std::vector<u64> h64(nrows * ct);
      HIP_TRY(ctx, hipMemcpy(h64.data(), dev, h64.size() * sizeof(u64), hipMemcpyDeviceToHost));
      for (size_t r = 0; r < nrows; r++)
        for (int p = 0; p < 2; p++)
          std::memcpy(out + (r * 2 + (swap ? 1 - p : p)) * KN, &h64[r * ct + (size_t)p * KN], KN * sizeof(u64));
      return OPIR_OK;
    }
    std::vector<u32> h32(nrows * kCtWords);
    HIP_TRY(ctx, hipMemcpy(h32.data(), dev, h32.size() * sizeof(u32), hipMemcpyDeviceToHost));
    for (size_t r = 0; r < nrows; r++)
      for (int p = 0; p < 2; p++) {
        const u32 *src = &h32[r * kCtWords + (size_t)p * 2 * kN];
        uint64_t *dst = out + (r * 2 + (swap ? 1 - p : p)) * KN;
        for (size_t j = 0; j < (size_t)kN; j++) {
          if (ctx->S.mp) {
            dst[j] = src[j];
            dst[kN + j] = src[kN + j];
          } else {
            const u64 q1 = ctx->S.limb[0].q, q2 = ctx->S.limb[1].q;
            u64 inv = 0;
            inv_h(q1 % q2, q2, inv);
            const u64 d = ((u64)src[kN + j] + q2 - (u64)src[j] % q2) % q2;
            dst[j] = (u64)src[j] + q1 * mulmod_h(d, inv, q2);
          }
        }
      }
    return OPIR_OK;
  };
  if (int rc = export_rows(k.bv, h * L, bv, true)) return rc;
  return export_rows(k.gsw, 2 * lk, gsw, false);
}

}  // extern "C"
