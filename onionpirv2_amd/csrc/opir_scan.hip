// First-dimension DB scan on the matrix cores (src/server.cpp:313-361,
// src/matrix.cpp:24-151,232-350):
//   C[limb][l][i][c] = sum_k A[limb][l][i][k] * Bq[limb][l][k][c] mod q_limb
// for every NTT level l, candidate row i < Nrest, first-dim column k < N0 and
// query column c = 2*query + poly.
//
// Exactness with int8 MFMA: every residue a < 2^(8P-2) is written in P
// balanced base-256 digits a = sum_s 256^s d_s, d_0..d_{P-2} in [-128, 128)
// and a small non-negative top digit. Then
//   sum_k a_k b_k = sum_{s,t} 256^(s+t) sum_k d_{s,k} e_{t,k},
// where each inner sum is one int8 GEMM with int32 accumulation. The P^2
// plane pairs accumulate into 2P-1 int32 tiles G_u (u = s+t); at most P pairs
// share a u, so |G_u| <= P * N0 * 2^14 < 2^31 for P * N0 <= 2^17. No
// rounding anywhere. P = 4 covers the 29-bit limbs of CONFIG_N2048_K1_COMP,
// P = 8 the 60-bit limbs of CONFIG_N4096_K2_MP; both store 1 byte per digit,
// i.e. exactly the bytes of the u32 / u64 residues.
//
// MFMA v_mfma_i32_16x16x64_i8 operand layout (gfx950), lane l, byte j:
//   A (16 x 64): row = l % 16, k = 16 * (l / 16) + j
//   B (64 x 16): col = l % 16, k = 16 * (l / 16) + j
//   D (16 x 16): col = l % 16, row = 4 * (l / 16) + r   (r = 0..3)
//
// Database layout ("fragment-native"): for each (limb, level, row tile rt of
// 16 rows, k block kb of 64, digit plane s) one 1 KiB block holding the A
// fragment exactly as the 64 lanes load it (16 bytes per lane), so every
// wave load is one contiguous 1 KiB run.
//
// Two workgroup shapes, both 4 waves = 4 row tiles, one limb, a run of 16
// levels, results of LV consecutive levels staged in LDS and written as
// level-contiguous runs of the output ciphertexts:
//  - direct (HBM-bound, small batches): one column tile, each wave loads its
//    own A and B fragments, no barrier in the main loop;
//  - LDS-shared (MFMA-bound, large batches): CB column tiles; per level and k
//    block the tiles' B fragments are staged once in LDS (double-buffered)
//    and shared by the 4 waves, each of which multiplies its A fragments
//    (prefetched one k block ahead) into CB x (2P-1) accumulator tiles;
//  - register-blocked direct (large batches): each wave multiplies R row
//    tiles by C column tiles from its own registers, R * C * P^2 MFMAs per
//    R * P + C * P fragments loaded.
// Beside them, for the 29-bit limbs, a VALU form (k_scan_valu: the
// reference's u32 x u32 -> u64 multiply-accumulate on the same layouts), so
// that the autotuner measures where the matrix cores pay off.
// The shape is picked per batch size from a table (launch_scan), or timed on
// the live operands in autotune mode.
#include <type_traits>

#include "opir_k2.h"
#include "opir_kernels.h"

namespace opir {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef unsigned __int128 u128;

namespace {

template <int P>
__device__ __forceinline__ void balanced_digits(u64 a, int (&d)[P]) {
  int64_t x = (int64_t)a;  // a < 2^(8P-2)
#pragma unroll
  for (int s = 0; s < P - 1; s++) {
    const int r = (int)(((x + 128) & 255) - 128);
    d[s] = r;
    x = (x - r) >> 8;
  }
  d[P - 1] = (int)x;
}

// Digit s of the same representation on its own: a + 128 * (1 + 256 + ...
// + 256^(P-2)) has the base-256 digits d_s + 128 (s < P - 1) and the top
// digit d_{P-1}, since every d_s + 128 lies in [0, 256).
template <int P>
__device__ __forceinline__ int balanced_digit(u64 a, int s) {
  u64 bias = 0;
#pragma unroll
  for (int i = 0; i < P - 1; i++) bias |= (u64)128 << (8 * i);
  const u64 y = a + bias;
  return s < P - 1 ? (int)((y >> (8 * s)) & 255) - 128 : (int)(y >> (8 * (P - 1)));
}

__device__ __forceinline__ size_t db_byte_offset(const DbGeom &g, int limb, int lv, size_t row, size_t k, int s) {
  const size_t rt = row >> 4, lr = row & 15, kb = k >> 6, grp = (k & 63) >> 4, j = k & 15;
  const size_t blk = ((((size_t)limb * g.nlv + lv) * g.nrt + rt) * g.nkb + kb) * g.P + s;
  return blk * 1024 + (grp * 16 + lr) * 16 + j;
}

template <int P>
__device__ __forceinline__ void store_digits(int8_t *db, const DbGeom &g, int limb, int lv, size_t row, size_t k,
                                             u64 a) {
  int d[P];
  balanced_digits<P>(a, d);
#pragma unroll
  for (int s = 0; s < P; s++) db[db_byte_offset(g, limb, lv, row, k, s)] = (int8_t)d[s];
}

// tmp [limb][count][nlv] (pt-major, NTT form) -> fragment-native DB.
// Tile: 64 pts x 32 levels staged through LDS for coalesced reads.
template <int P, typename V>
__global__ __launch_bounds__(256) void k_db_scatter(const V *__restrict__ tmp, int8_t *__restrict__ db,
                                                    size_t first_pt, size_t count, DbGeom g) {
  __shared__ V tile[32][65];
  const int limb = blockIdx.z;
  const size_t pt0 = (size_t)blockIdx.x * 64;
  const int lv0 = blockIdx.y * 32;
  const V *src = tmp + (size_t)limb * count * g.nlv;
  for (int e = threadIdx.x; e < 64 * 32; e += 256) {
    const int pl = e / 32, ll = e % 32;
    const size_t pt = pt0 + pl;
    tile[ll][pl] = pt < count ? src[pt * g.nlv + lv0 + ll] : 0;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 64 * 32; e += 256) {
    const int ll = e / 64, pl = e % 64;
    const size_t pt = pt0 + pl;
    if (pt >= count) continue;
    const size_t gpt = first_pt + pt;
    store_digits<P>(db, g, limb, lv0 + ll, gpt / g.N0, gpt % g.N0, (u64)tile[ll][pl]);
  }
}

// Reference composite layout ref [limb][coeff][num_pt] (server.cpp:162-170).
__global__ void k_db_from_ref(const u32 *__restrict__ ref, int8_t *__restrict__ db, DbGeom g) {
  const size_t num_pt = g.N0 * g.Nrest;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 2 * (size_t)g.nlv * num_pt) return;
  const size_t pt = i % num_pt, lv = (i / num_pt) % g.nlv, limb = i / (num_pt * g.nlv);
  store_digits<4>(db, g, (int)limb, (int)lv, pt / g.N0, pt % g.N0, ref[i]);
}

// Query columns -> B fragments. qntt [B][N0][2 polys][nl limbs][nlv] (NTT form)
// -> qfrag [limb][level][ct][kb][plane][grp][n < ncm][16 bytes], ct = column
// tile of 16 query columns (c = 2*query + poly), ncm = min(16, 2B); the 16
// bytes of (grp, n) are digit `plane` of k = 64*kb + 16*grp + j, j < 16
// (zero for k >= N0; columns c >= 2B of the last tile are zero).
// Workgroup: one (limb, column tile, 16-k group, block of kQfLv levels). It
// stages the 16 columns x 16 k x kQfLv levels through LDS (read as 128-byte
// level runs of each column's k), then each thread writes all P planes of
// one (level, column): for a fixed (level, plane) the tile's columns are one
// contiguous run of ncm * 16 bytes.
template <typename V>
constexpr int qf_levels() { return 128 / (int)sizeof(V); }  // levels per workgroup
template <int P, typename V>
__global__ __launch_bounds__(256) void k_qfrag(const V *__restrict__ qntt, int8_t *__restrict__ qfrag,
                                               size_t q_stride, QfragGeom qg, int c_end) {
  constexpr int kQfLv = qf_levels<V>();
  constexpr int kQfNs = 16 * (kQfLv + 1) + 4;  // LDS pitch of a column (conflict-free reads)
  __shared__ V tile[16 * kQfNs];
  const int ct = blockIdx.x / (qg.nkb * 4), kg = blockIdx.x % (qg.nkb * 4);  // kg = kb * 4 + grp
  const int kb = kg >> 2, grp = kg & 3;
  const int lv0 = blockIdx.y * kQfLv, limb = blockIdx.z;
  const int ncols = qg.ncm;  // columns stored per tile; c_end = 2B columns in all
  // Only the stored columns are staged (the writes below read no others), as
  // 16-byte pieces of each (column, k) level run, all of a thread's loads in
  // flight before the first LDS store.
  constexpr int VL = 16 / (int)sizeof(V);    // levels per piece
  constexpr int NCH = kQfLv / VL;            // pieces per run
  constexpr int NIT = 16 * 16 * NCH / 256;   // pieces per thread (16 columns)
  uint4 piece[NIT];
#pragma unroll
  for (int i = 0; i < NIT; i++) {
    const int e = threadIdx.x + 256 * i, ch = e % NCH, row = e / NCH, n = row >> 4, j = row & 15;
    const int c = ct * 16 + n, k = kg * 16 + j;
    piece[i] = make_uint4(0, 0, 0, 0);
    if (n < ncols && c < c_end && k < qg.N0)
      piece[i] = *reinterpret_cast<const uint4 *>(
          qntt + (size_t)(c >> 1) * q_stride + (((size_t)k * 2 + (c & 1)) * qg.nl + limb) * qg.nlv + lv0 + ch * VL);
  }
#pragma unroll
  for (int i = 0; i < NIT; i++) {
    const int e = threadIdx.x + 256 * i, ch = e % NCH, row = e / NCH, n = row >> 4, j = row & 15;
    if (n >= ncols) continue;
    V x[VL];
    __builtin_memcpy(x, &piece[i], 16);
#pragma unroll
    for (int v = 0; v < VL; v++) tile[n * kQfNs + j * (kQfLv + 1) + ch * VL + v] = x[v];
  }
  __syncthreads();
  const size_t blk = (((size_t)limb * qg.nlv + lv0) * qg.nct + ct) * qg.nkb + kb;  // level lv0
  const size_t lvstep = (size_t)qg.nct * qg.nkb;  // blocks between levels
  const size_t bsz = (size_t)P * 4 * ncols * 16;   // bytes of one block
  for (int e = threadIdx.x; e < kQfLv * ncols; e += 256) {
    const int n = e % ncols, lv = e / ncols;
    V x[16];
#pragma unroll
    for (int j = 0; j < 16; j++) x[j] = tile[n * kQfNs + j * (kQfLv + 1) + lv];
    int8_t *o = qfrag + (blk + lv * lvstep) * bsz + ((size_t)grp * ncols + n) * 16;
#pragma unroll
    for (int s = 0; s < P; s++) {
      u32 w[4];
#pragma unroll
      for (int q4 = 0; q4 < 4; q4++) {
        u32 packed = 0;
#pragma unroll
        for (int b = 0; b < 4; b++)
          packed |= ((u32)(uint8_t)(int8_t)balanced_digit<P>((u64)x[q4 * 4 + b], s)) << (8 * b);
        w[q4] = packed;
      }
      *reinterpret_cast<uint4 *>(o + (size_t)s * 4 * ncols * 16) = make_uint4(w[0], w[1], w[2], w[3]);
    }
  }
}

// ---------------------------------------------------------------------------
// Epilogues: sum_u (256^u mod q) G_u  ->  residue mod q
// ---------------------------------------------------------------------------
// The weights are held as int32 so that each term is one signed 32 x 32 ->
// 64-bit multiply-add (v_mad_i64_i32).
struct EpiK1 {
  static constexpr int P = 4;
  typedef u32 Res;
  int wu[2 * P - 1];  // 256^u mod q < 2^29
  u64 pos_off;  // q * ceil(2^59 / q): makes the int64 sum positive (|sum| < 2^56)
  const LimbConst *L;
  __device__ void init(const SchemeConst &S, int limb) {
    L = &S.limb[limb];
    u64 x = 1;
#pragma unroll
    for (int u = 0; u < 2 * P - 1; u++) {
      // readfirstlane (the value is uniform) hides the small known range of the
      // first weights, which would otherwise select an unsigned multiply with
      // a separate sign correction
      wu[u] = __builtin_amdgcn_readfirstlane((int)x);
      x = (x << 8) % L->q;
    }
    pos_off = ((((u64)1 << 59) / L->q) + 1) * L->q;
  }
  __device__ Res combine(const int (&g)[2 * P - 1]) const {
    int64_t v = 0;
#pragma unroll
    for (int u = 0; u < 2 * P - 1; u++) v += (int64_t)wu[u] * (int64_t)g[u];
    return reduce_u64((u64)(v + (int64_t)pos_off), *L);
  }
};

struct EpiK2 {
  static constexpr int P = 8;
  typedef u64 Res;
  int wlo[2 * P - 1], whi[2 * P - 1];  // 256^u mod q = whi * 2^30 + wlo, both < 2^30
  u64 off_lo, off_hi;                  // q * ceil(2^92 / q)
  const k2::Limb64 *L;
  __device__ void init(const k2::Scheme64 &S, int limb) {
    L = &S.limb[limb];
    const u64 q = L->q;
    u64 x = 1;
#pragma unroll
    for (int u = 0; u < 2 * P - 1; u++) {
      wlo[u] = __builtin_amdgcn_readfirstlane((int)(x & ((1u << 30) - 1)));  // as in EpiK1
      whi[u] = __builtin_amdgcn_readfirstlane((int)(x >> 30));
      x = (u64)(((u128)x << 8) % q);
    }
    const u128 off = ((((u128)1 << 92) / q) + 1) * (u128)q;
    off_lo = (u64)off;
    off_hi = (u64)(off >> 64);
  }
  __device__ Res combine(const int (&g)[2 * P - 1]) const {
    // |S_lo|, |S_hi| < 15 * 2^30 * 2^26 < 2^60; total < 2^91
    int64_t slo = 0, shi = 0;
#pragma unroll
    for (int u = 0; u < 2 * P - 1; u++) {
      slo += (int64_t)wlo[u] * (int64_t)g[u];
      shi += (int64_t)whi[u] * (int64_t)g[u];
    }
    const __int128 tot = (__int128)slo + ((__int128)shi << 30) + (__int128)(((u128)off_hi << 64) | off_lo);
    const u128 t = (u128)tot;
    const u64 hi = (u64)(t >> 64), lo = (u64)t;
    // (hi * 2^64 + lo) mod q, both halves by Shoup
    const u64 a = hi * L->r64 - __umul64hi(hi, L->r64s) * L->q;
    const u64 b = lo - __umul64hi(lo, L->one_s) * L->q;
    u64 r = a + b;
    r = r >= L->q2 ? r - L->q2 : r;
    return r >= L->q ? r - L->q : r;
  }
};

template <typename Epi>
struct ScanArgs {
  const int8_t *db;
  const int8_t *qfrag;
  typename Epi::Res *mid;
  size_t mid_stride;  // values between queries
  size_t ct_vals;     // values per output ciphertext (2 polys x nl limbs x nlv)
  DbGeom g;
  QfragGeom qg;
  int nrb, ncb, B;
};

constexpr int kLevelsPerWG = 16;

// Write one wave's staged results: for each (row rr < 16, column nn < ncols)
// of its tile, the LV consecutive levels st[0..LV)[rr][nn] go to
// mid[q][row][poly][limb][lvo ..] (ct_vals / 2 values per poly), c = col0 + nn = 2q + poly,
// row = 16 rt + rr.
// The stores are nontemporal: the next kernel reads the results, and
// keeping them out of the L2 leaves it to the database stream (the plain
// scan ran 0.707 against 0.734 ms with cached stores).
// With W = 16 / sizeof(Res) dividing LV, as 16-byte stores, consecutive
// lanes taking consecutive pieces of a run (its index, a multiple of W,
// makes each piece 16-byte aligned); rows past Nrest are not stored.
template <int LV, bool NTS = true, int SW, typename Res>
__device__ __forceinline__ void flush_stage(const Res (*st)[16][SW], int lane, int ncols, int rt, int col0, int limb,
                                            int lvo, Res *mid, size_t mid_stride, size_t ct_vals, int nlv,
                                            size_t Nrest) {
  constexpr int W = 16 / (int)sizeof(Res);
  constexpr int CH = LV % W == 0 ? LV / W : 0;  // 16-byte pieces per run
  if constexpr (CH > 0) {
    for (int e = lane; e < 16 * ncols * CH; e += 64) {
      const int run = e / CH, ch = e - run * CH;
      const int rr = run / ncols, nn = run % ncols;
      const size_t row = (size_t)rt * 16 + rr;
      if (row >= Nrest) continue;
      const int c = col0 + nn, qi = c >> 1, pp = c & 1;
      Res v[W];
#pragma unroll
      for (int w = 0; w < W; w++) v[w] = st[ch * W + w][rr][nn];
      v4i x;
      __builtin_memcpy(&x, v, 16);
      // nontemporal: the results are read by the next kernel, not by this one
      v4i *dst = reinterpret_cast<v4i *>(mid + qi * mid_stride + row * ct_vals + (size_t)pp * (ct_vals >> 1) +
                                         (size_t)limb * nlv + lvo + ch * W);
      if (NTS)
        __builtin_nontemporal_store(x, dst);
      else
        *dst = x;
    }
  } else {
    for (int e = lane; e < 16 * ncols; e += 64) {
      const int rr = e / ncols, nn = e % ncols;
      const size_t row = (size_t)rt * 16 + rr;
      if (row >= Nrest) continue;
      const int c = col0 + nn, qi = c >> 1, pp = c & 1;
      Res *dst = mid + qi * mid_stride + row * ct_vals + (size_t)pp * (ct_vals >> 1) + (size_t)limb * nlv + lvo;
#pragma unroll
      for (int li = 0; li < LV; li++) {
        if (NTS)
          __builtin_nontemporal_store(st[li][rr][nn], dst + li);
        else
          dst[li] = st[li][rr][nn];
      }
    }
  }
}

// NT: nontemporal A loads, for grids where each DB tile has one reader
// (one column block); with several, the tile's later readers find it in L2.
template <int CB, int LV, bool NT, typename Epi, typename SchemeT, int LW = kLevelsPerWG, bool NTS = true>
__global__ __launch_bounds__(256) void k_scan(ScanArgs<Epi> A, SchemeT S) {
  constexpr int P = Epi::P, NU = 2 * P - 1;
  constexpr int BCHUNK = CB * P * 1024;  // bytes of one k block's B fragments
  constexpr int BV = BCHUNK / 16 / 256;  // uint4 per thread to stage them
  typedef typename Epi::Res Res;
  __shared__ __align__(16) int8_t bsh[2][BCHUNK];
  __shared__ Res stage[4][LV][16][CB * 16 + 1];
  const DbGeom &g = A.g;
  const QfragGeom &qg = A.qg;
  const unsigned W = gridDim.x, pblk = blockIdx.x;
  const unsigned Lg = (W % 8 == 0) ? (pblk % 8) * (W / 8) + pblk / 8 : pblk;  // same-XCD neighbours
  // column blocks innermost: the workgroups sharing a DB tile run together
  const int cbk = Lg % A.ncb;
  unsigned rest = Lg / A.ncb;
  const int rb = rest % A.nrb;
  rest /= A.nrb;
  const int nlvb = g.nlv / LW;
  const int lvb = rest % nlvb;
  const int limb = rest / nlvb;
  Epi epi;
  epi.init(S, limb);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int rt = rb * 4 + wave;
  const bool row_ok = rt < (int)g.nrt;
  const int ct0 = cbk * CB;
  const size_t qtile = (size_t)(P * 4 * qg.ncm * 16);

  // cooperative load of one k block's B fragments (CB tiles, all 16 columns;
  // columns a partial last tile does not store are zero)
  auto load_b = [&](int lv, int kb, uint4 (&r)[BV]) {
#pragma unroll
    for (int i = 0; i < BV; i++) {
      const int e = threadIdx.x + 256 * i;  // 16-byte unit
      const int cb = e / (P * 64), rem = e % (P * 64);
      const int s = rem / 64, grp = (rem / 16) % 4, n = rem % 16;
      const int ct = ct0 + cb;
      r[i] = make_uint4(0, 0, 0, 0);
      if (ct < qg.nct && n < qg.ncm)
        r[i] = *reinterpret_cast<const uint4 *>(A.qfrag +
                                                ((((size_t)limb * g.nlv + lv) * qg.nct + ct) * qg.nkb + kb) * qtile +
                                                ((size_t)(s * 4 + grp) * qg.ncm + n) * 16);
    }
  };
  auto store_b = [&](int buf, const uint4 (&r)[BV]) {
#pragma unroll
    for (int i = 0; i < BV; i++) reinterpret_cast<uint4 *>(bsh[buf])[threadIdx.x + 256 * i] = r[i];
  };
  auto load_a = [&](int lv, int kb, v4i (&a)[P]) {
    const int8_t *p =
        A.db + ((((size_t)limb * g.nlv + lv) * g.nrt + rt) * g.nkb + kb) * (size_t)(P * 1024) + lane * 16;
#pragma unroll
    for (int s = 0; s < P; s++) {
      const v4i *q = reinterpret_cast<const v4i *>(p + s * 1024);
      a[s] = row_ok ? (NT ? __builtin_nontemporal_load(q) : *q) : (v4i){0, 0, 0, 0};
    }
  };

  for (int lsub = 0; lsub < LW; lsub += LV) {
#pragma unroll 1
    for (int li = 0; li < LV; li++) {
      const int lv = lvb * LW + lsub + li;
      v4i acc[CB][NU];
#pragma unroll
      for (int cb = 0; cb < CB; cb++)
#pragma unroll
        for (int u = 0; u < NU; u++) acc[cb][u] = (v4i){0, 0, 0, 0};
      uint4 breg[BV];
      v4i a_cur[P], a_nxt[P];
      load_b(lv, 0, breg);
      load_a(lv, 0, a_cur);
      __syncthreads();  // the previous level's readers of bsh[0] are done
      store_b(0, breg);
#pragma unroll 1
      for (int kb = 0; kb < (int)g.nkb; kb++) {
        __syncthreads();  // B(kb) is visible to all waves
        const bool more = kb + 1 < (int)g.nkb;
        if (more) {
          load_b(lv, kb + 1, breg);
          load_a(lv, kb + 1, a_nxt);
        }
        const int8_t *bb = bsh[kb & 1];
#pragma unroll
        for (int cb = 0; cb < CB; cb++) {
          v4i bf[P];
#pragma unroll
          for (int t = 0; t < P; t++) bf[t] = *reinterpret_cast<const v4i *>(bb + ((cb * P + t) * 64 + lane) * 16);
#pragma unroll
          for (int s = 0; s < P; s++)
#pragma unroll
            for (int t = 0; t < P; t++)
              acc[cb][s + t] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a_cur[s], bf[t], acc[cb][s + t], 0, 0, 0);
        }
        if (more) {
          // the other buffer: its last readers (k block kb-1) passed this k block's barrier
          store_b((kb + 1) & 1, breg);
#pragma unroll
          for (int s = 0; s < P; s++) a_cur[s] = a_nxt[s];
        }
      }
      // epilogue: D row = 4 * (lane / 16) + r, column = lane % 16
#pragma unroll
      for (int cb = 0; cb < CB; cb++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          int gs[NU];
#pragma unroll
          for (int u = 0; u < NU; u++) gs[u] = acc[cb][u][r];
          stage[wave][li][(lane >> 4) * 4 + r][cb * 16 + (lane & 15)] = epi.combine(gs);
        }
    }
    __syncthreads();
    // (row, column) pairs of this wave's row tile: LV consecutive levels each
    flush_stage<LV, NTS>(stage[wave], lane, min(CB * 16, 2 * A.B - ct0 * 16), rt, ct0 * 16, limb,
                         lvb * LW + lsub, A.mid, A.mid_stride, A.ct_vals, g.nlv, g.Nrest);
  }
}

// Direct variant: one column tile per workgroup, no LDS for B and no barrier
// in the main loop. Each wave streams its own row tile: per level it issues
// the A fragments of KC k blocks at once (nontemporal when the tile has no
// other reader: the database then passes through once and should not
// displace the query fragments in L2), with the tile's B fragments (small,
// L2-resident), then runs their MFMAs. Lanes
// past the stored query columns load column 0 (their result columns are
// never stored), so every load is unconditional and the compiler's in-order
// load counter lets the MFMAs of the first k block start before the last
// load has landed.
// SC: result columns staged per row (>= stored columns of the tile: 2 for
// one query). For the HBM-bound small batches.
// XPF: load the next batch of k blocks right after the current batch's MFMAs
// (across level boundaries, so a level's epilogue overlaps the next level's
// loads; the registers stay live through the epilogue). Otherwise each batch
// is loaded at the top of its own iteration.
template <int KC, int LV, int SC, bool NT, bool XPF, typename Epi, typename SchemeT, bool NTS = true,
          int LW = kLevelsPerWG>
__global__ __launch_bounds__(256) void k_scan_direct(ScanArgs<Epi> A, SchemeT S) {
  constexpr int P = Epi::P, NU = 2 * P - 1;
  typedef typename Epi::Res Res;
  __shared__ Res stage[4][LV][16][SC + 1];
  const DbGeom &g = A.g;
  const QfragGeom &qg = A.qg;
  const unsigned W = gridDim.x, pblk = blockIdx.x;
  const unsigned Lg = (W % 8 == 0) ? (pblk % 8) * (W / 8) + pblk / 8 : pblk;
  const int ct = Lg % A.ncb;
  unsigned rest = Lg / A.ncb;
  const int rb = rest % A.nrb;
  rest /= A.nrb;
  const int nlvb = g.nlv / LW;
  const int lvb = rest % nlvb;
  const int limb = rest / nlvb;
  Epi epi;
  epi.init(S, limb);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int rt = rb * 4 + wave;
  // a row tile past the last one re-reads the last tile; its rows are not stored
  const size_t rta = rt < (int)g.nrt ? (size_t)rt : g.nrt - 1;
  const int n = lane & 15, grp = lane >> 4;
  const bool has_col = n < qg.ncm;
  const size_t qtile = (size_t)(P * 4 * qg.ncm * 16);
  const int nkb = (int)g.nkb;
  const int lv0 = lvb * LW;
  const int8_t *abase = A.db + (((size_t)limb * g.nlv + lv0) * g.nrt + rta) * g.nkb * (size_t)(P * 1024) + lane * 16;
  const size_t astride = (size_t)g.nrt * g.nkb * (P * 1024);  // bytes between levels
  const int8_t *bbase = A.qfrag + (((size_t)limb * g.nlv + lv0) * qg.nct + ct) * qg.nkb * qtile +
                        ((size_t)grp * qg.ncm + (has_col ? n : 0)) * 16;
  const size_t bstride = (size_t)qg.nct * qg.nkb * qtile;

  // one register buffer for a batch of KC k blocks
  const int nch = nkb / KC;
  v4i a[KC][P], b[KC][P];
  auto load = [&](int lrel, int ch) {
    const int8_t *pa = abase + (size_t)lrel * astride + (size_t)ch * KC * (P * 1024);
    const int8_t *pb = bbase + (size_t)lrel * bstride + (size_t)ch * KC * qtile;
#pragma unroll
    for (int j = 0; j < KC; j++)
#pragma unroll
      for (int s = 0; s < P; s++) {
        const v4i *q = reinterpret_cast<const v4i *>(pa + (size_t)j * (P * 1024) + s * 1024);
        a[j][s] = NT ? __builtin_nontemporal_load(q) : *q;
      }
#pragma unroll
    for (int j = 0; j < KC; j++)
#pragma unroll
      for (int t = 0; t < P; t++)
        b[j][t] = *reinterpret_cast<const v4i *>(pb + (size_t)j * qtile + (size_t)t * 4 * qg.ncm * 16);
  };
  if (XPF) load(0, 0);
  for (int lsub = 0; lsub < LW; lsub += LV) {
#pragma unroll 1
    for (int li = 0; li < LV; li++) {
      v4i acc[NU];
#pragma unroll
      for (int u = 0; u < NU; u++) acc[u] = (v4i){0, 0, 0, 0};
#pragma unroll 1
      for (int ch = 0; ch < nch; ch++) {
        if (!XPF) load(lsub + li, ch);
#pragma unroll
        for (int j = 0; j < KC; j++)
#pragma unroll
          for (int s = 0; s < P; s++)
#pragma unroll
            for (int t = 0; t < P; t++)
              acc[s + t] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[j][s], b[j][t], acc[s + t], 0, 0, 0);
        if (XPF) {
          const bool last = ch + 1 == nch;
          const int nl = lsub + li + (last ? 1 : 0);
          if (nl < LW) load(nl, last ? 0 : ch + 1);
        }
      }
#pragma unroll
      for (int r = 0; r < 4; r++) {
        int gs[NU];
#pragma unroll
        for (int u = 0; u < NU; u++) gs[u] = acc[u][r];
        const Res v = epi.combine(gs);
        if (n < SC) stage[wave][li][grp * 4 + r][n] = v;
      }
    }
    __syncthreads();
    flush_stage<LV, NTS>(stage[wave], lane, min(16, 2 * A.B - ct * 16), rt, ct * 16, limb, lv0 + lsub, A.mid,
                    A.mid_stride, A.ct_vals, g.nlv, g.Nrest);
    __syncthreads();
  }
}

// Ping-pong variant of the direct shape: two register buffers of KC k blocks
// each. Step k of the workgroup's (level, batch) sequence first issues the
// loads of step k + 1 into the other buffer, then runs its MFMAs, so the
// wave's next loads are in flight while it multiplies and runs a level's
// epilogue (the direct shape with XPF issues them only after its MFMAs).
template <int KC, int LV, int SC, bool NT, typename Epi, typename SchemeT, bool NTS = true, int LW = kLevelsPerWG>
__global__ __launch_bounds__(256) void k_scan_pp(ScanArgs<Epi> A, SchemeT S) {
  constexpr int P = Epi::P, NU = 2 * P - 1;
  typedef typename Epi::Res Res;
  __shared__ Res stage[4][LV][16][SC + 1];
  const DbGeom &g = A.g;
  const QfragGeom &qg = A.qg;
  const unsigned W = gridDim.x, pblk = blockIdx.x;
  const unsigned Lg = (W % 8 == 0) ? (pblk % 8) * (W / 8) + pblk / 8 : pblk;
  const int ct = Lg % A.ncb;
  unsigned rest = Lg / A.ncb;
  const int rb = rest % A.nrb;
  rest /= A.nrb;
  const int nlvb = g.nlv / LW;
  const int lvb = rest % nlvb;
  const int limb = rest / nlvb;
  Epi epi;
  epi.init(S, limb);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int rt = rb * 4 + wave;
  const size_t rta = rt < (int)g.nrt ? (size_t)rt : g.nrt - 1;
  const int n = lane & 15, grp = lane >> 4;
  const bool has_col = n < qg.ncm;
  const size_t qtile = (size_t)(P * 4 * qg.ncm * 16);
  const int nkb = (int)g.nkb;
  const int lv0 = lvb * LW;
  const int8_t *abase = A.db + (((size_t)limb * g.nlv + lv0) * g.nrt + rta) * g.nkb * (size_t)(P * 1024) + lane * 16;
  const size_t astride = (size_t)g.nrt * g.nkb * (P * 1024);
  const int8_t *bbase = A.qfrag + (((size_t)limb * g.nlv + lv0) * qg.nct + ct) * qg.nkb * qtile +
                        ((size_t)grp * qg.ncm + (has_col ? n : 0)) * 16;
  const size_t bstride = (size_t)qg.nct * qg.nkb * qtile;
  const int nch = nkb / KC;                 // batches per level
  const int nsteps = LW * nch;              // even: LW is
  typedef v4i Buf[KC][P];
  auto load = [&](int k, Buf &a, Buf &b) {
    const int lrel = k / nch, ch = k - lrel * nch;
    const int8_t *pa = abase + (size_t)lrel * astride + (size_t)ch * KC * (P * 1024);
    const int8_t *pb = bbase + (size_t)lrel * bstride + (size_t)ch * KC * qtile;
#pragma unroll
    for (int j = 0; j < KC; j++)
#pragma unroll
      for (int s = 0; s < P; s++) {
        const v4i *q = reinterpret_cast<const v4i *>(pa + (size_t)j * (P * 1024) + s * 1024);
        a[j][s] = NT ? __builtin_nontemporal_load(q) : *q;
      }
#pragma unroll
    for (int j = 0; j < KC; j++)
#pragma unroll
      for (int t = 0; t < P; t++)
        b[j][t] = *reinterpret_cast<const v4i *>(pb + (size_t)j * qtile + (size_t)t * 4 * qg.ncm * 16);
  };
  const int ncols = min(16, 2 * A.B - ct * 16);
  v4i acc[NU];
#pragma unroll
  for (int u = 0; u < NU; u++) acc[u] = (v4i){0, 0, 0, 0};
  // loads are unconditional inside the loop (the last step is peeled), so
  // the compiler's load counter waits only for the step's own buffer
  auto step = [&](auto do_load, int k, const Buf &a, const Buf &b, Buf &an, Buf &bn) {
    if constexpr (decltype(do_load)::value) load(k + 1, an, bn);
#pragma unroll
    for (int j = 0; j < KC; j++)
#pragma unroll
      for (int s = 0; s < P; s++)
#pragma unroll
        for (int t = 0; t < P; t++)
          acc[s + t] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[j][s], b[j][t], acc[s + t], 0, 0, 0);
    const int lrel = k / nch;
    if (k - lrel * nch != nch - 1) return;
    // last batch of level lrel: epilogue, then a flush every LV levels
    const int li = lrel % LV;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      int gs[NU];
#pragma unroll
      for (int u = 0; u < NU; u++) gs[u] = acc[u][r];
      const Res v = epi.combine(gs);
      if (n < SC) stage[wave][li][grp * 4 + r][n] = v;
    }
#pragma unroll
    for (int u = 0; u < NU; u++) acc[u] = (v4i){0, 0, 0, 0};
    if (li != LV - 1) return;
    const int lsub = lrel - li;
    __syncthreads();
    flush_stage<LV, NTS>(stage[wave], lane, ncols, rt, ct * 16, limb, lv0 + lsub, A.mid, A.mid_stride, A.ct_vals,
                         g.nlv, g.Nrest);
    __syncthreads();
  };
  Buf a0, b0, a1, b1;
  load(0, a0, b0);
  const std::true_type next;
  const std::false_type last;
#pragma unroll 1
  for (int k = 0; k < nsteps - 2; k += 2) {
    step(next, k, a0, b0, a1, b1);
    step(next, k + 1, a1, b1, a0, b0);
  }
  step(next, nsteps - 2, a0, b0, a1, b1);
  step(last, nsteps - 1, a1, b1, a0, b0);
}

// Register-blocked direct variant for large batches: each wave multiplies R
// consecutive row tiles by C consecutive column tiles, so one k block's
// loads (R * P A fragments, C * P B fragments) feed R * C * P^2 MFMAs
// instead of P^2 (1 KiB per MFMA in the direct shape, 256 B at R = C = 2).
// The 4 waves form a (4 / WC) x WC grid over row and column blocks: a
// workgroup covers 4R / WC row tiles and WC * C column tiles, and waves of
// one row (column) block meet in L1 on the same A (B) fragments. The
// workgroups of one row block (column blocks innermost) are neighbours on
// one XCD, so the database tile passes through HBM once and is re-read from
// L2. Row or column tiles past the last one re-read the last tile and store
// nothing. Loads of the next (level, batch) are issued right after the
// current batch's MFMAs. Each workgroup takes LW consecutive levels, staged
// LV at a time: with cached stores (NTS false) the 4 * LV-byte pieces of one
// output run meet in L2 and leave it as whole lines.
template <int R, int C, int WC, int KC, int LV, int LW, bool NT, typename Epi, typename SchemeT, bool NTS>
__global__ __launch_bounds__(256) void k_scan_rb(ScanArgs<Epi> A, SchemeT S) {
  constexpr int P = Epi::P, NU = 2 * P - 1;
  constexpr int WR = 4 / WC;
  static_assert(WR * WC == 4 && LW % LV == 0, "4 waves; whole flushes");
  typedef typename Epi::Res Res;
  __shared__ Res stage[4][R][LV][16][16 * C + 1];
  const DbGeom &g = A.g;
  const QfragGeom &qg = A.qg;
  const unsigned W = gridDim.x, pblk = blockIdx.x;
  const unsigned Lg = (W % 8 == 0) ? (pblk % 8) * (W / 8) + pblk / 8 : pblk;
  const int cbk = Lg % A.ncb;
  unsigned rest = Lg / A.ncb;
  const int rb = rest % A.nrb;
  rest /= A.nrb;
  const int nlvb = g.nlv / LW;
  const int lvb = rest % nlvb;
  const int limb = rest / nlvb;
  Epi epi;
  epi.init(S, limb);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int rt0 = (rb * WR + wave / WC) * R;  // first row tile of this wave
  const int ct0 = (cbk * WC + wave % WC) * C;  // first column tile of this wave
  const int n = lane & 15, grp = lane >> 4;
  const bool has_col = n < qg.ncm;
  const size_t qtile = (size_t)(P * 4 * qg.ncm * 16);
  const int lv0 = lvb * LW;
  const size_t astride = (size_t)g.nrt * g.nkb * (P * 1024);  // bytes between levels
  const size_t bstride = (size_t)qg.nct * qg.nkb * qtile;
  const int8_t *abase[R];
  const int8_t *bbase[C];
#pragma unroll
  for (int r = 0; r < R; r++) {
    const size_t rta = rt0 + r < (int)g.nrt ? (size_t)(rt0 + r) : g.nrt - 1;
    abase[r] = A.db + (((size_t)limb * g.nlv + lv0) * g.nrt + rta) * g.nkb * (size_t)(P * 1024) + lane * 16;
  }
#pragma unroll
  for (int c = 0; c < C; c++) {
    const size_t cta = ct0 + c < qg.nct ? (size_t)(ct0 + c) : (size_t)qg.nct - 1;
    bbase[c] = A.qfrag + (((size_t)limb * g.nlv + lv0) * qg.nct + cta) * qg.nkb * qtile +
               ((size_t)grp * qg.ncm + (has_col ? n : 0)) * 16;
  }
  const int nch = (int)g.nkb / KC;
  v4i a[KC][R][P], b[KC][C][P];
  auto load = [&](int lrel, int ch) {
#pragma unroll
    for (int j = 0; j < KC; j++)
#pragma unroll
      for (int r = 0; r < R; r++)
#pragma unroll
        for (int s = 0; s < P; s++) {
          const v4i *q = reinterpret_cast<const v4i *>(abase[r] + (size_t)lrel * astride +
                                                       (size_t)(ch * KC + j) * (P * 1024) + s * 1024);
          a[j][r][s] = NT ? __builtin_nontemporal_load(q) : *q;
        }
#pragma unroll
    for (int j = 0; j < KC; j++)
#pragma unroll
      for (int c = 0; c < C; c++)
#pragma unroll
        for (int t = 0; t < P; t++)
          b[j][c][t] = *reinterpret_cast<const v4i *>(bbase[c] + (size_t)lrel * bstride +
                                                      (size_t)(ch * KC + j) * qtile + (size_t)t * 4 * qg.ncm * 16);
  };
  load(0, 0);
  for (int lsub = 0; lsub < LW; lsub += LV) {
#pragma unroll 1
    for (int li = 0; li < LV; li++) {
      v4i acc[R][C][NU];
#pragma unroll
      for (int r = 0; r < R; r++)
#pragma unroll
        for (int c = 0; c < C; c++)
#pragma unroll
          for (int u = 0; u < NU; u++) acc[r][c][u] = (v4i){0, 0, 0, 0};
#pragma unroll 1
      for (int ch = 0; ch < nch; ch++) {
#pragma unroll
        for (int j = 0; j < KC; j++)
#pragma unroll
          for (int r = 0; r < R; r++)
#pragma unroll
            for (int c = 0; c < C; c++)
#pragma unroll
              for (int s = 0; s < P; s++)
#pragma unroll
                for (int t = 0; t < P; t++)
                  acc[r][c][s + t] =
                      __builtin_amdgcn_mfma_i32_16x16x64_i8(a[j][r][s], b[j][c][t], acc[r][c][s + t], 0, 0, 0);
        const bool last = ch + 1 == nch;
        const int nl = lsub + li + (last ? 1 : 0);
        if (nl < LW) load(nl, last ? 0 : ch + 1);
      }
      // D row = 4 * (lane / 16) + q, column = lane % 16
#pragma unroll
      for (int r = 0; r < R; r++)
#pragma unroll
        for (int c = 0; c < C; c++)
#pragma unroll
          for (int q = 0; q < 4; q++) {
            int gs[NU];
#pragma unroll
            for (int u = 0; u < NU; u++) gs[u] = acc[r][c][u][q];
            stage[wave][r][li][grp * 4 + q][c * 16 + n] = epi.combine(gs);
          }
    }
    __syncthreads();
    const int ncols = min(C * 16, 2 * A.B - ct0 * 16);  // <= 0: a column block past the last
#pragma unroll
    for (int r = 0; r < R; r++)
      flush_stage<LV, NTS>(stage[wave][r], lane, ncols, rt0 + r, ct0 * 16, limb, lv0 + lsub, A.mid, A.mid_stride,
                      A.ct_vals, g.nlv, g.Nrest);
    __syncthreads();
  }
}

// Residues of 4 consecutive k from one 32-bit word of each digit plane
// (P = 4): y = a + 0x808080 has the base-256 digits d_s + 128 = byte ^ 0x80
// (s < 3) and the top digit d_3 (balanced_digit's identity), so the bytes of
// the four words regroup into y with one flip per word.
__device__ __forceinline__ void residues4(u32 w0, u32 w1, u32 w2, u32 w3, u32 (&a)[4]) {
  w0 ^= 0x80808080u;
  w1 ^= 0x80808080u;
  w2 ^= 0x80808080u;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const u32 y = ((w0 >> (8 * j)) & 255u) | (((w1 >> (8 * j)) & 255u) << 8) | (((w2 >> (8 * j)) & 255u) << 16) |
                  (((w3 >> (8 * j)) & 255u) << 24);
    a[j] = y - 0x808080u;
  }
}

// VALU form of the first-dimension scan for the 29-bit limbs: the
// reference's u32 x u32 -> u64 multiply-accumulate (matrix.cpp:24-151), one
// thread per candidate row, on the same database and query-fragment layouts
// as the MFMA shapes (residues rebuilt from their digit planes). The
// autotuner times it beside the MFMA shapes, so the matrix cores are used
// only for the batch sizes where they are faster.
// Workgroup: 256 rows x C query columns x kLevelsPerWG levels of one limb.
// Per level the C query columns are rebuilt into LDS (Q[k][c], read as
// broadcasts); each thread accumulates its row's C sums in u64, folding the
// high half back every 8 products, and keeps the C x 16 results in registers
// until they are written as 64-byte level runs.
// Bounds: residues a, b < 2^30, so each product is < 2^60; a fold leaves
// (x >> 32) * (2^32 mod q) + (x mod 2^32) < 2^61 + 2^32, and 8 more products
// keep the sum below 2^64. Columns past the batch's 2B read as zero and are
// not stored; rows past Nrest re-read the last row and are not stored.
template <int C>
__global__ __launch_bounds__(256) void k_scan_valu(ScanArgs<EpiK1> A, SchemeConst S) {
  constexpr int P = 4, LW = kLevelsPerWG;
  static_assert(C == 2 || C % 4 == 0, "query columns read as uint2 or uint4");
  static_assert(16 % C == 0, "a column block lies inside one column tile");
  extern __shared__ u32 qs[];  // [nkb * 64][C]
  const DbGeom &g = A.g;
  const QfragGeom &qg = A.qg;
  const unsigned W = gridDim.x, pblk = blockIdx.x;
  const unsigned Lg = (W % 8 == 0) ? (pblk % 8) * (W / 8) + pblk / 8 : pblk;
  const int cb = Lg % A.ncb;
  unsigned rest = Lg / A.ncb;
  const int rb = rest % A.nrb;
  rest /= A.nrb;
  const int nlvb = g.nlv / LW;
  const int lvb = rest % nlvb;
  const int limb = rest / nlvb;
  const LimbConst &L = S.limb[limb];
  const int t = threadIdx.x;
  const size_t row = (size_t)rb * 256 + t;
  const size_t rowa = row < g.nrt * 16 ? row : g.nrt * 16 - 1;
  const size_t rt = rowa >> 4;
  const int lr = (int)(rowa & 15);
  const int c0 = cb * C, ct = c0 >> 4, n0 = c0 & 15;
  const int nkb = (int)g.nkb;
  const size_t qtile = (size_t)(P * 4 * qg.ncm * 16);
  const size_t astride = (size_t)g.nrt * g.nkb * (P * 1024);  // bytes between levels
  const int8_t *abase = A.db + (((size_t)limb * g.nlv + (size_t)lvb * LW) * g.nrt + rt) * g.nkb * (size_t)(P * 1024);
  u32 res[LW][C];
  // unrolled so that res is indexed statically and stays in registers (each
  // level's body still loops over its k chunks)
#pragma unroll
  for (int li = 0; li < LW; li++) {
    const int lv = lvb * LW + li;
    __syncthreads();  // the previous level's reads of qs are done
    // rebuild Q[k][c] of this level: unit (kb, grp, cc) = 16 consecutive k of one column
    const int8_t *qb = A.qfrag + (((size_t)limb * g.nlv + lv) * qg.nct + ct) * qg.nkb * qtile;
    for (int u = t; u < nkb * 4 * C; u += 256) {
      const int cc = u % C, grp = (u / C) & 3, kb = u / (4 * C);
      const int n = n0 + cc;
      u32 v[16];
      if (n < qg.ncm) {
        const int8_t *p = qb + (size_t)kb * qtile + ((size_t)grp * qg.ncm + n) * 16;
        uint4 d[P];
#pragma unroll
        for (int s = 0; s < P; s++) d[s] = *reinterpret_cast<const uint4 *>(p + (size_t)s * 4 * qg.ncm * 16);
        const u32 w[P][4] = {{d[0].x, d[0].y, d[0].z, d[0].w}, {d[1].x, d[1].y, d[1].z, d[1].w},
                             {d[2].x, d[2].y, d[2].z, d[2].w}, {d[3].x, d[3].y, d[3].z, d[3].w}};
#pragma unroll
        for (int i = 0; i < 4; i++) {
          u32 a4[4];
          residues4(w[0][i], w[1][i], w[2][i], w[3][i], a4);
#pragma unroll
          for (int j = 0; j < 4; j++) v[4 * i + j] = a4[j];
        }
      } else {
#pragma unroll
        for (int j = 0; j < 16; j++) v[j] = 0;
      }
#pragma unroll
      for (int j = 0; j < 16; j++) qs[(kb * 64 + grp * 16 + j) * C + cc] = v[j];
    }
    __syncthreads();
    u64 acc[C];
#pragma unroll
    for (int c = 0; c < C; c++) acc[c] = 0;
    const int8_t *pa = abase + (size_t)li * astride + (size_t)lr * 16;
    // 16-k chunks (k block, group of 16): the next chunk's four plane words
    // are loaded while the current one multiplies
    const int nch = nkb * 4;
    auto ldchunk = [&](int ch, v4i(&d)[P]) {
      const int8_t *p = pa + (size_t)(ch >> 2) * (P * 1024) + (ch & 3) * 256;
#pragma unroll
      for (int s = 0; s < P; s++) d[s] = __builtin_nontemporal_load(reinterpret_cast<const v4i *>(p + s * 1024));
    };
    v4i dc[P];
    ldchunk(0, dc);
#pragma unroll 1
    for (int ch = 0; ch < nch; ch++) {
      v4i dn[P];
      ldchunk(ch + 1 < nch ? ch + 1 : ch, dn);
      const u32 *qk0 = qs + ch * 16 * C;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        // keeps the query-column reads of later steps from being hoisted
        // (they would hold 16 x C registers)
        asm volatile("" ::: "memory");
        u32 a4[4];
        residues4((u32)dc[0][i], (u32)dc[1][i], (u32)dc[2][i], (u32)dc[3][i], a4);
#pragma unroll
        for (int j = 0; j < 4; j++) {
          u32 qv[C];
          if constexpr (C % 4 == 0) {
#pragma unroll
            for (int h = 0; h < C / 4; h++) {
              const uint4 x = reinterpret_cast<const uint4 *>(qk0 + (4 * i + j) * C)[h];
              qv[4 * h] = x.x; qv[4 * h + 1] = x.y; qv[4 * h + 2] = x.z; qv[4 * h + 3] = x.w;
            }
          } else {
            const uint2 x = *reinterpret_cast<const uint2 *>(qk0 + (4 * i + j) * C);
            qv[0] = x.x; qv[1] = x.y;
          }
#pragma unroll
          for (int c = 0; c < C; c++) acc[c] += (u64)a4[j] * qv[c];
        }
        if (i & 1) {  // 8 products since the last fold
#pragma unroll
          for (int c = 0; c < C; c++) acc[c] = (acc[c] >> 32) * L.r32 + (u32)acc[c];
        }
      }
#pragma unroll
      for (int s = 0; s < P; s++) dc[s] = dn[s];
    }
#pragma unroll
    for (int c = 0; c < C; c++) res[li][c] = reduce_u64(acc[c], L);
  }
  if (row >= g.Nrest) return;
  const int ncols = min(C, 2 * A.B - c0);
#pragma unroll
  for (int c = 0; c < C; c++) {
    if (c >= ncols) break;
    const int col = c0 + c, qi = col >> 1, pp = col & 1;
    v4i *dst = reinterpret_cast<v4i *>(A.mid + qi * A.mid_stride + row * A.ct_vals + (size_t)pp * (A.ct_vals >> 1) +
                                       (size_t)limb * g.nlv + (size_t)lvb * LW);
#pragma unroll
    for (int h = 0; h < LW / 4; h++) {
      const v4i x = {(int)res[4 * h][c], (int)res[4 * h + 1][c], (int)res[4 * h + 2][c], (int)res[4 * h + 3][c]};
      __builtin_nontemporal_store(x, dst + h);
    }
  }
}

}  // namespace

static inline unsigned cdivu(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

DbGeom make_db_geom(size_t N0, size_t Nrest, int P, int nlv, int nl) {
  DbGeom g;
  g.nl = nl;
  g.N0 = N0;
  g.Nrest = Nrest;
  g.nrt = (Nrest + 15) / 16;
  g.nkb = (N0 + 63) / 64;
  g.P = P;
  g.nlv = nlv;
  return g;
}

size_t db_bytes(const DbGeom &g) { return (size_t)g.nl * g.nlv * g.nrt * g.nkb * (size_t)g.P * 1024; }

QfragGeom make_qfrag_geom(size_t N0, int B, int P, int nlv, int nl) {
  QfragGeom q;
  q.nl = nl;
  q.N0 = (int)N0;
  q.nkb = (int)((N0 + 63) / 64);
  q.nct = (2 * B + 15) / 16;
  q.ncm = 2 * B < 16 ? 2 * B : 16;
  q.P = P;
  q.nlv = nlv;
  return q;
}

size_t qfrag_bytes(const QfragGeom &q) { return (size_t)q.nl * q.nlv * q.nct * q.nkb * (size_t)(q.P * 4 * q.ncm * 16); }

void launch_db_scatter(const u32 *tmp, int8_t *db, size_t first_pt, size_t count, const DbGeom &g, hipStream_t st) {
  if (!count) return;
  dim3 grid(cdivu(count, 64), g.nlv / 32, 2);
  k_db_scatter<4, u32><<<grid, 256, 0, st>>>(tmp, db, first_pt, count, g);
}

void launch_db_scatter64(const u64 *tmp, int8_t *db, size_t first_pt, size_t count, const DbGeom &g,
                         hipStream_t st) {
  if (!count) return;
  dim3 grid(cdivu(count, 64), g.nlv / 32, g.nl);
  k_db_scatter<8, u64><<<grid, 256, 0, st>>>(tmp, db, first_pt, count, g);
}

void launch_db_from_ref(const u32 *ref, int8_t *db, const DbGeom &g, hipStream_t st) {
  const size_t total = 2 * (size_t)g.nlv * g.N0 * g.Nrest;
  k_db_from_ref<<<cdivu(total, 256), 256, 0, st>>>(ref, db, g);
}

void launch_qfrag(const u32 *qntt, int8_t *qfrag, size_t q_stride, int B, const DbGeom &g, hipStream_t st) {
  const QfragGeom qg = make_qfrag_geom(g.N0, B, 4, g.nlv);
  // every stored column and every k block is written (zeros for k >= N0)
  dim3 pg(qg.nct * qg.nkb * 4, g.nlv / qf_levels<u32>(), 2);
  k_qfrag<4, u32><<<pg, 256, 0, st>>>(qntt, qfrag, q_stride, qg, 2 * B);
}

void launch_qfrag64(const u64 *qntt, int8_t *qfrag, size_t q_stride, int B, const DbGeom &g, hipStream_t st) {
  const QfragGeom qg = make_qfrag_geom(g.N0, B, 8, g.nlv, g.nl);
  dim3 pg(qg.nct * qg.nkb * 4, g.nlv / qf_levels<u64>(), g.nl);
  k_qfrag<8, u64><<<pg, 256, 0, st>>>(qntt, qfrag, q_stride, qg, 2 * B);
}

// Scan variants: (kind, column tiles per workgroup CB, k blocks per load
// batch KC, levels staged per flush LV, staged columns SC, next loads issued
// early). A direct variant (k_scan_direct, or k_scan_pp with two register
// buffers, one column tile per workgroup; k_scan_rb, CB column tiles) needs
// nkb % KC == 0 and stages min(16, 2B) <= SC columns.
struct ScanShape {
  bool direct;
  int CB, KC, LV, SC;
  bool xpf;
  bool valu = false;  // k_scan_valu with CB query columns per workgroup
  // Largest batch offered (0: any). Shapes with few workgroups time well
  // alone but lose inside large batches, whose selector completion runs
  // beside the scan and holds CUs they would need.
  int max_b = 0;
};

template <typename Epi>
static ScanArgs<Epi> scan_args(const int8_t *db, const int8_t *qfrag, typename Epi::Res *mid, size_t mid_stride,
                               int B, const DbGeom &g, int CB) {
  ScanArgs<Epi> A;
  A.db = db; A.qfrag = qfrag; A.mid = mid; A.mid_stride = mid_stride; A.ct_vals = 2 * (size_t)g.nl * g.nlv;
  A.g = g;
  A.qg = make_qfrag_geom(g.N0, B, Epi::P, g.nlv, g.nl);
  A.nrb = (int)cdivu(g.nrt, 4);
  A.ncb = (int)cdivu(A.qg.nct, CB);
  A.B = B;
  return A;
}

template <typename Epi>
static unsigned scan_grid(const ScanArgs<Epi> &A, int LW = kLevelsPerWG) {
  return (unsigned)((size_t)A.g.nl * (A.g.nlv / LW) * A.nrb * A.ncb);
}

template <int CB, int LV, typename Epi, typename SchemeT, int LW = kLevelsPerWG, bool NTS = true>
static void run_lds(const int8_t *db, const int8_t *qfrag, typename Epi::Res *mid, size_t mid_stride, int B,
                    const DbGeom &g, const SchemeT &S, hipStream_t st) {
  const ScanArgs<Epi> A = scan_args<Epi>(db, qfrag, mid, mid_stride, B, g, CB);
  if (A.ncb == 1)
    k_scan<CB, LV, true, Epi, SchemeT, LW, NTS><<<scan_grid(A, LW), 256, 0, st>>>(A, S);
  else
    k_scan<CB, LV, false, Epi, SchemeT, LW, NTS><<<scan_grid(A, LW), 256, 0, st>>>(A, S);
}

// PAD: KiB of unused dynamic LDS per workgroup, to cap the workgroups a CU
// may hold (the staged results plus PAD above a third of the CU's 160 KiB
// leave room for two), so a grid of two workgroups per CU cannot land three
// on some CUs and one on others.
template <int KC, int LV, int SC, bool XPF, typename Epi, typename SchemeT, bool NTS = true, int LW = kLevelsPerWG,
          int PAD = 0>
static void run_direct(const int8_t *db, const int8_t *qfrag, typename Epi::Res *mid, size_t mid_stride, int B,
                       const DbGeom &g, const SchemeT &S, hipStream_t st) {
  const ScanArgs<Epi> A = scan_args<Epi>(db, qfrag, mid, mid_stride, B, g, 1);
  const size_t lds = (size_t)PAD * 1024;
  if (A.ncb == 1)
    k_scan_direct<KC, LV, SC, true, XPF, Epi, SchemeT, NTS, LW><<<scan_grid(A, LW), 256, lds, st>>>(A, S);
  else
    k_scan_direct<KC, LV, SC, false, XPF, Epi, SchemeT, NTS, LW><<<scan_grid(A, LW), 256, lds, st>>>(A, S);
}

template <int KC, int LV, int SC, typename Epi, typename SchemeT, bool NTS = true, int LW = kLevelsPerWG>
static void run_pp(const int8_t *db, const int8_t *qfrag, typename Epi::Res *mid, size_t mid_stride, int B,
                   const DbGeom &g, const SchemeT &S, hipStream_t st) {
  const ScanArgs<Epi> A = scan_args<Epi>(db, qfrag, mid, mid_stride, B, g, 1);
  if (A.ncb == 1)
    k_scan_pp<KC, LV, SC, true, Epi, SchemeT, NTS, LW><<<scan_grid(A, LW), 256, 0, st>>>(A, S);
  else
    k_scan_pp<KC, LV, SC, false, Epi, SchemeT, NTS, LW><<<scan_grid(A, LW), 256, 0, st>>>(A, S);
}

// VALU scan: 256-row blocks x C-column blocks (column blocks innermost, so
// the workgroups re-reading a row block's database bytes share an XCD's L2).
template <int C>
static void run_valu(const int8_t *db, const int8_t *qfrag, u32 *mid, size_t mid_stride, int B, const DbGeom &g,
                     const SchemeConst &S, hipStream_t st) {
  ScanArgs<EpiK1> A = scan_args<EpiK1>(db, qfrag, mid, mid_stride, B, g, 1);
  A.nrb = (int)cdivu(g.Nrest, 256);
  A.ncb = (int)cdivu(2 * (size_t)B, C);
  const unsigned grid = (unsigned)((size_t)g.nl * (g.nlv / kLevelsPerWG) * A.nrb * A.ncb);
  const size_t lds = g.nkb * 64 * C * sizeof(u32);
  k_scan_valu<C><<<grid, 256, lds, st>>>(A, S);
}

// NT A loads only where each A fragment has one reader (one column block,
// one wave per row block).
template <int R, int C, int WC, int KC, int LV, int LW, typename Epi, typename SchemeT, bool NTS>
static void run_rb(const int8_t *db, const int8_t *qfrag, typename Epi::Res *mid, size_t mid_stride, int B,
                   const DbGeom &g, const SchemeT &S, hipStream_t st) {
  ScanArgs<Epi> A = scan_args<Epi>(db, qfrag, mid, mid_stride, B, g, C * WC);
  A.nrb = (int)cdivu(g.nrt, 4 / WC * R);
  const unsigned grid = (unsigned)((size_t)g.nl * (g.nlv / LW) * A.nrb * A.ncb);
  if (A.ncb == 1 && WC == 1)
    k_scan_rb<R, C, WC, KC, LV, LW, true, Epi, SchemeT, NTS><<<grid, 256, 0, st>>>(A, S);
  else
    k_scan_rb<R, C, WC, KC, LV, LW, false, Epi, SchemeT, NTS><<<grid, 256, 0, st>>>(A, S);
}

template <typename Epi, typename SchemeT>
using ScanFn = void (*)(const int8_t *, const int8_t *, typename Epi::Res *, size_t, int, const DbGeom &,
                        const SchemeT &, hipStream_t);

template <typename Epi, typename SchemeT>
struct ScanEntry {
  ScanShape shape;
  ScanFn<Epi, SchemeT> fn;
};

typedef k2::Scheme64 S64;
static const ScanEntry<EpiK1, SchemeConst> kScanK1[] = {
    {{true, 1, 4, 16, 2, true}, run_direct<4, 16, 2, true, EpiK1, SchemeConst>},
    {{true, 1, 4, 16, 16, true}, run_direct<4, 16, 16, true, EpiK1, SchemeConst>},
    {{true, 1, 2, 16, 16, false}, run_direct<2, 16, 16, false, EpiK1, SchemeConst>},
    {{true, 1, 1, 16, 16, false}, run_direct<1, 16, 16, false, EpiK1, SchemeConst>},
    {{true, 1, 4, 16, 2, false}, run_direct<4, 16, 2, false, EpiK1, SchemeConst>},
    {{false, 1, 1, 4, 16, false}, run_lds<1, 4, EpiK1, SchemeConst>},
    {{false, 2, 1, 4, 32, false}, run_lds<2, 4, EpiK1, SchemeConst>},
    {{false, 2, 1, 2, 32, false}, run_lds<2, 2, EpiK1, SchemeConst>},
    {{true, 1, 2, 16, 2, true}, run_pp<2, 16, 2, EpiK1, SchemeConst>},
    {{true, 4, 4, 16, 16, true}, run_rb<1, 1, 4, 4, 16, 32, EpiK1, SchemeConst, false>},
    {{false, 2, 0, 16, 0, false, true}, run_valu<2>},
    {{false, 4, 0, 16, 0, false, true}, run_valu<4>},
    {{false, 8, 0, 16, 0, false, true}, run_valu<8>},
    // one query: more levels per workgroup, staged at once (longer result runs)
    {{true, 1, 4, 32, 2, true}, run_direct<4, 32, 2, true, EpiK1, SchemeConst, true, 32>},
    {{true, 1, 4, 64, 2, true}, run_direct<4, 64, 2, true, EpiK1, SchemeConst, true, 64>},
    // 512 workgroups for the 256 x 1024 layout, two per CU in one round (48 KiB
    // of staged results + 8 KiB of padding: a CU cannot take a third)
    {{true, 1, 4, 64, 2, true}, run_direct<4, 64, 2, true, EpiK1, SchemeConst, true, 128, 8>},
    {{true, 1, 4, 16, 16, true, false, 8}, run_direct<4, 16, 16, true, EpiK1, SchemeConst, true, 32>},
    {{true, 1, 4, 16, 16, true, false, 8}, run_direct<4, 16, 16, true, EpiK1, SchemeConst, true, 64>},
};

static const ScanEntry<EpiK2, S64> kScanK2[] = {
    {{true, 1, 2, 16, 2, false}, run_direct<2, 16, 2, false, EpiK2, S64>},
    {{true, 1, 1, 16, 2, false}, run_direct<1, 16, 2, false, EpiK2, S64>},
    {{true, 1, 2, 8, 16, false}, run_direct<2, 8, 16, false, EpiK2, S64>},
    {{true, 1, 1, 8, 16, false}, run_direct<1, 8, 16, false, EpiK2, S64>},
    {{false, 1, 1, 2, 16, false}, run_lds<1, 2, EpiK2, S64>},
    {{false, 2, 1, 2, 32, false}, run_lds<2, 2, EpiK2, S64>},
    {{false, 2, 1, 1, 32, false}, run_lds<2, 1, EpiK2, S64>},
    {{true, 1, 1, 16, 2, true}, run_direct<1, 16, 2, true, EpiK2, S64>},
    {{true, 1, 1, 16, 2, true}, run_pp<1, 16, 2, EpiK2, S64>},
    {{true, 4, 1, 8, 16, true}, run_rb<1, 1, 4, 1, 8, 16, EpiK2, S64, false>},
};

int scan_variant_count(int P) { return P == 4 ? (int)std::size(kScanK1) : (int)std::size(kScanK2); }

static bool applicable(const ScanShape &s, int B, const DbGeom &g) {
  if (s.max_b > 0 && B > s.max_b) return false;
  // VALU: LDS of the rebuilt query columns <= 64 KiB, and at most 16 column
  // blocks re-reading each row block (more only lengthen autotuning: 18 and
  // 33 ms at 32 and 64 blocks against 11 ms at 16 for 64 queries)
  if (s.valu) return g.nkb * s.CB <= 256 && (2 * B + s.CB - 1) / s.CB <= 16;
  if (!s.direct) return true;
  return g.nkb % s.KC == 0 && (2 * B < 16 ? 2 * B : 16) <= s.SC;
}

// Default shapes per batch-size class (2B <= 2, <= 16, larger), in order of
// preference, the fastest measured on MI355X at the 256 x 1024 and
// 512 x 4096 layouts (B = 1, 8 and 64 / 32). At 64 queries the
// register-blocked shape (one row tile x four column tiles per workgroup)
// takes 3.37 ms against the direct shape's 3.70. For one query the direct
// shape gains from more levels per workgroup, staged at once: 16 / 32 / 64
// levels (64- / 128- / 256-byte result runs) take 0.706 / 0.700 / 0.678 ms,
// and 128 levels per workgroup staged 64 at a time (512 workgroups, two per
// CU, one round) 0.671 ms. Staging fewer levels than a workgroup takes
// (16 of 32: 0.722 ms; 32 of 128: 0.690 ms) loses. The two-prime plain scan
// (u64 results, 128-byte runs at 16 levels) gains nothing from 32 of 64 or
// 128 (21.55-21.96 against 21.55 ms). For 8 queries 32 levels per
// workgroup take 0.993 against 1.009 ms. For 64 queries the register-blocked
// shape with 128 levels per workgroup timed 3.50 against 3.57 ms alone, but
// inside the batch, where the selector completion runs beside the scan, it
// took 3.50 against 3.35 ms: with fewer workgroups the scan cannot balance
// around the CUs the completion holds, so only the 32-level one is kept.
static const int kPrefK1[3][4] = {{15, 14, 3, 5}, {16, 1, 3, 5}, {9, 1, 6, 5}};
static const int kPrefK2[3][4] = {{8, 1, 3, 4}, {3, 4, -1, -1}, {3, 4, 5, -1}};

template <typename Epi, typename SchemeT, size_t NV>
static int dispatch(const ScanEntry<Epi, SchemeT> (&tab)[NV], const int (&pref)[3][4], int variant,
                    const int8_t *db, const int8_t *qfrag, typename Epi::Res *mid, size_t mid_stride, int B,
                    const DbGeom &g, const SchemeT &S, hipStream_t st) {
  if (variant < 0) {
    const int cls = 2 * B <= 2 ? 0 : (2 * B <= 16 ? 1 : 2);
    for (int i = 0; i < 4 && variant < 0; i++) {
      const int v = pref[cls][i];
      if (v >= 0 && applicable(tab[v].shape, B, g)) variant = v;
    }
  }
  if (variant < 0 || variant >= (int)NV || !applicable(tab[variant].shape, B, g)) return -1;
  tab[variant].fn(db, qfrag, mid, mid_stride, B, g, S, st);
  return 0;
}

bool scan_variant_ok(int P, int variant, int B, const DbGeom &g) {
  if (variant < 0 || variant >= scan_variant_count(P)) return false;
  return applicable(P == 4 ? kScanK1[variant].shape : kScanK2[variant].shape, B, g);
}

int launch_scan(const int8_t *db, const int8_t *qfrag, u32 *mid, size_t mid_stride, int B, const DbGeom &g,
                const SchemeConst &S, hipStream_t st, int variant) {
  return dispatch(kScanK1, kPrefK1, variant, db, qfrag, mid, mid_stride, B, g, S, st);
}

int launch_scan64(const int8_t *db, const int8_t *qfrag, u64 *mid, size_t mid_stride, int B, const DbGeom &g,
                  const k2::Scheme64 &S, hipStream_t st, int variant) {
  return dispatch(kScanK2, kPrefK2, variant, db, qfrag, mid, mid_stride, B, g, S, st);
}

}  // namespace opir
