// First-dimension DB scan on the matrix cores (src/server.cpp:313-361,
// src/matrix.cpp:232-350):
//   C[limb][l][i][c] = sum_k A[limb][l][i][k] * Bq[limb][l][k][c] mod q_limb
// for every NTT level l, candidate row i < Nrest, first-dim column k < N0 and
// query column c = 2*query + poly.
//
// Exactness with int8 MFMA: every residue a < 2^30 is written in balanced
// base-256 digits a = d0 + 2^8 d1 + 2^16 d2 + 2^24 d3 with d0..d2 in
// [-128, 128) and d3 in [0, 64]. Then
//   sum_k a_k b_k = sum_{s,t} 256^(s+t) sum_k d_{s,k} e_{t,k},
// where each inner sum is one int8 GEMM with int32 accumulation. The 16
// plane pairs accumulate into 7 int32 tiles G_u (u = s+t); |G_u| <= 4 * N0p
// * 2^14 < 2^31 for N0p <= 2^14. The epilogue forms sum_u (256^u mod q) G_u
// in int64 (|.| < 2^58) and reduces it mod q. No rounding anywhere.
//
// MFMA v_mfma_i32_16x16x64_i8 operand layout (gfx950), lane l, byte j:
//   A (16 x 64): row = l % 16, k = 16 * (l / 16) + j
//   B (64 x 16): col = l % 16, k = 16 * (l / 16) + j
//   D (16 x 16): col = l % 16, row = 4 * (l / 16) + r   (r = 0..3)
//
// Database layout ("fragment-native"): for each (limb, level, row tile rt of
// 16 rows, k block kb of 64, digit plane s) one 1 KiB block holding the A
// fragment exactly as the 64 lanes load it (16 bytes per lane), so every
// wave load is one contiguous 1 KiB line run.
#include "opir_kernels.h"

namespace opir {

typedef int v4i __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void balanced_digits(u32 a, int (&d)[4]) {
  int x = (int)a;  // a < 2^30
#pragma unroll
  for (int s = 0; s < 3; s++) {
    const int r = ((x + 128) & 255) - 128;
    d[s] = r;
    x = (x - r) >> 8;
  }
  d[3] = x;
}

__device__ __forceinline__ size_t db_byte_offset(const DbGeom &g, int limb, int lv, size_t row, size_t k, int s) {
  const size_t rt = row >> 4, lr = row & 15, kb = k >> 6, grp = (k & 63) >> 4, j = k & 15;
  const size_t blk = ((((size_t)limb * kN + lv) * g.nrt + rt) * g.nkb + kb) * 4 + s;
  return blk * 1024 + (grp * 16 + lr) * 16 + j;
}

__device__ __forceinline__ void store_digits(int8_t *db, const DbGeom &g, int limb, int lv, size_t row, size_t k,
                                             u32 a) {
  int d[4];
  balanced_digits(a, d);
#pragma unroll
  for (int s = 0; s < 4; s++) db[db_byte_offset(g, limb, lv, row, k, s)] = (int8_t)d[s];
}

// tmp [limb][count][N] (pt-major, NTT form) -> fragment-native DB.
// Tile: 64 pts x 32 levels staged through LDS for coalesced reads.
__global__ __launch_bounds__(256) void k_db_scatter(const u32 *__restrict__ tmp, int8_t *__restrict__ db,
                                                    size_t first_pt, size_t count, DbGeom g) {
  __shared__ u32 tile[32][65];
  const int limb = blockIdx.z;
  const size_t pt0 = (size_t)blockIdx.x * 64;
  const int lv0 = blockIdx.y * 32;
  const u32 *src = tmp + (size_t)limb * count * kN;
  for (int e = threadIdx.x; e < 64 * 32; e += 256) {
    const int pl = e / 32, ll = e % 32;
    const size_t pt = pt0 + pl;
    tile[ll][pl] = pt < count ? src[pt * kN + lv0 + ll] : 0;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 64 * 32; e += 256) {
    const int ll = e / 64, pl = e % 64;
    const size_t pt = pt0 + pl;
    if (pt >= count) continue;
    const size_t gpt = first_pt + pt;
    store_digits(db, g, limb, lv0 + ll, gpt / g.N0, gpt % g.N0, tile[ll][pl]);
  }
}

// Reference composite layout ref [limb][coeff][num_pt] (server.cpp:162-170).
__global__ void k_db_from_ref(const u32 *__restrict__ ref, int8_t *__restrict__ db, DbGeom g) {
  const size_t num_pt = g.N0 * g.Nrest;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 2 * (size_t)kN * num_pt) return;
  const size_t pt = i % num_pt, lv = (i / num_pt) % kN, limb = i / (num_pt * kN);
  store_digits(db, g, (int)limb, (int)lv, pt / g.N0, pt % g.N0, ref[i]);
}

// Query columns -> B fragments. qntt [B][N0][2 polys][2 limbs][N] (NTT form)
// -> qfrag [limb][level][ct][kb][t][grp][n < ncm][16 bytes], ct = column
// tile of 16 query columns (c = 2*query + poly), ncm = min(16, 2B).
// Workgroup: one (query, poly, limb, 16-k block, 64-level block).
__global__ __launch_bounds__(256) void k_qfrag(const u32 *__restrict__ qntt, int8_t *__restrict__ qfrag,
                                               size_t q_stride, QfragGeom qg) {
  __shared__ u32 tile[16][65];
  const int col = blockIdx.x;  // 2*query + poly
  const int qi = col >> 1, p = col & 1;
  const int k0 = blockIdx.y * 16;
  const int lv0 = (blockIdx.z >> 1) * 64, limb = blockIdx.z & 1;
  const u32 *src = qntt + (size_t)qi * q_stride;
  for (int e = threadIdx.x; e < 16 * 64; e += 256) {
    const int kk = e / 64, ll = e % 64;
    const int k = k0 + kk;
    tile[kk][ll] = k < qg.N0 ? src[(((size_t)k * 2 + p) * 2 + limb) * kN + lv0 + ll] : 0;
  }
  __syncthreads();
  const int ll = threadIdx.x & 63, t = threadIdx.x >> 6;
  const int lv = lv0 + ll;
  u32 w[4];
#pragma unroll
  for (int q4 = 0; q4 < 4; q4++) {
    u32 packed = 0;
#pragma unroll
    for (int b = 0; b < 4; b++) {
      int d[4];
      balanced_digits(tile[q4 * 4 + b][ll], d);
      packed |= ((u32)(uint8_t)(int8_t)d[t]) << (8 * b);
    }
    w[q4] = packed;
  }
  const int ct = col >> 4, n = col & 15, kb = k0 >> 6, grp = (k0 & 63) >> 4;
  const size_t off = ((((size_t)limb * kN + lv) * qg.nct + ct) * qg.nkb + kb) * (size_t)(4 * 4 * qg.ncm * 16) +
                     ((size_t)(t * 4 + grp) * qg.ncm + n) * 16;
  *reinterpret_cast<uint4 *>(qfrag + off) = make_uint4(w[0], w[1], w[2], w[3]);
}

// The scan. Workgroup = 4 waves = 4 row tiles (64 rows) x LV levels x one
// column tile, one limb. blockIdx.x is remapped so that consecutive logical
// blocks run on the same XCD: the column tiles of one DB tile, then the row
// blocks of one level block (which share the query fragments).
constexpr int kScanLV = 16;

__global__ __launch_bounds__(256) void k_scan_mfma(const int8_t *__restrict__ db, const int8_t *__restrict__ qfrag,
                                                   u32 *__restrict__ mid, size_t mid_stride, DbGeom g, QfragGeom qg,
                                                   int nrb, int B, SchemeConst S) {
  const unsigned W = gridDim.x, p = blockIdx.x;
  const unsigned L = (p % 8) * (W / 8) + p / 8;  // XCD-aware: consecutive L on one XCD
  // column tiles innermost: the WGs reading the same DB tile run together on
  // one XCD, so the tile comes from HBM once and from L2 for the others
  const int ct = L % qg.nct;
  unsigned rest = L / qg.nct;
  const int rb = rest % nrb;
  rest /= nrb;
  const int lb = rest % (kN / kScanLV);
  const int limb = rest / (kN / kScanLV);
  const LimbConst &Lc = S.limb[limb];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int rt = rb * 4 + wave;
  const int n = lane & 15, grp = lane >> 4;
  const int ncols_here = min(16, 2 * B - ct * 16);
  const bool has_col = n < ncols_here;
  // 256^u mod q and an offset C = q * ceil(2^59 / q) making the int64 sum positive
  u32 wu[7];
  {
    u64 x = 1;
#pragma unroll
    for (int u = 0; u < 7; u++) {
      wu[u] = (u32)x;
      x = (x << 8) % Lc.q;
    }
  }
  const u64 pos_off = ((((u64)1 << 59) / Lc.q) + 1) * Lc.q;
  const size_t qtile = (size_t)(4 * 4 * qg.ncm * 16);
  // per wave: [level][row 16][column 16] results, written out as level runs
  __shared__ u32 stage_lds[4][kScanLV][16][17];
#pragma unroll 1
  for (int lvl = 0; lvl < kScanLV; lvl++) {
    const int lv = lb * kScanLV + lvl;
    v4i acc[7];
#pragma unroll
    for (int u = 0; u < 7; u++) acc[u] = (v4i){0, 0, 0, 0};
    const int8_t *abase = db + ((((size_t)limb * kN + lv) * g.nrt + rt) * g.nkb) * 4096 + lane * 16;
    const int8_t *bbase = qfrag + (((size_t)limb * kN + lv) * qg.nct + ct) * qg.nkb * qtile;
    if (rt < (int)g.nrt) {
      for (int kb = 0; kb < (int)g.nkb; kb++) {
        v4i a[4], b[4];
#pragma unroll
        for (int s = 0; s < 4; s++) a[s] = *reinterpret_cast<const v4i *>(abase + (size_t)kb * 4096 + s * 1024);
#pragma unroll
        for (int t = 0; t < 4; t++)
          b[t] = has_col ? *reinterpret_cast<const v4i *>(bbase + kb * qtile + ((size_t)(t * 4 + grp) * qg.ncm + n) * 16)
                         : (v4i){0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < 4; s++)
#pragma unroll
          for (int t = 0; t < 4; t++) acc[s + t] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[s], b[t], acc[s + t], 0, 0, 0);
      }
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
      int64_t v = 0;
#pragma unroll
      for (int u = 0; u < 7; u++) v += (int64_t)wu[u] * (int64_t)acc[u][r];
      stage_lds[wave][lvl][grp * 4 + r][n] = reduce_u64((u64)(v + (int64_t)pos_off), Lc);
    }
  }
  __syncthreads();
  // (row, column) pairs of this wave's tile, LV consecutive levels each
  for (int e = lane; e < 16 * ncols_here; e += 64) {
    const int rr = e / ncols_here, nn = e % ncols_here;
    const size_t row = (size_t)rt * 16 + rr;
    if (row >= g.Nrest) continue;
    const int c = ct * 16 + nn, qi = c >> 1, pp = c & 1;
    uint4 *dst = reinterpret_cast<uint4 *>(mid + qi * mid_stride + row * kCtWords + (size_t)(pp * 2 + limb) * kN +
                                           lb * kScanLV);
#pragma unroll
    for (int v = 0; v < kScanLV / 4; v++)
      dst[v] = make_uint4(stage_lds[wave][4 * v][rr][nn], stage_lds[wave][4 * v + 1][rr][nn],
                          stage_lds[wave][4 * v + 2][rr][nn], stage_lds[wave][4 * v + 3][rr][nn]);
  }
}

static inline unsigned cdivu(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

DbGeom make_db_geom(size_t N0, size_t Nrest) {
  DbGeom g;
  g.N0 = N0;
  g.Nrest = Nrest;
  g.nrt = (Nrest + 15) / 16;
  g.nkb = (N0 + 63) / 64;
  return g;
}

size_t db_bytes(const DbGeom &g) { return 2 * (size_t)kN * g.nrt * g.nkb * 4096; }

QfragGeom make_qfrag_geom(size_t N0, int B) {
  QfragGeom q;
  q.N0 = (int)N0;
  q.nkb = (int)((N0 + 63) / 64);
  q.nct = (2 * B + 15) / 16;
  q.ncm = 2 * B < 16 ? 2 * B : 16;
  return q;
}

size_t qfrag_bytes(const QfragGeom &q) { return 2 * (size_t)kN * q.nct * q.nkb * (size_t)(4 * 4 * q.ncm * 16); }

void launch_db_scatter(const u32 *tmp, int8_t *db, size_t first_pt, size_t count, const DbGeom &g, hipStream_t st) {
  if (!count) return;
  dim3 grid(cdivu(count, 64), kN / 32, 2);
  k_db_scatter<<<grid, 256, 0, st>>>(tmp, db, first_pt, count, g);
}

void launch_db_from_ref(const u32 *ref, int8_t *db, const DbGeom &g, hipStream_t st) {
  const size_t total = 2 * (size_t)kN * g.N0 * g.Nrest;
  k_db_from_ref<<<cdivu(total, 256), 256, 0, st>>>(ref, db, g);
}

void launch_qfrag(const u32 *qntt, int8_t *qfrag, size_t q_stride, int B, const DbGeom &g, hipStream_t st) {
  const QfragGeom qg = make_qfrag_geom(g.N0, B);
  // k_qfrag writes every stored column and every k block (zeros for k >= N0);
  // the scan never reads the unstored columns of a partial last tile.
  dim3 pg(2 * B, cdivu(qg.nkb * 64, 16), (kN / 64) * 2);
  k_qfrag<<<pg, 256, 0, st>>>(qntt, qfrag, q_stride, qg);
}

void launch_scan(const int8_t *db, const int8_t *qfrag, u32 *mid, size_t mid_stride, int B, const DbGeom &g,
                 const SchemeConst &S, hipStream_t st) {
  const QfragGeom qg = make_qfrag_geom(g.N0, B);
  const int nrb = (int)cdivu(g.nrt, 4);
  const size_t W = (size_t)2 * (kN / kScanLV) * qg.nct * nrb;
  k_scan_mfma<<<(unsigned)W, 256, 0, st>>>(db, qfrag, mid, mid_stride, g, qg, nrb, B, S);
}

}  // namespace opir
