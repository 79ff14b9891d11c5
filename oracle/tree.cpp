// CPU restatement of the reference's tree PIR layer (test infrastructure):
// index math of src/tree_index.cpp, query packing and unpacking of
// src/tree_query.cpp, level selection of src/tree_select.cpp, on top of the
// oracle's expansion, completion and ext_prod_mux.
#include <bit>
#include <stdexcept>

#include "oracle.h"

namespace oracle {

static void treq(bool c, const char *m) {
  if (!c) throw std::invalid_argument(std::string("tree: ") + m);
}

TreeParams make_tree_params(size_t L, size_t a, const PirParams &scheme) {
  // tree_query.cpp:44-61 (ell_beta = ell_gamma = L_EP) over tree_index.cpp:53-118
  TreeParams p;
  p.L = L;
  p.N = size_t{1} << L;
  p.n = scheme.N();
  treq(std::has_single_bit(p.n), "ring degree must be a power of two");
  p.r = std::bit_width(p.n) - 1;
  p.a = a;
  p.N0 = size_t{1} << a;
  treq(L >= p.r + a, "tree height is below r + a; no beta bits remain");
  p.b = L - p.r - a;
  p.P = p.N / p.n;
  p.B = size_t{1} << p.b;
  p.ell = scheme.l_ep();
  p.w = p.N0 + p.ell * p.b + p.ell * p.r;
  treq(p.w <= p.n, "packed width exceeds the ring degree");
  p.W = std::bit_ceil(p.w);
  p.h_q = std::bit_width(p.W) - 1;
  p.t = scheme.plain_mod;
  treq(p.t % 2 == 1, "W must be invertible modulo t");
  return p;
}

TreeCoords tree_client_index(size_t leaf, const TreeParams &p) {
  // tree_index.cpp:150-170
  treq(leaf < p.N, "client_index leaf is out of range");
  const size_t gamma = leaf / p.P, pp = leaf % p.P;
  TreeCoords c;
  c.alpha = pp / p.B;
  const size_t beta = pp % p.B;
  for (size_t u = 0; u < p.b; u++) c.beta.push_back((beta >> u) & 1);
  for (size_t v = 0; v < p.r; v++) c.gamma.push_back((gamma >> v) & 1);
  return c;
}

std::vector<TreePlan> tree_level_plans(const TreeParams &p) {
  // tree_index.cpp:172-215
  std::vector<TreePlan> out;
  for (size_t level = 0; level <= p.L; level++) {
    TreePlan pl;
    pl.level = level;
    pl.R = level >= p.r ? size_t{1} << (level - p.r) : 1;
    if (pl.R == 1) {
      pl.select_case = 0;
      pl.coarsen = p.a;
    } else if (pl.R < p.N0) {
      pl.select_case = 1;
      pl.coarsen = p.a - (level - p.r);
    } else {
      pl.select_case = 2;
      pl.coarsen = 0;
    }
    if (pl.select_case == 2) {
      pl.beta_begin = p.L - level;
      pl.beta_count = level - p.r - p.a;
    } else {
      pl.beta_begin = p.b;
      pl.beta_count = 0;
    }
    out.push_back(pl);
  }
  return out;
}

u64 synthetic_tree_node(size_t level, size_t index, u64 t) {
  // tree_select.cpp:41-50
  u64 v = 0x74726565506972ULL ^ ((u64)level * 0xd6e8feb86659fd93ULL) ^ ((u64)index * 0xa5a3564e27f8862fULL);
  v += 0x9e3779b97f4a7c15ULL;
  v = (v ^ (v >> 30)) * 0xbf58476d1ce4e5b9ULL;
  v = (v ^ (v >> 27)) * 0x94d049bb133111ebULL;
  v ^= v >> 31;
  return v % t;
}

std::vector<std::vector<u64>> pack_tree_level(size_t level, const TreeParams &p) {
  // tree_select.cpp:52-74 with the synthetic node source
  treq(level <= p.L, "level is out of range");
  const size_t R = level >= p.r ? size_t{1} << (level - p.r) : 1, nodes = size_t{1} << level;
  std::vector<std::vector<u64>> d(R, std::vector<u64>(p.n, 0));
  for (size_t pp = 0; pp < R; pp++)
    for (size_t u = 0; u < p.n; u++) {
      const size_t node = pp + u * R;
      if (node >= nodes) break;
      d[pp][u] = synthetic_tree_node(level, node, p.t);
    }
  return d;
}

TreeLevelOracle tree_level_oracle(size_t leaf, size_t level, const TreeParams &p) {
  // tree_index.cpp:217-241
  TreeLevelOracle o;
  o.node = leaf >> (p.L - level);
  const size_t gamma = leaf / p.P, pp = leaf % p.P;
  if (level >= p.r) {
    o.plaintext = pp >> (p.L - level);
    o.position = gamma;
  } else {
    o.plaintext = 0;
    o.position = gamma >> (p.r - level);
  }
  return o;
}

RlweCt make_tree_query(PirClient &c, const TreeParams &tp, const PirParams &qview, size_t leaf) {
  // tree_query.cpp:140-196: one fresh encryption of zero, the alpha constant
  // W^-1 mod t lifted by Delta at BitRev(alpha), then the scaled gadget rows
  // of every beta / gamma bit that is 1 (groups b + v for gamma)
  const TreeCoords co = tree_client_index(leaf, tp);
  const size_t N = qview.N(), K = qview.K();
  const auto &qs = qview.rns_mods;
  const u64 t = qview.plain_mod;
  u64 inv = 0;
  treq(try_invert_uint_mod(tp.W, t, inv), "W is not invertible modulo t");
  RlweCt q;
  encrypt_zero_rns(c.sk, N, qs, qview.sigma(), c.rng, q, false);
  const size_t rev = bit_reverse(co.alpha, tp.h_q);
  if (K == 1) {
    const u64 Q = qs[0];
    const u64 scaled = (u64)(((u128)Q * inv + (t >> 1)) / t) % Q;
    q.c0[rev] = (u64)(((u128)q.c0[rev] + scaled) % Q);
  } else {
    const u128 Q = (u128)qs[0] * qs[1];
    const u128 Delta = Q / t;
    const u64 rem = (u64)(Q - Delta * t);
    const u64 rr = (u64)(((u128)rem * inv + (t >> 1)) / t);
    const u128 sm = Delta * inv + rr;
    for (size_t k = 0; k < K; k++) q.c0[k * N + rev] = (u64)(((u128)q.c0[k * N + rev] + (u64)(sm % qs[k])) % qs[k]);
  }
  std::vector<size_t> idx{co.alpha};
  for (auto bit : co.beta) idx.push_back(bit);
  for (auto bit : co.gamma) idx.push_back(bit);
  c.add_gsw_to_query(qview, q, idx);
  return q;
}

RlweCt tree_pt_ct_mul(const std::vector<u64> &pt, const RlweCt &ct, const PirParams &p) {
  // tree_select.cpp:86-119: per limb NTT(pt) * NTT(c_i), back to coefficients
  const size_t N = p.N(), K = p.K();
  RlweCt r;
  r.c0.assign(K * N, 0);
  r.c1.assign(K * N, 0);
  std::vector<u64> pn(N), comp(N);
  for (size_t k = 0; k < K; k++) {
    const u64 q = p.rns_mods[k];
    std::copy(pt.begin(), pt.end(), pn.begin());
    ntt_fwd(pn.data(), N, q);
    for (int second = 0; second < 2; second++) {
      const u64 *src = (second ? ct.c1.data() : ct.c0.data()) + k * N;
      u64 *dst = (second ? r.c1.data() : r.c0.data()) + k * N;
      std::copy(src, src + N, comp.begin());
      ntt_fwd(comp.data(), N, q);
      for (size_t j = 0; j < N; j++) comp[j] = (u64)((u128)comp[j] * pn[j] % q);
      ntt_inv(comp.data(), N, q);
      std::copy(comp.begin(), comp.end(), dst);
    }
  }
  return r;
}

static void ct_add(RlweCt &acc, const RlweCt &x, const PirParams &p) {
  const size_t N = p.N();
  for (size_t k = 0; k < p.K(); k++) {
    const u64 q = p.rns_mods[k];
    for (size_t j = 0; j < N; j++) {
      acc.c0[k * N + j] = (acc.c0[k * N + j] + x.c0[k * N + j]) % q;
      acc.c1[k * N + j] = (acc.c1[k * N + j] + x.c1[k * N + j]) % q;
    }
  }
}

std::vector<RlweCt> tree_select_all(PirServer &srv, const TreeParams &tp, size_t id, const RlweCt &query,
                                    const std::vector<std::vector<std::vector<u64>>> &levels) {
  const PirParams &p = srv.params;
  // unpack_tree_query (tree_query.cpp:198-240): expansion, then completion
  std::vector<RlweCt> exp = srv.fast_expand_qry(id, query);
  treq(exp.size() == tp.w, "expansion returned an unexpected leaf count");
  std::vector<GSWCt> sel = srv.complete_selectors(id, exp);
  treq(sel.size() == tp.b + tp.r, "completion returned an unexpected selector count");
  // build_alpha_pyramid (tree_select.cpp:121-146)
  std::vector<std::vector<RlweCt>> pyr;
  pyr.emplace_back(exp.begin(), exp.begin() + tp.N0);
  for (size_t c = 0; c < tp.a; c++) {
    std::vector<RlweCt> next;
    for (size_t j = 0; j < pyr.back().size() / 2; j++) {
      RlweCt s = pyr.back()[2 * j];
      ct_add(s, pyr.back()[2 * j + 1], p);
      next.push_back(std::move(s));
    }
    pyr.push_back(std::move(next));
  }
  std::vector<RlweCt> out;
  for (const TreePlan &pl : tree_level_plans(tp)) {
    const auto &db = levels[pl.level];
    if (pl.select_case == 0) {  // select_level, tree_select.cpp:208-211
      out.push_back(tree_pt_ct_mul(db[0], pyr[tp.a][0], p));
    } else if (pl.select_case == 1) {  // 212-225
      RlweCt acc = tree_pt_ct_mul(db[0], pyr[pl.coarsen][0], p);
      for (size_t k = 1; k < pl.R; k++) ct_add(acc, tree_pt_ct_mul(db[k], pyr[pl.coarsen][k], p), p);
      out.push_back(std::move(acc));
    } else {  // evaluate_alpha_dimension (148-173) + fold_beta_dimension (175-200)
      const size_t stride = size_t{1} << pl.beta_count;
      std::vector<RlweCt> cand;
      for (size_t delta = 0; delta < stride; delta++) {
        RlweCt acc = tree_pt_ct_mul(db[delta], pyr[0][0], p);
        for (size_t k = 1; k < tp.N0; k++) ct_add(acc, tree_pt_ct_mul(db[k * stride + delta], pyr[0][k], p), p);
        cand.push_back(std::move(acc));
      }
      for (size_t step = 0; step < pl.beta_count; step++) {
        const size_t u = pl.beta_begin + pl.beta_count - 1 - step, half = cand.size() / 2;
        for (size_t j = 0; j < half; j++) srv.ext_prod_mux(cand[j], cand[j + half], sel[u], cand[j]);
        cand.resize(half);
      }
      out.push_back(std::move(cand.front()));
    }
  }
  return out;
}

}  // namespace oracle
