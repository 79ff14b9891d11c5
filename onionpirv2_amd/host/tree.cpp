// Tree PIR of the host API (tree.h).
#include "tree.h"

#include <bit>
#include <cstring>
#include <stdexcept>

#include "onionpir_gpu.h"

namespace onionpir {

namespace {
void require(bool c, const char *m) {
  if (!c) throw std::invalid_argument(std::string("TreePirParams: ") + m);
}
}  // namespace

TreePirParams make_tree_pir_params_for_scheme(size_t L, size_t a, const PirParams &scheme) {
  TreePirParams p;
  p.L = L;
  require(L > 0 && L < 62, "tree height is out of range");
  p.N = size_t{1} << L;
  p.n = scheme.get_poly_degree();
  require(std::has_single_bit(p.n), "ring degree must be a power of two");
  p.rho = p.n;
  p.r = (size_t)std::bit_width(p.n) - 1;
  p.a = a;
  p.N0 = size_t{1} << a;
  require(L >= p.r + a, "tree height is below r + a; no beta bits remain");
  p.b = L - p.r - a;
  p.P = p.N / p.rho;
  p.B = size_t{1} << p.b;
  p.ell_beta = p.ell_gamma = scheme.get_l();
  p.w = p.N0 + p.ell_beta * p.b + p.ell_gamma * p.r;
  require(p.w <= p.n, "packed width exceeds the ring degree");
  p.W = std::bit_ceil(p.w);
  p.h_q = (size_t)std::bit_width(p.W) - 1;
  p.t = scheme.get_plain_mod();
  require(p.t % 2 == 1, "W must be invertible modulo t");
  p.rns_moduli = scheme.get_rns_mods();
  return p;
}

PirParams tree_query_expansion_params(const TreePirParams &tree, const PirParams &scheme) {
  return scheme.with_query_shape(tree.N0, tree.b + tree.r, tree.h_q);
}

ClientCoordinates client_index(size_t leaf, const TreePirParams &p) {
  if (leaf >= p.N) throw std::invalid_argument("client_index leaf is out of range");
  const size_t gamma = leaf / p.P, pp = leaf % p.P, beta = pp % p.B;
  ClientCoordinates c;
  c.alpha = pp / p.B;
  for (size_t u = 0; u < p.b; u++) c.beta_bits_le.push_back((uint8_t)((beta >> u) & 1));
  for (size_t v = 0; v < p.r; v++) c.gamma_bits_le.push_back((uint8_t)((gamma >> v) & 1));
  return c;
}

std::vector<LevelPlan> build_level_plans(const TreePirParams &p) {
  std::vector<LevelPlan> plans;
  for (size_t level = 0; level <= p.L; level++) {
    LevelPlan pl;
    pl.level = level;
    pl.R = level >= p.r ? size_t{1} << (level - p.r) : 1;
    if (pl.R == 1) {
      pl.select_case = SelectCase::Single;
      pl.coarsen_count = p.a;
    } else if (pl.R < p.N0) {
      pl.select_case = SelectCase::CoarsenedAlpha;
      pl.coarsen_count = p.a - (level - p.r);
    } else {
      pl.select_case = SelectCase::AlphaBeta;
    }
    if (pl.select_case == SelectCase::AlphaBeta) {
      pl.beta_begin = p.L - level;
      pl.beta_count = level - p.r - p.a;
    } else {
      pl.beta_begin = p.b;
    }
    if (level >= p.r) {
      pl.gamma_count = p.r;
    } else {
      pl.gamma_begin = p.r - level;
      pl.gamma_count = level;
    }
    pl.projection_depth = level < p.r ? level : p.r;
    plans.push_back(pl);
  }
  return plans;
}

uint64_t synthetic_tree_node_value(size_t level, size_t index, uint64_t t) {
  if (t == 0) throw std::invalid_argument("synthetic node values need a modulus");
  uint64_t v = 0x74726565506972ULL ^ ((uint64_t)level * 0xd6e8feb86659fd93ULL) ^ ((uint64_t)index * 0xa5a3564e27f8862fULL);
  v += 0x9e3779b97f4a7c15ULL;  // one splitmix64 step
  v = (v ^ (v >> 30)) * 0xbf58476d1ce4e5b9ULL;
  v = (v ^ (v >> 27)) * 0x94d049bb133111ebULL;
  return (v ^ (v >> 31)) % t;
}

std::vector<std::vector<uint64_t>> pack_tree_level(size_t level, const TreePirParams &tree,
                                                   const TreeNodeSource &source) {
  if (level > tree.L) throw std::invalid_argument("tree_select: level is out of range");
  const size_t R = level >= tree.r ? size_t{1} << (level - tree.r) : 1, nodes = size_t{1} << level;
  std::vector<std::vector<uint64_t>> d(R, std::vector<uint64_t>(tree.n, 0));
  for (size_t p = 0; p < R; p++)
    for (size_t u = 0; u < tree.n; u++) {
      const size_t node = p + u * R;
      if (node >= nodes) break;
      d[p][u] = source(level, node) % tree.t;
    }
  return d;
}

RlweCt make_tree_query(PirClient &client, const PirParams &scheme, const TreePirParams &tree, size_t leaf) {
  // the packed query of a PIR layout with N0 = tree.N0 first-dimension slots
  // and b + r selectors (beta bits, then gamma bits) at expansion height h_q
  const ClientCoordinates co = client_index(leaf, tree);
  std::vector<size_t> idx{co.alpha};
  for (auto bit : co.beta_bits_le) idx.push_back(bit);
  for (auto bit : co.gamma_bits_le) idx.push_back(bit);
  return client.generate_query_for_indices(tree_query_expansion_params(tree, scheme), idx);
}

TreePirServer::TreePirServer(const PirParams &scheme, const TreePirParams &tree, std::shared_ptr<GpuScheme> gpu)
    : scheme_(scheme), tree_(tree), gpu_(std::move(gpu)) {
  if (!gpu_) gpu_ = std::make_shared<GpuScheme>(scheme);
  gpu_->check(opir_tree_create(gpu_->ctx(), tree.L, tree.a, &t_));
}

TreePirServer::~TreePirServer() { opir_tree_destroy(t_); }

void TreePirServer::preprocess(const TreeNodeSource &source) {
  for (size_t level = 0; level <= tree_.L; level++) load_level(level, pack_tree_level(level, tree_, source));
}

void TreePirServer::load_level(size_t level, const std::vector<std::vector<uint64_t>> &pts) {
  const size_t n = tree_.n;
  std::vector<uint64_t> flat(pts.size() * n);
  for (size_t p = 0; p < pts.size(); p++) {
    if (pts[p].size() != n) throw std::invalid_argument("plaintext polynomial has the wrong degree");
    std::memcpy(&flat[p * n], pts[p].data(), n * 8);
  }
  const size_t R = level >= tree_.r ? size_t{1} << (level - tree_.r) : 1;
  if (pts.size() != R) throw std::invalid_argument("level database must hold R_l plaintexts");
  gpu_->check(opir_tree_load_level(t_, level, flat.data()));
}

std::vector<RlweCt> TreePirServer::select_path(size_t id, const RlweCt &query) {
  const size_t KN = scheme_.get_poly_degree() * scheme_.K();
  if (query.c0.size() != KN || query.c1.size() != KN) throw std::invalid_argument("query must hold K*N values per poly");
  std::vector<uint64_t> q(2 * KN), out((tree_.L + 1) * 2 * KN);
  std::memcpy(q.data(), query.c0.data(), KN * 8);
  std::memcpy(q.data() + KN, query.c1.data(), KN * 8);
  gpu_->check(opir_tree_select(t_, id, q.data(), out.data()));
  std::vector<RlweCt> r(tree_.L + 1);
  for (size_t l = 0; l <= tree_.L; l++) {
    r[l].c0.assign(out.begin() + l * 2 * KN, out.begin() + l * 2 * KN + KN);
    r[l].c1.assign(out.begin() + l * 2 * KN + KN, out.begin() + (l + 1) * 2 * KN);
  }
  return r;
}

}  // namespace onionpir
