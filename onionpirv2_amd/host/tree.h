// Tree PIR of the host API: the reference's full-binary-tree PIR layer
// (src/includes/tree_index.h, tree_query.h, tree_select.h) with its names
// and argument meaning; the server side runs on the GPU (opir_tree_*).
#pragma once
#include <cstddef>
#include <cstdint>
#include <functional>
#include <memory>
#include <vector>

#include "onionpir_host.h"

struct opir_tree;

namespace onionpir {

// src/includes/tree_index.h:23-56 (g = 1, response ring = main ring)
struct TreePirParams {
  size_t N = 0, L = 0, n = 0, g = 1, rho = 0, r = 0, N0 = 0, a = 0, P = 0, B = 0, b = 0;
  size_t ell_beta = 0, ell_gamma = 0, w = 0, h_q = 0, W = 0;
  uint64_t t = 0;
  std::vector<uint64_t> rns_moduli;
};

struct ClientCoordinates {  // tree_index.h:69-73
  size_t alpha = 0;
  std::vector<uint8_t> beta_bits_le, gamma_bits_le;
};

enum class SelectCase { Single, CoarsenedAlpha, AlphaBeta };

struct LevelPlan {  // tree_index.h:83-99
  size_t level = 0, R = 0;
  SelectCase select_case = SelectCase::Single;
  size_t coarsen_count = 0, beta_begin = 0, beta_count = 0, gamma_begin = 0, gamma_count = 0, projection_depth = 0;
};

using TreeNodeSource = std::function<uint64_t(size_t level, size_t index)>;

TreePirParams make_tree_pir_params_for_scheme(size_t L, size_t a, const PirParams &scheme);  // tree_query.cpp:44-61
PirParams tree_query_expansion_params(const TreePirParams &tree, const PirParams &scheme);   // tree_query.cpp:63-67
ClientCoordinates client_index(size_t leaf, const TreePirParams &params);                  // tree_index.cpp:150-170
std::vector<LevelPlan> build_level_plans(const TreePirParams &params);                     // tree_index.cpp:172-215
uint64_t synthetic_tree_node_value(size_t level, size_t index, uint64_t t);                // tree_select.cpp:41-50
// D_l[p][u] = node p + u R_l (tree_select.cpp:52-74)
std::vector<std::vector<uint64_t>> pack_tree_level(size_t level, const TreePirParams &tree,
                                                   const TreeNodeSource &source);
RlweCt make_tree_query(PirClient &client, const PirParams &scheme, const TreePirParams &tree, size_t leaf);  // 140-196

// Server side: every level's packed database resident on the GPU; one query
// selects one node per level (unpack_tree_query + build_alpha_pyramid +
// select_level, tree_query.cpp:198-240, tree_select.cpp:121-238).
class TreePirServer {
 public:
  TreePirServer(const PirParams &scheme, const TreePirParams &tree, std::shared_ptr<GpuScheme> gpu);
  ~TreePirServer();
  TreePirServer(const TreePirServer &) = delete;
  TreePirServer &operator=(const TreePirServer &) = delete;
  void preprocess(const TreeNodeSource &source);  // preprocess_tree_reference (tree_select.cpp:76-84)
  void load_level(size_t level, const std::vector<std::vector<uint64_t>> &plaintexts);
  // select_level for levels 0..L, coefficient form
  std::vector<RlweCt> select_path(size_t client_id, const RlweCt &query);
  const TreePirParams &tree() const { return tree_; }
  std::shared_ptr<GpuScheme> gpu_scheme() const { return gpu_; }

 private:
  PirParams scheme_;
  TreePirParams tree_;
  std::shared_ptr<GpuScheme> gpu_;
  opir_tree *t_ = nullptr;
};

}  // namespace onionpir
