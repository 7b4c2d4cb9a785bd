// Host-side tree model: preorder parent arrays, structural arrays of the
// Zhang-Shasha preorder formulation, and the seeded synthetic generators.
//
// A tree of n nodes is two int32 arrays in preorder:
//   parent[0] = -1, 0 <= parent[i] < i, and parent[i] is an ancestor-or-self
//   of node i-1 (this is exactly what makes the numbering a preorder);
//   label[i] is an int32 label id (equal ids == equal labels).
// The reference passes the same information as per-node children lists in
// preorder (X-TED_GPU/TED_C++.h:27-28, README.md:31-44).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace xted {

struct Tree {
  std::vector<int32_t> parent;
  std::vector<int32_t> label;
  int32_t size() const { return (int32_t)parent.size(); }
};

// Status codes shared with the C ABI (include/xted.h).
enum Status : int {
  OK = 0,
  ERR_INVALID_ARG = 1,
  ERR_NOT_PREORDER = 2,
  ERR_OUT_OF_MEMORY = 3,
  ERR_DEVICE = 4,
  ERR_TOO_LARGE = 5,
  ERR_PARSE = 6,
  ERR_BUSY = 7,
  ERR_TIMEOUT = 8,
  ERR_NOT_FOUND = 9,
  ERR_CANCELLED = 10,
};

// ---- validation / conversion ---------------------------------------------
// Returns OK when parent[] is a valid preorder parent array.
int validate_preorder(const int32_t* parent, int32_t n);
// Children lists (reference adjacency form) -> parent array. Fails with
// ERR_NOT_PREORDER when the lists do not describe a preorder-numbered tree.
int adj_to_parent(const std::vector<std::vector<int>>& adj, std::vector<int32_t>& parent);
// Parent array -> children lists, children in increasing preorder.
std::vector<std::vector<int>> parent_to_adj(const int32_t* parent, int32_t n);

// ---- structural arrays (host restatement, used for planning and checks) --
// orl[i] = last preorder node of subtree(i) = i + |subtree(i)| - 1.
void compute_orl(const int32_t* parent, int32_t n, int32_t* orl);
// Keyroots in descending preorder: {0} U {i : i is not the last child}.
std::vector<int32_t> keyroots_desc(const int32_t* parent, const int32_t* orl, int32_t n);
// Keyroot nesting height per node (only meaningful at keyroots): 0 when no
// other keyroot lies inside subtree(k), else 1 + max over nested keyroots.
void keyroot_height(const int32_t* parent, const int32_t* orl, int32_t n, int32_t* kd);
// S = sum over keyroots of |subtree(k)|; cells of a pair are S1*S2.
int64_t keyroot_size_sum(const int32_t* orl, const std::vector<int32_t>& kr);
// Deepest stack of pending sibling values in a right-to-left sweep over the
// tree (one entry pushed at each node with a previous sibling, popped at its
// previous sibling): what the cost form of the leaf lines needs per line.
int32_t line_stack_depth(const int32_t* parent, const int32_t* orl, int32_t n);

// ---- seeded generators -----------------------------------------------------
enum GenKind : int { GEN_RRT = 0, GEN_DEEP = 1, GEN_SHALLOW = 2 };
// Builds a preorder tree of exactly n nodes with labels uniform in
// [0, alphabet). Deterministic in (kind, n, alphabet, seed).
Tree generate(int kind, int32_t n, int32_t alphabet, uint64_t seed);
// generate() retried with derived seeds until depth_min <= depth <= depth_max.
Tree generate_banded(int kind, int32_t n, int32_t alphabet, uint64_t seed, int32_t depth_min,
                     int32_t depth_max);
// Depth (edges on the longest root-leaf path).
int32_t tree_depth(const int32_t* parent, int32_t n);
// Seed for (config id, index): splitmix64 mixing, stable across platforms.
uint64_t derive_seed(uint64_t config, uint64_t index);

// ---- reference text formats -------------------------------------------------
// Label line tokenisation of X-TED_GPU/dataset_processing.cu:69-82: a token is
// emitted only when a space is seen, so a trailing token without a space is
// dropped and two spaces yield an empty label.
std::vector<std::string> split_labels(const std::string& line);
// Adjacency line "[[1, 2], [], ...]" -> children lists
// (X-TED_GPU/dataset_processing.cu:7-67). Returns ERR_PARSE on malformed text.
int parse_adjacency(const std::string& line, std::vector<std::vector<int>>& adj);
// Inverse formatters used by the writer.
std::string format_labels(const std::vector<std::string>& labels);
std::string format_adjacency(const std::vector<std::vector<int>>& adj);
// Bracket notation "{a{b}{c}}" <-> labels + parent (preorder).
int parse_bracket(const std::string& text, std::vector<std::string>& labels,
                  std::vector<int32_t>& parent);
std::string format_bracket(const std::vector<std::string>& labels, const int32_t* parent,
                           int32_t n);

}  // namespace xted
