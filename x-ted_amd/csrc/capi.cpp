// C ABI implementation (include/xted.h) over the host engine.
#include "../../include/xted.h"

#include <algorithm>
#include <climits>
#include <cstring>
#include <mutex>
#include <exception>
#include <new>
#include <vector>

#include "engine.hpp"
#include "mapping.hpp"
#include "server.hpp"
#include "trees.hpp"

using namespace xted;

struct xted_context {
  Context ctx;
  // Serialises the entry points on one context: xted_ted and
  // xted_edit_mapping use the pool as scratch over several engine calls.
  std::recursive_mutex api;
  explicit xted_context(int d) : ctx(d) {}
};

namespace {

template <class F>
int guarded(F&& f) {
  try {
    return f();
  } catch (const std::bad_alloc&) {
    set_error("host allocation failed");
    return ERR_OUT_OF_MEMORY;
  } catch (const std::exception& e) {
    set_error(e.what());
    return ERR_INVALID_ARG;
  } catch (...) {
    set_error("unknown failure");
    return ERR_INVALID_ARG;
  }
}

void to_c(const BatchStats& s, xted_stats* out) {
  if (!out) return;
  out->plan_ms = s.plan_ms;
  out->dp_ms = s.dp_ms;
  out->total_ms = s.total_ms;
  out->cells = s.cells;
  out->tables = s.tables;
  out->levels = s.levels;
  out->launches = s.launches;
  out->xl_stale_reads = s.xl_stale;
  out->xl_waits = s.xl_waits;
}

}  // namespace

extern "C" {

int xted_create(int32_t device, xted_context** out) {
  return guarded([&]() -> int {
    if (!out) {
      set_error("xted_create: null output");
      return ERR_INVALID_ARG;
    }
    *out = nullptr;
    auto* c = new xted_context(device);
    int st = c->ctx.init();
    if (st != OK) {
      delete c;
      return st;
    }
    *out = c;
    return OK;
  });
}

int xted_destroy(xted_context* ctx) {
  delete ctx;
  return OK;
}

const char* xted_last_error(void) { return last_error(); }

const char* xted_status_string(int32_t status) {
  switch (status) {
    case OK: return "ok";
    case ERR_INVALID_ARG: return "invalid argument";
    case ERR_NOT_PREORDER: return "tree is not a preorder parent array";
    case ERR_OUT_OF_MEMORY: return "out of memory";
    case ERR_DEVICE: return "device error";
    case ERR_TOO_LARGE: return "input too large";
    case ERR_PARSE: return "parse error";
    case ERR_BUSY: return "busy";
    case ERR_TIMEOUT: return "timeout";
    case ERR_NOT_FOUND: return "not found";
    case ERR_CANCELLED: return "cancelled";
    default: return "unknown status";
  }
}

int xted_ted(xted_context* ctx, const int32_t* x_parent, const int32_t* x_label, int32_t m,
             const int32_t* y_parent, const int32_t* y_label, int32_t n, int32_t* td,
             int32_t* dist, xted_stats* stats) {
  return guarded([&]() -> int {
    if (!ctx) return ERR_INVALID_ARG;
    std::lock_guard<std::recursive_mutex> api(ctx->api);
    if (!ctx || !dist || m < 0 || n < 0) {
      set_error("xted_ted: invalid argument");
      return ERR_INVALID_ARG;
    }
    if (stats) *stats = xted_stats{};
    Context& c = ctx->ctx;
    // X-TED_GPU/TED.cu:119-124; under a cost model the distance to an empty
    // tree is the cost of inserting (deleting) every node of the other.
    if (m == 0 || n == 0) {
      if (!c.cost_mode()) {
        *dist = m == 0 ? n : m;
        return OK;
      }
      const int32_t* lab = m == 0 ? y_label : x_label;
      if (m + n > 0 && !lab) {
        set_error("xted_ted: null tree array");
        return ERR_INVALID_ARG;
      }
      const int64_t s = c.whole_tree_cost(lab, m + n, m > 0);
      if (s < 0 || s > INT32_MAX) {
        set_error(s < 0 ? "xted_ted: label outside the cost alphabet" : "xted_ted: cost overflow");
        return s < 0 ? ERR_INVALID_ARG : ERR_TOO_LARGE;
      }
      *dist = (int32_t)s;
      return OK;
    }
    if (!x_parent || !x_label || !y_parent || !y_label) {
      set_error("xted_ted: null tree array");
      return ERR_INVALID_ARG;
    }
    int st = c.pool_clear();
    if (st != OK) return st;
    std::vector<int32_t> sizes{m, n};
    std::vector<int32_t> par(m + n), lab(m + n);
    std::copy(x_parent, x_parent + m, par.begin());
    std::copy(y_parent, y_parent + n, par.begin() + m);
    std::copy(x_label, x_label + m, lab.begin());
    std::copy(y_label, y_label + n, lab.begin() + m);
    int32_t ids[2];
    st = c.pool_add(2, sizes.data(), par.data(), lab.data(), ids);
    if (st != OK) return st;
    int32_t* tds[1] = {td};
    BatchStats bs;
    st = c.run_batch(1, &ids[0], &ids[1], dist, td ? tds : nullptr, &bs);
    to_c(bs, stats);
    return st;
  });
}

int xted_edit_mapping(xted_context* ctx, const int32_t* x_parent, const int32_t* x_label,
                      int32_t m, const int32_t* y_parent, const int32_t* y_label, int32_t n,
                      int32_t* map_x, int32_t* map_y, int32_t* nmap, int32_t* dist) {
  return guarded([&]() -> int {
    if (!ctx) return ERR_INVALID_ARG;
    std::lock_guard<std::recursive_mutex> api(ctx->api);
    if (!ctx || !nmap || !dist || m < 0 || n < 0 || (m > 0 && n > 0 && (!map_x || !map_y))) {
      set_error("xted_edit_mapping: invalid argument");
      return ERR_INVALID_ARG;
    }
    *nmap = 0;
    if (ctx->ctx.cost_mode()) {
      set_error("xted_edit_mapping: unit costs only");
      return ERR_INVALID_ARG;
    }
    if (m == 0 || n == 0) {  // X-TED_GPU/TED.cu:119-124
      *dist = m + n;
      return OK;
    }
    if ((int64_t)m * n > kMaxMappingCells) {
      set_error("xted_edit_mapping: m*n above 2^26");
      return ERR_TOO_LARGE;
    }
    std::vector<int32_t> td((size_t)m * n);
    int st = xted_ted(ctx, x_parent, x_label, m, y_parent, y_label, n, td.data(), dist, nullptr);
    if (st != OK) return st;
    std::vector<std::pair<int32_t, int32_t>> map;
    const int64_t cost = edit_mapping(x_parent, x_label, m, y_parent, y_label, n, td.data(), map);
    if (cost != *dist) {
      set_error("xted_edit_mapping: traced mapping costs " + std::to_string(cost) +
                ", the table says " + std::to_string(*dist));
      return ERR_DEVICE;
    }
    for (size_t k = 0; k < map.size(); ++k) {
      map_x[k] = map[k].first;
      map_y[k] = map[k].second;
    }
    *nmap = (int32_t)map.size();
    return OK;
  });
}

int xted_mapping_from_td(const int32_t* x_parent, const int32_t* x_label, int32_t m,
                         const int32_t* y_parent, const int32_t* y_label, int32_t n,
                         const int32_t* td, int32_t* map_x, int32_t* map_y, int32_t* nmap,
                         int64_t* cost) {
  return guarded([&]() -> int {
    if (!nmap || !cost || m < 0 || n < 0 ||
        (m > 0 && n > 0 && (!td || !map_x || !map_y || !x_parent || !x_label || !y_parent ||
                            !y_label))) {
      set_error("xted_mapping_from_td: invalid argument");
      return ERR_INVALID_ARG;
    }
    if ((int64_t)m * n > kMaxMappingCells) {
      set_error("xted_mapping_from_td: m*n above 2^26");
      return ERR_TOO_LARGE;
    }
    int st;
    if ((m > 0 && (st = validate_preorder(x_parent, m)) != OK) ||
        (n > 0 && (st = validate_preorder(y_parent, n)) != OK))
      return st;
    std::vector<std::pair<int32_t, int32_t>> map;
    *cost = edit_mapping(x_parent, x_label, m, y_parent, y_label, n, td, map);
    for (size_t k = 0; k < map.size(); ++k) {
      map_x[k] = map[k].first;
      map_y[k] = map[k].second;
    }
    *nmap = (int32_t)map.size();
    return OK;
  });
}

int xted_tree_structure(xted_context* ctx, const int32_t* parent, int32_t n, int32_t* orl,
                        int32_t* kr, int32_t* K, int32_t* kr_depth) {
  return guarded([&]() -> int {
    if (!ctx) return ERR_INVALID_ARG;
    std::lock_guard<std::recursive_mutex> api(ctx->api);
    if (!ctx || !parent || n <= 0) {
      set_error("xted_tree_structure: invalid argument");
      return ERR_INVALID_ARG;
    }
    Context& c = ctx->ctx;
    int st = c.pool_clear();
    if (st != OK) return st;
    std::vector<int32_t> lab(n, 0);
    int32_t id = -1;
    st = c.pool_add(1, &n, parent, lab.data(), &id);
    if (st != OK) return st;
    return c.tree_arrays(id, orl, kr, K, kr_depth);
  });
}

int xted_pool_add(xted_context* ctx, int32_t ntrees, const int32_t* sizes,
                  const int32_t* parents, const int32_t* labels, int32_t* ids) {
  return guarded([&]() -> int {
    if (!ctx) return ERR_INVALID_ARG;
    std::lock_guard<std::recursive_mutex> api(ctx->api);
    if (!ctx) {
      set_error("xted_pool_add: null context");
      return ERR_INVALID_ARG;
    }
    return ctx->ctx.pool_add(ntrees, sizes, parents, labels, ids);
  });
}

int xted_pool_clear(xted_context* ctx) {
  if (!ctx) return ERR_INVALID_ARG;
  std::lock_guard<std::recursive_mutex> api(ctx->api);
  return ctx->ctx.pool_clear();
}

int xted_pool_size(xted_context* ctx, int32_t* ntrees) {
  if (!ctx || !ntrees) return ERR_INVALID_ARG;
  std::lock_guard<std::recursive_mutex> api(ctx->api);
  *ntrees = ctx->ctx.pool_size();
  return OK;
}

int xted_batch(xted_context* ctx, int32_t npairs, const int32_t* x_ids, const int32_t* y_ids,
               int32_t* dist, int32_t* const* td_out, xted_stats* stats) {
  return guarded([&]() -> int {
    if (!ctx) return ERR_INVALID_ARG;
    std::lock_guard<std::recursive_mutex> api(ctx->api);
    if (!ctx) {
      set_error("xted_batch: null context");
      return ERR_INVALID_ARG;
    }
    BatchStats bs;
    int st = ctx->ctx.run_batch(npairs, x_ids, y_ids, dist, td_out, &bs);
    to_c(bs, stats);
    return st;
  });
}

int xted_batch_noalloc(xted_context* ctx, int32_t npairs, const int32_t* x_ids,
                       const int32_t* y_ids, int32_t* dist) {
  return guarded([&]() -> int {
    if (!ctx) return ERR_INVALID_ARG;
    std::lock_guard<std::recursive_mutex> api(ctx->api);
    return ctx->ctx.run_batch_noalloc(npairs, x_ids, y_ids, dist);
  });
}

int xted_all_pairs(xted_context* ctx, int32_t ntrees, const int32_t* sizes,
                   const int32_t* parents, const int32_t* labels, int64_t max_bytes,
                   int32_t* dist, xted_stats* stats) {
  return guarded([&]() -> int {
    if (!ctx) return ERR_INVALID_ARG;
    std::lock_guard<std::recursive_mutex> api(ctx->api);
    if (ntrees < 0 || (ntrees > 0 && (!sizes || !parents || !labels || !dist))) {
      set_error("xted_all_pairs: invalid argument");
      return ERR_INVALID_ARG;
    }
    if (stats) *stats = xted_stats{};
    if (ntrees == 0) return OK;
    if (max_bytes <= 0) max_bytes = int64_t{16} << 30;
    Context& c = ctx->ctx;
    int st = c.pool_clear();
    if (st != OK) return st;
    std::vector<int32_t> ids(ntrees);
    st = c.pool_add(ntrees, sizes, parents, labels, ids.data());
    if (st != OK) return st;
    const bool sym = !c.cost_mode();
    const int64_t nn = (int64_t)ntrees * ntrees;
    std::fill(dist, dist + nn, 0);
    // Pairs in row order, cut into chunks by their device bytes (TD + TD^T
    // and the plan arrays).
    std::vector<int32_t> xs, ys, out;
    int64_t bytes = 0;
    auto flush = [&]() -> int {
      if (xs.empty()) return OK;
      out.resize(xs.size());
      BatchStats bs;
      const int r = c.run_batch((int32_t)xs.size(), xs.data(), ys.data(), out.data(), nullptr, &bs);
      if (r != OK) return r;
      for (size_t k = 0; k < xs.size(); ++k) {
        const int32_t i = xs[k] - ids[0], j = ys[k] - ids[0];
        dist[(int64_t)i * ntrees + j] = out[k];
        if (sym) dist[(int64_t)j * ntrees + i] = out[k];
      }
      if (stats) {
        stats->plan_ms += bs.plan_ms;
        stats->dp_ms += bs.dp_ms;
        stats->total_ms += bs.total_ms;
        stats->cells += bs.cells;
        stats->tables += bs.tables;
        stats->levels = std::max<int32_t>(stats->levels, bs.levels);
        stats->launches += bs.launches;
        stats->xl_stale_reads += bs.xl_stale;
        stats->xl_waits += bs.xl_waits;
      }
      xs.clear();
      ys.clear();
      bytes = 0;
      return OK;
    };
    for (int32_t i = 0; i < ntrees; ++i) {
      for (int32_t j = sym ? i + 1 : 0; j < ntrees; ++j) {
        const int64_t b = c.pair_device_bytes(ids[i], ids[j]);
        if (!xs.empty() && (bytes + b > max_bytes || xs.size() >= (size_t)(1 << 20))) {
          st = flush();
          if (st != OK) return st;
        }
        xs.push_back(ids[i]);
        ys.push_back(ids[j]);
        bytes += b;
      }
    }
    return flush();
  });
}

namespace {
// Copies s with its NUL into out when it fits; *need gets the size.
int put_text(const std::string& s, char* out, int64_t cap, int64_t* need) {
  const int64_t n = (int64_t)s.size() + 1;
  if (need) *need = n;
  if (!out || cap < n) return ERR_TOO_LARGE;
  std::memcpy(out, s.c_str(), (size_t)n);
  return OK;
}
}  // namespace

int xted_bracket_to_lines(const char* bracket, char* labels_line, int64_t labels_cap,
                          char* adj_line, int64_t adj_cap, int64_t* labels_need,
                          int64_t* adj_need) {
  return guarded([&]() -> int {
    if (!bracket) return ERR_INVALID_ARG;
    std::vector<std::string> labels;
    std::vector<int32_t> parent;
    if (parse_bracket(bracket, labels, parent) != OK) {
      set_error("xted_bracket_to_lines: not a bracket-notation tree");
      return ERR_PARSE;
    }
    const std::string ls = format_labels(labels);
    const std::string as =
        format_adjacency(parent_to_adj(parent.data(), (int32_t)parent.size()));
    if (labels_need) *labels_need = (int64_t)ls.size() + 1;
    if (adj_need) *adj_need = (int64_t)as.size() + 1;
    if (!labels_line || !adj_line || labels_cap < (int64_t)ls.size() + 1 ||
        adj_cap < (int64_t)as.size() + 1) {
      set_error("xted_bracket_to_lines: output buffer too small");
      return ERR_TOO_LARGE;
    }
    put_text(ls, labels_line, labels_cap, nullptr);
    put_text(as, adj_line, adj_cap, nullptr);
    return OK;
  });
}

int xted_lines_to_bracket(const char* labels_line, const char* adj_line, char* bracket,
                          int64_t cap, int64_t* need) {
  return guarded([&]() -> int {
    if (!labels_line || !adj_line) return ERR_INVALID_ARG;
    std::vector<std::vector<int>> adj;
    std::vector<int32_t> parent;
    const std::vector<std::string> labels = split_labels(labels_line);
    if (parse_adjacency(adj_line, adj) != OK || adj_to_parent(adj, parent) != OK ||
        labels.size() < adj.size() || adj.empty()) {
      set_error("xted_lines_to_bracket: malformed label or adjacency line");
      return ERR_PARSE;
    }
    for (size_t i = 0; i < adj.size(); ++i) {
      if (labels[i].find_first_of("{} ") != std::string::npos) {
        set_error("xted_lines_to_bracket: a label holds a brace");
        return ERR_PARSE;
      }
    }
    const std::string s = format_bracket(labels, parent.data(), (int32_t)adj.size());
    if (put_text(s, bracket, cap, need) != OK) {
      set_error("xted_lines_to_bracket: output buffer too small");
      return ERR_TOO_LARGE;
    }
    return OK;
  });
}

int xted_set_tuning(xted_context* ctx, int32_t lds_per_wave, int32_t max_cpl,
                    int32_t min_waves, int32_t wide_cols) {
  const bool pow2 = max_cpl == 1 || max_cpl == 2 || max_cpl == 4 || max_cpl == 8 || max_cpl == 16;
  if (!ctx || lds_per_wave < 0 || !pow2 || min_waves < 1 || min_waves > kMaxRowWaves ||
      wide_cols < 0) {
    set_error("xted_set_tuning: invalid argument");
    return ERR_INVALID_ARG;
  }
  ctx->ctx.rule.lds_per_wave = lds_per_wave;
  ctx->ctx.rule.max_cpl = max_cpl;
  ctx->ctx.rule.min_waves = min_waves;
  ctx->ctx.rule.wide_cols = wide_cols;
  return OK;
}

int xted_force_class(xted_context* ctx, int32_t cls, int32_t lane_enable) {
  if (!ctx || cls < -1 || cls == kClassXL || cls >= kNumClasses) {
    set_error("xted_force_class: invalid class");
    return ERR_INVALID_ARG;
  }
  ctx->ctx.force_class = cls;
  ctx->ctx.lane_enable = lane_enable ? 1 : 0;
  return OK;
}

int xted_set_xl_width(xted_context* ctx, int32_t max_block_cols) {
  if (!ctx || max_block_cols < xl_bw(kNumXLForms - 1) || max_block_cols > xl_bw(0)) {
    set_error("xted_set_xl_width: max_block_cols must be " + std::to_string(xl_bw(kNumXLForms - 1)) +
              " to " + std::to_string(xl_bw(0)));
    return ERR_INVALID_ARG;
  }
  int k = 0;
  while (k + 1 < kNumXLForms && xl_bw(k) > max_block_cols) ++k;
  ctx->ctx.rule.xl_first = k;
  return OK;
}

int xted_set_xl_rule(xted_context* ctx, int32_t seq_rows) {
  if (!ctx || seq_rows < 0) {
    set_error("xted_set_xl_rule: invalid argument");
    return ERR_INVALID_ARG;
  }
  ctx->ctx.rule.xl_seq_rows = seq_rows;
  return OK;
}

int xted_set_tail_rule(xted_context* ctx, int32_t tail_pairs) {
  if (!ctx || tail_pairs < 0) {
    set_error("xted_set_tail_rule: invalid argument");
    return ERR_INVALID_ARG;
  }
  ctx->ctx.tail_pairs = tail_pairs;
  return OK;
}

int xted_set_level_streams(xted_context* ctx, int32_t n) {
  if (!ctx || n < 1 || n > kMaxLevelStreams) {
    set_error("xted_set_level_streams: n must be 1 to " + std::to_string(kMaxLevelStreams));
    return ERR_INVALID_ARG;
  }
  ctx->ctx.level_streams = n;
  return OK;
}

int xted_set_batch_groups(xted_context* ctx, int32_t groups, int32_t split_min) {
  if (!ctx || groups < 1 || groups > kMaxLevelStreams || split_min < 1) {
    set_error("xted_set_batch_groups: groups must be 1 to " + std::to_string(kMaxLevelStreams) +
              " and split_min positive");
    return ERR_INVALID_ARG;
  }
  ctx->ctx.batch_groups = groups;
  ctx->ctx.batch_split_min = split_min;
  return OK;
}

int xted_set_size_bands(xted_context* ctx, int32_t mode) {
  if (!ctx) return ERR_INVALID_ARG;
  if (mode < 0 || mode > 2) {
    set_error("set_size_bands: mode must be 0, 1 or 2");
    return ERR_INVALID_ARG;
  }
  ctx->ctx.size_bands = mode;
  return OK;
}

int xted_set_light_lds(xted_context* ctx, int32_t bytes_per_wave) {
  if (!ctx) return ERR_INVALID_ARG;
  if (bytes_per_wave < 1024 || bytes_per_wave > kLdsMaxTable) {
    set_error("set_light_lds: bytes per wave must be in [1024, " + std::to_string(kLdsMaxTable) +
              "]");
    return ERR_INVALID_ARG;
  }
  ctx->ctx.light_per_wave = bytes_per_wave;
  return OK;
}

int xted_set_tiny_forms(xted_context* ctx, int32_t mask) {
  if (!ctx) return ERR_INVALID_ARG;
  if (mask < 0 || mask >= (1 << kNumTinyForms)) {
    set_error("set_tiny_forms: mask must be in [0, " + std::to_string((1 << kNumTinyForms) - 1) +
              "]");
    return ERR_INVALID_ARG;
  }
  ctx->ctx.rule.tiny = mask;
  return OK;
}

int xted_set_level_align(xted_context* ctx, int32_t mode) {
  if (!ctx || mode < 0 || mode > 2) {
    set_error("xted_set_level_align: mode must be 0, 1 or 2");
    return ERR_INVALID_ARG;
  }
  ctx->ctx.level_align = mode;
  return OK;
}

int xted_set_costs(xted_context* ctx, int32_t alphabet, const int32_t* del, const int32_t* ins,
                   const int32_t* ren) {
  return guarded([&]() -> int {
    if (!ctx) return ERR_INVALID_ARG;
    std::lock_guard<std::recursive_mutex> api(ctx->api);
    if (!ctx) return ERR_INVALID_ARG;
    return ctx->ctx.set_costs(alphabet, del, ins, ren);
  });
}

int xted_set_graphs(xted_context* ctx, int32_t on) {
  if (!ctx) return ERR_INVALID_ARG;
  ctx->ctx.use_graphs = on ? 1 : 0;
  return OK;
}

int xted_set_profile(xted_context* ctx, int32_t on) {
  if (!ctx) return ERR_INVALID_ARG;
  ctx->ctx.profile = on ? 1 : 0;
  return OK;
}

int xted_class_profile(xted_context* ctx, int32_t cls, int64_t* launches, int64_t* tables,
                       int64_t* cells, double* ms) {
  if (!ctx || cls < 0 || cls >= kNumClasses) {
    set_error("xted_class_profile: invalid class");
    return ERR_INVALID_ARG;
  }
  const ClassProfile& cp = ctx->ctx.class_profile[cls];
  if (launches) *launches = cp.launches;
  if (tables) *tables = cp.tables;
  if (cells) *cells = cp.cells;
  if (ms) *ms = cp.ms;
  return OK;
}

int xted_launch_profile(xted_context* ctx, int32_t i, int32_t* level, int32_t* cls,
                        int64_t* tables, int64_t* cells, double* start_ms, double* ms,
                        int32_t* count) {
  if (!ctx) {
    set_error("xted_launch_profile: null context");
    return ERR_INVALID_ARG;
  }
  const std::vector<LaunchRecord>& log = ctx->ctx.launch_log;
  if (count) *count = (int32_t)log.size();
  if (i < 0) return OK;
  if (i >= (int32_t)log.size()) {
    set_error("xted_launch_profile: launch index out of range");
    return ERR_INVALID_ARG;
  }
  const LaunchRecord& r = log[i];
  if (level) *level = r.level;
  if (cls) *cls = r.cls;
  if (tables) *tables = r.tables;
  if (cells) *cells = r.cells;
  if (start_ms) *start_ms = r.start_ms;
  if (ms) *ms = r.ms;
  return OK;
}

int xted_server_create(xted_context* ctx, int64_t hbm_budget, int32_t max_batch,
                       xted_server** out) {
  return guarded([&]() -> int {
    if (!ctx || !out || hbm_budget <= 0 || max_batch <= 0) {
      set_error("xted_server_create: invalid argument");
      return ERR_INVALID_ARG;
    }
    *out = reinterpret_cast<xted_server*>(new Server(&ctx->ctx, hbm_budget, max_batch));
    return OK;
  });
}

int xted_server_submit(xted_server* srv, int32_t x_id, int32_t y_id, uint64_t tag) {
  return guarded([&]() -> int {
    if (!srv) return ERR_INVALID_ARG;
    return reinterpret_cast<Server*>(srv)->submit(x_id, y_id, tag);
  });
}

int xted_server_poll(xted_server* srv, xted_result* out, int32_t cap, int32_t timeout_ms,
                     int32_t* count) {
  return guarded([&]() -> int {
    if (!srv || !count || (cap > 0 && !out)) return ERR_INVALID_ARG;
    std::vector<ServerResult> rs;
    int st = reinterpret_cast<Server*>(srv)->poll(cap, timeout_ms, rs);
    for (size_t i = 0; i < rs.size(); ++i) {
      out[i].tag = rs[i].tag;
      out[i].status = rs[i].status;
      out[i].dist = rs[i].dist;
      out[i].latency_ms = rs[i].latency_ms;
      out[i].queue_ms = rs[i].queue_ms;
    }
    *count = (int32_t)rs.size();
    return st;
  });
}

int xted_server_destroy(xted_server* srv) {
  delete reinterpret_cast<Server*>(srv);
  return OK;
}

int xted_generate(int32_t kind, int32_t n, int32_t alphabet, uint64_t seed, int32_t depth_min,
                  int32_t depth_max, int32_t* parent, int32_t* label) {
  return guarded([&]() -> int {
    if (n <= 0 || !parent || !label || kind < 0 || kind > 2) {
      set_error("xted_generate: invalid argument");
      return ERR_INVALID_ARG;
    }
    Tree t = (depth_min > 0 || depth_max > 0)
                 ? generate_banded(kind, n, alphabet, seed, depth_min,
                                   depth_max > 0 ? depth_max : n)
                 : generate(kind, n, alphabet, seed);
    std::copy(t.parent.begin(), t.parent.end(), parent);
    std::copy(t.label.begin(), t.label.end(), label);
    return OK;
  });
}

uint64_t xted_derive_seed(uint64_t config, uint64_t index) { return derive_seed(config, index); }

int xted_validate(const int32_t* parent, int32_t n) { return validate_preorder(parent, n); }

int xted_tree_stats(const int32_t* parent, int32_t n, int64_t* ksum, int32_t* depth,
                    int32_t* nkeyroots) {
  return guarded([&]() -> int {
    int st = validate_preorder(parent, n);
    if (st != OK) return st;
    std::vector<int32_t> orl(n);
    compute_orl(parent, n, orl.data());
    auto kr = keyroots_desc(parent, orl.data(), n);
    if (ksum) *ksum = keyroot_size_sum(orl.data(), kr);
    if (depth) *depth = tree_depth(parent, n);
    if (nkeyroots) *nkeyroots = (int32_t)kr.size();
    return OK;
  });
}

}  // extern "C"
