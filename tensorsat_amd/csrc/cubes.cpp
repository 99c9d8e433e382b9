// Batched SAT front-end: conditioned slicing of CNF instances into contraction jobs.
//
// Slicing a variable wire fixes its value in every leaf that holds it (the reference's
// Box.__getitem__, tensorsat/diagrams.py:972-996 -> FinRel._getitem fin_rel.py:342-343);
// slice results are OR-ed for decisions (Box.__add__, diagrams.py:965-970) and summed for
// counts.  Two exact tensor identities prune the slices: a clause whose literals are all
// false is the all-zero tensor, so its slice contributes nothing; a clause with one
// unassigned literal left is a one-hot rank-1 tensor, so only one value of that variable
// contributes and it is fixed too (unit propagation).  Slicing stops once the residual
// clause network's width estimate is <= W; the residual is planned (elimination order)
// and compiled into one job.  The enumeration does not depend on contraction results.

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cstring>
#include <stdexcept>
#include <thread>
#include <vector>

#include "planner_internal.h"

static_assert(sizeof(ts_step) == 128, "ts_step must be 128 bytes");
static_assert(sizeof(ts_job) == 56, "ts_job layout");

struct ts_program {
  std::vector<ts_job> jobs;
  // Job c's steps and leaf bytes, concatenated only by tsc_program_copy (straight into
  // the caller's arrays at jobs[c].step_begin / leaf_begin, on several threads).
  std::vector<tsp::Job> parts;
  std::vector<int32_t> stats;  // TSC_NSTAT per instance
  int64_t num_instances = 0;
  int64_t num_steps = 0, num_leaves = 0;
  int threads = 1;
};

namespace tsp {
namespace {

struct Formula {
  int n = 0;
  std::vector<std::vector<int>> clauses;  // literals +-(v+1), distinct vars, no tautologies
  bool has_empty = false;
};

Formula normalise(int n, const int32_t* lits, const int64_t* loff, int64_t c0, int64_t c1) {
  Formula f;
  f.n = n;
  for (int64_t c = c0; c < c1; ++c) {
    std::vector<int> cl(lits + loff[c], lits + loff[c + 1]);
    std::sort(cl.begin(), cl.end());
    cl.erase(std::unique(cl.begin(), cl.end()), cl.end());
    bool taut = false;
    for (size_t i = 0; i < cl.size(); ++i)
      for (size_t j = i + 1; j < cl.size(); ++j)
        if (cl[i] == -cl[j]) taut = true;
    if (taut) continue;
    if (cl.empty()) f.has_empty = true;
    // Keep the caller's literal order for the clause tensor axes.
    std::vector<int> ordered;
    for (int64_t i = loff[c]; i < loff[c + 1]; ++i) {
      int l = lits[i];
      if (std::find(ordered.begin(), ordered.end(), l) == ordered.end()) ordered.push_back(l);
    }
    f.clauses.push_back(ordered);
  }
  return f;
}

struct Enumerator {
  const Formula& f;
  const tsc_sat_opts& o;
  std::vector<int8_t> val;                 // per var: 0 free, +1 true, -1 false
  // Flat (CSR) copies of the occurrence lists (literal index 2v+(neg) -> clauses) and of
  // the clause literals, for the propagation and scoring loops.
  std::vector<int> occ_off, occ_cl, lit_off, lits;
  struct Span {
    const int* b;
    const int* e;
    const int* begin() const { return b; }
    const int* end() const { return e; }
  };
  Span occ(int li) const { return {occ_cl.data() + occ_off[li], occ_cl.data() + occ_off[li + 1]}; }
  Span clause(int c) const { return {lits.data() + lit_off[c], lits.data() + lit_off[c + 1]}; }
  std::vector<int> nsat, nfalse;
  std::vector<int> trail;
  std::vector<int> queue;
  int64_t nodes = 0, pruned = 0, empty_residual = 0;
  bool truncated = false;
  int emitted = 0;
  int seen = 0;  // cubes reached in depth-first order (compiled or skipped)
  int max_rank = 0;
  uint64_t seed;
  std::vector<Job>* jobs_out;
  std::vector<int32_t>* free_out;
  const float* prio = nullptr;  // per-variable priorities of this instance (learned planner)
  // Split mode (parallel enumeration): nodes that would branch at depth split_depth are
  // recorded as subtrees (their decision literals) instead of being searched; items
  // keeps the depth-first order of emitted jobs (kind 0) and subtrees (kind 1).
  int split_depth = -1;
  int depth = 0;
  std::vector<int> decisions;
  std::vector<std::vector<int>>* frontier = nullptr;
  std::vector<std::pair<int, int>>* items = nullptr;
  bool budget_hit = false, sat_found = false;
  // search() scratch (see search()).
  std::vector<int> s_rc, s_rv, s_score, s_pos, s_neg, s_vmap;
  std::vector<std::vector<int>> s_cl;
  int64_t learned_trees = 0, fallback_trees = 0;
  // emit() scratch.
  BitNet e_net;
  std::vector<int> e_vmap;
  std::vector<uint8_t> e_leaf;
  std::vector<size_t> e_leaf_off;
  std::vector<const uint8_t*> e_leaves;
  std::vector<int32_t> e_path;

  Enumerator(const Formula& f_, const tsc_sat_opts& o_, uint64_t sd, std::vector<Job>* jobs,
             std::vector<int32_t>* fr, const float* pr)
      : f(f_), o(o_), seed(sd), jobs_out(jobs), free_out(fr), prio(pr) {
    val.assign(f.n, 0);
    occ_off.assign(2 * f.n + 1, 0);
    lit_off.assign(1, 0);
    for (size_t c = 0; c < f.clauses.size(); ++c) {
      for (int l : f.clauses[c]) {
        occ_off[lit_index(l) + 1]++;
        lits.push_back(l);
      }
      lit_off.push_back((int)lits.size());
    }
    for (int i = 0; i < 2 * f.n; ++i) occ_off[i + 1] += occ_off[i];
    occ_cl.resize(occ_off[2 * f.n]);
    std::vector<int> fill(occ_off.begin(), occ_off.end() - 1);
    for (size_t c = 0; c < f.clauses.size(); ++c)
      for (int l : f.clauses[c]) occ_cl[fill[lit_index(l)]++] = (int)c;
    nsat.assign(f.clauses.size(), 0);
    nfalse.assign(f.clauses.size(), 0);
    s_score.assign(f.n, 0);
    s_pos.assign(f.n, 0);
    s_neg.assign(f.n, 0);
    s_vmap.assign(f.n, -1);
  }
  static int lit_index(int l) { return 2 * (std::abs(l) - 1) + (l < 0 ? 1 : 0); }
  bool lit_true(int l) const {
    int8_t v = val[std::abs(l) - 1];
    return l > 0 ? v > 0 : v < 0;
  }

  // Assign literal l true; push units; false on conflict.
  bool assign(int l) {
    int v = std::abs(l) - 1;
    val[v] = l > 0 ? 1 : -1;
    trail.push_back(l);
    bool ok = true;
    for (int c : occ(lit_index(l))) nsat[c]++;
    for (int c : occ(lit_index(-l))) {
      nfalse[c]++;
      if (nsat[c] == 0) {
        int len = lit_off[c + 1] - lit_off[c];
        if (nfalse[c] == len) ok = false;
        else if (nfalse[c] == len - 1) {
          for (int x : clause(c))
            if (val[std::abs(x) - 1] == 0) {
              queue.push_back(x);
              break;
            }
        }
      }
    }
    return ok;
  }
  bool propagate() {
    while (!queue.empty()) {
      int l = queue.back();
      queue.pop_back();
      int8_t cur = val[std::abs(l) - 1];
      if (cur != 0) {
        if (!lit_true(l)) {
          queue.clear();
          return false;
        }
        continue;
      }
      if (!assign(l)) {
        queue.clear();
        return false;
      }
    }
    return true;
  }
  void undo(size_t mark) {
    while (trail.size() > mark) {
      int l = trail.back();
      trail.pop_back();
      for (int c : occ(lit_index(l))) nsat[c]--;
      for (int c : occ(lit_index(-l))) nfalse[c]--;
      val[std::abs(l) - 1] = 0;
    }
  }
  bool initial_units() {
    for (size_t c = 0; c < f.clauses.size(); ++c)
      if (f.clauses[c].size() == 1) queue.push_back(f.clauses[c][0]);
    return propagate();
  }

  // Plan and compile the residual of the current node as one cube.  Returns false (the
  // caller branches further) when the planned tree's largest tensor exceeds 2^width or a
  // step does not fit a record.  Cubes before cube_begin are planned and counted but not
  // kept, so every pass numbers the cubes identically.
  bool emit(const std::vector<int>& rclauses_idx) {
    // Residual: unsatisfied clauses restricted to free variables.  The network, the leaf
    // tables and the path live in members reused by every cube of this enumerator.
    std::vector<int>& vmap = e_vmap;
    vmap.assign(f.n, -1);
    int nv = 0;
    BitNet& net = e_net;
    std::vector<uint8_t>& leaf_data = e_leaf;
    std::vector<size_t>& leaf_off = e_leaf_off;
    leaf_data.clear();
    leaf_off.clear();
    if (net.slot_wires.size() != rclauses_idx.size()) net.slot_wires.resize(rclauses_idx.size());
    for (size_t r = 0; r < rclauses_idx.size(); ++r) {
      std::vector<int>& ws = net.slot_wires[r];
      ws.clear();
      uint64_t zero = 0;
      for (int l : f.clauses[rclauses_idx[r]]) {
        int v = std::abs(l) - 1;
        if (val[v] != 0) continue;
        if (vmap[v] < 0) vmap[v] = nv++;
        ws.push_back(vmap[v]);
        zero = (zero << 1) | (l < 0 ? 1u : 0u);  // the one falsifying assignment, first wire highest
      }
      const size_t base = leaf_data.size();
      leaf_off.push_back(base);
      leaf_data.resize(base + ((size_t)1 << ws.size()), 1);
      leaf_data[base + zero] = 0;
    }
    int free_vars = 0;
    for (int v = 0; v < f.n; ++v)
      if (val[v] == 0 && vmap[v] < 0) ++free_vars;
    net.S = (int)net.slot_wires.size();
    net.W = nv;
    net.is_out.assign(nv, 0);
    std::vector<int32_t>& path = e_path;
    path.clear();
    TreeCost tc;
    if (prio) {
      // Learned proposal: one min-fill pass over the residual steered by the priorities,
      // gated by the exact cost model against one deterministic min-fill pass; the
      // plan_trials planner is the fallback.
      std::vector<float> rp(nv);
      for (int v = 0; v < f.n; ++v)
        if (vmap[v] >= 0) rp[vmap[v]] = prio[v];
      std::vector<int32_t> lpath, ppath;
      TreeCost lc, pc;
      prior_plan(net, rp.data(), o.prior_lam, lpath, &lc);
      elim_plan(net, 1, seed + (uint64_t)seen * 7919u, ppath, &pc);
      const bool take = lc.log2_max <= pc.log2_max && lc.log2_flops <= pc.log2_flops + o.gate_margin;
      if (take) {
        path = lpath;
        tc = lc;
        ++learned_trees;
      } else {
        elim_plan(net, o.plan_trials, seed + (uint64_t)seen * 7919u, path, &tc);
        ++fallback_trees;
      }
    } else {
      elim_plan(net, o.plan_trials, seed + (uint64_t)seen * 7919u, path, &tc);
    }
    if (!rclauses_idx.empty() && tc.log2_max > std::min(o.width, TS_MAX_BITS)) return false;
    std::vector<const uint8_t*>& leaves = e_leaves;
    leaves.clear();
    for (size_t o : leaf_off) leaves.push_back(leaf_data.data() + o);
    Job job;
    if (compile_job(net, path, leaves, job) != TSP_OK) return false;
    if (seen++ < o.cube_begin) return true;  // kept by an earlier pass
    max_rank = std::max(max_rank, job.max_rank);
    jobs_out->push_back(std::move(job));
    free_out->push_back(free_vars);
    if (items) items->push_back({0, (int)jobs_out->size() - 1});
    ++emitted;
    if (rclauses_idx.empty()) {
      ++empty_residual;
      if (o.stop_on_sat) {
        sat_found = true;
        truncated = true;
      }
    }
    if (o.max_cubes > 0 && emitted >= o.max_cubes) truncated = true;
    return true;
  }

  void search() {
    if (truncated) return;
    // Per-node scratch lives in members: everything below is consumed before the node
    // branches (only the first literal is kept across the recursion), so one set of
    // buffers serves every depth.
    std::vector<int>& rc = s_rc;
    std::vector<int>& rv = s_rv;
    rc.clear();
    rv.clear();
    // Residual clauses, variables and degrees in one pass (score/occ_pos/occ_neg are zero
    // outside rv).  A residual clause has no true literal, so its assigned literals are
    // all false and it has len - nfalse free literals.
    std::vector<int>& score = s_score;
    std::vector<int>& occ_pos = s_pos;
    std::vector<int>& occ_neg = s_neg;
    const int nclauses = (int)f.clauses.size();
    for (int c = 0; c < nclauses; ++c) {
      if (nsat[c] != 0) continue;
      rc.push_back(c);
      const int free_in = lit_off[c + 1] - lit_off[c] - nfalse[c];
      for (int l : clause(c)) {
        int v = std::abs(l) - 1;
        if (val[v] != 0) continue;
        if (score[v] == 0 && occ_pos[v] == 0 && occ_neg[v] == 0) rv.push_back(v);
        score[v] += free_in - 1;
        if (l > 0) occ_pos[v]++;
        else occ_neg[v]++;
      }
    }
    if (rc.empty()) {
      emit(rc);
      return;
    }
    bool leaf = (int)rv.size() <= o.width;
    if (!leaf && (int)rv.size() <= 4 * o.width + 8) {
      std::vector<int>& vmap = s_vmap;
      for (size_t i = 0; i < rv.size(); ++i) vmap[rv[i]] = (int)i;
      std::vector<std::vector<int>>& cl = s_cl;
      cl.resize(rc.size());
      for (size_t i = 0; i < rc.size(); ++i) {
        std::vector<int>& ws = cl[i];
        ws.clear();
        for (int l : clause(rc[i])) {
          int v = std::abs(l) - 1;
          if (val[v] == 0) ws.push_back(vmap[v]);
        }
      }
      for (int v : rv) vmap[v] = -1;
      // The estimate only filters; the planned tree decides (emit).
      leaf = mindeg_width((int)rv.size(), cl, o.width + 1) <= o.width + 1;
    }
    int bv = -1, first = 0;
    auto pick = [&]() {
      if (rv.empty()) return;
      bv = rv[0];
      if (prio && o.prior_branch) {
        for (int v : rv)
          if (prio[v] > prio[bv] || (prio[v] == prio[bv] && v < bv)) bv = v;
      } else {
        for (int v : rv)
          if (score[v] > score[bv] || (score[v] == score[bv] && v < bv)) bv = v;
      }
      first = occ_pos[bv] >= occ_neg[bv] ? bv + 1 : -(bv + 1);
      for (int v : rv) score[v] = occ_pos[v] = occ_neg[v] = 0;
    };
    if (leaf) {
      // emit() reuses none of this scratch, but it may return false (branch further):
      // take the branch literal first so the scratch is clean either way.
      pick();
      if (emit(rc)) return;
    } else {
      pick();
    }
    if (o.max_nodes > 0 && nodes >= o.max_nodes) {
      budget_hit = true;
      truncated = true;
      return;
    }
    if (split_depth >= 0 && depth >= split_depth) {
      items->push_back({1, (int)frontier->size()});
      frontier->push_back(decisions);
      return;
    }
    ++nodes;
    for (int l : {first, -first}) {
      size_t mark = trail.size();
      if (assign(l) && propagate()) {
        decisions.push_back(l);
        ++depth;
        search();
        --depth;
        decisions.pop_back();
      } else {
        ++pruned;
      }
      queue.clear();
      undo(mark);
      if (truncated) return;
    }
  }
};

}  // namespace
}  // namespace tsp

using namespace tsp;

extern "C" {

int tsp_elim(const tsp_network* n, int32_t trials, uint64_t seed, int32_t* path_out,
             tsp_cost* cost_out) {
  if (!n || n->num_slots < 0 || n->num_wires < 0) return TSP_EINVAL;
  BitNet net;
  net.S = n->num_slots;
  net.W = n->num_wires;
  net.is_out.assign(net.W, 0);
  for (int i = 0; i < n->num_out; ++i) {
    int w = n->out_wires[i];
    if (w < 0 || w >= net.W) return TSP_EINVAL;
    net.is_out[w] = 1;
  }
  for (int s = 0; s < net.S; ++s) {
    std::vector<int> ws(n->slot_wires + n->slot_offsets[s], n->slot_wires + n->slot_offsets[s + 1]);
    for (int w : ws)
      if (w < 0 || w >= net.W) return TSP_EINVAL;
    std::sort(ws.begin(), ws.end());
    ws.erase(std::unique(ws.begin(), ws.end()), ws.end());
    net.slot_wires.push_back(ws);
  }
  std::vector<int32_t> path;
  TreeCost tc;
  try {
    elim_plan(net, trials, seed, path, &tc);
  } catch (const std::bad_alloc&) {
    return TSP_ENOMEM;
  }
  if (!path.empty()) {
    if (!path_out) return TSP_EINVAL;
    std::memcpy(path_out, path.data(), path.size() * sizeof(int32_t));
  }
  if (cost_out) {
    cost_out->log2_flops = tc.log2_flops;
    cost_out->log2_write = tc.log2_write;
    cost_out->log2_max_size = tc.log2_max;
    cost_out->log2_peak = tc.log2_peak;
    cost_out->num_slices = 1;
  }
  return TSP_OK;
}

static int to_bitnet(const tsp_network* n, BitNet& net) {
  if (!n || n->num_slots < 0 || n->num_wires < 0 || n->num_out < 0) return TSP_EINVAL;
  net.S = n->num_slots;
  net.W = n->num_wires;
  net.is_out.assign(net.W, 0);
  for (int i = 0; i < n->num_out; ++i) {
    int w = n->out_wires[i];
    if (w < 0 || w >= net.W) return TSP_EINVAL;
    net.is_out[w] = 1;
  }
  for (int s = 0; s < net.S; ++s) {
    std::vector<int> ws(n->slot_wires + n->slot_offsets[s], n->slot_wires + n->slot_offsets[s + 1]);
    for (int w : ws)
      if (w < 0 || w >= net.W) return TSP_EINVAL;
    std::sort(ws.begin(), ws.end());
    ws.erase(std::unique(ws.begin(), ws.end()), ws.end());
    net.slot_wires.push_back(ws);
  }
  return TSP_OK;
}

static void fill_cost(const TreeCost& tc, tsp_cost* c) {
  if (!c) return;
  c->log2_flops = tc.log2_flops;
  c->log2_write = tc.log2_write;
  c->log2_max_size = tc.log2_max;
  c->log2_peak = tc.log2_peak;
  c->num_slices = 1;
}

int tsp_elim_order(const tsp_network* n, int32_t trials, uint64_t seed, int32_t* order_out, tsp_cost* cost_out) {
  BitNet net;
  int rc = to_bitnet(n, net);
  if (rc) return rc;
  if (!order_out) return TSP_EINVAL;
  std::vector<int32_t> path;
  std::vector<int> order;
  TreeCost tc;
  try {
    elim_plan(net, trials, seed, path, &tc, &order);
  } catch (const std::bad_alloc&) {
    return TSP_ENOMEM;
  }
  std::vector<char> in(net.W, 0);
  int k = 0;
  for (int w : order) {
    order_out[k++] = w;
    in[w] = 1;
  }
  for (int w = 0; w < net.W; ++w)
    if (!in[w]) order_out[k++] = w;
  fill_cost(tc, cost_out);
  return TSP_OK;
}

int tsp_elim_order_prior(const tsp_network* n, const float* prior, float lam, int32_t* order_out,
                         tsp_cost* cost_out) {
  BitNet net;
  int rc = to_bitnet(n, net);
  if (rc) return rc;
  if (!order_out || !prior) return TSP_EINVAL;
  std::vector<int32_t> path;
  std::vector<int> order;
  TreeCost tc;
  try {
    prior_plan(net, prior, lam, path, &tc, &order);
  } catch (const std::bad_alloc&) {
    return TSP_ENOMEM;
  }
  std::vector<char> in(net.W, 0);
  int k = 0;
  for (int w : order) {
    order_out[k++] = w;
    in[w] = 1;
  }
  for (int w = 0; w < net.W; ++w)
    if (!in[w]) order_out[k++] = w;
  fill_cost(tc, cost_out);
  return TSP_OK;
}

int tsp_order_cost(const tsp_network* n, const int32_t* order, int32_t* path_out, tsp_cost* cost_out) {
  BitNet net;
  int rc = to_bitnet(n, net);
  if (rc) return rc;
  if (!order) return TSP_EINVAL;
  std::vector<int> ord;
  for (int i = 0; i < net.W; ++i) {
    if (order[i] < 0 || order[i] >= net.W) return TSP_EINVAL;
    if (!net.is_out[order[i]]) ord.push_back(order[i]);
  }
  std::vector<int32_t> path;
  TreeCost tc;
  try {
    order_plan(net, ord, path, &tc);
  } catch (const std::bad_alloc&) {
    return TSP_ENOMEM;
  }
  if (path_out && !path.empty()) std::memcpy(path_out, path.data(), path.size() * sizeof(int32_t));
  fill_cost(tc, cost_out);
  return TSP_OK;
}

int tsc_compile(const tsp_network* n, const int32_t* path, const uint8_t* const* leaves,
                ts_program** out) {
  if (!n || !out || n->num_slots < 0 || n->num_wires < 0 || n->num_out < 0) return TSP_EINVAL;
  BitNet net;
  net.S = n->num_slots;
  net.W = n->num_wires;
  net.is_out.assign(net.W, 0);
  for (int w = 0; w < net.W; ++w)
    if (n->wire_dims[w] != 2) return TSP_EINVAL;
  for (int i = 0; i < n->num_out; ++i) {
    int w = n->out_wires[i];
    if (w < 0 || w >= net.W) return TSP_EINVAL;
    net.is_out[w] = 1;
    net.out_wires.push_back(w);
  }
  for (int s = 0; s < net.S; ++s) {
    std::vector<int> ws(n->slot_wires + n->slot_offsets[s], n->slot_wires + n->slot_offsets[s + 1]);
    for (int w : ws)
      if (w < 0 || w >= net.W) return TSP_EINVAL;
    net.slot_wires.push_back(ws);
  }
  std::vector<int32_t> p;
  if (net.S > 1) {
    if (!path) return TSP_EINVAL;
    p.assign(path, path + 2 * (net.S - 1));
  }
  std::vector<const uint8_t*> lv(leaves, leaves + net.S);
  auto* prog = new (std::nothrow) ts_program();
  if (!prog) return TSP_ENOMEM;
  try {
    Job job;
    int rc = compile_job(net, p, lv, job);
    if (rc) {
      delete prog;
      return rc;
    }
    ts_job j{};
    j.instance = 0;
    j.cube = 0;
    j.step_begin = 0;
    j.num_steps = (int32_t)job.steps.size();
    j.leaf_begin = 0;
    j.num_leaf_bytes = (int32_t)job.leaf_bytes.size();
    j.result_off = job.result_off;
    j.result_rank = job.result_rank;
    j.free_vars = 0;
    j.arena_elems = job.arena_elems;
    j.max_rank = job.max_rank;
    prog->jobs.push_back(j);
    prog->num_steps = (int64_t)job.steps.size();
    prog->num_leaves = (int64_t)job.leaf_bytes.size();
    prog->parts.push_back(std::move(job));
    prog->num_instances = 1;
    prog->stats.assign(TSC_NSTAT, 0);
    prog->stats[0] = 1;
    prog->stats[4] = j.max_rank;
    prog->stats[5] = -1;
  } catch (...) {
    delete prog;
    return TSP_ENOMEM;
  }
  *out = prog;
  return TSP_OK;
}

int tsc_sat_batch(int32_t num_inst, const int32_t* nvars, const int64_t* clause_offsets,
                  const int64_t* lit_offsets, const int32_t* lits, const tsc_sat_opts* opts,
                  ts_program** out) {
  if (num_inst < 0 || !opts || !out || (num_inst > 0 && (!nvars || !clause_offsets))) return TSP_EINVAL;
  if (opts->width < 0 || opts->width > TS_MAX_BITS) return TSP_EINVAL;
  struct PerInst {
    std::vector<Job> jobs;
    std::vector<int32_t> free_vars;
    int32_t stats[TSC_NSTAT] = {0};
    bool ok = true;
  };
  std::vector<PerInst> res(num_inst);
  std::vector<int64_t> var_offset(num_inst + 1, 0);
  for (int i = 0; i < num_inst; ++i) {
    if (nvars[i] < 0) return TSP_EINVAL;
    var_offset[i + 1] = var_offset[i] + nvars[i];
  }
  std::atomic<int> next{0};
  std::atomic<int> bad{0};
  int threads = opts->threads > 0 ? opts->threads : (int)std::thread::hardware_concurrency();
  threads = std::max(1, threads);
  // Few instances and a full enumeration: split each instance's search tree at a fixed
  // branch depth and enumerate the subtrees in parallel.  Concatenating the subtrees'
  // jobs in left-to-right order gives exactly the sequential depth-first order.
  const bool full = opts->max_cubes == 0 && opts->cube_begin == 0 && !opts->stop_on_sat && opts->max_nodes == 0;
  const bool intra = full && threads > 1 && num_inst * 2 <= threads;
  auto solve_one = [&](int i, int inner) {
    auto t0 = std::chrono::steady_clock::now();
    PerInst& r = res[i];
    try {
      Formula f = normalise(nvars[i], lits, lit_offsets, clause_offsets[i], clause_offsets[i + 1]);
      for (auto& c : f.clauses)
        for (int l : c)
          if (l == 0 || std::abs(l) > f.n) throw std::invalid_argument("literal");
      const float* pr = opts->priority ? opts->priority + var_offset[i] : nullptr;
      const uint64_t sd = opts->seed + (uint64_t)i * 1000003u;
      std::vector<Job> jobsA;
      std::vector<int32_t> freeA;
      std::vector<std::vector<int>> frontier;
      std::vector<std::pair<int, int>> items;
      Enumerator e(f, *opts, sd, inner > 1 ? &jobsA : &r.jobs, inner > 1 ? &freeA : &r.free_vars, pr);
      if (inner > 1) {
        int d = 0;
        while ((1 << d) < 8 * inner) ++d;
        e.split_depth = d;
        e.frontier = &frontier;
        e.items = &items;
      }
      bool alive = !f.has_empty && e.initial_units();
      if (alive) e.search();
      int64_t nodes = e.nodes, pruned = e.pruned, empty = e.empty_residual, lt = e.learned_trees,
              ft = e.fallback_trees;
      int max_rank = e.max_rank;
      int64_t seen = e.seen;
      if (inner > 1) {
        const int nf = (int)frontier.size();
        std::vector<std::vector<Job>> jobsB(nf);
        std::vector<std::vector<int32_t>> freeB(nf);
        std::vector<std::array<int64_t, 7>> st(nf);
        std::atomic<int> nx{0};
        auto sub = [&]() {
          for (;;) {
            const int k = nx.fetch_add(1);
            if (k >= nf) break;
            Enumerator eb(f, *opts, sd + 7919u * (uint64_t)(k + 1), &jobsB[k], &freeB[k], pr);
            bool ok = eb.initial_units();
            for (int l : frontier[k]) {
              if (!ok) break;
              ok = eb.assign(l) && eb.propagate();
            }
            if (ok) eb.search();
            st[k] = {eb.nodes, eb.pruned + (ok ? 0 : 1), eb.empty_residual, eb.learned_trees, eb.fallback_trees,
                     eb.max_rank, eb.seen};
          }
        };
        std::vector<std::thread> workers;
        for (int t = 1; t < std::min(inner, std::max(1, nf)); ++t) workers.emplace_back(sub);
        sub();
        for (auto& th : workers) th.join();
        for (const auto& it : items) {
          if (it.first == 0) {
            r.jobs.push_back(std::move(jobsA[it.second]));
            r.free_vars.push_back(freeA[it.second]);
          } else {
            auto& jb = jobsB[it.second];
            for (size_t c = 0; c < jb.size(); ++c) {
              r.jobs.push_back(std::move(jb[c]));
              r.free_vars.push_back(freeB[it.second][c]);
            }
            jb.clear();
          }
        }
        for (const auto& x : st) {
          nodes += x[0];
          pruned += x[1];
          empty += x[2];
          lt += x[3];
          ft += x[4];
          max_rank = std::max<int>(max_rank, (int)x[5]);
          seen += x[6];
        }
      }
      r.stats[0] = (int32_t)r.jobs.size();
      r.stats[1] = (int32_t)std::min<int64_t>(nodes, INT32_MAX);
      r.stats[2] = (int32_t)std::min<int64_t>(pruned + (alive ? 0 : 1), INT32_MAX);
      r.stats[3] = e.budget_hit ? 2 : (e.truncated ? 1 : 0);
      r.stats[4] = max_rank;
      // host verdict: 0 every branch pruned (UNSAT), 1 an all-satisfied cube was reached
      // with stop_on_sat, -1 undecided on the host
      r.stats[5] = e.sat_found ? 1 : (seen == 0 && !e.budget_hit ? 0 : -1);
      r.stats[7] = (int32_t)empty;
      r.stats[8] = (int32_t)lt;
      r.stats[9] = (int32_t)ft;
    } catch (const std::invalid_argument&) {
      r.ok = false;
      bad++;
    } catch (...) {
      r.ok = false;
      bad++;
    }
    auto t1 = std::chrono::steady_clock::now();
    r.stats[6] = (int32_t)std::chrono::duration_cast<std::chrono::microseconds>(t1 - t0).count();
  };
  if (intra) {
    for (int i = 0; i < num_inst; ++i) solve_one(i, threads);
  } else {
    auto worker = [&]() {
      for (;;) {
        int i = next.fetch_add(1);
        if (i >= num_inst) break;
        solve_one(i, 1);
      }
    };
    const int nt = std::max(1, std::min(threads, std::max(1, num_inst)));
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; ++t) pool.emplace_back(worker);
    worker();
    for (auto& th : pool) th.join();
  }
  if (bad) return TSP_EINVAL;
  auto* prog = new (std::nothrow) ts_program();
  if (!prog) return TSP_ENOMEM;
  try {
    prog->num_instances = num_inst;
    prog->stats.resize((size_t)num_inst * TSC_NSTAT);
    prog->threads = threads;
    size_t tot_jobs = 0;
    for (const auto& r : res) tot_jobs += r.jobs.size();
    prog->jobs.reserve(tot_jobs);
    prog->parts.reserve(tot_jobs);
    int64_t sb = 0, lb = 0;
    for (int i = 0; i < num_inst; ++i) {
      std::memcpy(&prog->stats[(size_t)i * TSC_NSTAT], res[i].stats, sizeof(res[i].stats));
      for (size_t c = 0; c < res[i].jobs.size(); ++c) {
        Job& jb = res[i].jobs[c];
        ts_job j{};
        j.instance = i;
        j.cube = (int32_t)(opts->cube_begin + (int)c);
        j.step_begin = (int32_t)sb;
        j.num_steps = (int32_t)jb.steps.size();
        j.leaf_begin = lb;
        j.num_leaf_bytes = (int32_t)jb.leaf_bytes.size();
        j.result_off = jb.result_off;
        j.result_rank = jb.result_rank;
        j.free_vars = res[i].free_vars[c];
        j.arena_elems = jb.arena_elems;
        j.max_rank = jb.max_rank;
        prog->jobs.push_back(j);
        sb += (int64_t)jb.steps.size();
        lb += (int64_t)jb.leaf_bytes.size();
        prog->parts.push_back(std::move(jb));
      }
      res[i].jobs.clear();
      res[i].jobs.shrink_to_fit();
    }
    prog->num_steps = sb;
    prog->num_leaves = lb;
  } catch (...) {
    delete prog;
    return TSP_ENOMEM;
  }
  *out = prog;
  return TSP_OK;
}

int tsc_program_sizes(const ts_program* p, int64_t* num_jobs, int64_t* num_steps,
                      int64_t* num_leaf_bytes, int64_t* num_instances) {
  if (!p) return TSP_EINVAL;
  if (num_jobs) *num_jobs = (int64_t)p->jobs.size();
  if (num_steps) *num_steps = p->num_steps;
  if (num_leaf_bytes) *num_leaf_bytes = p->num_leaves;
  if (num_instances) *num_instances = p->num_instances;
  return TSP_OK;
}

int tsc_program_copy(const ts_program* p, ts_job* jobs, ts_step* steps, uint8_t* leaf_bytes,
                     int32_t* inst_stats) {
  if (!p) return TSP_EINVAL;
  if (jobs && !p->jobs.empty()) std::memcpy(jobs, p->jobs.data(), p->jobs.size() * sizeof(ts_job));
  if (steps || leaf_bytes) {
    const int64_t nj = (int64_t)p->parts.size();
    auto copy_jobs = [&](int64_t b, int64_t e) {
      for (int64_t c = b; c < e; ++c) {
        const tsp::Job& jb = p->parts[c];
        const ts_job& j = p->jobs[c];
        if (steps && !jb.steps.empty())
          std::memcpy(steps + j.step_begin, jb.steps.data(), jb.steps.size() * sizeof(ts_step));
        if (leaf_bytes && !jb.leaf_bytes.empty())
          std::memcpy(leaf_bytes + j.leaf_begin, jb.leaf_bytes.data(), jb.leaf_bytes.size());
      }
    };
    // At least 16 MiB per thread: a small program is copied on the calling thread.
    const int64_t bytes = p->num_steps * (int64_t)sizeof(ts_step) + p->num_leaves;
    const int nt = (int)std::max<int64_t>(1, std::min<int64_t>({(int64_t)p->threads, bytes >> 24, nj}));
    if (nt <= 1) {
      copy_jobs(0, nj);
    } else {
      std::vector<std::thread> pool;
      bool spawned = true;
      try {
        for (int t = 1; t < nt; ++t) pool.emplace_back(copy_jobs, nj * t / nt, nj * (t + 1) / nt);
      } catch (...) {
        spawned = false;
      }
      if (!spawned) {  // no threads left: wait for the ones started, then copy everything here
        for (auto& th : pool) th.join();
        pool.clear();
        copy_jobs(0, nj);
      } else {
        copy_jobs(0, nj / nt);
      }
      for (auto& th : pool) th.join();
    }
  }
  if (inst_stats && !p->stats.empty())
    std::memcpy(inst_stats, p->stats.data(), p->stats.size() * sizeof(int32_t));
  return TSP_OK;
}

void tsc_program_free(ts_program* p) { delete p; }

}  // extern "C"
