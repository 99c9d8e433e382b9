// Shared host-side planner structures (libtsplan.so internals).
#pragma once

#include <cstdint>
#include <random>
#include <climits>
#include <vector>

#include "../../include/tensorsat_plan.h"
#include "../../include/tensorsat_prog.h"

namespace tsp {

// A network with extent-2 wires only ("bit network"), unique wires per slot.
struct BitNet {
  int S = 0, W = 0;
  std::vector<std::vector<int>> slot_wires;  // slot order = leaf layout before compilation
  std::vector<char> is_out;
  std::vector<int> out_wires;  // final layout of the result (unique)
};

struct TreeCost {
  double log2_flops = 0, log2_write = 0, log2_max = 0, log2_peak = 0;
};

// Merge-rule legs of every node of an SSA path (leaves 0..S-1, merge i -> S+i).
// Returns false if the path is malformed.
bool path_legs(const BitNet& net, const std::vector<int32_t>& path,
               std::vector<std::vector<int>>& legs);

TreeCost tree_cost(const BitNet& net, const std::vector<int32_t>& path);

// Elimination-order planning: `trials` randomised min-degree / min-fill orders
// (trial 0 deterministic min-fill), each turned into a tree by merging the tensors of
// each eliminated wire smallest-first.  Returns the cheapest (max size, then flops).
void elim_plan(const BitNet& net, int trials, uint64_t seed, std::vector<int32_t>& path,
               TreeCost* cost, std::vector<int>* best_order = nullptr);

// Tree and cost of a given elimination order (wires not listed are merged at the end).
void order_plan(const BitNet& net, const std::vector<int>& order, std::vector<int32_t>& path,
                TreeCost* cost);

// One adaptive min-fill pass steered by a per-wire prior (a learned proposal): the next
// wire minimises fill + lam * prior, ties to the smaller prior.
void prior_plan(const BitNet& net, const float* prior, float lam, std::vector<int32_t>& path, TreeCost* cost,
                std::vector<int>* order_out = nullptr);

// Width estimate (min-degree heuristic) of the primal graph of a clause list over
// variables 0..nv-1.  The pass stops once the width exceeds stop_above: results above it
// are only known to be above it.
int mindeg_width(int nv, const std::vector<std::vector<int>>& clauses_vars, int stop_above = INT_MAX);

// Compiled job.
struct Job {
  std::vector<ts_step> steps;
  std::vector<uint8_t> leaf_bytes;
  int64_t arena_elems = 0;
  int32_t result_off = 0, result_rank = 0, max_rank = 0;
};

// Compile a bit network + SSA path + leaf tensors (given in slot wire order, one byte
// per element) into a job.  Returns 0 or a negative TSP_E* code.
int compile_job(const BitNet& net, const std::vector<int32_t>& path,
                const std::vector<const uint8_t*>& leaves, Job& job);

}  // namespace tsp
