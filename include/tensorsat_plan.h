/*
 * tensorsat_plan.h — C-ABI of the host-side contraction planner (libtsplan.so).
 *
 * Replaces the planning half of the reference's cotengra-based contraction:
 *   tensorsat/contractions/cotengra.py:78-85  ctg.array_contract_tree(slot_wires_list,
 *       out_wires, {w: dim}, optimize=..., ...)            -> tsp_greedy / tsp_hyper
 *   notebooks/02-random-ksat.ipynb cell 32    tree.slice_and_reconfigure(target_size)
 *                                                          -> tsp_slice (+ re-plan)
 *   notebooks/02-random-ksat.ipynb cell 16    t.total_flops/total_write/max_size/peak_size
 *                                                          -> tsp_tree_cost
 *
 * A network is a flat wiring: slot s owns wires slot_wires[slot_offsets[s] ..
 * slot_offsets[s+1]); every wire is a hyperedge with extent wire_dims[w]; out wires
 * are kept to the end.  Paths are in SSA form: leaves are 0..S-1, the i-th merge
 * creates id S+i; path[2i], path[2i+1] are the merged ids.
 *
 * Every function returns 0 on success and a negative TSP_E* code on bad input.
 */
#ifndef TENSORSAT_PLAN_H
#define TENSORSAT_PLAN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TSP_OK 0
#define TSP_EINVAL (-1)   /* malformed network or path (maps to ValueError) */
#define TSP_ENOMEM (-2)

typedef struct {
  int32_t num_slots;
  const int32_t* slot_offsets; /* [num_slots + 1] */
  const int32_t* slot_wires;   /* [slot_offsets[num_slots]] */
  int32_t num_wires;
  const int32_t* wire_dims;    /* [num_wires], each >= 1 */
  int32_t num_out;
  const int32_t* out_wires;    /* [num_out] */
} tsp_network;

/* Cost statistics of a tree, cotengra's definitions (log2 of element counts):
 * flops = sum over merges of prod(dims of all wires of both operands);
 * write = sum of merge output sizes; max = largest merge output; peak = largest sum
 * of live tensor sizes (leaves included) over the merge sequence.  With sliced
 * wires, sliced wires are removed from every tensor and flops/write are multiplied
 * by the number of slices. */
typedef struct {
  double log2_flops;
  double log2_write;
  double log2_max_size;
  double log2_peak;
  double num_slices;
} tsp_cost;

/* One greedy run.  costmod weighs the input sizes in the merge score
 * size(out) - costmod * (size(a) + size(b)); temperature > 0 samples among the 8
 * best candidates with Boltzmann weights relative to the best score. */
int tsp_greedy(const tsp_network* net, double costmod, double temperature, uint64_t seed,
               int32_t* path_out /* [2*(S-1)] */);

typedef struct {
  int32_t trials;           /* random-greedy trials (>= 1); trial 0 is plain greedy */
  double costmod_min, costmod_max;
  double temperature_min, temperature_max;
  uint64_t seed;
  int32_t threads;          /* <= 0: hardware concurrency */
  double write_weight;      /* objective = flops + write_weight * write */
  double log2_size_limit;   /* trees whose max size exceeds this are ranked after all others */
} tsp_hyper_opts;

/* Hyper-optimised random greedy: best tree of `trials` runs by the objective. */
int tsp_hyper(const tsp_network* net, const tsp_hyper_opts* opts, int32_t* path_out,
              tsp_cost* cost_out);

/* Cost model of an SSA path with optional sliced wires. */
int tsp_tree_cost(const tsp_network* net, const int32_t* path, int32_t num_sliced,
                  const int32_t* sliced, tsp_cost* cost_out);

/* Greedy slicing: choose non-out wires to fix until the largest intermediate of the
 * path is <= 2^target_log2_size (or max_sliced wires are used).  Each choice minimises
 * the total sliced flops among the wires of the current largest intermediates. */
int tsp_slice(const tsp_network* net, const int32_t* path, double target_log2_size,
              int32_t max_sliced, int32_t* sliced_out, int32_t* num_sliced_out,
              tsp_cost* cost_out);

#ifdef __cplusplus
}
#endif

#endif /* TENSORSAT_PLAN_H */
