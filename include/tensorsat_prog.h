/*
 * tensorsat_prog.h — C-ABI of the host-side contraction-program compiler (libtsplan.so).
 *
 * A *program* is what the device executors run: one or more *jobs*, each a list of
 * leaf tensors plus a list of pairwise steps over an arena.  It is the compiled form of
 *   tensorsat/contractions/simple.py:134-194   SimpleContraction._new (contract2 args)
 *   tensorsat/contractions/cotengra.py:120-126 tree.contract(...) step loop
 * with every intermediate's axis order and arena offset fixed on the host.
 *
 * Tensors are Boolean/count tensors whose axes all have extent 2 ("bit" wires); the
 * flat index of a rank-r tensor with axes (a_0..a_{r-1}) is sum_i bit(a_i) << (r-1-i).
 * A step computes, for every out index o,
 *     acc(o) = sum over summed index c of lhs[L(o,c)] * rhs[R(o,c)]
 * where L/R place the bits of o and c at the positions given by the step record, and
 * writes epilogue(acc(o)): (acc > 0) for GT0, (acc mod 256) != 0 for MOD256, acc for
 * the counting semiring.
 */
#ifndef TENSORSAT_PROG_H
#define TENSORSAT_PROG_H

#include <stdint.h>

#include "tensorsat_plan.h"

#ifdef __cplusplus
extern "C" {
#endif

#define TS_MAX_BITS 27  /* max out bits and max summed bits of one step (128-byte records) */
#define TS_NOPOS 0xFF   /* operand does not hold this bit */

/* One pairwise step (128 bytes). Offsets are element offsets in the job arena. */
typedef struct {
  uint32_t lhs_off, rhs_off, out_off;
  uint8_t lhs_rank, rhs_rank, out_rank, sum_rank;
  uint8_t flags;   /* TS_STEP_* */
  uint8_t pad[3];
  /* For out bit j (LSB-first, j < out_rank): position of that bit in lhs / rhs. */
  uint8_t o_lpos[TS_MAX_BITS], o_rpos[TS_MAX_BITS];
  /* For summed bit j (j < sum_rank): position in lhs / rhs. */
  uint8_t c_lpos[TS_MAX_BITS], c_rpos[TS_MAX_BITS];
} ts_step;

#define TS_STEP_KINNER 1  /* summed bits are the lowest sum_rank bits of both operands, in order */
#define TS_STEP_NORHS 2   /* unary step (rewire/reduce): rhs is the constant 1 */
#define TS_STEP_PACKET 4  /* packet form (tsc_packetize): pad[0] = first lane of this step */
#define TS_STEP_PEND 8    /* packet form: last record of its packet */
#define TS_PACKET_LANES 64 /* lanes of one packet (one wavefront) */
#define TS_PACKET_CHUNK 16 /* packets never straddle a boundary of this many records of a job */

/* One job: a contraction of its leaves down to one tensor. */
typedef struct {
  int32_t instance;      /* instance index (batched decision: verdicts are OR-ed per instance) */
  int32_t cube;          /* index of this job among its instance's jobs */
  int32_t step_begin;    /* first step record */
  int32_t num_steps;
  int64_t leaf_begin;    /* first byte in the leaf buffer */
  int32_t num_leaf_bytes;/* leaves are copied to arena[0 .. num_leaf_bytes) at job start */
  int32_t result_off;    /* element offset of the result tensor in the arena */
  int32_t result_rank;
  int32_t free_vars;     /* counting: result is multiplied by 2^free_vars */
  int64_t arena_elems;   /* arena size in elements */
  int32_t max_rank;      /* largest tensor rank of the job */
  int32_t flags;
} ts_job;

/* ---- compiled programs (opaque handle) ---------------------------------------- */
typedef struct ts_program ts_program;

/* Compile one network + SSA path into a single job.  Wires must all have extent 2 and
 * be unique within each slot; leaves[s] holds 2^rank(s) bytes in slot wire order.
 * Out wires not held by any slot are added as all-ones rank-1 leaves (the
 * spider(t, 1) products of SimpleContraction.contract, simple.py:281-285). */
int tsc_compile(const tsp_network* net, const int32_t* path, const uint8_t* const* leaves,
                ts_program** out);

/* Elimination-order planner: `trials` randomised min-fill/min-degree orders turned into
 * trees; returns the cheapest SSA path (max size first, then flops). */
int tsp_elim(const tsp_network* net, int32_t trials, uint64_t seed, int32_t* path_out,
             tsp_cost* cost_out);

typedef struct {
  int32_t width;        /* cube width W: contract once the residual's width estimate <= W */
  int32_t max_cubes;    /* per-instance cap on emitted jobs (0 = unlimited) */
  int32_t plan_trials;  /* elimination-order trials per cube */
  int32_t threads;      /* host threads (<= 0: hardware concurrency) */
  uint64_t seed;
  int32_t cube_begin;   /* skip (do not compile) the first cube_begin cubes of each instance;
                           max_cubes then counts cubes from cube_begin on */
  const float* priority;/* optional per-variable priorities (learned scores), instances
                           concatenated (NULL: plain planner).  Each cube's tree is proposed
                           by one min-fill pass steered by the priorities (prior_lam), kept
                           only if the exact cost model rates it no larger than one
                           deterministic min-fill pass and no costlier in log2 flops by more
                           than gate_margin; otherwise the plan_trials planner is the
                           fallback.  With prior_branch, branching also takes the free
                           variable of highest priority. */
  float gate_margin;
  int32_t max_nodes;    /* branch-node budget per instance (0: none); exceeding it stops the
                           search with stats[3] = 2 */
  int32_t stop_on_sat;  /* stop an instance at its first cube whose residual is empty
                           (every clause satisfied): stats[5] = 1 */
  int32_t prior_branch; /* with priority: branch on the highest-priority free variable */
  float prior_lam;      /* with priority: weight of the priority against fill (fill units) */
} tsc_sat_opts;

/* Elimination order of the cheapest of `trials` min-fill/min-degree orders (labels for the
 * learned planner).  order_out has num_wires entries (out wires last, unordered). */
int tsp_elim_order(const tsp_network* net, int32_t trials, uint64_t seed, int32_t* order_out,
                   tsp_cost* cost_out);

/* One adaptive min-fill pass steered by a per-wire prior (a learned proposal; no
 * counterpart in the reference): the next wire minimises fill + lam * prior[w], ties to the
 * smaller prior.  Returns the order (all wires) and its tree cost. */
int tsp_elim_order_prior(const tsp_network* net, const float* prior, float lam, int32_t* order_out,
                         tsp_cost* cost_out);

/* Tree cost of an explicit elimination order (all num_wires wires, out wires ignored). */
int tsp_order_cost(const tsp_network* net, const int32_t* order, int32_t* path_out, tsp_cost* cost_out);

#define TSC_NSTAT 10
/* Per-instance statistics copied out by tsc_program_copy:
 * [0] jobs  [1] branch nodes  [2] pruned branches  [3] truncated (cap reached)
 * [4] max tensor rank  [5] decided on the host (-1 no, 0 UNSAT: every branch pruned)
 * [6] host planning microseconds  [7] cubes whose residual is empty (trivially SAT)
 * [8] cube trees taken from the learned proposal  [9] cube trees from the fallback planner */

/* Batched SAT front-end: conditioned slicing of each CNF instance into cubes (branch on
 * variables, propagate unit clauses, prune falsified branches) until the residual
 * formula's width estimate is <= width; each cube's residual clause network is planned
 * and compiled into one job.  Literals are +-(v+1).  Clauses of instance i are
 * clause_offsets[i] .. clause_offsets[i+1]; literals of clause c are lits[lit_offsets[c]
 * .. lit_offsets[c+1]]. */
int tsc_sat_batch(int32_t num_inst, const int32_t* nvars, const int64_t* clause_offsets,
                  const int64_t* lit_offsets, const int32_t* lits, const tsc_sat_opts* opts,
                  ts_program** out);

/* Packet form of a compiled program for one-wave workgroups (no counterpart in the
 * reference: an execution layout of this build).  In place, per job: the steps are
 * regrouped into packets of mutually independent steps sharing TS_PACKET_LANES lanes
 * (step k owns lanes pad[0] .. pad[0] + lanes_k - 1, lanes_k = 2^out_rank, or
 * 2^min(sum_rank, 6) if larger for a step that is neither K-innermost nor sum-free),
 * list-scheduled on true dependencies, packets in order and never straddling a
 * TS_PACKET_CHUNK-record boundary; intermediates are reallocated above the leaves by
 * first fit over their lifetimes in packets (result_off, arena_elems updated), so no
 * step of a packet touches memory another step of it writes.  Every record gets
 * TS_STEP_PACKET, the last of each packet TS_STEP_PEND, and the first of each packet
 * the 64-bit mask of first lanes in o_lpos[16..23] (little endian).  Requires every
 * step to have at most 6 out bits and no packet flags yet (TSP_EINVAL otherwise, jobs
 * already done stay packetized).  threads <= 0: hardware concurrency. */
int tsc_packetize(ts_job* jobs, int64_t num_jobs, ts_step* steps, int64_t num_steps, int32_t threads);

/* Lane form of a packet-form program (no counterpart in the reference: an execution
 * layout of this build): one 16-byte record per lane of every packet, all index math
 * done here, so a wave only loads, counts and stores.  Lane l of a packet runs out index
 * le = l - pad[0] of the step that owns it; its record holds absolute arena offsets. */
typedef struct {
  uint16_t lhs;    /* lhs byte of term 0 for this lane's out index */
  uint16_t rhs;    /* rhs byte of term 0 (lhs again for unary steps) */
  uint16_t out;    /* output byte, TS_LANE_NONE: this lane stores nothing */
  uint16_t tl, tr; /* general steps: offsets of summed term le within lhs / rhs (table entry) */
  uint8_t op;      /* bits 0-2: summed bits (0..6); TS_LANE_* flags */
  uint8_t first;   /* general steps: lane holding term 0 of the step's table */
  uint8_t count;   /* first record of a packet: its lanes (with TS_LANE_SKIP: records to skip) */
  uint8_t pad[3];
} ts_lane;

#define TS_LANE_NONE 0xFFFF
#define TS_LANE_GENERAL 8  /* summed terms through the table (not K-innermost) */
#define TS_LANE_NORHS 16   /* unary step: rhs is the constant 1 */
#define TS_LANE_SKIP 32    /* filler up to the next TS_LANE_CHUNK boundary */
#define TS_LANE_CHUNK 128  /* packets never straddle a boundary of this many lane records */

/* Lane form of packetized jobs (tsc_packetize output; every step at most 6 out and 6
 * summed bits, arenas below 64 KiB).  lane_begin[j] .. lane_begin[j+1] are job j's lane
 * records (lane_begin has num_jobs + 1 entries, lane_begin[0] = 0).  With lanes == NULL
 * only lane_begin is computed; otherwise lanes (cap records) is filled.  TSP_EINVAL for
 * programs outside these limits. */
int tsc_lanes(const ts_job* jobs, int64_t num_jobs, const ts_step* steps, int64_t num_steps, ts_lane* lanes,
              int64_t cap, int64_t* lane_begin, int32_t threads);

int tsc_program_sizes(const ts_program* p, int64_t* num_jobs, int64_t* num_steps,
                      int64_t* num_leaf_bytes, int64_t* num_instances);
int tsc_program_copy(const ts_program* p, ts_job* jobs, ts_step* steps, uint8_t* leaf_bytes,
                     int32_t* inst_stats);
void tsc_program_free(ts_program* p);

/* ---- packing for the device cube search (ts_search_run, tensorsat_amd.h) -------- */

/* Pack `num` CNF instances into the device search's clause words and instance records
 * (16-byte records {clause_begin, num_clauses, num_vars, 0}).  Instance i has
 * nvars[i] variables and nclauses[i] clauses; its clause lengths are lens[i][c] and
 * its literals (+-(v+1)) are lits[i][...] concatenated.  Duplicate literals are
 * dropped, tautologies removed, clause order kept.  words_out holds at least
 * sum(nclauses) words; *total_out receives the number written.  Returns TSP_EINVAL for
 * a literal out of range, TSC_EUNSUPPORTED for an instance outside the kernel's limits
 * (more than 128 variables, more than 768 clauses after normalisation, or a clause of
 * more than 3 distinct literals); *bad_out then names the first such instance.  No
 * counterpart in the reference (it builds Python diagrams, lib/sat.py:248-262). */
#define TSC_EUNSUPPORTED (-3)
int tsc_search_pack(int32_t num, const int32_t* nvars, const int64_t* nclauses, const int32_t* const* lens,
                    const int32_t* const* lits, uint32_t* words_out, int32_t* inst_out, int64_t* total_out,
                    int32_t* bad_out, int32_t threads /* <= 0: hardware concurrency */);

/* Start jobs of a batch for the device search (ts_search_job records, 64 bytes):
 * instance i gets 2^s_i jobs with prefixes 0 .. 2^s_i - 1, where s_i = clamp(ceil(e_i -
 * log2(split_target)), 0, min(max_split, 31)) and e_i = 2.5 + slope(m_i / n_i) n_i
 * estimates log2 of its branch nodes (slope 0.024 at clause ratio <= 3.5 rising linearly
 * to 0.054 at 4.26, then 0.054 * 4.26 / ratio); split_target <= 0: s_i = 0.  Jobs in
 * increasing -(e_i - s_i), ties in instance order, an instance's jobs in prefix order;
 * with split_target > 0, tail_boost > 0 and tail_fraction > 0, an instance whose e_i is at
 * least the k-th largest estimate of the batch, k = clamp(ceil(tail_fraction * num), 1,
 * num), gets s_i = min(s_i + tail_boost, min(max_split, 31)) (the longest searches set a
 * batch's time); tail_fraction outside [0, 1] or tail_boost outside 0..31: TSP_EINVAL;
 * own_from = 1 + the index of the prefix's highest set bit (0 for prefix 0).  inst:
 * `num` records of 4 int32 (clause_begin, num_clauses, num_vars, reserved), as
 * tsc_search_pack writes them.  *njobs_out = the job count; TSP_ENOMEM (nothing written)
 * when it exceeds `capacity`.  No counterpart in the reference. */
int tsc_search_plan(int32_t num, const int32_t* inst, double split_target, int32_t max_split,
                    double tail_fraction, int32_t tail_boost, void* jobs_out,
                    int64_t capacity, int64_t* njobs_out);

#ifdef __cplusplus
}
#endif

#endif /* TENSORSAT_PROG_H */
