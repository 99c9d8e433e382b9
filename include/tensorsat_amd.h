/*
 * tensorsat_amd.h — C-ABI of the MI355X (gfx950) executor, libtensorsat_amd.so.
 *
 * Device tensors are Boolean relations (uint8 0/1) or counts (uint64) whose axes all
 * have extent 2 ("bit" axes; wider FinSet axes are padded to powers of two by the host
 * layer).  A tensor of rank r is 2^r contiguous elements in row-major order; axis i is
 * bit (r-1-i) of the flat index.  Wires are small non-negative integers naming the axes
 * of an operation, as in the reference's einsum-style calls.
 *
 * Every function returns TS_OK (0) or a negative TS_E* code; the host layer maps
 *   TS_EINVAL -> ValueError, TS_EUNSUPPORTED -> NotImplementedError,
 *   TS_EOVERFLOW -> OverflowError, TS_EHIP -> RuntimeError.
 * Work is enqueued on `stream` (NULL = the null stream) and is asynchronous unless
 * stated otherwise.  Not thread-safe per stream.
 */
#ifndef TENSORSAT_AMD_H
#define TENSORSAT_AMD_H

#include <stddef.h>
#include <stdint.h>

#include "tensorsat_prog.h"

#ifdef __cplusplus
extern "C" {
#endif

#define TS_OK 0
#define TS_EINVAL (-1)
#define TS_ENOMEM (-2)
#define TS_EUNSUPPORTED (-3)
#define TS_EOVERFLOW (-4)
#define TS_EHIP (-5)

/* Epilogue of a Boolean step on the exact integer count c of an output entry. */
#define TS_MODE_GT0 0     /* c > 0 : relational composition (existential sum) */
#define TS_MODE_MOD256 1  /* (c mod 256) != 0 : what the reference's numpy code computes,
                             fin_rel.py:279,287,392 (np.sign(..., dtype=uint8) after a
                             wrapping integer einsum) */

typedef void* ts_stream; /* hipStream_t */

/* Library version / device check. Returns TS_OK if a gfx950 device is usable. */
int ts_init(int device);
const char* ts_version(void);

/* ---- Box algebra (one kernel launch each) --------------------------------------- */

/* FinRel._contract2 (tensorsat/lang/fin_rel.py:255-288), reached from
 * tensorsat/_autoray.py:31-62 (einsum) and Box.contract2 (diagrams.py:811-866).
 * out[o] = epilogue(sum_c lhs[..] * rhs[..]); out wires must be unique (the host layer
 * handles repeated out wires with ts_rewire); lhs/rhs wires may repeat (diagonals). */
int ts_contract2(const uint8_t* lhs, int32_t lhs_rank, const int32_t* lhs_wires,
                 const uint8_t* rhs, int32_t rhs_rank, const int32_t* rhs_wires,
                 uint8_t* out, int32_t out_rank, const int32_t* out_wires, int32_t mode,
                 ts_stream stream);

/* FinRel._rewire (fin_rel.py:379-393) and rewire_array (_utils/misc.py:104-123):
 * out axes are input ports, possibly repeated (diagonal embedding) or absent
 * (existential sum).  Pure permutations run the permute path. */
int ts_rewire(const uint8_t* in, int32_t in_rank, uint8_t* out, int32_t out_rank,
              const int32_t* out_ports, int32_t mode, ts_stream stream);

/* FinRel._spider (fin_rel.py:290-298): copy tensor of `num_ports` axes, each of
 * `bits` bits, ones where all ports hold the same value v < dim. */
int ts_spider(uint8_t* out, int32_t num_ports, int32_t bits, int32_t dim, ts_stream stream);

/* FinRel._union (fin_rel.py:345-346): out = a | b over n bytes. */
int ts_union(const uint8_t* a, const uint8_t* b, uint8_t* out, int64_t n, ts_stream stream);

/* FinRel._stack (fin_rel.py:304-308): out has rank in_rank + axis_bits; input k fills
 * the slice whose new axis (at position `port`) equals k; slices k >= num are zero. */
int ts_stack(const uint8_t* const* ins, int32_t num, int32_t in_rank, int32_t port,
             int32_t axis_bits, uint8_t* out, ts_stream stream);

/* FinRel._getitem (fin_rel.py:342-343): fix axes to values; fixed[i] is -1 (keep) or
 * the bit value of axis i.  out has rank = number of kept axes. */
int ts_select(const uint8_t* in, int32_t in_rank, const int32_t* fixed, uint8_t* out,
              ts_stream stream);

/* FinRel._as_scalar (fin_rel.py:472-473): all(tensor == 1) over n bytes; synchronous. */
int ts_as_scalar(const uint8_t* x, int64_t n, int32_t* result, ts_stream stream);

/* Tile of the dense int8 MFMA route for steps of one shape (bits of rows held by lhs
 * only, of columns held by rhs only, of shared batch, of the contraction): 1 64x64, 2
 * 128x64, 3 64x128, 4 128x128, 5 256x128, 6 128x256 (rows x columns), 0 the built-in
 * rule (128x256 when both sides have >= 10 bits, 128x128 when both have >= 7, else
 * 64x64).  Set by the tile autotuner (tensorsat_amd.dense_tune); no counterpart in the
 * reference. */
int ts_dense_tiles(int32_t mbits, int32_t nbits, int32_t bbits, int32_t kbits, int32_t tile);
/* The tile steps of that shape use now (1..6). */
int ts_dense_tile_of(int32_t mbits, int32_t nbits, int32_t bbits, int32_t kbits, int32_t* tile);

/* Routing of dense steps (ts_contract2 and ts_program_run) to the int8 MFMA kernels
 * (v_mfma_i32_16x16x64_i8, exact int32 counts): a binary step goes there when its summed
 * wires are innermost in both operands, number at least min_sum_bits, and at least 4 out
 * wires belong to lhs only and 4 to rhs only -- 6 or more summed wires to the pipelined
 * kernel, 0..5 to the one-stage kernel (K zero-padded to 64 bytes in LDS); other steps
 * run on the VALU path.  Default: enabled, min_sum_bits = 6 (the one-stage kernel is
 * opt-in: on the measured workloads its tiles' 64-byte output runs were no faster than
 * the VALU path's whole-line stores).  Results are identical on every path; the switch
 * exists to time one against the other.  Not thread-safe (process-wide setting). */
int ts_dense_config(int32_t enable, int32_t min_sum_bits);

/* Diagonal extraction for operands with repeated wires: out axis j of rank out_rank;
 * in axis i takes the bit of out axis axis_map[i] (several in axes may share one). */
int ts_diagonal(const uint8_t* in, int32_t in_rank, const int32_t* axis_map, uint8_t* out,
                int32_t out_rank, ts_stream stream);

/* ---- Counting semiring (uint64, exact; #SAT) ------------------------------------ */

/* Same index semantics as ts_contract2 with (+, x) over uint64.  Any entry >= 2^64 sets
 * *overflow (device int32, must be zeroed by the caller) to 1; values are then invalid. */
int ts_count_contract2(const uint64_t* lhs, int32_t lhs_rank, const int32_t* lhs_wires,
                       const uint64_t* rhs, int32_t rhs_rank, const int32_t* rhs_wires,
                       uint64_t* out, int32_t out_rank, const int32_t* out_wires,
                       int32_t* overflow, ts_stream stream);

/* Widen a 0/1 byte tensor to uint64 counts. */
int ts_bytes_to_u64(const uint8_t* in, uint64_t* out, int64_t n, ts_stream stream);

/* ---- Compiled programs (single instance, grid-wide steps) ----------------------- */

/* Run one job of a compiled program with every step launched grid-wide on `stream`.
 * steps_host is the host copy of the program's step array (records are passed by
 * value), leaves_dev the device copy of its leaf bytes; arena holds
 * jobs[job].arena_elems elements (uint8 for Boolean modes).  When `inter` is non-NULL,
 * step s's output is also copied to inter + inter_offsets[s] (intermediate dumps). */
int ts_program_run(const ts_job* job_host, const ts_step* steps_host, const uint8_t* leaves_dev,
                   uint8_t* arena, int32_t mode, uint8_t* inter, const int64_t* inter_offsets,
                   ts_stream stream);

/* As ts_program_run, with every run of consecutive steps whose operands and output all
 * have at most `small_bits` bits executed by one single-workgroup launch (the steps in
 * order, barriers between them; steps_dev is the device copy of the step array) and the
 * larger steps launched grid-wide: one launch per run instead of one per step.  No
 * intermediate dumps.  The per-node loop of cotengra.py:120-126 / simple.py:251-292 for
 * networks whose small steps would leave the device idle between launches. */
int ts_program_run_fused(const ts_job* job_host, const ts_step* steps_host, const ts_step* steps_dev,
                         const uint8_t* leaves_dev, uint8_t* arena, int32_t mode, int32_t small_bits,
                         ts_stream stream);

/* As ts_program_run_fused for `count` jobs at once (jobs_host: their records, contiguous
 * on the host), job g in arena + g * arena_stride (arena_stride a multiple of 256 and at
 * least every job's arena_elems): the jobs advance phase by phase -- phase p is each
 * job's p-th segment, a run of small steps or one large step -- with one launch per
 * phase for all the runs (one workgroup per job), one for the large steps (blocks shared
 * out among the jobs; steps with few outputs and many summed terms get one workgroup
 * per output) and one per MFMA-routed step.  Where the stride leaves room after a job's
 * arena (rounded up to 256 bytes), a large step whose rows of 2^4..2^10 bytes (summed
 * wires innermost in both operands) are read by many outputs has both operands' rows
 * packed into bit masks there first and runs on the masks (the same counts).  results[g]
 * (device) = 1 if any entry of job g's result is nonzero, else 0.  The slices of a sliced tree (cotengra.py:100-128) and
 * the cubes of one instance are gangs: their small steps, one workgroup each when run
 * alone, share the device.  Enqueued on `stream`; no allocation. */
int ts_program_run_gang(const ts_job* jobs_host, int32_t count, const ts_step* steps_host,
                        const ts_step* steps_dev, const uint8_t* leaves_dev, uint8_t* arena, int64_t arena_stride,
                        uint8_t* results, int32_t mode, int32_t small_bits, ts_stream stream);

/* Counting variant: arena of uint64; result scaled by 2^free_vars is added to *count
 * (device uint64); overflow sets *overflow. */
int ts_program_count(const ts_job* job_host, const ts_step* steps_host, const uint8_t* leaves_dev,
                     uint64_t* arena, uint64_t* count, int32_t* overflow, ts_stream stream);

/* ---- Batched persistent executor ------------------------------------------------ */

typedef struct {
  const ts_job* jobs;       /* device, num_jobs records */
  const int32_t* order;     /* device, dispatch order (num_jobs job indices); NULL: the jobs
                               array is already in dispatch order (position = job index) */
  int32_t num_jobs;
  const ts_step* steps;     /* device */
  const uint8_t* leaves;    /* device */
  uint8_t* arenas;          /* device, num_workgroups * arena_stride bytes */
  int64_t arena_stride;     /* bytes per workgroup arena (>= max arena_elems, 256-aligned) */
  int32_t num_workgroups;
  int32_t mode;             /* TS_MODE_GT0 / TS_MODE_MOD256 */
  uint32_t* verdicts;       /* device, per instance, zeroed by the caller: OR of job results */
  int32_t* counter;         /* device scratch: TS_BATCH_SEGMENTS * TS_BATCH_COUNTER_STRIDE
                               int32, zeroed by the caller; the queue past each workgroup's
                               first job is taken in interleaved segments, the counter of
                               segment x at counter[x * TS_BATCH_COUNTER_STRIDE] */
  int32_t early_exit;       /* 1: skip jobs of instances already decided SAT */
  uint32_t* job_done;       /* device, per job (optional): 1 ran, 2 skipped */
  const int32_t* num_jobs_dev; /* optional device count overriding num_jobs (compacted lists) */
  int32_t lds_arena;        /* 0: arenas in device memory (`arenas`); 1: each job's arena in LDS
                               (arena_stride bytes of LDS per workgroup, <= TS_BATCH_LDS_MAX;
                               `arenas` unused); 2: as 1 with one-wave workgroups, for programs
                               whose tensors all have at most 6 bits (every step <= 64 out
                               entries); 3: as 2 for packet-form programs (tsc_packetize):
                               one packet of independent steps per wave round; 4: as 3 with
                               `steps` holding the program's lane records (tsc_lanes; the
                               jobs' step_begin / num_steps count lane records).
                               num_workgroups from ts_batch_max_workgroups_lds. */
} ts_batch;

#define TS_BATCH_LDS_MAX 65536
#define TS_BATCH_SEGMENTS 8
#define TS_BATCH_COUNTER_STRIDE 64 /* int32s between queue counters: one 256-byte line each */

/* One persistent launch: each workgroup takes the next job in `order`, copies its leaves
 * into its arena (device memory, or LDS with lds_arena) and runs every step with the
 * workgroup, OR-ing the scalar result into verdicts[instance]. */
int ts_batch_run(const ts_batch* b, ts_stream stream);

/* Counting variant (#SAT): arenas hold uint64 entries (arena_stride in bytes, >= 8 x the
 * largest arena_elems), every job's scalar result times 2^free_vars is added to
 * counts[instance] (zeroed by the caller); a product, sum or shift past 2^64 sets
 * overflow[instance] to 1 (the count is then invalid).  mode/early_exit/verdicts are
 * ignored. */
int ts_batch_count(const ts_batch* b, uint64_t* counts, int32_t* overflow, ts_stream stream);

/* Keep the entries of order[0..n) whose instance is not yet decided SAT: they are written
 * to out[0..*out_count) (*out_count must be zeroed by the caller; relative order within a
 * wave is kept).  Used between dispatch rounds so the persistent kernel only walks live
 * jobs. */
int ts_batch_compact(const int32_t* order, int32_t n, const ts_job* jobs, const uint32_t* verdicts,
                     int32_t* out, int32_t* out_count, ts_stream stream);

/* Maximum resident workgroups of the batched kernel on the current device. */
int ts_batch_max_workgroups(int32_t* out);

/* Maximum resident workgroups of the LDS-arena batched kernel for arenas of arena_bytes
 * (a multiple of 256, at most TS_BATCH_LDS_MAX); one_wave = 1 selects the lds_arena = 2
 * variant, 2 the lds_arena = 3 (packet form) one, 3 the lds_arena = 4 (lane form) one. */
int ts_batch_max_workgroups_lds(int64_t arena_bytes, int32_t one_wave, int32_t* out);

/* ---- Device cube search (batched decisions without host planning) ---------------
 *
 * The batched counterpart of `bool(CotengraContraction(FinRel, wiring).contract(d))`
 * (examples/test_sat_solver.py:5-9, contractions/cotengra.py:100-128) for CNF networks,
 * with the conditioned slicing done on the device: a job is a depth-first branch over
 * the variable wires of one instance, each branch fixing a wire in every clause tensor
 * that holds it (Box.__getitem__, diagrams.py:972-996 -> FinRel._getitem,
 * fin_rel.py:342-343) and the branch results OR-ed (Box.__add__ -> FinRel._union,
 * fin_rel.py:345-346); all-false clause tensors prune a branch and one-hot ones fix
 * their variable (unit propagation); a residual network of at most `width` (<= 6) free
 * variables is contracted into one rank-|R| relation (FinRel._contract2 with every wire
 * kept, fin_rel.py:255-288) and reduced over all its ports (FinRel._rewire to no ports,
 * fin_rel.py:379-393; at most 64 terms meet, so GT0 and MOD256 agree).  One 64-lane
 * wave runs one job.  Search rules (branch variable, polarity, statistics) are those
 * restated by the oracle (oracle/search.c). */

#define TS_SEARCH_MAX_VARS 128  /* variables per instance */
#define TS_SEARCH_SLOTS 12      /* clauses per lane: at most 64 * 12 = 768 clauses */
#define TS_SEARCH_MAX_WIDTH 6   /* residuals of at most 2^6 entries are contracted */
/* Packed clause: bits 8i..8i+7 hold literal i (i < 3) as (variable << 1) | negated,
 * bits 24..31 the number of literals (0..3). */
#define TS_SEARCH_CLAUSE(len) ((uint32_t)(len) << 24)

typedef struct {
  int32_t clause_begin;   /* first packed clause of the instance */
  int32_t num_clauses;    /* <= 64 * TS_SEARCH_SLOTS */
  int32_t num_vars;       /* <= TS_SEARCH_MAX_VARS */
  int32_t reserved;       /* 0 */
} ts_search_inst;

typedef struct {
  int32_t instance;       /* < 0: an empty slot, skipped */
  int32_t depth;          /* depth of the job's start node */
  int32_t split;          /* branch levels below the start node fixed by prefix (0..31) */
  int32_t own_from;       /* the job counts (and budgets) the nodes at depth >= own_from;
                             shallower nodes on its prefix path belong to another job */
  uint32_t prefix;        /* bit k: child taken at depth `depth` + k, k < split
                             (0 = the first literal) */
  uint32_t reserved[3];   /* 0 */
  uint64_t assigned[2];   /* assignment of the start node before propagation: variable v */
  uint64_t value[2];      /* is bit v & 63 of word v >> 6 (the root: all zero) */
} ts_search_job;

typedef struct {
  int32_t status;         /* 0 complete, 1 stopped (instance decided by another job),
                             2 node budget exhausted (no continuation list), 3 handed its
                             remaining subtrees to the continuation list */
  int32_t max_depth;
  int64_t nodes, leaves, conflicts; /* over the nodes the job owns: the sums over all jobs
                             of an instance (continuations included) equal one unsplit
                             search's in exhaustive mode */
  uint64_t digest;        /* sum of mix(relation bits, rank) over owned leaves */
} ts_search_stat;

#define TS_SEARCH_SAT 1u          /* verdict bit: a satisfiable leaf was found */
#define TS_SEARCH_INCOMPLETE 2u   /* verdict bit: a job ran out of its node budget */

typedef struct {
  const uint32_t* clauses;         /* device, packed clauses */
  const ts_search_inst* instances; /* device */
  int32_t num_instances;
  const ts_search_job* jobs;       /* device, in launch order (one wave each) */
  int32_t num_jobs;
  uint32_t* verdicts;              /* device, per instance, zeroed by the caller */
  ts_search_stat* stats;           /* device, per job (optional) */
  int32_t width;                   /* 0..TS_SEARCH_MAX_WIDTH */
  int32_t exhaustive;              /* 1: visit every branch (statistics) */
  int32_t early_exit;              /* 1: stop a job once its instance is decided SAT */
  int32_t spill_capacity;          /* continuation list capacity (jobs) */
  int64_t max_nodes;               /* per job: owned nodes before it hands off (with a
                                      continuation list) or stops with status 2 */
  ts_search_job* spill;            /* device continuation list (optional): a job out of
                                      budget appends its current subtree and every untried
                                      sibling subtree below its split as new jobs; when they
                                      do not fit it writes empty slots and keeps searching */
  int32_t* spill_count;            /* device, zeroed by the caller: jobs appended (may
                                      exceed the capacity; only the first spill_capacity
                                      slots exist) */
  int32_t max_clauses;             /* largest num_clauses of the instances (0: unknown): the
                                      launch uses the kernel variant with just enough clause
                                      slots per lane; an instance beyond it is reported
                                      TS_SEARCH_INCOMPLETE, never decided */
  int32_t order;                   /* TS_SEARCH_ORDER_SCORE or TS_SEARCH_ORDER_LEX */
  uint64_t* witness;               /* device, 2 words per instance (optional): the assignment
                                      of the satisfiable leaf a job finds (variable v is bit
                                      v & 63 of word v >> 6; variables the leaf leaves free
                                      are 0, its relation's entry the lexicographically
                                      smallest), written by each job that finds one */
  uint64_t* counts;                /* device, per instance (optional, zeroed by the caller):
                                      with exhaustive = 1 and early_exit = 0, the exact model
                                      count -- each owned leaf adds its relation's entry count
                                      times 2 to the number of free variables outside it (no
                                      counterpart in the reference package; the #SAT
                                      prototype's counting einsum, notebooks/prototypes/
                                      2025-02-13-#sat-tensor-contraction.ipynb cell 12) */
  int32_t* count_overflow;         /* device int32 (required with counts, zeroed by the
                                      caller): set to 1 when a count does not fit in 64 bits */
} ts_search;

/* Branching orders.  SCORE: the free variable of largest (P + 1) * (N + 1), P and N its
 * occurrences in the residual clauses as a positive and as a negative literal, each
 * weighted 5 in a clause with two free literals and 1 in one with three (ties to the
 * smaller index), its more frequent polarity first (the decision default).  LEX: the
 * smallest variable of the residual, false first: with one job per instance, no hand-over
 * and no exhaustive mode, the first satisfiable leaf gives the lexicographically smallest
 * satisfying assignment (variables in index order, 0 before 1) -- the assignment the
 * reference's search over one contraction with 2n input selections returns
 * (lib/sat.py:231-246 contraction_wiring; notebooks/02-random-ksat.ipynb cell 36). */
#define TS_SEARCH_ORDER_SCORE 0
#define TS_SEARCH_ORDER_LEX 1

/* One launch of num_jobs one-wave workgroups.  Rejects (TS_EINVAL) null arrays and widths
 * out of range; job and instance records are the caller's to check (the host layer
 * validates them before upload, and the continuations the kernel writes are valid by
 * construction). */
int ts_search_run(const ts_search* s, ts_stream stream);

/* The whole search in one launch (queue mode): resident one-wave workgroups take job
 * records from a device queue in order, and a job that exhausts its s->max_nodes owned
 * nodes publishes its remaining subtrees (as in round mode) at the queue's end for any
 * idle workgroup to take -- no host round trip between hand-overs.  records holds
 * `capacity` jobs, the first s->num_jobs of them the start jobs (s->jobs, s->spill,
 * s->spill_count and s->spill_capacity are ignored); state (TS_SEARCH_QUEUE_STATE
 * int32) and ready (capacity uint32 flags) are device scratch that the call
 * initialises (a fill kernel, no memset).  After the launch completes: state[0] =
 * records reserved (those beyond capacity were searched in place instead),
 * state[192] != 0 if a workgroup gave up waiting for a record (a safety limit of 30 s;
 * the verdicts are then incomplete).  s->stats (optional, capacity entries) is indexed by record.  workers:
 * workgroups to launch (<= 0 or more than fit on the device at once: as many as fit);
 * records are claimed in order by whichever workgroup is free, so the search completes
 * however many of them the device runs at once.  Generalises
 * ts_search_run's hand-over rounds; no counterpart in the reference. */
#define TS_SEARCH_QUEUE_STATE 256
int ts_search_run_queue(const ts_search* s, ts_search_job* records, int32_t capacity, int32_t* state,
                        uint32_t* ready, int32_t workers, ts_stream stream);

/* Allocation-free variant of a whole search (no counterpart in the reference): for
 * callers that cannot call hipMalloc -- tight serving loops, harnesses that forbid device
 * allocation.  Runs ts_search_run on s->jobs, then on the jobs handed over, round after
 * round, until none are left, with the continuation lists in a static device array of
 * the library (two lists of TS_SEARCH_STATIC_JOBS jobs plus a counter, 8 MiB per GPU,
 * held until the process ends) instead of caller buffers.  s->spill and
 * s->spill_count must be NULL; s->spill_capacity is the list length to use (0: all of
 * it; above TS_SEARCH_STATIC_JOBS: TS_EINVAL); s->max_nodes is the hand-over budget;
 * s->stats (optional) receives the statistics of the caller's jobs only.  Hand-overs
 * that do not fit are searched in place, so no input is too large.  Synchronous: it
 * reads each round's job count on the host.  Calls that overlap on one GPU race on the
 * static lists.  *rounds_out (optional) receives the number of launches. */
#define TS_SEARCH_STATIC_JOBS 65536
int ts_search_decide_noalloc(const ts_search* s, int32_t* rounds_out, ts_stream stream);

/* ---- Learned planner: message-passing network kernels (bf16 in, fp32 accumulate) --
 * No counterpart in the reference: the learned planner is an addition of this build.
 * Matrices are row-major; bf16 buffers are passed as void*. */

/* C (+)= A x B (+ bias) (ReLU): A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn] (bf16),
 * C[m*ldc + n] fp32 or bf16 (out_bf16); v_mfma_f32_16x16x32_bf16 with fp32 accumulation. */
int ts_gnn_gemm(const void* A, int64_t sam, int64_t sak, const void* B, int64_t sbk, int64_t sbn, void* C,
                int64_t ldc, int32_t M, int32_t N, int32_t K, const float* bias, int32_t relu,
                int32_t out_bf16, int32_t accumulate, ts_stream stream);

/* Weighted signed neighbour sums over a CSR incidence (row r -> col[e], weight w[e] whose
 * sign is the literal's): forward writes sum |w| X[col, xoff..] over positive and negative
 * edges into bf16 column blocks off_pos / off_neg; with out_f32_accumulate both signs (read
 * from xoff_pos / xoff_neg) add into the fp32 block off_pos (the transpose, for the
 * backward pass, with the same symmetric weights).  F <= 128. */
int ts_gnn_segsum(const int32_t* rowptr, const int32_t* col, const float* weight, int32_t nrows, const void* X,
                  int64_t ldx, int32_t xoff_pos, int32_t xoff_neg, int32_t x_f32, void* out, int64_t ldo,
                  int32_t off_pos, int32_t off_neg, int32_t F, int32_t out_f32_accumulate, ts_stream stream);

/* dZ = dH * (H > 0) as bf16, column sums of dZ added to db (if non-NULL). */
int ts_gnn_relu_back(const float* dH, int64_t ldd, const void* H, int64_t ldh, int32_t rows, int32_t F, void* dZ,
                     int64_t ldz, float* db, ts_stream stream);

/* Scalar head y = H w + b; in training also the masked squared error against target
 * (added to *loss), dH = dy w^T (fp32 [rows, F]), dw, db. */
int ts_gnn_head(const void* H, int64_t ldh, int32_t rows, const float* w, const float* b, const float* target,
                const float* mask, float inv_count, float* y, float* loss, float* dH, float* dw, float* db,
                int32_t F, int32_t train, ts_stream stream);

/* Adam on fp32 master weights (bias-corrected; L2 term wd * p added to the gradient),
 * refreshing the bf16 copy pbf (if non-NULL). */
int ts_gnn_adam(float* p, const float* g, float* m, float* v, void* pbf, int64_t n, float lr, float b1, float b2,
                float eps, int32_t step, float wd, ts_stream stream);

int ts_gnn_cast_bf16(const float* x, void* y, int64_t n, ts_stream stream);

#ifdef __cplusplus
}
#endif

#endif /* TENSORSAT_AMD_H */
