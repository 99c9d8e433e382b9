// Device cube search (gfx950): batched SAT decisions with the conditioned slicing done on
// the device, one 64-lane wave per job.
//
// A job walks the depth-first branch tree of one instance below a fixed prefix of its
// first `split` branches.  Branching on a variable wire fixes it in every clause tensor
// that holds it (the reference's Box.__getitem__, tensorsat/diagrams.py:972-996 ->
// FinRel._getitem, lang/fin_rel.py:342-343); branch results are OR-ed (Box.__add__ ->
// FinRel._union, fin_rel.py:345-346).  A clause tensor whose literals are all false is
// all zero and prunes the branch; one with a single free literal is one-hot and fixes
// that variable (unit propagation).  A residual network with at most `width` (<= 6) free
// variables R is contracted: every residual clause tensor is multiplied into one
// rank-|R| relation over R (FinRel._contract2 keeping every wire, fin_rel.py:255-288),
// one lane per relation entry, and the relation is reduced over all its ports by a
// ballot (FinRel._rewire to no ports, fin_rel.py:379-393).
//
// Layout: lane l holds clauses l, 64 + l, ... (TS_SEARCH_SLOTS per lane) in registers as
// packed words (three literal bytes (v << 1) | negated and a length byte); the partial
// assignment is wave-uniform (two 128-bit masks: assigned, value, split into even and
// odd variables) and lives in scalar registers; the branch stack (assignment before the branch, branch literal, next child)
// and the branching histogram are in LDS.  A 256-byte table in LDS gives the status of
// every literal byte under the current assignment (free, false, true), rebuilt from the
// masks (one dword per lane) whenever they change, and a constant 256-entry table maps a
// clause's length and its three literal statuses to (satisfied, free literals, position
// of a free literal): a clause costs three byte extracts and four LDS reads.
// Propagation runs in rounds: every lane evaluates its clauses, a ballot finds
// conflicts, and the unit literals of the round (one per lane) are written into the
// literal table by all lanes at once, a read-back detecting a variable forced both
// ways; the masks are rebuilt from the table at the fixpoint, which is the same as any
// sequential order's.
//
// Every loop is bounded: propagation rounds by the number of variables (each round
// assigns at least one), the branch stack by the number of variables, and each job by
// its node budget.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>

#include "tensorsat_amd.h"

namespace {


constexpr int MAXV = TS_SEARCH_MAX_VARS;
constexpr uint32_t NO_CLAUSE = 0xFFu;  // length byte of an empty slot (always satisfied)
constexpr uint32_t NO_UNIT = 0xFFFFu;

__device__ __forceinline__ uint64_t splitmix(uint64_t x) {
  uint64_t z = x + 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ uint32_t uni(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ uint64_t uni64(uint64_t x) {
  return (uint64_t)uni((uint32_t)x) | ((uint64_t)uni((uint32_t)(x >> 32)) << 32);
}

// Reductions over the wave (every lane active): within rows of 16 lanes by DPP, then
// across the four rows by readlane.
template <typename Op>
__device__ __forceinline__ uint32_t wave_reduce(uint32_t x, Op op) {
  x = op(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0xB1, 0xF, 0xF, false));   // quad [1,0,3,2]
  x = op(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x4E, 0xF, 0xF, false));   // quad [2,3,0,1]
  x = op(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x141, 0xF, 0xF, false));  // row half mirror
  x = op(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x140, 0xF, 0xF, false));  // row mirror
  const uint32_t a = __builtin_amdgcn_readlane(x, 0), b = __builtin_amdgcn_readlane(x, 16);
  const uint32_t c = __builtin_amdgcn_readlane(x, 32), d = __builtin_amdgcn_readlane(x, 48);
  return op(op(a, b), op(c, d));
}
struct OpMax {
  __device__ uint32_t operator()(uint32_t a, uint32_t b) const { return a > b ? a : b; }
};
struct OpAnd {
  __device__ uint32_t operator()(uint32_t a, uint32_t b) const { return a & b; }
};

// A partial assignment in the kernel's layout: variable v is bit v >> 1 of the even word
// (v even) or of the odd word (v odd), so lane l's variables 2l and 2l + 1 are bit l of
// each.  Job records and witnesses use the natural layout (variable v = bit v & 63 of
// word v >> 6); the conversions below run on every lane (they ballot).
struct Assign {
  uint64_t ae, ao, ve, vo;  // assigned, value; even and odd variables
};

__device__ __forceinline__ uint32_t natural_bit(uint64_t n0, uint64_t n1, uint32_t v) {
  return (uint32_t)(((v < 64 ? n0 : n1) >> (v & 63)) & 1ull);
}
__device__ __forceinline__ uint32_t split_bit(uint64_t e, uint64_t o, uint32_t v) {
  return (uint32_t)((((v & 1u) ? o : e) >> (v >> 1)) & 1ull);
}
__device__ __forceinline__ void from_natural(uint64_t n0, uint64_t n1, uint64_t& e, uint64_t& o, int lane) {
  e = __ballot(natural_bit(n0, n1, 2u * (uint32_t)lane));
  o = __ballot(natural_bit(n0, n1, 2u * (uint32_t)lane + 1u));
}
__device__ __forceinline__ void to_natural(uint64_t e, uint64_t o, uint64_t& n0, uint64_t& n1, int lane) {
  n0 = __ballot(split_bit(e, o, (uint32_t)lane));
  n1 = __ballot(split_bit(e, o, (uint32_t)lane + 64u));
}
// Make literal byte b = (v << 1) | negated true.
__device__ __forceinline__ void set_literal(Assign& s, uint32_t b) {
  const uint32_t v = b >> 1;
  const uint64_t bit = 1ull << (v >> 1), val = (b & 1u) ? 0ull : bit;
  if (v & 1u) {
    s.ao |= bit;
    s.vo |= val;
  } else {
    s.ae |= bit;
    s.ve |= val;
  }
}

// The clause words never change during a job, so the compiler would decode all of them
// once, outside the search loop, into hundreds of registers; marking them as rewritten
// at each use keeps the decode next to its use (a few registers per slot).
template <int NS>
__device__ __forceinline__ void fresh(uint32_t (&cl)[NS]) {
#pragma unroll
  for (int k = 0; k < NS; ++k) asm volatile("" : "+v"(cl[k]));
}

// Literal statuses (byte b = (v << 1) | negated): 0 free, 1 false, 2 true.
constexpr uint32_t LIT_FREE = 0, LIT_FALSE = 1, LIT_TRUE = 2;
// Clause status byte: bit 0 satisfied, bits 1..2 free literals, bits 3..4 position of the
// last free literal.
__device__ __forceinline__ uint32_t cs_sat(uint32_t r) { return r & 1u; }
__device__ __forceinline__ uint32_t cs_nfree(uint32_t r) { return (r >> 1) & 3u; }
__device__ __forceinline__ uint32_t cs_pos(uint32_t r) { return (r >> 3) & 3u; }

// Rebuild the literal table from the masks: lane l writes the statuses of variables 2l
// and 2l + 1, both polarities (bytes 4l .. 4l + 3).
__device__ __forceinline__ void build_lit_table(uint32_t* lt, const Assign& s, int lane) {
  const uint32_t a = (uint32_t)((s.ae >> lane) & 1ull) | ((uint32_t)((s.ao >> lane) & 1ull) << 1);
  const uint32_t v = (uint32_t)((s.ve >> lane) & 1ull) | ((uint32_t)((s.vo >> lane) & 1ull) << 1);
  // per variable: assigned -> (positive literal, negative literal) = value ? (T, F) : (F, T)
  const uint32_t p0 = (a & 1u) ? ((v & 1u) ? LIT_TRUE : LIT_FALSE) : LIT_FREE;
  const uint32_t n0 = (a & 1u) ? ((v & 1u) ? LIT_FALSE : LIT_TRUE) : LIT_FREE;
  const uint32_t p1 = (a & 2u) ? ((v & 2u) ? LIT_TRUE : LIT_FALSE) : LIT_FREE;
  const uint32_t n1 = (a & 2u) ? ((v & 2u) ? LIT_FALSE : LIT_TRUE) : LIT_FREE;
  lt[lane] = p0 | (n0 << 8) | (p1 << 16) | (n1 << 24);
}

// The masks back from the literal table: lane l reads the statuses of the positive
// literals of variables 2l and 2l + 1.
__device__ __forceinline__ void masks_from_table(Assign& s, const uint32_t* lt, int lane) {
  const uint32_t w = lt[lane];
  const uint32_t pe = w & 0xFFu, po = (w >> 16) & 0xFFu;
  s.ae = __ballot(pe != LIT_FREE);
  s.ve = __ballot(pe == LIT_TRUE);
  s.ao = __ballot(po != LIT_FREE);
  s.vo = __ballot(po == LIT_TRUE);
}

// The clause-status table: entry ((len & 7) << 6 | x2 << 4 | x1 << 2 | x0) for literal
// statuses x_i of the first len literals; lengths 4..7 (the empty-slot length byte 0xFF
// lands on 7) read as satisfied.  Lane l writes entries 8l .. 8l + 7.
__device__ __forceinline__ void build_clause_table(uint32_t* ct, int lane) {
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    uint32_t word = 0;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const uint32_t idx = 8u * (uint32_t)lane + 4u * (uint32_t)h + (uint32_t)e, len = idx >> 6;
      uint32_t sat = len > 3 ? 1u : 0u, nfree = 0, pos = 0;
      for (uint32_t i = 0; i < len && i < 3; ++i) {
        const uint32_t x = (idx >> (2 * i)) & 3u;
        sat |= x == LIT_TRUE;
        if (x == LIT_FREE) {
          ++nfree;
          pos = i;
        }
      }
      word |= (sat | (nfree << 1) | (pos << 3)) << (8 * e);
    }
    ct[2 * lane + h] = word;
  }
}

// Status bytes of every slot's clause: all literal-status reads first, then the
// clause-status reads, so the LDS latencies overlap.
template <int NS>
__device__ __forceinline__ void clause_states(const uint32_t (&cl)[NS], const uint8_t* lt, const uint8_t* ct,
                                              uint32_t (&st)[NS]) {
  uint32_t x[NS][3];
#pragma unroll
  for (int k = 0; k < NS; ++k) {
#pragma unroll
    for (int i = 0; i < 3; ++i) x[k][i] = lt[(cl[k] >> (8 * i)) & 0xFFu];
  }
#pragma unroll
  for (int k = 0; k < NS; ++k)
    st[k] = ct[((cl[k] >> 18) & 0x1C0u) | (x[k][2] << 4) | (x[k][1] << 2) | x[k][0]];
}

// Unit propagation to the fixpoint.  Returns false on a conflict.  On success `s` is the
// final assignment, `st` holds every slot's clause status under it and the literal
// table matches it.  Each round every lane offers at most one unit literal (its first
// unit clause's, whose variable is free) and all lanes apply theirs at once: first each
// writes TRUE at its literal's byte, then FALSE at the opposite literal's byte.  One
// wave's LDS accesses complete in order, so a variable offered with both signs in the
// round ends with equal bytes at both literals (the later write's value at each) and
// every lane that offered it reads a mismatch back: a conflict.  A variable offered with
// one sign ends TRUE/FALSE however many lanes offered it.  Each round with a unit
// assigns at least one new variable.
template <int NS>
__device__ __forceinline__ bool propagate(uint32_t (&cl)[NS], Assign& s, uint32_t (&st)[NS], uint32_t* lt32,
                                          const uint32_t* ct32, int lane) {
  uint8_t* lt = reinterpret_cast<uint8_t*>(lt32);
  const uint8_t* ct = reinterpret_cast<const uint8_t*>(ct32);
  build_lit_table(lt32, s, lane);
  __syncthreads();
  for (int round = 0; round <= MAXV; ++round) {
    fresh(cl);
    clause_states(cl, lt, ct, st);
    bool conflict = false;
    uint32_t unit = NO_UNIT;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      const uint32_t r = st[k], nfree = cs_nfree(r);
      const bool open = !cs_sat(r);
      conflict |= open && nfree == 0;
      if (open && nfree == 1 && unit == NO_UNIT) unit = (cl[k] >> (8 * cs_pos(r))) & 0xFFu;
    }
    if (__ballot(conflict)) return false;
    const bool has = unit != NO_UNIT;
    if (!__ballot(has)) {
      masks_from_table(s, lt32, lane);
      return true;
    }
    // Compiler barriers keep the two write phases and the reads in this order (a read's
    // address differs from its lane's other write, so it could otherwise be hoisted).
    if (has) lt[unit] = (uint8_t)LIT_TRUE;
    asm volatile("" ::: "memory");
    if (has) lt[unit ^ 1u] = (uint8_t)LIT_FALSE;
    asm volatile("" ::: "memory");
    const bool clash = has && (lt[unit] != (uint8_t)LIT_TRUE || lt[unit ^ 1u] != (uint8_t)LIT_FALSE);
    if (__ballot(clash)) return false;  // a variable forced both ways
  }
  return true;  // unreachable: every round with a unit assigns a new variable
}

// Weight of a residual clause with two free literals in the branching histogram (one
// with three free literals weighs 1), and the score of a histogram entry.
constexpr uint32_t HIST_W2 = 5;
__device__ __forceinline__ uint32_t score_of(uint64_t h) {
  return ((uint32_t)(h & 0xFFFFu) + 1u) * ((uint32_t)((h >> 16) & 0xFFFFu) + 1u);
}

// Entries e of a rank <= 6 relation whose bit p is 1, as a 64-bit mask.
__device__ __forceinline__ uint64_t entries_with_bit(uint32_t p) {
  switch (p) {
    case 0: return 0xAAAAAAAAAAAAAAAAull;
    case 1: return 0xCCCCCCCCCCCCCCCCull;
    case 2: return 0xF0F0F0F0F0F0F0F0ull;
    case 3: return 0xFF00FF00FF00FF00ull;
    case 4: return 0xFFFF0000FFFF0000ull;
    default: return 0xFFFFFFFF00000000ull;
  }
}

struct Shared {
  unsigned long long hist[MAXV];  // per variable: weighted positive, weighted negative, counts
  uint64_t stk[4][MAXV + 1];    // assignment before the branch at each depth
  uint32_t stk_lit[MAXV + 1];   // branch literal byte | next child << 8
  uint32_t lit_table[64];       // status byte of each literal byte
  uint32_t clause_table[128];   // status byte of each (length, statuses) index
};

struct JobResult {
  int64_t nodes, leaves, conflicts;
  uint64_t digest;
  uint64_t count;  // models under the owned leaves (with p.counts)
  int status, max_depth, count_overflow;
};

// Queue mode (ts_search_run_queue): the continuation list is the queue itself; `state`
// holds, one 256-byte line each, the records reserved, the next record to claim, the
// records pending (published -- the start records from the outset -- and not finished)
// and an error flag; `ready` has one flag per record beyond the start records.  Null:
// round mode.
constexpr int Q_TAIL = 0, Q_HEAD = 64, Q_PENDING = 128, Q_ERROR = 192;
struct Queue {
  int32_t* state;
  uint32_t* ready;
};

// A continuation record written by lanes 0..15, one dword each, with system-scope
// stores: they reach memory past every cache, so a reader on another XCD never meets a
// stale copy of the line (records of different producers share 128-byte lines).
__device__ __forceinline__ void store_record(ts_search_job* rec, const ts_search_job& o, int lane) {
  uint32_t d;
  switch (lane) {
    case 0: d = (uint32_t)o.instance; break;
    case 1: d = (uint32_t)o.depth; break;
    case 2: d = (uint32_t)o.split; break;
    case 3: d = (uint32_t)o.own_from; break;
    case 4: d = o.prefix; break;
    case 5: d = o.reserved[0]; break;
    case 6: d = o.reserved[1]; break;
    case 7: d = o.reserved[2]; break;
    case 8: d = (uint32_t)o.assigned[0]; break;
    case 9: d = (uint32_t)(o.assigned[0] >> 32); break;
    case 10: d = (uint32_t)o.assigned[1]; break;
    case 11: d = (uint32_t)(o.assigned[1] >> 32); break;
    case 12: d = (uint32_t)o.value[0]; break;
    case 13: d = (uint32_t)(o.value[0] >> 32); break;
    case 14: d = (uint32_t)o.value[1]; break;
    default: d = (uint32_t)(o.value[1] >> 32); break;
  }
  if (lane < 16)
    __hip_atomic_store(reinterpret_cast<uint32_t*>(rec) + lane, d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Queue mode: make records [first, first + n) runnable -- one release fence for their
// stores, counted as pending (lane 0 adds), then flagged (a lane per record).
__device__ __forceinline__ void publish(const Queue& q, int first, int n, int lane) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  atomicAdd(&q.state[Q_PENDING], lane == 0 ? n : 0);
  for (int b = 0; b < n; b += 64)
    if (b + lane < n) __hip_atomic_store(&q.ready[first + b + lane], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One job's search with NS clause slots per lane (the instance has at most 64 * NS
// clauses).  The clause-status table and the zeroed histogram are in `sh` on entry.
template <int NS>
__device__ __forceinline__ void search_job(const ts_search& p, const ts_search_job& job, const ts_search_inst& inst,
                                           Shared& sh, int lane, JobResult& r, const Queue& q) {
  const int m = inst.num_clauses;
  int64_t owned = 0;
  const int base = job.depth, split = job.split, own_from = job.own_from;
  const uint32_t prefix = job.prefix;
  const int width = p.width;
  const uint8_t* lt8 = reinterpret_cast<const uint8_t*>(sh.lit_table);
  int64_t budget = p.max_nodes;

  uint32_t cl[NS];
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    const int i = k * 64 + lane;
    cl[k] = i < m ? p.clauses[inst.clause_begin + i] : (NO_CLAUSE << 24);
  }

  Assign s;
  from_natural(job.assigned[0], job.assigned[1], s.ae, s.ao, lane);
  from_natural(job.value[0], job.value[1], s.ve, s.vo, lane);
  uint32_t st[NS];
  // The loop alternates two phases.  Propagate: `s` is the assignment of the node at
  // depth `cd` (the start node, or the child of frame cd - 1) before propagation.
  // Next: try the next child of the branch frame at depth `fd`.
  int cd = base, fd = base - 1;
  bool prop = true;
  for (;;) {
    if (!prop) {
      const uint32_t fr = uni(sh.stk_lit[fd]);
      const uint32_t t = fr >> 8;
      if (t >= 2) {
        if (fd == base) break;
        --fd;
        continue;
      }
      sh.stk_lit[fd] = (fr & 0xFFu) | ((t + 1) << 8);
      __syncthreads();
      if (fd - base < split && t != ((prefix >> (fd - base)) & 1u)) continue;
      s.ae = uni64(sh.stk[0][fd]);
      s.ao = uni64(sh.stk[1][fd]);
      s.ve = uni64(sh.stk[2][fd]);
      s.vo = uni64(sh.stk[3][fd]);
      set_literal(s, (fr & 0xFFu) ^ t);  // the second child negates the literal
      cd = fd + 1;
      prop = true;
      continue;
    }
    // ---- propagate the node at depth cd ----
    if (!propagate<NS>(cl, s, st, sh.lit_table, sh.clause_table, lane)) {
      if (cd >= own_from) ++r.conflicts;
      if (cd == base) break;
      fd = cd - 1;
      prop = false;
      continue;
    }
    const int depth = cd;
    if (p.early_exit &&
        (__hip_atomic_load(&p.verdicts[job.instance], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) &
         TS_SEARCH_SAT)) {
      r.status = 1;
      break;
    }
    const bool own = depth >= own_from;
    if (own && owned >= budget) {
      if (!p.spill) {
        r.status = 2;
        break;
      }
      // Hand the rest over: this node's subtree, and the untried second child of
      // every frame the job explores freely (below its prefix levels).
      int n_out = 1;
      for (int k = base + split; k < depth; ++k) n_out += (uni(sh.stk_lit[k]) >> 8) == 1u;
      // one reservation for the wave (lane 0 adds, every lane takes part: no branch on
      // the lane inside the search loops)
      const int start = (int)uni((uint32_t)atomicAdd(p.spill_count, lane == 0 ? n_out : 0));
      if (start + n_out <= p.spill_capacity) {
        // Records are built on every lane (the layout conversion ballots) and
        // written by lane 0.
        ts_search_job o = {};
        o.instance = job.instance;
        int w = start;
        for (int k = base + split; k < depth; ++k) {
          const uint32_t fr = uni(sh.stk_lit[k]);
          if ((fr >> 8) != 1u) continue;
          Assign c{uni64(sh.stk[0][k]), uni64(sh.stk[1][k]), uni64(sh.stk[2][k]), uni64(sh.stk[3][k])};
          set_literal(c, (fr & 0xFFu) ^ 1u);
          o.depth = k + 1;
          o.own_from = k + 1;
          to_natural(c.ae, c.ao, o.assigned[0], o.assigned[1], lane);
          to_natural(c.ve, c.vo, o.value[0], o.value[1], lane);
          store_record(p.spill + w, o, lane);
          ++w;
        }
        // The subtree keeps what is left of the job's prefix levels (the other
        // children there belong to the jobs with the other prefixes).
        const int rel = depth - base;
        o.depth = depth;
        o.own_from = depth;
        o.split = rel < split ? split - rel : 0;
        o.prefix = rel < split ? prefix >> rel : 0u;
        to_natural(s.ae, s.ao, o.assigned[0], o.assigned[1], lane);
        to_natural(s.ve, s.vo, o.value[0], o.value[1], lane);
        store_record(p.spill + w, o, lane);
        if (q.ready) publish(q, start, n_out, lane);
        r.status = 3;
        break;
      }
      // No room: empty the reserved slots that exist and finish the subtree here.
      {
        ts_search_job o = {};
        o.instance = -1;
        const int end = start + n_out < p.spill_capacity ? start + n_out : p.spill_capacity;
        for (int w = start; w < end; ++w) store_record(p.spill + w, o, lane);
        if (q.ready && end > start) publish(q, start, end - start, lane);
      }
      budget = INT64_MAX;
    }
    if (own) {
      ++owned;
      ++r.nodes;
      r.max_depth = depth > r.max_depth ? depth : r.max_depth;
    }
    // Branching histogram over the residual clauses' free literals; it also names R.
    // Per variable, 16-bit fields: the weighted occurrences as a positive literal, as a
    // negative literal (weight HIST_W2 in a clause with two free literals, 1 with three),
    // and the two plain counts.  With at most 768 clauses no field overflows.
    fresh(cl);
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      const uint32_t c = cl[k], len = c >> 24;
      const bool open = !cs_sat(st[k]);
      const unsigned long long w = cs_nfree(st[k]) == 2u ? HIST_W2 : 1u;
      const unsigned long long inc_pos = w | (1ull << 32), inc_neg = (w << 16) | (1ull << 48);
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const uint32_t b = (c >> (8 * i)) & 0xFFu;
        if (open && (uint32_t)i < len && lt8[b] == LIT_FREE) atomicAdd(&sh.hist[b >> 1], (b & 1u) ? inc_neg : inc_pos);
      }
    }
    __syncthreads();
    const uint64_t h0 = sh.hist[lane], h1 = sh.hist[lane + 64];
    sh.hist[lane] = 0;
    sh.hist[lane + 64] = 0;
    const uint64_t r0 = __ballot(h0 != 0), r1 = __ballot(h1 != 0);
    const int rank = __popcll(r0) + __popcll(r1);
    if (rank <= width) {
      // ---- leaf: the residual clause tensors multiplied into the rank-|R| relation
      // (entry e: bit j of e = value of the j-th smallest variable of R), then reduced ----
      uint64_t bits = 1;  // no residual clause: the empty product
      if (rank > 0) {
        uint64_t mine = ~0ull;  // entries satisfying this lane's residual clauses
        fresh(cl);
#pragma unroll
        for (int k = 0; k < NS; ++k) {
          const uint32_t c = cl[k], len = c >> 24;
          uint64_t fals = cs_sat(st[k]) ? 0ull : ~0ull;  // a satisfied clause removes nothing
#pragma unroll
          for (int i = 0; i < 3; ++i) {
            const uint32_t b = (c >> (8 * i)) & 0xFFu, v = b >> 1;
            // assigned literals of a residual clause are false
            if ((uint32_t)i < len && lt8[b] == LIT_FREE) {
              const uint32_t pos = (v & 64) ? (uint32_t)(__popcll(r0) + __popcll(r1 & ((1ull << (v & 63)) - 1ull)))
                                            : (uint32_t)__popcll(r0 & ((1ull << v) - 1ull));
              const uint64_t ones = entries_with_bit(pos);
              fals &= (b & 1u) ? ones : ~ones;  // false where the variable equals the negation bit
            }
          }
          mine &= ~fals;
        }
        const uint32_t lo = wave_reduce((uint32_t)mine, OpAnd());
        const uint32_t hi = wave_reduce((uint32_t)(mine >> 32), OpAnd());
        const uint64_t valid = rank == 6 ? ~0ull : ((1ull << (1u << rank)) - 1ull);
        bits = (((uint64_t)hi << 32) | lo) & valid;
      }
      if (own) {
        ++r.leaves;
        r.digest += splitmix(splitmix(bits) + (uint64_t)rank);
        if (p.counts) {
          // models below this leaf: the relation's entries times 2^(free variables that
          // no residual clause holds); the assignment so far is propagated and consistent
          const uint64_t c = (uint64_t)__popcll(bits);
          const int freev = inst.num_vars - (__popcll(s.ae) + __popcll(s.ao)) - rank;
          if (c && (freev >= 64 || (freev > 0 && (c >> (64 - freev)) != 0))) {
            r.count_overflow = 1;
          } else if (c) {
            const uint64_t add = c << freev;
            if (r.count + add < r.count) r.count_overflow = 1;
            r.count += add;
          }
        }
        if (bits != 0) {
          uint64_t w0 = 0, w1 = 0;  // the leaf's assignment's values, natural layout
          if (p.witness) to_natural(s.ve, s.vo, w0, w1, lane);
          if (p.witness) {
            // computed on every lane (the values are uniform), stored by lanes 0 and 1
            // The satisfying entry that is smallest with the smallest variable of R
            // most significant (lexicographic order of the assignment), then the
            // assignment: its values, R's variables from that entry, 0 elsewhere.
            uint32_t best = 0, best_key = 0xFFFFFFFFu;
            for (uint64_t mbits = bits; mbits; mbits &= mbits - 1) {
              const uint32_t e = (uint32_t)__builtin_ctzll(mbits);
              const uint32_t key = __builtin_bitreverse32(e) >> (32 - (rank > 0 ? rank : 1));
              if (key < best_key) {
                best_key = key;
                best = e;
              }
            }
            uint64_t q0 = r0, q1 = r1;
            for (int j = 0; j < rank; ++j) {
              const bool in_lo = q0 != 0;
              const uint32_t v = in_lo ? (uint32_t)__builtin_ctzll(q0) : 64u + (uint32_t)__builtin_ctzll(q1);
              if (in_lo) q0 &= q0 - 1;
              else q1 &= q1 - 1;
              if ((best >> j) & 1u) {
                if (v < 64) w0 |= 1ull << v;
                else w1 |= 1ull << (v - 64);
              }
            }
            if (lane < 2) p.witness[2 * (size_t)job.instance + lane] = lane ? w1 : w0;
          }
          atomicOr(&p.verdicts[job.instance], lane == 0 ? TS_SEARCH_SAT : 0u);
          if (!p.exhaustive) break;
        }
      }
      if (depth == base) break;
      fd = depth - 1;
      prop = false;
      continue;
    }
    // ---- branch: free variable of largest score (weighted positive + 1) x (weighted
    // negative + 1) (ties to the smaller index), its more frequent polarity first; or
    // (LEX) the smallest variable of R, false first ----
    uint32_t first;
    if (p.order == TS_SEARCH_ORDER_LEX) {
      const uint32_t bv = r0 ? (uint32_t)__builtin_ctzll(r0) : 64u + (uint32_t)__builtin_ctzll(r1);
      first = (bv << 1) | 1u;
    } else {
      // scores < 3841^2 < 2^24, so (score << 8 | 255 - variable) orders scores, then indices
      const uint32_t k0 = h0 ? ((score_of(h0) << 8) | (255u - (uint32_t)lane)) : 0u;
      const uint32_t k1 = h1 ? ((score_of(h1) << 8) | (255u - (uint32_t)(lane + 64))) : 0u;
      const uint32_t key = wave_reduce(k0 > k1 ? k0 : k1, OpMax());
      const uint32_t bv = 255u - (key & 255u);
      const uint64_t neg0 = __ballot(((h0 >> 32) & 0xFFFFu) < (h0 >> 48));
      const uint64_t neg1 = __ballot(((h1 >> 32) & 0xFFFFu) < (h1 >> 48));
      const uint32_t negfirst = (uint32_t)(((bv < 64 ? neg0 : neg1) >> (bv & 63)) & 1ull);
      first = (bv << 1) | negfirst;
    }
    sh.stk[0][depth] = s.ae;
    sh.stk[1][depth] = s.ao;
    sh.stk[2][depth] = s.ve;
    sh.stk[3][depth] = s.vo;
    sh.stk_lit[depth] = first;  // next child 0
    __syncthreads();
    fd = depth;
    prop = false;
  }
}

// The job runs the first of the variants NS, Rest... (increasing clause slots per lane)
// that holds its clauses.
template <int NS, int... Rest>
__device__ __forceinline__ void search_variant(int need, const ts_search& p, const ts_search_job& job,
                                               const ts_search_inst& inst, Shared& sh, int lane, JobResult& r,
                                               const Queue& q) {
  if constexpr (sizeof...(Rest) == 0) {
    search_job<NS>(p, job, inst, sh, lane, r, q);
  } else {
    if (need <= NS)
      search_job<NS>(p, job, inst, sh, lane, r, q);
    else
      search_variant<Rest...>(need, p, job, inst, sh, lane, r, q);
  }
}

template <int... V>
constexpr int last_of() {
  int x = 0;
  ((x = V), ...);
  return x;
}

// One job record: its search (the smallest variant of V... that holds its instance),
// verdict bits and statistics in r.
template <int... V>
__device__ __forceinline__ void run_record(const ts_search& p, const ts_search_job& job, Shared& sh, int lane,
                                           JobResult& r, const Queue& q) {
  constexpr int NSMAX = last_of<V...>();
  if (job.instance < 0) return;
  const ts_search_inst inst = p.instances[job.instance];
  const int need = (int)uni((uint32_t)((inst.num_clauses + 63) / 64));
  if (inst.num_clauses > 64 * NSMAX) {
    // More clauses than this kernel holds (the host picks the kernel from the largest
    // instance): undecided, never a wrong verdict.
    r.status = 2;
  } else if (p.early_exit && (uni(__hip_atomic_load(&p.verdicts[job.instance], __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT)) & TS_SEARCH_SAT)) {
    r.status = 1;  // decided by another job already
  } else {
    sh.hist[lane] = 0;
    sh.hist[lane + 64] = 0;
    build_clause_table(sh.clause_table, lane);
    __syncthreads();
    search_variant<V...>(need, p, job, inst, sh, lane, r, q);
  }
  if (r.status == 2) atomicOr(&p.verdicts[job.instance], lane == 0 ? TS_SEARCH_INCOMPLETE : 0u);
  if (p.counts && (r.count || r.count_overflow)) {
    // whole-wave atomics, lane 0 carrying the value
    const unsigned long long add = lane == 0 ? (unsigned long long)r.count : 0ull;
    const unsigned long long old =
        atomicAdd(reinterpret_cast<unsigned long long*>(&p.counts[job.instance]), add);
    const bool wrapped = lane == 0 && old + add < old;
    atomicOr(p.count_overflow, (lane == 0 && (wrapped || r.count_overflow)) ? 1 : 0);
  }
}

// A job's statistics, written by lanes 0..9 (one dword each).
__device__ __forceinline__ void write_stat(ts_search_stat* out, const JobResult& r, int lane) {
  uint32_t d;
  switch (lane) {
    case 0: d = (uint32_t)r.status; break;
    case 1: d = (uint32_t)r.max_depth; break;
    case 2: d = (uint32_t)r.nodes; break;
    case 3: d = (uint32_t)((uint64_t)r.nodes >> 32); break;
    case 4: d = (uint32_t)r.leaves; break;
    case 5: d = (uint32_t)((uint64_t)r.leaves >> 32); break;
    case 6: d = (uint32_t)r.conflicts; break;
    case 7: d = (uint32_t)((uint64_t)r.conflicts >> 32); break;
    case 8: d = (uint32_t)r.digest; break;
    default: d = (uint32_t)(r.digest >> 32); break;
  }
  if (lane < 10) reinterpret_cast<uint32_t*>(out)[lane] = d;
}

}  // namespace

// Round mode: one workgroup per job of p.jobs.  Variants V... (increasing clause slots
// per lane); the host picks the kernel whose largest variant holds the batch's largest
// instance.
template <int... V>
__global__ __launch_bounds__(64) void k_search(ts_search p) {
  __shared__ Shared sh;
  const int lane = (int)threadIdx.x;
  const ts_search_job job = p.jobs[blockIdx.x];
  JobResult r{};
  run_record<V...>(p, job, sh, lane, r, Queue{nullptr, nullptr});
  if (p.stats) write_stat(&p.stats[blockIdx.x], r, lane);
}

namespace {

// A queue record as written by other workgroups in this launch: read past every cache
// (system-scope loads, lanes 0..15 one dword each) after its flag was acquired, then
// made uniform.
__device__ __forceinline__ ts_search_job load_record(const ts_search_job* rec, int lane) {
  const uint32_t* w = reinterpret_cast<const uint32_t*>(rec);
  const uint32_t x = lane < 16 ? __hip_atomic_load(w + lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) : 0u;
  uint32_t d[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) d[i] = (uint32_t)__builtin_amdgcn_readlane((int)x, i);
  ts_search_job j;
  j.instance = (int32_t)d[0];
  j.depth = (int32_t)d[1];
  j.split = (int32_t)d[2];
  j.own_from = (int32_t)d[3];
  j.prefix = d[4];
  j.reserved[0] = d[5];
  j.reserved[1] = d[6];
  j.reserved[2] = d[7];
  j.assigned[0] = (uint64_t)d[8] | ((uint64_t)d[9] << 32);
  j.assigned[1] = (uint64_t)d[10] | ((uint64_t)d[11] << 32);
  j.value[0] = (uint64_t)d[12] | ((uint64_t)d[13] << 32);
  j.value[1] = (uint64_t)d[14] | ((uint64_t)d[15] << 32);
  return j;
}

// Wall-clock ticks (100 MHz) a worker waits for a queue record before it gives up with
// the error flag: a safety net of 30 s, far beyond any job's length in this search.
constexpr uint64_t kWaitTicks = 3000000000ull;

}  // namespace

// Queue mode: the next record to take, claimed in order: a start record at once, a
// handed-over one once it is flagged (then after an acquire); -1 when the queue is
// drained (nothing pending: no job runs and none will ever be written at the claim) or
// the claim lies beyond the capacity.  Every lane takes part in each atomic (lane 0
// adds), so no branch in the worker loop depends on the lane; the result is
// wave-uniform.  Polls back off (sleeps growing to about 4 us, the pending count read
// every fourth poll) so that idle workers do not load the memory system the running
// ones use.
__device__ __forceinline__ int next_record(const ts_search& p, const Queue& q, int lane) {
  const int idx = (int)uni((uint32_t)atomicAdd(&q.state[Q_HEAD], lane == 0 ? 1 : 0));
  if (idx >= p.spill_capacity) return -1;
  if (idx < p.num_jobs) return idx;
  const uint64_t t0 = wall_clock64();
  int result = -1;
  for (uint32_t poll = 0;; ++poll) {
    if (uni(__hip_atomic_load(&q.ready[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
      result = idx;
      break;
    }
    if ((poll & 3u) == 3u) {
      if (uni((uint32_t)__hip_atomic_load(&q.state[Q_PENDING], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) == 0u)
        break;
      if (wall_clock64() - t0 > kWaitTicks) {
        atomicOr(&q.state[Q_ERROR], lane == 0 ? 1 : 0);
        break;
      }
    }
    if (poll < 8)
      __builtin_amdgcn_s_sleep(8);
    else if (poll < 32)
      __builtin_amdgcn_s_sleep(32);
    else
      __builtin_amdgcn_s_sleep(127);
  }
  if (result >= 0) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // the record's stores, after its flag
  return (int)uni((uint32_t)result);
}

// Queue mode: workers take records in claim order -- the start records (hardest first)
// and then the hand-overs in order of publication -- until the queue is drained.  No
// record belongs to a worker before the worker claims it, so the launch needs no
// assumption about how many workgroups are resident at once: a waiting worker waits for
// a record whose producer, or whose claimer, is running.  Every job finishes (its nodes
// are bounded), so every wait ends; a wall-clock limit turns a wait that would not into
// the error flag.
template <int... V>
__global__ __launch_bounds__(64) void k_search_queue(ts_search p, Queue q) {
  __shared__ Shared sh;
  const int lane = (int)threadIdx.x;
  for (int idx = next_record(p, q, lane); idx >= 0; idx = next_record(p, q, lane)) {
    const ts_search_job job = idx < p.num_jobs ? p.jobs[idx] : load_record(p.spill + idx, lane);
    JobResult r{};
    run_record<V...>(p, job, sh, lane, r, q);
    if (p.stats) write_stat(&p.stats[idx], r, lane);
    atomicAdd(&q.state[Q_PENDING], lane == 0 ? -1 : 0);  // continuations were published before
  }
}

__global__ void k_queue_init(int32_t* state, uint32_t* ready, int32_t capacity, int32_t num_jobs) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < capacity) ready[i] = 0u;
  if (i < TS_SEARCH_QUEUE_STATE) state[i] = (i == Q_TAIL || i == Q_PENDING) ? num_jobs : 0;
}

extern "C" int ts_search_run(const ts_search* p, ts_stream stream) {
  if (!p || p->num_jobs < 0 || p->num_instances < 0) return TS_EINVAL;
  if (p->width < 0 || p->width > TS_SEARCH_MAX_WIDTH || p->max_nodes <= 0) return TS_EINVAL;
  if (p->spill && (!p->spill_count || p->spill_capacity < 0)) return TS_EINVAL;
  if (p->max_clauses < 0 || p->max_clauses > 64 * TS_SEARCH_SLOTS) return TS_EINVAL;
  if (p->order != TS_SEARCH_ORDER_SCORE && p->order != TS_SEARCH_ORDER_LEX) return TS_EINVAL;
  if (p->num_jobs == 0) return TS_OK;
  if (!p->clauses || !p->instances || !p->jobs || !p->verdicts) return TS_EINVAL;
  // The kernel whose largest variant holds the largest instance; each job runs the
  // smallest of its variants that holds its own.
  const int need = p->max_clauses ? (p->max_clauses + 63) / 64 : TS_SEARCH_SLOTS;
  const dim3 grid((unsigned)p->num_jobs), block(64);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (need <= 2)
    hipLaunchKernelGGL((k_search<2>), grid, block, 0, st, *p);
  else if (need <= 4)
    hipLaunchKernelGGL((k_search<2, 3, 4>), grid, block, 0, st, *p);
  else if (need <= 6)
    hipLaunchKernelGGL((k_search<2, 3, 4, 5, 6>), grid, block, 0, st, *p);
  else if (need <= 8)
    hipLaunchKernelGGL((k_search<4, 6, 8>), grid, block, 0, st, *p);
  else
    hipLaunchKernelGGL((k_search<4, 8, TS_SEARCH_SLOTS>), grid, block, 0, st, *p);
  return hipGetLastError() == hipSuccess ? TS_OK : TS_EHIP;
}

namespace {

// Workers: the workgroups the device holds at once (more would only queue behind them;
// a worker that is not resident yet owns no record, so none waits for it).
template <int... V>
int launch_queue(const ts_search& p, const Queue& q, int workers, hipStream_t st) {
  int dev = 0, per_cu = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return TS_EHIP;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_search_queue<V...>, 64, 0) != hipSuccess)
    return TS_EHIP;
  const int resident = per_cu * cus;
  if (resident < 1) return TS_EHIP;
  if (workers <= 0 || workers > resident) workers = resident;
  workers = workers < p.spill_capacity ? workers : p.spill_capacity;
  if (workers < 1) workers = 1;
  const int cap = p.spill_capacity;
  const int n = cap > TS_SEARCH_QUEUE_STATE ? cap : TS_SEARCH_QUEUE_STATE;
  hipLaunchKernelGGL(k_queue_init, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, q.state, q.ready, cap,
                     p.num_jobs);
  hipLaunchKernelGGL((k_search_queue<V...>), dim3((unsigned)workers), dim3(64), 0, st, p, q);
  return hipGetLastError() == hipSuccess ? TS_OK : TS_EHIP;
}

}  // namespace

extern "C" int ts_search_run_queue(const ts_search* s, ts_search_job* records, int32_t capacity, int32_t* state,
                                   uint32_t* ready, int32_t workers, ts_stream stream) {
  if (!s || !records || !state || !ready || capacity <= 0) return TS_EINVAL;
  if (s->num_jobs < 0 || s->num_jobs > capacity || s->num_instances < 0) return TS_EINVAL;
  if (s->width < 0 || s->width > TS_SEARCH_MAX_WIDTH || s->max_nodes <= 0) return TS_EINVAL;
  if (s->max_clauses < 0 || s->max_clauses > 64 * TS_SEARCH_SLOTS) return TS_EINVAL;
  if (s->order != TS_SEARCH_ORDER_SCORE && s->order != TS_SEARCH_ORDER_LEX) return TS_EINVAL;
  if (s->num_jobs > 0 && (!s->clauses || !s->instances || !s->verdicts)) return TS_EINVAL;
  ts_search p = *s;
  p.jobs = records;
  p.spill = records;
  p.spill_count = state + Q_TAIL;
  p.spill_capacity = capacity;
  const Queue q{state, ready};
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const int need = p.max_clauses ? (p.max_clauses + 63) / 64 : TS_SEARCH_SLOTS;
  if (need <= 2) return launch_queue<2>(p, q, workers, st);
  if (need <= 4) return launch_queue<2, 3, 4>(p, q, workers, st);
  if (need <= 6) return launch_queue<2, 3, 4, 5, 6>(p, q, workers, st);
  if (need <= 8) return launch_queue<4, 6, 8>(p, q, workers, st);
  return launch_queue<4, 8, TS_SEARCH_SLOTS>(p, q, workers, st);
}

// ---- the allocation-free entry point ----------------------------------------------
//
// The continuation lists of ts_search_decide_noalloc live in this file-scope device
// array: a 256-byte counter line, then two lists of TS_SEARCH_STATIC_JOBS jobs used in
// turn by successive rounds.  Host code never uses the array's name as an address (it
// is a host-side stand-in there): each device's copy is located once with
// hipGetSymbolAddress and that address is passed to the kernels.
namespace {

constexpr size_t kScratchBytes = 256 + 2 * (size_t)TS_SEARCH_STATIC_JOBS * sizeof(ts_search_job);
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch[kMaxDevices] = {};

__global__ void k_zero_i32(int32_t* p) { *p = 0; }

// This device's copy of `scratch` (located on the first call made on the device).
int device_scratch(unsigned char** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return TS_EHIP;
  if (dev < 0 || dev >= kMaxDevices) return TS_EUNSUPPORTED;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return TS_EHIP;
    g_scratch[dev] = ptr;
  }
  *out = static_cast<unsigned char*>(g_scratch[dev]);
  return TS_OK;
}

}  // namespace

extern "C" int ts_search_decide_noalloc(const ts_search* s, int32_t* rounds_out, ts_stream stream) {
  if (!s || s->spill || s->spill_count) return TS_EINVAL;
  if (s->spill_capacity < 0 || s->spill_capacity > TS_SEARCH_STATIC_JOBS) return TS_EINVAL;
  if (s->num_jobs < 0 || s->width < 0 || s->width > TS_SEARCH_MAX_WIDTH || s->max_nodes <= 0) return TS_EINVAL;
  if (s->num_jobs > 0 && (!s->clauses || !s->instances || !s->jobs || !s->verdicts)) return TS_EINVAL;
  unsigned char* base = nullptr;
  const int rc = device_scratch(&base);
  if (rc != TS_OK) return rc;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  int32_t* count = reinterpret_cast<int32_t*>(base);
  ts_search_job* lists[2] = {reinterpret_cast<ts_search_job*>(base + 256),
                             reinterpret_cast<ts_search_job*>(base + 256) + TS_SEARCH_STATIC_JOBS};
  ts_search p = *s;
  p.spill_capacity = s->spill_capacity ? s->spill_capacity : TS_SEARCH_STATIC_JOBS;
  p.spill_count = count;
  int32_t rounds = 0;
  int32_t n = s->num_jobs;
  while (n > 0) {
    p.spill = lists[rounds & 1];
    hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(1), 0, st, count);
    if (hipGetLastError() != hipSuccess) return TS_EHIP;
    const int r = ts_search_run(&p, stream);
    if (r != TS_OK) return r;
    ++rounds;
    int32_t next = 0;
    if (hipMemcpyAsync(&next, count, sizeof(next), hipMemcpyDeviceToHost, st) != hipSuccess) return TS_EHIP;
    if (hipStreamSynchronize(st) != hipSuccess) return TS_EHIP;
    n = next < p.spill_capacity ? next : p.spill_capacity;
    p.jobs = lists[(rounds - 1) & 1];
    p.num_jobs = n;
    p.stats = nullptr;  // statistics cover the caller's own jobs only
  }
  if (rounds_out) *rounds_out = rounds;
  return TS_OK;
}

