/* ORACLE — TEST INFRASTRUCTURE ONLY.
 * CPU restatement of the device cube search (the batched decision path): conditioned
 * slicing of a CNF network by a depth-first branch over variables, each branch fixing a
 * variable wire in every leaf that holds it (the reference's Box.__getitem__,
 * tensorsat/diagrams.py:972-996 -> FinRel._getitem, lang/fin_rel.py:342-343) and the
 * branch results OR-ed (Box.__add__ -> FinRel._union, fin_rel.py:345-346).  Two exact
 * tensor identities prune branches: a clause tensor with every literal false is all
 * zero (the branch contributes nothing), and a clause with one free literal left is a
 * one-hot vector (only that value of the variable contributes; it is fixed too).  Once
 * the residual network has at most `width` free variables it is contracted: every
 * residual clause tensor is multiplied into one rank-|R| relation over the residual
 * variables (FinRel._contract2 with every wire kept, fin_rel.py:255-288) and the
 * relation is reduced over all its axes (FinRel._rewire to no ports, fin_rel.py:379-393);
 * for |R| <= 6 at most 64 terms meet, so the GT0 and MOD256 epilogues agree.
 *
 * The rules below define the search the device kernel runs (tensorsat_amd/csrc/
 * search.hip), so the statistics must agree exactly:
 *   - clauses: duplicate literals removed, tautologies dropped, clause order kept;
 *   - root: an empty clause, or a conflict of unit propagation from the unit clauses,
 *     decides UNSAT (conflicts = 1, nodes = 0);
 *   - node (after propagation without conflict): nodes += 1; residual = clauses with no
 *     true literal; R = their free variables; no residual clause -> leaf with result 1
 *     (the empty product); |R| <= width -> leaf, result = OR over the rank-|R| relation;
 *     else branch on the free variable of largest score = (P + 1) * (N + 1), P and N
 *     the sums over residual clauses holding it as a positive / negative literal of a
 *     weight 5 (two free literals) or 1 (three), ties to the smaller index, its more
 *     frequent polarity among residual clauses (by plain count) first, positive on a tie;
 *   - a branch assigns the literal and propagates to the fixpoint (conflict: conflicts
 *     += 1, the branch is skipped);
 *   - leaves += 1 at every leaf; digest += mix(relation bits, rank) (mod 2^64), the
 *     relation's entry e (bit j of e = value of the j-th smallest residual variable)
 *     being bit e of the 64-bit word;
 *   - without `exhaustive`, the search stops at the first leaf whose result is 1;
 *   - order LEX instead branches on the smallest residual variable, false first, and the
 *     first satisfiable leaf's witness (its assignment, R's variables from the relation
 *     entry that is smallest with R's smallest variable most significant, 0 for the other
 *     free variables) is the lexicographically smallest satisfying assignment.
 * Unit propagation to the fixpoint has one result whatever the order units are taken
 * in, so the device's parallel rounds and this sequential loop agree.
 *
 * Pinned by tests/test_search_oracle.py: verdicts against brute force (n <= 20) and the
 * independent DPLL (oracle/dpll.c); leaf relations and their reductions against
 * oracle.finrel's contract2 / rewire restatement.  The branch statistics (nodes, leaves,
 * conflicts, digest) have no counterpart in the reference: parity unpinned beyond the
 * verdicts and leaf relations (they check the device against this restatement). */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  int32_t verdict;   /* 1 SAT, 0 UNSAT, -1 node budget exhausted */
  int32_t max_depth;
  int64_t nodes, leaves, conflicts;
  uint64_t digest;
} search_stat;

typedef struct {
  int32_t instance;
  int32_t rank;
  uint64_t bits;
  int32_t vars[6];
  int32_t pad[2];
  uint64_t assigned[2], value[2]; /* the branch's assignment (variables < 128) */
} leaf_rec;

static uint64_t splitmix(uint64_t x) {
  uint64_t z = x + 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

uint64_t search_leaf_mix(uint64_t bits, int rank) { return splitmix(splitmix(bits) + (uint64_t)rank); }

typedef struct {
  int n, m;
  int* cl;   /* m x 3 literals (+-(v+1)), 0 = absent */
  int* len;
  int8_t* val;
  int width, exhaustive, order;
  uint64_t witness[2];
  int have_witness;
  int64_t max_nodes;
  search_stat st;
  leaf_rec* trace;
  int trace_cap, *trace_len, instance;
  int stop;
} Ctx;

static int lit_true(const Ctx* c, int l) {
  int8_t v = c->val[abs(l) - 1];
  return l > 0 ? v > 0 : v < 0;
}

/* Propagate to the fixpoint; 0 on conflict.  Assignments are recorded on `trail`. */
static int propagate(Ctx* c, int* trail, int* nt) {
  for (;;) {
    int changed = 0;
    for (int i = 0; i < c->m; ++i) {
      const int* q = c->cl + 3 * i;
      int sat = 0, nfree = 0, u = 0;
      for (int k = 0; k < c->len[i]; ++k) {
        if (lit_true(c, q[k])) {
          sat = 1;
          break;
        }
        if (c->val[abs(q[k]) - 1] == 0) {
          nfree++;
          u = q[k];
        }
      }
      if (sat) continue;
      if (nfree == 0) return 0;
      if (nfree == 1) {
        c->val[abs(u) - 1] = u > 0 ? 1 : -1;
        trail[(*nt)++] = abs(u) - 1;
        changed = 1;
      }
    }
    if (!changed) return 1;
  }
}

static void undo(Ctx* c, int* trail, int* nt, int mark) {
  while (*nt > mark) c->val[trail[--(*nt)]] = 0;
}

/* Returns the node's result (1: a leaf with result 1 below). */
static int dfs(Ctx* c, int* trail, int* nt, int depth) {
  if (c->stop) return 0;
  if (c->st.nodes >= c->max_nodes) {
    c->stop = 1;
    c->st.verdict = -1;
    return 0;
  }
  c->st.nodes++;
  if (depth > c->st.max_depth) c->st.max_depth = depth;
  int* pw = (int*)calloc((size_t)c->n, sizeof(int));
  int* nw = (int*)calloc((size_t)c->n, sizeof(int));
  int* pos = (int*)calloc((size_t)c->n, sizeof(int));
  int* neg = (int*)calloc((size_t)c->n, sizeof(int));
  char* inr = (char*)calloc((size_t)c->n, 1);
  int nres = 0;
  for (int i = 0; i < c->m; ++i) {
    const int* q = c->cl + 3 * i;
    int sat = 0, nfree = 0;
    for (int k = 0; k < c->len[i]; ++k) {
      if (lit_true(c, q[k])) sat = 1;
      else if (c->val[abs(q[k]) - 1] == 0) nfree++;
    }
    if (sat) continue;
    nres++;
    for (int k = 0; k < c->len[i]; ++k) {
      int v = abs(q[k]) - 1;
      if (c->val[v] != 0) continue;
      const int w = nfree == 2 ? 5 : 1;
      inr[v] = 1;
      if (q[k] > 0) {
        pos[v]++;
        pw[v] += w;
      } else {
        neg[v]++;
        nw[v] += w;
      }
    }
  }
  int r = 0;
  for (int v = 0; v < c->n; ++v)
    if (inr[v]) r++;
  int result = 0;
  if (nres == 0 || r <= c->width) {
    /* leaf: the rank-r relation over R (ascending), then OR over its entries */
    int vars[64];
    int k = 0;
    for (int v = 0; v < c->n && k < 64; ++v)
      if (inr[v]) vars[k++] = v;
    uint64_t bits = 0;
    if (nres == 0) {
      bits = 1;
      r = 0;
    } else {
      for (uint32_t e = 0; e < (1u << r); ++e) {
        for (int j = 0; j < r; ++j) c->val[vars[j]] = ((e >> j) & 1) ? 1 : -1;
        int ok = 1;
        for (int i = 0; i < c->m && ok; ++i) {
          const int* q = c->cl + 3 * i;
          int sat = 0;
          for (int t = 0; t < c->len[i]; ++t)
            if (lit_true(c, q[t])) sat = 1;
          ok = sat;
        }
        if (ok) bits |= 1ull << e;
      }
      for (int j = 0; j < r; ++j) c->val[vars[j]] = 0;
    }
    c->st.leaves++;
    c->st.digest += search_leaf_mix(bits, r);
    if (c->trace && *c->trace_len < c->trace_cap) {
      leaf_rec* L = &c->trace[(*c->trace_len)++];
      memset(L, 0, sizeof(*L));
      L->instance = c->instance;
      L->rank = r;
      L->bits = bits;
      for (int j = 0; j < r; ++j) L->vars[j] = vars[j];
      for (int v = 0; v < c->n && v < 128; ++v)
        if (c->val[v]) {
          L->assigned[v >> 6] |= 1ull << (v & 63);
          if (c->val[v] > 0) L->value[v >> 6] |= 1ull << (v & 63);
        }
    }
    result = bits != 0;
    if (result && !c->have_witness) {
      int best = 0;
      uint32_t best_key = 0xFFFFFFFFu;
      for (int e = 0; e < (1 << r); ++e) {
        if (!((bits >> e) & 1)) continue;
        uint32_t key = 0; /* e with bit 0 (R's smallest variable) most significant */
        for (int j = 0; j < r; ++j) key |= ((e >> j) & 1u) << (r - 1 - j);
        if (key < best_key) {
          best_key = key;
          best = e;
        }
      }
      c->witness[0] = c->witness[1] = 0;
      for (int v = 0; v < c->n && v < 128; ++v)
        if (c->val[v] > 0) c->witness[v >> 6] |= 1ull << (v & 63);
      for (int j = 0; j < r; ++j)
        if ((best >> j) & 1) c->witness[vars[j] >> 6] |= 1ull << (vars[j] & 63);
      c->have_witness = 1;
    }
    if (result && !c->exhaustive) c->stop = 1;
  } else {
    int bv = -1, first;
    if (c->order == 1) {
      for (int v = 0; v < c->n && bv < 0; ++v)
        if (inr[v]) bv = v;
      first = -(bv + 1);
    } else {
      int64_t best = -1;
      for (int v = 0; v < c->n; ++v) {
        const int64_t sc = (int64_t)(pw[v] + 1) * (nw[v] + 1);
        if (inr[v] && sc > best) {
          best = sc;
          bv = v;
        }
      }
      first = pos[bv] >= neg[bv] ? bv + 1 : -(bv + 1);
    }
    for (int t = 0; t < 2 && !c->stop; ++t) {
      int l = t ? -first : first;
      int mark = *nt;
      c->val[abs(l) - 1] = l > 0 ? 1 : -1;
      trail[(*nt)++] = abs(l) - 1;
      if (propagate(c, trail, nt)) {
        if (dfs(c, trail, nt, depth + 1)) result = 1;
      } else {
        c->st.conflicts++;
      }
      undo(c, trail, nt, mark);
    }
  }
  free(inr);
  free(neg);
  free(pos);
  free(nw);
  free(pw);
  return result;
}

static void run_one(int n, int m, const int64_t* loff, const int32_t* lits, int width, int exhaustive, int order,
                    int64_t max_nodes, search_stat* out, uint64_t* witness, leaf_rec* trace, int trace_cap,
                    int* trace_len, int instance) {
  Ctx c;
  memset(&c, 0, sizeof(c));
  c.n = n;
  c.cl = (int*)calloc((size_t)3 * (m > 0 ? m : 1), sizeof(int));
  c.len = (int*)calloc((size_t)(m > 0 ? m : 1), sizeof(int));
  c.val = (int8_t*)calloc((size_t)(n > 0 ? n : 1), 1);
  c.width = width;
  c.exhaustive = exhaustive;
  c.order = order;
  c.max_nodes = max_nodes;
  c.trace = trace;
  c.trace_cap = trace_cap;
  c.trace_len = trace_len;
  c.instance = instance;
  int empty = 0;
  for (int i = 0; i < m; ++i) {
    int tmp[3], k = 0, taut = 0;
    for (int64_t a = loff[i]; a < loff[i + 1]; ++a) {
      int l = lits[a], dup = 0;
      for (int j = 0; j < k; ++j) {
        if (tmp[j] == l) dup = 1;
        if (tmp[j] == -l) taut = 1;
      }
      if (!dup && k < 3) tmp[k++] = l;
    }
    if (taut) continue;
    if (k == 0) empty = 1;
    for (int j = 0; j < k; ++j) c.cl[3 * c.m + j] = tmp[j];
    c.len[c.m++] = k;
  }
  c.st.verdict = 0;
  int* trail = (int*)malloc(sizeof(int) * (size_t)(n + 1));
  int nt = 0;
  if (empty) {
    c.st.conflicts = 1;
  } else {
    /* unit clauses at the root */
    int ok = 1;
    for (int i = 0; i < c.m && ok; ++i)
      if (c.len[i] == 1) {
        int l = c.cl[3 * i], v = abs(l) - 1;
        int8_t want = l > 0 ? 1 : -1;
        if (c.val[v] == 0) {
          c.val[v] = want;
          trail[nt++] = v;
        } else if (c.val[v] != want) {
          ok = 0;
        }
      }
    if (ok) ok = propagate(&c, trail, &nt);
    if (!ok) {
      c.st.conflicts = 1;
    } else {
      int r = dfs(&c, trail, &nt, 0);
      if (c.st.verdict != -1) c.st.verdict = r;
    }
  }
  *out = c.st;
  if (witness) {
    witness[0] = c.have_witness ? c.witness[0] : 0;
    witness[1] = c.have_witness ? c.witness[1] : 0;
  }
  free(trail);
  free(c.val);
  free(c.len);
  free(c.cl);
}

typedef struct {
  int num, width, exhaustive, order, threads;
  int64_t max_nodes;
  const int32_t *nvars, *lits;
  const int64_t *coff, *loff;
  search_stat* out;
  uint64_t* witness;
  int next;
  pthread_mutex_t mu;
} Pool;

static void* worker(void* arg) {
  Pool* p = (Pool*)arg;
  for (;;) {
    pthread_mutex_lock(&p->mu);
    int i = p->next++;
    pthread_mutex_unlock(&p->mu);
    if (i >= p->num) break;
    int m = (int)(p->coff[i + 1] - p->coff[i]);
    run_one(p->nvars[i], m, p->loff + p->coff[i], p->lits, p->width, p->exhaustive, p->order, p->max_nodes,
            &p->out[i], p->witness ? p->witness + 2 * (size_t)i : NULL, NULL, 0, NULL, i);
  }
  return NULL;
}

/* Decide `num` instances: instance i has nvars[i] variables and clauses coff[i]..coff[i+1]
 * whose literals are lits[loff[c]..loff[c+1]) (+-(v+1), at most 3 per clause after
 * removing duplicates).  threads <= 1 runs on the calling thread; with `trace`, the
 * first trace_cap leaves (single-threaded runs only) are recorded. */
int search_batch(int num, const int32_t* nvars, const int64_t* coff, const int64_t* loff, const int32_t* lits,
                 int width, int exhaustive, int order, int64_t max_nodes, int threads, search_stat* out,
                 uint64_t* witness, leaf_rec* trace, int trace_cap, int* trace_len) {
  if (num < 0 || width < 0 || width > 6 || !out || order < 0 || order > 1) return -2;
  for (int i = 0; i < num; ++i) {
    if (nvars[i] < 0) return -2;
    for (int64_t c = coff[i]; c < coff[i + 1]; ++c) {
      int k = 0;
      for (int64_t a = loff[c]; a < loff[c + 1]; ++a) {
        if (lits[a] == 0 || abs(lits[a]) > nvars[i]) return -2;
        ++k;
      }
      if (k > 3) {
        /* more than 3 distinct literals cannot be packed */
        int d = 0;
        for (int64_t a = loff[c]; a < loff[c + 1]; ++a) {
          int dup = 0;
          for (int64_t b = loff[c]; b < a; ++b)
            if (lits[b] == lits[a]) dup = 1;
          d += !dup;
        }
        if (d > 3) return -3;
      }
    }
  }
  if (trace_len) *trace_len = 0;
  if (threads <= 1 || trace) {
    for (int i = 0; i < num; ++i) {
      int m = (int)(coff[i + 1] - coff[i]);
      run_one(nvars[i], m, loff + coff[i], lits, width, exhaustive, order, max_nodes, &out[i],
              witness ? witness + 2 * (size_t)i : NULL, trace, trace_cap, trace_len, i);
    }
    return 0;
  }
  Pool p;
  memset(&p, 0, sizeof(p));
  p.num = num;
  p.width = width;
  p.exhaustive = exhaustive;
  p.order = order;
  p.witness = witness;
  p.max_nodes = max_nodes;
  p.nvars = nvars;
  p.lits = lits;
  p.coff = coff;
  p.loff = loff;
  p.out = out;
  pthread_mutex_init(&p.mu, NULL);
  pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)threads);
  for (int t = 0; t < threads; ++t) pthread_create(&th[t], NULL, worker, &p);
  for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
  free(th);
  pthread_mutex_destroy(&p.mu);
  return 0;
}
