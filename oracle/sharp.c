/* ORACLE — TEST INFRASTRUCTURE ONLY.
 * Independent exact model counter (#SAT) for CNF formulas, used to pin the contraction
 * counts of the counting workload.  It shares no code and no algorithm with the
 * product's conditioned-slicing front-end: it is a DPLL model counter with connected
 * components and a component cache (the published scheme of Bayardo & Pehoushek 2000,
 * "Counting models using connected components", and Sang et al. 2004, Cachet):
 *
 *   #F = 2^(free variables in no clause) * prod over connected components C of #C
 *   #C = #(C | x) + #(C | not x)      for a branch variable x of C,
 *
 * each branch followed by unit propagation (a conflict contributes 0) and a fresh split
 * of the remaining clauses into components; a component's count is cached under its
 * residual clauses (the free literals of each clause, in clause order).  Counts are
 * unsigned 128-bit integers (exact up to 2^128; the callers' instances have at most 127
 * variables) and are returned as decimal strings.  Literals are +-(v+1). */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

typedef struct {
  int n, m;
  const int32_t* lits;
  const int64_t* off;
  int *occ_start, *occ; /* per literal index 2v+neg: clause ids */
  int *nsat, *nfalse, *len;
  int8_t* val;
  int *trail, ntrail;
  int *queue, nq;
  /* cache: open addressing, keys stored in an arena */
  uint64_t* h_hash;
  int64_t* h_key; /* offset into arena, -1 empty */
  int* h_klen;
  u128* h_val;
  int64_t h_cap, h_used;
  int32_t* arena;
  int64_t arena_len, arena_cap;
  int64_t nodes, hits;
} S;

static int lidx(int l) { return 2 * (abs(l) - 1) + (l < 0); }

static int assign_lit(S* s, int l) {
  int v = abs(l) - 1;
  s->val[v] = l > 0 ? 1 : -1;
  s->trail[s->ntrail++] = l;
  int ok = 1;
  int a = lidx(l), b = lidx(-l);
  for (int i = s->occ_start[a]; i < s->occ_start[a + 1]; ++i) s->nsat[s->occ[i]]++;
  for (int i = s->occ_start[b]; i < s->occ_start[b + 1]; ++i) {
    int c = s->occ[i];
    s->nfalse[c]++;
    if (s->nsat[c] == 0) {
      if (s->nfalse[c] == s->len[c])
        ok = 0;
      else if (s->nfalse[c] == s->len[c] - 1) {
        for (int64_t k = s->off[c]; k < s->off[c + 1]; ++k)
          if (s->val[abs(s->lits[k]) - 1] == 0) {
            s->queue[s->nq++] = s->lits[k];
            break;
          }
      }
    }
  }
  return ok;
}

static int propagate(S* s) {
  while (s->nq > 0) {
    int l = s->queue[--s->nq];
    int v = abs(l) - 1;
    if (s->val[v] != 0) {
      if ((l > 0) != (s->val[v] > 0)) {
        s->nq = 0;
        return 0;
      }
      continue;
    }
    if (!assign_lit(s, l)) {
      s->nq = 0;
      return 0;
    }
  }
  return 1;
}

static void undo(S* s, int mark) {
  while (s->ntrail > mark) {
    int l = s->trail[--s->ntrail];
    int a = lidx(l), b = lidx(-l);
    for (int i = s->occ_start[a]; i < s->occ_start[a + 1]; ++i) s->nsat[s->occ[i]]--;
    for (int i = s->occ_start[b]; i < s->occ_start[b + 1]; ++i) s->nfalse[s->occ[i]]--;
    s->val[abs(l) - 1] = 0;
  }
}

static uint64_t mix(uint64_t h, uint64_t x) {
  h ^= x + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
  h *= 0xff51afd7ed558ccdull;
  return h ^ (h >> 33);
}

/* key of a component: free literals of each residual clause (clause ids ascending),
 * each clause terminated by 0 */
static int build_key(S* s, const int* cl, int ncl, int32_t* key) {
  int k = 0;
  for (int i = 0; i < ncl; ++i) {
    int c = cl[i];
    for (int64_t t = s->off[c]; t < s->off[c + 1]; ++t)
      if (s->val[abs(s->lits[t]) - 1] == 0) key[k++] = s->lits[t];
    key[k++] = 0;
  }
  return k;
}

static int cache_find(S* s, const int32_t* key, int klen, uint64_t h, u128* out) {
  int64_t i = (int64_t)(h & (uint64_t)(s->h_cap - 1));
  while (s->h_key[i] >= 0) {
    if (s->h_hash[i] == h && s->h_klen[i] == klen &&
        memcmp(s->arena + s->h_key[i], key, (size_t)klen * sizeof(int32_t)) == 0) {
      *out = s->h_val[i];
      return 1;
    }
    i = (i + 1) & (s->h_cap - 1);
  }
  return 0;
}

static void cache_put(S* s, const int32_t* key, int klen, uint64_t h, u128 v) {
  if (s->h_used * 2 >= s->h_cap || s->arena_len + klen > s->arena_cap) {
    /* full: start over (the cache only saves work) */
    for (int64_t i = 0; i < s->h_cap; ++i) s->h_key[i] = -1;
    s->h_used = 0;
    s->arena_len = 0;
    if (klen > s->arena_cap) return;
  }
  int64_t i = (int64_t)(h & (uint64_t)(s->h_cap - 1));
  while (s->h_key[i] >= 0) i = (i + 1) & (s->h_cap - 1);
  memcpy(s->arena + s->arena_len, key, (size_t)klen * sizeof(int32_t));
  s->h_key[i] = s->arena_len;
  s->h_klen[i] = klen;
  s->h_hash[i] = h;
  s->h_val[i] = v;
  s->arena_len += klen;
  s->h_used++;
}

static u128 count_clauses(S* s, const int* cl, int ncl, int* overflow);

/* Split residual clauses (unsatisfied, ids in cl) into components by shared free
 * variables and multiply their counts; free_in_scope counts variables of the scope that
 * are free and occur in no residual clause (each doubles the count). */
static u128 count_split(S* s, const int* cl, int ncl, int free_outside, int* overflow) {
  u128 total = 1;
  if (free_outside >= 128) {
    *overflow = 1;
    return 0;
  }
  total <<= free_outside;
  if (ncl == 0) return total;
  /* union-find over clause indices via variables */
  int* parent = (int*)malloc(sizeof(int) * ncl);
  int* owner = (int*)malloc(sizeof(int) * s->n); /* var -> a clause index holding it */
  for (int v = 0; v < s->n; ++v) owner[v] = -1;
  for (int i = 0; i < ncl; ++i) parent[i] = i;
  for (int i = 0; i < ncl; ++i) {
    int c = cl[i];
    for (int64_t t = s->off[c]; t < s->off[c + 1]; ++t) {
      int v = abs(s->lits[t]) - 1;
      if (s->val[v] != 0) continue;
      if (owner[v] < 0) {
        owner[v] = i;
        continue;
      }
      int a = i, b = owner[v];
      while (parent[a] != a) a = parent[a] = parent[parent[a]];
      while (parent[b] != b) b = parent[b] = parent[parent[b]];
      if (a != b) parent[a < b ? b : a] = a < b ? a : b;
    }
  }
  int* comp = (int*)malloc(sizeof(int) * ncl);
  for (int i = 0; i < ncl; ++i) {
    int a = i;
    while (parent[a] != a) a = parent[a];
    comp[i] = a;
  }
  int* buf = (int*)malloc(sizeof(int) * ncl);
  for (int root = 0; root < ncl && total != 0 && !*overflow; ++root) {
    if (comp[root] != root) continue;
    int k = 0;
    for (int i = root; i < ncl; ++i)
      if (comp[i] == root) buf[k++] = cl[i];
    u128 c = count_clauses(s, buf, k, overflow);
    if (c != 0 && total > ((u128)0 - 1) / c) *overflow = 1;
    total *= c;
  }
  free(buf);
  free(comp);
  free(owner);
  free(parent);
  return total;
}

/* One connected component: residual clauses cl (ids ascending), all unsatisfied with at
 * least two free literals. */
static u128 count_clauses(S* s, const int* cl, int ncl, int* overflow) {
  s->nodes++;
  int klen_max = 0;
  for (int i = 0; i < ncl; ++i) klen_max += (int)(s->off[cl[i] + 1] - s->off[cl[i]]) + 1;
  int32_t* key = (int32_t*)malloc(sizeof(int32_t) * (size_t)(klen_max + 1));
  int klen = build_key(s, cl, ncl, key);
  uint64_t h = 1469598103934665603ull;
  for (int i = 0; i < klen; ++i) h = mix(h, (uint64_t)(uint32_t)key[i]);
  u128 cached;
  if (cache_find(s, key, klen, h, &cached)) {
    s->hits++;
    free(key);
    return cached;
  }
  /* variables of the component and their occurrence counts */
  int* vars = (int*)malloc(sizeof(int) * s->n);
  int* occn = (int*)calloc((size_t)s->n, sizeof(int));
  int nv = 0;
  for (int i = 0; i < ncl; ++i) {
    int c = cl[i];
    for (int64_t t = s->off[c]; t < s->off[c + 1]; ++t) {
      int v = abs(s->lits[t]) - 1;
      if (s->val[v] != 0) continue;
      if (occn[v]++ == 0) vars[nv++] = v;
    }
  }
  int x = vars[0];
  for (int i = 1; i < nv; ++i)
    if (occn[vars[i]] > occn[x] || (occn[vars[i]] == occn[x] && vars[i] < x)) x = vars[i];
  u128 total = 0;
  int* rest = (int*)malloc(sizeof(int) * (ncl > 0 ? ncl : 1));
  for (int pol = 0; pol < 2 && !*overflow; ++pol) {
    int mark = s->ntrail;
    s->nq = 0;
    int ok = assign_lit(s, pol == 0 ? x + 1 : -(x + 1)) && propagate(s);
    if (ok) {
      int k = 0;
      for (int i = 0; i < ncl; ++i)
        if (s->nsat[cl[i]] == 0) rest[k++] = cl[i];
      /* component variables now free and in no remaining clause */
      int* inrest = (int*)calloc((size_t)s->n, sizeof(int));
      for (int i = 0; i < k; ++i)
        for (int64_t t = s->off[rest[i]]; t < s->off[rest[i] + 1]; ++t) inrest[abs(s->lits[t]) - 1] = 1;
      int freeout = 0;
      for (int i = 0; i < nv; ++i)
        if (s->val[vars[i]] == 0 && !inrest[vars[i]]) freeout++;
      free(inrest);
      u128 c = count_split(s, rest, k, freeout, overflow);
      if (total + c < total) *overflow = 1;
      total += c;
    }
    s->nq = 0;
    undo(s, mark);
  }
  free(rest);
  free(occn);
  free(vars);
  if (!*overflow) cache_put(s, key, klen, h, total);
  free(key);
  return total;
}

static void u128_to_dec(u128 v, char* out) {
  char tmp[64];
  int k = 0;
  if (v == 0) tmp[k++] = '0';
  while (v > 0) {
    tmp[k++] = (char)('0' + (int)(v % 10));
    v /= 10;
  }
  for (int i = 0; i < k; ++i) out[i] = tmp[k - 1 - i];
  out[k] = 0;
}

/* Exact model count of the formula; writes the decimal count into out (>= 48 bytes).
 * Returns 0 on success, -2 on invalid input, -3 if the count reaches 2^128. */
int sharp_count(int n, int m, const int64_t* off, const int32_t* lits, char* out, int64_t* nodes) {
  if (n < 0 || m < 0 || !off || !out) return -2;
  for (int c = 0; c < m; ++c)
    for (int64_t k = off[c]; k < off[c + 1]; ++k)
      if (lits[k] == 0 || abs(lits[k]) > n) return -2;
  S s;
  memset(&s, 0, sizeof(s));
  s.n = n;
  s.m = m;
  s.lits = lits;
  s.off = off;
  s.occ_start = (int*)calloc((size_t)(2 * n + 1), sizeof(int));
  s.occ = (int*)malloc(sizeof(int) * (size_t)(off[m] + 1));
  s.nsat = (int*)calloc((size_t)m + 1, sizeof(int));
  s.nfalse = (int*)calloc((size_t)m + 1, sizeof(int));
  s.len = (int*)calloc((size_t)m + 1, sizeof(int));
  s.val = (int8_t*)calloc((size_t)n + 1, 1);
  s.trail = (int*)malloc(sizeof(int) * (size_t)(n + 1));
  s.queue = (int*)malloc(sizeof(int) * (size_t)(off[m] + n + 1));
  s.h_cap = 1 << 20;
  s.h_hash = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)s.h_cap);
  s.h_key = (int64_t*)malloc(sizeof(int64_t) * (size_t)s.h_cap);
  s.h_klen = (int*)malloc(sizeof(int) * (size_t)s.h_cap);
  s.h_val = (u128*)malloc(sizeof(u128) * (size_t)s.h_cap);
  for (int64_t i = 0; i < s.h_cap; ++i) s.h_key[i] = -1;
  s.arena_cap = (int64_t)1 << 26;
  s.arena = (int32_t*)malloc(sizeof(int32_t) * (size_t)s.arena_cap);
  /* clauses: drop duplicate literals; a tautology is always satisfied (skipped) */
  int* keep = (int*)malloc(sizeof(int) * ((size_t)m + 1));
  int nkeep = 0, empty = 0;
  for (int c = 0; c < m; ++c) {
    int taut = 0;
    for (int64_t a = off[c]; a < off[c + 1]; ++a)
      for (int64_t b = a + 1; b < off[c + 1]; ++b)
        if (lits[a] == -lits[b]) taut = 1;
    if (taut) {
      s.nsat[c] = 1 << 20; /* never unsatisfied */
      continue;
    }
    s.len[c] = (int)(off[c + 1] - off[c]);
    if (s.len[c] == 0) empty = 1;
    keep[nkeep++] = c;
  }
  int rc = 0;
  u128 total = 0;
  int overflow = 0;
  if (!empty) {
    /* duplicate literals inside a clause: count each literal once by occurrence */
    for (int c = 0; c < m; ++c)
      for (int64_t a = off[c]; a < off[c + 1]; ++a) {
        int dup = 0;
        for (int64_t b = off[c]; b < a; ++b)
          if (lits[b] == lits[a]) dup = 1;
        if (dup) {
          s.len[c]--;
          continue;
        }
        s.occ_start[lidx(lits[a]) + 1]++;
      }
    for (int i = 0; i < 2 * n; ++i) s.occ_start[i + 1] += s.occ_start[i];
    int* fill = (int*)malloc(sizeof(int) * (size_t)(2 * n + 1));
    memcpy(fill, s.occ_start, sizeof(int) * (size_t)(2 * n + 1));
    for (int c = 0; c < m; ++c)
      for (int64_t a = off[c]; a < off[c + 1]; ++a) {
        int dup = 0;
        for (int64_t b = off[c]; b < a; ++b)
          if (lits[b] == lits[a]) dup = 1;
        if (!dup) s.occ[fill[lidx(lits[a])]++] = c;
      }
    free(fill);
    /* unique literal lists are what the key and propagation read: rewrite duplicates
     * away by treating repeated literals as already counted (len adjusted above) */
    int ok = 1;
    for (int i = 0; i < nkeep && ok; ++i) {
      int c = keep[i];
      if (s.len[c] == 1) {
        for (int64_t a = off[c]; a < off[c + 1]; ++a) s.queue[s.nq++] = lits[a];
      }
    }
    ok = propagate(&s);
    if (ok) {
      int k = 0;
      for (int i = 0; i < nkeep; ++i)
        if (s.nsat[keep[i]] == 0) keep[k++] = keep[i];
      int* inrest = (int*)calloc((size_t)n + 1, sizeof(int));
      for (int i = 0; i < k; ++i)
        for (int64_t t = off[keep[i]]; t < off[keep[i] + 1]; ++t) inrest[abs(lits[t]) - 1] = 1;
      int freeout = 0;
      for (int v = 0; v < n; ++v)
        if (s.val[v] == 0 && !inrest[v]) freeout++;
      free(inrest);
      total = count_split(&s, keep, k, freeout, &overflow);
    }
  }
  if (overflow) rc = -3;
  u128_to_dec(total, out);
  if (nodes) *nodes = s.nodes;
  free(keep);
  free(s.arena);
  free(s.h_val);
  free(s.h_klen);
  free(s.h_key);
  free(s.h_hash);
  free(s.queue);
  free(s.trail);
  free(s.val);
  free(s.len);
  free(s.nfalse);
  free(s.nsat);
  free(s.occ);
  free(s.occ_start);
  return rc;
}
