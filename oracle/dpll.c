/* ORACLE — TEST INFRASTRUCTURE ONLY.
 * Independent SAT decision procedure used as a cross-check of contraction verdicts
 * (kissat, the reference's solver corpus source, is not available as code here).
 * DPLL: unit propagation over occurrence lists with counters, chronological
 * backtracking, branching on the most frequent unassigned variable in unsatisfied
 * clauses, positive polarity first.  Literals are +-(v+1). */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  int n, m;
  const int32_t* lits;
  const int64_t* off;
  int *occ_start, *occ;     /* per literal index 2v+neg: clause ids */
  int *nsat, *nfalse, *len;
  int8_t* val;
  int *trail, ntrail;
  int *queue, nq;
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
      if (s->nfalse[c] == s->len[c]) ok = 0;
      else if (s->nfalse[c] == s->len[c] - 1) {
        for (int64_t k = s->off[c]; k < s->off[c + 1]; ++k)
          if (s->val[abs(s->lits[k]) - 1] == 0) { s->queue[s->nq++] = s->lits[k]; break; }
      }
    }
  }
  return ok;
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

static int propagate(S* s) {
  while (s->nq > 0) {
    int l = s->queue[--s->nq];
    int8_t cur = s->val[abs(l) - 1];
    if (cur != 0) {
      if ((l > 0) != (cur > 0)) { s->nq = 0; return 0; }
      continue;
    }
    if (!assign_lit(s, l)) { s->nq = 0; return 0; }
  }
  return 1;
}

static int pick(S* s) {
  int best = -1, bc = -1;
  int* cnt = (int*)calloc((size_t)s->n, sizeof(int));
  for (int c = 0; c < s->m; ++c) {
    if (s->nsat[c]) continue;
    for (int64_t k = s->off[c]; k < s->off[c + 1]; ++k) {
      int v = abs(s->lits[k]) - 1;
      if (s->val[v] == 0) cnt[v]++;
    }
  }
  for (int v = 0; v < s->n; ++v)
    if (cnt[v] > bc) { bc = cnt[v]; best = v; }
  free(cnt);
  return bc > 0 ? best : -1;
}

static int solve(S* s, long* budget) {
  if (--(*budget) < 0) return -1;
  int v = pick(s);
  if (v < 0) return 1; /* every clause satisfied */
  for (int pol = 0; pol < 2; ++pol) {
    int mark = s->ntrail;
    int l = pol == 0 ? v + 1 : -(v + 1);
    int r = 0;
    if (assign_lit(s, l) && propagate(s)) r = solve(s, budget);
    s->nq = 0;
    undo(s, mark);
    if (r != 0) return r;
  }
  return 0;
}

/* Returns 1 SAT, 0 UNSAT, -1 budget exceeded, -2 bad input. */
int dpll_solve(int n, int m, const int64_t* off, const int32_t* lits, long budget) {
  if (n < 0 || m < 0) return -2;
  S s;
  memset(&s, 0, sizeof(s));
  s.n = n; s.m = m; s.lits = lits; s.off = off;
  s.occ_start = (int*)calloc((size_t)(2 * n + 1), sizeof(int));
  s.len = (int*)calloc((size_t)(m + 1), sizeof(int));
  for (int c = 0; c < m; ++c) {
    s.len[c] = (int)(off[c + 1] - off[c]);
    if (s.len[c] == 0) { free(s.occ_start); free(s.len); return 0; }
    for (int64_t k = off[c]; k < off[c + 1]; ++k) {
      int l = lits[k];
      if (l == 0 || abs(l) > n) { free(s.occ_start); free(s.len); return -2; }
      s.occ_start[lidx(l) + 1]++;
    }
  }
  for (int i = 0; i < 2 * n; ++i) s.occ_start[i + 1] += s.occ_start[i];
  int total = s.occ_start[2 * n];
  s.occ = (int*)malloc(sizeof(int) * (size_t)(total + 1));
  int* fill = (int*)calloc((size_t)(2 * n + 1), sizeof(int));
  for (int c = 0; c < m; ++c)
    for (int64_t k = off[c]; k < off[c + 1]; ++k) {
      int a = lidx(lits[k]);
      s.occ[s.occ_start[a] + fill[a]++] = c;
    }
  free(fill);
  s.nsat = (int*)calloc((size_t)(m + 1), sizeof(int));
  s.nfalse = (int*)calloc((size_t)(m + 1), sizeof(int));
  s.val = (int8_t*)calloc((size_t)(n + 1), 1);
  s.trail = (int*)malloc(sizeof(int) * (size_t)(n + 1));
  s.queue = (int*)malloc(sizeof(int) * (size_t)(n + total + 1));
  int res;
  for (int c = 0; c < m; ++c)
    if (s.len[c] == 1) s.queue[s.nq++] = lits[off[c]];
  if (!propagate(&s)) res = 0;
  else res = solve(&s, &budget);
  free(s.occ_start); free(s.occ); free(s.len); free(s.nsat); free(s.nfalse);
  free(s.val); free(s.trail); free(s.queue);
  return res;
}
