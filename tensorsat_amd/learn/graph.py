"""
Clause–variable incidence graphs for the learned planner.

A batch of CNF instances becomes one bipartite graph: variable rows (all instances
concatenated), clause rows, and two CSR incidences with literal signs: clause -> variables
and variable -> clauses.  Each incidence carries the weight +-1/sqrt(deg(c) deg(v)), signed
by the literal: the same value seen from either side, so one kernel computes the forward
neighbour sums and, over the transposed incidence, their backward pass.  Node features
are padded to 32 columns (one MFMA K step).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

FEAT = 32


@dataclass
class HostGraph:
    var_offset: np.ndarray      # int64 [I+1]
    clause_offset: np.ndarray   # int64 [I+1]
    c2v_rowptr: np.ndarray      # int32 [C+1]
    c2v_col: np.ndarray         # int32 [E]
    c2v_sign: np.ndarray        # int8  [E]
    c2v_w: np.ndarray           # float32 [E]
    v2c_rowptr: np.ndarray      # int32 [V+1]
    v2c_col: np.ndarray         # int32 [E]
    v2c_sign: np.ndarray        # int8  [E]
    v2c_w: np.ndarray           # float32 [E]
    var_feat: np.ndarray        # float32 [V, FEAT]
    clause_feat: np.ndarray     # float32 [C, FEAT]

    @property
    def num_vars(self) -> int:
        return int(self.var_offset[-1])

    @property
    def num_clauses(self) -> int:
        return int(self.clause_offset[-1])


def build(instances) -> HostGraph:
    """``instances``: sequence of (num_vars, clause list), objects with ``num_vars`` and
    ``clauses``, or objects with ``num_vars`` and a 2-D literal array ``lits`` (one row per
    clause, the vectorised path)."""
    nvs, lens_l, lits_l = [], [], []
    for inst in instances:
        if hasattr(inst, "lits"):
            a = np.asarray(inst.lits)
            nvs.append(int(inst.num_vars))
            lens_l.append(np.full(a.shape[0], a.shape[1], np.int64))
            lits_l.append(a.reshape(-1).astype(np.int64))
            continue
        if hasattr(inst, "num_vars"):
            n, cl = int(inst.num_vars), inst.clauses
        else:
            n, cl = int(inst[0]), inst[1]
        nvs.append(n)
        lens_l.append(np.fromiter((len(c) for c in cl), dtype=np.int64, count=len(cl)))
        lits_l.append(np.fromiter((l for c in cl for l in c), dtype=np.int64, count=int(lens_l[-1].sum())))
    return _build(nvs, lens_l, lits_l)


def _build(nvs, lens_l, lits_l) -> HostGraph:
    I = len(nvs)
    var_offset = np.zeros(I + 1, np.int64)
    var_offset[1:] = np.cumsum(nvs)
    ncl = [len(x) for x in lens_l]
    clause_offset = np.zeros(I + 1, np.int64)
    clause_offset[1:] = np.cumsum(ncl)
    lens = np.concatenate(lens_l) if lens_l else np.zeros(0, np.int64)
    lits = np.concatenate(lits_l) if lits_l else np.zeros(0, np.int64)
    inst_of_clause = np.repeat(np.arange(I), ncl)
    inst_of_lit = np.repeat(inst_of_clause, lens)
    var = (np.abs(lits) - 1 + var_offset[inst_of_lit]).astype(np.int32)
    sign = np.where(lits > 0, 1, -1).astype(np.int8)
    clause_of_lit = np.repeat(np.arange(len(lens)), lens).astype(np.int32)
    C = int(clause_offset[-1])
    V = int(var_offset[-1])
    c2v_rowptr = np.zeros(C + 1, np.int32)
    c2v_rowptr[1:] = np.cumsum(lens)
    order = np.argsort(var, kind="stable")
    v2c_col = clause_of_lit[order]
    v2c_sign = sign[order]
    vdeg = np.bincount(var, minlength=V)
    v2c_rowptr = np.zeros(V + 1, np.int32)
    v2c_rowptr[1:] = np.cumsum(vdeg)
    w = (sign / np.sqrt(lens[clause_of_lit] * vdeg[var])).astype(np.float32) if len(var) else np.zeros(0, np.float32)
    v2c_w = w[order]
    # features
    pos = np.bincount(var[sign > 0], minlength=V).astype(np.float32)
    neg = np.bincount(var[sign < 0], minlength=V).astype(np.float32)
    inst_of_var = np.repeat(np.arange(I), nvs)
    n_arr = np.asarray(nvs, np.float32)
    m_arr = np.asarray(ncl, np.float32)
    kmax = np.zeros(I, np.float32)
    np.maximum.at(kmax, inst_of_clause, lens.astype(np.float32)) if len(lens) else None
    ratio = m_arr / np.maximum(n_arr, 1)
    # second-order degree: number of incidences of a variable's clause neighbours
    deg2 = np.zeros(V, np.float32)
    if len(var):
        clause_len_of_lit = lens[clause_of_lit].astype(np.float32)
        np.add.at(deg2, var, clause_len_of_lit - 1)
    vf = np.zeros((V, FEAT), np.float32)
    vf[:, 0] = pos / 10.0
    vf[:, 1] = neg / 10.0
    vf[:, 2] = (pos + neg) / 20.0
    vf[:, 3] = deg2 / 60.0
    vf[:, 4] = kmax[inst_of_var] / 5.0
    vf[:, 5] = ratio[inst_of_var] / 20.0
    vf[:, 6] = np.log2(np.maximum(n_arr[inst_of_var], 1)) / 8.0
    vf[:, 7] = 1.0
    negs = np.bincount(clause_of_lit[sign < 0], minlength=C).astype(np.float32) if len(lens) else np.zeros(C, np.float32)
    cf = np.zeros((C, FEAT), np.float32)
    cf[:, 0] = lens / 5.0
    cf[:, 1] = negs / 5.0
    cf[:, 2] = (lens - negs) / 5.0
    cf[:, 3] = 1.0
    return HostGraph(var_offset, clause_offset, c2v_rowptr, var, sign, w, v2c_rowptr, v2c_col.astype(np.int32),
                     v2c_sign, v2c_w, vf, cf)


class DeviceGraph:
    """Device copies (torch) of a HostGraph; features in bf16."""

    def __init__(self, g: HostGraph, device: str = "cuda") -> None:
        import torch

        self.host = g
        self.V = g.num_vars
        self.C = g.num_clauses
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)
        self.c2v_rowptr = t(g.c2v_rowptr)
        self.c2v_col = t(g.c2v_col if g.c2v_col.size else np.zeros(1, np.int32))
        self.c2v_w = t(g.c2v_w if g.c2v_w.size else np.zeros(1, np.float32))
        self.v2c_rowptr = t(g.v2c_rowptr)
        self.v2c_col = t(g.v2c_col if g.v2c_col.size else np.zeros(1, np.int32))
        self.v2c_w = t(g.v2c_w if g.v2c_w.size else np.zeros(1, np.float32))
        self.var_feat = t(g.var_feat).to(torch.bfloat16)
        self.clause_feat = t(g.clause_feat).to(torch.bfloat16)
