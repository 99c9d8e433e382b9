"""
Plain PyTorch fp32 reference of the learned planner's network (CPU, autograd), used to
check the HIP forward and backward kernels.  Weights, input features and stored
activations are rounded to bf16 the way the device keeps them (straight-through in the
backward pass); sums and products run in fp32.
"""

from __future__ import annotations

import numpy as np
import torch

from tensorsat_amd.learn.model import HID, LAYERS, param_layout


def _bf(x: torch.Tensor) -> torch.Tensor:
    return x + (x.to(torch.bfloat16).to(torch.float32) - x).detach()


def _signed_sums(rows, cols, w, X, nrows):
    pos = torch.zeros((nrows, X.shape[1]), dtype=X.dtype)
    neg = torch.zeros((nrows, X.shape[1]), dtype=X.dtype)
    p = w > 0
    aw = w.abs()[:, None]
    pos = pos.index_add(0, rows[p], aw[p] * X[cols[p]])
    neg = neg.index_add(0, rows[~p], aw[~p] * X[cols[~p]])
    return pos, neg


def forward(params: np.ndarray, g, target=None, mask=None, layers: int = LAYERS):
    """Returns (y, loss, grads) with grads a flat fp32 vector in the parameter layout
    (None without a target)."""
    layout, total = param_layout(HID, layers)
    p = torch.tensor(params, dtype=torch.float32, requires_grad=True)

    def get(name, bf16):
        off, shp = layout[name]
        t = p[off : off + int(np.prod(shp))].reshape(shp)
        if bf16:
            t = t + (t.detach().to(torch.bfloat16).to(torch.float32) - t.detach())
        return t

    V, C = g.num_vars, g.num_clauses
    c_rows = torch.from_numpy(np.repeat(np.arange(C), np.diff(g.c2v_rowptr)).astype(np.int64))
    c_cols = torch.from_numpy(g.c2v_col.astype(np.int64))
    c_sign = torch.from_numpy(g.c2v_w)
    v_rows = torch.from_numpy(np.repeat(np.arange(V), np.diff(g.v2c_rowptr)).astype(np.int64))
    v_cols = torch.from_numpy(g.v2c_col.astype(np.int64))
    v_sign = torch.from_numpy(g.v2c_w)
    fv = _bf(torch.from_numpy(g.var_feat))
    fc = _bf(torch.from_numpy(g.clause_feat))
    hv = _bf(torch.relu(fv @ get("W_vin", True) + get("b_vin", False)))
    hc = _bf(torch.relu(fc @ get("W_cin", True) + get("b_cin", False)))
    for l in range(layers):
        sp, sn = _signed_sums(c_rows, c_cols, c_sign, hv, C)
        xc = torch.cat([hc, _bf(sp), _bf(sn)], 1)
        hc = _bf(torch.relu(xc @ get(f"Wc{l}", True) + get(f"bc{l}", False)))
        sp, sn = _signed_sums(v_rows, v_cols, v_sign, hc, V)
        xv = torch.cat([hv, _bf(sp), _bf(sn)], 1)
        hv = _bf(torch.relu(xv @ get(f"Wv{l}", True) + get(f"bv{l}", False)))
    y = hv @ get("w_out", False) + get("b_out", False)
    if target is None:
        return y.detach().numpy(), None, None
    t = torch.from_numpy(np.asarray(target, np.float32))
    m = torch.ones(V) if mask is None else torch.from_numpy(np.asarray(mask, np.float32))
    loss = (m * (y - t) ** 2).sum() / max(float(m.sum()), 1.0)
    loss.backward()
    return y.detach().numpy(), float(loss.detach()), p.grad.numpy().copy()


__all__ = ["forward", "HID", "LAYERS"]
