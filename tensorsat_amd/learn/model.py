"""
The learned planner's network: 6 message-passing layers of width 128 on the
clause–variable incidence graph, bf16 weights and activations, fp32 accumulation, fp32
master weights with Adam.  Every forward and backward operation runs in
``libtensorsat_amd.so`` (``ts_gnn_*``); PyTorch only holds the buffers.

Layer l (variables ``H_v``, clauses ``H_c``, ``+``/``-`` the literal signs,
``a_cv = 1/sqrt(deg(c) deg(v))``):

    H_c <- ReLU([H_c | sum_{v in c, +} a_cv H_v | sum_{v in c, -} a_cv H_v] Wc_l + bc_l)
    H_v <- ReLU([H_v | sum_{c ∋ v, +} a_cv H_c | sum_{c ∋ v, -} a_cv H_c] Wv_l + bv_l)

Inputs are embedded by one ReLU layer from 32 features; the head is ``y_v = H_v w + b``.
The target is a variable's normalised position in the elimination order of the planner,
so ascending ``y`` is a proposed elimination order and descending ``y`` a slicing order.
"""

from __future__ import annotations

import numpy as np

from .. import _native as N
from .graph import FEAT, DeviceGraph

HID = 128
LAYERS = 6


def param_layout(hidden: int = HID, layers: int = LAYERS):
    shapes = [("W_vin", (FEAT, hidden)), ("b_vin", (hidden,)), ("W_cin", (FEAT, hidden)), ("b_cin", (hidden,))]
    for l in range(layers):
        shapes += [(f"Wc{l}", (3 * hidden, hidden)), (f"bc{l}", (hidden,)),
                   (f"Wv{l}", (3 * hidden, hidden)), (f"bv{l}", (hidden,))]
    shapes += [("w_out", (hidden,)), ("b_out", (1,))]
    layout = {}
    off = 0
    for name, shp in shapes:
        n = int(np.prod(shp))
        layout[name] = (off, shp)
        off += n
    return layout, off


def init_params(seed: int = 0, hidden: int = HID, layers: int = LAYERS) -> np.ndarray:
    layout, total = param_layout(hidden, layers)
    rng = np.random.default_rng(seed)
    p = np.zeros(total, np.float32)
    for name, (off, shp) in layout.items():
        n = int(np.prod(shp))
        if name.startswith("W"):
            # He scale on an effective fan-in of 1.5 x hidden: the normalised neighbour
            # sums are smaller than the states, so the full 384 fan-in shrinks the
            # signal layer by layer and 128 grows it; 192 keeps the state scale
            # roughly constant over the 6 layers.
            fan_in = min(shp[0], 3 * hidden // 2)
            p[off : off + n] = rng.normal(0.0, np.sqrt(2.0 / fan_in), n).astype(np.float32)
        elif name == "w_out":
            p[off : off + n] = rng.normal(0.0, np.sqrt(1.0 / shp[0]), n).astype(np.float32)
        elif name.startswith("b") and name != "b_out":
            p[off : off + n] = 0.01
        elif name == "b_out":
            p[off] = 0.5
    return p


class GNN:
    def __init__(self, params: np.ndarray | None = None, *, seed: int = 0, device: str = "cuda",
                 layers: int = LAYERS) -> None:
        import torch

        N.lib()
        self.torch = torch
        self.device = device
        self.layers = layers
        self.layout, self.total = param_layout(HID, layers)
        if params is None:
            params = init_params(seed, HID, layers)
        if params.shape != (self.total,):
            raise ValueError("parameter vector has the wrong size")
        self.p = torch.from_numpy(params.astype(np.float32)).to(device)
        self.pbf = torch.empty(self.total, dtype=torch.bfloat16, device=device)
        N.check(N.lib().ts_gnn_cast_bf16(N.dptr(self.p), N.dptr(self.pbf), self.total, N.stream_ptr()), "cast")
        self.g = torch.zeros(self.total, dtype=torch.float32, device=device)
        self.m = torch.zeros(self.total, dtype=torch.float32, device=device)
        self.v = torch.zeros(self.total, dtype=torch.float32, device=device)
        self.steps = 0
        self.gemm_log = None  # a list to record every GEMM's flops, bytes and events into

    # -- views ---------------------------------------------------------------------
    def _ptr(self, buf, name: str, row: int = 0, elem: int = 4):
        off, shp = self.layout[name]
        cols = shp[1] if len(shp) == 2 else 1
        return buf.data_ptr() + (off + row * cols) * elem

    def fp(self, name):
        return self._ptr(self.p, name)

    def bf(self, name, row: int = 0):
        return self._ptr(self.pbf, name, row, 2)

    def gp(self, name):
        return self._ptr(self.g, name)

    # -- kernels ---------------------------------------------------------------------
    def _gemm(self, A, sam, sak, B, sbk, sbn, C, ldc, M, Nn, K, bias=None, relu=0, out_bf16=1, acc=0):
        log = self.gemm_log
        if log is not None:  # measurement: (flops, bytes, start, end), events on the current stream
            e0, e1 = self.torch.cuda.Event(enable_timing=True), self.torch.cuda.Event(enable_timing=True)
            e0.record()
        N.check(N.lib().ts_gnn_gemm(A, sam, sak, B, sbk, sbn, C, ldc, M, Nn, K, bias, relu, out_bf16, acc,
                                    N.stream_ptr()), "ts_gnn_gemm")
        if log is not None:
            e1.record()
            out_b = M * Nn * (2 if out_bf16 else 4) * (2 if acc else 1)  # read back when accumulating
            log.append((2.0 * M * Nn * K, 2.0 * (M * K + K * Nn) + out_b, e0, e1))

    def _segsum(self, rowptr, col, weight, nrows, X, ldx, xp, xn, xf32, out, ldo, op, on, acc):
        N.check(N.lib().ts_gnn_segsum(rowptr, col, weight, nrows, X, ldx, xp, xn, xf32, out, ldo, op, on, HID, acc,
                                      N.stream_ptr()), "ts_gnn_segsum")

    # -- forward / backward ------------------------------------------------------------
    def forward(self, dg: DeviceGraph, *, target=None, mask=None, train: bool = False):
        """Scores y (fp32 [V]); with ``train`` also the loss, gradients accumulated in
        ``self.g`` (zeroed first).  Returns (y, loss or None)."""
        torch = self.torch
        V, C = dg.V, dg.C
        bf16 = torch.bfloat16
        L = self.layers
        Xv = [torch.empty((max(V, 1), 3 * HID), dtype=bf16, device=self.device) for _ in range(L + 1)]
        Xc = [torch.empty((max(C, 1), 3 * HID), dtype=bf16, device=self.device) for _ in range(L + 1)]
        ld3 = 3 * HID
        self._gemm(dg.var_feat.data_ptr(), FEAT, 1, self.bf("W_vin"), HID, 1, Xv[0].data_ptr(), ld3, V, HID, FEAT,
                   bias=self.fp("b_vin"), relu=1)
        self._gemm(dg.clause_feat.data_ptr(), FEAT, 1, self.bf("W_cin"), HID, 1, Xc[0].data_ptr(), ld3, C, HID, FEAT,
                   bias=self.fp("b_cin"), relu=1)
        for l in range(L):
            self._segsum(dg.c2v_rowptr.data_ptr(), dg.c2v_col.data_ptr(), dg.c2v_w.data_ptr(), C,
                         Xv[l].data_ptr(), ld3, 0, 0, 0, Xc[l].data_ptr(), ld3, HID, 2 * HID, 0)
            self._gemm(Xc[l].data_ptr(), ld3, 1, self.bf(f"Wc{l}"), HID, 1, Xc[l + 1].data_ptr(), ld3, C, HID, ld3,
                       bias=self.fp(f"bc{l}"), relu=1)
            self._segsum(dg.v2c_rowptr.data_ptr(), dg.v2c_col.data_ptr(), dg.v2c_w.data_ptr(), V,
                         Xc[l + 1].data_ptr(), ld3, 0, 0, 0, Xv[l].data_ptr(), ld3, HID, 2 * HID, 0)
            self._gemm(Xv[l].data_ptr(), ld3, 1, self.bf(f"Wv{l}"), HID, 1, Xv[l + 1].data_ptr(), ld3, V, HID, ld3,
                       bias=self.fp(f"bv{l}"), relu=1)
        y = torch.empty(max(V, 1), dtype=torch.float32, device=self.device)
        if not train:
            N.check(N.lib().ts_gnn_head(Xv[L].data_ptr(), ld3, V, self.fp("w_out"), self.fp("b_out"), None, None,
                                        0.0, y.data_ptr(), None, None, None, None, HID, 0, N.stream_ptr()), "head")
            return y[:V], None
        self.g.zero_()
        loss = torch.zeros(1, dtype=torch.float32, device=self.device)
        dHv = torch.empty((max(V, 1), HID), dtype=torch.float32, device=self.device)
        cnt = float(mask.sum().item()) if mask is not None else float(V)
        N.check(N.lib().ts_gnn_head(Xv[L].data_ptr(), ld3, V, self.fp("w_out"), self.fp("b_out"),
                                    target.data_ptr(), mask.data_ptr() if mask is not None else None,
                                    1.0 / max(cnt, 1.0), y.data_ptr(), loss.data_ptr(), dHv.data_ptr(),
                                    self.gp("w_out"), self.gp("b_out"), HID, 1, N.stream_ptr()), "head")
        self._backward(dg, Xv, Xc, dHv)
        return y[:V], loss

    def _backward(self, dg, Xv, Xc, dHv):
        torch = self.torch
        V, C = dg.V, dg.C
        ld3 = 3 * HID
        dev = self.device
        bf16 = torch.bfloat16
        dHc = torch.zeros((max(C, 1), HID), dtype=torch.float32, device=dev)
        dZv = torch.empty((max(V, 1), HID), dtype=bf16, device=dev)
        dZc = torch.empty((max(C, 1), HID), dtype=bf16, device=dev)
        tmp_v = torch.empty((max(V, 1), 2 * HID), dtype=torch.float32, device=dev)
        tmp_c = torch.empty((max(C, 1), 2 * HID), dtype=torch.float32, device=dev)
        L = N.lib()
        s = N.stream_ptr()
        for l in reversed(range(self.layers)):
            # variable update of layer l
            N.check(L.ts_gnn_relu_back(dHv.data_ptr(), HID, Xv[l + 1].data_ptr(), ld3, V, HID, dZv.data_ptr(), HID,
                                       self.gp(f"bv{l}"), s), "relu_back")
            self._gemm(Xv[l].data_ptr(), 1, ld3, dZv.data_ptr(), HID, 1, self.gp(f"Wv{l}"), HID, ld3, HID, V,
                       out_bf16=0, acc=1)
            dHv_new = torch.empty((max(V, 1), HID), dtype=torch.float32, device=dev)
            self._gemm(dZv.data_ptr(), HID, 1, self.bf(f"Wv{l}"), 1, HID, dHv_new.data_ptr(), HID, V, HID, HID,
                       out_bf16=0)
            self._gemm(dZv.data_ptr(), HID, 1, self.bf(f"Wv{l}", HID), 1, HID, tmp_v.data_ptr(), 2 * HID, V, 2 * HID,
                       HID, out_bf16=0)
            self._segsum(dg.c2v_rowptr.data_ptr(), dg.c2v_col.data_ptr(), dg.c2v_w.data_ptr(), C,
                         tmp_v.data_ptr(), 2 * HID, 0, HID, 1, dHc.data_ptr(), HID, 0, 0, 1)
            # clause update of layer l
            N.check(L.ts_gnn_relu_back(dHc.data_ptr(), HID, Xc[l + 1].data_ptr(), ld3, C, HID, dZc.data_ptr(), HID,
                                       self.gp(f"bc{l}"), s), "relu_back")
            self._gemm(Xc[l].data_ptr(), 1, ld3, dZc.data_ptr(), HID, 1, self.gp(f"Wc{l}"), HID, ld3, HID, C,
                       out_bf16=0, acc=1)
            dHc_new = torch.empty((max(C, 1), HID), dtype=torch.float32, device=dev)
            self._gemm(dZc.data_ptr(), HID, 1, self.bf(f"Wc{l}"), 1, HID, dHc_new.data_ptr(), HID, C, HID, HID,
                       out_bf16=0)
            self._gemm(dZc.data_ptr(), HID, 1, self.bf(f"Wc{l}", HID), 1, HID, tmp_c.data_ptr(), 2 * HID, C, 2 * HID,
                       HID, out_bf16=0)
            self._segsum(dg.v2c_rowptr.data_ptr(), dg.v2c_col.data_ptr(), dg.v2c_w.data_ptr(), V,
                         tmp_c.data_ptr(), 2 * HID, 0, HID, 1, dHv_new.data_ptr(), HID, 0, 0, 1)
            dHv, dHc = dHv_new, dHc_new
        # input embeddings
        N.check(L.ts_gnn_relu_back(dHv.data_ptr(), HID, Xv[0].data_ptr(), ld3, V, HID, dZv.data_ptr(), HID,
                                   self.gp("b_vin"), s), "relu_back")
        self._gemm(dg.var_feat.data_ptr(), 1, FEAT, dZv.data_ptr(), HID, 1, self.gp("W_vin"), HID, FEAT, HID, V,
                   out_bf16=0, acc=1)
        N.check(L.ts_gnn_relu_back(dHc.data_ptr(), HID, Xc[0].data_ptr(), ld3, C, HID, dZc.data_ptr(), HID,
                                   self.gp("b_cin"), s), "relu_back")
        self._gemm(dg.clause_feat.data_ptr(), 1, FEAT, dZc.data_ptr(), HID, 1, self.gp("W_cin"), HID, FEAT, HID, C,
                   out_bf16=0, acc=1)

    def adam_step(self, lr: float = 1e-3, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8, wd: float = 0.0):
        self.steps += 1
        N.check(N.lib().ts_gnn_adam(N.dptr(self.p), N.dptr(self.g), N.dptr(self.m), N.dptr(self.v), N.dptr(self.pbf),
                                    self.total, lr, b1, b2, eps, self.steps, wd, N.stream_ptr()), "adam")

    def params(self) -> np.ndarray:
        return self.p.cpu().numpy().copy()

    def grads(self) -> np.ndarray:
        return self.g.cpu().numpy().copy()

    def save(self, path) -> None:
        np.savez(path, params=self.params(), steps=np.int64(self.steps), layers=np.int64(self.layers))

    @classmethod
    def load(cls, path, **kw) -> "GNN":
        with np.load(path, allow_pickle=False) as z:
            if "layers" in z:
                kw.setdefault("layers", int(z["layers"]))
            m = cls(z["params"], **kw)
            m.steps = int(z["steps"])
        return m
