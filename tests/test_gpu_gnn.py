"""
GPU checks of the learned planner's kernels: the MFMA bf16 GEMM (every stride form the
network uses, split-K, bias/ReLU/accumulate epilogues) against float64 numpy on
bf16-exact inputs, the signed neighbour sums, and the whole network's forward pass and
gradients against the PyTorch fp32 reference.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bf16_exact(rng, shape, scale=1.0):
    """fp32 values exactly representable in bf16."""
    x = (rng.standard_normal(shape) * scale).astype(np.float32)
    return (x.view(np.uint32) & np.uint32(0xFFFF0000)).view(np.float32)


def _gemm(torch, A32, sam, sak, B32, sbk, sbn, M, N, K, C=None, ldc=None, bias=None, relu=0, out_bf16=0, acc=0):
    from tensorsat_amd import _native as NT

    A = torch.from_numpy(A32).cuda().to(torch.bfloat16)
    B = torch.from_numpy(B32).cuda().to(torch.bfloat16)
    if C is None:
        C = torch.zeros((M, N), dtype=torch.bfloat16 if out_bf16 else torch.float32, device="cuda")
    ldc = N if ldc is None else ldc
    bt = torch.from_numpy(bias).cuda() if bias is not None else None
    NT.check(NT.lib().ts_gnn_gemm(A.data_ptr(), sam, sak, B.data_ptr(), sbk, sbn, C.data_ptr(), ldc, M, N, K,
                                  bt.data_ptr() if bt is not None else None, relu, out_bf16, acc,
                                  NT.stream_ptr()), "gemm")
    torch.cuda.synchronize()
    return C.float().cpu().numpy()


@pytest.mark.parametrize("M,N,K", [(16, 16, 32), (64, 64, 32), (77, 130, 45), (1000, 128, 384), (5, 3, 1)])
def test_gemm_row_major_matches_float64(M, N, K):
    import torch

    rng = np.random.default_rng(M * 7 + N * 13 + K)
    A = _bf16_exact(rng, (M, K))
    B = _bf16_exact(rng, (K, N))
    ref = A.astype(np.float64) @ B.astype(np.float64)
    print(f"gemm {M}x{N}x{K}", flush=True)
    got = _gemm(torch, A.reshape(-1), K, 1, B.reshape(-1), N, 1, M, N, K)
    np.testing.assert_allclose(got, ref, rtol=1e-5, atol=1e-4 * np.sqrt(K))


def test_gemm_transposed_operands_and_epilogues():
    """A read column-major (X^T dZ form) and B transposed (dZ W^T form); bias + ReLU and
    bf16 output; accumulate into an existing fp32 result."""
    import torch

    rng = np.random.default_rng(5)
    M, N, K = 96, 200, 70
    At = _bf16_exact(rng, (K, M))  # A(m, k) = At[k, m]
    Bt = _bf16_exact(rng, (N, K))  # B(k, n) = Bt[n, k]
    ref = At.T.astype(np.float64) @ Bt.T.astype(np.float64)
    got = _gemm(torch, At.reshape(-1), 1, M, Bt.reshape(-1), 1, K, M, N, K)
    np.testing.assert_allclose(got, ref, rtol=1e-5, atol=1e-3)
    bias = rng.standard_normal(N).astype(np.float32)
    got = _gemm(torch, At.reshape(-1), 1, M, Bt.reshape(-1), 1, K, M, N, K, bias=bias, relu=1, out_bf16=1)
    want = np.maximum(ref + bias, 0)
    np.testing.assert_allclose(got, want, rtol=1e-2, atol=1e-2)
    C0 = torch.from_numpy(_bf16_exact(rng, (M, N))).cuda()
    base = C0.cpu().numpy().astype(np.float64)
    got = _gemm(torch, At.reshape(-1), 1, M, Bt.reshape(-1), 1, K, M, N, K, C=C0, acc=1)
    np.testing.assert_allclose(got, base + ref, rtol=1e-5, atol=1e-3)


def test_gemm_split_k_weight_gradient_shape():
    """384 x 128 output over K = 9000 rows: split across blocks with fp32 atomics."""
    import torch

    rng = np.random.default_rng(6)
    M, N, K = 384, 128, 9000
    X = _bf16_exact(rng, (K, M))
    dZ = _bf16_exact(rng, (K, N))
    ref = X.T.astype(np.float64) @ dZ.astype(np.float64)
    got = _gemm(torch, X.reshape(-1), 1, M, dZ.reshape(-1), N, 1, M, N, K, acc=1)
    np.testing.assert_allclose(got, ref, rtol=1e-4, atol=2e-3 * np.sqrt(K))


def test_segsum_forward_and_transpose():
    import torch

    from tensorsat_amd import _native as NT
    from tensorsat_amd.learn import graph as G
    from tensorsat_amd.sat import CNFInstance

    cnfs = [CNFInstance.random(3, 40 + i, 150 + 3 * i, rng=i) for i in range(6)]
    g = G.build(cnfs)
    dg = G.DeviceGraph(g)
    rng = np.random.default_rng(1)
    F = 128
    Xv = _bf16_exact(rng, (g.num_vars, F))
    X = torch.from_numpy(Xv).cuda().to(torch.bfloat16)
    out = torch.zeros((g.num_clauses, 3 * F), dtype=torch.bfloat16, device="cuda")
    NT.check(NT.lib().ts_gnn_segsum(dg.c2v_rowptr.data_ptr(), dg.c2v_col.data_ptr(), dg.c2v_w.data_ptr(),
                                    g.num_clauses, X.data_ptr(), F, 0, 0, 0, out.data_ptr(), 3 * F, F, 2 * F, F, 0,
                                    NT.stream_ptr()), "segsum")
    torch.cuda.synchronize()
    got = out.float().cpu().numpy()
    pos = np.zeros((g.num_clauses, F))
    neg = np.zeros((g.num_clauses, F))
    for c in range(g.num_clauses):
        for e in range(g.c2v_rowptr[c], g.c2v_rowptr[c + 1]):
            (pos if g.c2v_sign[e] > 0 else neg)[c] += abs(float(g.c2v_w[e])) * Xv[g.c2v_col[e]]
    bf = lambda a: torch.from_numpy(a.astype(np.float32)).to(torch.bfloat16).float().numpy()
    np.testing.assert_allclose(got[:, F : 2 * F], bf(pos), rtol=1e-2, atol=1e-2)
    np.testing.assert_allclose(got[:, 2 * F :], bf(neg), rtol=1e-2, atol=1e-2)
    # Transpose (backward): fp32 sources from two column blocks added into an fp32 output.
    T = rng.standard_normal((g.num_clauses, 2 * F)).astype(np.float32)
    Tt = torch.from_numpy(T).cuda()
    base = rng.standard_normal((g.num_vars, F)).astype(np.float32)
    acc = torch.from_numpy(base).cuda()
    NT.check(NT.lib().ts_gnn_segsum(dg.v2c_rowptr.data_ptr(), dg.v2c_col.data_ptr(), dg.v2c_w.data_ptr(),
                                    g.num_vars, Tt.data_ptr(), 2 * F, 0, F, 1, acc.data_ptr(), F, 0, 0, F, 1,
                                    NT.stream_ptr()), "segsum")
    torch.cuda.synchronize()
    want = base.astype(np.float64).copy()
    for c in range(g.num_clauses):
        for e in range(g.c2v_rowptr[c], g.c2v_rowptr[c + 1]):
            v = g.c2v_col[e]
            want[v] += abs(float(g.c2v_w[e])) * (T[c, :F] if g.c2v_sign[e] > 0 else T[c, F:])
    np.testing.assert_allclose(acc.cpu().numpy(), want, rtol=1e-5, atol=1e-4)


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("layers,tol", [(1, 3e-3), (2, 3e-3), (6, 3e-2)])
def test_network_forward_and_gradients_match_torch_reference(layers, tol):
    """Forward scores and every parameter block's gradient against the fp32 reference.

    The tolerance is the reference's own conditioning: the network is differentiated
    through bf16 rounding and ReLU kinks, and perturbing the parameters of the 6-layer
    reference by 1e-7 relative already moves its gradient by about 1% (0.05% at 1 and
    2 layers), so the shallow networks carry the tight check and the deep one a looser
    bound with a cosine test.
    """
    import torch

    import gnn_ref
    from tensorsat_amd.learn import graph as G
    from tensorsat_amd.learn.model import GNN, init_params
    from tensorsat_amd.sat import CNFInstance

    cnfs = [CNFInstance.random(k, n, int(r * n), rng=10 + i)
            for i, (k, n, r) in enumerate([(3, 30, 4.0), (4, 25, 9.0), (5, 20, 19.0), (3, 50, 3.5)])]
    g = G.build(cnfs)
    rng = np.random.default_rng(3)
    target = rng.random(g.num_vars).astype(np.float32)
    mask = (rng.random(g.num_vars) < 0.8).astype(np.float32)
    params = init_params(7, layers=layers)
    y_ref, loss_ref, grad_ref = gnn_ref.forward(params, g, target, mask, layers=layers)
    model = GNN(params, layers=layers)
    dg = G.DeviceGraph(g)
    y, loss = model.forward(dg, target=torch.from_numpy(target).cuda(), mask=torch.from_numpy(mask).cuda(),
                            train=True)
    torch.cuda.synchronize()
    y = y.cpu().numpy()
    print(f"layers {layers}: y rel {_rel(y, y_ref):.2e} loss {float(loss.item()):.6f} vs {loss_ref:.6f}", flush=True)
    assert _rel(y, y_ref) < 5e-3
    assert abs(float(loss.item()) - loss_ref) <= 1e-2 * abs(loss_ref) + 1e-6
    grads = model.grads()
    worst = 0.0
    for name, (off, shp) in model.layout.items():
        n = int(np.prod(shp))
        r = _rel(grads[off : off + n], grad_ref[off : off + n])
        worst = max(worst, r)
        assert r < 2 * tol, f"{name}: {r:.3e}"
    total = _rel(grads, grad_ref)
    cos = float(grads @ grad_ref / (np.linalg.norm(grads) * np.linalg.norm(grad_ref)))
    print(f"grad rel {total:.2e} worst block {worst:.2e} cos {cos:.6f}", flush=True)
    assert total < tol and cos > 1 - tol**2
    # Inference path gives the same scores.
    y2, _ = model.forward(dg)
    np.testing.assert_allclose(y2.cpu().numpy(), y, rtol=1e-6, atol=1e-6)


def test_adam_matches_reference_update_and_training_lowers_loss():
    import torch

    from tensorsat_amd.learn import graph as G
    from tensorsat_amd.learn.model import GNN, init_params
    from tensorsat_amd.sat import CNFInstance

    params = init_params(seed=1)
    model = GNN(params)
    g_host = np.random.default_rng(0).standard_normal(model.total).astype(np.float32)
    model.g.copy_(torch.from_numpy(g_host))
    model.adam_step(lr=1e-2)
    torch.cuda.synchronize()
    # First Adam step: m = (1-b1) g, v = (1-b2) g^2, bias-corrected update = lr g / (|g| + eps).
    want = params - 1e-2 * g_host / (np.abs(g_host) + 1e-8)
    np.testing.assert_allclose(model.params(), want, rtol=1e-5, atol=1e-6)
    np.testing.assert_array_equal(model.pbf.float().cpu().numpy(),
                                  torch.from_numpy(model.params()).to(torch.bfloat16).float().numpy())
    # A few steps on one batch lower the loss.
    cnfs = [CNFInstance.random(3, 40, 160, rng=i) for i in range(8)]
    g = G.build(cnfs)
    dg = G.DeviceGraph(g)
    target = torch.from_numpy(np.random.default_rng(2).random(g.num_vars).astype(np.float32)).cuda()
    model = GNN(init_params(seed=2))
    losses = []
    for _ in range(30):
        _, loss = model.forward(dg, target=target, train=True)
        losses.append(float(loss.item()))
        model.adam_step(lr=3e-3)
    print("losses", losses[0], losses[-1], flush=True)
    assert losses[-1] < 0.7 * losses[0]
