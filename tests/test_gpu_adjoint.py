"""The adjoint kernels: the forward rollout and the compliance gradient of several uneven blocks in one launch equal,
bit for bit, the CPU oracle's FP32 adjoint of each block (checkpoint segments that do not divide the rollout
included); the torch.autograd path gives the same gradient; and the FP32 gradient agrees with FP64 central finite
differences on random tets of a block soft enough for FP32 to resolve its deformation."""
import ctypes as ct

import numpy as np
import pytest

from genesis_tissue_simulation_amd.workloads import adjoint_instance

pytestmark = pytest.mark.gpu

SIZES = [(7, 4, 4), (9, 5, 4), (5, 5, 6)]


def _bind(L):
    for suf in ("f64", "f32"):
        f = getattr(L, f"oracle_adjoint_{suf}")
        f.argtypes = [ct.c_void_p, ct.c_float, ct.c_int, ct.c_int] + [ct.c_void_p] * 5
        f.restype = ct.c_double
    return L


def P(a, dtype=np.float32):
    return np.ascontiguousarray(a, dtype=dtype).ctypes.data_as(ct.c_void_p)


def blocks(scale, seed=0, iterations=1):
    out = []
    for k, nodes in enumerate(SIZES):
        inst = adjoint_instance(nodes=nodes, seed=seed + k, substeps=10, iterations=iterations)
        inst.compliance_hidden = inst.compliance_hidden * scale
        inst.compliance_init = inst.compliance_init * scale
        out.append(inst)
    return out


def oracle_adjoint(L, Oracle, inst, nsub, c, target, double=False):
    """(loss, grad in generation order, end positions) of the oracle's adjoint for one block."""
    o = Oracle(inst.grid, inst.params, inst.initial_positions, double=double)
    grad = np.zeros(len(c), dtype=np.float64 if double else np.float32)
    f = L.oracle_adjoint_f64 if double else L.oracle_adjoint_f32
    loss = f(o.h, np.float32(inst.dt), nsub, inst.params.iterations, P(inst.force), P(c), P(inst.mask), P(target),
             grad.ctypes.data_as(ct.c_void_p))
    return loss, grad, o.state()[0]


# The blocks have 112, 180 and 150 nodes. lds_nodes: None stages every block in LDS in both passes; 0 runs every block
# from global memory; 160 stages the first and last block in the forward and none in the backward (which needs 2 N);
# 300 stages all three in the forward and the first and last in the backward.
@pytest.mark.parametrize("scale,nsub,ckpt,lds_nodes,iterations", [
    (1e4, 20, 7, None, 1), (1.0, 12, 4, None, 1), (1e4, 20, 7, 0, 1), (1e4, 9, 4, 160, 1), (1e4, 9, 4, 300, 1),
    (1e4, 20, 7, None, 2), (1e4, 9, 4, 0, 2), (1e4, 6, 6, 300, 3)])
def test_adjoint_matches_oracle_fp32(scale, nsub, ckpt, lds_nodes, iterations):
    import torch

    from genesis_tissue_simulation_amd.adjoint import ComplianceFit
    from oracle import oracle as orc

    L = _bind(orc.lib())
    insts = blocks(scale, iterations=iterations)
    fit = ComplianceFit(insts, nsub=nsub, ckpt=ckpt, block_size=256, lds_nodes=lds_nodes)
    b = fit.batch
    hidden = np.concatenate([i.compliance_hidden for i in insts]).astype(np.float32)
    init = np.concatenate([i.compliance_init for i in insts]).astype(np.float32)
    print(f"adjoint parity: blocks {SIZES} scale {scale} nsub {nsub} ckpt {ckpt} lds_nodes {lds_nodes} "
          f"iterations {iterations}", flush=True)
    target = fit.forward_raw(torch.as_tensor(hidden)).cpu().numpy().copy()
    x = fit.forward_raw(torch.as_tensor(init)).cpu().numpy().copy()
    mask = fit.mask.cpu().numpy()
    xbar = (x - target) * (np.float32(2) * mask)[:, None]
    grad = fit.backward_raw(torch.as_tensor(xbar).cuda()).cpu().numpy()
    torch.cuda.synchronize()
    for k, inst in enumerate(insts):
        n0, n1 = int(b.node_off[k]), int(b.node_off[k + 1])
        t0, t1 = int(b.tet_off[k]), int(b.tet_off[k + 1])
        c_k = init[t0:t1]
        # the oracle's target from the hidden compliance, then its adjoint at the initial compliance
        o = orc.Oracle(inst.grid, inst.params, inst.initial_positions)
        L.oracle_adjoint_f32(o.h, np.float32(inst.dt), nsub, iterations, P(inst.force), P(hidden[t0:t1]),
                             P(inst.mask), P(np.zeros((n1 - n0, 3))), None)
        tgt_o = o.state()[0]
        np.testing.assert_array_equal(target[n0:n1], tgt_o)
        loss_o, grad_o, x_o = oracle_adjoint(L, orc.Oracle, inst, nsub, c_k, tgt_o)
        np.testing.assert_array_equal(x[n0:n1], x_o)
        assert np.isfinite(grad_o).all() and np.abs(grad_o).max() > 0
        np.testing.assert_array_equal(grad[t0:t1], grad_o)


def test_autograd_function_matches_raw_backward():
    import torch

    from genesis_tissue_simulation_amd.adjoint import ComplianceFit

    insts = blocks(1e4, seed=7)
    fit = ComplianceFit(insts, nsub=10, ckpt=3, block_size=512)
    hidden = torch.as_tensor(np.concatenate([i.compliance_hidden for i in insts]).astype(np.float32))
    init = np.concatenate([i.compliance_init for i in insts]).astype(np.float32)
    target = fit.forward_raw(hidden).clone()
    c = torch.tensor(init, device="cuda", requires_grad=True)
    x = fit.rollout(c)
    loss = fit.loss(x, target)
    loss.backward()
    x2 = fit.forward_raw(torch.as_tensor(init)).clone()
    d = x2 - target
    g2 = fit.backward_raw(d * (2 * fit.mask[:, None]))
    assert bool(torch.equal(x.detach(), x2))
    torch.testing.assert_close(c.grad, g2, rtol=1e-6, atol=0)
    # a second rollout before backward invalidates the first one's checkpoints
    y = fit.rollout(c)
    fit.forward_raw(torch.as_tensor(init))
    with pytest.raises(RuntimeError):
        fit.loss(y, target).backward()


def test_fp32_gradient_against_fp64_finite_differences():
    """32 random tets of a soft block (compliance x1e4, strains ~1e-2), 20 substeps; the FP64 differences use 0.1%
    steps. Measured on the CPU oracle (whose FP32 adjoint the kernels reproduce bit for bit): cosine 0.998, largest
    component error 7.5% of the largest gradient; the bounds leave room for the seed."""
    import torch

    from genesis_tissue_simulation_amd.adjoint import ComplianceFit
    from oracle import oracle as orc

    L = _bind(orc.lib())
    inst = adjoint_instance(nodes=(7, 4, 4), seed=3, substeps=10)
    nsub, scale = 20, 1e4
    N = len(inst.grid.positions)
    hidden = (inst.compliance_hidden * scale).astype(np.float32)
    c = (inst.compliance_init * scale).astype(np.float32)
    oh = orc.Oracle(inst.grid, inst.params, inst.initial_positions, double=True)
    L.oracle_adjoint_f64(oh.h, np.float32(inst.dt), nsub, 1, P(inst.force), P(hidden), P(inst.mask),
                         P(np.zeros((N, 3))), None)
    target = oh.state()[0].astype(np.float32)
    fit = ComplianceFit([inst], nsub=nsub, ckpt=10)
    x = fit.forward_raw(torch.as_tensor(c)).cpu().numpy().copy()
    grad = fit.backward_raw(torch.as_tensor((x - target) * (2 * inst.mask)[:, None]).cuda()).cpu().numpy()

    def loss64(cc):
        o = orc.Oracle(inst.grid, inst.params, inst.initial_positions, double=True)
        return L.oracle_adjoint_f64(o.h, np.float32(inst.dt), nsub, 1, P(inst.force), P(cc), P(inst.mask),
                                    P(target), None)

    idx = np.random.default_rng(0).choice(len(c), 32, replace=False)
    fd = []
    for t in idx:
        cp, cm = c.copy(), c.copy()
        cp[t], cm[t] = np.float32(c[t] * 1.001), np.float32(c[t] * 0.999)
        fd.append((loss64(cp) - loss64(cm)) / (float(cp[t]) - float(cm[t])))
    fd = np.array(fd)
    g = grad[idx].astype(np.float64)
    cos = float(g @ fd / (np.linalg.norm(g) * np.linalg.norm(fd)))
    rel = float(np.linalg.norm(g - fd) / np.linalg.norm(fd))
    print(f"FP32 adjoint vs FP64 FD on 32 tets: cosine {cos:.5f}, relative L2 {rel:.3e}", flush=True)
    assert cos >= 0.99 and rel <= 0.15
    assert float(np.abs(g - fd).max()) <= 0.15 * float(np.abs(fd).max())


def test_full_size_block_at_the_stated_compliance():
    """One block of the compliance fit at full size and its stated compliance (20736 tets, 50 substeps, checkpoints
    every 10, median compliance 1e-7): the rollout and the gradient are bit-exact with the oracle's FP32 adjoint, and
    the gradient is finite everywhere."""
    import torch

    from genesis_tissue_simulation_amd.adjoint import ComplianceFit
    from oracle import oracle as orc

    L = _bind(orc.lib())
    inst = adjoint_instance(tets=20_000, seed=0)
    T = len(inst.compliance_init)
    print(f"full-size adjoint block: {T} tets, 50 substeps, checkpoints every 10", flush=True)
    fit = ComplianceFit([inst], nsub=50, ckpt=10)
    hidden = inst.compliance_hidden.astype(np.float32)
    init = inst.compliance_init.astype(np.float32)
    target = fit.forward_raw(torch.as_tensor(hidden)).cpu().numpy().copy()
    x = fit.forward_raw(torch.as_tensor(init)).cpu().numpy().copy()
    mask = fit.mask.cpu().numpy()
    grad = fit.backward_raw(torch.as_tensor((x - target) * (np.float32(2) * mask)[:, None]).cuda()).cpu().numpy()
    torch.cuda.synchronize()
    o = orc.Oracle(inst.grid, inst.params, inst.initial_positions)
    N = len(inst.grid.positions)
    L.oracle_adjoint_f32(o.h, np.float32(inst.dt), 50, 1, P(inst.force), P(hidden), P(inst.mask),
                         P(np.zeros((N, 3))), None)
    np.testing.assert_array_equal(target, o.state()[0])
    loss_o, grad_o, x_o = oracle_adjoint(L, orc.Oracle, inst, 50, init, target)
    np.testing.assert_array_equal(x, x_o)
    assert np.isfinite(grad).all() and np.abs(grad).max() > 0
    np.testing.assert_array_equal(grad, grad_o)
