"""GPU checks of the learned-start CNN kernels against the float64 CPU reference.

Tolerances: the GEMM is exact up to FP32 accumulation on BF16 operands (1e-5 relative);
forward and gradient carry BF16 rounding of hidden activations and of dZ (2e-2 relative
to the output / gradient scale of each layer).
"""

import ctypes

import numpy as np
import pytest
import torch

from imageprocessor_new_amd import synthetic
from oracle import cnn_ref

pytestmark = pytest.mark.gpu


def _dev():
    from imageprocessor_new_amd.backends import gpu_device

    return gpu_device(0)


def _bits_tensor(a, dev):
    return torch.from_numpy(cnn_ref.bf16_bits(a).view(np.int16)).to(dev)


@pytest.mark.parametrize("M,N,K,transA,transB", [(70, 45, 77, False, False), (64, 64, 32, True, False),
                                                 (130, 7, 300, False, True), (5, 129, 1000, True, True)])
def test_gemm_bf16_lane_layout(M, N, K, transA, transB):
    from imageprocessor_new_amd import _lib

    rng = np.random.default_rng(M + N + K)
    A = cnn_ref.bf16_round(rng.normal(size=(M, K)))
    B = cnn_ref.bf16_round(rng.normal(size=(K, N)))
    with _dev() as dev:
        Ad = A.T.copy() if transA else A
        Bd = B.T.copy() if transB else B
        At, Bt = _bits_tensor(Ad, dev), _bits_tensor(Bd, dev)
        sam, sak = (1, M) if transA else (K, 1)
        sbk, sbn = (1, K) if transB else (N, 1)
        C = torch.empty((M, N), dtype=torch.float32, device=dev)
        _lib.call("spx_gemm_bf16", At.data_ptr(), sam, sak, Bt.data_ptr(), sbk, sbn, C.data_ptr(), M, N, K,
                  int(torch.cuda.current_stream().cuda_stream))
        got = C.cpu().numpy()
    ref = A.astype(np.float64) @ B.astype(np.float64)
    np.testing.assert_allclose(got, ref, rtol=1e-5, atol=1e-4 * np.sqrt(K))


def test_prepare_matches_reference():
    from imageprocessor_new_amd.cnn import LearnedStart

    arr = synthetic.make_workload("hex16384", seed=31, H=300, W=300)
    seeds = np.concatenate([arr.centers_xy[:20], [[2.0, 3.0], [297.0, 150.0]]]).astype(np.float32)
    half = np.array([5, 7, 15, 20, 31, 12, 3, 9, 10, 11, 13, 14, 16, 17, 18, 19, 21, 22, 25, 30, 15, 15], np.int16)
    with _dev() as dev:
        net = LearnedStart(device=dev)
        x, norm = net.prepare(torch.from_numpy(arr.image).to(dev), torch.from_numpy(seeds).to(dev),
                              torch.from_numpy(half).to(dev))
        xg = (x.cpu().numpy().astype(np.int32).view(np.uint32) & 0xFFFF).astype(np.uint32) << 16
        xg = xg.view(np.float32).reshape(-1, 32, 32)
        ng = norm.cpu().numpy()
    for k in range(len(seeds)):
        xr, nr = cnn_ref.prepare(arr.image, seeds[k], half[k])
        np.testing.assert_allclose(ng[k], nr, rtol=1e-5)
        np.testing.assert_allclose(xg[k], cnn_ref.bf16_round(xr), rtol=1e-2, atol=1e-2)


def _theta(seed):
    from imageprocessor_new_amd.cnn import init_params

    th = init_params(seed)
    rng = np.random.default_rng(seed)
    # non-zero biases exercise the bias paths
    o = 0
    for k, n in cnn_ref.LAYERS:
        o += k * n
        th[o:o + n] = rng.normal(0, 0.05, size=n)
        o += n
    return th


def test_forward_matches_reference():
    from imageprocessor_new_amd.cnn import LearnedStart

    rng = np.random.default_rng(4)
    theta = _theta(4)
    x = cnn_ref.bf16_round(rng.normal(0, 1, size=(37, 32, 32)))
    with _dev() as dev:
        net = LearnedStart(theta, device=dev)
        got = net.forward_raw(_bits_tensor(x, dev)).cpu().numpy()
    th_bf = cnn_ref.bf16_round(theta).astype(np.float64)
    # biases stay FP32 on the GPU
    o = 0
    for k, n in cnn_ref.LAYERS:
        o += k * n
        th_bf[o:o + n] = theta[o:o + n]
        o += n
    ref = cnn_ref.forward(torch.tensor(th_bf), torch.tensor(x.astype(np.float64)), round_acts=True).numpy()
    scale = np.abs(ref).max()
    assert np.abs(got - ref).max() <= 2e-2 * scale


def test_gradient_matches_reference():
    from imageprocessor_new_amd.cnn import LearnedStart

    rng = np.random.default_rng(5)
    theta = _theta(5)
    B = 64
    x = cnn_ref.bf16_round(rng.normal(0, 1, size=(B, 32, 32)))
    t = rng.normal(0, 1, size=(B, 7)).astype(np.float32)
    with _dev() as dev:
        net = LearnedStart(theta, device=dev)
        g, loss = net.grad(_bits_tensor(x, dev), torch.from_numpy(t).to(dev))
        g = g.cpu().numpy()
    th_ref = cnn_ref.bf16_round(theta).astype(np.float64)
    o = 0
    for k, n in cnn_ref.LAYERS:
        o += k * n
        th_ref[o:o + n] = theta[o:o + n]
        o += n
    ref_loss, ref_g = cnn_ref.loss_and_grad(th_ref, x, t)
    assert abs(loss - ref_loss) <= 2e-2 * abs(ref_loss)
    # BF16 activations in the forward pass and BF16 dZ in every backward GEMM make the
    # gradient an approximation of the float64 one: each block must point the same way
    # (cosine >= 0.99) and agree in norm to 15 %.
    o = 0
    report = []
    for k, n in cnn_ref.LAYERS:
        for lo, hi in ((o, o + k * n), (o + k * n, o + k * n + n)):
            gs, rs = g[lo:hi].astype(np.float64), ref_g[lo:hi]
            err = np.linalg.norm(gs - rs) / max(np.linalg.norm(rs), 1e-12)
            cos = float(gs @ rs / max(np.linalg.norm(gs) * np.linalg.norm(rs), 1e-300))
            report.append((k, n, hi - lo, round(err, 4), round(cos, 5)))
        o += k * n + n
    print("gradient blocks (K, N, size, rel err, cosine):", report)
    for k, n, size, err, cos in report:
        assert cos >= 0.99 and err <= 0.15, report


def test_training_reduces_loss():
    from imageprocessor_new_amd.cnn import LearnedStart

    rng = np.random.default_rng(6)
    n = 4096
    x = rng.normal(0, 1, size=(n, 32, 32)).astype(np.float32)
    # a learnable target: fixed random linear functions of the patch means of four quadrants
    q = np.stack([x[:, :16, :16].mean((1, 2)), x[:, :16, 16:].mean((1, 2)), x[:, 16:, :16].mean((1, 2)),
                  x[:, 16:, 16:].mean((1, 2))], axis=1)
    Wt = rng.normal(0, 1, size=(4, 7))
    t = (q @ Wt * 10).astype(np.float32)
    with _dev() as dev:
        net = LearnedStart(device=dev, seed=1)
        xb = _bits_tensor(x, dev)
        tt = torch.from_numpy(t).to(dev)
        _, l0 = net.grad(xb, tt)
        log = net.train(xb, tt, epochs=15, batch=512, lr=3e-3)
        _, l1 = net.grad(xb, tt)
    assert l1 < 0.5 * l0, (l0, l1, log.losses)


def test_fused_proposal_equals_prepare_then_forward():
    from imageprocessor_new_amd.cnn import LearnedStart

    arr = synthetic.make_workload("hex16384", seed=32, H=300, W=300)
    seeds = np.concatenate([arr.centers_xy[:35], [[2.0, 3.0], [297.0, 150.0]]]).astype(np.float32)  # K = 37
    half = np.resize(np.array([5, 7, 15, 20, 31, 12, 3, 9], np.int16), len(seeds))
    with _dev() as dev:
        net = LearnedStart(_theta(6), device=dev)
        args = (torch.from_numpy(arr.image).to(dev), torch.from_numpy(seeds).to(dev), torch.from_numpy(half).to(dev))
        fused = net.propose(*args).cpu().numpy()
        two = net.propose_two_step(*args).cpu().numpy()
    assert np.isfinite(fused).all()
    np.testing.assert_array_equal(fused, two)
