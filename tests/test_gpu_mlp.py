"""BF16 matrix-core kernels of the learned guess (qbx_gemm_bf16 and the
elementwise/loss/Adam kernels) against numpy on the same BF16-rounded inputs.
Products of two BF16 values are exact in FP32, so the GEMM differs from the
FP64 reference only by FP32 summation rounding."""
import ctypes

import numpy as np
import pytest

from quiqbox_amd import _lib

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _bf16(a):
    """numpy float array -> (torch bf16 cuda tensor, float64 values it holds)"""
    t = torch.tensor(np.asarray(a, dtype=np.float32)).to(torch.bfloat16)
    return t.cuda(), t.float().numpy().astype(np.float64)


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_gemm_all_transposes_ragged_sizes(ta, tb):
    rng = np.random.default_rng(ta * 2 + tb)
    M, N, K = 70, 45, 37
    A = rng.normal(size=(K, M) if ta else (M, K))
    B = rng.normal(size=(N, K) if tb else (K, N))
    dA, hA = _bf16(A)
    dB, hB = _bf16(B)
    C0 = rng.normal(size=(M, N)).astype(np.float32)
    dC = torch.tensor(C0).cuda()
    L = _lib.load()
    print("[gpu-test] gemm ta=%d tb=%d %dx%dx%d" % (ta, tb, M, N, K), flush=True)
    _lib.check(L.qbx_gemm_bf16(ta, tb, M, N, K, 1.5, _p(dA), dA.shape[1], _p(dB), dB.shape[1], 0.5, _p(dC), N, 0,
                               None, None, None))
    torch.cuda.synchronize()
    opA = hA.T if ta else hA
    opB = hB.T if tb else hB
    ref = 1.5 * opA @ opB + 0.5 * C0.astype(np.float64)
    np.testing.assert_allclose(dC.cpu().numpy(), ref, rtol=2e-5, atol=2e-5)


def test_gemm_bias_epilogues():
    rng = np.random.default_rng(5)
    M, N, K = 130, 96, 64
    dA, hA = _bf16(rng.normal(size=(M, K)))
    dB, hB = _bf16(rng.normal(size=(K, N)))
    bias = rng.normal(size=N).astype(np.float32)
    dbias = torch.tensor(bias).cuda()
    dC = torch.zeros((M, N), dtype=torch.float32, device="cuda")
    dCb = torch.zeros((M, N), dtype=torch.bfloat16, device="cuda")
    L = _lib.load()
    _lib.check(L.qbx_gemm_bf16(0, 0, M, N, K, 1.0, _p(dA), K, _p(dB), N, 0.0, _p(dC), N, 1, _p(dbias), _p(dCb), None))
    torch.cuda.synchronize()
    ref = np.maximum(hA @ hB + bias, 0.0)
    np.testing.assert_allclose(dC.cpu().numpy(), ref, rtol=2e-5, atol=2e-5)
    np.testing.assert_allclose(dCb.float().cpu().numpy(), ref, rtol=1e-2, atol=1e-2)
    _lib.check(L.qbx_gemm_bf16(0, 0, M, N, K, 1.0, _p(dA), K, _p(dB), N, 0.0, _p(dC), N, 2, _p(dbias), None, None))
    torch.cuda.synchronize()
    np.testing.assert_allclose(dC.cpu().numpy(), hA @ hB + bias, rtol=2e-5, atol=2e-5)
    with pytest.raises(ValueError):
        _lib.check(L.qbx_gemm_bf16(0, 0, M, N, K, 1.0, _p(dA), K - 1, _p(dB), N, 0.0, _p(dC), N, 0, None, None, None))


def test_elementwise_loss_colsum_adam():
    rng = np.random.default_rng(9)
    L = _lib.load()
    x = rng.normal(size=(40, 24)).astype(np.float32)
    dx = torch.tensor(x).cuda()
    out = torch.zeros((40, 24), dtype=torch.bfloat16, device="cuda")
    _lib.check(L.qbx_mlp_elementwise(0, x.size, _p(dx), None, _p(out), None, None))
    H = torch.tensor(rng.normal(size=(40, 24)).astype(np.float32)).to(torch.bfloat16).cuda()
    dZb = torch.zeros_like(out)
    dZ = torch.zeros_like(dx)
    _lib.check(L.qbx_mlp_elementwise(1, x.size, _p(dx), _p(H), _p(dZb), _p(dZ), None))
    T = rng.normal(size=(40, 24)).astype(np.float32)
    rw = rng.uniform(0, 1, size=40).astype(np.float32)
    cm = (rng.uniform(size=24) > 0.3).astype(np.float32)
    dT, drw, dcm = (torch.tensor(v).cuda() for v in (T, rw, cm))
    dY = torch.zeros_like(dx)
    dYb = torch.zeros_like(out)
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    _lib.check(L.qbx_mlp_mse(40, 24, _p(dx), _p(dT), _p(drw), _p(dcm), 0.25, _p(dY), _p(dYb), _p(loss), None))
    cs = torch.zeros(24, dtype=torch.float32, device="cuda")
    _lib.check(L.qbx_mlp_colsum(40, 24, _p(dx), _p(cs), None))
    w = rng.normal(size=100).astype(np.float32)
    g = rng.normal(size=100).astype(np.float32)
    dw, dg = torch.tensor(w).cuda(), torch.tensor(g).cuda()
    dm, dv = torch.zeros(100, device="cuda"), torch.zeros(100, device="cuda")
    dwb = torch.zeros(100, dtype=torch.bfloat16, device="cuda")
    _lib.check(L.qbx_mlp_adam(100, _p(dw), _p(dm), _p(dv), _p(dg), _p(dwb), 1e-2, 0.9, 0.999, 1e-8, 1, None))
    torch.cuda.synchronize()
    np.testing.assert_allclose(out.float().cpu().numpy(), torch.tensor(x).to(torch.bfloat16).float().numpy())
    Hf = H.float().cpu().numpy()
    np.testing.assert_allclose(dZ.cpu().numpy(), np.where(Hf > 0, x, 0.0))
    w2 = (rw[:, None] * cm[None, :]).astype(np.float64)
    np.testing.assert_allclose(dY.cpu().numpy(), 2 * w2 * (x - T) * 0.25, rtol=1e-6, atol=1e-7)
    assert loss.item() == pytest.approx(float(np.sum(w2 * (x - T) ** 2) * 0.25), rel=1e-5)
    np.testing.assert_allclose(cs.cpu().numpy(), x.sum(0), rtol=1e-5, atol=1e-5)
    # first Adam step: m = 0.1 g, v = 0.001 g^2, bias-corrected step lr * sign(g) (up to eps)
    np.testing.assert_allclose(dw.cpu().numpy(), w - 1e-2 * g / (np.abs(g) + 1e-8), rtol=1e-5, atol=1e-6)
