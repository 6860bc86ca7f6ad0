"""Learned initial guess: the GPU backward pass of the BF16 network against a
numpy restatement with the same rounding points, and what the trained network
gains over the guesses without it on an unseen water cluster."""
import numpy as np
import pytest

import quiqbox_amd as qb
from quiqbox_amd import learned_guess as lg
from quiqbox_amd.systems import water_box

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _say(msg):
    print("[gpu-test] " + msg, flush=True)


def bf16(x):
    """round-to-nearest-even to BF16, returned as float64"""
    u = np.ascontiguousarray(np.asarray(x, dtype=np.float32)).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def test_backward_matches_numpy_restatement():
    model = lg.LearnedGuess(seed=3)
    rng = np.random.default_rng(4)
    n = 37
    X = rng.normal(size=(n, lg.N_IN)).astype(np.float32)
    T = rng.normal(size=(n, lg.N_OUT)).astype(np.float32)
    xb = torch.tensor(X, device="cuda").to(torch.bfloat16)
    yt = torch.tensor(T, device="cuda")
    rw = torch.ones(n, dtype=torch.float32, device="cuda")
    y, acts = model._forward(xb, n, keep=True)
    dY = torch.empty_like(y)
    dYb = torch.empty((n, lg.N_OUT), dtype=torch.bfloat16, device="cuda")
    loss = torch.zeros(1, device="cuda")
    cm = torch.ones(lg.N_OUT, device="cuda")
    from quiqbox_amd import _lib
    import ctypes
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    _lib.check(_lib.load().qbx_mlp_mse(n, lg.N_OUT, p(y), p(yt), p(rw), p(cm), 1.0 / n, p(dY), p(dYb), p(loss), None))
    grads = model.backward(acts, dY, dYb, n)
    torch.cuda.synchronize()

    W = [bf16(w.float().cpu().numpy()) for w in model.dev.Wb]
    b = [bb.cpu().numpy().astype(np.float64) for bb in model.dev.b]
    h0 = bf16(X)
    z1 = h0 @ W[0] + b[0]
    h1 = bf16(np.maximum(z1, 0))
    z2 = h1 @ W[1] + b[1]
    h2 = bf16(np.maximum(z2, 0))
    yref = h2 @ W[2] + b[2]
    np.testing.assert_allclose(y.cpu().numpy(), yref, rtol=1e-4, atol=1e-4)
    g3 = 2.0 * (y.cpu().numpy().astype(np.float64) - T) / n
    assert loss.item() == pytest.approx(float(np.sum((y.cpu().numpy() - T) ** 2) / n), rel=1e-4)
    g3b = bf16(g3)
    dW3, db3 = h2.T @ g3b, g3.sum(0)
    d2 = g3b @ W[2].T
    d2 = np.where(h2 > 0, d2, 0.0)
    d2b = bf16(d2)
    dW2, db2 = h1.T @ d2b, d2.sum(0)
    d1 = np.where(h1 > 0, d2b @ W[1].T, 0.0)
    d1b = bf16(d1)
    dW1, db1 = h0.T @ d1b, d1.sum(0)
    for (gW, gb), (rW, rb), name in zip(grads, [(dW1, db1), (dW2, db2), (dW3, db3)], ("1", "2", "3")):
        gw = gW.cpu().numpy()
        scale = np.abs(rW).max()
        _say("layer %s: max |dW gpu - numpy| %.2e of %.2e" % (name, np.abs(gw - rW).max(), scale))
        np.testing.assert_allclose(gw, rW, atol=2e-3 * scale, rtol=0)
        np.testing.assert_allclose(gb.cpu().numpy(), rb, atol=2e-3 * np.abs(rb).max() + 1e-6, rtol=0)


def _loss(model, xb, yt, n, scale):
    import ctypes
    from quiqbox_amd import _lib
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    y = model._forward(xb, n)
    no = model.layout.n_out
    dY = torch.empty_like(y)
    dYb = torch.empty((n, no), dtype=torch.bfloat16, device="cuda")
    loss = torch.zeros(1, device="cuda")
    rw = torch.ones(n, dtype=torch.float32, device="cuda")
    cm = torch.ones(no, device="cuda")
    _lib.check(_lib.load().qbx_mlp_mse(n, no, p(y), p(yt), p(rw), p(cm), scale, p(dY), p(dYb), p(loss), None))
    torch.cuda.synchronize()
    return float(loss.item()), y, dY, dYb


def test_backward_against_central_differences():
    """Each layer's weight and bias gradients from the GPU backward pass against
    central differences of the GPU loss along the gradient itself:
    L(W + h V) - L(W - h V) against g . (W+ - W-), V = g / max|g|, W+- the
    BF16 weights the forward pass actually uses.  Along the gradient the
    directional derivative has no cancellation, so the differences track it
    closely: the last layer's loss is quadratic in its weights and biases
    (exact up to FP32 rounding, 2e-3); the earlier layers pass through the BF16
    rounding of the hidden activations and the ReLU kinks (5%; the same check
    in FP64 without rounding agrees to 0.6% at this step)."""
    model = lg.LearnedGuess(seed=5, layout=lg.CHO_SP)
    rng = np.random.default_rng(6)
    n = 64
    lay = model.layout
    X = rng.normal(size=(n, lay.n_in)).astype(np.float32)
    T = rng.normal(size=(n, lay.n_out)).astype(np.float32)
    xb = torch.tensor(X, device="cuda").to(torch.bfloat16)
    yt = torch.tensor(T, device="cuda")
    scale = 1.0 / n
    _, y, dY, dYb = _loss(model, xb, yt, n, scale)
    _, acts = model._forward(xb, n, keep=True)
    grads = model.backward(acts, dY, dYb, n)
    torch.cuda.synchronize()
    d = model.dev
    for li in range(3):
        gW = grads[li][0].cpu().numpy().astype(np.float64)
        gb = grads[li][1].cpu().numpy().astype(np.float64)
        W0 = d.W[li].clone()
        Wb0 = d.Wb[li].clone()
        b0 = d.b[li].clone()
        tol = 2e-3 if li == 2 else 5e-2
        V = torch.tensor((gW / np.abs(gW).max()).astype(np.float32), device="cuda")
        h = 0.02 * float(W0.abs().mean())
        Wp = (W0 + h * V).to(torch.bfloat16)
        Wm = (W0 - h * V).to(torch.bfloat16)
        d.Wb[li].copy_(Wp)
        Lp = _loss(model, xb, yt, n, scale)[0]
        d.Wb[li].copy_(Wm)
        Lm = _loss(model, xb, yt, n, scale)[0]
        d.Wb[li].copy_(Wb0)
        pred = float(np.sum(gW * (Wp.float() - Wm.float()).cpu().numpy().astype(np.float64)))
        _say("layer %d weights: L+ - L- %.6e, g . dW %.6e (rel %.2e)" % (li + 1, Lp - Lm, pred,
                                                                       (Lp - Lm - pred) / pred))
        assert abs((Lp - Lm) - pred) <= tol * abs(pred)
        vb = torch.tensor((gb / np.abs(gb).max()).astype(np.float32), device="cuda")
        # biases start at zero: a step of 0.02 moves the hidden pre-activations
        # (|z| ~ 0.5) by several BF16 quanta; at 0.002 the rounding of the
        # activations alone puts the first layer's difference 12% off
        hb = 0.02 * max(float(b0.abs().mean()), 1.0)
        d.b[li].copy_(b0 + hb * vb)
        Lp = _loss(model, xb, yt, n, scale)[0]
        d.b[li].copy_(b0 - hb * vb)
        Lm = _loss(model, xb, yt, n, scale)[0]
        d.b[li].copy_(b0)
        pred = float(np.sum(gb * ((b0 + hb * vb) - (b0 - hb * vb)).cpu().numpy().astype(np.float64)))
        _say("layer %d biases: L+ - L- %.6e, g . db %.6e (rel %.2e)" % (li + 1, Lp - Lm, pred, (Lp - Lm - pred) / pred))
        assert abs((Lp - Lm) - pred) <= tol * abs(pred)


def test_trained_network_beats_the_core_guess_on_an_unseen_cluster():
    systems = [water_box(n, 100 + n * 7 + k, basis="cc-pVDZ") for n in (2, 3, 4) for k in range(4)]
    data = lg.training_set(systems, log=_say)
    assert len(data) >= 10
    model = lg.LearnedGuess(seed=1)
    hist = model.train(data, epochs=150, batch=1024, lr=1e-3, log=_say)
    assert hist[-1] < 0.2 * hist[0]
    nuc, bs = water_box(5, 999, basis="cc-pVDZ")
    gb = qb.GPUBasis(bs)
    strat = qb.SCFconfig(("ADIIS", "DIIS"), (1e-4, 1e-9))
    ref = qb.runHartreeFock(nuc, gb, qb.HFconfig(initial="CoreH", strategy=strat, maxStep=150), printInfo=False)
    assert ref.converged
    D_conv = 2.0 * ref.density[0]
    D_pred = model.predict_density(nuc, gb)
    assert abs(np.sum(D_pred * gb.overlap()) - 50.0) < 5.0
    from quiqbox_amd.hartree_fock import getC
    S = gb.overlap()
    H = gb.core_hamiltonian(nuc.charges, nuc.coord_array)
    C = getC(qb.getOrthonormalization(S), H)
    nocc = int(round(sum(nuc.charges))) // 2
    D_core = 2.0 * C[:, :nocc] @ C[:, :nocc].T
    err_pred = np.linalg.norm(D_pred - D_conv)
    err_core = np.linalg.norm(D_core - D_conv)
    learned = qb.runHartreeFock(nuc, gb, lg.guess_config(model, nuc, gb, strat, 150), printInfo=False)
    _say("unseen 5-water cluster: |D_pred - D| %.3f, |D_core - D| %.3f; iterations learned %d core %d" % (
        err_pred, err_core, len(learned.memory[0].shared.Etots), len(ref.memory[0].shared.Etots)))
    assert learned.converged
    assert abs(learned.energy[0] - ref.energy[0]) < 1e-7
    # the network's density is much closer to the converged one than the core guess's
    assert err_pred < 0.5 * err_core
    assert len(learned.memory[0].shared.Etots) <= len(ref.memory[0].shared.Etots)
