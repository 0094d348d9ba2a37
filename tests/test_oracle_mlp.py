"""Pins of the FP64 surrogate-MLP oracle (CPU): forward by brute-force loops,
backward by central finite differences, Adam against torch.optim.Adam, bf16
rounding against torch's conversion."""
import numpy as np
import torch

from oracle import mlp_oracle as mo


def _params(rng, sizes):
    return [(rng.normal(size=(o, i)) / np.sqrt(i), rng.normal(size=o) * 0.1) for i, o in zip(sizes[:-1], sizes[1:])]


def test_forward_matches_brute_force_loops():
    rng = np.random.default_rng(0)
    sizes = [3, 4, 2]
    P = _params(rng, sizes)
    X = rng.normal(size=(2, 3))
    Y, _ = mo.forward(P, X)
    for r in range(2):
        h = list(X[r])
        for l, (W, b) in enumerate(P):
            z = [sum(W[o][i] * h[i] for i in range(len(h))) + b[o] for o in range(len(b))]
            h = [max(v, 0.0) for v in z] if l < len(P) - 1 else z
        assert np.allclose(Y[r], h, rtol=0, atol=1e-14)


def test_backward_matches_central_differences():
    rng = np.random.default_rng(1)
    sizes = [5, 7, 6, 3]
    P = _params(rng, sizes)
    X = rng.normal(size=(4, 5))
    T = rng.normal(size=(4, 3))
    mask = (rng.random((4, 3)) > 0.3).astype(float)

    def loss(params):
        Y, _ = mo.forward(params, X)
        return mo.loss_and_grad(Y, T, mask, 0.7)[0]

    Y, hs = mo.forward(P, X)
    _, gY = mo.loss_and_grad(Y, T, mask, 0.7)
    grads = mo.backward(P, hs, gY)
    h = 1e-6
    for l in range(len(P)):
        for which in (0, 1):
            arr = P[l][which]
            num = np.zeros_like(arr)
            for idx in np.ndindex(arr.shape):
                old = arr[idx]
                arr[idx] = old + h
                lp = loss(P)
                arr[idx] = old - h
                lm = loss(P)
                arr[idx] = old
                num[idx] = (lp - lm) / (2 * h)
            np.testing.assert_allclose(grads[l][which], num, rtol=1e-6, atol=1e-8)


def test_planted_mistake_is_caught():
    """A backward that drops the ReLU mask fails the finite-difference check."""
    rng = np.random.default_rng(2)
    P = _params(rng, [4, 8, 2])
    X = rng.normal(size=(3, 4))
    T = rng.normal(size=(3, 2))
    Y, hs = mo.forward(P, X)
    _, gY = mo.loss_and_grad(Y, T)
    W1, _ = P[1]
    dh = gY @ W1
    wrong = dh.T @ hs[0]  # no mask
    right = mo.backward(P, hs, gY)[0][0]
    assert not np.allclose(wrong, right)


def test_adam_matches_torch():
    rng = np.random.default_rng(3)
    w = rng.normal(size=50)
    p = torch.nn.Parameter(torch.from_numpy(w.copy()))
    opt = torch.optim.Adam([p], lr=1e-2, betas=(0.9, 0.999), eps=1e-8)
    m = np.zeros(50)
    v = np.zeros(50)
    for t in range(1, 8):
        g = rng.normal(size=50)
        p.grad = torch.from_numpy(g.copy())
        opt.step()
        w, m, v = mo.adam(w, g, m, v, t, lr=1e-2)
    np.testing.assert_allclose(w, p.detach().numpy(), rtol=1e-12, atol=1e-14)


def test_bf16_round_matches_torch():
    x = np.concatenate([np.random.default_rng(4).normal(size=1000) * 10.0 ** np.arange(-3, 7).repeat(100),
                        [0.0, -0.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, np.inf, -np.inf]])
    want = torch.from_numpy(x.astype(np.float32)).to(torch.bfloat16).double().numpy()
    assert np.array_equal(mo.bf16_round(x), want)


def _gru_params(rng, I, H, O):
    s = 1 / np.sqrt(H)
    return dict(W_ih=rng.uniform(-s, s, (3 * H, I)), b_ih=rng.uniform(-s, s, 3 * H),
                W_hh=rng.uniform(-s, s, (3 * H, H)), b_hh=rng.uniform(-s, s, 3 * H),
                W_o=rng.normal(size=(O, H)) * s, b_o=rng.normal(size=O) * 0.1)


def _torch_gru(p, X, h0):
    I, H = p["W_ih"].shape[1], p["W_hh"].shape[1]
    gru = torch.nn.GRU(I, H, dtype=torch.float64)
    lin = torch.nn.Linear(H, p["W_o"].shape[0], dtype=torch.float64)
    with torch.no_grad():
        gru.weight_ih_l0.copy_(torch.from_numpy(p["W_ih"]))
        gru.bias_ih_l0.copy_(torch.from_numpy(p["b_ih"]))
        gru.weight_hh_l0.copy_(torch.from_numpy(p["W_hh"]))
        gru.bias_hh_l0.copy_(torch.from_numpy(p["b_hh"]))
        lin.weight.copy_(torch.from_numpy(p["W_o"]))
        lin.bias.copy_(torch.from_numpy(p["b_o"]))
    out, _ = gru(torch.from_numpy(X), torch.from_numpy(h0)[None])
    return lin(out), gru, lin


def test_gru_forward_and_backward_match_torch_gru_fp64():
    rng = np.random.default_rng(9)
    T, B, I, H, O = 5, 3, 4, 6, 2
    p = _gru_params(rng, I, H, O)
    X = rng.normal(size=(T, B, I))
    h0 = rng.normal(size=(B, H)) * 0.5
    Tg = rng.normal(size=(T, B, O))
    Y, cache = mo.gru_forward(p, X, h0)
    Yt, gru, lin = _torch_gru(p, X, h0)
    np.testing.assert_allclose(Y, Yt.detach().numpy(), rtol=1e-12, atol=1e-13)
    loss, gY = mo.loss_and_grad(Y, Tg, None, 0.3)
    g = mo.gru_backward(p, cache, gY)
    (0.5 * 0.3 * ((Yt - torch.from_numpy(Tg)) ** 2).sum()).backward()
    pairs = dict(W_ih=gru.weight_ih_l0, b_ih=gru.bias_ih_l0, W_hh=gru.weight_hh_l0, b_hh=gru.bias_hh_l0,
                 W_o=lin.weight, b_o=lin.bias)
    for k, t in pairs.items():
        np.testing.assert_allclose(g[k], t.grad.numpy(), rtol=1e-10, atol=1e-12, err_msg=k)


def test_gru_backward_central_differences():
    rng = np.random.default_rng(10)
    T, B, I, H, O = 4, 2, 3, 5, 2
    p = _gru_params(rng, I, H, O)
    X = rng.normal(size=(T, B, I))
    Tg = rng.normal(size=(T, B, O))

    def loss():
        return mo.loss_and_grad(mo.gru_forward(p, X)[0], Tg)[0]

    Y, cache = mo.gru_forward(p, X)
    g = mo.gru_backward(p, cache, mo.loss_and_grad(Y, Tg)[1])
    for k in ("W_hh", "b_ih", "W_o"):
        arr = p[k]
        for idx in list(np.ndindex(arr.shape))[:12]:
            old = arr[idx]
            arr[idx] = old + 1e-6
            lp = loss()
            arr[idx] = old - 1e-6
            lm = loss()
            arr[idx] = old
            assert abs((lp - lm) / 2e-6 - g[k][idx]) <= 1e-6 * max(1.0, abs(g[k][idx])), (k, idx)
