"""GPU tests of the surrogate's kernels through the C-ABI: the bf16 MFMA GEMM
(bit-exact on integer-valued operands, where FP32 accumulation is exact in
any order, so the check pins the matrix-core lane layout), and the
elementwise training kernels against PyTorch FP32."""
import numpy as np
import pytest
import torch

from digital_twin_brain_amd import _native as nat

pytestmark = pytest.mark.gpu


def ctypes_stream():
    return torch.cuda.current_stream().cuda_stream


def _bits(x):
    """bf16 tensor -> its raw uint16 storage (int16 view)."""
    return x.contiguous().view(torch.int16)


def gemm(A, B, bias=None, act=0, want_bf16=False, lda=None, ldb=None):
    """A [M,K] / B [N,K] bf16 on the device (may be column slices of wider buffers)."""
    L = nat.lib()
    M, K = A.shape
    N = B.shape[0]
    C = torch.empty((M, N), dtype=torch.float32, device="cuda")
    Cb = torch.empty((M, N), dtype=torch.bfloat16, device="cuda") if want_bf16 else None
    nat.check(L.dtb_mlp_gemm_bf16_nt(M, N, K, A.data_ptr(), lda or A.stride(0), B.data_ptr(), ldb or B.stride(0),
                                      nat.ptr(bias), act, C.data_ptr(), N, nat.ptr(Cb), N, ctypes_stream()))
    torch.cuda.synchronize()
    return (C, Cb) if want_bf16 else C


# the first four take 32x32 tiles with K split over the wavefronts (K = 16 gives
# three wavefronts an empty range); 1024x1024 takes 64x64 tiles; the next two
# take 128x128 tiles (at least 256 of them), with a 32-row and a 64-column edge band;
# the last three take LDS-staged 256x256 tiles (64, 72 and 81 tiles: the tile map
# deals workgroups eight at a time, so 81 leaves seven idle)
@pytest.mark.parametrize("M,N,K", [(32, 32, 16), (96, 160, 48), (256, 512, 384), (64, 96, 1040), (1024, 1024, 48),
                                   (2080, 2080, 64), (2048, 2112, 48), (2048, 2048, 128), (2304, 2048, 64),
                                   (2304, 2304, 192)])
def test_gemm_bit_exact_on_integer_operands(M, N, K):
    g = torch.Generator().manual_seed(M * 7 + N + K)
    A = torch.randint(-4, 5, (M, K), generator=g).to(torch.bfloat16)
    B = torch.randint(-4, 5, (N, K), generator=g).to(torch.bfloat16)
    bias = torch.randint(-8, 9, (N,), generator=g).float()
    want = A.double() @ B.double().T + bias.double()
    C = gemm(A.cuda(), B.cuda(), bias.cuda())
    assert torch.equal(C.cpu().double(), want)
    C, Cb = gemm(A.cuda(), B.cuda(), bias.cuda(), act=1, want_bf16=True)
    assert torch.equal(C.cpu().double(), want.clamp_min(0))
    assert torch.equal(Cb.cpu().float(), want.clamp_min(0).float().to(torch.bfloat16).float())


@pytest.mark.parametrize("M,N,K", [(128, 64, 96), (2048, 2048, 256)])
def test_gemm_strided_operands_and_random_values(M, N, K):
    g = torch.Generator().manual_seed(5)
    Abuf = torch.randn((M, K + 24), generator=g).to(torch.bfloat16).cuda()
    Bbuf = torch.randn((N, K + 8), generator=g).to(torch.bfloat16).cuda()
    A, B = Abuf[:, :K], Bbuf[:, :K]
    C = gemm(A, B)
    want = A.float().cpu() @ B.float().cpu().T
    torch.testing.assert_close(C.cpu(), want, rtol=1e-5, atol=1e-5 * K ** 0.5)


def test_gemm_rejects_bad_shapes():
    L = nat.lib()
    A = torch.zeros((32, 16), dtype=torch.bfloat16, device="cuda")
    C = torch.zeros((32, 32), dtype=torch.float32, device="cuda")
    assert L.dtb_mlp_gemm_bf16_nt(33, 32, 16, A.data_ptr(), 16, A.data_ptr(), 16, None, 0, C.data_ptr(), 32, None,
                                  0, None) == nat.DTB_INVALID_PARAMETER
    assert L.dtb_mlp_gemm_bf16_nt(32, 32, 12, A.data_ptr(), 16, A.data_ptr(), 16, None, 0, C.data_ptr(), 32, None,
                                  0, None) == nat.DTB_INVALID_PARAMETER
    assert L.dtb_mlp_gemm_bf16_nt(32, 32, 16, A.data_ptr(), 16, A.data_ptr(), 16, None, 2, C.data_ptr(), 32, None,
                                  0, None) == nat.DTB_INVALID_PARAMETER


def test_transpose_and_convert():
    L = nat.lib()
    g = torch.Generator().manual_seed(2)
    x = torch.randn((70, 45), generator=g)
    xb = x.to(torch.bfloat16).cuda()
    out = torch.empty((45, 70), dtype=torch.bfloat16, device="cuda")
    nat.check(L.dtb_mlp_transpose_bf16(70, 45, xb.data_ptr(), out.data_ptr(), ctypes_stream()))
    cb = torch.empty((70, 45), dtype=torch.bfloat16, device="cuda")
    xs = x.cuda()
    nat.check(L.dtb_mlp_f32_to_bf16(x.numel(), xs.data_ptr(), cb.data_ptr(), ctypes_stream()))
    torch.cuda.synchronize()
    assert torch.equal(_bits(out.cpu()), _bits(xb.cpu().T))
    # round to nearest even, as torch's conversion
    assert torch.equal(_bits(cb.cpu()), _bits(x.to(torch.bfloat16)))


def test_elementwise_training_kernels():
    L = nat.lib()
    g = torch.Generator().manual_seed(3)
    M, N = 64, 96
    dY = torch.randn((M, N), generator=g).cuda()
    Z = torch.randn((M, N), generator=g).cuda()
    dZ = torch.empty_like(dY)
    nat.check(L.dtb_mlp_relu_backward(M * N, dY.data_ptr(), Z.data_ptr(), dZ.data_ptr(), None, ctypes_stream()))
    db = torch.empty(N, device="cuda")
    nat.check(L.dtb_mlp_bias_grad(M, N, dZ.data_ptr(), db.data_ptr(), ctypes_stream()))
    out = torch.randn((M, N), generator=g).cuda()
    tgt = torch.randn((M, N), generator=g).cuda()
    mask = (torch.rand((M, N), generator=g) > 0.3).float().cuda()
    grad = torch.empty_like(out)
    nb = (M * N + 255) // 256
    lsum = torch.empty(nb, device="cuda")
    nat.check(L.dtb_mlp_mse_grad(M * N, out.data_ptr(), tgt.data_ptr(), mask.data_ptr(), 0.5, grad.data_ptr(), None,
                                 lsum.data_ptr(), ctypes_stream()))
    torch.cuda.synchronize()
    assert torch.equal(dZ, torch.where(Z > 0, dY, torch.zeros_like(dY)))
    torch.testing.assert_close(db, dZ.sum(0), rtol=1e-5, atol=1e-5)
    d = (out - tgt) * mask
    torch.testing.assert_close(grad, 0.5 * d, rtol=1e-6, atol=1e-7)
    torch.testing.assert_close(lsum.sum(), (d * d).sum(), rtol=1e-5, atol=1e-5)


def test_adam_matches_torch_optimizer():
    L = nat.lib()
    g = torch.Generator().manual_seed(4)
    n = 1000
    w0 = torch.randn(n, generator=g)
    p = torch.nn.Parameter(w0.clone())
    opt = torch.optim.Adam([p], lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    w = w0.clone().cuda()
    m = torch.zeros(n, device="cuda")
    v = torch.zeros(n, device="cuda")
    wb = torch.empty(n, dtype=torch.bfloat16, device="cuda")
    for t in range(1, 6):
        gr = torch.randn(n, generator=g)
        p.grad = gr.clone()
        opt.step()
        gc = gr.cuda()
        nat.check(L.dtb_mlp_adam(n, w.data_ptr(), gc.data_ptr(), m.data_ptr(), v.data_ptr(), 1e-3, 0.9, 0.999, 1e-8,
                                 t, wb.data_ptr(), ctypes_stream()))
    torch.cuda.synchronize()
    torch.testing.assert_close(w.cpu(), p.detach(), rtol=1e-5, atol=1e-6)
    assert torch.equal(_bits(wb.cpu()), _bits(w.cpu().to(torch.bfloat16)))


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_surrogate_gradients_match_fp64_oracle():
    """One forward + backward of a 3-hidden-layer MLP on the GPU against the
    FP64 oracle with bf16 rounding at the same points; remaining differences
    are FP32 accumulation order (and rare bf16 rounding flips)."""
    from digital_twin_brain_amd.surrogate import MLPSurrogate
    from oracle import mlp_oracle as mo

    rng = np.random.default_rng(7)
    n_in, n_out, r = 70, 45, 50
    sur = MLPSurrogate(n_in, n_out, hidden=(96, 128, 64), max_batch=64, seed=3)
    # non-zero biases so every term of the backward is exercised
    for b in sur.b:
        b.copy_(torch.from_numpy(rng.normal(size=b.shape[0]) * 0.1).float())
    X = rng.normal(size=(r, n_in))
    Y = rng.normal(size=(r, n_out))
    mask = (rng.random((r, n_out)) > 0.2).astype(np.float64)
    ls = sur.gradients(torch.from_numpy(X), torch.from_numpy(Y), torch.from_numpy(mask))
    torch.cuda.synchronize()
    P = sur.params_numpy()
    Xp = np.zeros((r, sur.sizes[0])); Xp[:, :n_in] = X
    Tp = np.zeros((r, sur.sizes[-1])); Tp[:, :n_out] = Y
    Mp = np.zeros((r, sur.sizes[-1])); Mp[:, :n_out] = mask
    Yo, hs = mo.forward(P, Xp, bf16=True)
    loss, gY = mo.loss_and_grad(Yo, Tp, Mp, 1.0 / r)
    grads = mo.backward(P, hs, gY, bf16=True)
    np.testing.assert_allclose(sur.H[-1][:r].cpu().double().numpy(), Yo, rtol=1e-4, atol=1e-4)
    assert abs(float(ls.sum()) / r / 2 - loss) <= 1e-4 * loss
    for l, (dW, db) in enumerate(grads):
        assert _rel(sur.dW[l].cpu().double().numpy(), dW) < 2e-3, l
        assert _rel(sur.db[l].cpu().double().numpy(), db) < 2e-3, l


def test_surrogate_predict_matches_oracle_and_training_fits():
    from digital_twin_brain_amd.surrogate import MLPSurrogate
    from oracle import mlp_oracle as mo

    rng = np.random.default_rng(8)
    n_in, n_out, n = 20, 12, 2048
    A = rng.normal(size=(n_in, n_out)) / np.sqrt(n_in)
    X = rng.normal(size=(n, n_in))
    Y = np.tanh(X @ A) + 0.5 * (X[:, :n_out] ** 2)
    sur = MLPSurrogate(n_in, n_out, hidden=(128, 128), max_batch=256, seed=1, lr=2e-3)
    sur.set_normalization(X, Y)
    hist = sur.fit(torch.from_numpy(X), torch.from_numpy(Y), epochs=40, batch=256)
    assert hist[-1] < 0.1 * hist[0], hist
    pred = sur.predict(torch.from_numpy(X[:300])).cpu().double().numpy()
    # the same prediction from the oracle with the trained weights
    P = sur.params_numpy()
    xs, xh = sur.x_scale.cpu().double().numpy(), sur.x_shift.cpu().double().numpy()
    Xn = np.zeros((300, sur.sizes[0]))
    Xn[:, :n_in] = (X[:300].astype(np.float32) * xs.astype(np.float32) + xh.astype(np.float32))
    Yn, _ = mo.forward(P, Xn, bf16=True)
    want = Yn[:, :n_out] * sur._y_std.cpu().double().numpy() + sur._y_mean.cpu().double().numpy()
    np.testing.assert_allclose(pred, want, rtol=1e-3, atol=1e-3 * np.abs(want).max())
    nrmse = np.linalg.norm(pred - Y[:300]) / np.sqrt(pred.size) / (Y.max() - Y.min())
    assert nrmse < 0.05, nrmse


def test_loss_partial_sums_cover_only_the_current_batch():
    """After a full-size batch fills the loss buffer, a smaller batch's
    returned partial sums add up to that batch's own squared error (the
    buffer's tail belongs to the earlier batch)."""
    from digital_twin_brain_amd.surrogate import GRUSurrogate, MLPSurrogate

    rng = np.random.default_rng(5)
    sur = MLPSurrogate(20, 10, hidden=(64, 64), max_batch=256, seed=1)
    sur.gradients(torch.randn(256, 20), torch.randn(256, 10) * 100.0)
    X, Y = rng.normal(size=(40, 20)), rng.normal(size=(40, 10)).astype(np.float32)
    ls = sur.gradients(torch.from_numpy(X), torch.from_numpy(Y))
    out = sur.H[-1][:40, :10].cpu().double().numpy()
    want = ((out - Y) ** 2).sum()
    assert abs(float(ls.sum()) - want) <= 1e-4 * want

    gru = GRUSurrogate(10, 6, hidden=32, max_batch=64, max_steps=8, seed=1)
    gru.gradients(torch.randn(8, 64, 10), torch.randn(8, 64, 6) * 100.0)
    T, r = 3, 20
    X, Y = rng.normal(size=(T, r, 10)), rng.normal(size=(T, r, 6)).astype(np.float32)
    ls = gru.gradients(torch.from_numpy(X), torch.from_numpy(Y))
    B = (r + 31) // 32 * 32  # rows padded to a multiple of 32
    out = gru.Y[:T * B].cpu().double().numpy().reshape(T, B, -1)[:, :r, :6]
    want = ((out - Y) ** 2).sum()
    assert abs(float(ls.sum()) - want) <= 1e-4 * want


def test_gru_gradients_match_fp64_oracle():
    """BPTT of the GRU surrogate on the GPU against the FP64 oracle (itself
    pinned to torch.nn.GRU) with bf16 rounding at the same points."""
    from digital_twin_brain_amd.surrogate import GRUSurrogate
    from oracle import mlp_oracle as mo

    rng = np.random.default_rng(11)
    T, r, n_in, n_out, H = 6, 40, 50, 30, 64
    sur = GRUSurrogate(n_in, n_out, hidden=H, max_batch=64, max_steps=8, seed=2)
    sur.p["b_o"].copy_(torch.from_numpy(rng.normal(size=sur.O) * 0.1).float())
    X = rng.normal(size=(T, r, n_in))
    Y = rng.normal(size=(T, r, n_out))
    h0 = rng.normal(size=(r, H)) * 0.3
    ls = sur.gradients(torch.from_numpy(X), torch.from_numpy(Y), h0=torch.from_numpy(h0))
    torch.cuda.synchronize()
    P = sur.params_numpy()
    Xp = np.zeros((T, r, sur.I)); Xp[:, :, :n_in] = X
    Tp = np.zeros((T, r, sur.O)); Tp[:, :, :n_out] = Y
    Mp = np.zeros((T, r, sur.O)); Mp[:, :, :n_out] = 1.0
    h0f = h0.astype(np.float32).astype(np.float64)
    Yo, cache = mo.gru_forward(P, Xp, h0f, bf16=True)
    loss, gY = mo.loss_and_grad(Yo, Tp, Mp, 1.0 / (T * r))
    g = mo.gru_backward(P, cache, gY, bf16=True)
    B = 64  # 40 rows padded to a multiple of 32
    Yg = sur.Y[:T * B].cpu().double().numpy().reshape(T, B, -1)[:, :r]
    np.testing.assert_allclose(Yg, Yo, rtol=1e-3, atol=1e-3)
    assert abs(float(ls.sum()) / (T * r) / 2 - loss) <= 1e-3 * loss
    for k in ("W_ih", "b_ih", "W_hh", "b_hh", "W_o", "b_o"):
        assert _rel(sur.grad[k].cpu().double().numpy(), g[k]) < 5e-3, k


def test_gru_training_fits_a_delayed_sequence_task():
    """y_t = x_{t-1} (a one-step memory): only the recurrence can fit it."""
    from digital_twin_brain_amd.surrogate import GRUSurrogate

    rng = np.random.default_rng(12)
    T, n, d = 8, 512, 8
    X = rng.normal(size=(T, n, d))
    Y = np.concatenate([np.zeros((1, n, d)), X[:-1]], axis=0)
    sur = GRUSurrogate(d, d, hidden=64, max_batch=64, max_steps=T, seed=1, lr=3e-3)
    hist = sur.fit(torch.from_numpy(X), torch.from_numpy(Y), epochs=30, batch=64)
    assert hist[-1] < 0.2 * hist[0], hist
    pred, hT = sur.predict(torch.from_numpy(X[:, :64]))
    err = pred.cpu().double().numpy()[1:] - Y[1:, :64]
    assert np.sqrt((err ** 2).mean()) < 0.4
    assert hT.shape == (64, 64)
