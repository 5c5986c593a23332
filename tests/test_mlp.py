"""The canopy surrogate's kernels (cable_amd/csrc/mlp.hip) against the CPU
oracle (oracle/mlp_oracle.py), and the oracle against finite differences."""
import numpy as np
import pytest

from oracle import mlp_oracle as O


def test_bf16_rounding_known_values():
    vals = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 2.0 ** -7, 1.0 + 3 * 2.0 ** -8, -2.5, 3.140625, 1e-3],
                    dtype=np.float32)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: ties go to the even mantissa (1)
    # 1 + 3*2^-8 is a tie between 1 + 2^-7 and 1 + 2^-6: even is 1 + 2^-6
    # 1e-3 is 0x3A83126F in FP32: the low half 0x126F rounds down to 0x3A83 = 0.00099945068359375
    exp = np.array([1.0, 1.0, 1.0 + 2.0 ** -7, 1.0 + 2.0 ** -6, -2.5, 3.140625, 0.00099945068359375])
    np.testing.assert_array_equal(O.bf16(vals), exp)


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = O.init_params(3)
    x = rng.normal(size=(7, O.NIN))
    t = rng.normal(size=(7, O.NOUT))
    _, g = O.loss_and_grad(p, x, t, bf16_ops=False)
    idx = rng.choice(O.NPARAM, 60, replace=False)
    idx = np.concatenate([idx, [O.NPARAM - 1, O.NPARAM - 9, 3072, 3200]])  # biases of every layer too
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (O.loss_and_grad(pp, x, t, False)[0] - O.loss_and_grad(pm, x, t, False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_oracle_matches_torch_forward_and_autograd():
    """The oracle's forward pass and gradient (FP64, no BF16 rounding)
    against torch's own Linear/Hardswish layers and autograd on the CPU: an
    implementation the oracle shares nothing with, so a transposed weight,
    a wrong activation or a dropped bias term in either pass shows."""
    import torch
    rng = np.random.default_rng(4)
    p = O.init_params(8)
    x = rng.normal(size=(33, O.NIN))
    t = rng.normal(size=(33, O.NOUT))
    Ws, bs = O.unpack(p)
    layers = []
    for l, (W, b) in enumerate(zip(Ws, bs)):
        lin = torch.nn.Linear(W.shape[0], W.shape[1]).double()
        with torch.no_grad():
            lin.weight.copy_(torch.from_numpy(W.T.copy()))
            lin.bias.copy_(torch.from_numpy(b))
        layers += [lin] + ([torch.nn.Hardswish()] if l < 3 else [])
    net = torch.nn.Sequential(*layers)
    y = net(torch.from_numpy(x))
    loss = torch.mean((y - torch.from_numpy(t)) ** 2)
    loss.backward()
    yo, _ = O.forward(p, x, bf16_ops=False)
    lo, go = O.loss_and_grad(p, x, t, bf16_ops=False)
    np.testing.assert_allclose(yo, y.detach().numpy(), rtol=1e-12, atol=1e-12)
    assert abs(lo - loss.item()) <= 1e-12 * loss.item()
    gt = np.concatenate([np.concatenate([m.weight.grad.numpy().T.reshape(-1), m.bias.grad.numpy()])
                         for m in net if isinstance(m, torch.nn.Linear)])
    np.testing.assert_allclose(go, gt, rtol=1e-10, atol=1e-14)


def test_oracle_adam_matches_closed_form_first_step():
    # the first Adam step moves every parameter by lr * sign(g) (up to eps)
    g = np.array([0.3, -2.0, 1e-3])
    p, m, v = O.adam(np.zeros(3), g, np.zeros(3), np.zeros(3), 1, 0.01)
    np.testing.assert_allclose(p, -0.01 * np.sign(g), rtol=1e-4)


def test_shuffled_order_is_a_permutation_of_every_epoch():
    """The order the device draws for training (cable_mlp_shuffled_order,
    the same function the rows kernel calls): every epoch of n positions is
    a permutation of range(n), for n at and around powers of two (where the
    cycle walk has the most to skip) and n = 1; consecutive epochs and
    different seeds give different orders, and few samples stay in place."""
    from cable_amd.surrogate import shuffled_order
    for n in (1, 2, 3, 5, 64, 65, 1000, 4096, 4097, 65543):
        two = shuffled_order(7, n, 0, 2 * n)
        for e in range(2):
            np.testing.assert_array_equal(np.sort(two[e * n:(e + 1) * n]), np.arange(n))
        if n >= 1000:
            assert (two[:n] != two[n:]).mean() > 0.99
            assert (shuffled_order(8, n, 0, n) != two[:n]).mean() > 0.99
            assert (two[:n] == np.arange(n)).mean() < 0.01
    # a window starting mid-epoch is the same sequence
    np.testing.assert_array_equal(shuffled_order(7, 1000, 1500, 700), shuffled_order(7, 1000, 0, 2200)[1500:])


def _layer_slices():
    o, out = 0, []
    for i, n in O.SHAPES:
        out.append(slice(o, o + i * n))
        o += i * n
        out.append(slice(o, o + n))
        o += n
    return out


@pytest.mark.gpu
def test_gpu_forward_matches_oracle():
    from cable_amd.surrogate import MLP
    print("[mlp] forward", flush=True)
    rng = np.random.default_rng(2)
    p = O.init_params(5).astype(np.float32)
    x = rng.normal(size=(300, O.NIN)).astype(np.float32)   # 300: not a multiple of 32
    net = MLP(batch=512, params=p)
    y = net.forward(x)
    ref, _ = O.forward(p.astype(np.float64), x, bf16_ops=True)
    err = np.abs(y - ref).max() / np.abs(ref).max()
    print(f"forward max error / max |y| = {err:.2e}", flush=True)
    assert err < 2e-3
    net.close()


@pytest.mark.gpu
def test_gpu_gradient_matches_oracle():
    from cable_amd.surrogate import MLP
    print("[mlp] gradient", flush=True)
    rng = np.random.default_rng(3)
    p = O.init_params(6).astype(np.float32)
    x = rng.normal(size=(2048, O.NIN)).astype(np.float32)
    t = rng.normal(size=(2048, O.NOUT)).astype(np.float32)
    net = MLP(batch=2048, params=p)
    loss, g = net.gradient(x, t)
    lref, gref = O.loss_and_grad(p.astype(np.float64), x, t, bf16_ops=True)
    assert abs(loss - lref) <= 1e-3 * lref
    for k, s in enumerate(_layer_slices()):
        a, b = g[s], gref[s]
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        print(f"{'Wb'[k % 2]}{k // 2 + 1}: relative gradient error {rel:.2e}", flush=True)
        assert rel < 5e-3, (k, rel)
    net.close()


@pytest.mark.gpu
def test_gpu_adam_step_matches_oracle():
    from cable_amd.surrogate import MLP
    print("[mlp] adam", flush=True)
    rng = np.random.default_rng(4)
    p = O.init_params(7).astype(np.float32)
    x = rng.normal(size=(256, O.NIN)).astype(np.float32)
    t = rng.normal(size=(256, O.NOUT)).astype(np.float32)
    net = MLP(batch=256, params=p)
    net.set_norm(np.concatenate([np.zeros(24), np.ones(24), np.zeros(8), np.ones(8)]))
    _, g = net.gradient(x, t)
    net.set_params(p)                       # reset Adam
    net.load_samples(x, t)
    from cable_amd import surrogate as S
    order = np.arange(256, dtype=np.int64)
    S._check(S.lib().cable_mlp_train(net.h, order.ctypes.data, 1, 1e-3, None))
    after = net.params()
    ref, _, _ = O.adam(p.astype(np.float64), g.astype(np.float64), 0.0, 0.0, 1, 1e-3)
    np.testing.assert_allclose(after, ref, rtol=1e-5, atol=1e-7)
    net.close()


@pytest.mark.gpu
def test_gpu_training_fits_a_smooth_map():
    """Training on the GPU drives the loss of a smooth 24 -> 8 map down."""
    from cable_amd.surrogate import MLP, norm_stats
    print("[mlp] training", flush=True)
    rng = np.random.default_rng(5)
    A = rng.normal(size=(O.NIN, O.NOUT)) / np.sqrt(O.NIN)
    x = rng.normal(size=(65536, O.NIN)).astype(np.float32)
    y = np.tanh(x @ A).astype(np.float32)
    net = MLP(batch=2048)
    net.set_norm(norm_stats(x, y))
    net.load_samples(x, y)
    losses = net.train(epochs=20, lr=2e-3)
    print(f"loss: first {losses[:8].mean():.4f} last {losses[-8:].mean():.5f}", flush=True)
    assert losses[-8:].mean() < 0.05 * losses[:8].mean()
    net.close()


@pytest.mark.gpu
def test_gpu_shuffled_training_is_training_in_that_order():
    """cable_mlp_train_shuffled draws the order on the device; the same
    batches through cable_mlp_train with the host's copy of that order give
    the same parameters and losses bit for bit (here across an epoch
    boundary: positions 3,584 .. 5,631 of 5,000 samples)."""
    from cable_amd.surrogate import MLP, shuffled_order
    print("[mlp] shuffled order", flush=True)
    rng = np.random.default_rng(8)
    p = O.init_params(9).astype(np.float32)
    x = rng.normal(size=(5000, O.NIN)).astype(np.float32)
    y = rng.normal(size=(5000, O.NOUT)).astype(np.float32)
    out = []
    for device_order in (True, False):
        net = MLP(batch=512, params=p)
        net.load_samples(x, y)
        if device_order:
            losses = np.zeros(4, dtype=np.float32)
            from cable_amd import surrogate as S
            S._check(S.lib().cable_mlp_train_shuffled(net.h, 11, 7, 4, 1e-3, losses.ctypes.data))
        else:
            losses = net.train_order(shuffled_order(11, 5000, 7 * 512, 4 * 512), 1e-3)
        out.append((net.params(), losses))
        net.close()
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][1], out[1][1])
