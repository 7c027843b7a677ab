"""The autoencoder oracle against torch.nn.functional (CPU, float64) for the
three layer kinds, forward and backward, and against finite differences of
the full network loss.  No GPU needed."""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from oracle import ae_oracle as ae


def _to_torch_w(kind, w, ci, co):
    if kind == "down":
        return torch.tensor(w.reshape(co, 2, 2, 2, ci).transpose(0, 4, 1, 2, 3).copy())
    if kind == "same":
        return torch.tensor(w.reshape(co, 3, 3, 3, ci).transpose(0, 4, 1, 2, 3).copy())
    # up: W[(tap, co)][ci] -> torch (ci, co, 2, 2, 2)
    return torch.tensor(w.reshape(2, 2, 2, co, ci).transpose(4, 3, 0, 1, 2).copy())


def _torch_layer(kind, x, w, b):
    if kind == "down":
        return F.conv3d(x, w, b, stride=2)
    if kind == "same":
        return F.conv3d(x, w, b, padding=1)
    return F.conv_transpose3d(x, w, b, stride=2)


@pytest.mark.parametrize("kind,ci,co", [("down", 1, 16), ("down", 16, 32), ("same", 8, 8),
                                        ("up", 16, 8), ("up", 8, 1)])
def test_layer_kinds_equal_torch(kind, ci, co):
    rng = np.random.default_rng(hash((kind, ci, co)) % 2**32)
    shape = (2, 4, 6, 8, ci)
    x = rng.standard_normal(shape)
    rows, cols = ae.kdim(kind, ci, co)
    w = rng.standard_normal((rows, cols))
    b = rng.standard_normal(co)
    y = ae.conv_fwd(kind, x, w, b, 0)
    xt = torch.tensor(x.transpose(0, 4, 1, 2, 3).copy(), requires_grad=True)
    wt = _to_torch_w(kind, w, ci, co).requires_grad_(True)
    bt = torch.tensor(b, requires_grad=True)
    yt = _torch_layer(kind, xt, wt, bt)
    assert np.allclose(y, yt.detach().numpy().transpose(0, 2, 3, 4, 1), rtol=1e-12, atol=1e-12)
    assert y.shape == ae.out_shape(kind, shape, co)
    # backward with a random upstream gradient
    g = rng.standard_normal(y.shape)
    yt.backward(torch.tensor(g.transpose(0, 4, 1, 2, 3).copy()))
    dx = ae.conv_bwd_data(kind, g, w, shape)
    dw, db = ae.conv_bwd_weight(kind, x, g)
    assert np.allclose(dx, xt.grad.numpy().transpose(0, 2, 3, 4, 1), rtol=1e-12, atol=1e-12)
    wg = wt.grad.numpy()
    if kind == "up":
        wg = wg.transpose(2, 3, 4, 1, 0).reshape(rows, cols)
    else:
        k = 3 if kind == "same" else 2
        wg = wg.transpose(0, 2, 3, 4, 1).reshape(co, k ** 3 * ci)
    assert np.allclose(dw, wg, rtol=1e-12, atol=1e-12)
    assert np.allclose(db, bt.grad.numpy(), rtol=1e-12, atol=1e-12)


def test_bf16_rounding_is_nearest_even():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 3.0e38, 1e-40], np.float32)
    got = ae.bf16(vals)
    # 1+2^-8 is a tie between 1 and 1+2^-7: even mantissa (1.0) wins;
    # 1+3*2^-8 ties between 1+2^-7 (odd) and 1+2^-6 (even): rounds up
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == np.float32(1.0 + 2 ** -6)
    assert got[3] == 1.0 and got[4] == -2.5
    ref = torch.tensor(vals).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(got, ref)


def test_network_shapes_and_compression():
    p = ae.init_params(0)
    x = np.random.default_rng(1).standard_normal((1, 16, 16, 16, 1)).astype(np.float32)
    outs = ae.forward(p, x)
    lat = outs[ae.N_ENC]
    assert lat.shape == (1, 2, 2, 2, 64)
    assert outs[-1].shape == x.shape
    # latent in bf16 (2 B) against the fp32 input (4 B): 16x fewer bytes
    assert x.size * 4 == 16 * lat.size * 2


def test_network_is_positively_homogeneous():
    """No biases: zero in, zero out, and f(a x) = a f(x) for a > 0 (a a
    power of two, so bf16 rounding commutes with the scaling)."""
    p = ae.init_params(2)
    x = np.random.default_rng(4).standard_normal((1, 16, 16, 16, 1)).astype(np.float32)
    assert not ae.forward(p, np.zeros_like(x))[-1].any()
    y1 = ae.forward(p, x)[-1]
    y4 = ae.forward(p, 4.0 * x)[-1]
    assert np.allclose(y4, 4.0 * y1, rtol=1e-6, atol=1e-12)


def test_network_gradient_matches_finite_differences():
    """Loss gradient of the oracle network in float64 (no bf16 rounding)
    against central differences along random directions."""
    rng = np.random.default_rng(3)
    p = [w.astype(np.float64) for w in ae.init_params(5)]
    x = rng.standard_normal((1, 8, 8, 8, 1))

    def loss64(params):
        h = x
        for li, (kind, ci, co, a) in enumerate(ae.NET):
            h = ae.conv_fwd(kind, h, params[li], np.zeros(co), a)
        return float(((h - x) ** 2).mean())

    # analytic gradient: the oracle's backward without rounding
    outs = [x]
    h = x
    for li, (kind, ci, co, a) in enumerate(ae.NET):
        h = ae.conv_fwd(kind, h, p[li], np.zeros(co), a)
        outs.append(h)
    g = 2.0 * (outs[-1] - x) / x.size
    grads = [None] * len(ae.NET)
    for li in range(len(ae.NET) - 1, -1, -1):
        kind = ae.NET[li][0]
        grads[li] = ae.conv_bwd_weight(kind, outs[li], g)[0]
        if li > 0:
            g = ae.act_grad(outs[li], ae.conv_bwd_data(kind, g, p[li], outs[li].shape), ae.NET[li - 1][3])
    for trial in range(4):
        dirs = [rng.standard_normal(w.shape) for w in p]
        ana = sum(float((gw * dw).sum()) for gw, dw in zip(grads, dirs))
        h = 1e-6
        num = (loss64([w + h * d for w, d in zip(p, dirs)]) - loss64([w - h * d for w, d in zip(p, dirs)])) / (2 * h)
        assert abs(num - ana) <= 1e-6 * max(1.0, abs(ana)), (trial, num, ana)


def test_adam_step_closed_form():
    """First Adam step moves every weight by lr * sign(g) (bias-corrected
    moments are g and g^2)."""
    w = np.array([1.0, -2.0, 0.5], np.float32)
    g = np.array([0.3, -1e-3, 2.0], np.float32)
    w1, m1, v1 = ae.adam(w, np.zeros(3), np.zeros(3), g, 1, 1e-2, eps=0.0)
    assert np.allclose(w1, w - 1e-2 * np.sign(g), rtol=0, atol=1e-7)
    assert np.allclose(m1, 0.1 * g) and np.allclose(v1, 0.001 * g * g)


def test_pca_weights_make_the_network_a_block_projection():
    """With pca_params(V) the oracle network maps every 8^3 block b of the
    input to V^T V b (block order of block_order_offsets): the weight
    layout, tap order and block ordering agree with a plain numpy
    projection."""
    rng = np.random.default_rng(8)
    Q, _ = np.linalg.qr(rng.standard_normal((512, 64)))
    V = Q.T.astype(np.float64)
    p = pca_params(V)
    x = rng.standard_normal((1, 16, 16, 16, 1))
    h = x
    for li, (kind, ci, co, a) in enumerate(ae.NET):
        h = ae.conv_fwd(kind, h, p[li], np.zeros(co), a)
    offs = ae.block_order_offsets()
    ref = np.zeros_like(x)
    for by in range(2):
        for bx in range(2):
            for bz in range(2):
                blk = np.array([x[0, 8 * by + y, 8 * bx + xx, 8 * bz + z, 0] for y, xx, z in offs])
                rec = V.T @ (V @ blk)
                for (y, xx, z), val in zip(offs, rec):
                    ref[0, 8 * by + y, 8 * bx + xx, 8 * bz + z, 0] = val
    assert np.allclose(h, ref, rtol=0, atol=1e-6)      # weights are held in fp32
    assert sorted(offs) == [(y, xx, z) for y in range(8) for xx in range(8) for z in range(8)]


def pca_params(V):
    return ae.pca_params(V)


def test_library_eigen_solver_matches_numpy():
    """The host Jacobi routine used by the fit against numpy.linalg.eigh on
    a Gram matrix of smooth 8^3 blocks (fast-decaying spectrum) and on a
    random symmetric matrix."""
    from geophysics_amd import vti
    rng = np.random.default_rng(9)
    g = np.arange(8.0)
    blocks = []
    for _ in range(600):
        k = rng.normal(size=3) * 0.6
        Y, X, Z = np.meshgrid(g, g, g, indexing="ij")
        blocks.append(np.cos(k[0] * Y + k[1] * X + k[2] * Z + rng.uniform(0, 6.28)).ravel())
    B = np.array(blocks)
    for A in (B.T @ B, (lambda M: M + M.T)(rng.standard_normal((40, 40)))):
        w, v = vti.sym_eigen(A)
        wr = np.linalg.eigvalsh(A)
        assert np.allclose(np.sort(w), wr, rtol=1e-9, atol=1e-9 * np.abs(wr).max())
        assert np.allclose(A @ v, v * w, atol=1e-8 * np.abs(wr).max())
        assert np.allclose(v.T @ v, np.eye(len(w)), atol=1e-10)
