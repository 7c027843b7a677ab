"""Snapshot autoencoder on the GPU (BF16 MFMA, FP32 accumulation) against
the CPU oracle (oracle/ae_oracle.py): layer outputs, gradients, the Adam
step, training progress and the device encode / decode path."""
import numpy as np
import pytest

from conftest import rel_l2
from geophysics_amd import vti
from oracle import ae_oracle as ae

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _flat(params):
    return np.concatenate([w.ravel() for w in params]).astype(np.float32)


def _unflat(flat):
    out, f = [], 0
    for kind, ci, co, a in ae.NET:
        r, c = ae.kdim(kind, ci, co)
        out.append(flat[f:f + r * c].reshape(r, c))
        f += r * c
    return out


def _wave_crops(nb, n, seed):
    """Smooth oscillatory test volumes: a few Gaussian-windowed plane waves."""
    rng = np.random.default_rng(seed)
    g = np.arange(n, dtype=np.float64)
    Y, X, Z = np.meshgrid(g, g, g, indexing="ij")
    out = np.zeros((nb, n, n, n, 1), np.float32)
    for b in range(nb):
        v = np.zeros((n, n, n))
        for _ in range(3):
            k = rng.normal(size=3)
            k *= rng.uniform(0.3, 0.9) / np.linalg.norm(k)
            c = rng.uniform(0, n, 3)
            r2 = (Y - c[0]) ** 2 + (X - c[1]) ** 2 + (Z - c[2]) ** 2
            v += np.cos(k[0] * Y + k[1] * X + k[2] * Z + rng.uniform(0, 6.28)) * np.exp(-r2 / (2 * (n / 3) ** 2))
        out[b, ..., 0] = (v / np.abs(v).max()).astype(np.float32)
    return out


def _bf16_close(gpu, ref, what):
    """gpu: bf16-stored values; ref: the oracle's unrounded output.  Each
    value is the bf16 rounding of ref up to one bf16 ulp (fp32 sums in a
    different order can move a value across a rounding boundary)."""
    r = ae.bf16(ref.astype(np.float32))
    ulp = np.maximum(np.abs(r) * 2.0 ** -7, 1e-30)
    bad = np.abs(gpu - r) > ulp * 1.001
    frac = bad.mean()
    _say(f"{what}: {int(bad.sum())} of {bad.size} beyond 1 bf16 ulp, rel-L2 {rel_l2(gpu, ref):.2e}")
    assert frac == 0.0
    assert rel_l2(gpu, ref) < 1e-2


def test_layer_outputs_match_oracle():
    """Each stored activation against the oracle layer applied to the GPU's
    own input of that layer (bf16 values), every kind and width."""
    n, nb = 16, 2
    with vti.Autoencoder(crop=n, batch=nb, seed=1) as A:
        p = ae.init_params(7)
        A.set_params(_flat(p))
        x = _wave_crops(nb, n, 3)
        A.train_batch(x, apply=False)
        shapes = [(nb, n, n, n, 1)]
        for kind, ci, co, a in ae.NET:
            shapes.append(ae.out_shape(kind, shapes[-1], co))
        acts = [A.activation(i, shapes[i]) for i in range(len(ae.NET) + 1)]
    assert np.array_equal(acts[0], ae.bf16(x))
    for li, (kind, ci, co, a) in enumerate(ae.NET):
        ref = ae.conv_fwd(kind, acts[li], ae.bf16(p[li]), np.zeros(co), a)
        if li == len(ae.NET) - 1:
            assert rel_l2(acts[li + 1], ref) < 1e-6, rel_l2(acts[li + 1], ref)
            assert np.allclose(acts[li + 1], ref, rtol=1e-5, atol=1e-6)
        else:
            _bf16_close(acts[li + 1], ref, f"layer {li} ({kind} {ci}->{co})")


def test_gradients_match_oracle():
    """Loss and every weight gradient of one step against the oracle's
    backward pass (same bf16 rounding points)."""
    n, nb = 16, 2
    with vti.Autoencoder(crop=n, batch=nb, seed=2) as A:
        p = ae.init_params(11)
        A.set_params(_flat(p))
        x = _wave_crops(nb, n, 5)
        loss = A.train_batch(x, apply=False)
        g = _unflat(A.grads())
    rl, rg = ae.loss_and_grads(p, x)
    _say(f"loss GPU {loss:.6e} oracle {rl:.6e}")
    assert abs(loss - rl) <= 1e-3 * rl
    for li, (gw, ow) in enumerate(zip(g, rg)):
        ew = rel_l2(gw, ow)
        _say(f"layer {li} {ae.NET[li][:3]}: dW rel-L2 {ew:.2e}")
        assert ew < 2e-2, li


def test_adam_step_matches_oracle():
    n, nb = 16, 1
    with vti.Autoencoder(crop=n, batch=nb, seed=3) as A:
        p = ae.init_params(13)
        A.set_params(_flat(p))
        A.set_lr(1e-3)
        x = _wave_crops(nb, n, 8)
        A.train_batch(x, apply=False)
        g = A.grads()
        A.train_batch(x, apply=True)        # recomputes the same gradients, then steps
        w1 = A.params()
    ref, _, _ = ae.adam(_flat(p), np.zeros_like(g), np.zeros_like(g), g, 1, 1e-3)
    assert np.allclose(w1, ref, rtol=0, atol=2e-6), np.abs(w1 - ref).max()


def test_training_reduces_loss():
    n, nb = 32, 4
    with vti.Autoencoder(crop=n, batch=nb, seed=4) as A:
        A.set_lr(2e-3)
        losses = []
        for it in range(300):
            losses.append(A.train_batch(_wave_crops(nb, n, 100 + it % 20), apply=True))
    _say("loss first / last:", losses[0], losses[-1], "mean of last 10:", np.mean(losses[-10:]))
    assert np.mean(losses[-10:]) < 0.25 * np.mean(losses[:5])


def test_device_encode_decode_equals_training_forward():
    """The full-volume path (strided fp32 field in, latent, fp32 field out)
    reproduces the training forward pass on the same volume bit for bit
    (the volume's max |value| is 1, so both paths see the same input)."""
    import torch
    n = 32
    x = _wave_crops(1, n, 21)
    with vti.Autoencoder(crop=n, batch=1, seed=5) as A:
        _, rec = A.train_batch(x, apply=False, want_recon=True)
        # the volume embedded in a padded field with strides
        pad = 5
        F = torch.zeros((n + 2 * pad, n + 2 * pad, n + 2 * pad + 3), dtype=torch.float32, device="cuda")
        F[pad:pad + n, pad:pad + n, pad:pad + n] = torch.from_numpy(x[0, ..., 0]).cuda()
        sy, sx = F.shape[1] * F.shape[2], F.shape[2]
        off = pad * sy + pad * sx + pad
        lat = torch.zeros(vti.Autoencoder.latent_elems(n), dtype=torch.int16, device="cuda")
        sc = torch.zeros(1, dtype=torch.float32, device="cuda")
        out = torch.zeros((n, n, n), dtype=torch.float32, device="cuda")
        A.encode(F.data_ptr(), off, sy, sx, n, lat.data_ptr(), sc.data_ptr())
        A.decode(lat.data_ptr(), sc.data_ptr(), n, out.data_ptr())
        torch.cuda.synchronize()
    assert float(sc.item()) == 1.0
    assert lat.numel() * 2 * 16 == n ** 3 * 4
    assert np.array_equal(out.cpu().numpy(), rec[0, ..., 0])
