"""CPU restatement of the snapshot autoencoder (numpy, float64 sums).

TEST INFRASTRUCTURE ONLY: the checker for the HIP network kernels in
tests/; never imported by the product.

The network has no counterpart in the reference (the reference stores no
source-wavefield snapshots; its RTM re-runs the forward operator,
Toa_gpu2DvtiFdRtmAdcigsLaplace.cu:1423-1437).  What is restated here is the
network of geophysics_amd/csrc/vti_nn.hip term by term:

* volumes are channels-last, (nb, ny, nx, nz, c), z fastest inside a channel
  vector;
* three layer kinds over 3D volumes:
  - "down": kernel 2, stride 2 (torch conv3d(k=2, s=2));
  - "same": kernel 3, stride 1, zero padding 1 (torch conv3d(k=3, p=1));
  - "up":   kernel 2, stride 2 transposed (torch conv_transpose3d(k=2, s=2));
* weights are the GEMM matrix of the layer, row n = output column:
  - down / same: W[co][k], k = tap * ci + c, tap = (dy*T + dx)*T + dz
    (T = 2 or 3, same-taps offset by -1);
  - up:          W[n][ci], n = tap * co + c, tap = (dy*2 + dx)*2 + dz;
* activation: leaky ReLU with slope 1/8 (exact in bf16 and fp32) or none
  (the network's layers are linear; the kernels implement both); no biases
  (homogeneous: zero in, zero out);
* bf16 (round to nearest even) at the points the GPU stores bf16: weights,
  activations, the latent, activation gradients; fp32 master weights and
  Adam moments; the reconstruction and the loss in fp32.

Pinned against torch.nn.functional (CPU, float64) for all three kinds in
forward and backward (tests/test_ae_oracle.py), and by finite differences.
"""
import numpy as np

LEAK = 0.125

# (kind, ci, co, act) of the 4-level encoder / decoder; latent after "E4"
NET = [
    ("down", 1, 8, 0),     # E1  N   -> N/2
    ("down", 8, 64, 0),    # E2  N/2 -> N/4
    ("down", 64, 64, 0),   # E3  N/4 -> N/8
    ("same", 64, 64, 0),   # E4  latent (64 ch at N/8: 16x fewer bytes in bf16 than fp32 input)
    ("same", 64, 64, 0),   # D4
    ("up", 64, 64, 0),     # D3  N/8 -> N/4
    ("up", 64, 8, 0),      # D2  N/4 -> N/2
    ("up", 8, 1, 0),       # D1  N/2 -> N
]
N_ENC = 4


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    nan = np.isnan(a)
    if nan.any():
        out = out.copy()
        out[nan] = np.nan
    return out.reshape(np.shape(x))


def taps(kind):
    T = 3 if kind == "same" else 2
    lo = -1 if kind == "same" else 0
    return [(dy + lo, dx + lo, dz + lo) for dy in range(T) for dx in range(T) for dz in range(T)]


def kdim(kind, ci, co):
    """(rows, cols) of the weight matrix."""
    if kind == "up":
        return 8 * co, ci
    return co, len(taps(kind)) * ci


def out_shape(kind, shape, co):
    nb, ny, nx, nz, _ = shape
    if kind == "down":
        return (nb, ny // 2, nx // 2, nz // 2, co)
    if kind == "up":
        return (nb, 2 * ny, 2 * nx, 2 * nz, co)
    return (nb, ny, nx, nz, co)


def act(y, a):
    return np.where(y > 0, y, LEAK * y) if a else y


def act_grad(y_out, g, a):
    """dL/dpre from dL/dy and the layer output (the sign of the output is
    the sign of the pre-activation)."""
    return np.where(y_out > 0, g, LEAK * g) if a else g


def _gather(x, kind):
    """(nb, oy, ox, oz, ntap*ci) patch matrix of a down / same layer."""
    nb, ny, nx, nz, ci = x.shape
    tl = taps(kind)
    if kind == "down":
        oy, ox, oz = ny // 2, nx // 2, nz // 2
        cols = [x[:, dy::2, dx::2, dz::2, :][:, :oy, :ox, :oz, :] for dy, dx, dz in tl]
    else:
        xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
        cols = [xp[:, 1 + dy:1 + dy + ny, 1 + dx:1 + dx + nx, 1 + dz:1 + dz + nz, :] for dy, dx, dz in tl]
    return np.concatenate(cols, axis=-1)


def conv_fwd(kind, x, w, b, a):
    """Layer output before rounding (float64): x channels-last, w the GEMM
    matrix, b the bias (co,)."""
    x = np.asarray(x, np.float64)
    w = np.asarray(w, np.float64)
    if kind == "up":
        nb, ny, nx, nz, ci = x.shape
        co = w.shape[0] // 8
        t = x @ w.T                                  # (nb,ny,nx,nz, 8*co)
        y = np.zeros((nb, 2 * ny, 2 * nx, 2 * nz, co))
        for d, (dy, dx, dz) in enumerate(taps("up")):
            y[:, dy::2, dx::2, dz::2, :] = t[..., d * co:(d + 1) * co]
    else:
        y = _gather(x, kind) @ w.T
    return act(y + np.asarray(b, np.float64), a)


def conv_bwd_data(kind, gpre, w, in_shape):
    """dL/dx (float64) from dL/dpre of the layer."""
    g = np.asarray(gpre, np.float64)
    w = np.asarray(w, np.float64)
    nb, ny, nx, nz, ci = in_shape
    if kind == "up":
        co = w.shape[0] // 8
        cols = [g[:, dy::2, dx::2, dz::2, :] for dy, dx, dz in taps("up")]
        return np.concatenate(cols, axis=-1) @ w     # (.., 8co) @ (8co, ci)
    t = g @ w                                        # (nb,oy,ox,oz, ntap*ci)
    dx_ = np.zeros((nb, ny, nx, nz, ci))
    if kind == "down":
        oy, ox, oz = ny // 2, nx // 2, nz // 2
        for d, (dy, dx, dz) in enumerate(taps("down")):
            dx_[:, dy:2 * oy:2, dx:2 * ox:2, dz:2 * oz:2, :] += t[..., d * ci:(d + 1) * ci]
        return dx_
    pad = np.zeros((nb, ny + 2, nx + 2, nz + 2, ci))
    for d, (dy, dx, dz) in enumerate(taps("same")):
        pad[:, 1 + dy:1 + dy + ny, 1 + dx:1 + dx + nx, 1 + dz:1 + dz + nz, :] += t[..., d * ci:(d + 1) * ci]
    return pad[:, 1:-1, 1:-1, 1:-1, :]


def conv_bwd_weight(kind, x, gpre):
    """(dW, db) in float64."""
    x = np.asarray(x, np.float64)
    g = np.asarray(gpre, np.float64)
    co = g.shape[-1]
    db = g.reshape(-1, co).sum(0)
    if kind == "up":
        ci = x.shape[-1]
        cols = [g[:, dy::2, dx::2, dz::2, :] for dy, dx, dz in taps("up")]
        gt = np.concatenate(cols, axis=-1).reshape(-1, 8 * co)
        return gt.T @ x.reshape(-1, ci), db
    A = _gather(x, kind).reshape(-1, len(taps(kind)) * x.shape[-1])
    return g.reshape(-1, co).T @ A, db


def block_order_offsets():
    """Voxel offset (y, x, z) inside an 8^3 block of each of the 512 inputs of
    the third encoder layer, k = (t2*8 + t1)*8 + t0, t = (dy*2 + dx)*2 + dz."""
    out = []
    for k in range(512):
        t2, t1, t0 = k >> 6, (k >> 3) & 7, k & 7
        b = lambda t, s: (t >> s) & 1
        out.append((4 * b(t2, 2) + 2 * b(t1, 2) + b(t0, 2), 4 * b(t2, 1) + 2 * b(t1, 1) + b(t0, 1),
                    4 * b(t2, 0) + 2 * b(t1, 0) + b(t0, 0)))
    return out


def pca_params(V):
    """Weights of the network whose encoder projects every 8^3 block on the
    rows of V (64 x 512, block order of block_order_offsets) and whose
    decoder reconstructs from them; latent convolutions identity."""
    p = []
    for kind, ci, co, a in NET:
        p.append(np.zeros(kdim(kind, ci, co), np.float32))
    p[0][:] = np.eye(8)
    p[1][:] = np.eye(64)
    p[2][:] = V
    for i in (3, 4):
        p[i][np.arange(64), 13 * 64 + np.arange(64)] = 1.0
    p[5][:] = V.T
    p[6][:] = np.eye(64)
    p[7][:] = np.eye(8)
    return p


def init_params(seed=0):
    """He-style initial fp32 master weights, one matrix per layer."""
    rng = np.random.default_rng(seed)
    params = []
    for kind, ci, co, a in NET:
        rows, cols = kdim(kind, ci, co)
        fan_in = cols if kind != "up" else ci
        params.append((rng.standard_normal((rows, cols)) * np.sqrt(2.0 / fan_in)).astype(np.float32))
    return params


def forward(params, x):
    """Full network on x (nb, n, n, n, 1) fp32: list of stored layer
    outputs (bf16 values; the last one fp32) and the latent."""
    h = bf16(x)
    outs = [h]
    for li, (kind, ci, co, a) in enumerate(NET):
        y = conv_fwd(kind, h, bf16(params[li]), np.zeros(co), a).astype(np.float32)
        h = y if li == len(NET) - 1 else bf16(y)
        outs.append(h)
    return outs


def loss_and_grads(params, x):
    """MSE loss (fp32 target x) and fp32 gradients of every weight matrix,
    following the GPU's rounding points: dL/dout, and every activation
    gradient handed to the next layer down, are rounded to bf16."""
    outs = forward(params, x)
    y = outs[-1].astype(np.float64)
    n = y.size
    diff = y - np.asarray(x, np.float64)
    loss = float((diff * diff).sum() / n)
    g = bf16((2.0 * diff / n).astype(np.float32))
    grads = [None] * len(NET)
    for li in range(len(NET) - 1, -1, -1):
        kind, ci, co, a = NET[li]
        gpre = g                                    # act' already folded into g below
        dW, _ = conv_bwd_weight(kind, outs[li], gpre)
        grads[li] = dW.astype(np.float32)
        if li > 0:
            dx = conv_bwd_data(kind, gpre, bf16(params[li]), outs[li].shape)
            pa = NET[li - 1][3]
            g = bf16(act_grad(outs[li], dx, pa).astype(np.float32))
    return loss, grads


def adam(w, m, v, g, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (fp32 state, t = step count from 1); returns new
    (w, m, v) in float32 (computed in float64, rounded once)."""
    w = np.asarray(w, np.float64); m = np.asarray(m, np.float64)
    v = np.asarray(v, np.float64); g = np.asarray(g, np.float64)
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    w = w - lr * mh / (np.sqrt(vh) + eps)
    return w.astype(np.float32), m.astype(np.float32), v.astype(np.float32)
