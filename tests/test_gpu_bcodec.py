"""Sparse block-transform codec (vti_bc_encode / vti_bc_decode) and the 3D
RTM that stores its source wavefield with it, against numpy (float64) and
the plain checkpointed RTM imaging the same steps."""
import numpy as np
import pytest

from conftest import rel_l2, smooth_model3d
from geophysics_amd import vti

pytestmark = pytest.mark.gpu

N, NPML, NT = 32, 6, 200


def _say(*a):
    print(*a, flush=True)


def _blocks(x):
    """(n^3) box -> [(n/8)^3 blocks][512], blocks z fastest, voxels (ty*8+tx)*8+tz."""
    nb = x.shape[0] // 8
    return x.reshape(nb, 8, nb, 8, nb, 8).transpose(0, 2, 4, 1, 3, 5).reshape(nb ** 3, 512)


def _unblocks(b, n):
    nb = n // 8
    return b.reshape(nb, nb, nb, 8, 8, 8).transpose(0, 3, 1, 4, 2, 5).reshape(n, n, n)


def _entry_value(e):
    """Top 23 bits of an entry = the kept value's fp32 pattern."""
    return (e & np.uint32(0xFFFFFE00)).view(np.float32)


def _orthonormal(seed):
    q, r = np.linalg.qr(np.random.default_rng(seed).standard_normal((512, 512)))
    return (q * np.sign(np.diag(r))).astype(np.float32)


def _field(seed=7, steps=120):
    vp, e, d = smooth_model3d(N, N, N, seed=seed)
    cfg = vti.cfg_ref3d(N, N, N, 10, 10, 10, NPML, 1, 25, NT, 800)
    with vti.Propagator(cfg, vp, e, d) as prop:
        prop.set_shot(vti.shot_ref3d(cfg, 0, 12, 0, 20, 0, 1, 0, rec_depth=1, rec_step=2))
        prop.step(0, steps)
        return prop.field("P")      # padded (nny, nnx, nnz)


SENTINEL = 0xDEADBEEF


def _codec(A, torch, P_pad, n, tol, cap, off_yxz=(NPML, NPML, NPML), floor=1e-8, guard=0):
    """Encode the n^3 box of the padded array at off_yxz and decode it.  The
    entry buffer holds cap entries and `guard` more, all SENTINEL before the
    call; the returned entries include the guard."""
    dev = torch.device("cuda:0")
    Pd = torch.from_numpy(np.ascontiguousarray(P_pad)).to(dev)
    ny, nx, nz = P_pad.shape
    off = off_yxz[2] + off_yxz[1] * nz + off_yxz[0] * nx * nz
    nblk = (n // 8) ** 3
    ent = torch.full((max(cap, 1) + guard,), SENTINEL - 2 ** 32, dtype=torch.int32, device=dev)
    boff = torch.zeros(nblk, dtype=torch.int32, device=dev)
    bcnt = torch.zeros(nblk, dtype=torch.int16, device=dev)
    torch.cuda.synchronize()         # the codec runs on the autoencoder's own stream
    kept, err = A.bc_encode(Pd.data_ptr(), off, nx * nz, nz, n, tol, ent.data_ptr(), cap, boff.data_ptr(),
                            bcnt.data_ptr(), floor=floor)
    out = None
    if kept <= cap:
        o = torch.zeros(n ** 3, dtype=torch.float32, device=dev)
        torch.cuda.synchronize()
        A.bc_decode(ent.data_ptr(), boff.data_ptr(), bcnt.data_ptr(), n, o.data_ptr())
        out = o.cpu().numpy().reshape(n, n, n)
    torch.cuda.synchronize()
    return kept, err, out, ent.cpu().numpy().view(np.uint32), boff.cpu().numpy().view(np.uint32), \
        bcnt.cpu().numpy().view(np.uint16)


@pytest.mark.parametrize("tol,floor", [(0.0, 1e-8), (0.01, 1e-8), (0.05, 1e-8), (0.2, 1e-8), (0.05, 1e-2)])
def test_codec_error_bound_and_entries(tol, floor):
    """Random orthonormal basis, an FD snapshot: every block's decoded error
    is within tol of max(its norm, sqrt(floor) x the strongest block's), the
    whole box within the reported bound, the entries of a block are its
    largest coefficients (the float64 projection to 2^-15), and the decode
    equals numpy's reconstruction from the same entries."""
    import torch
    V = _orthonormal(11)
    P = _field()
    box = P[NPML:NPML + N, NPML:NPML + N, NPML:NPML + N].astype(np.float64)
    with vti.Autoencoder(crop=16, batch=1, seed=1) as A:
        A.set_basis(V)
        assert np.array_equal(A.basis(), V)
        kept, err, out, ent, boff, bcnt = _codec(A, torch, P, N, tol, N ** 3, floor=floor)
    e = rel_l2(out, box)
    _say(f"tol {tol} floor {floor}: kept {kept} of {N ** 3}, bound {err:.4e}, error {e:.4e}")
    assert e <= err * (1 + 1e-3) + 1e-5
    assert int(bcnt.astype(np.int64).sum()) == kept
    xb = _blocks(box)
    eb = (xb ** 2).sum(axis=1)
    ref = np.maximum(eb, floor * eb.max())
    assert err <= max(tol, 2.0 ** -13) * np.sqrt(ref.sum() / eb.sum()) * (1 + 1e-5)
    errb = ((_blocks(out.astype(np.float64)) - xb) ** 2).sum(axis=1)
    assert np.all(errb <= max(tol * tol, 2.0 ** -26) * ref * 1.01 + 1e-9 * eb + 1e-30)
    # entries against the float64 projection
    c = xb @ V.astype(np.float64)
    bnorm = np.sqrt(eb)
    dense = np.zeros_like(c)
    ranges = []
    worst = 0.0
    for b in range(c.shape[0]):
        s0, k = int(boff[b]), int(bcnt[b])
        ranges.append((s0, s0 + k))
        ee = ent[s0:s0 + k]
        idx = (ee & 511).astype(np.int64)
        assert len(np.unique(idx)) == k
        vals = _entry_value(ee).astype(np.float64)
        dense[b, idx] = vals
        if not k:
            continue
        # value rounding (<= 2^-15 of the value) plus the transform's own
        # error, bounded by ~2^-16 of the block's norm (hi/lo splits)
        dev = np.abs(vals - c[b, idx]) - 2.0 ** -15 * np.abs(c[b, idx])
        worst = max(worst, dev.max() / max(bnorm[b], 1e-30))
        if k < 512:                          # a magnitude cut, up to the transform's error
            keep = np.zeros(512, bool)
            keep[idx] = True
            slack = 4e-5 * bnorm[b]
            assert np.abs(c[b, keep]).min() + slack >= np.abs(c[b, ~keep]).max() - slack
    _say(f"  largest coefficient deviation beyond the value rounding: {worst:.3e} of the block norm")
    assert worst < 4e-5
    ranges.sort()
    for (a0, a1), (b0, _) in zip(ranges, ranges[1:]):
        assert a1 <= b0                      # block ranges do not overlap
    rec = _unblocks(dense @ V.astype(np.float64).T, N)
    assert rel_l2(out, rec) < 1e-4


def test_codec_strided_box_and_capacity():
    """A box at an offset of a larger array (strides), and a capacity below
    what is needed: nothing past it is written."""
    import torch
    V = _orthonormal(5)
    rng = np.random.default_rng(3)
    n = 16
    big = rng.standard_normal((n + 5, n + 3, n + 9)).astype(np.float32)
    with vti.Autoencoder(crop=16, batch=1, seed=1) as A:
        A.set_basis(V)
        kept, err, out, ent, _, _ = _codec(A, torch, big, n, 0.1, n ** 3, off_yxz=(2, 1, 5))
        box = big[2:2 + n, 1:1 + n, 5:5 + n]
        assert rel_l2(out, box) <= err + 1e-6 and err <= 0.1
        cap, guard = kept - 1, 4096
        kept2, _, out2, ent2, boff2, bcnt2 = _codec(A, torch, big, n, 0.1, cap, off_yxz=(2, 1, 5), guard=guard)
    assert kept2 == kept and out2 is None
    assert len(ent2) == cap + guard
    assert np.all(ent2[cap:] == SENTINEL)    # nothing written past cap
    ends = boff2.astype(np.int64) + bcnt2.astype(np.int64)
    fit = ends <= cap
    assert fit.any()
    for b in np.nonzero(fit)[0]:             # the blocks that fit hold their entries
        e = ent2[boff2[b]:ends[b]]
        assert not np.any(e == SENTINEL)


@pytest.mark.parametrize("n", [24, 8])
def test_codec_block_counts_off_the_pack_workgroup(n):
    """27 and 1 blocks: the pack's last workgroup (8 blocks) has idle
    wavefronts.  Every block is within its bound, the entries and the block
    table cover exactly the kept count, and the decode is within the
    reported error."""
    import torch
    V = _orthonormal(9)
    rng = np.random.default_rng(n)
    big = rng.standard_normal((n + 5, n + 3, n + 9)).astype(np.float32)
    tol, floor = 0.05, 1e-8
    with vti.Autoencoder(crop=16, batch=1, seed=1) as A:
        A.set_basis(V)
        kept, err, out, ent, boff, bcnt = _codec(A, torch, big, n, tol, n ** 3, off_yxz=(2, 1, 5), floor=floor)
    box = big[2:2 + n, 1:1 + n, 5:5 + n].astype(np.float64)
    assert 0 < kept <= n ** 3 and err <= tol
    assert int(np.asarray(bcnt, dtype=np.int64).sum()) == kept
    starts = np.sort(np.asarray(boff, dtype=np.int64))
    assert starts[0] == 0 and starts[-1] <= kept
    assert rel_l2(out, box) <= err + 1e-6
    nb = n // 8
    for i in range(nb ** 3):
        bz, bx, by = i % nb, (i // nb) % nb, i // (nb * nb)
        sl = np.s_[8 * by:8 * by + 8, 8 * bx:8 * bx + 8, 8 * bz:8 * bz + 8]
        e_b = np.linalg.norm(out[sl] - box[sl])
        assert e_b <= tol * np.linalg.norm(box[sl]) * (1 + 1e-3) + 1e-6


def _setup(seed=7):
    vp, e, d = smooth_model3d(N, N, N, seed=seed)
    cfg = vti.cfg_ref3d(N, N, N, 10, 10, 10, NPML, 1, 25, NT, 800)
    return cfg, vp, e, d


def _shot(cfg, cx, cy):
    return vti.shot_ref3d(cfg, 0, cx, 0, cy, 0, 1, 0, rec_depth=1, rec_step=2)


def test_sparse_rtm_image_error_and_fallback():
    """Basis fitted on two shots, a third shot imaged from sparse
    coefficients at tol 1 % and 0.2 %, against the plain RTM image of the same
    steps; a budget below the need with checkpoints recomputes every
    snapshot (bit-identical image), without checkpoints it fails."""
    cfg, vp, e, d = _setup()
    with vti.Propagator(cfg, vp, e, d) as prop, vti.Autoencoder(crop=32, batch=1, seed=2) as A:
        for cx, cy in ((10, 12), (22, 20)):
            prop.set_shot(_shot(cfg, cx, cy))
            prop.ae_fit_shot(A, save_every=4, nblocks=64)
        A.fit_end()
        V = A.basis().astype(np.float64)
        assert np.allclose(V.T @ V, np.eye(512), atol=1e-5)
        prop.set_shot(_shot(cfg, 16, 16))
        o = prop.rtm_defaults()
        o.image_every, o.nsnap = 4, 6
        prop.rtm_reset()
        prop.rtm_shot(o, None)
        img_p, ill_p = prop.rtm_image(finalize=False)
        fin_p, _ = prop.rtm_image(finalize=True)
        res = {}
        for tol in (0.01, 0.002):
            lo = vti.learned_defaults()
            lo.codec, lo.tol = 1, tol
            lo.budget = (NT // 4) * N ** 3 * 4   # room for every coefficient: no recomputation
            prop.rtm_reset()
            st = prop.rtm_shot_learned(A, prop.rtm_defaults(), lo, None)
            img_l, ill_l = prop.rtm_image(finalize=False)
            fin_l, _ = prop.rtm_image(finalize=True)
            assert np.array_equal(ill_l, ill_p)
            res[tol] = (rel_l2(img_l, img_p), rel_l2(fin_l, fin_p), st)
            _say(f"tol {tol}: raw image {res[tol][0]:.3e}, final image {res[tol][1]:.3e}, {st}")
            assert st["recomputed"] == 0 and st["snapshots"] == NT // 4
            assert st["mean_checked_err"] <= tol
        lo = vti.learned_defaults()      # room for one entry: every non-zero snapshot is recomputed
        lo.codec, lo.tol, lo.budget, lo.ck_every = 1, 0.0, 4, 16
        prop.rtm_reset()
        st = prop.rtm_shot_learned(A, prop.rtm_defaults(), lo, None)
        img_r, _ = prop.rtm_image(finalize=False)
        lo.ck_every = 0
        prop.rtm_reset()
        with pytest.raises(vti.VtiError):
            prop.rtm_shot_learned(A, prop.rtm_defaults(), lo, None)
    assert res[0.01][0] < 5e-3 and res[0.01][1] < 3e-2       # measured 1.1e-3, 2.0e-2
    assert res[0.002][0] < 1e-3 and res[0.002][1] < 1e-2     # measured 2.1e-4, 5.1e-3
    assert res[0.002][0] < res[0.01][0] and res[0.002][1] < res[0.01][1]
    assert st["recomputed"] > 0
    assert np.array_equal(img_r, img_p), rel_l2(img_r, img_p)


def test_fitted_basis_keeps_fewer_coefficients_than_untrained_bases():
    """What the learning gains: the basis fitted on two shots' snapshots
    needs fewer coefficients for the same per-block error bound (1 %) on a
    third shot's snapshot than the identity basis (the raw samples) or a
    random orthonormal basis."""
    import torch
    cfg, vp, e, d = _setup()
    with vti.Propagator(cfg, vp, e, d) as prop, vti.Autoencoder(crop=32, batch=1, seed=5) as A:
        for cx, cy in ((10, 12), (22, 20)):
            prop.set_shot(_shot(cfg, cx, cy))
            prop.ae_fit_shot(A, save_every=4, nblocks=64, src_box=0)
        A.fit_end()
        fitted = A.basis()
        prop.set_shot(_shot(cfg, 16, 9))
        prop.reset()
        prop.step(0, 150)
        P = prop.field("P")
        kept = {}
        for name, V in (("fitted", fitted), ("identity", np.eye(512, dtype=np.float32)),
                        ("random", _orthonormal(3))):
            A.set_basis(V)
            k, err, out, *_ = _codec(A, torch, P, N, 0.01, N ** 3)
            box = P[NPML:NPML + N, NPML:NPML + N, NPML:NPML + N]
            assert rel_l2(out, box) <= err * (1 + 1e-3) + 1e-6
            kept[name] = k
    _say(f"coefficients kept at 1 % per block: {kept}")
    assert kept["fitted"] < kept["identity"] and kept["fitted"] < kept["random"]
