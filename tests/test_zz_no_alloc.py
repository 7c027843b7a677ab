"""vti_bc_encode_noalloc (the codec's encode with its scratch in a static
device array, no allocation) through the C ABI against vti_bc_encode: the
same entries per block, entry count, error bound and decoded box, on two
calls in a row (a 128^3 box, the maximum, then a smaller 64^3 box of
another array, so the second call starts from the first call's scratch),
and the maximum box side enforced."""
import ctypes as C

import numpy as np
import pytest

from geophysics_amd import vti

pytestmark = pytest.mark.gpu


def _orthonormal(seed):
    q, r = np.linalg.qr(np.random.default_rng(seed).standard_normal((512, 512)))
    return (q * np.sign(np.diag(r))).astype(np.float32)


def _encode(torch, A, fn, Pd, off, sy, sx, n, tol, floor):
    dev = torch.device("cuda:0")
    nblk = (n // 8) ** 3
    cap = n ** 3
    ent = torch.zeros(cap, dtype=torch.int32, device=dev)
    boff = torch.zeros(nblk, dtype=torch.int32, device=dev)
    bcnt = torch.zeros(nblk, dtype=torch.int16, device=dev)
    torch.cuda.synchronize()
    kept, err = fn(Pd.data_ptr(), off, sy, sx, n, tol, ent.data_ptr(), cap, boff.data_ptr(), bcnt.data_ptr(),
                   floor=floor)
    out = torch.zeros(n ** 3, dtype=torch.float32, device=dev)
    A.bc_decode(ent.data_ptr(), boff.data_ptr(), bcnt.data_ptr(), n, out.data_ptr())
    torch.cuda.synchronize()
    e = ent.cpu().numpy().view(np.uint32)
    bo = boff.cpu().numpy().view(np.uint32).astype(np.int64)
    bc = bcnt.cpu().numpy().view(np.uint16).astype(np.int64)
    blocks = [e[bo[b]:bo[b] + bc[b]] for b in range(nblk)]
    return kept, err, blocks, out.cpu().numpy()


def test_noalloc_encode_equals_encode_on_two_calls():
    import torch
    rng = np.random.default_rng(5)
    with vti.Autoencoder(crop=16, batch=1, seed=1) as A:
        A.set_basis(_orthonormal(3))
        for call, (n, pad, tol, floor) in enumerate(((128, (3, 5, 7), 0.01, 1e-8), (64, (9, 2, 4), 0.05, 1e-4))):
            big = rng.standard_normal((n + pad[0], n + pad[1], n + pad[2])).astype(np.float32)
            big[:, :, : n // 2] *= 1e-3                     # weak blocks too: the floor matters
            Pd = torch.from_numpy(big).to("cuda:0")
            ny, nx, nz = big.shape
            off = 2 + 1 * nz + 1 * nx * nz
            ref = _encode(torch, A, A.bc_encode, Pd, off, nx * nz, nz, n, tol, floor)
            got = _encode(torch, A, A.bc_encode_noalloc, Pd, off, nx * nz, nz, n, tol, floor)
            print(f"call {call}: n {n}, kept {got[0]} (encode {ref[0]}), error bound {got[1]:.4e}", flush=True)
            assert 0 < ref[0] < n ** 3
            # the bound sums block energies with atomics (order varies): equal to rounding
            assert got[0] == ref[0] and abs(got[1] - ref[1]) <= 1e-12 * max(1.0, ref[1])
            for b, (x, y) in enumerate(zip(got[2], ref[2])):
                assert np.array_equal(np.sort(x), np.sort(y)), (call, b)
            assert np.array_equal(got[3], ref[3])
            del Pd


def test_noalloc_encode_rejects_boxes_above_the_maximum():
    import torch
    with vti.Autoencoder(crop=16, batch=1, seed=1) as A:
        A.set_basis(_orthonormal(4))
        n = 136
        Pd = torch.zeros(n ** 3, dtype=torch.float32, device="cuda:0")
        ent = torch.zeros(16, dtype=torch.int32, device="cuda:0")
        boff = torch.zeros((n // 8) ** 3, dtype=torch.int32, device="cuda:0")
        bcnt = torch.zeros((n // 8) ** 3, dtype=torch.int16, device="cuda:0")
        with pytest.raises(vti.VtiError, match="maximum"):
            A.bc_encode_noalloc(Pd.data_ptr(), 0, n * n, n, n, 0.01, ent.data_ptr(), 16, boff.data_ptr(),
                                bcnt.data_ptr())
