"""Pins for the oracle's RTM pieces that no reference run can produce
(the reference RTM binary needs CUDA): closed forms of the adjoint
Laplacian and the imaging geometry of a flat reflector."""
import numpy as np

from oracle import pyoracle as po


def test_laplace_adj_closed_forms():
    # constant field -> 0 everywhere (each cell gets deg*c - sum of neighbours)
    c = np.full((7, 9), 2.5, np.float32)
    assert np.abs(po.laplace_adj(c)).max() == 0.0
    # delta -> degree at the centre, -1 on its neighbours
    d = np.zeros((7, 9), np.float32)
    d[3, 4] = 1.0
    out = po.laplace_adj(d)
    assert out[3, 4] == 4.0
    assert out[2, 4] == out[4, 4] == out[3, 3] == out[3, 5] == -1.0
    assert np.count_nonzero(out) == 5
    # corner delta has degree 2; 3D interior delta has degree 6
    d[:] = 0
    d[0, 0] = 1
    assert po.laplace_adj(d)[0, 0] == 2.0
    d3 = np.zeros((5, 6, 7), np.float32)
    d3[2, 3, 3] = 1
    o3 = po.laplace_adj(d3)
    assert o3[2, 3, 3] == 6.0 and o3.sum() == 0.0
    # it is the transpose of the forward 5-point operator: <L^T x, y> = <x, L y>
    rng = np.random.default_rng(0)
    x = rng.standard_normal((6, 8)).astype(np.float32)
    y = rng.standard_normal((6, 8)).astype(np.float32)
    def fwd(a):   # forward form (adj=0) of the same operator
        o = np.zeros_like(a, dtype=np.float64)
        nx, nz = a.shape
        for ix in range(nx):
            for iz in range(nz):
                for jx, jz in ((ix, iz - 1), (ix, iz + 1), (ix - 1, iz), (ix + 1, iz)):
                    if 0 <= jx < nx and 0 <= jz < nz:
                        o[ix, iz] += a[ix, iz] - a[jx, jz]
        return o
    lhs = float(np.sum(po.laplace_adj(x).astype(np.float64) * y))
    rhs = float(np.sum(x.astype(np.float64) * fwd(y)))
    assert abs(lhs - rhs) <= 1e-4 * max(1.0, abs(rhs))


def test_rtm_images_a_flat_reflector_at_its_depth():
    """Two-layer ISO model, reflector at z = 40 cells: the migrated image
    (divided by illumination, Laplacian-filtered) peaks at that depth."""
    nx, nz, npd, nt = 120, 80, 20, 900
    zr = 40
    vp = np.full((nx, nz), 2000.0, np.float32)
    vp[:, zr:] = 2800.0
    z = np.zeros_like(vp)
    p = po.params_gpu2d(nx, nz, 10, 10, npd, 0, 20, 1, 60, 0, 2, nt, 1000, 1)
    # observed data: the two-layer response minus the direct wave of a
    # homogeneous model (removes the direct arrival exactly)
    obs = po.Sim(p, vp, z, z).run()
    obs -= po.Sim(p, np.full_like(vp, 2000.0), z, z).run()
    # migrate in the smooth (homogeneous upper-layer) velocity
    img, ill, _ = po.rtm(p, np.full_like(vp, 2000.0), z, z, sv_fwd=0, sv_bwd=0, illum_mode=0, obs=obs)
    fin = po.laplace_adj(img / ill)
    prof = np.abs(fin[30:90, :]).sum(axis=0)
    peak = int(np.argmax(prof[10:])) + 10
    assert abs(peak - zr) <= 2, peak
