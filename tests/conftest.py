import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libvti.so")
    config.addinivalue_line("markers", "slow: longer CPU oracle runs")
    # Tests that hand torch device buffers to libvti.so need torch's HIP
    # runtime initialised before the library is loaded (one runtime per
    # process, the one torch ships), as bench.py does, whichever tests are
    # selected; on a machine without a device this is a no-op.
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.init()
    except Exception:
        pass


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def load_golden(name):
    d = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    return {k: d[k] for k in d.files}


def layered_model(nx, nz, seed, nlayer=4, vlo=1500.0, vhi=4000.0, elo=0.01, ehi=0.2,
                  dlo=0.01, dhi=0.1):
    """Flat layers (x-major, z fastest) as in the 2D benchmark configuration."""
    rng = np.random.default_rng(seed)
    vpl = np.sort(rng.uniform(vlo, vhi, nlayer))
    el = np.sort(rng.uniform(elo, ehi, nlayer))
    dl = np.sort(rng.uniform(dlo, dhi, nlayer))
    lay = np.minimum(np.arange(nz) * nlayer // nz, nlayer - 1)
    f = lambda v: np.tile(v[lay], (nx, 1)).astype(np.float32)
    return f(vpl), f(el), f(dl)


def smooth_model3d(nx, ny, nz, seed, vtop=1800.0, vbot=4200.0, noise=150.0, sigma=3.0,
                   emax=0.25, dmax=0.12):
    """Linear-gradient vp plus smoothed noise; smoothed eps/delta. Layout (y, x, z)."""
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(seed)
    z = np.linspace(vtop, vbot, nz, dtype=np.float64)
    n = gaussian_filter(rng.standard_normal((ny, nx, nz)), sigma)
    n /= max(np.abs(n).max(), 1e-12)
    vp = (z[None, None, :] + noise * n).astype(np.float32)

    def mm(a, hi):
        a = gaussian_filter(a, sigma)
        a = (a - a.min()) / max(a.max() - a.min(), 1e-12)
        return (hi * a).astype(np.float32)
    e = mm(rng.standard_normal((ny, nx, nz)), emax)
    d = mm(rng.standard_normal((ny, nx, nz)), dmax)
    return vp, e, d


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
