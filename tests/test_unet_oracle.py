"""The CPU oracle of the warm-start U-Net (oracle/unet_ref.py) is pinned by central finite
differences in float64, and its parameter layout matches the library's count."""
import numpy as np

from oracle.unet_ref import UNetRef, layout


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    L, C, cin = 3, 2, 2
    _, n = layout(L, C, cin)
    p = rng.normal(0, 0.5, n)
    x = rng.normal(size=(2, 8, 8, cin))
    t = rng.normal(size=(2, 8, 8))
    m = (rng.random((2, 8, 8)) > 0.3).astype(float)
    _, g = UNetRef(L, C, cin, p, np.float64).loss_grad(x, t, m)
    eps = 1e-6
    idx = rng.choice(n, 300, replace=False)
    fd = np.zeros(len(idx))
    for i, k in enumerate(idx):
        pp, pm = p.copy(), p.copy()
        pp[k] += eps
        pm[k] -= eps
        fd[i] = (UNetRef(L, C, cin, pp, np.float64).loss_grad(x, t, m)[0]
                 - UNetRef(L, C, cin, pm, np.float64).loss_grad(x, t, m)[0]) / (2 * eps)
    assert np.linalg.norm(fd - g[idx]) <= 1e-6 * np.linalg.norm(g[idx])


def test_layout_counts():
    # encoder 2 convs per level, decoder 2 per level below the bottom, one output conv
    convs, n = layout(4, 32, 4)
    assert len(convs) == 2 * 4 + 2 * 3 + 1
    assert n == sum(9 * ci * co + co for _, ci, co, _, _ in convs)
