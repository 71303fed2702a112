"""The warm-start network's GPU kernels against the CPU oracle: the bf16 matrix-core product
against numpy on bf16-rounded operands, and the U-Net's loss and parameter gradients
(forward and backward through every layer type) against oracle/unet_ref.py.  Tolerances:
products of bf16-rounded operands accumulated in fp32 agree to 1e-4 relative; the network
rounds every convolution input and gradient to bf16, so loss and gradients agree to 2 % and
5 % relative (norm-wise)."""
import numpy as np
import pytest

import randomtilings_amd as rt
from oracle.unet_ref import UNetRef

pytestmark = pytest.mark.gpu


def bf16(a):
    u = np.ascontiguousarray(a, np.float32).view(np.uint32)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.view(np.float32)


@pytest.fixture(scope="module")
def nctx():
    c = rt.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("layout", [0, 1, 2])
@pytest.mark.parametrize("M,N,K", [(64, 64, 32), (100, 70, 45), (32, 288, 9000), (5000, 32, 288), (7, 3, 1),
                                   (264, 136, 104), (1024, 512, 2056), (8, 8, 8)])
def test_gemm_against_numpy(nctx, M, N, K, layout):
    """Products through both kernels (128 x 128 tiles when every operand vector is aligned and
    whole, the general kernel otherwise), every operand layout, split K, partial tiles.  The
    tolerance is that of fp32 accumulation of bf16 products."""
    rng = np.random.default_rng(M * 7 + N + layout)
    A = rng.normal(size=(M, K)).astype(np.float32)
    B = rng.normal(size=(K, N)).astype(np.float32)
    print("gemm", M, N, K, layout, flush=True)
    C = rt.nn_gemm(nctx, A.T.copy() if layout == 1 else A, B.T.copy() if layout == 2 else B, layout)
    ref = bf16(A).astype(np.float64) @ bf16(B).astype(np.float64)
    assert np.abs(C - ref).max() <= 1e-4 * np.abs(ref).max() + 1e-5


@pytest.mark.parametrize("L,C,cin,H", [(2, 8, 3, 8), (3, 8, 4, 16), (4, 16, 4, 32)])
def test_unet_gradients_against_oracle(nctx, L, C, cin, H):
    rng = np.random.default_rng(L)
    net = rt.UNet(nctx, L, C, cin, seed=3)
    try:
        p = net.get_params()
        x = rng.normal(size=(2, H, H, cin)).astype(np.float32)
        t = rng.normal(size=(2, H, H)).astype(np.float32)
        m = (rng.random((2, H, H)) > 0.2).astype(np.float32)
        print("unet", L, C, cin, H, flush=True)
        loss, g = net.loss_grad(x, t, m)
        ref = UNetRef(L, C, cin, p, np.float64)
        rloss, rg = ref.loss_grad(x.astype(np.float64), t, m)
        y = net.forward(x)
        ry, _ = ref.forward(x.astype(np.float64))
        assert np.linalg.norm(y - ry) <= 2e-2 * np.linalg.norm(ry)
        assert abs(loss - rloss) <= 2e-2 * rloss
        assert np.linalg.norm(g - rg) <= 5e-2 * np.linalg.norm(rg)
    finally:
        net.close()


def test_unet_training_fits_a_batch(nctx):
    """Adam on one fixed batch drives the loss down by two orders of magnitude."""
    rng = np.random.default_rng(9)
    net = rt.UNet(nctx, 3, 8, 2, seed=1)
    try:
        H = 16
        x = rng.normal(size=(4, H, H, 2)).astype(np.float32)
        yy, xx = np.meshgrid(np.linspace(-1, 1, H), np.linspace(-1, 1, H), indexing="ij")
        t = np.broadcast_to((xx * xx - yy) * 0.5, (4, H, H)).astype(np.float32) + 0.1 * x[..., 0]
        m = np.ones((4, H, H), np.float32)
        first = net.train_step(x, t, m, 3e-3)
        for _ in range(300):
            last = net.train_step(x, t, m, 3e-3)
        assert last < first / 100
    finally:
        net.close()
