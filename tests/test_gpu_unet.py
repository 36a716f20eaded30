"""GPU checks of the U-Net kernels (learned warm start).

* one MFMA convolution (forward, with the fused upsample-concat input) and
  its weight gradient against the FP64 oracle on BF16-rounded operands
  (tolerance: 1% of the output scale for BF16 outputs, 1e-4 relative for the
  FP32-accumulated weight gradient);
* the whole network's backward pass against central finite differences of
  its own forward pass, layer by layer (5%: the forward runs in BF16;
  cubic fit over 8 steps);
* a short training run lowers the loss.
"""
import numpy as np
import pytest

import unet_oracle as U

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("C1,C2,Cout", [(16, 0, 16), (16, 0, 32), (32, 0, 64), (64, 0, 64), (32, 64, 32),
                                        (16, 32, 16), (64, 0, 96), (32, 0, 48)])
def test_conv3_forward_vs_oracle(ctx, C1, C2, Cout):
    import arec3d_amd as A

    print(f"\n[case] conv3 {C1}+{C2}->{Cout}", flush=True)
    rng = np.random.default_rng(C1 * 100 + C2 + Cout)
    D, H, W = 6, 8, 80
    a = U.bf16_round(rng.standard_normal((D, H, W, C1)))
    b = U.bf16_round(rng.standard_normal((D // 2, H // 2, W // 2, C2))) if C2 else None
    w = U.bf16_round(rng.standard_normal((27, Cout, C1 + C2)) / np.sqrt(27 * (C1 + C2)))
    bias = rng.standard_normal(Cout).astype(np.float32)
    ref = U.conv3(a, w, bias, in2=b, relu=True)
    da, dw, db, do = ctx.to_device(a), ctx.to_device(w), ctx.to_device(bias), ctx.empty((D, H, W, Cout))
    d2 = ctx.to_device(b) if C2 else None
    A._check(A.lib.arec_conv3_check(ctx.h, da.ptr, C1, None if d2 is None else d2.ptr, C2, dw.ptr, db.ptr, do.ptr,
                                    Cout, D, H, W, 1))
    got = do.download()
    assert np.abs(got - ref).max() <= 1e-2 * np.abs(ref).max()


@pytest.mark.parametrize("C1,C2,Cout", [(16, 0, 16), (16, 32, 16), (64, 0, 64), (32, 0, 48)])
@pytest.mark.parametrize("D,H,W", [(4, 10, 40), (2, 6, 70), (2, 2, 16), (36, 6, 70)])
def test_conv3_forward_row_blocks(ctx, C1, C2, Cout, D, H, W):
    """Heights that are not a multiple of the kernel's rows per workgroup
    (4, or 2 at 64+ output channels): the last block's extra rows are not
    written (arec_conv3_check fails if a guard past the output changes).
    D = 36 spans three of the 16-output kernel's runs of 16 D slices (its
    input ring and half-resolution ring restart at slices 16 and 32)."""
    import arec3d_amd as A

    print(f"\n[case] conv3 row blocks {C1}+{C2}->{Cout} {D}x{H}x{W}", flush=True)
    rng = np.random.default_rng(D * 1000 + H * 10 + Cout)
    a = U.bf16_round(rng.standard_normal((D, H, W, C1)))
    b = U.bf16_round(rng.standard_normal((D // 2, H // 2, W // 2, C2))) if C2 else None
    w = U.bf16_round(rng.standard_normal((27, Cout, C1 + C2)) / np.sqrt(27 * (C1 + C2)))
    bias = rng.standard_normal(Cout).astype(np.float32)
    ref = U.conv3(a, w, bias, in2=b, relu=False)
    da, dw, db = ctx.to_device(a), ctx.to_device(w), ctx.to_device(bias)
    do = ctx.empty((D, H, W, Cout))
    d2 = ctx.to_device(b) if C2 else None
    A._check(A.lib.arec_conv3_check(ctx.h, da.ptr, C1, None if d2 is None else d2.ptr, C2, dw.ptr, db.ptr, do.ptr,
                                    Cout, D, H, W, 0))
    assert np.abs(do.download() - ref).max() <= 1e-2 * np.abs(ref).max()


@pytest.mark.parametrize("C1,C2,Cout", [(16, 0, 16), (32, 0, 64), (32, 64, 32), (16, 32, 16)])
def test_conv3_wgrad_vs_oracle(ctx, C1, C2, Cout):
    import arec3d_amd as A

    print(f"\n[case] wgrad {C1}+{C2}->{Cout}", flush=True)
    rng = np.random.default_rng(7 + C1 + C2 + Cout)
    D, H, W = 16, 16, 32
    a = U.bf16_round(rng.standard_normal((D, H, W, C1)))
    b = U.bf16_round(rng.standard_normal((D // 2, H // 2, W // 2, C2))) if C2 else None
    dz = U.bf16_round(rng.standard_normal((D, H, W, Cout)))
    ref = U.conv3_wgrad(a, dz, in2=b)
    da, dd, out = ctx.to_device(a), ctx.to_device(dz), ctx.empty((27, Cout, C1 + C2))
    d2 = ctx.to_device(b) if C2 else None
    A._check(A.lib.arec_conv3_wgrad_check(ctx.h, da.ptr, C1, None if d2 is None else d2.ptr, C2, dd.ptr, Cout,
                                          D, H, W, out.ptr))
    got = out.download()
    assert np.abs(got - ref).max() <= 1e-4 * np.abs(ref).max()


def _loss(net, dx, dt, p, out):
    net.set_params(p)
    net.forward(dx, out)
    y = out.download().astype(np.float64)
    return np.mean((y - dt.download()) ** 2)


def test_network_gradient_finite_differences(ctx):
    from arec3d_amd.unet import UNet

    print("\n[case] unet gradient", flush=True)
    n = 64
    rng = np.random.default_rng(3)
    x = rng.standard_normal((n, n, n)).astype(np.float32)
    t = (x + 0.3 * rng.standard_normal((n, n, n))).astype(np.float32)
    net = UNet(ctx, n, n, n, training=True, seed=11)
    p0 = net.get_params()
    # give the zero-initialised last layer weight so every layer gets gradient
    p0[-17:-1] = rng.standard_normal(16).astype(np.float32) * 0.1
    net.set_params(p0)
    dx, dt, out = ctx.to_device(x), ctx.to_device(t), ctx.empty((n, n, n))
    loss, g = net.gradient(dx, dt)
    assert np.isfinite(loss) and np.all(np.isfinite(g))
    assert loss == pytest.approx(_loss(net, dx, dt, p0, out), rel=1e-3)
    # per-layer blocks of the parameter vector
    blocks = [(0, 27 * 16 + 16)]
    off = 27 * 16 + 16
    for cin, cout in [(16, 16), (16, 32), (32, 32), (32, 64), (64, 64), (96, 32), (32, 32), (48, 16), (16, 16)]:
        blocks.append((off, off + 27 * cout * cin + cout))
        off += 27 * cout * cin + cout
    blocks.append((off, off + 17))
    assert off + 17 == len(p0)
    for lo, hi in blocks:
        v = np.zeros_like(p0)
        v[lo:hi] = g[lo:hi]
        gn = np.linalg.norm(v)
        assert gn > 0
        v /= gn
        # The BF16 forward makes the loss noisy (~1e-3 relative: weight
        # changes flip activation roundings) and the ReLU network is curved
        # along large steps, so a two-point difference is unreliable either
        # way.  Fit a cubic to the loss at 8 steps of up to ~8% loss change;
        # its linear coefficient is the directional derivative, with the
        # curvature fitted out and the noise averaged.
        h = 0.02 * loss / gn
        ts = np.array([-4, -3, -2, -1, 1, 2, 3, 4], np.float64) * h
        ls = np.array([_loss(net, dx, dt, p0 + t * v, out) for t in ts])
        fd = np.polynomial.polynomial.polyfit(ts, ls, 3)[1]
        print(f"block {lo}:{hi} analytic {gn:.5e} finite-diff {fd:.5e}")
        assert fd == pytest.approx(gn, rel=0.05)


def test_training_lowers_loss(ctx):
    from arec3d_amd.unet import UNet

    n = 64
    rng = np.random.default_rng(4)
    zz, yy, xx = np.mgrid[0:n, 0:n, 0:n]
    x = np.sin(xx / 5.0) * np.cos(yy / 7.0) + 0.1 * rng.standard_normal((n, n, n))
    t = np.sin(xx / 5.0) * np.cos(yy / 7.0)
    net = UNet(ctx, n, n, n, training=True, seed=5)
    dx, dt = ctx.to_device(x.astype(np.float32)), ctx.to_device(t.astype(np.float32))
    losses = [net.train_step(dx, dt, 1e-3) for _ in range(60)]
    assert losses[-1] < 0.5 * losses[0]
