"""CPU: the FP64 convolution oracle of the U-Net agrees with scipy's
correlate and with brute force (pins it before it checks the MFMA kernels)."""
import numpy as np
from scipy.ndimage import correlate

import unet_oracle as U


def test_conv3_single_channel_matches_scipy():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((6, 5, 7, 1))
    k = rng.standard_normal((3, 3, 3))
    w = k.reshape(27, 1, 1)
    ref = correlate(x[..., 0], k, mode="constant", cval=0.0)
    np.testing.assert_allclose(U.conv3(x, w)[..., 0], ref, rtol=1e-12, atol=1e-12)


def test_conv3_channels_and_upsampled_concat_brute_force():
    rng = np.random.default_rng(1)
    a = rng.standard_normal((4, 4, 4, 2))
    b = rng.standard_normal((2, 2, 2, 3))
    w = rng.standard_normal((27, 2, 5))
    out = U.conv3(a, w, in2=b)
    z, h, x, co = 1, 2, 3, 1
    s = 0.0
    for t in range(27):
        zz, hh, xx = z + t // 9 - 1, h + (t // 3) % 3 - 1, x + t % 3 - 1
        if 0 <= zz < 4 and 0 <= hh < 4 and 0 <= xx < 4:
            v = np.concatenate([a[zz, hh, xx], b[zz // 2, hh // 2, xx // 2]])
            s += w[t, co] @ v
    assert abs(out[z, h, x, co] - s) < 1e-12


def test_wgrad_is_adjoint_of_conv_in_weights():
    # <dz, conv(x, w)> = <wgrad(x, dz), w> for every w: the wgrad is exact
    rng = np.random.default_rng(2)
    x = rng.standard_normal((4, 5, 6, 3))
    dz = rng.standard_normal((4, 5, 6, 2))
    w = rng.standard_normal((27, 2, 3))
    lhs = np.sum(dz * U.conv3(x, w))
    rhs = np.sum(U.conv3_wgrad(x, dz) * w)
    assert abs(lhs - rhs) < 1e-9 * abs(lhs)


def test_bf16_round():
    v = np.array([1.0, 1.00390625, 1.005859375, -3.14159], np.float32)
    r = U.bf16_round(v)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == np.float32(1.0078125)
    assert abs(r[3] + 3.140625) < 1e-6
