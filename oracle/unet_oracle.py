"""FP64 numpy restatement of the U-Net's convolution primitives (TEST
INFRASTRUCTURE ONLY: the checker for unet.hip's MFMA kernels).

Layout: channels-last [D][H][W][C]; weights [27][Cout][Cin] with tap
t = 9*(dz+1) + 3*(dh+1) + (dw+1); zero padding; input channels [C1, C1+C2)
read from a half-resolution tensor by nearest upsampling.
"""
import numpy as np


def bf16_round(a):
    """Round FP32 values to BF16 (nearest even), returned as FP32."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def concat_input(in1, in2):
    if in2 is None:
        return np.asarray(in1, np.float64)
    up = np.repeat(np.repeat(np.repeat(np.asarray(in2, np.float64), 2, 0), 2, 1), 2, 2)
    return np.concatenate([np.asarray(in1, np.float64), up], axis=3)


def conv3(in1, w, bias=None, in2=None, relu=False):
    x = concat_input(in1, in2)
    D, H, W, Ci = x.shape
    w = np.asarray(w, np.float64)
    Co = w.shape[1]
    xp = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    out = np.zeros((D, H, W, Co))
    for t in range(27):
        dz, dh, dw = t // 9, (t // 3) % 3, t % 3
        out += xp[dz:dz + D, dh:dh + H, dw:dw + W, :] @ w[t].T
    if bias is not None:
        out += np.asarray(bias, np.float64)
    return np.maximum(out, 0) if relu else out


def conv3_wgrad(in1, dz, in2=None):
    """dW[t][co][ci] = sum_s dz[s, co] * x[s + off(t), ci]"""
    x = concat_input(in1, in2)
    D, H, W, Ci = x.shape
    g = np.asarray(dz, np.float64).reshape(-1, dz.shape[-1])
    xp = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    out = np.zeros((27, g.shape[1], Ci))
    for t in range(27):
        a, b, c = t // 9, (t // 3) % 3, t % 3
        out[t] = g.T @ xp[a:a + D, b:b + H, c:c + W, :].reshape(-1, Ci)
    return out
