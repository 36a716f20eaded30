"""The no-allocation cross-correlation entry point against arec_align_cc.

arec_align_cc_noalloc runs arec_align_cc's kernels on a fixed device array
instead of allocated work buffers, so its shifts must equal arec_align_cc's
bit for bit.  Two calls in a row, the second on a smaller input (the array
keeps the first call's contents), a view count that needs several batches
of the fixed array, a non-power-of-two size (direct DFT with its temporary),
and the size limit refused on the host with an error.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _stack(rng, nv, ny, nx, shift=True):
    base = rng.random((ny, nx)).astype(np.float32)
    meas = np.empty((nv, ny, nx), np.float32)
    for v in range(nv):
        meas[v] = np.roll(base, (int(rng.integers(-3, 4)), int(rng.integers(-3, 4))), axis=(0, 1)) if shift else base
    meas += 0.05 * rng.standard_normal(meas.shape).astype(np.float32)
    return np.broadcast_to(base, (nv, ny, nx)).copy(), meas


def _both(ctx, proj, meas):
    import arec3d_amd as A

    nv = proj.shape[0]
    dp, dm = ctx.to_device(proj), ctx.to_device(meas)
    out = []
    for fn in (A.align_cc, A.align_cc_noalloc):
        sx, sy = ctx.empty(nv), ctx.empty(nv)
        fn(ctx, dp, dm, sx, sy)
        out.append((sx.download(), sy.download()))
    return out


@pytest.mark.parametrize("shape", [[(12, 64, 128), (3, 32, 64)], [(7, 48, 80), (2, 20, 36)]])
def test_noalloc_cc_equals_cc_twice(ctx, shape):
    rng = np.random.default_rng(17)
    for nv, ny, nx in shape:  # second call smaller than the first
        print(f"\n[case] noalloc cc {nv} views of {nx}x{ny}", flush=True)
        (sx, sy), (nsx, nsy) = _both(ctx, *_stack(rng, nv, ny, nx))
        assert np.array_equal(sx, nsx) and np.array_equal(sy, nsy)
        assert np.all(np.isfinite(sx)) and np.abs(sx).max() > 0.5  # real shifts were found


def test_noalloc_cc_several_batches(ctx):
    """1024 x 1024 views need 28 MiB of scratch each: 6 views run as
    batches of 4 and 2 in the 128 MiB array."""
    rng = np.random.default_rng(18)
    print("\n[case] noalloc cc 6 views of 1024x1024", flush=True)
    (sx, sy), (nsx, nsy) = _both(ctx, *_stack(rng, 6, 1024, 1024))
    assert np.array_equal(sx, nsx) and np.array_equal(sy, nsy)


def test_noalloc_cc_rejects_oversized_views(ctx):
    import arec3d_amd as A

    print("\n[case] noalloc cc size limit", flush=True)
    small = ctx.empty((1, 2, 2))
    sx, sy = ctx.empty(1), ctx.empty(1)
    # 4096 x 4096 needs 448 MiB: refused on the host before any launch (the
    # small buffers passed are never touched)
    rc = A.lib.arec_align_cc_noalloc(ctx.h, small.ptr, small.ptr, 4096, 4096, 1, sx.ptr, sy.ptr)
    assert rc == -103 and b"bytes of scratch" in A.lib.arec_last_error()
