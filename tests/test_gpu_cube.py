"""GPU parity of the cube <-> cylinder conversions (arec_cyl_to_cube,
arec_cube_to_cyl: ImageCyl2CB / ImageCB2Cyl's scatter and gather,
arecImage.cpp:88-232, with the detector rows as the cube's slowest axis)
against the oracle's restatement.  Copies: bit for bit."""
import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("radius,ny", [(1.0, 2), (1.5, 3), (31.5, 7), (32.0, 5), (255.5, 16)])
def test_cube_conversions_vs_oracle(ctx, radius, ny):
    from arec3d_amd import unet as U

    print(f"\n[case] cube conversions radius {radius}, {ny} rows", flush=True)
    rng = np.random.default_rng(int(radius * 10) + ny)
    n = int(round(2 * radius + 1))
    nr = len(O.cyl_geometry(radius))
    x = rng.standard_normal((nr, ny)).astype(np.float32)
    cube = ctx.empty((ny, n, n))
    U.cyl_to_cube(ctx, ctx.to_device(x), radius, ny, cube)
    assert np.array_equal(cube.download(), O.cyl_to_cube(x, radius))
    c = rng.standard_normal((ny, n, n)).astype(np.float32)
    xd = ctx.empty((nr, ny))
    U.cube_to_cyl(ctx, ctx.to_device(c), radius, ny, xd)
    assert np.array_equal(xd.download(), O.cube_to_cyl(c, radius))
