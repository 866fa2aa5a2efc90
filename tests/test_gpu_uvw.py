"""rotatedUVW / uvwRotationDelay on the device (ygrid_uvw_rotate; TableVisGridder.cc:481-495)
against the oracle's change of frame (oracle.uvw_rotate, pinned in test_oracle_pins), through
the host and the device paths, and end to end: raw phase-centre uvw rotated on the device feed
a device chunk whose gridding matches the oracle's gridding of the oracle-rotated chunk, with a
source at the image centre imaged at the centre pixel."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import announce
from tests.devmem import DevArray
from tests.test_gpu_parity import ARC, TOL, rel

pytestmark = pytest.mark.gpu

P = (1.1, -0.6)
T = (P[0] + math.radians(2.0), P[1] - math.radians(1.5))
C = (T[0] + math.radians(0.4), T[1] + math.radians(0.3))


def _uvw(nrow, seed, uv=3000.0, w=800.0):
    rng = np.random.default_rng(seed)
    return np.ascontiguousarray(rng.normal(size=(nrow, 3)) * np.array([uv, uv, w]))


@pytest.mark.parametrize("per_row", [False, True])
def test_uvw_rotate_host_and_device_match_oracle(gpu_ctx, per_row):
    announce(f"uvw rotate, per-row phase centres={per_row}")
    import yandasoft_amd as ya
    n = 5003
    uvw = _uvw(n, 3)
    pc = np.array(P)
    if per_row:
        pc = np.ascontiguousarray(np.tile(np.array(P), (n, 1)) + np.random.default_rng(4).normal(size=(n, 2)) * 0.01)
    ou, od = O.uvw_rotate(uvw, pc, T, C)
    gu, gd = ya.uvw_rotate(gpu_ctx, uvw, pc, T, C)
    assert np.max(np.abs(gu - ou)) < 1e-8 and np.max(np.abs(gd - od)) < 1e-8
    du, dd = DevArray(np.zeros((n, 3))), DevArray(np.zeros(n))
    uvw_d, pc_d = DevArray(uvw), DevArray(pc)   # held until the stream has finished with them
    ya.uvw_rotate(gpu_ctx, uvw_d, pc_d, T, C, du, dd)
    gpu_ctx.synchronize()
    assert np.array_equal(du.numpy(), gu) and np.array_equal(dd.numpy(), gd)
    # no delay wanted
    u2, _ = ya.uvw_rotate(gpu_ctx, uvw[:10], pc if not per_row else pc[:10], T, C)
    assert np.array_equal(u2, gu[:10])


def test_rotated_device_chunk_grids_as_the_oracle(gpu_ctx):
    announce("uvw rotate -> device chunk -> SphFunc grid and image")
    import yandasoft_amd as ya
    N, cell, nrow = 128, 60 * ARC, 2000
    freq = np.array([1.2e9, 1.25e9])
    rng = np.random.default_rng(9)
    r = 150.0 * np.sqrt(rng.random(nrow))
    th = 2 * np.pi * rng.random(nrow)
    uvw = np.ascontiguousarray(np.stack([r * np.cos(th), r * np.sin(th), rng.uniform(-100, 100, nrow)], 1))
    # a unit source at the image centre, phased to P (V = exp(2 pi i (u l + v m + w (n - 1)) / lambda))
    l, m = O.direction_lm(C, P)
    nn = math.sqrt(1 - l * l - m * m)
    ph = (uvw[:, 0] * l + uvw[:, 1] * m + uvw[:, 2] * (nn - 1))[None, :] * freq[:, None] / O.C_LIGHT
    vis = np.ascontiguousarray(np.exp(2j * np.pi * ph)[None].astype(np.complex64))   # (1, nchan, nrow)
    noise = np.ones_like(vis)
    du, dd = DevArray(np.zeros((nrow, 3))), DevArray(np.zeros(nrow))
    vis_d, noise_d, freq_d = DevArray(vis), DevArray(noise), DevArray(freq)
    uvw_d, pc_d = DevArray(uvw), DevArray(np.array(P))
    ya.uvw_rotate(gpu_ctx, uvw_d, pc_d, T, C, du, dd)
    g = ya.Gridder(gpu_ctx, "GPUSphFunc", {})
    g.initialise_grid(ya.ImageDesc(N, N, 1, 1, cell, cell, centre=C))
    g.grid(ya.Chunk(uvw=du, freq=freq_d, vis=vis_d, noise=noise_d, delay=dd, device=True))
    img = g.finalise_grid()
    ou, od = O.uvw_rotate(uvw, P, T, C)
    o = O.OracleGridder("SphFunc", "f32")
    o.initialise_grid((N, N, 1, 1), (cell, cell))
    o.grid(O.Chunk(ou, freq, np.transpose(vis, (2, 1, 0)), noise=np.transpose(noise, (2, 1, 0)), delay=od))
    gg = g.grid_array().reshape(1, 1, N, N)
    assert rel(gg, o.grid_) < TOL
    # the source at the image centre is phase-flat after the rotation delay: the image peaks at the
    # centre pixel with the sum of the weights (2 channels x nrow, unit noise)
    im = img.reshape(N, N)
    assert np.unravel_index(np.argmax(im), im.shape) == (N // 2, N // 2)
    assert abs(im[N // 2, N // 2] / (2 * nrow) - 1.0) < 2e-2
