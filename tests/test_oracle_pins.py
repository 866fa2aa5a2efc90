"""Pins of the CPU oracle against what the reference's tests and the mathematics fix.

Each check is chosen so that a plausible slip in the oracle (a dropped term, a wrong sign or
index, a transposed operand) fails it.  No GPU needed.
"""
import math

import numpy as np
import pytest
import scipy.special as sps

from oracle import oracle as O


# ---------------------------------------------------------------------------------------
# prolate spheroidal function (scimath::SpheroidalFunction, absent dependency)
# ---------------------------------------------------------------------------------------
def test_spheroidal_matches_schwab_rational_approximation():
    # SphFuncVisGridder.cc:268-333: the m=6, alpha=1 approximation "agreed to about the 7th
    # significant digit" with Schwab's sphfn; c = pi*support with support 3.
    sph = O.Spheroidal(math.pi * 3, 1.0)
    nu = np.linspace(0.0, 0.999, 200)
    ours = sph(nu)
    schwab = np.array([O.schwab_grdsf(x) for x in nu])
    assert np.max(np.abs(ours - schwab)) < 1e-6


@pytest.mark.parametrize("support,alpha", [(3, 1), (4, 1), (5, 2), (7, 0), (3, 2)])
def test_spheroidal_matches_scipy_pro_ang1(support, alpha):
    # psi_alpha(c, nu) = S_{alpha,alpha}(c, nu) (1-nu^2)^(-alpha/2) / S(0)
    # (apps/gridderExperiments.cc:232-233)
    c = math.pi * support
    sph = O.Spheroidal(c, alpha)
    nu = np.linspace(0.0, 0.98, 50)
    ref = np.array([sps.pro_ang1(alpha, alpha, c, x)[0] for x in nu])
    ref0 = sps.pro_ang1(alpha, alpha, c, 0.0)[0]
    ref = ref / ref0 * np.power(1 - nu * nu, -alpha / 2.0)
    assert np.max(np.abs(sph(nu) - ref)) < 1e-9


def test_spheroidal_zero_outside_unit_interval():
    sph = O.Spheroidal(math.pi * 3, 1.0)
    assert np.all(sph(np.array([1.0, 1.5, -1.0])) == 0.0)
    assert abs(sph(np.array([0.0]))[0] - 1.0) < 1e-14


# ---------------------------------------------------------------------------------------
# SupportSearcher (tests/gridding/SupportSearcherTest.h:52-115, restated)
# ---------------------------------------------------------------------------------------
def _gauss_buffer():
    xo, yo = -3, 7
    rows = np.arange(50)
    cols = np.arange(50)
    x = (rows[:, None] - 25 - xo) / 5.0
    y = (cols[None, :] - 25 - yo) / 5.0
    return (5.0 * np.exp(-(x * x + y * y) / 2)).astype(np.complex64), xo, yo


def test_support_searcher_peak_find():
    buf, xo, yo = _gauss_buffer()
    ss = O.support_search(buf, 1e-2)
    assert ss.peak == (25 + xo, 25 + yo)
    assert abs(ss.peak_val - 5.0) < 1e-6


@pytest.mark.parametrize("absolute", [False, True])
def test_support_searcher_known_answers(absolute):
    buf, xo, yo = _gauss_buffer()
    cutoff = 1e-2
    if absolute:
        ss = O.support_search(buf, cutoff * 5.0, 1.0)
    else:
        ss = O.support_search(buf, cutoff)
    hw = 5.0 * math.sqrt(-2.0 * math.log(cutoff))
    assert abs(ss.support() - 2 * hw) < 1.0
    assert abs(ss.symmetrical(buf.shape) - 2 * (hw + max(xo, yo))) < 1.0
    assert abs(ss.blc[0] - (25 + xo - hw)) < 1.0
    assert abs(ss.blc[1] - (25 + yo - hw)) < 1.0
    assert abs(ss.trc[0] - (25 + xo + hw)) < 1.0
    assert abs(ss.trc[1] - (25 + yo + hw)) < 1.0


# ---------------------------------------------------------------------------------------
# W sampling (tests/gridding/NonLinearWSamplingTest.h:50-90, restated) and getWPlane
# ---------------------------------------------------------------------------------------
def _round_trip(ws):
    mid = 0.0
    for i in range(101):
        plane = 2.0 * (i - 50) / 100.0
        m = ws.map(plane)
        if -0.5 <= m <= 0.5:
            mid += 1.0 / 101.0
        assert abs(ws.index(m) - plane) < 1e-6
    assert abs(ws.index(0.0)) < 1e-6
    assert abs(ws.index(-1.0) + 1.0) < 1e-6
    assert abs(ws.index(1.0) - 1.0) < 1e-6
    return mid


def test_power_law_sampling_round_trip():
    for i in range(5):
        exponent = 2.0 * (i - 5) / 10.0
        _round_trip(O.WSampling(1.0, 3, "powerlaw", exponent=exponent))


def test_gaussian_sampling_round_trip():
    for i in range(3, 11):
        w50 = i / 11.0 / math.sqrt(2)
        ws = O.WSampling(1.0, 101, "gaussian", nwplanes50=w50 * 101)
        mid = _round_trip(ws)
        assert abs(w50 - mid) < 0.02


def test_linear_wplane_mapping():
    ws = O.WSampling(1000.0, 9)  # wscale = 250
    assert ws.plane(0.0) == 4
    assert ws.plane(1000.0) == 8 and ws.plane(-1000.0) == 0
    assert ws.plane(124.9) == 4 and ws.plane(125.1) == 5 and ws.plane(-125.1) == 3
    assert ws.wterm(0) == -1000.0 and ws.wterm(8) == 1000.0
    with pytest.raises(ValueError):
        ws.plane(1200.0)
    assert O.WSampling(1000.0, 9, wmaxclip=True).plane(1200.0) == -1


# ---------------------------------------------------------------------------------------
# FrequencyMapper (tests/gridding/FrequencyMapperTest.h, restated)
# ---------------------------------------------------------------------------------------
def test_frequency_mapper():
    freqs = 1.308e9 + 1.6e7 * (np.arange(10) - 1.0)
    m = O.frequency_map(freqs, 8, 1.308e9, 1.42e9)
    assert m[0] == -1 and m[-1] == -1
    assert all(m[i] + 1 == i for i in range(1, 9))
    freqs = 1.436e9 - 1.6e7 * np.arange(10) + np.where(np.arange(10) % 2 == 0, 1.0, -1.0) * 1e6
    m = O.frequency_map(freqs, 8, 1.308e9, 1.42e9)
    assert m[0] == -1 and m[-1] == -1
    assert all(m[i] + 1 == 10 - i - 1 for i in range(1, 9))
    freqs = 1.304e9 + 8e6 * (np.arange(20) - 2.0)
    m = O.frequency_map(freqs, 8, 1.308e9, 1.42e9)
    assert all((m[i] == -1) if (i < 2 or i + 2 >= 20) else (m[i] + 1 == i // 2) for i in range(20))
    assert np.all(O.frequency_map(freqs, 1) == 0)


# ---------------------------------------------------------------------------------------
# convolution functions
# ---------------------------------------------------------------------------------------
def test_box_cf_closed_form():
    cf = O.box_cf()
    assert cf.oversample == 1 and len(cf.planes) == 1
    expect = np.zeros((3, 3))
    expect[1, 1] = 1
    assert np.array_equal(cf.planes[0], expect)


def test_sphfunc_cf_separable_and_normalised():
    cf = O.sphfunc_cf(1.0, 3, 8)
    sph = O.Spheroidal(3 * math.pi, 1.0)
    for fracu, fracv in [(0, 0), (3, 5), (7, 1)]:
        p = cf.planes[fracu + 8 * fracv]
        assert abs(np.sum(np.abs(p)) - 1.0) < 1e-12
        # separable closed form with nu = |os*(i-S)+frac|/(S*os)
        i = np.arange(7)
        nux = np.abs(8 * (i - 3) + fracu) / 24.0
        nuy = np.abs(8 * (i - 3) + fracv) / 24.0
        fx = sph(nux) * (1 - nux ** 2)
        fy = sph(nuy) * (1 - nuy ** 2)
        ref = np.outer(fy, fx)
        ref /= ref.sum()
        assert np.max(np.abs(p - ref)) < 1e-14
    # fracu shifts the kernel along u (columns), not v: a transposed operand would fail
    p = cf.planes[3]
    assert np.allclose(p, p[::-1, :])
    assert not np.allclose(p, p[:, ::-1])


def test_grid_correction_closed_form():
    n = 256
    c = O.sph_grid_correction(n)
    assert abs(c[n // 2] - 1.0) < 1e-12
    assert np.all(np.diff(c[n // 2:200]) > 0)  # correction grows away from the centre
    assert np.allclose(c[n // 2 + 1:], c[1:n // 2][::-1], rtol=1e-12)  # symmetric about n/2
    # direct evaluation of the 7 low-frequency terms about the centre (no FFT shifts)
    sph = O.Spheroidal(3 * math.pi, 1.0)
    b = sph(np.abs(np.arange(n) - n // 2) / (n // 2))
    b[0] = 3 * (b[1] - b[2]) + b[3]
    x = np.arange(n) - n // 2
    F = [np.sum(b * np.exp(-2j * np.pi * k * x / n)) for k in range(-3, 4)]
    f = sum(F[k + 3] * np.exp(2j * np.pi * k * x / n) for k in range(-3, 4)).real
    f *= b[n // 2] / f[n // 2]
    assert np.allclose(c, 1.0 / f, rtol=1e-10)


def test_wproject_cf_w0_plane_is_taper_transform_and_normalised():
    prm = O.WProjectParams(wmax=2000.0, nwplanes=5, cutoff=1e-3, oversample=4, maxsupport=128)
    N, cell = 256, math.radians(20 / 3600)
    uvcell = (1 / (cell * N), 1 / (cell * N))
    cf, S = O.wproject_cf(prm, N, N, uvcell, fp32_buffer=False)
    os_ = 4
    for z in range(5):
        for frac in range(16):
            p = cf.planes[frac + 16 * z]
            assert abs(np.sum(p.real) - 1.0) < 1e-12
    # brute-force DFT of the taper x phase screen, with coordinates relative to the centre:
    # C(ix, iy) ~ sum_xy t(x) t(y) exp(-i w (1 - sqrt(1 - r^2))) exp(-2 pi i (kx x + ky y)/n),
    # kx = ix*os + fracu (columns), ky = iy*os + fracv (rows), normalised by sum Re.
    n = prm.maxsupport
    q = n // os_
    sph = O.Spheroidal(3 * math.pi, 1.0)
    x = np.arange(q) - q // 2
    t = (sph(np.abs(x) / (q // 2)) / q).astype(np.float32)
    t[0] = 3 * (t[1] - t[2]) + t[3]
    ccell = 1.0 / (q * uvcell[0])
    r2 = (x[:, None] * ccell) ** 2 + (x[None, :] * ccell) ** 2
    ws = prm.sampling()
    for z, fu, fv in [(2, 0, 0), (0, 1, 3), (4, 3, 2)]:
        w = 2 * math.pi * ws.wterm(z)
        scr = np.where(r2 < 1, np.outer(t, t) * np.exp(-1j * w * (1 - np.sqrt(np.clip(1 - r2, 0, None)))), 0)
        k = np.arange(-S, S + 1) * os_
        ex = np.exp(-2j * np.pi * np.outer(k + fu, x) / n)  # (taps, x)
        ey = np.exp(-2j * np.pi * np.outer(k + fv, x) / n)
        ref = ey @ scr @ ex.T  # [iy, ix]
        ref /= np.sum(ref.real)
        assert np.max(np.abs(cf.planes[fu + os_ * (fv + os_ * z)] - ref)) < 1e-10


@pytest.mark.parametrize("cutoff,expect", [
    # w planes -10, -5, 0, 5, 10 wavelengths, du = 0.5: wThetaPix = |w| / du^2 = 40, 20, 0, 20, 40;
    # kernel widths worked by hand from the reference's fits (WProjectVisGridder.cc:405-418)
    (1e-2, [40.6078810, 21.1896201, 3.0, 21.1896201, 40.6078810]),   # sqrt(49 + wThetaPix^2)
    (1e-3, [51.6, 28.8, 3.0, 28.8, 51.6]),                            # 6 + 1.14 * wThetaPix
])
def test_wproject_pcf_kernel_closed_form(cutoff, expect):
    prm = O.WProjectParams(wmax=10.0, nwplanes=5, cutoff=cutoff, oversample=2)
    cf = O.wproject_pcf_cf(prm, (0.5, 0.5))
    assert len(cf.planes) == 5 * 4
    for z in range(5):
        for frac in range(4):
            p = cf.planes[frac + 4 * z]
            assert p.shape == (3, 3)
            assert p[1, 1].real == 1.0
            assert abs(p[1, 1].imag - expect[z]) < 1e-5 * expect[z]
            q = p.copy()
            q[1, 1] = 0
            assert not np.any(q)
    # below one pixel of w.theta^2 the nominal 3 pixels (|w| = 0.2: 0.8), above it the fit (|w| = 0.3: 1.2)
    small = O.WProjectParams(wmax=0.2, nwplanes=3, cutoff=1e-3, oversample=1)
    assert O.wproject_pcf_kernel_pix(small, (0.5, 0.5), 0) == np.float32(3.0)
    above = O.WProjectParams(wmax=0.3, nwplanes=3, cutoff=1e-3, oversample=1)
    assert abs(O.wproject_pcf_kernel_pix(above, (0.5, 0.5), 2) - 7.368) < 1e-5


# ---------------------------------------------------------------------------------------
# the gridding loop: closed forms, brute force, adjoint identity
# ---------------------------------------------------------------------------------------
def _disc(nrow, N, cell, f, rng, frac=0.4):
    lam = O.C_LIGHT / f
    uvmax = frac * (N / 2) * (1 / (cell * N)) * lam
    r = uvmax * np.sqrt(rng.random(nrow))
    th = 2 * np.pi * rng.random(nrow)
    return np.stack([r * np.cos(th), r * np.sin(th), np.zeros(nrow)], 1)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_point_source_degrids_to_flux(prec):
    # ImageFFTEquationTest.h:124-175: a model pixel of flux F at the image centre predicts F.
    rng = np.random.default_rng(1)
    N, cell, f = 128, math.radians(10 / 3600), np.array([1.4e9])
    uvw = _disc(300, N, cell, f[0], rng)
    vis = np.zeros((300, 1, 1), np.complex64)
    model = np.zeros((1, 1, N, N))
    model[0, 0, N // 2, N // 2] = 2.5
    g = O.OracleGridder("SphFunc", prec, sph_oversample=16)
    g.initialise_degrid((N, N, 1, 1), (cell, cell), model)
    v = g.degrid(O.Chunk(uvw, f, vis))
    assert np.max(np.abs(v - 2.5)) < 1e-5


def test_unit_visibilities_image_peak_is_sum_of_weights():
    rng = np.random.default_rng(2)
    N, cell, f = 128, math.radians(10 / 3600), np.array([1.4e9])
    uvw = _disc(400, N, cell, f[0], rng)
    vis = np.ones((400, 1, 1), np.complex64)
    noise = np.full_like(vis, 0.5)  # weight 1/0.25 = 4
    g = O.OracleGridder("SphFunc", "f64", sph_oversample=16)
    g.initialise_grid((N, N, 1, 1), (cell, cell))
    g.grid(O.Chunk(uvw, f, vis, noise=noise))
    img = g.finalise_grid()
    w = g.finalise_weights()
    assert abs(w[0, 0, 0, 0] - 1600.0) < 1e-9
    assert abs(img[0, 0, N // 2, N // 2] - 1600.0) < 1e-6
    # the unit-visibility image is the PSF: its centre is the maximum of the inner quarter
    inner = img[0, 0, N // 4:3 * N // 4, N // 4:3 * N // 4]
    assert np.unravel_index(np.argmax(inner), inner.shape) == (N // 4, N // 4)


def test_grid_brute_force_single_sample():
    # one visibility, weight 1: grid(iu+du, iv+dv) = conj(V) * C(du, dv) on the chosen plane
    N, cell, f = 64, math.radians(30 / 3600), np.array([1.0e9])
    lam = O.C_LIGHT / f[0]
    du = 1 / (cell * N)
    os_ = 8
    u = (5 + 3.0 / os_ + 0.01) * du * lam  # fractional offset avoids exact ties
    v = (-7 - 2.0 / os_ - 0.01) * du * lam
    uvw = np.array([[u, v, 0.0]])
    vis = np.array([[[0.3 - 0.8j]]], np.complex64)
    g = O.OracleGridder("SphFunc", "f64", sph_oversample=os_)
    g.initialise_grid((N, N, 1, 1), (cell, cell))
    g.grid(O.Chunk(uvw, f, vis))
    uS, vS = u / lam / du, v / lam / du
    iu, iv = O.nint(uS), O.nint(vS)
    fu, fv = O.nint(os_ * (iu - uS)), O.nint(os_ * (iv - vS))
    if fu < 0:
        iu, fu = iu + 1, fu + os_
    if fv < 0:
        iv, fv = iv + 1, fv + os_
    plane = g.cf.planes[fu + os_ * fv]
    expect = np.zeros((N, N), complex)
    expect[iv + N // 2 - 3:iv + N // 2 + 4, iu + N // 2 - 3:iu + N // 2 + 4] = np.conj(vis[0, 0, 0]) * plane
    assert np.max(np.abs(g.grid_[0, 0] - expect)) < 1e-12


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_polarisation_and_channel_planes_are_independent(prec):
    # every (pol, image chan) plane of Array(nx, ny, npol, nchan) is gridded on its own: two
    # pols holding V and 2V give planes G and 2G of the single-pol grid; two image channels
    # each receive the samples of their own frequency
    rng = np.random.default_rng(4)
    N, cell = 64, math.radians(40 / 3600)
    f = np.array([1.0e9, 1.2e9])
    uvw = _disc(300, N, cell, f[1], rng, 0.4)
    v = (rng.normal(size=(300, 2, 1)) + 1j * rng.normal(size=(300, 2, 1))).astype(np.complex64)
    g1 = O.OracleGridder("Box", prec)
    g1.initialise_grid((N, N, 1, 1), (cell, cell))
    g1.grid(O.Chunk(uvw, f, v))
    g2 = O.OracleGridder("Box", prec)
    g2.initialise_grid((N, N, 2, 1), (cell, cell))
    g2.grid(O.Chunk(uvw, f, np.concatenate([v, 2 * v], axis=2)))
    assert np.array_equal(g2.grid_[0, 0], g1.grid_[0, 0])
    assert np.allclose(g2.grid_[0, 1], 2 * g1.grid_[0, 0], rtol=1e-6, atol=1e-6)
    g3 = O.OracleGridder("Box", prec)
    g3.initialise_grid((N, N, 1, 2), (cell, cell), freq_axis=(1.0e9, 1.2e9))
    g3.grid(O.Chunk(uvw, f, v))
    ga = O.OracleGridder("Box", prec)
    ga.initialise_grid((N, N, 1, 1), (cell, cell))
    ga.grid(O.Chunk(uvw, f[:1], np.ascontiguousarray(v[:, :1])))
    assert np.array_equal(g3.grid_[0, 0], ga.grid_[0, 0])


def test_adjoint_identity_of_grid_and_degrid():
    # sum_s D_s(M) conj(W_s) = sum_p G(W)[p] conj(M_p)   (GridKernel.cc:79-80 vs :121-122)
    rng = np.random.default_rng(3)
    N, cell, f = 96, math.radians(15 / 3600), np.array([1.2e9, 1.3e9])
    nrow = 200
    uvw = _disc(nrow, N, cell, f[1], rng, 0.5)
    W = (rng.normal(size=(nrow, 2, 1)) + 1j * rng.normal(size=(nrow, 2, 1))).astype(np.complex64)
    g = O.OracleGridder("SphFunc", "f64", sph_oversample=8)
    g.initialise_grid((N, N, 1, 1), (cell, cell))
    g.grid(O.Chunk(uvw, f, W))
    G = g.grid_[0, 0].copy()
    M = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
    d = O.OracleGridder("SphFunc", "f64", sph_oversample=8)
    d.initialise_degrid((N, N, 1, 1), (cell, cell), np.ones((1, 1, N, N)))
    d.grid_[0, 0] = M
    D = d.degrid(O.Chunk(uvw, f, np.zeros_like(W)))
    lhs = np.sum(D * np.conj(W))
    rhs = np.sum(G * np.conj(M))
    assert abs(lhs - rhs) < 1e-9 * abs(rhs)


def test_degrid_offgrid_sample_reference_adds_stale_value():
    # TableVisGridder.cc:763-861: an off-grid sample leaves itsImagePolFrameVis unchanged and
    # the stale value of the previous on-grid sample is still added to the visibility.
    N, cell, f = 64, math.radians(30 / 3600), np.array([1.0e9])
    lam = O.C_LIGHT / f[0]
    du = 1 / (cell * N)
    uvw = np.array([[3.1 * du * lam, 2.2 * du * lam, 0], [40.0 * du * lam, 0.0, 0]])
    model = np.zeros((1, 1, N, N))
    model[0, 0, N // 2, N // 2] = 1.0
    d = O.OracleGridder("SphFunc", "f64", sph_oversample=4)
    d.initialise_degrid((N, N, 1, 1), (cell, cell), model)
    v = d.degrid(O.Chunk(uvw, f, np.zeros((2, 1, 1), np.complex64)))
    assert abs(v[0, 0, 0] - 1.0) < 1e-9 and abs(v[1, 0, 0] - v[0, 0, 0]) < 1e-12
    # without the stale add the off-grid sample stays zero and the on-grid one is unchanged
    d.stale_offgrid = False
    v2 = d.degrid(O.Chunk(uvw, f, np.zeros((2, 1, 1), np.complex64)))
    assert v2[1, 0, 0] == 0 and v2[0, 0, 0] == v[0, 0, 0]


# ---------------------------------------------------------------------------------------
# learned-CF network oracle: analytic gradient vs central differences, SiLU' closed form
# ---------------------------------------------------------------------------------------
def test_mlp_oracle_gradient_matches_finite_differences():
    from oracle import mlp as M
    rng = np.random.default_rng(3)
    hidden, layers = 6, 2
    n = sum(a * b + b for a, b in zip(M.dims(hidden, layers)[:-1], M.dims(hidden, layers)[1:]))
    p = rng.normal(0, 0.5, n)
    x = rng.uniform(-1, 1, (7, 5))
    t = rng.normal(0, 1, (7, 2))
    loss, g = M.loss_grad(p, x, t, hidden, layers)
    assert loss == pytest.approx(np.sum((M.forward(p, x, hidden, layers) - t) ** 2) / 14, rel=1e-14)
    h = 1e-6
    fd = np.zeros(n)
    for i in range(n):
        e = np.zeros(n)
        e[i] = h
        fd[i] = (M.loss_grad(p + e, x, t, hidden, layers)[0] - M.loss_grad(p - e, x, t, hidden, layers)[0]) / (2 * h)
    assert np.max(np.abs(fd - g)) <= 1e-7 * max(1.0, np.max(np.abs(g)))


def test_mlp_oracle_fp16_rounding_form():
    """forward_f16 is forward with FP16 rounding of weights, inputs and hidden activations: with the
    rounding replaced by the identity it is forward exactly, and with FP16 rounding it stays within
    the bound that rounding steps of 2^-11 propagated through the layers allow."""
    from oracle import mlp as M
    rng = np.random.default_rng(3)
    d = M.dims(32, 3)
    p = rng.normal(0, 0.15, sum(a * b + b for a, b in zip(d[:-1], d[1:])))
    x = rng.uniform(-1, 1, (64, 5))
    exact = M.forward(p, x, 32, 3)
    assert np.array_equal(M.forward_f16(p, x, 32, 3, rnd=lambda a: np.asarray(a, np.float64)), exact)
    d = np.max(np.abs(M.forward_f16(p, x, 32, 3) - exact)) / np.max(np.abs(exact))
    assert 0 < d < 3 * 2.0 ** -11 * 4


def test_mlp_oracle_silu_derivative():
    from oracle import mlp as M
    z = np.linspace(-8, 8, 101)
    h = 1e-6
    assert np.max(np.abs((M.silu(z + h) - M.silu(z - h)) / (2 * h) - M.silu_grad(z))) < 1e-8
    assert M.silu(0.0) == 0.0 and M.silu_grad(0.0) == 0.5


# ---- polarisation conversion (oracle.pol_matrix / pol_to_image / pol_to_vis) -----------------
XX, XY, YX, YY, RR, RL, LR, LL, SI, SQ, SU, SV = 9, 10, 11, 12, 5, 6, 7, 8, 1, 2, 3, 4


def test_pol_matrix_closed_forms():
    """Stokes from the linear and circular products by the textbook inversions
    (I = (XX+YY)/2, Q = (XX-YY)/2, U = (XY+YX)/2, V = (XY-YX)/(2i); I = (RR+LL)/2,
    V = (RR-LL)/2, Q = (RL+LR)/2, U = (RL-LR)/(2i)), and the forward products of a source
    with known Stokes parameters."""
    rng = np.random.default_rng(1)
    xx, xy, yx, yy = rng.normal(size=4) + 1j * rng.normal(size=4)
    lin = np.array([xx, xy, yx, yy])
    got = O.pol_matrix([XX, XY, YX, YY], [SI, SQ, SU, SV]) @ lin
    want = [(xx + yy) / 2, (xx - yy) / 2, (xy + yx) / 2, (xy - yx) / 2j]
    assert np.allclose(got, want, atol=1e-14)
    rr, rl, lr, ll = lin
    got = O.pol_matrix([RR, RL, LR, LL], [SI, SQ, SU, SV]) @ np.array([rr, rl, lr, ll])
    assert np.allclose(got, [(rr + ll) / 2, (rl + lr) / 2, (rl - lr) / 2j, (rr - ll) / 2], atol=1e-14)
    # I only from the parallel hands
    assert np.allclose(O.pol_matrix([XX, YY], [SI]), [[0.5, 0.5]], atol=1e-15)
    # a source (I, Q, U, V) = (3, 1, 0.5, 0.25): the products by their definitions
    s = np.array([3.0, 1.0, 0.5, 0.25])
    assert np.allclose(O.pol_matrix([SI, SQ, SU, SV], [XX, XY, YX, YY]) @ s,
                       [4.0, 0.5 + 0.25j, 0.5 - 0.25j, 2.0], atol=1e-14)
    assert np.allclose(O.pol_matrix([SI, SQ, SU, SV], [RR, RL, LR, LL]) @ s,
                       [3.25, 1.0 + 0.5j, 1.0 - 0.5j, 2.75], atol=1e-14)


def test_pol_matrix_round_trips_and_undetermined_outputs():
    lin, circ, stk = [XX, XY, YX, YY], [RR, RL, LR, LL], [SI, SQ, SU, SV]
    for a, b in ((lin, stk), (circ, stk), (lin, circ)):
        assert np.allclose(O.pol_matrix(b, a) @ O.pol_matrix(a, b), np.eye(4), atol=1e-14)
    with pytest.raises(ValueError):
        O.pol_matrix([XX, YY], [SV])  # V needs the cross hands
    with pytest.raises(ValueError):
        O.pol_matrix([RR, LL], [SQ])  # Q needs the cross hands
    # predicting from an I-only image: unpolarised products, zero cross hands
    m = np.ones((1, 2, 3), np.complex128) * (2 + 1j)
    v = O.pol_to_vis(m, [SI], lin)
    assert np.allclose(v[0], 2 + 1j) and np.allclose(v[3], 2 + 1j) and np.allclose(v[1:3], 0)


def test_pol_noise_closed_forms():
    """Independent errors: I = (XX+YY)/2 with sigma (s_re, s_im) on both hands has sigma /
    sqrt(2) on each part; V = (XY-YX)/(2i) = (-i/2) XY + (i/2) YX moves the imaginary parts'
    errors into the real part and vice versa."""
    n = np.empty((4, 1, 1), np.complex128)
    n[:, 0, 0] = [0.4 + 0.2j, 1.0 + 3.0j, 2.0 + 5.0j, 0.4 + 0.2j]
    _, nI, _ = O.pol_to_image(None, n, None, [XX, XY, YX, YY], [SI])
    assert np.allclose(nI[0, 0, 0], (0.4 + 0.2j) / np.sqrt(2))
    _, nV, _ = O.pol_to_image(None, n, None, [XX, XY, YX, YY], [SV])
    assert np.allclose(nV[0, 0, 0], 0.5 * np.sqrt(3.0 ** 2 + 5.0 ** 2) + 0.5j * np.sqrt(1.0 + 2.0 ** 2))
    f = np.zeros((4, 1, 2), bool)
    f[2, 0, 1] = True
    _, _, fo = O.pol_to_image(None, None, f, [XX, XY, YX, YY], [SI, SQ])
    assert fo.shape == (2, 1, 2) and not fo[:, 0, 0].any() and fo[:, 0, 1].all()


# ---- w stacking (oracle.wstack_screen) ----------------------------------------------------------
def test_wstack_screen_closed_forms():
    """w = 0: no change; |screen| = 1; the phase 2 pi w (1 - sqrt(1 - r^2)) at a pixel computed
    by hand; pixels with r^2 >= 1 left unchanged."""
    nx, ny, uvcell = 16, 8, (1.0 / (16 * 0.1), 1.0 / (8 * 0.15))  # image cells 0.1 and 0.15 (radians)
    assert np.allclose(O.wstack_screen(nx, ny, uvcell, 0.0), 1.0)
    s = O.wstack_screen(nx, ny, uvcell, 40.0)
    assert np.allclose(np.abs(s), 1.0, atol=1e-6)
    x, y = 0.1 * (11 - 8), 0.15 * (2 - 4)  # pixel (ix 11, iy 2)
    p = 2 * np.pi * 40.0 * (1 - np.sqrt(1 - x * x - y * y))
    assert abs(s[2, 11] - np.exp(-1j * p)) < 2e-5
    # corner (0, 0): r^2 = 0.8^2 + 0.6^2 = 1 -> outside, unchanged
    assert s[0, 0] == 1.0
    assert np.angle(s[4, 8]) == 0.0  # the centre


# ---- A-projection: DiskIllumination, AWProjectVisGridder CFs and weights, fftPad ----------------
def _lattice_points_in_disk(r2max, nu, nv):
    """Gauss circle count of integer offsets (du, dv), -n/2 <= d < n/2, with du^2 + dv^2 <= r2max,
    by integer square roots row by row (no distance arrays)."""
    count = 0
    for dv in range(-nv // 2, nv - nv // 2):
        rem = r2max - dv * dv
        if rem < 0:
            continue
        k = math.isqrt(int(math.floor(rem)))
        lo, hi = max(-k, -nu // 2), min(k, nu - nu // 2 - 1)
        count += max(0, hi - lo + 1)
    return count


@pytest.mark.parametrize("diameter,blockage", [(12.0, 0.0), (25.0, 6.0)])
def test_disk_pattern_count_normalisation_and_modulus(diameter, blockage):
    """DiskIllumination.cc:154-176: the pattern is nU*nV/count on every lattice point of the
    annulus rmin^2 <= r^2 <= rmax^2 and 0 elsewhere; with a beam offset only its phase changes."""
    nu = nv = 64
    ucell, os_, freq = 8.0, 2, 1.0e9
    cell_m = ucell / os_ * (O.C_LIGHT / freq)
    r2max = (diameter / (2 * cell_m)) ** 2
    r2min = (blockage / (2 * cell_m)) ** 2
    count = _lattice_points_in_disk(r2max, nu, nv) - (
        _lattice_points_in_disk(r2min - 1e-12, nu, nv) if blockage > 0 else 0)
    pat = O.disk_pattern(nu, nv, ucell, ucell, os_, freq, diameter, blockage)
    assert np.count_nonzero(pat) == count
    assert abs(pat.sum() - nu * nv) < 1e-3 * nu * nv            # nU nV / count on count points
    assert np.allclose(pat[pat != 0], nu * nv / count, rtol=1e-6)
    off = O.disk_pattern(nu, nv, ucell, ucell, os_, freq, diameter, blockage, l=0.01, m=-0.02)
    assert np.array_equal(off != 0, pat != 0)
    assert np.allclose(np.abs(off[off != 0]), nu * nv / count, rtol=1e-6)
    with pytest.raises(ValueError):
        O.disk_pattern(nu, nv, ucell, ucell, os_, freq, 10.0, 12.0)   # hole larger than the dish


@pytest.mark.parametrize("kx,ky", [(3, 0), (-2, 5)])
def test_disk_pattern_beam_points_at_its_offset(kx, ky):
    """The phase slope of an offset beam (DiskIllumination.cc:105-110,166-167) steers the voltage
    pattern (its inverse transform, AWProjectVisGridder.cc:383) to the offset: with
    l = kx / (n cellU) the peak of |A| lies kx pixels from the centre, on the same side."""
    n, ucell, os_, freq = 64, 8.0, 2, 1.0e9
    cu = ucell / os_
    pat = O.disk_pattern(n, n, ucell, ucell, os_, freq, 12.0, 0.0, l=kx / (n * cu), m=ky / (n * cu))
    A = np.abs(O.fft_centred(pat.astype(np.complex128), False))
    iy, ix = np.unravel_index(np.argmax(A), A.shape)
    assert (ix - n // 2, iy - n // 2) == (kx, ky)


def test_direction_lm_is_an_orthographic_projection():
    """(l, m, n) of a direction about the image centre is a rotation of its unit vector:
    l^2 + m^2 + n^2 = 1 with n = cos(angular distance); the centre maps to (0, 0); a direction
    on the centre's meridian has l = 0 and m = sin(dlat)."""
    rng = np.random.default_rng(5)
    for _ in range(20):
        cl, cb = rng.uniform(0, 2 * np.pi), rng.uniform(-1.2, 1.2)
        bl, bb = cl + rng.uniform(-0.3, 0.3), np.clip(cb + rng.uniform(-0.3, 0.3), -1.5, 1.5)
        l, m = O.direction_lm((bl, bb), (cl, cb))
        nn = math.sin(bb) * math.sin(cb) + math.cos(bb) * math.cos(cb) * math.cos(bl - cl)
        assert abs(l * l + m * m + nn * nn - 1.0) < 1e-12
    assert O.direction_lm((1.0, 0.4), (1.0, 0.4)) == (0.0, 0.0)
    l, m = O.direction_lm((1.0, 0.45), (1.0, 0.4))
    assert abs(l) < 1e-15 and abs(m - math.sin(0.05)) < 1e-12


def _aw_small():
    prm = O.AWProjectParams(wmax=300.0, nwplanes=3, oversample=2, maxsupport=64, cutoff=1e-2, diameter=12.0,
                            maxfeeds=2, freqdep=True)
    freqs = np.array([1.0e9, 1.05e9])
    lam = O.C_LIGHT / freqs[0]
    uvcell = (12.0 / lam) / 8.3 * prm.oversample   # dish radius 4.15 pattern cells: no lattice ties
    cu = uvcell / prm.oversample
    lm = {0: (0.0, 0.0), 1: (2 / (prm.maxsupport * cu), -1 / (prm.maxsupport * cu))}
    return prm, freqs, uvcell, lm


def test_awproject_planes_match_brute_force_transforms():
    """AWProjectVisGridder.cc:370-440: per (feed, chan, w) the CF is os^2 x the forward transform,
    divided by nx ny, of |A|^2 x the w screen inside r^2 < 1, where A is the inverse transform of
    the disk pattern; planes are cut out at k = t os + frac about the centre.  Restated here with
    explicit centre-relative DFT sums (no shifts, no FFT), including the beam-offset phase."""
    prm, freqs, uvcell, lm = _aw_small()
    st = O.awproject_planes(prm, 64, 64, (uvcell, uvcell), freqs, lm)
    S, os_, n, nw = st["support"], prm.oversample, prm.maxsupport, prm.nwplanes
    assert S >= 3
    ws = prm.sampling()
    c = np.arange(n) - n / 2.0
    x = c / ((n // os_) * uvcell)
    r2 = x[None, :] ** 2 + x[:, None] ** 2
    E = np.exp(2j * np.pi * np.outer(c, c) / n)
    for feed, (l, m) in lm.items():
        for ch, f in enumerate(freqs):
            cu = uvcell / os_
            R2 = (prm.diameter / (2 * cu * O.C_LIGHT / f)) ** 2
            disk = (c[None, :] ** 2 + c[:, None] ** 2 <= R2)
            pat = disk * np.exp(-2j * np.pi * cu * (l * c[None, :] + m * c[:, None])) * (n * n / disk.sum())
            wt = np.abs(E @ pat @ E.T / (n * n)) ** 2
            for iw in range(nw):
                ph = 2 * np.pi * ws.wterm(iw) * (1 - np.sqrt(np.clip(1 - r2, 0, None)))
                scr = np.where(r2 < 1, wt * np.exp(-1j * ph), 0)
                z = iw + nw * (ch + len(freqs) * feed)
                for fu in range(os_):
                    for fv in range(os_):
                        Eu = np.exp(-2j * np.pi * np.outer(np.arange(-S, S + 1) * os_ + fu, c) / n)
                        Ev = np.exp(-2j * np.pi * np.outer(np.arange(-S, S + 1) * os_ + fv, c) / n)
                        ref = Ev @ scr @ Eu.T / (n * n) * os_ * os_
                        got = st["planes"][fu + os_ * (fv + os_ * z)]
                        assert np.max(np.abs(got - ref)) < 1e-6 * np.max(np.abs(ref)), (feed, ch, iw, fu, fv)


def test_awproject_variable_and_offset_support_planes_match_brute_force():
    """variablesupport + offsetsupport with power-law w planes (AWProjectVisGridder.cc:444-503,
    WProjectVisGridder.cc:233-263): every plane is cut out at its own support about its own offset,
    the offset being the plane's peak (in grid pixels, truncated).  For a beam offset (l0, m0) the
    w screen exp(-i pi w l^2) of the beam-centred aperture power, |A(l - l0)|^2, carries the slope
    exp(-2 pi i w l0 l), so the forward transform puts the kernel's peak at -w (l0, m0)
    wavelengths and the offsets follow that closed form; the planes themselves are checked against the same centre-relative DFT sums as above."""
    prm, freqs, uvcell, _ = _aw_small()
    prm.wmax, prm.variablesupport, prm.offsetsupport = 2000.0, True, True
    prm.wsampling, prm.exponent = "powerlaw", 2.0
    os_, n, nw = prm.oversample, prm.maxsupport, prm.nwplanes
    cu = uvcell / os_
    lm = {0: (0.0, 0.0), 1: (4 / (n * cu), -2 / (n * cu))}
    st = O.awproject_planes(prm, 64, 64, (uvcell, uvcell), freqs, lm)
    ws = prm.sampling()
    c = np.arange(n) - n / 2.0
    x = c / ((n // os_) * uvcell)
    r2 = x[None, :] ** 2 + x[:, None] ** 2
    E = np.exp(2j * np.pi * np.outer(c, c) / n)
    sizes = set()
    for feed, (l, m) in lm.items():
        for ch, f in enumerate(freqs):
            R2 = (prm.diameter / (2 * cu * O.C_LIGHT / f)) ** 2
            disk = (c[None, :] ** 2 + c[:, None] ** 2 <= R2)
            pat = disk * np.exp(-2j * np.pi * cu * (l * c[None, :] + m * c[:, None])) * (n * n / disk.sum())
            wt = np.abs(E @ pat @ E.T / (n * n)) ** 2
            for iw in range(nw):
                w = ws.wterm(iw)
                ph = 2 * np.pi * w * (1 - np.sqrt(np.clip(1 - r2, 0, None)))
                scr = np.where(r2 < 1, wt * np.exp(-1j * ph), 0)
                z = iw + nw * (ch + len(freqs) * feed)
                ou, ov = st["offsets"][z]
                expect = (O.cdiv(round(-w * l / cu), os_), O.cdiv(round(-w * m / cu), os_))
                assert (ou, ov) == expect, (feed, ch, iw, (ou, ov), expect)
                S = (st["planes"][os_ * os_ * z].shape[0] - 1) // 2
                sizes.add(S)
                for fu in range(os_):
                    for fv in range(os_):
                        Eu = np.exp(-2j * np.pi * np.outer((np.arange(-S, S + 1) + ou) * os_ + fu, c) / n)
                        Ev = np.exp(-2j * np.pi * np.outer((np.arange(-S, S + 1) + ov) * os_ + fv, c) / n)
                        ref = Ev @ scr @ Eu.T / (n * n) * os_ * os_
                        got = st["planes"][fu + os_ * (fv + os_ * z)]
                        assert got.shape == (2 * S + 1, 2 * S + 1)
                        assert np.max(np.abs(got - ref)) < 1e-6 * np.max(np.abs(ref)), (feed, ch, iw, fu, fv)
    assert len(sizes) > 1 and any(o != (0, 0) for o in st["offsets"])
    # the w = 0 plane (the smallest kernel) keeps the minimum support of 3 pixels or more
    assert st["support"] == (st["planes"][0].shape[0] - 1) // 2


def test_awproject_w0_plane_of_a_centred_disk_is_real_and_symmetric():
    """With no beam offset |A|^2 is real and symmetric, so at w = 0 the CF is real and has the
    square's symmetries (x -> -x, y -> -y, x <-> y) on the unshifted oversampling plane."""
    prm, freqs, uvcell, lm = _aw_small()
    st = O.awproject_planes(prm, 64, 64, (uvcell, uvcell), freqs, {0: (0.0, 0.0)})
    nw, os_ = prm.nwplanes, prm.oversample
    p = st["planes"][os_ * os_ * (nw // 2)]           # fracu = fracv = 0, w plane 0, chan 0, feed 0
    peak = np.max(np.abs(p))
    assert np.max(np.abs(p.imag)) < 1e-6 * peak
    assert np.max(np.abs(p - p[::-1, :])) < 1e-6 * peak
    assert np.max(np.abs(p - p[:, ::-1])) < 1e-6 * peak
    assert np.max(np.abs(p - p.T)) < 1e-6 * peak
    # an offset beam changes the w = 0 CF by a phase slope only (|CF| unchanged)
    st2 = O.awproject_planes(prm, 64, 64, (uvcell, uvcell), freqs, {0: lm[1]})
    p2 = st2["planes"][os_ * os_ * (nw // 2)]
    assert np.max(np.abs(np.abs(p2) - np.abs(p))) < 1e-3 * peak
    assert np.max(np.abs(p2 - p)) > 1e-2 * peak


def test_aprojectwstack_planes_match_brute_force_transforms():
    """AProjectWStackVisGridder.cc:222-370: per (chan, feed) the pattern's image-plane power |A|^2,
    divided by its peak, transformed forward and divided by nx ny, cut out at k = t os + frac for
    t in [-S, S-1] times os^2 (the last row and column zero); no w term.  Restated with explicit
    centre-relative DFT sums, the beam-offset phase included.  The support is 1 + 2 int(R) / os
    with R the dish radius in pattern cells (DiskIllumination.cc:139): R = 4.15 here, S = 5, set
    from the first channel and kept for the second (R = 4.36 there)."""
    prm, freqs, uvcell, lm = _aw_small()
    st = O.aprojectwstack_planes(prm, 64, 64, (uvcell, uvcell), freqs, lm)
    S, os_, n = st["support"], prm.oversample, prm.maxsupport
    assert S == 1 + (2 * 4) // os_ == 5
    c = np.arange(n) - n / 2.0
    E = np.exp(2j * np.pi * np.outer(c, c) / n)
    cu = uvcell / os_
    for feed, (l, m) in lm.items():
        for ch, f in enumerate(freqs):
            R2 = (prm.diameter / (2 * cu * O.C_LIGHT / f)) ** 2
            disk = (c[None, :] ** 2 + c[:, None] ** 2 <= R2)
            pat = disk * np.exp(-2j * np.pi * cu * (l * c[None, :] + m * c[:, None]))
            p = np.abs(E @ pat @ E.T) ** 2
            p /= p.max()
            z = ch + len(freqs) * feed
            t = np.arange(-S, S)
            for fu in range(os_):
                for fv in range(os_):
                    Eu = np.exp(-2j * np.pi * np.outer(t * os_ + fu, c) / n)
                    Ev = np.exp(-2j * np.pi * np.outer(t * os_ + fv, c) / n)
                    ref = Ev @ p @ Eu.T / (n * n) * os_ * os_
                    got = st["planes"][fu + os_ * (fv + os_ * z)]
                    assert got.shape == (2 * S + 1, 2 * S + 1)
                    assert not np.any(got[2 * S, :]) and not np.any(got[:, 2 * S])
                    assert np.max(np.abs(got[:2 * S, :2 * S] - ref)) < 1e-6 * np.max(np.abs(ref)), (feed, ch, fu, fv)
    # the peak-normalised power transforms to a kernel whose taps, summed over the oversampling
    # planes, give os^2 / (nx ny) x the power at the origin (1 for the centred feed, where it peaks)
    z0 = 0
    tot = sum(st["planes"][k + os_ * os_ * z0][:2 * S, :2 * S].sum() for k in range(os_ * os_))
    assert abs(tot - os_ * os_) < 1e-3 * os_ * os_


def test_aprojectwstack_weights_closed_form():
    """AProjectWStackVisGridder.cc:377-470: one CF with one nonzero tap at (dx, dy) from the centre
    (the cut-out's taps -S..S-1) transforms to a pure phase ramp of modulus 1 after the peak
    normalisation, so the weights image is the sum of weights everywhere; a CF with taps a at
    the centre and b at (dx, 0) gives |a + b e^{2 pi i dx X / cnx}|^2 over its largest value on the
    cnx pixels (cnx = grid here, so the padding leaves it as is)."""
    prm = O.AWProjectParams(wmax=300.0, nwplanes=1, oversample=2, maxsupport=32)
    g = O.OracleGridder("AProjectWStack", "f64", wproject=prm)
    ngrid = 32
    g._setup((ngrid, ngrid, 1, 1), (1e-3, 1e-3))
    S = 3
    a, b, dx = 0.8 - 0.1j, 0.3 + 0.2j, 2
    cf = np.zeros((2 * S + 1, 2 * S + 1), np.complex128)
    cf[S, S], cf[S, S + dx] = a, b
    planes = [cf if k == 0 else np.zeros_like(cf) for k in range(4)]
    g.cf = O.CFCache(planes, 2, None)
    g.sumw = np.zeros((1, 1, 1))
    g.sumw[0, 0, 0] = 2.5
    w = g._aws_weights()[0, 0]
    X = np.arange(ngrid) - ngrid // 2
    e = np.abs(a + b * np.exp(2j * np.pi * dx * X / ngrid)) ** 2
    e /= e.max()                                   # the peak over the sampled pixels
    assert np.max(np.abs(w - 2.5 * e[None, :])) < 1e-6      # the CF cache holds complex float
    # the last row and column of the cut-out are left out of the placement
    cf2 = np.zeros_like(cf)
    cf2[S, S], cf2[2 * S, S] = 1.0, 5.0
    g.cf = O.CFCache([cf2] + [np.zeros_like(cf)] * 3, 2, None)
    assert np.max(np.abs(g._aws_weights()[0, 0] - 2.5)) < 1e-6


def _aw_weights_gridder(ngrid, taper=False):
    prm = O.AWProjectParams(wmax=300.0, nwplanes=1, oversample=2, maxsupport=32, spheroidaltaper=taper,
                            spheroidalweightscutoff=0.05)
    g = O.OracleGridder("AWProject", "f64", wproject=prm)
    g._setup((ngrid, ngrid, 1, 1), (1e-3, 1e-3))
    g.track_per_os = False
    return g


@pytest.mark.parametrize("ngrid", [32, 96])
def test_aw_weights_two_tap_closed_form(ngrid):
    """AWProjectVisGridder.cc:618-770: weights = sum_z wt_z |F_z|^2, F_z = (inverse transform x
    cnx cny) of CF plane z placed about the centre of a cnx x cny buffer, then Fourier-padded to
    the grid.  A plane with taps a (centre) and b (offset (dx, dy)) gives
    |a + b exp(2 pi i (dx X + dy Y) / cnx)|^2, a band-limited function that the padding samples
    exactly on the finer grid; a second plane with a single tap adds a constant."""
    g = _aw_weights_gridder(ngrid)
    S, os_, cn = 3, 2, 32
    a, b, dx, dy = 0.7 + 0.1j, 0.2 - 0.3j, 2, -1
    planes = []
    for k in range(2 * os_ * os_):
        c = np.zeros((2 * S + 1, 2 * S + 1), complex)
        if k < os_ * os_:
            c[S, S], c[S + dy, S + dx] = a, b
        else:
            c[S, S] = 0.5
        planes.append(c)
    g.cf = O.CFCache(planes, os_)
    g.sumw = np.array([[[3.0, 2.0]]])
    w = g._aw_weights()[0, 0]
    X = (np.arange(ngrid) - ngrid // 2) * (cn / ngrid)
    ph = 2 * np.pi * (dx * X[None, :] + dy * X[:, None]) / cn
    exp = 3.0 * np.abs(a + b * np.exp(1j * ph)) ** 2 + 2.0 * 0.25
    assert np.max(np.abs(w - exp)) < 1e-6 * np.max(exp)
    wrong = 3.0 * np.abs(a + b * np.exp(-1j * ph)) ** 2 + 2.0 * 0.25
    assert np.max(np.abs(wrong - exp)) > 1e-1 * np.max(exp)   # the check can see the transform direction


def test_aw_weights_spheroidal_taper_is_divided_out():
    """With the spheroidal taper the weights divide |F|^2 by the taper squared where it exceeds
    the cutoff and are 0 elsewhere (AWProjectVisGridder.cc:638-667,722-737): a CF whose transform
    is the taper itself gives flat weights on the taper's support."""
    g = _aw_weights_gridder(32, taper=True)
    cn, c0 = 32, 16
    sph = O.Spheroidal(math.pi * 3, g.wp.alpha)
    f = sph(np.abs(np.arange(cn) - c0) / c0)
    t1 = O.interpolate_edge(np.where(f > 0.05, f, 0.0))
    t = np.outer(t1, t1)
    plane = np.zeros((cn + 1, cn + 1), complex)
    plane[:cn, :cn] = O.fft_centred(t.astype(complex), True) / (cn * cn)
    g.cf = O.CFCache([plane] * 4, 2)
    g.sumw = np.full((1, 1, 1), 5.0)
    w = g._aw_weights()[0, 0]
    assert 0 < np.count_nonzero(t > 0) < t.size
    assert np.max(np.abs(w - np.where(t > 0, 5.0, 0.0))) < 1e-5 * 5.0


def test_fft_pad_reproduces_a_band_limited_function():
    """Fourier padding (PaddingUtils::fftPad as used at AWProjectVisGridder.cc:765) samples the
    band-limited function of its input on the finer grid, centre to centre."""
    ny_in, nx_in, ny_out, nx_out = 16, 24, 40, 60

    def f(x, y):
        return 1.5 + np.cos(2 * np.pi * (3 * x / nx_in - 2 * y / ny_in) + 0.4)

    a = f((np.arange(nx_in) - nx_in // 2)[None, :], (np.arange(ny_in) - ny_in // 2)[:, None])
    out = O.fft_pad(a[None], ny_out, nx_out)[0]
    xo = (np.arange(nx_out) - nx_out // 2) * (nx_in / nx_out)
    yo = (np.arange(ny_out) - ny_out // 2) * (ny_in / ny_out)
    assert np.max(np.abs(out - f(xo[None, :], yo[:, None]))) < 1e-12
    assert np.array_equal(O.fft_pad(a, ny_in, nx_in), a)


# ---- w stacking composition (OracleWStack): the measurement equation with its w term ------------
def _wstack_case():
    N, cell, f = 128, math.radians(60 / 3600), np.array([1.0e9])
    lam, uvcell = O.C_LIGHT / f[0], 1 / (cell * N)
    ws = O.WSampling(2000.0, 9)
    ks = [(5, -7, 2), (-11, 3, 7), (20, 14, 4)]   # (pixel u, pixel v, w plane): on grid and on planes
    uvw = np.array([[ku * uvcell * lam, kv * uvcell * lam, ws.wterm(iw) * lam] for ku, kv, iw in ks])
    V = np.array([1.0 + 0.5j, -0.3 + 0.8j, 0.7 - 0.2j])
    return N, cell, f, uvcell, ws, ks, uvw, V


def test_wstack_image_is_the_direct_transform_with_the_w_term():
    """Each visibility goes to the grid of its w plane; the image sums the w-screened inverse
    transforms (WStackVisGridder.cc:134-171,255-289).  For visibilities on grid pixels and on
    plane w values the dirty image is Re sum V exp(-2 pi i (u l + v m + w (sqrt(1-l^2-m^2) - 1))),
    up to the spheroidal kernel's correction accuracy: the uv and w terms carry the same sign
    (a flipped screen fails at O(1))."""
    N, cell, f, uvcell, ws, ks, uvw, V = _wstack_case()
    o = O.OracleWStack(2000.0, 9, precision="f64")
    o.grid((N, N, 1, 1), (cell, cell), O.Chunk(uvw, f, V.reshape(3, 1, 1).astype(np.complex64)))
    img = o.finalise_grid()[0, 0]
    X = (np.arange(N) - N // 2) * cell
    l, m = X[None, :], X[:, None]
    n = np.sqrt(1 - l * l - m * m)
    exp = sum((v * np.exp(-2j * np.pi * (ku * uvcell * l + kv * uvcell * m + ws.wterm(iw) * (n - 1)))).real
              for (ku, kv, iw), v in zip(ks, V))
    q = slice(N // 4, 3 * N // 4)
    assert np.max(np.abs(img - exp)[q, q]) < 2e-3 * np.max(np.abs(exp))


def test_wstack_degrid_predicts_the_measurement_equation():
    """Degridding a point source of flux F at (l0, m0) predicts
    F exp(2 pi i (u l0 + v m0 + w (n0 - 1))) (WStackVisGridder.cc:314-327), the adjoint of the
    imaging above, to the kernel's correction accuracy."""
    N, cell, f, uvcell, ws, ks, uvw, V = _wstack_case()
    model = np.zeros((1, 1, N, N))
    X0, Y0, F = 17, -9, 2.0
    model[0, 0, N // 2 + Y0, N // 2 + X0] = F
    pv = O.OracleWStack(2000.0, 9, precision="f64").degrid(
        (N, N, 1, 1), (cell, cell), model, O.Chunk(uvw, f, V.reshape(3, 1, 1).astype(np.complex64)))
    l0, m0 = X0 * cell, Y0 * cell
    n0 = math.sqrt(1 - l0 * l0 - m0 * m0)
    for i, (ku, kv, iw) in enumerate(ks):
        e = F * np.exp(2j * np.pi * (ku * uvcell * l0 + kv * uvcell * m0 + ws.wterm(iw) * (n0 - 1)))
        assert abs(pv[i, 0, 0] - e) < 2e-3 * F


# ---- row selection: MaxPointingSeparation and the representative PSF feed / field -------------
def _two_field_chunk(seed=3, nrow=200):
    rng = np.random.default_rng(seed)
    N, cell, f = 128, math.radians(20 / 3600), np.array([1.4e9, 1.41e9])
    uvw = _disc(nrow, N, cell, f[0], rng)
    vis = (rng.normal(size=(nrow, 2, 1)) + 1j * rng.normal(size=(nrow, 2, 1))).astype(np.complex64)
    centre = (1.0, -0.5)
    far = rng.random(nrow) < 0.4                       # rows pointing 2 deg away
    point = np.tile(np.array(centre), (nrow, 1))
    point[far, 0] += math.radians(2.0) / math.cos(centre[1])
    feed = (rng.random(nrow) < 0.5).astype(np.int32)
    return N, cell, f, uvw, vis, centre, far, point, feed


def test_max_pointing_separation_rejects_whole_rows():
    """TableVisGridder.cc:571-578: a row whose pointing lies farther than the threshold from the
    image centre is skipped before anything else; gridding the chunk equals gridding the kept
    rows alone, and the rejected rows are counted once each."""
    N, cell, f, uvw, vis, centre, far, point, feed = _two_field_chunk()
    g = O.OracleGridder("SphFunc", "f64", sph_oversample=16)
    g.max_sep, g.centre = math.radians(1.0), centre
    g.initialise_grid((N, N, 1, 1), (cell, cell))
    g.grid(O.Chunk(uvw, f, vis, pointing1=point))
    keep = ~far
    h = O.OracleGridder("SphFunc", "f64", sph_oversample=16)
    h.initialise_grid((N, N, 1, 1), (cell, cell))
    h.grid(O.Chunk(np.ascontiguousarray(uvw[keep]), f, np.ascontiguousarray(vis[keep])))
    assert np.array_equal(g.grid_, h.grid_)
    assert g.rows.rows_rejected == np.count_nonzero(far)
    assert g.counters.samples_gridded == h.counters.samples_gridded
    # a threshold beyond the offset keeps every row
    g2 = O.OracleGridder("SphFunc", "f64", sph_oversample=16)
    g2.max_sep, g2.centre = math.radians(3.0), centre
    g2.initialise_grid((N, N, 1, 1), (cell, cell))
    g2.grid(O.Chunk(uvw, f, vis, pointing1=point))
    assert g2.rows.rows_rejected == 0


def test_psf_uses_the_first_rows_feed_and_field_across_calls():
    """TableVisGridder.cc:581-593,809-812: without alldatapsf the PSF grids only the rows with the
    feed and the dish pointing of the first row it reaches, and that choice persists over the
    following calls until the next initialiseGrid."""
    N, cell, f, uvw, vis, centre, far, point, feed = _two_field_chunk(seed=4)
    n = uvw.shape[0] // 2
    g = O.OracleGridder("SphFunc", "f64", sph_oversample=16, use_all_psf=False)
    g.initialise_grid((N, N, 1, 1), (cell, cell), dopsf=True)
    for sl in (slice(0, n), slice(n, None)):
        g.grid(O.Chunk(np.ascontiguousarray(uvw[sl]), f, np.ascontiguousarray(vis[sl]), feed1=feed[sl],
                       pointing1=np.ascontiguousarray(point[sl])))
    sel = (feed == feed[0]) & (far == far[0])
    h = O.OracleGridder("SphFunc", "f64", sph_oversample=16, use_all_psf=True)
    h.initialise_grid((N, N, 1, 1), (cell, cell), dopsf=True)
    h.grid(O.Chunk(np.ascontiguousarray(uvw[sel]), f, np.ascontiguousarray(vis[sel])))
    assert 0 < np.count_nonzero(sel) < sel.size
    assert np.array_equal(g.grid_, h.grid_)


def test_wstack_pcf_on_one_w0_plane_is_the_single_grid_preconditioner_over_n():
    # WStackVisGridder.cc:222-262 against TableVisGridder.cc:1109-1236 for the preconditioner:
    # with every sample at w = 0 (one plane, updatew's factor 3 / 3 = 1) the w-stacking image
    # is the real part of the normalised inverse transform of the one plane, where the single
    # grid's is that times nx * ny (neither corrects the preconditioner)
    rng = np.random.default_rng(8)
    N, cell, f = 64, math.radians(40 / 3600), np.array([1.0e9, 1.05e9])
    lam = O.C_LIGHT / f[-1]
    uvmax = 0.4 * (N / 2) / (cell * N) * lam
    uvw = np.zeros((40, 3))
    uvw[:, 0] = rng.uniform(-uvmax, uvmax, 40)
    uvw[:, 1] = rng.uniform(-uvmax, uvmax, 40)
    vis = np.zeros((40, 2, 1), np.complex64)
    noise = (0.5 + rng.random((40, 2, 1))).astype(np.float32).astype(np.complex64)
    ch = O.Chunk(uvw, f, vis, noise=noise)
    ws = O.OracleWStack(2000.0, 9, precision="f64")
    grids = ws.grid((N, N, 1, 1), (cell, cell), ch, dopcf=True)
    assert [iw for iw in range(9) if np.abs(grids[iw]).max() > 0] == [4] and ws.pcf_scale(4) == 1.0
    one = O.OracleGridder("SphFunc", "f64")
    one.initialise_grid((N, N, 1, 1), (cell, cell), dopcf=True)
    one.grid(ch)
    assert np.abs(one.grid_).max() > 0
    a, b = ws.finalise_grid(), one.finalise_grid() / (N * N)
    assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(b))


# ---- ATCAIllumination and the parallactic-angle / frequency cache policy ------------------------
def _atca(**kw):
    base = dict(illumination="ATCA", diameter=22.0, blockage=2.0)
    base.update(kw)
    return O.AWProjectParams(**base)


_ATCA_GEOM = dict(nu=256, nv=256, ucell=8.0, vcell=8.0, os_=4, freq=1.4e9)  # 0.428 m cells, R = 25.7


def _atca_pattern(prm, **kw):
    g = dict(_ATCA_GEOM)
    g.update(kw)
    return O.atca_pattern(g["nu"], g["nv"], g["ucell"], g["vcell"], g["os_"], g["freq"], prm,
                          kw.get("l", 0.0), kw.get("m", 0.0), kw.get("pa", 0.0))


def _atca_cell_m():
    g = _ATCA_GEOM
    return g["ucell"] / g["os_"] * (O.C_LIGHT / g["freq"])


def test_atca_without_taper_and_legs_is_the_disk_with_root_normalisation():
    """ATCAIllumination.cc:253-330 with every effect off: 1 on the annulus lattice (the disk's),
    scaled by nU nV / sqrt(count) (the disk divides by the count itself); a beam offset changes
    only the phase."""
    prm = _atca(tapering=False, feedlegs=False)
    g = _ATCA_GEOM
    disk = O.disk_pattern(g["nu"], g["nv"], g["ucell"], g["vcell"], g["os_"], g["freq"], 22.0, 2.0)
    pat = _atca_pattern(prm)
    count = np.count_nonzero(disk)
    assert np.array_equal(pat != 0, disk != 0)
    assert np.allclose(pat[pat != 0], g["nu"] * g["nv"] / math.sqrt(count), rtol=1e-6)
    off = _atca_pattern(prm, l=0.003, m=-0.001)
    assert np.allclose(np.abs(off[pat != 0]), g["nu"] * g["nv"] / math.sqrt(count), rtol=1e-6)
    assert not np.allclose(off, pat)


@pytest.mark.parametrize("kw", [dict(), dict(defocusing=1.3), dict(feedlegs=False, defocusing=-0.4),
                                dict(wedges_shadowing=(0.2,), pa=0.7, l=0.002, m=0.001)])
def test_atca_pattern_is_normalised_to_unit_power(kw):
    """The normalisation divides by sqrt(sum |value|^2) (ATCAIllumination.cc:320-330): whatever the
    taper, defocusing, shadows and beam offset, sum |pattern|^2 = (nU nV)^2."""
    geo = {k: kw.pop(k) for k in ("pa", "l", "m") if k in kw}
    pat = _atca_pattern(_atca(**kw), **geo).astype(np.complex128)
    n2 = float(_ATCA_GEOM["nu"] * _ATCA_GEOM["nv"]) ** 2
    assert abs(float(np.sum(np.abs(pat) ** 2)) / n2 - 1.0) < 1e-5


def test_atca_taper_is_continuous_and_the_defocusing_phase_is_quadratic():
    """simulateTapering (ATCAIllumination.h): the defocusing phase is the given value at the dish
    edge and grows quadratically with the radius; the Jamesian taper joins its inner and outer
    forms continuously at 0.4 of the radius (both normalised to 1 there, times 1 - 0.3 x 0.6) and
    is zero beyond the rim."""
    j = O.atca_jamesian(np.array([0.4 - 1e-12, 0.4, 0.4 + 1e-12, 1.0 + 1e-9, 3.0]))
    assert abs(j[0] - 0.82) < 1e-9 and abs(j[1] - 0.82) < 1e-12 and abs(j[2] - 0.82) < 1e-9
    assert j[3] == 0.0 and j[4] == 0.0
    phi = 1.3
    pat = _atca_pattern(_atca(feedlegs=False, defocusing=phi))
    n = _ATCA_GEOM["nu"]
    R = 11.0 / _atca_cell_m()
    for r in (3, 10, 18, 25):   # along the u axis; no beam offset, so the phase is the defocusing's
        v = pat[n // 2, n // 2 + r]
        expect = phi * r * r / (R * R)
        assert abs(np.angle(v) - expect) < 2e-6, r
    # the rim cell of the u axis carries (nearly) the full defocusing phase
    rim = int(math.floor(R))
    assert abs(np.angle(pat[n // 2, n // 2 + rim]) - phi) < 0.1


def _leg_geometry(ou, ov, t, half, r0, tan_a):
    """The shadow model in polar form, from the documented geometry: two legs through the centre at
    angles t (from u towards v) and t + 90 deg, each shadow a band of half width ``half``; a wedge
    on each leg arm from radius r0 outwards, the arm plus or minus the opening angle about its
    apex.  Returns (in a band, in an arm-t wedge, in an arm-(t+90) wedge)."""
    r = math.hypot(ou, ov)
    psi = math.atan2(ov, ou)
    along1, across1 = r * math.cos(psi - t), r * math.sin(psi - t)   # along / across the leg at t
    along2, across2 = r * math.cos(psi - t - math.pi / 2), r * math.sin(psi - t - math.pi / 2)
    band = abs(across1) < half or abs(across2) < half
    w1 = abs(along1) > r0 and abs(across1) < tan_a * (abs(along1) - r0)
    w2 = abs(along2) > r0 and abs(across2) < tan_a * (abs(along2) - r0)
    return band, w1, w2


@pytest.mark.parametrize("pa", [0.0, 0.37, -1.1])
def test_atca_feed_leg_shadows_follow_the_documented_geometry(pa):
    """simulateFeedLegShadows / simulateFeedLegWedges (ATCAIllumination.h): shadows of the given
    full width along the legs at rotation + pa, attenuated by the shadowing factor; wedge 1 (on
    the legs at that angle) and wedge 2 (on the orthogonal pair) multiply by shadowing x their
    own factor.  Without the taper every cell of the dish carries the product of the factors that
    cover it."""
    s, w1, w2 = 0.75, 0.6, 0.5
    rot = 0.3
    prm = _atca(tapering=False, feedlegs_rotation=rot, feedlegs_shadowing=s, wedges_shadowing=(w1, w2),
                feedlegs_width=1.8, wedges_startradius=3.5, wedges_angle=math.radians(15.0))
    pat = _atca_pattern(prm, pa=pa).astype(np.complex128)
    cell = _atca_cell_m()
    half, r0, tan_a = 0.9 / cell, 3.5 / cell, math.tan(math.radians(15.0))
    n = _ATCA_GEOM["nu"]
    unit = None
    counts = [0, 0, 0]
    for iv in range(n):
        for iu in range(n):
            v = pat[iv, iu]
            if v == 0:
                continue
            band, a1, a2 = _leg_geometry(iu - n / 2, iv - n / 2, rot + pa, half, r0, tan_a)
            f = (s if band else 1.0) * ((s * w1) if a1 else 1.0) * ((s * w2) if a2 else 1.0)
            counts[0] += band
            counts[1] += a1
            counts[2] += a2
            if unit is None and f == 1.0:
                unit = v
            assert abs(v.imag) < 1e-9 * abs(v.real)
            if unit is not None:
                assert abs(v.real / unit.real - f) < 1e-6, (iu, iv)
    # each kind of region is populated: bands of ~2 x (2 half) x 2 R cells, four wedges
    R = 11.0 / cell
    assert abs(counts[0] / (2 * 2 * half * 2 * R) - 1.0) < 0.2
    assert counts[1] > 50 and counts[2] > 50


def test_atca_pattern_rotates_with_the_parallactic_angle():
    """The pattern is defined in a frame turned by the parallactic angle: increasing pa by 90 deg
    turns the whole (asymmetric, s1 != s2) shadow pattern by 90 deg from u towards v, so that
    P(pa + 90)[(-dv, du)] = P(pa)[(du, dv)] on the lattice."""
    prm = _atca(defocusing=0.8, wedges_shadowing=(0.9, 0.2))
    n = _ATCA_GEOM["nu"]
    a = _atca_pattern(prm, pa=0.4)
    b = _atca_pattern(prm, pa=0.4 + math.pi / 2)
    iv, iu = np.nonzero(a)
    assert iv.min() > 0 and iu.min() > 0
    # cell (du, dv) = (iu - n/2, iv - n/2) -> (-dv, du): column n - iv, row iu
    rot = b[iu, n - iv]
    assert np.allclose(rot, a[iv, iu], rtol=1e-5, atol=0)
    assert not np.allclose(b[iv, iu], a[iv, iu])          # the pattern itself is not 4-fold symmetric
    sym = _atca_pattern(_atca(feedlegs=False), pa=0.4)
    assert np.array_equal(sym, _atca_pattern(_atca(feedlegs=False), pa=1.9))


def _aw_oracle(prm, pas=None, freq=(1.4e9,)):
    o = O.OracleGridder("AWProject", "f32", wproject=prm)
    return o, O.Chunk(np.zeros((len(pas), 3)), np.asarray(freq, dtype=np.float64),
                      np.zeros((len(pas), len(freq), 1), np.complex64), feed1_pa=np.asarray(pas, np.float32))


def test_cf_cache_validation_by_parallactic_angle_and_frequency():
    """AProjectGridderBase::validateCFCache (AProjectGridderBase.cc:202-253) as restated: the cache
    is dropped at the first row whose angle is off the cached one by more than patolerance, and
    that row's angle is cached; a symmetric illumination never checks angles; the frequency check
    runs only when the angle did not drop the cache, and frequencies are cached on every drop."""
    prm = _atca(patolerance=0.1, freqtolerance=1e-6)
    o, ch = _aw_oracle(prm, pas=[0.05, -0.08])
    o._validate_aw(ch)
    assert o.aw_invalid == (False, True) and o.aw_cf_pa == 0.0          # first chunk: frequencies
    ch.feed1_pa = np.array([0.05, 0.3, 0.9], np.float32)
    o.aw_done = {0}
    o._validate_aw(ch)
    assert o.aw_invalid == (True, False) and o.aw_cf_pa == np.float32(0.3) and not o.aw_done
    ch.feed1_pa = np.array([0.35, 0.25], np.float32)
    o.aw_done = {0}
    o._validate_aw(ch)                                                     # within tolerance of 0.3
    assert o.aw_invalid == (False, False) and o.aw_done == {0}
    ch.freq = np.array([1.4e9 * (1 + 5e-7)])
    o._validate_aw(ch)
    assert o.aw_invalid == (False, False)
    ch.freq = np.array([1.4e9 * (1 + 5e-6)])
    o._validate_aw(ch)
    assert o.aw_invalid == (False, True) and not o.aw_done
    ch.freq = np.array([1.5e9])
    ch.feed1_pa = np.array([1.0], np.float32)
    o._validate_aw(ch)                                                     # angle first; frequency cached
    assert o.aw_invalid == (True, False) and o.aw_cached_freq[0] == 1.5e9
    disk = O.AWProjectParams(diameter=12.0, blockage=1.0)
    od, chd = _aw_oracle(disk, pas=[0.0, 2.0])
    od._validate_aw(chd)
    assert od.aw_invalid == (False, True)
    chd.feed1_pa = np.array([3.0], np.float32)
    od._validate_aw(chd)
    assert od.aw_invalid == (False, False)
    nofreq = _atca(freqtolerance=None)
    on, chn = _aw_oracle(nofreq, pas=[0.0])
    on._validate_aw(chn)
    assert on.aw_invalid == (False, False) and on.aw_cached_freq is None


# ---- rotatedUVW / uvwRotationDelay (oracle.uvw_rotate) ------------------------------------------
def _uvw_case(seed=12, nrow=300):
    rng = np.random.default_rng(seed)
    uvw = rng.normal(size=(nrow, 3)) * np.array([3000.0, 3000.0, 800.0])
    p = (1.1, -0.6)
    t = (p[0] + math.radians(2.0), p[1] - math.radians(1.5))     # tangent point 2.5 deg away
    c = (t[0] + math.radians(0.4), t[1] + math.radians(0.3))     # image centre off the tangent point
    return uvw, p, t, c


def _lmn(d, ref):
    l, m = O.direction_lm(d, ref)
    return l, m, math.sqrt(1.0 - l * l - m * m)


def test_uvw_rotation_is_a_change_of_frame():
    """TableVisGridder.cc:481-495: the rotation keeps each baseline's length, is the identity when
    the tangent point is the phase centre, and composes (p -> t -> q equals p -> q)."""
    uvw, p, t, c = _uvw_case()
    rt, _ = O.uvw_rotate(uvw, p, t, c)
    assert np.allclose(np.linalg.norm(rt, axis=1), np.linalg.norm(uvw, axis=1), rtol=0, atol=1e-9)
    same, d0 = O.uvw_rotate(uvw, p, p, p)
    assert np.max(np.abs(same - uvw)) < 1e-9 and np.max(np.abs(d0)) < 1e-9
    q = (p[0] - 0.3, p[1] + 0.2)
    a, _ = O.uvw_rotate(rt, t, q, q)
    b, _ = O.uvw_rotate(uvw, p, q, q)
    assert np.max(np.abs(a - b)) < 1e-9
    # per-row phase centres: each row as its own one-row chunk
    pcs = np.stack([np.full(len(uvw), p[0]), np.full(len(uvw), p[1])], 1)
    pcs[::2, 0] += 0.01
    rows, drows = O.uvw_rotate(uvw, pcs, t, c)
    for i in (0, 1, 7, 8):
        one, d1 = O.uvw_rotate(uvw[i:i + 1], tuple(pcs[i]), t, c)
        assert np.allclose(rows[i], one[0], rtol=0, atol=1e-12) and abs(drows[i] - d1[0]) < 1e-12


def test_uvw_rotation_delay_rephases_to_the_image_centre():
    """The gridder multiplies conj(V) by exp(2 pi i f delay / c) (TableVisGridder.cc:652-655).  With
    visibilities phased to p, V = exp(2 pi i (u l + v m + w (n - 1)) / lambda) for a source at
    direction cosines (l, m, n) about p (the sign the degridding pins fix), a source at the image
    centre becomes exactly 1 and a source at d becomes
    exp(-2 pi i (u' (l_d - l_c) + v' (m_d - m_c) + w' (n_d - n_c)) / lambda) in the rotated uvw
    (u', v', w') with direction cosines about the tangent point: the delay rephases to the image
    centre and the rotation carries the tangent point's projection.  Direction cosines here come
    from the spherical-trigonometry formula, not from the frame vectors."""
    uvw, p, t, c = _uvw_case()
    lam = O.C_LIGHT / 1.2e9
    rt, delay = O.uvw_rotate(uvw, p, t, c)
    phasor = np.exp(2j * np.pi * delay / lam)

    def vis(d):
        l, m, n = _lmn(d, p)
        return np.exp(2j * np.pi * (uvw[:, 0] * l + uvw[:, 1] * m + uvw[:, 2] * (n - 1)) / lam)
    assert np.max(np.abs(phasor * np.conj(vis(c)) - 1.0)) < 1e-8
    d = (c[0] - math.radians(0.7), c[1] + math.radians(0.5))
    lc, mc, nc = _lmn(c, t)
    ld, md, nd = _lmn(d, t)
    expect = np.exp(-2j * np.pi * (rt[:, 0] * (ld - lc) + rt[:, 1] * (md - mc) + rt[:, 2] * (nd - nc)) / lam)
    assert np.max(np.abs(phasor * np.conj(vis(d)) - expect)) < 1e-8
    # a flipped delay fails at O(1)
    assert np.max(np.abs(np.exp(-2j * np.pi * delay / lam) * np.conj(vis(c)) - 1.0)) > 0.5


# ---- snapshot imaging (oracle.OracleSnapshot and its pieces) ------------------------------------
def test_wplane_fit_recovers_a_coplanar_snapshot_and_refits_only_beyond_tolerance():
    rng = np.random.default_rng(21)
    uv = rng.normal(size=(500, 2)) * 2000.0
    A, B = 1.37, -0.62
    uvw = np.column_stack([uv, A * uv[:, 0] + B * uv[:, 1]])
    a, b = O.wplane_fit(uvw)
    assert abs(a - A) < 1e-12 and abs(b - B) < 1e-12
    # least squares: the residual is orthogonal to u and v
    noisy = uvw.copy()
    noisy[:, 2] += rng.normal(size=500) * 30.0
    a, b = O.wplane_fit(noisy)
    r = noisy[:, 2] - a * noisy[:, 0] - b * noisy[:, 1]
    assert abs(r @ noisy[:, 0]) < 1e-6 * np.abs(noisy[:, 0]).sum() and abs(r @ noisy[:, 1]) < 1e-6 * np.abs(noisy[:, 1]).sum()
    assert O.wplane_fit(np.zeros((3, 3))) == (0.0, 0.0)
    s = O.OracleSnapshot(lambda: None, wtolerance=100.0)
    ch = O.Chunk(uvw, np.array([1e9]), np.zeros((500, 1, 1), np.complex64))
    assert s._update_plane(ch) == (True, pytest.approx((A, B)))
    s.A, s.B = A + 1e-3, B                                   # 2 m off at most: inside the tolerance
    changed, ab = s._update_plane(ch)
    assert not changed and ab == (A + 1e-3, B)


def test_slant_projection_round_trip_and_closed_form():
    """slant_inverse undoes slant_forward; at the tangent point nothing moves; far out the shift is
    -A (1 - n), -B (1 - n)."""
    rng = np.random.default_rng(2)
    l, m = rng.uniform(-0.2, 0.2, 200), rng.uniform(-0.2, 0.2, 200)
    for A, B in ((0.0, 0.0), (1.7, -0.4), (-2.5, 3.1)):
        x, y = O.slant_forward(l, m, A, B)
        n = np.sqrt(1 - l * l - m * m)
        assert np.allclose(x - l, -A * (1 - n), atol=1e-15) and np.allclose(y - m, -B * (1 - n), atol=1e-15)
        l2, m2 = O.slant_inverse(x, y, A, B)
        assert np.max(np.abs(l2 - l)) < 1e-12 and np.max(np.abs(m2 - m)) < 1e-12
        assert O.slant_forward(0.0, 0.0, A, B) == (0.0, 0.0)


def test_interpolation_reproduces_polynomials_and_zeroes_outside():
    """Bilinear reproduces linear functions; Keys' cubic convolution (a = -1/2) reproduces
    polynomials up to degree 2 (Keys 1981); stencils that leave the image give 0."""
    ny, nx = 20, 24
    yy, xx = np.mgrid[0:ny, 0:nx].astype(float)
    lin = 0.3 + 1.1 * xx - 0.7 * yy
    quad = lin + 0.05 * xx * xx - 0.02 * xx * yy + 0.03 * yy * yy
    rng = np.random.default_rng(4)
    px, py = rng.uniform(1.0, nx - 3.0, 300), rng.uniform(1.0, ny - 3.0, 300)
    fl = 0.3 + 1.1 * px - 0.7 * py
    fq = fl + 0.05 * px * px - 0.02 * px * py + 0.03 * py * py
    assert np.max(np.abs(O.interp2(lin, px, py, "linear") - fl)) < 1e-12
    assert np.max(np.abs(O.interp2(lin, px, py, "cubic") - fl)) < 1e-12
    assert np.max(np.abs(O.interp2(quad, px, py, "cubic") - fq)) < 1e-11
    assert O.interp2(lin, np.array([0.5]), np.array([5.0]), "cubic")[0] == 0.0
    assert O.interp2(lin, np.array([nx - 1.0]), np.array([5.0]), "linear")[0] == 0.0
    assert O.interp2(lin, np.array([3.0]), np.array([4.0]), "cubic")[0] == pytest.approx(lin[4, 3])


def test_snapshot_imaging_puts_a_wide_field_source_at_its_true_position():
    """Two coplanar snapshots (w = A u + B v exactly, different planes) of a 1 Jy source 500 pixels
    off the tangent point, imaged with the SphFunc kernel on the plane-relative w: each snapshot's
    image shows the source displaced by -(A, B) (1 - n) / cell (about 18 and 12 pixels here), and
    the regridded sum peaks at the true pixel with the summed weight (the reference's own bar,
    wtermtest.py: position within one cell, flux within 10%)."""
    N, cell, f = 1024, math.radians(20 / 3600), np.array([1.2e9])
    lam = O.C_LIGHT / f[0]
    kx, ky = 400, -300
    l0, m0 = kx * cell, ky * cell
    n0 = math.sqrt(1 - l0 * l0 - m0 * m0)
    rng = np.random.default_rng(17)
    uvmax = 0.4 * (N / 2) / (N * cell) * lam
    chunks, planes = [], [(1.5, 0.5), (-1.0, 1.2)]
    for A, B in planes:
        r = uvmax * np.sqrt(rng.random(800))
        th = 2 * np.pi * rng.random(800)
        u, v = r * np.cos(th), r * np.sin(th)
        uvw = np.column_stack([u, v, A * u + B * v])
        V = np.exp(2j * np.pi * (u * l0 + v * m0 + uvw[:, 2] * (n0 - 1)) / lam)
        chunks.append(O.Chunk(uvw, f, V.reshape(-1, 1, 1).astype(np.complex64)))
    s = O.OracleSnapshot(lambda: O.OracleGridder("SphFunc", "f64", sph_oversample=32), wtolerance=1.0)
    s.initialise_grid((N, N, 1, 1), (cell, cell))
    shifts = []
    for ch, (A, B) in zip(chunks, planes):
        s.grid(ch)
        raw = s.inner.finalise_grid()[0, 0]
        y, x = np.unravel_index(np.argmax(raw), raw.shape)
        shifts.append((x - (N // 2 + kx), y - (N // 2 + ky)))
        assert abs((x - N // 2) - (kx - A * (1 - n0) / cell)) <= 0.5 and abs((y - N // 2) - (ky - B * (1 - n0) / cell)) <= 0.5
    assert min(abs(d) for sh in shifts for d in sh) >= 5             # the slant is several pixels
    img = s.finalise_grid()[0, 0]
    assert np.unravel_index(np.argmax(img), img.shape) == (N // 2 + ky, N // 2 + kx)
    assert abs(img.max() / 1600.0 - 1.0) < 0.1
    assert s.fits == 2 and s.regrids == 4
    w = s.finalise_weights()[0, 0]
    assert abs(w[N // 2, N // 2] - 1600.0) < 1e-6
