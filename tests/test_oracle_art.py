"""The ART oracle (CPU): HEALPix directions against the ring-scheme
definition and known pixel centres, the recombination fits against textbook
coefficients, the ray tracer against closed forms (one column through a
uniform medium; total path length per direction), the step-function factors
against quadrature, and the library's own direction set and factors against
the oracle's (host code only, no GPU)."""
import ctypes as C
import math

import numpy as np
import pytest
from scipy import integrate

import oracle as O

EV = 1.6022e-12


def test_healpix_known_centres():
    # nside = 1: pixel 0 at (acos(2/3), pi/4), pixel 4 on the equator at phi = 0,
    # pixel 11 at (acos(-2/3), 7 pi/4)
    th, ph = O.pix2ang_ring(1, 0)
    assert abs(th - math.acos(2 / 3)) < 1e-15 and abs(ph - math.pi / 4) < 1e-15
    th, ph = O.pix2ang_ring(1, 4)
    assert abs(th - math.pi / 2) < 1e-15 and abs(ph) < 1e-15
    th, ph = O.pix2ang_ring(1, 11)
    assert abs(th - math.acos(-2 / 3)) < 1e-15 and abs(ph - 7 * math.pi / 4) < 1e-15
    # nside = 2: first pixel of the north cap, z = 1 - 1/(3 nside^2)
    th, ph = O.pix2ang_ring(2, 0)
    assert abs(math.cos(th) - (1 - 1 / 12)) < 1e-15 and abs(ph - math.pi / 4) < 1e-15


@pytest.mark.parametrize("nside", [1, 2, 4, 8])
def test_healpix_ring_structure(nside):
    npix = 12 * nside * nside
    ang = np.array([O.pix2ang_ring(nside, i) for i in range(npix)])
    z = np.cos(ang[:, 0])
    # rings: 4i pixels at z = 1 - i^2/(3 N^2) in the caps, 4N per ring in the belt
    zr, counts = np.unique(np.round(z, 12), return_counts=True)
    assert len(zr) == 4 * nside - 1
    for i in range(1, nside):
        zi = 1 - i * i / (3 * nside * nside)
        assert counts[np.argmin(abs(zr - zi))] == 4 * i
        assert counts[np.argmin(abs(zr + zi))] == 4 * i
    for i in range(nside, 3 * nside + 1):
        zi = (2 * nside - i) * 2 / (3 * nside)
        assert counts[np.argmin(abs(zr - zi))] == 4 * nside
    # equal areas: the directions are balanced, sum of unit vectors = 0
    v = np.stack([np.sin(ang[:, 0]) * np.cos(ang[:, 1]),
                  np.sin(ang[:, 0]) * np.sin(ang[:, 1]), z], axis=1)
    assert np.abs(v.sum(0)).max() < 1e-12
    # the second moment of a uniform sphere: <z^2> = 1/3 for equal-area pixels
    assert abs(np.mean(z * z) - 1 / 3) < 0.04 / nside ** 2
    assert ang[:, 1].min() >= 0 and ang[:, 1].max() < 2 * math.pi


def test_angle_classes():
    ovr, base, corner, cnum = O.art_angles(8, sorted_=False)
    assert ovr.shape == (768, 3)
    np.testing.assert_allclose(np.linalg.norm(ovr.astype(np.float64), axis=1), 1.0, atol=1e-6)
    a = np.abs(ovr)
    top = a.max(1) - 1e-6          # ties (|x| = |y| at phi = 3 pi/4 ...) count for both
    # base 0/1: z dominant (+/-), 2/3: x, 4/5: y
    assert np.all((base <= 1) <= (a[:, 2] >= top))
    assert np.all((base == 0) <= (ovr[:, 2] > 0)) and np.all((base == 1) <= (ovr[:, 2] < 0))
    assert np.all(((base == 2) | (base == 3)) <= (a[:, 0] >= top))
    assert np.all((base == 2) <= (ovr[:, 0] > 0)) and np.all((base == 3) <= (ovr[:, 0] < 0))
    assert np.all(((base == 4) | (base == 5)) <= (a[:, 1] >= top))
    assert np.all((base == 4) <= (ovr[:, 1] > 0)) and np.all((base == 5) <= (ovr[:, 1] < 0))
    assert cnum.sum() == 768 and np.all(np.bincount(corner, minlength=8) == cnum)
    # sorting keeps the set and orders by corner
    ovr_s, base_s, corner_s, _ = O.art_angles(8, sorted_=True)
    assert np.all(np.diff(corner_s) >= 0)
    key = lambda o, b: sorted(map(tuple, np.c_[o, b]))
    assert key(ovr, base) == key(ovr_s, base_s)


def rmesh_single(T=1.0e4, x=0.5, a=0.5, b=0.5, dens=1.0, helium_fact=0.0811258278):
    uenetok = 1.0
    fHII = x
    fHeII, fHeIII = a * (1 - b), a * b
    felec = fHII + helium_fact * (fHeII + 2 * fHeIII)
    wmol = np.float32(1.0 / ((1.0 + np.float32(felec)) * 0.755 + 0.25 * 0.245))
    uene = np.float32(T / (uenetok * float(wmol)))
    chem = np.array([[felec, 1 - x, x, 1 - a, fHeII, fHeIII, uene]], np.float32)
    kap, src = O.art_rmesh(np.array([dens], np.float32), chem, 1e-3, uenetok, helium_fact)
    nH = dens * 1e-3
    ne = nH * (np.float32(x) + helium_fact * (np.float32(fHeII) + 2 * np.float32(fHeIII)))
    return kap[0], src[0], nH, ne, chem[0]


def test_recombination_fits_match_textbook_coefficients():
    """Source function S_b = alpha_1 n_e n_ion h nu_b / (4 pi kappa_b): the
    ground-state recombination coefficient alpha_A - alpha_B of the fits at
    1e4 K against Osterbrock & Ferland (2006) tables 2.1/2.4: H 4.18e-13 -
    2.59e-13, He+ 4.27e-13 - 2.72e-13 (the fits differ by a few percent)."""
    hf = 0.0811258278
    kap, src, nH, ne, ch = rmesh_single(helium_fact=hf)
    alpha_H = src[0] * kap[0] * 4 * math.pi / (ne * ch[2] * nH * 13.6 * EV)
    assert abs(alpha_H / (4.18e-13 - 2.59e-13) - 1) < 0.03, alpha_H
    alpha_He = src[1] * kap[1] * 4 * math.pi / (ne * ch[4] * hf * nH * 13.6 * 1.8125 * EV)
    assert abs(alpha_He / (4.27e-13 - 2.72e-13) - 1) < 0.12, alpha_He
    # absorption: sigma at the band edges times the absorber densities
    assert abs(kap[0] / (6.3e-18 * ch[1] * nH) - 1) < 1e-5


def test_recombination_fits_equal_the_reference_build():
    """The six fits of src/diffuse_photon/diffuse_chemistry.c:13-97 (k02,
    k04, k06, case A and B) bit for bit against the reference's own file
    compiled unmodified (oracle/_ref/libref_art.so), over 10 K to 1e9 K
    including both sides of k04_A's 0.8 eV and k04_B's 500.3 K branches."""
    ref = O.ref_art()
    if ref is None:
        pytest.skip("oracle/_ref/libref_art.so not built")
    ts = np.concatenate([np.logspace(1.0, 9.0, 4001),
                         [500.3, np.nextafter(500.3, 0), np.nextafter(500.3, 1e9),
                          0.8 / 8.61623e-5, np.nextafter(0.8 / 8.61623e-5, 0),
                          np.nextafter(0.8 / 8.61623e-5, 1e9), 1.0e4, 157807.0]])
    out = np.zeros(6)
    for t in ts:
        ref.ref_art_recomb(float(t), out)
        mine = O.art_recomb(t)
        assert np.array_equal(mine.view(np.uint64), out.view(np.uint64)), (t, mine, out)


@pytest.mark.parametrize("heatcool", [False, True])
def test_step_factors_against_quadrature(heatcool):
    f = O.art_step_factors(heatcool)
    h = 6.626e-27
    for s, (unit, nuL, cs) in enumerate([(1.0, 3.2885e15, O.lib().orc_csect_HI),
                                         (1.8125, 5.9483e15, O.lib().orc_csect_HeI),
                                         (4.0, 1.3154e16, O.lib().orc_csect_HeII)]):
        g = lambda nu: 4 * math.pi * cs(nu) / (h * nu * nuL)
        if not heatcool:
            # one bin at the edge: the ionization factor is the edge value
            assert abs(f[2 * s] / g(unit + 1e-7) - 1) < 1e-6
            continue
        w = 0.0632353
        ion = integrate.quad(g, unit, unit + w)[0] / w
        heat = h * nuL * integrate.quad(lambda nu: g(nu) * (nu - unit), unit, unit + w)[0] / w
        assert abs(f[2 * s] / ion - 1) < 2e-3, (s, f[2 * s], ion)
        assert abs(f[2 * s + 1] / heat - 1) < 0.05, (s, f[2 * s + 1], heat)


def test_single_column_uniform_medium_closed_form():
    """Rays along +z through a uniform medium: per cell I_cell = S (1 -
    e^{-k l}) (1 - e^{-k l})^... exactly: entering intensity after k cells is
    S (1 - e^{-k kappa l}), absorbed fraction per cell 1 - e^{-kappa l}."""
    n = 8
    g = O.art_grid(n, 2.0e22, 1e-3, 0.08)
    ncell = n ** 3
    kap = np.full((ncell, 3), [3e-22, 1e-21, 5e-23], np.float32)
    src = np.full((ncell, 3), [2.0, 0.5, 7.0], np.float32)
    length, I, nc = O.art_direction(g, np.float32([1e-31, 1e-31, 1.0]), 0, kap, src)
    ell = np.float32(np.float32(1 / n) * np.float32(2.0e22))
    assert nc == n * n * n
    np.testing.assert_array_equal(length, np.full(ncell, ell, np.float32))
    k = np.arange(n)
    for b in range(3):
        t = float(kap[0, b]) * float(ell)
        expect = src[0, b] * (1 - np.exp(-k * t)) * (1 - np.exp(-t))
        got = I.reshape(n, n, n, 3)[:, :, :, b]
        np.testing.assert_allclose(got, np.broadcast_to(expect, (n, n, n)), rtol=2e-5,
                                   atol=1e-12 * src[0, b])


def test_total_path_length_per_direction():
    """Every ray crosses the box along its dominant axis exactly once
    (side faces wrap), so the path lengths sum to N_face L / |dominant|."""
    n = 8
    g = O.art_grid(n, 1.0, 1e-3, 0.08)
    kap = np.full((n ** 3, 3), 1e-3, np.float32)
    src = np.zeros((n ** 3, 3), np.float32)
    ovr, base, _, _ = O.art_angles(2)
    for d in range(len(base)):
        length, _, nc = O.art_direction(g, ovr[d], base[d], kap, src)
        dom = [2, 2, 0, 0, 1, 1][base[d]]
        expect = n * n / abs(float(ovr[d, dom]))
        assert abs(length.astype(np.float64).sum() / expect - 1) < 2e-5, (d, base[d])
        assert length.min() > 0            # every cell is visited in every direction
        assert nc >= n ** 3


def test_library_directions_and_factors_match_oracle():
    """setup_diffuse_photon is host code: the library's sorted directions and
    factors equal the oracle's bit for bit."""
    from argot_amd import abi
    from argot_amd._lib import lib
    L = lib()
    L.argot_hip_set_diffuse_nside(8)
    L.argot_hip_set_diffuse_heatcool(1)
    hd = abi.HostDiffuseParam()
    L.init_hd_param(C.byref(hd))
    L.setup_diffuse_photon(C.byref(hd), C.byref(abi.RunParam()))
    ovr, base, corner, cnum = O.art_angles(8)
    got = np.array([[hd.angle[i].xovr, hd.angle[i].yovr, hd.angle[i].zovr] for i in range(768)],
                   np.float32)
    np.testing.assert_array_equal(got, ovr)
    np.testing.assert_array_equal([hd.angle[i].base_id for i in range(768)], base)
    np.testing.assert_array_equal([hd.angle[i].corner_id for i in range(768)], corner)
    np.testing.assert_array_equal(hd.corner_id_num[:], cnum)
    sf = hd.step_fact[0]
    np.testing.assert_array_equal(
        np.float32([sf.HI[0], sf.HI[1], sf.HeI[0], sf.HeI[1], sf.HeII[0], sf.HeII[1]]),
        O.art_step_factors(True))
    L.free_hd_param(C.byref(hd))
