"""Frequency grid, blackbody photon rates and cross sections (CPU).

Pinned three ways: against the reference's own table sources compiled
unmodified into oracle/_ref (bit for bit), against independent quadrature of
the Planck photon spectrum, and against textbook threshold cross sections.
The library's tables (setup_freq_param / setup_photon_rate /
setup_cross_section of libargot_hip.so) are held to the same bits."""
import ctypes as C
import os

import numpy as np
import pytest
from scipy import integrate

import oracle as O
from argot_amd import abi
from argot_amd import radiation as R
from argot_amd._lib import lib

REF_IEEE = os.path.join(O.REF_DIR, "libref_tables_ieee.so")
REF_FAST = os.path.join(O.REF_DIR, "libref_tables_fastmath.so")


def ref_tables(path, temps_rates):
    """Run the reference's setup_freq_param / setup_photon_rate /
    setup_cross_section from a compiled oracle/_ref library."""
    L = C.CDLL(path)
    L.setup_freq_param.argtypes = [C.POINTER(abi.FreqParam)]
    L.setup_photon_rate.argtypes = [C.POINTER(abi.FreqParam), C.POINTER(abi.RadiationSrc),
                                    C.c_double]
    L.setup_cross_section.argtypes = [C.POINTER(abi.CrossSection), C.POINTER(abi.FreqParam)]
    freq = abi.FreqParam()
    L.setup_freq_param(C.byref(freq))
    cs = (abi.CrossSection * abi.NGRID_NU)()
    L.setup_cross_section(cs, C.byref(freq))
    spectra = []
    for t, q in temps_rates:
        s = abi.RadiationSrc()
        s.type = 0
        s.param = t
        L.setup_photon_rate(C.byref(freq), C.byref(s), q)
        spectra.append(np.ctypeslib.as_array(s.photon_rate).copy())
    nu = np.ctypeslib.as_array(freq.nu).copy()
    csa = np.array([[c.csect_HI, c.csect_HeI, c.csect_HeII] for c in cs], np.float32)
    return nu, csa, spectra, freq.dlog_nu


CASES = [(1.0e5, 5.0e48), (1.0e5, 1.0e47), (3.0e4, 2.5e49), (1.0e5, 1.0e50)]


def need_ref(path):
    if not os.path.exists(path):
        pytest.skip(f"{path} not built (needs the reference sources: make -C oracle ref)")


def oracle_tables():
    nu, dlog = O.freq_grid()
    cs = O.cross_sections(nu)
    spectra = [O.photon_rate_bb(t, q) for t, q in CASES]
    return nu, cs, spectra, dlog


def product_tables():
    freq, cs = R.freq_and_csect()
    nu = np.ctypeslib.as_array(freq.nu).copy()
    csa = np.array([[c.csect_HI, c.csect_HeI, c.csect_HeII] for c in cs], np.float32)
    spectra = []
    for t, q in CASES:
        s = abi.RadiationSrc()
        s.type = 0
        s.param = t
        lib().setup_photon_rate(C.byref(freq), C.byref(s), q)
        spectra.append(np.ctypeslib.as_array(s.photon_rate).copy())
    return nu, csa, spectra, freq.dlog_nu


def bits_equal(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return a.dtype == b.dtype and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def test_oracle_tables_match_reference_build():
    need_ref(REF_IEEE)
    rnu, rcs, rspec, rd = ref_tables(REF_IEEE, CASES)
    nu, cs, spec, d = oracle_tables()
    assert rd == d
    assert bits_equal(rnu, nu)
    assert bits_equal(rcs, cs)
    for a, b in zip(rspec, spec):
        assert bits_equal(a, b)


def test_library_tables_match_reference_build():
    need_ref(REF_IEEE)
    rnu, rcs, rspec, rd = ref_tables(REF_IEEE, CASES)
    nu, cs, spec, d = product_tables()
    assert rd == d
    assert bits_equal(rnu, nu)
    assert bits_equal(rcs, cs)
    for a, b in zip(rspec, spec):
        assert bits_equal(a, b)


def test_reference_fastmath_build_differs_only_in_rounding():
    """The reference's own -O3 -ffast-math build of the same sources."""
    need_ref(REF_FAST)
    rnu, rcs, rspec, _ = ref_tables(REF_FAST, CASES)
    nu, cs, spec, _ = oracle_tables()
    np.testing.assert_allclose(rnu, nu, rtol=1e-14)
    np.testing.assert_allclose(rcs, cs, rtol=1e-6)
    for a, b in zip(rspec, spec):
        np.testing.assert_allclose(a, b, rtol=1e-9)


def test_frequency_grid_closed_form():
    nu, d = O.freq_grid()
    assert d == 1.5 / 256
    k = np.arange(256)
    np.testing.assert_allclose(np.log10(nu), (k + 0.5) * d, rtol=0, atol=1e-15)


def test_blackbody_bins_against_quadrature():
    """Each bin holds Q * int_bin B / int_[1, 10^1.5] B of the photon-number
    Planck shape nu^2/(exp(h nu/kT) - 1).  Against adaptive quadrature
    (scipy): bin-to-bin ratios to 1e-7 (narrow bins, where the midpoint rule
    converges), and the normalisation to the few-1e-5 error of the
    reference's at-most-five-level midpoint rule over the whole range."""
    h, k, nuL = 6.626e-27, 1.38066e-16, 3.28e15
    # (temperature, bin-ratio tolerance, normalisation tolerance): the
    # whole-range midpoint rule stops after five levels, which is 2e-5 off at
    # 1e5 K and several percent off at 3e4 K, where the spectrum is steeper.
    for temp, rtol, ntol in ((1.0e5, 5e-6, 1e-4), (3.0e4, 5e-5, 0.1)):
        q = 5.0e48
        x = h * nuL / (k * temp)
        f = lambda nu: nu * nu / np.expm1(x * nu)
        total, _ = integrate.quad(f, 1.0, 10 ** 1.5, epsabs=0, epsrel=1e-12, limit=200)
        spec = O.photon_rate_bb(temp, q)
        edges = 10 ** (np.arange(257) * (1.5 / 256))
        b0, _ = integrate.quad(f, edges[0], edges[1], epsabs=0, epsrel=1e-12)
        for i in (17, 100, 200, 255):
            b, _ = integrate.quad(f, edges[i], edges[i + 1], epsabs=0, epsrel=1e-12)
            assert abs((spec[i] / spec[0]) / (b / b0) - 1) < rtol, (temp, i)
        assert abs(spec[0] / (q * b0 / total) - 1) < ntol
        # sum of the bins = Q up to the same normalisation error
        assert abs(spec.sum() / q - 1) < ntol


def test_threshold_cross_sections():
    """Hydrogenic sigma_HI(13.6 eV) = 6.30e-18 cm^2, sigma_HeII(54.4 eV) =
    6.3e-18/Z^2 = 1.575e-18 cm^2; the HeI fit gives 7.83e-18 cm^2 at 24.6 eV;
    all vanish below their edges and the hydrogenic form falls off as nu^-7/2
    far above the edge."""
    eps = 1e-9
    assert abs(O.lib().orc_csect_HI(1.0 + eps) / 6.3e-18 - 1) < 1e-3
    assert abs(O.lib().orc_csect_HeII(4.0 + eps) / 1.575e-18 - 1) < 1e-3
    assert abs(O.lib().orc_csect_HeI(1.8125) / 7.83e-18 - 1) < 1e-12
    assert O.lib().orc_csect_HI(0.999) == 0.0
    assert O.lib().orc_csect_HeI(1.8) == 0.0
    assert O.lib().orc_csect_HeII(3.99) == 0.0
    s1, s2 = O.lib().orc_csect_HI(1.0e6), O.lib().orc_csect_HI(2.0e6)
    assert abs((s2 / s1) / 2.0 ** -3.5 - 1) < 0.01
