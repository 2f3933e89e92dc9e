"""Oracle sampling pinned against the reference's known-answer tests.

Known answers restated as data from the reference tests:
Sobol (PBRT v4 b0bdf461): tests/sampling_tests.cpp:26-151;
Philox4x32-10 (Random123 kat_vectors): tests/render_tests.cpp:59-72;
CIE 1931 row 479 nm and Y normaliser: tests/render_tests.cpp:92-104.
"""
import math
import struct

import numpy as np
import pytest

from tests.golden.sobol_kat import KATS


def f(h):
    return float.fromhex(h)


@pytest.mark.parametrize("kat", KATS, ids=lambda k: f"m{k['m']}-dim{k['dimension']}")
def test_sobol_known_answers(orc, kat):
    seed_bits = kat["seed"] & 0xFFFFFFFF
    idx = orc.sobol_index(kat["m"], kat["x"], kat["y"], kat["sample"])
    assert idx == kat["index"]
    assert orc.L.or_dimension_hash(kat["dimension"], seed_bits) == kat["hash"]
    for mode in range(4):
        assert orc.sobol_word(idx, kat["dimension"], mode, seed_bits) == kat["words"][mode]
    for dim, pixel in enumerate((kat["x"], kat["y"])):
        g = orc.sobol_word(idx, dim, 0, seed_bits)
        assert orc.L.or_camera_uniform(g, kat["m"], pixel) == f(kat["camera"][dim])


def test_sobol_uniform_conversion(orc):
    # tests/sampling_tests.cpp:139-144
    assert orc.L.or_sobol_uniform(0xFFFFFFFF) == f("0x1.fffffep-1")
    assert orc.L.or_sobol_uniform(0) == 0.0
    assert orc.L.or_sobol_uniform(0x80000081) == f("0x1.000002p-1")
    assert orc.L.or_camera_uniform(0x7FFFFFFF, 1, 0) == f("0x1.fffffep-1")


def test_sobol_interval_properties(orc):
    # Interval index keeps the absolute frame and maps into the requested pixel
    # (tests/sampling_tests.cpp:153-172).
    for m in (1, 3, 7, 11, 16, 25):
        scale = 1 << m
        last = (1 << (52 - 2 * m)) - 1
        for x in (0, scale // 2, scale - 1):
            for y in (0, scale // 2, scale - 1):
                for frame in (0, last // 2, last):
                    i = orc.sobol_index(m, x, y, frame)
                    assert i >> (2 * m) == frame
                    assert orc.sobol_word(i, 0, 0, 0) >> (32 - m) == x
                    assert orc.sobol_word(i, 1, 0, 0) >> (32 - m) == y


@pytest.mark.parametrize("mode", range(4))
def test_sobol_power_of_two_prefix_stratified(orc, mode):
    # tests/sampling_tests.cpp:174-188: 256 points fill every 8-bit bin once.
    for dim in (2, 163, 1023):
        bins = np.zeros(256, int)
        for i in range(256):
            bins[orc.sobol_word(i, dim, mode, 0xFFFFFFFF) >> 24] += 1
        assert (bins == 1).all()


def test_philox_known_answers(orc):
    assert orc.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    m = 0xFFFFFFFF
    assert orc.philox([m, m, m, m], [m, m]) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert orc.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == [
        0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]
    u = orc.L.or_philox_u01
    assert u(0) == 0 and u(255) == 0 and u(256) == 2.0 ** -24 and u(0xFFFFFFFF) == 1 - 2.0 ** -24


def test_cie_table(orc):
    import ctypes as C

    row = (C.c_float * 3)()
    orc.L.or_cie_row(479 - 360, row)
    for got, want in zip(row, (0.1042979, 0.1334528, 0.8566193)):
        assert got == np.float32(want)
    ys = np.array([_cie_y(orc, i) for i in range(471)], dtype=np.float64)
    # Y normaliser of the reference (tests/render_tests.cpp:94), trapezoid over 1 nm
    trap = ys.sum() - 0.5 * (ys[0] + ys[-1])
    assert abs(trap - 106.8569149167669) < 2e-5


def _cie_y(orc, i):
    import ctypes as C

    row = (C.c_float * 3)()
    orc.L.or_cie_row(i, row)
    return float(row[1])


def test_hero_uniform_wavelengths(orc):
    # src/hero_spectrum_shared.h:225-244: four lanes cyclically shifted by 1/4
    for u in (0.0, 0.1, 0.24999, 0.5, 0.75, f("0x1.fffffep-1")):
        lam, pdf = orc.wavelengths(np.float32(u), 0)
        shifted = sorted(((np.float32(u) + np.float32(k * 0.25)) % 1.0) for k in range(4))
        assert all(360 <= l < 830 for l in lam)
        assert all(p == np.float32(1 / 470) for p in pdf)
        assert lam[0] == np.float32(360 + 470 * np.float32(u)) or lam[0] == np.nextafter(np.float32(830), 0)
    lam, pdf = orc.wavelengths(float("nan"), 0)
    assert all(math.isnan(x) for x in lam)


def test_hero_visible_wavelengths_pdf_normalised(orc):
    # PBRT visible-wavelength pdf integrates to ~1 over [360, 830]
    lam = np.linspace(360, 830, 47001)
    pdf = 0.0039398042 / np.cosh(0.0072 * (lam - 538)) ** 2
    assert abs(np.trapezoid(pdf, lam) - 1) < 2e-3
    for u in (0.01, 0.3, 0.77):
        l, p = orc.wavelengths(u, 1)
        for lk, pk in zip(l, p):
            assert 360 <= lk <= 830
            assert abs(pk - 0.0039398042 / math.cosh(0.0072 * (lk - 538)) ** 2) < 1e-6
