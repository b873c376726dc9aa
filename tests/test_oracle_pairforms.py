"""The short-range pair forms of the ionic field in the oracle (TwoBody.h
restated in oracle/akmc_oracle.c), pinned on a dimer: energies against
their closed forms at chosen separations, and forces against central
differences of the energy -- which pins the sign and size of each
reference force expression, including where it is not -dV/dr:
Lennard-Jones forces come out as +dV/dr (TwoBody.h:240-254), and the
energy-only Morse form adds C12/r^12 that the force form leaves out
(TwoBody.h:107-115 vs :229-238).  The 7-6 force expression
(TwoBody.h:257-273) is neither -dV/dr nor its negative and has no closed
form to pin: parity unpinned beyond this restatement."""
import math

import numpy as np
import pytest

LAT = np.diag([30.0, 30.0, 30.0]).reshape(9)


def _dimer(oracle, pot, r, energy_path=False):
    f = oracle.ion_field(["A"], [pot], rc=10.0, ewald=None, lat=LAT)
    f.energy_path = int(energy_path)
    pos = np.array([[5.0, 5.0 + r], [5.0, 5.0], [5.0, 5.0]])
    out = oracle.ion_forces(f, pos, np.zeros(2, np.int32), np.zeros(2), LAT)
    return out["e_vdw"], out["force"][0, 1]


def _fd(oracle, pot, r, h=1e-6):
    return (_dimer(oracle, pot, r + h)[0] - _dimer(oracle, pot, r - h)[0]) / (2 * h)


def test_bhm_closed_form_and_gradient(oracle):
    pot = ("bhm", "A", "A", 1500.0, 3.2, 0.9, 25.0, 11.0)
    r = 2.3
    e, f = _dimer(oracle, pot, r)
    assert e == pytest.approx(1500.0 * math.exp(3.2 * (0.9 - r)) - 25.0 / r ** 6 - 11.0 / r ** 8, rel=1e-13)
    assert f == pytest.approx(-_fd(oracle, pot, r), rel=1e-7)


def test_morse_force_path_and_energy_path(oracle):
    pot = ("morse", "A", "A", 0.35, 1.4, 2.6, 7.0)
    for r in (2.2, 2.6, 3.4):
        e, f = _dimer(oracle, pot, r)
        assert e == pytest.approx(0.35 * (1 - math.exp(-1.4 * (r - 2.6))) ** 2 - 0.35, abs=1e-14)
        assert f == pytest.approx(-_fd(oracle, pot, r), rel=1e-7, abs=1e-10)
        ep, _ = _dimer(oracle, pot, r, energy_path=True)
        assert ep - e == pytest.approx(7.0 / r ** 12, rel=1e-12)


def test_lennard_jones_minimum_and_reference_sign(oracle):
    eps, sig = 0.0104, 3.4
    pot = ("ljones", "A", "A", eps, sig)
    rmin = 2 ** (1 / 6) * sig
    e, f = _dimer(oracle, pot, rmin)
    assert e == pytest.approx(-eps, rel=1e-13) and abs(f) < 1e-14
    for r in (3.3, 4.5):
        e, f = _dimer(oracle, pot, r)
        assert f == pytest.approx(+_fd(oracle, pot, r), rel=1e-7)   # the reference's sign


def test_seven_six_energy(oracle):
    eps, sig = 0.2, 2.8
    pot = ("severnsix", "A", "A", eps, sig)
    e, _ = _dimer(oracle, pot, sig)
    assert e == pytest.approx(-eps, rel=1e-14)
    e, _ = _dimer(oracle, pot, 3.5)
    assert e == pytest.approx(eps * (sig / 3.5) ** 7 - 2 * eps * (sig / 3.5) ** 6, rel=1e-13)


def _dsf_dimer(oracle, r, alpha=0.3, rc=9.0, q=(1.5, -2.0)):
    f = oracle.ion_field(["A", "B"], [("A", "B", 0.0, 1.0, 0.0)], rc=rc, ewald=None, lat=LAT)
    oracle.set_damped(f, alpha, rc)
    pos = np.array([[5.0, 5.0 + r], [5.0, 5.0], [5.0, 5.0]])
    out = oracle.ion_forces(f, pos, np.array([0, 1], np.int32), np.array(q), LAT)
    return out["e_real"], out["force"][0, 1]


def test_damped_mesh_is_abramowitz_stegun_erfc(oracle):
    """the mesh holds erfc(alpha r)/r to the A&S 7.1.26 bound (1.5e-7 on erfc)"""
    from scipy.special import erfc
    e, f = oracle.set_damped(oracle.ion_field(["A"], [("A", "A", 0.0, 1.0, 0.0)], rc=9.0, ewald=None, lat=LAT),
                             0.3, 9.0)
    r = np.arange(1, e.size) * 9.0 / (e.size - 4)
    assert np.all(np.abs(e[1:] * r - erfc(0.3 * r)) <= 1.5e-7)
    ex = 2 * 0.3 / math.sqrt(math.pi) * np.exp(-(0.3 * r) ** 2)
    assert np.allclose(f[1:], (e[1:] + ex) / r ** 2, rtol=1e-14)


def test_damped_shifted_coulomb_closed_form(oracle):
    """E = k qi qj [erfc(a r)/r - erfc(a rc)/rc + (r - rc)(erfc(a rc)/rc^2 + 2a/sqrt(pi) e^{-a^2 rc^2}/rc)]
    (the damped-shifted-force Coulomb), its force -dE/dr, and both vanish at rc"""
    from scipy.special import erfc
    a, rc, k = 0.3, 9.0, 14.3997584 * 1.5 * -2.0

    def exact(r):
        s = erfc(a * rc) / rc ** 2 + 2 * a / math.sqrt(math.pi) * math.exp(-(a * rc) ** 2) / rc
        return k * (erfc(a * r) / r - erfc(a * rc) / rc + (r - rc) * s)

    for r in (2.1, 4.37, 7.9):
        e, f = _dsf_dimer(oracle, r)
        assert e == pytest.approx(exact(r), abs=2e-6 * abs(k))
        h = 1e-5
        assert f == pytest.approx(-(exact(r + h) - exact(r - h)) / (2 * h), abs=1e-5 * abs(k))
    e, f = _dsf_dimer(oracle, rc - 1e-6)
    assert abs(e) < 1e-6 and abs(f) < 1e-5
