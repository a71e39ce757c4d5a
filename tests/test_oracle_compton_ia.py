"""Pin the oracle's impulse-approximation Compton sampler (L=2,
comptonScatterNRC, reference transport.hpp:342-483) where no reference build
exists (its shell table comes from the xraylib-backed Material).

The sampler is run on explicit shell tables (oracle.compton_ia). Each kept
sample gives E' and the scattering cosine; the electron's momentum
projection pz follows from them by energy-momentum conservation of the
impulse approximation (Ribberfors 1975):

    pz = (k k' (1 - cos) - (k - k')) / sqrt(k^2 + k'^2 - 2 k k' cos)

with k, k' the photon energies over the electron rest energy. The
reference draws pz by inverting the CDF of the analytic Compton profile in
v = 2 J0 m c^2 pz (transport.hpp:437-446),

    N(v) = 1/2 exp(1/2 - 1/2 (1 - v)^2)   (v <= 0),
    J(v) = dN/dv = 1/2 (1 + |v|) exp(1/2 - 1/2 (1 + |v|)^2),

at r4 exp(-b), b = ((1 + 2 J0 |pi|)^2 + 1)/2, so v is truncated at
N(v_t) = exp(-b), and keeps it with probability F(pz)/Fmax,
F = 1 + alpha clamp(pz, -p, p) (transport.hpp:424-458). Given the
cosine, the kept v therefore has density J(v) F(v) on v <= v_t. Each
sample's conditional CDF value is computed here by closed-form integration
of that density (not by the sampler's inverse), and must be uniform.

Also pinned: the shell choice in proportion to the electrons of the shells
bound below E with J0 > 0 (transport.hpp:348-360), the return of 0 with the
photon unchanged when the chosen shell is bound above E (Q9,
transport.hpp:367-371), the Klein-Nishina core in the limit of a narrow
profile, and the distribution of the scattering angle, which weights the
Klein-Nishina proposal with the reference's S(pi) (transport.hpp:391-422,
restated here in numpy) and the pz acceptance probability.
"""
import math

import numpy as np
import pytest
from scipy.special import erfc

import oracle

MEC2 = 510.9989461


def _recover(out, E):
    """cosine, k, k', pz of each sample (photons start along +z)."""
    mu = np.clip(out[:, 4], -1, 1)
    k = E / MEC2
    kp = out[:, 1] / MEC2
    pz = (k * kp * (1 - mu) - (k - kp)) / np.sqrt(k * k + kp * kp - 2 * k * kp * mu)
    return mu, k, kp, pz


def _profile_cdf(v):
    """N(v) of the analytic profile, v <= 0."""
    y = 1 - v
    return 0.5 * math.exp(0.5) * np.exp(-0.5 * y * y)


def _profile_first_moment(v):
    """M(v) = integral of x J(x) over (-inf, v], v <= 0 (closed form)."""
    y = 1 - v
    return 0.5 * math.exp(0.5) * ((1 - y) * np.exp(-0.5 * y * y) - math.sqrt(math.pi / 2) * erfc(y / math.sqrt(2)))


def _kinematics(mu, k, U, J0):
    """pi, alpha, exp(-b), truncation v_t and Fmax of the reference at each cosine."""
    pb = math.sqrt(2 * U + U * U)
    pi = (k * (k - U) * (1 - mu) - U) / np.sqrt(2 * k * (k - U) * (1 - mu) + U * U)
    e = 1 / (1 + k * (1 - mu))
    kc = k * e
    qc = np.sqrt(k * k + kc * kc - 2 * k * kc * mu)
    alpha = qc * (1 + kc * (kc - k * mu) / (qc * qc)) / k
    bp = 1 + 2 * J0 * np.abs(pi)
    eb = np.exp(-(bp * bp + 1) / 2)
    vt = 1 - np.sqrt(1 - 2 * np.log(2 * eb))
    fmax = np.where(pi <= -pb, 1 - alpha * pb, np.where(pi >= pb, 1 + alpha * pb, 1 + alpha * pi))
    return pb, pi, alpha, eb, vt, fmax


def _kept_mass(v, alpha, pb, c):
    """integral of J(x) F(x) over (-inf, v], v <= 0, F = 1 + alpha clamp(x/c, -pb, pb)."""
    xl = -pb * c
    low = (1 - alpha * pb) * _profile_cdf(np.minimum(v, xl))
    mid = (_profile_cdf(v) - _profile_cdf(xl)) + alpha / c * (_profile_first_moment(v) - _profile_first_moment(xl))
    return np.where(v <= xl, low, low + mid)


def _ks_uniform(g):
    g = np.sort(g)
    n = len(g)
    i = np.arange(1, n + 1)
    return max(np.max(i / n - g), np.max(g - (i - 1) / n)) * math.sqrt(n)


def test_profile_closed_forms():
    """N and M of the test against quadrature of J (guards the test itself)."""
    x = np.linspace(-40, -0.3, 400001)
    j = 0.5 * (1 + np.abs(x)) * np.exp(0.5 - 0.5 * (1 + np.abs(x)) ** 2)
    assert abs(np.trapezoid(j, x) - _profile_cdf(-0.3)) < 1e-9
    assert abs(np.trapezoid(x * j, x) - _profile_first_moment(-0.3)) < 1e-8
    assert abs(_profile_cdf(0.0) - 0.5) < 1e-15


SHELL = (4.0, 2.0, 0.02)  # binding keV, electrons, J0: broad profile, large F tilt
E0 = 60.0


@pytest.fixture(scope="module")
def single_shell():
    out = oracle.compton_ia([SHELL], E0, 300000, seed=91)
    mu, k, kp, pz = _recover(out, E0)
    return out, mu, k, kp, pz


def test_ia_energy_and_direction(single_shell):
    """returned energy = E - E', unit direction."""
    out = single_shell[0]
    assert np.all(out[:, 0] == E0 - out[:, 1])
    assert np.max(np.abs(np.linalg.norm(out[:, 2:5], axis=1) - 1)) < 1e-12


def test_ia_pz_profile(single_shell):
    """Given the cosine, v = 2 J0 mc^2 pz has density J(v) F(v) on v <= v_t:
    the conditional CDF values are uniform (Kolmogorov-Smirnov). Planted
    mistakes fail: the sign of pz, J0 off by 2 either way, F dropped (or an
    Fmax below F, which keeps every draw), and the truncation at N(v_t) =
    N(2 J0 pi) of the profile instead of exp(-b)."""
    out, mu, k, kp, pz = single_shell
    binding, _, J0 = SHELL
    U = binding / MEC2
    c = 2 * J0 * MEC2
    pb, pi, alpha, eb, vt, _ = _kinematics(mu, k, U, J0)
    v = c * pz
    assert np.all(v <= vt * (1 - 1e-9)), "every draw lies below the truncation"

    def ks(v_, alpha_, vt_):
        return _ks_uniform(_kept_mass(np.minimum(v_, vt_), alpha_, pb, c) / _kept_mass(vt_, alpha_, pb, c))

    d = ks(v, alpha, vt)
    assert d < 1.95, d  # p ~ 0.001
    assert ks(-v, alpha, vt) > 20  # sign of pz
    assert ks(v / 2, alpha, vt) > 20  # J0 doubled in the sampler
    assert ks(np.maximum(2 * v, 1.5 * vt), alpha, vt) > 20  # J0 halved
    assert ks(v, 0 * alpha, vt) > 5  # F tilt dropped / Fmax too small
    vt_wrong = 1 - np.sqrt(1 - 2 * np.log(2 * _profile_cdf(np.minimum(-c * np.abs(pi), 0)) + 1e-300))
    assert ks(v, alpha, np.minimum(vt_wrong, 0)) > 5


def test_ia_angle_distribution(single_shell):
    """The cosine follows the Klein-Nishina proposal (e uniform on
    [1/(1+2k), 1], kept with g) times S(pi) times the pz acceptance
    probability  int J F / (Fmax exp(-b)) (transport.hpp:375-458). Without S,
    or without the acceptance factor, it fails."""
    out, mu, k, kp, pz = single_shell
    binding, _, J0 = SHELL
    U = binding / MEC2
    c = 2 * J0 * MEC2
    x = np.linspace(-1, 1 - 1e-9, 20001)  # alpha is 0/0 at cos = 1
    pb, pi, alpha, eb, vt, fmax = _kinematics(x, k, U, J0)
    e = 1 / (1 + k * (1 - x))
    emin = 1 / (1 + 2 * k)
    g = (1 / e + e - (1 - x) * (1 + x)) / (1 / emin + emin)
    # S(pi) of transport.hpp:391-422
    sp2 = math.sqrt(2 * math.pi)
    tp = 1 / (1 + 0.332673 * (1 + 2 * J0 * pb))
    tq = 1 / (1 + 0.332673 * (1 + 2 * J0 * np.abs(pi)))
    poly = lambda t: t * (0.34802 - 0.0958798 * t + 0.7478556 * t * t)  # noqa: E731
    part = alpha * sp2 / (4 * J0) * ((math.sqrt(math.e) - eb * poly(tp)) - (math.sqrt(math.e) - eb * poly(tq)))
    S = np.where(pi <= -pb, (1 - alpha * pb) * eb / 2,
                 np.where(pi <= 0, (1 - alpha * pi) * eb / 2 - part,
                          np.where(pi < pb, 1 - (1 - alpha * pi) * eb / 2 - part, 1 - (1 - alpha * pb) * eb / 2)))
    accept = _kept_mass(vt, alpha, pb, c) / (fmax * eb)
    jac = k * e * e  # de/dcos
    edges = np.linspace(-1, 1, 41)

    def chi2(pdf):
        cdf = np.concatenate([[0], np.cumsum((pdf[1:] + pdf[:-1]) / 2 * np.diff(x))])
        cdf /= cdf[-1]
        expected = np.diff(np.interp(edges, x, cdf)) * len(mu)
        observed, _ = np.histogram(mu, edges)
        return np.sum((observed - expected) ** 2 / expected)

    assert chi2(jac * g * S * accept) < 80
    assert chi2(jac * g * accept) > 300
    assert chi2(jac * g * S) > 300


def test_ia_shell_choice():
    """Shells bound below E with J0 > 0 are chosen in proportion to their
    electrons; one bound above E and one with J0 = 0 never are. The shells'
    profiles differ 100-fold in width, so each sample's |pz| names its shell."""
    shells = [(4.0, 2.0, 0.05), (0.3, 6.0, 5.0), (80.0, 10.0, 0.05), (0.5, 20.0, 0.0)]
    n = 100000
    out = oracle.compton_ia(shells, E0, n, seed=92)
    assert np.all(np.isfinite(out))
    _, _, _, pz = _recover(out, E0)
    broad = pz < -0.003
    narrow = pz > -0.0015
    assert np.all(broad | narrow)
    sigma = math.sqrt(n * 0.25 * 0.75)
    assert abs(broad.sum() - 0.25 * n) < 5 * sigma, broad.sum()
    # if the shell bound above E counted, the broad share would be 12/18
    assert abs(broad.sum() - 12 / 18 * n) > 50 * sigma


def test_ia_binding_above_energy():
    """Q9: no shell is bound below E, the first shell is chosen, its binding
    exceeds E: the sampler returns 0 and leaves the photon unchanged."""
    out = oracle.compton_ia([(80.0, 2.0, 0.05), (70.0, 6.0, 0.5)], E0, 1000, seed=93)
    assert np.all(out[:, 0] == 0)
    assert np.all(out[:, 1] == E0)
    assert np.all(out[:, 2:5] == np.array([0.0, 0.0, 1.0]))


def test_ia_klein_nishina_limit():
    """A narrow profile (J0 = 30) on a nearly free electron: beyond the
    forward cone (1 - cos > 0.2) S = 1 and F/Fmax = 1 within 1e-4 and pz is
    below 1e-3, so E' equals the Compton line E/(1 + k(1 - cos)) within
    1e-3 and eps = E'/E follows Klein-Nishina there (chi-square); a uniform
    eps, or Klein-Nishina without the sin^2 term, fails."""
    E = 100.0
    n = 300000
    out = oracle.compton_ia([(1e-6, 1.0, 30.0)], E, n, seed=94)
    mu, k, kp, _ = _recover(out, E)
    line = E / (1 + k * (1 - mu))
    assert np.max(np.abs(out[:, 1] / line - 1)) < 1e-3
    eps = 1 / (1 + k * (1 - mu))
    emin = 1 / (1 + 2 * k)
    ecut = 1 / (1 + k * 0.2)
    eps = eps[eps < ecut]
    x = np.linspace(emin, ecut, 20001)
    t = (1 - x) / (k * x)
    kn = 1 / x + x - t * (2 - t)
    edges = np.linspace(emin, ecut, 41)

    def chi2(pdf):
        cdf = np.concatenate([[0], np.cumsum((pdf[1:] + pdf[:-1]) / 2 * np.diff(x))])
        cdf /= cdf[-1]
        expected = np.diff(np.interp(edges, x, cdf)) * len(eps)
        observed, _ = np.histogram(eps, edges)
        return np.sum((observed - expected) ** 2 / expected)

    assert chi2(kn) < 80
    assert chi2(np.ones_like(x)) > 1000
    assert chi2(1 / x + x) > 1000
