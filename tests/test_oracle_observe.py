"""Pins of the Observe flux restatement (oracle/observe.c) on the CPU.

The reference ships the fluxes this stage writes (175706/Run-04/*_flux(*).txt)
but not the temperature and visibility files they were made from, so the
restatement is pinned on what the mathematics fixes: the flux of a uniform
blackbody polyhedron (projected area x emissivity x Planck / d^2, computed
independently), Planck's Wien and Rayleigh-Jeans limits, the colour
temperature fit recovering a blackbody's temperature (and the table index it
selects), the smooth fit's reference quirk (band 1 alone), and a flat crater
giving the smooth flux exactly.
"""
import math
import os

import numpy as np
import pytest

from binarytpm_amd import inputs
from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
H, CV, K = 6.626068963e-34, 299792458.0, 1.380650424e-23    # observe_generator.cpp:10-12
EPS = 0.95
# the system 150 m from the observer (shape units are metres here): no
# cancellation against astronomical offsets, so FP64 checks can be tight
HEL = (1e-9, 0.1, 0.2)
OBS = (1e-9, -0.3, 1.1)
ROT = (0.4, 1.3)


def _sphere(subdiv=3):
    V, F = inputs.icosphere(subdiv)
    return V, F.astype(np.int32)


def _observer_dir_and_pos():
    def cart(p):
        return O.AU_M * p[0] * np.array([math.cos(p[1]) * math.cos(p[2]),
                                         math.cos(p[1]) * math.sin(p[2]), math.sin(p[1])])
    acen = cart(HEL)
    ocen = acen - cart(OBS)
    return acen, ocen


def _run(V, F, T, ntau=4, wl=O.WISE_UM, colour=None, crater=None, vis=None, steps=None):
    nf = len(F)
    st = np.broadcast_to(np.asarray(T, np.float32).reshape(-1, 1), (nf, ntau)).copy() \
        if np.ndim(T) else np.full((nf, ntau), T, np.float32)
    if vis is None:
        vis = np.ones((nf, ntau), np.int32)
    if colour is None:
        colour = np.ones((3, 467))
    return O.observe(V, F, ntau, HEL, OBS, ROT, vis, st, colour, crater=crater, eps=EPS,
                     wavelengths=wl, steps=steps)


def _planck(lam_um, T):
    lam = lam_um * 1e-6
    return (2.0 * H * CV * CV) / lam ** 5 / (math.exp(H * CV / (lam * K * T)) - 1.0)


def test_uniform_polyhedron_flux():
    """A uniform blackbody polyhedron: flux = eps B(T) sum_f max(0, A_f n_f.u_f) / d_f^2
    with u_f and d_f the direction and distance from each facet to the observer
    (independent numpy geometry; the body turns, so step 0 only)."""
    V, F = _sphere(3)
    sm, _ = _run(V, F, 300.0, steps=(0, 1))
    acen, ocen = _observer_dir_and_pos()
    # the body frame at step 0: the pole matrix of the restatement applied to V
    r0, r1 = ROT
    g = math.acos(math.cos(r0) * math.sin(abs(r1)))
    tm = np.array([
        [math.sin(r0) / math.sin(g), -math.cos(r0) ** 2 * math.sin(r1) * math.cos(r1) / math.sin(g),
         math.cos(r0) * math.cos(r1)],
        [0.0, (math.sin(r0) ** 2 + math.cos(r0) ** 2 * math.cos(r1) ** 2) / math.sin(g),
         math.cos(r0) * math.sin(r1)],
        [-math.cos(r0) * math.cos(r1) / math.sin(g), -math.sin(r0) * math.cos(r0) * math.sin(r1) /
         math.sin(g), math.sin(r0)]])
    P = V @ tm.T + acen
    p0, p1, p2 = P[F[:, 0]], P[F[:, 1]], P[F[:, 2]]
    n = np.cross(p1 - p0, p2 - p0)
    A = np.linalg.norm(n, axis=1) / 2.0
    n /= 2.0 * A[:, None]
    mid = (p0 + p1 + p2) / 3.0
    to_obs = ocen - mid
    d = np.linalg.norm(to_obs, axis=1)
    proj = np.clip(np.einsum("ij,ij->i", n, to_obs / d[:, None]), 0.0, None)
    for i, lam in enumerate(O.WISE_UM):
        want = EPS * _planck(lam, 300.0) * (A * proj / d ** 2).sum()
        assert abs(sm[0, i] - want) <= 1e-10 * want, (i, sm[0, i], want)


def test_planck_limits():
    """Wien's law where hc/(lambda k T) >> 1 and Rayleigh-Jeans where << 1,
    through the whole synthesis (a sphere at uniform T: the flux ratio of two
    temperatures at one wavelength is the ratio of their intensities)."""
    V, F = _sphere(2)
    # Wien: 2 um at 150 / 200 K, x = 48 / 36: B ratio = exp(-x1 + x2) to 1e-15
    sm1, _ = _run(V, F, 150.0, wl=(2.0, 2.0, 2.0), steps=(0, 1))
    sm2, _ = _run(V, F, 200.0, wl=(2.0, 2.0, 2.0), steps=(0, 1))
    x = lambda T: H * CV / (2e-6 * K * T)  # noqa: E731
    assert abs(sm1[0, 0] / sm2[0, 0] / math.exp(x(200.0) - x(150.0)) - 1.0) < 1e-12
    # Rayleigh-Jeans: 10 mm at 500 / 1000 K, x = 2.9e-3 / 1.4e-3: ratio = T1/T2 to x/2
    sm1, _ = _run(V, F, 500.0, wl=(1e4, 1e4, 1e4), steps=(0, 1))
    sm2, _ = _run(V, F, 1000.0, wl=(1e4, 1e4, 1e4), steps=(0, 1))
    assert abs(sm1[0, 0] / sm2[0, 0] - 0.5) < 2e-3 and sm1[0, 0] / sm2[0, 0] != 0.5


@pytest.mark.parametrize("T0", [150, 273, 400])
def test_colour_fit_recovers_blackbody_temperature(T0):
    """A uniform blackbody: both colour-temperature fits land on T0, so each
    band is scaled by the WISE correction of that temperature."""
    V, F = _sphere(2)
    cc = np.load(os.path.join(GOLD, "fg3_observe.npz"))["colour"]
    raw_s, raw_r = _run(V, F, float(T0), steps=(0, 1))
    sm, _ = _run(V, F, float(T0), colour=cc, steps=(0, 1))
    np.testing.assert_allclose(sm[0], raw_s[0] * cc[:, T0 - 100], rtol=1e-14)


def test_smooth_fit_uses_band_one_only():
    """Two temperatures on one body: the rough fit minimises the sum over three
    bands, the smooth fit only its first term (the reference's
    `((b-m)/m, 2.0)` comma terms add 2.0 each); the two fits then pick
    different colour temperatures, and each output band carries the
    correction of its own fit."""
    V, F = _sphere(2)
    nf = len(F)
    T = np.where(np.arange(nf) % 2 == 0, 150.0, 380.0)
    cvx = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]], np.float32)
    cfv = np.array([[0, 1, 2], [1, 3, 2]], np.int32)      # flat crater: rough = smooth
    ntau = 4
    cv = np.ones((nf, ntau, 2), np.float32)
    ct = np.repeat(T.astype(np.float32)[:, None, None], ntau, 1).repeat(2, 2)
    crater = (cvx, cfv, cv, ct)
    raw_s, raw_r = _run(V, F, T, crater=crater, steps=(0, 1))
    np.testing.assert_allclose(raw_r, raw_s, rtol=1e-12)
    temps = np.arange(100, 567, dtype=np.float64)
    b = np.array([[_planck(lam, t) for lam in O.WISE_UM] for t in temps])
    b /= b.mean(axis=1, keepdims=True)
    m = raw_s[0] / raw_s[0].mean()
    t_smooth = int(temps[np.argmin(((b[:, 0] - m[0]) / m[0]) ** 2)])
    t_rough = int(temps[np.argmin((((b - m) / m) ** 2).sum(axis=1))])
    assert t_smooth != t_rough
    cc = np.load(os.path.join(GOLD, "fg3_observe.npz"))["colour"]
    sm, ro = _run(V, F, T, crater=crater, colour=cc, steps=(0, 1))
    np.testing.assert_allclose(sm[0], raw_s[0] * cc[:, t_smooth - 100], rtol=1e-13)
    np.testing.assert_allclose(ro[0], raw_r[0] * cc[:, t_rough - 100], rtol=1e-13)


def test_flat_crater_equals_smooth_and_visibility_masks():
    """A flat crater (every sub-facet facing +z) reproduces the smooth flux:
    sum_j a_j cos(view_j) / sum_j a_j n_jz = cos(view); halving every
    sub-facet's visible fraction halves the rough flux; facets flagged
    invisible drop out of both."""
    V, F = _sphere(2)
    nf, ntau = len(F), 6
    T = 200.0 + 100.0 * np.random.default_rng(0).random(nf)
    g = np.linspace(0, 1, 5)
    cvx = np.array([(x, y, 0.0) for y in g for x in g], np.float32)
    cfv = []
    for j in range(4):
        for i in range(4):
            a = j * 5 + i
            cfv += [(a, a + 1, a + 6), (a, a + 6, a + 5)]
    cfv = np.array(cfv, np.int32)
    ct = np.repeat(np.repeat(T.astype(np.float32)[:, None, None], ntau, 1), len(cfv), 2)
    full = (cvx, cfv, np.ones_like(ct), ct)
    half = (cvx, cfv, np.full_like(ct, 0.5), ct)
    s1, r1 = _run(V, F, T, ntau=ntau, crater=full)
    np.testing.assert_allclose(r1, s1, rtol=1e-12)
    _, r2 = _run(V, F, T, ntau=ntau, crater=half)
    np.testing.assert_allclose(r2, 0.5 * r1, rtol=1e-12)
    vis = np.ones((nf, ntau), np.int32)
    vis[::3] = 0
    s3, r3 = _run(V, F, T, ntau=ntau, crater=full, vis=vis)
    assert (s3 < s1).all() and (r3 < r1).all()
