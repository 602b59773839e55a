"""Pins the FP64 oracle (oracle/btpm_oracle.c) before it is trusted as the
parity reference.  Each check is chosen so a plausible mistake (dropped term,
wrong sign or index, transposed operand) fails it:

  * Newton SBC root        -> bisection on the monotone residual
  * FTCS + zero-flux bottom -> exact discrete eigenmode decay
  * steady state           -> uniform T* = ((1-A)F/es)^(1/4) is a fixed point
  * periodic energy balance -> sum_s[(1-A)F + Q - es T0^4] = 0 at convergence
  * TI -> 0                -> instantaneous equilibrium
  * sparse self-heating    -> dense numpy product (equivalent source)
  * crater contraction     -> dense numpy product (equivalent source)
  * crater init            -> closed form under uniform forcing
  * facet / crater scattering fixed points -> direct linear solves
  * RGI / RGH              -> dense einsum
  * freeze / rotation loop -> replay with the rotation primitive
  * reference data files   -> crater view factors vs the spherical-cavity
                              identity; mesh statistics of the shipped shape
"""
import json
import math
import os

import numpy as np
import pytest

from binarytpm_amd import formats, inputs
from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")
ES = 0.95 * 5.67e-8
P = 4.3 * 3600.0


def bisect_root(src, m1, t1, es):
    g = lambda x: src + m1 * (t1 - x) - es * x ** 4  # noqa: E731
    lo, hi = 0.0, 2000.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if g(mid) > 0:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def test_newton_matches_bisection():
    rng = np.random.default_rng(3)
    for _ in range(200):
        src = rng.uniform(0, 1400)
        m1 = rng.choice([0.1, 1.0, 30.0, 400.0])
        t1 = rng.uniform(50, 400)
        t0 = rng.uniform(50, 400)
        x, it = O.newton_sbc(src, m1, t1, ES, t0, 1e-4)
        assert it < 50
        assert abs(x - bisect_root(src, m1, t1, ES)) < 1e-4


def _uniform_case(nf=3, F0=800.0, A=0.1):
    return np.full((nf, 650), F0), A, ((1 - A) * F0 / ES) ** 0.25


@pytest.mark.parametrize("modes", [(0,), (1,), (0, 1)])
def test_ftcs_discrete_eigenmode(modes):
    """Dirichlet-like surface (TI -> 0 pins T0 = T*) and zero-flux bottom:
    T_i - T* = a cos(theta (i - NX + 1/2)) with theta = (2k+1) pi / (2 NX - 1)
    decays by exactly (1 - 4 r sin^2(theta/2)) per step."""
    nx, nt = 56, 650
    F, A, Ts = _uniform_case(nf=1)
    p = O.params(nt, P, tacc=0.01, newton_tol=1e-12)
    r = (2 * math.pi / nt) / (8.0 / nx) ** 2
    i = np.arange(nx + 1)
    temp = np.full((1, nx + 1), Ts)
    expect = np.zeros(nx + 1)
    for k in modes:
        th = (2 * k + 1) * math.pi / (2 * nx - 1)
        shape = np.cos(th * (i - nx + 0.5))
        temp[0] += 2.0 * shape
        expect += 2.0 * shape * (1 - 4 * r * math.sin(th / 2) ** 2) ** nt
    stemp = np.full((1, nt), Ts)
    t2, s2, _ = O.smooth_rotation(p, F, [A], [1e-9], [1.0], temp, stemp)
    np.testing.assert_allclose(t2[0] - Ts, expect, atol=1e-9)


def test_uniform_state_is_fixed_point():
    F, A, Ts = _uniform_case()
    p = O.params(650, P)
    temp = np.full((3, 57), Ts)
    stemp = np.full((3, 650), Ts)
    t2, s2, dmax = O.smooth_rotation(p, F, [A] * 3, [10.0, 200.0, 2500.0], [1.0] * 3, temp, stemp)
    assert np.abs(t2 - Ts).max() < 1e-9 and np.abs(s2 - Ts).max() < 1e-9
    assert dmax.max() < 1e-9


@pytest.fixture(scope="module")
def small_body():
    case = inputs.config0(subdiv=1)          # 80 facets
    return case


@pytest.mark.parametrize("ti", [10.0, 200.0, 2500.0])
def test_periodic_energy_balance(small_body, ti):
    """Zero-flux bottom: the interior heat content is periodic, so the
    surface balance sums to zero over a converged rotation."""
    case = small_body
    p = O.params(650, case.rot_period, tacc=0.001)
    r = O.smooth_solve(p, case.illum, 0.1, ti)
    assert r.converged
    absorbed = (0.9 * case.illum.astype(np.float64)).sum(1)
    emitted = (ES * r.stemp ** 4).sum(1)
    rel = np.abs(absorbed - emitted) / absorbed
    assert rel.max() < 5e-3


def test_small_inertia_tends_to_equilibrium(small_body):
    case = small_body
    p = O.params(650, case.rot_period, tacc=0.001)
    inst = O.smooth_solve(p, case.illum, 0.1, 0.0)
    cond = O.smooth_solve(p, case.illum, 0.1, 1e-3)
    day = case.illum > 20.0
    assert np.abs(inst.stemp - cond.stemp)[day].max() < 0.05
    # instantaneous equilibrium itself: es T^4 = (1-A) F exactly
    np.testing.assert_allclose(ES * inst.stemp ** 4, 0.9 * case.illum.astype(np.float64),
                               rtol=1e-9, atol=1e-6)


def _dense(csr, n):
    rp, ci, vv = csr
    M = np.zeros((n, n))
    for f in range(n):
        M[f, ci[rp[f]:rp[f + 1]]] = vv[rp[f]:rp[f + 1]]
    return M


def test_self_heating_equals_dense_source(small_body):
    """A rotation with sparse view factors equals a rotation without them
    whose illumination carries Q/(1-A), Q = VF_dense @ (es T_prev^4)."""
    case = small_body
    nf = case.illum.shape[0]
    csr, _ = inputs.sparse_viewfactors(nf, 0.2, 0.2, seed=5)
    p = O.params(650, case.rot_period)
    rng = np.random.default_rng(1)
    temp, stemp = O.smooth_init(p, case.illum, np.full(nf, 0.1))
    temp += rng.uniform(-5, 5, temp.shape)
    stemp += rng.uniform(-20, 20, stemp.shape)
    A = np.full(nf, 0.1)
    ti = np.full(nf, 200.0)
    nr = np.ones(nf)
    t_a, s_a, _ = O.smooth_rotation(p, case.illum, A, ti, nr, temp, stemp, csr=csr)
    Q = _dense(csr, nf) @ (ES * stemp ** 4)
    t_b, s_b, _ = O.smooth_rotation(p, case.illum + Q / 0.9, A, ti, nr, temp, stemp)
    assert np.abs(s_a - s_b).max() < 1e-8 and np.abs(t_a - t_b).max() < 1e-8


def test_crater_contraction_equals_dense_source():
    rng = np.random.default_rng(2)
    nfac, nt, ns = 2, 650, 100
    cg = inputs.crater_cap(45.0)
    vf = cg.vf.astype(np.float64)
    illum = rng.uniform(0, 900, (nfac, nt, ns))
    rgh = rng.uniform(0, 50, (nfac, nt, ns))
    p = O.params(nt, P)
    temp, stemp = O.crater_init(p, vf, [0.1] * nfac, illum, rgh)
    stemp += rng.uniform(-10, 10, stemp.shape)
    import ctypes as C
    def rot(vfm, rg):
        t, s = temp.copy(), stemp.copy()
        dmax = np.zeros(nfac)
        O.lib().orc_crater_rotation(C.byref(p), nfac, ns, O._ptr(O._d(vfm), O._dp),
                                    O._ptr(O._d(np.full(nfac, 0.1)), O._dp),
                                    O._ptr(O._d(np.full(nfac, 300.0)), O._dp),
                                    O._ptr(O._d(np.ones(nfac)), O._dp), O._ptr(O._d(illum), O._dp),
                                    O._ptr(O._d(rg), O._dp), None, O._ptr(t, O._dp),
                                    O._ptr(s, O._dp), O._ptr(dmax, O._dp), None)
        return t, s
    t_a, s_a = rot(vf, rgh)
    V = vf.reshape(ns, ns)                       # V[i][j] = VF(j->i)
    Q = np.einsum("ij,fsi->fsj", V, ES * stemp ** 4)
    t_b, s_b = rot(np.zeros_like(vf), rgh + Q)
    assert np.abs(s_a - s_b).max() < 1e-8


def test_crater_init_closed_form():
    """Uniform absorbed flux S on every sub-facet: pass 1 gives
    T1 = (S/es)^(1/4); pass 2 adds sum_i VF(j->i) es T1^4 = S rowsum_j."""
    cg = inputs.crater_cap(45.0)
    V = cg.vf.astype(np.float64).reshape(100, 100)
    S = 500.0
    illum = np.full((1, 650, 100), S / 0.9)
    rgh = np.zeros_like(illum)
    p = O.params(650, P)
    temp, stemp = O.crater_init(p, cg.vf, [0.1], illum, rgh)
    rowsum = V.sum(axis=0)                       # over emitters i, per receiver j
    expect = (S * (1 + rowsum) / ES) ** 0.25
    np.testing.assert_allclose(stemp[0, 0], expect, rtol=1e-12)
    np.testing.assert_allclose(temp[0, :, :], np.broadcast_to(expect, (57, 100)), rtol=1e-12)


def test_facet_scattering_fixed_point():
    rng = np.random.default_rng(4)
    nf, nt = 60, 650
    csr, _ = inputs.sparse_viewfactors(nf, 0.2, 0.2, seed=9)
    D = rng.uniform(0, 1000, (nf, nt))
    A = rng.uniform(0.05, 0.3, nf)
    M = _dense(csr, nf)
    F1 = O.facet_scatter(D, A, csr, multiple=False)
    np.testing.assert_allclose(F1, D + M @ (A[:, None] * D), rtol=1e-12)
    Fm = O.facet_scatter(D, A, csr, multiple=True, acc=1e-9)
    p = np.linalg.solve(np.eye(nf) - A[:, None] * M, A[:, None] * D)
    np.testing.assert_allclose(Fm, p / A[:, None], atol=1e-6)


def test_crater_scattering_fixed_point():
    rng = np.random.default_rng(6)
    cg = inputs.crater_cap(45.0)
    V = cg.vf.astype(np.float64).reshape(100, 100)
    nfac, nt = 2, 650
    sh = rng.uniform(0, 1, (nfac, nt, 100))
    cs = rng.uniform(0, 1, (nfac, nt, 100))
    rgi = rng.uniform(0, 30, (nfac, nt, 100))
    A = np.array([0.1, 0.3])
    solar = 1129.75
    il = solar * sh * cs + rgi
    single = O.crater_scatter(cg.vf, A, sh, cs, rgi, solar, multiple=False)
    expect1 = il + A[:, None, None] * np.einsum("ij,fsi->fsj", V, il)
    np.testing.assert_allclose(single, np.maximum(expect1, 0), rtol=1e-12)
    multi = O.crater_scatter(cg.vf, A, sh, cs, rgi, solar, multiple=True, acc=1e-10)
    for f in range(nfac):
        Mf = np.eye(100) - A[f] * V.T           # p_j = A (sp_j + sum_i V[i][j] p_i)
        pf = np.linalg.solve(Mf, (A[f] * il[f]).T).T
        np.testing.assert_allclose(multi[f], pf / A[f], atol=1e-6)


def test_rgi_rgh_dense():
    rng = np.random.default_rng(7)
    nf, nt = 20, 650
    csr, dirs = inputs.sparse_viewfactors(nf, 0.3, 0.2, seed=3)
    drf = rng.uniform(0, 0.01, (len(csr[1]), 100))
    F = rng.uniform(0, 1000, (nf, nt))
    A = rng.uniform(0.05, 0.2, nf)
    T = rng.uniform(100, 400, (nf, nt))
    rp, ci, _ = csr
    exp_i = np.zeros((nf, nt, 100))
    exp_h = np.zeros((nf, nt, 100))
    for f in range(nf):
        for k in range(rp[f], rp[f + 1]):
            exp_i[f] += np.outer(F[ci[k]] * A[ci[k]], drf[k])
            exp_h[f] += np.outer(ES * T[ci[k]] ** 4, drf[k])
    np.testing.assert_allclose(O.rgi(csr, drf, F, A, nf), exp_i, rtol=1e-10)
    np.testing.assert_allclose(O.rgh(csr, drf, T, ES, nf), exp_h, rtol=1e-10)


def test_freeze_replay_matches_solve(small_body):
    """The solve loop equals a replay: rotate, and from rev > 5 freeze each
    facet whose rotation-to-rotation change is below tacc."""
    case = small_body
    nf = case.illum.shape[0]
    p = O.params(650, case.rot_period, tacc=0.01)
    A, ti, nr = np.full(nf, 0.1), np.full(nf, 300.0), np.ones(nf)
    ref = O.smooth_solve(p, case.illum, A, ti)
    temp, stemp = O.smooth_init(p, case.illum, A)
    frozen = np.zeros(nf, np.uint8)
    rev = 0
    while rev < p.max_rev:
        temp, stemp, dmax = O.smooth_rotation(p, case.illum, A, ti, nr, temp, stemp, frozen=frozen)
        if rev > p.check_after:
            frozen |= (dmax < p.tacc).astype(np.uint8)
            if frozen.all():
                rev += 1
                break
        rev += 1
    assert rev == ref.rotations >= p.check_after + 2
    np.testing.assert_array_equal(stemp, ref.stemp)
    assert ref.frozen_rev.min() >= p.check_after + 1
    assert ref.frozen_rev.max() == ref.rotations - 1


def test_rotation_cap():
    case = inputs.config0(subdiv=0)
    p = O.params(650, case.rot_period, tacc=1e-7, max_rev=9)
    r = O.smooth_solve(p, case.illum, 0.1, 2500.0)
    assert r.rotations == 9 and not r.converged


# ---- reference data files ----------------------------------------------------

def test_reference_crater_viewfactors_follow_cavity_identity():
    """The shipped hemispherical crater (new_crater2) has VF(i->j) close to
    A_j / (4 pi R^2) with R = 10 and row sums near 1/2, and our view-factor
    reader/transposition maps its rows to receivers."""
    z = np.load(os.path.join(GOLD, "new_crater2.npz"))
    V, F = z["V"], z["F"]
    csr = (z["rowptr"], z["colidx"], z["vf"])
    M = formats.crater_vf_matrix(csr).reshape(100, 100)      # M[i][j] = VF(j->i)
    area = 0.5 * np.linalg.norm(np.cross(V[F[:, 1]] - V[F[:, 0]], V[F[:, 2]] - V[F[:, 0]]), axis=1)
    R = np.linalg.norm(V, axis=1).max()
    assert abs(R - 10.0) < 1e-3
    rows = M.sum(axis=0)                                      # per receiver j
    assert 0.47 < rows.min() and rows.max() < 0.51
    ident = np.repeat((area / (4 * np.pi * R * R))[:, None], 100, axis=1)
    np.fill_diagonal(ident, 0)
    rel = np.abs(M - ident)[ident > 0] / ident[ident > 0]
    assert np.median(rel) < 0.1


def test_cap_generator_matches_shipped_topology():
    """Our ring-built cap has the shipped crater's counts (61 vertices,
    100 facets) and, at 90 degrees, its row sums (1 - cos g)/2 = 1/2."""
    z = np.load(os.path.join(GOLD, "new_crater2.npz"))
    cg = inputs.crater_cap(90.0)
    assert cg.V.shape == z["V"].shape and cg.F.shape == z["F"].shape
    rows = cg.vf.reshape(100, 100).sum(axis=0)
    full = cg.area.sum() / (4 * np.pi)
    assert abs(full - 0.5) < 0.02
    assert np.all(rows < full)
    cg45 = inputs.crater_cap(45.0)
    assert abs(cg45.area.sum() / (4 * np.pi) - (1 - math.cos(math.pi / 4)) / 2) < 0.01


def test_reference_mesh_statistics():
    """Shape reader + area/volume code reproduce doppler_binary_properties.txt."""
    z = np.load(os.path.join(GOLD, "doppler_binary.npz"))
    props = json.load(open(os.path.join(GOLD, "doppler_binary_properties.json")))
    area, vol, deq = formats.shape_properties(z["V"], z["F"])
    assert abs(area - props["Total Area"][0]) < 1e-3 * props["Total Area"][0]
    assert abs(vol - props["Total Volume"][0]) < 1e-3 * props["Total Volume"][0]
    assert abs(deq - props["Effective Diameter"][0]) < 1e-3 * props["Effective Diameter"][0]


def test_refcpu_baseline_agrees_with_oracle(small_body):
    """The float OpenMP restatement (timed CPU baseline) stays within the
    parity tolerance of the FP64 oracle."""
    case = small_body
    p = O.params(650, case.rot_period, tacc=0.01)
    r = O.smooth_solve(p, case.illum, 0.1, 500.0)
    st, rev, _ = O.refcpu_smooth_solve(case.illum, 0.1, 500.0, 650, case.rot_period, tacc=0.01)
    assert np.abs(st - r.stemp).max() < 0.05
    assert abs(rev - r.rotations) <= 2


# ---- the TI sweep's carried state (warm start) -----------------------------

def _small_rough():
    case, crater = inputs.rough_primary(subdiv=0, nfac=None, density=0.2)   # 20 facets
    nf = case.illum.shape[0]
    F = O.facet_scatter(case.illum, 0.1, case.csr, multiple=True, acc=1e-4)
    rgi = O.rgi(case.csr, crater["drf"], F, np.full(nf, 0.1), nf)
    il = O.crater_scatter(crater["vf"], 0.1, crater["shadow"], crater["cosang"], rgi,
                          crater["solar"], multiple=True, acc=1e-4)
    return case, dict(vf=crater["vf"], illum=il, drf=crater["drf"]), F


def test_ti_sweep_first_solve_is_the_cold_solve():
    """The sweep's first TI starts from the initial state: it equals the
    single solve bit for bit, smooth facets and craters."""
    case, cr, F = _small_rough()
    nf = F.shape[0]
    p = O.params(650, case.rot_period, tacc=0.01)
    (sm, c), = O.ti_sweep(p, F, 0.1, [120.0], csr=case.csr, crater=cr)
    cold = O.smooth_solve(p, F, 0.1, 120.0, csr=case.csr)
    np.testing.assert_array_equal(sm.stemp, cold.stemp)
    rgh = O.rgh(case.csr, cr["drf"], cold.stemp, ES, nf)
    cc = O.crater_solve(p, cr["vf"], 0.1, 120.0, cr["illum"], rgh)
    np.testing.assert_array_equal(c.stemp, cc.stemp)


def test_ti_sweep_warm_states_reach_the_cold_periodic_state():
    """The periodic state does not depend on where the iteration starts:
    every warm-started TI of a chain matches its own cold solve.  The freeze
    test stops a column once it changes by less than tacc per rotation, which
    leaves it up to tacc / (1 - lambda) from the periodic state (lambda its
    slowest decay per rotation), on the side it came from: 0.05 K for the
    smooth facets here, 0.15 K for the slower craters.  And a carried state
    is really used: starting from the state of the same TI converges at the
    first allowed test (rev > 5, 7 rotations), moving less than tacc per
    rotation."""
    case, cr, F = _small_rough()
    nf = F.shape[0]
    p = O.params(650, case.rot_period, tacc=0.01)
    tis = [40.0, 60.0, 90.0, 130.0]
    chain = O.ti_sweep(p, F, 0.1, tis, csr=case.csr, crater=cr)
    for ti, (sm, c) in zip(tis, chain):
        cold = O.smooth_solve(p, F, 0.1, ti, csr=case.csr)
        assert np.abs(sm.stemp - cold.stemp).max() < 0.05
        rgh = O.rgh(case.csr, cr["drf"], cold.stemp, ES, nf)
        cc = O.crater_solve(p, cr["vf"], 0.1, ti, cr["illum"], rgh)
        assert np.abs(c.stemp - cc.stemp).max() < 0.15
        if ti != tis[0]:                                 # the carried state is nearer
            assert c.column_steps < cc.column_steps
    sm, c = chain[-1]
    again = O.ti_sweep(p, F, 0.1, [130.0, 130.0], csr=case.csr, crater=cr)
    r2s, r2c = again[1]
    assert r2s.rotations == p.check_after + 2 and r2c.rotations == p.check_after + 2
    # each of those rotations moves the state by less than tacc (the decay
    # that froze it continues)
    assert np.abs(r2s.stemp - again[0][0].stemp).max() < (p.check_after + 2) * p.tacc
    assert np.abs(r2c.stemp - again[0][1].stemp).max() < (p.check_after + 2) * p.tacc


def test_ti_sweep_restarts_after_instantaneous_equilibrium():
    """The 3905 order: TI 0 first (no depth profile; craters from zero
    temperatures), then the first TI > 0 from the initial state, so it
    equals its cold solve bit for bit."""
    case, cr, F = _small_rough()
    nf = F.shape[0]
    p = O.params(650, case.rot_period, tacc=0.01)
    (s0, c0), (s1, c1) = O.ti_sweep(p, F, 0.1, [0.0, 50.0], csr=case.csr, crater=cr)
    inst = O.smooth_solve(p, F, 0.1, 0.0, csr=case.csr)
    np.testing.assert_array_equal(s0.stemp, inst.stemp)
    rgh0 = O.rgh(case.csr, cr["drf"], inst.stemp, ES, nf)
    z = O.crater_solve(p, cr["vf"], 0.1, 0.0, cr["illum"], rgh0, temp=np.zeros((nf, 57, 100)),
                       stemp=np.zeros(cr["illum"].shape))
    np.testing.assert_array_equal(c0.stemp, z.stemp)
    cold = O.smooth_solve(p, F, 0.1, 50.0, csr=case.csr)
    np.testing.assert_array_equal(s1.stemp, cold.stemp)


def test_refcpu_warm_sweep_agrees_with_oracle(small_body):
    """refcpu's carried state (the timed CPU baseline of the warm sweep)
    against the oracle's chained sweep: temperatures within the parity
    tolerance, rotations within 2."""
    case = small_body
    p = O.params(650, case.rot_period, tacc=0.01)
    tis = [80.0, 110.0, 150.0]
    ref = O.ti_sweep(p, case.illum, 0.1, tis)
    state = {}
    for ti, (r, _) in zip(tis, ref):
        st, rev, _ = O.refcpu_smooth_solve(case.illum, 0.1, ti, 650, case.rot_period, tacc=0.01,
                                           state=state)
        assert np.abs(st - r.stemp).max() < 0.05
        assert abs(rev - r.rotations) <= 2


# --------------------------------------------------------------------------
# illumination geometry and shadows (geometry.c)
# --------------------------------------------------------------------------
POS = (1.16, math.radians(-1.95), math.radians(198.31))   # r AU, ecliptic lat, lon
POLE_REF = (math.radians(-83.0), math.radians(266.0))      # the reference's pole


def _body_frame(rot):
    """Equatorial -> ecliptic rotation built from its definition: the body's
    z axis along the pole (lat, lon); used to place vertices independently."""
    lat, lon = rot
    z = np.array([math.cos(lat) * math.cos(lon), math.cos(lat) * math.sin(lon), math.sin(lat)])
    return z


def test_geometry_convex_body_only_faces_away():
    """On a convex body nothing shadows anything: a facet is unlit exactly
    when the Sun is behind its plane, and illum = solar cos shadow with
    solar = 1367 / r^2."""
    V, F = inputs.ellipsoid(2, (0.6, 0.5, 0.45))
    for rot in [(math.pi / 2, 0.0), POLE_REF]:
        g = O.geometry(V, F, 96, POS, rot, rev_primary=2.0)
        assert np.array_equal(g.shadow, (g.cosang >= 0.0).astype(np.int32))
        assert 0.45 < g.shadow.mean() < 0.55
        assert g.solar == np.float32(1367.0 / POS[0] ** 2)
        ref = (np.float32(g.solar) * g.cosang.astype(np.float32)) * g.shadow.astype(np.float32)
        assert np.array_equal(g.illum, ref)


def test_geometry_subsolar_point_turns_with_the_body():
    """Closed form for a sphere with its pole along ecliptic north: at step s
    the body has turned by 2 pi s n / ntau, so the facet nearest the Sun sits
    at body longitude lon_sun - 2 pi s n / ntau, lon_sun = system longitude +
    pi; the facet of largest cos(angle) must be within a facet of it."""
    V, F = inputs.ellipsoid(3, (1.0, 1.0, 1.0))
    n, c, _, _ = inputs.facet_geometry(V, F)
    ntau, nrev = 48, 2.0
    g = O.geometry(V, F, ntau, (1.0, 0.0, math.radians(40.0)), (math.pi / 2, 0.0), rev_primary=nrev)
    for s in range(0, ntau, 5):
        want = math.radians(40.0) + math.pi - 2 * math.pi * s * nrev / ntau
        d = np.array([math.cos(want), math.sin(want), 0.0])
        best = int(np.argmax(g.cosang[:, s]))
        assert float(n[best] @ d) > math.cos(math.radians(8.0))
        assert g.cosang[best, s] > math.cos(math.radians(8.0))


def test_pole_frame_maps_z_to_the_pole():
    """The equatorial -> ecliptic matrix of the reference (thermal.cpp:1286-1296)
    maps the body's z axis onto the pole (lat, lon).  The six facets around
    the +z apex of a hexagonal bipyramid have a mean normal k z (k < 1), so
    their mean cos(angle) is k (z . sun) at every step as the body turns about
    the pole; it changes sign with z . sun."""
    # hexagonal bipyramid: apexes on the z axis, the top fan symmetric about it
    ang = np.arange(6) * (math.pi / 3)
    V = np.vstack([[0.0, 0.0, 1.0], np.stack([np.cos(ang), np.sin(ang), np.zeros(6)], 1),
                   [0.0, 0.0, -1.0]])
    F = np.array([[0, 1 + i, 1 + (i + 1) % 6] for i in range(6)] +
                 [[7, 1 + (i + 1) % 6, 1 + i] for i in range(6)], np.int32)
    ring = np.arange(6)
    n, _, _, _ = inputs.facet_geometry(V, F)
    assert np.allclose(n[ring].mean(axis=0)[:2], 0.0) and n[ring].mean(axis=0)[2] > 0.5
    k = float(np.linalg.norm(n[ring].mean(axis=0)))
    sun = -np.array([1.0, 0.0, 0.0])            # system on the +x axis
    for rot in [POLE_REF, (math.radians(35.0), math.radians(120.0)),
                (math.radians(-20.0), math.radians(10.0))]:
        z = _body_frame(rot)
        g = O.geometry(V, F, 12, (1.0, 0.0, 0.0), rot, rev_primary=1.0)
        m = g.cosang[ring].mean(axis=0)
        assert np.allclose(m, k * float(z @ sun), atol=1e-9)


def _brute_force_shadow(equ, fv, npv, ntau, pos, rot, revs):
    """Independent shadow flags: vertices placed with Rodrigues-free matrices
    built from the pole, Moller-Trumbore line/triangle tests in the body
    frame (Sun at -centre), the reference's rule that only facets nearer the
    Sun (ties to the lower index) can shadow."""
    lat, lon = rot
    z = _body_frame(rot)
    # the reference's x axis: column 0 of its matrix, from its closed form
    gam = math.acos(math.cos(lat) * math.sin(abs(lon)))
    x = np.array([math.sin(lat), 0.0, -math.cos(lat) * math.cos(lon)]) / math.sin(gam)
    y = np.cross(z, x)
    M = np.stack([x, y, z], axis=1)
    r, blat, blon = pos
    cen = O.AU_KM * r * np.array([math.cos(blat) * math.cos(blon), math.cos(blat) * math.sin(blon),
                                  math.sin(blat)])
    nf = len(fv)
    out = np.zeros((nf, ntau), np.int32)
    for s in range(ntau):
        P = np.empty_like(equ)
        for body, sl in ((0, slice(0, npv)), (1, slice(npv, None))):
            a = 2 * math.pi * s * revs[body] / ntau
            Rz = np.array([[math.cos(a), -math.sin(a), 0], [math.sin(a), math.cos(a), 0], [0, 0, 1]])
            P[sl] = equ[sl] @ Rz.T @ M.T
        v0, v1, v2 = P[fv[:, 0]], P[fv[:, 1]], P[fv[:, 2]]
        mid = (v0 + v1 + v2) / 3.0
        nrm = np.cross(v1 - v0, v2 - v0)
        dist = np.linalg.norm(mid + cen, axis=1)
        lit = (np.einsum("ij,ij->i", nrm, -(mid + cen)) >= 0.0)
        e1, e2 = v1 - v0, v2 - v0
        for G in np.where(lit)[0]:
            nearer = (dist < dist[G]) | ((dist == dist[G]) & (np.arange(nf) < G))
            D = mid[G] + cen                      # Sun -> centroid, body frame
            O_ = -cen
            pv = np.cross(D, e2[nearer])
            det = np.einsum("ij,ij->i", e1[nearer], pv)
            T = O_ - v0[nearer]
            with np.errstate(divide="ignore", invalid="ignore"):
                u = np.einsum("ij,ij->i", T, pv) / det
                q = np.cross(T, e1[nearer])
                v = (q @ D) / det
            if np.any((u >= 0) & (v >= 0) & (u + v <= 1)):
                lit[G] = False
        out[:, s] = lit
    return out


def test_geometry_shadows_match_brute_force():
    """A dented primary (concave troughs shadow their neighbours) and a close
    secondary in the Sun's plane (mutual eclipses): the flags equal an
    independent line/triangle test wherever the Sun is more than 1e-6 rad
    above the facet's plane (at grazing incidence the Sun->centroid line runs
    along the neighbouring facets and the two constructions may round either
    way; the illumination at stake there is below 2e-3 W/m^2), and the pairs
    test shadows more than the back faces alone."""
    equ, fv, npv = inputs.binary_shape(2, 1, sep=1.2, dent=0.35)
    ntau, revs = 40, (3.0, 1.0)
    g = O.geometry(equ, fv, ntau, (1.0, 0.0, 0.0), (math.pi / 2, 0.0), nv_primary=npv,
                   rev_primary=revs[0], rev_secondary=revs[1])
    bf = _brute_force_shadow(equ, fv, npv, ntau, (1.0, 0.0, 0.0), (math.pi / 2, 0.0), revs)
    backface_only = int((g.cosang >= 0.0).sum())
    print(f"shadowed beyond back faces: {backface_only - int(g.shadow.sum())} facet-steps")
    assert backface_only - int(g.shadow.sum()) > 50
    clear = np.abs(g.cosang) > 1e-6
    print(f"grazing facet-steps excluded: {int((~clear).sum())}, disagreeing there: "
          f"{int(((g.shadow != bf) & ~clear).sum())}")
    assert np.array_equal(g.shadow[clear], bf[clear])
