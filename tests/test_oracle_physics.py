"""Pins the oracle's step physics against closed forms and textbook solvers:
exact strain rate of a linear velocity field under kernel gradient
correction, Shepard/no-slip reproduction of uniform states, the discrete SPH
Laplacian of x^2 on a lattice, BiCGSTAB against scipy's sparse LU, the
predictor/corrector identity, the MOST stable-branch formula and the max
statistics."""
import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import abi
import cases
import oracle


def _sorted_input(o, P):
    keys, perm, _, _ = o.nnps()
    return P[perm]


def test_strain_rate_of_linear_field_is_exact():
    # dyadic spacing and coefficients: u = A x is exact in float32 storage
    P, _ = cases.box(nf=10, walls=3, jitter=0.0, dx=2.0 ** -7)
    A = np.array([[0.25, -0.75, 0.125], [0.5, 0.125, -0.375], [-0.625, 0.875, -0.375]])
    pos = np.stack([P["x"], P["y"], P["z"]], 1).astype(np.float64)
    u = pos @ A.T
    P["ux"], P["uy"], P["uz"] = u[:, 0], u[:, 1], u[:, 2]
    prm = oracle.default_params()
    prm.update(turbulence_model=abi.SSM, fv_solve=abi.FV_CLEARY, ppe_solver=abi.PPE_REF_JACOBI)
    o = oracle.Oracle(prm, len(P))
    o.upload(P)
    o.step(0, 1e-5)
    S = _sorted_input(o, P)
    Ssym = 0.5 * (A + A.T)
    expect = math.sqrt(2.0 * (Ssym * Ssym).sum())
    SR = o.field("SR")
    pt = o.ifield("P_TYPE")
    off0, nb0 = o.neighbors(0)
    off1, nb1 = o.neighbors(1)
    sel = []
    for i in np.nonzero(pt == 1)[0]:
        a = nb0[off0[i]:off0[i + 1]]
        a = set(a[(pt[a] == 1) & (a != i)].tolist())
        b = nb1[off1[i]:off1[i + 1]]
        b = set(b[pt[b] == 1].tolist())
        if a == b:
            sel.append(i)
    sel = np.array(sel)
    assert len(sel) > 300
    np.testing.assert_allclose(SR[sel], expect, rtol=1e-11)
    # Smagorinsky viscosity from the same strain: Cs^2 l^2 |S| with l >= h
    vis = o.field("VIS_T")[sel]
    assert np.all(vis >= 0.15 ** 2 * (S["h"][sel].astype(np.float64)) ** 2 * expect * (1 - 1e-9))


def _lidded_box(U, walls_move=False):
    P, _ = cases.box(nf=10, walls=3, jitter=0.05)
    top = P["z"] > P["z"][P["p_type"] == 1].max() + 1e-6
    P["p_type"][top] = -1
    move = (P["p_type"] != 0) | walls_move
    P["ux"] = np.where(move, U[0], 0.0)
    P["uy"] = np.where(move, U[1], 0.0)
    P["uz"] = np.where(move, U[2], 0.0)
    return P


def test_uniform_flow_is_reproduced():
    U = (0.7, -0.2, 0.1)
    P = _lidded_box(U, walls_move=True)
    prm = oracle.default_params()
    prm.update(ppe_solver=abi.PPE_REF_JACOBI, noslip_bc=0)
    o = oracle.Oracle(prm, len(P))
    o.upload(P)
    o.step(0, 1e-4)
    pt = o.ifield("P_TYPE")
    fl = pt == 1
    for c, name in enumerate(("UX", "UY", "UZ")):
        v = o.field(name)
        assert np.all(v[fl] == np.float32(U[c]))            # no force on a uniform state
        np.testing.assert_allclose(v[pt == -1], np.float32(U[c]), rtol=1e-13)  # free-slip Shepard
        # free-slip velocities are Shepard averages (rounding ~1e-16)
        assert np.abs(o.field("FTOTAL" + name[1])[fl]).max() < 1e-12
    assert np.abs(o.field("PRES")).max() < 1e-9


def test_noslip_wall_velocity():
    U = (0.7, -0.2, 0.1)
    P = _lidded_box(U)
    prm = oracle.default_params()
    prm.update(ppe_solver=abi.PPE_REF_JACOBI)
    o = oracle.Oracle(prm, len(P))
    o.upload(P)
    o.step(0, 1e-4)
    pt = o.ifield("P_TYPE")
    # functions_BC.cuh:184: u_w = -sum_{fluid} u_j W / (1e-10 + sum W); in a
    # uniform flow that is -U wherever the wall sees fluid (free-slip
    # particles have p_type -1 and are not averaged)
    w = pt == 0
    ux0 = o.field("UX0")[w]
    ratio = ux0 / -np.float32(U[0])
    assert np.all((ratio >= -1e-12) & (ratio <= 1 + 1e-12))
    pos = np.stack([o.field("X"), o.field("Y"), o.field("Z")], 1)
    fl = pos[pt == 1]
    dmin = np.array([np.sqrt(((fl - q) ** 2).sum(1)).min() for q in pos[w]])
    near = dmin < 0.015
    assert near.sum() > 300
    np.testing.assert_allclose(ux0[near], -np.float32(U[0]), rtol=1e-12)


def test_uniform_pressure_is_preserved():
    P = _lidded_box((0.0, 0.0, 0.0))
    P["pres"] = 123.5
    prm = oracle.default_params()
    prm.update(ppe_solver=abi.PPE_REF_JACOBI)
    o = oracle.Oracle(prm, len(P))
    o.upload(P)
    o.step(0, 1e-4)
    p = o.field("PRES")
    np.testing.assert_allclose(p, 123.5, rtol=1e-13)


def test_laplacian_of_x_squared_on_lattice():
    # dyadic lattice: positions exact in float32, so neighbourhoods are
    # exactly symmetric and r = 3 dx lies exactly on the support radius
    dx = 2.0 ** -7
    P, _ = cases.box(nf=14, walls=3, jitter=0.0, dx=dx)
    rng = np.random.default_rng(5)
    P["ux"] = rng.standard_normal(len(P)) * 1e-3
    prm = cases.box_params(oracle.default_params(), tol=1e-10)
    o = oracle.Oracle(prm, len(P))
    o.upload(P)
    o.step(0, 1e-4)
    sysm = o.ppe_system()
    x = o.field("X")
    rows = sysm["rows"]
    q = x[rows] ** 2
    rp, cols, vals, diag = sysm["rowptr"], sysm["cols"], sysm["vals"], sysm["diag"]
    Lq = diag * q - np.array([np.dot(vals[rp[r]:rp[r + 1]], q[cols[rp[r]:rp[r + 1]]]) for r in range(len(rows))])
    # interior rows: every PPE pair is an unknown, so diag == sum of the row
    rowsum = np.array([vals[rp[r]:rp[r + 1]].sum() for r in range(len(rows))])
    interior = np.abs(diag - rowsum) <= 1e-12 * np.abs(diag)
    # independent lattice sums: 2 * sum r_x^2 dW6(r)/r / sum r_x^2 dW2(r)/r
    h, kap = 1.5 * dx, 2.0
    g = np.arange(-3, 4)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    r = dx * np.sqrt(X ** 2 + Y ** 2 + Z ** 2).ravel()
    rx2 = (dx * X.ravel()) ** 2
    m = (r > 0) & (r < kap * h)
    d2 = np.array([oracle.kernel_dw(abi.WENDLAND2, kap, h, v) for v in r[m]])
    d6 = np.array([oracle.kernel_dwp(abi.WENDLAND2, kap, h, v) for v in r[m]])
    expect = 2.0 * (rx2[m] * d6 / r[m]).sum() / (rx2[m] * d2 / r[m]).sum()
    # rows whose lattice neighbourhood is complete and symmetric
    full = interior & (np.array([rp[r + 1] - rp[r] for r in range(len(rows))]) == m.sum())
    assert full.sum() > 50
    np.testing.assert_allclose(Lq[full], expect, rtol=1e-8)
    assert expect == pytest.approx(2.0 * 3.342253804929802 / 6.788953041263660, rel=0.05)


def test_bicgstab_matches_sparse_lu():
    P, _ = cases.box(nf=10, walls=3)
    prm = cases.box_params(oracle.default_params(), tol=1e-13)
    o = oracle.Oracle(prm, len(P))
    o.upload(P)
    o.step(0, 1e-4)
    rep = o.ppe_report()
    assert rep.converged and rep.iterations > 5
    s = o.ppe_system()
    n = len(s["rows"])
    rows = np.repeat(np.arange(n), np.diff(s["rowptr"]))
    A = sp.csr_matrix((-s["vals"], (rows, s["cols"])), shape=(n, n)) + sp.diags(s["diag"])
    xs = spla.spsolve(A.tocsc(), s["rhs"])
    np.testing.assert_allclose(s["x"], xs, rtol=1e-9, atol=1e-9 * np.abs(xs).max())
    # the solved pressure is what the step keeps for interior fluid
    p = o.field("PRES")
    np.testing.assert_allclose(p[s["rows"]][:10], xs[:10], rtol=1e-9, atol=1e-9 * np.abs(xs).max())


def test_predictor_corrector_identity():
    P, _ = cases.box(nf=8, walls=3)
    prm = cases.box_params(oracle.default_params(), tol=1e-10)
    o = oracle.Oracle(prm, len(P))
    o.upload(P)
    dt = 1e-4
    o.step(0, dt)
    pt = o.ifield("P_TYPE")
    fl = pt == 1
    for a in ("X", "Y", "Z"):
        ux, u0 = o.field("U" + a), o.field("U" + a + "0")
        ft, fp = o.field("FTOTAL" + a), o.field("FP" + a)
        np.testing.assert_array_equal(ux[fl], (u0[fl] + ft[fl] * dt) + fp[fl] * dt)


def test_most_prescribed_length_closed_form():
    P, nrm, g = cases.small_canyon(dx=0.02)
    prm = cases.canyon_params(g, oracle.default_params(), L=50.0, z0=0.001)
    prm.update(ppe_solver=abi.PPE_REF_JACOBI)
    o = oracle.Oracle(prm, len(P))
    o.upload(P, None, nrm)
    o.step(0, 5e-5)
    S = _sorted_input(o, P)
    most = S["MOST_buffer"] == 1
    assert most.sum() > 10
    z = S["z"][most].astype(np.float64)
    ux = S["ux"][most].astype(np.float64)
    uy = S["uy"][most].astype(np.float64)
    U = np.clip(np.sqrt(ux ** 2 + uy ** 2), 0.001, 100.0)
    phim = np.log(z / 0.001) + 5.0 * (z - 0.001) / 50.0
    ustar = 0.4 * U / phim
    space = S["h"][most].astype(np.float64) / 1.5
    fmx = -S["rho"][most] * ustar * (ux * ustar / U) * space * space / S["m"][most]
    np.testing.assert_allclose(o.field("FLOAT1")[most], ustar, rtol=1e-13)
    np.testing.assert_allclose(o.field("FMX")[most], fmx, rtol=1e-12)
    np.testing.assert_allclose(o.field("FLOAT3")[most], 50.0)


def test_stats_are_field_maxima():
    P, _ = cases.box(nf=8, walls=3)
    prm = cases.box_params(oracle.default_params(), tol=1e-8)
    o = oracle.Oracle(prm, len(P))
    o.upload(P)
    o.step(0, 1e-4)
    umax, drho, fmax = o.stats()
    pt, it = o.ifield("P_TYPE"), o.ifield("I_TYPE")
    ok = ~((pt == 0) | (it >= 2))
    u = np.sqrt(o.field("UX") ** 2 + o.field("UY") ** 2 + o.field("UZ") ** 2)
    assert umax == pytest.approx(float(np.where(ok, u, 0).max()), rel=1e-7)
    assert fmax == pytest.approx(float(np.where(ok, o.field("FTOTAL"), 0).max()), rel=1e-7)
    assert drho == 0.0


def test_ppe_jacobi_is_one_sweep_of_the_operator():
    """REF_JACOBI applies q <- q + (rhs - A q)/diag once (the reference's
    sweep with the race removed); check against the exported operator."""
    P, _ = cases.box(nf=8, walls=3)
    prm = cases.box_params(oracle.default_params(), tol=1e-12)
    o1 = oracle.Oracle(prm, len(P))
    o1.upload(P)
    o1.step(0, 1e-4)
    s = o1.ppe_system()
    prm2 = cases.box_params(oracle.default_params())
    prm2.update(ppe_solver=abi.PPE_REF_JACOBI)
    o2 = oracle.Oracle(prm2, len(P))
    o2.upload(P)
    o2.step(0, 1e-4)
    # one Jacobi sweep from x0 on the BiCGSTAB system
    n = len(s["rows"])
    rows = np.repeat(np.arange(n), np.diff(s["rowptr"]))
    A = sp.csr_matrix((-s["vals"], (rows, s["cols"])), shape=(n, n)) + sp.diags(s["diag"])
    x1 = s["x0"] + (s["rhs"] - A @ s["x0"]) / s["diag"]
    # compare only rows not overwritten later by the boundary kernels
    p = o2.field("PRES")[s["rows"]]
    np.testing.assert_allclose(p, x1, rtol=1e-10, atol=1e-10 * np.abs(x1).max())


def test_nonfinite_state_ends_the_solve_at_once():
    """A NaN in the state makes the PPE right-hand side NaN: the solve stops
    without iterating and the step reports SPH_E_NONFINITE."""
    P, _ = cases.box(nf=6, walls=3)
    fl = np.nonzero(P["p_type"] == 1)[0]
    P["ux"][fl[len(fl) // 2]] = np.nan
    prm = cases.box_params(oracle.default_params(), tol=1e-8)
    prm.update(ppe_solver=abi.PPE_BICGSTAB)
    o = oracle.Oracle(prm, len(P))
    o.upload(P)
    assert o.step(0, 1e-4) == abi.E_NONFINITE
    r = o.ppe_report()
    assert r.converged == 0 and r.iterations == 0 and not np.isfinite(r.rel_residual)
