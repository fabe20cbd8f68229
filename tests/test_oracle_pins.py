"""Pins of the oracle passes that the whole-step checks left open, each run
on its own through the oracle's per-pass hooks, and a mutation test that
plants one-line mistakes in the oracle and shows that the pin of the
affected pass fails.

Two kinds of pin:

* closed forms on a dyadic lattice (positions exact in float32, so every
  interior neighbourhood is exactly symmetric): the SPH sum of a field whose
  continuous answer is known reduces to a moment of the kernel over the
  lattice offsets, summed here independently of the particle arrays;
* pair sums over all particle pairs written in numpy from the reference
  kernels' text (the window rule of the cell search, the per-kernel radius,
  the reference's coefficients), on jittered clouds with varying density,
  mass, smoothing length and viscosity, so that operand swaps that a uniform
  lattice cannot see show up;
* the MOST Obukhov length against scipy's brentq root of the bulk
  Richardson relation with Businger-Dyer stability functions written out
  from the textbook (Paulson 1970, Dyer 1974).

The mutants are compiled into the oracle behind oracle_set_mutant(); every
other run has none.
"""
import functools
import math

import numpy as np
import pytest
from scipy.optimize import brentq

import abi
import cases
import oracle
from test_oracle_nnps import _axis

DX = 2.0 ** -7           # dyadic lattice spacing
KAP = 2.0
A_W2 = 21.0 / (2.0 * math.pi)


# ---------------------------------------------------------------- helpers --
def w2(r, h):
    """Wendland C2 in 3-D (Wendland 1995) with support kappa*h."""
    q = np.asarray(r, np.float64) / (KAP * h)
    return np.where(q < 1, A_W2 / (KAP * h) ** 3 * (1 - q) ** 4 * (1 + 4 * q), 0.0)


def dw2(r, h):
    q = np.asarray(r, np.float64) / (KAP * h)
    return np.where(q < 1, -20.0 * q * (1 - q) ** 3 * A_W2 / (KAP * h) ** 4, 0.0)


def lattice_offsets(rmax, spacing=DX, self_too=False):
    g = np.arange(-5, 6)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    v = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1) * spacing
    r = np.sqrt((v ** 2).sum(1))
    keep = (r < rmax) & ((r > 0) | self_too)
    return v[keep], r[keep]


def _searched(P, prm, normals=None):
    o = oracle.Oracle(prm, len(P))
    o.upload(P, None, normals)
    o.search()
    return o


def _state(o, P):
    """Working state after the search, sorted order; eli (no field id) from
    the deck through the search permutation."""
    S = {k: o.work(k) for k in ("X", "Y", "Z", "UX", "UY", "UZ", "M", "RHO", "H", "PRES", "PRES0", "VIS_T",
                                "TEMP", "FBX", "FBY", "FBZ", "NX", "NY", "NZ")}
    for k in ("P_TYPE", "BUFFER_TYPE", "I_TYPE", "NCELL", "MOST_BUFFER"):
        S[k] = o.ifield(k)
    S["ELI"] = P["eli"][o.nnps()[1]].astype(np.float64)
    S["pos"] = np.stack([S["X"], S["Y"], S["Z"]], 1)
    return S


class Grid:
    """The reference's two cell formulas (binning and dcell query) in float32
    and the window rule: j is scanned from centre cell c with reach nc when
    its binning cell is within nc of c on every axis; cells are scanned in
    ascending key, which is ascending sorted index."""

    def __init__(self, o, S):
        d = o.domain()
        self.lo = (d.x_min, d.y_min, d.z_min)
        self.hi = (d.x_max, d.y_max, d.z_max)
        self.N = (d.NI, d.NJ, d.NK)
        self.dcell = d.dcell
        self.bin = self.cells(S["pos"], query=False)
        self.key = self.bin[:, 0] + self.N[0] * self.bin[:, 1] + self.N[0] * self.N[1] * self.bin[:, 2]

    def cells(self, pts, query):
        pts = np.atleast_2d(pts)
        return np.stack([_axis(pts[:, k], self.lo[k], self.hi[k], self.N[k], self.dcell if query else None)
                         for k in range(3)], 1)

    def window(self, centre, nc):
        return np.all(np.abs(self.bin - centre) <= nc, axis=1)


def _interior(S, lo_idx, hi_idx, dx):
    """Particles whose lattice indices lie in [lo_idx, hi_idx] on each axis."""
    ijk = np.rint(S["pos"] / dx).astype(np.int64)
    return np.all((ijk >= lo_idx) & (ijk <= hi_idx), axis=1)


def _close(got, exp, rtol, what, atol_frac=0.0):
    got, exp = np.asarray(got, np.float64), np.asarray(exp, np.float64)
    scale = np.abs(exp).max() if exp.size else 0.0
    err = np.abs(got - exp)
    bad = err > rtol * np.abs(exp) + atol_frac * scale
    assert not bad.any(), f"{what}: {bad.sum()} of {bad.size} off, worst {err.max():.3e} (scale {scale:.3e})"


# ------------------------------------------------------- pressure force --
def pin_pressure_force_lattice():
    """p = a.x, uniform rho, m, h: interior force is
    f = (m / rho^2) a sum_lattice r_x^2 dW(r)/r (functions_PPE.cuh:482-647),
    which tends to -grad p / rho."""
    P, _ = cases.box(nf=12, walls=3, jitter=0.0, dx=DX)
    o = _searched(P, cases.box_params(oracle.default_params()))
    S = _state(o, P)
    a = np.array([3.0e3, -1.25e3, 0.5e3])
    o.set_work("PRES", S["pos"] @ a)
    o.set_work("PRES0", 0.0)
    o.run_pass("pressure_force")
    v, r = lattice_offsets(KAP * 1.5 * DX)
    s2 = (v[:, 0] ** 2 * dw2(r, 1.5 * DX) / r).sum()
    m, rho = 1000.0 * DX ** 3, 1000.0
    sel = _interior(S, 3, 14, DX) & (S["P_TYPE"] == 1)
    assert sel.sum() >= 200
    for k, f in enumerate(("FPX", "FPY", "FPZ")):
        _close(o.work(f)[sel], m / rho ** 2 * a[k] * s2, 1e-10, f)
    assert m / rho ** 2 * s2 == pytest.approx(-1.0 / rho, rel=0.05)


@functools.lru_cache(maxsize=None)
def _cloud(seed=11):
    """Jittered 16^3 box (dx = 0.01 m) with h in [0.8, 1] h_ref, rho within
    10 %, m = rho dx^3, and per-particle pressure, pres0, velocity, eddy
    viscosity and eli."""
    P, _ = cases.box(nf=10, walls=3, jitter=0.05)
    n = len(P)
    u = cases.uniform(seed, 9 * n).reshape(9, n)
    P["h"] = P["h_ref"] * (0.8 + 0.2 * u[0])
    P["rho"] = 1000.0 * (1.0 + 0.1 * u[1])
    P["m"] = P["rho"] * 0.01 ** 3
    P["pres"] = 50.0 * (u[2] - 0.5)
    P["pres0"] = 5.0 * (u[3] - 0.5)
    fl = P["p_type"] == 1
    P["ux"] = np.where(fl, 0.3 * (u[4] - 0.5), 0.0)
    P["uy"] = np.where(fl, 0.3 * (u[5] - 0.5), 0.0)
    P["uz"] = np.where(fl, 0.3 * (u[6] - 0.5), 0.0)
    P["vis_t"] = 1e-4 * u[7]
    P["eli"] = np.where(u[8] < 0.3, 0.25, 0.0)
    return P


def _pf_expected(S, g, prm):
    n = len(S["X"])
    out = np.zeros((n, 3))
    pq = S["PRES"] - S["PRES0"]
    for i in range(n):
        if S["I_TYPE"][i] > 2 or S["P_TYPE"][i] >= 1000 or S["P_TYPE"][i] <= 0 or S["BUFFER_TYPE"][i] != 0:
            continue
        c = g.cells(S["pos"][i], query=True)[0]
        d = S["pos"][i] - S["pos"]
        td = np.sqrt((d ** 2).sum(1)) - 1e-20
        hi, hj = S["H"][i], S["H"]
        ok = g.window(c, S["NCELL"][i]) & (S["P_TYPE"] < 1000) & (td < KAP * np.maximum(hi, hj))
        j = np.nonzero(ok)[0]
        hij = 0.5 * (hi + hj[j])
        dw = dw2(td[j], hij)
        rhoi, rhoj = S["RHO"][i], S["RHO"][j]
        cp = -2.0 * S["M"][j] / (rhoi * rhoj) * (pq[i] * rhoj + pq[j] * rhoi) / (rhoi + rhoj)
        out[i] = ((cp * dw / td[j])[:, None] * d[j]).sum(0)
    return out


@functools.lru_cache(maxsize=None)
def _pf_case():
    P = _cloud()
    prm = cases.box_params(oracle.default_params())
    o = _searched(P, prm)
    S = _state(o, P)
    return P, prm, _pf_expected(S, Grid(o, S), prm)


def pin_pressure_force_pairs():
    P, prm, exp = _pf_case()
    o = _searched(P, prm)
    o.run_pass("pressure_force")
    got = np.stack([o.work(f) for f in ("FPX", "FPY", "FPZ")], 1)
    _close(got, exp, 1e-10, "pressure force", atol_frac=1e-12)


# ------------------------------------------------------------ advection --
def _lattice_adv(eli, **upd):
    P, _ = cases.box(nf=14, walls=3, jitter=0.0, dx=DX)
    P["eli"] = eli
    prm = cases.box_params(oracle.default_params())
    prm.update(**upd)
    return _searched(P, prm), prm, P


def pin_cleary_lattice():
    """u = (c y^2, 0, 0), uniform mu, rho, m: the Cleary sum
    (functions_PPE.cuh:326-343) is -2 c mu m / rho^2 sum r_y^2 dW/r -> 2 c nu."""
    o, prm, P = _lattice_adv(1.0, fv_solve=abi.FV_CLEARY, artificial_viscosity=0)   # eli = 1: no Euler term
    S = _state(o, P)
    c = 0.75
    o.set_work("UX", c * S["Y"] ** 2)
    o.set_work("UY", 0.0)
    o.set_work("UZ", 0.0)
    o.run_pass("advection")
    v, r = lattice_offsets(KAP * 1.5 * DX)
    s2 = (v[:, 1] ** 2 * dw2(r, 1.5 * DX) / r).sum()
    mu, m, rho = prm.mu_fluid, 1000.0 * DX ** 3, 1000.0
    sel = _interior(S, 5, 14, DX) & (S["P_TYPE"] == 1)   # every neighbour is fluid
    assert sel.sum() >= 500
    _close(o.work("FTOTALX")[sel], -2.0 * c * mu * m / rho ** 2 * s2, 1e-9, "Cleary x")
    assert np.abs(o.work("FTOTALY")[sel]).max() == 0.0
    assert -2.0 * c * mu * m / rho ** 2 * s2 == pytest.approx(2.0 * c * mu / rho, rel=0.05)


def pin_euler_lattice():
    """u = A x: the Euler convection sum (functions_PPE.cuh:372-386) is
    -V S2 A u_i, so ftotal = V S2 A u_i -> -(u.grad)u."""
    o, prm, P = _lattice_adv(0.0, fv_solve=abi.FV_NONE, artificial_viscosity=0)
    S = _state(o, P)
    A = np.array([[0.5, -0.25, 0.125], [0.375, 0.25, -0.5], [-0.125, 0.75, 0.25]])
    u = S["pos"] @ A.T
    for k, f in enumerate(("UX", "UY", "UZ")):
        o.set_work(f, u[:, k])
    o.run_pass("advection")
    v, r = lattice_offsets(KAP * 1.5 * DX)
    s2 = (v[:, 0] ** 2 * dw2(r, 1.5 * DX) / r).sum()
    V = DX ** 3
    sel = _interior(S, 3, 16, DX) & (S["P_TYPE"] == 1)
    assert sel.sum() >= 1000
    exp = V * s2 * (u @ A.T)
    for k, f in enumerate(("FTOTALX", "FTOTALY", "FTOTALZ")):
        _close(o.work(f)[sel], exp[sel, k], 1e-9, f, atol_frac=1e-12)


def _adv_expected(S, g, prm):
    """KERNEL_advection_force3D (functions_PPE.cuh:191-480) over all pairs:
    Cleary or Monaghan viscosity, Euler convection, artificial viscosity."""
    n = len(S["X"])
    out = np.zeros((n, 3))
    mu_of = lambda pt: np.where(pt == 1, prm.mu_fluid, np.where((pt == 0) | (pt == 1000), prm.mu_wall,
                                                                 prm.mu_fluid))
    for i in range(n):
        pti = S["P_TYPE"][i]
        if S["I_TYPE"][i] > 2 or pti == 1000 or pti <= 0:
            continue
        c = g.cells(S["pos"][i], query=True)[0]
        d = S["pos"][i] - S["pos"]                       # x_i - x_j
        td = np.sqrt((d ** 2).sum(1)) + 1e-20
        hi, hj = S["H"][i], S["H"]
        ok = g.window(c, S["NCELL"][i]) & (S["P_TYPE"] < 1000) & (td < KAP * np.maximum(hi, hj))
        j = np.nonzero(ok)[0]
        dj, tj = d[j], td[j]
        dw = dw2(tj, hi)
        grad = (dw / tj)[:, None] * dj                   # dW (x_i - x_j)/r
        ui = np.array([S["UX"][i], S["UY"][i], S["UZ"][i]])
        uj = np.stack([S["UX"][j], S["UY"][j], S["UZ"][j]], 1)
        mi, mj = S["M"][i], S["M"][j]
        rhoi, rhoj = S["RHO"][i], S["RHO"][j]
        visi = mu_of(np.array([pti]))[0] + S["VIS_T"][i] * rhoi
        visj = mu_of(S["P_TYPE"][j]) + S["VIS_T"][j] * rhoj
        f = np.zeros(3)
        if prm.fv_solve == abi.FV_CLEARY:
            cv = (dj * grad).sum(1) * (visi * visj) / (visi + visj) * 4 * (mj / (rhoi * rhoj)) / tj / tj
            f += (cv[:, None] * (ui - uj)).sum(0)
        elif prm.fv_solve == abi.FV_MONAGHAN:
            uxx = ((ui - uj) * dj).sum(1)
            pij = -2.0 * (3 + 2.0) * (uxx / (tj * tj + 0.01 * hi * hi)) * mj / rhoj
            pij = pij * (2.0 * (visi * visj) / (visi + visj - 1e-20) / rhoi)
            pij = np.where(S["P_TYPE"][j] <= 0, 0.0, pij)
            f += (-pij[:, None] * grad).sum(0)
        if prm.artificial_viscosity:
            uxx = ((ui - uj) * dj).sum(1)
            hij = 0.5 * (hi + hj[j])
            phi = hij * uxx / (tj * tj + 0.01 * hij * hij)
            pij = mi * phi * (-prm.av_alpha * prm.av_c + prm.av_beta * phi) / (rhoi + rhoj) * 0.5
            pij = np.where(uxx < 0, pij, 0.0)
            f += (-pij[:, None] * grad).sum(0)
        eli = S["ELI"][i]
        if eli < 1.0:
            conv = (grad * ui).sum(1)                    # u_i . grad W
            f -= ((conv * mj / rhoj * (1 - eli))[:, None] * (uj - ui)).sum(0)
        out[i] = f
    return out


@functools.lru_cache(maxsize=None)
def _adv_case(fv):
    P = _cloud()
    prm = cases.box_params(oracle.default_params())
    prm.update(fv_solve=fv, artificial_viscosity=1)
    o = _searched(P, prm)
    S = _state(o, P)
    return P, prm, _adv_expected(S, Grid(o, S), prm)


def _pin_adv(fv):
    P, prm, exp = _adv_case(fv)
    o = _searched(P, prm)
    o.run_pass("advection")
    got = np.stack([o.work(f) for f in ("FTOTALX", "FTOTALY", "FTOTALZ")], 1)
    _close(got, exp, 1e-9, "advection", atol_frac=1e-12)


def pin_advection_pairs_cleary():
    _pin_adv(abi.FV_CLEARY)


def pin_advection_pairs_monaghan():
    _pin_adv(abi.FV_MONAGHAN)


# ------------------------------------------------------------------ IBM --
@functools.lru_cache(maxsize=None)
def _ibm_deck():
    """Dyadic fluid lattice with a square plate of IBM markers across its
    middle (normals mixed) and a uniform stream."""
    P, _ = cases.box(nf=14, walls=3, jitter=0.0, dx=DX)
    ijk = np.rint(np.stack([P["x"], P["y"], P["z"]], 1) / DX).astype(int)
    plate = (ijk[:, 2] == 9) & np.all((ijk[:, :2] >= 6) & (ijk[:, :2] <= 13), axis=1)
    P["p_type"][plate] = 1000
    nrm = np.zeros((len(P), 3), np.float32)
    nrm[plate] = (0.0, 0.0, 1.0)
    tilt = plate & (ijk[:, 0] >= 10)
    nrm[tilt] = (0.6, 0.0, 0.8)
    side = plate & (ijk[:, 1] >= 11)
    nrm[side] = (0.0, -2.0, 0.0)      # not unit: the pass normalises
    U = (0.3, -0.2, 0.1)
    fl = P["p_type"] == 1
    P["ux"], P["uy"], P["uz"] = np.where(fl, U[0], 0.0), np.where(fl, U[1], 0.0), np.where(fl, U[2], 0.0)
    return P, nrm, np.array(U, np.float32).astype(np.float64)


def pin_ibm_force():
    """Marker force (functions_IBM.cuh:342-490) in a uniform stream: Shepard
    velocity U filt/(filt + 1e-10) with filt a lattice sum of W at h/2 over
    the fluid neighbours; force rho_air (-(u.n) n + U_target - u_t) / dt."""
    P, nrm, U = _ibm_deck()
    prm = cases.box_params(oracle.default_params())
    prm.update(ibm_ut_factor=0.2, rho_air=1.225)
    o = _searched(P, prm, nrm)
    S = _state(o, P)
    dt = 1e-4
    o.run_pass("ibm_force", dt=dt)
    mk = np.nonzero(S["P_TYPE"] == 1000)[0]
    fl = S["P_TYPE"] == 1
    h = 1.5 * DX
    filt = np.array([(DX ** 3 * w2(np.sqrt(((S["pos"][fl] - S["pos"][i]) ** 2).sum(1)) + 1e-20, 0.5 * h)).sum()
                     for i in mk])
    _close(o.work("CONCENTRATION")[mk], filt, 1e-12, "marker Shepard sum")
    uf = np.outer(filt / (filt + 1e-10), U)
    n = np.stack([S["NX"][mk], S["NY"][mk], S["NZ"][mk]], 1)
    n = n / (np.linalg.norm(n, axis=1) + 1e-20)[:, None]
    un = (uf * n).sum(1)
    ut = uf - un[:, None] * n
    tgt = np.zeros_like(uf)
    tgt[:, 0] = 0.2 * np.array([oracle.lib.oracle_interp_ux_caseH(z) for z in S["Z"][mk]])
    fb = 1.225 * (-un[:, None] * n + (tgt - ut)) / dt
    for k, f in enumerate(("FBX", "FBY", "FBZ")):
        _close(o.work(f)[mk], fb[:, k], 1e-12, f, atol_frac=1e-14)


def pin_ibm_spread():
    """Fluid force (functions_IBM.cuh:492-613): Shepard average of the
    markers' forces with W at h/2, zero where no marker is in reach."""
    P, nrm, _ = _ibm_deck()
    prm = cases.box_params(oracle.default_params())
    o = _searched(P, prm, nrm)
    S = _state(o, P)
    mkm = S["P_TYPE"] == 1000
    fb = np.stack([1e3 * S["X"] - 2.0, -5e2 * S["Y"] + 3.0 * S["Z"], 7.0 + 1e2 * S["X"] * S["Y"]], 1)
    for k, f in enumerate(("FBX", "FBY", "FBZ")):
        o.set_work(f, np.where(mkm, fb[:, k], 0.0))
    o.run_pass("ibm_spread")
    h = 1.5 * DX
    mk = np.nonzero(mkm)[0]
    fl = np.nonzero((S["P_TYPE"] == 1) & (S["BUFFER_TYPE"] != 2))[0]
    near = 0
    for i in fl:
        w = DX ** 3 * w2(np.sqrt(((S["pos"][mk] - S["pos"][i]) ** 2).sum(1)) + 1e-20, 0.5 * h)
        filt = w.sum()
        exp = (w[:, None] * fb[mk]).sum(0) / (filt + 1e-10) if filt >= 1e-12 else np.zeros(3)
        near += filt >= 1e-12
        S.setdefault("_exp", {})[i] = (exp, filt)
    got = np.stack([o.work(f) for f in ("FBX", "FBY", "FBZ")], 1)
    conc = o.work("CONCENTRATION")
    ids = np.array(sorted(S["_exp"]))
    exp = np.array([S["_exp"][i][0] for i in ids])
    filt = np.array([S["_exp"][i][1] for i in ids])
    assert near >= 100
    _close(got[ids], exp, 1e-11, "spread force", atol_frac=1e-13)
    _close(conc[ids], np.where(filt >= 1e-12, filt, 0.0), 1e-12, "spread Shepard sum")


# ------------------------------------------------- periodic and outlet ----
def _copy_deck(extra_marker=False):
    """Fluid lattice (no walls) with periodic side buffers, an outlet slab
    and duplicate particles 0.3 mm from some copy sources, so that the
    reference's last-match rule decides which value is copied."""
    nf = 16
    P, _ = cases.box(nf=nf, walls=0, jitter=0.0, dx=DX)
    n = len(P)
    u = cases.uniform(23, 5 * n).reshape(5, n)
    P["ux"], P["uy"], P["uz"] = u[0], u[1], u[2]
    P["pres"], P["temp"] = 100.0 * u[3], 290.0 + u[4]
    ijk = np.rint(np.stack([P["x"], P["y"], P["z"]], 1) / DX).astype(int)
    src = (ijk[:, 1] >= nf - 6) & (ijk[:, 1] <= nf - 4) & (ijk[:, 0] % 3 == 0)
    src |= (ijk[:, 1] >= 3) & (ijk[:, 1] <= 5) & (ijk[:, 2] % 2 == 0)
    src |= (ijk[:, 0] == nf - 5) & (ijk[:, 1] % 2 == 1)
    dup = P[src].copy()
    dup["x"] = dup["x"] + 0.0003
    v = cases.uniform(29, 2 * len(dup)).reshape(2, len(dup))
    dup["ux"], dup["pres"] = 10.0 + v[0], 500.0 + v[1]
    if extra_marker:
        mk = dup[: len(dup) // 3].copy()
        mk["p_type"] = 1000
        mk["x"] = mk["x"] - 0.0001
        dup = np.concatenate([dup, mk])
    P = np.concatenate([P, dup])
    ijk = np.rint(np.stack([P["x"], P["y"], P["z"]], 1) / DX).astype(int)
    P["buffer_type"][ijk[:, 1] <= 2] = 3
    P["buffer_type"][ijk[:, 1] >= nf - 3] = 4
    outl = (ijk[:, 0] >= nf - 3) & (P["buffer_type"] == 0)
    P["buffer_type"][outl] = 2
    # keep the last particle (skipped by the domain bounds) inside the box
    P = np.concatenate([P[-1:], P[:-1]])
    return P, nf


def _first_cell_last_match(g, S, centre, nc, ghost, tol, filt=None):
    ok = g.window(centre, nc) & (np.sqrt(((S["pos"] - ghost) ** 2).sum(1)) - 1e-20 < tol)
    if filt is not None:
        ok &= filt
    j = np.nonzero(ok)[0]
    if not len(j):
        return -1
    kmin = g.key[j].min()
    return j[g.key[j] == kmin].max()


def pin_periodic():
    """Periodic copy (functions_BC.cuh:1015-1154): ghost at y +- (L_right -
    L_left); the values of the last match within 1 mm in the first window
    cell (scan order) that holds one."""
    P, nf = _copy_deck()
    prm = oracle.default_params()
    prm.update(L_left=2.5 * DX, L_right=(nf - 3.5) * DX, copy_tol=0.001)
    o = _searched(P, prm)
    S = _state(o, P)
    g = Grid(o, S)
    o.run_pass("periodic")
    L = prm.L_right - prm.L_left
    bt = S["BUFFER_TYPE"]
    fields = ("UX", "UY", "UZ", "PRES", "TEMP")
    got = {f: o.work(f) for f in fields}
    checked = dups = 0
    for i in np.nonzero(((bt == 3) | (bt == 4)) & (S["P_TYPE"] < 1000))[0]:
        ghost = S["pos"][i] + np.array([0.0, L if bt[i] == 3 else -L, 0.0])
        c = g.cells(ghost, query=True)[0]
        j = _first_cell_last_match(g, S, c, S["NCELL"][i], ghost, 0.001)
        assert j >= 0
        for f in fields:
            assert got[f][i] == S[f][j], (f, i, j)
        checked += 1
        dups += S["UX"][j] >= 10.0
    assert checked > 500 and dups > 20


def pin_outlet():
    """Outlet copy (functions_BC.cuh:673-761): ghost (L3 - 0.008, y, z),
    window of 2 ncell around the particle's own cell, the last match within
    1 mm among particles below p_type 1000, zeros when none is in reach."""
    P, nf = _copy_deck(extra_marker=True)
    prm = oracle.default_params()
    prm.update(L3=(nf - 3.5) * DX, outlet_copy_dx=1.5 * DX, copy_tol=0.001)
    o = _searched(P, prm)
    S = _state(o, P)
    g = Grid(o, S)
    o.run_pass("outlet")
    fields = ("UX", "UY", "UZ", "PRES", "TEMP")
    got = {f: o.work(f) for f in fields}
    checked = dups = 0
    for i in np.nonzero(S["BUFFER_TYPE"] == 2)[0]:
        ghost = np.array([prm.L3 - prm.outlet_copy_dx, S["Y"][i], S["Z"][i]])
        c = g.cells(S["pos"][i], query=True)[0]
        ok = (g.window(c, 2 * S["NCELL"][i]) & (S["P_TYPE"] < 1000)
              & (np.sqrt(((S["pos"] - ghost) ** 2).sum(1)) - 1e-20 < 0.001))
        j = np.nonzero(ok)[0]
        for f in fields:
            assert got[f][i] == (S[f][j.max()] if len(j) else 0.0), (f, i)
        checked += 1
        dups += len(j) > 1
    assert checked > 300 and dups > 20


# -------------------------------------------------------------- Neumann --
@functools.lru_cache(maxsize=None)
def _neumann_case():
    P, _ = cases.box(nf=12, walls=3, jitter=0.05)
    n = len(P)
    u = cases.uniform(31, 4 * n).reshape(4, n)
    P["pres"] = 40.0 * (u[0] - 0.5) + 500.0 * P["x"]
    P["h"] = P["h_ref"] * (0.85 + 0.15 * u[1])
    P["rho"] = 1000.0 * (1.0 + 0.05 * u[2])
    P["m"] = P["rho"] * 0.01 ** 3
    ijk = np.rint(np.stack([P["x"], P["y"], P["z"]], 1) / 0.01).astype(int)
    wall = P["p_type"] == 0
    P["buffer_type"][wall & (ijk[:, 0] <= 2)] = 1            # inlet walls: mirrored ghost
    P["p_type"][wall & (ijk[:, 2] >= 15) & (u[3] < 0.5)] = -1   # some free-slip lid particles
    fl = P["p_type"] == 1
    P["buffer_type"][fl & (ijk[:, 1] == 3) & (u[3] < 0.3)] = 3  # fluid outside buffer 0: not averaged
    P["ux"] = np.where(fl, 0.2 * u[3], 0.0)
    P["uy"] = np.where(fl, -0.1 * u[2], 0.0)
    P["uz"] = np.where(fl, 0.05 * u[1], 0.0)
    prm = oracle.default_params()
    prm.update(L2=0.045, neumann_ghost_r=0.1)
    o = _searched(P, prm)
    S = _state(o, P)
    g = Grid(o, S)
    exp = {}
    for i in np.nonzero(S["P_TYPE"] != 1)[0]:
        xi = S["pos"][i]
        ghost = xi.copy()
        bt = S["BUFFER_TYPE"][i]
        if bt in (1, -1):
            ghost[0] = 2.0 * prm.L2 - xi[0]
        c = g.cells(ghost, query=True)[0]
        dist = np.sqrt(((S["pos"] - ghost) ** 2).sum(1)) - 1e-20
        cand = np.nonzero(g.window(c, 1) & (dist < prm.neumann_ghost_r))[0]
        pg = S["PRES"][cand.max()] if len(cand) else 0.0
        h2 = 2.0 * S["H"][i]
        j = np.nonzero(g.window(c, S["NCELL"][i]) & (S["P_TYPE"] < 1000) & (dist < KAP * h2))[0]
        nb = ((S["P_TYPE"][j] == 1) & (S["BUFFER_TYPE"][j] == 0)).astype(float)
        V = S["M"][j] / S["RHO"][j]
        w = w2(dist[j], h2)
        gw = (dw2(dist[j], h2) / dist[j])[:, None] * (ghost - S["pos"][j])
        flt = (V * w * nb).sum()
        if flt < 1e-6:
            exp[i] = (0.0, None)
            continue
        p = (V * S["PRES"][j] * w * nb).sum() / flt
        vel = None
        if S["P_TYPE"][i] == -1:
            vel = tuple((V * S[f][j] * w * nb).sum() / flt for f in ("UX", "UY", "UZ"))
        if bt in (1, -1):
            grad = ((V * (S["PRES"][j] - pg) * nb)[:, None] * gw).sum(0)
            p = pg + ((xi - ghost) * grad).sum()
        exp[i] = (p, vel)
    return P, prm, exp


def pin_neumann():
    """Neumann pass (functions_BC.cuh:290-446): Shepard pressure of fluid,
    buffer-0 neighbours with W at 2h over the ncell window; free-slip
    particles also take the velocity; inlet particles (buffer +-1) mirror to
    x_g = 2 L2 - x, take the ghost pressure from the last particle within
    0.1 m of the ghost in its 3x3x3 cells and extrapolate with the SPH
    gradient about it."""
    P, prm, exp = _neumann_case()
    o = _searched(P, prm)
    o.run_pass("neumann")
    pres = o.work("PRES")
    vel = np.stack([o.work(f) for f in ("UX", "UY", "UZ")], 1)
    ids = np.array(sorted(exp))
    e = np.array([exp[i][0] for i in ids])
    _close(pres[ids], e, 1e-10, "Neumann pressure", atol_frac=1e-12)
    vi = [i for i in ids if exp[i][1] is not None]
    assert len(vi) > 20
    _close(vel[vi], np.array([exp[i][1] for i in vi]), 1e-10, "free-slip velocity", atol_frac=1e-12)
    bt = o.ifield("BUFFER_TYPE")
    assert (bt[ids] == 1).sum() > 100


# ----------------------------------------------------------------- MOST --
def _psi_m(z):
    """Businger-Dyer momentum stability function (Paulson 1970)."""
    if z < 0:
        x = (1.0 - 16.0 * z) ** 0.25
        return 2.0 * math.log((1 + x) / 2) + math.log((1 + x * x) / 2) - 2.0 * math.atan(x) + math.pi / 2
    return -5.0 * z


def _psi_h(z):
    if z < 0:
        y = (1.0 - 16.0 * z) ** 0.5
        return 2.0 * math.log((1 + y) / 2)
    return -5.0 * z


def _phi(psi, L, z, z0):
    return math.log(z / z0) - psi(z / L) + psi(z0 / L)


def pin_most_newton():
    """Obukhov length of the reference's Newton iteration
    (functions_MOST.cuh:214-249) against the root of
    Ri_b = (z/L) Phi_h(L) / Phi_m(L)^2 found by brentq, and u*, theta* and
    the wall stress from that root; the neutral branch at |Ri_b| <= 1e-3."""
    P, _ = cases.box(nf=8, walls=3, jitter=0.0, dx=0.01)
    fl = np.nonzero(P["p_type"] == 1)[0]
    pick = fl[::23][:14]
    P["MOST_buffer"][pick] = 1
    k = np.arange(len(pick))
    P["ux"][pick] = 0.15 + 0.05 * (k % 5)
    P["uy"][pick] = 0.02 * (k % 3) - 0.02
    dT = np.array([0.6, -0.8, 0.3, -2.0, 1.5, -0.2, 0.0004, 0.9, -1.2, 2.5, -3.0, 0.05, -0.05, 1.0])
    P["temp"][pick] = 300.0 + dT
    prm = oracle.default_params()
    prm.update(most_mode=abi.MOST_REFERENCE, most_z0=0.01, most_z0h=0.001, gravity=9.81, T_ini=300.0,
               T_start_t=1.0, T_con_t=2.0, T_con=300.0)
    o = _searched(P, prm)
    S = _state(o, P)
    o.run_pass("most", count=0, dt=1e-4)
    L_got, us_got, ts_got = o.work("FLOAT3"), o.work("FLOAT1"), o.work("FLOAT2")
    fmx, hmz, dtemp = o.work("FMX"), o.work("HMZ"), o.work("DTEMP")
    newton = neutral = 0
    for i in np.nonzero(S["MOST_BUFFER"] == 1)[0]:
        z, T = S["Z"][i], S["TEMP"][i]
        U = min(max(math.hypot(S["UX"][i], S["UY"][i]), 0.001), 100.0)
        rib = (9.81 / T) * (T - 300.0) * z / (U * U + 1e-10)
        z0, z0h = 0.01, 0.001
        if abs(rib) > 1e-3:
            f = lambda L: rib - (z / L) * _phi(_psi_h, L, z, z0h) / _phi(_psi_m, L, z, z0) ** 2
            L = brentq(f, 1e-3 * z, 1e7, xtol=1e-14, rtol=1e-14) if rib > 0 else \
                brentq(f, -1e7, -1e-3 * z, xtol=1e-14, rtol=1e-14)
            us = 0.4 * U / _phi(_psi_m, L, z, z0)
            ts = 0.4 * (T - 300.0) / _phi(_psi_h, L, z, z0h)
            assert L_got[i] == pytest.approx(L, rel=1e-4), (i, rib)
            newton += 1
        else:
            L, us, ts = 1e6, 0.4 * U / math.log(z / z0), 0.4 * (T - 300.0) / math.log(z / z0h)
            assert L_got[i] == L
            neutral += 1
        assert us_got[i] == pytest.approx(us, rel=1e-5)
        assert ts_got[i] == pytest.approx(ts, rel=1e-5, abs=1e-12)
        space = S["H"][i] / 1.5
        assert fmx[i] == pytest.approx(-S["RHO"][i] * us * (S["UX"][i] * us / U) * space * space / S["M"][i],
                                       rel=2e-5)
        # sensible heat flux -rho cp u* theta* 100, half of it deposited over
        # one spacing (functions_MOST.cuh:268-303); the ghost 5 m above has no
        # neighbours here, so the particle keeps its own deposit
        dep = 0.5 * (-S["RHO"][i] * prm.air_cp * us * ts * 100.0) / (space * S["RHO"][i] * prm.air_cp)
        assert hmz[i] == pytest.approx(dep, rel=2e-5, abs=1e-12)
        assert dtemp[i] == pytest.approx(dep, rel=2e-5, abs=1e-12)
    assert newton >= 10 and neutral >= 1


# ---------------------------------------------------------- prep and DDF --
def _prep_lattice(model, k0=None, nu0=0.0, dt=0.05):
    """Dyadic lattice with u = A x and particle volume 10 dx^3 (so that the
    kernel length sigma exceeds h and decides l = max(sigma, h))."""
    P, _ = cases.box(nf=14, walls=3, jitter=0.0, dx=DX)
    P["m"] = 10.0 * P["rho"] * DX ** 3
    prm = cases.box_params(oracle.default_params())
    prm.update(turbulence_model=model, fv_solve=abi.FV_CLEARY, Cs_sgs=0.12, Ck_sgs=0.1, Ce_sgs=0.93)
    o = _searched(P, prm)
    S = _state(o, P)
    A = np.array([[0.25, -0.75, 0.125], [0.5, 0.125, -0.375], [-0.625, 0.875, -0.375]])
    u = S["pos"] @ A.T
    for k, f in enumerate(("UX", "UY", "UZ")):
        o.set_work(f, u[:, k])
    if k0 is not None:
        o.set_work("K_TURB", k0)
        o.set_work("VIS_T", nu0)
    o.run_pass("grad_corr")
    o.run_pass("prep", count=0, dt=dt)
    Ssym = 0.5 * (A + A.T)
    SR = math.sqrt(2.0 * (Ssym * Ssym).sum())
    V = 10.0 * DX ** 3
    h = 1.5 * DX
    v, r = lattice_offsets(KAP * h, self_too=True)
    rr = r + 1e-20
    flt = (V * w2(rr, h)).sum()
    sigma = math.sqrt(0.3333 * (rr * rr * w2(rr, h) * V).sum())
    sel = _interior(S, 5, 14, DX) & (S["P_TYPE"] == 1)
    return o, sel, SR, flt, sigma, h


def pin_prep_lattice():
    """Shepard sum flt_s = V sum W (self included), kernel length
    sigma = sqrt(0.3333 V sum r^2 W), l = max(sigma, h) and the Smagorinsky
    viscosity Cs^2 l^2 |S| (functions_PREP.cuh:395-460)."""
    o, sel, SR, flt, sigma, h = _prep_lattice(abi.SSM)
    assert sel.sum() >= 500 and sigma > h
    _close(o.work("FLT_S")[sel], flt, 1e-12, "flt_s")
    _close(o.work("SR")[sel], SR, 1e-10, "|S|")
    _close(o.work("VIS_T")[sel], 0.12 ** 2 * sigma ** 2 * SR, 1e-10, "SSM viscosity")


def pin_ddf():
    """Deardorff one-equation model (functions_PREP.cuh:462-478):
    P = nu |S|^2, eps = Ce k^1.5 / l with the step's initial k,
    k' = max(k + dt (P - eps), 1e-4), nu_t = Ck l sqrt(k')."""
    for k0, nu0, dt in ((0.02, 3e-4, 0.05), (2e-4, 0.0, 1.0)):
        o, sel, SR, flt, sigma, h = _prep_lattice(abi.DDF, k0=k0, nu0=nu0, dt=dt)
        l = max(sigma, h)
        eps = 0.93 * k0 ** 1.5 / l
        k1 = max(k0 + dt * (nu0 * SR * SR - eps), 1e-4)
        _close(o.work("K_TURB")[sel], k1, 1e-12, "k")
        _close(o.work("E_TURB")[sel], eps, 1e-12, "eps")
        _close(o.work("VIS_T")[sel], 0.1 * l * math.sqrt(k1), 1e-12, "nu_t")
        if nu0 == 0.0:
            assert k0 + dt * (nu0 * SR * SR - eps) < 1e-4     # the floor case


# ------------------------------------------------------- IB normals --
def _face_deck(mode):
    P, nrm, geom, faces = cases.face_marker_canyon()
    prm = cases.canyon_params(geom, oracle.default_params(), tol=1e-8)
    prm.update(ib_normals=mode, **faces)
    return P, nrm, geom, faces, prm


def pin_ib_normals():
    """Marker normals of the face-plane rule (functions_PREP.cuh:923-1015):
    on a face the outward unit normal of that face, on an edge the outward
    bisector, on a corner the outward diagonal, each of unit length to float
    precision; zero off the planes and on every non-marker."""
    P, nrm, geom, faces, prm = _face_deck(abi.IBN_REFERENCE)
    o = _searched(P, prm, nrm * 0.0)
    assert o.run_pass("ib_normals", count=0) == 0
    S = _state(o, P)
    pos, mk = S["pos"], S["P_TYPE"] >= 1000
    n = np.stack([o.work("NX"), o.work("NY"), o.work("NZ")], 1)
    assert np.all(n[~mk] == 0.0)
    lo = np.array([faces["ib_face_xlo"], faces["ib_face_ylo"], faces["ib_face_zlo"]])
    hi = np.array([faces["ib_face_xhi"], faces["ib_face_yhi"], faces["ib_face_zhi"]])
    on_lo = np.abs(pos - lo) < 1e-5
    on_hi = np.abs(pos - hi) < 1e-5
    out = on_hi.astype(float) - on_lo.astype(float)        # outward direction, per axis
    k = np.abs(out).sum(1)
    assert (k[mk] == 1).sum() > 40 and (k[mk] == 2).sum() >= 10 and (k[mk] == 3).sum() == 4
    assert (k[mk] == 0).sum() == 1
    for m in (1, 2, 3):
        sel = mk & (k == m)
        _close(n[sel], out[sel] / math.sqrt(m), 1e-7, f"normal on {m} planes")
        _close(np.linalg.norm(n[sel], axis=1), np.ones(sel.sum()), 1e-7, f"unit length on {m} planes")
    assert np.all(n[mk & (k == 0)] == 0.0)


PINS = {f.__name__: f for f in (
    pin_ib_normals,
    pin_pressure_force_lattice, pin_pressure_force_pairs, pin_cleary_lattice, pin_euler_lattice,
    pin_advection_pairs_cleary, pin_advection_pairs_monaghan, pin_ibm_force, pin_ibm_spread, pin_periodic,
    pin_outlet, pin_neumann, pin_most_newton, pin_prep_lattice, pin_ddf)}

# planted mistake -> pins that must catch it (oracle g_mutant values)
MUTANTS = {
    1: ("pin_pressure_force_lattice", "pin_pressure_force_pairs"),   # -2 -> -1
    2: ("pin_pressure_force_pairs",),                                 # h_ij = max(hi, hj)
    3: ("pin_pressure_force_pairs",),                                 # rho_i / rho_j swapped
    4: ("pin_cleary_lattice", "pin_advection_pairs_cleary"),          # Cleary 4 -> 2
    5: ("pin_cleary_lattice", "pin_advection_pairs_cleary"),          # arithmetic for harmonic mean
    6: ("pin_advection_pairs_cleary",),                               # vis_t of i for j
    7: ("pin_advection_pairs_cleary",),                               # AV h_ij = h_i
    8: ("pin_advection_pairs_cleary",),                               # AV 0.01 h^2 -> 0.1 h^2
    9: ("pin_advection_pairs_cleary",),                               # AV factor 0.5 dropped
    10: ("pin_advection_pairs_cleary", "pin_advection_pairs_monaghan"),  # convecting velocity u_j
    11: ("pin_euler_lattice", "pin_advection_pairs_cleary"),          # Euler term sign
    12: ("pin_ibm_force",),                                           # marker kernel at h
    13: ("pin_ibm_force",),                                           # target speed along y
    14: ("pin_ibm_force",),                                           # normal correction sign
    15: ("pin_ibm_spread",),                                          # spread not normalised
    16: ("pin_ibm_spread",),                                          # spread kernel at h
    17: ("pin_periodic",),                                            # ghost offset sign
    18: ("pin_periodic",),                                            # first match kept
    19: ("pin_outlet",),                                              # copy plane L3 + dx
    20: ("pin_outlet",),                                              # first match kept
    21: ("pin_neumann",),                                             # inlet mirror sign
    22: ("pin_neumann",),                                             # first ghost match
    23: ("pin_neumann",),                                             # extrapolation sign
    24: ("pin_neumann",),                                             # kernel at h, not 2h
    25: ("pin_most_newton",),                                         # psi_h constant 16 -> 9
    26: ("pin_most_newton",),                                         # z0 for z0h in the iteration
    27: ("pin_most_newton",),                                         # Newton step sign
    28: ("pin_prep_lattice", "pin_ddf"),                              # sigma factor 0.3333 -> 0.5
    29: ("pin_prep_lattice", "pin_ddf"),                              # l = min(sigma, h)
    30: ("pin_prep_lattice",),                                        # Shepard without m/rho
    31: ("pin_ddf",),                                                 # eps from the updated k
    32: ("pin_ddf",),                                                 # nu_t without sqrt
    33: ("pin_ddf",),                                                 # k floor 1e-4 -> 0
    34: ("pin_ib_normals",),                                          # upper x face points inward
    35: ("pin_ib_normals",),                                          # normalised by the 1-norm
}


@pytest.mark.parametrize("name", sorted(PINS))
def test_pin(name):
    assert oracle.set_mutant(0) == 0
    PINS[name]()


@pytest.mark.parametrize("mutant,name", [(m, p) for m, ps in sorted(MUTANTS.items()) for p in ps])
def test_planted_mistake_is_caught(mutant, name):
    prev = oracle.set_mutant(mutant)
    try:
        with pytest.raises(AssertionError):
            PINS[name]()
    finally:
        oracle.set_mutant(prev)


@pytest.mark.parametrize("mode", [abi.IBN_REFERENCE, abi.IBN_PLANES, abi.IBN_UPLOADED])
def test_ib_normal_modes_over_steps(mode):
    """Which normals each mode leaves after the count == 0 step and after a
    later one: the reference's rule sets them at count == 0 and its part3
    reset zeroes them at later steps (ISPH_Calc.cuh:61,70-74); the plane mode
    keeps them; the uploaded mode keeps the caller's."""
    P, nrm, geom, faces, prm = _face_deck(mode)
    up = nrm.copy()
    up[P["p_type"] == 1000] = (0.0, 0.6, 0.8)      # differs from every plane normal
    o = oracle.Oracle(prm, len(P))
    o.upload(P, None, up)
    plane = None
    for count in range(2):
        assert o.step(count, 5e-5) == 0
        n = np.stack([o.field("NX"), o.field("NY"), o.field("NZ")], 1)
        mk = o.ifield("P_TYPE") >= 1000
        if mode == abi.IBN_UPLOADED:
            assert np.allclose(n[mk], np.array([0.0, 0.6, 0.8], np.float32).astype(np.float64), rtol=0, atol=0)
        elif count == 0:
            plane = n
            assert (np.abs(n[mk]).sum(1) > 0).sum() == mk.sum() - 1
        elif mode == abi.IBN_REFERENCE:
            assert np.all(n == 0.0)
        else:
            assert np.array_equal(n, plane)
