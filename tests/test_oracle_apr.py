"""Adaptive particle refinement in the oracle (functions_APR.cuh:270-500,
functions_APR_etc.cuh:354-375), pinned by what the splitting rule fixes:
which cells split in a pass, conservation of mass and volume, the children's
geometry, and convergence of repeated passes to the target levels."""
import numpy as np
import pytest

import abi
import cases
import oracle


def _box_searched(nf=8):
    P, _ = cases.box(nf=nf, walls=3)
    prm = cases.box_params(oracle.default_params())
    o = oracle.Oracle(prm, 40 * len(P))
    o.upload(P)
    assert o.step(0, 1e-4) == 0
    return o


def _state(o):
    s = {k: o.field(k) for k in ("X", "Y", "Z", "M", "VOL", "H", "RHO")}
    s.update(IT=o.ifield("I_TYPE"), PT=o.ifield("P_TYPE"))
    # the box deck gives every particle the same mass, which is also its
    # reference mass m_ref (the refinement level is log8(m_ref / m))
    s["M_REF"] = np.full(len(s["M"]), M_REF)
    return s


M_REF = None


@pytest.fixture(autouse=True)
def _mref():
    global M_REF
    P, _ = cases.box(nf=8, walls=3)
    assert np.all(P["m"] == P["m"][0]) and np.all(P["m_ref"] == P["m"])
    M_REF = float(P["m"][0])


def test_one_pass_splits_the_first_eligible_particle_of_each_cell():
    o = _box_searched()
    s0 = _state(o)
    keys, _, _, _ = o.nnps()
    fluid = (s0["PT"] == 1) & (s0["IT"] == 1)
    xc = 0.5 * (s0["X"][fluid].min() + s0["X"][fluid].max())
    target = np.where(fluid & (s0["X"] < xc), 1.0, 0.0)
    # expected splits: per occupied cell, the first particle in cell order
    # with target - level >= 0.75 (levels are 0 before refinement)
    level = np.log(s0["M_REF"] / s0["M"]) / np.log(8.0)
    expect = []
    for cell in np.unique(keys):
        members = np.nonzero(keys == cell)[0]
        hit = [j for j in members if target[j] - level[j] >= 0.75]
        if hit:
            expect.append(hit[0])
    n0 = o.n
    n1 = o.apr_refine(target, max_passes=1)
    assert n1 == n0 + 8 * len(expect)
    s1 = _state(o)
    # parents are dummies now; dummies sort to the end of the cell order
    dummies = np.nonzero(s1["IT"] == 3)[0]
    assert len(dummies) == len(expect)
    live = s1["IT"] != 3
    # mass and volume are conserved exactly in the live particles
    assert np.isclose(np.sum(s1["M"][live]), np.sum(s0["M"]), rtol=1e-13, atol=0)
    assert np.isclose(np.sum(s1["VOL"][live]), np.sum(s0["VOL"]), rtol=1e-13, atol=0)
    # every parent has 8 children: m/8, h/2, at +-cbrt(V)/4 around it, the
    # positions rounded to float as the reference's float records hold them
    # (functions_APR.cuh:544-590)
    for d in dummies:
        px, py, pz = s1["X"][d], s1["Y"][d], s1["Z"][d]
        a = 0.25 * np.cbrt(s1["M"][d] / s1["RHO"][d])
        dist = np.abs(s1["X"] - px) + np.abs(s1["Y"] - py) + np.abs(s1["Z"] - pz)
        kids = np.nonzero(live & (dist < 3.01 * a) & np.isclose(s1["M"], s1["M"][d] / 8, rtol=1e-15))[0]
        assert len(kids) == 8
        offs = np.stack([s1["X"][kids] - px, s1["Y"][kids] - py, s1["Z"][kids] - pz], 1)
        assert len({tuple(np.sign(v)) for v in offs}) == 8       # the 8 octants
        for k, (sx, sy, sz) in zip(kids, np.sign(offs)):
            assert s1["X"][k] == float(np.float32(px + sx * a))
            assert s1["Y"][k] == float(np.float32(py + sy * a))
            assert s1["Z"][k] == float(np.float32(pz + sz * a))
        ulp = np.spacing(np.float32(max(abs(px), abs(py), abs(pz)) + a))
        np.testing.assert_allclose(np.abs(offs), a, atol=float(ulp))
        np.testing.assert_array_equal(s1["H"][kids], 0.5 * s1["H"][d])


def test_passes_converge_to_the_target_levels():
    o = _box_searched()
    s0 = _state(o)
    fluid = (s0["PT"] == 1) & (s0["IT"] == 1)
    zc = 0.5 * (s0["Z"][fluid].min() + s0["Z"][fluid].max())
    target = np.where(fluid & (s0["Z"] > zc), 2.0, 0.0)
    o.apr_refine(target, max_passes=200)
    s1 = _state(o)
    live = s1["IT"] != 3
    level = np.log(s1["M_REF"] / s1["M"]) / np.log(8.0)
    # nothing left to split: fluid above the plane at level 2, below it
    # untouched
    assert o.apr_refine(None, max_passes=1) == o.n
    assert np.all(np.round(level[live]) <= 2)
    fl = live & (s1["PT"] == 1) & (s1["IT"] == 1)
    top = fl & (s1["Z"] > zc + 0.01)
    assert np.all(np.isclose(level[top], 2.0, atol=1e-12))
    low = fl & (s1["Z"] < zc - 0.01)
    assert np.all(np.isclose(level[low], 0.0, atol=1e-12))
    assert np.isclose(np.sum(s1["M"][live]), np.sum(s0["M"]), rtol=1e-12, atol=0)
    # the refined case still steps
    assert o.step(1, 1e-4) == 0
    assert np.all(np.isfinite(o.field("PRES")[live]))


def test_capacity_is_checked():
    P, _ = cases.box(nf=6, walls=3)
    prm = cases.box_params(oracle.default_params())
    o = oracle.Oracle(prm, len(P) + 8)
    o.upload(P)
    assert o.step(0, 1e-4) == 0
    with pytest.raises(RuntimeError):
        o.apr_refine(np.ones(o.n), max_passes=1)
