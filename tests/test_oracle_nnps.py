"""Pins the oracle's neighbour search (functions_NNPS.cuh:534-564,813-955,
ISPH_Calc.cuh:23-52) against an independent numpy restatement: float32
binning and query formulas, numpy's stable argsort, run-length cell bounds
and a brute-force O(n^2) neighbour set under each kernel's window rule."""
import numpy as np
import pytest

import abi
import cases
import oracle

f32 = np.float32


def _domain_np(P, kappa):
    x, y, z = P["x"], P["y"], P["z"]
    lo = [a[:-1].min() if len(a) > 1 else a[0] for a in (x, y, z)]
    hi = [a[:-1].max() if len(a) > 1 else a[0] for a in (x, y, z)]
    dcell = f32(f32(f32(1.1) * f32(kappa)) * f32(P["h_ref"][0])) / f32(2.0)
    N = [int(f32(f32(b) - f32(a)) / dcell) + 1 for a, b in zip(lo, hi)]
    return lo, hi, dcell, N


def _axis(v, mn, mx, N, dcell=None):
    v = v.astype(f32)
    if mx == mn:
        return np.zeros(v.shape, np.int64)
    if dcell is None:
        t = np.floor((v - f32(mn)) / (f32(mx) - f32(mn)) * f32(N))
    else:
        t = np.floor((v - f32(mn)) / f32(dcell))
    t = np.minimum(t, f32(N - 1))
    c = t.astype(np.int64)
    return np.maximum(c, 0)


def _cells(P, lo, hi, N, dcell=None):
    return np.stack([_axis(P[a], lo[k], hi[k], N[k], dcell) for k, a in enumerate("xyz")], 1)


def _run(P, prm, normals=None):
    o = oracle.Oracle(prm, len(P))
    o.upload(P, None, normals)
    o.step(0, 1e-4)
    return o


def _random_cloud(n=3000, seed=3):
    rng = np.random.default_rng(seed)
    P = np.zeros(n, abi.PART1)
    P["i_type"] = 1
    P["p_type"] = rng.choice([1, 1, 1, 0, -1], n)
    P["x"] = rng.uniform(0, 0.2, n)
    P["y"] = rng.uniform(0, 0.15, n)
    P["z"] = rng.uniform(0, 0.1, n)
    # exact duplicates and a last particle far outside (find_minmax skips it)
    P[10] = P[20]
    P["x"][-1] = -0.05
    P["h"] = 0.012
    P["h_ref"] = 0.012
    P["m"] = 1e-3
    P["rho"] = 1000.0
    P["temp"] = 300.0
    return P


@pytest.fixture(scope="module", params=["box", "cloud", "canyon"])
def searched(request):
    prm = oracle.default_params()
    prm.update(ppe_solver=abi.PPE_REF_JACOBI)
    if request.param == "box":
        P, nrm = cases.box(nf=10, walls=3)
    elif request.param == "cloud":
        P, nrm = _random_cloud(), None
    else:
        P, nrm, g = cases.small_canyon(dx=0.02)
        prm = cases.canyon_params(g, prm)
        prm.update(ppe_solver=abi.PPE_REF_JACOBI)
    return P, prm, _run(P, prm, nrm)


def test_domain_and_keys(searched):
    P, prm, o = searched
    lo, hi, dcell, N = _domain_np(P, prm.kappa)
    d = o.domain()
    assert (d.NI, d.NJ, d.NK) == tuple(N)
    assert f32(d.dcell) == dcell
    assert d.num_cells == N[0] * N[1] * N[2] + 1
    c = _cells(P, lo, hi, N)
    key = c[:, 0] + N[0] * c[:, 1] + N[0] * N[1] * c[:, 2]
    perm = np.argsort(key, kind="stable")
    keys, operm, gs, ge = o.nnps()
    np.testing.assert_array_equal(operm, perm)
    np.testing.assert_array_equal(keys, key[perm])
    # cell bounds by run length
    exp_s = np.full(d.num_cells, -1)
    exp_e = np.zeros(d.num_cells, np.int64)
    u, first, counts = np.unique(key[perm], return_index=True, return_counts=True)
    exp_s[u] = first
    exp_e[u] = first + counts
    np.testing.assert_array_equal(gs, exp_s)
    np.testing.assert_array_equal(ge, exp_e)


def test_neighbour_sets_brute_force(searched):
    P, prm, o = searched
    lo, hi, dcell, N = _domain_np(P, prm.kappa)
    keys, perm, _, _ = o.nnps()
    S = P[perm]
    n = len(S)
    pos = np.stack([S["x"], S["y"], S["z"]], 1).astype(np.float64)
    h = S["h"].astype(np.float64)
    bin_c = _cells(S, lo, hi, N)
    qry_c = _cells(S, lo, hi, N, dcell)
    ncell = np.ceil(f32(2.0) / (S["h_ref"].astype(np.float64) / h).astype(f32)).astype(np.int64)
    for kind in (0, 1):
        off, nbr = o.neighbors(kind)
        for i in range(0, n, max(1, n // 400)):
            centre = qry_c[i] if kind == 0 else bin_c[i]
            inwin = np.all(np.abs(bin_c - centre) <= ncell[i], axis=1)
            r = np.sqrt(((pos[i] - pos) ** 2).sum(1))
            if kind == 0:
                ok = inwin & (r + 1e-20 < prm.kappa * np.maximum(h[i], h))
            else:
                ok = inwin & (r > 0) & (r < prm.kappa * h[i])
            expect = np.nonzero(ok)[0]
            got = nbr[off[i]:off[i + 1]]
            np.testing.assert_array_equal(got, expect, err_msg=f"kind {kind} particle {i}")


def test_window_rule_misses_far_cell_neighbours():
    """T1: binning cells are narrower than dcell, so the query window can miss
    a particle that lies within the support radius.  The oracle must keep
    that behaviour (it is the reference algorithm)."""
    P, _ = cases.box(nf=20, walls=3)
    prm = oracle.default_params()
    o = _run(P, prm)
    keys, perm, _, _ = o.nnps()
    S = P[perm]
    pos = np.stack([S["x"], S["y"], S["z"]], 1).astype(np.float64)
    off, nbr = o.neighbors(0)
    missing = 0
    for i in range(0, len(S), 37):
        r = np.sqrt(((pos[i] - pos) ** 2).sum(1))
        full = set(np.nonzero(r + 1e-20 < 2.0 * S["h"][i])[0])
        got = set(nbr[off[i]:off[i + 1]].tolist())
        assert got <= full
        missing += len(full - got)
    assert missing > 0
