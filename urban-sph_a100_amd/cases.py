"""Synthetic SOPHIA cases in the reference's part1 layout.

The reference reads its particles from a tab-separated deck
(function_init.cuh:550-676) that it does not ship, and sets the Case-H inflow
in c_initial_inner_outer_particle_single (functions_NNPS.cuh:196-373).  These
generators build equivalent decks directly:

* ``box``    -- a jittered fluid cube inside fixed walls (lid-free cavity);
* ``canyon`` -- an urban street canyon: ground walls, a free-slip lid, an
  inlet/outlet buffer pair, periodic side buffers, cuboid buildings carried
  by IBM markers (p_type 1000) with outward normals, and a MOST layer;
* ``ensemble`` -- many small canyons of uneven size for batched runs.

All randomness comes from SplitMix64 counters, so a seed gives the same case
on every machine and numpy version.
"""
import math

import numpy as np

import abi

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, n, offset=0):
    """n SplitMix64 outputs of stream ``seed`` starting at counter ``offset``."""
    with np.errstate(over="ignore"):
        k = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed, n, offset=0):
    """Doubles in [0, 1) with 53 random bits."""
    return (splitmix64(seed, n, offset) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def normal(seed, n, sigma=1.0):
    """Box-Muller (cosine branch): one normal per pair of uniforms."""
    u = uniform(seed, 2 * n)
    u0, u1 = u[0::2], u[1::2]
    return sigma * np.sqrt(-2.0 * np.log(1.0 - u0)) * np.cos(2.0 * math.pi * u1)


def _blank(n):
    p = np.zeros(n, dtype=abi.PART1)
    p["i_type"] = 1
    p["temp"] = 300.0
    return p


def _fill_common(p, dx, rho, h_factor=1.5):
    p["rho"] = rho
    p["m"] = rho * dx ** 3
    p["m_ref"] = p["m"]
    p["h"] = h_factor * dx
    p["h_ref"] = p["h"]
    p["vol"] = dx ** 3


def box(nf=32, walls=3, dx=0.01, rho=1000.0, sigma_u=0.01, jitter=0.05,
        seed_pos=1, seed_vel=2):
    """Fluid cube of nf^3 sites on a cubic lattice, jittered by +-jitter*dx
    (seed_pos), velocities N(0, sigma_u) (seed_vel), inside ``walls`` layers of
    fixed wall particles (p_type 0).  x varies fastest.  nf=32, walls=3 gives
    the 54,872-particle box of the plain FP64 configuration."""
    nt = nf + 2 * walls
    k, j, i = np.meshgrid(np.arange(nt), np.arange(nt), np.arange(nt), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    n = nt ** 3
    p = _blank(n)
    fluid = ((i >= walls) & (i < walls + nf) & (j >= walls) & (j < walls + nf)
             & (k >= walls) & (k < walls + nf))
    pos = np.stack([i, j, k], axis=1).astype(np.float64) * dx
    nfl = int(fluid.sum())
    jit = (2.0 * uniform(seed_pos, 3 * nfl) - 1.0) * (jitter * dx)
    pos[fluid] += jit.reshape(nfl, 3)
    p["x"], p["y"], p["z"] = pos[:, 0], pos[:, 1], pos[:, 2]
    p["p_type"] = np.where(fluid, 1, 0)
    vel = normal(seed_vel, 3 * nfl, sigma_u).reshape(nfl, 3)
    for c, name in enumerate(("ux", "uy", "uz")):
        col = np.zeros(n)
        col[fluid] = vel[:, c]
        p[name] = col
    _fill_common(p, dx, rho)
    return p, None


def box_params(P=None, dt=1e-4, tol=1e-8, precision=abi.FP64):
    """Parameters of the plain FP64 box: Wendland-2, h = 1.5 dx, Cleary
    viscosity with nu = 1e-6 (mu = 1e-3 Pa s), laminar, no gravity,
    predictor-corrector, no-slip walls, BiCGSTAB PPE."""
    prm = abi.Params() if P is None else P
    prm.update(kernel_type=abi.WENDLAND2, kappa=2.0, fv_solve=abi.FV_CLEARY,
               turbulence_model=abi.LAMINAR, fg_solve=0, time_type=abi.PRE_COR,
               noslip_bc=1, mu_fluid=1.0e-3, u_limit=1000.0,
               ppe_solver=abi.PPE_BICGSTAB, ppe_tol=tol, precision=precision)
    return prm


def log_law(z, u_star=0.3, z0=0.001, kappa_vk=0.4):
    return (u_star / kappa_vk) * np.log(np.maximum(z, z0) / z0)


def canyon(dx=0.008, Lx=2.0, Ly=1.0, Lz=0.5, buildings=None, bsize=(0.1, 0.1, 0.15),
           walls=3, lid=3, buffers=3, u_star=0.3, z0=0.001, rho=1.225, u_scale=1.0, fill_buildings=True):
    """Urban canyon deck.  Fluid sites at ((i+.5)dx, (j+.5)dx, (k+.5)dx) fill
    [0,Lx)x[0,Ly)x(0,Lz) minus the buildings; ``walls`` ground layers
    (p_type 0) below z=0; ``lid`` free-slip layers (p_type -1) above Lz;
    buildings as IBM marker shells (p_type 1000) on their four sides and roof,
    outward normals returned separately.  The building volume keeps its fluid
    sites at rest (fill_buildings, the immersed-boundary setting: the markers'
    forcing, not a void, carries the wall); fill_buildings=False cuts them
    out, which leaves an unconstrained void that the projection cannot
    handle (the flow next to the faces blows up within a few steps).  Inlet (buffer 1) is x < L2, outlet
    (buffer 2) x >= L3, periodic side buffers (3/4) y < L_left and y > L_right,
    each ``buffers`` layers wide (functions_NNPS.cuh:253-275).  The first fluid
    layer outside the buildings is the MOST layer (MOST_buffer = 1).  Velocity: log law with u*, z0
    (scaled by u_scale) for fluid and lid particles.

    Returns (part1 array, normals (n,3) float32, geometry dict)."""
    if buildings is None:
        buildings = [(0.7 + 0.2 * a, 0.4 + 0.2 * b) for a in range(4) for b in range(2)]
    bx, by, bh = bsize
    nx, ny, nz = int(round(Lx / dx)), int(round(Ly / dx)), int(round(Lz / dx))
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    x = (i.ravel() + 0.5) * dx
    y = (j.ravel() + 0.5) * dx
    z = (k.ravel() + 0.5) * dx
    keep = np.ones(x.shape, dtype=bool)
    solid = np.zeros(x.shape, dtype=bool)
    for (xc, yc) in buildings:
        inside = (np.abs(x - xc) < 0.5 * bx) & (np.abs(y - yc) < 0.5 * by) & (z < bh)
        solid |= inside
    if not fill_buildings:
        keep &= ~solid
    fx, fy, fz = x[keep], y[keep], z[keep]
    fsolid = solid[keep]
    first_layer = (k.ravel()[keep] == 0)
    # ground walls and lid
    kw, jw, iw = np.meshgrid(np.arange(walls), np.arange(ny), np.arange(nx), indexing="ij")
    wx, wy, wz = (iw.ravel() + 0.5) * dx, (jw.ravel() + 0.5) * dx, -(kw.ravel() + 0.5) * dx
    kl, jl, il = np.meshgrid(np.arange(lid), np.arange(ny), np.arange(nx), indexing="ij")
    lx, ly, lz = (il.ravel() + 0.5) * dx, (jl.ravel() + 0.5) * dx, nz * dx + (kl.ravel() + 0.5) * dx
    # markers
    mk, nrm = [], []
    for (xc, yc) in buildings:
        nyf, nzf, nxf = int(math.ceil(by / dx)), int(math.ceil(bh / dx)), int(math.ceil(bx / dx))
        ys = yc - 0.5 * by + (np.arange(nyf) + 0.5) * (by / nyf)
        xs = xc - 0.5 * bx + (np.arange(nxf) + 0.5) * (bx / nxf)
        zs = (np.arange(nzf) + 0.5) * (bh / nzf)
        for sgn in (-1.0, 1.0):
            Z, Y = np.meshgrid(zs, ys, indexing="ij")
            mk.append(np.stack([np.full(Y.size, xc + sgn * 0.5 * bx), Y.ravel(), Z.ravel()], 1))
            nrm.append(np.tile([sgn, 0.0, 0.0], (Y.size, 1)))
            Z, X = np.meshgrid(zs, xs, indexing="ij")
            mk.append(np.stack([X.ravel(), np.full(X.size, yc + sgn * 0.5 * by), Z.ravel()], 1))
            nrm.append(np.tile([0.0, sgn, 0.0], (X.size, 1)))
        Y, X = np.meshgrid(ys, xs, indexing="ij")
        mk.append(np.stack([X.ravel(), Y.ravel(), np.full(X.size, bh)], 1))
        nrm.append(np.tile([0.0, 0.0, 1.0], (X.size, 1)))
    mk = np.concatenate(mk) if mk else np.zeros((0, 3))
    nrm = np.concatenate(nrm) if nrm else np.zeros((0, 3))

    nfl, nw, nl, nm = fx.size, wx.size, lx.size, mk.shape[0]
    n = nfl + nw + nl + nm
    p = _blank(n)
    p["x"] = np.concatenate([fx, wx, lx, mk[:, 0]])
    p["y"] = np.concatenate([fy, wy, ly, mk[:, 1]])
    p["z"] = np.concatenate([fz, wz, lz, mk[:, 2]])
    p["p_type"] = np.concatenate([np.ones(nfl, np.int32), np.zeros(nw, np.int32),
                                  -np.ones(nl, np.int32), np.full(nm, 1000, np.int32)])
    _fill_common(p, dx, rho)
    # velocity: log law for fluid and lid
    u = np.zeros(n)
    u[:nfl] = np.where(fsolid, 0.0, u_scale * log_law(fz, u_star, z0))
    u[nfl + nw:nfl + nw + nl] = u_scale * log_law(lz, u_star, z0)
    p["ux"] = u
    # buffers (open_boundary > 0 branch, functions_NNPS.cuh:251-276)
    L2 = buffers * dx
    L3 = (nx - buffers + 0.5) * dx
    L_left = buffers * dx
    L_right = (ny - buffers) * dx
    px, py = p["x"].astype(np.float64), p["y"].astype(np.float64)
    pt = p["p_type"]
    bt = np.zeros(n, np.int32)
    it = np.ones(n, np.int32)
    cand = (pt == 1) | (pt == -1)
    left = cand & (py < L_left)
    right = cand & ~left & (py > L_right)
    outl = cand & ~left & ~right & (px >= L3 - 1e-6)
    inl = cand & ~left & ~right & ~outl & (px < L2 - 1e-6)
    bt[left], bt[right], bt[outl], bt[inl] = 3, 4, 2, 1
    it[left | right | outl | inl] = 2
    p["buffer_type"] = bt
    p["i_type"] = it
    most = np.zeros(n, np.int32)
    most[:nfl][first_layer & (bt[:nfl] < 3) & ~fsolid] = 1   # open ground only, not under a building
    p["MOST_buffer"] = most
    normals = np.zeros((n, 3), np.float32)
    normals[nfl + nw + nl:] = nrm
    geom = dict(L2=L2, L3=L3, L_left=L_left, L_right=L_right, outlet_copy_dx=dx,
                n_fluid=nfl, n_wall=nw, n_lid=nl, n_marker=nm, dx=dx,
                buildings=[(xc, yc, bx, by, bh) for (xc, yc) in buildings])
    return p, normals, geom


def canyon_params(geom, P=None, dt=5e-5, tol=1e-6, precision=abi.FP64, Cs=0.12,
                  L=50.0, z0=0.001):
    """Street-canyon parameters: Smagorinsky LES (Cs), MOST with prescribed
    Obukhov length L, log-law roughness z0, air properties, BiCGSTAB PPE."""
    prm = abi.Params() if P is None else P
    prm.update(kernel_type=abi.WENDLAND2, kappa=2.0, fv_solve=abi.FV_CLEARY,
               turbulence_model=abi.SSM, Cs_sgs=Cs, fg_solve=0,
               time_type=abi.PRE_COR, noslip_bc=1, u_limit=50.0,
               most_mode=abi.MOST_PRESCRIBED_L, most_L=L, most_z0=z0,
               most_z0h=0.1 * z0, L2=geom["L2"], L3=geom["L3"],
               L_left=geom["L_left"], L_right=geom["L_right"],
               outlet_copy_dx=geom["outlet_copy_dx"],
               ppe_solver=abi.PPE_BICGSTAB, ppe_tol=tol, precision=precision,
               ib_normals=abi.IBN_UPLOADED)   # outward normals of every building, from the generator
    return prm


def small_canyon(dx=0.01, **kw):
    """A ~20k-particle canyon with one building for parity tests."""
    args = dict(dx=dx, Lx=0.40, Ly=0.24, Lz=0.16, buildings=[(0.2, 0.12)],
                bsize=(0.05, 0.05, 0.08))
    args.update(kw)
    return canyon(**args)


def face_marker_canyon(dx=0.02):
    """small_canyon with markers added on the building's vertical edges, roof
    edges and roof corners and one marker off every face plane (at mid-height
    inside the building), for the face-plane rule of the marker normals.
    Returns (part1, generator normals, geometry, faces) with faces the
    ib_face_* parameters of the building (no marker lies on its z = 0 plane)."""
    P, nrm, geom = small_canyon(dx=dx)
    xc, yc, bx, by, bh = geom["buildings"][0]
    xlo, xhi, ylo, yhi = xc - 0.5 * bx, xc + 0.5 * bx, yc - 0.5 * by, yc + 0.5 * by
    extra = []
    for zz in (0.25 * bh, 0.5 * bh):              # vertical edges
        for xx in (xlo, xhi):
            for yy in (ylo, yhi):
                extra.append((xx, yy, zz))
    for xx in (xlo, xhi):                         # roof edges along y, and corners
        for yy in (yc, ylo, yhi):
            extra.append((xx, yy, bh))
    for yy in (ylo, yhi):                         # roof edges along x
        extra.append((xc, yy, bh))
    extra.append((xc, yc, 0.5 * bh))              # on no face plane
    extra = np.array(extra)
    add = np.zeros(len(extra), P.dtype)
    proto = P[np.flatnonzero(P["p_type"] == 1000)[0]]
    for name in P.dtype.names:
        add[name] = proto[name]
    add["x"], add["y"], add["z"] = extra[:, 0], extra[:, 1], extra[:, 2]
    P2 = np.concatenate([P, add])
    nrm2 = np.concatenate([nrm, np.zeros((len(extra), 3), np.float32)])
    faces = dict(ib_face_xlo=xlo, ib_face_xhi=xhi, ib_face_ylo=ylo, ib_face_yhi=yhi,
                 ib_face_zlo=0.0, ib_face_zhi=bh)
    return P2, nrm2, geom, faces


def apr_targets(x, y, z, p_type, i_type, buildings, r1, r2, most=None):
    """Refinement target levels of the multi-resolution canyon: 2 for active
    fluid outside the buildings within r2 of a building wall or roof, 1
    within r1, 0 elsewhere.  ``buildings`` as in the canyon geometry
    (xc, yc, bx, by, bh).  With ``most`` (the MOST_buffer flags) the MOST
    wall-model layer stays at its resolution: its children inherit the flag,
    and two levels down the lowest ones sit at z = z0, where the log law's
    log(z / z0) vanishes."""
    x, y, z = (np.asarray(v, np.float64) for v in (x, y, z))
    d = np.full(x.shape, np.inf)
    inside = np.zeros(x.shape, bool)
    for (xc, yc, bx, by, bh) in buildings:
        ex = np.maximum(np.abs(x - xc) - 0.5 * bx, 0.0)
        ey = np.maximum(np.abs(y - yc) - 0.5 * by, 0.0)
        ez = np.maximum(z - bh, 0.0)
        d = np.minimum(d, np.sqrt(ex * ex + ey * ey + ez * ez))
        inside |= (np.abs(x - xc) < 0.5 * bx) & (np.abs(y - yc) < 0.5 * by) & (z < bh)
    active = (np.asarray(p_type) == 1) & (np.asarray(i_type) == 1) & ~inside
    if most is not None:
        active &= np.asarray(most) == 0
    t = np.zeros(x.shape)
    t[active & (d < r1)] = 1.0
    t[active & (d < r2)] = 2.0
    return t


def sized_canyon(target, u_scale=1.0, bh=0.15, dx=0.008):
    """A one-building canyon whose domain length is chosen so that it holds
    about ``target`` particles at spacing dx (cross-section 0.24 m wide,
    0.06 m of air above the roof)."""
    Ly, Lz = 0.24, max(0.16, round((bh + 0.06) / dx) * dx)
    per_x = (Ly / dx) * (Lz / dx + 6)
    Lx = max(0.2, round(target / per_x) * dx)
    return canyon(dx=dx, Lx=Lx, Ly=Ly, Lz=Lz, buildings=[(0.5 * Lx, 0.5 * Ly)],
                  bsize=(0.05, 0.05, bh), u_scale=u_scale / 4.0)


def ensemble(n_cases=64, lo=2e4, hi=3e5, seed=7, dx=0.008):
    """Independent canyon cases of uneven size: particle counts log-uniform in
    [lo, hi], inflow scale U(0.5, 2.0) and building height U(0.05, 0.2), all
    from one seed.  The domain length is chosen to hit the target count at
    the given dx.  Returns a list of (part1, normals, geom, info)."""
    r = uniform(seed, 3 * n_cases).reshape(n_cases, 3)
    out = []
    for c in range(n_cases):
        target = math.exp(math.log(lo) + r[c, 0] * (math.log(hi) - math.log(lo)))
        uscale = 0.5 + 1.5 * r[c, 1]
        bh = 0.05 + 0.15 * r[c, 2]
        p, nrm, g = sized_canyon(target, uscale, bh, dx)
        out.append((p, nrm, g, dict(target=target, u_scale=uscale, bh=bh)))
    return out


STABILITY_L = (-50.0, 50.0, 1.0e6)   # unstable, stable, neutral (|L| -> inf)


def serving_requests(n_req=24, rate=4.0, lo=1e4, hi=2e5, steps_lo=10, steps_hi=40, seed=9):
    """A Poisson stream of canyon scenario requests: exponential
    inter-arrival times at ``rate`` per second, particle counts log-uniform in
    [lo, hi], step counts uniform in [steps_lo, steps_hi], inflow scale
    U(0.5, 2.0) and Obukhov length drawn from STABILITY_L, all from one seed.
    Returns a list of dicts (t_arrival, target, steps, u_scale, L, bh)."""
    r = uniform(seed, 6 * n_req).reshape(n_req, 6)
    t = 0.0
    out = []
    for k in range(n_req):
        t += -math.log(1.0 - r[k, 0]) / rate
        target = math.exp(math.log(lo) + r[k, 1] * (math.log(hi) - math.log(lo)))
        steps = steps_lo + int(r[k, 2] * (steps_hi - steps_lo + 1))
        out.append(dict(t_arrival=t, target=target, steps=min(steps, steps_hi), u_scale=0.5 + 1.5 * r[k, 3],
                        L=STABILITY_L[min(2, int(3 * r[k, 4]))], bh=0.05 + 0.15 * r[k, 5]))
    return out
