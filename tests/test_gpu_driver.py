"""The C++ host driver (urban-sph_a100_amd/sophia_run, the counterpart of the
reference's time loop ISPH_Calc, ISPH_Calc.cuh:238-577) on the GPU: it reads
a solv.txt and a particle deck with the library's readers, steps through the
C ABI, prints the reference's statistics line, writes VTK frames and dumps
the final part1 records.  The dump must equal the oracle stepped from the
same deck and settings, rounded to the records' floats."""
import os
import subprocess

import numpy as np
import pytest

import abi
import cases
import oracle
import urbansph
from test_io import _deck_text, _read_vtk

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "urban-sph_a100_amd", "sophia_run")

SOLV = """\
solver_type : ISPH
dimension(1/2/3): 3
kernel type : Wendland2
time integration : Predictor_Corrector
pressure-force : YES
viscous-force : Cleary
turbulence-model : Laminar
gravitational-force : NO
noslip_boundary : YES
kappa: 2.0
time-step : 1e-4
start-time : 0.0
end-time : 1.0
filtering : 1
plot-output : 2
plot-variables: 3 p_type velocity pressure
velocity-limit : 1000
"""

# settings the solv.txt has no keyword for (the driver's key value files)
EXTRA = {"ppe_solver": abi.PPE_BICGSTAB, "ppe_tol": 1e-10, "mu_fluid": 1e-3}


def test_driver_steps_a_deck_like_the_oracle(tmp_path):
    assert os.path.exists(DRIVER), "build() makes the driver"
    P, _ = cases.box(nf=8, walls=3)
    codes = [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 26, 27]
    deck = tmp_path / "deck.txt"
    deck.write_text(_deck_text(P, codes))
    solv = tmp_path / "solv.txt"
    solv.write_text(SOLV)
    extra = tmp_path / "params.txt"
    extra.write_text("".join(f"{k} {v!r}\n" for k, v in EXTRA.items()))
    vtk = tmp_path / "vtk"
    vtk.mkdir()
    dump = tmp_path / "out.bin"
    steps = 4
    cmd = [DRIVER, str(deck), "--solv", str(solv), str(extra), "--steps", str(steps), "--vtk", str(vtk),
           "--dump", str(dump)]
    print("[driver] " + " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0, r.stderr
    assert f"steps={steps}" in r.stdout
    assert "\t tu_max=" in r.stdout and "Total number of particles=" in r.stdout
    out = np.fromfile(dump, dtype=abi.PART1)
    # the same deck and settings through the oracle
    Pd = urbansph.read_deck(deck)
    prm, run = urbansph.read_solv(solv, oracle.default_params())
    prm.update(**EXTRA)
    assert run["dt"] == 1e-4 and run["freq_output"] == 2
    o = oracle.Oracle(prm, len(Pd))
    o.upload(Pd)
    for c in range(steps):
        assert o.step(c, run["dt"], c * run["dt"]) == 0
    assert len(out) == len(Pd)
    for name, f in (("x", "X"), ("ux", "UX"), ("uy", "UY"), ("uz", "UZ"), ("pres", "PRES"), ("rho", "RHO")):
        ref = o.field(f).astype(np.float32)
        got = out[name]
        # the FP64 state agrees with the oracle to 1e-10; rounding the two to
        # float may still land on neighbouring floats
        np.testing.assert_allclose(got, ref, rtol=2e-7, atol=1e-7 * float(np.abs(ref).max() + 1e-30), err_msg=name)
    np.testing.assert_array_equal(out["p_type"], o.ifield("P_TYPE"))
    # VTK frames at the plot-output steps, with the requested arrays
    frames = sorted(os.listdir(vtk))
    assert frames == ["fluid_2stp.vtk", "fluid_4stp.vtk"], frames
    head, pts, arrays = _read_vtk(vtk / "fluid_4stp.vtk")
    keep = (out["i_type"] == 1) | (out["i_type"] == 2)
    np.testing.assert_array_equal(pts, np.stack([out["x"], out["y"], out["z"]], 1)[keep])
    pres = dict(arrays)["pressure"][:, 0]
    np.testing.assert_array_equal(pres, out["pres"][keep])
