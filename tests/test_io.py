"""File formats of the library (host C++, no GPU): the solv.txt keyword
reader (function_init.cuh:166-463), the particle deck reader
(function_init.cuh:550-676), the legacy VTK writer
(functions_OUTPUT.cuh:121-616), the property-table reader
(function_init.cuh:724-970) and its lookup (functions_PROP.cuh:8-40).
Expected values follow from the formats themselves: each keyword's
filler-token count, the deck's column codes, the VTK layout parsed back
independently here, and numpy's piecewise-linear interpolation."""
import numpy as np
import pytest

import abi
import cases
import urbansph

# Each keyword with the number of filler tokens the reference reads before
# its value (function_init.cuh:193-458), written out as a solv.txt would be.
SOLV = """\
solver_type : ISPH
dimension(1/2/3): 3
property_table(YES/NO): YES
atmospheric-stability : Stable
terrain-type : City
kernel type : Wendland6
time integration : Predictor_Corrector
structure : NO
open boundary : PERIODIC
pressure-force : YES
viscous-force : Monaghan
turbulence-model : DDF
gravitational-force : YES
Conduction(YES/NO): YES
Boussinesq-natural-convection : NO
Concentration-diffusion(YES/NO): YES
kernel-gradient-correction : NO
switch_ptype : YES
noslip_boundary : YES
APR_solve : Cell
Block_configuration NBI 4 NBJ 5 NBK 6
Variable : NO
Buffersize: 2.5
reference-pressure : 101325
kappa: 2.2
minimum-iteration: 3
maximum-iteration: 77
pressure-convergence : 0.01
density-convergence : 0.02
pressure-relaxation : 0.5
time-step : 2.5e-5
start-time : 0.5
end-time : 12.5
neighbor search cell : 5X5
timestep update : YES
structure-timestep : 1e-6
filtering : 7
temperature-filtering : 9
plot-output : 40
APR_frequency: 11
h-reset : 13
BAPR_frequency: 17
plot-variables: 4 p_type velocity pressure rho_ref
velocity-limit : 30
"""


def test_solv_reads_every_keyword(tmp_path):
    f = tmp_path / "solv.txt"
    f.write_text(SOLV)
    base = urbansph.default_params()
    base.update(kgc_solve=1, u_limit=1.0)
    prm, run = urbansph.read_solv(f, base)
    p = prm.as_dict()
    assert p["dim"] == 3
    assert p["kernel_type"] == abi.WENDLAND6
    assert p["time_type"] == abi.PRE_COR
    assert p["fp_solve"] == 1 and p["fg_solve"] == 1 and p["con_solve"] == 1 and p["concn_solve"] == 1
    assert p["fv_solve"] == abi.FV_MONAGHAN
    assert p["turbulence_model"] == abi.DDF
    assert p["boussinesq_solve"] == 0 and p["kgc_solve"] == 0
    assert p["noslip_bc"] == 1
    assert p["apr_h_change"] == 0
    assert p["kappa"] == 2.2
    assert p["freq_filt"] == 7
    assert p["u_limit"] == 30.0
    assert run["solver_type"] == 1 and run["prop_table"] == 1
    assert run["atmos_stab"] == 2 and run["terrain_type"] == 3
    assert run["structure"] == 0 and run["open_boundary"] == 2 and run["switch_ptype"] == 1
    assert run["apr_solve"] == 1 and run["block_config"] == [4, 5, 6]
    assert run["buffersize"] == 2.5 and run["p_ref"] == 101325.0
    assert run["min_iteration"] == 3 and run["max_iteration"] == 77
    assert (run["dp_th"], run["drho_th"], run["p_relaxation"]) == (0.01, 0.02, 0.5)
    assert (run["dt"], run["time"], run["time_end"]) == (2.5e-5, 0.5, 12.5)
    assert run["nb_cell_type"] == 1 and run["timestep_update"] == 1
    assert run["dt_structure"] == 1e-6
    assert (run["freq_temp"], run["freq_output"], run["freq_apr"]) == (9, 40, 11)
    assert (run["freq_h_reset"], run["freq_bapr"]) == (13, 17)
    assert run["plot"] == ["p_type", "velocity", "pressure", "rho_ref"]
    assert run["unknown"] == []
    # members no keyword names keep their values
    assert p["ppe_tol"] == base.ppe_tol and p["mu_fluid"] == base.mu_fluid


def test_solv_clears_the_solve_switches_and_keeps_unknown_values(tmp_path):
    f = tmp_path / "solv.txt"
    f.write_text("kernel type : Cubic\nviscous-force : Violeau\nVariable : Smooth\ntime-step : 1e-3\n")
    base = urbansph.default_params()
    base.update(fp_solve=1, fv_solve=1, fg_solve=1, kernel_type=abi.QUARTIC)
    prm, run = urbansph.read_solv(f, base)
    # cleared before reading (function_init.cuh:168-181), then set
    assert prm.fp_solve == 0 and prm.fg_solve == 0
    # reference codes without an enum here are stored as coded: 3 = Violeau,
    # 2 = Smooth (sph_create rejects what the step does not implement)
    assert prm.fv_solve == 3 and prm.apr_h_change == 2
    # an unknown option value leaves the member as it was
    assert prm.kernel_type == abi.QUARTIC and run["unknown"] == ["kernel"]
    assert run["dt"] == 1e-3 and run["solver_type"] == 0


def test_solv_missing_file_is_an_error(tmp_path):
    with pytest.raises(urbansph.SPHError) as e:
        urbansph.read_solv(tmp_path / "none.txt")
    assert e.value.rc == abi.E_ARG


def _deck_text(P, codes):
    cols = {1: "x", 2: "y", 3: "z", 4: "ux", 5: "uy", 6: "uz", 7: "m", 8: "p_type", 9: "h", 10: "temp",
            11: "pres", 12: "rho", 15: "concn", 26: "h_ref", 27: "m_ref"}
    lines = ["\t".join(str(c) for c in codes)]
    for r in P:
        vals = []
        for c in codes:
            if c in cols:
                v = r[cols[c]]
                vals.append(str(int(v)) if c == 8 else repr(float(v)))
            else:
                vals.append("-7.5")       # a skipped column
        lines.append("\t".join(vals))
    return "\n".join(lines) + "\n"


def test_deck_round_trip(tmp_path):
    P, _ = cases.box(nf=4, walls=2)
    codes = [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 21, 26, 27]
    f = tmp_path / "deck.txt"
    f.write_text(_deck_text(P, codes))
    Q = urbansph.read_deck(f)
    assert len(Q) == len(P)
    for k in ("x", "y", "z", "ux", "uy", "uz", "m", "p_type", "h", "temp", "pres", "rho", "concn", "h_ref",
              "m_ref"):
        np.testing.assert_array_equal(Q[k], P[k], err_msg=k)
    assert np.all(Q["i_type"] == 1)
    # columns the deck does not carry are zero
    for k in ("eli", "vol", "buffer_type", "k_turb", "source"):
        assert np.all(Q[k] == 0), k


def test_deck_header_quirks_and_partial_record(tmp_path):
    f = tmp_path / "deck.txt"
    # a trailing tab adds a column of code 0, read and skipped; the last,
    # incomplete record is dropped
    f.write_text("1\t2\t3\t8\t\n0.5 1.5 2.5 1 99\n3.0 4.0 5.0 0 99\n6.0 7.0\n")
    Q = urbansph.read_deck(f)
    assert len(Q) == 2
    assert list(Q["x"]) == [0.5, 3.0] and list(Q["z"]) == [2.5, 5.0] and list(Q["p_type"]) == [1, 0]
    g = tmp_path / "deck2.txt"
    g.write_text("1\t2\t3\t8\n0.5 1.5 2.5 1 99 3.0 4.0 5.0\n")
    Q = urbansph.read_deck(g)
    # without the extra column the values shift: 4 per record
    assert len(Q) == 2 and Q["p_type"][0] == 1 and Q["x"][1] == 99.0 and Q["p_type"][1] == 5


def test_deck_capacity(tmp_path):
    import ctypes as C
    f = tmp_path / "deck.txt"
    f.write_text("1\t2\t3\n1 2 3\n4 5 6\n7 8 9\n")
    L = urbansph.lib()
    n = C.c_int()
    assert L.sph_read_deck(str(f).encode(), None, 0, C.byref(n)) == 0 and n.value == 3
    P = np.zeros(2, abi.PART1)
    assert L.sph_read_deck(str(f).encode(), abi.ptr(P), 2, C.byref(n)) == abi.E_NOMEM
    assert L.sph_read_deck(str(tmp_path / "none").encode(), None, 0, C.byref(n)) == abi.E_ARG


def _read_vtk(path):
    """Independent parser of the legacy binary VTK the writer emits."""
    b = path.read_bytes()
    pos = 0

    def line():
        nonlocal pos
        e = b.index(b"\n", pos)
        s = b[pos:e].decode()
        pos = e + 1
        return s

    head = [line() for _ in range(4)]
    tag, np_, ty = line().split("\t")
    assert tag == "POINTS" and ty == "float"
    npts = int(np_)
    pts = np.frombuffer(b, ">f4", 3 * npts, pos).reshape(npts, 3)
    pos += 12 * npts
    assert line() == f"POINT_DATA\t{npts}"
    tag, nf = line().split("\t")
    assert tag == "FIELD FieldData"
    arrays = []
    for _ in range(int(nf)):
        name, nc, cnt, ty = line().split("\t")
        nc, cnt = int(nc), int(cnt)
        assert cnt == npts
        dt = ">i4" if ty == "int" else ">f4"
        a = np.frombuffer(b, dt, nc * cnt, pos).reshape(cnt, nc)
        pos += 4 * nc * cnt
        arrays.append((name, a))
    assert pos == len(b)
    return head, pts, arrays


def test_vtk_layout(tmp_path):
    rng = np.random.default_rng(3)
    n = 50
    P1 = np.zeros(n, abi.PART1)
    P2 = np.zeros(n, abi.PART2)
    P3 = np.zeros(n, abi.PART3)
    for k in ("x", "y", "z", "ux", "uy", "uz", "pres", "rho", "h", "vis_t", "fmx", "fmy"):
        P1[k] = rng.standard_normal(n).astype(np.float32)
    P1["i_type"] = rng.choice([0, 1, 2, 3], n)
    P1["p_type"] = rng.integers(-2, 3, n)
    P1["ncell"] = rng.integers(1, 5, n)
    P2["rho_ref"] = rng.standard_normal(n).astype(np.float32)
    for k in ("nx", "ny", "nz"):
        P3[k] = rng.standard_normal(n).astype(np.float32)
    labels = ["p_type", "velocity", "pressure", "smoothing", "MOST_force", "rho_ref", "ppe1", "normal", "ncell"]
    f = tmp_path / "fluid_0stp.vtk"
    urbansph.write_vtk(f, P1, labels, P2, P3)
    head, pts, arrays = _read_vtk(f)
    assert head == ["# vtk DataFile Version 3.0", "Print out results in vtk format", "BINARY", "DATASET POLYDATA"]
    keep = (P1["i_type"] == 1) | (P1["i_type"] == 2)
    np.testing.assert_array_equal(pts, np.stack([P1["x"], P1["y"], P1["z"]], 1)[keep])
    names = [a[0] for a in arrays]
    # rho_ref matches the 3-character "rho" entry too; ppe1 matches nothing
    assert names == ["p_type", "velocity", "pressure", "h", "fm", "rho", "rho_ref", "normal", "ncell"]
    A = dict(arrays)
    np.testing.assert_array_equal(A["p_type"][:, 0], P1["p_type"][keep])
    np.testing.assert_array_equal(A["velocity"], np.stack([P1["ux"], P1["uy"], P1["uz"]], 1)[keep])
    np.testing.assert_array_equal(A["pressure"][:, 0], P1["pres"][keep])
    np.testing.assert_array_equal(A["h"][:, 0], P1["h"][keep])
    np.testing.assert_array_equal(A["fm"], np.stack([P1["fmx"], P1["fmy"]], 1)[keep])
    np.testing.assert_array_equal(A["rho"][:, 0], P1["rho"][keep])
    np.testing.assert_array_equal(A["rho_ref"][:, 0], P2["rho_ref"][keep])
    np.testing.assert_array_equal(A["normal"], np.stack([P3["nx"], P3["ny"], P3["nz"]], 1)[keep])
    np.testing.assert_array_equal(A["ncell"][:, 0], P1["ncell"][keep])


def test_vtk_needs_the_records_its_labels_read(tmp_path):
    P1 = np.zeros(4, abi.PART1)
    P1["i_type"] = 1
    with pytest.raises(urbansph.SPHError):
        urbansph.write_vtk(tmp_path / "a.vtk", P1, ["normal"])
    with pytest.raises(urbansph.SPHError):
        urbansph.write_vtk(tmp_path / "b.vtk", P1, ["rho_ref"], P3=np.zeros(4, abi.PART3))
    urbansph.write_vtk(tmp_path / "c.vtk", P1, [])
    _, pts, arrays = _read_vtk(tmp_path / "c.vtk")
    assert len(pts) == 4 and arrays == []


TABLES = """\
# material tables
#p_type 0
//T
280,1537,2300,2450,2650
//h
-554932,215789,757894,905263,1268421
//mu
1e-3,2e-3,3e-3,4e-3,5e-3
#end
#p_type 1
//T
300,400,500
//k
0.5,0.6
#end
"""


def test_tables_blocks_and_lengths(tmp_path):
    f = tmp_path / "tables.txt"
    f.write_text(TABLES)
    T = urbansph.read_tables(f)
    assert len(T) == 2
    np.testing.assert_array_equal(T[0]["T"], np.float32([280, 1537, 2300, 2450, 2650]))
    np.testing.assert_array_equal(T[0]["h"], np.float32([-554932, 215789, 757894, 905263, 1268421]))
    np.testing.assert_array_equal(T[0]["mu"], np.float32([1e-3, 2e-3, 3e-3, 4e-3, 5e-3]))
    # a block is as long as the last array read in it (here k, 2 values)
    np.testing.assert_array_equal(T[1]["T"], np.float32([300, 400]))
    np.testing.assert_array_equal(T[1]["k"], np.float32([0.5, 0.6]))
    t = abi.PropTables()
    assert urbansph.lib().sph_read_tables(str(f).encode(), t) == 0
    assert (t.ntables, t.ndata) == (2, 7)
    assert list(t.index[:2]) == [0, 5] and list(t.size[:2]) == [5, 2]
    # the third T value of block 1 sits past its length, where block 2 would start
    assert t.T[7] == 500


def test_tables_errors(tmp_path):
    f = tmp_path / "t.txt"
    f.write_text("#p_type 0\n//T\n1,2\n")        # no #end
    with pytest.raises(urbansph.SPHError):
        urbansph.read_tables(f)
    f.write_text("#p_type 0\n//T\n" + ",".join(["1"] * 600) + "\n#end\n")
    with pytest.raises(urbansph.SPHError) as e:
        urbansph.read_tables(f)
    assert e.value.rc == abi.E_NOMEM


def test_table_interp():
    x = np.float32([280, 1537, 2300, 2450, 2650])
    y = np.float32([-554932, 215789, 757894, 905263, 1268421])
    # inside: the piecewise-linear interpolant
    for v in (300.0, 1000.0, 1537.0, 2299.5, 2450.0, 2600.0, 2650.0):
        assert np.isclose(urbansph.table_interp(x, y, v), np.interp(v, x, y), rtol=1e-6, atol=0.5), v
    # the first knot takes the left branch: exactly y[0]
    assert urbansph.table_interp(x, y, 280.0) == y[0]
    # outside: the end segments extended
    lo = y[0] + (y[1] - y[0]) / (x[1] - x[0]) * (100.0 - x[0])
    hi = y[4] + (y[4] - y[3]) / (x[4] - x[3]) * (3000.0 - x[4])
    assert np.isclose(urbansph.table_interp(x, y, 100.0), lo, rtol=1e-6)
    assert np.isclose(urbansph.table_interp(x, y, 3000.0), hi, rtol=1e-6)
    assert np.isnan(urbansph.table_interp(x, y, float("nan")))
    assert np.isnan(urbansph.table_interp(x[:1], y[:1], 1.0))
