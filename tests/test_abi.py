"""CPU checks of the C ABI: the library loads, exports every function that
include/urbansph.h declares, record layouts match the reference's structs and
the default parameters equal the reference values restated by the oracle."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import abi
import oracle
import urbansph

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "urbansph.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sph_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = urbansph.lib()
    names = _declared()
    assert len(names) >= 15
    for name in names:
        assert hasattr(L, name), name
    assert set(urbansph.EXPORTED) <= set(names)


def test_record_layouts():
    assert abi.PART1.itemsize == 152
    assert abi.PART2.itemsize == 52
    assert abi.PART3.itemsize == 164
    assert abi.PART1.fields["x"][1] == 16
    assert abi.PART1.fields["ncell"][1] == 144
    assert abi.PART3.fields["inv_cm_xx"][1] == 4 * 27


def test_default_params_match_reference_restatement():
    a = urbansph.default_params().as_dict()
    b = oracle.default_params().as_dict()
    assert a == b
    assert a["Cs_sgs"] == 0.15 and a["mu_fluid"] == 1.8e-5 and a["L3"] == 13.304


def test_params_struct_layout_matches_c(tmp_path):
    """The ctypes mirror must agree with the C compiler's layout of every
    sph_params member and of the records."""
    members = [n for n, _ in abi.Params._fields_]
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "urbansph.h"', 'int main(void){',
             'printf("%zu %zu %zu %zu\\n", sizeof(sph_params), sizeof(sph_part1), sizeof(sph_part2), sizeof(sph_part3));']
    for m in members:
        lines.append(f'printf("%zu\\n", offsetof(sph_params, {m}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    import subprocess
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)])
    out = subprocess.check_output([str(exe)]).decode().split()
    assert int(out[0]) == C.sizeof(abi.Params)
    assert [int(v) for v in out[1:4]] == [152, 52, 164]
    for m, off in zip(members, out[4:]):
        assert getattr(abi.Params, m).offset == int(off), m


@pytest.mark.parametrize("struct,cname", [("RunConfig", "sph_run_config"), ("PropTables", "sph_prop_tables"),
                                          ("ServerInfo", "sph_server_info"), ("RequestInfo", "sph_request_info"),
                                          ("PPEReport", "sph_ppe_report")])
def test_aux_struct_layouts_match_c(tmp_path, struct, cname):
    S = getattr(abi, struct)
    members = [n for n, _ in S._fields_]
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "urbansph.h"', 'int main(void){',
             f'printf("%zu\\n", sizeof({cname}));']
    for m in members:
        lines.append(f'printf("%zu\\n", offsetof({cname}, {m}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    import subprocess
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)])
    out = [int(v) for v in subprocess.check_output([str(exe)]).decode().split()]
    assert out[0] == C.sizeof(S)
    for m, off in zip(members, out[1:]):
        assert getattr(S, m).offset == off, m


def test_enum_values_match_c(tmp_path):
    """The Python constants of the solver options and report codes equal the
    header's enumerators."""
    pairs = [("PRECOND_JACOBI", "SPH_PRECOND_JACOBI"), ("PRECOND_MULTIGRID", "SPH_PRECOND_MULTIGRID"),
             ("ORDER_CELL", "SPH_ORDER_CELL"), ("ORDER_LATTICE", "SPH_ORDER_LATTICE"),
             ("FP64", "SPH_FP64"), ("MIXED", "SPH_MIXED"),
             ("MG_OK", "SPH_MG_OK"), ("MG_UNSUITED", "SPH_MG_UNSUITED"), ("MG_TERMS", "SPH_MG_TERMS"),
             ("MG_COLUMNS", "SPH_MG_COLUMNS"), ("MG_SPAN", "SPH_MG_SPAN"), ("MG_PIVOT", "SPH_MG_PIVOT"),
             ("MG_STALLED", "SPH_MG_STALLED"), ("MG_LEVELS", "SPH_MG_LEVELS"),
             ("MG_STALLED_SOLVE", "SPH_MG_STALLED_SOLVE"), ("IBN_REFERENCE", "SPH_IBN_REFERENCE"),
             ("IBN_PLANES", "SPH_IBN_PLANES"), ("IBN_UPLOADED", "SPH_IBN_UPLOADED")]
    lines = ['#include <stdio.h>', '#include "urbansph.h"', 'int main(void){']
    lines += [f'printf("%d\\n", (int){c});' for _, c in pairs]
    lines.append("return 0;}")
    src = tmp_path / "enums.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "enums"
    import subprocess
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)])
    out = [int(v) for v in subprocess.check_output([str(exe)]).decode().split()]
    assert out == [getattr(abi, p) for p, _ in pairs]


def test_case_presets_normal_sources():
    """The canyon generator supplies every building's outward normals, so its
    parameters keep the uploaded normals; the library default is the
    reference's face-plane rule."""
    import cases
    _, _, geom = cases.small_canyon(dx=0.02)
    assert cases.canyon_params(geom, urbansph.default_params()).ib_normals == abi.IBN_UPLOADED
    assert urbansph.default_params().ib_normals == abi.IBN_REFERENCE
    assert cases.box_params(urbansph.default_params()).ib_normals == abi.IBN_REFERENCE
