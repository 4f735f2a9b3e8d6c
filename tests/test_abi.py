"""C-ABI library checks that need no GPU: the library loads, exports every
function include/texo_gp.h declares, and reports the typed set-up failures of
eplace::place (eplace.rs:255-336, electrostatic_placement.rs:139-240) before
any device work."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import texo_amd
from texo_amd import abi, synth
from tests import problems

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for fname in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if fname.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fname)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names += re.findall(r"\b(texo_gp_\w+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = texo_amd.load()
    names = declared_functions()
    assert "texo_gp_place" in names and len(names) >= 14
    for n in names:
        assert hasattr(lib, n), n


def test_struct_layouts_match_the_header(tmp_path):
    # the ctypes mirror must agree with what a C compiler makes of the header
    structs = {"texo_gp_problem": abi.texo_gp_problem, "texo_gp_options": abi.texo_gp_options,
               "texo_gp_report": abi.texo_gp_report, "texo_gp_field_eval": abi.texo_gp_field_eval,
               "texo_gp_schedule_state": abi.texo_gp_schedule_state,
               "texo_gp_server_stats": abi.texo_gp_server_stats}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "texo_gp.h"', "int main(void) {"]
    for name, cls in structs.items():
        lines.append(f'printf("{name} %zu\\n", sizeof({name}));')
        for field, _ in cls._fields_:
            lines.append(f'printf("{name}.{field} %zu\\n", offsetof({name}, {field}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = dict(line.split() for line in out if line)
    for name, cls in structs.items():
        assert int(got[name]) == C.sizeof(cls), name
        for field, _ in cls._fields_:
            assert int(got[f"{name}.{field}"]) == getattr(cls, field).offset, (name, field)


def create(problem, **kw):
    lib = texo_amd.load()
    rep = abi.texo_gp_report()
    h = lib.texo_gp_session_create(C.byref(problem.struct()), C.byref(abi.options(**kw)), C.byref(rep))
    return h, rep


def test_fixed_capacity_exhausted_is_reported_before_device_work():
    W, H, cap = problems.mixed_device_capacity()
    p = problems.density_problem(W, H, cap, [(1.0, 0.0, [(problems.LUT4, 0.0, 0.0)])],
                                 fixed_occupancy=[(problems.LUT4, 0, 0)] * 2)
    h, rep = create(p)
    assert not h and abi.STATUS[rep.status] == "FixedCapacityExhausted"
    assert rep.err_kind == problems.LUT4 and (rep.err_x, rep.err_y) == (0, 0)


def test_point_outside_grid_is_typed():
    W, H, cap = problems.mixed_device_capacity()
    p = problems.density_problem(W, H, cap, [(1.0, 0.0, [(problems.LUT4, 0.0, 0.0)])],
                                 fixed_occupancy=[(problems.LUT4, 9, 0)])
    h, rep = create(p)
    assert not h and abi.STATUS[rep.status] == "PointOutsideGrid" and rep.err_x == 9


def test_insufficient_capacity_is_typed():
    W, H, cap = problems.mixed_device_capacity()
    p = problems.density_problem(W, H, cap, [(0.0, 0.0, [(abi.KIND_MEMORY, 0.0, 0.0)])])
    h, rep = create(p)
    assert not h and abi.STATUS[rep.status] == "InsufficientCapacity"
    assert rep.err_kind == abi.KIND_MEMORY and tuple(rep.err_value) == (1.0, 0.0)


def test_non_finite_initial_coordinate_is_typed():
    W, H, cap = problems.mixed_device_capacity()
    p = problems.density_problem(W, H, cap, [(float("nan"), 0.0, [(problems.LUT4, 0.0, 0.0)])])
    h, rep = create(p)
    assert not h and abi.STATUS[rep.status] == "NonFiniteUnitCoordinate"


def test_duplicate_register_controls_are_invalid_normalization():
    W, H, cap = problems.mixed_device_capacity()
    p = problems.density_problem(W, H, cap, [(1.0, 1.0, [(problems.FF, 0.0, 0.0)])])
    p.register_cell = np.array([0, 0], np.uint32)
    p.register_clock_lsr = np.array([1, 2, 1, 2], np.uint64)
    p.register_ce = np.array([3, 4], np.uint64)
    h, rep = create(p)
    assert not h and abi.STATUS[rep.status] == "InvalidNormalization"


def test_bad_abi_version_and_options_are_rejected():
    p = synth.plain_1k()
    s = p.struct()
    s.abi_version = 99
    lib = texo_amd.load()
    rep = abi.texo_gp_report()
    assert not lib.texo_gp_session_create(C.byref(s), C.byref(abi.options()), C.byref(rep))
    assert abi.STATUS[rep.status] == "InvalidArgument"
    h, rep = create(synth.plain_1k(), dct_precision=7)
    assert not h and abi.STATUS[rep.status] == "InvalidArgument"


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(texo_amd, "_lib", None)
    monkeypatch.setattr(texo_amd, "LIB_PATH", str(tmp_path / "absent.so"))
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        texo_amd.load()
