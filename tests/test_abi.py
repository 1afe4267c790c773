"""The C-ABI library loads, exports every symbol declared in
``include/beamz_hip.h``, and its constants agree with ``abi_layout.json``
(the counterpart of ``tests/unit/test_cuda_abi_schema.py:16-76``).  No compute
call is made here."""

from __future__ import annotations

import ctypes
import json
import os
import re

import pytest

from beamz_amd import _abi, native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "beamz_hip.h")


def _declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bzh_[a-z_0-9]+)\s*\(", text)))


def _macros():
    text = open(HEADER).read()
    out = {}
    for name, value in re.findall(r"#define\s+(BZH_[A-Z_0-9]+)\s+(.+)", text):
        value = value.split("/*")[0].strip()
        out[name] = value
    return out


def test_library_loads_and_exports_every_declared_symbol():
    ok, why = native.status()
    assert ok, why
    lib = ctypes.CDLL(native.LIB_PATH)
    declared = _declared_functions()
    assert set(declared) == set(native.SIGNATURES), declared
    for name in declared:
        assert hasattr(lib, name), name


def test_module_queries():
    lib = native.library()
    assert lib.bzh_abi_version() == _abi.ABI_VERSION
    assert lib.bzh_component_version().decode() == _abi.COMPONENT_VERSION
    assert lib.bzh_targets().decode().split(",") == [_abi.STREAMED_TARGET, _abi.PROGRAM_TARGET]
    assert lib.bzh_last_error().decode() == ""


def test_header_constants_match_layout_json():
    m = _macros()
    layout = json.load(open(os.path.join(ROOT, "beamz_amd", "abi_layout.json")))
    assert int(m["BZH_ABI_VERSION"]) == layout["abi_version"] == 12
    assert m["BZH_COMPONENT_VERSION"].strip('"') == layout["component_version"]
    for name, bit in layout["flags"].items():
        assert m[f"BZH_FLAG_{name.upper()}"] == f"(1 << {bit})"
    for name, value in layout["program_layouts"].items():
        assert int(m[f"BZH_LAYOUT_{name.upper()}"]) == value
    for name, value in layout["layout"].items():
        assert int(m[f"BZH_{name.upper()}"]) == value


def test_layout_relations_match_reference_schema():
    # tests/unit/test_cuda_abi_schema.py:24-47 relations
    L = _abi
    assert L.CPML_E_INPUT_OFFSET == L.FIELD_COUNT + L.CPML_PHASE_INPUT_COUNT
    assert L.CPML_H_PSI_INPUT_OFFSET == L.FIELD_COUNT + 6 + 1 + 3 * L.CPML_TERM_COUNT
    assert L.CPML_E_PSI_INPUT_OFFSET == L.CPML_E_INPUT_OFFSET + 6 + 1 + 3 * L.CPML_TERM_COUNT
    assert L.CPML_H_METRIC_OFFSET == L.CPML_E_INPUT_OFFSET + L.CPML_PHASE_INPUT_COUNT
    assert L.CPML_E_METRIC_OFFSET == L.CPML_H_METRIC_OFFSET + 3
    assert L.CPML_GRAPH_INPUT_COUNT == L.CPML_E_METRIC_OFFSET + 3
    assert L.TEMPORAL_FIELD_WORKSPACE_INPUT == L.CPML_GRAPH_INPUT_COUNT
    assert L.TEMPORAL_PSI_WORKSPACE_INPUT == L.TEMPORAL_FIELD_WORKSPACE_INPUT + L.FIELD_COUNT
    assert L.YEE_GRAPH_INPUT_COUNT == L.FIELD_COUNT + 12 + 6


def test_bad_arguments_are_rejected_before_any_launch():
    lib = native.library()
    attrs = native.ProgramAttrs(abi_version=999, nsteps=1, program_layout=0)
    arr = (native.Buffer * 1)()
    assert lib.bzh_program(None, arr, 1, ctypes.byref(attrs)) == 1
    assert "ABI" in lib.bzh_last_error().decode()
    attrs = native.ProgramAttrs(abi_version=_abi.ABI_VERSION, nsteps=1, program_layout=0)
    assert lib.bzh_program(None, arr, 1, ctypes.byref(attrs)) == 1
    assert "missing" in lib.bzh_last_error().decode()


_STRUCTS = {
    "bzh_buffer": native.Buffer, "bzh_phase_attrs": native.PhaseAttrs,
    "bzh_program_attrs": native.ProgramAttrs, "bzh_sim2d": native.Sim2d,
    "bzh_batch_attrs": native.BatchAttrs, "bzh_adjoint2d": native.Adjoint2d,
    "bzh_launch": native.Launch, "bzh_source_group": native.SourceGroup,
    "bzh_dft_group": native.DftGroup, "bzh_program_launch": native.ProgramLaunch,
}


def test_ctypes_structs_match_the_header_layout(tmp_path):
    """Every descriptor the ctypes binding passes has the size and field
    offsets the C compiler gives the header's struct (host gcc, no GPU)."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no host C compiler")
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "beamz_hip.h"', "int main(void) {"]
    for cname, cls in _STRUCTS.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'  printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for line in out:
        if line:
            *key, val = line.split()
            got[tuple(key)] = int(val)
    for cname, cls in _STRUCTS.items():
        assert got[(cname, "size")] == ctypes.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert got[(cname, fname)] == getattr(cls, fname).offset, (cname, fname)
