"""The C-ABI library loads and exports every entry point include/storm_gpu.h declares.

No compute call is made (no GPU here): ctypes resolves the symbols only.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import pytest

from storm_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    if not os.path.exists(capi.LIB_PATH):
        pytest.fail("libstorm_amd.so is not built (run __graft_entry__.build())")
    lib = C.CDLL(capi.LIB_PATH)
    declared = capi.exported_symbols_declared()
    assert len(declared) >= 20
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_signatures_cover_declarations():
    lib = capi.load_library()
    for name in capi.exported_symbols_declared():
        assert getattr(lib, name).restype is not None or name.endswith("last_error"), name


def test_struct_layouts_match_header_sizes():
    # storm_packet_record: 8 doubles + 2 u64 = 80 bytes; Packet<float> = 8 floats + 1 double + 4 u32 + energy, padded = 64 bytes
    assert C.sizeof(capi.PacketRecord) == 80
    assert C.sizeof(capi.PacketF32) == 64
    hdr = open(os.path.join(ROOT, "include", "storm_problem.h")).read()
    for field in ("x_edges", "sigma_a0", "face_kind", "new_photons_per_cell", "comb_nmin", "T0", "seed"):
        assert re.search(r"\b" + field + r"\b", hdr), field
        assert any(field in [f[0] for f in s._fields_] for s in
                   (capi.MeshDesc, capi.MaterialDesc, capi.BoundaryDesc, capi.ImcParams, capi.Problem)), field


def test_abi_version_without_gpu():
    lib = capi.load_library()
    assert lib.storm_gpu_abi_version() == 8


def test_cpp_manager_program_fails_loudly_without_engine():
    """tests/cpp/manager_slab (user code over MonteCarloManagerGPU.hpp) links
    the engine library; without a GPU it must stop with the engine's error,
    never fall back to anything."""
    import json
    import subprocess
    exe = os.path.join(ROOT, "tests", "cpp", "_build", "manager_slab")
    if not os.path.exists(exe):
        pytest.fail("tests/cpp/_build/manager_slab is not built (run __graft_entry__.build())")
    out = subprocess.run([exe, "0", "0", "1", "0", "1", "10", "10", "2", "4"], capture_output=True, text=True,
                         timeout=120)
    if out.returncode == 0:  # a GPU is present: the run must be real
        assert len(json.loads(out.stdout)["runs"]) == 1
    else:
        assert out.returncode == 1 and "gfx950 engine" in out.stderr, (out.returncode, out.stderr)

