"""ctypes mirror of include/storm_problem.h and include/storm_gpu.h.

The engine library is loaded from storm_amd/_lib/libstorm_amd.so (built in
tree by storm_amd/build_lib.sh).  There is no fallback: if the library is
missing or fails to load, `load_library()` raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libstorm_amd.so")

XLO, XHI, YLO, YHI, ZLO, ZHI = range(6)
FACE_REFLECT, FACE_ESCAPE = 0, 1
MIXED, FP64 = 0, 1
SCHEDULE_PASS, SCHEDULE_REFILL = 0, 1
MATERIAL_DEVICE, MATERIAL_HOST = 0, 1

_dp = C.POINTER(C.c_double)


class MeshDesc(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("reserved", C.c_int32),
                ("x_edges", _dp), ("y_edges", _dp), ("z_edges", _dp)]


class MaterialDesc(C.Structure):
    _fields_ = [("sigma_a0", _dp), ("sigma_s0", _dp), ("density", _dp),
                ("opacity_T_ref", C.c_double), ("opacity_T_floor", C.c_double),
                ("a_temp_pow", C.c_double), ("a_rho_pow", C.c_double),
                ("s_temp_pow", C.c_double), ("s_rho_pow", C.c_double),
                ("eos_f", C.c_double), ("eos_beta", C.c_double), ("eos_mu", C.c_double),
                ("eos_cv_T_floor", C.c_double)]


class BoundaryDesc(C.Structure):
    _fields_ = [("face_kind", C.c_int32 * 6), ("source_face", C.c_int32),
                ("freeze_source_T4", C.c_int32), ("source_temperature", C.c_double),
                ("source_packets_per_face", C.c_uint64), ("source_xhi", C.c_int32), ("reserved", C.c_int32),
                ("source_temperature_xhi", C.c_double)]


class ImcParams(C.Structure):
    _fields_ = [("new_photons_per_cell", C.c_uint64), ("with_random_walk", C.c_int32),
                ("with_ddmc", C.c_int32), ("rw_min_cell_optical_depth", C.c_double),
                ("rw_min_particle_optical_depth", C.c_double),
                ("ddmc_min_cell_optical_depth", C.c_double),
                ("ddmc_min_particle_optical_depth", C.c_double),
                ("comb", C.c_int32), ("ddmc_multigroup_pgrw", C.c_int32),
                ("comb_nmin", C.c_uint64), ("comb_factor", C.c_double), ("comb_min_per_group", C.c_uint64)]


class GroupDesc(C.Structure):
    _fields_ = [("num_groups", C.c_int32), ("with_eg_time_avg", C.c_int32), ("bounds", _dp),
                ("nu_pow", C.c_double), ("nu_ref", C.c_double)]


_u32p = C.POINTER(C.c_uint32)
_i32p = C.POINTER(C.c_int32)


class UnstructuredMesh(C.Structure):
    _fields_ = [("ncells", C.c_uint32), ("nfaces", C.c_uint32), ("box_lo", C.c_double * 3),
                ("box_hi", C.c_double * 3), ("volume", _dp), ("mesh_point", _dp), ("face_start", _u32p),
                ("face_neighbor", _i32p), ("face_normal", _dp), ("face_centroid", _dp), ("face_area", _dp),
                ("face_vertex_start", _u32p), ("vertices", _dp), ("cell_instance", _u32p)]


class InstanceDesc(C.Structure):
    _fields_ = [("source_temperature", C.c_double), ("source_packets_per_face", C.c_uint64),
                ("new_photons_per_cell", C.c_uint64), ("comb_nmin", C.c_uint64), ("comb_factor", C.c_double)]


class Problem(C.Structure):
    _fields_ = [("mesh", MeshDesc), ("material", MaterialDesc), ("boundary", BoundaryDesc),
                ("imc", ImcParams), ("T0", _dp), ("seed", C.c_uint64), ("groups", GroupDesc),
                ("unstructured", C.POINTER(UnstructuredMesh)), ("num_instances", C.c_uint32),
                ("reserved2", C.c_uint32), ("instances", C.POINTER(InstanceDesc))]


class PacketRecord(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double),
                ("vx", C.c_double), ("vy", C.c_double), ("vz", C.c_double),
                ("weight", C.c_double), ("frequency", C.c_double),
                ("cell", C.c_uint64), ("id", C.c_uint64)]


class Counters(C.Structure):
    _fields_ = [("initial_particles", C.c_uint64), ("prestep_particles", C.c_uint64),
                ("start_particles", C.c_uint64), ("end_particles", C.c_uint64),
                ("emitted_particles", C.c_uint64), ("boundary_particles", C.c_uint64),
                ("census_before_comb", C.c_uint64), ("leaving_particles", C.c_uint64),
                ("segments", C.c_uint64), ("transport_passes", C.c_uint64),
                ("transport_launches", C.c_uint64), ("random_walk_steps", C.c_uint64),
                ("ddmc_steps", C.c_uint64), ("energy_quantum_exp", C.c_int32),
                ("transport_block", C.c_int32), ("seconds_step", C.c_double),
                ("seconds_transport", C.c_double), ("seconds_transport_max_launch", C.c_double),
                ("max_packet_blocks", C.c_uint64),
                ("transport_chunks", C.c_int32),
                ("transport_schedule", C.c_int32)]


class U128(C.Structure):
    _fields_ = [("lo", C.c_uint64), ("hi", C.c_uint64)]

    def value(self) -> int:
        return int(self.lo) + (int(self.hi) << 64)


class Ledger(C.Structure):
    _fields_ = [("census_in", U128), ("emitted", U128), ("source", U128), ("deposited", U128),
                ("escaped", U128), ("census_pre", U128), ("census_out", U128), ("comb_rem", U128),
                ("comb_dropped", U128),
                ("eq", C.c_int32), ("eq_prev", C.c_int32), ("requant_rem", C.c_double),
                ("clamp", C.c_double), ("material_before", C.c_double),
                ("material_after", C.c_double)]


class PacketF32(C.Structure):
    """storm::Packet<float> (storm_amd/csrc/common/transport.h)."""
    _fields_ = [("u", C.c_float * 3), ("s", C.c_float), ("o", C.c_float * 3), ("wmin", C.c_float),
                ("w", C.c_double), ("ijk", C.c_uint32), ("id", C.c_uint32), ("ctr", C.c_uint32),
                ("flags", C.c_uint32), ("nu", C.c_float), ("pad", C.c_uint32)]


_LIB = None


def load_library(path: str | None = None):
    """Load libstorm_amd.so (raises if absent: there is no CPU fallback)."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"storm_amd engine library not built: {p} (run storm_amd/build_lib.sh)")
    lib = C.CDLL(p)
    vp = C.c_void_p
    sigs = {
        "storm_gpu_last_error": ([], C.c_char_p),
        "storm_gpu_abi_version": ([], C.c_int),
        "storm_gpu_device_info": ([C.c_char_p, C.c_size_t], C.c_int),
        "storm_gpu_create": ([C.POINTER(Problem), C.c_int32, C.c_int32, C.POINTER(vp)], C.c_int),
        "storm_gpu_destroy": ([vp], C.c_int),
        "storm_gpu_step": ([vp, C.c_double], C.c_int),
        "storm_gpu_step_noalloc": ([vp, C.c_double], C.c_int),
        "storm_gpu_set_source_temperature": ([vp, C.c_double], C.c_int),
        "storm_gpu_set_cell_state": ([vp, _dp, _dp, _dp, _dp, _dp], C.c_int),
        "storm_gpu_get_cells": ([vp, _dp, _dp, _dp, _dp, _dp, _dp], C.c_int),
        "storm_gpu_get_segments": ([vp, C.POINTER(C.c_uint64)], C.c_int),
        "storm_gpu_get_beginning_counts": ([vp, C.POINTER(C.c_uint64)], C.c_int),
        "storm_gpu_get_counters": ([vp, C.POINTER(Counters)], C.c_int),
        "storm_gpu_get_ledger": ([vp, C.POINTER(Ledger)], C.c_int),
        "storm_gpu_get_instance_ledger": ([vp, C.c_uint32, C.POINTER(Ledger)], C.c_int),
        "storm_gpu_get_group_erad": ([vp, _dp], C.c_int),
        "storm_gpu_set_chunks_per_wave": ([vp, C.c_int32], C.c_int),
        "storm_gpu_get_launch": ([vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                  C.POINTER(C.c_int32)], C.c_int),
        "storm_gpu_queue_source_packets": ([vp, C.POINTER(PacketRecord), _dp, C.c_uint64], C.c_int),
        "storm_gpu_census_size": ([vp, C.POINTER(C.c_uint64)], C.c_int),
        "storm_gpu_export_census": ([vp, C.POINTER(PacketRecord), C.c_uint64, C.POINTER(C.c_uint64)], C.c_int),
        "storm_gpu_import_census": ([vp, C.POINTER(PacketRecord), C.c_uint64], C.c_int),
        "storm_gpu_set_tuning": ([vp, C.c_int32, C.c_int32, C.c_int32], C.c_int),
        "storm_gpu_debug_keep_transport_inputs": ([vp, C.c_int32], C.c_int),
        "storm_gpu_debug_transport_inputs": ([vp, C.POINTER(C.c_uint64), vp, vp, C.POINTER(C.c_int64)], C.c_int),
        "storm_gpu_debug_tallies": ([vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                     C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)], C.c_int),
        "storm_gpu_debug_accel": ([vp, vp, C.POINTER(C.c_uint64)], C.c_int),
        "storm_gpu_debug_spectrum": ([vp, C.POINTER(C.c_int32), C.POINTER(C.c_double), vp, vp, vp], C.c_int),
        "storm_gpu_calibration_copy": ([C.c_uint64, C.POINTER(C.c_double)], C.c_int),
        "storm_voronoi_build": ([_dp, C.c_uint32, _dp, _dp, C.POINTER(vp)], C.c_int),
        "storm_voronoi_mesh": ([vp, C.POINTER(UnstructuredMesh)], C.c_int),
        "storm_voronoi_free": ([vp], C.c_int),
    }
    for name, (args, res) in sigs.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    if path is None:
        _LIB = lib
    return lib


def exported_symbols_declared() -> list[str]:
    """Function names declared in the library's C headers (include/storm_gpu.h, include/storm_mesh.h)."""
    import re
    names = set()
    for h in ("storm_gpu.h", "storm_mesh.h"):
        with open(os.path.join(HERE, "..", "include", h)) as f:
            names |= set(re.findall(r"\b(storm_(?:gpu|voronoi)_[a-z0-9_]+)\s*\(", f.read()))
    return sorted(names)
