"""Python handle over the C ABI (include/storm_gpu.h).

Mirrors the reference's manager surface for one problem:
    eng = GpuEngine(spec, precision=capi.MIXED)
    eng.step(dt)                  # MonteCarloManagerSerial::step
    eng.cells()                   # T, internal energy, Erad time average, census Erad, Fleck, sigma_P
    eng.segments()                # GetCellsStepsCounters
    eng.counters()                # particle counts, segments, timings
Every call goes to libstorm_amd.so; a failing status raises StormError with
the library's message (the reference raises StormError, StormError.hpp:13-57).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi
from .problems import ProblemSpec


class StormError(RuntimeError):
    pass


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class GpuEngine:
    def __init__(self, spec: ProblemSpec, precision: int = capi.MIXED,
                 material: int = capi.MATERIAL_DEVICE):
        self.lib = capi.load_library()
        self.spec = spec
        self.N = spec.ncells
        self.precision = precision
        self._prob = spec.to_c()
        h = C.c_void_p()
        self._check(self.lib.storm_gpu_create(C.byref(self._prob), precision, material, C.byref(h)))
        self.h = h

    def _check(self, rc):
        if rc != 0:
            raise StormError(f"storm_gpu error {rc}: {self.lib.storm_gpu_last_error().decode()}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.storm_gpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- stepping
    def step(self, dt: float | None = None):
        self._check(self.lib.storm_gpu_step(self.h, float(self.spec.dt if dt is None else dt)))

    def step_noalloc(self, dt: float | None = None):
        """storm_gpu_step_noalloc: the same step without any device allocation."""
        self._check(self.lib.storm_gpu_step_noalloc(self.h, float(self.spec.dt if dt is None else dt)))

    def set_source_temperature(self, T: float):
        self._check(self.lib.storm_gpu_set_source_temperature(self.h, float(T)))

    def set_tuning(self, block: int = 256, segments_per_pass: int = 1 << 20, schedule: int | None = None):
        """Transport launch shape; `schedule` is capi.SCHEDULE_PASS or capi.SCHEDULE_REFILL (None: keep
        the engine's); a block size fixes the launch (no schedule watch), block 0 tunes the schedule,
        the block size and the chunks per wavefront on line over the next 20 steps, then watches."""
        if schedule is None:
            schedule = -1
        self._check(self.lib.storm_gpu_set_tuning(self.h, int(block), int(segments_per_pass), int(schedule)))

    def set_cell_state(self, T=None, sigma_a=None, sigma_s=None, cv=None, e=None):
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (T, sigma_a, sigma_s, cv, e)]
        ptrs = [C.POINTER(C.c_double)() if a is None else _dp(a) for a in arrs]
        self._check(self.lib.storm_gpu_set_cell_state(self.h, *ptrs))

    # -- results
    def cells(self) -> dict:
        out = {k: np.empty(self.N) for k in ("T", "e", "erad", "census", "fleck", "sigma_p")}
        self._check(self.lib.storm_gpu_get_cells(self.h, *(_dp(out[k]) for k in
                                                            ("T", "e", "erad", "census", "fleck", "sigma_p"))))
        return out

    def group_erad(self, groups: int) -> np.ndarray:
        """getEgTimeAvg: time-averaged radiation energy density per cell and
        group of the last step, [cell, group] (problems with eg_time_avg)."""
        out = np.empty(self.N * groups)
        self._check(self.lib.storm_gpu_get_group_erad(self.h, _dp(out)))
        return out.reshape(self.N, groups)

    def set_chunks_per_wave(self, chunks: int):
        self._check(self.lib.storm_gpu_set_chunks_per_wave(self.h, int(chunks)))

    def launch(self) -> dict:
        """The transport launch in use (block, chunks per wavefront, schedule)
        and whether on-line tuning is still running."""
        v = [C.c_int32() for _ in range(4)]
        self._check(self.lib.storm_gpu_get_launch(self.h, *(C.byref(x) for x in v)))
        return dict(block=v[0].value, chunks=v[1].value, schedule=v[2].value, tuning=bool(v[3].value))

    def segments(self) -> np.ndarray:
        a = np.empty(self.N, dtype=np.uint64)
        self._check(self.lib.storm_gpu_get_segments(self.h, a.ctypes.data_as(C.POINTER(C.c_uint64))))
        return a

    def beginning_counts(self) -> np.ndarray:
        a = np.empty(self.N, dtype=np.uint64)
        self._check(self.lib.storm_gpu_get_beginning_counts(self.h, a.ctypes.data_as(C.POINTER(C.c_uint64))))
        return a

    def counters(self) -> dict:
        c = capi.Counters()
        self._check(self.lib.storm_gpu_get_counters(self.h, C.byref(c)))
        return {name: getattr(c, name) for name, _ in c._fields_}

    def ledger(self, instance: int | None = None) -> dict:
        """Energy ledger of the last step; `instance`: that instance of a batch."""
        L = capi.Ledger()
        if instance is None:
            self._check(self.lib.storm_gpu_get_ledger(self.h, C.byref(L)))
        else:
            self._check(self.lib.storm_gpu_get_instance_ledger(self.h, int(instance), C.byref(L)))
        out = {}
        for name, typ in L._fields_:
            v = getattr(L, name)
            out[name] = v.value() if typ is capi.U128 else v
        return out

    # -- census exchange
    def census_size(self) -> int:
        n = C.c_uint64()
        self._check(self.lib.storm_gpu_census_size(self.h, C.byref(n)))
        return int(n.value)

    def export_census(self):
        n = self.census_size()
        buf = (capi.PacketRecord * max(n, 1))()
        w = C.c_uint64()
        self._check(self.lib.storm_gpu_export_census(self.h, buf, n, C.byref(w)))
        return np.ctypeslib.as_array(buf)[: w.value].copy() if n else np.zeros(0, dtype=np.ctypeslib.as_array(buf).dtype)

    def queue_source_packets(self, records, time_left):
        """New packets made on the host for the next step (a user boundary's
        generateNewBoundaryParticles), each with its own time left (s)."""
        recs = np.ascontiguousarray(records)
        tl = np.ascontiguousarray(time_left, dtype=np.float64)
        assert len(tl) == len(recs)
        ptr = recs.ctypes.data_as(C.POINTER(capi.PacketRecord))
        self._check(self.lib.storm_gpu_queue_source_packets(self.h, ptr, _dp(tl), len(recs)))

    def import_census(self, records):
        recs = np.ascontiguousarray(records)
        n = len(recs)
        ptr = recs.ctypes.data_as(C.POINTER(capi.PacketRecord))
        self._check(self.lib.storm_gpu_import_census(self.h, ptr, n))

    # -- replay support
    def keep_transport_inputs(self, enable: bool = True):
        self._check(self.lib.storm_gpu_debug_keep_transport_inputs(self.h, 1 if enable else 0))

    def transport_inputs(self):
        n = C.c_uint64()
        consts = (C.c_int64 * 4)()
        self._check(self.lib.storm_gpu_debug_transport_inputs(self.h, C.byref(n), None, None, consts))
        pk = (capi.PacketF32 * max(n.value, 1))()
        xs = np.empty(2 * self.N, dtype=np.float32)
        self._check(self.lib.storm_gpu_debug_transport_inputs(self.h, C.byref(n), C.cast(pk, C.c_void_p),
                                                              xs.ctypes.data_as(C.c_void_p), consts))
        return pk, int(n.value), xs, [int(v) for v in consts]

    def spectrum_data(self):
        """Multigroup data of the last step (float32 arrays for the mixed
        engine): dict(groups, nu_pow, bounds_kev, cdf, sig_s), or None when gray."""
        G = C.c_int32()
        npow = C.c_double()
        self._check(self.lib.storm_gpu_debug_spectrum(self.h, C.byref(G), C.byref(npow), None, None, None))
        if G.value < 2:
            return None
        dt = np.float32 if self.precision == capi.MIXED else np.float64
        b = np.empty(G.value + 1, dtype=dt)
        cdf = np.empty(self.N * (G.value + 1), dtype=dt)
        ss = np.empty(self.N, dtype=dt)
        vp = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
        self._check(self.lib.storm_gpu_debug_spectrum(self.h, C.byref(G), C.byref(npow), vp(b), vp(cdf), vp(ss)))
        return dict(groups=G.value, nu_pow=npow.value, bounds_kev=b, cdf=cdf, sig_s=ss)

    def accel_data(self) -> bytes | None:
        n = C.c_uint64()
        self._check(self.lib.storm_gpu_debug_accel(self.h, None, C.byref(n)))
        if n.value == 0:
            return None
        buf = C.create_string_buffer(n.value)
        self._check(self.lib.storm_gpu_debug_accel(self.h, C.cast(buf, C.c_void_p), C.byref(n)))
        return buf.raw

    def tallies(self):
        """Integer tallies of the last step: dict of dep, rdep, pint, esc (2 limbs
        per cell) and segs."""
        t = {k: np.empty(2 * self.N, dtype=np.uint64) for k in ("dep", "rdep", "pint", "esc")}
        t["segs"] = np.empty(self.N, dtype=np.uint32)
        u64 = lambda a: a.ctypes.data_as(C.POINTER(C.c_uint64))
        self._check(self.lib.storm_gpu_debug_tallies(self.h, u64(t["dep"]), u64(t["rdep"]), u64(t["pint"]),
                                                     u64(t["esc"]), t["segs"].ctypes.data_as(C.POINTER(C.c_uint32))))
        return t


def device_info() -> str:
    lib = capi.load_library()
    buf = C.create_string_buffer(256)
    rc = lib.storm_gpu_device_info(buf, 256)
    if rc != 0:
        raise StormError(lib.storm_gpu_last_error().decode())
    return buf.value.decode()


def calibration_copy(nbytes: int) -> float:
    """One copy of `nbytes` by the library's plain copy kernel (reads and
    writes each byte once); returns its time in ms."""
    lib = capi.load_library()
    ms = C.c_double(0.0)
    rc = lib.storm_gpu_calibration_copy(C.c_uint64(nbytes), C.byref(ms))
    if rc != 0:
        raise StormError(lib.storm_gpu_last_error().decode())
    return ms.value
