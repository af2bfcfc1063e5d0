"""Problem descriptions (storm_problem.h) for the gray IMC workloads.

Each builder returns a `ProblemSpec` holding numpy arrays and scalars; call
`.to_c()` for the ctypes `Problem` (the spec keeps the arrays alive).

Workloads (all CGS, T in K):
  c1_slab   gray slab 100x1x1 over [0,1] cm, sigma_a = 1, cv = 1e15 erg/(cm^3 keV),
            T0 = 1e-3 keV, SideTemperature 1 keV at x- (escape there, reflect
            elsewhere), comb population control, dt = 1e-11 s.
  c2_cube   gray 64^3 unit cube, sigma_a = sigma0_c * max(T, 0.1 keV)^-3 (T in keV),
            sigma0_c = 10^(2U) per cell, rho = 1, cv = 2 a (1 keV)^2 T (so that
            e(1 keV) = a (1 keV)^4), T0 = 1e-3 keV, 1 keV source at x-, every box
            face escapes, ~1e7 packets per step, dt = 1e-11 s.
  marshak   the power-law Marshak-wave family of the reference examples
            (examples/marshak_wave/MarshakCommon.hpp:48-172), gray, no random walk.
The 0.1 keV opacity floor of c2_cube is a recorded choice: at the literal
1e-3 keV start the opacity sigma0 * 1e9 cm^-1 makes every cell an effective
scatterer with ~1e4 events per absorption length.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import capi

CLIGHT = 2.99792458e10
K_BOLTZ = 1.380649e-16
KEV = 1.602176634e-9
KEV_K = KEV / K_BOLTZ
SIGMA_SB = 5.670374419e-5
ARAD = 4.0 * SIGMA_SB / CLIGHT


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _ptr(a):
    if a is None:
        return C.POINTER(C.c_double)()
    return a.ctypes.data_as(C.POINTER(C.c_double))


@dataclass
class UMesh:
    """A polyhedral mesh in the flattened face tables of storm_unstructured_mesh
    (include/storm_problem.h): per cell its volume and mesh point, per half
    face its neighbour (cell, or -1 - box side), outward unit normal,
    centroid, area and polygon."""
    volume: np.ndarray
    mesh_point: np.ndarray         # (N, 3)
    face_start: np.ndarray         # N + 1, uint32
    face_neighbor: np.ndarray      # F, int32
    face_normal: np.ndarray        # (F, 3)
    face_centroid: np.ndarray      # (F, 3)
    face_area: np.ndarray          # F
    face_vertex_start: np.ndarray  # F + 1, uint32
    vertices: np.ndarray           # (V, 3)
    box_lo: tuple = (0.0, 0.0, 0.0)
    box_hi: tuple = (1.0, 1.0, 1.0)
    cell_instance: np.ndarray | None = None
    _keep: list = field(default_factory=list, repr=False)

    @property
    def ncells(self) -> int:
        return len(self.volume)

    @property
    def nfaces(self) -> int:
        return len(self.face_neighbor)

    def cell_of_face(self) -> np.ndarray:
        return np.repeat(np.arange(self.ncells), np.diff(self.face_start))

    def to_c(self) -> capi.UnstructuredMesh:
        f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64).reshape(-1)  # noqa: E731
        arrs = dict(volume=f64(self.volume), mesh_point=f64(self.mesh_point),
                    face_start=np.ascontiguousarray(self.face_start, dtype=np.uint32),
                    face_neighbor=np.ascontiguousarray(self.face_neighbor, dtype=np.int32),
                    face_normal=f64(self.face_normal), face_centroid=f64(self.face_centroid),
                    face_area=f64(self.face_area),
                    face_vertex_start=np.ascontiguousarray(self.face_vertex_start, dtype=np.uint32),
                    vertices=f64(self.vertices))
        if self.cell_instance is not None:
            arrs["cell_instance"] = np.ascontiguousarray(self.cell_instance, dtype=np.uint32)
        self._keep = list(arrs.values())
        m = capi.UnstructuredMesh()
        m.ncells, m.nfaces = self.ncells, self.nfaces
        for a in range(3):
            m.box_lo[a], m.box_hi[a] = float(self.box_lo[a]), float(self.box_hi[a])
        ptype = dict(face_start=C.POINTER(C.c_uint32), face_vertex_start=C.POINTER(C.c_uint32),
                     face_neighbor=C.POINTER(C.c_int32), cell_instance=C.POINTER(C.c_uint32))
        for k, v in arrs.items():
            setattr(m, k, v.ctypes.data_as(ptype.get(k, C.POINTER(C.c_double))))
        return m


def voronoi_mesh(points, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0)) -> UMesh:
    """Voronoi tessellation of `points` (n x 3) in the box, built by the
    engine library (storm_voronoi_build); cells in Z order of the generators."""
    lib = capi.load_library()
    pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1)
    lo_a = np.ascontiguousarray(lo, dtype=np.float64)
    hi_a = np.ascontiguousarray(hi, dtype=np.float64)
    h = C.c_void_p()
    rc = lib.storm_voronoi_build(_ptr(pts), len(pts) // 3, _ptr(lo_a), _ptr(hi_a), C.byref(h))
    if rc != 0:
        raise RuntimeError(lib.storm_gpu_last_error().decode())
    try:
        v = capi.UnstructuredMesh()
        if lib.storm_voronoi_mesh(h, C.byref(v)) != 0:
            raise RuntimeError(lib.storm_gpu_last_error().decode())
        N, F = v.ncells, v.nfaces
        nv = int(np.ctypeslib.as_array(v.face_vertex_start, shape=(F + 1,))[-1])
        cp = lambda ptr, n: np.ctypeslib.as_array(ptr, shape=(n,)).copy()  # noqa: E731
        return UMesh(volume=cp(v.volume, N), mesh_point=cp(v.mesh_point, 3 * N).reshape(N, 3),
                     face_start=cp(v.face_start, N + 1), face_neighbor=cp(v.face_neighbor, F),
                     face_normal=cp(v.face_normal, 3 * F).reshape(F, 3),
                     face_centroid=cp(v.face_centroid, 3 * F).reshape(F, 3), face_area=cp(v.face_area, F),
                     face_vertex_start=cp(v.face_vertex_start, F + 1), vertices=cp(v.vertices, 3 * nv).reshape(nv, 3),
                     box_lo=tuple(lo_a), box_hi=tuple(hi_a))
    finally:
        lib.storm_voronoi_free(h)


def cartesian_umesh(xe, ye, ze) -> UMesh:
    """A Cartesian mesh written as polyhedral cells: cell (i*ny + j)*nz + k,
    faces x-, x+, y-, y+, z-, z+, the mesh point at the cell centre."""
    xe, ye, ze = (np.asarray(a, dtype=np.float64) for a in (xe, ye, ze))
    nx, ny, nz = len(xe) - 1, len(ye) - 1, len(ze) - 1
    N = nx * ny * nz
    vol, mp, nbr, nrm, fcs, area, vstart, verts = [], [], [], [], [], [], [0], []
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                lo = np.array([xe[i], ye[j], ze[k]])
                hi = np.array([xe[i + 1], ye[j + 1], ze[k + 1]])
                vol.append(float(np.prod(hi - lo)))
                mp.append(0.5 * (lo + hi))
                idx = (i, j, k)
                n3 = (nx, ny, nz)
                stride = (ny * nz, nz, 1)
                c = (i * ny + j) * nz + k
                for ax in range(3):
                    for up in (0, 1):
                        side = 2 * ax + up
                        inside = (idx[ax] + 1 < n3[ax]) if up else (idx[ax] > 0)
                        nbr.append(c + (stride[ax] if up else -stride[ax]) if inside else -1 - side)
                        n = np.zeros(3)
                        n[ax] = 1.0 if up else -1.0
                        nrm.append(n)
                        b, a2 = (ax + 1) % 3, (ax + 2) % 3
                        x0 = hi[ax] if up else lo[ax]
                        # counter-clockwise seen from outside: (b, a2) order for +n, reversed for -n
                        corners = [(lo[b], lo[a2]), (hi[b], lo[a2]), (hi[b], hi[a2]), (lo[b], hi[a2])]
                        if not up:
                            corners = corners[::-1]
                        for cb, ca in corners:
                            v = np.zeros(3)
                            v[ax], v[b], v[a2] = x0, cb, ca
                            verts.append(v)
                        vstart.append(len(verts))
                        fc = 0.5 * (lo + hi)
                        fc[ax] = x0
                        fcs.append(fc)
                        area.append(float((hi[b] - lo[b]) * (hi[a2] - lo[a2])))
    return UMesh(volume=np.array(vol), mesh_point=np.array(mp), face_start=np.arange(0, 6 * N + 1, 6, dtype=np.uint32),
                 face_neighbor=np.array(nbr, dtype=np.int32), face_normal=np.array(nrm), face_centroid=np.array(fcs),
                 face_area=np.array(area), face_vertex_start=np.array(vstart, dtype=np.uint32),
                 vertices=np.array(verts), box_lo=(xe[0], ye[0], ze[0]), box_hi=(xe[-1], ye[-1], ze[-1]))


@dataclass
class ProblemSpec:
    name: str
    x_edges: np.ndarray
    y_edges: np.ndarray
    z_edges: np.ndarray
    sigma_a0: np.ndarray
    T0: np.ndarray
    dt: float
    steps: int
    sigma_s0: np.ndarray | None = None
    density: np.ndarray | None = None
    T_ref: float = 1.0
    T_floor: float = 0.0
    a_pow: float = 0.0
    a_rho: float = 0.0
    s_pow: float = 0.0
    s_rho: float = 0.0
    eos_f: float = 1.0
    eos_beta: float = 1.0
    eos_mu: float = 0.0
    eos_cv_floor: float = 0.0
    face_kind: tuple = (capi.FACE_ESCAPE,) + (capi.FACE_REFLECT,) * 5
    source_face: int = capi.XLO
    source_T: float = KEV_K
    source_xhi: int = 0          # a second source on x+ (TwoSidesTemperature)
    source_T_hi: float = 0.0
    freeze_T4: int = 1
    source_per_face: int = 100
    new_photons_per_cell: int = 10
    comb: int = 1
    comb_nmin: int = 20
    comb_factor: float = 2.0
    comb_min_per_group: int = 2        # stratified comb (comb=2): minParticlesPerGroup
    with_rw: int = 0
    with_ddmc: int = 0
    ddmc_pgrw: int = 0           # ddmcUseMultigroupPGRW (multigroup DDMC on the diffusive groups)
    rw_cell_tau: float = 25.0
    rw_particle_tau: float = 5.0
    ddmc_cell_tau: float = 15.0
    ddmc_particle_tau: float = 5.0
    seed: int = 42
    group_bounds: np.ndarray | None = None  # photon energies, erg (multigroup when >= 3 entries)
    nu_pow: float = 0.0
    nu_ref: float = 1.0
    eg_time_avg: bool = False           # withEgTimeAvg: per-group time-averaged radiation energy
    umesh: UMesh | None = None          # polyhedral mesh (the edges above are then unused)
    instances: list | None = None       # batches: dicts of storm_instance_desc fields
    _keep: list = field(default_factory=list, repr=False)

    @property
    def shape(self):
        if self.umesh is not None:
            return (self.umesh.ncells, 1, 1)
        return (len(self.x_edges) - 1, len(self.y_edges) - 1, len(self.z_edges) - 1)

    @property
    def ncells(self) -> int:
        nx, ny, nz = self.shape
        return nx * ny * nz

    def volumes(self) -> np.ndarray:
        if self.umesh is not None:
            return np.asarray(self.umesh.volume, dtype=np.float64)
        dx = np.diff(self.x_edges)[:, None, None]
        dy = np.diff(self.y_edges)[None, :, None]
        dz = np.diff(self.z_edges)[None, None, :]
        return (dx * dy * dz).reshape(-1)

    def centers(self):
        if self.umesh is not None:
            mp = np.asarray(self.umesh.mesh_point)
            return mp[:, 0].copy(), mp[:, 1].copy(), mp[:, 2].copy()
        xc = 0.5 * (self.x_edges[1:] + self.x_edges[:-1])
        yc = 0.5 * (self.y_edges[1:] + self.y_edges[:-1])
        zc = 0.5 * (self.z_edges[1:] + self.z_edges[:-1])
        X, Y, Z = np.meshgrid(xc, yc, zc, indexing="ij")
        return X.reshape(-1), Y.reshape(-1), Z.reshape(-1)

    def with_seed(self, seed: int) -> "ProblemSpec":
        import copy
        s = copy.copy(self)
        s.seed = int(seed)
        s._keep = []
        return s

    def to_c(self) -> capi.Problem:
        arrs = [_f64(self.x_edges), _f64(self.y_edges), _f64(self.z_edges), _f64(self.sigma_a0),
                None if self.sigma_s0 is None else _f64(self.sigma_s0),
                None if self.density is None else _f64(self.density), _f64(self.T0),
                None if self.group_bounds is None else _f64(self.group_bounds)]
        self._keep = arrs
        xe, ye, ze, sa, ss, rho, t0, gb = arrs
        p = capi.Problem()
        nx, ny, nz = self.shape
        p.mesh = capi.MeshDesc(nx, ny, nz, 0, _ptr(xe), _ptr(ye), _ptr(ze))
        m = capi.MaterialDesc()
        m.sigma_a0, m.sigma_s0, m.density = _ptr(sa), _ptr(ss), _ptr(rho)
        m.opacity_T_ref, m.opacity_T_floor = self.T_ref, self.T_floor
        m.a_temp_pow, m.a_rho_pow, m.s_temp_pow, m.s_rho_pow = self.a_pow, self.a_rho, self.s_pow, self.s_rho
        m.eos_f, m.eos_beta, m.eos_mu, m.eos_cv_T_floor = self.eos_f, self.eos_beta, self.eos_mu, self.eos_cv_floor
        p.material = m
        b = capi.BoundaryDesc()
        for f in range(6):
            b.face_kind[f] = int(self.face_kind[f])
        b.source_face = int(self.source_face)
        b.freeze_source_T4 = int(self.freeze_T4)
        b.source_temperature = float(self.source_T)
        b.source_packets_per_face = int(self.source_per_face)
        b.source_xhi = int(self.source_xhi)
        b.source_temperature_xhi = float(self.source_T_hi)
        p.boundary = b
        q = capi.ImcParams()
        q.new_photons_per_cell = int(self.new_photons_per_cell)
        q.with_random_walk, q.with_ddmc = int(self.with_rw), int(self.with_ddmc)
        q.rw_min_cell_optical_depth, q.rw_min_particle_optical_depth = self.rw_cell_tau, self.rw_particle_tau
        q.ddmc_min_cell_optical_depth, q.ddmc_min_particle_optical_depth = self.ddmc_cell_tau, self.ddmc_particle_tau
        q.comb, q.comb_nmin, q.comb_factor = int(self.comb), int(self.comb_nmin), float(self.comb_factor)
        q.comb_min_per_group = int(self.comb_min_per_group)
        q.ddmc_multigroup_pgrw = int(self.ddmc_pgrw)
        p.imc = q
        p.T0 = _ptr(t0)
        p.seed = int(self.seed)
        g = capi.GroupDesc()
        g.num_groups = 0 if gb is None else len(gb) - 1
        g.bounds = _ptr(gb)
        g.nu_pow, g.nu_ref = float(self.nu_pow), float(self.nu_ref)
        g.with_eg_time_avg = int(bool(self.eg_time_avg))
        p.groups = g
        if self.umesh is not None:
            um = self.umesh.to_c()
            self._keep.append(um)
            p.unstructured = C.pointer(um)
        if self.instances:
            arr = (capi.InstanceDesc * len(self.instances))()
            for q, d in enumerate(self.instances):
                arr[q].source_temperature = float(d["source_temperature"])
                arr[q].source_packets_per_face = int(d["source_packets_per_face"])
                arr[q].new_photons_per_cell = int(d["new_photons_per_cell"])
                arr[q].comb_nmin = int(d["comb_nmin"])
                arr[q].comb_factor = float(d["comb_factor"])
            self._keep.append(arr)
            p.num_instances = len(self.instances)
            p.instances = arr
        return p


def c1_slab(seed: int = 0, packets_scale: float = 1.0, steps: int = 100, nx: int = 100) -> ProblemSpec:
    """100x1x1 gray slab, ~1e5 packets per step at packets_scale = 1."""
    n = nx
    xe = np.linspace(0.0, 1.0, n + 1)
    h = 1.0 / n
    ye = np.array([0.0, h])
    ze = np.array([0.0, h])
    s = max(packets_scale, 1e-6)
    return ProblemSpec(
        name="c1_slab", x_edges=xe, y_edges=ye, z_edges=ze,
        sigma_a0=np.ones(n), T0=np.full(n, 1e-3 * KEV_K), dt=1e-11, steps=steps,
        eos_f=1e15 / KEV_K, eos_beta=1.0, eos_mu=0.0,
        face_kind=(capi.FACE_ESCAPE,) + (capi.FACE_REFLECT,) * 5,
        source_face=capi.XLO, source_T=KEV_K, freeze_T4=1,
        source_per_face=max(1, int(round(25000 * s))),
        new_photons_per_cell=max(1, int(round(25 * s))),
        comb=1, comb_nmin=max(1, int(round(250 * s))), comb_factor=2.0, seed=seed)


def c2_cube(seed: int = 0, n: int = 64, packets_scale: float = 1.0, steps: int = 50, literal: bool = False,
            source_per_face: int = 500, comb_nmin: int = 10) -> ProblemSpec:
    """n^3 gray cube with sigma_a = sigma0 T^-3; ~1e7 packets per step at n = 64.

    literal=False keeps the recorded 0.1 keV opacity floor; literal=True runs
    sigma0 T^-3 as stated (only the examples' 1 K floor, MarshakOpacity.hpp:36)
    with the random walk on at the reference's default thresholds (cells of
    optical depth >= 25 over the mean chord, packets >= 5 mean free paths from
    every face), as the reference's drivers switch it on for opaque media."""
    xe = np.linspace(0.0, 1.0, n + 1)
    N = n ** 3
    rng = np.random.default_rng(np.random.SeedSequence([int(seed), 0xC2]))
    sigma0 = 10.0 ** (2.0 * rng.random(N))
    s = max(packets_scale, 1e-6)
    scale_cells = (64.0 / n) ** 3  # keep packets per cell comparable on smaller meshes
    return ProblemSpec(
        name="c2_cube", x_edges=xe, y_edges=xe.copy(), z_edges=xe.copy(),
        sigma_a0=sigma0, T0=np.full(N, 1e-3 * KEV_K), dt=1e-11, steps=steps,
        T_ref=KEV_K, T_floor=1.0 if literal else 0.1 * KEV_K, a_pow=3.0,
        eos_f=ARAD * KEV_K ** 2, eos_beta=2.0, eos_mu=0.0,
        face_kind=(capi.FACE_ESCAPE,) * 6, source_face=capi.XLO, source_T=KEV_K, freeze_T4=0,
        source_per_face=max(1, int(round(source_per_face * s * min(scale_cells, 64.0) ** (1 / 3)))),
        new_photons_per_cell=max(1, int(round(2 * s))),
        comb=1, comb_nmin=max(1, int(round(comb_nmin * s))), comb_factor=1.0, seed=seed,
        with_rw=int(literal), rw_cell_tau=25.0, rw_particle_tau=5.0)


def c3_step(seed: int = 0, nx: int = 256, nyz: int = 16, packets_scale: float = 1.0, steps: int = 100,
            rw: bool = True, ddmc: bool = True) -> ProblemSpec:
    """Two-material step-opacity slab, nx x nyz x nyz cells of 1/nx cm.

    x < L/2: sigma_a = 1; x >= L/2: sigma_a = 1e4 and sigma_s = 1e4 (scattering
    ratio 0.5).  T0 = 0.1 keV, cv = 1e15 erg/(cm^3 keV), 1 keV SideTemperature
    on x- (escape there, reflect elsewhere).  Random walk and DDMC in cells of
    optical thickness >= 10; the random walk needs a particle depth of 10 and
    DDMC of 5, so packets at depths 5..10 take DDMC steps (with equal
    thresholds DDMC is never reached: the reference tries the random walk
    first, RadiationIMC.hpp:1723-1738)."""
    h = 1.0 / nx
    xe = np.linspace(0.0, 1.0, nx + 1)
    ye = np.linspace(0.0, nyz * h, nyz + 1)
    N = nx * nyz * nyz
    xc = np.repeat(0.5 * (xe[1:] + xe[:-1]), nyz * nyz)
    thick = xc >= 0.5
    sa = np.where(thick, 1e4, 1.0)
    ss = np.where(thick, 1e4, 0.0)
    s = max(packets_scale, 1e-6)
    return ProblemSpec(
        name="c3_step", x_edges=xe, y_edges=ye, z_edges=ye.copy(),
        sigma_a0=sa, sigma_s0=ss, T0=np.full(N, 0.1 * KEV_K), dt=1e-11, steps=steps,
        eos_f=1e15 / KEV_K, eos_beta=1.0, eos_mu=0.0,
        face_kind=(capi.FACE_ESCAPE,) + (capi.FACE_REFLECT,) * 5, source_face=capi.XLO, source_T=KEV_K,
        freeze_T4=1, source_per_face=max(1, int(round(4000 * s))),
        new_photons_per_cell=max(1, int(round(20 * s))),
        comb=1, comb_nmin=max(1, int(round(20 * s))), comb_factor=1.0,
        with_rw=int(rw), with_ddmc=int(ddmc), rw_cell_tau=10.0, rw_particle_tau=10.0,
        ddmc_cell_tau=10.0, ddmc_particle_tau=5.0, seed=seed)


# Marshak-wave problems 1-4 (examples/marshak_wave/MarshakCommon.hpp:48-112):
# kappa_P = kP0 (keV/T)^alpha rho^beta, scattering kappa_R - kappa_P,
# e = f T^eb rho^(1 - mu) per volume, T_bath = Tc (t/ns)^Te keV.
MARSHAK = {
    1: dict(L=0.2, x0=0.0, kP0=0.1, kR0=100.0, alpha=3.0, beta=0.0, f=6.860085e14 / KEV_K ** 4, eb=4.0, mu=0.0,
            Tc=1.008038, Te=1.0 / 3.0, dt0=1e-15),
    2: dict(L=0.2, x0=0.0, kP0=100.0, kR0=100.0, alpha=3.0, beta=0.0, f=6.860085e14 / KEV_K ** 4, eb=4.0, mu=0.0,
            Tc=1.014565, Te=1.0 / 3.0, dt0=1e-15),
    3: dict(L=1.0, x0=0.0, kP0=0.1, kR0=40.0, alpha=1.5, beta=1.2, f=1e14 / KEV_K ** 3.4, eb=3.4, mu=0.14,
            Tc=1.0470478, Te=86.0 / 57.0, dt0=1e-15),
    4: dict(L=1.0, x0=1e-5, kP0=0.001, kR0=2.0, alpha=4.5, beta=1.9, f=1e14 / KEV_K ** 6.0, eb=6.0, mu=0.3,
            Tc=1.01008116, Te=14.0 / 139.0, dt0=1e-17),
}
MARSHAK_T_FINAL = 1e-9


def marshak_density(problem: int, x):
    """ComputeDensity (MarshakCommon.hpp:114-123)."""
    x = np.maximum(np.asarray(x, dtype=np.float64), 1e-30)
    if problem == 3:
        return x ** (20.0 / 19.0)
    if problem == 4:
        return x ** (-40.0 / 139.0)
    return np.ones_like(x)


def marshak4_edges(x0: float, xmax: float) -> np.ndarray:
    """BuildMarshak4MeshEdges (MarshakCommon.hpp:256-271): 512 cells growing by 1.00969..."""
    e = np.empty(513)
    e[0] = x0
    w = 7.0e-5
    for i in range(512):
        e[i + 1] = xmax if i == 511 else e[i] + w
        w *= 1.0096903847403653
    return e


def marshak(problem: int = 2, nx: int = 256, seed: int = 0, new_photons: int = 15, boundary_photons: int = 100,
            steps: int = 10, dt: float | None = None, random_walk: bool = True) -> ProblemSpec:
    """Marshak wave 1-4 as RunMarshakWave sets it up (MarshakCommon.hpp:273-330): gray,
    random walk on, comb(15, 6.0), x- drive with escape, reflecting elsewhere."""
    P = MARSHAK[problem]
    xmax = P["x0"] + P["L"]
    if problem == 4:
        xe = marshak4_edges(P["x0"], xmax)
        h = xmax / 512.0
    else:
        xe = np.linspace(0.0, xmax, nx + 1)
        h = xmax / nx
    n = len(xe) - 1
    xc = 0.5 * (xe[1:] + xe[:-1])
    rho = marshak_density(problem, xc)
    T0 = 1e-3 * KEV_K
    return ProblemSpec(
        name=f"marshak{problem}", x_edges=xe, y_edges=np.array([0.0, h]), z_edges=np.array([0.0, h]),
        sigma_a0=np.full(n, P["kP0"]), sigma_s0=np.full(n, max(P["kR0"] - P["kP0"], 0.0)),
        density=None if problem in (1, 2) else rho,
        T0=np.full(n, T0), dt=P["dt0"] if dt is None else dt, steps=steps,
        T_ref=KEV_K, T_floor=1.0, a_pow=P["alpha"], s_pow=P["alpha"], a_rho=P["beta"], s_rho=P["beta"],
        eos_f=P["f"], eos_beta=P["eb"], eos_mu=P["mu"], eos_cv_floor=1.0,
        face_kind=(capi.FACE_ESCAPE,) + (capi.FACE_REFLECT,) * 5, source_face=capi.XLO,
        source_T=marshak_bath_T(problem, P["dt0"]), freeze_T4=0,
        source_per_face=boundary_photons, new_photons_per_cell=new_photons,
        comb=1, comb_nmin=15, comb_factor=6.0, with_rw=int(random_walk), seed=seed)


def marshak_schedule(problem: int):
    """The reference's time steps to 1 ns (MarshakCommon.hpp:343-357):
    yields (t_now, dt) with the bath temperature taken at t_now."""
    dt0 = MARSHAK[problem]["dt0"]
    t, dt = 0.0, dt0
    while t < MARSHAK_T_FINAL:
        t_now = max(t, dt0)
        dt = min(dt, MARSHAK_T_FINAL - t)
        yield t_now, dt
        t += dt
        dt = min(max(dt0, t * 1e-3), 5e-11)


def marshak_l1(x_sim, T_sim, ref) -> float:
    """ComputeL1 (MarshakCommon.hpp:192-228): mean relative |T_sim - T_ref| over
    the reference points with T_ref > 1e-4 keV, T_sim linearly interpolated."""
    l1, cnt, j = 0.0, 0, 0
    for xr, Tr in zip(ref[:, 0], ref[:, 1]):
        while j + 1 < len(x_sim) and x_sim[j + 1] < xr:
            j += 1
        if j + 1 >= len(x_sim):
            break
        fr = (xr - x_sim[j]) / (x_sim[j + 1] - x_sim[j])
        Ts = T_sim[j] + fr * (T_sim[j + 1] - T_sim[j])
        if Tr / KEV_K > 1e-4:
            l1 += abs(Ts / KEV_K - Tr / KEV_K) / (Tr / KEV_K)
            cnt += 1
    return l1 / cnt if cnt else -1.0


def marshak_shape(x_sim, T_sim, ref):
    """Shape checks of a Marshak wave that do not hinge on the cold floor.

    The reference's relative L1 (marshak_l1) divides by T_ref down to its
    1e-3 keV floor, so a front foot one cell ahead of the diffusion profile
    weighs tens of times more than the whole heated region; the reference's
    own regression accepts L1 >= 0.10 as a warning for that reason
    (regression_tests/lib/regression_checks.sh:71-95).  Returns the mean
    relative |T_sim - T_ref| over the reference points hotter than a tenth of
    the hottest, and the front shift: the position where T falls through half
    its maximum, simulation minus reference (cm), with the local cell width
    there."""
    xr, Tr = ref[:, 0], ref[:, 1]
    hot = Tr > 0.1 * Tr.max()
    Ts = np.interp(xr, x_sim, T_sim)
    l1_hot = float(np.mean(np.abs(Ts[hot] - Tr[hot]) / Tr[hot]))

    def front(x, T):
        thr = 0.5 * T.max()
        i = int(np.nonzero(T > thr)[0][-1])
        if i + 1 >= len(x):
            return float(x[-1])
        return float(x[i] + (thr - T[i]) / (T[i + 1] - T[i]) * (x[i + 1] - x[i]))

    xf = front(xr, Tr)
    j = int(np.clip(np.searchsorted(x_sim, xf), 1, len(x_sim) - 1))
    return l1_hot, front(np.asarray(x_sim), np.asarray(T_sim)) - xf, float(x_sim[j] - x_sim[j - 1])


def densmore(nx: int = 512, seed: int = 0, new_photons: int = 16, boundary_photons: int = 100, steps: int = 200,
             groups: int = 30, comb_nmin: int = 200, comb_factor: float = 5.0, random_walk: bool = True) -> ProblemSpec:
    """Densmore et al. (2012) step-opacity benchmark as examples/densmore2012/main.cpp:128-200
    sets it up: slab [0, 3] cm, sigma(E) = sigma0 / (sqrt(kT) E^3) with sigma0 = 10
    keV^3.5/cm for x < 2 cm and 1000 beyond, cv = 1e15 erg/(cm^3 keV), T0 = 1 eV,
    a 1 keV x- source (escape there, reflect elsewhere), 30 log-spaced groups over
    1e-4..1e2 keV, dt = 5e-12 s, comb(200, 5), and the random walk on (partial-gray
    in multigroup, main.cpp:208-209) with the default thresholds."""
    L, xstep = 3.0, 2.0
    xe = np.linspace(0.0, L, nx + 1)
    h = L / nx
    xc = 0.5 * (xe[1:] + xe[:-1])
    sa0 = np.where(xc < xstep, 10.0, 1000.0)
    bounds = KEV * 1e-4 * (1e6 ** (np.arange(groups + 1) / groups))
    bounds[-1] = KEV * 1e2
    return ProblemSpec(
        name="densmore", x_edges=xe, y_edges=np.array([0.0, h]), z_edges=np.array([0.0, h]),
        sigma_a0=sa0, T0=np.full(nx, KEV_K / 1000.0), dt=5e-12, steps=steps,
        T_ref=KEV_K, T_floor=1.0, a_pow=0.5,
        eos_f=1e15 / KEV_K, eos_beta=1.0, eos_mu=0.0,
        face_kind=(capi.FACE_ESCAPE,) + (capi.FACE_REFLECT,) * 5, source_face=capi.XLO, source_T=KEV_K,
        freeze_T4=0, source_per_face=boundary_photons, new_photons_per_cell=new_photons,
        comb=1, comb_nmin=comb_nmin, comb_factor=comb_factor, seed=seed, with_rw=int(random_walk),
        group_bounds=bounds, nu_pow=3.0, nu_ref=KEV)


def marshak_bath_T(problem: int, t: float) -> float:
    P = MARSHAK[problem]
    t_ns = max(t, 1e-20) * 1e9
    return P["Tc"] * t_ns ** P["Te"] * KEV_K


def voronoi_points(n: int, seed: int, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0)) -> np.ndarray:
    """n generators uniform in the box, from a seeded generator."""
    rng = np.random.default_rng(np.random.SeedSequence([int(seed), 0xC4]))
    lo, hi = np.asarray(lo, dtype=np.float64), np.asarray(hi, dtype=np.float64)
    return lo + (hi - lo) * rng.random((n, 3))


def voronoi_slab(seed: int = 0, n: int = 200, mesh_seed: int = 4, steps: int = 3, thick: bool = False,
                 packets_scale: float = 1.0, groups: int = 0) -> ProblemSpec:
    """A Voronoi tessellation of n seeded points in [0,1] x [0,0.5] x [0,0.5]
    cm: sigma_a = 3 cm^-1 (x < 0.5) and 30 cm^-1 beyond (thick=True: 2e3
    cm^-1 with sigma_s = 2e3 and the random walk and DDMC on, cell depth 10,
    particle depths 10 and 5), cv = 1e15 erg/(cm^3 keV), T0 = 1e-3 keV (0.05
    keV when thick), a 1 keV x- source, escape on x+, reflecting elsewhere,
    comb.  groups > 0: that many log-spaced groups over 0.01-20 keV with
    sigma(E) proportional to E^-3 (1 keV reference)."""
    lo, hi = (0.0, 0.0, 0.0), (1.0, 0.5, 0.5)
    m = voronoi_mesh(voronoi_points(n, mesh_seed, lo, hi), lo, hi)
    N = m.ncells
    x = m.mesh_point[:, 0]
    s = max(packets_scale, 1e-6)
    sa = np.where(x < 0.5, 3.0, 2e3 if thick else 30.0)
    ss = np.where(x < 0.5, 0.0, 2e3) if thick else None
    spec = ProblemSpec(
        name="voronoi_slab", x_edges=np.array([0.0, 1.0]), y_edges=np.array([0.0, 0.5]), z_edges=np.array([0.0, 0.5]),
        sigma_a0=sa, sigma_s0=ss, T0=np.full(N, (0.05 if thick else 1e-3) * KEV_K), dt=1e-11, steps=steps,
        eos_f=1e15 / KEV_K, eos_beta=1.0, eos_mu=0.0,
        face_kind=(capi.FACE_REFLECT, capi.FACE_ESCAPE) + (capi.FACE_REFLECT,) * 4, source_face=capi.XLO,
        source_T=KEV_K, freeze_T4=1, source_per_face=max(1, int(round(400 * s))),
        new_photons_per_cell=max(1, int(round(8 * s))), comb=1, comb_nmin=max(1, int(round(12 * s))),
        comb_factor=2.0, seed=seed, umesh=m,
        with_rw=int(thick), with_ddmc=int(thick), rw_cell_tau=10.0, rw_particle_tau=10.0, ddmc_cell_tau=10.0,
        ddmc_particle_tau=5.0)
    if groups:
        b = KEV * 0.01 * (2000.0 ** (np.arange(groups + 1) / groups))
        b[-1] = KEV * 20.0
        spec.group_bounds = b
        spec.nu_pow, spec.nu_ref = 3.0, KEV
        spec.T0 = np.full(N, 0.3 * KEV_K)
    return spec


def c4_voronoi(seed: int = 0, n: int = 200000, groups: int = 64, steps: int = 20, packets_scale: float = 1.0,
               mesh: UMesh | None = None) -> ProblemSpec:
    """Multigroup IMC on the Voronoi tessellation of n uniform seeded points
    in the unit cube: 64 log-spaced groups over 0.01-20 keV, sigma(E, T) =
    10 (E/keV)^-3 (T/keV)^-1/2 cm^-1 (the Densmore form,
    examples/densmore2012/DensmoreOpacity.hpp:58-66), density
    1 + 0.5 sin(2 pi x), e = 1e15 erg/(g keV) * rho * T, T0 = 0.1 keV, a 1 keV
    source on the x- faces, escape on every face, comb; ~5e7 packets per step
    at packets_scale = 1: 1e7 from the source, 2e7 emitted (the emission
    budget is 10 NewPhotonsPerCell per cell), a census combed to 100-200 per
    cell."""
    m = mesh if mesh is not None else voronoi_mesh(voronoi_points(n, 0xC4))
    N = m.ncells
    rho = 1.0 + 0.5 * np.sin(2.0 * np.pi * m.mesh_point[:, 0])
    b = KEV * 0.01 * (2000.0 ** (np.arange(groups + 1) / groups))
    b[-1] = KEV * 20.0
    s = max(packets_scale, 1e-6)
    # the x- box faces: about n^(2/3) of them
    nsrc = int(np.count_nonzero(m.face_neighbor == -1))
    return ProblemSpec(
        name="c4_voronoi", x_edges=np.array([0.0, 1.0]), y_edges=np.array([0.0, 1.0]), z_edges=np.array([0.0, 1.0]),
        sigma_a0=np.full(N, 10.0), density=rho, T0=np.full(N, 0.1 * KEV_K), dt=1e-11, steps=steps,
        T_ref=KEV_K, T_floor=1.0, a_pow=0.5, eos_f=1e15 / KEV_K, eos_beta=1.0, eos_mu=0.0,
        face_kind=(capi.FACE_ESCAPE,) * 6, source_face=capi.XLO, source_T=KEV_K, freeze_T4=1,
        source_per_face=max(1, int(round(1e7 * s / max(nsrc, 1)))),
        new_photons_per_cell=max(1, int(round(10 * s))), comb=1, comb_nmin=max(1, int(round(100 * s))),
        comb_factor=1.0, seed=seed, umesh=m, group_bounds=b, nu_pow=3.0, nu_ref=KEV)


def c5_instances(count: int = 256, seed: int = 0, packets_scale: float = 1.0,
                 cells_range=(50, 2000), packets_range=(1e4, 1e6)) -> list:
    """Parameters of independent 1-D slab gray IMC instances: cells
    log-uniform in cells_range, sigma_a log-uniform in [0.1, 1e3] cm^-1,
    boundary temperature uniform in [0.3, 1.5] keV, packets per step
    log-uniform in packets_range (half from the x- source, a quarter emitted,
    a quarter the comb's floor)."""
    rng = np.random.default_rng(np.random.SeedSequence([int(seed), 0xC5]))
    out = []
    lc = np.log10(np.asarray(cells_range, dtype=np.float64))
    lp = np.log10(np.asarray(packets_range, dtype=np.float64))
    for _ in range(count):
        n = int(round(10.0 ** rng.uniform(lc[0], lc[1])))
        sigma = 10.0 ** rng.uniform(-1.0, 3.0)
        Tb = rng.uniform(0.3, 1.5)
        P = 10.0 ** rng.uniform(lp[0], lp[1]) * packets_scale
        out.append(dict(cells=n, sigma_a=sigma, T_keV=Tb, packets=P,
                        source_packets_per_face=max(1, int(round(0.5 * P))),
                        new_photons_per_cell=max(1, int(round(0.25 * P / n))),
                        comb_nmin=max(1, int(round(0.25 * P / n))), comb_factor=1.0))
    return out


def c5_instance_spec(d: dict, seed: int = 0, steps: int = 10) -> ProblemSpec:
    """One instance of the batch as a stand-alone Cartesian problem: nx x 1 x 1
    cubic cells over [0, 1] cm, C1's material (cv = 1e15 erg/(cm^3 keV), T0 =
    1e-3 keV), x- source at the instance's temperature (escape there, reflect
    elsewhere), DDMC in cells of optical depth >= 10 (particle depth 5)."""
    n = d["cells"]
    h = 1.0 / n
    return ProblemSpec(
        name="c5_instance", x_edges=np.linspace(0.0, 1.0, n + 1), y_edges=np.array([0.0, h]),
        z_edges=np.array([0.0, h]), sigma_a0=np.full(n, d["sigma_a"]), T0=np.full(n, 1e-3 * KEV_K), dt=1e-11,
        steps=steps, eos_f=1e15 / KEV_K, eos_beta=1.0, eos_mu=0.0,
        face_kind=(capi.FACE_ESCAPE,) + (capi.FACE_REFLECT,) * 5, source_face=capi.XLO, source_T=d["T_keV"] * KEV_K,
        freeze_T4=1, source_per_face=d["source_packets_per_face"], new_photons_per_cell=d["new_photons_per_cell"],
        comb=1, comb_nmin=d["comb_nmin"], comb_factor=d["comb_factor"], seed=seed,
        with_ddmc=1, ddmc_cell_tau=10.0, ddmc_particle_tau=5.0)


def c5_batch(instances: list, seed: int = 0, steps: int = 10) -> ProblemSpec:
    """The instances as one batched problem: their slabs, written as polyhedral
    cells (cartesian_umesh), concatenated with per-cell instance indices; no
    face joins two instances."""
    meshes = [cartesian_umesh(np.linspace(0.0, 1.0, d["cells"] + 1), np.array([0.0, 1.0 / d["cells"]]),
                              np.array([0.0, 1.0 / d["cells"]])) for d in instances]
    vol, mp, fs, nb, nrm, fc, fa, fvs, vs, ci = [], [], [0], [], [], [], [], [0], [], []
    cell0 = 0
    for k, m in enumerate(meshes):
        vol.append(m.volume)
        mp.append(m.mesh_point)
        fs.extend((np.asarray(m.face_start[1:], np.int64) + fs[-1]).tolist())
        nbk = np.asarray(m.face_neighbor, dtype=np.int64)
        nb.append(np.where(nbk >= 0, nbk + cell0, nbk))
        nrm.append(m.face_normal)
        fc.append(m.face_centroid)
        fa.append(m.face_area)
        fvs.extend((np.asarray(m.face_vertex_start[1:], np.int64) + fvs[-1]).tolist())
        vs.append(m.vertices)
        ci.append(np.full(m.ncells, k, dtype=np.uint32))
        cell0 += m.ncells
    um = UMesh(volume=np.concatenate(vol), mesh_point=np.concatenate(mp), face_start=np.array(fs, dtype=np.uint32),
               face_neighbor=np.concatenate(nb).astype(np.int32), face_normal=np.concatenate(nrm),
               face_centroid=np.concatenate(fc), face_area=np.concatenate(fa),
               face_vertex_start=np.array(fvs, dtype=np.uint32), vertices=np.concatenate(vs),
               box_lo=(0.0, 0.0, 0.0), box_hi=(1.0, 1.0, 1.0), cell_instance=np.concatenate(ci))
    N = um.ncells
    sa = np.concatenate([np.full(d["cells"], d["sigma_a"]) for d in instances])
    spec = c5_instance_spec(instances[0], seed=seed, steps=steps)
    spec.name = "c5_batch"
    spec.umesh = um
    spec.sigma_a0 = sa
    spec.T0 = np.full(N, 1e-3 * KEV_K)
    spec.instances = [dict(source_temperature=d["T_keV"] * KEV_K, source_packets_per_face=d["source_packets_per_face"],
                           new_photons_per_cell=d["new_photons_per_cell"], comb_nmin=d["comb_nmin"],
                           comb_factor=d["comb_factor"]) for d in instances]
    return spec
