/*
 * storm_problem.h -- plain-C description of one gray Implicit Monte Carlo
 * problem on a Cartesian mesh: mesh, material models, boundary and IMC
 * parameters.  Shared by the gfx950 engine (include/storm_gpu.h) and by the
 * FP64 CPU oracle used as the correctness checker.
 *
 * Units are the reference's: CGS, temperature in K, opacities in cm^-1,
 * energies in erg, times in s.
 *
 * What each block stands for in the reference:
 *   storm_mesh_desc      MadCart::CartesianMesh3D(xEdges, yLo, yHi, ny, zLo, zHi, nz)
 *                        as used by examples/marshak_wave/MarshakCommon.hpp:280-290;
 *                        cell index (i*ny + j)*nz + k
 *                        (examples/cartesian_parallel_check/main.cpp:54).
 *   storm_material_desc  RadiationOpacityModel::CalcPlanckOpacity /
 *                        CalcScatteringOpacity (radiation/RadiationOpacityModel.hpp:30-48)
 *                        for the power-law family of
 *                        examples/marshak_wave/MarshakOpacity.hpp, and the
 *                        duck-typed EOS dT2cv / de2T (radiation/RadiationIMC.hpp:848,1648;
 *                        examples/marshak_wave/MarshakCommon.hpp:128-172).
 *   storm_boundary_desc  BoundaryCondition::apply per box face
 *                        (boundary/SideTemperature.hpp:84-112, boundary/Vacuum.hpp:18-22,
 *                        boundary/RigidBoundary.hpp:42-58) and the Planck surface
 *                        source of SideTemperature::generateNewBoundaryParticles
 *                        (boundary/SideTemperature.hpp:114-163).
 *   storm_imc_params     RadiationIMCParameters (radiation/RadiationIMCParameters.hpp:12-56)
 *                        plus CombPopulationControl(grid, Nmin, factor)
 *                        (population/CombPopulationControl.hpp:75-77).
 *   storm_group_desc     multigroup opacity (withMultigroupOpacity, energyBoundaries):
 *                        CalcAbsorptionOpacity(cell, frequency),
 *                        CalcPlanckOpacity and GetThermalEnergy of the
 *                        opacity-weighted Planck family of
 *                        examples/densmore2012/DensmoreOpacity.hpp:38-110, and
 *                        the Planck group sampling of multigroup boundary
 *                        sources (boundary/SideTemperature.hpp:55-80, 150-153).
 */
#ifndef STORM_PROBLEM_H
#define STORM_PROBLEM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Box faces, in this order everywhere: x-, x+, y-, y+, z-, z+. */
enum { STORM_XLO = 0, STORM_XHI = 1, STORM_YLO = 2, STORM_YHI = 3, STORM_ZLO = 4, STORM_ZHI = 5 };

/* What happens to a packet that reaches a box face. */
enum {
  STORM_FACE_REFLECT = 0, /* specular reflection, the packet stays (REFLECT) */
  STORM_FACE_ESCAPE = 1   /* the packet leaves the problem (REMOVE)        */
};

typedef struct storm_mesh_desc {
  int32_t nx, ny, nz, reserved;
  const double *x_edges; /* nx + 1 ascending */
  const double *y_edges; /* ny + 1 ascending */
  const double *z_edges; /* nz + 1 ascending */
} storm_mesh_desc;

/*
 * sigma_a(c, T) = sigma_a0[c] * (max(T, T_floor) / T_ref)^(-a_temp_pow) * rho[c]^a_rho_pow
 * sigma_s(c, T) = sigma_s0[c] * (max(T, T_floor) / T_ref)^(-s_temp_pow) * rho[c]^s_rho_pow
 * EOS, energy per volume   e(T, rho) = eos_f * T^eos_beta * rho^(1 - eos_mu)
 *      temperature         T(e, rho) = (max(e, 0) / (eos_f * rho^(1 - eos_mu)))^(1/eos_beta)
 *      dT2cv (per mass)    eos_f * eos_beta * max(T, cv_T_floor)^(eos_beta - 1) * rho^(1 - eos_mu) / rho
 * The Fleck factor uses dT2cv as returned, exactly as radiation/RadiationIMC.hpp:1648-1649 does.
 */
typedef struct storm_material_desc {
  const double *sigma_a0; /* per cell, required */
  const double *sigma_s0; /* per cell, NULL = no scattering */
  const double *density;  /* per cell, NULL = 1 g/cm^3 */
  double opacity_T_ref;
  double opacity_T_floor;
  double a_temp_pow, a_rho_pow;
  double s_temp_pow, s_rho_pow;
  double eos_f, eos_beta, eos_mu, eos_cv_T_floor;
} storm_material_desc;

typedef struct storm_boundary_desc {
  int32_t face_kind[6];             /* STORM_FACE_* per box face */
  int32_t source_face;              /* -1: none; STORM_XLO: Planck source entering along +x */
  int32_t freeze_source_T4;         /* 1: T^4 taken once at the first step (SideTemperature.hpp:117) */
  double source_temperature;        /* K */
  uint64_t source_packets_per_face; /* Npercell of SideTemperature */
  int32_t source_xhi;               /* 1: a second source on x+ entering along -x, same Npercell
                                       (TwoSidesTemperature, boundary/TwoSidesTemperature.hpp:156-199) */
  int32_t reserved;
  double source_temperature_xhi;    /* K */
} storm_boundary_desc;

typedef struct storm_imc_params {
  uint64_t new_photons_per_cell;      /* RadiationIMCParameters::newPhotonsPerCell */
  int32_t with_random_walk;           /* withRandomWalk */
  int32_t with_ddmc;                  /* withDDMC */
  double rw_min_cell_optical_depth;   /* default 25 */
  double rw_min_particle_optical_depth; /* default 5 */
  double ddmc_min_cell_optical_depth; /* default 15 */
  double ddmc_min_particle_optical_depth; /* default 5 */
  int32_t comb;                       /* 0: NoPopulationControl, 1: CombPopulationControl,
                                         2: StratifiedCombPopulationControl classified by
                                         photon group (multigroup; one group when gray) */
  int32_t ddmc_multigroup_pgrw;       /* ddmcUseMultigroupPGRW (default 0) */
  uint64_t comb_nmin;                 /* Nmin (default 20) */
  double comb_factor;                 /* totalParticlesFactor (default 2.0) */
  uint64_t comb_min_per_group;        /* stratified comb: minParticlesPerGroup (default 2) */
} storm_imc_params;

/*
 * Multigroup (num_groups >= 2; 0 or 1 = gray).  Packets carry a photon
 * energy E (erg) within [bounds[0], bounds[G]].
 *   sigma_a(c, T, E) = sigma_a(c, T) * (max(E, bounds[0]) / nu_ref)^(-nu_pow)
 *                      (sigma_a(c, T) of storm_material_desc; scattering is gray)
 *   Planck opacity   = sum_g sigma_a(c, T, Ec_g) B_g / sum_g B_g, Ec_g the group
 *                      centre, B_g the Planck fraction of group g at
 *                      kT = k max(T, opacity_T_floor)   (1e-20 if sum_g B_g = 0)
 *   emission / effective-scatter energy: group g with probability
 *                      proportional to sigma_a(c, T, Ec_g) B_g, E uniform in the
 *                      group (uniform over the whole range if the sum is 0)
 *   boundary source energy: group g with probability B_g(T_source), E
 *                      uniform in the group.
 * With imc.with_random_walk, packets in the optically thick low groups take
 * the partial-gray random walk (RadiationIMC.hpp:1008-1061, 1065-1276).  With
 * imc.with_ddmc, DDMC runs on the Planck-mean opacity and keeps the photon
 * energy (the reference without ddmcUseMultigroupPGRW, :1343-1350); with
 * imc.ddmc_multigroup_pgrw as well, on the diffusive groups' averages, and
 * only packets below their cutoff energy take DDMC steps (:1301-1342,
 * :1437-1450).
 */
typedef struct storm_group_desc {
  int32_t num_groups;
  int32_t with_eg_time_avg; /* RadiationIMCParameters::withEgTimeAvg: tally the
                               time-averaged radiation energy per cell and group
                               (storm_gpu_get_group_erad) */
  const double *bounds; /* num_groups + 1 ascending photon energies, erg */
  double nu_pow;
  double nu_ref;        /* erg */
} storm_group_desc;

/*
 * A mesh of convex polyhedral cells flattened to per-cell face tables: a
 * Voronoi tessellation (the reference's MadVoro meshes) or any other.  It
 * carries what the reference's Grid concept answers
 * (GetVolume, GetMeshPoint, GetCellFaces, GetFaceNeighbors, Normal, FaceCM,
 * GetArea, GetPointsInFace / GetFacePoints; particle/Particle.hpp:188-279,
 * radiation/RadiationIMC.hpp:1364-1404, utils/RandomInCell.hpp:10-67,
 * utils/RandomOnFace.hpp:14-66).  Faces are stored per cell ("half faces"):
 * a face between two cells appears in each, with opposite normals.  A face on
 * the bounding box has neighbour -1 - side (sides in the STORM_XLO..STORM_ZHI
 * order) and takes that side's storm_boundary_desc.face_kind; Planck sources
 * enter through the faces of side x- (and x+ with source_xhi).
 * Cells of a batch of independent problems (cell_instance set) never share a
 * face; each instance then has its own sources, photon budget and
 * population control (storm_instance_desc).
 */
typedef struct storm_unstructured_mesh {
  uint32_t ncells;
  uint32_t nfaces;                   /* half faces: sum over cells */
  double box_lo[3], box_hi[3];
  const double *volume;              /* ncells, cm^3 */
  const double *mesh_point;          /* 3 ncells: the generator, strictly inside its cell */
  const uint32_t *face_start;        /* ncells + 1: cell c owns faces [face_start[c], face_start[c + 1]) */
  const int32_t *face_neighbor;      /* nfaces: the cell across the face, or -1 - box side */
  const double *face_normal;         /* 3 nfaces: outward unit normal */
  const double *face_centroid;       /* 3 nfaces */
  const double *face_area;           /* nfaces, cm^2 */
  const uint32_t *face_vertex_start; /* nfaces + 1 */
  const double *vertices;            /* 3 per polygon vertex, counter-clockwise seen from outside */
  const uint32_t *cell_instance;     /* ncells, or NULL: one problem */
} storm_unstructured_mesh;

/* Per instance of a batch (storm_unstructured_mesh.cell_instance): what
 * differs between independent problems that share the material laws. */
typedef struct storm_instance_desc {
  double source_temperature;         /* K, on the instance's x- box faces */
  uint64_t source_packets_per_face;
  uint64_t new_photons_per_cell;
  uint64_t comb_nmin;
  double comb_factor;
} storm_instance_desc;

typedef struct storm_problem {
  storm_mesh_desc mesh;              /* ignored when `unstructured` is set */
  storm_material_desc material;
  storm_boundary_desc boundary;
  storm_imc_params imc;
  const double *T0; /* initial material temperature per cell, K */
  uint64_t seed;
  storm_group_desc groups;
  const storm_unstructured_mesh *unstructured; /* NULL: the Cartesian mesh */
  uint32_t num_instances;            /* batches: instances 0 .. num_instances - 1 */
  uint32_t reserved2;
  const storm_instance_desc *instances;
} storm_problem;

#ifdef __cplusplus
}
#endif

#endif /* STORM_PROBLEM_H */
