/*
 * storm_mesh.h -- polyhedral meshes for the engine of include/storm_gpu.h.
 *
 * The reference runs its IMC path on MadVoro Voronoi meshes through the Grid
 * concept; MadVoro is an external library, so this library builds the
 * tessellation itself and hands it over as the flattened face tables of
 * storm_unstructured_mesh (include/storm_problem.h).  Cells are numbered in Z
 * order of their generators.  No GPU is used.
 */
#ifndef STORM_MESH_H
#define STORM_MESH_H

#include <stdint.h>

#include "storm_problem.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct storm_voronoi storm_voronoi;

/* Voronoi tessellation of `n` distinct generators (x, y, z interleaved)
 * strictly inside the box [lo, hi]: each cell is the box clipped by the
 * bisector half-spaces of its neighbours.  Status 0 or an error code with
 * storm_gpu_last_error(). */
int storm_voronoi_build(const double *points, uint32_t n, const double lo[3], const double hi[3],
                        storm_voronoi **out);

/* The tables of a built tessellation (valid until storm_voronoi_free);
 * cell_instance is left NULL. */
int storm_voronoi_mesh(const storm_voronoi *v, storm_unstructured_mesh *out);

int storm_voronoi_free(storm_voronoi *v);

#ifdef __cplusplus
}
#endif

#endif /* STORM_MESH_H */
