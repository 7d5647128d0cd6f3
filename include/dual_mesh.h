/* Dual regions of the quad mesh, drop-in for the reference's
 * include/dual_mesh.h:25-121 (identical layouts). */
#ifndef PEEC_DUAL_MESH_H
#define PEEC_DUAL_MESH_H

#include "mesh.h"

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One quadrilateral piece of a dual region: 4 corners x (x,y,z). */
typedef struct
{
    double points[12];
    int owner_id; /* node or edge that owns the piece */
    int cell_id;  /* source quad */
} QuadPatch;

typedef struct
{
    QuadPatch *patches;
    size_t count;
    size_t capacity;
} PatchList;

typedef struct
{
    size_t n_nodes;
    size_t n_edges;
    PatchList *node_patches;     /* quarter-quads around each node */
    PatchList *edge_patches;     /* half-quads along each edge */
    double *node_region_areas;
    double *edge_region_areas;
    double *node_region_centers;
    double *edge_region_centers;
} DualMesh;

/* dual_mesh.h:95-121 */
void dual_mesh_init(DualMesh *dual);
void dual_mesh_free(DualMesh *dual);
int dual_mesh_build(const Mesh *mesh, DualMesh *dual);
void dual_mesh_print_info(const DualMesh *dual);

#ifdef __cplusplus
}
#endif

#endif
