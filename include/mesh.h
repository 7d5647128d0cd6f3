/* Surface quad mesh, drop-in for the reference's include/mesh.h:13-133.
 * Field order and types are identical so a host built against the
 * reference header can pass its Mesh to this library unchanged. */
#ifndef PEEC_MESH_H
#define PEEC_MESH_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct
{
    size_t n_nodes;
    size_t n_quads;
    size_t n_edges;
    double *xyz;             /* 3*n_nodes, metres */
    int *quads;              /* 4*n_quads node indices */
    int *edges;              /* 2*n_edges: current flows edges[2e] -> edges[2e+1] */
    int *cell_edges;         /* 4*n_quads global edge ids */
    int *cell_edge_signs;    /* 4*n_quads, +1 / -1 against the global edge */
    double *edge_vectors;    /* 3*n_edges, r_b - r_a */
    double *edge_lengths;    /* n_edges */
    double *edge_directions; /* 3*n_edges unit vectors */
    double *edge_centers;    /* 3*n_edges */
    double *cell_centers;    /* 3*n_quads */
    double *cell_areas;      /* n_quads, split (0,1,2)+(0,2,3) */
    double *cell_normals;    /* 3*n_quads */
} Mesh;

/* mesh.h:142 */
void mesh_init(Mesh *mesh);
/* mesh.h:148 */
void mesh_free(Mesh *mesh);
/* mesh.h:157; Gmsh 4.x ASCII, quads (type 3) only, node index = $Nodes order */
int mesh_load_gmsh(const char *filename, Mesh *mesh);
/* mesh.h:169; unique edges, oriented and numbered by first appearance in
 * cell order (topology.c:160-304) */
int mesh_build_topology(Mesh *mesh);
/* mesh.h:175 */
void mesh_print_info(const Mesh *mesh);

#ifdef __cplusplus
}
#endif

#endif
