/* Oracle-internal mesh layout (mesh.rs:19-59).  TEST INFRASTRUCTURE ONLY. */
#ifndef OR_MESH_INTERNAL_H
#define OR_MESH_INTERNAL_H
#include "or_vec.h"
#include "oracle.h"

typedef struct {
    v3 a, e1, e2;
    uint32_t tri;
    uint16_t material;
} or_prim;

typedef struct {
    v3 bmin, bmax;
    uint32_t a, b;
} or_node;

struct or_mesh {
    uint32_t nv, nt, nm, nbands;
    v3 *positions;
    uint32_t *indices;
    uint16_t *tri_material;
    uint16_t *tri_surface;
    or_material *materials;
    v3 *normals;
    or_node *nodes;
    uint32_t n_nodes, cap_nodes;
    or_prim *packed;
    uint32_t n_prims;
    uint32_t max_depth;
};

int or_ray_tri(v3 a, v3 e1, v3 e2, v3 o, v3 d, float *t_out, float *u_out, float *v_out);

#endif
