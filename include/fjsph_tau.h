/*
 * TAU mesh and solution ingestion for the aerodynamic mesh of the FJSPH
 * step engine: the NetCDF readers of the reference (src/CDFIO.cpp) turned
 * into the flat arrays fjsph_set_mesh takes.
 *
 *   3D: Read_tau_mesh_FACE + Place_Faces (src/CDFIO.cpp:1103-1356):
 *       faces = triangles then quadrilaterals, left/right elements per face.
 *   2D: Read_tau_mesh_EDGE + Place_Edges (src/CDFIO.cpp:830-990): faces are
 *       the element edges; coordinates are the two of points_xc/yc/zc the
 *       file holds (Get_Coordinates, src/CDFIO.cpp:462-628).
 *   Both: Read_SOLUTION (src/CDFIO.cpp:655-822): density, pressure and the
 *       velocity components at the points; in 2D through vertices_in_use
 *       and the offset axis ("2D offset vector": 1 drops x, 2 drops y,
 *       3 drops z).  Cells list their faces in face order.
 *   Read_BMAP / Find_Angle_Alpha (src/CDFIO.cpp:198-316): the angle of
 *   attack and the gravity rotation it implies.
 *
 * The reader is the engine library's own (NetCDF classic, CDF-1/2/5); no
 * NetCDF library is needed.  Host code, no GPU.
 */
#ifndef FJSPH_TAU_H
#define FJSPH_TAU_H

#include "fjsph_engine.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fjsph_tau_mesh fjsph_tau_mesh;

typedef struct fjsph_tau_sizes {
    int32_t dim;
    int32_t ignored_axis;   /* 2D: the coordinate the mesh file lacks (1 x, 2 y, 3 z); 3D: 0 */
    int64_t nverts, nfaces, nface_verts, ncells, ncell_faces;
    int64_t nsurf;          /* boundary faces (right element < 0)                    */
    int64_t n_sol_points;   /* points in the solution file (2D: both planes)          */
} fjsph_tau_sizes;

/* Read a TAU mesh (primary grid: face file in 3D, edge file in 2D) and its
 * solution.  offset_axis: the 2D "2D offset vector" that picks the velocity
 * components (ignored in 3D); scale: "Grid scale" (vertices are
 * multiplied by it).  sol_path may be null: the vertex solution is then
 * zero. */
fjsph_status fjsph_tau_open(const char* mesh_path, const char* sol_path, int32_t dim, int32_t offset_axis,
                            double scale, fjsph_tau_mesh** out);
void fjsph_tau_close(fjsph_tau_mesh* m);
const char* fjsph_tau_last_error(void);
fjsph_status fjsph_tau_get_sizes(const fjsph_tau_mesh* m, fjsph_tau_sizes* out);
/* The arrays in the layout of fjsph_set_mesh (null pointers are skipped):
 * verts nverts*dim, face_off nfaces+1, face_verts nface_verts, face_lr
 * 2*nfaces (left, right; right < 0 on the boundary, as the file has it),
 * cell_off ncells+1, cell_faces ncell_faces, vert_vel nverts*dim, vert_p and
 * vert_rho nverts. */
fjsph_status fjsph_tau_get_arrays(const fjsph_tau_mesh* m, double* verts, int64_t* face_off, int64_t* face_verts,
                                  int64_t* face_lr, int64_t* cell_off, int64_t* cell_faces, double* vert_vel,
                                  double* vert_p, double* vert_rho);
/* Boundary faces in face order and their markers (boundarymarker_of_surfaces;
 * where the file has none, the right element value of the face). */
fjsph_status fjsph_tau_get_surfaces(const fjsph_tau_mesh* m, int64_t* faces, int32_t* markers);
/* fjsph_set_mesh with the mesh's arrays (cell averages of the solution,
 * Average_Point_Data_to_Cell, src/CDFIO.cpp:112-155, are formed there). */
fjsph_status fjsph_set_mesh_tau(fjsph_ctx* ctx, const fjsph_tau_mesh* m);

/* Read_BMAP's angle of attack ("Angle alpha (degree)", in radians on
 * return) and the gravity rotation of Read_BMAP / Find_Angle_Alpha applied
 * to s->grav when s is not null. */
fjsph_status fjsph_tau_read_bmap(const char* bmap_path, fjsph_settings* s, double* angle_alpha);

#ifdef __cplusplus
}
#endif

#endif /* FJSPH_TAU_H */
