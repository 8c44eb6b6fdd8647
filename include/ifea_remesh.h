/* ifea_remesh.h — host geometry of the update and remesh steps of the 3D forging solver (no GPU work).
 *
 * The reference remeshes inside Forging3DSolver.RemeshCore (src/ImpressionFEA.Core/ThreeD/Forging3DSolver.cs:1958-2102):
 * a point locator over the old mesh gives the inside test of the workpiece, isosurface stuffing meshes that domain
 * anew, and the same locator remaps element history and nodal velocities onto the new mesh. These entry points are
 * those two tools and the interior smoothing that TryUpdate and RemeshCore both run; the rest of the per-node
 * placement logic (contact, volume compensation, placement repair) stays with the caller.
 *
 * Arrays are caller-owned and read during the call only: node positions x[3 nn] (x, y, z per node), tets[4 ne]
 * (positively oriented node ids), query points p[3 np]. Return 0 on success, -1 for invalid arguments, -6 when the
 * lattice would be too large or memory runs out; ifea_remesh_last_error() gives the message (per thread).
 */
#ifndef IFEA_REMESH_H
#define IFEA_REMESH_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ifea_locator* ifea_locator_t;
typedef struct ifea_stuffing* ifea_stuffing_t;

/* TetLocator(mesh, x, cell) (TetMesh.cs:128-185): a uniform bucket grid of the tets' bounding boxes. The locator
 * keeps its own copy of x and tets. */
int ifea_locator_create(int nn, const double* x, int ne, const int* tets, double cell, ifea_locator_t* out);

/* TetLocator.Find (TetMesh.cs:188-215): per point the first tet of its cell (ascending) whose barycentric weights are
 * all >= -1e-9, or -1; the weights into w[4 np] when w is not NULL. */
int ifea_locator_find(ifea_locator_t h, int np, const double* p, int* tet, double* w);

/* TetLocator.FindNearest (TetMesh.cs:237-262): the containing tet, else the tet of the point's cell whose centroid is
 * nearest, else the nearest centroid over all tets. */
int ifea_locator_find_nearest(ifea_locator_t h, int np, const double* p, int* tet);

/* TetLocator.Interpolate (TetMesh.cs:220-234): linear interpolation of a nodal 3-vector field field[3 nn]; the value at
 * the nearest node when a point lies outside the mesh. Output out[3 np]. */
int ifea_locator_interpolate(ifea_locator_t h, int np, const double* p, const double* field, double* out);

void ifea_locator_destroy(ifea_locator_t h);

/* IsosurfaceStuffing.Stuff (IsosurfaceStuffing.cs:17-88, StuffLattice :115-294) with the locator's Contains as the
 * inside test: a body-centred-cubic lattice of spacing h over [bmin, bmax], cut points by 16 bisections, warping with
 * alpha_long 0.24999 / alpha_short 0.41189, clipped tets coned from their smallest vertex. symX / symY: the model is a
 * quarter or half with symmetry planes x = 0 / y = 0. Query the result with ifea_stuffing_sizes / ifea_stuffing_get. */
int ifea_stuff(ifea_locator_t inside, const double* bmin, const double* bmax, double h, int symX, int symY,
               ifea_stuffing_t* out);
int ifea_stuffing_sizes(ifea_stuffing_t s, int* nn, int* ne);
int ifea_stuffing_get(ifea_stuffing_t s, double* x, int* tets);
void ifea_stuffing_destroy(ifea_stuffing_t s);

/* SmoothInterior (Forging3DSolver.cs:1776-1838) on x[3 nn] in place, for TryUpdate and the remesh: each interior node
 * (is_surface[n] == 0) of a tet below q_target moves towards its neighbours' centroid at a full, half or quarter step
 * when that raises the worst quality of its tets; nodes ascending, up to `sweeps` sweeps, symmetry-plane nodes stay on
 * their planes. node_tet_start[nn+1] / node_tet_list: each node's tets, ascending (TetMesh.cs:18-26). *moves: the
 * number of node moves. */
int ifea_smooth_interior(int nn, int ne, const int* tets, const int* node_tet_start, const int* node_tet_list,
                         const unsigned char* is_surface, const unsigned char* on_x0, const unsigned char* on_y0,
                         double* x, double q_target, int sweeps, int* moves);

const char* ifea_remesh_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* IFEA_REMESH_H */
