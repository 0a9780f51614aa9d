/*
 * fe_mesh.h -- host-side C ABI of libfe_host.so: meshes, edges, DOF numbering,
 * element coloring and synthetic coefficient fields for the feNG
 * assemble-and-solve path.
 *
 * All entry points return an int status with the reference's feStatus values
 * (feMessage.h:44-62): 0 = FE_STATUS_OK, -1 = FE_STATUS_ERROR, ...
 * Arrays are caller-owned, 0-based, int32 indices and float64 coordinates.
 * Handles are opaque; one thread per handle.
 */
#ifndef FE_MESH_H
#define FE_MESH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes: the values of the reference's feStatus (feMessage.h:44-62).
 * A translation unit that includes the reference's headers first gets the
 * reference's enumerators (same values) and no macros; include these headers
 * after the reference's, since the macros would rename its enumerators. */
#ifndef _FEMESSAGE_
#define FE_STATUS_OK 0
#define FE_STATUS_ERROR (-1)
#define FE_STATUS_FAILED (-2)
#define FE_STATUS_ASSERTION_FAILED (-3)
#define FE_STATUS_OUT_OF_MEMORY (-4)
#define FE_STATUS_FILE_CANNOT_BE_OPENED (-5)
#define FE_STATUS_POINTER_ERROR (-6)
#define FE_STATUS_READ_ERROR (-7)
#define FE_STATUS_WRITE_ERROR (-8)
#define FE_STATUS_RANGE_ERROR (-9)
#define FE_STATUS_FORMAT_ERROR (-10)
#endif

/* Element families: the local layouts of feSpaceTriP1 (feSpace_2D.cpp:154-159),
 * feSpaceTriP2 (:772-786), feSpaceVecTriP2<2> (:1017-1041), feSpaceTetPn
 * with n = 1 (feSpace_3D.cpp:293-345), feSpaceTriP3 (feSpace_2D.cpp:1315-1327,
 * :1450-1480: vertices, two DOFs per edge, one element DOF) and feSpaceTriP4
 * (:1700-1715, :1876-1912: vertices, three per edge, three element DOFs). */
#define FE_TRI_P1 1
#define FE_TRI_P2 2
#define FE_TRI_VP2 3
#define FE_TET_P1 4
#define FE_TRI_P3 5
#define FE_TRI_P4 6

typedef struct fe_mesh fe_mesh;
typedef struct fe_numbering fe_numbering;

/* ---- meshes ------------------------------------------------------------- */

/* Unit square, N x N cells, each split along its (0,0)-(1,1) diagonal into
 * (v00,v10,v11) and (v00,v11,v01); vertex id = i + (N+1) j; interior vertices
 * moved by uniform(-jitter*h, +jitter*h) per coordinate from a SplitMix64
 * stream seeded with `seed` (jitter = 0: no motion).  The boundary ("Bord")
 * is the counter-clockwise loop of 4N segments starting at (0,0).
 * Replaces reading a Gmsh file with feMesh2DP1 (feMesh.cpp:434,
 * feMeshRead.cpp:645-1622) for generated inputs. */
int fe_mesh_square(int N, double jitter, uint64_t seed, fe_mesh **out);

/* Unit cube, N^3 hexes, each split into 6 Kuhn tetrahedra along the
 * (0,0,0)-(1,1,1) diagonal, positively oriented; vertex id = i+(N+1)(j+(N+1)k).
 * The boundary is the 12 N^2 triangles of the 6 faces. */
int fe_mesh_cube(int N, fe_mesh **out);

/* Mesh from caller arrays: dim 2 (triangles, nvpe 3, boundary segments bnv 2)
 * or dim 3 (tets, nvpe 4, boundary triangles bnv 3).  Vertex ids follow the
 * order of xyz (feMeshRead.cpp:933-958); elements and boundary elements keep
 * the given order. */
int fe_mesh_from_arrays(int dim, int V, const double *xyz, int E, int nvpe, const int *conn,
                        int nB, int bnv, const int *bconn, fe_mesh **out);

/* Gmsh .msh v4.1 ASCII writer: physical "Bord" (boundary, dim-1) and "Domaine"
 * (dim), boundary elements written first.  feMeshRead.cpp:645-1622 reads it
 * back with vertex ids = file order, element order = file order. */
int fe_mesh_write_msh(const fe_mesh *m, const char *path);
/* Gmsh .msh v4.1 ASCII reader (feMeshRead.cpp:1709-2005): the domain is every
 * element of the highest dimension in a named physical group, the boundary
 * every element one dimension below in a named physical group, each tagged
 * with its group (fe_mesh_boundary_groups). */
int fe_mesh_read_msh(const char *path, fe_mesh **out);

int fe_mesh_destroy(fe_mesh *m);

/* Sizes: dim, V, E, nvpe, nB, bnv, nEdges (2D: unique triangle edges; 3D: 0). */
int fe_mesh_info(const fe_mesh *m, int *dim, int *V, int *E, int *nvpe, int *nB, int *bnv,
                 int *nEdges);

/* Copy out arrays (any pointer may be NULL): xyz[3V], conn[E nvpe],
 * bconn[nB bnv], edges[3E] (2D: signed 1-based tags by first encounter,
 * feMeshRead.cpp:1412-1456), bedge[nB] (2D: edge index of each boundary
 * segment, feMeshRead.cpp:1836-1869). */
int fe_mesh_get(const fe_mesh *m, double *xyz, int *conn, int *bconn, int *edges, int *bedge);

/* ---- numbering ---------------------------------------------------------- */

/* DOF numbering of feMetaNumber (feNumber.cpp:677-843, numberUnknowns
 * :370-426, numberEssential :428-483) for nfields fields of kind FE_TRI_P1,
 * FE_TRI_P2, FE_TRI_P3, FE_TRI_P4, FE_TRI_VP2 or FE_TET_P1 defined on the whole mesh; essential[f] != 0 makes the field's
 * DOFs on the boundary essential (a Dirichlet "Bord" space).  Fields are
 * numbered in the given order (first appearance in the reference's space
 * vector). */
int fe_numbering_create(const fe_mesh *m, int nfields, const int *kinds, const int *essential,
                        fe_numbering **out);
/* fe_numbering_create with the essential boundary of each field given per
 * boundary physical group: bit g of essentialGroups[f] makes the DOFs of
 * field f on the boundary elements of group g essential (the reference's
 * essential spaces, one per field and physical group, feNumber.cpp:677-843);
 * fe_numbering_create is this with every group for essential[f] != 0. */
int fe_numbering_create_groups(const fe_mesh *m, int nfields, const int *kinds, const unsigned *essentialGroups,
                               fe_numbering **out);
/* Boundary physical groups of a mesh (dimension dim - 1; a mesh read from a
 * .msh file has its named groups in physical-tag order, a generated mesh the
 * one group "Bord"): their count, their names separated by '\n' in names
 * (capacity len, NULL: not wanted) and the group of each boundary element in
 * bgroup[nB] (NULL: not wanted). */
int fe_mesh_boundary_groups(const fe_mesh *m, int *ngroups, char *names, int len, int *bgroup);
/* Orientation check the reference runs when it builds a space on a mesh
 * (feCncGeo::computeJacobians, feCncGeo.cpp:346-353, :388-395, :410-412):
 * FE_STATUS_ERROR when any element has a Jacobian determinant <= 0, with the
 * first such element and its determinant in *elem / *det (either may be
 * NULL); FE_STATUS_OK and *elem = -1 otherwise.  fe_numbering_create and
 * fe_hip_upload_mesh refuse such meshes the same way. */
int fe_mesh_check_jacobians(const fe_mesh *m, int *elem, double *det);
/* feNumber::applyPeriodicity (feNumber.cpp:217-307): the (master, slave)
 * DOF pairs of field `field` between boundary groups masterGroup and
 * slaveGroup (indices of fe_mesh_boundary_groups) of a triangle mesh: every
 * master-group vertex whose translate by offset[3] is a slave-group vertex
 * (distance < 1e-8) pairs its DOFs with that vertex's, every master-group edge
 * whose two translated vertices are a slave-group edge's pairs its edge DOFs
 * (components in order; for P3/P4 the positions along the edge that match
 * after the translation).  Pairs sorted by master DOF, one per master (the
 * reference's std::map).  *count = the number of pairs; master / slave
 * (cap entries) filled when both are non-NULL (FE_STATUS_ERROR when cap is
 * smaller).  FE_STATUS_ERROR when nothing matches, as the reference's error
 * exit.  Feed the pairs to fe_hip_set_periodic_dofs. */
int fe_numbering_periodic(const fe_mesh *m, const fe_numbering *num, int field, int masterGroup, int slaveGroup,
                          const double *offset, int *master, int *slave, int cap, int *count);
int fe_numbering_destroy(fe_numbering *num);
/* nInc (unknowns, the system size) and nDofs (unknowns + essential). */
int fe_numbering_info(const fe_numbering *num, int *nInc, int *nDofs);
/* Element -> global DOF map of field f: E * nloc ints
 * (initializeAddressingVector order). */
int fe_numbering_adr(const fe_numbering *num, int field, int *adr);
/* Physical coordinates of every DOF (3 nDofs doubles) and its field index
 * (nDofs ints) and vector component (nDofs ints) -- the points at which
 * feSolution::initializeEssentialBC (feSolution.cpp:246-377) evaluates the
 * boundary callbacks. */
int fe_numbering_dof_info(const fe_numbering *num, double *xyz, int *field, int *component);

/* ---- coloring ----------------------------------------------------------- */

/* Element coloring with no two elements of a color sharing a vertex
 * (the property feCncGeo::colorElements, feCncGeo.cpp:735-794, provides).
 * colorOf[E]; returns the number of colors in *ncolors. */
int fe_color_elements(const fe_mesh *m, int *colorOf, int *ncolors);

/* ---- synthetic coefficient fields --------------------------------------- */

/* Piecewise-constant field on B x B blocks of the unit square, log-uniform in
 * [lo, hi] (SplitMix64 stream from seed), evaluated per element at its
 * centroid: k[E]. */
int fe_coeff_blocks(const fe_mesh *m, int B, double lo, double hi, uint64_t seed, double *k);

/* k = exp(g) per tetrahedron at its centroid, g a Gaussian random field with
 * squared-exponential covariance sigma^2 exp(-|r|^2 / (2 ell^2)) sampled on
 * the (N+1)^3 vertex grid by a seeded random-Fourier-feature expansion with
 * nfeat features, interpolated linearly to the centroid. */
int fe_coeff_grf_cube(const fe_mesh *m, double ell, double sigma, int nfeat, uint64_t seed,
                      double *k);

#ifdef __cplusplus
}
#endif

#endif
