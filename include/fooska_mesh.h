/*
 * fooska_mesh.h -- host-side mesh and case library (C ABI).
 *
 * Builds the unstructured mesh the finite-volume step runs on (linear
 * tetrahedra, pyramids, prisms and hexahedra, the element types of the
 * reference's translator.cu:122-293) and returns it in the reference solver's
 * MESH field semantics:
 *   fc2el[nFcs][2]   owner, neighbour (-1 on boundary faces), owner < neighbour
 *   boundCond[nFcs]  0 interior, 1 symmetry, 2 outlet, 3 inlet, 4 slip, 5 wall
 *   n, nt1, nt2      [nFcs][3] unit normal (outward from the owner) and tangents
 *   area[nFcs], fcCenter[nFcs][3], cellCenter[nElem][3], volume[nElem]
 * plus one array the reference has no counterpart for:
 *   shift[nFcs][3]   periodic translation: the face centre seen from the
 *                    neighbour cell is fcCenter - shift (all zero if no face
 *                    of the mesh is periodic).
 *
 * Replaces (one rank, no METIS):
 *   translator_ptw      mulitphase Solver/doublePrecision/translator.cu:3
 *   meshConnectivity    mulitphase Solver/doublePrecision/meshConnectivity.cu:175
 *   faceConnectivity    mulitphase Solver/doublePrecision/buildLocalMesh.cu:244
 *   meshProperties      mulitphase Solver/doublePrecision/meshProperties.cu:139
 *   initialize/initVec  mulitphase Solver/doublePrecision/initialize.cu:18,61
 * Cells may be renumbered (reverse Cuthill-McKee or Morton order of the
 * centroids, see fkm_box) before faces are built.
 *
 * Every function returns 0 on success and a negative code on error;
 * fkm_last_error() gives the message of the last error on the calling thread.
 */
#ifndef FOOSKA_MESH_H
#define FOOSKA_MESH_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fkm_mesh fkm_mesh;

/* Error codes shared by both libraries. */
#define FK_OK 0
#define FK_EINVAL -1
#define FK_ENOMEM -2
#define FK_EIO -3
#define FK_EDEVICE -4
#define FK_ESTATE -5
#define FK_ERANGE -6

const char *fkm_last_error(void);

/* Box [x0,x1]x[y0,y1]x[z0,z1] as nx*ny*nz hexes, each split into 6 tets that
 * share the hex diagonal.  Interior nodes (every node along a periodic axis)
 * are moved by a uniform random offset in [-jitter, +jitter] * spacing per
 * coordinate, drawn from splitmix64(seed).  periodic_mask bit 0/1/2 = x/y/z.
 * bc[6] = boundary code of the faces at x0, x1, y0, y1, z0, z1.
 * order: 0 = generation order (hex-major), 1 = reverse Cuthill-McKee,
 * 2 = Morton (Z-order) of the cell centroids. */
int fkm_box(int nx, int ny, int nz, const double lo[3], const double hi[3],
            double jitter, uint64_t seed, int periodic_mask, const int bc[6],
            int order, fkm_mesh **out);

/* As fkm_box without periodic axes, with the element kind of every hex
 * (kinds[nz][ny][nx]): 0 six tets sharing the hex diagonal, 1 one hexahedron,
 * 2 two prisms along z (split through hex vertices 0, 3, 4, 7), 3 a centre
 * node with a pyramid on every quadrilateral face and two tets on every face
 * that meets triangles.  Neighbouring hexes must present the same face shape
 * (kind 3 adapts to its neighbours); FK_EINVAL otherwise.  No reference
 * counterpart (the reference reads such meshes from Gambit files). */
int fkm_box_kinds(int nx, int ny, int nz, const double lo[3], const double hi[3],
                  double jitter, uint64_t seed, const int bc[6], int order,
                  const int32_t *kinds, fkm_mesh **out);

/* Gambit neutral file, node coordinates multiplied by scale; order as for
 * fkm_box.  Elements by node count as translator.cu:122-293: 4 tetrahedron,
 * 5 pyramid, 6 prism, 8 hexahedron, in Gambit's node order.  BC tags
 * 45/26/13/42/51 map to codes 1/2/3/4/5 and the element face number of a
 * boundary entry is the local face + 1, as in translator.cu:295-349. */
int fkm_read_neu(const char *path, double scale, int order, fkm_mesh **out);
/* Writes the mesh (generation-order-independent) as a Gambit neutral file. */
int fkm_write_neu(const fkm_mesh *m, const char *path);

void fkm_free(fkm_mesh *m);

/* Sizes: cells, faces, nodes, boundary faces. */
int fkm_sizes(const fkm_mesh *m, int64_t *n_cells, int64_t *n_faces,
              int64_t *n_nodes, int64_t *n_bfaces);

/* Borrowed pointer to an internal array (valid until fkm_free).  Names:
 *  int32:  "fc2el" [F][2], "bc" [F], "el2nd" [C][8] (-1 past the element's
 *          nodes), "nnode" [C] (4, 5, 6 or 8)
 *  double: "n" "nt1" "nt2" "fc_center" "shift" [F][3], "area" [F],
 *          "cell_center" [C][3], "volume" [C], "nodes" [N][3] */
int fkm_array(const fkm_mesh *m, const char *name, void **ptr);

/* ---- initial conditions (replaces initialize.cu / SP initialize.cu) ----
 * U is written in the reference's AoS layout U[cell][NEQ]. */

/* Single-phase (NEQ = 5, U = [rho, rho u, rho v, rho w, rho E]); ideal or
 * stiffened gas (gamma, pinf).  Cells with centre x < x_split get the left
 * state (rho, u, v, w, p), the others the right state. */
int fkm_init_sod(const fkm_mesh *m, double gamma, double pinf, double x_split,
                 const double left[5], const double right[5], double *U);

/* Taylor-Green vortex on [0,2pi]^3 with speed U0, density rho0, Mach number
 * mach (p0 = rho0 U0^2 / (gamma Ma^2)). */
int fkm_init_tgv(const fkm_mesh *m, double gamma, double rho0, double U0,
                 double mach, double *U);

/* Two-phase (NEQ = 10) regions.  Each region gives the liquid volume
 * fraction alpha_L, pressure p, temperature T and velocity; both phases are
 * set to (p, T) through rho_k(p, T), e_k(p, rho_k) of their stiffened-gas
 * EOS, as initVec does (initialize.cu:37-54). */
typedef struct fkm_phase_eos {
  double gamma[2], pinf[2], cp[2], eta[2];
} fkm_phase_eos;
typedef struct fkm_region {
  int kind;            /* 0 background, 1 sphere, 2 half-space x < x0 */
  double c[3];         /* sphere centre / c[0] = plane position */
  double r;            /* sphere radius */
  double alphaL, p, T; /* liquid fraction, pressure, temperature */
  double u[3];
} fkm_region;
/* Regions are applied in order; a later region overrides an earlier one on
 * the cells it contains.  region[0] must be the background (kind 0). */
int fkm_init_regions(const fkm_mesh *m, const fkm_phase_eos *eos,
                     const fkm_region *regions, int n_regions, double *U);

/* Seeded non-overlapping sphere placement inside [lo,hi] (rejection
 * sampling, splitmix64(seed)); radii uniform in [rmin, rmax]; centres
 * restricted to x >= xmin.  Writes centres[n][3], radii[n]. */
int fkm_place_bubbles(int n, const double lo[3], const double hi[3],
                      double rmin, double rmax, double xmin, uint64_t seed,
                      double *centres, double *radii);

#ifdef __cplusplus
}
#endif
#endif
