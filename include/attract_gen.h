/*
 * attract_gen.h — C ABI of the seeded synthetic-workload generator (host code).
 *
 * The reference reads proteins, pair tables, grids and poses from files
 * (fileIO/readFile.cpp) and has no grid generator at all (it only reads
 * ATTRACT grid binaries, fileIO/grid_orig.cpp:85-301).  This library builds the
 * same in-memory objects from a seed so that every consumer (the HIP library's
 * tests and benchmark, the CPU oracle, the compiled-reference harness) sees
 * bit-identical inputs.  Everything here is deterministic for a given seed and
 * independent of the thread count.
 *
 * Layouts follow the reference's data model:
 *   body coordinates   SoA [x: n][y: n][z: n]                (model/Protein.h:55-65)
 *   mode vectors       [x: n*M][y: n*M][z: n*M], index i*M+m (model/Protein.h:88-98)
 *   potential grids    float4 {fx,fy,fz,E}, [grid][z][y][x], grid 0 = electrostatics
 *                      (model/IntrplGrid.h:185-198, fileIO/readFile.cpp:308-319)
 *   neighbour grid     uint2 {numEl, idx} per voxel [z][y][x] + concatenated u32 list
 *                      (model/NLGrid.h:24-28, 150-186)
 *   pair table         numTypes^2 entries, index numTypes*(recType-1)+(ligType-1)
 *                      (model/ParamTable.h:80-82, model/neighborlist.h:108)
 *   DOF                {pos xyz, ang phi ssi rot} 6 floats   (model/Types_6D.h:31-37)
 */
#ifndef ATTRACT_GEN_H_
#define ATTRACT_GEN_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Entity ids for independent random substreams of one master seed. */
enum {
    ATTG_ENT_RECEPTOR = 0,
    ATTG_ENT_LIGAND = 1,
    ATTG_ENT_TYPES = 2,
    ATTG_ENT_TABLE = 3,
    ATTG_ENT_POSES = 4,
    ATTG_ENT_MODES = 5,
    ATTG_ENT_PAIRS = 6,
    ATTG_ENT_ARRIVALS = 7
};

/* Geometry of one regular grid: number of points per axis, position of point
 * (0,0,0) and spacing.  Matches Grid<REAL> (model/Grid.h:32-80). */
typedef struct {
    unsigned nx, ny, nz;
    float ox, oy, oz;
    float spacing;
} attg_geom;

/* Uniform double in [0,1) from the substream (seed, entity) at position `counter`. */
double attg_uniform(uint64_t seed, unsigned entity, uint64_t counter);

/* Beads placed uniformly in a sphere of `radius` (by rejection) with pairwise
 * distance >= min_spacing, centred so that their centroid is the origin.
 * Types uniform in 1..num_types (a fraction `type0_frac` gets type 0, which the
 * reference treats as non-interacting), a fraction `charged_frac` gets charge
 * +1 or -1.  Returns 0, or -1 when the sphere cannot hold the beads. */
int attg_make_body(uint64_t seed, unsigned entity, unsigned n, double radius,
                   double min_spacing, unsigned num_types, double charged_frac,
                   double type0_frac, float* xyz, unsigned* type, float* charge);

/* Raw ATTRACT pair parameters as read from attract.par (rbc, abc, ipon), built
 * from rmin ~ U(rmin_lo, rmin_hi), emin ~ U(emin_lo, emin_hi) and ipon = -1 with
 * probability repulsive_frac.  Symmetric in the two types. */
void attg_make_raw_params(uint64_t seed, unsigned num_types, double rmin_lo, double rmin_hi,
                          double emin_lo, double emin_hi, double repulsive_frac,
                          float* rbc, float* abc, int* ipon);

/* Derived table {rc, ac, ipon, rmin2, emin} per type pair for the 8-6 (shape=8)
 * or 12-6 (shape=12) potential, restating fileIO/readFile.cpp:479-513. */
int attg_derive_params(unsigned num_types, int shape, const float* rbc, const float* abc,
                       const int* ipon_in, float* rc, float* ac, int* ipon, float* rmin2,
                       float* emin);

/* Inner-grid geometry: bounding box of the beads plus `margin` on each side. */
void attg_inner_geom(const float* xyz, unsigned n, float spacing, float margin, attg_geom* g);

/* Cubic grid of half-width `halfwidth` centred at (cx,cy,cz). */
void attg_cube_geom(float cx, float cy, float cz, float halfwidth, float spacing, attg_geom* g);

/* Potential grids for `num_types` probe types plus electrostatics (grid 0).
 * Source beads are the body the grid belongs to.  For every grid point p and
 * probe type t the value is the sum over source beads j of the pair potential
 * at r' = max(|p - x_j|, plateau) and the force -dV/dr(r') along (p - x_j)/r,
 * which is what the neighbour-list term subtracts (model/neighborlist.h:121-157),
 * so grid + neighbour list reproduce the pair potential.  Accumulated in double.
 * src_is_ligand selects the table index convention: receptor grid probes are
 * ligand types (table[(src-1)*T + (t-1)]), ligand grid probes are receptor types
 * (table[(t-1)*T + (src-1)]).  dielec: 0 constant, 1 variable.
 * out holds (num_types+1)*nx*ny*nz float4 (16 bytes each). */
int attg_build_grid(const attg_geom* g, const float* xyz, const unsigned* type,
                    const float* charge, unsigned n, unsigned num_types, int shape,
                    const float* rc, const float* ac, const int* ipon, const float* rmin2,
                    const float* emin, int src_is_ligand, float plateau, int dielec,
                    float ffelec, float* out, int nthreads);

/* Basis grids of the pair-parameter derivative: for every source type a
 * (1..num_types) the energy sums S_a = sum_j 1/r'^s (s = 8 or 12) and
 * S6_a = sum_j 1/r'^6 over source beads of type a, r' = max(r, plateau), so
 * that a probe grid's energy is sum_a ipon(rc S_a - ac S6_a) whenever every
 * rmin^2 <= plateau^2.  out holds 2*num_types*nz*ny*nx floats, ordered
 * [S_1, S6_1, S_2, S6_2, ...][z][y][x].  Returns -1 when some pair has
 * rmin^2 > plateau^2 (the grid is then not linear in rc, ac). */
int attg_build_basis(const attg_geom* g, const float* xyz, const unsigned* type, unsigned n,
                     unsigned num_types, int shape, const float* rmin2, float plateau,
                     float* out, int nthreads);

/* Neighbour lists over the voxels of `g`: source beads with type != 0 within
 * `neighbourdis` of the voxel point, ascending bead index.  Call with list ==
 * NULL to get the total count in *total (desc is filled either way). */
int attg_build_nl(const attg_geom* g, const float* xyz, const unsigned* type, unsigned n,
                  float neighbourdis, uint32_t* desc, uint32_t* list, uint64_t* total,
                  int nthreads);

/* Poses: Euler angles Haar-distributed for the reference's ZYZ form
 * (phi, rot ~ U[0,2pi), ssi = acos U[-1,1]), translation uniform in direction
 * with radius ~ U[r_lo, r_hi] about (cx,cy,cz).  dof holds n*6 floats. */
void attg_make_poses(uint64_t seed, unsigned entity, uint64_t first, uint64_t n, double r_lo,
                     double r_hi, float cx, float cy, float cz, float* dof);

/* n*M Gaussian amplitudes with standard deviation sigma. */
void attg_make_amplitudes(uint64_t seed, unsigned entity, uint64_t first, uint64_t n,
                          unsigned M, double sigma, float* out);

/* M orthonormal random 3n-vectors in the reference mode layout, and the
 * per-mode force constants k_m = h_m^2 with h_m ~ U(h_lo, h_hi)
 * (fileIO/readFile.cpp:218-221 squares the mode-file value). */
void attg_make_modes(uint64_t seed, unsigned entity, unsigned n, unsigned M, double h_lo,
                     double h_hi, float* modes, float* force_const);

#ifdef __cplusplus
}
#endif

#endif /* ATTRACT_GEN_H_ */
