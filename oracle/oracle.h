/*
 * oracle.h — CPU restatement of gpuATTRACT's pose-scoring path.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker for the HIP library and
 * the CPU baseline of the benchmark; no product code links or calls it.
 *
 * Follows the reference's CPU service (service/CPUEnergyService6D.tpp:62-204,
 * service/CPUEnergyService6DModes.tpp:121-327) and the model headers it calls,
 * in FP32 with the reference's double promotions, plus an FP64 recompute of
 * the same algorithm.  Sine and cosine come from the C library's float
 * functions, as the reference's std::sin/std::cos(float) do.
 */
#ifndef ATTRACT_ORACLE_H_
#define ATTRACT_ORACLE_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Interpolation grid (model/IntrplGrid.h:34-209). data = ngrids*nz*ny*nx float4. */
typedef struct {
    unsigned nx, ny, nz, ngrids;
    float ox, oy, oz, spacing;
    const float* data;
} orc_igrid;

/* Neighbour-list grid (model/NLGrid.h:36-202). desc = {numEl, idx} per voxel. */
typedef struct {
    unsigned nx, ny, nz;
    float ox, oy, oz, spacing, plateau;
    const uint32_t* desc;
    const uint32_t* list;
} orc_nlgrid;

/* Protein (model/Protein.h:38-167).  xyz SoA; modes [x:n*M][y:n*M][z:n*M]. */
typedef struct {
    unsigned n, M;
    const float* xyz;
    const unsigned* type;
    const unsigned* mtype;
    const float* charge;
    const float* modes;
    const float* modeForce;
} orc_body;

/* Pair table (model/ParamTable.h:18-101) + sim params (model/SimParam.h:39-54). */
typedef struct {
    unsigned T;
    int shape; /* 8 or 12 */
    int dielec; /* 0 constant, 1 variable */
    float ffelec;
    const float* rc;
    const float* ac;
    const int* ipon;
    const float* rmin2;
    const float* emin;
} orc_params;

/* Per-bead diagnostics of one pose (all arrays sized by the body). */
typedef struct {
    float* xyz;        /* transformed coordinates, SoA */
    float* pot;        /* per-bead fx,fy,fz,E after grid + NL, SoA [4][n] */
    int* cell;         /* per bead: inner idx xyz, outer idx xyz, NL idx xyz (9 ints), -1 if unused */
    unsigned* nl_pairs; /* per bead: number of listed neighbours inside the plateau */
    float* rec_pot;    /* modes only: receptor forces [3][n_rec] after grid rotate + NL */
    float* rec_xyz;    /* modes only: receptor coordinates in the ligand frame */
} orc_beads;

/* FP32 plateau-ratio table (model/NLGrid.tpp:44-51); size int(1e4*P^2)+1. */
unsigned orc_ratio_size(float plateau);
void orc_ratio_lut(float plateau, float* out);

void orc_euler2rotmat(float phi, float ssi, float rot, float R[9]);
void orc_rotmats(const float* ang /* 3n: phi, ssi, rot */, size_t n, float* R /* 9n */);
void orc_euler2torquemat(float phi, float ssi, float rot, float tm[27]);

/* Rigid scoring: dof = n*6 {pos, ang}, res = n*7 {E, pos, ang} (Result_6D).
 * The _f64 variants optionally return `scale` (same shape as res): the sum of
 * |terms| behind each output component, the denominator of the parity bound. */
int orc_score_rigid(const orc_body* rec, const orc_body* lig, const orc_igrid* inner,
                    const orc_igrid* outer, const orc_nlgrid* nl, const orc_params* par,
                    const float* dof, size_t n, float* res, int nthreads);
int orc_score_rigid_f64(const orc_body* rec, const orc_body* lig, const orc_igrid* inner,
                        const orc_igrid* outer, const orc_nlgrid* nl, const orc_params* par,
                        const float* dof, size_t n, double* res, double* scale, int nthreads);
int orc_beads_rigid(const orc_body* rec, const orc_body* lig, const orc_igrid* inner,
                    const orc_igrid* outer, const orc_nlgrid* nl, const orc_params* par,
                    const float* dof, orc_beads* out);

/* Modes scoring in the reference CPU semantics.  dof = n*26
 * {pos, ang, modesRec[10], modesLig[10]} (DOF_6D_Modes), res = n*27
 * {E, pos, ang, modesRec[10], modesLig[10]} (Result_6D_Modes). */
int orc_score_modes(const orc_body* rec, const orc_body* lig, const orc_igrid* recInner,
                    const orc_igrid* recOuter, const orc_nlgrid* recNL,
                    const orc_igrid* ligInner, const orc_igrid* ligOuter, const orc_params* par,
                    const float* dof, size_t n, float* res, int nthreads);
int orc_score_modes_f64(const orc_body* rec, const orc_body* lig, const orc_igrid* recInner,
                        const orc_igrid* recOuter, const orc_nlgrid* recNL,
                        const orc_igrid* ligInner, const orc_igrid* ligOuter,
                        const orc_params* par, const float* dof, size_t n, double* res,
                        double* scale, int nthreads);
int orc_beads_modes(const orc_body* rec, const orc_body* lig, const orc_igrid* recInner,
                    const orc_igrid* recOuter, const orc_nlgrid* recNL,
                    const orc_igrid* ligInner, const orc_igrid* ligOuter, const orc_params* par,
                    const float* dof, orc_beads* out);

/* Exact gradient of the energy the scoring functions return (FP64; the _f32
 * variant computes in FP32 from bit-identical bead positions):
 *   E = sum over ligand beads of the interpolated grid energy (type grid plus
 *       charge * electrostatic grid) + sum over listed pairs inside the plateau
 *       of [LJ(r) - LJ(P)] + [Coulomb(r) - Coulomb(P)].
 * dof = n*26 (the DOF_6D_Modes layout; amplitudes beyond a body's modes are
 * ignored), w = per-pose upstream weights (NULL = 1).  out = n*27
 * {E, dE/dpos, dE/dang, dE/dmodesRec[10], dE/dmodesLig[10]} (gradients times
 * w); scale (may be NULL, n*27) the sums of |terms| behind each component.
 * pgrad (may be NULL) = T*T*3 {dE/drc, dE/dac, dE/demin} summed over
 * poses with weights; the grid part needs the basis grids (attg_build_basis,
 * 2T floats per grid point [S_1, S6_1, ...][z][y][x]); pass NULL to omit it. */
int orc_adjoint_f64(const orc_body* rec, const orc_body* lig, const orc_igrid* inner,
                    const orc_igrid* outer, const orc_nlgrid* nl, const orc_params* par,
                    const float* inner_basis, const float* outer_basis, const float* dof,
                    size_t n, const double* w, double* out, double* scale, double* pgrad);
int orc_adjoint_f32(const orc_body* rec, const orc_body* lig, const orc_igrid* inner,
                    const orc_igrid* outer, const orc_nlgrid* nl, const orc_params* par,
                    const float* inner_basis, const float* outer_basis, const float* dof,
                    size_t n, const double* w, double* out, double* scale, double* pgrad);

/* VA13 variable-metric minimizer (apps/em/minfor.f, mc11a.f), any dimension.
 * fg(ctx, x, &f, g) evaluates the objective and its gradient.  Reports the
 * last evaluated point and value (as the reference's solver does) and the
 * lowest value seen; nfun = line-search evaluations used (at most ivmax). */
typedef void (*orc_fg)(void* ctx, const double* x, double* f, double* g);
int orc_va13(int n, const double* x0, int ivmax, orc_fg fg, void* ctx, double* xlast,
             double* flast, double* fbest, int* nfun);

/* mc11a alone (packed factors a, n(n+1)/2 values); z and w may alias. */
void orc_va13_update(double* a, int n, double* z, double sig, double* w, int* ir, int mk, double eps);

/* Minimizes n poses (26-float DOF_6D_Modes layout in and out) with VA13 on the
 * FP32 exact adjoint; the state is (angles, position, receptor modes, ligand
 * modes). */
/* The FP32 adjoint evaluated in the HIP kernel's summation order for `waves`
 * (1, 2 or 4) waves per pose; dof = 26 floats per pose, w = float weights or
 * NULL, out = 27 per pose (no parameter gradient).  -1 for other waves. */
int orc_adjoint_kernel_order(const orc_body* rec, const orc_body* lig, const orc_igrid* inner,
                             const orc_igrid* outer, const orc_nlgrid* nl, const orc_params* par,
                             const float* dof, size_t n, const float* w, int waves, double* out);
/* VA13 on that adjoint (the device minimizer's trajectories). */
int orc_minimize_kernel_order(const orc_body* rec, const orc_body* lig, const orc_igrid* inner,
                              const orc_igrid* outer, const orc_nlgrid* nl, const orc_params* par,
                              const float* dof, size_t n, int ivmax, int waves, float* dof_out,
                              double* e_last, double* e_best, int* nfun);
int orc_minimize_f32(const orc_body* rec, const orc_body* lig, const orc_igrid* inner,
                     const orc_igrid* outer, const orc_nlgrid* nl, const orc_params* par,
                     const float* dof, size_t n, int ivmax, float* dof_out, double* e_last,
                     double* e_best, int* nfun);
/* The same with the FP64 adjoint (requests still rounded to float DOFs). */
int orc_minimize_f64(const orc_body* rec, const orc_body* lig, const orc_igrid* inner,
                     const orc_igrid* outer, const orc_nlgrid* nl, const orc_params* par,
                     const float* dof, size_t n, int ivmax, float* dof_out, double* e_last,
                     double* e_best, int* nfun);

#ifdef __cplusplus
}
#endif

#endif
