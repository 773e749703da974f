/*
 * oracle.c — CPU restatement of gpuATTRACT's pose-scoring path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library and the CPU
 * baseline of the benchmark.  Only tests/, the smoke check and the benchmark's
 * CPU-baseline leg load it; the product never does.
 *
 * Build: gcc -O3 -DNDEBUG -ffp-contract=off (no -march), as the reference
 * (Makefile:53,63) builds its CPU path; see oracle/Makefile.
 *
 * Parity pins: oracle/ref_harness.cpp compiles the reference's own model
 * headers into oracle/_ref/libattract_ref.so; tests/golden/ holds the vectors
 * it produced, and tests/test_oracle.py checks the restatement against them
 * and against closed forms (rotation/torque-matrix identities, trilinear
 * exactness on affine fields, LJ minimum, plateau continuity, Coulomb cap).
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define REAL float
#define SFX _f32
#include "oracle_impl.inc"
#include "oracle_adj.inc"
#undef REAL
#undef SFX

#define REAL double
#define SFX _f64
#include "oracle_impl.inc"
#include "oracle_adj.inc"
#undef REAL
#undef SFX


/* Exact adjoint (oracle_adj.inc), FP32 and FP64 arithmetic. */
int orc_adjoint_f32(const orc_body* rec, const orc_body* lig, const orc_igrid* inner,
                    const orc_igrid* outer, const orc_nlgrid* nl, const orc_params* par,
                    const float* inner_basis, const float* outer_basis, const float* dof,
                    size_t n, const double* w, double* out, double* scale, double* pgrad) {
    adjoint_f32(rec, lig, inner, outer, nl, par, inner_basis, outer_basis, dof, n, w, out, scale, pgrad);
    return 0;
}

int orc_adjoint_f64(const orc_body* rec, const orc_body* lig, const orc_igrid* inner,
                    const orc_igrid* outer, const orc_nlgrid* nl, const orc_params* par,
                    const float* inner_basis, const float* outer_basis, const float* dof,
                    size_t n, const double* w, double* out, double* scale, double* pgrad) {
    adjoint_f64(rec, lig, inner, outer, nl, par, inner_basis, outer_basis, dof, n, w, out, scale, pgrad);
    return 0;
}

unsigned orc_ratio_size(float plateau) {
    const float P2 = plateau * plateau;
    return (unsigned)((int)(10000 * P2) + 1);
}

void orc_ratio_lut(float plateau, float* out) {
    float P2;
    float* r = make_ratio_f32(plateau, &P2);
    memcpy(out, r, sizeof(float) * orc_ratio_size(plateau));
    free(r);
}

void orc_euler2rotmat(float phi, float ssi, float rot, float R[9]) { rotmat_f32(phi, ssi, rot, R); }

void orc_euler2torquemat(float phi, float ssi, float rot, float tm[27]) {
    torquemat_f32(phi, ssi, rot, tm);
}

#define SETUP(SUF)                                                        \
    ctx##SUF x;                                                           \
    memset(&x, 0, sizeof(x));                                             \
    x.rec = rec; x.lig = lig; x.par = par;

int orc_score_rigid(const orc_body* rec, const orc_body* lig, const orc_igrid* inner,
                    const orc_igrid* outer, const orc_nlgrid* nl, const orc_params* par,
                    const float* dof, size_t n, float* res, int nthreads) {
    SETUP(_f32)
    x.inner = inner; x.outer = outer; x.nl = nl;
    ctx_init_f32(&x);
    run_f32(&x, dof, n, res, 0, nthreads, 0);
    ctx_free_f32(&x);
    return 0;
}

int orc_score_rigid_f64(const orc_body* rec, const orc_body* lig, const orc_igrid* inner,
                        const orc_igrid* outer, const orc_nlgrid* nl, const orc_params* par,
                        const float* dof, size_t n, double* res, double* scale, int nthreads) {
    SETUP(_f64)
    x.inner = inner; x.outer = outer; x.nl = nl;
    ctx_init_f64(&x);
    run_f64(&x, dof, n, res, scale, nthreads, 0);
    ctx_free_f64(&x);
    return 0;
}

int orc_beads_rigid(const orc_body* rec, const orc_body* lig, const orc_igrid* inner,
                    const orc_igrid* outer, const orc_nlgrid* nl, const orc_params* par,
                    const float* dof, orc_beads* out) {
    SETUP(_f32)
    x.inner = inner; x.outer = outer; x.nl = nl;
    ctx_init_f32(&x);
    scratch_f32 s;
    scratch_alloc_f32(&s, lig->n, rec->n, 0);
    float res[7];
    pose_rigid_f32(&x, &s, dof, res, out, 0);
    scratch_free_f32(&s);
    ctx_free_f32(&x);
    return 0;
}

int orc_score_modes(const orc_body* rec, const orc_body* lig, const orc_igrid* recInner,
                    const orc_igrid* recOuter, const orc_nlgrid* recNL,
                    const orc_igrid* ligInner, const orc_igrid* ligOuter, const orc_params* par,
                    const float* dof, size_t n, float* res, int nthreads) {
    SETUP(_f32)
    x.inner = recInner; x.outer = recOuter; x.nl = recNL;
    x.ligInner = ligInner; x.ligOuter = ligOuter;
    ctx_init_f32(&x);
    run_f32(&x, dof, n, res, 0, nthreads, 1);
    ctx_free_f32(&x);
    return 0;
}

int orc_score_modes_f64(const orc_body* rec, const orc_body* lig, const orc_igrid* recInner,
                        const orc_igrid* recOuter, const orc_nlgrid* recNL,
                        const orc_igrid* ligInner, const orc_igrid* ligOuter,
                        const orc_params* par, const float* dof, size_t n, double* res,
                        double* scale, int nthreads) {
    SETUP(_f64)
    x.inner = recInner; x.outer = recOuter; x.nl = recNL;
    x.ligInner = ligInner; x.ligOuter = ligOuter;
    ctx_init_f64(&x);
    run_f64(&x, dof, n, res, scale, nthreads, 1);
    ctx_free_f64(&x);
    return 0;
}

int orc_beads_modes(const orc_body* rec, const orc_body* lig, const orc_igrid* recInner,
                    const orc_igrid* recOuter, const orc_nlgrid* recNL,
                    const orc_igrid* ligInner, const orc_igrid* ligOuter, const orc_params* par,
                    const float* dof, orc_beads* out) {
    SETUP(_f32)
    x.inner = recInner; x.outer = recOuter; x.nl = recNL;
    x.ligInner = ligInner; x.ligOuter = ligOuter;
    ctx_init_f32(&x);
    scratch_f32 s;
    scratch_alloc_f32(&s, lig->n, rec->n, 1);
    float res[27];
    pose_modes_f32(&x, &s, dof, res, out, 0);
    scratch_free_f32(&s);
    ctx_free_f32(&x);
    return 0;
}
