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

/* Sine and cosine as model/matrixFunctions.h:51-60 evaluates them: the C
 * library's single-precision functions in the FP32 build (glibc turns each
 * pair into one sincosf call with the same values), double in the FP64
 * recompute. */
#define REAL float
#define SFX _f32
#define RCOS(x) cosf(x)
#define RSIN(x) sinf(x)
#include "oracle_impl.inc"
#include "oracle_adj.inc"
#undef REAL
#undef SFX
#undef RCOS
#undef RSIN
#include "oracle_kadj.inc"

#define REAL double
#define SFX _f64
#define RCOS(x) cos(x)
#define RSIN(x) sin(x)
#include "oracle_impl.inc"
#include "oracle_adj.inc"
#undef REAL
#undef SFX
#undef RCOS
#undef RSIN


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

/* The FP32 adjoint in the HIP kernel's summation order (oracle_kadj.inc) for
 * `waves` waves per pose; w (float weights) may be NULL. */
int orc_adjoint_kernel_order(const orc_body* rec, const orc_body* lig, const orc_igrid* inner,
                             const orc_igrid* outer, const orc_nlgrid* nl, const orc_params* par,
                             const float* dof, size_t n, const float* w, int waves, double* out) {
    if (waves != 1 && waves != 2 && waves != 4) return -1;
    ctx_f32 x;
    memset(&x, 0, sizeof(x));
    x.rec = rec; x.lig = lig; x.par = par;
    x.inner = inner; x.outer = outer; x.nl = nl;
    ctx_init_f32(&x);
    for (size_t p = 0; p < n; ++p) kadj_pose(&x, dof + 26 * p, w ? w[p] : 1.f, waves, out + 27 * p);
    ctx_free_f32(&x);
    return 0;
}

/* Pose minimization with VA13 (oracle_va13.c) on the FP32 adjoint: the state
 * is (angles, position, receptor modes, ligand modes) in double
 * (apps/em/meta.h:90-170); each request is rounded to float DOFs, as the
 * reference builds its DOF_6D<float> requests, and scored by the FP32 exact
 * adjoint (energy and dE/dstate). */
typedef struct {
    const ctx_f32* x32;
    const ctx_f64* x64;
    int waves;  /* > 0: the FP32 adjoint in the kernel's order for this many waves */
    unsigned Mr, Ml;
    float dof[26];
    double res[27];
} min_ctx;

static void min_fg(void* p, const double* s, double* f, double* g) {
    min_ctx* m = (min_ctx*)p;
    memset(m->dof, 0, sizeof(m->dof));
    for (int k = 0; k < 3; ++k) {
        m->dof[3 + k] = (float)s[k];  /* angles */
        m->dof[k] = (float)s[3 + k];  /* position */
    }
    for (unsigned k = 0; k < m->Mr; ++k) m->dof[6 + k] = (float)s[6 + k];
    for (unsigned k = 0; k < m->Ml; ++k) m->dof[16 + k] = (float)s[6 + m->Mr + k];
    if (m->x32 && m->waves > 0) kadj_pose(m->x32, m->dof, 1.f, m->waves, m->res);
    else if (m->x32) adj_pose_f32(m->x32, 0, 0, m->dof, 1.0, m->res, 0, 0);
    else adj_pose_f64(m->x64, 0, 0, m->dof, 1.0, m->res, 0, 0);
    *f = m->res[0];
    for (int k = 0; k < 3; ++k) {
        g[k] = m->res[4 + k];
        g[3 + k] = m->res[1 + k];
    }
    for (unsigned k = 0; k < m->Mr; ++k) g[6 + k] = m->res[7 + k];
    for (unsigned k = 0; k < m->Ml; ++k) g[6 + m->Mr + k] = m->res[17 + k];
}

static void minimize_poses(min_ctx* m, const float* dof, size_t n, int ivmax, float* dof_out,
                           double* e_last, double* e_best, int* nfun) {
    const int dim = 6 + (int)(m->Mr + m->Ml);
    double s0[26], sl[26];
    for (size_t p = 0; p < n; ++p) {
        const float* d = dof + 26 * p;
        for (int k = 0; k < 3; ++k) {
            s0[k] = d[3 + k];
            s0[3 + k] = d[k];
        }
        for (unsigned k = 0; k < m->Mr; ++k) s0[6 + k] = d[6 + k];
        for (unsigned k = 0; k < m->Ml; ++k) s0[6 + m->Mr + k] = d[16 + k];
        double fl, fb;
        int nf;
        orc_va13(dim, s0, ivmax, min_fg, m, sl, &fl, &fb, &nf);
        float* o = dof_out + 26 * p;
        memset(o, 0, sizeof(float) * 26);
        for (int k = 0; k < 3; ++k) {
            o[3 + k] = (float)sl[k];
            o[k] = (float)sl[3 + k];
        }
        for (unsigned k = 0; k < m->Mr; ++k) o[6 + k] = (float)sl[6 + k];
        for (unsigned k = 0; k < m->Ml; ++k) o[16 + k] = (float)sl[6 + m->Mr + k];
        if (e_last) e_last[p] = fl;
        if (e_best) e_best[p] = fb;
        if (nfun) nfun[p] = nf;
    }
}

int orc_minimize_f32(const orc_body* rec, const orc_body* lig, const orc_igrid* inner,
                     const orc_igrid* outer, const orc_nlgrid* nl, const orc_params* par,
                     const float* dof, size_t n, int ivmax, float* dof_out, double* e_last,
                     double* e_best, int* nfun) {
    ctx_f32 x;
    memset(&x, 0, sizeof(x));
    x.rec = rec; x.lig = lig; x.par = par;
    x.inner = inner; x.outer = outer; x.nl = nl;
    ctx_init_f32(&x);
    min_ctx m;
    memset(&m, 0, sizeof(m));
    m.x32 = &x;
    m.Mr = rec->M;
    m.Ml = lig->M;
    minimize_poses(&m, dof, n, ivmax, dof_out, e_last, e_best, nfun);
    ctx_free_f32(&x);
    return 0;
}

/* The same on the kernel-order FP32 adjoint for `waves` waves per pose: the
 * trajectories of the device minimizer, bit for bit. */
int orc_minimize_kernel_order(const orc_body* rec, const orc_body* lig, const orc_igrid* inner,
                              const orc_igrid* outer, const orc_nlgrid* nl, const orc_params* par,
                              const float* dof, size_t n, int ivmax, int waves, float* dof_out,
                              double* e_last, double* e_best, int* nfun) {
    if (waves != 1 && waves != 2 && waves != 4) return -1;
    ctx_f32 x;
    memset(&x, 0, sizeof(x));
    x.rec = rec; x.lig = lig; x.par = par;
    x.inner = inner; x.outer = outer; x.nl = nl;
    ctx_init_f32(&x);
    min_ctx m;
    memset(&m, 0, sizeof(m));
    m.x32 = &x;
    m.waves = waves;
    m.Mr = rec->M;
    m.Ml = lig->M;
    minimize_poses(&m, dof, n, ivmax, dof_out, e_last, e_best, nfun);
    ctx_free_f32(&x);
    return 0;
}

int orc_minimize_f64(const orc_body* rec, const orc_body* lig, const orc_igrid* inner,
                     const orc_igrid* outer, const orc_nlgrid* nl, const orc_params* par,
                     const float* dof, size_t n, int ivmax, float* dof_out, double* e_last,
                     double* e_best, int* nfun) {
    ctx_f64 x;
    memset(&x, 0, sizeof(x));
    x.rec = rec; x.lig = lig; x.par = par;
    x.inner = inner; x.outer = outer; x.nl = nl;
    ctx_init_f64(&x);
    min_ctx m;
    memset(&m, 0, sizeof(m));
    m.x64 = &x;
    m.Mr = rec->M;
    m.Ml = lig->M;
    minimize_poses(&m, dof, n, ivmax, dof_out, e_last, e_best, nfun);
    ctx_free_f64(&x);
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
void orc_rotmats(const float* ang, size_t n, float* R) {
    for (size_t i = 0; i < n; ++i) rotmat_f32(ang[3 * i], ang[3 * i + 1], ang[3 * i + 2], R + 9 * i);
}

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
