/*
 * attract_hip.h — C ABI of the MI355X (gfx950) pose-scoring library.
 *
 * The library replaces the reference's GPU energy services
 * (service/GPUEnergyService6D.tpp, service/GPUEnergyService6DModes.tpp) and the
 * device-data upload they rely on (server/DeviceDataConfigurator.tpp).  All
 * types are plain C; no exceptions cross the boundary.  Every call returns an
 * att_status (ATT_OK = 0); att_last_error() gives the message of the last
 * failing call on that context.
 *
 * Threading: one context per host thread (the reference binds one worker
 * thread per device, service/GPUEnergyService6D.tpp:469-477).  Calls on one
 * context must not overlap; several contexts may share a device.  `stream` is
 * a hipStream_t (NULL = the context's own stream).  Device-resident calls on
 * different streams of one context may run concurrently; where they use the
 * context's device scratch (spatial pose order from 4096 poses on, adjoint
 * parameter-gradient rows) they take turns on it in stream order.
 */
#ifndef ATTRACT_HIP_H_
#define ATTRACT_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct att_ctx att_ctx;
typedef int att_status;
typedef uint64_t att_handle;

enum {
    ATT_OK = 0,
    ATT_ERR_INVALID = 1,     /* bad argument (the reference throws std::invalid_argument) */
    ATT_ERR_HIP = 2,         /* HIP runtime failure (the reference's CUDA_CHECK, commons/macros.h:62-72) */
    ATT_ERR_NOMEM = 3,
    ATT_ERR_HANDLE = 4,      /* unknown or wrong-kind handle (DataManager::get, server/DataManager.cpp:42-49) */
    ATT_ERR_UNSUPPORTED = 5, /* input outside what the kernels support (sizes, types) */
    ATT_NOT_READY = 6        /* att_event_query: the work before the event is still running */
};

#define ATT_MODES_MAX 10 /* MODES_MAX_NUMBER, model/Types_6D_Config.h:11 */

/* Protein (model/Protein.h:38-167), coordinates pivotized. */
typedef struct {
    unsigned n;                /* numAtoms */
    unsigned numModes;         /* 0..ATT_MODES_MAX */
    const float* xyz;          /* SoA [x: n][y: n][z: n] */
    const unsigned* type;      /* raw atom types (neighbour-list term), 0 = none */
    const unsigned* mappedType;/* grid types (interpolation term), 0 = none */
    const float* charge;
    const float* modes;        /* [x: n*M][y: n*M][z: n*M], element i*M+m (Protein.h:88-98) */
    const float* modeForce;    /* numModes force constants (readFile.cpp:218-221) */
} att_protein_desc;

/* Interpolation grid (IntrplGrid::Desc, model/IntrplGrid.h:165-199). */
typedef struct {
    unsigned nx, ny, nz;
    unsigned numGrids;         /* ligand-type grids + 1; grid 0 is electrostatics */
    float ox, oy, oz;          /* position of grid point (0,0,0) */
    float spacing;
    const float* data;         /* numGrids*nz*ny*nx float4 {fx,fy,fz,E} */
} att_grid_desc;

/* Neighbour-list grid (NLGrid::Desc, model/NLGrid.h:146-187). */
typedef struct {
    unsigned nx, ny, nz;
    float ox, oy, oz;
    float spacing;
    float plateau;             /* dPlateau */
    const uint32_t* desc;      /* nz*ny*nx {numEl, idx} */
    const uint32_t* list;      /* concatenated receptor indices */
    size_t numEl;
} att_nlgrid_desc;

/* ParamTable::attractFFParams_t (model/ParamTable.h:26-32). */
typedef struct {
    float rc, ac;
    int ipon;
    float rmin2, emin;
} att_pair_params;

/* DOF_6D<float> / Result_6D<float> (model/Types_6D.h:31-54). */
typedef struct { float pos[3]; float ang[3]; } att_dof6;
typedef struct { float E; float pos[3]; float ang[3]; } att_result6;

/* DOF_6D_Modes<float> / Result_6D_Modes<float> (model/Types_6D_Modes.h:33-60). */
typedef struct { att_dof6 d6; float modesRec[ATT_MODES_MAX]; float modesLig[ATT_MODES_MAX]; } att_dof6m;
typedef struct { att_result6 r6; float modesRec[ATT_MODES_MAX]; float modesLig[ATT_MODES_MAX]; } att_result6m;

/* --- context ---------------------------------------------------------------
 * Replaces the per-worker CUDA state of GPUEnergyService6D::Private
 * (service/GPUEnergyService6D.tpp:94-115) and cudaSetDevice per item (:474). */
att_status att_ctx_create(int hip_device, att_ctx** out);
/* Spatial pose order (default on) of the single-pair adjoint (when every
 * resident workgroup gets at least four poses), of the forward scoring calls
 * from 4096 poses on, and of the pairs of a batch created while it is on
 * (pairs of 256 poses or more): poses are sorted by the Morton code of their
 * translation and each XCD scores a contiguous range of that order, so its
 * L2 holds the grid lines its poses share.  Per-pose results do not depend on
 * it; the parameter-gradient sum runs in a different (still fixed) order. */
att_status att_set_pose_order(att_ctx* ctx, int enabled);
/* Waves per pose of the flexible forward kernel (att_score_modes[_dev]):
 * 1, 2, 4 or 8, or 0 (default: 1 below 384 ligand beads, else 4).  Energy, forces,
 * torques and ligand mode gradients do not depend on it (bead-order sums);
 * the receptor mode gradients are summed per wave, then over the waves in
 * order, so their last bits do. */
att_status att_set_modes_waves(att_ctx* ctx, int waves);
/* Diagnostics (no counterpart in the reference): while d_counts is set, every
 * att_score_rigid[_dev], att_score_modes[_dev] and att_score_adjoint[_dev]
 * call writes, per pose, two uint32 {neighbour-list entries of the pose's
 * ligand beads walked (the list lengths of their cells), pairs inside the
 * plateau} to d_counts (device memory of `capacity` poses; a call with more
 * poses fails).  NULL turns it off.  Integer sums: the same in every kernel
 * and every order. */
att_status att_debug_pair_counts(att_ctx* ctx, void* d_counts, size_t capacity);
void att_ctx_destroy(att_ctx* ctx);
const char* att_last_error(const att_ctx* ctx);

/* --- device data -----------------------------------------------------------
 * Replace DeviceDataConfigurator::attach for Protein (DeviceDataConfigurator.tpp:27-95),
 * GridUnion (:117-326) and ParamTable/SimParam (:337-362); att_release replaces
 * DeviceManager::detachFromDevice (server/DeviceManager.cpp:61-111). */
att_status att_upload_protein(att_ctx* ctx, const att_protein_desc* desc, att_handle* out);
att_status att_upload_grid(att_ctx* ctx, const att_grid_desc* inner, const att_grid_desc* outer,
                           const att_nlgrid_desc* nl /* may be NULL (ligand grids) */,
                           att_handle* out);
att_status att_upload_params(att_ctx* ctx, const att_pair_params* table, unsigned numTypes,
                             int potShape /* 8 or 12 */, int dielec /* 0 constant, 1 variable */,
                             float epsilon, float ffelec, att_handle* out);
att_status att_release(att_ctx* ctx, att_handle h);

/* --- rigid scoring (Types_6D<float>) --------------------------------------
 * att_score_rigid replaces one call of GPUEnergyService6D's item processor
 * (service/GPUEnergyService6D.tpp:469-512: H2D of the DOFs, transform,
 * interpolation, neighbour list, reduction, D2H and the host final reduce
 * h_finalReduce, model/reduction.h:159-207).  Host buffers in and out;
 * returns when `out` holds the n results. */
att_status att_score_rigid(att_ctx* ctx, att_handle rec, att_handle lig, att_handle grid,
                           att_handle params, const att_dof6* dofs, size_t n, att_result6* out,
                           void* stream);
/* Same computation on device-resident DOFs/results; asynchronous on `stream`. */
att_status att_score_rigid_dev(att_ctx* ctx, att_handle rec, att_handle lig, att_handle grid,
                               att_handle params, const att_dof6* d_dofs, size_t n,
                               att_result6* d_out, void* stream);

/* --- flexible scoring (Types_6D_Modes<float>) -----------------------------
 * Replaces GPUEnergyService6DModes's item processor
 * (service/GPUEnergyService6DModes.tpp:634-676) with the semantics of the
 * reference CPU path (service/CPUEnergyService6DModes.tpp:121-327): receptor
 * moved into the ligand frame with R(-ang) (transform_modes.h:26-41), scored
 * on the ligand grid, forces rotated back, neighbour list on the deformed
 * receptor, torque with undeformed ligand coordinates, correctModeForce. */
att_status att_score_modes(att_ctx* ctx, att_handle rec, att_handle lig, att_handle recGrid,
                           att_handle ligGrid, att_handle params, const att_dof6m* dofs,
                           size_t n, att_result6m* out, void* stream);
att_status att_score_modes_dev(att_ctx* ctx, att_handle rec, att_handle lig, att_handle recGrid,
                               att_handle ligGrid, att_handle params, const att_dof6m* d_dofs,
                               size_t n, att_result6m* d_out, void* stream);

/* --- grid generation on the device ------------------------------------------
 * No reference counterpart: the reference reads grids made by ATTRACT's tools
 * (fileIO/grid_orig.cpp:85-301, IntrplGrid::Desc / NLGrid::Desc).  These build
 * them from a body on the device with the semantics the scoring path assumes:
 * grid t>0 at point p = sum over beads of type 1..T of the LJ potential of
 * (bead type, probe t) at r' = max(r, plateau) (model/forcefield.h:28-83),
 * force channel -dV/dr' along (p - x)/r; grid 0 = Coulomb at r'
 * (forcefield.h:85-146); neighbour lists = beads with type != 0 within
 * neighbourdis of p, ascending index.  Sums in double, rounded once to float.
 * srcIsLigand selects the table index: 0 (receptor grids, probes are ligand
 * types) table[(src-1)*T + (probe-1)], 1 (ligand grids) table[(probe-1)*T +
 * (src-1)].  Geometry: point (ix,iy,iz) = origin + index * spacing. */
typedef struct {
    unsigned nx, ny, nz;
    float ox, oy, oz;
    float spacing;
} att_grid_geom;
/* The (numTypes+1) point grids into host memory: float4 [grid][z][y][x]. */
att_status att_generate_grid(att_ctx* ctx, const att_protein_desc* body,
                             const att_pair_params* table, unsigned numTypes, int potShape,
                             int dielec, float ffelec, float plateau, int srcIsLigand,
                             const att_grid_geom* geom, float* out);
/* Neighbour lists into host memory: desc {count, start} per point; list may
 * be NULL to learn *total first. */
att_status att_generate_nl(att_ctx* ctx, const att_protein_desc* body, const att_grid_geom* geom,
                           float neighbourdis, uint32_t* desc, uint32_t* list,
                           size_t listCapacity, uint64_t* total);
/* Inner and outer grids and (neighbourdis > 0) the neighbour-list grid over
 * the inner geometry, built and kept on the device: the handle is used like
 * one from att_upload_grid. */
att_status att_build_grid(att_ctx* ctx, const att_protein_desc* body, const att_pair_params* table,
                          unsigned numTypes, int potShape, int dielec, float ffelec, float plateau,
                          int srcIsLigand, const att_grid_geom* inner, const att_grid_geom* outer,
                          float neighbourdis, att_handle* out);

/* --- hand-written adjoint -----------------------------------------------------
 * No reference counterpart (the reference has no gradient beyond its force
 * reduction; service/CPUEnergyService6D.tpp:100-195 defines the energy).
 * Exact gradient of the returned energy E = grid term of the ligand beads +
 * neighbour-list pair terms:
 *   out per pose (stride 7 rigid, 27 modes): {E, w*dE/dpos[3], w*dE/dang[3],
 *   w*dE/dmodesRec[10], w*dE/dmodesLig[10]} -- gradients, not forces, with the
 *   rotation derivative over deformed coordinates.  Per-bead forces are
 *   recomputed, never stored.
 *   pgrad (NULL to skip): numTypes^2 x 3 {dE/drc, dE/dac, dE/demin} per table
 *   entry, summed over poses with weights w (NULL = 1); needs basis grids.
 * `modes` selects the DOF layout: 0 = att_dof6, 1 = att_dof6m. */
att_status att_upload_basis(att_ctx* ctx, att_handle grid, unsigned numTypes,
                            const float* inner /* 2T*nz*ny*nx */, const float* outer);
att_status att_score_adjoint(att_ctx* ctx, att_handle rec, att_handle lig, att_handle grid,
                             att_handle params, const float* dofs, int modes, size_t n,
                             const float* w, float* out, float* pgrad, void* stream);
att_status att_score_adjoint_dev(att_ctx* ctx, att_handle rec, att_handle lig, att_handle grid,
                                 att_handle params, const float* d_dofs, int modes, size_t n,
                                 const float* d_w, float* d_out, float* d_pgrad, void* stream);

/* att_score_adjoint_dev without any allocation: no reference counterpart; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops that forbid hipMalloc).  The parameter-gradient scratch is a
 * fixed device array of the library (13 MB per GPU, kept until the process
 * ends): at most 16 pair types and 4096 resident workgroups, else
 * ATT_ERR_UNSUPPORTED; poses are taken in input order (no spatial order,
 * which needs sort scratch), so the per-pose outputs equal
 * att_score_adjoint_dev's and the parameter gradient equals it when that
 * call runs in input order too (att_set_pose_order(ctx, 0)).  All calls on
 * one GPU share the scratch: calls with parameter gradients that overlap on
 * one GPU (other streams, other contexts) race. */
att_status att_score_adjoint_dev_noalloc(att_ctx* ctx, att_handle rec, att_handle lig,
                                         att_handle grid, att_handle params, const float* d_dofs,
                                         int modes, size_t n, const float* d_w, float* d_out,
                                         float* d_pgrad, void* stream);

/* The parameter gradient with respect to the pair parameters the table is
 * derived from (fileIO/readFile.cpp:479-513): per table entry rmin =
 * sqrt(rmin2) and emin (the table's own field, minus the well depth), of
 * which rc and ac are functions for the 8-6 shape (rc = -3 emin rmin^8,
 * ac = -4 emin rmin^6) and the 12-6 shape (rbc^2 = rmin^2 / 1.25992105,
 * abc = -4 emin, rc = abc rbc^12, ac = abc rbc^6).
 *   pgrad: numTypes^2 x 3 {dE/drc, dE/dac, dE/demin} (att_score_adjoint);
 *   out:   numTypes^2 x 2 {dE/drmin, dE/demin}, in double, rounded to float.
 * The rmin2 branch point of the pair term contributes no derivative (E is
 * continuous there).  Host arithmetic only; no context, no device.  Returns
 * ATT_ERR_INVALID for null pointers or a shape other than 8 and 12. */
att_status att_param_grad_rmin_emin(const att_pair_params* table, unsigned numTypes, int potShape,
                                    const float* pgrad, float* out);

/* --- diagnostics -------------------------------------------------------------
 * Per-bead intermediates of one pose, computed by the same device functions
 * as the scoring kernels: transformed coordinates (SoA 3n), per-bead
 * {fx,fy,fz,E} after grid + neighbour list (SoA 4n), cell indices per bead
 * (inner xyz, outer xyz, neighbour-list xyz; -1 when not used) and the number
 * of listed neighbours inside the plateau.  Mirrors the intermediate buffers
 * d_trafoLig / d_potLig of service/GPUEnergyService6D.tpp:117-128. */
att_status att_debug_beads_rigid(att_ctx* ctx, att_handle rec, att_handle lig, att_handle grid,
                                 att_handle params, const att_dof6* dof, float* xyz, float* pot,
                                 int* cell, unsigned* nl_pairs);
/* Modes variant; additionally the receptor forces (SoA 3*n_rec, after rotation
 * back and neighbour-list reaction) and receptor coordinates in the ligand frame. */
att_status att_debug_beads_modes(att_ctx* ctx, att_handle rec, att_handle lig, att_handle recGrid,
                                 att_handle ligGrid, att_handle params, const att_dof6m* dof,
                                 float* xyz, float* pot, int* cell, unsigned* nl_pairs,
                                 float* rec_pot, float* rec_xyz);

/* Rotation matrices R(phi, ssi, rot) of n angle triples as the kernels compute
 * them (matrixFunctions.h:38-77), for checking them against the host. */
att_status att_debug_rotmats(att_ctx* ctx, const float* angles /* 3n */, size_t n, float* R /* 9n */);

/* Single-precision sine and cosine of n floats as the kernels evaluate them
 * (the values of the C library's sincosf, which the reference's
 * std::sin/std::cos(float) calls in matrixFunctions.h:51-60 reach), for
 * checking them against the host library. */
att_status att_debug_sincosf(att_ctx* ctx, const float* x, size_t n, float* s, float* c);

/* Entries 0..n-1 of the neighbour-list plateau-ratio table of plateau^2 = P2
 * as the kernels evaluate them per pair (model/NLGrid.tpp:44-51 builds the
 * table, model/NLGrid.h:141-143 reads entry int(dr2 * 10000)), for checking
 * them against the table. */
att_status att_debug_plateau_ratio(att_ctx* ctx, float P2, size_t n, float* ratio);

att_status att_sync(att_ctx* ctx, void* stream);

/* A plain float4 copy of `bytes` (multiple of 16) from d_src to d_dst on
 * `stream`: the known-size kernel that calibrates hardware byte counters
 * (FETCH_SIZE/WRITE_SIZE) against the scoring kernels.  No reference
 * counterpart. */
att_status att_debug_copy(att_ctx* ctx, void* d_dst, const void* d_src, size_t bytes, void* stream);

/* --- device buffers, copies and timing -------------------------------------
 * Replace the reference's DeviceAllocator / HostPinnedAllocator
 * (allocator/DeviceAllocator.tpp:19-30, allocator/HostPinnedAllocator.tpp:19-30)
 * and the CUDA events of GPUEnergyService6D::Private (:101-114), so that a
 * caller needs no other GPU runtime in its process. */
att_status att_dev_alloc(att_ctx* ctx, size_t bytes, void** d_ptr);
att_status att_dev_free(att_ctx* ctx, void* d_ptr);
att_status att_host_alloc(att_ctx* ctx, size_t bytes, void** h_ptr); /* pinned */
att_status att_host_free(att_ctx* ctx, void* h_ptr);
att_status att_memcpy_h2d(att_ctx* ctx, void* d_dst, const void* src, size_t bytes, void* stream);
att_status att_memcpy_d2h(att_ctx* ctx, void* dst, const void* d_src, size_t bytes, void* stream);
att_status att_get_stream(att_ctx* ctx, void** stream);
att_status att_event_create(att_ctx* ctx, void** ev);
att_status att_event_destroy(att_ctx* ctx, void* ev);
att_status att_event_record(att_ctx* ctx, void* ev, void* stream);
att_status att_event_elapsed_ms(att_ctx* ctx, void* ev_start, void* ev_stop, float* ms);

/* --- streams, asynchronous copies, completion ------------------------------
 * The pipeline of GPUEnergyService6D::Private (:101-114, :163-432: its own
 * streams, asynchronous copies from pinned buffers, events polled by the
 * worker) expressed over the context's device.  Streams are created at the
 * default priority.  The async copies return once queued; `src`/`dst` must be
 * pinned (att_host_alloc) for the copy to overlap.  att_event_query returns
 * ATT_OK when the work before the event has finished, ATT_NOT_READY when not. */
att_status att_stream_create(att_ctx* ctx, void** stream);
/* Free and total memory of the context's device (hipMemGetInfo). */
att_status att_device_memory(att_ctx* ctx, size_t* free_bytes, size_t* total_bytes);
att_status att_stream_destroy(att_ctx* ctx, void* stream);
att_status att_memcpy_h2d_async(att_ctx* ctx, void* d_dst, const void* src, size_t bytes, void* stream);
att_status att_memcpy_d2h_async(att_ctx* ctx, void* dst, const void* d_src, size_t bytes, void* stream);
att_status att_event_query(att_ctx* ctx, void* ev);
att_status att_event_sync(att_ctx* ctx, void* ev);

/* --- batched adjoint over many pairs --------------------------------------
 * No reference counterpart: the reference scores one receptor-ligand pair per
 * request (GPUEnergyService6D::createItemProcessor, GPUEnergyService6D.tpp:
 * 468-520, items of 100 poses).  A batch scores many pairs of uneven bead and
 * pose counts with a few persistent launches: pairs are grouped into buckets
 * by mode capacity and ligand size, each bucket's launch shape (waves per
 * pose, poses per work tile) is picked by att_batch_tune, and the launches are
 * captured into one hipGraph that att_batch_run replays.  Per pose the output
 * is the adjoint's (att_score_adjoint_dev without parameter gradients).  The
 * device buffers named by the pairs must stay valid while the batch exists.
 * att_batch_create sorts the poses of every pair of 256 or more poses into a
 * spatial order from the DOFs its buffer holds at that moment (synchronously);
 * later runs keep that order whatever the buffer then holds, which changes
 * only their speed, never a result. */
typedef struct att_batch att_batch;
typedef struct {
    att_handle rec, lig, grid, params;  /* uploaded receptor, ligand, receptor grid, table */
    const float* d_dofs;                /* n * (modes ? 26 : 6) floats */
    float* d_out;                       /* n * (modes ? 27 : 7) floats */
    const float* d_w;                   /* n upstream weights, or NULL */
    size_t n;
    int modes;
} att_batch_pair;
typedef struct {
    unsigned modeCapacity;   /* 0, 5 or 10 */
    int sizeClass;           /* ligand beads: 0 (<192), 1 (<384), 2 (<640), 3 (<1024), 4 */
    unsigned numPairs, numTiles;
    int wavesPerPose;        /* 1, 2 or 4 */
    unsigned posesPerTile;
    float ms;                /* time of the bucket's launch when tuned, else -1 */
} att_batch_bucket;
att_status att_batch_create(att_ctx* ctx, const att_batch_pair* pairs, unsigned npairs, att_batch** out);
att_status att_batch_destroy(att_ctx* ctx, att_batch* b);
unsigned att_batch_buckets(const att_batch* b);
att_status att_batch_bucket_info(att_ctx* ctx, const att_batch* b, unsigned bucket, att_batch_bucket* info);
/* The bucket that scores pair `pair` (its index in att_batch_create's list). */
att_status att_batch_pair_bucket(att_ctx* ctx, const att_batch* b, unsigned pair, unsigned* bucket);
att_status att_batch_set_shape(att_ctx* ctx, att_batch* b, unsigned bucket, int wavesPerPose,
                               unsigned posesPerTile);
att_status att_batch_tune(att_ctx* ctx, att_batch* b, void* stream);
att_status att_batch_run(att_ctx* ctx, att_batch* b, void* stream);

/* --- device-resident minimization -----------------------------------------
 * Replaces the reference's minimization loop (apps/em/RequestHandler.tpp:
 * 101-234 with VA13Solver, apps/em/VA13Solver.cpp:36-69, minfor.f): every
 * pose runs VA13 with at most maxFunEval line-search evaluations (the
 * reference's default is 500, VA13Solver.h:66) on the variables (angles,
 * position, receptor modes, ligand modes); each round scores all running
 * poses with the adjoint (energy and exact gradient).  Results per pose, as
 * the reference's solver reports them: the last evaluated DOFs (same layout as
 * the input) and their energy; optionally the lowest energy seen (ebest) and
 * the number of line-search evaluations (nfun).  The _dev form takes device
 * buffers and launches on `stream`; its host loop reads the number of running
 * poses every eight rounds. */
att_status att_minimize_dev(att_ctx* ctx, att_handle rec, att_handle lig, att_handle grid,
                            att_handle params, const float* d_dofs, int modes, size_t n,
                            int maxFunEval, float* d_dofs_out, float* d_energy, float* d_ebest,
                            int* d_nfun, void* stream);
att_status att_minimize(att_ctx* ctx, att_handle rec, att_handle lig, att_handle grid,
                        att_handle params, const float* dofs, int modes, size_t n, int maxFunEval,
                        float* dofs_out, float* energy, float* ebest, int* nfun, void* stream);

/* Minimization jobs: the pieces of att_minimize_dev for callers that
 * interleave several (the serving engine).  A job owns att_minjob_bytes of
 * device memory; start, then issue rounds (one scoring launch and one
 * minimizer step each, no host synchronization); att_minjob_running copies the
 * number of running poses to host memory asynchronously (pinned memory lets
 * the copy overlap); after att_minjob_round_limit rounds every pose has
 * finished.  Results land in the buffers given at creation. */
typedef struct att_minjob att_minjob;
size_t att_minjob_bytes(size_t n, int dim, int modes);
att_status att_minjob_create(att_ctx* ctx, att_handle rec, att_handle lig, att_handle grid,
                             att_handle params, int modes, size_t n, int maxFunEval,
                             float* d_dofs_out, float* d_energy, float* d_ebest, int* d_nfun,
                             att_minjob** out);
att_status att_minjob_start(att_ctx* ctx, att_minjob* job, const float* d_dofs, void* stream);
att_status att_minjob_rounds(att_ctx* ctx, att_minjob* job, int rounds, void* stream);
int att_minjob_round_limit(const att_minjob* job);
att_status att_minjob_running(att_ctx* ctx, att_minjob* job, unsigned* h_count, void* stream);
att_status att_minjob_destroy(att_ctx* ctx, att_minjob* job);

/* Library build identification (arch, flags). */
const char* att_build_info(void);

#ifdef __cplusplus
}
#endif

#endif /* ATTRACT_HIP_H_ */
