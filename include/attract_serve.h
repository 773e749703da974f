/*
 * attract_serve.h — docking service with continuous batching on one GPU.
 *
 * No reference counterpart for the engine itself: it replaces the
 * reference's dispatch of minimization requests (apps/em/RequestHandler.tpp:
 * 101-234 feeding server/Dispatcher.tpp:64-113 and the polling
 * Server::wait, server/Server.tpp:168-180) with one request loop that keeps
 * receptor grids and minimizer state resident on the device.
 *
 * - Receptors (protein + grids), ligands and parameter tables are registered
 *   once and uploaded on first use; receptor grids stay cached across jobs
 *   and are evicted least recently used when their memory is needed and no
 *   running job uses them.
 * - A job is a set of starting poses of one receptor-ligand pair, each
 *   minimized by VA13 (at most maxFunEval line-search evaluations) on the
 *   adjoint (att_minimize semantics).  Jobs are admitted in arrival order as
 *   soon as their HBM footprint (state + buffers + grids not yet resident)
 *   fits the budget; admitted jobs run concurrently, each on its own stream,
 *   and new jobs join while others are running (continuous batching).
 * - A job's results depend only on its own inputs: alone or in any mix they
 *   are the same, bit for bit.
 *
 * Thread safety: every function may be called from any thread; the device
 * is driven by the engine's own thread.
 */
#ifndef ATTRACT_SERVE_H_
#define ATTRACT_SERVE_H_

#include <stddef.h>
#include <stdint.h>

#include "attract_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct att_server att_server;

typedef struct {
    double queued_ms;   /* submission to admission */
    double run_ms;      /* admission to results on the host */
    double latency_ms;  /* submission to results on the host */
    int grid_cached;    /* the receptor grid was resident at admission */
    unsigned rounds;    /* scoring rounds issued */
    size_t bytes;       /* device memory the job held (state + buffers) */
    uint64_t admit_seq; /* admission order: 1 for the server's first admitted job */
} att_job_info;

typedef struct {
    uint64_t jobs_done, poses_done, evaluations;
    uint64_t grid_hits, grid_misses, evictions;
    size_t budget_bytes, peak_bytes;
    unsigned max_running;  /* most jobs running at once */
    double p50_ms, p99_ms; /* job latency percentiles over finished jobs */
    double busy_s;         /* time with at least one job on the device */
} att_server_stats;

/* hbm_budget = 0: 80 % of the device memory free at creation.  Destroy drops
 * queued jobs, lets running ones finish and frees every upload. */
att_status att_server_create(int hip_device, size_t hbm_budget, att_server** out);
void att_server_destroy(att_server* s);
/* Message of the calling thread's last failed att_server_* call (a failed
 * job's reason after att_server_wait returned an error for it). */
const char* att_server_last_error(const att_server* s);

att_status att_server_add_receptor(att_server* s, const att_protein_desc* rec,
                                   const att_grid_desc* inner, const att_grid_desc* outer,
                                   const att_nlgrid_desc* nl, uint64_t* id);
/* A receptor given as a body only: its grids (att_build_grid semantics, for the
 * table `params`, neighbour lists over the inner geometry) are built on the
 * device when a job first needs them and cached like uploaded ones.  Jobs on
 * it must use the same table. */
att_status att_server_add_receptor_body(att_server* s, const att_protein_desc* rec, uint64_t params,
                                        const att_grid_geom* inner, const att_grid_geom* outer,
                                        float plateau, float neighbourdis, uint64_t* id);
att_status att_server_add_ligand(att_server* s, const att_protein_desc* lig, uint64_t* id);
att_status att_server_add_params(att_server* s, const att_pair_params* table, unsigned numTypes,
                                 int potShape, int dielec, float epsilon, float ffelec,
                                 uint64_t* id);

/* Copies the n starting poses (26 floats each with modes, else 6) and queues
 * the job; returns at once. */
att_status att_server_submit(att_server* s, uint64_t receptor, uint64_t ligand, uint64_t params,
                             const float* dofs, int modes, size_t n, int maxFunEval,
                             uint64_t* job);
/* Waits up to timeout_ms (< 0: no limit) for the job; ATT_NOT_READY on
 * time-out.  Copies the results (any output may be NULL) and forgets the job. */
att_status att_server_wait(att_server* s, uint64_t job, int timeout_ms, float* dofs_out,
                           float* energy, float* ebest, int* nfun, att_job_info* info);
att_status att_server_get_stats(att_server* s, att_server_stats* out);

#ifdef __cplusplus
}
#endif

#endif
