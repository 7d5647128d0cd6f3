/* Request loop over device-resident solves (no counterpart in the reference,
 * which solves one point per process run in main.c:1249-1568).
 *
 * A server owns `slots` worker threads. A request names a registered mesh, a
 * frequency, a mode and 1..n port pairs; it is queued on submission and
 * admitted as soon as a worker and device memory are free, while earlier
 * requests are still running (each engine has its own HIP stream, so
 * admitted requests overlap on the GPU). Admission is by HBM budget: a
 * request runs on an engine (the working set of one mesh: static Lp/P, the
 * retarded matrices, the FP64 MNA matrix and its copy, the FP32 factor and
 * the vectors) that is idle, or on a new one when its estimated size fits
 * under the budget, after idle engines of other meshes are evicted, least
 * recently used first. A request that does not fit waits and later requests
 * that fit are admitted past it. Factorizations are cached by (mesh,
 * frequency, mode): an idle engine that still holds the factor of the
 * request's point is preferred, and the request then only solves its
 * right-hand sides.
 *
 * Continuous batching: when a worker admits a mixed-precision request, the
 * mixed requests of the same mesh waiting in the queue at that moment (up to
 * max_batch in all, and as many as the budget allows) join it, and the whole
 * group runs as one peec_engine_solve_batch on one engine: their FP32
 * factors share every launch of the ordered factor. Requests that arrive
 * meanwhile form the next batch. A member's answer is that of the request
 * solved alone in the same row order.
 *
 * The static Lp and P of a mesh are assembled once when the mesh is added;
 * further engines of that mesh start from copies of them. Mesh and DualMesh
 * must stay valid while the server exists. Calls on one server may come from
 * any host thread. */
#ifndef PEEC_SERVE_H
#define PEEC_SERVE_H

#include <stddef.h>

#include "peec_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct
{
    PeecEngineOptions engine;  /* options of every engine (retarded, refinement, ...);
                                  smooth_moments and shifted_moments are set to 0
                                  (no stored moments) */
    int slots;                 /* worker threads = requests running at once, default 4 */
    size_t hbm_budget_bytes;   /* device memory the engines may hold; 0 = 90% of free memory
                                  when the server is created */
    int max_batch;             /* mixed requests of one mesh solved together, default 8;
                                  1 = one request at a time */
} PeecServerOptions;

typedef struct PeecServer PeecServer;

typedef struct
{
    int status;                /* as peec_engine_solve */
    PeecPointReport report;
    double queue_s;            /* submission to admission */
    double latency_s;          /* submission to completion */
    int engine;                /* index of the engine that ran it (for tests) */
} PeecRequestResult;

void peec_server_options_default(PeecServerOptions *options);

/* 0, or -1 (bad options / no GPU). */
int peec_server_create(const PeecServerOptions *options, PeecServer **server);

/* Registers a mesh: assembles its static Lp and P on the first engine.
 * *mesh_id receives its handle. 0 or -1. */
int peec_server_add_mesh(PeecServer *server, const Mesh *mesh, const DualMesh *dual, int *mesh_id);

/* Queues a request; returns at once. port_nodes: 2 * n_ports node indices
 * (copied). *ticket identifies it for peec_server_wait. 0 or -1 (bad input,
 * or a mesh whose working set exceeds the whole budget). */
int peec_server_submit(PeecServer *server, int mesh_id, double frequency, int mode,
                       const int *port_nodes, size_t n_ports, long long *ticket);

/* Waits for a request and releases it. z_in: n_ports values (may be NULL).
 * Returns the request's status, or -1 for an unknown ticket. */
int peec_server_wait(PeecServer *server, long long ticket, Complex *z_in, PeecRequestResult *result);

/* Estimated device bytes of one engine of a mesh with n_edges edges and
 * n_nodes nodes (the unit of admission). */
size_t peec_server_engine_bytes(size_t n_edges, size_t n_nodes, int retarded);

/* Device memory currently held by the server's engines, and their count. */
size_t peec_server_memory(const PeecServer *server, int *n_engines);

/* Finishes the queued requests, stops the workers, frees the engines. */
void peec_server_destroy(PeecServer *server);

#ifdef __cplusplus
}
#endif

#endif
