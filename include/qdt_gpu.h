/*
 * qdt_gpu.h -- GPU-specific entry points beside the drop-in qdt.h.
 *
 * None of these exist in the reference; they are for callers that keep their
 * points in GPU memory, that want the results in bulk, or that need the
 * pipeline's counters.  Plain C types only: device pointers are passed as
 * ordinary pointers to memory of the current HIP device.
 */
#ifndef QDT_MI355X_QDT_GPU_H
#define QDT_MI355X_QDT_GPU_H

#include <stddef.h>
#include <stdint.h>

#include "qdt.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Counters of the last batch call on a context. */
typedef struct qdt_gpu_stats
{
    uint64_t orient_calls;        /* orient2d evaluations in leaves and merges */
    uint64_t orient_exact;        /* ... that the FP64 filter could not decide */
    uint64_t incircle_calls;
    uint64_t incircle_exact;
    uint64_t extract_exact;       /* face tests queued for the exact kernel */
    uint64_t unique_points;
    uint64_t triangles;
    uint64_t kd_levels;           /* levels of the median-cut tree */
    double device_ms;             /* time of the device pipeline (events) */
    /* filled only with profiling on (qdt_gpu_set_profiling): */
    double stage_ms[8];           /* keys+y sort, x sort, dedupe, frame-1 sort, kd levels, leaves, merges, extract */
    double merge_launch_ms_sum;   /* sum of the merge kernel launches' durations (one launch per
                                     kd level merged in one piece per node) */
    uint64_t merge_launches;
    uint64_t merge_edges;         /* quad-edge records those launches created */
    double seam_level_ms_sum;     /* kd levels cut into seam pieces, all their kernels */
    uint64_t seam_retries;        /* runs repeated without seam pieces */
    uint64_t seam_counters[8];    /* anchors certified, anchors inserted, pieces run, pieces
                                     resumed, pieces stopped by a claimed vertex, anchors with a
                                     clean dry run, anchors selected, vertex rings rewritten */
    uint64_t err_bits;            /* device error flags of the last run (0 when it succeeded) */
    uint64_t cells;               /* LDS cells (one wavefront builds each cell's kd subtree) */
    uint64_t cell_level;          /* kd level of the cells */
    uint64_t spare_records;       /* quad-edge records taken from the spare region above 3 * unique
                                     points (demand; a run that needs more than the region holds
                                     is repeated without seam pieces) */
    uint64_t cells_exact;         /* LDS cells rebuilt with the exact predicates: the first pass
                                     runs the floating-point filters alone and lists every cell
                                     where one could not decide */
    /* batched form: */
    uint64_t graph_replay;        /* 1: the call was one launch of the shape's captured HIP graph */
    uint64_t graph_captured;      /* 1: the call was the shape's first (run, then captured) */
    uint64_t graph_fallback;      /* 1: the graph's guards failed (the data needed other launches
                                     than the recorded run): the call ran without the graph */
    uint64_t cell_points;         /* largest kd node built in one LDS cell (128, 256 or 512) */
    uint64_t merge_lanes;         /* merges per wavefront of the one-thread-per-node merges */
    uint64_t tune_source;         /* 0 default, 1 timed on this call, 2 tuning cache, 3 environment */
    double tune_ms;               /* the chosen pair's device time when it was timed */
} qdt_gpu_stats;

/* Like qdt_execute(), but d_pts points to n points in device memory.  The
 * triangles stay on the device (see qdt_gpu_device_triangles); the host-side
 * copies that the query functions read are made on the first query.  The
 * same checks and status codes as qdt_execute() apply.  The points are
 * copied to the host only when qdt_get_point() is first called. */
QDT_API qdt_status qdt_gpu_execute_device(qdt_t* ptr, const qdt_vec2d* d_pts, size_t n);

/* Device pointer to the 3 * T uint32 vertex indices of the last batch call,
 * valid until the next call on the context; returns T. */
QDT_API size_t qdt_gpu_device_triangles(const qdt_t* ptr, const uint32_t** d_tris);

/* Copies all triangles (3 * T uint32, counter-clockwise) to host memory out,
 * which must hold 3 * qdt_triangles_size() entries; returns T. */
QDT_API size_t qdt_gpu_copy_triangles(const qdt_t* ptr, uint32_t* out);

/* Counters of the last batch call; returns 1, or 0 for a NULL argument. */
QDT_API int qdt_gpu_get_stats(const qdt_t* ptr, qdt_gpu_stats* out);

/* Turns per-stage HIP event timing on or off for a context (default off). */
QDT_API void qdt_gpu_set_profiling(qdt_t* ptr, int enable);

/* Cuts large merges into seam pieces that run in parallel (default 1); 0
 * merges every kd node with one thread.  Results are identical either way. */
QDT_API void qdt_gpu_set_seams(qdt_t* ptr, int enable);

/* Builds the bottom kd levels in LDS cells, one wavefront per cell (default
 * 1); 0 uses one launch per level there too.  Results are identical. */
QDT_API void qdt_gpu_set_cells(qdt_t* ptr, int enable);

/* ---- Exact verifier ----
 *
 * No counterpart in the reference's library: its test harness checks results
 * on the CPU (test/qdt_verify.cpp:622-696).  These run the same kind of
 * checks as device kernels: every triangle strictly counter-clockwise (exact
 * orient2d); every edge in one triangle (boundary) or in two with opposite
 * directions; every interior edge locally Delaunay (exact incircle of the
 * opposite vertex <= 0); the boundary turning counter-clockwise (or straight)
 * at every vertex and continuing at every boundary vertex; T = 2V - 2 - b;
 * with duplicate information, only group representatives used, and all of
 * them; with no triangles, the distinct points collinear.  Together these make the list a Delaunay triangulation of the
 * distinct points (not necessarily the same choice on cocircular ties). */
typedef struct qdt_gpu_verify_report
{
    uint64_t triangles;
    uint64_t boundary_edges;      /* b */
    uint64_t vertices;            /* V: distinct vertices used */
    uint64_t bad_index;           /* vertex index out of range or repeated in a triangle */
    uint64_t not_representative;  /* vertex that is not the lowest index of its duplicate group */
    uint64_t bad_orientation;     /* triangles not strictly counter-clockwise */
    uint64_t bad_edges;           /* non-manifold edges, repeated directed edges, broken boundary */
    uint64_t non_delaunay;        /* interior edges failing the exact incircle test */
    uint64_t concave_boundary;    /* boundary vertices turning clockwise */
    uint64_t unused_points;       /* representatives used by no triangle */
    uint64_t off_line;            /* with no triangles: distinct points off the line through the
                                     first two (a set without triangles must be collinear) */
    int euler_ok;                 /* T = 2V - 2 - b (or T = 0) */
    int ok;                       /* every check passed */
    double device_ms;             /* kernel time of the check */
} qdt_gpu_verify_report;

/* Checks the context's last batch result (triangles, the duplicate rule
 * and every distinct point used).  QDT_ERR_NULL_CONTEXT / QDT_ERR_NULL_ARGUMENT
 * for NULL arguments; out-of-memory / internal codes otherwise.  An empty
 * context reports ok. */
QDT_API qdt_status qdt_gpu_verify(qdt_t* ptr, qdt_gpu_verify_report* out);

/* Checks T triangles (3 * T uint32 input indices, device memory) over n
 * points (device memory).  d_rep (n uint32, device memory: the lowest index
 * of each point's duplicate group) may be NULL, which skips the two
 * representative checks. */
QDT_API qdt_status qdt_gpu_verify_triangles(const qdt_vec2d* d_pts, size_t n, const uint32_t* d_tris, size_t T,
                                            const uint32_t* d_rep, qdt_gpu_verify_report* out);

/* ---- Half-edge mesh ----
 *
 * The half-edge mesh libqdt builds from a batch result before its first
 * insertion or constraint (GuibasStolfiDwyer_Wrapper.cpp:99-202, ensure_mesh_;
 * fields of wrapper.hxx:44-71), built on the device.  The reference keeps it
 * private; here it is an export.  For T triangles (3 T input indices < n,
 * device memory):
 *   d_he   (3 T x 4 int32, device): half-edge 3t+i = {origin, twin, next,
 *          triangle}: it runs from corner i of triangle t (origin: its local
 *          vertex) to corner (i+1)%3; next = 3t+(i+1)%3; twin = the half-edge
 *          running the other way along the same edge, -1 on the boundary.
 *   d_vert (up to min(n, 3 T) x 2 int32, device): local vertex k = {input
 *          index, half-edge}; local vertices are numbered in order of first
 *          appearance (triangle t, then corner i), and a vertex's half-edge is
 *          that first appearance.
 *   d_orig_to_local (n int32, device): the local vertex of each input
 *          index, -1 for indices no triangle uses.
 * The reference pairs twins by scanning each vertex's out-going half-edges;
 * here twins come from a sort of the undirected edge keys, with the same
 * result on any triangulation (each edge in one triangle, or in two with
 * opposite directions).  Edges that are neither stay without twins and are
 * counted in `unmatched`.  An index >= n: QDT_ERR_INVALID_INDEX. */
typedef struct qdt_gpu_mesh_info
{
    uint64_t vertices;   /* local vertices */
    uint64_t halfedges;  /* 3 T */
    uint64_t boundary;   /* half-edges without a twin (one per boundary edge) */
    uint64_t unmatched;  /* half-edges of edges in more than two triangles or in two with one direction */
    double device_ms;    /* kernel time of the build */
} qdt_gpu_mesh_info;

QDT_API qdt_status qdt_gpu_build_mesh(const uint32_t* d_tris, size_t T, size_t n, int32_t* d_he, int32_t* d_vert,
                                      int32_t* d_orig_to_local, qdt_gpu_mesh_info* out);

/* For tests of the fallback paths (no counterpart in the reference).
 * device_bytes: the most device memory the context's workspace may hold; a
 * batch call that needs more fails with QDT_ERR_OUT_OF_MEMORY and empties
 * the context, as a failed allocation does (0 = no limit).  spare_records:
 * the most quad-edge records the seam pieces may take beyond each node's own
 * range (0 = the default region); a run that needs more is repeated without
 * seam pieces (qdt_gpu_stats.seam_retries), with the same result. */
QDT_API void qdt_gpu_set_test_limits(qdt_t* ptr, size_t device_bytes, size_t spare_records);

/* ---- No-allocation form ----
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the same triangulation as qdt_gpu_execute_device of n
 * device-resident points, with every buffer of the pipeline taken from a
 * fixed per-GPU array inside the library (112 MiB, held by each GPU the
 * library runs on until the process ends) instead of allocated.  The call
 * allocates and frees no device or pinned memory; it runs on the null
 * stream and returns when the results are in place.
 *   d_tris: room for tri_capacity >= 2 n triangles (3 uint32 each); receives
 *           the triangles as input indices, counter-clockwise;
 *   d_rep:  n uint32, the lowest index with the same coordinates per point;
 *   *triangles: the triangle count.
 * At most 262,144 points (QDT_ERR_TOO_MANY_POINTS beyond); NULL pointers:
 * QDT_ERR_NULL_ARGUMENT; tri_capacity below 2 n: QDT_ERR_OUT_OF_MEMORY; a
 * NaN or infinite coordinate: QDT_ERR_NOT_FINITE.  Calls on one GPU share
 * its array: they are serialised by a lock inside the library, and two
 * processes never share one (each loads its own copy). */
QDT_API qdt_status qdt_gpu_execute_noalloc(const qdt_vec2d* d_pts, size_t n, uint32_t* d_tris, size_t tri_capacity,
                                           uint32_t* d_rep, size_t* triangles);

/* Number of HIP devices visible to the library (0 without a GPU). */
QDT_API int qdt_gpu_device_count(void);

/* Frees the device workspace a context keeps between calls. */
QDT_API void qdt_gpu_release_workspace(qdt_t* ptr);

/* ---- Batched form: many independent point sets in one pipeline pass ----
 *
 * No counterpart in the reference (whose callers run one qdt_execute per
 * set, each on its own context).  Set i is points [offsets[i], offsets[i+1])
 * of the input; every set is triangulated exactly as qdt_execute would do
 * it alone (same triangle set, same duplicate rule), but all sets share one
 * sort, one kd forest and one merge launch per level. */
typedef struct qdt_gpu_batch qdt_gpu_batch;

QDT_API qdt_gpu_batch* qdt_gpu_batch_create(void);
QDT_API void qdt_gpu_batch_destroy(qdt_gpu_batch* b);

/* Host input.  offsets holds k + 1 non-decreasing values starting at 0.
 * Returns QDT_ERR_NULL_CONTEXT (b NULL), QDT_ERR_NULL_ARGUMENT (offsets NULL
 * with k > 0, or pts NULL with points), QDT_ERR_INVALID_INDEX (offsets not
 * starting at 0 or decreasing), QDT_ERR_TOO_MANY_POINTS (more than 1e8 points
 * in all), or the out-of-memory / internal codes.  A set holding a NaN or
 * infinite coordinate does not fail the call: its own status is
 * QDT_ERR_NOT_FINITE and it has no triangles.  A failed call empties b. */
QDT_API qdt_status qdt_gpu_batch_execute(qdt_gpu_batch* b, const qdt_vec2d* pts, const size_t* offsets, size_t k);

/* Device input (d_pts in device memory, offsets on the host). */
QDT_API qdt_status qdt_gpu_batch_execute_device(qdt_gpu_batch* b, const qdt_vec2d* d_pts, const size_t* offsets,
                                                size_t k);

/* Number of sets of the last call. */
QDT_API size_t qdt_gpu_batch_sets(const qdt_gpu_batch* b);
/* Status of one set: QDT_OK or QDT_ERR_NOT_FINITE (QDT_ERR_INVALID_INDEX
 * for a set index out of range). */
QDT_API qdt_status qdt_gpu_batch_set_status(const qdt_gpu_batch* b, size_t set);
QDT_API size_t qdt_gpu_batch_triangles_size(const qdt_gpu_batch* b, size_t set);
QDT_API size_t qdt_gpu_batch_unique_points(const qdt_gpu_batch* b, size_t set);
/* Triangle `id` of a set, as set-local input indices (the set's first point
 * is 0), counter-clockwise; 1 on success, 0 (T untouched) otherwise. */
QDT_API size_t qdt_gpu_batch_get_triangle(const qdt_gpu_batch* b, size_t set, size_t T[3], size_t id);
/* Where a set's triangles lie in the bulk array: first triangle and count. */
QDT_API int qdt_gpu_batch_set_range(const qdt_gpu_batch* b, size_t set, size_t* first, size_t* count);
/* All triangles (3 * total uint32, set-local indices, each set's triangles
 * contiguous at its range), to host memory; returns the total. */
QDT_API size_t qdt_gpu_batch_copy_triangles(const qdt_gpu_batch* b, uint32_t* out);
/* The same array on the device, valid until the next call on b. */
QDT_API size_t qdt_gpu_batch_device_triangles(const qdt_gpu_batch* b, const uint32_t** d_tris);
/* Lowest set-local index with the same coordinates, for every input point
 * (n uint32, in input order); returns n. */
QDT_API size_t qdt_gpu_batch_copy_rep(const qdt_gpu_batch* b, uint32_t* out);
QDT_API int qdt_gpu_batch_get_stats(const qdt_gpu_batch* b, qdt_gpu_stats* out);
QDT_API void qdt_gpu_batch_set_profiling(qdt_gpu_batch* b, int enable);
QDT_API void qdt_gpu_batch_set_seams(qdt_gpu_batch* b, int enable);
/* HIP graphs (default 1): the first batch of a shape (the sets' sizes) runs
 * as usual while every value the pipeline sizes its launches by is recorded;
 * the same pipeline is then captured in a graph, and later batches of that
 * shape are one launch of it.  Device-side guards compare the recorded
 * values with the new batch's (distinct counts, digit ranges of the sorts,
 * triangle counts, ...); a batch that differs runs without the graph.
 * Results are identical either way. */
QDT_API void qdt_gpu_batch_set_graphs(qdt_gpu_batch* b, int enable);
/* Launch parameters: LDS cell size in points (128, 256, 512) and merges per
 * wavefront (16, 32, 64); 0 (the default) picks both by timing every pair on
 * the first batch of each shape bucket (cached in memory and in the file
 * $QDT_TUNE_FILE, default ~/.cache/qdt_amd/tune.txt; QDT_CELL_POINTS and
 * QDT_MERGE_LANES override).  Results are identical for every choice. */
QDT_API void qdt_gpu_batch_set_tuning(qdt_gpu_batch* b, int cell_points, int merge_lanes);

/* ---- Serving form: a request loop with continuous batching ----
 *
 * No counterpart in the reference.  Requests (host point sets) can be
 * submitted at any time from any thread.  `workers` threads, each with its
 * own HIP stream and resident device workspace, take every queued request
 * that fits beside the running batches within hbm_budget_bytes (each
 * worker is charged its engine's exact workspace for the batch plus its input
 * buffer) and triangulate them together in one
 * batch, so requests that arrive while others run are admitted as memory
 * frees up and their batches overlap on the device.  A request's result is
 * the same as qdt_execute gives for it alone. */
typedef struct qdt_gpu_server qdt_gpu_server;

/* NULL without a GPU or memory. */
QDT_API qdt_gpu_server* qdt_gpu_server_create(size_t hbm_budget_bytes, int workers);
/* Finishes the queued requests, then stops the workers. */
QDT_API void qdt_gpu_server_destroy(qdt_gpu_server* s);
/* Copies the n points and queues them; returns the request id (0 on a NULL
 * argument or no memory). */
QDT_API uint64_t qdt_gpu_server_submit(qdt_gpu_server* s, const qdt_vec2d* pts, size_t n);
/* 0 queued, 1 running, 2 done, -1 unknown id. */
QDT_API int qdt_gpu_server_poll(qdt_gpu_server* s, uint64_t id);
/* Blocks until the request is done; returns its status (QDT_OK,
 * QDT_ERR_NOT_FINITE, QDT_ERR_TOO_MANY_POINTS, QDT_ERR_OUT_OF_MEMORY when it
 * cannot fit the budget alone, ...) and the submit-to-done latency. */
QDT_API qdt_status qdt_gpu_server_wait(qdt_gpu_server* s, uint64_t id, double* latency_ms);
/* Submit, start and done times (ms, steady clock) of a done request. */
QDT_API int qdt_gpu_server_times(qdt_gpu_server* s, uint64_t id, double* submit_ms, double* start_ms,
                                 double* done_ms);
QDT_API size_t qdt_gpu_server_triangles_size(qdt_gpu_server* s, uint64_t id);
QDT_API size_t qdt_gpu_server_unique_points(qdt_gpu_server* s, uint64_t id);
/* 3 * T uint32 request-local indices, counter-clockwise; returns T. */
QDT_API size_t qdt_gpu_server_copy_triangles(qdt_gpu_server* s, uint64_t id, uint32_t* out);
/* Forgets a done request. */
QDT_API void qdt_gpu_server_release(qdt_gpu_server* s, uint64_t id);
/* Device bytes the workers hold now (each worker's engine workspace and
 * input buffer, kept between batches; a running batch is charged the larger
 * of that and its estimate) and the largest such sum so far, counting both
 * the estimates and the memory really held.  Admission keeps the charges
 * within hbm_budget_bytes, releasing idle workers' memory when needed. */
QDT_API int qdt_gpu_server_memory(qdt_gpu_server* s, size_t* held, size_t* peak);
/* Batches run so far and the requests they held. */
QDT_API int qdt_gpu_server_counters(qdt_gpu_server* s, uint64_t* batches, uint64_t* requests);

#ifdef __cplusplus
}
#endif

#endif
