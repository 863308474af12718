// engine.hpp -- host-side interface of the GPU batch Delaunay engine.
//
// One Engine owns a HIP stream and a device workspace; it triangulates one
// point set at a time (qdt_execute semantics) from device-resident input and
// leaves the results on the device.  The C ABI (qdt_capi.cpp) drives it.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <string>
#include <vector>

namespace qdt
{

// Counters of one run, for reports and tests.
struct RunStats
{
    uint64_t orient, orient_exact, incircle, incircle_exact; // predicate calls / filter failures
    uint64_t extract_queue;                                   // extraction tests sent to the exact kernel
    uint32_t unique;                                          // distinct points
    uint32_t triangles;
    uint32_t levels;      // kd levels with merges
    uint32_t leaf_levels; // kd levels built inside LDS cells
    uint32_t block_level; // level of the LDS cells (0 = a whole set per cell)
    uint64_t cells;       // LDS cells of the run
    uint64_t cell_points; // distinct points inside them
    double device_ms;     // first kernel start .. last kernel end
    // with profiling on: time per stage (ms) and the merge launches
    double stage_ms[8];   // keys+sort y, sort x, dedupe, frame-1 sort, kd levels, leaves, merges, extract
    double merge_launch_ms_sum; // the merge kernel's launches (levels merged in one piece)
    uint32_t merge_launches;
    uint64_t merge_edges;       // edges those launches created
    double seam_level_ms_sum;   // levels cut into seam pieces (all their kernels)
    uint32_t seam_retries; // runs repeated with unsplit merges
    // seam pieces (of the first run): anchors certified, anchors inserted,
    // pieces run, pieces resumed, pieces stopped by a claimed vertex,
    // anchors with a clean dry run, anchors selected, rings rewritten
    uint64_t seam[8];
    uint32_t err_bits;
    uint64_t spare_records; // records drawn from the spare region (demand, may exceed it)
    uint64_t cells_exact;   // cells rebuilt with the exact predicates (a filter could not decide)
    // batched form with graphs on: this call was a replay of a captured
    // graph / the first of its shape (run eagerly, then captured) / a
    // replay whose guards failed (run eagerly instead)
    uint32_t graph_replay, graph_captured, graph_fallback;
    uint32_t cell_points_max, merge_lanes; // the tuned parameters of the run
};

enum class Status : int
{
    ok = 0,
    not_finite = -4,
    out_of_memory = -10,
    internal = -11,
};

// Results of a run over several point sets (Engine::run_batch).
struct BatchResult
{
    int k = 0;
    std::vector<int> status;          // per set: 0, or -4 for a set with a non-finite point
    std::vector<uint32_t> unique;     // per set: distinct points
    std::vector<uint64_t> tri_start;  // per set: first triangle in `tris`
    std::vector<uint64_t> tri_count;  // per set: triangles
    const uint32_t* tris = nullptr;   // device: 3 * triangles set-local input indices, grouped by set
    const uint32_t* rep = nullptr;    // device: n entries, set-local representative of every input point
    size_t n = 0;
    uint64_t triangles = 0;
};

struct DeviceResult
{
    const uint32_t* tris = nullptr; // 3 * triangles original indices
    const uint32_t* rep = nullptr;  // n entries
    uint32_t triangles = 0;
    uint32_t unique = 0;
    size_t n = 0;
};

// Result of the exact check of a triangle list (verify_triangulation).
struct VerifyReport
{
    uint64_t triangles = 0;
    uint64_t boundary_edges = 0;    // edges of one triangle only
    uint64_t vertices = 0;          // distinct vertices used
    uint64_t bad_index = 0;         // vertex index >= n, or repeated inside a triangle
    uint64_t not_representative = 0; // a vertex that is not its duplicate group's lowest index
    uint64_t bad_orientation = 0;   // triangles not strictly counter-clockwise (exact orient2d)
    uint64_t bad_edges = 0;         // edges of three or more triangles, twice the same direction, or a
                                    // boundary vertex with two outgoing boundary edges or none
    uint64_t non_delaunay = 0;      // interior edges whose opposite vertex lies strictly inside the
                                    // other triangle's circumcircle (exact incircle)
    uint64_t concave_boundary = 0;  // boundary vertices where the boundary turns clockwise
    uint64_t unused_points = 0;     // duplicate-group representatives used by no triangle
    uint64_t off_line = 0;          // no triangles: representatives off the line of the first two
    bool euler_ok = false;          // T = 2V - 2 - b
    bool ok = false;                // all of the above clean
    double device_ms = 0;
};

// Exact check, on the device, that tris (T triangles, 3 uint32 input indices
// each) is a Delaunay triangulation of the distinct points of pts (n
// points): orientation, edge manifoldness, every interior edge locally
// Delaunay, a convex boundary, the Euler count, every representative used
// (rep: lowest index of each point's duplicate group, or null to skip the
// representative checks).  Temporary device memory is allocated and freed.
Status verify_triangulation(const double* pts, uint32_t n, const uint32_t* tris, uint32_t T, const uint32_t* rep,
                            hipStream_t s, VerifyReport& out);

// Counts of a half-edge mesh build (build_mesh).
struct MeshReport
{
    uint64_t vertices = 0;  // local vertices (input indices some triangle uses)
    uint64_t halfedges = 0; // 3 T
    uint64_t boundary = 0;  // half-edges without a twin
    uint64_t unmatched = 0; // half-edges of edges in more than two triangles, or in two with the same direction
    uint64_t bad_index = 0; // triangles with an index >= n
    double device_ms = 0;
};

// Half-edge mesh of T triangles (3 T input indices < n, device memory), as
// built before the first insertion (GuibasStolfiDwyer_Wrapper.cpp:99-202).
// he: 3 T int4 {origin (local vertex), twin (-1: none), next, triangle};
// vert: int2 {input index, first half-edge out of it} per local vertex, in
// order of first appearance (triangle, then corner); orig_to_local: n ints
// (-1: unused).  Temporary device memory is allocated and freed.
Status build_mesh(const uint32_t* tris, uint32_t T, uint32_t n, int4* he, int2* vert, int* orig_to_local,
                  hipStream_t s, MeshReport& out);

// The twin of every half-edge of T triangles (3 T input indices, device
// memory), half-edge 3t+i running from corner i to corner (i+1)%3: twin
// (device, 3 T ints) gets the opposite half-edge, -1 on the boundary.
// unmatched counts half-edges of edges that are in more than two triangles
// or twice in the same direction.  Temporary device memory is allocated.
Status mesh_twins(const uint32_t* tris, uint32_t T, int* twin, hipStream_t s, uint64_t& unmatched);

class Engine
{
public:
    explicit Engine(hipStream_t stream = nullptr);
    // External workspace: every buffer is carved from [ext, ext + ext_bytes)
    // (device memory the caller owns), host staging lives in the object, no
    // events are recorded (device_ms stays 0) and work runs on `stream`
    // (null: the null stream), so that a run allocates and frees no memory;
    // a run that needs more than ext_bytes fails as out of memory.
    Engine(hipStream_t stream, void* ext, size_t ext_bytes);
    ~Engine();
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;

    // Triangulates n points at d_pts (device memory, (x, y) doubles).
    Status run(const double* d_pts, size_t n);
    const DeviceResult& result() const { return res_; }
    // Triangulates k independent point sets in one pass: set i is points
    // [set_off[i], set_off[i+1]) of d_pts (set_off on the host).
    Status run_batch(const double* d_pts, const size_t* set_off, int k);
    const BatchResult& batch() const { return batch_; }
    const RunStats& stats() const { return st_; }
    hipStream_t stream() const { return stream_; }
    // Frees the workspace (the next run allocates again).
    void release();
    // Tests only: cap the LDS leaf-block size (0 = default).
    void set_block_points(int c) { block_points_ = c; }
    // Records HIP events around every stage and every merge launch.
    void set_profiling(bool on) { profiling_ = on; }
    // Cuts large merges into seam pieces (default on).
    void set_seams(bool on) { seams_ = on; }
    // Batched form: capture each shape's pipeline in a HIP graph and replay
    // it for later batches of the same shape (default off).
    void set_graphs(bool on) { graphs_ = on; }
    bool graphs() const { return graphs_; }
    // The tunable launch parameters: the largest kd node built in one LDS
    // cell (128, 256 or 512 points; the cell kernel's LDS per wavefront and
    // so its wavefronts per CU follow from it) and the merges per wavefront
    // of the one-thread-per-node merge kernel (16, 32 or 64).
    void set_tuning(int cell_points, int merge_lanes);
    int cell_points() const { return cell_max_; }
    int merge_lanes() const { return merge_lanes_; }
    // Builds the bottom kd levels in LDS cells (default on).
    void set_cells(bool on) { cells_ = on; }
    // Device bytes currently held.
    size_t workspace_bytes() const;
    // Device bytes held after a run over n points in k sets (k = 0: a
    // single set), for admission by memory budget.
    size_t bytes_after(size_t n, size_t k) const;
    // Tests of the fallback paths: the device bytes the workspace may hold
    // (a larger request fails as out of memory; 0 = no limit) and the spare
    // quad-edge records seam pieces may draw (0 = the default region; too
    // few make a run repeat without seam pieces).
    void set_limits(size_t mem_bytes, size_t spare_records)
    {
        mem_cap_ = mem_bytes;
        spare_limit_ = spare_records;
    }

    struct Plan;

private:
    Status ensure(size_t n);
    Status ensure_ws(size_t n);
    Status ensure_nodes(const Plan& P);
    Status ensure_anchors(size_t seg_nodes, size_t anchor_cap);
    Status ensure_sets(size_t k, size_t n);
    Status run_impl(const double* d_pts, size_t n, const size_t* set_off, int k, bool seams, bool& redo);
    Status run_small(const double* d_pts, size_t n);
    Status batch_output(const uint32_t* dT, uint32_t T, int k, int kg, const std::vector<int>& set_rank,
                        const std::vector<uint32_t>& vstart);
    Status finish_batch(const uint32_t* toff, uint32_t T, int k, int kg, const std::vector<int>& set_rank,
                        const std::vector<uint32_t>& vstart);
    Status run_batch_eager(const double* d_pts, const size_t* set_off, int k);
    Status run_batch_impl(const double* d_pts, const size_t* set_off, int k);

    // Graph capture of the batched form.  A batch of a new shape runs
    // eagerly in record mode, which keeps every value a host decision was
    // taken on (a trace); the same run is then captured in replay mode,
    // where those values come from the trace and device guards compare them
    // with the values of the data at hand.  A replay whose guards fail is
    // run eagerly.
    enum class Mode
    {
        eager,
        record,
        replay
    };
    struct Trace;
    struct GraphEntry;
    Mode mode_ = Mode::eager;
    Trace* tr_ = nullptr;
    size_t tr_cursor_ = 0;
    bool graphs_ = false;
    std::vector<GraphEntry*> graphs_cache_;
    uint64_t ws_gen_ = 0; // bumped whenever a region is carved (cached graphs hold its pointers)
    int cell_max_ = 256, merge_lanes_ = 16;
    void drop_graphs();
    Status sort_keys(uint64_t*& kA, uint32_t*& vA, uint64_t*& kB, uint32_t*& vB, uint32_t n, uint32_t must = 0);
    Status upload(void* dst, const void* src, size_t bytes);
    Status replay_graph(GraphEntry* g, const double* d_pts, const size_t* set_off, int k, bool& ok);
    Status capture(GraphEntry* e, const double* d_pts, const size_t* set_off, int k);
    Status split_capture();
    GraphEntry* cap_ = nullptr; // the entry being captured
    char* ws_gin_ = nullptr; // the graphs' input buffer
    size_t ws_gin_bytes_ = 0;
    double* gin_ = nullptr;
    size_t gin_cap_ = 0;

    hipStream_t stream_;
    bool own_stream_;
    char* ws_ = nullptr;       // point-proportional buffers
    char* ws_nodes_ = nullptr; // kd forest nodes
    char* ws_anchors_ = nullptr; // seam pieces
    char* ws_sets_ = nullptr;  // per-set buffers of batches
    size_t ws_bytes_ = 0, ws_nodes_bytes_ = 0, ws_anchors_bytes_ = 0, ws_sets_bytes_ = 0;
    size_t cap_n_ = 0;
    int block_points_ = 0;
    bool profiling_ = false;
    bool seams_ = true;
    bool split_ = true; // eager sorts may take their top passes and a fix-up (cleared for a repeat)
    bool cells_ = true;
    size_t mem_cap_ = 0, spare_limit_ = 0;
    size_t alloc_budget() const;
    // external workspace (see the constructor): bump allocation, reset at
    // the start of every run
    char* ext_ = nullptr;
    size_t ext_bytes_ = 0, ext_used_ = 0, ext_ws_end_ = 0;
    uint32_t ext_scalars_[64] = {};
    int ext_table_[4 * 64] = {};
    uint32_t ext_hist_[2 * 8 * 256] = {}; // the sorts' global digit histograms, read back (external workspace)
    template <size_t N>
    Status carve_region(const struct Item (&items)[N], char*& region, size_t& bytes);
    void free_region(char*& region, size_t& bytes);
    std::vector<hipEvent_t> evs_; // profiling events, created on demand
    hipEvent_t prof_event(size_t i);
    DeviceResult res_;
    BatchResult batch_;
    RunStats st_{};
    hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
    // carved buffers (see ensure())
    struct Bufs;
    Bufs* b_ = nullptr;
};

} // namespace qdt
