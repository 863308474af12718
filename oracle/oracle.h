/*
 * oracle.h -- CPU restatement of libqdt's batch Delaunay path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links or loads
 * this code.  It exists so that tests/, the smoke check and the bench's CPU
 * baseline leg have an independent checker for the GPU path.
 *
 * What it restates (reference file:line in each function's comment):
 *   - robust predicates: libqdt/robust.hxx:281-396 (FP64 filter with the same
 *     static bounds and floor, power-of-two normalisation, exact fallback),
 *   - duplicate removal: libqdt/builder.hxx:356-390,
 *   - Guibas-Stolfi divide and conquer with Dwyer's alternating cuts:
 *     libqdt/builder.hxx:71-348,
 *   - triangle extraction: libqdt/builder.hxx:98-116.
 * Plus an exact checker (restated from the checks of test/qdt_verify.cpp:535-721)
 * and canonicalisation / hashing helpers.
 *
 * Pinning (see tests/test_oracle.py): predicate signs against integer closed
 * forms and against the reference's own predicates (oracle/_ref shim); the
 * triangulation against the reference's published triangle counts, its README
 * examples, and the compiled reference's canonical triangle sets.
 */
#ifndef QDT_ORACLE_H
#define QDT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Predicates, argument order (ax, ay, bx, by, ...).  The return value's sign
 * is the answer, as in the reference (filter value or +-1/0 from exact). */
double orc_orient2d(double ax, double ay, double bx, double by, double cx, double cy);
double orc_incircle(double ax, double ay, double bx, double by, double cx, double cy,
                    double dx, double dy);
/* Exact path only (always runs the expansion arithmetic). */
int orc_orient2d_exact(double ax, double ay, double bx, double by, double cx, double cy);
int orc_incircle_exact(double ax, double ay, double bx, double by, double cx, double cy,
                       double dx, double dy);

/* Batch triangulation of n points (xy interleaved).  Returns NULL only when
 * memory runs out. */
typedef struct orc_result orc_result;
orc_result* orc_triangulate(const double* xy, size_t n);
size_t orc_ntri(const orc_result* r);
const uint32_t* orc_tris(const orc_result* r); /* 3 * ntri original indices, CCW */
const uint32_t* orc_rep(const orc_result* r);  /* n entries: lowest index with equal coordinates */
size_t orc_unique(const orc_result* r);
/* stats[0..5]: orient2d calls, orient2d exact, incircle calls, incircle exact,
 * quad-edge records allocated, records live at the end */
void orc_stats(const orc_result* r, uint64_t stats[6]);
void orc_free(orc_result* r);

/* Rotate every triangle to its smallest index (orientation kept) and sort the
 * list, in place (test/qdt_verify.cpp:408-427). */
void orc_canonicalize(uint32_t* tris, size_t ntri);
/* FNV-1a 64 over the canonical list, each index as 8 little-endian bytes. */
uint64_t orc_canon_hash(const uint32_t* tris, size_t ntri);

/* Convex hull of the distinct points, strictly convex (no collinear points),
 * counter-clockwise, as lowest-duplicate indices.  Returns the count. */
size_t orc_hull(const double* xy, size_t n, uint32_t* out);

/* Exact validity check of a triangulation of xy (restates the checks of
 * test/qdt_verify.cpp:593-696): strict CCW, lowest-duplicate indices, no
 * repeated directed edge, every distinct point used, boundary edges on the
 * convex hull, Euler count T = 2m - 2 - b, every interior edge locally
 * Delaunay (exact incircle <= 0).  Returns 0 when valid, else a nonzero code
 * and a message.  info[0..3] = m distinct, b boundary vertices, T, E. */
int orc_check(const double* xy, size_t n, const uint32_t* tris, size_t ntri,
              char* msg, size_t msglen, int64_t info[4]);

/* Half-edge mesh of a triangle list, restating Triangulation::ensure_mesh_
 * (libqdt/GuibasStolfiDwyer_Wrapper.cpp:99-202): local vertices in order of
 * first appearance (triangle, then corner), half-edge 3t+i from corner i to
 * corner i+1 with next 3t+(i+1)%3 and triangle t, twins found by scanning the
 * out-going half-edges of the far vertex (CSR adjacency by origin).
 * he: 4 * 3 ntri int32 {origin, twin, next, triangle}; vert: 2 * (vertices)
 * int32 {input index, half-edge}, room for min(n, 3 ntri); o2l: n int32
 * (-1: unused).  Returns the number of local vertices, or -1 for an index
 * >= n. */
long orc_build_mesh(const uint32_t* tris, size_t ntri, size_t n, int32_t* he, int32_t* vert, int32_t* o2l);

#ifdef __cplusplus
}
#endif

#endif
