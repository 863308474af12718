/*
 * qdt.h -- C interface of the MI355X (gfx950) batch Delaunay library.
 *
 * Drop-in for libqdt's public header: the same types, status codes and
 * fourteen functions with the same meaning (reference: libqdt/qdt.h:38-401).
 * The batch calls run on the GPU; results live in host memory owned by the
 * context, so the query functions are plain O(1) reads.
 *
 * Each declaration names the reference declaration it replaces.
 */
#ifndef QDT_MI355X_QDT_H
#define QDT_MI355X_QDT_H

#include <stddef.h>

#define QDT_API __attribute__((visibility("default")))

#ifdef __cplusplus
extern "C" {
#endif

/* A 2D point with double coordinates; coordinates must be finite.
 * Replaces libqdt/qdt.h:38-43. */
struct qdt_vec2d
{
    double x;
    double y;
};
typedef struct qdt_vec2d qdt_vec2d;

/* End points of a constraint, as point indices.  Replaces libqdt/qdt.h:48-53. */
struct qdt_edge
{
    size_t v0;
    size_t v1;
};
typedef struct qdt_edge qdt_edge;

/* Status of a call that can fail: 0 on success, negative on error.
 * Replaces libqdt/qdt.h:79-93 (same names and values). */
typedef enum qdt_status
{
    QDT_OK = 0,
    QDT_ERR_NULL_CONTEXT = -1,
    QDT_ERR_NULL_ARGUMENT = -2,
    QDT_ERR_TOO_MANY_POINTS = -3,
    QDT_ERR_NOT_FINITE = -4,
    QDT_ERR_INVALID_INDEX = -5,
    QDT_ERR_SAME_POINT = -6,
    QDT_ERR_NO_TRIANGLES = -7,
    QDT_ERR_CROSSES_CONSTRAINT = -8,
    QDT_ERR_THROUGH_POINT = -9,
    QDT_ERR_OUT_OF_MEMORY = -10,
    QDT_ERR_INTERNAL = -11
} qdt_status;

/* Short description of a status ("unknown status" for other values).
 * Replaces libqdt/qdt.h:102 (texts of libqdt/qdt.cpp:145-163). */
QDT_API const char* qdt_status_string(qdt_status status);

/* Opaque context.  One context must not be used by two threads at once;
 * distinct contexts are independent.  Replaces libqdt/qdt.h:125. */
typedef struct qdt_t qdt_t;

/* New empty context, or NULL when memory runs out.  Touches no GPU.
 * Replaces libqdt/qdt.h:133. */
QDT_API qdt_t* qdt_create(void);

/* Delaunay triangulation of n points (copied).  Checks, in this order:
 * n > 100000000 -> QDT_ERR_TOO_MANY_POINTS, pts NULL with n > 0 ->
 * QDT_ERR_NULL_ARGUMENT, a NaN/infinite coordinate -> QDT_ERR_NOT_FINITE;
 * every rejection empties the context.  n = 0 empties it and returns QDT_OK.
 * Replaces libqdt/qdt.h:160. */
QDT_API qdt_status qdt_execute(qdt_t* ptr, const qdt_vec2d* pts, size_t n);

/* Same result as qdt_execute() (the GPU path is already parallel).
 * Replaces libqdt/qdt.h:181. */
QDT_API qdt_status qdt_execute_parallel(qdt_t* ptr, const qdt_vec2d* pts, size_t n);

/* Appends one point and updates the triangulation.  Replaces libqdt/qdt.h:219. */
QDT_API qdt_status qdt_insert_point(qdt_t* ptr, qdt_vec2d pt, size_t* index);

/* Same as qdt_insert_point().  Replaces libqdt/qdt.h:235. */
QDT_API qdt_status qdt_insert_point_parallel(qdt_t* ptr, qdt_vec2d pt, size_t* index);

/* Forces the segment v0-v1 into the triangulation.  Replaces libqdt/qdt.h:284. */
QDT_API qdt_status qdt_insert_constraint(qdt_t* ptr, size_t v0, size_t v1);

/* Inserts constraints one after another.  Replaces libqdt/qdt.h:314. */
QDT_API qdt_status qdt_insert_constraints_batch(qdt_t* ptr, const qdt_edge* edges, size_t num_edges,
                                                qdt_status* results);

/* Number of triangles (0 for NULL).  Replaces libqdt/qdt.h:326. */
QDT_API size_t qdt_triangles_size(const qdt_t* ptr);

/* Vertex indices of triangle id, counter-clockwise; 1 on success, 0 (T
 * untouched) otherwise.  Replaces libqdt/qdt.h:348. */
QDT_API size_t qdt_get_triangle(const qdt_t* ptr, size_t T[3], size_t id);

/* Number of points, duplicates included (0 for NULL).  Replaces libqdt/qdt.h:362. */
QDT_API size_t qdt_points_size(const qdt_t* ptr);

/* Coordinates of point id; 1 on success, 0 (pt untouched) otherwise.
 * Replaces libqdt/qdt.h:375. */
QDT_API size_t qdt_get_point(const qdt_t* ptr, qdt_vec2d* pt, size_t id);

/* 1 when a triangle references point id.  Replaces libqdt/qdt.h:394. */
QDT_API size_t qdt_point_is_used(const qdt_t* ptr, size_t id);

/* Releases a context; NULL is allowed.  Replaces libqdt/qdt.h:401. */
QDT_API void qdt_destroy(qdt_t* ptr);

#ifdef __cplusplus
}
#endif

#endif
