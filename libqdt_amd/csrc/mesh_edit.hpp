// mesh_edit.hpp -- host-side edits of a triangulation: point insertion and
// constrained edges (qdt_insert_point*, qdt_insert_constraint*).
//
// The batch result (built on the GPU) is turned into a mesh of triangles
// whose half-edges are implicit: half-edge 3t+i runs from corner i of
// triangle t to corner (i+1)%3, so the corner array V is at the same time the
// output triangle list (original point indices, counter-clockwise) and the
// half-edges' origins.  Beside it: the twin of every half-edge (-1 on the
// convex hull), a constrained flag per half-edge, and one out-going
// half-edge per vertex.  Triangle slots are only ever rewritten or appended
// (an insertion replaces the one or two triangles it splits and appends the
// rest; a flip rewrites its two), so triangle ids stay dense without moving
// triangles around.
//
// The edits follow the reference's semantics (libqdt/GuibasStolfiDwyer_
// Wrapper.cpp): insertion locates the point by a walk, splits the triangle
// or the edge it lies on (or joins it to the visible hull edges), and
// restores the Delaunay property by Lawson flips that never flip a
// constrained edge (:543-876); a constraint is forced in by flipping away
// the edges it crosses (Sloan's method) and then legalised (:1050-1244).
// All geometric decisions use the exact predicates of predicates.hpp.
// Serial and latency-bound: this is host code.
#pragma once

#include "../../include/qdt.h"

#include <stddef.h>
#include <stdint.h>

#include <vector>

namespace qdt
{

class MeshEdit
{
public:
    // Takes a context's contents: n points, rep[i] = lowest index with the
    // same coordinates as point i, and T > 0 triangles (3 T indices, every
    // distinct point a corner).  twins (3 T entries) may come from the
    // device (the twin of every half-edge, -1 on the hull); when null they
    // are computed here.  Returns false when the triangles are not a
    // manifold triangulation (an edge in more than two triangles).
    bool build(const qdt_vec2d* pts, size_t n, const uint32_t* rep, const uint32_t* tris, size_t T,
               const int32_t* twins);

    // Inserts p (finite) as point number points(); sets index.  Returns
    // QDT_OK, or QDT_ERR_INTERNAL when the mesh is inconsistent (the caller
    // resets the context).  A duplicate of an existing vertex gets an index
    // but no triangle.
    qdt_status insert(const qdt_vec2d& p, size_t& index);

    // Forces the segment v0-v1 (point indices) into the triangulation as a
    // constrained edge, or rejects it (QDT_ERR_INVALID_INDEX ..
    // QDT_ERR_THROUGH_POINT, nothing changed).  QDT_ERR_INTERNAL: the
    // flipping did not end with the edge in place (the caller resets).
    qdt_status constrain(size_t v0, size_t v1);

    size_t points() const { return pts_.size(); }
    size_t triangles() const { return V_.size() / 3; }
    const uint32_t* tris() const { return V_.data(); }
    const qdt_vec2d* pts() const { return pts_.data(); }
    const uint32_t* rep() const { return rep_.data(); }
    // constrained edges (each once, as a half-edge id), for tests
    size_t constrained_edges() const;
    uint64_t flips() const { return flips_; }

private:
    std::vector<qdt_vec2d> pts_;
    std::vector<uint32_t> rep_;
    std::vector<uint32_t> V_; // corners = half-edge origins
    std::vector<int32_t> N_;  // twin half-edge, -1 on the hull
    std::vector<uint8_t> C_;  // constrained
    std::vector<int32_t> VH_; // one half-edge out of each vertex, -1: not a vertex
    int32_t last_ = 0;        // walk start
    uint64_t rng_ = 0x9e3779b97f4a7c15ull;
    uint64_t flips_ = 0;

    static int32_t nxt(int32_t h) { return h - h % 3 + (h % 3 + 1) % 3; }
    static int32_t prv(int32_t h) { return h - h % 3 + (h % 3 + 2) % 3; }
    uint32_t org(int32_t h) const { return V_[h]; }
    uint32_t dst(int32_t h) const { return V_[nxt(h)]; }
    int orient(uint32_t a, uint32_t b, const qdt_vec2d& c) const;
    int orient(uint32_t a, uint32_t b, uint32_t c) const { return orient(a, b, pts_[c]); }
    int incircle(uint32_t a, uint32_t b, uint32_t c, uint32_t d) const;

    void link(int32_t h, int32_t g);
    int32_t new_triangle();
    void write_tri(int32_t t, uint32_t a, uint32_t b, uint32_t c);
    int32_t locate(const qdt_vec2d& p, int32_t& hull_he);
    int32_t locate_scan(const qdt_vec2d& p, int32_t& hull_he) const;
    int32_t next_hull(int32_t h) const;
    int32_t prev_hull(int32_t h) const;
    template <class F>
    void for_each_out(uint32_t v, F&& f) const;
    int32_t find_halfedge(uint32_t u, uint32_t v) const;
    bool flip(int32_t h);
    void mark(int32_t h);
    qdt_status trace(uint32_t u, uint32_t v, std::vector<int32_t>& crossed) const;
    void legalize(std::vector<int32_t>& stack);
    bool segments_cross(uint32_t a, uint32_t b, uint32_t c, uint32_t d) const;
};

} // namespace qdt
