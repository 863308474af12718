// seam.hpp -- cutting one large Guibas-Stolfi merge into independent pieces.
//
// A merge of two triangulations L and R (separated in the node's frame)
// produces a chain of cross edges from the lower to the upper common tangent.
// The rising bubble walks that chain one edge at a time, so a top-level merge
// of millions of points is a single thread chasing ~sqrt(n) pointers.  Any
// edge that is in *every* Delaunay triangulation of L u R cuts the chain: the
// bubble below it and the bubble above it touch disjoint parts of the store.
//
//  1. anchors   one thread per (node, k): a point q on the cut line at the
//               k-th height; greedy walks find q's nearest vertices u in L and
//               v in R (floating point is fine here); q moves to where the
//               bisector of u and v meets its height until (u, v) is stable.
//  2. certify   exactly: circles through u and v are parametrised by their
//               centre on the bisector; each Delaunay neighbour w of u (in L)
//               or of v (in R) forbids the centres on one side of the
//               circumcentre of (u, v, w).  (u, v) is in every Delaunay
//               triangulation of L u R iff every neighbour left of u->v and
//               every neighbour right of it are strictly separated:
//               incircle(u, v, wl, wr) < 0 for all such pairs (u, v, wl
//               counter-clockwise), and no neighbour is collinear with u, v.
//               (Vor_L(u) is cut out by u's Delaunay neighbours in L, Vor_R(v)
//               by v's in R, so these constraints are all there are.)  Only
//               the extreme neighbour of each side needs testing.
//  3. insert    the edges of L and of R crossing each anchor are found by a
//               straight walk (dry run; any degeneracy drops the anchor).
//               Selected anchors (ordered, vertex-disjoint) are inserted in
//               parallel by rewriting the ring of every vertex they touch.
//  4. segments  one thread per piece, all pieces of a node in one
//               workgroup, in two phases (even pieces, then odd, a barrier
//               between): the rising bubble from anchor k stops when anchor
//               k+1 is the next cross edge (QuadEdges::bubble).  A piece
//               claims each vertex before touching its ring and stops at a
//               vertex another piece holds; stopped pieces are finished
//               afterwards by one thread per node.
#pragma once

#include "quadedge.hpp"

namespace qdt
{

// frame coordinates: frame 0 is (x, y); frame 1 is (y, -x), a rotation that
// turns the frame-1 order of builder.hxx:75 into lexicographic (X, Y) order
__device__ __forceinline__ double frame_x(double2 p, int axis) { return axis ? p.y : p.x; }
__device__ __forceinline__ double frame_y(double2 p, int axis) { return axis ? -p.x : p.y; }
__device__ __forceinline__ double2 from_frame(double X, double Y, int axis)
{
    return axis ? make_double2(-Y, X) : make_double2(X, Y);
}

template <class Q>
__device__ __forceinline__ double dist2(Q& q, int v, double2 p)
{
    const double2 a = q.pt(v);
    const double dx = a.x - p.x, dy = a.y - p.y;
    return dx * dx + dy * dy;
}

// Greedy walk to the vertex nearest to p (floating point).  Returns the
// vertex and, in e_out, an edge out of it; -1 when the start has no edge.
template <class Q>
__device__ int nn_walk(Q& q, int s, double2 p, int budget, int& e_out)
{
    int e0 = q.VE[s];
    if (e0 < 0) return -1;
    int best = s;
    double bd = dist2(q, s, p);
    for (int step = 0; step < budget && !q.err; ++step)
    {
        int nb = -1, ne = -1;
        double nbd = bd;
        int guard = 0;
        // ring scan, pipelined: the next edge and its far end are requested
        // before the current end's coordinates are used, so the scan waits
        // for one load per edge instead of two
        int e = e0, w = q.dest(e0), en = q.onext(e0);
        for (;;)
        {
            const double2 W = q.pt(w);
            const int wn = q.dest(en), enn = q.onext(en);
            const double dx = W.x - p.x, dy = W.y - p.y;
            const double d = dx * dx + dy * dy;
            if (d < nbd)
            {
                nbd = d;
                nb = w;
                ne = e ^ 2;
            }
            e = en;
            if (e == e0 || ++guard >= 4096 || q.err) break;
            w = wn;
            en = enn;
        }
        if (nb < 0)
        {
            e_out = e0;
            return best;
        }
        best = nb;
        bd = nbd;
        e0 = ne;
    }
    e_out = e0;
    return best;
}

// Exact certificate that (u, v) is an edge of every Delaunay triangulation of
// L u R (u in L, v in R).  The circles through u and v form a one-parameter
// family; every neighbour w has the member c(w) through it.  The condition
// "incircle(u, v, wl, wr) < 0 for every left neighbour wl and right
// neighbour wr" is max c(right) < min c(left), so it suffices to find the
// extreme neighbour of each side by a streaming scan (the side's point inside
// the circles of all the others; comparisons are the same exact incircle
// tests) and test that one pair.  A circle through u and v meets their line
// at u and v only, so a neighbour on that line lies strictly inside every
// such circle when it is between u and v (the anchor is rejected) and
// strictly outside every one otherwise (it constrains nothing; lattices
// give such neighbours at every edge).  Why the neighbours suffice: a circle
// through u that holds a point of L inside (or on it) shrunk toward u keeps
// u on it and first meets a Delaunay neighbour of u in L, which is then
// inside (or on) the original circle; the same holds for v and R.
// certify_ring scans the ring of e0's origin (u's ring or v's ring; both are
// needed), certify_join combines the two scans.
template <class Q>
__device__ bool certify_ring(Q& q, int u, int v, int e0, int& wl, int& wr)
{
    constexpr int kMax = 64; // ring length bound
    int e = e0, guard = 0;
    int w = q.dest(e0), en = q.onext(e0);
    do
    {
        // the next edge and its far end are requested before this one's
        // tests (pipelined ring scan)
        const int wn = q.dest(en), enn = q.onext(en);
        if (w != u && w != v)
        {
            const double o = q.orient(u, v, w);
            if (o == 0)
            {
                // on the line of u and v: between them (in the coordinate
                // that varies along the line) rejects, outside is no constraint
                const double2 U = q.pt(u), V = q.pt(v), X = q.pt(w);
                const bool byx = U.x != V.x;
                const double a = byx ? U.x : U.y, b = byx ? V.x : V.y, c = byx ? X.x : X.y;
                if ((a < c && c < b) || (b < c && c < a)) return false;
            }
            else if (o > 0)
            {
                // left: keep the one inside the others' circles (u, v, w ccw)
                if (wl < 0 || q.incircle_v(u, v, wl, w) > 0) wl = w;
            }
            else
            {
                // right: (v, u, w) is ccw
                if (wr < 0 || q.incircle_v(v, u, wr, w) > 0) wr = w;
            }
        }
        e = en;
        w = wn;
        en = enn;
    } while (e != e0 && ++guard < kMax && !q.err);
    return e == e0;
}

// The extreme neighbours of two scans (of u's ring and of v's ring) combined
// with the same test, then the one pair that decides.  The circles through u
// and v are ordered by their centres along the bisector, so the extreme of
// the union is the more extreme of the two partial extremes; where two
// neighbours tie (cocircular with u and v) either one gives the same answer.
template <class Q>
__device__ bool certify_join(Q& q, int u, int v, int wl0, int wr0, int wl1, int wr1)
{
    int wl = wl0, wr = wr0;
    if (wl1 >= 0 && (wl < 0 || q.incircle_v(u, v, wl, wl1) > 0)) wl = wl1;
    if (wr1 >= 0 && (wr < 0 || q.incircle_v(v, u, wr, wr1) > 0)) wr = wr1;
    if (wl >= 0 && wr >= 0 && !(q.incircle_v(u, v, wl, wr) < 0)) return false;
    return !q.err;
}

// The edge x out of u such that direction u->p lies strictly inside the
// counter-clockwise wedge from x to onext(x); -1 when p's direction
// coincides with an edge or the ring is malformed.
template <class Q>
__device__ int find_wedge(Q& q, int u, int e0, int p)
{
    const int n1 = q.onext(e0);
    if (n1 == e0) return e0; // single edge: every other direction is inside
    const double2 U = q.pt(u), P = q.pt(p);
    // the ring in order, each edge's far end and coordinates carried to the
    // next step; the onext after that is requested before the tests
    int e = e0, n = n1, guard = 0;
    int a = q.dest(e0), b = q.dest(n1);
    double2 A = q.pt(a), B = q.pt(b);
    do
    {
        const int nn = q.onext(n);
        const double oap = q.pred_orient(U, A, P);
        const double opb = q.pred_orient(U, P, B);
        if (oap == 0 && q.pred_orient(U, A, B) != 0)
        {
            // collinear with edge e: same direction is a coincidence, the
            // opposite direction lies in some wedge (handled below)
            if ((A.x - U.x) * (P.x - U.x) + (A.y - U.y) * (P.y - U.y) > 0) return -1;
        }
        const double oab = q.pred_orient(U, A, B);
        bool inside;
        if (oab > 0)
            inside = oap > 0 && opb > 0;
        else
            inside = oap > 0 || opb > 0; // reflex or straight wedge
        if (inside) return e;
        e = n;
        a = b;
        A = B;
        n = nn;
        b = q.dest(n);
        B = q.pt(b);
    } while (e != e0 && ++guard < 256 && !q.err);
    return -1;
}

// Kind of the face left of e: 1 a counter-clockwise triangle, 0 the outer
// face (its boundary runs clockwise, or is longer than any face the anchor
// insertions can create), -1 a larger inner face (left by an earlier
// anchor's deletions) or a degenerate one.
template <class Q>
__device__ int face_kind(Q& q, int e)
{
    constexpr int kLimit = 48;
    const int f = q.lnext(e), g = q.lnext(f);
    if (q.lnext(g) == e)
    {
        const double o = q.orient(q.org(e), q.org(f), q.org(g));
        return o > 0 ? 1 : (o < 0 ? 0 : -1);
    }
    // shoelace area of the boundary, relative to its first vertex
    const double2 p0 = q.pt(q.org(e));
    double area = 0;
    int h = f, n = 1;
    double2 prev = q.pt(q.org(f));
    while (h != e && n < kLimit && !q.err)
    {
        h = q.lnext(h);
        const double2 cur = q.pt(q.org(h));
        area += (prev.x - p0.x) * (cur.y - p0.y) - (prev.y - p0.y) * (cur.x - p0.x);
        prev = cur;
        ++n;
    }
    if (h != e) return 0; // longer than any inner face an insertion leaves
    return area < 0 ? 0 : -1;
}

// Edges of one triangulation crossed by the segment u -> p (u a vertex of it,
// p outside its hull).  x is u's wedge edge for direction p.  Dry run: fills
// out[] and returns the count, or -1 on any degeneracy (the segment through a
// vertex, p inside a face, a face enlarged by an earlier anchor, too many
// crossings).
template <class Q>
__device__ int collect_crossings(Q& q, int u, int x, int p, int* out, int cap)
{
    int n = 0;
    const int ox = q.onext(x);
    if (ox == x) return 0; // u has one edge: nothing around it to cross
    const double2 U = q.pt(u), P = q.pt(p);
    // the face left of x, between x and onext(x): a triangle (u, a, b).  Its
    // corners are read as dest of its edges, so that the corner loads and
    // the closing test of the 3-cycle are in flight together.
    int f = q.lnext(x);
    int g = q.lnext(f);
    int a = q.dest(x), b = q.dest(f);
    double2 A = q.pt(a), B = q.pt(b);
    {
        const bool tri = q.lnext(g) == x;
        const double o = q.pred_orient(U, A, B);
        // Not a 3-cycle: the halves are complete triangulations whose bounded
        // faces are counter-clockwise triangles, so this is the outer face,
        // whose boundary turns clockwise at a corner of the hull; only a
        // straight turn (collinear points) needs the walk of face_kind.
        if (!tri) return (o < 0 || face_kind(q, x) == 0) ? 0 : -1;
        if (o < 0) return 0;  // the outer face
        if (o == 0) return -1;
        if (ox != (g ^ 2)) return -1;
        // f runs a -> b opposite u; the segment leaves through it
        if (q.pred_orient(A, B, P) >= 0) return -1; // p inside the triangle or on its edge
    }
    int e = f; // runs a -> b with a right of u->p and b left of it
    for (int guard = 0; guard < cap + 4; ++guard)
    {
        if (n >= cap) return -1;
        out[n++] = e;
        // the face beyond e, left of h = b -> a: a triangle (b, a, c) with
        // f = a -> c and g = c -> b, or the outer face (a hull edge: done)
        const int h = e ^ 2;
        f = q.lnext(h);
        g = q.lnext(f);
        const int c = q.dest(f);
        const double2 C = q.pt(c);
        const bool tri = q.lnext(g) == h;
        const double o = q.pred_orient(B, A, C);
        if (!tri) return (o < 0 || face_kind(q, h) == 0) ? n : -1; // as for the first face
        if (o < 0) return n; // the outer face
        if (o == 0) return -1;
        const double oc = q.pred_orient(U, P, C);
        if (oc == 0) return -1; // the segment passes through c
        if (oc > 0)
        {
            e = f; // a -> c crosses (a right, c left)
            b = c;
            B = C;
        }
        else
        {
            e = g; // c -> b crosses (c right, b left)
            a = c;
            A = C;
        }
        if (q.pred_orient(A, B, P) >= 0) return -1; // p inside this triangle
    }
    return -1;
}

} // namespace qdt
