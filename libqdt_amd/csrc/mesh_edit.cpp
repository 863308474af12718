// mesh_edit.cpp -- point insertion and constrained edges on the host (see
// mesh_edit.hpp for the layout).
#include "mesh_edit.hpp"

#include "predicates.hpp"

#include <algorithm>
#include <deque>
#include <utility>

namespace qdt
{

namespace
{

int sgn(double d) { return d > 0 ? 1 : (d < 0 ? -1 : 0); }

// Do the collinear vectors a->b and a->c point the same way?  Signs of the
// coordinate differences only (a rounded difference has the exact one's
// sign), so nothing can underflow.
bool same_direction(const qdt_vec2d& a, const qdt_vec2d& b, const qdt_vec2d& c)
{
    const int sx = sgn(b.x - a.x) * sgn(c.x - a.x);
    const int sy = sgn(b.y - a.y) * sgn(c.y - a.y);
    return sx >= 0 && sy >= 0 && (sx > 0 || sy > 0);
}

// one edge of the polygon around a new point: from -> to, the half-edge on
// its far side (-1: hull) and its constrained flag
struct RingEdge
{
    uint32_t from, to;
    int32_t twin;
    uint8_t cons;
};

} // namespace

int MeshEdit::orient(uint32_t a, uint32_t b, const qdt_vec2d& c) const
{
    const qdt_vec2d& A = pts_[a];
    const qdt_vec2d& B = pts_[b];
    unsigned ne = 0;
    return sgn(qdt::orient2d(A.x, A.y, B.x, B.y, c.x, c.y, ne));
}

int MeshEdit::incircle(uint32_t a, uint32_t b, uint32_t c, uint32_t d) const
{
    const qdt_vec2d &A = pts_[a], &B = pts_[b], &C = pts_[c], &D = pts_[d];
    unsigned ne = 0;
    return sgn(qdt::incircle(A.x, A.y, B.x, B.y, C.x, C.y, D.x, D.y, ne));
}

bool MeshEdit::build(const qdt_vec2d* pts, size_t n, const uint32_t* rep, const uint32_t* tris, size_t T,
                     const int32_t* twins)
{
    pts_.assign(pts, pts + n);
    rep_.assign(rep, rep + n);
    V_.assign(tris, tris + 3 * T);
    C_.assign(3 * T, 0);
    VH_.assign(n, -1);
    const int32_t E = (int32_t)(3 * T);
    for (int32_t h = 0; h < E; ++h)
    {
        if (V_[h] >= n) return false;
        VH_[V_[h]] = h;
    }
    if (twins)
        N_.assign(twins, twins + 3 * T);
    else
    {
        // half-edges grouped by origin (CSR); the twin of a -> b is the
        // half-edge out of b that ends at a
        std::vector<uint32_t> off(n + 1, 0);
        for (int32_t h = 0; h < E; ++h) ++off[V_[h] + 1];
        for (size_t v = 0; v < n; ++v) off[v + 1] += off[v];
        std::vector<int32_t> out(E);
        {
            std::vector<uint32_t> at(off.begin(), off.end() - 1);
            for (int32_t h = 0; h < E; ++h) out[at[V_[h]]++] = h;
        }
        N_.assign(E, -1);
        for (int32_t h = 0; h < E; ++h)
        {
            const uint32_t a = org(h), b = dst(h);
            int found = 0;
            for (uint32_t k = off[b]; k < off[b + 1]; ++k)
            {
                const int32_t g = out[k];
                if (dst(g) != a) continue;
                if (++found > 1) return false; // the same edge in three or more triangles
                N_[h] = g;
            }
            if (N_[h] >= 0 && N_[N_[h]] >= 0 && N_[N_[h]] != h) return false;
        }
    }
    last_ = 0;
    return true;
}

void MeshEdit::link(int32_t h, int32_t g)
{
    N_[h] = g;
    if (g >= 0) N_[g] = h;
}

int32_t MeshEdit::new_triangle()
{
    const int32_t t = (int32_t)triangles();
    V_.resize(V_.size() + 3, 0);
    N_.resize(N_.size() + 3, -1);
    C_.resize(C_.size() + 3, 0);
    return t;
}

void MeshEdit::write_tri(int32_t t, uint32_t a, uint32_t b, uint32_t c)
{
    V_[3 * t] = a;
    V_[3 * t + 1] = b;
    V_[3 * t + 2] = c;
    VH_[a] = 3 * t;
    VH_[b] = 3 * t + 1;
    VH_[c] = 3 * t + 2;
}

// Walk through the triangles towards p, leaving each one through an edge
// that has p strictly on its outer side, tried from a random first edge
// (a stochastic visibility walk: it ends in any triangulation, constrained
// ones included).  Returns a half-edge of the triangle holding p (boundary
// included), or -1 with hull_he set to a hull half-edge that has p strictly
// outside.
int32_t MeshEdit::locate(const qdt_vec2d& p, int32_t& hull_he)
{
    hull_he = -1;
    const int32_t T = (int32_t)triangles();
    if (!T) return -1;
    int32_t t = last_ >= 0 && last_ < T ? last_ : 0;
    int32_t from = -1; // the half-edge the walk came in through
    const size_t limit = 4 * (size_t)T + 64;
    for (size_t step = 0; step < limit; ++step)
    {
        rng_ ^= rng_ << 13;
        rng_ ^= rng_ >> 7;
        rng_ ^= rng_ << 17;
        const int r = (int)(rng_ % 3);
        bool moved = false;
        for (int k = 0; k < 3; ++k)
        {
            const int32_t h = 3 * t + (r + k) % 3;
            if (h == from) continue;
            if (orient(org(h), dst(h), p) < 0)
            {
                const int32_t g = N_[h];
                if (g < 0)
                {
                    hull_he = h;
                    return -1;
                }
                from = g;
                t = g / 3;
                moved = true;
                break;
            }
        }
        if (!moved)
        {
            last_ = t;
            return 3 * t;
        }
    }
    return locate_scan(p, hull_he);
}

int32_t MeshEdit::locate_scan(const qdt_vec2d& p, int32_t& hull_he) const
{
    hull_he = -1;
    const int32_t T = (int32_t)triangles();
    for (int32_t t = 0; t < T; ++t)
        if (orient(V_[3 * t], V_[3 * t + 1], p) >= 0 && orient(V_[3 * t + 1], V_[3 * t + 2], p) >= 0 &&
            orient(V_[3 * t + 2], V_[3 * t], p) >= 0)
            return 3 * t;
    for (int32_t h = 0; h < 3 * T; ++h)
        if (N_[h] < 0 && orient(org(h), dst(h), p) < 0)
        {
            hull_he = h;
            return -1;
        }
    return -1;
}

// the hull half-edge that starts where hull half-edge h ends / that ends
// where h starts (rotating around the shared vertex through the interior)
int32_t MeshEdit::next_hull(int32_t h) const
{
    int32_t x = nxt(h);
    for (size_t guard = 0; N_[x] >= 0 && guard < N_.size(); ++guard) x = nxt(N_[x]);
    return x;
}

int32_t MeshEdit::prev_hull(int32_t h) const
{
    int32_t x = prv(h);
    for (size_t guard = 0; N_[x] >= 0 && guard < N_.size(); ++guard) x = prv(N_[x]);
    return x;
}

qdt_status MeshEdit::insert(const qdt_vec2d& p, size_t& index)
{
    int32_t hull = -1;
    const int32_t h0 = locate(p, hull);
    if (h0 < 0 && hull < 0) return QDT_ERR_INTERNAL;
    const uint32_t id = (uint32_t)pts_.size();
    if (h0 >= 0)
    {
        // a vertex is never inside a triangle or inside one of its edges,
        // so a copy of a vertex finds that vertex among the corners
        for (int k = 0; k < 3; ++k)
        {
            const uint32_t v = V_[h0 + k];
            if (pts_[v].x == p.x && pts_[v].y == p.y)
            {
                pts_.push_back(p);
                rep_.push_back(v);
                VH_.push_back(-1);
                index = id;
                return QDT_OK;
            }
        }
    }
    pts_.push_back(p);
    rep_.push_back(id);
    VH_.push_back(-1);
    index = id;

    // the polygon around p (counter-clockwise) and the slots it frees
    std::vector<RingEdge> ring;
    int32_t slots[2];
    int nslots = 0;
    bool closed = true;
    uint32_t split_a = ~0u, split_b = ~0u; // a constrained edge that p splits
    auto edge_of = [&](int32_t h) { return RingEdge{ org(h), dst(h), N_[h], C_[h] }; };
    if (h0 >= 0)
    {
        const int32_t t = h0 / 3;
        int z = -1;
        for (int k = 0; k < 3; ++k)
            if (orient(V_[3 * t + k], V_[3 * t + (k + 1) % 3], p) == 0) z = k;
        slots[nslots++] = t;
        if (z < 0)
        {
            for (int k = 0; k < 3; ++k) ring.push_back(edge_of(3 * t + k));
        }
        else
        {
            const int32_t e = 3 * t + z, g = N_[e];
            if (C_[e])
            {
                split_a = org(e);
                split_b = dst(e);
            }
            ring.push_back(edge_of(nxt(e)));
            ring.push_back(edge_of(prv(e)));
            if (g >= 0)
            {
                ring.push_back(edge_of(nxt(g)));
                ring.push_back(edge_of(prv(g)));
                slots[nslots++] = g / 3;
            }
            else
                closed = false;
        }
    }
    else
    {
        // outside the hull: the hull edges p sees, in hull order
        std::vector<int32_t> fwd, back;
        for (int32_t x = next_hull(hull); x != hull && orient(org(x), dst(x), p) < 0; x = next_hull(x))
        {
            fwd.push_back(x);
            if (fwd.size() > N_.size()) return QDT_ERR_INTERNAL;
        }
        for (int32_t x = prev_hull(hull); x != hull && orient(org(x), dst(x), p) < 0; x = prev_hull(x))
        {
            back.push_back(x);
            if (back.size() > N_.size()) return QDT_ERR_INTERNAL;
        }
        std::vector<int32_t> chain(back.rbegin(), back.rend());
        chain.push_back(hull);
        chain.insert(chain.end(), fwd.begin(), fwd.end());
        // around p the polygon runs against the hull's direction
        for (size_t j = chain.size(); j-- > 0;)
            ring.push_back(RingEdge{ dst(chain[j]), org(chain[j]), chain[j], C_[chain[j]] });
        closed = false;
    }

    // the fan of triangles (ring edge k, p)
    const size_t K = ring.size();
    std::vector<int32_t> tri(K);
    for (size_t k = 0; k < K; ++k) tri[k] = (int)k < nslots ? slots[k] : new_triangle();
    for (size_t k = 0; k < K; ++k)
    {
        const int32_t t = tri[k];
        write_tri(t, ring[k].from, ring[k].to, id);
        link(3 * t, ring[k].twin);
        C_[3 * t] = ring[k].cons;
        C_[3 * t + 1] = C_[3 * t + 2] = 0;
        N_[3 * t + 1] = N_[3 * t + 2] = -1;
    }
    for (size_t k = 0; k + 1 < K; ++k) link(3 * tri[k] + 1, 3 * tri[k + 1] + 2);
    if (closed) link(3 * tri[K - 1] + 1, 3 * tri[0] + 2);
    if (split_a != ~0u)
        for (size_t k = 0; k < K; ++k)
        {
            if (ring[k].to == split_a || ring[k].to == split_b) mark(3 * tri[k] + 1);
            if (ring[k].from == split_a || ring[k].from == split_b) mark(3 * tri[k] + 2);
        }
    last_ = tri[0];

    // Lawson flips around p: half-edge 0 of each triangle at p is the edge
    // facing it, and every flip keeps p at corner 2 of both new triangles
    std::vector<int32_t> stack;
    for (size_t k = 0; k < K; ++k) stack.push_back(3 * tri[k]);
    while (!stack.empty())
    {
        const int32_t h = stack.back();
        stack.pop_back();
        if (V_[prv(h)] != id) continue;
        const int32_t g = N_[h];
        if (g < 0 || C_[h]) continue;
        if (incircle(org(h), dst(h), id, V_[prv(g)]) <= 0) continue;
        const int32_t t1 = h / 3, t2 = g / 3;
        if (flip(h))
        {
            stack.push_back(3 * t1);
            stack.push_back(3 * t2);
        }
    }
    return QDT_OK;
}

template <class F>
void MeshEdit::for_each_out(uint32_t v, F&& f) const
{
    const int32_t h0 = VH_[v];
    if (h0 < 0) return;
    const size_t limit = N_.size();
    size_t steps = 0;
    // counter-clockwise: the twin of the half-edge that ends at v
    int32_t h = h0;
    for (;;)
    {
        if (f(h)) return;
        if (++steps > limit) return;
        const int32_t g = N_[prv(h)];
        if (g < 0) break; // the hull
        if (g == h0) return;
        h = g;
    }
    // clockwise from h0 (hull vertices only)
    h = h0;
    for (;;)
    {
        const int32_t g = N_[h];
        if (g < 0) return;
        h = nxt(g);
        if (h == h0) return;
        if (f(h)) return;
        if (++steps > limit) return;
    }
}

int32_t MeshEdit::find_halfedge(uint32_t u, uint32_t v) const
{
    int32_t found = -1;
    for_each_out(u, [&](int32_t h) {
        if (dst(h) != v) return false;
        found = h;
        return true;
    });
    return found;
}

void MeshEdit::mark(int32_t h)
{
    C_[h] = 1;
    if (N_[h] >= 0) C_[N_[h]] = 1;
}

// Flips the edge of half-edge h (a -> b, apex c) and its twin (b -> a, apex
// d) to the diagonal c-d when the quadrilateral a, d, b, c is strictly
// convex: the triangles become (a, d, c) and (d, b, c) in the same slots.
bool MeshEdit::flip(int32_t h)
{
    const int32_t g = N_[h];
    if (g < 0 || C_[h]) return false;
    const uint32_t a = org(h), b = dst(h), c = V_[prv(h)], d = V_[prv(g)];
    if (orient(a, d, c) <= 0 || orient(d, b, c) <= 0) return false;
    const int32_t t1 = h / 3, t2 = g / 3;
    const int32_t bc = N_[nxt(h)], ca = N_[prv(h)], ad = N_[nxt(g)], db = N_[prv(g)];
    const uint8_t cbc = C_[nxt(h)], cca = C_[prv(h)], cad = C_[nxt(g)], cdb = C_[prv(g)];
    write_tri(t1, a, d, c);
    write_tri(t2, d, b, c);
    link(3 * t1, ad);
    link(3 * t1 + 2, ca);
    link(3 * t2, db);
    link(3 * t2 + 1, bc);
    link(3 * t1 + 1, 3 * t2 + 2);
    C_[3 * t1] = cad;
    C_[3 * t1 + 2] = cca;
    C_[3 * t2] = cdb;
    C_[3 * t2 + 1] = cbc;
    C_[3 * t1 + 1] = C_[3 * t2 + 2] = 0;
    ++flips_;
    return true;
}

bool MeshEdit::segments_cross(uint32_t a, uint32_t b, uint32_t c, uint32_t d) const
{
    return orient(a, b, c) * orient(a, b, d) < 0 && orient(c, d, a) * orient(c, d, b) < 0;
}

// The edges the open segment u-v crosses, in order from u, each as the
// half-edge running from its end right of u->v to its end left of it; or
// why the segment cannot become an edge: it crosses a constrained edge, or
// it passes exactly through a vertex.
qdt_status MeshEdit::trace(uint32_t u, uint32_t v, std::vector<int32_t>& crossed) const
{
    const qdt_vec2d& U = pts_[u];
    const qdt_vec2d& Vp = pts_[v];
    int32_t start = -1;
    bool blocked = false;
    for_each_out(u, [&](int32_t h) {
        const int32_t hn = nxt(h);
        const uint32_t y = org(hn), z = dst(hn);
        const int oy = orient(u, y, Vp), oz = orient(u, z, Vp);
        if (oy > 0 && oz < 0)
        {
            start = hn;
            return true;
        }
        if ((oy == 0 && same_direction(U, pts_[y], Vp)) || (oz == 0 && same_direction(U, pts_[z], Vp)))
        {
            blocked = true;
            return true;
        }
        return false;
    });
    if (blocked) return QDT_ERR_THROUGH_POINT;
    if (start < 0) return QDT_ERR_INTERNAL;
    int32_t c = start;
    for (size_t guard = 0; guard <= triangles(); ++guard)
    {
        if (C_[c]) return QDT_ERR_CROSSES_CONSTRAINT;
        crossed.push_back(c);
        const int32_t tw = N_[c];
        if (tw < 0) return QDT_ERR_INTERNAL;
        const uint32_t w = V_[prv(tw)];
        if (w == v) return QDT_OK;
        const int o = orient(u, v, pts_[w]);
        if (o == 0) return QDT_ERR_THROUGH_POINT;
        c = o > 0 ? nxt(tw) : prv(tw);
    }
    return QDT_ERR_INTERNAL;
}

// Lawson's flips over the edges on the stack (vertex pairs: a flip moves
// half-edges between slots, so edges are named by their ends): every edge
// that is not constrained and not locally Delaunay is flipped, and the four
// edges around it are examined in turn.
void MeshEdit::legalize(std::vector<int32_t>& stack)
{
    while (stack.size() >= 2)
    {
        const uint32_t b0 = (uint32_t)stack.back();
        stack.pop_back();
        const uint32_t a0 = (uint32_t)stack.back();
        stack.pop_back();
        int32_t h = find_halfedge(a0, b0);
        if (h < 0) h = find_halfedge(b0, a0);
        if (h < 0) continue;
        const int32_t g = N_[h];
        if (g < 0 || C_[h]) continue;
        const uint32_t a = org(h), b = dst(h), c = V_[prv(h)], d = V_[prv(g)];
        if (incircle(a, b, c, d) <= 0) continue;
        if (!flip(h)) continue;
        const uint32_t around[8] = { a, d, d, b, b, c, c, a };
        stack.insert(stack.end(), around, around + 8);
    }
}

qdt_status MeshEdit::constrain(size_t v0, size_t v1)
{
    if (v0 >= pts_.size() || v1 >= pts_.size()) return QDT_ERR_INVALID_INDEX;
    const uint32_t u = rep_[v0], v = rep_[v1];
    if (u == v) return QDT_ERR_SAME_POINT;
    if (!triangles() || VH_[u] < 0 || VH_[v] < 0) return QDT_ERR_NO_TRIANGLES;
    int32_t e = find_halfedge(u, v);
    if (e < 0) e = find_halfedge(v, u);
    if (e >= 0)
    {
        mark(e);
        return QDT_OK;
    }
    std::vector<int32_t> crossed;
    const qdt_status traced = trace(u, v, crossed);
    if (traced != QDT_OK) return traced;

    // Sloan: flip every crossed edge whose quadrilateral is strictly convex;
    // a new diagonal that still crosses u-v goes back into the queue
    std::deque<std::pair<uint32_t, uint32_t>> queue;
    for (int32_t h : crossed) queue.emplace_back(org(h), dst(h));
    std::vector<std::pair<uint32_t, uint32_t>> fresh;
    size_t stalled = 0;
    while (!queue.empty())
    {
        const auto ab = queue.front();
        queue.pop_front();
        int32_t h = find_halfedge(ab.first, ab.second);
        if (h < 0) h = find_halfedge(ab.second, ab.first);
        if (h < 0) return QDT_ERR_INTERNAL;
        const uint32_t c = V_[prv(h)], d = N_[h] >= 0 ? V_[prv(N_[h])] : c;
        if (flip(h))
        {
            stalled = 0;
            if (segments_cross(u, v, c, d))
                queue.emplace_back(c, d);
            else
                fresh.emplace_back(c, d);
        }
        else
        {
            queue.push_back(ab);
            // a whole round without a flip cannot happen when no vertex lies
            // on the segment
            if (++stalled > queue.size()) break;
        }
    }
    bool ok = false;
    if (queue.empty())
    {
        e = find_halfedge(u, v);
        if (e < 0) e = find_halfedge(v, u);
        if (e >= 0)
        {
            mark(e);
            ok = true;
        }
    }
    // restore the (constrained) Delaunay property around what changed
    std::vector<int32_t> stack;
    auto around = [&](uint32_t a, uint32_t b) {
        int32_t h = find_halfedge(a, b);
        if (h < 0) h = find_halfedge(b, a);
        if (h < 0) return;
        const uint32_t x = org(h), y = dst(h), c = V_[prv(h)];
        const uint32_t q[6] = { x, y, y, c, c, x };
        stack.insert(stack.end(), q, q + 6);
        if (N_[h] >= 0)
        {
            const uint32_t d = V_[prv(N_[h])];
            const uint32_t r[4] = { x, d, d, y };
            stack.insert(stack.end(), r, r + 4);
        }
    };
    for (const auto& ab : fresh) around(ab.first, ab.second);
    for (const auto& ab : queue) around(ab.first, ab.second);
    legalize(stack);
    return ok ? QDT_OK : QDT_ERR_INTERNAL;
}

size_t MeshEdit::constrained_edges() const
{
    size_t c = 0;
    for (size_t h = 0; h < C_.size(); ++h)
        if (C_[h] && (N_[h] < 0 || (int32_t)h < N_[h])) ++c;
    return c;
}

} // namespace qdt
