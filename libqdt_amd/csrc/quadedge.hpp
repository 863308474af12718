// quadedge.hpp -- structure-of-arrays quad-edge store and the Guibas-Stolfi
// operations on it, for one thread of a merge kernel.
//
// Layout.  A quad-edge record k owns the four directed edges 4k+r (r = 0..3,
// r odd = dual edges), as in libqdt/builder.hxx:78-86.  Instead of an array
// of records the store keeps two flat arrays:
//   NX[e]      onext of directed edge e   (int32, 4 per record, 16 B)
//   OR[e >> 1] origin of primal edge e    (int32, 2 per record, 8 B)
// so every operation is one 4-byte access, records can be initialised with
// one 16-byte store, and the same code runs on HBM or on LDS pointers.
// VE[v] keeps one edge out of every vertex (-1 while it has none), updated by
// every connect and delete; the seam kernels start their walks from it.
//
// Storage of a kd node.  A node holding points [o, o+s) owns records
// [3o, 3o+3s).  A planar graph on s >= 3 vertices has at most 3s-6 edges, so
// a node never needs more records than its own range; the records a node is
// not using sit on its free list (linked through NX[4k], marked OR[2k] = -1),
// which the parent inherits.  Seam segments that run concurrently inside one
// merge take extra records from a spare region above 3m with an atomic
// counter when their private lists run dry.
//
// Claims.  Concurrent segments must never touch the ring of one vertex at
// the same time.  A segment claims every vertex whose ring it is about to
// read or change.  All segments that can meet at a vertex run in one
// workgroup, so the claims live in an LDS hash table of (vertex, segment)
// words, cleared between phases.  A segment that finds a vertex held by
// another one (or the table full) stops before touching it; the operation
// sequence is a pure function of the store, so the segment can be resumed
// later from the cross edge where it stopped.
#pragma once

#include "predicates.hpp"

namespace qdt
{

__device__ __forceinline__ int e_rot(int e) { return (e & ~3) | ((e + 1) & 3); }
__device__ __forceinline__ int e_invrot(int e) { return (e & ~3) | ((e + 3) & 3); }
__device__ __forceinline__ int e_sym(int e) { return e ^ 2; }

constexpr unsigned long long kNoClaim = ~0ull; // empty claim-table entry
constexpr int kSpareChunk = 8;                 // spare records taken at once

struct Pool
{
    int head, tail; // free records of a node, -1 = empty
};

struct PredCount
{
    unsigned orient, orient_exact, incircle, incircle_exact;
    unsigned connects; // edges created
};

// Error codes (`err`): 1 index out of range, 2 no free record, 3 loop
// budget exhausted, 4 a seam segment ended before its stop edge, 5 a
// predicate's filter could not decide in an instance without the exact stage
// (its caller redoes the work with one that has it).  Every access is bounds-checked (a compare next to a
// dependent load costs nothing measurable) and every loop has a budget: a
// logic error unwinds instead of touching memory outside the store or
// spinning forever.
// kChecked = false drops the index checks: for stores in LDS, where an index
// out of range cannot fault (the hardware discards such writes and returns 0
// for such reads) and the checks cost a quarter of the merge instructions.
// The loop budgets stay, so a damaged store still cannot spin forever.  Such
// stores (the LDS cells) also take no claims and keep no vertex-edge entries.
// kExact = false leaves out the exact predicate stage (and with it the calls
// that raise every kernel's register count): an undecided filter sets err 5.
// kStatic = true (without the exact stage): the static bounds are always set
// (the caller makes sure), so the filters compare against them alone.
template <class IntPtr, class PtPtr, bool kChecked = true, bool kExact = true, bool kStatic = false>
struct QuadEdges
{
    static constexpr bool kHasExact = kExact;
    static constexpr bool kStaticBounds = kStatic && !kExact;
    IntPtr NX;
    IntPtr OR;
    PtPtr PT;
    IntPtr VE;          // vertex -> an edge out of it (may be null)
    int lim_e;          // directed edge ids are in [0, lim_e)
    int lim_v;          // vertex ids are in [0, lim_v)
    unsigned* spare;    // spare record counter (null: no spare region)
    int spare_base;     // first spare record
    int spare_cap;      // spare records available
    unsigned long long* claim; // LDS claim table (null: single owner, no claims)
    int claim_mask;            // table entries - 1 (a power of two)
    int claim_me;
    int h0, h1, h2, h3; // recently claimed vertices (named registers: no
                        // indexed array, which would live in scratch memory)
    int err;
    PredCount pc;
    double sb_orient, sb_incircle; // static filter bounds (0: run-time bounds), used without the exact stage

    __device__ __forceinline__ void init(IntPtr nx, IntPtr orr, PtPtr pt, IntPtr ve, int le, int lv)
    {
        NX = nx;
        OR = orr;
        PT = pt;
        VE = ve;
        lim_e = le;
        lim_v = lv;
        spare = nullptr;
        spare_base = spare_cap = 0;
        claim = nullptr;
        claim_mask = 0;
        claim_me = 0;
        h0 = h1 = h2 = h3 = -1;
        err = 0;
        pc = PredCount{ 0, 0, 0, 0, 0 };
        sb_orient = sb_incircle = 0.0;
    }

    __device__ __forceinline__ int chk(int e)
    {
        if (!kChecked) return e;
        if ((unsigned)e >= (unsigned)lim_e)
        {
            err = err ? err : 1;
            return 0;
        }
        return e;
    }
    __device__ __forceinline__ int onext(int e) { return NX[chk(e)]; }
    __device__ __forceinline__ void set_next(int e, int v) { NX[chk(e)] = v; }
    __device__ __forceinline__ int org(int e) { return OR[chk(e) >> 1]; }
    __device__ __forceinline__ int dest(int e) { return OR[chk(e ^ 2) >> 1]; }
    __device__ __forceinline__ void set_org(int e, int v) { OR[chk(e) >> 1] = v; }
    __device__ __forceinline__ int oprev(int e) { return e_rot(onext(e_rot(e))); }
    __device__ __forceinline__ int lnext(int e) { return e_rot(onext(e_invrot(e))); }
    __device__ __forceinline__ int rprev(int e) { return onext(e ^ 2); }
    __device__ __forceinline__ double2 pt(int v)
    {
        if (!kChecked) return PT[v];
        if ((unsigned)v >= (unsigned)lim_v)
        {
            err = err ? err : 1;
            v = 0;
        }
        return PT[v];
    }
    // (stores in LDS, the cells, keep no vertex-edge entries and no claims:
    // their entries are derived from the final records, and one thread owns
    // each merge)
    __device__ __forceinline__ void set_ve(int v, int e)
    {
        if constexpr (!kChecked) return;
        if (VE && (unsigned)v < (unsigned)lim_v) VE[v] = e;
    }
    // true when this segment may touch v's ring (claims it on first use)
    __device__ __forceinline__ bool take(int v)
    {
        if constexpr (!kChecked) return true;
        if (!claim) return true;
        if (v == h0 || v == h1 || v == h2 || v == h3) return true;
        if ((unsigned)v >= (unsigned)lim_v)
        {
            err = err ? err : 1;
            return false;
        }
        const unsigned long long mine = ((unsigned long long)(unsigned)v << 32) | (unsigned)claim_me;
        unsigned x = ((unsigned)v * 2654435761u) & (unsigned)claim_mask;
        for (int probe = 0; probe <= claim_mask; ++probe)
        {
            unsigned long long cur = claim[x];
            if (cur == kNoClaim)
            {
                cur = atomicCAS(&claim[x], kNoClaim, mine);
                if (cur == kNoClaim) cur = mine;
            }
            if ((unsigned)(cur >> 32) == (unsigned)v)
            {
                if (cur != mine) return false;
                h3 = h2;
                h2 = h1;
                h1 = h0;
                h0 = v;
                return true;
            }
            x = (x + 1) & (unsigned)claim_mask;
        }
        return false; // table full
    }

    __device__ __forceinline__ int pop(Pool& p)
    {
        const int k = p.head;
        if (k < 0)
        {
            if (spare)
            {
                // kSpareChunk records per request (one device-wide atomic
                // per chunk): the first is returned, the rest become the pool
                const unsigned s = atomicAdd(spare, (unsigned)kSpareChunk);
                if ((int)s + kSpareChunk <= spare_cap)
                {
                    const int r0 = spare_base + (int)s;
                    for (int i = 1; i < kSpareChunk; ++i)
                    {
                        NX[4 * (r0 + i)] = i + 1 < kSpareChunk ? r0 + i + 1 : -1;
                        OR[2 * (r0 + i)] = -1;
                    }
                    p.head = r0 + 1;
                    p.tail = r0 + kSpareChunk - 1;
                    return r0;
                }
            }
            err = err ? err : 2;
            return 0;
        }
        if (4 * k >= lim_e)
        {
            err = err ? err : 1;
            return 0;
        }
        const int nx = NX[4 * k];
        p.head = nx;
        if (nx < 0) p.tail = -1;
        return k;
    }

    __device__ __forceinline__ void push(Pool& p, int k)
    {
        OR[2 * k] = -1;
        NX[4 * k] = p.head;
        if (p.head < 0) p.tail = k;
        p.head = k;
    }

    __device__ __forceinline__ Pool cat(Pool a, Pool b)
    {
        if (a.head < 0) return b;
        if (b.head < 0) return a;
        set_next(4 * a.tail, b.head);
        Pool r = { a.head, b.tail };
        return r;
    }

    // builder.hxx:131-142
    __device__ __forceinline__ int make_edge(Pool& p)
    {
        const int k = pop(p);
        const int e = 4 * k;
        NX[e] = e;
        NX[e + 1] = e + 3;
        NX[e + 2] = e + 2;
        NX[e + 3] = e + 1;
        OR[2 * k] = -1;
        OR[2 * k + 1] = -1;
        return e;
    }

    // builder.hxx:144-149
    __device__ __forceinline__ void splice(int a, int b)
    {
        const int na = onext(a), nb = onext(b);
        const int alpha = e_rot(na), beta = e_rot(nb);
        set_next(a, nb);
        set_next(b, na);
        const int xa = onext(alpha), xb = onext(beta);
        set_next(alpha, xb);
        set_next(beta, xa);
    }

    // builder.hxx:151-159
    __device__ __forceinline__ int connect(int a, int b, Pool& p)
    {
        const int e = make_edge(p);
        ++pc.connects;
        const int o = dest(a), d = org(b);
        set_org(e, o);
        set_org(e ^ 2, d);
        splice(e, lnext(a));
        splice(e ^ 2, b);
        set_ve(o, e);
        set_ve(d, e ^ 2);
        return e;
    }

    // builder.hxx:161-166
    __device__ __forceinline__ void delete_edge(int e, Pool& p)
    {
        const int o = org(e), d = dest(e);
        const int po = oprev(e);
        splice(e, po);
        const int pd = oprev(e ^ 2);
        splice(e ^ 2, pd);
        set_ve(o, po != e ? po : -1);
        set_ve(d, pd != (e ^ 2) ? pd : -1);
        push(p, chk(e) >> 2);
    }

    __device__ __forceinline__ double pred_orient(double2 A, double2 B, double2 C)
    {
        ++pc.orient;
        if constexpr (kExact) return orient2d(A.x, A.y, B.x, B.y, C.x, C.y, pc.orient_exact);
        if constexpr (kStaticBounds)
        {
            const double det = orient2d_det(A.x, A.y, B.x, B.y, C.x, C.y);
            if (!(fabs(det) > sb_orient)) err = err ? err : 5;
            return det;
        }
        double det;
        if (sb_orient > 0)
        {
            det = orient2d_det(A.x, A.y, B.x, B.y, C.x, C.y);
            if (!(det > sb_orient || -det > sb_orient)) err = err ? err : 5;
            return det;
        }
        if (!orient2d_filter(A.x, A.y, B.x, B.y, C.x, C.y, det)) err = err ? err : 5;
        return det;
    }
    __device__ __forceinline__ double pred_incircle(double2 A, double2 B, double2 C, double2 D)
    {
        ++pc.incircle;
        if constexpr (kExact) return incircle(A.x, A.y, B.x, B.y, C.x, C.y, D.x, D.y, pc.incircle_exact);
        if constexpr (kStaticBounds)
        {
            const double det = incircle_det(A.x, A.y, B.x, B.y, C.x, C.y, D.x, D.y);
            if (!(fabs(det) > sb_incircle)) err = err ? err : 5;
            return det;
        }
        double det;
        if (sb_incircle > 0)
        {
            det = incircle_det(A.x, A.y, B.x, B.y, C.x, C.y, D.x, D.y);
            if (!(det > sb_incircle || -det > sb_incircle)) err = err ? err : 5;
            return det;
        }
        if (!incircle_filter(A.x, A.y, B.x, B.y, C.x, C.y, D.x, D.y, det)) err = err ? err : 5;
        return det;
    }
    __device__ __forceinline__ double orient(int a, int b, int c) { return pred_orient(pt(a), pt(b), pt(c)); }
    __device__ __forceinline__ bool ccw(int a, int b, int c) { return orient(a, b, c) > 0; }
    __device__ __forceinline__ bool right_of(int x, int e) { return ccw(x, dest(e), org(e)); }
    __device__ __forceinline__ bool left_of(int x, int e) { return ccw(x, org(e), dest(e)); }

    // builder.hxx:171-175: distinct vertices, strictly inside
    __device__ __forceinline__ bool in_circle(int a, int b, int c, int d)
    {
        if (a == d || b == d || c == d || a == b || a == c || b == c) return false;
        return pred_incircle(pt(a), pt(b), pt(c), pt(d)) > 0;
    }
    // the same tests on coordinates already loaded
    __device__ __forceinline__ bool ccw_p(double2 A, double2 B, double2 C) { return pred_orient(A, B, C) > 0; }
    __device__ __forceinline__ bool in_circle_p(int a, int b, int c, int d, double2 A, double2 B, double2 C,
                                                double2 D)
    {
        if (a == d || b == d || c == d || a == b || a == c || b == c) return false;
        return pred_incircle(A, B, C, D) > 0;
    }
    __device__ __forceinline__ double incircle_v(int a, int b, int c, int d)
    {
        return pred_incircle(pt(a), pt(b), pt(c), pt(d));
    }

    // frame order of builder.hxx:71-76
    __device__ __forceinline__ bool less_f(int u, int v, int frame)
    {
        const double2 a = pt(u), b = pt(v);
        if (frame == 0) return a.x < b.x || (a.x == b.x && a.y < b.y);
        return a.y < b.y || (a.y == b.y && a.x > b.x);
    }

    // Leaf of two or three vertices s[0..cnt) given in the frame's order,
    // builder.hxx:177-201.  Records come from [rec0, rec0 + 3*cnt); the
    // unused ones go to the returned pool.
    __device__ __forceinline__ void leaf(const int* s, int cnt, int rec0, int& first, int& second, Pool& pool)
    {
        const int nrec = 3 * cnt;
        // chain all records of the leaf into its pool, then allocate from it
        for (int i = 0; i < nrec; ++i)
        {
            const int k = rec0 + i;
            NX[4 * k] = (i + 1 < nrec) ? k + 1 : -1;
            OR[2 * k] = -1;
        }
        pool.head = rec0;
        pool.tail = rec0 + nrec - 1;
        if (cnt == 1)
        {
            first = second = -1; // a point set of one distinct point
            return;
        }
        if (cnt == 2)
        {
            const int a = make_edge(pool);
            set_org(a, s[0]);
            set_org(a ^ 2, s[1]);
            set_ve(s[0], a);
            set_ve(s[1], a ^ 2);
            first = a;
            second = a ^ 2;
            return;
        }
        const int a = make_edge(pool), b = make_edge(pool);
        splice(a ^ 2, b);
        set_org(a, s[0]);
        set_org(a ^ 2, s[1]);
        set_org(b, s[1]);
        set_org(b ^ 2, s[2]);
        set_ve(s[0], a);
        set_ve(s[1], b);
        set_ve(s[2], b ^ 2);
        const double o = orient(s[0], s[1], s[2]);
        if (o > 0)
        {
            connect(b, a, pool);
            first = a;
            second = b ^ 2;
        }
        else if (o < 0)
        {
            const int c = connect(b, a, pool);
            first = c ^ 2;
            second = c;
        }
        else
        {
            first = a;
            second = b ^ 2;
        }
    }

    // Lower common tangent and first cross edge of a merge (builder.hxx:253-261):
    // the walk (reads only) and the connect.
    __device__ __forceinline__ void tangent_walk(int& ldi, int& rdi, int& budget)
    {
        for (;;)
        {
            if (err || --budget < 0) break;
            if (left_of(org(rdi), ldi))
                ldi = lnext(ldi);
            else if (right_of(org(ldi), rdi))
                rdi = rprev(rdi);
            else
                break;
        }
    }
    __device__ __forceinline__ int tangent_connect(int& ldo, int ldi, int rdi, int& rdo, Pool& pool)
    {
        const int basel = connect(rdi ^ 2, ldi, pool);
        if (org(ldi) == org(ldo)) ldo = basel ^ 2;
        if (org(rdi) == org(rdo)) rdo = basel;
        return basel;
    }
    __device__ __forceinline__ int tangent(int& ldo, int ldi, int rdi, int& rdo, Pool& pool, int& budget)
    {
        tangent_walk(ldi, rdi, budget);
        return tangent_connect(ldo, ldi, rdi, rdo, pool);
    }

    // Deletion of e when oprev(e) = po and onext(e) = na are known (eo, ed:
    // its ends).  Same result as delete_edge: the two splices of
    // builder.hxx:161-166, with every word they read back fixed by
    // onext(oprev(x)) = x and rot(x).onext = invrot(oprev(x)); only the ring
    // at the far end (ed) is read.
    __device__ __forceinline__ void delete_known(int e, int po, int na, int eo, int ed, Pool& p)
    {
        const int s = e ^ 2;
        const int na2 = onext(s);
        const int r2 = onext(e_rot(s));
        const int pd = e_rot(r2);
        set_next(e, e);
        set_next(po, na);
        set_next(e_rot(na), e_invrot(po));
        set_next(e_rot(e), e_invrot(e));
        set_next(s, s);
        set_next(pd, na2);
        set_next(e_rot(na2), r2);
        set_next(e_rot(s), e_invrot(s));
        set_ve(eo, po != e ? po : -1);
        set_ve(ed, pd != s ? pd : -1);
        push(p, chk(e) >> 2);
    }

    // New edge o -> d placed in o's ring right after x (whose onext was nx)
    // and in d's ring right after y (whose onext was ny): the result of
    // connect (builder.hxx:151-159) with every word it reads known.
    __device__ __forceinline__ int connect_known(int x, int nx, int y, int ny, int o, int d, Pool& p)
    {
        const int e = 4 * pop(p);
        ++pc.connects;
        set_org(e, o);
        set_org(e ^ 2, d);
        set_next(e, nx);
        set_next(x, e);
        set_next(e + 1, e_invrot(x));
        set_next(e_rot(nx), e + 3);
        set_next(e + 2, ny);
        set_next(y, e + 2);
        set_next(e + 3, e_invrot(y));
        set_next(e_rot(ny), e + 1);
        set_ve(o, e);
        set_ve(d, e ^ 2);
        return e;
    }

    // The rising bubble of builder.hxx:263-283 from cross edge `basel`.  With
    // stop < 0 it runs to the upper common tangent.  With a stop edge (a
    // cross edge already in the store, directed right -> left like basel) it
    // ends when the next cross edge to connect is the stop edge.  Where the
    // stop edge is the first candidate around a base vertex, that side has no
    // candidate inside this piece (between two cross edges sharing a vertex
    // only further cross edges can lie: an edge of one side there would cross
    // an edge of the other side).  Returns 1 at the stop edge, 0 at the top,
    // 2 when a vertex it needs is claimed by another segment: `basel` is then
    // the cross edge to resume from.
    //
    // The left scan deletes only edges at bd between left vertices and the
    // right scan only edges at bo between right vertices, so neither changes
    // what the other side reads.  Each step's candidates are carried over
    // instead of read again, so that a step waits on as few dependent loads
    // (each a trip to L2 or HBM, or an LDS access in a cell) as possible:
    //  * The side not chosen keeps its candidate, the candidate after it and
    //    their ends and coordinates: the new edge changes neither ring.
    //  * The chosen side's new candidate comes from the connect: on the
    //    right oprev(new base) = lnext(rc), on the left onext(new base.sym)
    //    = the old onext(lc.sym).
    //  * Deleting lc: oprev(lc) = basel.sym, onext(lc) = ln; deleting rc:
    //    oprev(rc) = rp, onext(rc) = basel; the next candidate's ring word is
    //    read before the deletion's stores.
    // Where a carried word is one the step itself rewrote (an end vertex of
    // degree one or two), the value written is used instead.
    // (kStop = false: no stop edge, the run ends at the upper tangent; the
    // stop tests are compiled out)
    template <bool kStop = true>
    __device__ int bubble(int& basel, int stop, Pool& pool, int& budget)
    {
        const int srec = kStop && stop >= 0 ? (stop >> 2) : -1;
        const int so = kStop && stop >= 0 ? org(stop) : -1;
        const int sd = kStop && stop >= 0 ? dest(stop) : -1;
        int bo = org(basel), bd = dest(basel);
        double2 Po = pt(bo), Pd = pt(bd);
        bool fresh = true;
        int lc = 0, ln = 0, lcd = 0, lnd = 0, rc = 0, rp = 0, rcd = 0, rpd = 0;
        double2 Pl = Po, Pln = Po, Pr = Po, Prp = Po;
        for (;;)
        {
            if (err) return 0;
            if (--budget < 0)
            {
                err = err ? err : 3;
                return 0;
            }
            if (!take(bo) || !take(bd)) return 2;
            if (fresh)
            {
                lc = onext(basel ^ 2);
                rc = oprev(basel);
                ln = onext(lc);
                rp = oprev(rc);
                lcd = dest(lc);
                rcd = dest(rc);
                lnd = dest(ln);
                rpd = dest(rp);
                Pl = pt(lcd);
                Pr = pt(rcd);
                Pln = pt(lnd);
                Prp = pt(rpd);
                fresh = false;
            }
            const int b2 = basel ^ 2;
            bool lv = false;
            if (!kStop || (lc >> 2) != srec)
            {
                lv = ccw_p(Pl, Pd, Po);
                if (lv)
                {
                    bool deleted = false;
                    for (;;)
                    {
                        if (err || --budget < 0) break;
                        if (kStop && (ln >> 2) == srec) break;
                        if (!in_circle_p(bd, bo, lcd, lnd, Pd, Po, Pl, Pln)) break;
                        if (!take(lcd)) return 2;
                        const int nl = onext(ln);
                        delete_known(lc, b2, ln, bd, lcd, pool);
                        lc = ln;
                        lcd = lnd;
                        Pl = Pln;
                        ln = lc == b2 ? b2 : nl; // bd's ring was {basel.sym, lc}
                        lnd = dest(ln);
                        Pln = pt(lnd);
                        deleted = true;
                    }
                    if (deleted) lv = ccw_p(Pl, Pd, Po);
                }
            }
            bool rv = false;
            if (!kStop || (rc >> 2) != srec)
            {
                rv = ccw_p(Pr, Pd, Po);
                if (rv)
                {
                    bool deleted = false;
                    for (;;)
                    {
                        if (err || --budget < 0) break;
                        if (kStop && (rp >> 2) == srec) break;
                        if (!in_circle_p(bd, bo, rcd, rpd, Pd, Po, Pr, Prp)) break;
                        if (!take(rcd)) return 2;
                        const int pr = onext(e_rot(rp));
                        delete_known(rc, rp, basel, bo, rcd, pool);
                        rc = rp;
                        rcd = rpd;
                        Pr = Prp;
                        rp = rc == basel ? basel : e_rot(pr); // bo's ring was {basel, rc}
                        rpd = dest(rp);
                        Prp = pt(rpd);
                        deleted = true;
                    }
                    if (deleted) rv = ccw_p(Pr, Pd, Po);
                }
            }
            if (!lv && !rv) return 0;
            const bool right = !lv || (rv && in_circle_p(lcd, bd, bo, rcd, Pl, Pd, Po, Pr));
            // the new cross edge would run rcd -> bd or bo -> lcd
            if (kStop && stop >= 0)
            {
                const int no = right ? rcd : bo;
                const int nd = right ? bd : lcd;
                if (no == so && nd == sd) return 1;
            }
            if (!take(right ? rcd : lcd)) return 2;
            if (right)
            {
                // e = rcd -> bd after lnext(rc) at rcd and after basel.sym at bd
                const int x = lnext(rc);
                const int nx = onext(x), xr = onext(e_rot(x)), xd = dest(x);
                const int e = connect_known(x, nx, b2, lc, rcd, bd, pool);
                if (lc == b2) // bd's ring was {basel.sym}: now {basel.sym, e.sym}
                {
                    ln = e ^ 2;
                    lnd = rcd;
                    Pln = Pr;
                }
                basel = e;
                bo = rcd;
                Po = Pr;
                rc = x;
                rcd = xd;
                rp = nx == x ? e : e_rot(xr); // rcd's ring was {x}: now {x, e}
                rpd = dest(rp);
                Pr = pt(rcd);
                Prp = pt(rpd);
            }
            else
            {
                // e = bo -> lcd after rc (= lnext(basel.sym)) at bo and after lc.sym at lcd
                const int ls = lc ^ 2;
                const int ny = onext(ls);
                const int nl = onext(ny), nd = dest(ny);
                const int e = connect_known(rc, basel, ls, ny, bo, lcd, pool);
                if (rc == basel) // bo's ring was {basel}: now {basel, e}
                {
                    rp = e;
                    rpd = lcd;
                    Prp = Pl;
                }
                basel = e;
                bd = lcd;
                Pd = Pl;
                lc = ny;
                lcd = nd;
                ln = ny == ls ? (e ^ 2) : nl; // lcd's ring was {lc.sym}: now {lc.sym, e.sym}
                lnd = dest(ln);
                Pl = pt(lcd);
                Pln = pt(lnd);
            }
        }
    }

    // Merge of two triangulations separated in the current frame,
    // builder.hxx:205-285.  Left/right candidate scans run one after the
    // other; they touch disjoint edges, so the order does not matter.
    __device__ void merge(int& ldo, int ldi, int rdi, int& rdo, Pool& pool, int budget)
    {
        int basel = tangent(ldo, ldi, rdi, rdo, pool, budget);
        bubble<false>(basel, -1, pool, budget);
    }

    // builder.hxx:287-301
    __device__ int hull_walk(int e, int frame, bool want_max, int budget)
    {
        for (;;)
        {
            if (err) return e;
            if (--budget < 0)
            {
                err = err ? err : 3;
                return e;
            }
            const int nx = rprev(e);
            const int pv = oprev(e) ^ 2;
            const int oe = org(e);
            const int on = org(nx);
            if (want_max ? less_f(oe, on, frame) : less_f(on, oe, frame))
            {
                e = nx;
                continue;
            }
            const int op = org(pv);
            if (want_max ? less_f(oe, op, frame) : less_f(op, oe, frame))
            {
                e = pv;
                continue;
            }
            return e;
        }
    }

    // The hull edge out of v with the outer face on its right (the class of
    // ldo and of hull_walk's result), for v a vertex of the convex hull that
    // is not inside a straight stretch of it: the ring's faces are strictly
    // counter-clockwise triangles except the outer face, whose sweep around
    // v is the one with orient(v, dest(a), dest(onext(a))) <= 0 (a single
    // edge sweeps 360 degrees); the edge after that sweep.  -1 if the ring
    // shows no such sweep within the budget.
    __device__ int hull_edge_out(int v, int budget)
    {
        if (!VE || (unsigned)v >= (unsigned)lim_v) return -1;
        const int e0 = VE[v];
        if (e0 < 0) return -1;
        const double2 V = pt(v);
        int a = e0, da = dest(e0);
        double2 A = pt(da);
        for (int i = 0; i < budget && !err; ++i)
        {
            const int b = onext(a), db = dest(b);
            const double2 B = pt(db);
            if (!(pred_orient(V, A, B) > 0)) return b;
            a = b;
            A = B;
            if (a == e0) break;
        }
        return -1;
    }

    // re-walk of a merged node's extreme hull edges to the parent's frame
    // (builder.hxx:340-346) when the extreme points in that frame are known
    // (vmin, vmax; < 0: not known): the hull edges out of them directly, in
    // O(degree) instead of a walk along the hull (which points in convex
    // position make as long as the node).  hull_walk finds the same edges:
    // it follows the hull in the class of its start edge to the vertex that
    // is extreme in the frame (the order is unimodal along a convex hull),
    // and the frame's extreme point is a hull vertex whose interior angle is
    // below 180 degrees.
    __device__ __forceinline__ void rewalk_ext(int ldo, int rdo, int axis, int out_axis, int npts, int vmin,
                                               int vmax, int& first, int& second)
    {
        if (out_axis != axis && vmin >= 0 && vmax >= 0)
        {
            const int f = hull_edge_out(vmin, npts + 8);
            const int g = hull_edge_out(vmax, npts + 8);
            if (f >= 0 && g >= 0)
            {
                first = f;
                second = oprev(g);
                return;
            }
        }
        rewalk(ldo, rdo, axis, out_axis, npts, first, second);
    }

    // re-walk of a merged node's extreme hull edges to the parent's frame
    // (builder.hxx:340-346)
    __device__ __forceinline__ void rewalk(int ldo, int rdo, int axis, int out_axis, int npts, int& first,
                                           int& second)
    {
        if (out_axis != axis)
        {
            first = hull_walk(ldo, out_axis, false, npts + 8);
            const int re = hull_walk(onext(rdo), out_axis, true, npts + 8);
            second = oprev(re);
        }
        else
        {
            first = ldo;
            second = rdo;
        }
    }

    // npts = points of the node (bounds every loop); vmin, vmax: the node's
    // extreme points in the parent's frame when known (else -1)
    __device__ __forceinline__ void merge_node(int lfirst, int lsecond, int rfirst, int rsecond, int axis,
                                               int out_axis, int npts, int& first, int& second, Pool& pool,
                                               int vmin = -1, int vmax = -1)
    {
        int ldo = lfirst, rdo = rsecond;
        const int budget = 16 * npts + 64;
        merge(ldo, lsecond, rfirst, rdo, pool, budget);
        rewalk_ext(ldo, rdo, axis, out_axis, npts, vmin, vmax, first, second);
    }
};

} // namespace qdt
