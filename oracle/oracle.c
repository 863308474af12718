/*
 * oracle.c -- CPU restatement of libqdt's batch Delaunay path.
 *
 * TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
 * bench's CPU baseline leg, never by the product library.  Build with strict
 * IEEE arithmetic (-ffp-contract=off, no fast-math), as the reference demands
 * (CMakeLists.txt:52-59): the filter bounds assume unfused operations.
 */
#include "oracle.h"

#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Expansion arithmetic (non-overlapping expansions, smallest term first).   */
/* The reference's exact path (robust.hxx:29-279) keeps expansions in        */
/* increasing magnitude and takes the sign of the last term; the routines    */
/* here are the textbook ones (Shewchuk 1997) with zero elimination and      */
/* compression.  Any exact method gives the same sign.                       */
/* ------------------------------------------------------------------------ */

#define XCAP 1600 /* 2*16*16*3 = 1536 terms is the largest uncompressed size */

static void two_sum(double a, double b, double* x, double* y)
{
    const double s = a + b;
    const double bv = s - a;
    const double av = s - bv;
    *x = s;
    *y = (a - av) + (b - bv);
}

static void two_prod(double a, double b, double* x, double* y)
{
    const double p = a * b;
    *x = p;
    *y = fma(a, b, -p);
}

/* h = e + b */
static int x_grow(const double* e, int ne, double b, double* h)
{
    double q = b;
    int nh = 0;
    for (int i = 0; i < ne; ++i)
    {
        double s, t;
        two_sum(q, e[i], &s, &t);
        q = s;
        if (t != 0.0) h[nh++] = t;
    }
    if (q != 0.0 || nh == 0) h[nh++] = q;
    if (nh == 1 && h[0] == 0.0) nh = 0;
    return nh;
}

/* h = e + f (repeated growth; valid for any non-overlapping inputs) */
static int x_add(const double* e, int ne, const double* f, int nf, double* h)
{
    double a[XCAP], b[XCAP];
    memcpy(a, e, sizeof(double) * (size_t)ne);
    int na = ne;
    double* cur = a;
    double* nxt = b;
    for (int i = 0; i < nf; ++i)
    {
        na = x_grow(cur, na, f[i], nxt);
        double* t = cur;
        cur = nxt;
        nxt = t;
    }
    memcpy(h, cur, sizeof(double) * (size_t)na);
    return na;
}

static int x_neg(const double* e, int ne, double* h)
{
    for (int i = 0; i < ne; ++i) h[i] = -e[i];
    return ne;
}

/* h = e * b */
static int x_scale(const double* e, int ne, double b, double* h)
{
    if (ne == 0 || b == 0.0) return 0;
    double q, t;
    int nh = 0;
    two_prod(e[0], b, &q, &t);
    if (t != 0.0) h[nh++] = t;
    for (int i = 1; i < ne; ++i)
    {
        double p1, p0, s;
        two_prod(e[i], b, &p1, &p0);
        two_sum(q, p0, &s, &t);
        if (t != 0.0) h[nh++] = t;
        two_sum(p1, s, &q, &t);
        if (t != 0.0) h[nh++] = t;
    }
    if (q != 0.0) h[nh++] = q;
    return nh;
}

/* Shewchuk's compress: same value, fewer terms. */
static int x_compress(const double* e, int ne, double* h)
{
    if (ne == 0) return 0;
    double g[XCAP];
    int bottom = ne - 1;
    double q = e[bottom];
    for (int i = ne - 2; i >= 0; --i)
    {
        const double enow = e[i];
        const double qn = q + enow;
        const double small = enow - (qn - q);
        q = qn;
        if (small != 0.0)
        {
            g[bottom--] = q;
            q = small;
        }
    }
    g[bottom] = q;
    int top = 0;
    for (int i = bottom + 1; i < ne; ++i)
    {
        const double hnow = g[i];
        const double qn = hnow + q;
        const double small = q - (qn - hnow);
        q = qn;
        if (small != 0.0) h[top++] = small;
    }
    h[top++] = q;
    if (top == 1 && h[0] == 0.0) top = 0;
    return top;
}

/* h = e * f, compressed */
static int x_mul(const double* e, int ne, const double* f, int nf, double* h)
{
    if (ne == 0 || nf == 0) return 0;
    const double* big = ne >= nf ? e : f;
    const double* sml = ne >= nf ? f : e;
    const int nb = ne >= nf ? ne : nf;
    const int ns = ne >= nf ? nf : ne;
    double acc[XCAP], t[XCAP], s[XCAP];
    int na = 0;
    for (int i = 0; i < ns; ++i)
    {
        const int nt = x_scale(big, nb, sml[i], t);
        na = x_add(acc, na, t, nt, s);
        na = x_compress(s, na, acc);
    }
    memcpy(h, acc, sizeof(double) * (size_t)na);
    return na;
}

static int x_sign(const double* e, int ne)
{
    if (ne == 0) return 0;
    return e[ne - 1] > 0.0 ? 1 : (e[ne - 1] < 0.0 ? -1 : 0);
}

/* a - b as an expansion of up to two terms */
static int x_diff(double a, double b, double* h)
{
    double x, y;
    two_sum(a, -b, &x, &y);
    int n = 0;
    if (y != 0.0) h[n++] = y;
    if (x != 0.0) h[n++] = x;
    return n;
}

/* ------------------------------------------------------------------------ */
/* Predicates (robust.hxx:281-396)                                           */
/* ------------------------------------------------------------------------ */

static const double EPS = 1.1102230246251565e-16;           /* robust.hxx:281 */
#define ERR_ORIENT ((3.0 + 16.0 * EPS) * EPS)               /* robust.hxx:282 */
#define ERR_INCIRCLE ((10.0 + 96.0 * EPS) * EPS)            /* robust.hxx:283 */
static const double FILTER_FLOOR = 1e-250;                  /* robust.hxx:289 */

/* Power of two that brings the largest coordinate near 1, or 0 when the
 * largest exponent already lies in [-64, 64] (robust.hxx:297-306). */
static int norm_exponent(const double* v, int n)
{
    double m = 0.0;
    for (int i = 0; i < n; ++i)
    {
        const double a = fabs(v[i]);
        if (a > m) m = a;
    }
    if (m == 0.0 || !isfinite(m)) return 0;
    const int e = ilogb(m);
    return (e >= -64 && e <= 64) ? 0 : e;
}

/* robust.hxx:321-336 */
int orc_orient2d_exact(double ax, double ay, double bx, double by, double cx, double cy)
{
    double v[6] = { ax, ay, bx, by, cx, cy };
    const int e = norm_exponent(v, 6);
    if (e)
        for (int i = 0; i < 6; ++i) v[i] = ldexp(v[i], -e);
    double acx[2], acy[2], bcx[2], bcy[2];
    const int nacx = x_diff(v[0], v[4], acx), nacy = x_diff(v[1], v[5], acy);
    const int nbcx = x_diff(v[2], v[4], bcx), nbcy = x_diff(v[3], v[5], bcy);
    double m1[XCAP], m2[XCAP], n2[XCAP], s[XCAP];
    const int nm1 = x_mul(acx, nacx, bcy, nbcy, m1);
    const int nm2 = x_mul(acy, nacy, bcx, nbcx, m2);
    x_neg(m2, nm2, n2);
    const int ns = x_add(m1, nm1, n2, nm2, s);
    return x_sign(s, ns);
}

/* robust.hxx:337-347 */
double orc_orient2d(double ax, double ay, double bx, double by, double cx, double cy)
{
    const double detleft = (ax - cx) * (by - cy);
    const double detright = (ay - cy) * (bx - cx);
    const double det = detleft - detright;
    const double mag = fabs(detleft) + fabs(detright);
    const double bound = ERR_ORIENT * mag;
    if ((det > bound || -det > bound) && mag > FILTER_FLOOR) return det;
    return (double)orc_orient2d_exact(ax, ay, bx, by, cx, cy);
}

/* lift * cross where lift = dx^2 + dy^2 and cross = p*q - r*s, all compressed */
static int lifted_term(const double* dx, int ndx, const double* dy, int ndy,
                       const double* p, int np, const double* q, int nq,
                       const double* r, int nr, const double* s, int ns, double* out)
{
    double m1[XCAP], m2[XCAP], t[XCAP], lift[XCAP], cross[XCAP];
    int n1 = x_mul(dx, ndx, dx, ndx, m1);
    int n2 = x_mul(dy, ndy, dy, ndy, m2);
    int nt = x_add(m1, n1, m2, n2, t);
    const int nl = x_compress(t, nt, lift);
    n1 = x_mul(p, np, q, nq, m1);
    n2 = x_mul(r, nr, s, ns, m2);
    x_neg(m2, n2, m2);
    nt = x_add(m1, n1, m2, n2, t);
    const int nc = x_compress(t, nt, cross);
    return x_mul(lift, nl, cross, nc, out);
}

/* robust.hxx:348-376 */
int orc_incircle_exact(double ax, double ay, double bx, double by, double cx, double cy,
                       double dx, double dy)
{
    double v[8] = { ax, ay, bx, by, cx, cy, dx, dy };
    const int e = norm_exponent(v, 8);
    if (e)
        for (int i = 0; i < 8; ++i) v[i] = ldexp(v[i], -e);
    double adx[2], ady[2], bdx[2], bdy[2], cdx[2], cdy[2];
    const int nadx = x_diff(v[0], v[6], adx), nady = x_diff(v[1], v[7], ady);
    const int nbdx = x_diff(v[2], v[6], bdx), nbdy = x_diff(v[3], v[7], bdy);
    const int ncdx = x_diff(v[4], v[6], cdx), ncdy = x_diff(v[5], v[7], cdy);
    double t1[XCAP], t2[XCAP], t3[XCAP], s12[XCAP], tot[XCAP];
    /* alift * (bdx*cdy - cdx*bdy) */
    const int n1 = lifted_term(adx, nadx, ady, nady, bdx, nbdx, cdy, ncdy, cdx, ncdx, bdy, nbdy, t1);
    /* blift * (cdx*ady - adx*cdy) */
    const int n2 = lifted_term(bdx, nbdx, bdy, nbdy, cdx, ncdx, ady, nady, adx, nadx, cdy, ncdy, t2);
    /* clift * (adx*bdy - bdx*ady) */
    const int n3 = lifted_term(cdx, ncdx, cdy, ncdy, adx, nadx, bdy, nbdy, bdx, nbdx, ady, nady, t3);
    int n = x_add(t1, n1, t2, n2, s12);
    n = x_add(s12, n, t3, n3, tot);
    return x_sign(tot, n);
}

/* robust.hxx:377-396 */
double orc_incircle(double ax, double ay, double bx, double by, double cx, double cy,
                    double dx, double dy)
{
    const double adx = ax - dx, ady = ay - dy;
    const double bdx = bx - dx, bdy = by - dy;
    const double cdx = cx - dx, cdy = cy - dy;
    const double bdxcdy = bdx * cdy, cdxbdy = cdx * bdy;
    const double cdxady = cdx * ady, adxcdy = adx * cdy;
    const double adxbdy = adx * bdy, bdxady = bdx * ady;
    const double alift = adx * adx + ady * ady;
    const double blift = bdx * bdx + bdy * bdy;
    const double clift = cdx * cdx + cdy * cdy;
    const double det = alift * (bdxcdy - cdxbdy) + blift * (cdxady - adxcdy) + clift * (adxbdy - bdxady);
    const double permanent = (fabs(bdxcdy) + fabs(cdxbdy)) * alift +
        (fabs(cdxady) + fabs(adxcdy)) * blift + (fabs(adxbdy) + fabs(bdxady)) * clift;
    const double bound = ERR_INCIRCLE * permanent;
    if ((det > bound || -det > bound) && permanent > FILTER_FLOOR) return det;
    return (double)orc_incircle_exact(ax, ay, bx, by, cx, cy, dx, dy);
}

/* ------------------------------------------------------------------------ */
/* Builder state                                                            */
/* ------------------------------------------------------------------------ */

typedef struct
{
    double x, y;
    uint32_t id; /* lowest input index with these coordinates */
} vtx_t;

typedef struct
{
    int next[4]; /* onext of the four directed edges 4k+r */
    int org[2];  /* origin of 4k and of 4k+2 (-1: record unused) */
} qrec_t;

typedef struct
{
    vtx_t* V;
    qrec_t* q;
    int nq, capq;
    int free_head;
    int oom;
    uint64_t st[6];
} bld_t;

/* builder.hxx:71-76: frame 0 orders by x then y; frame 1 by y, then x descending */
static int less_f(const vtx_t* a, const vtx_t* b, int frame)
{
    if (frame == 0) return a->x < b->x || (a->x == b->x && a->y < b->y);
    return a->y < b->y || (a->y == b->y && a->x > b->x);
}

/* builder.hxx:78-80 */
static int e_rot(int e) { return (e & ~3) | ((e + 1) & 3); }
static int e_invrot(int e) { return (e & ~3) | ((e + 3) & 3); }
static int e_sym(int e) { return e ^ 2; }

/* builder.hxx:123-129 */
static int onext(const bld_t* B, int e) { return B->q[e >> 2].next[e & 3]; }
static int org(const bld_t* B, int e) { return B->q[e >> 2].org[(e >> 1) & 1]; }
static int dest(const bld_t* B, int e) { return org(B, e_sym(e)); }
static void set_org(bld_t* B, int e, int v) { B->q[e >> 2].org[(e >> 1) & 1] = v; }
static int oprev(const bld_t* B, int e) { return e_rot(onext(B, e_rot(e))); }
static int lnext(const bld_t* B, int e) { return e_rot(onext(B, e_invrot(e))); }
static int rprev(const bld_t* B, int e) { return onext(B, e_sym(e)); }

/* Serial edge storage with a free list (GuibasStolfiDwyer_Wrapper.cpp:29-54) */
static int rec_alloc(bld_t* B)
{
    if (B->free_head >= 0)
    {
        const int k = B->free_head;
        B->free_head = B->q[k].next[0];
        return k;
    }
    if (B->nq == B->capq)
    {
        const int nc = B->capq * 2 + 16;
        qrec_t* nq = (qrec_t*)realloc(B->q, sizeof(qrec_t) * (size_t)nc);
        if (!nq)
        {
            B->oom = 1;
            return 0;
        }
        B->q = nq;
        B->capq = nc;
    }
    B->st[4]++;
    return B->nq++;
}

static void rec_release(bld_t* B, int k)
{
    B->q[k].org[0] = -1;
    B->q[k].next[0] = B->free_head;
    B->free_head = k;
}

/* builder.hxx:131-142 */
static int make_edge(bld_t* B)
{
    const int k = rec_alloc(B);
    const int e = 4 * k;
    qrec_t* r = &B->q[k];
    r->next[0] = e;
    r->next[1] = e + 3;
    r->next[2] = e + 2;
    r->next[3] = e + 1;
    r->org[0] = r->org[1] = -1;
    return e;
}

/* builder.hxx:144-149 */
static void splice(bld_t* B, int a, int b)
{
    const int alpha = e_rot(onext(B, a)), beta = e_rot(onext(B, b));
    int* pa = &B->q[a >> 2].next[a & 3];
    int* pb = &B->q[b >> 2].next[b & 3];
    int t = *pa;
    *pa = *pb;
    *pb = t;
    int* qa = &B->q[alpha >> 2].next[alpha & 3];
    int* qb = &B->q[beta >> 2].next[beta & 3];
    t = *qa;
    *qa = *qb;
    *qb = t;
}

/* builder.hxx:151-159 */
static int connect(bld_t* B, int a, int b)
{
    const int e = make_edge(B);
    set_org(B, e, dest(B, a));
    set_org(B, e_sym(e), org(B, b));
    splice(B, e, lnext(B, a));
    splice(B, e_sym(e), b);
    return e;
}

/* builder.hxx:161-166 */
static void delete_edge(bld_t* B, int e)
{
    splice(B, e, oprev(B, e));
    splice(B, e_sym(e), oprev(B, e_sym(e)));
    rec_release(B, e >> 2);
}

static double orient_v(bld_t* B, int a, int b, int c)
{
    const vtx_t *A = &B->V[a], *Bv = &B->V[b], *C = &B->V[c];
    B->st[0]++;
    const double detleft = (A->x - C->x) * (Bv->y - C->y);
    const double detright = (A->y - C->y) * (Bv->x - C->x);
    const double det = detleft - detright;
    const double mag = fabs(detleft) + fabs(detright);
    const double bound = ERR_ORIENT * mag;
    if ((det > bound || -det > bound) && mag > FILTER_FLOOR) return det;
    B->st[1]++;
    return (double)orc_orient2d_exact(A->x, A->y, Bv->x, Bv->y, C->x, C->y);
}

/* builder.hxx:168-175 */
static int ccw(bld_t* B, int a, int b, int c) { return orient_v(B, a, b, c) > 0; }
static int right_of(bld_t* B, int x, int e) { return ccw(B, x, dest(B, e), org(B, e)); }
static int left_of(bld_t* B, int x, int e) { return ccw(B, x, org(B, e), dest(B, e)); }
/* counted variant of orc_incircle for the stats (same filter) */
static int in_circle_counted(bld_t* B, int a, int b, int c, int d)
{
    if (a == d || b == d || c == d || a == b || a == c || b == c) return 0;
    const vtx_t *A = &B->V[a], *Bv = &B->V[b], *C = &B->V[c], *D = &B->V[d];
    B->st[2]++;
    const double adx = A->x - D->x, ady = A->y - D->y;
    const double bdx = Bv->x - D->x, bdy = Bv->y - D->y;
    const double cdx = C->x - D->x, cdy = C->y - D->y;
    const double bdxcdy = bdx * cdy, cdxbdy = cdx * bdy;
    const double cdxady = cdx * ady, adxcdy = adx * cdy;
    const double adxbdy = adx * bdy, bdxady = bdx * ady;
    const double alift = adx * adx + ady * ady;
    const double blift = bdx * bdx + bdy * bdy;
    const double clift = cdx * cdx + cdy * cdy;
    const double det = alift * (bdxcdy - cdxbdy) + blift * (cdxady - adxcdy) + clift * (adxbdy - bdxady);
    const double permanent = (fabs(bdxcdy) + fabs(cdxbdy)) * alift +
        (fabs(cdxady) + fabs(adxcdy)) * blift + (fabs(adxbdy) + fabs(bdxady)) * clift;
    const double bound = ERR_INCIRCLE * permanent;
    if ((det > bound || -det > bound) && permanent > FILTER_FLOOR) return det > 0;
    B->st[3]++;
    return orc_incircle_exact(A->x, A->y, Bv->x, Bv->y, C->x, C->y, D->x, D->y) > 0;
}

typedef struct
{
    int first, second;
} epair;

static void sort_range(vtx_t* V, int lo, int hi, int frame)
{
    /* insertion sort: leaves hold two or three points */
    for (int i = lo + 1; i < hi; ++i)
    {
        vtx_t t = V[i];
        int j = i - 1;
        while (j >= lo && less_f(&t, &V[j], frame))
        {
            V[j + 1] = V[j];
            --j;
        }
        V[j + 1] = t;
    }
}

/* builder.hxx:177-201 */
static epair leaf(bld_t* B, int lo, int hi, int frame)
{
    sort_range(B->V, lo, hi, frame);
    const int n = hi - lo;
    const int s1 = lo, s2 = lo + 1;
    epair r;
    if (n == 2)
    {
        const int a = make_edge(B);
        set_org(B, a, s1);
        set_org(B, e_sym(a), s2);
        r.first = a;
        r.second = e_sym(a);
        return r;
    }
    const int s3 = lo + 2;
    const int a = make_edge(B), b = make_edge(B);
    splice(B, e_sym(a), b);
    set_org(B, a, s1);
    set_org(B, e_sym(a), s2);
    set_org(B, b, s2);
    set_org(B, e_sym(b), s3);
    const double o = orient_v(B, s1, s2, s3);
    if (o > 0)
    {
        connect(B, b, a);
        r.first = a;
        r.second = e_sym(b);
        return r;
    }
    if (o < 0)
    {
        const int c = connect(B, b, a);
        r.first = e_sym(c);
        r.second = c;
        return r;
    }
    r.first = a;
    r.second = e_sym(b);
    return r;
}

/* builder.hxx:251-285 (with the two scans of :205-249 run one after the other) */
static epair merge(bld_t* B, int ldo, int ldi, int rdi, int rdo)
{
    for (;;)
    {
        if (left_of(B, org(B, rdi), ldi)) ldi = lnext(B, ldi);
        else if (right_of(B, org(B, ldi), rdi)) rdi = rprev(B, rdi);
        else break;
    }
    int basel = connect(B, e_sym(rdi), ldi);
    if (org(B, ldi) == org(B, ldo)) ldo = e_sym(basel);
    if (org(B, rdi) == org(B, rdo)) rdo = basel;

    for (;;)
    {
        const int bo = org(B, basel);
        const int bd = dest(B, basel);
        /* left scan */
        int lc = onext(B, e_sym(basel));
        int lv = ccw(B, dest(B, lc), bd, bo);
        if (lv)
        {
            int deleted = 0;
            while (in_circle_counted(B, bd, bo, dest(B, lc), dest(B, onext(B, lc))))
            {
                const int t = onext(B, lc);
                delete_edge(B, lc);
                lc = t;
                deleted = 1;
            }
            if (deleted) lv = ccw(B, dest(B, lc), bd, bo);
        }
        /* right scan */
        int rc = oprev(B, basel);
        int rv = ccw(B, dest(B, rc), bd, bo);
        if (rv)
        {
            int deleted = 0;
            while (in_circle_counted(B, bd, bo, dest(B, rc), dest(B, oprev(B, rc))))
            {
                const int t = oprev(B, rc);
                delete_edge(B, rc);
                rc = t;
                deleted = 1;
            }
            if (deleted) rv = ccw(B, dest(B, rc), bd, bo);
        }
        if (!lv && !rv) break;
        if (!lv || (rv && in_circle_counted(B, dest(B, lc), bd, bo, dest(B, rc))))
            basel = connect(B, rc, e_sym(basel));
        else
            basel = connect(B, e_sym(basel), e_sym(lc));
    }
    epair r = { ldo, rdo };
    return r;
}

/* builder.hxx:287-301 */
static int hull_walk(bld_t* B, int e, int frame, int want_max)
{
    for (;;)
    {
        const int nx = rprev(B, e);
        const int pv = e_sym(oprev(B, e));
        const vtx_t* ve = &B->V[org(B, e)];
        const vtx_t* vn = &B->V[org(B, nx)];
        const vtx_t* vp = &B->V[org(B, pv)];
        const int better_n = want_max ? less_f(ve, vn, frame) : less_f(vn, ve, frame);
        if (better_n)
        {
            e = nx;
            continue;
        }
        const int better_p = want_max ? less_f(ve, vp, frame) : less_f(vp, ve, frame);
        if (better_p)
        {
            e = pv;
            continue;
        }
        return e;
    }
}

/* Partition V[lo, hi) so that V[mid] is the element of rank mid - lo in the
 * given frame (std::nth_element's contract, builder.hxx:309-325).  Points are
 * distinct, so the set that ends up in [lo, mid) is unique. */
static void select_nth(vtx_t* V, int lo, int hi, int mid, int frame)
{
    while (hi - lo > 8)
    {
        /* median of three as pivot */
        const int a = lo, b = lo + (hi - lo) / 2, c = hi - 1;
        int p;
        if (less_f(&V[a], &V[b], frame))
            p = less_f(&V[b], &V[c], frame) ? b : (less_f(&V[a], &V[c], frame) ? c : a);
        else
            p = less_f(&V[a], &V[c], frame) ? a : (less_f(&V[b], &V[c], frame) ? c : b);
        const vtx_t pivot = V[p];
        /* three-way partition: < pivot | == pivot | > pivot */
        int lt = lo, i = lo, gt = hi;
        while (i < gt)
        {
            if (less_f(&V[i], &pivot, frame))
            {
                vtx_t t = V[lt];
                V[lt] = V[i];
                V[i] = t;
                ++lt;
                ++i;
            }
            else if (less_f(&pivot, &V[i], frame))
            {
                --gt;
                vtx_t t = V[gt];
                V[gt] = V[i];
                V[i] = t;
            }
            else
                ++i;
        }
        if (mid < lt) hi = lt;
        else if (mid >= gt) lo = gt;
        else return;
    }
    sort_range(V, lo, hi, frame);
}

/* builder.hxx:303-348 */
static epair build(bld_t* B, int lo, int hi, int axis, int out_axis, int presorted)
{
    const int n = hi - lo;
    if (n <= 3) return leaf(B, lo, hi, out_axis);
    const int mid = lo + n / 2;
    if (!presorted) select_nth(B->V, lo, hi, mid, axis);
    const int child = 1 - axis;
    const epair L = build(B, lo, mid, child, axis, 0);
    const epair R = build(B, mid, hi, child, axis, 0);
    epair res = merge(B, L.first, L.second, R.first, R.second);
    if (out_axis != axis)
    {
        const int le = hull_walk(B, res.first, out_axis, 0);
        const int re_ccw = hull_walk(B, onext(B, res.second), out_axis, 1);
        res.first = le;
        res.second = oprev(B, re_ccw);
    }
    return res;
}

/* ------------------------------------------------------------------------ */
/* Batch entry                                                              */
/* ------------------------------------------------------------------------ */

struct orc_result
{
    uint32_t* tris;
    size_t ntri;
    uint32_t* rep;
    size_t n;
    size_t unique;
    uint64_t st[6];
};

static int cmp_xyid(const void* pa, const void* pb)
{
    const vtx_t* a = (const vtx_t*)pa;
    const vtx_t* b = (const vtx_t*)pb;
    if (a->x != b->x) return a->x < b->x ? -1 : 1;
    if (a->y != b->y) return a->y < b->y ? -1 : 1;
    return a->id < b->id ? -1 : (a->id > b->id ? 1 : 0);
}

/* GuibasStolfiDwyer_Wrapper.cpp:1287-1305 with builder.hxx:356-390 and :98-116 */
orc_result* orc_triangulate(const double* xy, size_t n)
{
    orc_result* R = (orc_result*)calloc(1, sizeof(orc_result));
    if (!R) return NULL;
    R->n = n;
    R->rep = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    vtx_t* V = (vtx_t*)malloc(sizeof(vtx_t) * (n ? n : 1));
    if (!R->rep || !V)
    {
        free(V);
        orc_free(R);
        return NULL;
    }
    /* load_unique_points: sort by (x, y, index); the first of each run of equal
     * coordinates is the representative of the run (builder.hxx:356-390) */
    for (size_t i = 0; i < n; ++i)
    {
        V[i].x = xy[2 * i];
        V[i].y = xy[2 * i + 1];
        V[i].id = (uint32_t)i;
        R->rep[i] = (uint32_t)i;
    }
    qsort(V, n, sizeof(vtx_t), cmp_xyid);
    size_t m = 0;
    for (size_t i = 0; i < n; ++i)
    {
        if (m > 0 && V[m - 1].x == V[i].x && V[m - 1].y == V[i].y)
        {
            R->rep[V[i].id] = V[m - 1].id;
            continue;
        }
        V[m++] = V[i];
    }
    R->unique = m;

    if (m >= 3)
    {
        bld_t B;
        memset(&B, 0, sizeof(B));
        B.V = V;
        B.capq = (int)(3 * m + 16);
        B.q = (qrec_t*)malloc(sizeof(qrec_t) * (size_t)B.capq);
        B.free_head = -1;
        if (!B.q)
        {
            free(V);
            orc_free(R);
            return NULL;
        }
        build(&B, 0, (int)m, 0, 0, 1);
        if (B.oom)
        {
            free(B.q);
            free(V);
            orc_free(R);
            return NULL;
        }
        /* extract (builder.hxx:98-116) */
        size_t cap = 2 * m + 8;
        R->tris = (uint32_t*)malloc(sizeof(uint32_t) * 3 * cap);
        size_t nt = 0;
        uint64_t live = 0;
        for (int k = 0; k < B.nq; ++k)
        {
            if (B.q[k].org[0] < 0) continue;
            ++live;
            for (int r = 0; r < 4; r += 2)
            {
                const int e = 4 * k + r;
                const int f = lnext(&B, e), g = lnext(&B, f);
                if (lnext(&B, g) != e) continue;
                if (e > f || e > g) continue;
                const int a = org(&B, e), b = org(&B, f), c = org(&B, g);
                if (orient_v(&B, a, b, c) > 0)
                {
                    if (nt == cap)
                    {
                        cap *= 2;
                        uint32_t* t = (uint32_t*)realloc(R->tris, sizeof(uint32_t) * 3 * cap);
                        if (!t)
                        {
                            free(B.q);
                            free(V);
                            orc_free(R);
                            return NULL;
                        }
                        R->tris = t;
                    }
                    R->tris[3 * nt] = V[a].id;
                    R->tris[3 * nt + 1] = V[b].id;
                    R->tris[3 * nt + 2] = V[c].id;
                    ++nt;
                }
            }
        }
        R->ntri = nt;
        B.st[5] = live;
        memcpy(R->st, B.st, sizeof(R->st));
        free(B.q);
    }
    free(V);
    return R;
}

size_t orc_ntri(const orc_result* r) { return r ? r->ntri : 0; }
const uint32_t* orc_tris(const orc_result* r) { return r ? r->tris : NULL; }
const uint32_t* orc_rep(const orc_result* r) { return r ? r->rep : NULL; }
size_t orc_unique(const orc_result* r) { return r ? r->unique : 0; }
void orc_stats(const orc_result* r, uint64_t stats[6])
{
    for (int i = 0; i < 6; ++i) stats[i] = r ? r->st[i] : 0;
}
void orc_free(orc_result* r)
{
    if (!r) return;
    free(r->tris);
    free(r->rep);
    free(r);
}

/* ------------------------------------------------------------------------ */
/* Canonical form and hash (test/qdt_verify.cpp:408-427)                      */
/* ------------------------------------------------------------------------ */

static int cmp_tri(const void* pa, const void* pb)
{
    const uint32_t* a = (const uint32_t*)pa;
    const uint32_t* b = (const uint32_t*)pb;
    for (int i = 0; i < 3; ++i)
        if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
    return 0;
}

void orc_canonicalize(uint32_t* t, size_t ntri)
{
    for (size_t i = 0; i < ntri; ++i)
    {
        uint32_t* p = t + 3 * i;
        int k = 0;
        if (p[1] < p[k]) k = 1;
        if (p[2] < p[k]) k = 2;
        const uint32_t a = p[k], b = p[(k + 1) % 3], c = p[(k + 2) % 3];
        p[0] = a;
        p[1] = b;
        p[2] = c;
    }
    qsort(t, ntri, 3 * sizeof(uint32_t), cmp_tri);
}

uint64_t orc_canon_hash(const uint32_t* tris, size_t ntri)
{
    uint32_t* t = (uint32_t*)malloc(sizeof(uint32_t) * 3 * (ntri ? ntri : 1));
    if (!t) return 0;
    memcpy(t, tris, sizeof(uint32_t) * 3 * ntri);
    orc_canonicalize(t, ntri);
    uint64_t h = 0xcbf29ce484222325ULL;
    for (size_t i = 0; i < 3 * ntri; ++i)
    {
        uint64_t v = t[i];
        for (int b = 0; b < 8; ++b)
        {
            h ^= (v >> (8 * b)) & 0xff;
            h *= 0x100000001b3ULL;
        }
    }
    free(t);
    return h;
}

/* ------------------------------------------------------------------------ */
/* Exact checker                                                            */
/* ------------------------------------------------------------------------ */

static const double* g_xy; /* for qsort comparators (single-threaded use) */

static int cmp_idx_xy(const void* pa, const void* pb)
{
    const uint32_t a = *(const uint32_t*)pa, b = *(const uint32_t*)pb;
    const double ax = g_xy[2 * a], ay = g_xy[2 * a + 1], bx = g_xy[2 * b], by = g_xy[2 * b + 1];
    if (ax != bx) return ax < bx ? -1 : 1;
    if (ay != by) return ay < by ? -1 : 1;
    return a < b ? -1 : (a > b ? 1 : 0);
}

static int orient_idx(const double* xy, uint32_t a, uint32_t b, uint32_t c)
{
    return orc_orient2d_exact(xy[2 * a], xy[2 * a + 1], xy[2 * b], xy[2 * b + 1], xy[2 * c], xy[2 * c + 1]);
}

/* distinct points in (x, y) order, as lowest indices; returns m */
static size_t distinct_sorted(const double* xy, size_t n, uint32_t* order, uint32_t* rep)
{
    for (size_t i = 0; i < n; ++i) order[i] = (uint32_t)i;
    g_xy = xy;
    qsort(order, n, sizeof(uint32_t), cmp_idx_xy);
    size_t m = 0;
    for (size_t i = 0; i < n; ++i)
    {
        const uint32_t o = order[i];
        if (m > 0 && xy[2 * order[m - 1]] == xy[2 * o] && xy[2 * order[m - 1] + 1] == xy[2 * o + 1])
        {
            if (rep) rep[o] = order[m - 1];
            continue;
        }
        if (rep) rep[o] = o;
        order[m++] = o;
    }
    return m;
}

/* monotone chain on distinct sorted points, collinear points dropped
 * (test/qdt_verify.cpp:440-462) */
static size_t hull_of_sorted(const double* xy, const uint32_t* p, size_t m, uint32_t* h)
{
    if (m < 3)
    {
        for (size_t i = 0; i < m; ++i) h[i] = p[i];
        return m;
    }
    size_t k = 0;
    for (size_t i = 0; i < m; ++i)
    {
        while (k >= 2 && orient_idx(xy, h[k - 2], h[k - 1], p[i]) <= 0) --k;
        h[k++] = p[i];
    }
    for (size_t i = m - 1, t = k + 1; i > 0; --i)
    {
        while (k >= t && orient_idx(xy, h[k - 2], h[k - 1], p[i - 1]) <= 0) --k;
        h[k++] = p[i - 1];
    }
    return k - 1;
}

size_t orc_hull(const double* xy, size_t n, uint32_t* out)
{
    uint32_t* order = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t* h = (uint32_t*)malloc(sizeof(uint32_t) * (2 * n + 2));
    if (!order || !h)
    {
        free(order);
        free(h);
        return 0;
    }
    const size_t m = distinct_sorted(xy, n, order, NULL);
    const size_t k = hull_of_sorted(xy, order, m, h);
    memcpy(out, h, sizeof(uint32_t) * k);
    free(order);
    free(h);
    return k;
}

typedef struct
{
    uint32_t a, b, c; /* directed edge a->b of the triangle whose third vertex is c */
} dedge_t;

static int cmp_dedge(const void* pa, const void* pb)
{
    const dedge_t* a = (const dedge_t*)pa;
    const dedge_t* b = (const dedge_t*)pb;
    if (a->a != b->a) return a->a < b->a ? -1 : 1;
    if (a->b != b->b) return a->b < b->b ? -1 : 1;
    return 0;
}

static const dedge_t* find_dedge(const dedge_t* es, size_t ne, uint32_t a, uint32_t b)
{
    size_t lo = 0, hi = ne;
    while (lo < hi)
    {
        const size_t mid = lo + (hi - lo) / 2;
        if (es[mid].a < a || (es[mid].a == a && es[mid].b < b)) lo = mid + 1;
        else hi = mid;
    }
    if (lo < ne && es[lo].a == a && es[lo].b == b) return &es[lo];
    return NULL;
}

/* exact twice-area accumulator: acc += orient(a, b, c) as an expansion */
static int area_add(double* acc, int nacc, const double* xy, uint32_t a, uint32_t b, uint32_t c)
{
    /* (bx - ax)(cy - ay) - (by - ay)(cx - ax), all exact */
    double bax[2], cay[2], bay[2], cax[2], m1[XCAP], m2[XCAP], s[XCAP], t[XCAP];
    const int n1 = x_diff(xy[2 * b], xy[2 * a], bax), n2 = x_diff(xy[2 * c + 1], xy[2 * a + 1], cay);
    const int n3 = x_diff(xy[2 * b + 1], xy[2 * a + 1], bay), n4 = x_diff(xy[2 * c], xy[2 * a], cax);
    const int k1 = x_mul(bax, n1, cay, n2, m1);
    int k2 = x_mul(bay, n3, cax, n4, m2);
    x_neg(m2, k2, m2);
    const int ks = x_add(m1, k1, m2, k2, s);
    const int kt = x_add(acc, nacc, s, ks, t);
    return x_compress(t, kt, acc);
}

#define FAIL(code, ...)                                  \
    do                                                   \
    {                                                    \
        if (msg && msglen) snprintf(msg, msglen, __VA_ARGS__); \
        rc = (code);                                     \
        goto done;                                       \
    } while (0)

static int check_impl(const double* xy, size_t n, const uint32_t* tris, size_t ntri,
                      char* msg, size_t msglen, int64_t info[4]);

/* The checker's area sum multiplies coordinate differences without the
 * predicates' normalisation, so a point set whose largest coordinate is far
 * from 1 is first scaled by a power of two (as robust.hxx:297-314 scales the
 * predicates' operands): exact unless a nonzero coordinate drops below the
 * normal range, which the sets checked here (spread < 2^600) never do. */
int orc_check(const double* xy, size_t n, const uint32_t* tris, size_t ntri,
              char* msg, size_t msglen, int64_t info[4])
{
    int emax = INT_MIN;
    for (size_t i = 0; i < 2 * n; ++i)
        if (xy[i] != 0.0)
        {
            const int e = ilogb(xy[i]);
            if (e > emax) emax = e;
        }
    if (emax == INT_MIN || (emax >= -400 && emax <= 400)) return check_impl(xy, n, tris, ntri, msg, msglen, info);
    double* sc = (double*)malloc(sizeof(double) * 2 * n);
    if (!sc) return 99;
    for (size_t i = 0; i < 2 * n; ++i) sc[i] = ldexp(xy[i], -emax);
    const int rc = check_impl(sc, n, tris, ntri, msg, msglen, info);
    free(sc);
    return rc;
}

static int check_impl(const double* xy, size_t n, const uint32_t* tris, size_t ntri,
                      char* msg, size_t msglen, int64_t info[4])
{
    int rc = 0;
    uint32_t* order = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t* rep = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t* hull = (uint32_t*)malloc(sizeof(uint32_t) * (2 * n + 2));
    unsigned char* used = (unsigned char*)calloc(n ? n : 1, 1);
    dedge_t* es = (dedge_t*)malloc(sizeof(dedge_t) * (3 * ntri + 1));
    uint32_t* bnext = NULL;
    if (msg && msglen) msg[0] = 0;
    if (info) info[0] = info[1] = info[2] = info[3] = 0;
    if (!order || !rep || !hull || !used || !es) FAIL(99, "out of memory");

    const size_t m = distinct_sorted(xy, n, order, rep);
    if (info)
    {
        info[0] = (int64_t)m;
        info[2] = (int64_t)ntri;
    }
    int degenerate = m < 3;
    if (!degenerate)
    {
        degenerate = 1;
        for (size_t i = 2; i < m && degenerate; ++i)
            if (orient_idx(xy, order[0], order[1], order[i]) != 0) degenerate = 0;
    }
    if (degenerate)
    {
        if (ntri) FAIL(1, "degenerate input produced %zu triangles", ntri);
        goto done;
    }
    if (ntri == 0) FAIL(2, "non-degenerate input produced no triangles");

    for (size_t t = 0; t < ntri; ++t)
    {
        const uint32_t* p = tris + 3 * t;
        for (int k = 0; k < 3; ++k)
        {
            if (p[k] >= n) FAIL(3, "triangle %zu: index out of range", t);
            if (rep[p[k]] != p[k]) FAIL(4, "triangle %zu references a duplicate index %u", t, p[k]);
            used[p[k]] = 1;
        }
        if (p[0] == p[1] || p[1] == p[2] || p[0] == p[2]) FAIL(5, "triangle %zu has a repeated vertex", t);
        if (orient_idx(xy, p[0], p[1], p[2]) <= 0) FAIL(6, "triangle %zu is not strictly counter-clockwise", t);
        for (int k = 0; k < 3; ++k)
        {
            es[3 * t + k].a = p[k];
            es[3 * t + k].b = p[(k + 1) % 3];
            es[3 * t + k].c = p[(k + 2) % 3];
        }
    }
    const size_t ne = 3 * ntri;
    qsort(es, ne, sizeof(dedge_t), cmp_dedge);
    for (size_t i = 1; i < ne; ++i)
        if (es[i].a == es[i - 1].a && es[i].b == es[i - 1].b)
            FAIL(7, "directed edge %u->%u used twice", es[i].a, es[i].b);
    for (size_t i = 0; i < m; ++i)
        if (!used[order[i]]) FAIL(8, "distinct point %u not referenced", order[i]);

    /* boundary edges (no reverse) must form one loop along the convex hull */
    bnext = (uint32_t*)malloc(sizeof(uint32_t) * n);
    if (!bnext) FAIL(99, "out of memory");
    for (size_t i = 0; i < n; ++i) bnext[i] = UINT32_MAX;
    size_t nb = 0, interior = 0;
    for (size_t i = 0; i < ne; ++i)
    {
        const dedge_t* e = &es[i];
        const dedge_t* rv = find_dedge(es, ne, e->b, e->a);
        if (!rv)
        {
            if (bnext[e->a] != UINT32_MAX) FAIL(9, "vertex %u starts two boundary edges", e->a);
            bnext[e->a] = e->b;
            ++nb;
        }
        else if (e->a < e->b)
        {
            ++interior;
            const uint32_t a = e->a, b = e->b, c = e->c, d = rv->c;
            if (orc_incircle_exact(xy[2 * a], xy[2 * a + 1], xy[2 * b], xy[2 * b + 1], xy[2 * c], xy[2 * c + 1],
                                   xy[2 * d], xy[2 * d + 1]) > 0)
                FAIL(10, "interior edge %u-%u is not locally Delaunay", a, b);
        }
    }
    const size_t h = hull_of_sorted(xy, order, m, hull);
    if (info)
    {
        info[1] = (int64_t)nb;
        info[3] = (int64_t)(nb + interior);
    }
    /* Walk the boundary loop from the first hull vertex: it must visit the hull
     * vertices in order, and every vertex between two of them must lie on the
     * hull edge joining them. */
    {
        uint32_t cur = hull[0];
        size_t hi = 0, steps = 0;
        do
        {
            const uint32_t nx = bnext[cur];
            if (nx == UINT32_MAX) FAIL(11, "hull vertex %u has no outgoing boundary edge", cur);
            const uint32_t ha = hull[hi], hb = hull[(hi + 1) % h];
            if (orient_idx(xy, ha, hb, nx) != 0) FAIL(12, "boundary edge %u->%u is not on the convex hull", cur, nx);
            if (nx == hb) hi = (hi + 1) % h;
            cur = nx;
            if (++steps > nb) FAIL(13, "boundary loop longer than the boundary edge count");
        } while (cur != hull[0]);
        if (steps != nb) FAIL(14, "boundary edges form more than one loop (%zu of %zu)", steps, nb);
        if (hi != 0) FAIL(15, "boundary loop does not go around the hull");
    }
    /* Euler: T = 2m - 2 - b for a triangulation of the hull with b boundary vertices */
    if (ntri != 2 * m - 2 - nb) FAIL(16, "triangle count %zu != 2m-2-b = %zu", ntri, 2 * m - 2 - nb);
    /* exact area: sum of triangle areas == hull area */
    {
        double acc[XCAP];
        int nacc = 0;
        for (size_t t = 0; t < ntri; ++t)
            nacc = area_add(acc, nacc, xy, tris[3 * t], tris[3 * t + 1], tris[3 * t + 2]);
        double hac[XCAP];
        int nh = 0;
        for (size_t i = 1; i + 1 < h; ++i) nh = area_add(hac, nh, xy, hull[0], hull[i], hull[i + 1]);
        double neg[XCAP], d[XCAP];
        x_neg(hac, nh, neg);
        const int nd = x_add(acc, nacc, neg, nh, d);
        if (x_sign(d, nd) != 0) FAIL(17, "triangle areas do not add up to the hull area");
    }
done:
    free(order);
    free(rep);
    free(hull);
    free(used);
    free(es);
    free(bnext);
    return rc;
}

/* Triangulation::ensure_mesh_, libqdt/GuibasStolfiDwyer_Wrapper.cpp:99-202. */
long orc_build_mesh(const uint32_t* tris, size_t ntri, size_t n, int32_t* he, int32_t* vert, int32_t* o2l)
{
    const size_t nhe = 3 * ntri;
    for (size_t i = 0; i < n; ++i) o2l[i] = -1;
    for (size_t i = 0; i < nhe; ++i)
        if (tris[i] >= n) return -1;
    /* local vertices by first appearance (:111-131) */
    long nv = 0;
    for (size_t i = 0; i < nhe; ++i)
    {
        const uint32_t orig = tris[i];
        if (o2l[orig] < 0)
        {
            o2l[orig] = (int32_t)nv;
            vert[2 * nv] = (int32_t)orig;
            vert[2 * nv + 1] = -1;
            ++nv;
        }
    }
    /* half-edges of each triangle, first half-edge of each vertex (:133-154) */
    for (size_t t = 0; t < ntri; ++t)
        for (int i = 0; i < 3; ++i)
        {
            const size_t h = 3 * t + (size_t)i;
            const int32_t v = o2l[tris[h]];
            he[4 * h] = v;
            he[4 * h + 1] = -1;
            he[4 * h + 2] = (int32_t)(3 * t + (size_t)((i + 1) % 3));
            he[4 * h + 3] = (int32_t)t;
            if (vert[2 * v + 1] < 0) vert[2 * v + 1] = (int32_t)h;
        }
    /* out-going half-edges of each vertex, CSR (:156-172) */
    int32_t* off = (int32_t*)calloc((size_t)nv + 1, sizeof(int32_t));
    int32_t* adj = (int32_t*)malloc((nhe ? nhe : 1) * sizeof(int32_t));
    int32_t* cur = (int32_t*)malloc(((size_t)nv + 1) * sizeof(int32_t));
    if (!off || !adj || !cur)
    {
        free(off);
        free(adj);
        free(cur);
        return -2;
    }
    for (size_t i = 0; i < nhe; ++i) ++off[he[4 * i] + 1];
    for (long v = 0; v < nv; ++v) off[v + 1] += off[v];
    for (long v = 0; v < nv; ++v) cur[v] = off[v];
    for (size_t i = 0; i < nhe; ++i) adj[cur[he[4 * i]]++] = (int32_t)i;
    /* twins: the half-edge out of the far end that comes back (:174-195) */
    for (size_t i = 0; i < nhe; ++i)
    {
        if (he[4 * i + 1] >= 0) continue;
        const int32_t u = he[4 * i], v = he[4 * he[4 * i + 2]];
        for (int32_t k = off[v]; k < off[v + 1]; ++k)
        {
            const int32_t h2 = adj[k];
            if (he[4 * h2] != v) continue;
            if (he[4 * he[4 * h2 + 2]] == u)
            {
                he[4 * i + 1] = h2;
                he[4 * h2 + 1] = (int32_t)i;
                break;
            }
        }
    }
    free(off);
    free(adj);
    free(cur);
    return nv;
}
