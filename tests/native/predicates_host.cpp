// predicates_host.cpp -- host build of the GPU predicate code, so that its
// exact arithmetic can be checked on the CPU against the oracle and against
// integer closed forms.  Test infrastructure (built by __graft_entry__.build).
#include "../../libqdt_amd/csrc/predicates.hpp"

extern "C" {

int ph_orient2d_exact(double ax, double ay, double bx, double by, double cx, double cy)
{
    return qdt::orient2d_exact(ax, ay, bx, by, cx, cy);
}

int ph_incircle_exact(double ax, double ay, double bx, double by, double cx, double cy, double dx, double dy)
{
    return qdt::incircle_exact(ax, ay, bx, by, cx, cy, dx, dy);
}

// the general (wide accumulator) path, bypassing the 62-bit fast path
int ph_orient2d_general(double ax, double ay, double bx, double by, double cx, double cy)
{
    double hi[4], lo[4];
    qdt::exact_diff(ax, cx, hi[0], lo[0]);
    qdt::exact_diff(ay, cy, hi[1], lo[1]);
    qdt::exact_diff(bx, cx, hi[2], lo[2]);
    qdt::exact_diff(by, cy, hi[3], lo[3]);
    return qdt::orient_general(hi, lo);
}

int ph_incircle_general(double ax, double ay, double bx, double by, double cx, double cy, double dx, double dy)
{
    double hi[6], lo[6];
    qdt::exact_diff(ax, dx, hi[0], lo[0]);
    qdt::exact_diff(ay, dy, hi[1], lo[1]);
    qdt::exact_diff(bx, dx, hi[2], lo[2]);
    qdt::exact_diff(by, dy, hi[3], lo[3]);
    qdt::exact_diff(cx, dx, hi[4], lo[4]);
    qdt::exact_diff(cy, dy, hi[5], lo[5]);
    return qdt::incircle_general(hi, lo);
}

double ph_orient2d(double ax, double ay, double bx, double by, double cx, double cy, unsigned* nexact)
{
    return qdt::orient2d(ax, ay, bx, by, cx, cy, *nexact);
}

double ph_incircle(double ax, double ay, double bx, double by, double cx, double cy, double dx, double dy,
                   unsigned* nexact)
{
    return qdt::incircle(ax, ay, bx, by, cx, cy, dx, dy, *nexact);
}

// the filters with bounds fixed from an extent (wx, wy), as the LDS cells use them
void ph_static_bounds(double wx, double wy, double* b_orient, double* b_incircle)
{
    qdt::static_bounds(wx, wy, *b_orient, *b_incircle);
}

double ph_orient2d_det(double ax, double ay, double bx, double by, double cx, double cy)
{
    return qdt::orient2d_det(ax, ay, bx, by, cx, cy);
}

double ph_incircle_det(double ax, double ay, double bx, double by, double cx, double cy, double dx, double dy)
{
    return qdt::incircle_det(ax, ay, bx, by, cx, cy, dx, dy);
}
}
