// ref_shim.cpp -- exports the reference's own predicates (libqdt/robust.hxx)
// so that the oracle's restatement can be checked against them.  Compiled only
// into oracle/_ref/libqdt_ref.so, together with the reference's libqdt/*.cpp
// taken from where they lie; test infrastructure, never part of the product.
#include "robust.hxx"

extern "C" double ref_orient2d(double ax, double ay, double bx, double by, double cx, double cy)
{
    const wrapper::Point a = { ax, ay }, b = { bx, by }, c = { cx, cy };
    return wrapper::robust::orient2d(a, b, c);
}

extern "C" double ref_incircle(double ax, double ay, double bx, double by, double cx, double cy,
                               double dx, double dy)
{
    const wrapper::Point a = { ax, ay }, b = { bx, by }, c = { cx, cy }, d = { dx, dy };
    return wrapper::robust::incircle(a, b, c, d);
}
