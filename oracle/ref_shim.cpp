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

// Bulk read of a reference context's triangles (test speed only): reads the
// public result fields of the context type declared in libqdt/wrapper.hxx.
#include "qdt.h"
#include "wrapper.hxx"
#include <cstdint>

extern "C" size_t ref_copy_triangles(const qdt_t* q, uint32_t* out)
{
    const wrapper::Triangulation* T = reinterpret_cast<const wrapper::Triangulation*>(q);
    const size_t n = T->triangles.size();
    if (out)
        for (size_t i = 0; i < n; ++i)
            for (int k = 0; k < 3; ++k) out[3 * i + k] = static_cast<uint32_t>(T->triangles[i][k]);
    return n;
}
