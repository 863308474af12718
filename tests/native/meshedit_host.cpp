// meshedit_host.cpp -- the library's host-side mesh editor (point insertion
// and constrained edges, libqdt_amd/csrc/mesh_edit.cpp) built on its own for
// the CPU tests, seeded with triangulations from the oracle.  Test
// infrastructure (built by __graft_entry__.build).
#include "../../libqdt_amd/csrc/mesh_edit.hpp"

#include <new>

extern "C" {

void* me_create() { return new (std::nothrow) qdt::MeshEdit(); }

void me_destroy(void* m) { delete static_cast<qdt::MeshEdit*>(m); }

int me_build(void* m, const double* xy, size_t n, const uint32_t* rep, const uint32_t* tris, size_t T)
{
    return static_cast<qdt::MeshEdit*>(m)->build(reinterpret_cast<const qdt_vec2d*>(xy), n, rep, tris, T, nullptr)
               ? 1
               : 0;
}

int me_insert(void* m, double x, double y, size_t* index)
{
    return static_cast<qdt::MeshEdit*>(m)->insert(qdt_vec2d{ x, y }, *index);
}

int me_constrain(void* m, size_t v0, size_t v1) { return static_cast<qdt::MeshEdit*>(m)->constrain(v0, v1); }

size_t me_triangles(void* m) { return static_cast<qdt::MeshEdit*>(m)->triangles(); }

size_t me_points(void* m) { return static_cast<qdt::MeshEdit*>(m)->points(); }

void me_copy(void* m, uint32_t* tris, uint32_t* rep)
{
    const qdt::MeshEdit* e = static_cast<qdt::MeshEdit*>(m);
    if (tris)
        for (size_t i = 0; i < 3 * e->triangles(); ++i) tris[i] = e->tris()[i];
    if (rep)
        for (size_t i = 0; i < e->points(); ++i) rep[i] = e->rep()[i];
}

size_t me_constrained_edges(void* m) { return static_cast<qdt::MeshEdit*>(m)->constrained_edges(); }
}
