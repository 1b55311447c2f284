// kifi/Mesher.h -- marching cubes over the sparse volume (reference:
// KinectFission/kifi/Mesher.h:13-40, Mesher.cpp:14-229).
#pragma once

#include <cstdio>
#include <vector>

#include "kifi/Volume.h"
#include "kifi/util/math.h"

namespace kifi {

class Mesher {
public:
    // vertices in the reference's generation order
    void Mesh(const Volume &volume, std::vector<util::float3> &outVertices)
    {
        std::size_t n = 0;
        kifi_status s = kifi_mesh(volume.Handle(), nullptr, 0, &n);
        if (s != KIFI_OK && s != KIFI_ERR_CAPACITY)
            detail::check(s, "kifi_mesh");
        outVertices.resize(n);
        detail::check(kifi_mesh(volume.Handle(), n ? &outVertices[0].x : nullptr, n, &n), "kifi_mesh");
    }

    // vertices ordered by vertex ID, 3 indices per triangle
    void Mesh(const Volume &volume, std::vector<util::float3> &outVertices, std::vector<unsigned> &outIndices)
    {
        std::size_t nv = 0, ni = 0;
        kifi_status s = kifi_mesh_indexed(volume.Handle(), nullptr, 0, nullptr, 0, &nv, &ni, nullptr);
        if (s != KIFI_OK && s != KIFI_ERR_CAPACITY)
            detail::check(s, "kifi_mesh_indexed");
        outVertices.resize(nv);
        outIndices.resize(ni);
        detail::check(kifi_mesh_indexed(volume.Handle(), nv ? &outVertices[0].x : nullptr, nv,
                                        ni ? outIndices.data() : nullptr, ni, &nv,
                                        &ni, nullptr),
                      "kifi_mesh_indexed");
    }

    // Wavefront OBJ: "v x y z" lines, then 1-based "f a b c" lines (Mesher.cpp:50-78)
    static void Mesh2Obj(const std::vector<util::float3> &vertices, const std::vector<unsigned> &indices,
                         const char *outObjFileName)
    {
        FILE *f = std::fopen(outObjFileName, "w");
        if (!f) {
            std::fprintf(stderr, "kifi: cannot write %s\n", outObjFileName);
            std::abort();
        }
        for (const util::float3 &v : vertices)
            std::fprintf(f, "v %f %f %f\n", v.x, v.y, v.z);
        for (std::size_t i = 0; i + 2 < indices.size(); i += 3)
            std::fprintf(f, "f %u %u %u\n", indices[i] + 1, indices[i + 1] + 1, indices[i + 2] + 1);
        std::fclose(f);
    }
};

} // namespace kifi
