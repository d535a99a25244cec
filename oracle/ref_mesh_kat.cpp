// ORACLE -- test infrastructure only.
//
// Mesh-loader checker for the reference's own loaders.  Compiled in the
// authoring container from the reference's sources (parsers/load_serialized.cpp
// with its vendored miniz, parsers/parse_ply.cpp with its vendored tinyply,
// transform.cpp; see oracle/Makefile, target `ref`) into oracle/_ref/.  It
// loads one mesh with the reference's loader, applies the reference parser's
// post-processing (vertex normals from compute_normal when the file has none,
// parse_scene.cpp:1294-1301), and writes a binary dump:
//   u64 nv, nn, nuv, nt; f64 positions[3 nv], normals[3 nn], uvs[2 nuv];
//   i32 indices[3 nt]
//
//   ref_mesh_kat serialized <file> <shape_index> <out> [16 row-major to_world]
//   ref_mesh_kat ply <file> <out> [16 row-major to_world]
#include "parsers/load_serialized.h"
#include "parsers/parse_ply.h"
#include "parsers/shape_utils.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

int main(int argc, char **argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: ref_mesh_kat serialized <file> <shape_index> <out> [m00 .. m33]\n"
                             "       ref_mesh_kat ply <file> <out> [m00 .. m33]\n");
        return 2;
    }
    std::string kind = argv[1];
    const bool ser = kind == "serialized";
    const int xf_at = ser ? 5 : 4;
    if (argc < xf_at) return 2;
    Matrix4x4 xf = Matrix4x4::identity();
    if (argc >= xf_at + 16)
        for (int i = 0; i < 16; i++) xf(i / 4, i % 4) = std::atof(argv[xf_at + i]);
    TriangleMesh m = ser ? load_serialized(argv[2], std::atoi(argv[3]), xf) : parse_ply(argv[2], xf);
    if (m.normals.empty()) m.normals = compute_normal(m.positions, m.indices);
    FILE *f = std::fopen(argv[ser ? 4 : 3], "wb");
    if (!f) return 1;
    uint64_t head[4] = {m.positions.size(), m.normals.size(), m.uvs.size(), m.indices.size()};
    std::fwrite(head, 8, 4, f);
    for (auto &p : m.positions) std::fwrite(&p.x, 8, 3, f);
    for (auto &n : m.normals) std::fwrite(&n.x, 8, 3, f);
    for (auto &t : m.uvs) std::fwrite(&t.x, 8, 2, f);
    for (auto &i : m.indices) std::fwrite(&i.x, 4, 3, f);
    std::fclose(f);
    return 0;
}
