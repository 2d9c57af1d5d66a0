// hpGEM .hpgem mesh files: a host-side reader (no device needed).
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace hgdg {

// Reference geometries a file may hold (values of the HGDG_SHAPE_* enum).
enum MeshShape {
    MESH_SHAPE_LINE = 1,
    MESH_SHAPE_SQUARE = 2,
    MESH_SHAPE_TRIANGLE = 3,
    MESH_SHAPE_CUBE = 4,
    MESH_SHAPE_TETRAHEDRON = 5,
    MESH_SHAPE_PRISM = 6,
    MESH_SHAPE_PYRAMID = 7
};

struct MeshReadError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

struct MeshFile {
    int version = 0, dim = 0, shape = 0, n_vertices = 0;
    int64_t n_nodes = 0, n_elements = 0, n_faces = 0, n_edges = 0, n_boundary_faces = 0;
    std::vector<std::string> zone_names;
    std::vector<double> coords;          // [element][vertex][dim], the element's coordinate copies
    std::vector<int64_t> node_ids;       // [element][vertex]
    std::vector<int32_t> zones;          // [element]
    std::vector<int64_t> face_elements;  // [face][2]: element on each side, -1 for none
    std::vector<int32_t> face_local;     // [face][2]: local face index on each side, -1 for none
};

// Parse a version-1 or version-2 file with a single partition; throws
// MeshReadError with file:line on malformed input.
MeshFile read_mesh_file(const std::string& path);

// The text of a version-2, single-partition file for a mesh given as the
// reader returns it (coords, node_ids, zones, zone_names; shape and
// n_vertices set; faces are not needed): nodes 0 .. n-1 with their distinct
// coordinate copies in ascending order, elements, then the codim-1 and
// codim-2 entities numbered by first appearance, element by element, with
// the Preprocessor's local entity order. Throws MeshReadError on
// inconsistent input.
std::string format_mesh_file(const MeshFile& m);

}  // namespace hgdg
