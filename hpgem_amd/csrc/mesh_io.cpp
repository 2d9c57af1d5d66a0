// Reader and writer of hpGEM's .hpgem mesh files (format versions 1 and 2;
// the writer produces version 2), host only.
//
// The format, as MeshFileInformation::readInformation
// (kernel/Base/MeshFileInformation.cpp:48-103) and MeshManipulator::readMesh
// (kernel/Base/MeshManipulator_Impl.h:1505-1800) read it, and the
// Preprocessor writes it (applications/Preprocessor/lib/output_impl.h:91-208):
//
//   mesh <version>
//   <#nodes> <#elements> <dim> <#codim1> ... <#codim(dim-1)>
//   <#partitions> <#nodes of each partition>
//   zones / <#zones> / one name per line                  (version 2 only)
//   nodes                                                 (version 2 only)
//   per node:    <#partitions> <partition ids>
//                <#coordinates> <coordinates, dim per copy, hex floats>
//   elements                                              (version 2 only)
//   per element: <#nodes> (<node id> <coordinate index>)* <owner> <#shadow> <shadow ids>
//                <zone index>                             (version 2 only)
//   codim1 (dim > 1), then codim2 (dim > 2)               (headers: version 2 only)
//   per entity:  <#elements> (<element> <local index>)* <#partitions> <partition ids>
//
// A node carries one coordinate per periodic copy; an element names the copy
// it uses. Only single-partition files are accepted (one device, one rank),
// which is what the reference requires of a run on one process
// (MeshManipulator_Impl.h:1538-1542).
#include <algorithm>
#include <array>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <set>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "mesh_io.hpp"

namespace hgdg {
namespace {

struct Cursor {
    std::string text;
    size_t pos = 0;
    const std::string& file;

    explicit Cursor(const std::string& f) : file(f) {}

    [[noreturn]] void fail(const std::string& what) const {
        size_t line = 1;
        for (size_t i = 0; i < pos && i < text.size(); ++i) line += text[i] == '\n';
        throw MeshReadError(file + ":" + std::to_string(line) + ": " + what);
    }
    void skip_space() {
        while (pos < text.size() && std::isspace(static_cast<unsigned char>(text[pos]))) ++pos;
    }
    std::string token(const char* what) {
        skip_space();
        const size_t start = pos;
        while (pos < text.size() && !std::isspace(static_cast<unsigned char>(text[pos]))) ++pos;
        if (start == pos) fail(std::string("unexpected end of file, expected ") + what);
        return text.substr(start, pos - start);
    }
    // The rest of the current line (without the newline), then step past it.
    std::string line() {
        const size_t start = pos;
        while (pos < text.size() && text[pos] != '\n') ++pos;
        std::string s = text.substr(start, pos - start);
        if (pos < text.size()) ++pos;
        while (!s.empty() && (s.back() == '\r' || s.back() == ' ' || s.back() == '\t')) s.pop_back();
        return s;
    }
    int64_t integer(const char* what) {
        const std::string t = token(what);
        char* end = nullptr;
        const long long v = std::strtoll(t.c_str(), &end, 10);
        if (*end != '\0' || v < 0) fail(std::string("bad ") + what + " '" + t + "'");
        return int64_t(v);
    }
    double real(const char* what) {
        const std::string t = token(what);
        char* end = nullptr;
        const double v = std::strtod(t.c_str(), &end);  // decimal or C99 hex float
        if (*end != '\0') fail(std::string("bad ") + what + " '" + t + "'");
        return v;
    }
    void header(const char* expected) {
        const std::string t = token(expected);
        if (t != expected) fail(std::string("expected segment '") + expected + "', got '" + t + "'");
    }
    void single_partition(const char* what) {
        const int64_t n = integer("partition count");
        for (int64_t j = 0; j < n; ++j)
            if (integer("partition id") != 0) fail(std::string(what) + " lies in a partition other than 0");
    }
};

// hpGEM reference geometry from (dimension, vertices per element).
int shape_of(int dim, int nv) {
    if (dim == 1 && nv == 2) return MESH_SHAPE_LINE;
    if (dim == 2 && nv == 3) return MESH_SHAPE_TRIANGLE;
    if (dim == 2 && nv == 4) return MESH_SHAPE_SQUARE;
    if (dim == 3 && nv == 4) return MESH_SHAPE_TETRAHEDRON;
    if (dim == 3 && nv == 5) return MESH_SHAPE_PYRAMID;
    if (dim == 3 && nv == 6) return MESH_SHAPE_PRISM;
    if (dim == 3 && nv == 8) return MESH_SHAPE_CUBE;
    return 0;
}

// Number of codim-1 faces of each reference geometry.
int faces_of(int shape) {
    switch (shape) {
        case MESH_SHAPE_LINE: return 2;
        case MESH_SHAPE_TRIANGLE: return 3;
        case MESH_SHAPE_SQUARE: return 4;
        case MESH_SHAPE_TETRAHEDRON: return 4;
        case MESH_SHAPE_PYRAMID: return 5;
        case MESH_SHAPE_PRISM: return 5;
        case MESH_SHAPE_CUBE: return 6;
        default: return 0;
    }
}

// Local codim-1 and codim-2 entities of each geometry as vertex lists, in the
// order the Preprocessor numbers them (applications/Preprocessor/lib/mesh/
// ElementShapes.cpp:74-169): the file lists entities in order of first
// appearance under this numbering.
struct LocalEntities {
    std::vector<std::vector<int>> faces, edges;  // edges: only for 3D shapes
};

LocalEntities local_entities(int shape) {
    switch (shape) {
        case MESH_SHAPE_LINE: return {{{0}, {1}}, {}};
        case MESH_SHAPE_TRIANGLE: return {{{0, 1}, {0, 2}, {1, 2}}, {}};
        case MESH_SHAPE_SQUARE: return {{{0, 1}, {0, 2}, {1, 3}, {2, 3}}, {}};
        case MESH_SHAPE_TETRAHEDRON:
            return {{{0, 3, 2}, {0, 1, 3}, {0, 2, 1}, {1, 2, 3}}, {{0, 1}, {0, 2}, {0, 3}, {2, 3}, {1, 3}, {1, 2}}};
        case MESH_SHAPE_CUBE:
            return {{{0, 1, 2, 3}, {0, 1, 4, 5}, {0, 2, 4, 6}, {1, 3, 5, 7}, {2, 3, 6, 7}, {4, 5, 6, 7}},
                    {{0, 1}, {2, 3}, {4, 5}, {6, 7}, {0, 2}, {1, 3}, {4, 6}, {5, 7}, {0, 4}, {1, 5}, {2, 6}, {3, 7}}};
        case MESH_SHAPE_PRISM:
            return {{{0, 2, 1}, {3, 4, 5}, {2, 0, 5, 3}, {0, 1, 3, 4}, {1, 2, 4, 5}},
                    {{0, 1}, {0, 2}, {1, 2}, {3, 4}, {3, 5}, {4, 5}, {0, 3}, {1, 4}, {2, 5}}};
        case MESH_SHAPE_PYRAMID:
            return {{{3, 4, 1, 2}, {3, 1, 0}, {2, 4, 0}, {1, 2, 0}, {4, 3, 0}},
                    {{0, 1}, {0, 2}, {0, 3}, {0, 4}, {1, 2}, {2, 4}, {4, 3}, {3, 1}}};
        default: return {};
    }
}

// Entities of one codimension: key = sorted node ids, value = the list of
// (element, local index) in the order elements meet it.
struct EntityTable {
    std::map<std::vector<int64_t>, int64_t> id;
    std::vector<std::vector<std::pair<int64_t, int>>> uses;
    void add(std::vector<int64_t> key, int64_t element, int local) {
        std::sort(key.begin(), key.end());
        auto it = id.find(key);
        if (it == id.end()) {
            id.emplace(std::move(key), int64_t(uses.size()));
            uses.push_back({{element, local}});
        } else {
            uses[size_t(it->second)].push_back({element, local});
        }
    }
};

std::string hexfloat(double v) {
    char buf[64];
    std::snprintf(buf, sizeof buf, "%a", v);
    return buf;
}

}  // namespace

std::string format_mesh_file(const MeshFile& m) {
    const int dim = m.dim, nv = m.n_vertices;
    const int64_t ne = m.n_elements;
    if (shape_of(dim, nv) != m.shape || m.shape == 0) throw MeshReadError("unknown element geometry");
    if (m.coords.size() != size_t(ne) * nv * dim || m.node_ids.size() != size_t(ne) * nv ||
        (!m.zones.empty() && m.zones.size() != size_t(ne)))
        throw MeshReadError("array sizes do not match the element count");
    const std::vector<std::string> zone_names = m.zone_names.empty() ? std::vector<std::string>{"Main"} : m.zone_names;
    int64_t n_nodes = 0;
    for (int64_t id : m.node_ids) {
        if (id < 0) throw MeshReadError("negative node id");
        n_nodes = std::max(n_nodes, id + 1);
    }
    // distinct coordinate copies of each node, ascending (lexicographic <)
    using Coord = std::array<double, 3>;
    std::vector<std::set<Coord>> copies(static_cast<size_t>(n_nodes));
    auto coord_of = [&](int64_t e, int v) {
        Coord c{0.0, 0.0, 0.0};
        for (int d = 0; d < dim; ++d) c[d] = m.coords[(size_t(e) * nv + v) * dim + d];
        return c;
    };
    for (int64_t e = 0; e < ne; ++e)
        for (int v = 0; v < nv; ++v) copies[size_t(m.node_ids[size_t(e) * nv + v])].insert(coord_of(e, v));
    for (const auto& c : copies)
        if (c.empty()) throw MeshReadError("node ids must number the nodes 0 .. n-1 without gaps");
    // entities by first appearance
    const LocalEntities loc = local_entities(m.shape);
    EntityTable faces, edges;
    for (int64_t e = 0; e < ne; ++e) {
        const int64_t* nid = m.node_ids.data() + size_t(e) * nv;
        for (int codim = 1; codim <= 2 && codim < dim; ++codim) {
            const auto& list = codim == 1 ? loc.faces : loc.edges;
            EntityTable& table = codim == 1 ? faces : edges;
            for (size_t i = 0; i < list.size(); ++i) {
                std::vector<int64_t> key;
                for (int v : list[i]) key.push_back(nid[v]);
                table.add(std::move(key), e, int(i));
            }
        }
    }
    std::ostringstream out;
    out << "mesh 2\n" << n_nodes << " " << ne << " " << dim;
    if (dim > 1) out << " " << faces.uses.size();
    if (dim > 2) out << " " << edges.uses.size();
    out << "\n1 " << n_nodes << " \n";
    out << "zones\n" << zone_names.size() << "\n";
    for (const auto& z : zone_names) out << z << "\n";
    out << "nodes\n";
    for (const auto& c : copies) {
        out << "1 0 \n" << c.size() << " ";
        for (const Coord& x : c)
            for (int d = 0; d < dim; ++d) out << hexfloat(x[d]) << " ";
        out << "\n";
    }
    out << "elements\n";
    for (int64_t e = 0; e < ne; ++e) {
        out << nv << " ";
        for (int v = 0; v < nv; ++v) {
            const int64_t node = m.node_ids[size_t(e) * nv + v];
            const auto& c = copies[size_t(node)];
            out << node << " " << std::distance(c.begin(), c.find(coord_of(e, v))) << " ";
        }
        out << "0 0 \n" << (m.zones.empty() ? 0 : m.zones[size_t(e)]) << "\n";
    }
    for (int codim = 1; codim <= 2 && codim < dim; ++codim) {
        out << "codim" << codim << "\n";
        for (const auto& u : (codim == 1 ? faces : edges).uses) {
            out << u.size() << " ";
            for (const auto& el : u) out << el.first << " " << el.second << " ";
            out << "1 0 \n";
        }
    }
    return out.str();
}

MeshFile read_mesh_file(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw MeshReadError("cannot open mesh file '" + path + "'");
    std::ostringstream buf;
    buf << in.rdbuf();
    Cursor c(path);
    c.text = buf.str();

    MeshFile m;
    // Header: 'mesh <version>' (MeshFileInformation.cpp:50-61).
    const std::string first = c.line();
    if (first.compare(0, 4, "mesh") != 0) c.fail("not an hpGEM mesh file (header must start with 'mesh')");
    {
        std::istringstream v(first.substr(4));
        if (!(v >> m.version) || m.version < 1 || m.version > 2)
            c.fail("mesh format version must be 1 or 2");
    }
    // Entity counts: nodes, elements, dimension, then codim 1 .. dim-1 (:64-75).
    m.n_nodes = c.integer("node count");
    m.n_elements = c.integer("element count");
    m.dim = int(c.integer("dimension"));
    if (m.dim < 1 || m.dim > 3) c.fail("dimension must be 1, 2 or 3");
    std::vector<int64_t> entity(size_t(m.dim) + 1, 0);
    for (int i = m.dim - 1; i > 0; --i) entity[size_t(i)] = c.integer("entity count");
    m.n_faces = m.dim >= 2 ? entity[size_t(m.dim - 1)] : 0;
    m.n_edges = m.dim >= 3 ? entity[1] : 0;
    // Partitions (:77-85); the rest of the line is padding.
    const int64_t partitions = c.integer("partition count");
    if (partitions != 1) c.fail("mesh has " + std::to_string(partitions) + " partitions; one is supported");
    const int64_t local_nodes = c.integer("partition node count");
    if (local_nodes != m.n_nodes) c.fail("partition node count differs from the node count");
    c.line();
    // Zones (:87-103); version 1 has the single zone "Main".
    if (m.version == 1) {
        m.zone_names.push_back("Main");
    } else {
        if (c.line() != "zones") c.fail("expected the 'zones' line");
        const int64_t nz = c.integer("zone count");
        c.line();
        for (int64_t z = 0; z < nz; ++z) {
            if (c.pos >= c.text.size()) c.fail("missing zone name");
            m.zone_names.push_back(c.line());
        }
    }
    if (m.n_elements > 0 && m.zone_names.empty()) c.fail("elements but no zones");

    // Nodes (MeshManipulator_Impl.h:1577-1626).
    if (m.version > 1) c.header("nodes");
    std::vector<int64_t> first_coord(size_t(m.n_nodes) + 1, 0);
    std::vector<double> coords;
    for (int64_t i = 0; i < m.n_nodes; ++i) {
        c.single_partition("node");
        const int64_t nc = c.integer("coordinate count");
        if (nc < 1) c.fail("node without coordinates");
        for (int64_t k = 0; k < nc * m.dim; ++k) coords.push_back(c.real("coordinate"));
        first_coord[size_t(i) + 1] = first_coord[size_t(i)] + nc;
    }

    // Elements (:1631-1700).
    if (m.version > 1) c.header("elements");
    m.node_ids.reserve(size_t(m.n_elements) * 8);
    for (int64_t e = 0; e < m.n_elements; ++e) {
        const int nv = int(c.integer("nodes per element"));
        const int shape = shape_of(m.dim, nv);
        if (shape == 0) c.fail("element with " + std::to_string(nv) + " nodes in dimension " + std::to_string(m.dim));
        if (e == 0) {
            m.shape = shape;
            m.n_vertices = nv;
        } else if (shape != m.shape) {
            c.fail("mixed element types are not supported");
        }
        for (int j = 0; j < nv; ++j) {
            const int64_t node = c.integer("node id");
            const int64_t copy = c.integer("coordinate index");
            if (node >= m.n_nodes) c.fail("node id out of range");
            if (first_coord[size_t(node)] + copy >= first_coord[size_t(node) + 1]) c.fail("coordinate index out of range");
            m.node_ids.push_back(node);
            const size_t at = size_t(first_coord[size_t(node)] + copy) * size_t(m.dim);
            m.coords.insert(m.coords.end(), coords.begin() + long(at), coords.begin() + long(at + size_t(m.dim)));
        }
        if (c.integer("owning partition") != 0) c.fail("element owned by a partition other than 0");
        const int64_t shadows = c.integer("shadow partition count");
        for (int64_t j = 0; j < shadows; ++j) c.integer("shadow partition");
        const int64_t zone = m.version == 1 ? 0 : c.integer("zone index");
        if (zone >= int64_t(m.zone_names.size())) c.fail("zone index out of range");
        m.zones.push_back(int32_t(zone));
    }

    // Faces (codim 1) (:1708-1770): one element on the boundary, two inside.
    const int n_local = faces_of(m.shape);
    if (m.dim > 1) {
        if (m.version > 1) c.header("codim1");
        m.face_elements.reserve(size_t(m.n_faces) * 2);
        for (int64_t f = 0; f < m.n_faces; ++f) {
            const int64_t k = c.integer("elements of a face");
            if (k < 1 || k > 2) c.fail("a face must have 1 or 2 elements");
            for (int64_t j = 0; j < 2; ++j) {
                int64_t el = -1, lf = -1;
                if (j < k) {
                    el = c.integer("element index");
                    lf = c.integer("local face index");
                    if (el >= m.n_elements || lf >= n_local) c.fail("face refers to a missing element or face");
                }
                m.face_elements.push_back(el);
                m.face_local.push_back(int32_t(lf));
            }
            c.single_partition("face");
            if (k == 1) ++m.n_boundary_faces;
        }
    }
    // Edges (codim 2) are read for validation and not kept (:1775-1800).
    if (m.dim > 2) {
        if (m.version > 1) c.header("codim2");
        for (int64_t i = 0; i < m.n_edges; ++i) {
            const int64_t k = c.integer("elements of an edge");
            for (int64_t j = 0; j < k; ++j) {
                if (c.integer("element index") >= m.n_elements) c.fail("edge refers to a missing element");
                c.integer("local edge index");
            }
            c.single_partition("edge");
        }
    }
    return m;
}

}  // namespace hgdg
