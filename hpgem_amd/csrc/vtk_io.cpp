// VTK output in the layout of hpGEM's VTKSpecificTimeWriter
// (kernel/Output/VTKSpecificTimeWriter_Impl.h) at order 1: a .pvtu master
// file naming one .0.vtu piece; the piece holds the points (each element's
// vertices, in VTK's vertex order), the cells (connectivity, offsets, types)
// and one Float64 point array per field, every array written in VTK's
// "binary" form: base64 of a uint32 byte count, then base64 of the data.
#include "vtk_io.hpp"

#include <fstream>
#include <sstream>
#include <stdexcept>

#include "host_tables.hpp"
#include "simplex_tables.hpp"

namespace hgdg {

std::string to_base64(const void* data, size_t len) {
    static const char abc[] = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
    const unsigned char* b = static_cast<const unsigned char*>(data);
    std::string out;
    out.reserve((len + 2) / 3 * 4);
    size_t i = 0;
    for (; i + 2 < len; i += 3) {
        const unsigned v = (unsigned(b[i]) << 16) | (unsigned(b[i + 1]) << 8) | b[i + 2];
        out += abc[(v >> 18) & 63];
        out += abc[(v >> 12) & 63];
        out += abc[(v >> 6) & 63];
        out += abc[v & 63];
    }
    if (len - i == 2) {
        const unsigned v = (unsigned(b[i]) << 16) | (unsigned(b[i + 1]) << 8);
        out += abc[(v >> 18) & 63];
        out += abc[(v >> 12) & 63];
        out += abc[(v >> 6) & 63];
        out += '=';
    } else if (len - i == 1) {
        const unsigned v = unsigned(b[i]) << 16;
        out += abc[(v >> 18) & 63];
        out += abc[(v >> 12) & 63];
        out += "==";
    }
    return out;
}

namespace {

template <typename T>
std::string binary_array(const std::vector<T>& data) {
    const uint32_t bytes = uint32_t(data.size() * sizeof(T));
    return to_base64(&bytes, sizeof bytes) + to_base64(data.data(), bytes);
}

// VTK cell of each geometry at order 1: type id and, per VTK vertex, the
// element vertex (hpGEM numbering) it is. VTK numbers the quadrilateral and
// hexahedron vertices around each face (hpGEM's 2/3 and 6/7 swapped).
struct VtkCell {
    uint8_t type;
    std::vector<int> vertex;
};

VtkCell vtk_cell(int shape) {
    switch (shape) {
        case SHAPE_SQUARE: return {9, {0, 1, 3, 2}};
        case SHAPE_CUBE: return {12, {0, 1, 3, 2, 4, 5, 7, 6}};
        case SHAPE_TETRAHEDRON: return {10, {0, 1, 2, 3}};
        default: throw std::runtime_error("VTK output is available for squares, cubes and tetrahedra");
    }
}

}  // namespace

void write_vtk(const std::string& base, int shape, int64_t n_elements, const double* coords, const int32_t* orders,
               int n_fields, const double* state, const std::vector<std::string>& names) {
    const VtkCell cell = vtk_cell(shape);
    const int dim = shape_dim(shape), nv = int(cell.vertex.size());
    if (int(names.size()) != n_fields) throw std::runtime_error("one name per field");
    const int64_t n_points = n_elements * nv;
    std::vector<double> points(static_cast<size_t>(n_points) * 3, 0.0);
    std::vector<uint32_t> connectivity(static_cast<size_t>(n_points)), offsets(static_cast<size_t>(n_elements));
    std::vector<uint8_t> types(static_cast<size_t>(n_elements), cell.type);
    std::vector<std::vector<double>> values(static_cast<size_t>(n_fields), std::vector<double>(static_cast<size_t>(n_points)));
    int64_t off = 0;  // first coefficient of the element in `state`
    for (int64_t e = 0; e < n_elements; ++e) {
        const int p = orders[e];
        int nb;
        if (shape == SHAPE_TETRAHEDRON) {
            if (p < 1 || p > kMaxSimplexOrder) throw std::runtime_error("tetrahedral orders must be in 1..4");
            nb = tet_nb(p);
        } else {
            if (p < 1 || p > 5) throw std::runtime_error("orders must be in 1..5");
            nb = 1;
            for (int d = 0; d < dim; ++d) nb *= p + 1;
        }
        std::vector<int> perm;
        if (shape != SHAPE_TETRAHEDRON) perm = tensor_permutation(dim, p);
        std::vector<double> phi(static_cast<size_t>(nb));
        for (int k = 0; k < nv; ++k) {
            const int v = cell.vertex[size_t(k)];
            const int64_t pt = e * nv + k;
            connectivity[size_t(pt)] = uint32_t(pt);
            for (int d = 0; d < dim; ++d) points[size_t(pt) * 3 + d] = coords[(size_t(e) * nv + v) * dim + d];
            // basis functions at the reference vertex v
            if (shape == SHAPE_TETRAHEDRON) {
                long double x[3] = {0, 0, 0};
                if (v > 0) x[v - 1] = 1;
                std::vector<long double> val(static_cast<size_t>(nb)), grad(static_cast<size_t>(nb) * 3);
                tet_basis(p, x, val.data(), grad.data());
                for (int i = 0; i < nb; ++i) phi[size_t(i)] = double(val[size_t(i)]);
            } else {
                double xi[3];
                for (int d = 0; d < dim; ++d) xi[d] = (v >> d) & 1 ? 1.0 : -1.0;
                for (int i = 0; i < nb; ++i) {
                    int t = perm[size_t(i)];
                    double val = 1;
                    for (int d = 0; d < dim; ++d) {
                        val *= psi(t % (p + 1), xi[d]);
                        t /= p + 1;
                    }
                    phi[size_t(i)] = val;
                }
            }
            for (int f = 0; f < n_fields; ++f) {
                double s = 0;
                for (int i = 0; i < nb; ++i) s += state[off + int64_t(f) * nb + i] * phi[size_t(i)];
                values[size_t(f)][size_t(pt)] = s;
            }
        }
        offsets[size_t(e)] = uint32_t((e + 1) * nv);
        off += int64_t(n_fields) * nb;
    }
    if (n_points > int64_t(UINT32_MAX)) throw std::runtime_error("too many points for UInt32 connectivity");

    std::string file = base;
    if (file.find('/') != std::string::npos) file = file.substr(file.find_last_of('/') + 1);
    std::ostringstream master;
    master << "<?xml version=\"1.0\"?>\n"
           << "<VTKFile type=\"PUnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n"
           << "  <PUnstructuredGrid GhostLevel=\"0\">\n"
           << "    <Piece Source=\"" << file << ".0.vtu\"/>\n"
           << "    <PPointData>\n";
    for (const auto& n : names) master << "      <PDataArray type=\"Float64\" Name=\"" << n << "\"/>\n";
    master << "    </PPointData>\n"
           << "    <PPoints>\n"
           << "      <PDataArray type=\"Float64\" NumberOfComponents=\"3\"/>\n"
           << "    </PPoints>\n"
           << "  </PUnstructuredGrid>\n"
           << "</VTKFile>\n";
    std::ostringstream piece;
    piece << "<?xml version=\"1.0\"?>\n"
          << "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n"
          << "  <UnstructuredGrid>\n"
          << "    <Piece NumberOfPoints=\"" << n_points << "\" NumberOfCells=\"" << n_elements << "\">\n"
          << "      <Points>\n"
          << "        <DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"binary\">\n"
          << "          " << binary_array(points) << "\n"
          << "        </DataArray>\n"
          << "      </Points>\n"
          << "      <Cells>\n"
          << "        <DataArray type=\"UInt32\" Name=\"connectivity\" format=\"binary\">\n"
          << "          " << binary_array(connectivity) << "\n"
          << "        </DataArray>\n"
          << "        <DataArray type=\"UInt32\" Name=\"offsets\" format=\"binary\">\n"
          << "          " << binary_array(offsets) << "\n"
          << "        </DataArray>\n"
          << "        <DataArray type=\"UInt8\" Name=\"types\" format=\"binary\">\n"
          << "          " << binary_array(types) << "\n"
          << "        </DataArray>\n"
          << "      </Cells>\n"
          << "      <PointData>\n";
    for (int f = 0; f < n_fields; ++f)
        piece << "      <DataArray type=\"Float64\" Name=\"" << names[size_t(f)] << "\" format=\"binary\">\n"
              << "        " << binary_array(values[size_t(f)]) << "\n"
              << "      </DataArray>\n";
    piece << "      </PointData>\n"
          << "    </Piece>\n"
          << "  </UnstructuredGrid>\n"
          << "</VTKFile>\n";
    for (const auto& [path, text] : {std::make_pair(base + ".pvtu", master.str()),
                                     std::make_pair(base + ".0.vtu", piece.str())}) {
        std::ofstream out(path, std::ios::binary);
        if (!out) throw std::runtime_error("cannot open " + path);
        out << text;
        out.close();
        if (!out) throw std::runtime_error("cannot write " + path);
    }
}

}  // namespace hgdg
