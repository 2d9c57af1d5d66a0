// Checker harness built against hpGEM's own table layer (FE bases, reference
// geometries, face mappings, Gauss rules) compiled from the reference source
// tree. It prints those tables as JSON so the oracle's restatement can be
// pinned against what the reference itself computes. Test infrastructure only:
// nothing in the product links or loads it.
//
// Reference entry points exercised:
//   kernel/FE/BasisFunctions2DH1ConformingSquare.cpp:164  createDGBasisFunctionSet2DH1Square
//   kernel/FE/BasisFunctions3DH1ConformingCube.cpp:434    createDGBasisFunctionSet3DH1Cube
//   kernel/FE/BasisFunctions2DH1ConformingTriangle.cpp:203 createDGBasisFunctionSet2DH1Triangle
//   kernel/FE/BasisFunctions3DH1ConformingTetrahedron.cpp:366 createDGBasisFunctionSet3DH1Tetrahedron
//   kernel/Integration/QuadratureRules/AllGaussQuadratureRules.cpp:139 getRule
//   kernel/Geometry/Reference{Square,Cube,Triangle,Tetrahedron,Line}.cpp (nodes,
//     face node lists, codim-1 face maps, codim-0 orientation maps and indices)
#include "FE/BasisFunctionSet.h"
#include "FE/BasisFunctions2DH1ConformingSquare.h"
#include "FE/BasisFunctions2DH1ConformingTriangle.h"
#include "FE/BasisFunctions3DH1ConformingCube.h"
#include "FE/BasisFunctions3DH1ConformingTetrahedron.h"
#include "FE/BasisFunctions3DAinsworthCoyle.h"
#include "Geometry/PointReference.h"
#include "Geometry/ReferenceCube.h"
#include "Geometry/ReferenceLine.h"
#include "Geometry/ReferenceSquare.h"
#include "Geometry/ReferenceTetrahedron.h"
#include "Geometry/ReferenceTriangle.h"
#include "Geometry/Mappings/MappingReferenceToReference.h"
#include "Integration/QuadratureRules/AllGaussQuadratureRules.h"
#include "Integration/QuadratureRules/GaussQuadratureRule.h"

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

using namespace hpgem;

namespace {

bool first_item = true;

void sep() {
    if (!first_item) std::printf(",\n");
    first_item = false;
}

void print_vec(const std::vector<double>& v) {
    std::printf("[");
    for (std::size_t i = 0; i < v.size(); ++i)
        std::printf("%s%.17g", i ? "," : "", v[i]);
    std::printf("]");
}

template <std::size_t D>
std::vector<double> coords(const Geometry::PointReference<D>& p) {
    std::vector<double> r(D);
    for (std::size_t i = 0; i < D; ++i) r[i] = p[i];
    return r;
}

// Deterministic sample points strictly inside each reference shape.
std::vector<std::vector<double>> sample_points(int dim, bool simplex) {
    std::vector<std::vector<double>> pts;
    std::uint64_t s = 0x9E3779B97F4A7C15ull;
    auto next = [&]() {
        s ^= s << 13;
        s ^= s >> 7;
        s ^= s << 17;
        return double(s >> 11) * (1.0 / 9007199254740992.0);
    };
    while (pts.size() < 6) {
        std::vector<double> p(dim);
        double sum = 0;
        for (int d = 0; d < dim; ++d) {
            p[d] = simplex ? next() : 2.0 * next() - 1.0;
            sum += p[d];
        }
        if (simplex && sum >= 1.0) continue;
        pts.push_back(p);
    }
    return pts;
}

template <std::size_t D>
void dump_rules(const char* shape, const Geometry::ReferenceGeometry* geo,
                std::size_t max_order) {
    for (std::size_t q = 0; q <= max_order; ++q) {
        auto* rule = QuadratureRules::AllGaussQuadratureRules::instance()
                         .getRule(geo, q);
        sep();
        std::printf("{\"kind\":\"rule\",\"shape\":\"%s\",\"requested\":%zu,"
                    "\"name\":\"%s\",\"order\":%zu,\"points\":[",
                    shape, q, rule->getName().c_str(), rule->order());
        for (std::size_t i = 0; i < rule->getNumberOfPoints(); ++i) {
            const auto& p = static_cast<const Geometry::PointReference<D>&>(
                rule->getPoint(i));
            if (i) std::printf(",");
            print_vec(coords<D>(p));
        }
        std::printf("],\"weights\":[");
        for (std::size_t i = 0; i < rule->getNumberOfPoints(); ++i)
            std::printf("%s%.17g", i ? "," : "", rule->weight(i));
        std::printf("]}");
    }
}

template <std::size_t D>
void dump_basis(const char* shape, FE::BasisFunctionSet* (*make)(std::size_t),
                std::size_t p, bool simplex) {
    FE::BasisFunctionSet* set = make(p);
    auto pts = sample_points(int(D), simplex);
    sep();
    std::printf("{\"kind\":\"basis\",\"shape\":\"%s\",\"p\":%zu,\"n\":%zu,"
                "\"points\":[",
                shape, p, set->size());
    for (std::size_t k = 0; k < pts.size(); ++k) {
        if (k) std::printf(",");
        print_vec(pts[k]);
    }
    std::printf("],\"values\":[");
    for (std::size_t k = 0; k < pts.size(); ++k) {
        Geometry::PointReference<D> pt;
        for (std::size_t d = 0; d < D; ++d) pt[d] = pts[k][d];
        std::vector<double> v(set->size());
        for (std::size_t i = 0; i < set->size(); ++i) v[i] = set->eval(i, pt);
        if (k) std::printf(",");
        print_vec(v);
    }
    std::printf("],\"grads\":[");
    for (std::size_t k = 0; k < pts.size(); ++k) {
        Geometry::PointReference<D> pt;
        for (std::size_t d = 0; d < D; ++d) pt[d] = pts[k][d];
        std::vector<double> g;
        for (std::size_t i = 0; i < set->size(); ++i)
            for (std::size_t d = 0; d < D; ++d)
                g.push_back(set->evalDeriv(i, d, pt));
        if (k) std::printf(",");
        print_vec(g);
    }
    std::printf("]}");
    delete set;
}

// Vector-valued H(curl) set on the reference tetrahedron: values and curls
// at the sample points.
void dump_hcurl_basis(const char* name, FE::BasisFunctionSet* (*make)(std::size_t), std::size_t p) {
    FE::BasisFunctionSet* set = make(p);
    auto pts = sample_points(3, true);
    sep();
    std::printf("{\"kind\":\"hcurl_basis\",\"set\":\"%s\",\"p\":%zu,\"n\":%zu,\"points\":[", name, p,
                set->size());
    for (std::size_t k = 0; k < pts.size(); ++k) {
        if (k) std::printf(",");
        print_vec(pts[k]);
    }
    std::printf("],\"values\":[");
    for (std::size_t k = 0; k < pts.size(); ++k) {
        Geometry::PointReference<3> pt;
        for (std::size_t d = 0; d < 3; ++d) pt[d] = pts[k][d];
        std::vector<double> v;
        for (std::size_t i = 0; i < set->size(); ++i) {
            LinearAlgebra::SmallVector<3> r;
            set->eval(i, pt, r);
            for (std::size_t d = 0; d < 3; ++d) v.push_back(r[d]);
        }
        if (k) std::printf(",");
        print_vec(v);
    }
    std::printf("],\"curls\":[");
    for (std::size_t k = 0; k < pts.size(); ++k) {
        Geometry::PointReference<3> pt;
        for (std::size_t d = 0; d < 3; ++d) pt[d] = pts[k][d];
        std::vector<double> v;
        for (std::size_t i = 0; i < set->size(); ++i) {
            LinearAlgebra::SmallVector<3> r = set->evalCurl(i, pt);
            for (std::size_t d = 0; d < 3; ++d) v.push_back(r[d]);
        }
        if (k) std::printf(",");
        print_vec(v);
    }
    std::printf("]}");
    delete set;
}

template <std::size_t D>
void dump_geometry(const char* shape, const Geometry::ReferenceGeometry& geo,
                   bool simplex) {
    // nodes
    sep();
    std::printf("{\"kind\":\"nodes\",\"shape\":\"%s\",\"nodes\":[", shape);
    for (std::size_t i = 0; i < geo.getNumberOfNodes(); ++i) {
        const auto& p = static_cast<const Geometry::PointReference<D>&>(
            geo.getReferenceNodeCoordinate(i));
        if (i) std::printf(",");
        print_vec(coords<D>(p));
    }
    std::printf("]}");
    // faces: node lists, map of sample face points, outward sign
    auto face_pts = sample_points(int(D) - 1, simplex);
    if (D == 1) face_pts = {std::vector<double>{}};
    for (std::size_t f = 0; f < geo.getNumberOfCodim1Entities(); ++f) {
        auto nodes = geo.getCodim1EntityLocalIndices(f);
        const Geometry::BoundaryFaceMapping* map = geo.getCodim1MappingPtr(f);
        sep();
        std::printf("{\"kind\":\"face\",\"shape\":\"%s\",\"face\":%zu,"
                    "\"nodes\":[",
                    shape, f);
        for (std::size_t i = 0; i < nodes.size(); ++i)
            std::printf("%s%zu", i ? "," : "", nodes[i]);
        std::printf("],\"normal_sign\":%.17g,\"points_in\":[",
                    map->outwardNormalSign());
        for (std::size_t k = 0; k < face_pts.size(); ++k) {
            if (k) std::printf(",");
            print_vec(face_pts[k]);
        }
        std::printf("],\"points_out\":[");
        for (std::size_t k = 0; k < face_pts.size(); ++k) {
            Geometry::PointReference<D - 1> pin;
            for (std::size_t d = 0; d + 1 < D; ++d) pin[d] = face_pts[k][d];
            Geometry::PointReference<D> pout = map->transform(pin);
            if (k) std::printf(",");
            print_vec(coords<D>(pout));
        }
        std::printf("]}");
    }
}

// Codim-0 maps of a face shape (the face-to-face orientation maps) and the
// orientation index chosen for every consistent pair of node lists.
template <std::size_t D>
void dump_orientations(const char* shape, const Geometry::ReferenceGeometry& geo,
                       bool simplex, std::size_t n_maps) {
    auto pts = sample_points(int(D), simplex);
    for (std::size_t m = 0; m < n_maps; ++m) {
        const Geometry::MappingReferenceToReference<0>* map =
            geo.getCodim0MappingPtr(m);
        sep();
        std::printf("{\"kind\":\"orientation_map\",\"shape\":\"%s\","
                    "\"index\":%zu,\"points_in\":[",
                    shape, m);
        for (std::size_t k = 0; k < pts.size(); ++k) {
            if (k) std::printf(",");
            print_vec(pts[k]);
        }
        std::printf("],\"points_out\":[");
        for (std::size_t k = 0; k < pts.size(); ++k) {
            Geometry::PointReference<D> pin;
            for (std::size_t d = 0; d < D; ++d) pin[d] = pts[k][d];
            Geometry::PointReference<D> pout = map->transform(pin);
            if (k) std::printf(",");
            print_vec(coords<D>(pout));
        }
        std::printf("]}");
    }
    // Every node relabelling that is a symmetry of the shape: list2 is the
    // image of list1 = (10, 11, ...) under a vertex permutation that maps
    // the reference shape onto itself.
    std::size_t nn = geo.getNumberOfNodes();
    std::vector<std::size_t> perm(nn);
    for (std::size_t i = 0; i < nn; ++i) perm[i] = i;
    do {
        // keep permutations that preserve the face-adjacency structure:
        // for squares, neighbours along edges stay neighbours.
        bool ok = true;
        if (!simplex && nn == 4) {
            // square node i at ((i%2),(i/2)); an isometry keeps Hamming distance
            for (std::size_t a = 0; a < 4 && ok; ++a)
                for (std::size_t b = 0; b < 4 && ok; ++b) {
                    int da = __builtin_popcount(unsigned(a ^ b));
                    int db = __builtin_popcount(unsigned(perm[a] ^ perm[b]));
                    if (da != db) ok = false;
                }
        }
        if (!ok) continue;
        std::vector<std::size_t> l1(nn), l2(nn);
        for (std::size_t i = 0; i < nn; ++i) {
            l1[i] = 10 + i;
            l2[perm[i]] = 10 + i;
        }
        std::size_t idx = geo.getCodim0MappingIndex(l1, l2);
        sep();
        std::printf("{\"kind\":\"orientation_index\",\"shape\":\"%s\","
                    "\"left\":[",
                    shape);
        for (std::size_t i = 0; i < nn; ++i)
            std::printf("%s%zu", i ? "," : "", l1[i]);
        std::printf("],\"right\":[");
        for (std::size_t i = 0; i < nn; ++i)
            std::printf("%s%zu", i ? "," : "", l2[i]);
        std::printf("],\"index\":%zu}", idx);
    } while (std::next_permutation(perm.begin(), perm.end()));
}

}  // namespace

int main() {
    std::printf("{\"tables\":[\n");
    dump_rules<1>("line", &Geometry::ReferenceLine::Instance(), 11);
    dump_rules<2>("square", &Geometry::ReferenceSquare::Instance(), 11);
    dump_rules<2>("triangle", &Geometry::ReferenceTriangle::Instance(), 11);
    dump_rules<3>("cube", &Geometry::ReferenceCube::Instance(), 11);
    dump_rules<3>("tetrahedron", &Geometry::ReferenceTetrahedron::Instance(),
                  10);
    for (std::size_t p = 1; p <= 4; ++p) {
        dump_basis<2>("square", &FE::createDGBasisFunctionSet2DH1Square, p,
                      false);
        dump_basis<3>("cube", &FE::createDGBasisFunctionSet3DH1Cube, p, false);
        dump_basis<2>("triangle", &FE::createDGBasisFunctionSet2DH1Triangle, p,
                      true);
        dump_basis<3>("tetrahedron",
                      &FE::createDGBasisFunctionSet3DH1Tetrahedron, p, true);
    }
    for (std::size_t p = 1; p <= 3; ++p)
        dump_hcurl_basis("ainsworth_coyle", &FE::createDGBasisFunctionSet3DAinsworthCoyle, p);
    dump_geometry<2>("square", Geometry::ReferenceSquare::Instance(), false);
    dump_geometry<3>("cube", Geometry::ReferenceCube::Instance(), false);
    dump_geometry<2>("triangle", Geometry::ReferenceTriangle::Instance(), true);
    dump_geometry<3>("tetrahedron", Geometry::ReferenceTetrahedron::Instance(),
                     true);
    dump_orientations<1>("line", Geometry::ReferenceLine::Instance(), false, 2);
    dump_orientations<2>("square", Geometry::ReferenceSquare::Instance(), false,
                         8);
    dump_orientations<2>("triangle", Geometry::ReferenceTriangle::Instance(),
                         true, 6);
    std::printf("\n]}\n");
    return 0;
}
