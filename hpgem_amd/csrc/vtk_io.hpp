// VTK output of a DG state, host only.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace hgdg {

// One piece of hpGEM's VTK output (order 1: standard cells, one cell per
// element, the solution evaluated at the element's own vertices, so the
// field is discontinuous as the DG solution is). `state` is in hpGEM's
// layout: element after element, field-major inside an element, nb(p_e)
// coefficients per field in hpGEM's basis order. Writes <base>.pvtu and
// <base>.0.vtu; returns nothing, throws std::runtime_error on failure.
void write_vtk(const std::string& base, int shape, int64_t n_elements, const double* coords, const int32_t* orders,
               int n_fields, const double* state, const std::vector<std::string>& names);

// Base64 of a byte range (RFC 4648 alphabet, '=' padding).
std::string to_base64(const void* data, size_t len);

}  // namespace hgdg
