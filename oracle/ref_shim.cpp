// Exposes the reference's self-contained pieces through a C ABI for pinning the oracle:
// src/core/pointwise.hpp (hash RNG, update coefficients, fftfreq), src/core/layout.hpp
// (fftshift rolls) and src/core/recipe.cpp (expansion, iteration count, validation).
// Compiled from the reference tree by `make ref` into _ref/libref.so; test-only.
#include <cstdint>
#include <stdexcept>

#include "core/layout.hpp"
#include "core/pointwise.hpp"
#include "qoraphasecu/recipe.hpp"

extern "C" {

std::uint64_t ref_mix64(std::uint64_t z) { return qpcu::detail::mix64(z); }

float ref_uniform(std::uint64_t seed, std::uint64_t start, std::uint64_t index) {
  return qpcu::detail::uniform(seed, start, index);
}

float ref_fftfreq(std::uint32_t k, std::uint32_t n) { return qpcu::detail::fftfreq(k, n); }

// kind 0 ER, 1 HIO, 2 RAAR
void ref_coefficients(int kind, float beta, float* a, float* b) {
  const auto c = kind == 0 ? qpcu::detail::er_coefficients()
                 : kind == 1 ? qpcu::detail::hio_coefficients(beta)
                             : qpcu::detail::raar_coefficients(beta);
  *a = c.a;
  *b = c.b;
}

// roll a uint8 array (any payload can be encoded as indices)
void ref_roll_u32(const std::uint32_t* in, std::uint32_t* out, int ndim, const std::size_t* shape,
                  int to_fft_order) {
  qpcu::Shape s(shape, shape + ndim);
  std::size_t n = 1;
  for (int a = 0; a < ndim; ++a) n *= shape[a];
  const auto r = to_fft_order ? qpcu::detail::to_fft_order<std::uint32_t>(std::span(in, n), s)
                              : qpcu::detail::to_centered<std::uint32_t>(std::span(in, n), s);
  for (std::size_t i = 0; i < n; ++i) out[i] = r[i];
}

int ref_default_recipe_iterations() { return qpcu::default_recipe().iterations(); }

// kinds of the default recipe's flat op list (0 ER 1 HIO 2 RAAR 3 SW 4 DB 5 DE); returns count
int ref_default_recipe_kinds(int* kinds, int cap) {
  const auto r = qpcu::default_recipe();
  int i = 0;
  for (const auto& op : r.ops) {
    if (i < cap) kinds[i] = static_cast<int>(op.index());
    ++i;
  }
  return i;
}

// 1 when validate() throws std::invalid_argument for the recipe of `case_id`
int ref_validate_rejects(int case_id) {
  using namespace qpcu;
  try {
    switch (case_id) {
      case 0: validate(Recipe{}); break;
      case 1: validate(recipe(ER{0})); break;
      case 2: validate(recipe(DetwinBegin{}, ER{1})); break;
      case 3: validate(recipe(DetwinEnd{})); break;
      case 4: validate(recipe(detwin(detwin(ER{1})))); break;
      case 5: validate(recipe(Shrinkwrap{.threshold = 0.0f})); break;
      case 6: validate(default_recipe()); break;
      default: return -1;
    }
  } catch (const std::invalid_argument&) {
    return 1;
  }
  return 0;
}
}
