// Recipes (reference: include/qoraphasecu/recipe.hpp:6-88, same API).
#pragma once

#include <variant>
#include <vector>

namespace qpcu {

// x <- P_S P_M x
struct ER {
  int n;
};

// x <- P_M x inside the support, x - beta P_M x outside.
struct HIO {
  int n;
  float beta = 0.9f;
};

// x <- (beta / 2)(R_S R_M x + x) + (1 - beta) P_M x
struct RAAR {
  int n;
  float beta = 0.9f;
};

enum class Threshold { max, rms };

// New support from the Gaussian-blurred magnitude, thresholded against its maximum
// (`max`) or its rms over the current support (`rms`).
struct Shrinkwrap {
  float threshold = 0.3f;
  float sigma = 1.0f;
  Threshold method = Threshold::max;
};

// Halves the support (half_support) until the matching DetwinEnd restores it.
struct DetwinBegin {
  int axis = 0;
};

struct DetwinEnd {};

using Op = std::variant<ER, HIO, RAAR, Shrinkwrap, DetwinBegin, DetwinEnd>;

// A flat operation list; repeat() and detwin() expand in place.
struct Recipe {
  std::vector<Op> ops;

  // Iterations that each produce one error value.
  int iterations() const;
};

namespace detail {
inline void append(Recipe& r, const Recipe& part) {
  r.ops.insert(r.ops.end(), part.ops.begin(), part.ops.end());
}
inline void append(Recipe& r, const Op& op) { r.ops.push_back(op); }
}  // namespace detail

template <class... Parts>
Recipe recipe(const Parts&... parts) {
  Recipe r;
  (detail::append(r, parts), ...);
  return r;
}

template <class... Parts>
Recipe repeat(int n, const Parts&... parts) {
  const Recipe body = recipe(parts...);
  Recipe r;
  for (int i = 0; i < n; ++i) detail::append(r, body);
  return r;
}

template <class... Parts>
Recipe detwin(const Parts&... parts) {
  return recipe(DetwinBegin{}, parts..., DetwinEnd{});
}

// RAAR(50), detwin(HIO(20)), RAAR(50), repeat(8, RAAR(50), Shrinkwrap()), ER(200).
Recipe default_recipe();

// std::invalid_argument on non-positive counts, bad Shrinkwrap parameters or an
// unbalanced detwin.
void validate(const Recipe& recipe);

}  // namespace qpcu
