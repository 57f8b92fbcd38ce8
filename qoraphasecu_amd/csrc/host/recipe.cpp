// Reference behaviour: src/core/recipe.cpp:16-61.
#include "qoraphasecu/recipe.hpp"

#include <stdexcept>
#include <type_traits>

namespace qpcu {

int Recipe::iterations() const {
  int total = 0;
  for (const Op& op : ops) {
    std::visit(
        [&](const auto& o) {
          using T = std::decay_t<decltype(o)>;
          if constexpr (std::is_same_v<T, ER> || std::is_same_v<T, HIO> || std::is_same_v<T, RAAR>) {
            total += o.n;
          }
        },
        op);
  }
  return total;
}

Recipe default_recipe() {
  return recipe(RAAR{50}, detwin(HIO{20}), RAAR{50}, repeat(8, RAAR{50}, Shrinkwrap{}), ER{200});
}

void validate(const Recipe& r) {
  if (r.ops.empty()) throw std::invalid_argument("recipe is empty");
  bool open = false;
  for (const Op& op : r.ops) {
    std::visit(
        [&](const auto& o) {
          using T = std::decay_t<decltype(o)>;
          if constexpr (std::is_same_v<T, ER>) {
            if (o.n < 1) throw std::invalid_argument("ER needs n >= 1");
          } else if constexpr (std::is_same_v<T, HIO>) {
            if (o.n < 1) throw std::invalid_argument("HIO needs n >= 1");
          } else if constexpr (std::is_same_v<T, RAAR>) {
            if (o.n < 1) throw std::invalid_argument("RAAR needs n >= 1");
          } else if constexpr (std::is_same_v<T, Shrinkwrap>) {
            if (!(o.threshold > 0.0f) || !(o.sigma >= 0.0f)) {
              throw std::invalid_argument("Shrinkwrap needs threshold > 0, sigma >= 0");
            }
          } else if constexpr (std::is_same_v<T, DetwinBegin>) {
            if (open) throw std::invalid_argument("nested detwin");
            open = true;
          } else {
            if (!open) throw std::invalid_argument("DetwinEnd without DetwinBegin");
            open = false;
          }
        },
        op);
  }
  if (open) throw std::invalid_argument("DetwinBegin without DetwinEnd");
}

}  // namespace qpcu
