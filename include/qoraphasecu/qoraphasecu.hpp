#pragma once

#include "qoraphasecu/metrics.hpp"
#include "qoraphasecu/operators.hpp"
#include "qoraphasecu/problem.hpp"
#include "qoraphasecu/recipe.hpp"
#include "qoraphasecu/solver.hpp"
