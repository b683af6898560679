// Interval bounds of law expressions, for skipping idle terms (the reference's interface, src/field/LawBounds.hpp).
//
// A term whose every equation's rate is exactly zero over the whole lattice, given what is known of the fields'
// ranges, is not dispatched: the reference's CPU engine skips it (src/field/Solver.cpp:341-347) and so does its CUDA
// backend (src/field/cuda/CudaSolver.cpp:586-595). The same decisions need the same interval rules, which these
// functions restate: bounds_of (the range a rate can take), always_zero (provably zero from fields known to be zero)
// and flow_bounds (what explicit substeps of a term may do to the ranges of the fields it writes).
#pragma once

#include "field/Law.hpp"
#include "field/LawNode.hpp"

#include <limits>
#include <vector>

namespace fieldengine::field {

struct Bounds {
    double lo = -std::numeric_limits<double>::infinity();
    double hi = std::numeric_limits<double>::infinity();
    [[nodiscard]] bool none() const { return lo == 0.0 && hi == 0.0; }
};

// a * b, where a zero factor wins over an infinite one.
double times(double a, double b);

// Zero at every cell whatever the other fields hold, given which fields are zero at every cell.
bool always_zero(const Node& n, const std::vector<char>& zero, const std::vector<double>& params);

// The range of a node's values over the lattice, given each field's range (unbounded where unknown).
Bounds bounds_of(const Node& n, const std::vector<Bounds>& fields, const std::vector<double>& params, double h);

// The ranges after `substeps` explicit Euler substeps of dt / substeps of `equations`, carried substep by substep;
// never narrower than the cells can be. Clipped to the declared range only for a clamped field.
void flow_bounds(const std::vector<Equation>& equations, double dt, int substeps, const Law& law, double h, std::vector<Bounds>& flowing);

} // namespace fieldengine::field
