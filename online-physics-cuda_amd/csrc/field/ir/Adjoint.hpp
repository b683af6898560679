// The adjoint of a term, as a law: reverse mode over the term's SSA graph, written back as law text that this
// compiler and backend run like any other law (the reference's `lawc --adjoint`, src/field/ir/Adjoint.cpp:27-225).
//
// One step of the adjoint law is one vector-Jacobian product of one step of the term: for the explicit step
// F' = F + dt R(F, p), every field F the term touches gets F__adj += dt (dR/dF)^T A, and every param p the term
// reads gets a field p__adj += dt (dR/dp)^T A, whose total over the lattice is dL/dp. The primal fields are static
// inputs holding the state the primal step began from. Run backwards over the primal's states, seeded with dL/dF at
// the end, the adjoint fields finish as dL/dF at the start.
#pragma once

#include "field/Law.hpp"

#include <string>

namespace fieldengine::field::ir {

// The adjoint law text of `term` (empty: the first), or "" with `error` (an op without an adjoint here, a
// diffusivity that is a field, no such term).
std::string adjoint_law(const Law& law, const std::string& term, std::string& error);

} // namespace fieldengine::field::ir
