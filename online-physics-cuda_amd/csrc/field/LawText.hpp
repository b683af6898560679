// A Law written back as law text that compiles to the same expression trees.
//
// Used to emit generated laws (the adjoint of a term) and to check the front end (text -> law -> text -> law gives the
// same IR). A subtree used more than once within a term becomes a `let`, so a generated law with heavy sharing stays
// linear in size; parentheses appear only where precedence needs them.
#pragma once

#include "field/Law.hpp"

#include <string>

namespace fieldengine::field {

// The whole law: fields (in id order, with their contracts), params, nets (declarations; weights stay in their files),
// terms with lets, derives. `header` lines are written first as comments.
std::string law_text(const Law& law, const std::string& header = {});

// One expression, every shared subtree written out in full.
std::string expression_text(const Node& node, const Law& law);

} // namespace fieldengine::field
