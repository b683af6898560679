// What the C ABI's opaque handles hold (include/fe_field.h).
#pragma once

#include "field/Law.hpp"
#include "field/ir/LawIR.hpp"

struct fe_law {
    fieldengine::field::Law law;
    fieldengine::field::ir::Module module;
};
