// The law compiler behind include/fe_field.h's fe_law_* entry points.
#include "fe_field.h"

#include "field/Law.hpp"
#include "field/LawText.hpp"
#include "field/capi_internal.hpp"
#include "field/ir/LawIR.hpp"

#include <cstdlib>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

using namespace fieldengine::field;

namespace {

char* dup(const std::string& s)
{
    char* p = static_cast<char*>(std::malloc(s.size() + 1));
    if (p != nullptr) {
        std::memcpy(p, s.c_str(), s.size() + 1);
    }
    return p;
}

void put(char* buf, size_t bytes, const std::string& s)
{
    if (buf == nullptr || bytes == 0) {
        return;
    }
    const size_t n = s.size() < bytes - 1 ? s.size() : bytes - 1;
    std::memcpy(buf, s.data(), n);
    buf[n] = '\0';
}

} // namespace

extern "C" {

fe_law* fe_law_compile(const char* const* names, const char* const* texts, int count, char* errors, size_t error_bytes)
{
    std::vector<LawSource> sources;
    for (int i = 0; i < count; ++i) {
        sources.push_back(LawSource{names != nullptr && names[i] != nullptr ? names[i] : "law", texts[i] != nullptr ? texts[i] : ""});
    }
    CompileResult r = compile_law(sources);
    std::string all;
    for (const std::string& e : r.errors) {
        all += e + "\n";
    }
    put(errors, error_bytes, all);
    if (!r.law) {
        return nullptr;
    }
    auto* law = new fe_law{};
    law->law = std::move(*r.law);
    law->module = ir::lower(law->law);
    return law;
}

void fe_law_free(fe_law* law)
{
    delete law;
}

int fe_law_field_count(const fe_law* law)
{
    return law == nullptr ? 0 : static_cast<int>(law->law.fields.size());
}

int fe_law_field_id(const fe_law* law, const char* name)
{
    return law == nullptr || name == nullptr ? -1 : law->law.field_id(name);
}

const char* fe_law_field_name(const fe_law* law, int field)
{
    if (law == nullptr || field < 0 || field >= static_cast<int>(law->law.fields.size())) {
        return nullptr;
    }
    return law->law.fields[static_cast<std::size_t>(field)].name.c_str();
}

int fe_law_param_count(const fe_law* law)
{
    return law == nullptr ? 0 : static_cast<int>(law->law.param_names.size());
}

int fe_law_term_count(const fe_law* law)
{
    return law == nullptr ? 0 : static_cast<int>(law->law.terms.size());
}

char* fe_law_ir_json(const fe_law* law)
{
    return law == nullptr ? nullptr : dup(ir::to_json(law->module, law->law));
}

char* fe_law_describe(const fe_law* law)
{
    return law == nullptr ? nullptr : dup(ir::describe(law->module, law->law));
}

char* fe_law_text(const fe_law* law)
{
    if (law == nullptr) {
        return nullptr;
    }
    try {
        return dup(law_text(law->law));
    } catch (const std::exception&) {
        return nullptr; // a law holding something law text cannot state (a creature's state, a non-finite number)
    }
}

void fe_string_free(char* s)
{
    std::free(s);
}

} // extern "C"
