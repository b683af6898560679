// lawc: the law compiler on the command line (the reference's tools/lawc/main.cpp:1-21, with --hip for --cuda).
//
//   lawc --ir LAW...              the IR as a listing
//   lawc --ir-json OUT LAW...     the IR as JSON (the reference's format)
//   lawc --hip OUT LAW...         the generated gfx950 translation unit (the reference's --cuda, main.cpp:221-224)
#include "fe_field.h"

#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

int main(int argc, char** argv)
{
    bool listing = false;
    std::string ir_json;
    std::string hip_out;
    std::vector<std::string> paths;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--ir") {
            listing = true;
        } else if (a == "--ir-json" && i + 1 < argc) {
            ir_json = argv[++i];
        } else if (a == "--hip" && i + 1 < argc) {
            hip_out = argv[++i];
        } else if (a == "--help" || a == "-h") {
            std::printf("lawc [--ir] [--ir-json F] [--hip F] LAW...\n");
            return 0;
        } else {
            paths.push_back(a);
        }
    }
    if (paths.empty()) {
        std::fprintf(stderr, "no law: give LAW files (--help)\n");
        return 2;
    }
    std::vector<std::string> texts;
    for (const std::string& p : paths) {
        std::ifstream in(p);
        if (!in) {
            std::fprintf(stderr, "no law at %s\n", p.c_str());
            return 1;
        }
        std::ostringstream s;
        s << in.rdbuf();
        texts.push_back(s.str());
    }
    std::vector<const char*> names, bodies;
    for (std::size_t i = 0; i < paths.size(); ++i) {
        names.push_back(paths[i].c_str());
        bodies.push_back(texts[i].c_str());
    }
    std::vector<char> errors(1 << 16);
    fe_law* law = fe_law_compile(names.data(), bodies.data(), static_cast<int>(paths.size()), errors.data(), errors.size());
    if (errors[0] != '\0') {
        std::fprintf(stderr, "%s", errors.data());
    }
    if (law == nullptr) {
        return 1;
    }
    if (listing) {
        char* s = fe_law_describe(law);
        std::printf("%s", s);
        fe_string_free(s);
    }
    if (!ir_json.empty()) {
        char* s = fe_law_ir_json(law);
        std::ofstream(ir_json) << s;
        fe_string_free(s);
        std::printf("IR written to %s\n", ir_json.c_str());
    }
    int status = 0;
    if (!hip_out.empty()) {
        std::vector<char> why(4096);
        char* src = fe_law_hip_source(law, why.data(), why.size());
        if (src == nullptr) {
            std::fprintf(stderr, "no HIP source: %s\n", why.data());
            status = 1;
        } else {
            std::ofstream(hip_out) << src;
            fe_string_free(src);
            std::printf("HIP written to %s\n", hip_out.c_str());
        }
    }
    fe_law_free(law);
    return status;
}
