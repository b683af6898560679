// Runs a law on the REFERENCE's CPU engine (linked from its objects, oracle/Makefile.ref) and prints the contract
// violations of its last step as JSON lines: {"term","field","kind","cells"} (cells: the count a range report starts
// with, -1 otherwise). Test infrastructure: tests/golden/make_golden.py stores them beside a case's fields.
//
//   ref_violations LAW NX NY NZ H INIT_DIR STEPS DT
#include "field/Law.hpp"
#include "field/Scene.hpp"
#include "field/Solver.hpp"

#include <cstdio>
#include <cstdlib>
#include <string>

using namespace fieldengine::field;

int main(int argc, char** argv)
{
    if (argc != 9) {
        std::fprintf(stderr, "usage: ref_violations LAW NX NY NZ H INIT_DIR STEPS DT\n");
        return 2;
    }
    const std::optional<LawSource> source = read_law_file(argv[1]);
    if (!source) {
        return 1;
    }
    CompileResult compiled = compile_law({*source});
    if (!compiled.law) {
        return 1;
    }
    const GridShape grid{std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4]), std::atof(argv[5])};
    Solver solver(*compiled.law, grid);
    std::string error;
    if (!load_fields(solver.fields(), argv[6], error)) {
        std::fprintf(stderr, "%s\n", error.c_str());
        return 1;
    }
    solver.refresh_derived();
    const int steps = std::atoi(argv[7]);
    for (int s = 0; s < steps; ++s) {
        solver.step(std::atof(argv[8]));
    }
    for (const Violation& v : solver.violations()) {
        const long cells = v.kind == Violation::Kind::Conservation || v.kind == Violation::Kind::SubstepCap ? -1 : std::atol(v.detail.c_str());
        std::printf("{\"term\": \"%s\", \"field\": \"%s\", \"kind\": \"%s\", \"cells\": %ld}\n", v.term.c_str(), v.field.c_str(),
            to_string(v.kind), cells);
    }
    return 0;
}
