// lawc: the law compiler and runner on the command line, with the reference's flags (tools/lawc/main.cpp:95-445),
// --hip in place of --cuda and the HIP backend as the one that steps.
//
//   lawc --ir LAW...                       the IR as a listing
//   lawc --ir-json OUT LAW...              the IR as JSON (the reference's format)
//   lawc --hip OUT LAW...                  the generated gfx950 translation unit (the reference's --cuda)
//   lawc --adjoint OUT [--term T] LAW...   the adjoint law
//   lawc [--scene S] [--grid NX NY NZ H] [--init DIR] [--param NAME VALUE]... --steps N [--dt X] [--learned] [--eager]
//        [--profile] [--dump DIR [--every K] [--dump-only F,G]] [--json FILE [--warmup N]] [LAW...]
//                                          runs the law on the card: a scene (or the grid and the laws' initial values),
//                                          then a state from DIR (--init), N steps; dumps states as DIR/step_S (the
//                                          state format of lawc --dump: meta.json and FIELD.bin, raw float32)
// The reference's --backend names its CPU, CUDA and D3D11 engines; this build steps on the card only ("hip"), and its
// --compare and --vjp (CPU beside GPU, a backward pass over dumped states) are left to the reference's own lawc.
#include "field/Law.hpp"
#include "field/Scene.hpp"
#include "field/Solver.hpp"
#include "field/hip/HipSolver.hpp"
#include "field/ir/Adjoint.hpp"
#include "field/ir/LawIR.hpp"

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <optional>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

using namespace fieldengine::field;

namespace {

struct Options {
    bool listing = false, learned = false, eager = false, profile = false;
    std::optional<std::string> ir_json, hip_out, adjoint_out, scene, dump, init, json;
    std::string term, dump_only, backend = "hip";
    std::optional<GridShape> grid;
    int steps = 0, every = 0, warmup = 0;
    double dt = 0.05;
    std::vector<std::pair<std::string, double>> params;
    std::vector<std::string> laws;
};

const char* kUsage =
    "lawc [--ir] [--ir-json F] [--hip F] [--adjoint F [--term T]] [--scene S] [--grid NX NY NZ H] [--steps N] [--dt X]\n"
    "     [--backend hip] [--learned] [--profile] [--eager] [--dump DIR [--every K] [--dump-only F,G]] [--init DIR]\n"
    "     [--param NAME VALUE]... [--json FILE [--warmup N]] LAW...\n";

// Flags into options; exits with 2 on a flag without its value or one this build does not take.
Options parse(int argc, char** argv)
{
    Options o;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto value = [&]() -> std::string {
            if (i + 1 >= argc) {
                std::cerr << a << " needs a value\n";
                std::exit(2);
            }
            return argv[++i];
        };
        auto number = [&]() { return std::atof(value().c_str()); };
        if (a == "--ir") o.listing = true;
        else if (a == "--ir-json") o.ir_json = value();
        else if (a == "--hip") o.hip_out = value();
        else if (a == "--adjoint") o.adjoint_out = value();
        else if (a == "--term") o.term = value();
        else if (a == "--scene") o.scene = value();
        else if (a == "--steps") o.steps = std::atoi(value().c_str());
        else if (a == "--dt") o.dt = number();
        else if (a == "--backend") o.backend = value();
        else if (a == "--learned") o.learned = true;
        else if (a == "--eager") o.eager = true;
        else if (a == "--profile") o.profile = true;
        else if (a == "--dump") o.dump = value();
        else if (a == "--every") o.every = std::atoi(value().c_str());
        else if (a == "--dump-only") o.dump_only = value();
        else if (a == "--init") o.init = value();
        else if (a == "--json") o.json = value();
        else if (a == "--warmup") o.warmup = std::atoi(value().c_str());
        else if (a == "--param") {
            const std::string name = value();
            o.params.emplace_back(name, number());
        } else if (a == "--grid") {
            GridShape g;
            g.nx = std::atoi(value().c_str());
            g.ny = std::atoi(value().c_str());
            g.nz = std::atoi(value().c_str());
            g.cell = number();
            o.grid = g;
        } else if (a == "--help" || a == "-h") {
            std::cout << kUsage;
            std::exit(0);
        } else if (a == "--cuda" || a == "--compare" || a == "--vjp") {
            std::cerr << a << " is the reference's lawc's (this build's backend is HIP: --hip; no CPU engine to compare with)\n";
            std::exit(2);
        } else if (a.rfind("--", 0) == 0) {
            std::cerr << "unknown flag " << a << " (--help)\n";
            std::exit(2);
        } else {
            o.laws.push_back(a);
        }
    }
    return o;
}

// The fields named in `only` ("a,b"; empty: all) of `fields`, dumped into dir.
bool dump_some(const FieldSet& fields, const std::string& only, const std::string& dir, std::string& error)
{
    if (only.empty()) {
        return dump_fields(fields, dir, error);
    }
    FieldSet some(fields.shape());
    std::istringstream names(only);
    for (std::string name; std::getline(names, name, ',');) {
        const int id = fields.find(name);
        if (id < 0) {
            error = "--dump-only: the law has no field '" + name + "'";
            return false;
        }
        some.values(some.add(name)) = fields.values(id);
    }
    return dump_fields(some, dir, error);
}

void print_profile(const StepProfile& p)
{
    std::cout << "  profile: " << p.backend << ", " << p.launches << " launches, card " << p.gpu_ms << " ms, step " << p.total_ms << " ms\n";
    for (const TermProfile& t : p.terms) {
        std::cout << "    " << std::left << std::setw(16) << t.name << std::right << std::setw(9) << t.ms << " ms  " << t.substeps
                  << " substep(s)" << (t.idle ? "  idle" : "") << (t.learned ? "  learned" : "") << (t.fell_back ? "  fell back" : "")
                  << "\n";
    }
}

} // namespace

int main(int argc, char** argv)
{
    Options o = parse(argc, argv);
    Scene scene;
    std::string error;
    if (o.scene && !load_scene(*o.scene, scene, error)) {
        std::cerr << error << "\n";
        return 1;
    }
    if (o.laws.empty()) {
        o.laws = scene.laws;
    }
    if (o.grid) {
        scene.grid = *o.grid;
    }
    std::vector<LawSource> sources;
    for (const std::string& path : o.laws) {
        std::optional<LawSource> source = read_law_file(path);
        if (!source) {
            std::cerr << "no law at " << path << "\n";
            return 1;
        }
        sources.push_back(std::move(*source));
    }
    if (sources.empty()) {
        std::cerr << "no law: give LAW files or a scene with 'law' lines (--help)\n";
        return 2;
    }
    CompileResult compiled = compile_law(sources);
    for (const std::string& e : compiled.errors) {
        std::cerr << e << "\n";
    }
    if (!compiled.law) {
        return 1;
    }
    Law& law = *compiled.law;
    for (const NetDecl& net : law.nets) {
        if (!net.loaded) {
            std::cerr << "note: net '" << net.name << "' has no weights at " << net.file << " (it starts untrained: its output is 0)\n";
        }
    }
    if (o.listing || o.ir_json) {
        const ir::Module m = ir::lower(law);
        if (o.listing) {
            std::cout << ir::describe(m, law);
        }
        if (o.ir_json) {
            std::ofstream(*o.ir_json) << ir::to_json(m, law);
            std::cout << "IR written to " << *o.ir_json << "\n";
        }
    }
    if (o.hip_out) {
        std::string why;
        const std::string src = HipSolver::source(law, why);
        if (src.empty()) {
            std::cerr << "no HIP source: " << why << "\n";
            return 1;
        }
        std::ofstream(*o.hip_out) << src;
        std::cout << "HIP written to " << *o.hip_out << "\n";
    }
    if (o.adjoint_out) {
        const std::string text = ir::adjoint_law(law, o.term, error);
        if (text.empty()) {
            std::cerr << "no adjoint: " << error << "\n";
            return 1;
        }
        std::ofstream(*o.adjoint_out) << text;
        std::cout << "adjoint law written to " << *o.adjoint_out << "\n";
    }
    if (o.steps <= 0 && !o.dump) {
        return 0;
    }
    if (o.backend != "hip") {
        std::cerr << "--backend " << o.backend << ": this build steps laws on the card only (hip)\n";
        return 2;
    }

    // The scene's start, the overrides and a state from disk, on the host; then the card.
    Solver solver(law, scene.grid);
    if (!apply_scene(scene, solver, error)) {
        std::cerr << error << "\n";
        return 1;
    }
    for (const auto& [name, value] : o.params) {
        if (!solver.set_param(name, value)) {
            std::cerr << "the law has no param '" << name << "'\n";
            return 1;
        }
    }
    if (o.init && !load_fields(solver.fields(), *o.init, error)) {
        std::cerr << error << "\n";
        return 1;
    }
    if (o.steps <= 0) {
        // The start alone, before any derive (no card needed): what a scene makes of the fields.
        if (!dump_some(solver.fields_const(), o.dump_only, *o.dump + "/step_0", error)) {
            std::cerr << error << "\n";
            return 1;
        }
        std::cout << "start written to " << *o.dump << "/step_0\n";
        return 0;
    }
    std::string why;
    if (!solver.use_gpu(why)) {
        std::cerr << "no hip backend: " << why << "\n";
        return 1;
    }
    solver.refresh_derived();
    solver.set_learned(o.learned);
    solver.check_contracts_every(1);
    solver.set_lazy_readback(!o.eager); // lazy: the cells come back when they are dumped, not every step
    std::cout << law.fields.size() << " fields, " << law.terms.size() << " terms, " << law.nets.size() << " nets; grid " << scene.grid.nx
              << "x" << scene.grid.ny << "x" << scene.grid.nz << " at " << scene.grid.cell << " m; " << o.steps << " steps of " << o.dt
              << " s on " << solver.backend_api() << " (" << solver.gpu_name() << ")\n";
    auto dump_at = [&](int step) {
        if (!dump_some(solver.fields_const(), o.dump_only, *o.dump + "/step_" + std::to_string(step), error)) {
            std::cerr << error << "\n";
            return false;
        }
        return true;
    };
    if (o.dump && !dump_at(0)) {
        return 1;
    }
    double total_ms = 0.0, timed_ms = 0.0, timed_gpu_ms = 0.0;
    std::size_t violations = 0;
    int fallbacks = 0;
    std::vector<double> step_ms, term_ms(law.terms.size(), 0.0);
    std::vector<int> last_substeps;
    for (int s = 1; s <= o.steps; ++s) {
        const auto start = std::chrono::steady_clock::now();
        solver.step(o.dt);
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - start).count();
        total_ms += ms;
        const StepProfile& p = solver.profile();
        fallbacks += p.fallbacks;
        if (s > o.warmup) {
            timed_ms += ms;
            timed_gpu_ms += p.gpu_ms;
            step_ms.push_back(ms);
            for (std::size_t t = 0; t < p.terms.size() && t < term_ms.size(); ++t) {
                term_ms[t] += p.terms[t].ms;
            }
        }
        for (const Violation& v : solver.violations()) {
            if (++violations <= 12) {
                std::cerr << "  step " << s << ": [" << v.term << "] " << v.field << " " << to_string(v.kind) << ": " << v.detail << "\n";
            }
        }
        if (o.dump && o.every > 0 && s % o.every == 0 && !dump_at(s)) {
            return 1;
        }
        if (o.profile && (s == 1 || s == o.steps)) {
            print_profile(p);
        }
        last_substeps = solver.last_substeps();
    }
    if (o.dump && (o.every <= 0 || o.steps % o.every != 0) && !dump_at(o.steps)) {
        return 1;
    }
    std::cout << "mean step " << total_ms / o.steps << " ms, " << violations << " contract violation(s)\n";
    if (o.json) {
        std::sort(step_ms.begin(), step_ms.end());
        const int timed = static_cast<int>(step_ms.size());
        std::ofstream out(*o.json);
        out << std::setprecision(9) << "{\"backend\":\"" << solver.backend_api() << "\",\"cells\":"
            << static_cast<long long>(scene.grid.cells()) << ",\"steps\":" << timed << ",\"mean_ms\":" << (timed ? timed_ms / timed : 0.0)
            << ",\"median_ms\":" << (timed ? step_ms[static_cast<std::size_t>(timed / 2)] : 0.0) << ",\"gpu_ms\":"
            << (timed ? timed_gpu_ms / timed : 0.0) << ",\"violations\":" << violations << ",\"fallbacks\":" << fallbacks << ",\"substeps\":[";
        for (std::size_t t = 0; t < last_substeps.size(); ++t) {
            out << (t ? "," : "") << last_substeps[t];
        }
        out << "],\"terms\":[";
        for (std::size_t t = 0; t < law.terms.size(); ++t) {
            out << (t ? "," : "") << "{\"name\":\"" << law.terms[t].name << "\",\"ms\":" << (timed ? term_ms[t] / timed : 0.0) << "}";
        }
        out << "]}\n";
    }
    return 0;
}
