// The HIP backend behind the reference's own engine: built against the reference's unchanged headers (field/Law.hpp,
// field/Backend.hpp, field/Solver.hpp, field/ir/LawIR.hpp) and linked with its Solver, Law and IR objects, the way a
// maintainer adds the backend to that tree. Compiles a law with the reference's front end, builds the reference's
// Solver, and asks it for the card with FIELD_BACKEND=hip: Solver::use_gpu must reach HipSolver::create, which answers
// with the card's backend where there is one and with nullptr and the reason where there is none. Prints what it saw.
#include "field/Law.hpp"
#include "field/Solver.hpp"
#include "field/hip/HipSolver.hpp"

#include <cstdio>
#include <cstdlib>
#include <string>

using namespace fieldengine::field;

int main()
{
    const CompileResult compiled = compile_law({LawSource{"heat.law",
        "field T init 0 range -100 100\nparam k = 0.1\nterm conduction\n  d T = diffuse(k, T)\n  conserves T\n"
        "  max_dt = 0.15 * h * h / k\n"}});
    if (!compiled.law) {
        std::printf("compile failed: %s\n", compiled.errors.empty() ? "?" : compiled.errors[0].c_str());
        return 2;
    }
    Solver solver(*compiled.law, GridShape{16, 12, 8, 0.25});
    setenv("FIELD_BACKEND", "hip", 1);
    std::string why;
    const bool on_card = solver.use_gpu(why);
    std::printf("use_gpu=%d\n", on_card ? 1 : 0);
    std::printf("why=%s\n", why.c_str());
    std::printf("api=%s\n", solver.backend_api());
    std::printf("unavailable=%s\n", HipSolver::unavailable().c_str());
    const std::string source = HipSolver::source(*compiled.law);
    std::printf("source_bytes=%zu\n", source.size());
    std::printf("gather=%d\n", source.find("void g_t0(K k") != std::string::npos ? 1 : 0);
    if (on_card) {
        // A step on the card through the reference's Solver, against its own CPU step of the same start.
        Solver cpu(*compiled.law, GridShape{16, 12, 8, 0.25});
        for (std::size_t i = 0; i < solver.fields().values(0).size(); ++i) {
            const float v = static_cast<float>((i * 2654435761u) % 1000) / 1000.0f;
            solver.fields().values(0)[i] = v;
            cpu.fields().values(0)[i] = v;
        }
        solver.step(0.01);
        cpu.step(0.01);
        std::printf("same_bits=%d\n", solver.fields().digest() == cpu.fields().digest() ? 1 : 0);
    }
    return 0;
}
