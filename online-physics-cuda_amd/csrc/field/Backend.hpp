// Where a law runs: the engine-facing interface a backend implements, and what a step reports.
//
// This is the reference's interface unchanged (src/field/Backend.hpp:16-84): the same Violation, CellBox, FieldEdits,
// TermProfile and StepProfile members and the same LawBackend virtuals, so the HIP backend drops in behind the
// reference's Solver and the reference's headers compile the backend as well as these do. What only the HIP backend
// offers (replaying captured runs of steps, the names of the kernels a step launched) is declared beside it, in
// field/hip/HipSolver.hpp.
#pragma once

#include "field/Grid.hpp"
#include "field/Law.hpp"

#include <cstddef>
#include <string>
#include <vector>

namespace fieldengine::field {

// A declared contract that a step broke, in words: which term, which field, what happened.
struct Violation {
    enum class Kind { Range, NotFinite, Conservation, SubstepCap };
    Kind kind = Kind::Range;
    std::string term; // empty for a derive
    std::string field;
    std::string detail;
};

const char* to_string(Violation::Kind kind);

// Cells [x0, x1) x [y0, y1) x [z0, z1).
struct CellBox {
    int x0 = 0, y0 = 0, z0 = 0, x1 = 0, y1 = 0, z1 = 0;
};

// What the host wrote into the lattice since the backend last looked.
struct FieldEdits {
    bool all = true;
    std::vector<std::vector<char>> planes; // per field, per z-plane
    // The window moved: the cell at c now holds what was at c + shift; only `boxes` hold cells new to the lattice.
    int shift[3] = {0, 0, 0};
    std::vector<CellBox> boxes;
    [[nodiscard]] bool shifted() const { return shift[0] != 0 || shift[1] != 0 || shift[2] != 0 || !boxes.empty(); }
};

struct TermProfile {
    std::string name;
    double ms = 0.0; // card time by events
    int launches = 0;
    int substeps = 0;
    bool idle = false;      // not dispatched: nothing to do
    bool learned = false;   // its learned scheme ran and held
    bool fell_back = false; // its learned scheme broke a contract: undone, the classical scheme ran
};

struct StepProfile {
    std::string backend;
    double total_ms = 0.0; // the step as the host waited for it
    double gpu_ms = 0.0;   // the card's time summed over what it ran
    double derive_ms = 0.0;
    double upload_ms = 0.0;
    double download_ms = 0.0;
    std::size_t upload_bytes = 0;
    std::size_t download_bytes = 0;
    int launches = 0;
    int fallbacks = 0;
    std::vector<TermProfile> terms;
};

class LawBackend {
public:
    virtual ~LawBackend() = default;
    // One step of every term; `changed` says per field whether any cell moved (only those come back into `fields`).
    virtual void step(double dt, bool respect_max_dt, bool learned, Law& law, FieldSet& fields, FieldEdits& edits,
        std::vector<char>& changed, std::vector<Violation>& violations, std::vector<int>& substeps) = 0;
    virtual void derive(Law& law, FieldSet& fields, FieldEdits& edits, std::vector<char>& changed) = 0;
    // Contracts are checked on one step of this many (a guarded learned term is checked on every one).
    virtual void check_contracts_every(int steps) = 0;
    // Lazy readback: results stay on the card until fetch() asks for them.
    virtual void set_lazy(bool on) { (void)on; }
    virtual void fetch(FieldSet& fields, const std::vector<char>& which) { (void)fields; (void)which; }
    [[nodiscard]] virtual const std::string& adapter() const = 0;
    [[nodiscard]] virtual const char* api() const = 0;
    [[nodiscard]] virtual const StepProfile& profile() const = 0;
};

} // namespace fieldengine::field
