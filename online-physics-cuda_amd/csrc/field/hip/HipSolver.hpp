// THE LAW ON MI355X: the HIP backend, the drop-in for the reference's CudaSolver (src/field/cuda/CudaSolver.hpp:24-32).
//
// HipSolver has the reference backend's three statics; what create() returns is a LawBackend (the reference's
// interface, unchanged), which is also a HipBackend: the calls only this backend offers, for callers that know it
// (the C-ABI in include/fe_field.h, this build's Solver).
//
// When a law is built, its IR is turned into one HIP translation unit (HipCodegen.hpp), compiled by hipRTC for the
// card (cached on disk by text, options, arch and runtime) and loaded with hipModule. Each term is one fused gather
// kernel (ping-pong when a field it writes is read off its own cell) after the stages the IR plans; contracts (range,
// finite, conserved totals) run as kernels with fixed-order reductions and are read back once a step when checked;
// a learned scheme runs under guard (snapshot, step, check, undo and the classical substeps on a breach). Every launch
// is on the runtime's one stream; terms are timed with events. No result depends on thread timing.
#pragma once

#include "field/Backend.hpp"

#include <memory>
#include <string>
#include <vector>

namespace fieldengine::field {

class HipSolver {
public:
    // Null with the reason when there is no HIP device, or the law uses an op this backend does not run, or the
    // generated source does not compile.
    static std::shared_ptr<LawBackend> create(Law& law, const GridShape& shape, std::string& why);
    // Why no law can run on the card here; empty when one can.
    static std::string unavailable();
    // The generated translation unit of a law (lawc --hip prints it); empty when it cannot be generated.
    static std::string source(const Law& law);
    static std::string source(const Law& law, std::string& why);
};

// What the HIP backend adds to LawBackend. None of it has a counterpart in the reference; a caller holding a
// LawBackend from HipSolver::create reaches it with dynamic_cast.
class HipBackend : public LawBackend {
public:
    // `steps` steps of dt, the state exactly that of as many step() calls: changed is the union over the steps,
    // violations all of theirs, substeps the last step's. Runs of steps with no contract due and nothing read back
    // replay captured graphs.
    virtual void run(double dt, int steps, bool respect_max_dt, bool learned, Law& law, FieldSet& fields,
        FieldEdits& edits, std::vector<char>& changed, std::vector<Violation>& violations, std::vector<int>& substeps) = 0;
    // Per term, the gather kernel its last substep launched (the name profilers list it under).
    [[nodiscard]] virtual const std::vector<std::string>& gather_kernels() const = 0;
    // A field's current buffer on the card; whoever holds it may write through it (the backend forgets what it knew
    // of the field's cells, e.g. a cached total).
    virtual void* device_field(int field) = 0;
    // The caller wrote a field's cells on the card: the host copy is stale.
    virtual void device_written(int field) = 0;
    // A static field's cells are read from `buffer` (the caller's device memory, a whole lattice of floats) from now
    // on, without a copy; nullptr goes back to the solver's own array. False (nothing changes) for a field a program
    // writes. While bound, whatever writes the field's cells (an upload, a window shift) writes into `buffer`.
    virtual bool bind_field(int field, void* buffer) = 0;
    // A field's fixed-order compensated total and total magnitude, on the card, waited for.
    virtual void field_total(int field, double& sum, double& magnitude) = 0;
    // Adds a field's total to two doubles on the card (a running sum and its compensation), without waiting.
    virtual void field_total_accumulate(int field, double* device_pair) = 0;
};

} // namespace fieldengine::field
