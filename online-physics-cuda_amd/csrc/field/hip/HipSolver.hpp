// THE LAW ON MI355X: the HIP backend. The drop-in for the reference's CudaSolver (src/field/cuda/CudaSolver.hpp:24-32):
// the same three statics, and a LawBackend that runs the law's generated module.
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

namespace fieldengine::field {

class HipSolver {
public:
    // Null with the reason when there is no HIP device, or the law uses an op this backend does not run, or the
    // generated source does not compile.
    static std::shared_ptr<LawBackend> create(Law& law, const GridShape& shape, std::string& why);
    // Why no law can run on the card here; empty when one can.
    static std::string unavailable();
    // The generated translation unit of a law (lawc --hip prints it); empty with a reason when it cannot be generated.
    static std::string source(const Law& law, std::string& why);
};

} // namespace fieldengine::field
