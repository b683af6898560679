// Integrates a compiled law on one lattice, on the card.
//
// The reference's Solver API (src/field/Solver.hpp:103-215) for the calls that drive a backend: fields are the host's
// truth between steps (edits go up, results come back, lazily if asked), a step is every term in declaration order
// then every derive, contracts are reported as Violations. This build has no CPU integrator: a Solver runs only once
// use_gpu() has succeeded, and step() refuses to run otherwise (the CPU path is the oracle's, under oracle/).
#pragma once

#include "field/Backend.hpp"
#include "field/Law.hpp"
#include "field/hip/HipSolver.hpp"

#include <cstdint>
#include <memory>
#include <optional>
#include <string>
#include <string_view>
#include <vector>

namespace fieldengine::field {

class Solver {
public:
    Solver(const Law& law, GridShape shape);

    [[nodiscard]] FieldSet& fields(); // every field may change: all go up before the next step
    [[nodiscard]] const FieldSet& fields_const();
    [[nodiscard]] const std::vector<float>& values(int field); // fetches the one field when lazy
    float& edit_cell(int field, std::size_t cell);
    void mark_field_edited(int field);
    // The window moved by (dx, dy, dz) cells (src/field/Solver.hpp:40-42): the caller slides the host cells (what is now
    // at c was at c + d) and fills `arriving`; the backend slides its own copy on the card and uploads only those boxes.
    [[nodiscard]] FieldSet& shift_window(int dx, int dy, int dz, const std::vector<CellBox>& arriving);

    bool set_param(std::string_view name, double value);
    [[nodiscard]] std::optional<double> param(std::string_view name) const;
    [[nodiscard]] const Law& law() const { return law_; }
    [[nodiscard]] Law& law_for_training() { return law_; }

    bool use_gpu(std::string& why);
    [[nodiscard]] bool on_gpu() const { return gpu_ != nullptr; }
    [[nodiscard]] std::string gpu_name() const { return gpu_ ? gpu_->adapter() : std::string{}; }
    [[nodiscard]] const char* backend_api() const { return gpu_ ? gpu_->api() : "CPU"; }
    [[nodiscard]] int field(std::string_view name) const { return law_.field_id(name); }
    [[nodiscard]] bool learned() const { return learned_; }
    // A trainer's note that it rewrote net weights (the backend also notices by hashing them every step).
    void nets_changed() { ++nets_version_; }
    [[nodiscard]] std::uint64_t nets_version() const { return nets_version_; }
    [[nodiscard]] HipBackend* backend() const { return gpu_.get(); }
    // A field's cells were written on the card by the caller: the host copy is fetched before it is next read.
    void mark_device_written(int field) { stale_[static_cast<std::size_t>(field)] = 1; }
    [[nodiscard]] std::size_t fields_const_shape_cells() const { return fields_.shape().cells(); }

    void refresh_derived();
    void step(double dt);
    // `steps` steps of dt in one call (the backend may replay them from a captured graph); the same state as that many
    // step() calls, violations gathered over all of them.
    void run(double dt, int steps);
    void set_lazy_readback(bool on);
    [[nodiscard]] bool lazy_readback() const { return lazy_; }
    void check_contracts_every(int steps);
    void set_learned(bool on) { learned_ = on; }
    void set_respect_max_dt(bool respect) { respect_max_dt_ = respect; }

    [[nodiscard]] const std::vector<Violation>& violations() const { return violations_; }
    [[nodiscard]] const std::vector<int>& last_substeps() const { return substeps_; }
    [[nodiscard]] const std::vector<char>& last_changed() const { return changed_; }
    [[nodiscard]] const StepProfile& profile() const;

private:
    void fetch(const std::vector<char>& which);

    Law law_;
    FieldSet fields_;
    std::vector<char> stale_;
    FieldEdits edits_;
    bool lazy_ = false;
    bool learned_ = false;
    bool respect_max_dt_ = true;
    std::uint64_t nets_version_ = 0;
    std::vector<Violation> violations_;
    std::vector<int> substeps_;
    std::vector<char> changed_;
    std::shared_ptr<HipBackend> gpu_;
    StepProfile empty_profile_;
};

} // namespace fieldengine::field
