// Where a law runs, and what a step reports: the engine-facing boundary the HIP backend implements.
//
// The same interface as the reference's (src/field/Backend.hpp:16-84, src/field/Grid.hpp:24-77), so a backend written
// against one drops into the other: LawBackend's virtuals, Violation, CellBox, FieldEdits, TermProfile and StepProfile
// keep their members and meaning; GridShape indexes (z * ny + y) * nx + x; FieldSet holds the host's float fields in
// declaration order with an FNV-1a digest of their bits.
#pragma once

#include "field/Law.hpp"

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace fieldengine::field {

struct GridShape {
    int nx = 0;
    int ny = 0;
    int nz = 0;
    double cell = 1.0; // metres per voxel edge (h)
    [[nodiscard]] std::size_t cells() const
    {
        return static_cast<std::size_t>(nx) * static_cast<std::size_t>(ny) * static_cast<std::size_t>(nz);
    }
    [[nodiscard]] std::size_t index(int x, int y, int z) const
    {
        return (static_cast<std::size_t>(z) * static_cast<std::size_t>(ny) + static_cast<std::size_t>(y)) * static_cast<std::size_t>(nx) +
            static_cast<std::size_t>(x);
    }
};

struct Field {
    std::string name;
    std::vector<float> values;
};

class FieldSet {
public:
    FieldSet() = default;
    explicit FieldSet(GridShape shape) : shape_(shape) {}
    [[nodiscard]] const GridShape& shape() const { return shape_; }
    int add(const std::string& name, float initial = 0.0f);
    [[nodiscard]] int find(const std::string& name) const;
    [[nodiscard]] std::size_t count() const { return fields_.size(); }
    [[nodiscard]] Field& at(int id) { return fields_[static_cast<std::size_t>(id)]; }
    [[nodiscard]] const Field& at(int id) const { return fields_[static_cast<std::size_t>(id)]; }
    [[nodiscard]] std::vector<float>& values(int id) { return at(id).values; }
    [[nodiscard]] const std::vector<float>& values(int id) const { return at(id).values; }
    // FNV-1a over every field's bits in declaration order (src/field/Grid.cpp:60-75).
    [[nodiscard]] std::uint64_t digest() const;

private:
    GridShape shape_{};
    std::vector<Field> fields_;
};

struct Violation {
    enum class Kind { Range, NotFinite, Conservation, SubstepCap };
    Kind kind = Kind::Range;
    std::string term;
    std::string field;
    std::string detail;
};

const char* to_string(Violation::Kind kind);

struct CellBox {
    int x0 = 0, y0 = 0, z0 = 0, x1 = 0, y1 = 0, z1 = 0;
};

struct FieldEdits {
    bool all = true;
    std::vector<std::vector<char>> planes; // per field, per z-plane: written from outside since the backend last saw it
    int shift[3] = {0, 0, 0};
    std::vector<CellBox> boxes;
    [[nodiscard]] bool shifted() const { return shift[0] != 0 || shift[1] != 0 || shift[2] != 0 || !boxes.empty(); }
};

struct TermProfile {
    std::string name;
    double ms = 0.0;
    int launches = 0;
    int substeps = 0;
    bool idle = false;
    bool learned = false;
    bool fell_back = false;
    std::string kernel; // the gather kernel the term's last substep launched (for profilers' names)
};

struct StepProfile {
    std::string backend;
    double total_ms = 0.0;
    double gpu_ms = 0.0;
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
    virtual void step(double dt, bool respect_max_dt, bool learned, Law& law, FieldSet& fields, FieldEdits& edits,
        std::vector<char>& changed, std::vector<Violation>& violations, std::vector<int>& substeps) = 0;
    virtual void derive(Law& law, FieldSet& fields, FieldEdits& edits, std::vector<char>& changed) = 0;
    // `steps` steps of dt, as that many calls of step() take them: changed is the union over the steps, violations all
    // of theirs, substeps the last step's. No counterpart in the reference: a backend may replay what it captured.
    virtual void run(double dt, int steps, bool respect_max_dt, bool learned, Law& law, FieldSet& fields, FieldEdits& edits,
        std::vector<char>& changed, std::vector<Violation>& violations, std::vector<int>& substeps)
    {
        std::vector<char> any(law.fields.size(), 0);
        for (int i = 0; i < steps; ++i) {
            step(dt, respect_max_dt, learned, law, fields, edits, changed, violations, substeps);
            for (std::size_t f = 0; f < changed.size() && f < any.size(); ++f) {
                any[f] = any[f] || changed[f];
            }
        }
        changed = any;
    }
    virtual void check_contracts_every(int steps) = 0;
    virtual void set_lazy(bool on) { (void)on; }
    virtual void fetch(FieldSet& fields, const std::vector<char>& which) { (void)fields; (void)which; }
    [[nodiscard]] virtual const std::string& adapter() const = 0;
    [[nodiscard]] virtual const char* api() const = 0;
    [[nodiscard]] virtual const StepProfile& profile() const = 0;
};

} // namespace fieldengine::field
