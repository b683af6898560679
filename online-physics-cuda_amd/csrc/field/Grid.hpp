// The lattice and the host's copy of its fields: the types a backend is handed.
//
// Member for member the reference's interface (src/field/Grid.hpp:14-75), so a backend compiled against either header
// set sees the same layout: cells are indexed (z * ny + y) * nx + x, every field is one float per cell, fields are
// addressed by their declaration order, and the digest is FNV-1a over every field's bytes in that order.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace fieldengine::field {

struct GridShape {
    int nx = 0;
    int ny = 0;
    int nz = 0;
    double cell = 1.0; // h: the edge of a cell in metres

    [[nodiscard]] std::size_t cells() const
    {
        return static_cast<std::size_t>(nx) * static_cast<std::size_t>(ny) * static_cast<std::size_t>(nz);
    }
    [[nodiscard]] std::size_t index(int x, int y, int z) const
    {
        const auto row = static_cast<std::size_t>(z) * static_cast<std::size_t>(ny) + static_cast<std::size_t>(y);
        return row * static_cast<std::size_t>(nx) + static_cast<std::size_t>(x);
    }
    [[nodiscard]] bool inside(int x, int y, int z) const
    {
        return x >= 0 && y >= 0 && z >= 0 && x < nx && y < ny && z < nz;
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

    // A new field holding `initial` everywhere; the id of the one already there when the name is taken.
    int add(const std::string& name, float initial = 0.0f);
    [[nodiscard]] int find(const std::string& name) const; // -1 when absent
    [[nodiscard]] std::size_t count() const { return fields_.size(); }

    [[nodiscard]] Field& at(int id) { return fields_[static_cast<std::size_t>(id)]; }
    [[nodiscard]] const Field& at(int id) const { return fields_[static_cast<std::size_t>(id)]; }
    [[nodiscard]] std::vector<float>& values(int id) { return at(id).values; }
    [[nodiscard]] const std::vector<float>& values(int id) const { return at(id).values; }

    // The cell at (x, y, z) with each coordinate clamped into the lattice (past an edge the edge cell answers).
    [[nodiscard]] float get_clamped(int id, int x, int y, int z) const;
    // Trilinear interpolation at a point in metres; cell centres sit at (i + 1/2) h.
    [[nodiscard]] float sample(int id, double px, double py, double pz) const;

    // FNV-1a 64 over the bytes of every field, in declaration order: equal digests, equal bits.
    [[nodiscard]] std::uint64_t digest() const;

private:
    GridShape shape_{};
    std::vector<Field> fields_;
};

} // namespace fieldengine::field
