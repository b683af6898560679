// The host's fields. See Grid.hpp.
#include "field/Grid.hpp"

#include <algorithm>
#include <cmath>

namespace fieldengine::field {

int FieldSet::add(const std::string& name, float initial)
{
    if (const int existing = find(name); existing >= 0) {
        return existing;
    }
    fields_.push_back(Field{name, std::vector<float>(shape_.cells(), initial)});
    return static_cast<int>(fields_.size() - 1);
}

int FieldSet::find(const std::string& name) const
{
    const auto it = std::find_if(fields_.begin(), fields_.end(), [&](const Field& f) { return f.name == name; });
    return it == fields_.end() ? -1 : static_cast<int>(it - fields_.begin());
}

float FieldSet::get_clamped(int id, int x, int y, int z) const
{
    const int cx = std::clamp(x, 0, shape_.nx - 1);
    const int cy = std::clamp(y, 0, shape_.ny - 1);
    const int cz = std::clamp(z, 0, shape_.nz - 1);
    return values(id)[shape_.index(cx, cy, cz)];
}

float FieldSet::sample(int id, double px, double py, double pz) const
{
    // Lattice coordinates of the point relative to the centres: cell i's centre is at (i + 0.5) h.
    const double g[3] = {px / shape_.cell - 0.5, py / shape_.cell - 0.5, pz / shape_.cell - 0.5};
    int base[3];
    double w[3];
    for (int a = 0; a < 3; ++a) {
        const double f = std::floor(g[a]);
        base[a] = static_cast<int>(f);
        w[a] = g[a] - f;
    }
    double acc = 0.0;
    for (int corner = 0; corner < 8; ++corner) {
        double weight = 1.0;
        int at[3];
        for (int a = 0; a < 3; ++a) {
            const int up = (corner >> a) & 1;
            at[a] = base[a] + up;
            weight *= up ? w[a] : 1.0 - w[a];
        }
        if (weight != 0.0) {
            acc += weight * static_cast<double>(get_clamped(id, at[0], at[1], at[2]));
        }
    }
    return static_cast<float>(acc);
}

std::uint64_t FieldSet::digest() const
{
    std::uint64_t h = 0xcbf29ce484222325ull; // FNV-1a 64 offset basis
    for (const Field& f : fields_) {
        const auto* bytes = reinterpret_cast<const unsigned char*>(f.values.data());
        for (std::size_t i = 0, n = f.values.size() * sizeof(float); i < n; ++i) {
            h = (h ^ bytes[i]) * 0x100000001b3ull; // FNV prime
        }
    }
    return h;
}

} // namespace fieldengine::field
