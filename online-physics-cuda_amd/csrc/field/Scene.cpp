// Scenes and states on disk. See Scene.hpp; the formats are the reference's (src/field/Scene.cpp:9-167).
#include "field/Scene.hpp"

#include <cstdint>
#include <filesystem>
#include <fstream>
#include <sstream>

namespace fieldengine::field {

namespace {

// How many numbers follow FIELD VALUE for each shape.
int arity(const std::string& kind)
{
    if (kind == "fill") return 0;
    if (kind == "box") return 6;
    if (kind == "ball") return 4;
    if (kind == "noise") return 1;
    return -1;
}

// The noise of cell `i` for `seed`, uniform in [0, 1): a 64-bit mix of the two (splitmix64's finaliser over
// seed * golden ratio xor i * its second constant), so no generator state is carried between cells.
double cell_noise(std::uint64_t seed, std::uint64_t i)
{
    std::uint64_t h = seed * 0x9e3779b97f4a7c15ull ^ i * 0xbf58476d1ce4e5b9ull;
    h ^= h >> 31;
    h *= 0x94d049bb133111ebull;
    h ^= h >> 29;
    return static_cast<double>(h >> 11) * 0x1.0p-53;
}

bool inside(const Scene::Op& op, int x, int y, int z)
{
    if (op.kind == "box") {
        return x >= op.args[0] && y >= op.args[1] && z >= op.args[2] && x < op.args[3] && y < op.args[4] && z < op.args[5];
    }
    const double dx = x - op.args[0], dy = y - op.args[1], dz = z - op.args[2];
    return dx * dx + dy * dy + dz * dz <= op.args[3] * op.args[3]; // ball
}

} // namespace

bool load_scene(const std::string& path, Scene& scene, std::string& error)
{
    std::ifstream in(path);
    if (!in) {
        error = "no scene at " + path;
        return false;
    }
    std::string line;
    for (int number = 1; std::getline(in, line); ++number) {
        line = line.substr(0, line.find('#'));
        std::istringstream words(line);
        std::string kind;
        if (!(words >> kind)) {
            continue;
        }
        const std::string where = path + ":" + std::to_string(number) + ": ";
        if (kind == "grid") {
            GridShape g;
            if (!(words >> g.nx >> g.ny >> g.nz >> g.cell) || g.nx < 1 || g.ny < 1 || g.nz < 1 || !(g.cell > 0.0)) {
                error = where + "expected 'grid NX NY NZ METRES'";
                return false;
            }
            scene.grid = g;
        } else if (kind == "law") {
            std::string law;
            words >> law;
            scene.laws.push_back(law);
        } else if (kind == "param") {
            std::pair<std::string, double> p;
            if (!(words >> p.first >> p.second)) {
                error = where + "expected 'param NAME VALUE'";
                return false;
            }
            scene.params.push_back(p);
        } else if (arity(kind) >= 0) {
            Scene::Op op;
            op.kind = kind;
            if (!(words >> op.field >> op.value)) {
                error = where + "expected '" + kind + " FIELD VALUE ...'";
                return false;
            }
            const int count = arity(kind);
            for (int a = 0; a < count; ++a) {
                if (!(words >> op.args[a])) {
                    error = where + kind + " needs " + std::to_string(count) + " numbers after the value";
                    return false;
                }
            }
            scene.ops.push_back(op);
        } else {
            error = where + "unknown line '" + kind + "' (grid, law, param, fill, box, ball, noise)";
            return false;
        }
    }
    return true;
}

bool apply_scene(const Scene& scene, Solver& solver, std::string& error)
{
    for (const auto& [name, value] : scene.params) {
        if (!solver.set_param(name, value)) {
            error = "the law has no param '" + name + "'";
            return false;
        }
    }
    FieldSet& set = solver.fields();
    const GridShape& g = set.shape();
    for (const Scene::Op& op : scene.ops) {
        const int id = set.find(op.field);
        if (id < 0) {
            error = "the law has no field '" + op.field + "'";
            return false;
        }
        std::vector<float>& cells = set.values(id);
        const float v = static_cast<float>(op.value);
        if (op.kind == "fill") {
            std::fill(cells.begin(), cells.end(), v);
            continue;
        }
        for (int z = 0; z < g.nz; ++z) {
            for (int y = 0; y < g.ny; ++y) {
                for (int x = 0; x < g.nx; ++x) {
                    const std::size_t i = g.index(x, y, z);
                    if (op.kind == "noise") {
                        cells[i] += static_cast<float>(op.value * cell_noise(static_cast<std::uint64_t>(op.args[0]), i));
                    } else if (inside(op, x, y, z)) {
                        cells[i] = v;
                    }
                }
            }
        }
    }
    if (solver.on_gpu()) {
        solver.refresh_derived(); // off the card the derives are taken when the solver reaches it
    }
    return true;
}

bool dump_fields(const FieldSet& fields, const std::string& dir, std::string& error)
{
    std::error_code ec;
    std::filesystem::create_directories(dir, ec);
    const GridShape& g = fields.shape();
    std::ostringstream names;
    for (std::size_t i = 0; i < fields.count(); ++i) {
        const Field& f = fields.at(static_cast<int>(i));
        names << (i ? "," : "") << '"' << f.name << '"';
        std::ofstream out(dir + "/" + f.name + ".bin", std::ios::binary);
        out.write(reinterpret_cast<const char*>(f.values.data()), static_cast<std::streamsize>(f.values.size() * sizeof(float)));
        if (!out.flush()) {
            error = "cannot write " + dir + "/" + f.name + ".bin (disk full?)";
            return false;
        }
    }
    std::ofstream meta(dir + "/meta.json");
    meta << "{\"nx\":" << g.nx << ",\"ny\":" << g.ny << ",\"nz\":" << g.nz << ",\"h\":" << g.cell << ",\"fields\":[" << names.str() << "]}\n";
    if (!meta.flush()) {
        error = "cannot write " + dir;
        return false;
    }
    return true;
}

bool load_fields(FieldSet& fields, const std::string& dir, std::string& error)
{
    if (!std::filesystem::is_directory(dir)) {
        error = "no state at " + dir;
        return false;
    }
    for (std::size_t i = 0; i < fields.count(); ++i) {
        std::vector<float>& values = fields.values(static_cast<int>(i));
        const std::string path = dir + "/" + fields.at(static_cast<int>(i)).name + ".bin";
        std::error_code ec;
        const std::uintmax_t bytes = std::filesystem::file_size(path, ec);
        if (ec) {
            continue; // not in this state: the field keeps its values
        }
        if (bytes != values.size() * sizeof(float)) {
            error = path + " holds " + std::to_string(bytes / sizeof(float)) + " floats, the lattice " + std::to_string(values.size());
            return false;
        }
        std::ifstream in(path, std::ios::binary);
        in.read(reinterpret_cast<char*>(values.data()), static_cast<std::streamsize>(bytes));
        if (!in) {
            error = "cannot read " + path;
            return false;
        }
    }
    return true;
}

} // namespace fieldengine::field
