// Starting states as text, and states on disk: the reference's scene and state formats (src/field/Scene.hpp:9-44), so
// scenes, dumps and the tools that read them (its trainers, its PDEBench exporter) carry over unchanged.
//
// A scene, one statement a line ('#' starts a comment):
//   grid NX NY NZ METRES         the lattice
//   law PATH                     a law file (several may follow each other)
//   param NAME VALUE
//   fill FIELD V                 every cell
//   box FIELD V X0 Y0 Z0 X1 Y1 Z1    cells of [X0, X1) x [Y0, Y1) x [Z0, Z1)
//   ball FIELD V X Y Z R         cells within R cells of (X, Y, Z)
//   noise FIELD AMP SEED         adds a hash of (seed, cell) in [0, AMP): the same text, the same start
// A state on disk: DIR/meta.json ({"nx","ny","nz","h","fields":[...]}) and DIR/FIELD.bin, raw float32 in cell order
// (x fastest).
#pragma once

#include "field/Grid.hpp"
#include "field/Solver.hpp"

#include <string>
#include <utility>
#include <vector>

namespace fieldengine::field {

struct Scene {
    GridShape grid{64, 64, 1, 1.0};
    std::vector<std::string> laws;
    struct Op {
        std::string kind; // fill, box, ball, noise
        std::string field;
        double value = 0.0;
        double args[6] = {0, 0, 0, 0, 0, 0};
    };
    std::vector<Op> ops;
    std::vector<std::pair<std::string, double>> params;
};

bool load_scene(const std::string& path, Scene& scene, std::string& error);
// The scene's params, then its shapes in order into the solver's fields (by name: a name the law lacks is an error),
// then the derives (on the card; a solver not yet on it takes them when it gets there).
bool apply_scene(const Scene& scene, Solver& solver, std::string& error);

bool dump_fields(const FieldSet& fields, const std::string& dir, std::string& error);
// Each DIR/FIELD.bin of a field the set has replaces that field (it must hold exactly the lattice's cells); fields the
// directory lacks keep their values.
bool load_fields(FieldSet& fields, const std::string& dir, std::string& error);

} // namespace fieldengine::field
