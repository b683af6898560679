// From the law's IR to one HIP translation unit for gfx950, compiled at run time by hipRTC.
//
// The reference's generator (src/field/cuda/CudaCodegen.hpp:12-55) emits float arithmetic; this one emits the reference
// CPU's: every cell-shaped instruction is a device function that computes in double from its arguments (floats widened,
// scalars kept in double) and rounds once to float, the way the CPU evaluator stores each node (src/field/Law.cpp:757-
// 1466); scalar instructions stay double. The gather kernel of a term applies each equation as the CPU integrator does:
// d = float(double(rate) * sub_dt), T = T + d (src/field/Solver.cpp:391-427). So a generated step is the CPU's step to
// the bit, on every law whose values are not identically zero over the whole lattice (the CPU turns such an array into
// the scalar 0, src/field/Law.cpp:554-563, which changes nothing but signed zeros) and whose ops are exact in IEEE
// arithmetic or correctly rounded in both libraries (exp, log and pow are within one double ulp on both sides, so their
// float results agree unless a value falls within that of a float rounding boundary).
//
// Kernel arguments are one struct by value (the field table of now/next buffers, the lattice, the device buffers),
// so a step has no constant-memory writes between launches and can be captured in a graph.
#pragma once

#include "field/Law.hpp"
#include "field/ir/LawIR.hpp"

#include <map>
#include <string>
#include <vector>

namespace fieldengine::field::hip {

struct StageCode {
    // Distance: the signed distance to the surface of where the argument is >= 0.5, recomputed only when that set
    // changed. Slots from the instruction's slot s: s the answer, s+1 the set (1 inside, 0 outside), s+2 / s+3 the
    // squared distances to the nearest inside / outside cell, s+4 / s+5 scratch. Kernels: <kernel>_mark (the set and
    // whether it changed), <kernel>_scan (one axis of the squared distance transform), <kernel>_signed.
    // Poisson: Jacobi sweeps warm-started from the last answer, skipped when the source is what it was at the last
    // solve. Slots: s the answer, s+1 the other half of the ping-pong, s+2 the source (a cell source). Kernels:
    // <kernel>_mark, <kernel>_sweep, <kernel>_copy.
    // Both keep four words of state (words .. words+3 of the solver's stage words, zero at the start): [0] changed,
    // [1] / [2] Distance: some cell inside / outside, Poisson with a scalar source: its double's bits; [3] solved
    // once. A solve begins by clearing [0] (Distance: [0..2]) with k_zero.
    // NetHalf: a `half` net's op (mlp, flux, ldiffuse) on the matrix cores, a chunk of cells at a time: <kernel>_pack
    // writes each evaluation's normalised inputs as an fp16 row (rows_per_cell a cell: 1 for mlp; flux and ldiffuse
    // 6, the cell's +x, +y, +z faces both ways round), the library's MFMA forward turns rows into outputs, and
    // <kernel>_face makes the op's value (mlp) or each face's flux (into a 3-a-cell face buffer); for flux and
    // ldiffuse <kernel> then sums every cell's faces (the divergence).
    enum class Kind { Stored, Reduce, Tridiag, Distance, Poisson, NetHalf };
    Kind kind = Kind::Stored;
    int inst = -1;
    std::string kernel;     // the stage's kernel (Reduce: its per-block pass; the final pass is k_reduce_final)
    int axis = 0;           // Tridiag: the axis its lines run along
    bool uniform = false;   // Tridiag: a, b and u one number each, so c' is every line's: kept in LDS (see below)
    Op reduce = Op::SumOf;
    int iterations = 0;     // Poisson: sweeps a solve
    int words = -1;         // Distance, Poisson: the first of its four state words
    bool scalar_source = false; // Poisson: the source is one number (its _mark kernel runs as one thread)
    int net = -1;               // NetHalf: the net, its op, rows a cell, input row stride (halves), outputs a row
    Op net_op = Op::Mlp;
    int rows_per_cell = 0;
    int in_pad = 0;
    int outs = 0;
};

// k_bounds spreads its integer atomics over this many slots a field (block b uses slot b % kBoundSlots), so no single
// address takes every block's; the host folds the slots. Keys: [lo: field][slot], then [hi: field][slot].
constexpr int kBoundSlots = 32;

// A uniform Tridiag stage keeps c' in LDS for lines up to this long (the launch gives it 8 bytes a cell of the line);
// longer lines take its <kernel>_s, which keeps c' in the scratch with d'.
constexpr int kTridiagLdsLine = 4096;

// A window gather (z-marching tiles, off-cell field reads from an LDS ring of planes): launched as a 1-D grid of
// tiles_x * tiles_y * tiles_z blocks of 64 x block_y threads with (K, zchunk, tiles_x, tiles_y, tiles_z, mode: bit 0
// track changes, bit 1 spread consecutive tiles over the XCDs in runs); a tile is
// 64 * vx cells by ty rows by zchunk planes (at most kMaxZchunk: a lane keeps one bit a plane). vx = 4 needs nx a
// multiple of 4.
constexpr int kMaxZchunk = 64;
struct WindowKernel {
    std::string name;
    int vx = 1;
    int ty = 0;
    int block_y = 0;
    // A conduction term's flux-carrying window (vx = 4; preferred where nx is a multiple of 4): 1 with waves that only
    // load the halo rows, 2 with the halo rows loaded by the tile's edge rows; 0 the generic window.
    int flux = 0;
    int zchunk = 0; // planes a tile; 0: the program's
};

struct ProgramCode {
    std::string gather;     // empty: no equations. The direct gather: a thread a cell, blocks of 64 x 4 over (x, y),
                            // one z-plane a block row; always there.
    bool window = false;    // window gathers exist (then one of them runs)
    int radius = 0;
    int zchunk = 0;
    std::vector<WindowKernel> windows;
    bool in_place = false;
    std::vector<int> swaps; // ping-pong: the fields whose two buffers swap after the kernel
    std::vector<int> stages;
    std::vector<int> limit_stages;
    int max_dt = -1;
};

struct LawLayout {
    int fields = 0;
    int aux_slots = 0;      // float slots of A, one lattice each
    int uniforms = 0;       // doubles of U (reductions)
    int stage_words = 0;    // u32 state words of the Distance and Poisson stages, zero at the start
    std::vector<int> slot;  // per instruction: its slot in A, or -1
    std::vector<int> uslot; // per instruction: its value in U, or -1
    std::map<int, StageCode> stages;
    std::vector<ProgramCode> terms;
    std::vector<ProgramCode> learned;
    std::vector<ProgramCode> derives;
    int kernel_arg_bytes = 0; // sizeof the generated argument struct
};

// The translation unit of a law, or "" with `why` when it uses an op this backend does not run yet.
std::string generate(const ir::Module& m, const Law& law, LawLayout& layout, std::string& why);

// hipRTC options every generated source is compiled with (see the prelude for why each is there).
std::vector<std::string> compile_options();

// The law's nets as one float blob for the card, net after net: in_shift[inputs] (0 where not given),
// in_scale[inputs] (1 where not given), out_scale, then the weights (per layer W (out x in, row-major), b);
// offsets[n] is where net n starts.
std::vector<float> net_blob(const Law& law, std::vector<std::size_t>& offsets);

} // namespace fieldengine::field::hip
