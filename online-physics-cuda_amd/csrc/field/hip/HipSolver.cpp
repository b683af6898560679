// The HIP backend. See HipSolver.hpp.
#include "field/hip/HipSolver.hpp"

#include "fe_hip.h"
#include "field/LawBounds.hpp"
#include "field/hip/HipCodegen.hpp"
#include "field/ir/LawIR.hpp"
#include "kernels/net_mfma.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <set>
#include <mutex>
#include <sstream>
#include <stdexcept>

namespace fieldengine::field {

namespace {

// One process-wide lock around every backend's create/step/derive/fetch, as the reference's backends take
// (src/field/cuda/CudaSolver.cpp:40-44): they share one device and one stream.
std::mutex& device_mutex()
{
    static std::mutex m;
    return m;
}

void check(int status, const char* what)
{
    if (status != 0) {
        throw std::runtime_error(std::string(what) + ": " + (status > 0 ? fe_hip_error_string(status) : "invalid arguments"));
    }
}

struct Buffer {
    void* p = nullptr;
    std::size_t bytes = 0;
    Buffer() = default;
    Buffer(const Buffer&) = delete;
    Buffer& operator=(const Buffer&) = delete;
    ~Buffer()
    {
        if (p != nullptr) {
            fe_hip_mem_free(p);
        }
    }
    void allocate(std::size_t n)
    {
        if (p != nullptr) {
            fe_hip_mem_free(p);
            p = nullptr;
        }
        bytes = n;
        check(fe_hip_mem_alloc(&p, n == 0 ? 8 : n), "device allocation");
    }
    template <typename T>
    [[nodiscard]] T* as() const
    {
        return static_cast<T*>(p);
    }
};

constexpr unsigned kBlockX = 64;
constexpr unsigned kBlockY = 4;
constexpr int kReduceBlocks = 1024; // blocks of a reduction's first pass (its partials fold in one block)
constexpr int kTotals = 1024;      // doubles of conserved totals judged at once
constexpr int kFillWaves = 2048;    // two waves a SIMD of MI355X: a generic window's launch shortens its z-chunks to reach it
constexpr int kRanges = 1024;      // range checks judged at once (a record each: two counts and the first bad cell)
constexpr int kMaxWindowRegisters = 168; // a lane's registers that still leave a window gather three waves a SIMD

struct Impl final : HipBackend {
    Law* law = nullptr;
    GridShape shape;
    ir::Module module;
    hip::LawLayout layout;
    std::string name = "HIP";
    std::string adapter_name;
    fe_hip_module mod = nullptr;
    std::map<std::string, fe_hip_function> fns;
    mutable StepProfile prof;
    mutable bool timings_pending = false;
    std::size_t derive_event = 0;
    bool derive_timed = false;
    int check_every = 1;
    long long step_count = 0;
    bool lazy = false;
    std::vector<char> stale; // per field: the card holds newer cells than the host
    // What is known of each field's range, for skipping terms with nothing to do (the reference's rule: a term whose
    // every rate is exactly 0 is not dispatched, src/field/cuda/CudaSolver.cpp:586-595). `known`: at the start of the
    // next step, measured on the card (k_bounds) or carried; `flowing`: carried through the terms of the step in
    // progress (flow_bounds: as wide as the cells or wider, so a term with work is never skipped). Measured after
    // every eager step, and every kTightenEvery lazy steps; carried otherwise.
    std::vector<Bounds> known, flowing;
    long long lazy_steps = 0;
    static constexpr int kTightenEvery = 8;
    // Launch-shape overrides, read once from the environment at construction (for measuring shapes against each
    // other): FIELD_HIP_TILE="<cells a lane>x<rows>", FIELD_HIP_ZCHUNK=<planes a tile>, FIELD_HIP_DIRECT=1,
    // FIELD_HIP_REMAP=0|1 (tiles in XCD runs), FIELD_HIP_STREAM=0|1 (streaming stores).
    int tile_vx = 0, tile_ty = 0, zchunk_override = 0, remap_override = -1, stream_override = -1;
    int tile_flux = 0; // FIELD_HIP_TILE named a flux-carrying window: 1 "f4x..", 2 "h4x.."
    bool force_direct = false;
    std::string last_gather;
    std::vector<void*> own; // per field bound to a caller's buffer: its own array (else null)
    std::vector<std::string> gathers; // per term: the gather kernel its last substep launched
    bool tracking = true; // the gathers mark change flags (the step reads them back)

    // Device state.
    std::vector<std::unique_ptr<Buffer>> now_buf, next_buf;
    std::vector<std::unique_ptr<Buffer>> guard_copy; // a guarded term's fields before its learned scheme ran
    std::vector<void*> now, nxt;
    Buffer P, U, A, S, L, N, W;
    Buffer tile_sums;   // a flux window's per-tile totals on a contract step (2 doubles a tile)
    // The contract records, one allocation read back in one copy: range counts (two u32 a field), the first bad
    // cell (a u64 a field), then the conserved totals (kTotals doubles); `judged` the pinned host copy.
    struct View {
        void* p = nullptr;
        template <typename T>
        [[nodiscard]] T* as() const
        {
            return static_cast<T*>(p);
        }
    };
    // The front of that allocation holds the change flags (a u32 a field) and the bounds keys (two u32 a field), so that a
    // step's readback is one copy of a prefix: [flags][keys][records][firsts][totals].
    Buffer contract;
    View C, bounds_keys, records, firsts, totals;
    void* judged = nullptr;
    bool judged_has_flags = false; // the last copy into `judged` took the change flags and bounds keys as they are now
    // `half` nets on the matrix cores: their packed weights, and the row, output and face buffers their stages share
    // (a chunk of kHalfChunk cells at a time).
    static constexpr long long kHalfChunk = 1LL << 20;
    std::map<int, std::unique_ptr<hip::HalfNetCard>> half_nets;
    Buffer half_rows, half_outs, half_faces;
    Buffer spare, box_stage; // a moved window: the lattice a field slides into, the packed arriving box
    std::vector<float> net_host; // the law's nets as the kernels read them (HipCodegen.cpp: net_blob)
    std::vector<double> uploaded_params;
    bool params_valid = false;
    std::vector<fe_hip_event> events;

    ~Impl() override
    {
        drop_pairs();
        if (judged != nullptr) {
            fe_hip_host_free(judged);
        }
        for (fe_hip_event e : events) {
            fe_hip_event_destroy(e);
        }
        if (mod != nullptr) {
            fe_hip_module_unload(mod);
        }
    }

    [[nodiscard]] const std::string& adapter() const override { return adapter_name; }
    [[nodiscard]] const char* api() const override { return "HIP"; }
    [[nodiscard]] const StepProfile& profile() const override
    {
        resolve_timings();
        return prof;
    }
    void check_contracts_every(int steps) override { check_every = std::max(1, steps); }
    void set_lazy(bool on) override { lazy = on; }

    fe_hip_function fn(const std::string& n)
    {
        const auto it = fns.find(n);
        if (it != fns.end()) {
            return it->second;
        }
        fe_hip_function f = nullptr;
        check(fe_hip_module_function(mod, n.c_str(), &f), ("kernel " + n).c_str());
        fns.emplace(n, f);
        return f;
    }

    // The generated module's argument struct, by value (HipCodegen.cpp: struct K).
    std::vector<std::uint64_t> args(double dt) const
    {
        const int nf = std::max(1, static_cast<int>(law->fields.size()));
        const int np = std::max(1, static_cast<int>(law->param_values.size()));
        std::vector<std::uint64_t> k(static_cast<std::size_t>(2 * nf + 6 + 6 + 2 + np), 0);
        std::size_t w = 0;
        for (int f = 0; f < nf; ++f) {
            k[w++] = reinterpret_cast<std::uint64_t>(f < static_cast<int>(now.size()) ? now[static_cast<std::size_t>(f)] : nullptr);
        }
        for (int f = 0; f < nf; ++f) {
            k[w++] = reinterpret_cast<std::uint64_t>(f < static_cast<int>(nxt.size()) ? nxt[static_cast<std::size_t>(f)] : nullptr);
        }
        k[w++] = reinterpret_cast<std::uint64_t>(U.p);
        k[w++] = reinterpret_cast<std::uint64_t>(A.p);
        k[w++] = reinterpret_cast<std::uint64_t>(S.p);
        k[w++] = reinterpret_cast<std::uint64_t>(L.p);
        k[w++] = reinterpret_cast<std::uint64_t>(C.p);
        k[w++] = reinterpret_cast<std::uint64_t>(N.p);
        const long long nxy = static_cast<long long>(shape.nx) * shape.ny;
        const long long cells = nxy * shape.nz;
        std::memcpy(&k[w++], &nxy, 8);
        std::memcpy(&k[w++], &cells, 8);
        std::memcpy(&k[w++], &shape.cell, 8);
        std::memcpy(&k[w++], &dt, 8);
        // 1 / h and 1 / h^2, used only where h (h^2) is a power of two: then a product with them is the quotient's bits.
        int e = 0;
        const bool h_pow2 = std::frexp(shape.cell, &e) == 0.5;
        const double hh = shape.cell * shape.cell;
        const bool hh_pow2 = std::frexp(hh, &e) == 0.5;
        const double inv_h = h_pow2 ? 1.0 / shape.cell : 0.0;
        const double inv_hh = hh_pow2 ? 1.0 / hh : 0.0;
        std::memcpy(&k[w++], &inv_h, 8);
        std::memcpy(&k[w++], &inv_hh, 8);
        const int ints[4] = {shape.nx, shape.ny, shape.nz, (h_pow2 ? 1 : 0) | (hh_pow2 ? 2 : 0)};
        std::memcpy(&k[w], ints, 16);
        w += 2;
        for (std::size_t p = 0; p < law->param_values.size(); ++p) {
            std::memcpy(&k[w++], &law->param_values[p], 8); // the params by value (struct K's last member)
        }
        return k;
    }

    void launch_cells(const std::string& kernel, double dt, std::vector<void*> extra = {})
    {
        std::vector<std::uint64_t> k = args(dt);
        std::vector<void*> argv;
        argv.push_back(k.data());
        argv.insert(argv.end(), extra.begin(), extra.end());
        const unsigned gx = (static_cast<unsigned>(shape.nx) + kBlockX - 1) / kBlockX;
        const unsigned gy = (static_cast<unsigned>(shape.ny) + kBlockY - 1) / kBlockY;
        check(fe_hip_launch(fn(kernel), gx, gy, static_cast<unsigned>(shape.nz), kBlockX, kBlockY, 1, 0, argv.data()),
            ("launch " + kernel).c_str());
        ++prof.launches;
    }

    // A conserved field's "after" total and range taken by a flux-carrying window as it writes the field (the
    // contract step's last gather): where the tile sums go, and the range record and the field's bounds.
    struct FusedTotal {
        int field = -1;
        int record = -1;
        long long tiles = 0;
    };

    // The flux-carrying window the gather would launch for `code`, when it can take the totals of `field` on the way:
    // the program writes that field alone, and the field is not clamped (a clamp would have to come between the total
    // and the store). nullptr otherwise.
    const hip::WindowKernel* fusable(const ir::Program& p, const hip::ProgramCode& code, int field) const
    {
        const hip::WindowKernel* w = window_for(code);
        if (w == nullptr || w->flux == 0 || p.writes.size() != 1 || p.writes[0] != field || law->fields[static_cast<std::size_t>(field)].clamp) {
            return nullptr;
        }
        return w;
    }

    // A term's gather: a thread a cell, or (window kernels) z-marching tiles of 64 x ty cells. With `fused` (a flux
    // window the caller checked with fusable) the gather also leaves the tiles' totals in tile_sums and counts the
    // range into its record.
    void launch_gather(const hip::ProgramCode& code, double dt, FusedTotal* fused = nullptr)
    {
        const hip::WindowKernel* w = window_for(code);
        if (w == nullptr) {
            // the direct gather, with 32-bit reads while the lattice allows them
            last_gather = shape.cells() < (std::size_t{1} << 31) ? code.gather + "_32" : code.gather;
            launch_cells(last_gather, dt);
            return;
        }
        last_gather = w->name;
        int zchunk = zchunk_override > 0 ? zchunk_override : (w->zchunk > 0 ? w->zchunk : code.zchunk);
        int tx = (shape.nx + 64 * w->vx - 1) / (64 * w->vx);
        int ty = (shape.ny + w->ty - 1) / w->ty;
        if (w->flux == 0 && zchunk_override <= 0) {
            // A flat lattice (a few rows of tiles, e.g. ny = 1) leaves few tiles: shorter z-chunks, until the card has
            // two waves a SIMD (or the chunks are 4 planes).
            while (zchunk > 4 && 1LL * tx * ty * ((shape.nz + zchunk - 1) / zchunk) * w->block_y < kFillWaves) {
                zchunk /= 2;
            }
        }
        int tz = (shape.nz + zchunk - 1) / zchunk;
        const long long total = 1LL * tx * ty * tz;
        if (total > 0x7fffffffLL) {
            throw std::runtime_error("lattice too large for one launch");
        }
        std::vector<std::uint64_t> k = args(dt);
        const bool fits = fits_llc();
        const bool remap = remap_override >= 0 ? remap_override != 0 : fits;
        const bool stream = stream_override >= 0 ? stream_override != 0 : !fits;
        int mode = (tracking ? 1 : 0) | (remap ? 2 : 0) | (stream ? 4 : 0);
        if (w->flux == 0) {
            void* argv[] = {k.data(), &zchunk, &tx, &ty, &tz, &mode};
            check(fe_hip_launch(fn(w->name), static_cast<unsigned>(total), 1, 1, 64, static_cast<unsigned>(w->block_y), 1, 0, argv),
                ("launch " + w->name).c_str());
            ++prof.launches;
            return;
        }
        // Flux windows: rows spanning several tiles take the kernel that loads halo columns; a contract step's may take
        // the totals.
        double* tsum = nullptr;
        unsigned* rec = nullptr;
        unsigned long long* first = nullptr;
        double lo = 0.0, hi = 0.0;
        if (fused != nullptr) {
            if (tile_sums.bytes < static_cast<std::size_t>(total) * 16) {
                tile_sums.allocate(static_cast<std::size_t>(total) * 16);
            }
            tsum = tile_sums.as<double>();
            rec = records.as<unsigned>() + 2 * fused->record;
            first = firsts.as<unsigned long long>() + fused->record;
            const FieldDecl& d = law->fields[static_cast<std::size_t>(fused->field)];
            lo = d.lo;
            hi = d.hi;
            fused->tiles = total;
        }
        const std::string name = w->name + (tx > 1 ? "x" : "") + (fused != nullptr ? "_t" : "");
        void* argv[] = {k.data(), &zchunk, &tx, &ty, &tz, &mode, &tsum, &rec, &first, &lo, &hi};
        check(fe_hip_launch(fn(name), static_cast<unsigned>(total), 1, 1, 64, static_cast<unsigned>(w->block_y), 1, 0, argv),
            ("launch " + name).c_str());
        ++prof.launches;
    }

    // Whether the fields a step touches (two copies of each) fit the 256 MB last-level cache. While they do, window
    // tiles go out in XCD runs (block b on XCD b % 8 takes the b / 8-th tile of that XCD's run, so neighbouring tiles
    // share an L2) and stores are ordinary; past it, tiles in launch order and streaming (nontemporal) stores. Measured
    // on MI355X, conduction: 256^3 32.7 us in runs against 38.4 in launch order, streaming stores 34-35 against 33;
    // 512^3 235 us in launch order against 240 in runs, streaming stores 225 against 231.
    bool fits_llc() const
    {
        return 8.0 * static_cast<double>(shape.nx) * shape.ny * shape.nz * layout.fields <= 256.0 * 1024 * 1024;
    }

    // The window gather for this lattice. Where rows are whole float4s: a conduction term's flux-carrying window (the
    // form with halo rows loaded by the edge rows while the fields fit the last-level cache, the form with halo waves
    // past it), else four cells a lane; else a cell a lane. Among windows of one kind, the tallest tile that still
    // leaves at least 512 tiles (two per CU) to spread (measured on MI355X, the generic window at z-chunks of 32: 256^3
    // 29.1 us with 256x4 tiles against 31.2 with 256x2 and 39.9 with 256x6; 512^3 219 us with 256x6 against 251 and
    // 253). FIELD_HIP_TILE (e.g. "4x2", "1x6", "f4x4" or "h4x8" for the two flux-carrying forms) picks one. None: the
    // direct gather.
    const hip::WindowKernel* window_for(const hip::ProgramCode& code) const
    {
        // A lattice one row high leaves a window's other rows (and its halo-row waves) idle: the direct gather (measured
        // on MI355X, Black-Scholes 4096 x 1 x 512: 16.8 us a pass against 23.8 through 256x2 tiles).
        if (!code.window || code.windows.empty() || force_direct || (shape.ny == 1 && tile_vx == 0)) {
            return nullptr;
        }
        const hip::WindowKernel* best = nullptr;
        const bool cached = fits_llc();
        auto rank = [&](const hip::WindowKernel& w) {
            if (w.flux != 0) {
                return (w.flux == 2) == cached ? 3 : 2;
            }
            return w.vx == 4 ? 1 : 0;
        };
        // A window's tile is 64 vx cells wide; a row that fills its tiles badly (144 cells in a 256-wide tile, or in three
        // 64-wide ones) leaves lanes idle in every plane of every tile, and the direct gather was faster there
        // (the valley law on 144 x 144 x 48: 0.66 ms a step against 0.82 through 256x2 and 64x6 tiles), so a window is
        // taken only when the row fills at least 90% of its tiles' width.
        auto fills = [&](const hip::WindowKernel& w) {
            const long long width = 64LL * w.vx, tiles_x = (shape.nx + width - 1) / width;
            return 10LL * shape.nx >= 9LL * width * tiles_x;
        };
        for (const hip::WindowKernel& w : code.windows) {
            if (w.vx == 4 && shape.nx % 4 != 0) {
                continue;
            }
            if (tile_vx > 0) {
                if (w.vx == tile_vx && w.ty == tile_ty && w.flux == tile_flux) {
                    return &w;
                }
                continue;
            }
            if (heavy.count(w.name) != 0 || !fills(w)) {
                continue;
            }
            const int zc = zchunk_override > 0 ? zchunk_override : (w.zchunk > 0 ? w.zchunk : code.zchunk);
            const long long tz = (shape.nz + zc - 1) / zc;
            const long long tiles = 1LL * ((shape.nx + 64 * w.vx - 1) / (64 * w.vx)) * ((shape.ny + w.ty - 1) / w.ty) * tz;
            if (best == nullptr || rank(w) > rank(*best) || (rank(w) == rank(*best) && w.ty > best->ty && tiles >= 512)) {
                best = &w;
            }
        }
        return best;
    }

    void read_overrides()
    {
        if (const char* t = std::getenv("FIELD_HIP_TILE")) {
            tile_flux = t[0] == 'f' ? 1 : (t[0] == 'h' ? 2 : 0);
            if (std::sscanf(t + (tile_flux != 0 ? 1 : 0), "%dx%d", &tile_vx, &tile_ty) != 2) {
                tile_vx = tile_ty = 0;
                tile_flux = 0;
            }
        }
        if (const char* z = std::getenv("FIELD_HIP_ZCHUNK")) {
            zchunk_override = std::min(std::max(0, std::atoi(z)), hip::kMaxZchunk);
        }
        if (const char* m = std::getenv("FIELD_HIP_REMAP")) {
            remap_override = std::atoi(m) != 0 ? 1 : 0;
        }
        if (const char* m = std::getenv("FIELD_HIP_STREAM")) {
            stream_override = std::atoi(m) != 0 ? 1 : 0;
        }
        if (const char* d = std::getenv("FIELD_HIP_DIRECT")) {
            force_direct = std::atoi(d) != 0;
        }
    }

    void build(std::string& why)
    {
        read_overrides();
        module = ir::lower(*law);
        const std::string src = hip::generate(module, *law, layout, why);
        if (src.empty()) {
            throw std::runtime_error(why);
        }
        std::vector<std::string> opts = hip::compile_options();
        std::vector<const char*> argv;
        for (const std::string& o : opts) {
            argv.push_back(o.c_str());
        }
        void* code = nullptr;
        std::size_t bytes = 0;
        char* log = nullptr;
        int cached = 0;
        const int r = fe_hip_compile(src.c_str(), "law.hip", argv.data(), static_cast<int>(argv.size()), &code, &bytes, &log, &cached);
        const std::string log_text = log != nullptr ? log : "";
        fe_hip_free(log);
        if (r != 0) {
            throw std::runtime_error("the generated source does not compile: " + log_text);
        }
        const int lr = fe_hip_module_load(code, bytes, &mod);
        fe_hip_free(code);
        check(lr, "module load");
        weigh_windows();
        allocate();
    }

    // Window gathers that would run badly: private (scratch) memory, or more registers a lane than leave three waves a
    // SIMD (four cells a lane of a law with many fields). The automatic choice passes them over.
    std::set<std::string> heavy;
    void weigh_windows()
    {
        heavy.clear();
        for (const auto* progs : {&layout.terms, &layout.learned}) {
            for (const hip::ProgramCode& p : *progs) {
                for (const hip::WindowKernel& w : p.windows) {
                    int regs = 0, priv = 0;
                    if (fe_hip_function_attributes(fn(w.name), &regs, &priv) == 0 && (priv > 0 || regs > kMaxWindowRegisters)) {
                        heavy.insert(w.name);
                    }
                }
            }
        }
    }

    void allocate()
    {
        const std::size_t cells = shape.cells();
        const std::size_t nf = law->fields.size();
        own.clear(); // bindings end with the arrays they stood in for
        std::vector<char> swapped(nf, 0);
        for (const auto* progs : {&layout.terms, &layout.learned, &layout.derives}) {
            for (const hip::ProgramCode& p : *progs) {
                for (const int f : p.swaps) {
                    swapped[static_cast<std::size_t>(f)] = 1;
                }
            }
        }
        now_buf.clear();
        next_buf.clear();
        now.assign(nf, nullptr);
        nxt.assign(nf, nullptr);
        for (std::size_t f = 0; f < nf; ++f) {
            now_buf.push_back(std::make_unique<Buffer>());
            now_buf.back()->allocate(cells * 4);
            now[f] = now_buf.back()->p;
            next_buf.push_back(std::make_unique<Buffer>());
            if (swapped[f]) {
                next_buf.back()->allocate(cells * 4);
                nxt[f] = next_buf.back()->p;
            }
        }
        P.allocate(std::max<std::size_t>(1, law->param_values.size()) * 8);
        U.allocate(static_cast<std::size_t>(std::max(1, layout.uniforms)) * 8);
        A.allocate(static_cast<std::size_t>(std::max(1, layout.aux_slots)) * cells * 4);
        bool tridiag = false;
        for (const auto& [id, s] : layout.stages) {
            tridiag = tridiag || s.kind == hip::StageCode::Kind::Tridiag;
        }
        S.allocate(tridiag ? 2 * cells * 8 : 8);
        L.allocate(static_cast<std::size_t>(3 * kReduceBlocks) * 8);
        std::vector<std::size_t> net_offsets;
        net_host = hip::net_blob(*law, net_offsets);
        N.allocate(std::max<std::size_t>(1, net_host.size()) * 4);
        if (!net_host.empty()) {
            check(fe_hip_memcpy_htod_async(N.p, net_host.data(), net_host.size() * 4), "nets upload");
        }
        nets_key = nets_hash(*law);
        // The Distance and Poisson stages' state words start at zero, and a Poisson stage's answer (its warm start) too.
        W.allocate(static_cast<std::size_t>(std::max(1, layout.stage_words)) * 4);
        check(fe_hip_memset_d32_async(W.p, 0u, static_cast<std::size_t>(std::max(1, layout.stage_words))), "stage words reset");
        for (const auto& [id, s] : layout.stages) {
            if (s.kind == hip::StageCode::Kind::Poisson) {
                char* answer = A.as<char>() + static_cast<std::size_t>(layout.slot[static_cast<std::size_t>(id)]) * cells * 4;
                check(fe_hip_memset_d32_async(answer, 0u, cells), "poisson warm start reset");
            }
        }
        std::size_t row_bytes = 0, out_bytes = 0;
        bool faces = false;
        half_nets.clear();
        for (const auto& [id, st] : layout.stages) {
            if (st.kind != hip::StageCode::Kind::NetHalf) {
                continue;
            }
            if (half_nets.count(st.net) == 0) {
                half_nets.emplace(st.net, hip::half_net_create(law->nets[static_cast<std::size_t>(st.net)]));
            }
            const std::size_t chunk = static_cast<std::size_t>(std::min<long long>(kHalfChunk, static_cast<long long>(cells)));
            row_bytes = std::max(row_bytes, chunk * static_cast<std::size_t>(st.rows_per_cell * st.in_pad) * 2);
            out_bytes = std::max(out_bytes, chunk * static_cast<std::size_t>(st.rows_per_cell * st.outs) * 4);
            faces = faces || st.net_op != Op::Mlp;
        }
        half_rows.allocate(row_bytes);
        half_outs.allocate(out_bytes);
        half_faces.allocate(faces ? cells * 3 * 4 : 0);
        const std::size_t rec_bytes = static_cast<std::size_t>(kRanges) * 8, first_bytes = static_cast<std::size_t>(kRanges) * 8;
        const std::size_t flag_bytes = (std::max<std::size_t>(1, nf) * 4 + 7) / 8 * 8, key_bytes = std::max<std::size_t>(1, nf) * 8 * hip::kBoundSlots;
        contract.allocate(flag_bytes + key_bytes + rec_bytes + first_bytes + static_cast<std::size_t>(kTotals) * 8);
        C.p = contract.p;
        bounds_keys.p = static_cast<char*>(contract.p) + flag_bytes;
        records.p = static_cast<char*>(contract.p) + flag_bytes + key_bytes;
        firsts.p = static_cast<char*>(records.p) + rec_bytes;
        totals.p = static_cast<char*>(firsts.p) + first_bytes;
        if (judged != nullptr) {
            fe_hip_host_free(judged);
            judged = nullptr;
        }
        check(fe_hip_host_alloc(&judged, contract.bytes), "pinned allocation");

        stale.assign(nf, 0);
        known.assign(nf, Bounds{});
    }

    // A trainer may write a law's net weights between steps (Solver::law_for_training, fe_solver_set_net_weights):
    // their bits are hashed every step, as the reference's backend does (src/field/cuda/CudaSolver.cpp:282-311), and
    // a change goes up again, in place (the same addresses, so captured graphs stay valid).
    std::uint64_t nets_key = 0;
    static std::uint64_t nets_hash(const Law& l)
    {
        std::uint64_t key = 1469598103934665603ull;
        auto mix = [&](const std::vector<float>& v) {
            for (const float w : v) {
                std::uint32_t bits = 0;
                std::memcpy(&bits, &w, 4);
                key = (key ^ bits) * 1099511628211ull;
            }
        };
        for (const NetDecl& net : l.nets) {
            mix(net.weights);
            mix(net.in_shift);
            mix(net.in_scale);
            mix({net.out_scale});
        }
        return key;
    }
    void upload_nets()
    {
        if (law->nets.empty()) {
            return;
        }
        const std::uint64_t key = nets_hash(*law);
        if (key == nets_key) {
            return;
        }
        std::vector<std::size_t> offsets;
        std::vector<float> blob = hip::net_blob(*law, offsets);
        if (blob.size() != net_host.size()) {
            throw std::runtime_error("a net changed its shape after the law was built");
        }
        net_host = std::move(blob);
        check(fe_hip_memcpy_htod_async(N.p, net_host.data(), net_host.size() * 4), "nets upload");
        for (auto& [n, card] : half_nets) {
            hip::half_net_update(*card, law->nets[static_cast<std::size_t>(n)]);
        }
        check(fe_hip_stream_sync(), "sync"); // the host copies the uploads read live until here
        nets_key = key;
        drop_pairs(); // captured launches carry a half net's out_scale by value
    }

    void upload_params()
    {
        upload_nets();
        if (params_valid && uploaded_params == law->param_values) {
            return;
        }
        if (!law->param_values.empty()) {
            check(fe_hip_memcpy_htod_async(P.p, law->param_values.data(), law->param_values.size() * 8), "params upload");
        }
        uploaded_params = law->param_values;
        params_valid = true;
    }

    // The window moved (edits.shift, edits.boxes; src/field/cuda/CudaSolver.cpp:895-920 keeps a ring and moves its
    // origin): here every field's cells slide on the card into a spare lattice and back (the kernels keep plain
    // row-major indexing), then only the arriving boxes go up, packed, and scatter into place. Captured step graphs are
    // dropped. The stages' caches (a poisson warm start, the last set or source they compare with) stay where they are,
    // as the CPU's caches in the law do: a moved set or source differs from the one held, and is solved again.
    void slide(FieldSet& fields, const FieldEdits& edits)
    {
        const std::size_t cells = shape.cells();
        std::fill(known.begin(), known.end(), Bounds{}); // measured again after the step
        if (edits.shift[0] != 0 || edits.shift[1] != 0 || edits.shift[2] != 0) {
            if (spare.p == nullptr) {
                spare.allocate(cells * 4);
            }
            for (std::size_t f = 0; f < law->fields.size(); ++f) {
                void* src = now[f];
                void* dst = spare.p;
                int dx = edits.shift[0], dy = edits.shift[1], dz = edits.shift[2];
                launch_cells("k_slide", 0.0, {&src, &dst, &dx, &dy, &dz});
                check(fe_hip_memcpy_dtod_async(now[f], spare.p, cells * 4), "slide");
            }
        }
        for (const CellBox& b0 : edits.boxes) {
            const CellBox b{std::max(0, b0.x0), std::max(0, b0.y0), std::max(0, b0.z0), std::min(shape.nx, b0.x1),
                std::min(shape.ny, b0.y1), std::min(shape.nz, b0.z1)};
            int bx = b.x1 - b.x0, by = b.y1 - b.y0, bz = b.z1 - b.z0;
            if (bx <= 0 || by <= 0 || bz <= 0) {
                continue;
            }
            const std::size_t n = static_cast<std::size_t>(bx) * by * bz;
            if (box_stage.bytes < n * 4) {
                box_stage.allocate(n * 4);
            }
            std::vector<float> packed(n);
            for (std::size_t f = 0; f < law->fields.size(); ++f) {
                const std::vector<float>& v = fields.values(static_cast<int>(f));
                std::size_t q = 0;
                for (int z = b.z0; z < b.z1; ++z) {
                    for (int y = b.y0; y < b.y1; ++y) {
                        const std::size_t row = (static_cast<std::size_t>(z) * shape.ny + y) * shape.nx;
                        for (int x = b.x0; x < b.x1; ++x) {
                            const float c = v[row + static_cast<std::size_t>(x)];
                            packed[q++] = c != 0.0f && std::fabs(c) < 1.17549435e-38f ? std::copysign(0.0f, c) : c;
                        }
                    }
                }
                check(fe_hip_memcpy_htod_async(box_stage.p, packed.data(), n * 4), "box upload");
                // the host buffer is reused for the next field: wait for the copy
                check(fe_hip_stream_sync(), "sync");
                int field = static_cast<int>(f), x0 = b.x0, y0 = b.y0, z0 = b.z0;
                void* src = box_stage.p;
                std::vector<std::uint64_t> k = args(0.0);
                void* argv[] = {k.data(), &field, &src, &x0, &y0, &z0, &bx, &by, &bz};
                const unsigned blocks = static_cast<unsigned>((n + 255) / 256);
                check(fe_hip_launch(fn("k_scatter"), blocks, 1, 1, 256, 1, 1, 0, argv), "launch k_scatter");
                ++prof.launches;
                prof.upload_bytes += n * 4;
            }
        }
        for (std::size_t f = 0; f < law->fields.size(); ++f) {
            touched(static_cast<int>(f));
        }
        drop_pairs();
    }

    void upload(FieldSet& fields, FieldEdits& edits)
    {
        const std::size_t cells = shape.cells();
        if (!edits.all && edits.shifted()) {
            slide(fields, edits);
        }
        for (std::size_t f = 0; f < law->fields.size(); ++f) {
            bool dirty = edits.all;
            if (!dirty && f < edits.planes.size()) {
                dirty = std::any_of(edits.planes[f].begin(), edits.planes[f].end(), [](char c) { return c != 0; });
            }
            if (!dirty) {
                continue;
            }
            // The card only ever holds flushed floats: what the CPU reads of a denormal is its zero (DAZ). The same pass
            // takes the field's range (NaN skipped).
            std::vector<float>& v = fields.values(static_cast<int>(f));
            float lo = std::numeric_limits<float>::infinity(), hi = -std::numeric_limits<float>::infinity();
            for (float& x : v) {
                if (x != 0.0f && std::fabs(x) < 1.17549435e-38f) {
                    x = std::copysign(0.0f, x);
                }
                lo = x < lo ? x : lo;
                hi = hi < x ? x : hi;
            }
            known[f] = Bounds{lo, hi};
            check(fe_hip_memcpy_htod_async(now[f], v.data(), cells * 4), "field upload");
            prof.upload_bytes += cells * 4;
            stale[f] = 0;
            touched(static_cast<int>(f));
        }
        edits.all = false;
        edits.planes.assign(law->fields.size(), std::vector<char>(static_cast<std::size_t>(shape.nz), 0));
        edits.boxes.clear();
        edits.shift[0] = edits.shift[1] = edits.shift[2] = 0;
    }

    // A scalar instruction on the host, in double, with reductions read from U.
    double scalar(int id, const std::vector<double>& u) const
    {
        const ir::Inst& in = module.insts[static_cast<std::size_t>(id)];
        auto a = [&](std::size_t i) { return scalar(in.args[i], u); };
        switch (in.op) {
        case Op::Const: return in.number;
        case Op::Param: return law->param_values[static_cast<std::size_t>(in.index)];
        case Op::Cell: return shape.cell;
        case Op::Dt: return 0.0;
        case Op::Local: return 0.0;
        case Op::MaxOf: case Op::MinOf: case Op::SumOf: {
            const int s = layout.uslot[static_cast<std::size_t>(id)];
            return s >= 0 ? u[static_cast<std::size_t>(s)] : 0.0;
        }
        case Op::Neg: return -a(0);
        case Op::Abs: return std::abs(a(0));
        case Op::Sqrt: { const double q = a(0); return std::sqrt(0.0 < q ? q : 0.0); }
        case Op::Exp: { const double q = a(0); return std::exp(80.0 < q ? 80.0 : q); }
        case Op::Log: { const double q = a(0); return std::log(q < 1e-30 ? 1e-30 : q); }
        case Op::Step: return a(0) > 0.0 ? 1.0 : 0.0;
        case Op::Pos: { const double q = a(0); return q < 0.0 ? 0.0 : q; }
        case Op::Add: return a(0) + a(1);
        case Op::Sub: return a(0) - a(1);
        case Op::Mul: { const double p = a(0); if (p == 0.0) return 0.0; const double q = a(1); return q == 0.0 ? 0.0 : p * q; }
        case Op::Div: { const double q = a(1); return q == 0.0 ? 0.0 : a(0) / q; }
        case Op::Pow: return std::pow(a(0), a(1));
        case Op::Min: { const double p = a(0), q = a(1); return q < p ? q : p; }
        case Op::Max: { const double p = a(0), q = a(1); return p < q ? q : p; }
        case Op::Clamp: {
            const double v = a(0), lo = a(1), h = a(2);
            const double hi = lo < h ? h : lo;
            return v < lo ? lo : (hi < v ? hi : v);
        }
        default: return 0.0;
        }
    }

    void run_stage(int id, double dt)
    {
        const hip::StageCode& s = layout.stages.at(id);
        switch (s.kind) {
        case hip::StageCode::Kind::Stored:
            launch_cells(s.kernel, dt);
            break;
        case hip::StageCode::Kind::Tridiag: {
            const long long lines = s.axis == 0 ? 1LL * shape.ny * shape.nz : (s.axis == 1 ? 1LL * shape.nx * shape.nz : 1LL * shape.nx * shape.ny);
            std::vector<std::uint64_t> k = args(dt);
            void* argv[] = {k.data()};
            // A uniform stage keeps c' (8 bytes a cell of a line) in LDS when the line fits; <kernel>_s otherwise.
            const int n = s.axis == 0 ? shape.nx : (s.axis == 1 ? shape.ny : shape.nz);
            const bool lds = s.uniform && n <= hip::kTridiagLdsLine;
            check(fe_hip_launch(fn(s.uniform && !lds ? s.kernel + "_s" : s.kernel), static_cast<unsigned>((lines + 63) / 64), 1, 1, 64,
                      1, 1, lds ? static_cast<unsigned>(n) * 8u : 0u, argv), "tridiag launch");
            ++prof.launches;
            break;
        }
        case hip::StageCode::Kind::Distance: {
            unsigned* w = W.as<unsigned>() + s.words;
            const int slot = layout.slot[static_cast<std::size_t>(id)];
            int clear = 3;
            void* zargs[] = {&w, &clear};
            check(fe_hip_launch(fn("k_zero"), 1, 1, 1, 64, 1, 1, 0, zargs), "distance reset");
            launch_cells(s.kernel + "_mark", dt, {&w});
            // x from the set into (s+2, s+3), y into (s+4, s+5), z back into (s+2, s+3)
            const int passes[3][4] = {{-1, -1, slot + 2, slot + 3}, {slot + 2, slot + 3, slot + 4, slot + 5}, {slot + 4, slot + 5, slot + 2, slot + 3}};
            for (int axis = 0; axis < 3; ++axis) {
                int a = axis, fi = passes[axis][0], fo = passes[axis][1], ti = passes[axis][2], to = passes[axis][3];
                // A line and set a thread (the envelopes' stacks, 12 bytes a cell of a line for both sets, in LDS) while
                // a block's lines fit 64 KB and the squared distances stay below 2^24; the outward search from every
                // cell otherwise.
                const int n = axis == 0 ? shape.nx : (axis == 1 ? shape.ny : shape.nz);
                int lines_a_block = 64;
                while (lines_a_block > 1 && 12LL * lines_a_block * n > 65536) {
                    lines_a_block /= 2;
                }
                if (n <= 2048 && 12LL * lines_a_block * n <= 65536) {
                    std::vector<std::uint64_t> k = args(dt);
                    void* argv[] = {k.data(), &w, &a, &fi, &fo, &ti, &to};
                    const long long lines = static_cast<long long>(shape.cells()) / n;
                    check(fe_hip_launch(fn(s.kernel + "_lines"), static_cast<unsigned>((lines + lines_a_block - 1) / lines_a_block), 1, 1,
                              static_cast<unsigned>(2 * lines_a_block), 1, 1, static_cast<unsigned>(12 * lines_a_block * n), argv),
                        "distance lines launch");
                    ++prof.launches;
                } else {
                    launch_cells(s.kernel + "_scan", dt, {&w, &a, &fi, &fo, &ti, &to});
                }
            }
            launch_cells(s.kernel + "_signed", dt, {&w});
            ++prof.launches;
            break;
        }
        case hip::StageCode::Kind::Poisson: {
            unsigned* w = W.as<unsigned>() + s.words;
            const int slot = layout.slot[static_cast<std::size_t>(id)];
            int clear = 1;
            void* zargs[] = {&w, &clear};
            check(fe_hip_launch(fn("k_zero"), 1, 1, 1, 64, 1, 1, 0, zargs), "poisson reset");
            if (s.scalar_source) {
                std::vector<std::uint64_t> k = args(dt);
                void* margs[] = {k.data(), &w};
                check(fe_hip_launch(fn(s.kernel + "_mark"), 1, 1, 1, 1, 1, 1, 0, margs), "poisson mark");
                ++prof.launches;
            } else {
                launch_cells(s.kernel + "_mark", dt, {&w});
            }
            int from = slot, to = slot + 1;
            for (int it = 0; it < s.iterations; ++it) {
                launch_cells(s.kernel + "_sweep", dt, {&w, &from, &to});
                std::swap(from, to);
            }
            launch_cells(s.kernel + "_copy", dt, {&w, &from});
            ++prof.launches;
            break;
        }
        case hip::StageCode::Kind::NetHalf: {
            const long long cells = static_cast<long long>(shape.cells());
            const hip::HalfNetCard& net = *half_nets.at(s.net);
            void* rows = half_rows.p;
            float* outs = half_outs.as<float>();
            float* faces = half_faces.as<float>();
            std::vector<std::uint64_t> k = args(dt);
            for (long long c0 = 0; c0 < cells; c0 += kHalfChunk) {
                long long count = std::min(kHalfChunk, cells - c0);
                const unsigned blocks = static_cast<unsigned>((count + 255) / 256);
                void* pargs[] = {k.data(), &rows, &c0, &count};
                check(fe_hip_launch(fn(s.kernel + "_pack"), blocks, 1, 1, 256, 1, 1, 0, pargs), "net pack launch");
                check(hip::half_net_run(net, rows, count * s.rows_per_cell, outs), "half net launch");
                void* fargs[] = {k.data(), &outs, &faces, &c0, &count};
                check(fe_hip_launch(fn(s.kernel + "_face"), blocks, 1, 1, 256, 1, 1, 0, fargs), "net face launch");
                prof.launches += 3;
            }
            if (s.net_op != Op::Mlp) {
                launch_cells(s.kernel, dt, {&faces});
            }
            break;
        }
        case hip::StageCode::Kind::Reduce: {
            const long long cells = static_cast<long long>(shape.cells());
            long long chunk = (cells + kReduceBlocks - 1) / kReduceBlocks;
            chunk = std::max(1LL, chunk);
            const int blocks = static_cast<int>((cells + chunk - 1) / chunk);
            std::vector<std::uint64_t> k = args(dt);
            void* argv1[] = {k.data(), &chunk};
            check(fe_hip_launch(fn(s.kernel), static_cast<unsigned>(blocks), 1, 1, 256, 1, 1, 0, argv1), "reduce launch");
            int kind = s.reduce == Op::SumOf ? 0 : (s.reduce == Op::MaxOf ? 1 : 2);
            int parts = blocks;
            int uslot = layout.uslot[static_cast<std::size_t>(id)];
            void* argv2[] = {k.data(), &kind, &parts, &uslot};
            check(fe_hip_launch(fn("k_reduce_final"), 1, 1, 1, 256, 1, 1, 0, argv2), "reduce final launch");
            prof.launches += 2;
            break;
        }
        }
    }

    void swap(const std::vector<int>& fields)
    {
        for (const int f : fields) {
            std::swap(now[static_cast<std::size_t>(f)], nxt[static_cast<std::size_t>(f)]);
        }
    }

    // Which fields may have changed since: a counter a field, bumped by every write to it (gathers, derives, uploads,
    // device writes, restores, clamps, replays). A conserved total taken at a generation stands for the field until
    // the next bump, so a judged step's "before" total is the last judged step's "after" one, copied.
    std::vector<long long> generation;
    struct Cached {
        int slot;
        long long gen;
    };
    std::map<int, Cached> cached_totals;

    void touched(int f)
    {
        if (generation.size() <= static_cast<std::size_t>(f)) {
            generation.resize(static_cast<std::size_t>(f) + 1, 0);
        }
        ++generation[static_cast<std::size_t>(f)];
    }
    long long gen_of(int f)
    {
        touched(f);
        --generation[static_cast<std::size_t>(f)];
        return generation[static_cast<std::size_t>(f)];
    }
    void note_total(int field, int slot)
    {
        for (auto it = cached_totals.begin(); it != cached_totals.end();) {
            it = it->second.slot == slot && it->first != field ? cached_totals.erase(it) : std::next(it);
        }
        cached_totals[field] = Cached{slot, gen_of(field)};
    }
    // The total of a field into totals[slot], from the last one taken when the field has not changed since.
    void total_before(int field, int slot)
    {
        const auto it = cached_totals.find(field);
        if (it != cached_totals.end() && it->second.gen == gen_of(field)) {
            if (it->second.slot != slot) {
                check(fe_hip_memcpy_dtod_async(totals.as<double>() + slot, totals.as<double>() + it->second.slot, 16), "total copy");
            }
        } else {
            total(field, slot);
        }
        note_total(field, slot);
    }

    // The conserved total (sum, sum of magnitudes) of a field into totals[slot], [slot + 1]; with a range record
    // (`record` >= 0) the same pass checks (and clamps) the field's range into it.
    void total(int field, int slot, int record = -1)
    {
        const bool ranged = record >= 0;
        const long long cells = static_cast<long long>(shape.cells());
        long long chunk = std::max(1LL, (cells + kReduceBlocks - 1) / kReduceBlocks);
        const int blocks = static_cast<int>((cells + chunk - 1) / chunk);
        std::vector<std::uint64_t> k = args(0.0);
        double* parts = L.as<double>();
        const FieldDecl& d = law->fields[static_cast<std::size_t>(field)];
        double lo = d.lo, hi = d.hi;
        int clamp = d.clamp ? 1 : 0;
        unsigned* rec = ranged ? records.as<unsigned>() + 2 * record : nullptr;
        unsigned long long* first = firsts.as<unsigned long long>() + (ranged ? record : 0);
        void* argv1[] = {k.data(), &field, &chunk, &parts, &lo, &hi, &clamp, &rec, &first};
        check(fe_hip_launch(fn("k_total"), static_cast<unsigned>(blocks), 1, 1, 256, 1, 1, 0, argv1), "total launch");
        double* dst = totals.as<double>() + slot;
        int count = blocks;
        void* argv2[] = {&parts, &count, &dst};
        check(fe_hip_launch(fn("k_total_final"), 1, 1, 1, 256, 1, 1, 0, argv2), "total final launch");
        prof.launches += 2;
        if (ranged && clamp) {
            touched(field);
        }
    }

    // A range check of a field into a record of its own (the counts and first bad cell of this check alone, as the
    // CPU reports each term's check, src/field/Solver.cpp:463-511).
    void range(int field, int record)
    {
        const FieldDecl& d = law->fields[static_cast<std::size_t>(field)];
        double lo = d.lo, hi = d.hi;
        int clamp = d.clamp ? 1 : 0;
        unsigned* rec = records.as<unsigned>() + 2 * record;
        unsigned long long* first = firsts.as<unsigned long long>() + record;
        launch_cells("k_range", 0.0, {&field, &lo, &hi, &clamp, &rec, &first});
        if (clamp) {
            touched(field);
        }
    }

    // A clamped field's out-of-range cells onto its range (the range check does it on a step that checks contracts).
    void clamp(int field)
    {
        const FieldDecl& d = law->fields[static_cast<std::size_t>(field)];
        if (!d.clamp) {
            return;
        }
        double lo = d.lo, hi = d.hi;
        launch_cells("k_clamp", 0.0, {&field, &lo, &hi});
    }

    void reset_records()
    {
        fill_words(records.as<unsigned>(), 2 * kRanges, 2 * kRanges);
        next_record = 0;
    }
    // `zeros` words from p to 0, then `ones` to all ones, in one launch.
    void fill_words(unsigned* p, int zeros, int ones)
    {
        void* argv[] = {&p, &zeros, &ones};
        check(fe_hip_launch(fn("k_fill"), static_cast<unsigned>((zeros + ones + 255) / 256), 1, 1, 256, 1, 1, 0, argv), "launch k_fill");
        ++prof.launches;
    }
    int next_record = 0; // the next free range record since the records were last reset

    struct Judge {
        std::string term;
        std::vector<std::pair<int, int>> ranged; // (field, its range record) for every field the program wrote
        std::vector<std::vector<int>> groups;
        int slot = 0; // totals[slot + 4g ...]: before sum, before mag, after sum, after mag
        int substeps = 1;
    };

    // Runs one program for a step of dt: its substeps, stages and gather; queues contract kernels when checking.
    // Returns the substeps taken.
    int run_program(const ir::Program& p, const hip::ProgramCode& code, const std::string& term, double dt, bool respect,
        bool contracts, std::vector<Violation>& violations, std::vector<Judge>& judges, int& total_slot)
    {
        if (code.gather.empty()) {
            return 1;
        }
        int substeps = 1;
        if (respect && code.max_dt >= 0) {
            std::vector<double> u(static_cast<std::size_t>(std::max(1, layout.uniforms)), 0.0);
            if (!code.limit_stages.empty()) {
                for (const int s : code.limit_stages) {
                    run_stage(s, 0.0);
                }
                check(fe_hip_memcpy_dtoh_async(u.data(), U.p, u.size() * 8), "uniforms readback");
                check(fe_hip_stream_sync(), "sync");
            }
            const double limit = scalar(code.max_dt, u);
            if (limit > 0.0 && std::isfinite(limit) && dt > limit) {
                const double wanted = std::ceil(dt / limit);
                if (wanted > 4096.0) {
                    Violation v;
                    v.kind = Violation::Kind::SubstepCap;
                    v.term = term;
                    std::ostringstream d;
                    d << "max_dt " << limit << " asks for " << wanted << " substeps; capped at 4096";
                    v.detail = d.str();
                    violations.push_back(v);
                    substeps = 4096;
                } else {
                    substeps = static_cast<int>(wanted);
                }
            }
        }
        const double sub_dt = dt / substeps;
        Judge j;
        j.term = term;

        j.groups = p.conserves;
        j.substeps = substeps;
        j.slot = total_slot;
        // Slots: per group of m fields, m (sum, magnitude) pairs before the substeps, then m after.
        if (contracts) {
            for (const std::vector<int>& group : p.conserves) {
                for (std::size_t q = 0; q < group.size(); ++q) {
                    total_before(group[q], total_slot + 2 * static_cast<int>(q));
                }
                total_slot += 4 * static_cast<int>(group.size());
            }
        }
        // A conduction term conserving the one field it writes: its window takes the "after" total on the last substep.
        FusedTotal fused;
        const bool fuse = contracts && p.conserves.size() == 1 && p.conserves[0].size() == 1 &&
            fusable(p, code, p.conserves[0][0]) != nullptr;
        for (int s = 0; s < substeps; ++s) {
            for (const int st : code.stages) {
                run_stage(st, sub_dt);
            }
            if (fuse && s + 1 == substeps) {
                fused.field = p.conserves[0][0];
                fused.record = next_record++;
                launch_gather(code, sub_dt, &fused);
            } else {
                launch_gather(code, sub_dt);
            }
            swap(code.swaps);
        }
        for (const int f : p.writes) {
            touched(f);
            if (!contracts) {
                clamp(f);
            }
        }
        if (contracts) {
            // a written field's range is checked in the pass that takes its "after" total
            std::set<int> ranged;
            int slot = j.slot;
            for (const std::vector<int>& group : p.conserves) {
                const int m = static_cast<int>(group.size());
                for (int q = 0; q < m; ++q) {
                    const int f = group[static_cast<std::size_t>(q)];
                    const bool writes = std::find(p.writes.begin(), p.writes.end(), f) != p.writes.end() && ranged.count(f) == 0;
                    if (f == fused.field) {
                        // taken by the window: its tiles' totals folded in order
                        double* parts = tile_sums.as<double>();
                        int count = static_cast<int>(fused.tiles);
                        double* dst = totals.as<double>() + slot + 2 * m + 2 * q;
                        void* fargs[] = {&parts, &count, &dst};
                        check(fe_hip_launch(fn("k_total_final"), 1, 1, 1, 256, 1, 1, 0, fargs), "total final launch");
                        ++prof.launches;
                        note_total(f, slot + 2 * m + 2 * q);
                        j.ranged.emplace_back(f, fused.record);
                        ranged.insert(f);
                        continue;
                    }
                    const int record = writes ? next_record++ : -1;
                    total(f, slot + 2 * m + 2 * q, record);
                    note_total(f, slot + 2 * m + 2 * q);
                    if (writes) {
                        j.ranged.emplace_back(f, record);
                        ranged.insert(f);
                        if (law->fields[static_cast<std::size_t>(f)].clamp) {
                            cached_totals.erase(f); // the clamp changed it after its total
                        }
                    }
                }
                slot += 4 * m;
            }
            for (const int f : p.writes) {
                if (ranged.count(f) == 0) {
                    j.ranged.emplace_back(f, next_record);
                    range(f, next_record++);
                }
            }
            judges.push_back(j);
        }
        return substeps;
    }

    // Totals slots one program needs (kTotals doubles in all).
    static int total_slots(const ir::Program& p)
    {
        int n = 0;
        for (const std::vector<int>& g : p.conserves) {
            n += 4 * static_cast<int>(g.size());
        }
        return n;
    }

    void judge(const std::vector<Judge>& judges, std::vector<Violation>& violations)
    {
        if (judges.empty()) {
            return;
        }
        // One copy into pinned memory: the records and the totals slots these judges use.
        int used = 0;
        for (const Judge& j : judges) {
            int end = j.slot;
            for (const std::vector<int>& g : j.groups) {
                end += 4 * static_cast<int>(g.size());
            }
            used = std::max(used, end);
        }
        const std::size_t head = static_cast<std::size_t>(static_cast<char*>(totals.p) - static_cast<char*>(contract.p));
        check(fe_hip_memcpy_dtoh_async(judged, contract.p, head + static_cast<std::size_t>(used) * 8), "records readback");
        judged_has_flags = true; // the copy began at the change flags and bounds keys
        check(fe_hip_stream_sync(), "sync");
        const unsigned* rec = reinterpret_cast<const unsigned*>(static_cast<const char*>(judged) + (static_cast<char*>(records.p) - static_cast<char*>(contract.p)));
        const unsigned long long* first =
            reinterpret_cast<const unsigned long long*>(static_cast<const char*>(judged) + (static_cast<char*>(firsts.p) - static_cast<char*>(contract.p)));
        const double* tot = reinterpret_cast<const double*>(static_cast<const char*>(judged) + head);
        for (const Judge& j : judges) {
            int slot = j.slot;
            for (std::size_t g = 0; g < j.groups.size(); ++g) {
                const int m = static_cast<int>(j.groups[g].size());
                double before = 0.0, scale = 0.0, after = 0.0;
                std::string names;
                for (int q = 0; q < m; ++q) {
                    before += tot[static_cast<std::size_t>(slot + 2 * q)];
                    scale += tot[static_cast<std::size_t>(slot + 2 * q + 1)];
                    after += tot[static_cast<std::size_t>(slot + 2 * m + 2 * q)];
                    names += (names.empty() ? "" : " + ") + law->fields[static_cast<std::size_t>(j.groups[g][static_cast<std::size_t>(q)])].name;
                }
                slot += 4 * m;
                // Float storage rounds every cell every substep: the budget scales with the amount and the work
                // (src/field/Solver.cpp:433-461).
                const double tolerance = 1e-5 * (scale + 1.0) * std::sqrt(static_cast<double>(j.substeps));
                if (std::abs(after - before) > tolerance) {
                    Violation v;
                    v.kind = Violation::Kind::Conservation;
                    v.term = j.term;
                    v.field = names;
                    std::ostringstream d;
                    d << "total went " << before << " -> " << after << " (tolerance " << tolerance << ")";
                    v.detail = d.str();
                    violations.push_back(v);
                }
            }
            for (const auto& [f, r] : j.ranged) {
                const unsigned out = rec[2 * static_cast<std::size_t>(r)];
                const unsigned nan = rec[2 * static_cast<std::size_t>(r) + 1];
                if (out + nan == 0) {
                    continue;
                }
                const FieldDecl& d = law->fields[static_cast<std::size_t>(f)];
                Violation v;
                v.kind = nan > 0 ? Violation::Kind::NotFinite : Violation::Kind::Range;
                v.term = j.term;
                v.field = d.name;
                const unsigned long long i = first[static_cast<std::size_t>(r)];
                const std::size_t nx = static_cast<std::size_t>(shape.nx), ny = static_cast<std::size_t>(shape.ny);
                std::ostringstream det;
                det << out + nan << " cell(s); first at (" << i % nx << "," << (i / nx) % ny << "," << i / (nx * ny)
                    << "), declared range [" << d.lo << ", " << d.hi << "]";
                v.detail = det.str();
                violations.push_back(v);
            }
        }
    }

    // A guarded term's written fields are copied aside before its learned scheme runs, into one lattice a field that
    // is allocated the first time that field is guarded and kept for the solver's life (no allocation on a step).
    void snapshot(const std::vector<int>& fields)
    {
        const std::size_t bytes = shape.cells() * 4;
        guard_copy.resize(now.size());
        for (const int f : fields) {
            std::unique_ptr<Buffer>& copy = guard_copy[static_cast<std::size_t>(f)];
            if (!copy) {
                copy = std::make_unique<Buffer>();
                copy->allocate(bytes);
            }
            check(fe_hip_memcpy_dtod_async(copy->p, now[static_cast<std::size_t>(f)], bytes), "snapshot");
        }
    }

    void restore(const std::vector<int>& fields)
    {
        const std::size_t bytes = shape.cells() * 4;
        for (const int f : fields) {
            touched(f);
            check(fe_hip_memcpy_dtod_async(now[static_cast<std::size_t>(f)], guard_copy[static_cast<std::size_t>(f)]->p, bytes),
                "restore");
        }
    }

    void run_derives(bool contracts, std::vector<Judge>& judges)
    {
        for (std::size_t d = 0; d < module.derives.size(); ++d) {
            const hip::ProgramCode& code = layout.derives[d];
            if (code.gather.empty()) {
                continue;
            }
            for (const int st : code.stages) {
                run_stage(st, 0.0);
            }
            launch_cells(code.gather, 0.0);
            swap(code.swaps);
            for (const int f : module.derives[d].writes) {
                touched(f);
            }
        }
        if (!contracts) {
            std::set<int> done;
            for (const ir::Program& p : module.derives) {
                for (const int t : p.writes) {
                    if (done.insert(t).second) {
                        clamp(t);
                    }
                }
            }
        }
        if (contracts && !module.derives.empty()) {
            Judge j;
            std::vector<int> targets;
            for (const ir::Program& p : module.derives) {
                for (const int t : p.writes) {
                    if (std::find(targets.begin(), targets.end(), t) == targets.end()) {
                        targets.push_back(t);
                    }
                }
            }
            for (const int f : targets) {
                j.ranged.emplace_back(f, next_record);
                range(f, next_record++);
            }
            judges.push_back(j);
        }
    }

    void download_changed(FieldSet& fields, const std::vector<char>& which)
    {
        const std::size_t cells = shape.cells();
        for (std::size_t f = 0; f < which.size(); ++f) {
            if (!which[f]) {
                continue;
            }
            check(fe_hip_memcpy_dtoh_async(fields.values(static_cast<int>(f)).data(), now[f], cells * 4), "field download");
            prof.download_bytes += cells * 4;
            stale[f] = 0;
        }
        check(fe_hip_stream_sync(), "sync");
    }

    void read_changed(std::vector<char>& changed)
    {
        const std::size_t nf = law->fields.size();
        std::vector<unsigned> c(std::max<std::size_t>(1, nf));
        check(fe_hip_memcpy_dtoh_async(c.data(), C.p, c.size() * 4), "change flags readback");
        check(fe_hip_stream_sync(), "sync");
        changed.assign(nf, 0);
        for (std::size_t f = 0; f < nf; ++f) {
            changed[f] = c[f] != 0 ? 1 : 0;
        }
    }

    fe_hip_event event(std::size_t i)
    {
        while (events.size() <= i) {
            fe_hip_event e = nullptr;
            check(fe_hip_event_create(&e), "event");
            events.push_back(e);
        }
        return events[i];
    }

    void step(double dt, bool respect_max_dt, bool learned, Law& l, FieldSet& fields, FieldEdits& edits,
        std::vector<char>& changed, std::vector<Violation>& violations, std::vector<int>& substeps) override
    {
        const std::lock_guard<std::mutex> lock(device_mutex());
        step_locked(dt, respect_max_dt, learned, l, fields, edits, changed, violations, substeps);
    }

    // Captured runs of plain steps: a graph of an even number of whole steps (every ping-pong field back where it
    // began), keyed by where the fields lie when it starts, its length, dt, max_dt's use and the params its launches
    // carry by value. Runs of kLongRun steps where no contract falls due, pairs otherwise.
    static constexpr int kLongRun = 32;
    struct Pair {
        std::vector<void*> start;
        int steps = 2;
        double dt = 0.0;
        bool respect = true;
        std::vector<double> params;
        fe_hip_graph graph = nullptr;
        std::vector<int> substeps;
    };
    std::vector<Pair> pairs;

    void drop_pairs()
    {
        for (Pair& q : pairs) {
            if (q.graph != nullptr) {
                fe_hip_graph_destroy(q.graph);
            }
        }
        pairs.clear();
    }

    // A step the card can take from a graph: lazy, no contracts due, no guarded learned term, and substeps that need
    // no reduction read back (max_dt of params, h and constants only, or not respected).
    bool replayable(bool respect, bool learned, const FieldEdits& edits)
    {
        if (!lazy || edits.all || edits.shifted()) {
            return false;
        }
        for (const std::vector<char>& planes : edits.planes) {
            if (std::any_of(planes.begin(), planes.end(), [](char c) { return c != 0; })) {
                return false;
            }
        }
        for (std::size_t t = 0; t < module.terms.size(); ++t) {
            if (learned && !module.learned[t].equations.empty()) {
                return false;
            }
            if (respect && layout.terms[t].max_dt >= 0 && !layout.terms[t].limit_stages.empty()) {
                return false;
            }
        }
        return !any_idle(); // a graph runs every term; a step that skips one is taken eagerly
    }

    [[nodiscard]] const std::vector<std::string>& gather_kernels() const override { return gathers; }

    void run(double dt, int steps, bool respect, bool learned, Law& l, FieldSet& fields, FieldEdits& edits,
        std::vector<char>& changed, std::vector<Violation>& violations, std::vector<int>& substeps) override
    {
        const std::lock_guard<std::mutex> lock(device_mutex());
        law = &l;
        const std::size_t nf = law->fields.size();
        std::vector<char> any(nf, 0);
        auto merge = [&] {
            for (std::size_t f = 0; f < nf && f < changed.size(); ++f) {
                any[f] = any[f] || changed[f];
            }
        };
        int i = 0;
        while (i < steps) {
            // steps from here before the next one that checks contracts (step_count % check_every == 0)
            const long long every = check_every;
            const long long due = ((step_count + every - 1) / every) * every;
            const long long free_steps = std::min<long long>(due - step_count, steps - i);
            const int length = free_steps >= kLongRun ? kLongRun : (free_steps >= 2 ? 2 : 0);
            if (length > 0 && replayable(respect, learned, edits)) {
                replay_pair(dt, respect, length, substeps);
                step_count += length;
                carry_bounds(dt, length, substeps);
                for (const auto* progs : {&module.terms, &module.derives}) {
                    for (const ir::Program& p : *progs) {
                        for (const int f : p.writes) {
                            touched(f);
                        }
                    }
                }
                changed.assign(nf, 0);
                for (const auto* progs : {&module.terms, &module.derives}) {
                    for (const ir::Program& p : *progs) {
                        for (const int f : p.writes) {
                            changed[static_cast<std::size_t>(f)] = 1;
                        }
                    }
                }
                for (std::size_t f = 0; f < nf; ++f) {
                    stale[f] = stale[f] || changed[f];
                }
                merge();
                i += length;
                continue;
            }
            step_locked(dt, respect, learned, l, fields, edits, changed, violations, substeps);
            merge();
            i += 1;
        }
        changed = any;
    }

    // `length` (even) plain steps from the graph captured for where the fields lie now (captured first if there is
    // none).
    void replay_pair(double dt, bool respect, int length, std::vector<int>& substeps)
    {
        upload_params();
        Pair* hit = nullptr;
        for (Pair& q : pairs) {
            if (q.start == now && q.steps == length && q.dt == dt && q.respect == respect && q.params == law->param_values) {
                hit = &q;
            }
        }
        if (hit == nullptr) {
            if (pairs.size() >= 8) {
                drop_pairs();
            }
            Pair q;
            q.start = now;
            q.steps = length;
            q.dt = dt;
            q.respect = respect;
            q.params = law->param_values;
            const std::vector<void*> now0 = now, nxt0 = nxt;
            check(fe_hip_graph_begin_capture(), "graph capture");
            try {
                for (int s = 0; s < length; ++s) {
                    plain_step(dt, respect, q.substeps);
                }
            } catch (...) {
                fe_hip_graph end = nullptr;
                if (fe_hip_graph_end_capture(&end) == 0 && end != nullptr) {
                    fe_hip_graph_destroy(end);
                }
                now = now0;
                nxt = nxt0;
                throw;
            }
            check(fe_hip_graph_end_capture(&q.graph), "graph capture");
            if (now != now0) {
                throw std::runtime_error("a captured run of steps did not bring the fields back where they began");
            }
            pairs.push_back(std::move(q));
            hit = &pairs.back();
        }
        check(fe_hip_graph_launch(hit->graph), "graph launch");
        substeps = hit->substeps;
        prof = StepProfile{};
        prof.backend = "HIP";
        prof.terms.resize(module.terms.size());
        for (std::size_t t = 0; t < module.terms.size(); ++t) {
            prof.terms[t].name = module.terms[t].name;
            prof.terms[t].substeps = t < substeps.size() ? substeps[t] : 1;
            prof.terms[t].ms = -1.0; // not timed: the steps ran as one graph
        }
        timings_pending = false;
    }

    // One step's launches and nothing else: no contracts, no events, no waits (what a graph captures).
    void plain_step(double dt, bool respect, std::vector<int>& substeps)
    {
        std::vector<Violation> none;
        std::vector<Judge> judges;
        int slot = 0;
        tracking = false;
        substeps.assign(module.terms.size(), 1);
        for (std::size_t t = 0; t < module.terms.size(); ++t) {
            substeps[t] = run_program(module.terms[t], layout.terms[t], module.terms[t].name, dt, respect, false, none, judges, slot);
        }
        run_derives(false, judges);
    }

    // ---- ranges and idle terms ----

    // Whether a program has nothing to do this step: every equation's rate is exactly 0 at every cell, provably from
    // the fields known to be zero (always_zero) or from the carried ranges (bounds_of is [0, 0]); the reference's test
    // (src/field/cuda/CudaSolver.cpp:590-592). An empty program is idle.
    bool idle(const std::vector<Equation>& equations) const
    {
        std::vector<char> zero(flowing.size(), 0);
        for (std::size_t f = 0; f < flowing.size(); ++f) {
            zero[f] = flowing[f].none() ? 1 : 0;
        }
        for (const Equation& eq : equations) {
            if (!always_zero(*eq.rhs, zero, law->param_values) && !bounds_of(*eq.rhs, flowing, law->param_values, shape.cell).none()) {
                return false;
            }
        }
        return true;
    }

    // A derive recomputes its target from other fields: its range is unknown until measured.
    void derives_unbounded()
    {
        for (const Derive& d : law->derives) {
            flowing[static_cast<std::size_t>(d.target)] = Bounds{};
        }
    }

    // Queues the measurement of the fields' ranges (k_bounds, one pass over each) into bounds_keys.
    void queue_measure(const std::vector<int>& which)
    {
        int nf = static_cast<int>(law->fields.size());
        unsigned* keys = bounds_keys.as<unsigned>();
        void* rargs[] = {&keys, &nf};
        check(fe_hip_launch(fn("k_bounds_reset"), static_cast<unsigned>((2 * nf * hip::kBoundSlots + 255) / 256), 1, 1, 256, 1, 1, 0, rargs),
            "launch k_bounds_reset");
        std::vector<std::uint64_t> k = args(0.0);
        for (int f : which) {
            void* bargs[] = {k.data(), &f, &keys, &nf};
            check(fe_hip_launch(fn("k_bounds"), 2048, 1, 1, 256, 1, 1, 0, bargs), "launch k_bounds");
            ++prof.launches;
        }
        prof.launches += 1;
    }

    // The measured ranges into `known` (after the card has finished the copy into `host`).
    void take_measure(const std::vector<int>& which, const std::vector<unsigned>& host)
    {
        const std::size_t nf = law->fields.size();
        auto value = [](unsigned key) {
            const unsigned b = (key & 0x80000000u) ? (key & 0x7fffffffu) : ~key;
            float v = 0.0f;
            std::memcpy(&v, &b, 4);
            return static_cast<double>(v);
        };
        const std::size_t slots = hip::kBoundSlots;
        for (const int f : which) {
            unsigned lo = 0xffffffffu, hi = 0u;
            for (std::size_t q = 0; q < slots; ++q) {
                lo = std::min(lo, host[static_cast<std::size_t>(f) * slots + q]);
                hi = std::max(hi, host[(nf + static_cast<std::size_t>(f)) * slots + q]);
            }
            if (lo == 0xffffffffu && hi == 0u) { // every cell NaN: the empty range, as a look over the cells finds
                known[static_cast<std::size_t>(f)] = Bounds{std::numeric_limits<double>::infinity(), -std::numeric_limits<double>::infinity()};
            } else {
                known[static_cast<std::size_t>(f)] = Bounds{value(lo), value(hi)};
            }
        }
    }

    // `steps` replayed steps carried through flow_bounds on the host (no measurement), with each term's substeps.
    void carry_bounds(double dt, int steps, const std::vector<int>& substeps)
    {
        for (int s = 0; s < steps; ++s) {
            flowing = known;
            for (std::size_t t = 0; t < law->terms.size(); ++t) {
                flow_bounds(law->terms[t].equations, dt, t < substeps.size() ? substeps[t] : 1, *law, shape.cell, flowing);
            }
            derives_unbounded();
            bool same = true;
            for (std::size_t f = 0; f < known.size(); ++f) {
                same = same && flowing[f].lo == known[f].lo && flowing[f].hi == known[f].hi;
            }
            known = flowing;
            if (same) {
                break; // a fixed point: further steps carry nothing further
            }
        }
    }

    // Whether any term would be skipped at the next step (then the step is taken eagerly, not from a graph that runs
    // every term).
    bool any_idle()
    {
        flowing = known;
        for (std::size_t t = 0; t < law->terms.size(); ++t) {
            if (!layout.terms[t].gather.empty() && idle(law->terms[t].equations)) {
                return true;
            }
            flow_bounds(law->terms[t].equations, 0.0, 1, *law, shape.cell, flowing);
        }
        return false;
    }

    void step_locked(double dt, bool respect_max_dt, bool learned, Law& l, FieldSet& fields, FieldEdits& edits,
        std::vector<char>& changed, std::vector<Violation>& violations, std::vector<int>& substeps)
    {
        const auto t0 = std::chrono::steady_clock::now();
        law = &l;
        prof = StepProfile{};
        prof.backend = "HIP";
        prof.terms.resize(module.terms.size());
        upload_params();
        upload(fields, edits);
        const std::size_t nf = law->fields.size();
        const bool contracts = (step_count % check_every) == 0;
        ++step_count;
        // The change flags are read back unless the step is lazy and unjudged (then every field a program writes counts
        // as changed, and the flags are not cleared either).
        const bool read_flags = !lazy || contracts;
        tracking = read_flags;
        if (read_flags) {
            fill_words(C.as<unsigned>(), static_cast<int>(std::max<std::size_t>(1, nf)), 0);
        }
        std::vector<Judge> judges;
        int total_slot = 0;
        if (contracts) {
            reset_records();
        }
        substeps.assign(module.terms.size(), 0);
        flowing = known;
        std::vector<char> wrote(nf, 0); // fields a program that ran wrote this step
        auto note_writes = [&](const ir::Program& p) {
            for (const int f : p.writes) {
                wrote[static_cast<std::size_t>(f)] = 1;
            }
        };
        std::size_t ev = 0;
        for (std::size_t t = 0; t < module.terms.size(); ++t) {
            TermProfile& tp = prof.terms[t];
            tp.name = module.terms[t].name;
            const int launches0 = prof.launches;
            check(fe_hip_event_record(event(ev)), "event");
            const bool guarded = learned && !module.learned[t].equations.empty();
            // Judge what is pending before the records are reused: a guarded term judges its own at once, the totals
            // buffer holds kTotals doubles and the range records kRanges checks.
            const int need = std::max(total_slots(module.terms[t]), total_slots(module.learned[t]));
            const int checks = static_cast<int>(std::max(module.terms[t].writes.size(), module.learned[t].writes.size()));
            if (contracts && !judges.empty() && (guarded || total_slot + need > kTotals || next_record + checks > kRanges)) {
                judge(judges, violations);
                judges.clear();
                total_slot = 0;
                reset_records();
            }
            bool done = false;
            if (guarded && idle(law->terms[t].learned)) {
                tp.idle = true; // the learned scheme has nothing to do, so the term holds as it is
                substeps[t] = 1;
                done = true;
            } else if (guarded) {
                // Snapshot what the term writes, run the learned scheme and judge it now; a breach is undone and the
                // classical scheme runs (src/field/cuda/CudaSolver.cpp:673-705 is the reference's guard).
                const std::vector<Bounds> flowing_before = flowing;
                snapshot(module.terms[t].writes);
                std::vector<Judge> own;
                int slot = 0;
                reset_records();
                std::vector<Violation> caught;
                const int n = run_program(module.learned[t], layout.learned[t], module.terms[t].name, dt, respect_max_dt, true,
                    caught, own, slot);
                judge(own, caught);
                reset_records();
                if (caught.empty()) {
                    tp.learned = true;
                    substeps[t] = n;
                    done = true;
                    flow_bounds(law->terms[t].learned, dt, n, *law, shape.cell, flowing);
                    note_writes(module.learned[t]);
                } else {
                    restore(module.terms[t].writes);
                    flowing = flowing_before;
                    tp.fell_back = true;
                    ++prof.fallbacks;
                    check(fe_hip_stream_sync(), "sync");
                }
            }
            if (!done && idle(law->terms[t].equations)) {
                // Nothing to do: not dispatched, and its contracts hold without counting (src/field/Solver.cpp:341-347).
                tp.idle = true;
                substeps[t] = 1;
            } else if (!done) {
                substeps[t] = run_program(module.terms[t], layout.terms[t], module.terms[t].name, dt, respect_max_dt, contracts,
                    violations, judges, total_slot);
                flow_bounds(law->terms[t].equations, dt, substeps[t], *law, shape.cell, flowing);
                note_writes(module.terms[t]);
            }
            check(fe_hip_event_record(event(ev + 1)), "event");
            gathers.resize(module.terms.size());
            gathers[t] = last_gather;
            last_gather.clear();
            tp.substeps = substeps[t];
            tp.launches = prof.launches - launches0;
            ev += 2;
        }
        derive_event = ev;
        derive_timed = !module.derives.empty();
        if (derive_timed) {
            check(fe_hip_event_record(event(ev++)), "event");
        }
        if (contracts && !judges.empty() && next_record + static_cast<int>(law->fields.size()) > kRanges) {
            judge(judges, violations);
            judges.clear();
            total_slot = 0;
            reset_records();
        }
        run_derives(contracts, judges);
        if (derive_timed) {
            check(fe_hip_event_record(event(ev++)), "event");
        }
        derives_unbounded();
        for (const ir::Program& p : module.derives) {
            note_writes(p);
        }
        // The ranges the next step starts from: measured on the card after an eager step and every kTightenEvery lazy
        // steps (one pass over each field written, read back with what else this step reads), carried otherwise.
        known = flowing;
        const bool measure = !lazy || (lazy_steps++ % kTightenEvery) == 0;
        std::vector<int> measured;
        if (measure) {
            for (std::size_t f = 0; f < nf; ++f) {
                if (wrote[f]) {
                    measured.push_back(static_cast<int>(f));
                }
            }
            if (!measured.empty()) {
                queue_measure(measured);
            }
        }
        // One copy and one wait for all the step reads back: the change flags and the bounds keys lead the contract
        // buffer, so the judge's copy of the records and totals brings them along; without contracts, that prefix alone.
        judged_has_flags = false;
        if (contracts) {
            judge(judges, violations);
        }
        if ((read_flags || !measured.empty()) && !judged_has_flags) {
            const std::size_t prefix = static_cast<std::size_t>(static_cast<char*>(records.p) - static_cast<char*>(contract.p));
            check(fe_hip_memcpy_dtoh_async(judged, contract.p, prefix), "change flags readback");
            check(fe_hip_stream_sync(), "sync");
        }
        const unsigned* back = static_cast<const unsigned*>(judged);
        const unsigned* keys_back = reinterpret_cast<const unsigned*>(static_cast<const char*>(judged) +
            (static_cast<char*>(bounds_keys.p) - static_cast<char*>(contract.p)));
        if (!read_flags) {
            // Nothing comes back and nothing is waited for: every field a program writes is taken as changed, and the
            // card's time is read from the events when the profile is asked for.
            changed.assign(nf, 0);
            for (const auto* progs : {&module.terms, &module.learned, &module.derives}) {
                for (const ir::Program& p : *progs) {
                    for (const int f : p.writes) {
                        changed[static_cast<std::size_t>(f)] = 1;
                    }
                }
            }
        } else {
            changed.assign(nf, 0);
            for (std::size_t f = 0; f < nf; ++f) {
                changed[f] = back[f] != 0 ? 1 : 0;
            }
        }
        if (!measured.empty()) {
            take_measure(measured, std::vector<unsigned>(keys_back, keys_back + 2 * nf * hip::kBoundSlots));
        }
        for (std::size_t f = 0; f < nf; ++f) {
            stale[f] = stale[f] || changed[f];
        }
        if (!lazy) {
            download_changed(fields, changed);
        }
        timings_pending = true;
        if (!lazy) {
            resolve_timings();
        }
        prof.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }

    // Card time per term and for the derives, from the last step's events (waits for them).
    void resolve_timings() const
    {
        if (!timings_pending) {
            return;
        }
        timings_pending = false;
        if (derive_event == 0 && !derive_timed) {
            return; // no term and no derive: nothing was timed
        }
        fe_hip_event_sync(events[derive_timed ? derive_event + 1 : derive_event - 1]);
        double gpu = 0.0;
        for (std::size_t t = 0; t < prof.terms.size(); ++t) {
            float ms = 0.0f;
            fe_hip_event_elapsed(events[2 * t], events[2 * t + 1], &ms);
            prof.terms[t].ms = ms;
            gpu += ms;
        }
        float dms = 0.0f;
        if (derive_timed) {
            fe_hip_event_elapsed(events[derive_event], events[derive_event + 1], &dms);
        }
        prof.derive_ms = dms;
        prof.gpu_ms = gpu + dms;
    }

    void derive(Law& l, FieldSet& fields, FieldEdits& edits, std::vector<char>& changed) override
    {
        const std::lock_guard<std::mutex> lock(device_mutex());
        law = &l;
        upload_params();
        upload(fields, edits);
        check(fe_hip_memset_d32_async(C.p, 0u, std::max<std::size_t>(1, law->fields.size())), "change flags reset");
        std::vector<Judge> none;
        run_derives(false, none);
        std::vector<int> measured;
        for (const Derive& d : law->derives) {
            if (std::find(measured.begin(), measured.end(), d.target) == measured.end()) {
                measured.push_back(d.target);
            }
        }
        std::vector<unsigned> keys(2 * law->fields.size() * hip::kBoundSlots, 0u);
        if (!measured.empty()) {
            queue_measure(measured);
            check(fe_hip_memcpy_dtoh_async(keys.data(), bounds_keys.p, keys.size() * 4), "bounds readback");
        }
        read_changed(changed);
        take_measure(measured, keys);
        for (std::size_t f = 0; f < law->fields.size(); ++f) {
            stale[f] = stale[f] || changed[f];
        }
        if (!lazy) {
            download_changed(fields, changed);
        }
    }

    void fetch(FieldSet& fields, const std::vector<char>& which) override
    {
        const std::lock_guard<std::mutex> lock(device_mutex());
        std::vector<char> need(law->fields.size(), 0);
        for (std::size_t f = 0; f < need.size() && f < which.size(); ++f) {
            need[f] = which[f] && stale[f] ? 1 : 0;
        }
        download_changed(fields, need);
    }

    // Device pointers of a field's current cells (for callers that keep data on the card).
    [[nodiscard]] void* field_pointer(int f) const { return now[static_cast<std::size_t>(f)]; }
    void* device_field(int f) override
    {
        touched(f); // whoever holds the address may write through it: a cached total no longer stands for it
        known[static_cast<std::size_t>(f)] = Bounds{}; // and its range is no longer known
        return field_pointer(f);
    }
    // A caller wrote field f's cells on the card: the host copy is stale until fetched.
    void device_written(int f) override
    {
        stale[static_cast<std::size_t>(f)] = 1;
        touched(f);
        known[static_cast<std::size_t>(f)] = Bounds{};
    }

    // A static field read from the caller's buffer (its own array kept for when it is unbound).
    bool bind_field(int f, void* buffer) override
    {
        const std::size_t i = static_cast<std::size_t>(f);
        for (const auto* progs : {&module.terms, &module.learned, &module.derives}) {
            for (const ir::Program& p : *progs) {
                if (std::find(p.writes.begin(), p.writes.end(), f) != p.writes.end()) {
                    return false;
                }
            }
        }
        if (own.size() != now.size()) {
            own.assign(now.size(), nullptr);
        }
        if (own[i] == nullptr) {
            own[i] = now[i];
        }
        now[i] = buffer != nullptr ? buffer : own[i];
        if (buffer == nullptr) {
            own[i] = nullptr;
        }
        device_written(f);
        return true;
    }

    // The field's fixed-order compensated total and total magnitude, from the card (waits for it).
    // The field's total added into acc[0] (sum) and acc[1] (magnitude) on the card, queued.
    void field_total_accumulate(int f, double* acc) override
    {
        const std::lock_guard<std::mutex> lock(device_mutex());
        total(f, kTotals - 2);
        note_total(f, kTotals - 2);
        double* src = totals.as<double>() + (kTotals - 2);
        void* argv[] = {&src, &acc};
        check(fe_hip_launch(fn("k_acc"), 1, 1, 1, 64, 1, 1, 0, argv), "total accumulate");
    }

    void field_total(int f, double& sum, double& mag) override
    {
        const std::lock_guard<std::mutex> lock(device_mutex());
        total(f, kTotals - 2);
        note_total(f, kTotals - 2);
        double out[2] = {0.0, 0.0};
        check(fe_hip_memcpy_dtoh_async(out, totals.as<double>() + (kTotals - 2), sizeof(out)), "totals readback");
        check(fe_hip_stream_sync(), "sync");
        sum = out[0];
        mag = out[1];
    }
};

} // namespace

std::shared_ptr<LawBackend> HipSolver::create(Law& law, const GridShape& shape, std::string& why)
{
    const std::string no = unavailable();
    if (!no.empty()) {
        why = no;
        return nullptr;
    }
    if (shape.nx < 1 || shape.ny < 1 || shape.nz < 1 || !(shape.cell > 0.0)) {
        why = "the lattice must have at least one cell and a positive cell size";
        return nullptr;
    }
    const std::lock_guard<std::mutex> lock(device_mutex());
    auto impl = std::make_shared<Impl>();
    impl->law = &law;
    impl->shape = shape;
    fe_hip_device_info info{};
    fe_hip_device_open(&info);
    impl->adapter_name = std::string(info.name) + " (" + info.arch + ")";
    try {
        impl->build(why);
    } catch (const std::exception& e) {
        why = e.what();
        return nullptr;
    }
    return impl;
}

std::string HipSolver::unavailable()
{
    return fe_hip_unavailable();
}

std::string HipSolver::source(const Law& law)
{
    std::string why;
    return source(law, why);
}

std::string HipSolver::source(const Law& law, std::string& why)
{
    hip::LawLayout layout;
    return hip::generate(ir::lower(law), law, layout, why);
}

} // namespace fieldengine::field
