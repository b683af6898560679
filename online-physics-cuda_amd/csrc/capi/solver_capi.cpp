// The solver behind include/fe_field.h's fe_solver_* entry points: the reference's Solver calls
// (src/field/Solver.hpp:103-215) over the HIP backend, with errors returned instead of thrown.
#include "fe_field.h"

#include "field/Solver.hpp"
#include "field/capi_internal.hpp"
#include "field/hip/HipSolver.hpp"
#include "field/ir/Adjoint.hpp"
#include "fe_hip.h"

#include <cstdlib>
#include <cstring>
#include <exception>
#include <sstream>
#include <string>

using namespace fieldengine::field;

struct fe_solver {
    std::unique_ptr<Solver> solver;
    std::string error;
};

namespace {

void put(char* buf, size_t bytes, const std::string& s)
{
    if (buf == nullptr || bytes == 0) {
        return;
    }
    const size_t n = s.size() < bytes - 1 ? s.size() : bytes - 1;
    std::memcpy(buf, s.data(), n);
    buf[n] = '\0';
}

char* dup(const std::string& s)
{
    char* p = static_cast<char*>(std::malloc(s.size() + 1));
    if (p != nullptr) {
        std::memcpy(p, s.c_str(), s.size() + 1);
    }
    return p;
}

void json_string(std::ostringstream& out, const std::string& s)
{
    out << '"';
    for (const char c : s) {
        if (c == '"' || c == '\\') {
            out << '\\' << c;
        } else if (static_cast<unsigned char>(c) < 0x20) {
            out << ' ';
        } else {
            out << c;
        }
    }
    out << '"';
}

template <typename Fn>
int guarded(fe_solver* s, Fn&& fn)
{
    if (s == nullptr || !s->solver) {
        return -1;
    }
    try {
        fn();
        s->error.clear();
        return 0;
    } catch (const std::exception& e) {
        s->error = e.what();
        return 1;
    }
}

bool valid_field(fe_solver* s, int field)
{
    return s != nullptr && s->solver && field >= 0 && field < static_cast<int>(s->solver->law().fields.size());
}

} // namespace

extern "C" {

fe_solver* fe_solver_create(const fe_law* law, int nx, int ny, int nz, double h, char* why, size_t why_bytes)
{
    if (law == nullptr || nx < 1 || ny < 1 || nz < 1 || !(h > 0.0)) {
        put(why, why_bytes, "fe_solver_create: a law and a lattice of at least one cell with h > 0");
        return nullptr;
    }
    auto* s = new fe_solver{};
    try {
        s->solver = std::make_unique<Solver>(law->law, GridShape{nx, ny, nz, h});
        std::string reason;
        if (!s->solver->use_gpu(reason)) {
            put(why, why_bytes, reason);
            delete s;
            return nullptr;
        }
    } catch (const std::exception& e) {
        put(why, why_bytes, e.what());
        delete s;
        return nullptr;
    }
    put(why, why_bytes, "");
    return s;
}

void fe_solver_free(fe_solver* s)
{
    delete s;
}

int fe_solver_set_field(fe_solver* s, int field, const float* values)
{
    if (!valid_field(s, field) || values == nullptr) {
        return -1;
    }
    return guarded(s, [&] {
        std::vector<float>& v = s->solver->fields().values(field);
        std::memcpy(v.data(), values, v.size() * sizeof(float));
    });
}

int fe_solver_get_field(fe_solver* s, int field, float* out)
{
    if (!valid_field(s, field) || out == nullptr) {
        return -1;
    }
    return guarded(s, [&] {
        const std::vector<float>& v = s->solver->values(field);
        std::memcpy(out, v.data(), v.size() * sizeof(float));
    });
}

int fe_solver_set_param(fe_solver* s, const char* name, double value)
{
    if (s == nullptr || !s->solver || name == nullptr) {
        return -1;
    }
    return s->solver->set_param(name, value) ? 0 : -1;
}

int fe_solver_step(fe_solver* s, double dt)
{
    return guarded(s, [&] { s->solver->step(dt); });
}

int fe_solver_run(fe_solver* s, double dt, int steps)
{
    return guarded(s, [&] { s->solver->run(dt, steps); });
}

int fe_solver_refresh_derived(fe_solver* s)
{
    return guarded(s, [&] { s->solver->refresh_derived(); });
}

int fe_solver_set_lazy_readback(fe_solver* s, int on)
{
    return guarded(s, [&] { s->solver->set_lazy_readback(on != 0); });
}

int fe_solver_check_contracts_every(fe_solver* s, int steps)
{
    return guarded(s, [&] { s->solver->check_contracts_every(steps); });
}

int fe_solver_set_learned(fe_solver* s, int on)
{
    return guarded(s, [&] { s->solver->set_learned(on != 0); });
}

int fe_solver_set_respect_max_dt(fe_solver* s, int on)
{
    return guarded(s, [&] { s->solver->set_respect_max_dt(on != 0); });
}

char* fe_solver_violations_json(fe_solver* s)
{
    if (s == nullptr || !s->solver) {
        return nullptr;
    }
    std::ostringstream out;
    out << '[';
    bool first = true;
    for (const Violation& v : s->solver->violations()) {
        out << (first ? "" : ",") << "{\"kind\":\"" << to_string(v.kind) << "\",\"term\":";
        json_string(out, v.term);
        out << ",\"field\":";
        json_string(out, v.field);
        out << ",\"detail\":";
        json_string(out, v.detail);
        out << '}';
        first = false;
    }
    out << ']';
    return dup(out.str());
}

int fe_solver_last_substeps(fe_solver* s, int* out, int count)
{
    if (s == nullptr || !s->solver || out == nullptr) {
        return -1;
    }
    const std::vector<int>& v = s->solver->last_substeps();
    for (int i = 0; i < count && i < static_cast<int>(v.size()); ++i) {
        out[i] = v[static_cast<std::size_t>(i)];
    }
    return static_cast<int>(v.size());
}

int fe_solver_last_changed(fe_solver* s, unsigned char* out, int count)
{
    if (s == nullptr || !s->solver || out == nullptr) {
        return -1;
    }
    const std::vector<char>& v = s->solver->last_changed();
    for (int i = 0; i < count && i < static_cast<int>(v.size()); ++i) {
        out[i] = static_cast<unsigned char>(v[static_cast<std::size_t>(i)]);
    }
    return static_cast<int>(v.size());
}

char* fe_solver_profile_json(fe_solver* s)
{
    if (s == nullptr || !s->solver) {
        return nullptr;
    }
    const StepProfile& p = s->solver->profile();
    const HipBackend* backend = s->solver->backend();
    std::ostringstream out;
    out.precision(9);
    out << "{\"backend\":";
    json_string(out, p.backend);
    out << ",\"total_ms\":" << p.total_ms << ",\"gpu_ms\":" << p.gpu_ms << ",\"derive_ms\":" << p.derive_ms
        << ",\"upload_bytes\":" << p.upload_bytes << ",\"download_bytes\":" << p.download_bytes << ",\"launches\":" << p.launches
        << ",\"fallbacks\":" << p.fallbacks << ",\"terms\":[";
    for (std::size_t t = 0; t < p.terms.size(); ++t) {
        const TermProfile& tp = p.terms[t];
        out << (t ? "," : "") << "{\"name\":";
        json_string(out, tp.name);
        out << ",\"kernel\":";
        const std::vector<std::string>* kernels = backend != nullptr ? &backend->gather_kernels() : nullptr;
        json_string(out, kernels != nullptr && t < kernels->size() ? (*kernels)[t] : std::string());
        out << ",\"ms\":" << tp.ms << ",\"launches\":" << tp.launches << ",\"substeps\":" << tp.substeps
            << ",\"idle\":" << (tp.idle ? "true" : "false") << ",\"learned\":" << (tp.learned ? "true" : "false")
            << ",\"fell_back\":" << (tp.fell_back ? "true" : "false") << '}';
    }
    out << "]}";
    return dup(out.str());
}

uint64_t fe_solver_digest(fe_solver* s)
{
    if (s == nullptr || !s->solver) {
        return 0;
    }
    try {
        return s->solver->fields_const().digest();
    } catch (const std::exception& e) {
        s->error = e.what();
        return 0;
    }
}

const char* fe_solver_error(fe_solver* s)
{
    return s == nullptr ? "no solver" : s->error.c_str();
}

void* fe_solver_field_device(fe_solver* s, int field)
{
    if (!valid_field(s, field) || s->solver->backend() == nullptr) {
        return nullptr;
    }
    return s->solver->backend()->device_field(field);
}

int fe_solver_set_field_device(fe_solver* s, int field, const void* src)
{
    if (!valid_field(s, field) || src == nullptr || s->solver->backend() == nullptr) {
        return -1;
    }
    void* dst = s->solver->backend()->device_field(field);
    const std::size_t bytes = s->solver->fields_const_shape_cells() * sizeof(float);
    const int r = fe_hip_memcpy_dtod_async(dst, src, bytes);
    if (r != 0) {
        s->error = std::string("device copy: ") + fe_hip_error_string(r);
        return r;
    }
    s->solver->backend()->device_written(field);
    s->solver->mark_device_written(field);
    return 0;
}

int fe_solver_bind_field_device(fe_solver* s, int field, void* buffer)
{
    if (!valid_field(s, field) || s->solver->backend() == nullptr) {
        return -1;
    }
    if (!s->solver->backend()->bind_field(field, buffer)) {
        s->error = "field " + std::to_string(field) + " is written by the law: only a static field can be bound";
        return -1;
    }
    s->solver->mark_device_written(field);
    return 0;
}

int fe_solver_set_net_weights(fe_solver* s, const char* net, const float* weights, size_t count)
{
    if (s == nullptr || net == nullptr || weights == nullptr) {
        return -1;
    }
    Law& law = s->solver->law_for_training();
    const int n = law.net_id(net);
    if (n < 0) {
        s->error = std::string("no net '") + net + "'";
        return -1;
    }
    NetDecl& d = law.nets[static_cast<std::size_t>(n)];
    if (count != d.parameter_count()) {
        s->error = "net '" + d.name + "' has " + std::to_string(d.parameter_count()) + " weights, not " + std::to_string(count);
        return -1;
    }
    d.weights.assign(weights, weights + count);
    return 0;
}

int fe_solver_shift_window(fe_solver* s, int dx, int dy, int dz, const int* boxes, int n_boxes, const float* arriving)
{
    if (s == nullptr || !s->solver || n_boxes < 0 || (n_boxes > 0 && (boxes == nullptr || arriving == nullptr))) {
        return -1;
    }
    const fieldengine::field::GridShape g = s->solver->fields_const().shape();
    std::vector<fieldengine::field::CellBox> arriving_boxes;
    for (int b = 0; b < n_boxes; ++b) {
        const int* q = boxes + 6 * b;
        const fieldengine::field::CellBox box{q[0], q[1], q[2], q[3], q[4], q[5]};
        if (box.x0 < 0 || box.y0 < 0 || box.z0 < 0 || box.x1 > g.nx || box.y1 > g.ny || box.z1 > g.nz || box.x0 > box.x1 ||
            box.y0 > box.y1 || box.z0 > box.z1) {
            s->error = "shift_window: a box outside the lattice";
            return -1;
        }
        arriving_boxes.push_back(box);
    }
    return guarded(s, [&] {
        fieldengine::field::FieldSet& f = s->solver->shift_window(dx, dy, dz, arriving_boxes);
        const std::size_t nx = static_cast<std::size_t>(g.nx), nxy = nx * static_cast<std::size_t>(g.ny);
        const float* in = arriving;
        for (int field = 0; field < static_cast<int>(f.count()); ++field) {
            std::vector<float>& v = f.values(field);
            // what is now at c was at c + d (cells whose source is outside keep their value until a box fills them)
            const std::vector<float> old = v;
            for (int z = 0; z < g.nz; ++z) {
                for (int y = 0; y < g.ny; ++y) {
                    for (int x = 0; x < g.nx; ++x) {
                        const int sx = x + dx, sy = y + dy, sz = z + dz;
                        if (sx >= 0 && sx < g.nx && sy >= 0 && sy < g.ny && sz >= 0 && sz < g.nz) {
                            v[z * nxy + y * nx + x] = old[sz * nxy + sy * nx + sx];
                        }
                    }
                }
            }
            for (const fieldengine::field::CellBox& b : arriving_boxes) {
                for (int z = b.z0; z < b.z1; ++z) {
                    for (int y = b.y0; y < b.y1; ++y) {
                        for (int x = b.x0; x < b.x1; ++x) {
                            v[z * nxy + y * nx + x] = *in++;
                        }
                    }
                }
            }
        }
    });
}

int fe_solver_fill_field(fe_solver* s, int field, float value)
{
    if (!valid_field(s, field) || s->solver->backend() == nullptr) {
        return -1;
    }
    void* dst = s->solver->backend()->device_field(field);
    std::uint32_t bits = 0;
    std::memcpy(&bits, &value, 4);
    const int r = fe_hip_memset_d32_async(dst, bits, s->solver->fields_const_shape_cells());
    if (r != 0) {
        s->error = std::string("device fill: ") + fe_hip_error_string(r);
        return r;
    }
    s->solver->backend()->device_written(field);
    s->solver->mark_device_written(field);
    return 0;
}

int fe_solver_total_accumulate(fe_solver* s, int field, double* device_acc)
{
    if (!valid_field(s, field) || device_acc == nullptr || s->solver->backend() == nullptr) {
        return -1;
    }
    return guarded(s, [&] { s->solver->backend()->field_total_accumulate(field, device_acc); });
}

int fe_solver_total(fe_solver* s, int field, double* sum, double* magnitude)
{
    if (!valid_field(s, field) || sum == nullptr || s->solver->backend() == nullptr) {
        return -1;
    }
    return guarded(s, [&] {
        double m = 0.0;
        s->solver->backend()->field_total(field, *sum, m);
        if (magnitude != nullptr) {
            *magnitude = m;
        }
    });
}

char* fe_law_adjoint(const fe_law* law, const char* term, char* why, size_t why_bytes)
{
    if (law == nullptr) {
        return nullptr;
    }
    std::string reason;
    const std::string text = ir::adjoint_law(law->law, term != nullptr ? term : "", reason);
    put(why, why_bytes, reason);
    return text.empty() ? nullptr : dup(text);
}

char* fe_law_hip_source(const fe_law* law, char* why, size_t why_bytes)
{
    if (law == nullptr) {
        return nullptr;
    }
    std::string reason;
    const std::string src = HipSolver::source(law->law, reason);
    put(why, why_bytes, reason);
    return src.empty() ? nullptr : dup(src);
}

} // extern "C"
