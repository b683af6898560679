/* fe_field.h - the Field Engine's MI355X backend as a C ABI: handwritten kernels, the law compiler and the solver.
 *
 * The reference's engine API is C++ (src/field/Solver.hpp, src/field/Backend.hpp); this header is the same surface with
 * plain pointers and sizes, for bindings (ctypes here; cgo/JNI would bind the same symbols) and for the tests. All device
 * work is queued on the library's one stream (fe_hip.h). Entry points return 0 on success, a positive HIP status for a
 * device failure, or -1 for arguments they do not take; none throws.
 */
#ifndef FE_FIELD_H
#define FE_FIELD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- handwritten kernels ------------------------------------------------------------------------------------------ */

/* One explicit substep of `d T = diffuse(k, T)` (one k everywhere), in -> out, device arrays of nx*ny*nz floats (x
 * fastest), with the reference CPU's arithmetic (rate in double per cell, rounded to float; float store; denormals
 * flushed). The reference's bit-identity kernel `heat` (bench/stencil_bench.cpp:47-65), whose result its engine is held
 * to cell by cell (bench/stencil_bench.cpp:286-294). */
int fe_heat_handwritten(const float* in, float* out, int nx, int ny, int nz, double k, double h, double sub_dt);
/* The same with the launch shape given. variant 1: a lane per column, ty rows of 64 columns a block (1, 2, 4, 8);
 * variant 2 (nx a multiple of 4): four cells a lane, ty rows of 256 columns a block (2, 6, 14); 0 picks. zchunk planes a
 * block; remap = 1 for the XCD-aware block order. Every shape and variant gives the same bits. */
int fe_heat_handwritten_cfg(const float* in, float* out, int nx, int ny, int nz, double k, double h, double sub_dt, int ty,
    int zchunk, int remap, int variant);
/* b[i] = a[i], n floats: the bandwidth roof the reference measures beside its stencil (bench/stencil_bench.cpp:41-45). */
int fe_copy_f32(const float* a, float* b, size_t n);
/* No counterpart in the reference: the sweep that establishes the attainable HBM rate. b[i] = a[i] (read_only = 0) or a
 * read of a alone (read_only = 1; b may then be NULL), n floats (a multiple of 4, 16-byte aligned), `inflight` float4
 * loads in flight a lane (1, 2, 4 or 8), nontemporal = 1 for nontemporal loads and stores, `blocks` blocks of 256
 * threads (grid-stride). Queued on the library's stream; -1 for arguments it does not take. */
int fe_roof_f32(const float* a, float* b, size_t n, int inflight, int nontemporal, int read_only, int blocks);

/* ---- the law compiler ------------------------------------------------------------------------------------------- */

typedef struct fe_law fe_law;

/* Compiles law sources (names label errors "name:line: message"; texts are the law files' contents). NULL when the law
 * does not compile, with every error, one a line, in `errors`. The front end the reference's compile_law is
 * (src/field/Law.hpp:204, src/field/Law.cpp:1662), lowered at once to the IR (src/field/ir/LawIR.hpp:537). */
fe_law* fe_law_compile(const char* const* names, const char* const* texts, int count, char* errors, size_t error_bytes);
void fe_law_free(fe_law* law);
int fe_law_field_count(const fe_law* law);
int fe_law_field_id(const fe_law* law, const char* name); /* -1 when absent (Law::field_id, src/field/Law.hpp:193) */
const char* fe_law_field_name(const fe_law* law, int field);
int fe_law_param_count(const fe_law* law);
int fe_law_term_count(const fe_law* law);
/* The IR as JSON in the reference's format (ir::to_json, src/field/ir/LawIR.hpp:542) and as a listing
 * (ir::describe, src/field/ir/LawIR.hpp:544); free with fe_string_free. */
char* fe_law_ir_json(const fe_law* law);
char* fe_law_describe(const fe_law* law);
/* The law written back as law text that compiles to the same IR (shared subtrees as lets). No counterpart in the
 * reference (its lawc prints IR, not law text); the adjoint generator emits its laws through it. */
char* fe_law_text(const fe_law* law);
void fe_string_free(char* s);
/* The generated gfx950 translation unit of a law (CudaSolver::source, src/field/cuda/CudaSolver.hpp:30; lawc --cuda,
 * tools/lawc/main.cpp:221-224); NULL with the reason in `why` when the law uses an op the backend does not run. */
char* fe_law_hip_source(const fe_law* law, char* why, size_t why_bytes);
/* The adjoint of a term (NULL or "": the first) as law text: one step of it is one vector-Jacobian product of one
 * step of the term (ir::adjoint_law, src/field/ir/Adjoint.cpp:27-225; lawc --adjoint, tools/lawc/main.cpp:274-361);
 * NULL with the reason in `why`. Free with fe_string_free. */
char* fe_law_adjoint(const fe_law* law, const char* term, char* why, size_t why_bytes);

/* ---- the solver on the card ----------------------------------------------------------------------------------------- */

typedef struct fe_solver fe_solver;

/* A solver of `law` on an nx*ny*nz lattice of cell size h, already on the card: Solver(law, shape) + use_gpu()
 * (src/field/Solver.hpp:111,156), the backend being HipSolver::create (the drop-in for CudaSolver::create,
 * src/field/cuda/CudaSolver.hpp:24). NULL with the reason in `why` when there is no device or the law cannot run here.
 * Fields start at their declared init. */
fe_solver* fe_solver_create(const fe_law* law, int nx, int ny, int nz, double h, char* why, size_t why_bytes);
void fe_solver_free(fe_solver* s);
/* A whole field from the host (nx*ny*nz floats, x fastest): Solver::fields() written (src/field/Solver.hpp:119-124). */
int fe_solver_set_field(fe_solver* s, int field, const float* values);
/* A whole field to the host, fetched from the card when the readback is lazy (Solver::values, Solver.hpp:135). */
int fe_solver_get_field(fe_solver* s, int field, float* out);
int fe_solver_set_param(fe_solver* s, const char* name, double value); /* -1: no such param (Solver.hpp:143) */
/* One step of dt: every term in order (substeps from max_dt), then every derive (Solver::step, Solver.hpp:149). 0, or
 * nonzero with the reason in fe_solver_error(). Contract breaches are not errors: see fe_solver_violations_json. */
int fe_solver_step(fe_solver* s, double dt);
/* `steps` steps of dt in one call: the state that many fe_solver_step calls leave, violations gathered over all of
 * them. No counterpart in the reference (its callers loop over Solver::step): with lazy readback, pairs of steps that
 * need no contract check and no read-back run as one captured graph, replayed. */
int fe_solver_run(fe_solver* s, double dt, int steps);
/* The forward pass of a `half` law net on the matrix cores (fp16 operands, fp32 accumulation), for a batch of input
 * rows already normalised: sizes[n_sizes] (inputs, hidden..., outputs), act in the order tanh, relu, silu, linear,
 * gelu, weights in the .net layout (per layer W (out x in, row-major) then b), rows n_rows x inputs, out n_rows x
 * outputs (times out_scale). Replaces the reference's tensor-core net kernels (src/field/cuda/CudaNet.cpp:189-277) as
 * a testable entry point; laws reach the same kernel through the solver. Waits for the card. */
int fe_net_half_forward(const int* sizes, int n_sizes, int act, float out_scale, const float* weights, const float* rows,
    long long n_rows, float* out);
int fe_solver_refresh_derived(fe_solver* s);               /* Solver::refresh_derived (Solver.hpp:148) */
int fe_solver_set_lazy_readback(fe_solver* s, int on);     /* Solver::set_lazy_readback (Solver.hpp:138) */
int fe_solver_check_contracts_every(fe_solver* s, int steps); /* Solver::check_contracts_every (Solver.hpp:171) */
int fe_solver_set_learned(fe_solver* s, int on);           /* Solver::set_learned (Solver.hpp:165) */
int fe_solver_set_respect_max_dt(fe_solver* s, int on);    /* Solver::set_respect_max_dt (Solver.hpp:179) */
/* The last step's violations as a JSON array of {kind, term, field, detail} (Solver::violations, Solver.hpp:173). */
char* fe_solver_violations_json(fe_solver* s);
int fe_solver_last_substeps(fe_solver* s, int* out, int count);          /* Solver.hpp:175; returns the term count */
int fe_solver_last_changed(fe_solver* s, unsigned char* out, int count); /* Solver.hpp:151; returns the field count */
/* The last step's StepProfile (src/field/Backend.hpp:43-65) as JSON: card ms per term from events, launches. */
char* fe_solver_profile_json(fe_solver* s);
/* FNV-1a over every field's bits (FieldSet::digest, src/field/Grid.hpp:341-343), fetching first when lazy. */
uint64_t fe_solver_digest(fe_solver* s);
const char* fe_solver_error(fe_solver* s);
/* The device array holding a field's current cells (valid until the next step swaps it). */
void* fe_solver_field_device(fe_solver* s, int field);
/* A whole field from device memory (nx*ny*nz floats) into the solver's current cells on the card, queued on the
 * library's stream; the host copy is fetched before it is next read. For callers that keep states on the card
 * (checkpoints, adjoint sweeps); call it once the solver has uploaded its fields (a step or refresh_derived). No
 * counterpart in the reference, whose states travel through the host (tools/lawc/main.cpp:310-323). */
int fe_solver_set_field_device(fe_solver* s, int field, const void* src);
/* A static field (one no term, learned scheme or derive writes) read from `buffer` (device memory, nx*ny*nz floats)
 * from now on, without a copy; NULL goes back to the solver's own array. For callers that hold states on the card and
 * feed them to a solver in turn (adjoint sweeps over checkpointed states). While bound, whatever writes the field's
 * cells (fe_solver_set_field, a window shift) writes into `buffer`. -1 for a field the law writes. No counterpart in
 * the reference. */
int fe_solver_bind_field_device(fe_solver* s, int field, void* buffer);
/* The window moved by (dx, dy, dz) cells (Solver::shift_window, src/field/Solver.hpp:40-42): every field's host cells
 * slide (what is now at c was at c + d) and the arriving boxes (n_boxes of x0 y0 z0 x1 y1 z1, half-open, inside the
 * lattice) take `arriving`: for each field in order, each box in order, its cells x fastest. Before the next step the
 * card slides its own copy and receives only those boxes (the reference's CUDA backend moves a ring's origin,
 * src/field/cuda/CudaSolver.cpp:895-920). -1: bad arguments or a box outside the lattice. */
int fe_solver_shift_window(fe_solver* s, int dx, int dy, int dz, const int* boxes, int n_boxes, const float* arriving);
/* Every cell of a field set to `value` on the card (no host copy; the host copy goes stale as after a device write).
 * No counterpart in the reference: the reverse-mode driver clears its per-step parameter adjoints with it. */
int fe_solver_fill_field(fe_solver* s, int field, float value);
/* A law net's weights replaced (count = the net's parameter count, the .net layout): what a trainer writes through
 * Solver::law_for_training() (src/field/Solver.hpp:82-84); the next step takes them up (the backend hashes the
 * weights every step, src/field/cuda/CudaSolver.cpp:282-311). -1: no such net or another count. */
int fe_solver_set_net_weights(fe_solver* s, const char* net, const float* weights, size_t count);
/* A field's total and total magnitude as the contracts compute them on the card: fixed-order compensated sums in
 * double (a block a fixed range of cells, the blocks folded in order), so two runs give the same bits; waits for the
 * card. The reference reports the totals its conservation judge reads (src/field/Solver.cpp:31-39,
 * src/field/cuda/CudaSolver.cpp:515-571). magnitude may be NULL. */
int fe_solver_total(fe_solver* s, int field, double* sum, double* magnitude);
/* The same total added on the card into device_acc[0] (sum) and device_acc[1] (magnitude), two doubles in device
 * memory, queued without waiting: a caller summing a field's total over many steps reads the sum once. */
int fe_solver_total_accumulate(fe_solver* s, int field, double* device_acc);

#ifdef __cplusplus
}
#endif

#endif
