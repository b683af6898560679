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
void fe_string_free(char* s);

#ifdef __cplusplus
}
#endif

#endif
