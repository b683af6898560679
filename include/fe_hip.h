/* fe_hip.h - the device runtime of the Field Engine's MI355X backend, as a C ABI.
 *
 * This is the layer the reference keeps in src/field/cuda/CudaApi.hpp (a dlopen'd CUDA driver + NVRTC). Here it is a thin
 * C-ABI over HIP, hipRTC and hipModule: one device, one non-blocking stream, run-time compilation of generated sources
 * to gfx950 code objects with an on-disk cache, modules, memory, launches, events and graph capture. Every entry point
 * returns 0 on success or a nonzero HIP/hipRTC status; fe_hip_error_string() names it. No entry point throws.
 *
 * Plain pointers and sizes only: a binding (ctypes, cgo, JNI) needs nothing but this header.
 */
#ifndef FE_HIP_H
#define FE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fe_hip_module_st* fe_hip_module;
typedef struct fe_hip_function_st* fe_hip_function;
typedef struct fe_hip_event_st* fe_hip_event;
typedef struct fe_hip_graph_st* fe_hip_graph;

/* What the opened device is. Replaces the fields of `struct Device` (CudaApi.hpp:84-100). */
typedef struct fe_hip_device_info {
    int ok;              /* 1 when a device is open */
    char why[512];       /* when not ok: the reason, in words */
    char name[256];      /* marketing name */
    char arch[64];       /* gcnArchName, e.g. "gfx950:sramecc+:xnack-" */
    int compute_units;   /* CUs (CudaApi.hpp:92 `sms`) */
    size_t hbm_bytes;    /* global memory */
    int lds_bytes;       /* LDS per workgroup */
    int runtime_version; /* HIP runtime version */
} fe_hip_device_info;

/* Opens the first device, once per process, with one non-blocking stream. Replaces device() (CudaApi.hpp:102,
 * CudaApi.cpp:168-271). Safe to call repeatedly; fills `out` (may be NULL). Returns 0 when a device is open. */
int fe_hip_device_open(fe_hip_device_info* out);
/* Makes the device current on the calling thread (Device::bind, CudaApi.hpp:97-98). */
int fe_hip_bind(void);
/* Why no law can run on this machine; "" when one can (CudaSolver::unavailable, CudaSolver.hpp:28). Static storage. */
const char* fe_hip_unavailable(void);
const char* fe_hip_error_string(int status);
/* The library's stream as a hipStream_t (for interop with a caller's own HIP work). */
void* fe_hip_stream(void);

/* Compiles HIP C++ to a gfx950 code object, or takes it from the disk cache keyed by FNV-1a of text + options + arch +
 * runtime version. Replaces compile() (CudaApi.hpp:104-107, CudaApi.cpp:273-329). On success *code is malloc'd
 * (free with fe_hip_free). On failure returns nonzero and *log (malloc'd, may be NULL) holds the compiler's log.
 * The options are passed to hipRTC as given; the runtime adds --offload-arch for the open device (gfx950 when none). */
int fe_hip_compile(const char* source, const char* name, const char* const* options, int n_options, void** code,
    size_t* code_bytes, char** log, int* from_cache);
void fe_hip_free(void* p);

/* Modules (Module, CudaApi.hpp:110-126). */
int fe_hip_module_load(const void* code, size_t bytes, fe_hip_module* out);
int fe_hip_module_unload(fe_hip_module module);
int fe_hip_module_function(fe_hip_module module, const char* name, fe_hip_function* out);
/* What a loaded kernel holds a lane: registers (hipFuncGetAttribute NUM_REGS) and private (scratch) bytes
 * (LOCAL_SIZE_BYTES). No counterpart in the reference: the backend uses it to choose among a term's gathers. */
int fe_hip_function_attributes(fe_hip_function f, int* registers, int* private_bytes);
int fe_hip_module_global(fe_hip_module module, const char* name, void** dptr, size_t* bytes);

/* Device memory (Buffer, CudaApi.hpp:129-145) and page-locked host memory (Pinned, CudaApi.hpp:148-161). */
int fe_hip_mem_alloc(void** dptr, size_t bytes);
int fe_hip_mem_free(void* dptr);
int fe_hip_host_alloc(void** ptr, size_t bytes);
int fe_hip_host_free(void* ptr);

/* Copies and fills on the library's stream (cuMemcpy*Async / cuMemsetD32Async, CudaApi.hpp:52-55). */
int fe_hip_memcpy_htod_async(void* dst, const void* src, size_t bytes);
int fe_hip_memcpy_dtoh_async(void* dst, const void* src, size_t bytes);
int fe_hip_memcpy_dtod_async(void* dst, const void* src, size_t bytes);
int fe_hip_memset_d32_async(void* dst, uint32_t value, size_t count);
int fe_hip_stream_sync(void);

/* Launches f over a grid on the library's stream; args[i] points at the i-th argument's value (launch,
 * CudaApi.hpp:163-165). */
int fe_hip_launch(fe_hip_function f, unsigned gx, unsigned gy, unsigned gz, unsigned bx, unsigned by, unsigned bz,
    unsigned shared_bytes, void** args);

/* Events on the library's stream (cuEvent*, CudaApi.hpp:60-63). */
int fe_hip_event_create(fe_hip_event* out);
int fe_hip_event_destroy(fe_hip_event e);
int fe_hip_event_record(fe_hip_event e);
int fe_hip_event_sync(fe_hip_event e);
int fe_hip_event_elapsed(fe_hip_event start, fe_hip_event end, float* ms);

/* Graphs: everything sent to the library's stream between begin and end becomes one instantiated graph. No
 * counterpart in the reference (it launches every kernel from the host, CudaSolver.cpp:586-661). */
int fe_hip_graph_begin_capture(void);
int fe_hip_graph_end_capture(fe_hip_graph* out);
int fe_hip_graph_launch(fe_hip_graph g);
int fe_hip_graph_destroy(fe_hip_graph g);

/* fp32 -> bf16 bits, round to nearest even (the host side of MFMA weights; replaces to_half, CudaApi.hpp:167-168). */
uint16_t fe_hip_to_bf16(float v);

#ifdef __cplusplus
}
#endif

#endif
