// The device runtime behind include/fe_hip.h: HIP + hipRTC + hipModule, one device, one non-blocking stream.
//
// The reference loads the CUDA driver and NVRTC with dlopen (src/field/cuda/CudaApi.cpp:168-271) so it starts on machines
// without CUDA. A ROCm build links libamdhip64 and libhiprtc directly: both ship with every ROCm install, and without a
// device the runtime answers hipErrorNoDevice, which fe_hip_unavailable() reports in words.
#include "fe_hip.h"

#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <sys/stat.h>
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace {

struct DeviceState {
    std::once_flag once;
    fe_hip_device_info info{};
    hipStream_t stream = nullptr;
    int device = 0;
    std::string arch = "gfx950"; // the code-object target; the device's own when one is open
};

DeviceState& state()
{
    static DeviceState s;
    return s;
}

void open_once()
{
    DeviceState& s = state();
    std::memset(&s.info, 0, sizeof(s.info));
    int count = 0;
    hipError_t r = hipGetDeviceCount(&count);
    if (r != hipSuccess || count < 1) {
        std::snprintf(s.info.why, sizeof(s.info.why), "no HIP device (%s)",
            r != hipSuccess ? hipGetErrorString(r) : "the runtime sees none");
        return;
    }
    r = hipSetDevice(0);
    if (r != hipSuccess) {
        std::snprintf(s.info.why, sizeof(s.info.why), "hipSetDevice(0): %s", hipGetErrorString(r));
        return;
    }
    hipDeviceProp_t prop{};
    r = hipGetDeviceProperties(&prop, 0);
    if (r != hipSuccess) {
        std::snprintf(s.info.why, sizeof(s.info.why), "hipGetDeviceProperties: %s", hipGetErrorString(r));
        return;
    }
    std::snprintf(s.info.name, sizeof(s.info.name), "%s", prop.name);
    std::snprintf(s.info.arch, sizeof(s.info.arch), "%s", prop.gcnArchName);
    s.info.compute_units = prop.multiProcessorCount;
    s.info.hbm_bytes = prop.totalGlobalMem;
    s.info.lds_bytes = static_cast<int>(prop.sharedMemPerBlock);
    (void)hipRuntimeGetVersion(&s.info.runtime_version);
    std::string arch = prop.gcnArchName;
    const std::size_t colon = arch.find(':');
    s.arch = colon == std::string::npos ? arch : arch.substr(0, colon);
    // Default priority on purpose: streams of different priorities have hung work on this hardware.
    r = hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking);
    if (r != hipSuccess) {
        std::snprintf(s.info.why, sizeof(s.info.why), "hipStreamCreate: %s", hipGetErrorString(r));
        return;
    }
    s.info.ok = 1;
}

bool ready()
{
    DeviceState& s = state();
    std::call_once(s.once, open_once);
    return s.info.ok != 0;
}

constexpr int kNoDevice = 100; // hipErrorNoDevice's value in HIP

std::uint64_t fnv1a(const void* data, std::size_t bytes, std::uint64_t h)
{
    const unsigned char* p = static_cast<const unsigned char*>(data);
    for (std::size_t i = 0; i < bytes; ++i) {
        h = (h ^ p[i]) * 1099511628211ull;
    }
    return h;
}

std::string cache_dir()
{
    if (const char* d = std::getenv("FIELD_HIP_CACHE")) {
        return d;
    }
    return "/tmp/field-hip-cache-" + std::to_string(static_cast<long long>(getuid()));
}

bool read_file(const std::string& path, std::vector<char>& out)
{
    FILE* f = std::fopen(path.c_str(), "rb");
    if (f == nullptr) {
        return false;
    }
    std::fseek(f, 0, SEEK_END);
    const long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    out.resize(n > 0 ? static_cast<std::size_t>(n) : 0);
    const bool ok = n > 0 && std::fread(out.data(), 1, out.size(), f) == out.size();
    std::fclose(f);
    return ok;
}

void write_file_atomic(const std::string& path, const std::vector<char>& data)
{
    const std::string tmp = path + ".tmp" + std::to_string(static_cast<long long>(getpid()));
    FILE* f = std::fopen(tmp.c_str(), "wb");
    if (f == nullptr) {
        return;
    }
    const bool ok = std::fwrite(data.data(), 1, data.size(), f) == data.size();
    std::fclose(f);
    if (ok) {
        std::rename(tmp.c_str(), path.c_str());
    } else {
        std::remove(tmp.c_str());
    }
}

char* dup_string(const std::string& s)
{
    char* p = static_cast<char*>(std::malloc(s.size() + 1));
    if (p != nullptr) {
        std::memcpy(p, s.c_str(), s.size() + 1);
    }
    return p;
}

} // namespace

extern "C" {

int fe_hip_device_open(fe_hip_device_info* out)
{
    const bool ok = ready();
    if (out != nullptr) {
        *out = state().info;
    }
    return ok ? 0 : kNoDevice;
}

int fe_hip_bind(void)
{
    if (!ready()) {
        return kNoDevice;
    }
    return static_cast<int>(hipSetDevice(state().device));
}

const char* fe_hip_unavailable(void)
{
    ready();
    return state().info.ok ? "" : state().info.why;
}

const char* fe_hip_error_string(int status)
{
    return hipGetErrorString(static_cast<hipError_t>(status));
}

void* fe_hip_stream(void)
{
    return ready() ? static_cast<void*>(state().stream) : nullptr;
}

int fe_hip_compile(const char* source, const char* name, const char* const* options, int n_options, void** code,
    size_t* code_bytes, char** log, int* from_cache)
{
    if (code == nullptr || code_bytes == nullptr || source == nullptr) {
        return static_cast<int>(hipErrorInvalidValue);
    }
    *code = nullptr;
    *code_bytes = 0;
    if (log != nullptr) {
        *log = nullptr;
    }
    if (from_cache != nullptr) {
        *from_cache = 0;
    }
    ready(); // the device's arch, when there is one
    std::vector<std::string> opts;
    opts.push_back("--offload-arch=" + state().arch);
    for (int i = 0; i < n_options; ++i) {
        opts.emplace_back(options[i]);
    }
    int rt = 0;
    (void)hipRuntimeGetVersion(&rt);
    std::uint64_t key = fnv1a(source, std::strlen(source), 1469598103934665603ull);
    for (const std::string& o : opts) {
        key = fnv1a(o.data(), o.size(), key);
        key = fnv1a("\n", 1, key);
    }
    key = fnv1a(&rt, sizeof(rt), key);
    char hex[32];
    std::snprintf(hex, sizeof(hex), "%016llx", static_cast<unsigned long long>(key));
    const std::string dir = cache_dir();
    const std::string path = dir + "/" + hex + ".hsaco";
    std::vector<char> blob;
    if (read_file(path, blob)) {
        *code = std::malloc(blob.size());
        std::memcpy(*code, blob.data(), blob.size());
        *code_bytes = blob.size();
        if (from_cache != nullptr) {
            *from_cache = 1;
        }
        return 0;
    }
    hiprtcProgram prog = nullptr;
    hiprtcResult r = hiprtcCreateProgram(&prog, source, name != nullptr ? name : "law.hip", 0, nullptr, nullptr);
    if (r != HIPRTC_SUCCESS) {
        if (log != nullptr) {
            *log = dup_string(std::string("hiprtcCreateProgram: ") + hiprtcGetErrorString(r));
        }
        return 1000 + static_cast<int>(r);
    }
    std::vector<const char*> argv;
    for (const std::string& o : opts) {
        argv.push_back(o.c_str());
    }
    r = hiprtcCompileProgram(prog, static_cast<int>(argv.size()), argv.data());
    std::size_t log_size = 0;
    hiprtcGetProgramLogSize(prog, &log_size);
    std::string text(log_size, '\0');
    if (log_size > 0) {
        hiprtcGetProgramLog(prog, text.data());
    }
    if (r != HIPRTC_SUCCESS) {
        if (log != nullptr) {
            *log = dup_string(text.empty() ? std::string(hiprtcGetErrorString(r)) : text);
        }
        hiprtcDestroyProgram(&prog);
        return 1000 + static_cast<int>(r);
    }
    std::size_t size = 0;
    hiprtcGetCodeSize(prog, &size);
    blob.resize(size);
    hiprtcGetCode(prog, blob.data());
    hiprtcDestroyProgram(&prog);
    mkdir(dir.c_str(), 0700);
    write_file_atomic(path, blob);
    *code = std::malloc(blob.size());
    std::memcpy(*code, blob.data(), blob.size());
    *code_bytes = blob.size();
    if (log != nullptr && !text.empty()) {
        *log = dup_string(text);
    }
    return 0;
}

void fe_hip_free(void* p)
{
    std::free(p);
}

int fe_hip_module_load(const void* code, size_t bytes, fe_hip_module* out)
{
    (void)bytes;
    if (!ready()) {
        return kNoDevice;
    }
    hipModule_t m = nullptr;
    const hipError_t r = hipModuleLoadData(&m, code);
    *out = reinterpret_cast<fe_hip_module>(m);
    return static_cast<int>(r);
}

int fe_hip_module_unload(fe_hip_module module)
{
    return module == nullptr ? 0 : static_cast<int>(hipModuleUnload(reinterpret_cast<hipModule_t>(module)));
}

int fe_hip_module_function(fe_hip_module module, const char* name, fe_hip_function* out)
{
    hipFunction_t f = nullptr;
    const hipError_t r = hipModuleGetFunction(&f, reinterpret_cast<hipModule_t>(module), name);
    *out = reinterpret_cast<fe_hip_function>(f);
    return static_cast<int>(r);
}

int fe_hip_function_attributes(fe_hip_function f, int* registers, int* private_bytes)
{
    if (f == nullptr || registers == nullptr || private_bytes == nullptr) {
        return static_cast<int>(hipErrorInvalidValue);
    }
    const hipFunction_t h = reinterpret_cast<hipFunction_t>(f);
    hipError_t r = hipFuncGetAttribute(registers, HIP_FUNC_ATTRIBUTE_NUM_REGS, h);
    if (r == hipSuccess) {
        r = hipFuncGetAttribute(private_bytes, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, h);
    }
    return static_cast<int>(r);
}

int fe_hip_module_global(fe_hip_module module, const char* name, void** dptr, size_t* bytes)
{
    hipDeviceptr_t p = nullptr;
    std::size_t n = 0;
    const hipError_t r = hipModuleGetGlobal(&p, &n, reinterpret_cast<hipModule_t>(module), name);
    *dptr = p;
    if (bytes != nullptr) {
        *bytes = n;
    }
    return static_cast<int>(r);
}

int fe_hip_mem_alloc(void** dptr, size_t bytes)
{
    if (!ready()) {
        return kNoDevice;
    }
    return static_cast<int>(hipMalloc(dptr, bytes == 0 ? 4 : bytes));
}

int fe_hip_mem_free(void* dptr)
{
    return dptr == nullptr ? 0 : static_cast<int>(hipFree(dptr));
}

int fe_hip_host_alloc(void** ptr, size_t bytes)
{
    if (!ready()) {
        return kNoDevice;
    }
    return static_cast<int>(hipHostMalloc(ptr, bytes == 0 ? 4 : bytes, hipHostMallocDefault));
}

int fe_hip_host_free(void* ptr)
{
    return ptr == nullptr ? 0 : static_cast<int>(hipHostFree(ptr));
}

int fe_hip_memcpy_htod_async(void* dst, const void* src, size_t bytes)
{
    return static_cast<int>(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, state().stream));
}

int fe_hip_memcpy_dtoh_async(void* dst, const void* src, size_t bytes)
{
    return static_cast<int>(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, state().stream));
}

int fe_hip_memcpy_dtod_async(void* dst, const void* src, size_t bytes)
{
    return static_cast<int>(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, state().stream));
}

int fe_hip_memset_d32_async(void* dst, uint32_t value, size_t count)
{
    return static_cast<int>(hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(dst), static_cast<int>(value), count,
        state().stream));
}

int fe_hip_stream_sync(void)
{
    if (!ready()) {
        return kNoDevice;
    }
    return static_cast<int>(hipStreamSynchronize(state().stream));
}

int fe_hip_launch(fe_hip_function f, unsigned gx, unsigned gy, unsigned gz, unsigned bx, unsigned by, unsigned bz,
    unsigned shared_bytes, void** args)
{
    return static_cast<int>(hipModuleLaunchKernel(reinterpret_cast<hipFunction_t>(f), gx, gy, gz, bx, by, bz, shared_bytes,
        state().stream, args, nullptr));
}

int fe_hip_event_create(fe_hip_event* out)
{
    if (!ready()) {
        return kNoDevice;
    }
    hipEvent_t e = nullptr;
    const hipError_t r = hipEventCreate(&e);
    *out = reinterpret_cast<fe_hip_event>(e);
    return static_cast<int>(r);
}

int fe_hip_event_destroy(fe_hip_event e)
{
    return e == nullptr ? 0 : static_cast<int>(hipEventDestroy(reinterpret_cast<hipEvent_t>(e)));
}

int fe_hip_event_record(fe_hip_event e)
{
    return static_cast<int>(hipEventRecord(reinterpret_cast<hipEvent_t>(e), state().stream));
}

int fe_hip_event_sync(fe_hip_event e)
{
    return static_cast<int>(hipEventSynchronize(reinterpret_cast<hipEvent_t>(e)));
}

int fe_hip_event_elapsed(fe_hip_event start, fe_hip_event end, float* ms)
{
    return static_cast<int>(hipEventElapsedTime(ms, reinterpret_cast<hipEvent_t>(start), reinterpret_cast<hipEvent_t>(end)));
}

int fe_hip_graph_begin_capture(void)
{
    if (!ready()) {
        return kNoDevice;
    }
    return static_cast<int>(hipStreamBeginCapture(state().stream, hipStreamCaptureModeRelaxed));
}

int fe_hip_graph_end_capture(fe_hip_graph* out)
{
    hipGraph_t graph = nullptr;
    hipError_t r = hipStreamEndCapture(state().stream, &graph);
    if (r != hipSuccess) {
        return static_cast<int>(r);
    }
    hipGraphExec_t exec = nullptr;
    r = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    *out = reinterpret_cast<fe_hip_graph>(exec);
    return static_cast<int>(r);
}

int fe_hip_graph_launch(fe_hip_graph g)
{
    return static_cast<int>(hipGraphLaunch(reinterpret_cast<hipGraphExec_t>(g), state().stream));
}

int fe_hip_graph_destroy(fe_hip_graph g)
{
    return g == nullptr ? 0 : static_cast<int>(hipGraphExecDestroy(reinterpret_cast<hipGraphExec_t>(g)));
}

uint16_t fe_hip_to_bf16(float v)
{
    std::uint32_t bits = 0;
    std::memcpy(&bits, &v, 4);
    if ((bits & 0x7f800000u) == 0x7f800000u && (bits & 0x007fffffu) != 0u) {
        return static_cast<std::uint16_t>((bits >> 16) | 0x0040u); // a NaN stays a (quiet) NaN
    }
    const std::uint32_t lsb = (bits >> 16) & 1u;
    bits += 0x7fffu + lsb;
    return static_cast<std::uint16_t>(bits >> 16);
}

} // extern "C"
