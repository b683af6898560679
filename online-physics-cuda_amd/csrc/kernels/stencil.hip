// Handwritten stencil kernels for gfx950: the bandwidth roof (a copy) and the 7-point conduction step.
//
// The conduction step is the reference's bit-identity kernel (bench/stencil_bench.cpp:47-65) written for CDNA4 and for
// the reference CPU's arithmetic: `d T = diffuse(k, T)` with one k everywhere, one explicit substep, exactly as the CPU
// solver computes it (src/field/Law.cpp:986-1018 for the rate, src/field/Solver.cpp:391-403,419-427 for the store):
//
//   acc  = sum over the faces a cell has, in the order -x +x -y +y -z +z, of kf * (double(T[nb]) - double(T[c]))
//   rate = float(acc / (h*h))                  kf = 2 k k / (k + k), the harmonic mean of one k with itself
//   d    = float(double(rate) * sub_dt)
//   T'   = T + d                               (float, denormals flushed as the CPU's FTZ/DAZ threads do)
//
// Layout: one thread per (x, y) column walks a chunk of z-planes, keeping T at z-1, z, z+1 in registers; the x and y
// neighbours of plane z come from an LDS tile of (TY+2) x (TX+2) floats, double buffered so one barrier a plane
// suffices. A wave is 64 consecutive x: every global load and store of a plane is one coalesced 256-byte row. Blocks
// are numbered so that consecutive tiles land on the same XCD (the hardware deals workgroups round-robin over the 8
// XCDs, each with its own L2), so the halo a tile shares with its neighbour is more often in that XCD's L2.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "exact.hpp"

FE_EXACT_HELPERS

namespace {

constexpr int kXcds = 8;

__device__ __forceinline__ int xcd_tile(int b, int total)
{
    const int per = total / kXcds;
    const int rem = total % kXcds;
    const int xcd = b % kXcds;
    const int local = b / kXcds;
    return xcd < rem ? xcd * (per + 1) + local : rem * (per + 1) + (xcd - rem) * per + local;
}

struct HeatArgs {
    const float* in;
    float* out;
    int nx, ny, nz;
    int zchunk;
    int tiles_x, tiles_y, tiles_z;
    int remap;
    double kf;     // the face conductance: the harmonic mean of k with itself, as the CPU computes it
    double hh;     // h * h
    double inv_hh; // 1 / (h * h) when that is exact (h*h a power of two), else unused
    int hh_pow2;
    double sub_dt;
};

template <int TX, int TY>
__global__ __launch_bounds__(TX * TY) void heat_exact(HeatArgs a)
{
    __shared__ float tile[2][TY + 2][TX + 2];
    const int total = a.tiles_x * a.tiles_y * a.tiles_z;
    const int t = a.remap ? xcd_tile(static_cast<int>(blockIdx.x), total) : static_cast<int>(blockIdx.x);
    const int txi = t % a.tiles_x;
    const int tyi = (t / a.tiles_x) % a.tiles_y;
    const int tzi = t / (a.tiles_x * a.tiles_y);
    const int tx = static_cast<int>(threadIdx.x);
    const int ty = static_cast<int>(threadIdx.y);
    const int x = txi * TX + tx;
    const int y = tyi * TY + ty;
    const int z0 = tzi * a.zchunk;
    const int z1 = min(z0 + a.zchunk, a.nz);
    const bool live = x < a.nx && y < a.ny;
    const int xc = min(x, a.nx - 1);
    const int yc = min(y, a.ny - 1);
    const std::size_t plane = static_cast<std::size_t>(a.nx) * static_cast<std::size_t>(a.ny);
    const std::size_t col0 = static_cast<std::size_t>(yc) * static_cast<std::size_t>(a.nx) + static_cast<std::size_t>(xc);
    const float* col = a.in + col0;
    float lo = z0 > 0 ? col[static_cast<std::size_t>(z0 - 1) * plane] : 0.0f;
    float c = col[static_cast<std::size_t>(z0) * plane];
    float hi = z0 + 1 < a.nz ? col[static_cast<std::size_t>(z0 + 1) * plane] : 0.0f;
    const double kf = a.kf;
    int buf = 0;
    for (int z = z0; z < z1; ++z) {
        const float next_hi = (z + 2 < a.nz && z + 1 < z1) ? col[static_cast<std::size_t>(z + 2) * plane] : 0.0f;
        const std::size_t here = static_cast<std::size_t>(z) * plane;
        tile[buf][ty + 1][tx + 1] = c;
        if (tx == 0) {
            tile[buf][ty + 1][0] = xc > 0 ? col[here - 1] : c;
        }
        if (tx == TX - 1) {
            tile[buf][ty + 1][TX + 1] = xc + 1 < a.nx ? col[here + 1] : c;
        }
        if (ty == 0) {
            tile[buf][0][tx + 1] = yc > 0 ? col[here - static_cast<std::size_t>(a.nx)] : c;
        }
        if (ty == TY - 1) {
            tile[buf][TY + 1][tx + 1] = yc + 1 < a.ny ? col[here + static_cast<std::size_t>(a.nx)] : c;
        }
        __syncthreads();
        if (live) {
            const double cc = static_cast<double>(c);
            const float xm = tile[buf][ty + 1][tx];
            const float xp = tile[buf][ty + 1][tx + 2];
            const float ym = tile[buf][ty][tx + 1];
            const float yp = tile[buf][ty + 2][tx + 1];
            // acc starts at +0 and can never become -0, so adding an exact +0 for a missing face is the same as
            // skipping it: the selects keep the lanes converged.
            double acc = 0.0;
            acc += x > 0 ? kf * (static_cast<double>(xm) - cc) : 0.0;
            acc += x + 1 < a.nx ? kf * (static_cast<double>(xp) - cc) : 0.0;
            acc += y > 0 ? kf * (static_cast<double>(ym) - cc) : 0.0;
            acc += y + 1 < a.ny ? kf * (static_cast<double>(yp) - cc) : 0.0;
            acc += z > 0 ? kf * (static_cast<double>(lo) - cc) : 0.0;
            acc += z + 1 < a.nz ? kf * (static_cast<double>(hi) - cc) : 0.0;
            const double q = a.hh_pow2 ? acc * a.inv_hh : acc / a.hh;
            const float rate = fe_rf(q);
            const float d = fe_rf(static_cast<double>(rate) * a.sub_dt);
            a.out[here + static_cast<std::size_t>(y) * static_cast<std::size_t>(a.nx) + static_cast<std::size_t>(x)] =
                fe_ftz(c + d);
        }
        buf ^= 1;
        lo = c;
        c = hi;
        hi = next_hi;
    }
}

__global__ __launch_bounds__(256) void copy_f32x4(const float4* __restrict__ a, float4* __restrict__ b, std::size_t n4)
{
    for (std::size_t i = blockIdx.x * static_cast<std::size_t>(blockDim.x) + threadIdx.x; i < n4;
         i += static_cast<std::size_t>(gridDim.x) * blockDim.x) {
        b[i] = a[i];
    }
}

__global__ void copy_f32_tail(const float* __restrict__ a, float* __restrict__ b, std::size_t from, std::size_t n)
{
    const std::size_t i = from + threadIdx.x;
    if (i < n) {
        b[i] = a[i];
    }
}

bool pow2(double v)
{
    int e = 0;
    return v > 0.0 && __builtin_frexp(v, &e) == 0.5;
}

} // namespace

extern "C" {

void* fe_hip_stream(void);

// One substep of the handwritten conduction kernel on the library's stream. in/out are device arrays of nx*ny*nz
// floats (x fastest), distinct. ty: rows per block (1, 2, 4 or 8; 64 columns each); zchunk: planes per block;
// remap: XCD-aware block order. Returns a HIP status, or -1 for arguments the kernel does not take.
int fe_heat_handwritten_cfg(const float* in, float* out, int nx, int ny, int nz, double k, double h, double sub_dt, int ty,
    int zchunk, int remap)
{
    if (in == nullptr || out == nullptr || in == out || nx < 1 || ny < 1 || nz < 1 || zchunk < 1 || !(h > 0.0)) {
        return -1;
    }
    HeatArgs a{};
    a.in = in;
    a.out = out;
    a.nx = nx;
    a.ny = ny;
    a.nz = nz;
    a.zchunk = zchunk;
    constexpr int TX = 64;
    a.tiles_x = (nx + TX - 1) / TX;
    a.tiles_y = (ny + ty - 1) / ty;
    a.tiles_z = (nz + zchunk - 1) / zchunk;
    a.remap = remap;
    a.kf = (k + k) > 0.0 ? 2.0 * k * k / (k + k) : 0.0;
    a.hh = h * h;
    a.hh_pow2 = pow2(a.hh) ? 1 : 0;
    a.inv_hh = a.hh_pow2 ? 1.0 / a.hh : 0.0;
    a.sub_dt = sub_dt;
    const long long total = static_cast<long long>(a.tiles_x) * a.tiles_y * a.tiles_z;
    if (total > 0x7fffffffLL) {
        return -1;
    }
    hipStream_t s = static_cast<hipStream_t>(fe_hip_stream());
    const dim3 grid(static_cast<unsigned>(total));
    switch (ty) {
    case 1: hipLaunchKernelGGL((heat_exact<TX, 1>), grid, dim3(TX, 1), 0, s, a); break;
    case 2: hipLaunchKernelGGL((heat_exact<TX, 2>), grid, dim3(TX, 2), 0, s, a); break;
    case 4: hipLaunchKernelGGL((heat_exact<TX, 4>), grid, dim3(TX, 4), 0, s, a); break;
    case 8: hipLaunchKernelGGL((heat_exact<TX, 8>), grid, dim3(TX, 8), 0, s, a); break;
    default: return -1;
    }
    return static_cast<int>(hipGetLastError());
}

int fe_heat_handwritten(const float* in, float* out, int nx, int ny, int nz, double k, double h, double sub_dt)
{
    return fe_heat_handwritten_cfg(in, out, nx, ny, nz, k, h, sub_dt, 4, 32, 1);
}

// b[i] = a[i] for n floats: the bandwidth a plain copy reaches, the roof a stencil's step is measured against.
int fe_copy_f32(const float* a, float* b, size_t n)
{
    if (a == nullptr || b == nullptr) {
        return -1;
    }
    hipStream_t s = static_cast<hipStream_t>(fe_hip_stream());
    const std::size_t n4 = n / 4;
    if (n4 > 0) {
        const unsigned blocks = static_cast<unsigned>(n4 / 256 < 4096 ? (n4 + 255) / 256 : 4096);
        hipLaunchKernelGGL(copy_f32x4, dim3(blocks), dim3(256), 0, s, reinterpret_cast<const float4*>(a),
            reinterpret_cast<float4*>(b), n4);
    }
    if (n4 * 4 < n) {
        hipLaunchKernelGGL(copy_f32_tail, dim3(1), dim3(4), 0, s, a, b, n4 * 4, n);
    }
    return static_cast<int>(hipGetLastError());
}

} // extern "C"
