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

// Four cells a lane: the kernel the headline step runs when nx is a multiple of 4.
//
// A block is TY + 2 waves over a 256 x TY tile of (x, y): waves 1..TY own a row each (64 lanes x 4 cells, one 128-bit
// load and store a plane), waves 0 and TY + 1 only load the halo rows. Every wave keeps its column's next planes in
// flight: a ring of four plane slots, the loop unrolled four times so the slots are renamed rather than copied (a copy
// of a register a load is still filling would wait for the load). Face fluxes are shared where IEEE arithmetic allows
// it exactly: kf * (a - b) = -(kf * (b - a)) to the bit, so the flux a cell sends through its +x (or +z) face is the
// one its neighbour receives, negated. A face past the lattice's edge is computed against the cell itself:
// kf * (c - c) = +0, and adding +0 to a sum that starts at +0 changes no bit, so the edges need no branches. The sums
// keep the CPU's order: -x, +x, -y, +y, -z, +z.
struct Plane4 {
    float4 v; // this lane's four cells of one plane
    float l;  // the cell left of them (lane 0 only, inside the lattice)
    float r;  // the cell right of them (lane 63 only, inside the lattice)
};

template <int TY, bool POW2, bool XH>
__global__ __launch_bounds__(64 * (TY + 2)) void heat_exact_x4(HeatArgs a)
{
    constexpr int W = 256;
    constexpr int ROW = W + 8; // [3] the left halo, [4, 4 + W) the tile's cells, [4 + W] the right halo
    __shared__ __attribute__((aligned(16))) float tile[2][TY + 2][ROW];
    const int total = a.tiles_x * a.tiles_y * a.tiles_z;
    const int t = a.remap ? xcd_tile(static_cast<int>(blockIdx.x), total) : static_cast<int>(blockIdx.x);
    const int txi = t % a.tiles_x;
    const int tyi = (t / a.tiles_x) % a.tiles_y;
    const int tzi = t / (a.tiles_x * a.tiles_y);
    const int lane = static_cast<int>(threadIdx.x);
    const int r = static_cast<int>(threadIdx.y);
    const int x0 = txi * W + 4 * lane;
    const int y = tyi * TY + r - 1;
    const int z0 = tzi * a.zchunk;
    const int z1 = min(z0 + a.zchunk, a.nz);
    const bool live = x0 < a.nx && r >= 1 && r <= TY && y < a.ny;
    const int xs = min(x0, a.nx - 4);
    const int ys = min(max(y, 0), a.ny - 1);
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const std::size_t plane = static_cast<std::size_t>(nx) * static_cast<std::size_t>(ny);
    const float* base = a.in + static_cast<std::size_t>(ys) * static_cast<std::size_t>(nx) + static_cast<std::size_t>(xs);
    float* outp = a.out + static_cast<std::size_t>(ys) * static_cast<std::size_t>(nx) + static_cast<std::size_t>(xs);
    // Every load is unconditional, from a clamped plane: a load under a branch makes the compiler wait for all
    // loads in flight. Past the top of the lattice the "next" plane is the plane itself, which is the zero-flux
    // edge (the face against itself is +0). Halo columns (only when a row spans several tiles) likewise: lanes
    // other than 0 and 63, and lanes at the lattice's edge, load their own cell, which nothing reads.
    const int zlim = min(z1, nz - 1);
    const int lofs = XH && lane == 0 && xs > 0 ? -1 : 0;
    const int rofs = XH && lane == 63 && xs + 4 < nx ? 4 : 0;
    const double kf = a.kf;

    auto load = [&](Plane4& p, int z) {
        const float* q = base + static_cast<std::size_t>(min(z, zlim)) * plane;
        p.v = *reinterpret_cast<const float4*>(q);
        if (XH) {
            p.l = q[lofs];
            p.r = q[rofs];
        }
    };

    Plane4 A, B, C, D;
    load(A, z0);
    load(B, z0 + 1);
    load(C, z0 + 2);
    const float4 lo = *reinterpret_cast<const float4*>(base + static_cast<std::size_t>(max(z0 - 1, 0)) * plane);
    double c0 = static_cast<double>(A.v.x), c1 = static_cast<double>(A.v.y), c2 = static_cast<double>(A.v.z),
           c3 = static_cast<double>(A.v.w);
    // The flux each cell receives through its -z face at the chunk's first plane (+0 at the bottom of the lattice).
    double fzm0 = kf * (static_cast<double>(lo.x) - c0), fzm1 = kf * (static_cast<double>(lo.y) - c1),
           fzm2 = kf * (static_cast<double>(lo.z) - c2), fzm3 = kf * (static_cast<double>(lo.w) - c3);
    int buf = 0;

    // One plane: `cur` is plane z, `nxt` plane z + 1 (loaded), `far` receives plane z + 3.
    auto step = [&](int z, const Plane4& cur, const Plane4& nxt, Plane4& far) {
        load(far, z + 3);
        const float4 c = cur.v;
        *reinterpret_cast<float4*>(&tile[buf][r][4 + 4 * lane]) = c;
        if (XH) {
            if (lane == 0) {
                tile[buf][r][3] = cur.l;
            }
            if (lane == 63) {
                tile[buf][r][4 + W] = cur.r;
            }
        }
        __syncthreads();
        const float4 hi = nxt.v;
        const double d0 = static_cast<double>(hi.x), d1 = static_cast<double>(hi.y), d2 = static_cast<double>(hi.z),
                     d3 = static_cast<double>(hi.w);
        if (live) {
            // Values first, then the edge selects: a select between an LDS lvalue and a register would make the
            // compiler take addresses (flat loads and scratch).
            const float4 tym = *reinterpret_cast<const float4*>(&tile[buf][r - 1][4 + 4 * lane]);
            const float4 typ = *reinterpret_cast<const float4*>(&tile[buf][r + 1][4 + 4 * lane]);
            const float txm = tile[buf][r][3 + 4 * lane];
            const float txp = tile[buf][r][8 + 4 * lane];
            // The halo waves load a clamped row, so past the lattice's y edges the row beside is the row itself:
            // ym and yp need no select. A row spans every lane only when nx is a whole number of tiles, so x does.
            const float4 ym = tym;
            const float4 yp = typ;
            const float xme = x0 > 0 ? txm : c.x;
            const float xpe = x0 + 4 < nx ? txp : c.w;
            const double fxm = kf * (static_cast<double>(xme) - c0);
            const double f01 = kf * (c1 - c0);
            const double f12 = kf * (c2 - c1);
            const double f23 = kf * (c3 - c2);
            const double fxp = kf * (static_cast<double>(xpe) - c3);
            const double fzp0 = kf * (d0 - c0), fzp1 = kf * (d1 - c1), fzp2 = kf * (d2 - c2), fzp3 = kf * (d3 - c3);
            double q0 = 0.0 + fxm;
            q0 += f01;
            q0 += kf * (static_cast<double>(ym.x) - c0);
            q0 += kf * (static_cast<double>(yp.x) - c0);
            q0 += fzm0;
            q0 += fzp0;
            double q1 = 0.0 + (-f01);
            q1 += f12;
            q1 += kf * (static_cast<double>(ym.y) - c1);
            q1 += kf * (static_cast<double>(yp.y) - c1);
            q1 += fzm1;
            q1 += fzp1;
            double q2 = 0.0 + (-f12);
            q2 += f23;
            q2 += kf * (static_cast<double>(ym.z) - c2);
            q2 += kf * (static_cast<double>(yp.z) - c2);
            q2 += fzm2;
            q2 += fzp2;
            double q3 = 0.0 + (-f23);
            q3 += fxp;
            q3 += kf * (static_cast<double>(ym.w) - c3);
            q3 += kf * (static_cast<double>(yp.w) - c3);
            q3 += fzm3;
            q3 += fzp3;
            if (POW2) {
                q0 *= a.inv_hh;
                q1 *= a.inv_hh;
                q2 *= a.inv_hh;
                q3 *= a.inv_hh;
            } else {
                q0 /= a.hh;
                q1 /= a.hh;
                q2 /= a.hh;
                q3 /= a.hh;
            }
            float4 o;
            o.x = fe_ftz(c.x + fe_rf(static_cast<double>(fe_rf(q0)) * a.sub_dt));
            o.y = fe_ftz(c.y + fe_rf(static_cast<double>(fe_rf(q1)) * a.sub_dt));
            o.z = fe_ftz(c.z + fe_rf(static_cast<double>(fe_rf(q2)) * a.sub_dt));
            o.w = fe_ftz(c.w + fe_rf(static_cast<double>(fe_rf(q3)) * a.sub_dt));
            *reinterpret_cast<float4*>(outp + static_cast<std::size_t>(z) * plane) = o;
            fzm0 = -fzp0;
            fzm1 = -fzp1;
            fzm2 = -fzp2;
            fzm3 = -fzp3;
        }
        buf ^= 1;
        c0 = d0;
        c1 = d1;
        c2 = d2;
        c3 = d3;
    };

    for (int z = z0; z < z1; z += 4) {
        step(z, A, B, D);
        if (z + 1 >= z1) {
            break;
        }
        step(z + 1, B, C, A);
        if (z + 2 >= z1) {
            break;
        }
        step(z + 2, C, D, B);
        if (z + 3 >= z1) {
            break;
        }
        step(z + 3, D, A, C);
    }
}

// The attainable HBM rate: a copy (or a read) of float4s, D loads in flight a lane (grid-stride), ordinary or
// nontemporal accesses. The read form keeps its loads alive through a sum that is stored only in a case that cannot
// occur for the sweep's data (finite floats sum to a finite value, never NaN).
typedef float roof_f4 __attribute__((ext_vector_type(4)));
template <int D, bool NT, bool RO>
__global__ __launch_bounds__(256) void roof_f32x4(const roof_f4* __restrict__ a, roof_f4* __restrict__ b, std::size_t n4,
    float* __restrict__ sink)
{
    const std::size_t stride = static_cast<std::size_t>(gridDim.x) * blockDim.x;
    std::size_t i = blockIdx.x * static_cast<std::size_t>(blockDim.x) + threadIdx.x;
    float acc = 0.0f;
    for (; i + (D - 1) * stride < n4; i += D * stride) {
        roof_f4 v[D];
        for (int u = 0; u < D; ++u) {
            v[u] = NT ? __builtin_nontemporal_load(a + i + u * stride) : a[i + u * stride];
        }
        for (int u = 0; u < D; ++u) {
            if (RO) {
                acc += v[u].x + v[u].y + v[u].z + v[u].w;
            } else if (NT) {
                __builtin_nontemporal_store(v[u], b + i + u * stride);
            } else {
                b[i + u * stride] = v[u];
            }
        }
    }
    for (; i < n4; i += stride) {
        const roof_f4 v = a[i];
        if (RO) {
            acc += v.x + v.y + v.z + v.w;
        } else {
            b[i] = v;
        }
    }
    if (RO && acc != acc) {
        sink[0] = acc;
    }
}

__global__ __launch_bounds__(256) void copy_f32x4(const float4* __restrict__ a, float4* __restrict__ b, std::size_t n4)
{
    // Four 128-bit loads in flight a lane before the first store.
    const std::size_t stride = static_cast<std::size_t>(gridDim.x) * blockDim.x;
    std::size_t i = blockIdx.x * static_cast<std::size_t>(blockDim.x) + threadIdx.x;
    for (; i + 3 * stride < n4; i += 4 * stride) {
        const float4 v0 = a[i];
        const float4 v1 = a[i + stride];
        const float4 v2 = a[i + 2 * stride];
        const float4 v3 = a[i + 3 * stride];
        b[i] = v0;
        b[i + stride] = v1;
        b[i + 2 * stride] = v2;
        b[i + 3 * stride] = v3;
    }
    for (; i < n4; i += stride) {
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
// floats (x fastest), distinct. variant 1: a lane a column, ty rows of 64 columns a block (1, 2, 4 or 8); variant 2
// (nx a multiple of 4): four cells a lane, ty rows of 256 columns a block (2, 6 or 14); variant 0 picks. zchunk:
// planes per block; remap: XCD-aware block order. Returns a HIP status, or -1 for arguments the kernel does not take.
int fe_heat_handwritten_cfg(const float* in, float* out, int nx, int ny, int nz, double k, double h, double sub_dt, int ty,
    int zchunk, int remap, int variant)
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
    if (variant == 2 || (variant == 0 && nx % 4 == 0)) {
        if (nx % 4 != 0) {
            return -1;
        }
        a.tiles_x = (nx + 255) / 256;
        a.tiles_y = (ny + ty - 1) / ty;
        const long long total4 = static_cast<long long>(a.tiles_x) * a.tiles_y * a.tiles_z;
        if (total4 > 0x7fffffffLL) {
            return -1;
        }
        const dim3 grid4(static_cast<unsigned>(total4));
#define FE_X4_NT(TYV, P2, XHV) hipLaunchKernelGGL((heat_exact_x4<TYV, P2, XHV>), grid4, dim3(64, TYV + 2), 0, s, a);
#define FE_X4(TYV)                                                                                                     \
    if (a.tiles_x > 1) {                                                                                               \
        if (a.hh_pow2) {                                                                                               \
            FE_X4_NT(TYV, true, true)                                                                                  \
        } else {                                                                                                       \
            FE_X4_NT(TYV, false, true)                                                                                 \
        }                                                                                                              \
    } else if (a.hh_pow2) {                                                                                            \
        FE_X4_NT(TYV, true, false)                                                                                     \
    } else {                                                                                                           \
        FE_X4_NT(TYV, false, false)                                                                                    \
    }
        switch (ty) {
        case 2: FE_X4(2) break;
        case 6: FE_X4(6) break;
        case 14: FE_X4(14) break;
        default: return -1;
        }
#undef FE_X4
#undef FE_X4_NT
        return static_cast<int>(hipGetLastError());
    }
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
    if (nx % 4 == 0) {
        // Launch shapes measured on MI355X: small rows like many short blocks, long rows fewer, taller ones.
        if (nx <= 256) {
            return fe_heat_handwritten_cfg(in, out, nx, ny, nz, k, h, sub_dt, 2, 32, 1, 2);
        }
        return fe_heat_handwritten_cfg(in, out, nx, ny, nz, k, h, sub_dt, 14, 32, 0, 2);
    }
    return fe_heat_handwritten_cfg(in, out, nx, ny, nz, k, h, sub_dt, 4, 32, 1, 1);
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
        const std::size_t want = (n4 + 1023) / 1024;
        const unsigned blocks = static_cast<unsigned>(want < 1 ? 1 : (want > 65536 ? 65536 : want));
        hipLaunchKernelGGL(copy_f32x4, dim3(blocks), dim3(256), 0, s, reinterpret_cast<const float4*>(a),
            reinterpret_cast<float4*>(b), n4);
    }
    if (n4 * 4 < n) {
        hipLaunchKernelGGL(copy_f32_tail, dim3(1), dim3(4), 0, s, a, b, n4 * 4, n);
    }
    return static_cast<int>(hipGetLastError());
}

int fe_roof_f32(const float* a, float* b, size_t n, int inflight, int nontemporal, int read_only, int blocks)
{
    if (a == nullptr || (b == nullptr && !read_only) || n % 4 != 0 || blocks < 1 || blocks > (1 << 20) ||
        (inflight != 1 && inflight != 2 && inflight != 4 && inflight != 8)) {
        return -1;
    }
    hipStream_t s = static_cast<hipStream_t>(fe_hip_stream());
    const roof_f4* a4 = reinterpret_cast<const roof_f4*>(a);
    roof_f4* b4 = reinterpret_cast<roof_f4*>(b);
    float* sink = b; // written only if a sum is NaN (never for finite data); a read-only run may pass b = NULL
    const std::size_t n4 = n / 4;
    const dim3 g(static_cast<unsigned>(blocks)), t(256);
#define FE_ROOF(D)                                                                                                      \
    if (inflight == D) {                                                                                                \
        if (read_only) {                                                                                                \
            if (nontemporal) hipLaunchKernelGGL((roof_f32x4<D, true, true>), g, t, 0, s, a4, b4, n4, sink);             \
            else hipLaunchKernelGGL((roof_f32x4<D, false, true>), g, t, 0, s, a4, b4, n4, sink);                        \
        } else {                                                                                                        \
            if (nontemporal) hipLaunchKernelGGL((roof_f32x4<D, true, false>), g, t, 0, s, a4, b4, n4, sink);            \
            else hipLaunchKernelGGL((roof_f32x4<D, false, false>), g, t, 0, s, a4, b4, n4, sink);                       \
        }                                                                                                               \
    }
    FE_ROOF(1)
    FE_ROOF(2)
    FE_ROOF(4)
    FE_ROOF(8)
#undef FE_ROOF
    return static_cast<int>(hipGetLastError());
}

} // extern "C"
