// A learned flux correction for a coarse 2-D two-field lattice, inference and training on gfx950 bf16 MFMA.
//
// The net: 18 inputs (the 3x3 patch of both fields around a cell, in the frame of one of its faces) -> 64 -> 64 -> 2,
// GELU (tanh form: 0.5 z (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3)))) after each hidden layer; bf16 weights and
// activations into the MFMAs, fp32 accumulation, fp32 biases. The flux of each field from cell c across its +d face
// (d = x, y) into c' is
//     F = s * (g(P) - g(mirror P')) / 2
// with P the patch around c in the face's frame (a along d, b across), P' the patch around c' and mirror the flip
// a -> -a: a field mirrored about the face flips F, and a flat field gives F = 0 (g(P) = g(mirror P') there). A cell's
// correction is (inflow - outflow) / h over its faces, the lattice's edges closed: conservative by construction.
// This architecture has no counterpart in the reference (its `flux` op sees the two cells of a face and uses output 0,
// src/field/Law.cpp:691-707); the reference's nets are inference-only fp16 mma (src/field/cuda/CudaNet.cpp:88-277).
//
// MFMA layout (v_mfma_f32_16x16x32_bf16, wave64): A is 16 x 32 (lane l: row l % 16), B is 32 x 16 (lane l: column
// l % 16), each lane 8 bf16 of k from lane group g = l / 16; D is 16 x 16, lane l: column l % 16, rows 4g .. 4g + 3.
// The product only needs A and B to give a lane's 8 k-values the same features, so k-chunk c is taken as features
//     F(c, g, e) = 32c + 4g + (e & 3) + 16 (e >> 2)
// which is exactly what a lane holds of the D blocks 2c and 2c + 1 of the layer before (rows 4g..4g+3 of each): the
// activations go from one layer's D registers into the next layer's B operand without moving. Layers run transposed
// (H^T = W^T X^T: A = weights, B = activations, columns = 16 evaluation rows).
#include "fe_learned.h"
#include "fe_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

constexpr int kIn = 18;
constexpr int kHid = 64;
constexpr int kOut = 2;
// parameters, in this order: W1 (kIn x kHid, row-major: [in][out]), b1, W2 (kHid x kHid), b2, W3 (kHid x kOut), b3
constexpr int kW1 = 0, kB1 = kW1 + kIn * kHid, kW2 = kB1 + kHid, kB2 = kW2 + kHid * kHid, kW3 = kB2 + kHid, kB3 = kW3 + kHid * kOut;
constexpr int kParams = kB3 + kOut; // 5506

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;

// Fragments (A operands, bf16x8 a lane, 64 lanes a fragment):
//   forward  W1^T: 4 (out blocks) x 1 (k chunk);  W2^T: 4 x 2;  W3^T: 1 x 2        -> 14
//   backward W3:   4 (in blocks) x 1 (k chunk of outputs); W2: 4 x 2              -> 12
constexpr int kFragF1 = 0, kFragF2 = 4, kFragF3 = 12, kFragB3 = 14, kFragB2 = 18, kFrags = 26;

__device__ __host__ __forceinline__ int kfeat(int c, int g, int e) { return 32 * c + 4 * g + (e & 3) + 16 * (e >> 2); }

__device__ __host__ __forceinline__ unsigned short to_bf16_bits(float f)
{
    unsigned int u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) {
        return static_cast<unsigned short>((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u)); // inf stays, NaN stays quiet
    }
    u += 0x7fffu + ((u >> 16) & 1u); // round to nearest even
    return static_cast<unsigned short>(u >> 16);
}

// Round to nearest even: the hardware conversion (v_cvt_pk_bf16_f32 on gfx950), what to_bf16_bits computes by bits.
__device__ __forceinline__ __bf16 bf(float f) { return static_cast<__bf16>(f); }
__device__ __forceinline__ float unbf(__bf16 b) { return __builtin_bit_cast(float, static_cast<unsigned int>(__builtin_bit_cast(unsigned short, b)) << 16); }

// GELU, tanh form: 0.5 z (1 + tanh u), u = sqrt(2 / pi) (z + 0.044715 z^3), computed as z s with s = 1 / (1 + e^(-2u))
// (the same function: 0.5 (1 + tanh u) is the logistic of 2u), e^(-2u) as 2^(z (a + b z^2)) with the constants folded,
// through the hardware exponential and reciprocal (v_exp_f32, v_rcp_f32: a few ulp; the activations it feeds are rounded
// to bf16, 8 bits). Eight instructions where the library tanhf is a long polynomial sequence: the net's forward and
// backward passes are bound by these, not by the matrix cores. Saturates to z (e^(-2u) -> 0) and to -0 (-> inf).
constexpr float kGeluA = -2.0f * 0.7978845608028654f * 1.4426950408889634f;             // -2 sqrt(2/pi) log2(e)
constexpr float kGeluB = -2.0f * 0.7978845608028654f * 0.044715f * 1.4426950408889634f; // the same times 0.044715

__device__ __forceinline__ float gelu_logistic(float z)
{
    const float e = __builtin_amdgcn_exp2f(z * (kGeluA + kGeluB * (z * z)));
    return __builtin_amdgcn_rcpf(1.0f + e);
}

__device__ __forceinline__ float gelu(float z) { return z * gelu_logistic(z); }

// d/dz z s(z) = s + z s (1 - s) 2 u'(z), u' = sqrt(2 / pi) (1 + 3 0.044715 z^2).
__device__ __forceinline__ float gelu_grad(float z)
{
    const float k0 = 0.7978845608028654f;
    const float sg = gelu_logistic(z);
    return sg + z * sg * (1.0f - sg) * (2.0f * k0) * (1.0f + 3.0f * 0.044715f * (z * z));
}

__device__ __forceinline__ floatx4 mfma(const bf16x8& a, const bf16x8& b, const floatx4& c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Packs the fp32 master parameters into bf16 fragments (one block of 64 lanes a fragment).
__global__ void k_pack(const float* __restrict__ p, bf16x8* __restrict__ frags)
{
    const int f = (int)blockIdx.x;
    const int l = (int)threadIdx.x;
    const int g = l >> 4, m = l & 15;
    bf16x8 out;
    for (int e = 0; e < 8; ++e) {
        float v = 0.0f;
        if (f < kFragF2) { // W1^T: row = out 16 ob + m, k = input F(0, g, e)
            const int ob = f - kFragF1, in = kfeat(0, g, e), o = 16 * ob + m;
            v = in < kIn ? p[kW1 + in * kHid + o] : 0.0f;
        } else if (f < kFragF3) { // W2^T: row = out 16 ob + m, k = hidden F(c, g, e)
            const int q = f - kFragF2, ob = q >> 1, c = q & 1, in = kfeat(c, g, e), o = 16 * ob + m;
            v = p[kW2 + in * kHid + o];
        } else if (f < kFragB3) { // W3^T: row = output m, k = hidden F(c, g, e)
            const int c = f - kFragF3, in = kfeat(c, g, e);
            v = m < kOut ? p[kW3 + in * kOut + m] : 0.0f;
        } else if (f < kFragB2) { // W3: row = hidden 16 ib + m, k = output F(0, g, e)
            const int ib = f - kFragB3, o = kfeat(0, g, e), in = 16 * ib + m;
            v = o < kOut ? p[kW3 + in * kOut + o] : 0.0f;
        } else { // W2: row = hidden-in 16 ib + m, k = hidden-out F(c, g, e)
            const int q = f - kFragB2, ib = q >> 1, c = q & 1, o = kfeat(c, g, e), in = 16 * ib + m;
            v = p[kW2 + in * kHid + o];
        }
        out[e] = bf(v);
    }
    frags[f * 64 + l] = out;
}

struct Lattice {
    int nx, ny, count; // count lattices of nx x ny, u then v, each field count * nx * ny floats
    float h;
};

// (lattice, x, y) of cell index i of count lattices of nx x ny, in 32-bit arithmetic (a 64-bit division is a long
// software sequence on the card; the host keeps 2 * count * nx * ny below 2^31).
__device__ __forceinline__ void cell_of(long long i, int nx, int ny, long long& lat, int& x, int& y)
{
    const unsigned q = (unsigned)i, cells = (unsigned)nx * (unsigned)ny;
    const unsigned l = q / cells, c = q - l * cells;
    const unsigned yy = c / (unsigned)nx;
    lat = (long long)l;
    y = (int)yy;
    x = (int)(c - yy * (unsigned)nx);
}

// Input k of evaluation row `row` (a face evaluation: (lattice, cell, direction) * 2 + side) at feature in (< 18).
__device__ __forceinline__ float patch_value(const float* __restrict__ u, const float* __restrict__ v, const Lattice& L, long long fe,
    int side, int in)
{
    const long long cells = (long long)L.nx * L.ny;
    const int dir = (int)(fe & 1);
    long long lat;
    int x, y;
    cell_of(fe >> 1, L.nx, L.ny, lat, x, y);
    const int field = in / 9, a = (in % 9) / 3 - 1, b = in % 3 - 1;
    // the anchor: the cell itself (side 0) or the one across the face, read mirrored (a -> -a)
    int da = a, db = b;
    if (side) {
        if (dir == 0) {
            x += 1;
        } else {
            y += 1;
        }
        da = -a;
    }
    int xx = dir == 0 ? x + da : x + db;
    int yy = dir == 0 ? y + db : y + da;
    xx = min(max(xx, 0), L.nx - 1);
    yy = min(max(yy, 0), L.ny - 1);
    const float* f = field == 0 ? u : v;
    return f[lat * cells + (long long)yy * L.nx + xx];
}

// The forward pass of 16 evaluation rows (8 faces, direct and mirror) held by one wave. Returns, in lane n < 16 of
// group 0, outputs 0 and 1 of row n (o0, o1); keeps what the backward pass needs when asked.
struct Acts {
    floatx4 z1[4], z2[4];
    bf16x8 x, h1[2], h2[2];
};

__device__ __forceinline__ void forward16(const bf16x8* __restrict__ frags, const float* __restrict__ params, const float* u,
    const float* v, const Lattice& L, long long fe0, long long fe_end, Acts& s, float& o0, float& o1)
{
    const int l = (int)(threadIdx.x & 63u);
    const int g = l >> 4, n = l & 15;
    const long long fe = fe0 + (n >> 1);
    const int side = n & 1;
    // B operand of layer 1: inputs F(0, g, e) of row n (0 past the batch's end or past 18)
    for (int e = 0; e < 8; ++e) {
        const int in = kfeat(0, g, e);
        const float val = (in < kIn && fe < fe_end) ? patch_value(u, v, L, fe, side, in) : 0.0f;
        s.x[e] = bf(val);
    }
    const floatx4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int ob = 0; ob < 4; ++ob) {
        floatx4 z = mfma(frags[(kFragF1 + ob) * 64 + l], s.x, zero);
        for (int r = 0; r < 4; ++r) {
            z[r] += params[kB1 + 16 * ob + 4 * g + r];
        }
        s.z1[ob] = z;
    }
    for (int c = 0; c < 2; ++c) {
        for (int e = 0; e < 4; ++e) {
            s.h1[c][e] = bf(gelu(s.z1[2 * c][e]));
            s.h1[c][e + 4] = bf(gelu(s.z1[2 * c + 1][e]));
        }
    }
    for (int ob = 0; ob < 4; ++ob) {
        floatx4 z = mfma(frags[(kFragF2 + 2 * ob) * 64 + l], s.h1[0], zero);
        z = mfma(frags[(kFragF2 + 2 * ob + 1) * 64 + l], s.h1[1], z);
        for (int r = 0; r < 4; ++r) {
            z[r] += params[kB2 + 16 * ob + 4 * g + r];
        }
        s.z2[ob] = z;
    }
    for (int c = 0; c < 2; ++c) {
        for (int e = 0; e < 4; ++e) {
            s.h2[c][e] = bf(gelu(s.z2[2 * c][e]));
            s.h2[c][e + 4] = bf(gelu(s.z2[2 * c + 1][e]));
        }
    }
    floatx4 y = mfma(frags[(kFragF3 + 0) * 64 + l], s.h2[0], zero);
    y = mfma(frags[(kFragF3 + 1) * 64 + l], s.h2[1], y);
    o0 = y[0] + params[kB3 + 0];
    o1 = y[1] + params[kB3 + 1];
}

// Face fluxes of every face evaluation: F[fe][field] = s (g_direct - g_mirror) / 2 (0 across the lattice's edge).
// A wave takes 16 rows (8 faces) at a time.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_faces(const bf16x8* __restrict__ frags, const float* __restrict__ params, const float* __restrict__ u,
    const float* __restrict__ v, Lattice L, float scale, long long faces, float* __restrict__ F)
{
    const int l = (int)(threadIdx.x & 63u);
    const int n = l & 15, g = l >> 4;
    const long long wave = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long long waves = ((long long)gridDim.x * blockDim.x) >> 6;
    for (long long t = wave; t * 8 < faces; t += waves) {
        Acts s;
        float o0, o1;
        forward16(frags, params, u, v, L, t * 8, faces, s, o0, o1);
        const float m0 = __shfl_down(o0, 1, 64), m1 = __shfl_down(o1, 1, 64);
        const long long fe = t * 8 + (n >> 1);
        if (g == 0 && (n & 1) == 0 && fe < faces) {
            const int dir = (int)(fe & 1);
            long long lat;
            int x, y;
            cell_of(fe >> 1, L.nx, L.ny, lat, x, y);
            const bool open = dir == 0 ? x + 1 < L.nx : y + 1 < L.ny;
            F[2 * fe] = open ? 0.5f * scale * (o0 - m0) : 0.0f;
            F[2 * fe + 1] = open ? 0.5f * scale * (o1 - m1) : 0.0f;
        }
    }
}

// A cell's correction from its faces' fluxes: (inflow - outflow) / h, field by field.
__device__ __forceinline__ void correction(const float* __restrict__ F, const Lattice& L, long long lat, int x, int y, float& cu, float& cv)
{
    const long long cells = (long long)L.nx * L.ny;
    const long long c = lat * cells + (long long)y * L.nx + x;
    float au = -F[2 * (2 * c)] - F[2 * (2 * c + 1)];
    float av = -F[2 * (2 * c) + 1] - F[2 * (2 * c + 1) + 1];
    if (x > 0) {
        au += F[2 * (2 * (c - 1))];
        av += F[2 * (2 * (c - 1)) + 1];
    }
    if (y > 0) {
        au += F[2 * (2 * (c - L.nx) + 1)];
        av += F[2 * (2 * (c - L.nx) + 1) + 1];
    }
    cu = au / L.h;
    cv = av / L.h;
}

struct Reaction {
    float du, dv, k; // du = Du lap u + u - u^3 - k - v, dv = Dv lap v + u - v
};

// The plain law's rates at a cell (5-point Laplacian, the edge cell repeating: zero flux), in fp32.
__device__ __forceinline__ void plain(const float* __restrict__ u, const float* __restrict__ v, const Lattice& L, const Reaction& R,
    long long lat, int x, int y, float& ru, float& rv)
{
    const long long base = lat * (long long)L.nx * L.ny;
    auto at = [&](const float* f, int xx, int yy) {
        xx = min(max(xx, 0), L.nx - 1);
        yy = min(max(yy, 0), L.ny - 1);
        return f[base + (long long)yy * L.nx + xx];
    };
    const float uc = at(u, x, y), vc = at(v, x, y);
    const float ih2 = 1.0f / (L.h * L.h);
    const float lu = (at(u, x - 1, y) + at(u, x + 1, y) + at(u, x, y - 1) + at(u, x, y + 1) - 4.0f * uc) * ih2;
    const float lv = (at(v, x - 1, y) + at(v, x + 1, y) + at(v, x, y - 1) + at(v, x, y + 1) - 4.0f * vc) * ih2;
    ru = R.du * lu + uc - uc * uc * uc - R.k - vc;
    rv = R.dv * lv + uc - vc;
}

// One step: u' = u + dt (plain + correction) (correction off: the plain law), out of place. A cell that comes out
// non-finite or beyond `bound` raises the breach counter.
__global__ void k_step(const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ F, Lattice L, Reaction R, float dt,
    int learned, float bound, float* __restrict__ u2, float* __restrict__ v2, unsigned* __restrict__ breach)
{
    const long long cells = (long long)L.nx * L.ny;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cells * L.count) {
        return;
    }
    long long lat;
    int x, y;
    cell_of(i, L.nx, L.ny, lat, x, y);
    float ru, rv;
    plain(u, v, L, R, lat, x, y, ru, rv);
    if (learned) {
        float cu, cv;
        correction(F, L, lat, x, y, cu, cv);
        ru += cu;
        rv += cv;
    }
    const float nu = u[i] + dt * ru, nv = v[i] + dt * rv;
    u2[i] = nu;
    v2[i] = nv;
    const bool bad = !(fabsf(nu) <= bound) || !(fabsf(nv) <= bound);
    if (bad && breach != nullptr) {
        atomicAdd(breach, 1u);
    }
}

// The fallback: where the learned step breached, the plain law's step replaces it (the breach count is left for the
// host to read); a graph can hold it, it decides on the card.
__global__ void k_fallback(const float* __restrict__ u, const float* __restrict__ v, Lattice L, Reaction R, float dt, const unsigned* __restrict__ breach,
    float* __restrict__ u2, float* __restrict__ v2, unsigned* __restrict__ fallbacks)
{
    if (*breach == 0u) {
        return;
    }
    const long long cells = (long long)L.nx * L.ny;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        atomicAdd(fallbacks, 1u);
    }
    if (i >= cells * L.count) {
        return;
    }
    long long lat;
    int x, y;
    cell_of(i, L.nx, L.ny, lat, x, y);
    float ru, rv;
    plain(u, v, L, R, lat, x, y, ru, rv);
    u2[i] = u[i] + dt * ru;
    v2[i] = v[i] + dt * rv;
}

__global__ void k_clear(unsigned* p, int n)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) {
        p[i] = 0u;
    }
}

// Training, 1: the loss L = mean over cells and fields of (correction - target)^2 with target = (s1 - s0) / dt -
// plain(s0); e = dL/dcorrection into E (two a cell), and the loss's terms into loss_parts (one a cell, summed in order
// by k_sum).
__global__ void k_residual(const float* __restrict__ u0, const float* __restrict__ v0, const float* __restrict__ u1, const float* __restrict__ v1,
    const float* __restrict__ F, Lattice L, Reaction R, float dt, float weight, float* __restrict__ E, float* __restrict__ loss_parts)
{
    const long long cells = (long long)L.nx * L.ny;
    const long long n = cells * L.count;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) {
        return;
    }
    long long lat;
    int x, y;
    cell_of(i, L.nx, L.ny, lat, x, y);
    float ru, rv, cu, cv;
    plain(u0, v0, L, R, lat, x, y, ru, rv);
    correction(F, L, lat, x, y, cu, cv);
    const float tu = (u1[i] - u0[i]) / dt - ru, tv = (v1[i] - v0[i]) / dt - rv;
    const float du = cu - tu, dv = cv - tv;
    const float inv = weight / (2.0f * (float)n);
    E[2 * i] = 2.0f * du * inv;
    E[2 * i + 1] = 2.0f * dv * inv;
    loss_parts[i] = (du * du + dv * dv) * inv;
}

// Training, 2: dL/dg of every evaluation row from E (through the faces' divergence and the antisymmetric pair).
__device__ __forceinline__ void row_grad(const float* __restrict__ E, const Lattice& L, float scale, long long fe, int side, float& d0, float& d1)
{
    const int dir = (int)(fe & 1);
    const long long c = fe >> 1;
    long long lat;
    int x, y;
    cell_of(c, L.nx, L.ny, lat, x, y);
    const bool open = dir == 0 ? x + 1 < L.nx : y + 1 < L.ny;
    if (!open) {
        d0 = d1 = 0.0f;
        return;
    }
    const long long c2 = c + (dir == 0 ? 1 : L.nx);
    // F leaves c (-F/h) and enters c2 (+F/h); F = s (g_d - g_m) / 2
    const float df0 = (E[2 * c2] - E[2 * c]) / L.h, df1 = (E[2 * c2 + 1] - E[2 * c + 1]) / L.h;
    const float k = side ? -0.5f * scale : 0.5f * scale;
    d0 = k * df0;
    d1 = k * df1;
}

// Training, 3: forward again and backward, 32 rows (two 16-row tiles) at a time a wave; weight gradients by MFMA over
// the rows (LDS-transposed tiles), summed in a fixed order into this wave's partial (G[wave][kParams]).
// One wave a block: the block barriers below order a wave's own LDS tiles, so every wave of a block must take the same
// number of trips; with one wave a block that holds by construction.
constexpr int kWavesTrain = 1;

__global__ __launch_bounds__(64 * kWavesTrain) void k_backward(const bf16x8* __restrict__ frags, const float* __restrict__ params,
    const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ E, Lattice L, float scale, long long faces,
    float* __restrict__ G)
{
    // A wave's LDS: rows 0..31 of the pair, features transposed ([feature][row], bf16).
    struct Tiles {
        __bf16 x[32][32];   // inputs (padded to 32)
        __bf16 h1[64][32];
        __bf16 dz1[64][32];
        __bf16 h2[64][32];
        __bf16 dz2[64][32];
        __bf16 dy[16][32];  // output gradients (padded to 16)
    };
    __shared__ Tiles tiles[kWavesTrain];
    const int w = (int)(threadIdx.x >> 6);
    Tiles& T = tiles[w];
    const int l = (int)(threadIdx.x & 63u);
    const int g = l >> 4, n = l & 15;
    const long long wave = (long long)blockIdx.x * kWavesTrain + w;
    const long long waves = (long long)gridDim.x * kWavesTrain;
    const floatx4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
    floatx4 gw1[2][4], gw2[4][4], gw3[4];
    for (int a = 0; a < 4; ++a) {
        gw3[a] = zero;
        for (int b = 0; b < 4; ++b) {
            gw2[a][b] = zero;
        }
        gw1[0][a] = zero;
        gw1[1][a] = zero;
    }
    float gb1 = 0.0f, gb2 = 0.0f, gb3 = 0.0f; // lane l: feature l (b3: output l < 2)
    for (long long t = wave; t * 16 < faces; t += waves) {
        // rows 32 of the pair: tiles 2t and 2t + 1 (a face is 2 rows: 8 faces a tile)
        for (int half = 0; half < 2; ++half) {
            const long long fe0 = (2 * t + half) * 8;
            Acts s;
            float o0, o1;
            forward16(frags, params, u, v, L, fe0, faces, s, o0, o1);
            (void)o0;
            (void)o1;
            // output gradients of the 16 rows: B operand of dH2 = W3 dY (lane group 0, elements 0, 1)
            const long long fe = fe0 + (n >> 1);
            float d0 = 0.0f, d1 = 0.0f;
            if (fe < faces) {
                row_grad(E, L, scale, fe, n & 1, d0, d1);
            }
            bf16x8 dyb;
            for (int e = 0; e < 8; ++e) {
                dyb[e] = bf(0.0f);
            }
            if (g == 0) {
                dyb[0] = bf(d0);
                dyb[1] = bf(d1);
            }
            floatx4 dz2[4];
            for (int ib = 0; ib < 4; ++ib) {
                const floatx4 dh = mfma(frags[(kFragB3 + ib) * 64 + l], dyb, zero);
                for (int r = 0; r < 4; ++r) {
                    dz2[ib][r] = dh[r] * gelu_grad(s.z2[ib][r]);
                }
            }
            bf16x8 dz2b[2];
            for (int c = 0; c < 2; ++c) {
                for (int e = 0; e < 4; ++e) {
                    dz2b[c][e] = bf(dz2[2 * c][e]);
                    dz2b[c][e + 4] = bf(dz2[2 * c + 1][e]);
                }
            }
            floatx4 dz1[4];
            for (int ib = 0; ib < 4; ++ib) {
                floatx4 dh = mfma(frags[(kFragB2 + 2 * ib) * 64 + l], dz2b[0], zero);
                dh = mfma(frags[(kFragB2 + 2 * ib + 1) * 64 + l], dz2b[1], dh);
                for (int r = 0; r < 4; ++r) {
                    dz1[ib][r] = dh[r] * gelu_grad(s.z1[ib][r]);
                }
            }
            // Transposed into LDS: row = 16 half + n; features of this lane: F(c, g, e) of the packed registers.
            const int row = 16 * half + n;
            for (int e = 0; e < 8; ++e) {
                const int f0 = kfeat(0, g, e);
                T.x[f0][row] = s.x[e];
            }
            for (int c = 0; c < 2; ++c) {
                for (int e = 0; e < 8; ++e) {
                    const int f = kfeat(c, g, e);
                    T.h1[f][row] = s.h1[c][e];
                    T.h2[f][row] = s.h2[c][e];
                    T.dz2[f][row] = dz2b[c][e];
                }
            }
            for (int ib = 0; ib < 4; ++ib) {
                for (int r = 0; r < 4; ++r) {
                    T.dz1[16 * ib + 4 * g + r][row] = bf(dz1[ib][r]);
                }
            }
            for (int o = 0; o < 16; ++o) {
                if (g == 0) {
                    T.dy[o][row] = o == 0 ? dyb[0] : (o == 1 ? dyb[1] : bf(0.0f));
                }
            }
        }
        __syncthreads(); // the tiles written above, read across lanes below
        // Weight gradients over the 32 rows: A = act^T (16 features x 32 rows), B = grad (32 rows x 16 features); a
        // lane's 8 rows are 8g .. 8g + 7 in both.
        auto rows8 = [&](const __bf16* rowp) {
            bf16x8 r8;
            for (int e = 0; e < 8; ++e) {
                r8[e] = rowp[8 * g + e];
            }
            return r8;
        };
        for (int ib = 0; ib < 4; ++ib) {
            const bf16x8 a2 = rows8(T.h1[16 * ib + n]);
            const bf16x8 a3 = rows8(T.h2[16 * ib + n]);
            for (int ob = 0; ob < 4; ++ob) {
                gw2[ib][ob] = mfma(a2, rows8(T.dz2[16 * ob + n]), gw2[ib][ob]);
            }
            gw3[ib] = mfma(a3, rows8(T.dy[n]), gw3[ib]);
        }
        for (int ib = 0; ib < 2; ++ib) {
            const bf16x8 a1 = rows8(T.x[16 * ib + n]);
            for (int ob = 0; ob < 4; ++ob) {
                gw1[ib][ob] = mfma(a1, rows8(T.dz1[16 * ob + n]), gw1[ib][ob]);
            }
        }
        // biases: lane l sums feature l over the 32 rows, in row order
        for (int r = 0; r < 32; ++r) {
            gb1 += unbf(T.dz1[l][r]);
            gb2 += unbf(T.dz2[l][r]);
            if (l < kOut) {
                gb3 += unbf(T.dy[l][r]);
            }
        }
        __syncthreads();
    }
    // D of a gradient block: lane l holds column (output feature) 16 ob + n, rows (input features) 16 ib + 4g + r.
    float* out = G + wave * kParams;
    for (int ib = 0; ib < 4; ++ib) {
        for (int ob = 0; ob < 4; ++ob) {
            for (int r = 0; r < 4; ++r) {
                out[kW2 + (16 * ib + 4 * g + r) * kHid + 16 * ob + n] = gw2[ib][ob][r];
            }
        }
        for (int r = 0; r < 4; ++r) {
            if (n < kOut) {
                out[kW3 + (16 * ib + 4 * g + r) * kOut + n] = gw3[ib][r];
            }
        }
    }
    for (int ib = 0; ib < 2; ++ib) {
        for (int ob = 0; ob < 4; ++ob) {
            for (int r = 0; r < 4; ++r) {
                const int in = 16 * ib + 4 * g + r;
                if (in < kIn) {
                    out[kW1 + in * kHid + 16 * ob + n] = gw1[ib][ob][r];
                }
            }
        }
    }
    out[kB1 + l] = gb1;
    out[kB2 + l] = gb2;
    if (l < kOut) {
        out[kB3 + l] = gb3;
    }
}

// The first level of k_sum's fixed order: column c of rows [chunk * j, chunk * (j + 1)) summed in row order into
// out[j][c] (grid: columns x chunks, so the rows' loads spread over many waves instead of one wave a 64 columns).
__global__ void k_sum_chunks(const float* __restrict__ parts, int rows, int width, int chunk, float* __restrict__ out)
{
    const int c = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int j = (int)blockIdx.y;
    if (c >= width) {
        return;
    }
    const int r1 = (j + 1) * chunk < rows ? (j + 1) * chunk : rows;
    float s = 0.0f;
#pragma unroll 8
    for (int r = j * chunk; r < r1; ++r) { // eight loads in flight, added in order
        s += parts[(long long)r * width + c];
    }
    out[(long long)j * width + c] = s;
}

// Sums `parts` rows of `width` floats column by column, in row order (fixed: the same bits every run).
__global__ void k_sum(const float* __restrict__ parts, int rows, int width, float* __restrict__ out)
{
    const int c = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (c >= width) {
        return;
    }
    float s = 0.0f;
#pragma unroll 8
    for (int r = 0; r < rows; ++r) { // eight loads in flight, added in order
        s += parts[(long long)r * width + c];
    }
    out[c] = s;
}

// Sums n floats in two fixed-order levels: block b its chunk (each lane a fixed stride, then a fixed tree) into
// parts[b], then k_total_fold the blocks' parts in order (one block over all n waited on memory a lane at a time).
constexpr int kTotalBlocks = 256;
__global__ __launch_bounds__(256) void k_total(const float* __restrict__ v, long long n, double* __restrict__ parts)
{
    __shared__ double s[256];
    const int t = (int)threadIdx.x;
    const long long chunk = (n + gridDim.x - 1) / gridDim.x;
    const long long begin = (long long)blockIdx.x * chunk;
    const long long end = begin + chunk < n ? begin + chunk : n;
    double a = 0.0;
    for (long long i = begin + t; i < end; i += 256) {
        a += (double)v[i];
    }
    s[t] = a;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            s[t] += s[t + w];
        }
        __syncthreads();
    }
    if (t == 0) {
        parts[blockIdx.x] = s[0];
    }
}

__global__ void k_total_fold(const double* __restrict__ parts, int count, float* __restrict__ out)
{
    if (threadIdx.x == 0) {
        double a = 0.0;
        for (int b = 0; b < count; ++b) {
            a += parts[b];
        }
        *out = (float)a;
    }
}

// Adam (Kingma & Ba 2014), the update clipped to +-clip a step: m, v the moments, step the 1-based step count.
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ grad, float* __restrict__ m, float* __restrict__ v, int n, float lr,
    float b1, float b2, float eps, int step, float clip)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) {
        return;
    }
    const float gr = grad[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gr;
    const float vi = b2 * v[i] + (1.0f - b2) * gr * gr;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / (1.0f - powf(b1, (float)step));
    const float vh = vi / (1.0f - powf(b2, (float)step));
    const float d = lr * mh / (sqrtf(vh) + eps);
    p[i] -= fminf(fmaxf(d, -clip), clip);
}

// Block averages of a fine lattice (f x f cells a coarse cell), lattice by lattice.
__global__ void k_coarsen(const float* __restrict__ fine, int fnx, int fny, int count, int f, float* __restrict__ coarse)
{
    const int cnx = fnx / f, cny = fny / f;
    const long long cells = (long long)cnx * cny;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cells * count) {
        return;
    }
    long long lat;
    int x, y;
    cell_of(i, cnx, cny, lat, x, y);
    const float* src = fine + lat * (long long)fnx * fny;
    double s = 0.0;
    for (int b = 0; b < f; ++b) {
        for (int a = 0; a < f; ++a) {
            s += (double)src[(long long)(y * f + b) * fnx + x * f + a];
        }
    }
    coarse[i] = (float)(s / (double)(f * f));
}

void hcheck(hipError_t e, const char* what)
{
    if (e != hipSuccess) {
        throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
    }
}

template <typename T>
struct Dev {
    T* p = nullptr;
    std::size_t n = 0;
    Dev() = default;
    Dev(const Dev&) = delete;
    Dev& operator=(const Dev&) = delete;
    ~Dev()
    {
        if (p != nullptr) {
            (void)hipFree(p);
        }
    }
    void alloc(std::size_t count)
    {
        if (count <= n && p != nullptr) {
            return;
        }
        if (p != nullptr) {
            (void)hipFree(p);
            p = nullptr;
        }
        n = count;
        hcheck(hipMalloc(reinterpret_cast<void**>(&p), std::max<std::size_t>(1, count) * sizeof(T)), "hipMalloc");
    }
};

unsigned blocks_for(long long n, int threads) { return static_cast<unsigned>((n + threads - 1) / threads); }

} // namespace

// The scratch of fe_lnet_step_noalloc, one copy on every device the library runs on, held until the process ends: the
// face fluxes of up to kNoAllocCells cells (two faces a cell, two fields a face) and, after them, the step's breach
// counter. Host code never passes `scratch` itself (on the host it names a stand-in, not device memory): it passes the
// address hipGetSymbolAddress gives for the current device, kept per device in noalloc_scratch.
constexpr long long kNoAllocCells = 1LL << 20;
constexpr long long kNoAllocFloats = 4 * kNoAllocCells + 64; // faces, then 256 bytes holding the breach counter
static __device__ __align__(256) float scratch[kNoAllocFloats];

namespace {

constexpr int kMaxDevices = 64;
std::mutex noalloc_mutex;
float* noalloc_scratch[kMaxDevices] = {};

// Resolves the current device's copy of `scratch` once (at fe_lnet_create, with the device current); 0 or a reason.
const char* resolve_noalloc_scratch()
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        return "no current HIP device for the no-allocation scratch";
    }
    const std::lock_guard<std::mutex> lock(noalloc_mutex);
    if (noalloc_scratch[dev] == nullptr) {
        void* ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || ptr == nullptr) {
            return "hipGetSymbolAddress(scratch) failed";
        }
        noalloc_scratch[dev] = static_cast<float*>(ptr);
    }
    return nullptr;
}

float* noalloc_scratch_here()
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        return nullptr;
    }
    const std::lock_guard<std::mutex> lock(noalloc_mutex);
    return noalloc_scratch[dev];
}

} // namespace

struct fe_lnet {
    Dev<float> params, grad, m, v, partials, chunk_sums, loss_parts, loss, F, E;
    Dev<double> loss_blocks;
    Dev<bf16x8> frags;
    Dev<unsigned> flags; // [0] breaches of the last step, [1] fallbacks so far
    float scale = 1.0f;
    int adam_step = 0;
    std::string error;
    hipStream_t stream() const { return static_cast<hipStream_t>(fe_hip_stream()); }
    static constexpr int kFaceBlocks = 1024; // face kernels: waves in flight (a grid-stride loop over 16-row tiles)
    static constexpr int kTrainBlocks = 1024; // one wave a SIMD: the training pass holds 512 registers a lane
};

namespace {

int fail(fe_lnet* n, const std::string& what)
{
    if (n != nullptr) {
        n->error = what;
    }
    return -1;
}

Lattice lattice(int nx, int ny, int count, double h)
{
    if (2.0 * nx * ny * count >= 2147483648.0) {
        throw std::runtime_error("batch too large: 2 * count * nx * ny must stay below 2^31 (32-bit cell indexing)");
    }
    return Lattice{nx, ny, count, static_cast<float>(h)};
}

} // namespace

extern "C" {

fe_lnet* fe_lnet_create(const float* params, float scale, char* why, size_t why_bytes)
{
    try {
        fe_hip_device_info info{};
        if (fe_hip_device_open(&info) != 0) {
            if (why != nullptr && why_bytes > 0) {
                std::snprintf(why, why_bytes, "no HIP device: %s", fe_hip_unavailable());
            }
            return nullptr;
        }
        auto n = std::make_unique<fe_lnet>();
        n->scale = scale;
        n->params.alloc(kParams);
        n->grad.alloc(kParams);
        n->m.alloc(kParams);
        n->v.alloc(kParams);
        n->frags.alloc(kFrags * 64);
        n->flags.alloc(2);
        n->loss.alloc(1);
        hcheck(hipMemcpy(n->params.p, params, sizeof(float) * kParams, hipMemcpyHostToDevice), "upload");
        hcheck(hipMemset(n->m.p, 0, sizeof(float) * kParams), "memset");
        hcheck(hipMemset(n->v.p, 0, sizeof(float) * kParams), "memset");
        hcheck(hipMemset(n->flags.p, 0, sizeof(unsigned) * 2), "memset");
        hipLaunchKernelGGL(k_pack, dim3(kFrags), dim3(64), 0, n->stream(), n->params.p, n->frags.p);
        hcheck(hipStreamSynchronize(n->stream()), "sync");
        if (const char* bad = resolve_noalloc_scratch()) {
            throw std::runtime_error(bad);
        }
        return n.release();
    } catch (const std::exception& e) {
        if (why != nullptr && why_bytes > 0) {
            std::snprintf(why, why_bytes, "%s", e.what());
        }
        return nullptr;
    }
}

void fe_lnet_free(fe_lnet* n) { delete n; }
const char* fe_lnet_error(fe_lnet* n) { return n != nullptr ? n->error.c_str() : "no net"; }
int fe_lnet_param_count(void) { return kParams; }

int fe_lnet_get_params(fe_lnet* n, float* out)
{
    if (n == nullptr || out == nullptr) {
        return -1;
    }
    hcheck(hipMemcpyAsync(out, n->params.p, sizeof(float) * kParams, hipMemcpyDeviceToHost, n->stream()), "download");
    return hipStreamSynchronize(n->stream()) == hipSuccess ? 0 : fail(n, "sync");
}

int fe_lnet_set_params(fe_lnet* n, const float* in)
{
    if (n == nullptr || in == nullptr) {
        return -1;
    }
    try {
        hcheck(hipMemcpyAsync(n->params.p, in, sizeof(float) * kParams, hipMemcpyHostToDevice, n->stream()), "upload");
        hipLaunchKernelGGL(k_pack, dim3(kFrags), dim3(64), 0, n->stream(), n->params.p, n->frags.p);
        hcheck(hipStreamSynchronize(n->stream()), "sync");
        return 0;
    } catch (const std::exception& e) {
        return fail(n, e.what());
    }
}

int fe_lnet_faces(fe_lnet* n, const float* u, const float* v, int nx, int ny, int count, double h, float* faces_out)
{
    if (n == nullptr || u == nullptr || v == nullptr || nx < 2 || ny < 2 || count < 1) {
        return -1;
    }
    try {
        const Lattice L = lattice(nx, ny, count, h);
        const long long faces = 2LL * nx * ny * count;
        float* F = faces_out;
        if (F == nullptr) {
            n->F.alloc(static_cast<std::size_t>(2 * faces));
            F = n->F.p;
        }
        hipLaunchKernelGGL(k_faces, dim3(fe_lnet::kFaceBlocks), dim3(256), 0, n->stream(), n->frags.p, n->params.p, u, v, L, n->scale, faces, F);
        hcheck(hipGetLastError(), "faces");
        return 0;
    } catch (const std::exception& e) {
        return fail(n, e.what());
    }
}

int fe_lnet_step(fe_lnet* n, const float* u, const float* v, int nx, int ny, int count, double h, double du, double dv, double k,
    double dt, int learned, double bound, float* u2, float* v2)
{
    if (u == nullptr || v == nullptr || u2 == nullptr || v2 == nullptr || nx < 2 || ny < 2 || count < 1) {
        return fail(n, "invalid arguments");
    }
    if (learned && n == nullptr) {
        return -1;
    }
    try {
        const Lattice L = lattice(nx, ny, count, h);
        const Reaction R{static_cast<float>(du), static_cast<float>(dv), static_cast<float>(k)};
        const long long cells = static_cast<long long>(nx) * ny * count;
        const hipStream_t s = static_cast<hipStream_t>(fe_hip_stream());
        if (learned) {
            n->F.alloc(static_cast<std::size_t>(4 * cells));
            hipLaunchKernelGGL(k_clear, dim3(1), dim3(64), 0, s, n->flags.p, 1);
            hipLaunchKernelGGL(k_faces, dim3(fe_lnet::kFaceBlocks), dim3(256), 0, s, n->frags.p, n->params.p, u, v, L, n->scale, 2 * cells, n->F.p);
            hipLaunchKernelGGL(k_step, dim3(blocks_for(cells, 256)), dim3(256), 0, s, u, v, n->F.p, L, R, static_cast<float>(dt), 1,
                static_cast<float>(bound), u2, v2, n->flags.p);
            hipLaunchKernelGGL(k_fallback, dim3(blocks_for(cells, 256)), dim3(256), 0, s, u, v, L, R, static_cast<float>(dt), n->flags.p, u2, v2,
                n->flags.p + 1);
        } else {
            hipLaunchKernelGGL(k_step, dim3(blocks_for(cells, 256)), dim3(256), 0, s, u, v, nullptr, L, R, static_cast<float>(dt), 0,
                static_cast<float>(bound), u2, v2, nullptr);
        }
        hcheck(hipGetLastError(), "step");
        return 0;
    } catch (const std::exception& e) {
        return fail(n, e.what());
    }
}

int fe_lnet_step_noalloc(fe_lnet* n, const float* u, const float* v, int nx, int ny, int count, double h, double du, double dv,
    double k, double dt, double bound, float* u2, float* v2)
{
    if (n == nullptr || u == nullptr || v == nullptr || u2 == nullptr || v2 == nullptr || nx < 2 || ny < 2 || count < 1) {
        return fail(n, "invalid arguments");
    }
    const long long cells = static_cast<long long>(nx) * ny * count;
    if (cells > kNoAllocCells) {
        return fail(n, "fe_lnet_step_noalloc: " + std::to_string(cells) + " cells, above its scratch of " +
            std::to_string(kNoAllocCells) + " cells (use fe_lnet_step)");
    }
    float* faces = noalloc_scratch_here();
    if (faces == nullptr) {
        return fail(n, "fe_lnet_step_noalloc: the scratch was not resolved on this device (fe_lnet_create does it)");
    }
    unsigned* breach = reinterpret_cast<unsigned*>(faces + 4 * kNoAllocCells);
    try {
        const Lattice L = lattice(nx, ny, count, h);
        const Reaction R{static_cast<float>(du), static_cast<float>(dv), static_cast<float>(k)};
        const hipStream_t s = static_cast<hipStream_t>(fe_hip_stream());
        // The kernels of fe_lnet_step's learned path, with the faces and the breach counter in the scratch: the counter
        // is cleared by a kernel (no memset node), the fallback count stays the net's own.
        hipLaunchKernelGGL(k_clear, dim3(1), dim3(64), 0, s, breach, 1);
        hipLaunchKernelGGL(k_faces, dim3(fe_lnet::kFaceBlocks), dim3(256), 0, s, n->frags.p, n->params.p, u, v, L, n->scale, 2 * cells, faces);
        hipLaunchKernelGGL(k_step, dim3(blocks_for(cells, 256)), dim3(256), 0, s, u, v, faces, L, R, static_cast<float>(dt), 1,
            static_cast<float>(bound), u2, v2, breach);
        hipLaunchKernelGGL(k_fallback, dim3(blocks_for(cells, 256)), dim3(256), 0, s, u, v, L, R, static_cast<float>(dt), breach, u2, v2,
            n->flags.p + 1);
        hcheck(hipGetLastError(), "step_noalloc");
        return 0;
    } catch (const std::exception& e) {
        return fail(n, e.what());
    }
}

int fe_lnet_fallbacks(fe_lnet* n, unsigned* out)
{
    if (n == nullptr || out == nullptr) {
        return -1;
    }
    hcheck(hipMemcpyAsync(out, n->flags.p + 1, sizeof(unsigned), hipMemcpyDeviceToHost, n->stream()), "download");
    return hipStreamSynchronize(n->stream()) == hipSuccess ? 0 : fail(n, "sync");
}

int fe_lnet_gradient(fe_lnet* n, const float* u0, const float* v0, const float* u1, const float* v1, int nx, int ny, int count, double h,
    double du, double dv, double k, double dt, double weight, float* grad_out, float* loss_out)
{
    if (n == nullptr || u0 == nullptr || v0 == nullptr || u1 == nullptr || v1 == nullptr || nx < 2 || ny < 2 || count < 1 ||
        !(weight > 0.0)) {
        return -1;
    }
    try {
        const Lattice L = lattice(nx, ny, count, h);
        const Reaction R{static_cast<float>(du), static_cast<float>(dv), static_cast<float>(k)};
        const long long cells = static_cast<long long>(nx) * ny * count;
        const long long faces = 2 * cells;
        const hipStream_t s = n->stream();
        n->F.alloc(static_cast<std::size_t>(2 * faces));
        n->E.alloc(static_cast<std::size_t>(2 * cells));
        n->loss_parts.alloc(static_cast<std::size_t>(cells));
        const int waves = fe_lnet::kTrainBlocks * kWavesTrain;
        n->partials.alloc(static_cast<std::size_t>(waves) * kParams);
        hipLaunchKernelGGL(k_faces, dim3(fe_lnet::kFaceBlocks), dim3(256), 0, s, n->frags.p, n->params.p, u0, v0, L, n->scale, faces, n->F.p);
        hipLaunchKernelGGL(k_residual, dim3(blocks_for(cells, 256)), dim3(256), 0, s, u0, v0, u1, v1, n->F.p, L, R, static_cast<float>(dt),
            static_cast<float>(weight), n->E.p, n->loss_parts.p);
        hipLaunchKernelGGL(k_backward, dim3(fe_lnet::kTrainBlocks), dim3(64 * kWavesTrain), 0, s, n->frags.p, n->params.p, u0, v0, n->E.p, L,
            n->scale, faces, n->partials.p);
        // the waves' partials summed in a fixed order: chunks of 32 waves in wave order, then the chunks in order
        constexpr int kChunk = 32;
        const int chunks = (waves + kChunk - 1) / kChunk;
        n->chunk_sums.alloc(static_cast<std::size_t>(chunks) * kParams);
        hipLaunchKernelGGL(k_sum_chunks, dim3(blocks_for(kParams, 256), chunks), dim3(256), 0, s, n->partials.p, waves, kParams, kChunk,
            n->chunk_sums.p);
        hipLaunchKernelGGL(k_sum, dim3(blocks_for(kParams, 256)), dim3(256), 0, s, n->chunk_sums.p, chunks, kParams, n->grad.p);
        if (loss_out != nullptr) { // the loss's total only when it is asked for
            n->loss_blocks.alloc(kTotalBlocks);
            hipLaunchKernelGGL(k_total, dim3(kTotalBlocks), dim3(256), 0, s, n->loss_parts.p, cells, n->loss_blocks.p);
            hipLaunchKernelGGL(k_total_fold, dim3(1), dim3(64), 0, s, n->loss_blocks.p, kTotalBlocks, n->loss.p);
        }
        hcheck(hipGetLastError(), "gradient");
        if (grad_out != nullptr) {
            hcheck(hipMemcpyAsync(grad_out, n->grad.p, sizeof(float) * kParams, hipMemcpyDeviceToHost, s), "download");
        }
        if (loss_out != nullptr) {
            hcheck(hipMemcpyAsync(loss_out, n->loss.p, sizeof(float), hipMemcpyDeviceToHost, s), "download");
        }
        if (grad_out != nullptr || loss_out != nullptr) {
            hcheck(hipStreamSynchronize(s), "sync");
        }
        return 0;
    } catch (const std::exception& e) {
        return fail(n, e.what());
    }
}

int fe_lnet_adam(fe_lnet* n, double lr, double clip)
{
    if (n == nullptr) {
        return -1;
    }
    try {
        ++n->adam_step;
        hipLaunchKernelGGL(k_adam, dim3(blocks_for(kParams, 256)), dim3(256), 0, n->stream(), n->params.p, n->grad.p, n->m.p, n->v.p, kParams,
            static_cast<float>(lr), 0.9f, 0.999f, 1e-8f, n->adam_step, static_cast<float>(clip));
        hipLaunchKernelGGL(k_pack, dim3(kFrags), dim3(64), 0, n->stream(), n->params.p, n->frags.p);
        hcheck(hipGetLastError(), "adam");
        return 0;
    } catch (const std::exception& e) {
        return fail(n, e.what());
    }
}

int fe_lnet_coarsen(const float* fine, int fnx, int fny, int count, int factor, float* coarse)
{
    if (fine == nullptr || coarse == nullptr || factor < 1 || fnx % factor != 0 || fny % factor != 0 || count < 1) {
        return -1;
    }
    const long long cells = static_cast<long long>(fnx / factor) * (fny / factor) * count;
    hipLaunchKernelGGL(k_coarsen, dim3(blocks_for(cells, 256)), dim3(256), 0, static_cast<hipStream_t>(fe_hip_stream()), fine, fnx, fny, count,
        factor, coarse);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

} // extern "C"
