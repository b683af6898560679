// The --hip generator. See HipCodegen.hpp.
#include "field/hip/HipCodegen.hpp"

#include "kernels/exact.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <set>
#include <functional>
#include <sstream>

namespace fieldengine::field::hip {

namespace {

// The generic kernels every law module has: contracts (range and finite, clamping, conserved totals), the reduction
// passes' final fold and the change flags. All reductions are fixed-order: a block sums a fixed range of cells, each
// lane a fixed stride of it, the lanes in a fixed tree; the partials fold in one block, in order. No float atomics.
const char* kPrelude = R"HIP(
typedef long long i64;
typedef float f4v __attribute__((ext_vector_type(4)));
typedef unsigned long long u64;
#define IDX(x, y, z) ((i64)(z) * k.nxy + (i64)(y) * k.nx + (i64)(x))
#define AT(slot, x, y, z) k.A[(i64)(slot) * k.cells + IDX(x, y, z)]
#define CELL_HEAD const int x = (int)(blockIdx.x * blockDim.x + threadIdx.x); const int y = (int)(blockIdx.y * blockDim.y + threadIdx.y); \
    const int z = (int)blockIdx.z; if (x >= k.nx || y >= k.ny) return;
// A cell's coordinates from its index: 32-bit divisions while the lattice has fewer than 2^31 cells (a 64-bit division
// is a long software sequence on the card).
__device__ __forceinline__ void cell_xyz(const K& k, i64 i, int& x, int& y, int& z)
{
    if (k.cells < 2147483648LL) {
        const unsigned u = (unsigned)i, nx = (unsigned)k.nx, ny = (unsigned)k.ny;
        const unsigned r = u / nx;
        x = (int)(u - r * nx); y = (int)(r % ny); z = (int)(r / ny);
    } else {
        x = (int)(i % k.nx); y = (int)((i / k.nx) % k.ny); z = (int)(i / k.nxy);
    }
}
// Clamps written as max(0, min(v, n - 1)): one v_med3_i32.
__device__ __forceinline__ int ex(const K& k, int v) { return max(0, min(v, k.nx - 1)); }
__device__ __forceinline__ int ey(const K& k, int v) { return max(0, min(v, k.ny - 1)); }
__device__ __forceinline__ int ez(const K& k, int v) { return max(0, min(v, k.nz - 1)); }
// A field read at a neighbour's coordinate (in a function templated on its memory M): clamped into the lattice where M
// reads HBM; a window's LDS holds every coordinate within its reach at the clamped cell's value, so there the
// coordinate goes unclamped and the neighbours' addresses stay constant offsets of the cell's.
#define LX(v) (M::kClamps ? ex(k, (v)) : (v))
#define LY(v) (M::kClamps ? ey(k, (v)) : (v))
#define LZ(v) (M::kClamps ? ez(k, (v)) : (v))
__device__ __forceinline__ bool finite_d(double v) { return __builtin_isfinite(v); }
__device__ __forceinline__ int clampi(int v, int lo, int hi) { return max(lo, min(v, hi)); }
// a / h, a / h^2, a / (h * span): when the divisor is a power of two (flags, set by the host) the quotient is the exact
// product with its reciprocal, so the multiply gives the division's bits without the division's cost.
// M::kPow2: the caller has checked both flags (a kernel instantiated for that case), so no branch is left.
template <class M> __device__ __forceinline__ double div_h(const K& k, double a)
{
    return (M::kPow2 || (k.flags & 1)) ? a * k.inv_h : a / k.h;
}
template <class M> __device__ __forceinline__ double div_hh(const K& k, double a)
{
    return (M::kPow2 || (k.flags & 2)) ? a * k.inv_hh : a / (k.h * k.h);
}
template <class M> __device__ __forceinline__ double div_hs(const K& k, double a, int span)
{
    return (M::kPow2 || (k.flags & 1)) && span > 0 ? a * (span == 1 ? k.inv_h : 0.5 * k.inv_h) : a / (k.h * (double)span);
}
// Where a value is read from (struct Direct, generated after this prelude: the field's array in HBM; a window kernel
// has its own, an LDS ring of planes).
// Workgroups are dealt round-robin over the 8 XCDs (each with its own L2): consecutive tiles to one XCD keep the halo
// a tile shares with its neighbour in that XCD's L2 more often.
__device__ __forceinline__ int xcd_tile(int b, int total)
{
    const int per = total / 8, rem = total % 8, xcd = b % 8, local = b / 8;
    return xcd < rem ? xcd * (per + 1) + local : rem * (per + 1) + (xcd - rem) * per + local;
}
// The cells i along an axis whose clamp(i + d) is j (one run: j - d inside, and what is clamped onto an edge cell).
__device__ __forceinline__ void run_to(int j, int d, int n, int& lo, int& hi)
{
    lo = j == 0 ? 0 : j - d; if (lo < 0) lo = 0;
    hi = j == n - 1 ? n - 1 : j - d; if (hi > n - 1) hi = n - 1;
}
// "Field t changed": one store a wave at most, by its first lane that saw a change.
__device__ __forceinline__ void mark(const K& k, int t, bool changed)
{
    const u64 m = __ballot(changed);
    if (m != 0ull && (int)(threadIdx.x & 63u) == __ffsll((long long)m) - 1) k.C[t] = 1u;
}
// Neumaier's compensated sum in double.
struct Sum { double s; double c; };
__device__ __forceinline__ void add(Sum& a, double v)
{
    const double t = a.s + v;
    a.c += fabs(a.s) >= fabs(v) ? (a.s - t) + v : (v - t) + a.s;
    a.s = t;
}
__device__ __forceinline__ void merge(Sum& a, const Sum& b) { add(a, b.s); a.c += b.c; }
// A pair folded across a wave's 64 lanes by a butterfly (lane i with lane i ^ o, o = 32, 16, ..., 1: cross-lane moves,
// no LDS): the same tree every run, so the same bits; lane 0 holds the wave's total.
__device__ __forceinline__ void wave_fold(Sum& a)
{
    for (int o = 32; o > 0; o >>= 1) {
        const Sum b = {__shfl_xor(a.s, o), __shfl_xor(a.c, o)};
        merge(a, b);
    }
}
// The least and greatest across a wave, the same butterfly (a comparison that a NaN never wins, as the host's).
__device__ __forceinline__ void wave_bounds(double& lo, double& hi)
{
    for (int o = 32; o > 0; o >>= 1) {
        const double a = __shfl_xor(lo, o), b = __shfl_xor(hi, o);
        lo = a < lo ? a : lo;
        hi = hi < b ? b : hi;
    }
}
// A 256-thread block's four waves, in order: lanes 0 leave their wave's values in w[4][4]; thread 0 folds them.
__device__ __forceinline__ void block_pairs(Sum& a, Sum& b, double (*w)[4])
{
    wave_fold(a); wave_fold(b);
    const int t = (int)threadIdx.x;
    if ((t & 63) == 0) { w[t >> 6][0] = a.s; w[t >> 6][1] = a.c; w[t >> 6][2] = b.s; w[t >> 6][3] = b.c; }
    __syncthreads();
    if (t == 0) {
        a = Sum{w[0][0], w[0][1]}; b = Sum{w[0][2], w[0][3]};
        for (int q = 1; q < 4; ++q) { merge(a, Sum{w[q][0], w[q][1]}); merge(b, Sum{w[q][2], w[q][3]}); }
    }
}
__device__ __forceinline__ void block_sum_bounds(Sum& a, double& lo, double& hi, double (*w)[4])
{
    wave_fold(a); wave_bounds(lo, hi);
    const int t = (int)threadIdx.x;
    if ((t & 63) == 0) { w[t >> 6][0] = a.s; w[t >> 6][1] = a.c; w[t >> 6][2] = lo; w[t >> 6][3] = hi; }
    __syncthreads();
    if (t == 0) {
        a = Sum{w[0][0], w[0][1]}; lo = w[0][2]; hi = w[0][3];
        for (int q = 1; q < 4; ++q) {
            merge(a, Sum{w[q][0], w[q][1]});
            lo = w[q][2] < lo ? w[q][2] : lo;
            hi = hi < w[q][3] ? w[q][3] : hi;
        }
    }
}

// A block's range counts (out of range, not finite) and first bad cell into its record: the lanes' by wave folds, the
// waves' in LDS, then at most three atomics a block (an atomic a bad cell would queue on one address when a field is
// broken across the lattice). Every thread of the block calls it.
__device__ __forceinline__ void range_fold(unsigned nout, unsigned nnan, u64 fb, unsigned* rec, u64* first)
{
    __shared__ unsigned rn[16], rb[16];
    __shared__ u64 rf[16];
    for (int o = 32; o > 0; o >>= 1) {
        nout += (unsigned)__shfl_xor((int)nout, o);
        nnan += (unsigned)__shfl_xor((int)nnan, o);
        const u64 other = __shfl_xor(fb, o);
        fb = other < fb ? other : fb;
    }
    const int t = (int)(threadIdx.y * blockDim.x + threadIdx.x), w = t >> 6, nw = (int)((blockDim.x * blockDim.y + 63) >> 6);
    if ((t & 63) == 0) { rn[w] = nout; rb[w] = nnan; rf[w] = fb; }
    __syncthreads();
    if (t == 0) {
        unsigned a = 0, b = 0; u64 f = ~0ull;
        for (int q = 0; q < nw; ++q) { a += rn[q]; b += rb[q]; f = rf[q] < f ? rf[q] : f; }
        if (a != 0u) atomicAdd(&rec[0], a);
        if (b != 0u) atomicAdd(&rec[1], b);
        if (a != 0u || b != 0u) atomicMin(first, f);
    }
}

// A field's range contract (src/field/Solver.cpp:463-511): counts of non-finite and out-of-range cells and the least
// index of a bad one (integer atomics: the order does not matter). A clamped field's finite strays are clamped instead.
extern "C" __global__ void k_range(K k, int field, double lo, double hi, int clamp, unsigned* rec, u64* first)
{
    const int x = (int)(blockIdx.x * blockDim.x + threadIdx.x); const int y = (int)(blockIdx.y * blockDim.y + threadIdx.y);
    const int z = (int)blockIdx.z;
    unsigned nout = 0, nnan = 0; u64 fb = ~0ull;
    if (x < k.nx && y < k.ny) {
        const i64 i = IDX(x, y, z);
        const float v = k.now[field][i];
        const double d = (double)v;
        const bool fin = finite_d(d);
        bool bad = !fin;
        if (fin && (d < lo || d > hi)) {
            if (clamp) k.now[field][i] = fe_rf(d < lo ? lo : (hi < d ? hi : d));
            else bad = true;
        }
        if (bad) { if (fin) nout = 1; else nnan = 1; fb = (u64)i; }
    }
    range_fold(nout, nnan, fb, rec, first);
}

// A clamped field's finite strays into its range, on a step that checks no contract (src/field/Solver.cpp:471-476
// clamps every step; the reference's CUDA backend launches its k_clamp after every term that writes the field).
extern "C" __global__ void k_clamp(K k, int field, double lo, double hi)
{
    CELL_HEAD
    const i64 i = IDX(x, y, z);
    const double d = (double)k.now[field][i];
    if (finite_d(d) && (d < lo || d > hi)) k.now[field][i] = fe_rf(d < lo ? lo : (hi < d ? hi : d));
}

// The window moved by d (src/field/Backend.hpp:33-40): dst gets what src held at c + d, where that is inside the lattice
// (the rest arrives in boxes from the host).
extern "C" __global__ void k_slide(K k, const float* src, float* dst, int dx, int dy, int dz)
{
    CELL_HEAD
    const int sx = x + dx, sy = y + dy, sz = z + dz;
    if (sx >= 0 && sx < k.nx && sy >= 0 && sy < k.ny && sz >= 0 && sz < k.nz) dst[IDX(x, y, z)] = src[IDX(sx, sy, sz)];
}

// A box of arriving cells, packed x fastest on the host, into field `field`: [x0, x0 + bx) x [y0, y0 + by) x [z0, z0 + bz).
extern "C" __global__ void k_scatter(K k, int field, const float* packed, int x0, int y0, int z0, int bx, int by, int bz)
{
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (i64)bx * by * bz) return;
    const int x = x0 + (int)(i % bx), y = y0 + (int)((i / bx) % by), z = z0 + (int)(i / ((i64)bx * by));
    k.now[field][IDX(x, y, z)] = packed[i];
}

// Conserved totals, pass 1: block b sums cells [b * chunk, (b + 1) * chunk), lane t every 256th from t; then the lanes
// fold in a fixed tree. out[2b], out[2b+1]: the sum and the sum of magnitudes (each a compensated pair, rounded).
// With rec not null the same pass is also the field's range contract (k_range's test and clamp, on the value the
// total took).
extern "C" __global__ __launch_bounds__(256) void k_total(K k, int field, i64 chunk, double* out, double lo, double hi, int clamp,
    unsigned* rec, u64* first)
{
    __shared__ double wv[4][4];
    const int t = (int)threadIdx.x;
    const i64 begin = (i64)blockIdx.x * chunk;
    i64 end = begin + chunk; if (end > k.cells) end = k.cells;
    Sum s = {0.0, 0.0}; Sum m = {0.0, 0.0};
    unsigned nout = 0, nnan = 0; u64 fb = ~0ull;
    const float* v = k.now[field];
    // Eight cells' loads (clamped into the range, unconditional) ahead of their compensated sums, which keep their
    // order: the pass waits on memory once in eight cells, not every cell.
    for (i64 i0 = begin + t; i0 < end; i0 += 8 * 256) {
        float vv[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const i64 i = i0 + q * 256;
            vv[q] = v[i < end ? i : end - 1];
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const i64 i = i0 + q * 256;
            if (i >= end) break;
            const double d = (double)vv[q]; add(s, d); add(m, fabs(d));
            if (rec != nullptr) {
                const bool fin = finite_d(d);
                bool bad = !fin;
                if (fin && (d < lo || d > hi)) {
                    if (clamp) k.now[field][i] = fe_rf(d < lo ? lo : (hi < d ? hi : d));
                    else bad = true;
                }
                if (bad) { if (fin) ++nout; else ++nnan; fb = (u64)i < fb ? (u64)i : fb; }
            }
        }
    }
    if (rec != nullptr) range_fold(nout, nnan, fb, rec, first);
    block_pairs(s, m, wv);
    if (t == 0) { out[2 * blockIdx.x] = s.s + s.c; out[2 * blockIdx.x + 1] = m.s + m.c; }
}

// Pass 2: one block folds `parts` pairs in order into dst[0], dst[1].
extern "C" __global__ __launch_bounds__(256) void k_total_final(const double* parts, int count, double* dst)
{
    __shared__ double wv[4][4];
    const int t = (int)threadIdx.x;
    Sum s = {0.0, 0.0}; Sum m = {0.0, 0.0};
    for (int i = t; i < count; i += 256) { add(s, parts[2 * i]); add(m, parts[2 * i + 1]); }
    block_pairs(s, m, wv);
    if (t == 0) { dst[0] = s.s + s.c; dst[1] = m.s + m.c; }
}

// acc[0..1] += src[0..1] in double (a field's total added on the card, in call order: no wait for the host).
extern "C" __global__ void k_acc(const double* src, double* acc)
{
    if (threadIdx.x == 0) { acc[0] = acc[0] + src[0]; acc[1] = acc[1] + src[1]; }
}

// Words [0, zeros) to 0 and [zeros, zeros + ones) to all ones: the change flags, range counts and first bad cells of a
// step cleared in one launch.
extern "C" __global__ void k_fill(unsigned* p, int zeros, int ones)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < zeros) p[i] = 0u; else if (i < zeros + ones) p[i] = 0xffffffffu;
}

// Clears n words (the change flags) with a kernel: a graph replays kernels, and fills recorded as memset nodes have
// given wrong results on replay on this hardware.
extern "C" __global__ void k_zero(unsigned* p, int n)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) p[i] = 0u;
}

// A field's least and greatest value (NaN skipped, as the host's look over a lattice skips it), for deciding which terms
// have nothing to do: floats as order keys (a key's unsigned order is the float's order), so the least and greatest are
// integer atomics and the answer does not depend on the order. out[field * FE_BOUND_SLOTS + s] the least key,
// out[(fields + field) * FE_BOUND_SLOTS + s] the greatest, s a block's slot (the host folds the slots); k_bounds_reset
// starts them at the empty range.
__device__ __forceinline__ unsigned order_key(float v)
{
    const unsigned b = __float_as_uint(v);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ void bound_key(float x, unsigned& lo, unsigned& hi)
{
    if (x == x) { const unsigned key = order_key(x); lo = min(lo, key); hi = max(hi, key); }
}
extern "C" __global__ __launch_bounds__(256) void k_bounds(K k, int field, unsigned* out, int fields)
{
    unsigned lo = 0xffffffffu, hi = 0u;
    const float* v = k.now[field];
    const i64 stride = (i64)gridDim.x * blockDim.x;
    const i64 quads = k.cells / 4; // fields are 256-byte aligned: whole float4s, then the tail
    const float4* v4 = reinterpret_cast<const float4*>(v);
    i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + 3 * stride < quads; i += 4 * stride) { // four loads in flight a lane
        float4 q[4];
        for (int u = 0; u < 4; ++u) q[u] = v4[i + u * stride];
        for (int u = 0; u < 4; ++u) { bound_key(q[u].x, lo, hi); bound_key(q[u].y, lo, hi); bound_key(q[u].z, lo, hi); bound_key(q[u].w, lo, hi); }
    }
    for (; i < quads; i += stride) {
        const float4 q = v4[i];
        bound_key(q.x, lo, hi); bound_key(q.y, lo, hi); bound_key(q.z, lo, hi); bound_key(q.w, lo, hi);
    }
    for (i64 i = 4 * quads + (i64)blockIdx.x * blockDim.x + threadIdx.x; i < k.cells; i += stride) bound_key(v[i], lo, hi);
    for (int o = 32; o > 0; o >>= 1) { lo = min(lo, (unsigned)__shfl_xor((int)lo, o)); hi = max(hi, (unsigned)__shfl_xor((int)hi, o)); }
    // the block's four waves through LDS, then one pair of atomics a block (fewer on the one contended address)
    __shared__ unsigned wl[4], wh[4];
    if ((threadIdx.x & 63u) == 0u) { wl[threadIdx.x >> 6] = lo; wh[threadIdx.x >> 6] = hi; }
    __syncthreads();
    if (threadIdx.x == 0u) {
        lo = min(min(wl[0], wl[1]), min(wl[2], wl[3])); hi = max(max(wh[0], wh[1]), max(wh[2], wh[3]));
        const int slot = (int)(blockIdx.x % FE_BOUND_SLOTS);
        atomicMin(out + field * FE_BOUND_SLOTS + slot, lo); atomicMax(out + (fields + field) * FE_BOUND_SLOTS + slot, hi);
    }
}
extern "C" __global__ void k_bounds_reset(unsigned* out, int fields)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < 2 * fields * FE_BOUND_SLOTS) out[i] = i < fields * FE_BOUND_SLOTS ? 0xffffffffu : 0u;
}

// A reduction's partials (sum, least, most per block, from k_reduce passes) folded in order into U[uslot].
extern "C" __global__ __launch_bounds__(256) void k_reduce_final(K k, int kind, int parts, int uslot)
{
    __shared__ double wv[4][4];
    const int t = (int)threadIdx.x;
    Sum s = {0.0, 0.0}; double a = k.L[1]; double b = k.L[2];
    for (int i = t; i < parts; i += 256) {
        add(s, k.L[3 * i]);
        a = k.L[3 * i + 1] < a ? k.L[3 * i + 1] : a;
        b = b < k.L[3 * i + 2] ? k.L[3 * i + 2] : b;
    }
    block_sum_bounds(s, a, b, wv);
    if (t == 0) ((double*)k.U)[uslot] = kind == 0 ? s.s + s.c : (kind == 1 ? b : a);
}
// The lower envelope of parabolas f(j) + (q - j)^2 along a line (the distance stage's line kernel). push adds the
// parabola at q (at infinity: never lowest, left out), popping those it hides, by exact cross-multiplied comparisons of
// the intersections (whole numbers below 2^24, products far below 2^53); the stack's top two entries stay in
// registers, so LDS is read only when one is popped. start / at then walk the envelope from the left, giving the
// value of the parabola lowest at q (LDS read only when the walk moves on).
struct FeEnv {
    unsigned short* st; float* fs;
    int kk = -1;
    int tv = 0, pv = 0; double tF = 0.0, pF = 0.0;
    int j = 0, v0 = 0, v1 = 0; float f0 = 0.0f; double F0 = 0.0, F1 = 0.0;
    __device__ __forceinline__ void push(float fq, int q)
    {
        if (!(fq < __builtin_inff())) return;
        const double Fq = (double)fq + (double)q * (double)q;
        while (kk >= 1 && (Fq - tF) * (double)(tv - pv) <= (tF - pF) * (double)(q - tv)) {
            --kk; tv = pv; tF = pF;
            if (kk >= 1) { pv = st[kk - 1]; pF = (double)fs[kk - 1] + (double)pv * (double)pv; }
        }
        if (kk >= 0) { pv = tv; pF = tF; }
        ++kk; st[kk] = (unsigned short)q; fs[kk] = fq; tv = q; tF = Fq;
    }
    __device__ __forceinline__ void start()
    {
        j = 0;
        if (kk < 0) return;
        v0 = st[0]; f0 = fs[0]; F0 = (double)f0 + (double)v0 * (double)v0;
        if (kk >= 1) { v1 = st[1]; F1 = (double)fs[1] + (double)v1 * (double)v1; }
    }
    __device__ __forceinline__ float at(int q)
    {
        if (kk < 0) return __builtin_inff();
        while (j < kk && F1 - F0 < 2.0 * (double)q * (double)(v1 - v0)) {
            ++j; v0 = v1; F0 = F1; f0 = fs[j];
            if (j < kk) { v1 = st[j + 1]; F1 = (double)fs[j + 1] + (double)v1 * (double)v1; }
        }
        const double dq = (double)(q - v0);
        return (float)(dq * dq + (double)f0);
    }
};
)HIP";

std::string literal(double v)
{
    if (std::isnan(v)) {
        return "__builtin_nan(\"\")";
    }
    if (std::isinf(v)) {
        return v > 0 ? "__builtin_inf()" : "(-__builtin_inf())";
    }
    char text[64];
    std::snprintf(text, sizeof(text), "%a", v); // exact: a hexadecimal double literal
    return std::string("(") + text + ")";
}

class Generator {
public:
    Generator(const ir::Module& m, const Law& law, LawLayout& lay) : m_(m), law_(law), lay_(lay) {}

    std::string run(std::string& why)
    {
        (void)why; // every op of the IR has a form here
        allocate();
        lay_.fields = static_cast<int>(law_.fields.size());
        const int nf = std::max(1, lay_.fields);
        out_ << "// generated by the Field Engine's --hip emitter for gfx950\n";
        out_ << FE_EXACT_HELPERS_TEXT;
        // The params travel by value in the arguments (the kernarg segment is read-only: the compiler hoists every read
        // of them out of loops), so a step's launches carry the values they were queued with.
        const int np = std::max(1, static_cast<int>(law_.param_values.size()));
        const int nu = std::max(1, lay_.uniforms);
        out_ << "struct K {\n    float* now[" << nf << "];\n    float* nxt[" << nf << "];\n"
             << "    const double* U;\n    float* A;\n    double* S;\n    double* L;\n    unsigned* C;\n    const float* N;\n"
             << "    long long nxy;\n    long long cells;\n    double h;\n    double dt;\n    double inv_h;\n    double inv_hh;\n"
             << "    int nx;\n    int ny;\n    int nz;\n    int flags;\n    double P[" << np << "];\n};\n";
        lay_.kernel_arg_bytes = (2 * nf + 6) * 8 + 6 * 8 + 4 * 4 + 8 * np;
        out_ << "#define FE_BOUND_SLOTS " << kBoundSlots << "\n";
        out_ << kPrelude;
        // Reductions' results are read once a kernel into registers (a load inside the z loop would be repeated and,
        // waited for, would drain the loads in flight).
        out_ << "struct Direct {\n    static constexpr bool kClamps = true;\n    static constexpr bool kExact = true;\n"
             << "    static constexpr bool kPow2 = false;\n    __device__ void flag(bool) const {}\n    double u[" << nu << "];\n"
             << "    __device__ explicit Direct(const K& k) { for (int q = 0; q < " << nu << "; ++q) u[q] = " << (lay_.uniforms > 0 ? "k.U[q]" : "0.0")
             << "; }\n"
             << "    template <int F> __device__ __forceinline__ float fld(const K& k, int x, int y, int z) const { return k.now[F][IDX(x, y, z)]; }\n"
             << "    template <int F> __device__ __forceinline__ float fld_any(const K& k, int x, int y, int z) const { return k.now[F][IDX(x, y, z)]; }\n"
             << "};\n"
             // The direct view specialised: W32, reads with 32-bit index arithmetic (lattices of fewer than 2^31 cells: a
             // 64-bit product a read is several instructions, and a law's gather makes dozens of reads a cell); P2, h a
             // power of two, so a division by h or h^2 is the exact product with the reciprocal and no division is left
             // in the code (a run-time choice between the two would have the card compute both).
             << "template <bool W32, bool P2> struct DirectT : Direct {\n    static constexpr bool kPow2 = P2;\n"
             << "    __device__ explicit DirectT(const K& k) : Direct(k) {}\n"
             << "    template <int F> __device__ __forceinline__ float fld(const K& k, int x, int y, int z) const\n"
             << "    { return W32 ? k.now[F][(unsigned)z * (unsigned)k.nxy + (unsigned)y * (unsigned)k.nx + (unsigned)x] : k.now[F][IDX(x, y, z)]; }\n"
             << "    template <int F> __device__ __forceinline__ float fld_any(const K& k, int x, int y, int z) const { return fld<F>(k, x, y, z); }\n"
             << "};\n"
             << "#define FE_POW2(k) (((k).flags & 3) == 3)\n";
        std::vector<std::size_t> offsets;
        (void)net_blob(law_, offsets);
        if (!law_.nets.empty()) {
            // std::min / std::max on floats
            out_ << "__device__ __forceinline__ float fe_minf(float a, float b) { return b < a ? b : a; }\n"
                 << "__device__ __forceinline__ float fe_maxf(float a, float b) { return a < b ? b : a; }\n";
        }
        for (std::size_t n = 0; n < law_.nets.size(); ++n) {
            net_function(static_cast<int>(n), offsets[n]);
        }
        for (std::size_t id = 0; id < m_.insts.size(); ++id) {
            if (used(static_cast<int>(id))) {
                emit(static_cast<int>(id));
            }
        }
        for (const auto& [id, stage] : lay_.stages) {
            stage_kernel(id);
        }
        for (std::size_t t = 0; t < m_.terms.size(); ++t) {
            lay_.terms.push_back(program(m_.terms[t], "t" + std::to_string(t), false));
            lay_.learned.push_back(program(m_.learned[t], "l" + std::to_string(t), false));
        }
        for (std::size_t d = 0; d < m_.derives.size(); ++d) {
            lay_.derives.push_back(program(m_.derives[d], "d" + std::to_string(d), true));
        }
        return out_.str();
    }

private:
    [[nodiscard]] const ir::Inst& at(int id) const { return m_.insts[static_cast<std::size_t>(id)]; }
    [[nodiscard]] bool scalar(int id) const { return at(id).shape == ir::Shape::Scalar; }

    // Whether any program reaches the instruction (folding leaves dead ones behind).
    bool used(int id)
    {
        if (live_.empty()) {
            live_.assign(m_.insts.size(), 0);
            std::vector<int> todo;
            auto root = [&](int v) {
                if (v >= 0) {
                    todo.push_back(v);
                }
            };
            for (const auto* progs : {&m_.terms, &m_.learned, &m_.derives}) {
                for (const ir::Program& p : *progs) {
                    for (const ir::Equation& e : p.equations) {
                        root(e.value);
                    }
                    root(p.max_dt);
                }
            }
            while (!todo.empty()) {
                const int v = todo.back();
                todo.pop_back();
                if (live_[static_cast<std::size_t>(v)]) {
                    continue;
                }
                live_[static_cast<std::size_t>(v)] = 1;
                for (const int a : at(v).args) {
                    todo.push_back(a);
                }
            }
        }
        return live_[static_cast<std::size_t>(id)] != 0;
    }

    void allocate()
    {
        lay_.slot.assign(m_.insts.size(), -1);
        lay_.uslot.assign(m_.insts.size(), -1);
        std::set<int> stages;
        for (const auto* progs : {&m_.terms, &m_.learned, &m_.derives}) {
            for (const ir::Program& p : *progs) {
                stages.insert(p.stages.begin(), p.stages.end());
                stages.insert(p.limit_stages.begin(), p.limit_stages.end());
            }
        }
        for (const int id : stages) {
            const ir::Inst& in = at(id);
            StageCode s;
            s.inst = id;
            s.kernel = "st" + std::to_string(id);
            if (in.op == Op::MaxOf || in.op == Op::MinOf || in.op == Op::SumOf) {
                s.kind = StageCode::Kind::Reduce;
                s.reduce = in.op;
                lay_.uslot[static_cast<std::size_t>(id)] = lay_.uniforms++;
            } else if (in.op == Op::Tridiag) {
                s.kind = StageCode::Kind::Tridiag;
                s.axis = in.index;
                s.uniform = scalar(in.args[0]) && scalar(in.args[1]) && scalar(in.args[2]);
                lay_.slot[static_cast<std::size_t>(id)] = lay_.aux_slots++;
            } else if ((in.op == Op::Mlp || in.op == Op::Flux || in.op == Op::LDiffuse) && in.shape == ir::Shape::Cell &&
                law_.nets[static_cast<std::size_t>(in.index)].half) {
                const NetDecl& net = law_.nets[static_cast<std::size_t>(in.index)];
                s.kind = StageCode::Kind::NetHalf;
                s.net = in.index;
                s.net_op = in.op;
                s.rows_per_cell = in.op == Op::Mlp ? 1 : 6;
                s.in_pad = (net.inputs() + 15) / 16 * 16;
                s.outs = net.outputs();
                lay_.slot[static_cast<std::size_t>(id)] = lay_.aux_slots++;
            } else if (in.op == Op::Distance) {
                s.kind = StageCode::Kind::Distance;
                lay_.slot[static_cast<std::size_t>(id)] = lay_.aux_slots;
                lay_.aux_slots += 6;
                s.words = lay_.stage_words;
                lay_.stage_words += 4;
            } else if (in.op == Op::Poisson) {
                s.kind = StageCode::Kind::Poisson;
                s.iterations = in.iterations;
                s.scalar_source = scalar(in.args[0]);
                lay_.slot[static_cast<std::size_t>(id)] = lay_.aux_slots;
                lay_.aux_slots += 3;
                s.words = lay_.stage_words;
                lay_.stage_words += 4;
            } else {
                s.kind = StageCode::Kind::Stored;
                lay_.slot[static_cast<std::size_t>(id)] = lay_.aux_slots++;
            }
            lay_.stages.emplace(id, s);
        }
    }

    // A neighbour's coordinate along axis a (0 x, 1 y, 2 z) for reading argument `arg` at `v`: clamped into the lattice,
    // or through LX/LY/LZ when the argument is a field itself (the memory it is read from decides).
    std::string cl(int arg, int a, const std::string& v) const
    {
        static const char* direct[3] = {"ex(k, ", "ey(k, ", "ez(k, "};
        static const char* leaf[3] = {"LX(", "LY(", "LZ("};
        const bool is_leaf = at(arg).op == Op::Field && lay_.stages.count(arg) == 0;
        return std::string(is_leaf ? leaf[a] : direct[a]) + v + ")";
    }

    // A value at a point. Where the point is a runtime coordinate (ops that walk or choose their cells), a field is
    // read through vr<id>: straight from the window, without the checks that serve known offsets from registers.
    std::string v(int id, const std::string& x = "x", const std::string& y = "y", const std::string& z = "z") const
    {
        const bool leaf = at(id).op == Op::Field && lay_.stages.count(id) == 0;
        return std::string(leaf && runtime_ > 0 ? "vr" : "v") + std::to_string(id) + "(k, m, " + x + ", " + y + ", " + z + ")";
    }

    struct Runtime {
        int& n;
        explicit Runtime(int& c) : n(c) { ++n; }
        ~Runtime() { --n; }
    };
    int runtime_ = 0;

    // An argument as a double at a point: a scalar as computed, a cell value widened from its float.
    std::string d(int id, const std::string& x = "x", const std::string& y = "y", const std::string& z = "z") const
    {
        if (scalar(id)) {
            return "s" + std::to_string(id) + "(k, m)";
        }
        return "(double)" + v(id, x, y, z);
    }

    void emit(int id)
    {
        const ir::Inst& in = at(id);
        const std::string n = std::to_string(id);
        std::ostringstream f;
        if (in.shape == ir::Shape::Scalar) {
            f << "template <class M> __device__ __forceinline__ double s" << n << "(const K& k, const M& m) { (void)m; ";
            switch (in.op) {
            case Op::Const: f << "return " << literal(in.number) << ";"; break;
            case Op::Param: f << "return k.P[" << in.index << "];"; break;
            case Op::Cell: f << "return k.h;"; break;
            case Op::Dt: f << "return k.dt;"; break;
            case Op::Local: f << "return 0.0;"; break;
            case Op::MaxOf: case Op::MinOf: case Op::SumOf: {
                const int u = lay_.uslot[static_cast<std::size_t>(id)];
                f << "return " << (u >= 0 ? "m.u[" + std::to_string(u) + "]" : std::string("0.0")) << ";";
                break;
            }
            case Op::Mlp: {
                // every input one number: the net once (src/field/Law.cpp:1310-1318)
                const NetDecl& net = law_.nets[static_cast<std::size_t>(in.index)];
                f << "float xin[" << in.args.size() << "]; ";
                for (std::size_t q = 0; q < in.args.size(); ++q) {
                    f << "xin[" << q << "] = (float)s" << in.args[q] << "(k, m); ";
                }
                f << "float yout[" << net.outputs() << "]; net_" << in.index << "(k, xin, yout); return (double)yout[0];";
                break;
            }
            default: f << "return " << arith(in, id) << ";"; break;
            }
            f << " }\n";
            out_ << f.str();
            return;
        }
        const auto stage = lay_.stages.find(id);
        if (stage != lay_.stages.end() && stage->second.kind == StageCode::Kind::NetHalf) {
            f << "template <class M> __device__ __forceinline__ float v" << n << "(const K& k, const M& m, int x, int y, int z)\n{\n"
              << "    (void)m;\n    return AT(" << lay_.slot[static_cast<std::size_t>(id)] << ", x, y, z);\n}\n";
            out_ << f.str();
            return;
        }
        const bool stored = stage != lay_.stages.end() && stage->second.kind == StageCode::Kind::Stored;
        const std::string head = "template <class M> __device__ __forceinline__ float " + std::string(stored ? "b" : "v") + n +
            "(const K& k, const M& m, int x, int y, int z)";
        f << head << "\n{\n";
        switch (in.op) {
        case Op::Field:
            f << "    return m.template fld<" << in.index << ">(k, x, y, z);\n}\n"
              << "template <class M> __device__ __forceinline__ float vr" << n << "(const K& k, const M& m, int x, int y, int z)\n{\n"
              << "    return m.template fld_any<" << in.index << ">(k, x, y, z);\n";
            break;
        case Op::Rim:
            f << "    return (x == 0 || y == 0 || z == 0 || x + 1 == k.nx || y + 1 == k.ny || z + 1 == k.nz) ? 1.0f : 0.0f;\n";
            break;
        case Op::Shift: {
            const int a = in.args[0];
            f << "    return " << v(a, cl(a, 0, "x + (" + std::to_string(in.offset[0]) + ")"), cl(a, 1, "y + (" + std::to_string(in.offset[1]) + ")"),
                cl(a, 2, "z + (" + std::to_string(in.offset[2]) + ")"))
              << ";\n";
            break;
        }
        case Op::Mul: {
            // Zero times anything is zero on the CPU, without looking at the anything (src/field/Law.cpp:804-843).
            const int a = in.args[0];
            const int b = in.args[1];
            if (scalar(a) || scalar(b)) {
                const int s = scalar(a) ? a : b;
                f << "    const double s = s" << s << "(k, m);\n    if (s == 0.0) return 0.0f;\n";
            }
            f << "    return fe_rf(" << d(a) << " * " << d(b) << ");\n";
            break;
        }
        case Op::Lap: {
            // (xm + xp + ym + yp + zm + zp - 6 c) / h^2 in double, left to right; not finite -> 0.
            const int a = in.args[0];
            f << "    const double c = " << d(a) << ";\n"
              << "    double s = " << d(a, cl(a, 0, "x - 1")) << " + " << d(a, cl(a, 0, "x + 1")) << ";\n"
              << "    s = s + " << d(a, "x", cl(a, 1, "y - 1")) << ";\n"
              << "    s = s + " << d(a, "x", cl(a, 1, "y + 1")) << ";\n"
              << "    s = s + " << d(a, "x", "y", cl(a, 2, "z - 1")) << ";\n"
              << "    s = s + " << d(a, "x", "y", cl(a, 2, "z + 1")) << ";\n"
              << "    const double o = div_hh<M>(k, s - 6.0 * c);\n"
              << "    return fe_rf(finite_d(o) ? o : 0.0);\n";
            break;
        }
        case Op::Gx: case Op::Gy: case Op::Gz: {
            const int axis = in.op == Op::Gx ? 0 : (in.op == Op::Gy ? 1 : 2);
            const char* c = axis == 0 ? "x" : (axis == 1 ? "y" : "z");
            const char* cnt = axis == 0 ? "k.nx" : (axis == 1 ? "k.ny" : "k.nz");
            const std::string hi = std::string(c) + " + 1 < " + cnt + " ? " + c + " + 1 : " + c;
            const std::string lo = std::string(c) + " > 0 ? " + c + " - 1 : " + c;
            f << "    const int hi = " << hi << ";\n    const int lo = " << lo << ";\n";
            const std::string ahi = axis == 0 ? d(in.args[0], "hi") : (axis == 1 ? d(in.args[0], "x", "hi") : d(in.args[0], "x", "y", "hi"));
            const std::string alo = axis == 0 ? d(in.args[0], "lo") : (axis == 1 ? d(in.args[0], "x", "lo") : d(in.args[0], "x", "y", "lo"));
            const std::string body = "div_hs<M>(k, " + ahi + " - " + alo + ", hi - lo)";
            if (axis == 0) {
                f << "    const double o = " << body << ";\n";
            } else {
                f << "    const double o = " << cnt << " > 1 ? " << body << " : 0.0;\n";
            }
            f << "    return fe_rf(finite_d(o) ? o : 0.0);\n";
            break;
        }
        case Op::Diffuse: {
            // div(k grad F) in flux form: faces -x +x -y +y -z +z, each kf (F[nb] - F[c]) with the harmonic mean
            // kf = 2 ka kb / (ka + kb) (src/field/Law.cpp:986-1018), a missing face skipped.
            const int kk = in.args[0];
            const int fa = in.args[1];
            // Exact form (M::kExact): the CPU's, missing faces skipped. Fast form: every neighbour is read at its clamped
            // coordinate, so across a missing face the neighbour is the cell and the face's term is kf * (c - c), +0 or
            // -0 when c and kf are finite and NaN otherwise. The sum starts at +0 and can never become -0, so adding a
            // zero changes no bit: when the six-term sum is not NaN it is the CPU's sum bit for bit. A NaN sum flags
            // the cell (m.flag) and its caller takes it again in the exact form. No selects and no branches: the reads
            // of neighbouring cells merge. The -x face is written -(kf * (c - n0)), the same bits as kf * (n0 - c), so
            // it is literally the +x flux of the cell before it (shared when a lane computes both cells).
            f << "    const double c = " << d(fa) << ";\n";
            if (scalar(kk)) {
                f << "    const double kv = s" << kk << "(k, m);\n"
                  << "    const double kf = (kv + kv) > 0.0 ? 2.0 * kv * kv / (kv + kv) : 0.0;\n";
            } else {
                f << "    const double ka = " << d(kk) << ";\n";
            }
            const char* cond[6] = {"x > 0", "x + 1 < k.nx", "y > 0", "y + 1 < k.ny", "z > 0", "z + 1 < k.nz"};
            const int axis[6] = {0, 0, 1, 1, 2, 2};
            const char* step[6] = {"x - 1", "x + 1", "y - 1", "y + 1", "z - 1", "z + 1"};
            auto at_face = [&](int arg, int q) {
                const std::string c = cl(arg, axis[q], step[q]);
                return axis[q] == 0 ? d(arg, c) : (axis[q] == 1 ? d(arg, "x", c) : d(arg, "x", "y", c));
            };
            for (int q = 0; q < 6; ++q) {
                f << "    const double n" << q << " = " << at_face(fa, q) << ";\n";
                if (!scalar(kk)) {
                    f << "    const double kb" << q << " = " << at_face(kk, q) << ";\n"
                      << "    const double kf" << q << " = (ka + kb" << q << ") > 0.0 ? 2.0 * ka * kb" << q << " / (ka + kb" << q
                      << ") : 0.0;\n";
                }
            }
            auto term = [&](int q) {
                const std::string kf = scalar(kk) ? std::string("kf") : "kf" + std::to_string(q);
                const std::string n = "n" + std::to_string(q);
                return q == 0 && scalar(kk) ? "-(" + kf + " * (c - " + n + "))" : kf + " * (" + n + " - c)";
            };
            f << "    double acc = 0.0;\n    if constexpr (M::kExact) {\n";
            for (int q = 0; q < 6; ++q) {
                f << "        if (" << cond[q] << ") acc += " << term(q) << ";\n";
            }
            f << "    } else {\n";
            for (int q = 0; q < 6; ++q) {
                f << "        acc += " << term(q) << ";\n";
            }
            f << "        m.flag(acc != acc);\n    }\n    return fe_rf(div_hh<M>(k, acc));\n";
            break;
        }
        case Op::Adv: {
            // -(u . grad F), upwind (src/field/Law.cpp:1020-1040); neighbours read first, the upwind choice a select.
            const int fa = in.args[0];
            f << "    const double c = " << d(fa) << ";\n"
              << "    const double u = " << d(in.args[1]) << ";\n    const double w = " << d(in.args[2]) << ";\n"
              << "    const double s = " << d(in.args[3]) << ";\n"
              << "    const double xl = " << d(fa, cl(fa, 0, "x - 1")) << "; const double xh = " << d(fa, cl(fa, 0, "x + 1")) << ";\n"
              << "    const double yl = " << d(fa, "x", cl(fa, 1, "y - 1")) << "; const double yh = " << d(fa, "x", cl(fa, 1, "y + 1"))
              << ";\n"
              << "    const double zl = " << d(fa, "x", "y", cl(fa, 2, "z - 1")) << "; const double zh = "
              << d(fa, "x", "y", cl(fa, 2, "z + 1")) << ";\n";
            auto up = [&](const char* speed, const char* c, const char* cnt, const char* lo, const char* hi) {
                return std::string("(") + speed + " > 0.0 ? (" + c + " > 0 ? div_h<M>(k, " + speed + " * (c - " + lo + ")) : 0.0) : (" + c +
                    " + 1 < " + cnt + " ? div_h<M>(k, " + speed + " * (" + hi + " - c)) : 0.0))";
            };
            f << "    double o = " << up("u", "x", "k.nx", "xl", "xh") << ";\n"
              << "    o = o + " << up("w", "y", "k.ny", "yl", "yh") << ";\n"
              << "    o = o + " << up("s", "z", "k.nz", "zl", "zh") << ";\n"
              << "    return fe_rf(-o);\n";
            break;
        }
        case Op::ShiftT: {
            const Runtime walk(runtime_);
            f << "    int lx, hx, ly, hy, lz, hz;\n"
              << "    run_to(x, " << in.offset[0] << ", k.nx, lx, hx); run_to(y, " << in.offset[1] << ", k.ny, ly, hy); run_to(z, "
              << in.offset[2] << ", k.nz, lz, hz);\n"
              << "    double s = 0.0;\n"
              << "    for (int zz = lz; zz <= hz; ++zz) for (int yy = ly; yy <= hy; ++yy) for (int xx = lx; xx <= hx; ++xx) s += "
              << d(in.args[0], "xx", "yy", "zz") << ";\n"
              << "    return fe_rf(s);\n";
            break;
        }
        case Op::GxT: case Op::GyT: case Op::GzT: {
            const Runtime walk(runtime_);
            const int axis = in.op == Op::GxT ? 0 : (in.op == Op::GyT ? 1 : 2);
            const char* c = axis == 0 ? "x" : (axis == 1 ? "y" : "z");
            const char* cnt = axis == 0 ? "k.nx" : (axis == 1 ? "k.ny" : "k.nz");
            const std::string aq = axis == 0 ? d(in.args[0], "q") : (axis == 1 ? d(in.args[0], "x", "q") : d(in.args[0], "x", "y", "q"));
            f << "    double s = 0.0;\n"
              << "    for (int dd = 1; dd >= -1; dd -= 2) { int lo, hi; run_to(" << c << ", dd, " << cnt << ", lo, hi);\n"
              << "        for (int q = lo; q <= hi; ++q) {\n"
              << "            const int span = (q + 1 < " << cnt << " - 1 ? q + 1 : " << cnt << " - 1) - (q - 1 > 0 ? q - 1 : 0);\n"
              << "            const double w = span > 0 ? " << aq << " / (k.h * span) : 0.0;\n"
              << "            s += dd > 0 ? w : -w; } }\n"
              << "    return fe_rf(s);\n";
            break;
        }
        case Op::Tridiag: case Op::Distance: case Op::Poisson: // computed over the lattice by their stage kernels
            f << "    return AT(" << lay_.slot[static_cast<std::size_t>(id)] << ", x, y, z);\n";
            break;
        case Op::Fall: {
            const Runtime walk(runtime_);
            // F moving down into the room below (src/field/Law.cpp:1183-1210): the flux out of a cell through its
            // floor, in double; the cell loses its own and gains the one above's.
            if (scalar(in.args[0])) {
                f << "    return 0.0f;\n";
                break;
            }
            const int fa = in.args[0], ba = in.args[1], sa = in.args[2];
            f << "    if (k.nz < 2) return 0.0f;\n"
              << "    auto down = [&](int zz) -> double {\n"
              << "        if (zz == 0) return 0.0;\n"
              << "        const double room = fe_clamp(1.0 - " << d(ba, "x", "y", "zz - 1") << " - " << d(fa, "x", "y", "zz - 1") << ", 0.0, 1.0);\n"
              << "        return div_h<M>(k, fe_max(0.0, " << d(sa, "x", "y", "zz") << ")) * fe_max(0.0, " << d(fa, "x", "y", "zz")
              << ") * room;\n    };\n"
              << "    float out = fe_rf(0.0 - down(z));\n"
              << "    if (z + 1 < k.nz) out = fe_rf((double)out + down(z + 1));\n"
              << "    return out;\n";
            break;
        }
        case Op::Spread: {
            const Runtime walk(runtime_);
            // Sideways, downhill, into the room on the receiving side (src/field/Law.cpp:1212-1244): the flux across
            // each x and y face in double, summed into the cell in the CPU's order, rounding to float at each step.
            if (scalar(in.args[0])) {
                f << "    return 0.0f;\n";
                break;
            }
            const int fa = in.args[0], ba = in.args[1], ka = in.args[2];
            f << "    auto across = [&](int ax, int ay, int cx, int cy) -> double {\n"
              << "        const double fa = " << d(fa, "ax", "ay", "z") << ", fc = " << d(fa, "cx", "cy", "z") << ";\n"
              << "        const bool to_c = fa > fc;\n        const int rx = to_c ? cx : ax, ry = to_c ? cy : ay;\n"
              << "        const double room = fe_clamp(1.0 - " << d(ba, "rx", "ry", "z") << " - " << d(fa, "rx", "ry", "z") << ", 0.0, 1.0);\n"
              << "        const double kf = 0.5 * (fe_max(0.0, " << d(ka, "ax", "ay", "z") << ") + fe_max(0.0, " << d(ka, "cx", "cy", "z")
              << "));\n"
              << "        return div_hh<M>(k, kf) * (fa - fc) * room;\n    };\n"
              << "    float out = 0.0f;\n"
              << "    if (y > 0) out = fe_rf((double)out + across(x, y - 1, x, y));\n"
              << "    if (x > 0) out = fe_rf((double)out + across(x - 1, y, x, y));\n"
              << "    out = fe_rf((double)out - (x + 1 < k.nx ? across(x, y, x + 1, y) : 0.0));\n"
              << "    out = fe_rf((double)out - (y + 1 < k.ny ? across(x, y, x, y + 1) : 0.0));\n"
              << "    return out;\n";
            break;
        }
        case Op::Topple: {
            const Runtime walk(runtime_);
            // A pile slides diagonally down off what holds it (src/field/Law.cpp:1246-1296): what the cell sends down
            // its four sides, and what the four cells above-and-beside send toward this column, in double.
            if (scalar(in.args[0])) {
                f << "    return 0.0f;\n";
                break;
            }
            const int fa = in.args[0], ba = in.args[1], ra = in.args[2];
            f << "    if (k.nz < 2) return 0.0f;\n"
              << "    auto room = [&](int qx, int qy, int qz) -> double {\n"
              << "        return fe_clamp(1.0 - " << d(ba, "qx", "qy", "qz") << " - " << d(fa, "qx", "qy", "qz") << ", 0.0, 1.0);\n    };\n"
              << "    auto send = [&](int qx, int qy, int qz, int side) -> double {\n"
              << "        const double here = " << d(fa, "qx", "qy", "qz") << ";\n"
              << "        if (here <= 0.0 || qz == 0) return 0.0;\n"
              << "        if (room(qx, qy, qz - 1) > 0.01) return 0.0;\n"
              << "        int jx = qx, jy = qy;\n"
              << "        if (side == 0) { if (qx == 0) return 0.0; jx = qx - 1; }\n"
              << "        else if (side == 1) { if (qx + 1 >= k.nx) return 0.0; jx = qx + 1; }\n"
              << "        else if (side == 2) { if (qy == 0) return 0.0; jy = qy - 1; }\n"
              << "        else { if (qy + 1 >= k.ny) return 0.0; jy = qy + 1; }\n"
              << "        return div_h<M>(k, 0.25 * fe_max(0.0, " << d(ra, "qx", "qy", "qz") << ")) * here * fe_min(room(jx, jy, qz), room(jx, jy, qz - 1));\n"
              << "    };\n"
              << "    double out = 0.0;\n"
              << "    out -= send(x, y, z, 0);\n    out -= send(x, y, z, 1);\n    out -= send(x, y, z, 2);\n    out -= send(x, y, z, 3);\n"
              << "    if (z + 1 < k.nz) {\n"
              << "        if (x + 1 < k.nx) out += send(x + 1, y, z + 1, 0);\n"
              << "        if (x > 0) out += send(x - 1, y, z + 1, 1);\n"
              << "        if (y + 1 < k.ny) out += send(x, y + 1, z + 1, 2);\n"
              << "        if (y > 0) out += send(x, y - 1, z + 1, 3);\n"
              << "    }\n"
              << "    return fe_rf(out);\n";
            break;
        }
        case Op::Mlp: {
            // the net at the cell, inputs the arguments there as floats (src/field/Law.cpp:1320-1333)
            const NetDecl& net = law_.nets[static_cast<std::size_t>(in.index)];
            f << "    float xin[" << in.args.size() << "];\n";
            for (std::size_t q = 0; q < in.args.size(); ++q) {
                f << "    xin[" << q << "] = (float)" << d(in.args[q]) << ";\n";
            }
            f << "    float yout[" << net.outputs() << "];\n    net_" << in.index << "(k, xin, yout);\n    return yout[0];\n";
            break;
        }
        case Op::Flux: case Op::LDiffuse: {
            const Runtime walk(runtime_);
            net_faces(in, f);
            break;
        }
        default:
            f << "    return fe_rf(" << arith(in, id) << ");\n";
            break;
        }
        f << "}\n";
        out_ << f.str();
        if (stored) {
            out_ << "template <class M> __device__ __forceinline__ float v" << n
                 << "(const K& k, const M& m, int x, int y, int z) { (void)m; return AT(" << lay_.slot[static_cast<std::size_t>(id)]
                 << ", x, y, z); }\n";
        }
    }

    // net_<n>(k, in, out): the CPU's forward pass (src/field/nn/Net.cpp:96-152) in fp32 in its order: inputs
    // normalised ((x - shift) * scale), each layer acc = b[o] then += W[o][i] x[i] for i in order, the activation on
    // every layer but the last, outputs times out_scale. No multiply-adds are fused (the CPU has none), so nets of
    // relu and linear layers give the CPU's bits; tanh and silu go through the card's own tanhf and expf (within an ulp
    // of the C library's).
    void net_function(int n, std::size_t off)
    {
        const NetDecl& net = law_.nets[static_cast<std::size_t>(n)];
        const int nin = net.inputs();
        std::ostringstream f;
        f << "__device__ __forceinline__ void net_" << n << "(const K& k, const float* in, float* out)\n{\n"
          << "    const float* P = k.N + " << off << ";\n"
          << "    float h0[" << nin << "];\n"
          << "#pragma unroll\n    for (int i = 0; i < " << nin << "; ++i) h0[i] = (in[i] - P[i]) * P[" << nin << " + i];\n";
        std::size_t w = static_cast<std::size_t>(2 * nin + 1);
        for (int l = 0; l < net.layers(); ++l) {
            const int ni = net.sizes[static_cast<std::size_t>(l)], no = net.sizes[static_cast<std::size_t>(l) + 1];
            const bool last = l + 1 == net.layers();
            f << "    float h" << l + 1 << "[" << no << "];\n"
              << "#pragma unroll\n    for (int o = 0; o < " << no << "; ++o) {\n"
              << "        const float* row = P + " << w << " + o * " << ni << ";\n"
              << "        float acc = P[" << w + static_cast<std::size_t>(ni) * static_cast<std::size_t>(no) << " + o];\n"
              << "#pragma unroll\n        for (int i = 0; i < " << ni << "; ++i) acc += row[i] * h" << l << "[i];\n";
            if (last) {
                f << "        h" << l + 1 << "[o] = acc;\n";
            } else {
                // By the activation's name: the reference's Activation has no gelu, and this generator compiles
                // against either header set.
                const std::string act = to_string(net.act);
                f << "        h" << l + 1 << "[o] = ";
                if (act == "tanh") {
                    f << "tanhf(acc);\n";
                } else if (act == "relu") {
                    f << "acc > 0.0f ? acc : 0.0f;\n";
                } else if (act == "silu") {
                    f << "acc / (1.0f + expf(-acc));\n";
                } else if (act == "gelu") { // the tanh form
                    f << "0.5f * acc * (1.0f + tanhf(0.7978845608028654f * (acc + 0.044715f * acc * acc * acc)));\n";
                } else {
                    f << "acc;\n";
                }
            }
            f << "    }\n";
            w += static_cast<std::size_t>(ni + 1) * static_cast<std::size_t>(no);
        }
        f << "#pragma unroll\n    for (int o = 0; o < " << net.outputs() << "; ++o) out[o] = h" << net.layers() << "[o] * P[" << 2 * nin
          << "];\n}\n";
        out_ << f.str();
    }

    // flux(NET, F, a..., lo, hi) and ldiffuse(NET, k, F) at a cell: its six faces' fluxes in the CPU's order
    // (src/field/Law.cpp:1379-1389), each computed here (a face's flux comes out the same in both its cells: the same
    // inputs, the same arithmetic), out = sum (inflow - outflow) in float, over h.
    void net_faces(const ir::Inst& in, std::ostringstream& f)
    {
        const NetDecl& net = law_.nets[static_cast<std::size_t>(in.index)];
        const bool flux = in.op == Op::Flux;
        if (flux) {
            // face i -> j: the net both ways round, half the difference, limited (src/field/Law.cpp:673-707)
            const std::size_t mm = in.args.size() - 2;
            const double lo = at(in.args[mm]).number, hi = at(in.args[mm + 1]).number;
            f << "    auto face = [&](int xi, int yi, int zi, int xj, int yj, int zj) -> float {\n"
              << "        float xij[" << 2 * mm << "], xji[" << 2 * mm << "], y[" << net.outputs() << "];\n";
            for (std::size_t q = 0; q < mm; ++q) {
                f << "        xij[" << q << "] = (float)" << d(in.args[q], "xi", "yi", "zi") << "; xij[" << mm + q << "] = (float)"
                  << d(in.args[q], "xj", "yj", "zj") << ";\n"
                  << "        xji[" << q << "] = xij[" << mm + q << "]; xji[" << mm + q << "] = xij[" << q << "];\n";
            }
            f << "        net_" << in.index << "(k, xij, y); const float gij = y[0];\n"
              << "        net_" << in.index << "(k, xji, y); const float gji = y[0];\n"
              << "        float r = 0.5f * (gij - gji);\n"
              << "        if (k.dt > 0.0) {\n"
              << "            const float per = (float)(0.999 * k.h / (6.0 * k.dt));\n"
              << "            const float fi = xij[0], fj = xij[" << mm << "];\n";
            const bool has_lo = lo > -1e29, has_hi = hi < 1e29;
            f << "            if (r > 0.0f) {\n";
            if (has_lo) {
                f << "                r = fe_minf(r, fe_maxf(0.0f, (fi - (float)" << literal(lo) << ") * per));\n";
            }
            if (has_hi) {
                f << "                r = fe_minf(r, fe_maxf(0.0f, ((float)" << literal(hi) << " - fj) * per));\n";
            }
            f << "            } else if (r < 0.0f) {\n";
            if (has_lo) {
                f << "                r = fe_maxf(r, -fe_maxf(0.0f, (fj - (float)" << literal(lo) << ") * per));\n";
            }
            if (has_hi) {
                f << "                r = fe_maxf(r, -fe_maxf(0.0f, ((float)" << literal(hi) << " - fi) * per));\n";
            }
            f << "            }\n        }\n        return r;\n    };\n";
        } else {
            // ldiffuse: the face's window F[i-1] F[i] | F[j] F[j+1] weighted by what the net sees (src/field/Law.cpp:709-738)
            const int ka = in.args[0], fa = in.args[1];
            f << "    auto raw = [&](const float* w, float kf) -> float {\n"
              << "        const float centre = 0.5f * (w[1] + w[2]);\n        float spread = 1e-6f;\n"
              << "        for (int a = 0; a < 4; ++a) spread = fe_maxf(spread, fabsf(w[a] - centre));\n"
              << "        float x[5];\n        for (int a = 0; a < 4; ++a) x[a] = (w[a] - centre) / spread;\n"
              << "        const float dtf = (float)k.dt, hf = (float)k.h;\n"
              << "        x[4] = kf * dtf / (hf * hf);\n"
              << "        float y[4];\n        net_" << in.index << "(k, x, y);\n"
              << "        const float mean = 0.25f * (y[0] + y[1] + y[2] + y[3]);\n"
              << "        const float base[4] = {0.0f, -1.0f, 1.0f, 0.0f};\n        float acc = 0.0f;\n"
              << "        for (int a = 0; a < 4; ++a) acc += (base[a] + (y[a] - mean)) * w[a];\n"
              << "        return -kf / hf * acc;\n    };\n"
              << "    auto face = [&](int axis, int xi, int yi, int zi) -> float {\n"
              << "        const int cnt = axis == 0 ? k.nx : (axis == 1 ? k.ny : k.nz);\n"
              << "        const int ci = axis == 0 ? xi : (axis == 1 ? yi : zi);\n"
              << "        const int sx = axis == 0, sy = axis == 1, sz = axis == 2;\n"
              << "        const int xj = xi + sx, yj = yi + sy, zj = zi + sz;\n"
              << "        const int xm = ci > 0 ? xi - sx : xi, ym = ci > 0 ? yi - sy : yi, zm = ci > 0 ? zi - sz : zi;\n"
              << "        const int xp = ci + 2 < cnt ? xj + sx : xj, yp = ci + 2 < cnt ? yj + sy : yj, zp = ci + 2 < cnt ? zj + sz : zj;\n"
              << "        const float w[4] = {(float)" << d(fa, "xm", "ym", "zm") << ", (float)" << d(fa, "xi", "yi", "zi") << ", (float)"
              << d(fa, "xj", "yj", "zj") << ", (float)" << d(fa, "xp", "yp", "zp") << "};\n"
              << "        const double ka = " << d(ka, "xi", "yi", "zi") << ", kb = " << d(ka, "xj", "yj", "zj") << ";\n"
              << "        const float kf = (float)((ka + kb) > 0.0 ? 2.0 * ka * kb / (ka + kb) : 0.0);\n"
              << "        const float mirror[4] = {w[3], w[2], w[1], w[0]};\n"
              << "        const float ij = raw(w, kf), ji = raw(mirror, kf);\n"
              << "        return 0.5f * (ij - ji);\n    };\n";
        }
        // the cell: in through its -x, -y, -z faces, out through +x, +y, +z, in that order
        if (flux) {
            f << "    float out = 0.0f;\n"
              << "    if (x > 0) out += face(x - 1, y, z, x, y, z);\n"
              << "    if (x + 1 < k.nx) out -= face(x, y, z, x + 1, y, z);\n"
              << "    if (y > 0) out += face(x, y - 1, z, x, y, z);\n"
              << "    if (y + 1 < k.ny) out -= face(x, y, z, x, y + 1, z);\n"
              << "    if (z > 0) out += face(x, y, z - 1, x, y, z);\n"
              << "    if (z + 1 < k.nz) out -= face(x, y, z, x, y, z + 1);\n";
        } else {
            f << "    float out = 0.0f;\n"
              << "    if (x > 0) out += face(0, x - 1, y, z);\n"
              << "    if (x + 1 < k.nx) out -= face(0, x, y, z);\n"
              << "    if (y > 0) out += face(1, x, y - 1, z);\n"
              << "    if (y + 1 < k.ny) out -= face(1, x, y, z);\n"
              << "    if (z > 0) out += face(2, x, y, z - 1);\n"
              << "    if (z + 1 < k.nz) out -= face(2, x, y, z);\n";
        }
        f << "    return out / (float)k.h;\n";
    }

    // An element-wise op in double over its arguments at (x, y, z), as the CPU's lambdas compute it
    // (src/field/Law.cpp:782-857): std::min/std::max/std::clamp comparisons, not fmin/fmax.
    std::string arith(const ir::Inst& in, int id) const
    {
        auto a = [&](std::size_t i) { return d(in.args[i]); };
        switch (in.op) {
        case Op::Neg: return "(-" + a(0) + ")";
        case Op::Abs: return "fabs(" + a(0) + ")";
        case Op::Sqrt: return "([](double q) { return sqrt(0.0 < q ? q : 0.0); })(" + a(0) + ")";
        case Op::Exp: return "([](double q) { return exp(80.0 < q ? 80.0 : q); })(" + a(0) + ")";
        case Op::Log: return "([](double q) { return log(q < 1e-30 ? 1e-30 : q); })(" + a(0) + ")";
        case Op::Step: return "(" + a(0) + " > 0.0 ? 1.0 : 0.0)";
        case Op::Pos: return "([](double q) { return q < 0.0 ? 0.0 : q; })(" + a(0) + ")";
        case Op::Add: return "(" + a(0) + " + " + a(1) + ")";
        case Op::Sub: return "(" + a(0) + " - " + a(1) + ")";
        case Op::Mul: return "([](double p, double q) { return p == 0.0 || q == 0.0 ? 0.0 : p * q; })(" + a(0) + ", " + a(1) + ")";
        case Op::Div: {
            // A division by h or h * h (the lattice spacing): the product with the reciprocal where that is a power of
            // two (the quotient's bits, as div_h and div_hh give the stencils), no division on the card.
            const ir::Inst& q = at(in.args[1]);
            const bool by_h = q.op == Op::Cell;
            const bool by_hh = q.op == Op::Mul && at(q.args[0]).op == Op::Cell && at(q.args[1]).op == Op::Cell;
            if (by_h || by_hh) {
                return std::string("(k.h == 0.0 ? 0.0 : ") + (by_h ? "div_h" : "div_hh") + "<M>(k, " + a(0) + "))";
            }
            return "([](double p, double q) { return q == 0.0 ? 0.0 : p / q; })(" + a(0) + ", " + a(1) + ")";
        }
        case Op::Pow: return "pow(" + a(0) + ", " + a(1) + ")";
        case Op::Min: return "fe_min(" + a(0) + ", " + a(1) + ")";
        case Op::Max: return "fe_max(" + a(0) + ", " + a(1) + ")";
        case Op::Clamp: return "([](double q, double lo, double hi) { return fe_clamp(q, lo, fe_max(lo, hi)); })(" + a(0) + ", " +
            a(1) + ", " + a(2) + ")";
        default:
            throw std::runtime_error(std::string("no element-wise form for ") + ir::op_name(in.op) + " (instruction " +
                std::to_string(id) + ")");
        }
    }

    void stage_kernel(int id)
    {
        const StageCode& s = lay_.stages.at(id);
        const ir::Inst& in = at(id);
        const int slot = lay_.slot[static_cast<std::size_t>(id)];
        switch (s.kind) {
        case StageCode::Kind::Stored:
            out_ << "extern \"C\" __global__ void " << s.kernel << "(K k) { CELL_HEAD\n"
                 << "    if (FE_POW2(k)) { const DirectT<false, true> m(k); AT(" << slot << ", x, y, z) = b" << id << "(k, m, x, y, z); }\n"
                 << "    else { const Direct m(k); AT(" << slot << ", x, y, z) = b" << id << "(k, m, x, y, z); }\n}\n";
            break;
        case StageCode::Kind::Reduce: {
            // Pass 1: block b sums (or bounds) cells [b * chunk, ...), lane t every 256th; a fixed tree; L[3b..3b+2].
            const int arg = in.args[0];
            const int kind = in.op == Op::SumOf ? 0 : (in.op == Op::MaxOf ? 1 : 2);
            out_ << "extern \"C\" __global__ __launch_bounds__(256) void " << s.kernel << "(K k, i64 chunk) {\n"
                 << "    __shared__ double wv[4][4];\n"
                 << "    const int t = (int)threadIdx.x; const i64 begin = (i64)blockIdx.x * chunk; i64 end = begin + chunk;\n"
                 << "    if (end > k.cells) end = k.cells;\n";
            out_ << "    const Direct m(k);\n";
            if (scalar(arg)) {
                out_ << "    const double first = s" << arg << "(k, m);\n";
            } else {
                out_ << "    const double first = (double)v" << arg << "(k, m, 0, 0, 0);\n";
            }
            out_ << "    Sum s = {0.0, 0.0}; double a = first; double b = first;\n"
                 << "    for (i64 i = begin + t; i < end; i += 256) {\n"
                 << "        int x, y, z; cell_xyz(k, i, x, y, z);\n";
            if (scalar(arg)) {
                out_ << "        const double q = s" << arg << "(k, m);\n";
            } else {
                out_ << "        const double q = (double)v" << arg << "(k, m, x, y, z);\n";
            }
            out_ << "        add(s, q); a = q < a ? q : a; b = b < q ? q : b; }\n"
                 << "    block_sum_bounds(s, a, b, wv);\n"
                 << "    if (t == 0) { k.L[3 * blockIdx.x] = s.s + s.c; k.L[3 * blockIdx.x + 1] = a; k.L[3 * blockIdx.x + 2] = b; }\n"
                 << "}\n";
            (void)kind;
            break;
        }
        case StageCode::Kind::Tridiag:
            tridiag_kernels(s, in, slot);
            break;
        case StageCode::Kind::Distance:
            distance_kernels(s, in, slot);
            break;
        case StageCode::Kind::Poisson:
            poisson_kernels(s, in, slot);
            break;
        case StageCode::Kind::NetHalf:
            net_half_kernels(s, in, slot);
            break;
        }
    }

    // tridiag_x|y|z: a thread a line, Thomas in double with the edge cell repeating past the ends (its neighbour's
    // coefficient on the diagonal), as the CPU solves it (src/field/Law.cpp:1098-1138): the same operations in the same
    // order, so each line's answer is the CPU's to the bit. 64 lines a block, so a 256^3 lattice has one wave a SIMD
    // and nothing else to hide memory latency behind: every sweep fetches its next operands into registers while it
    // runs the current ones (double buffering).
    //  - Lines along y or z: neighbouring threads take neighbouring x, so each step's reads and writes are one coalesced
    //    row; runs of kTridiagRun steps; d' (and c') in the scratch S at the cells' own indices.
    //  - Lines along x: staged through LDS 32 cells at a time. The block's 64 threads read the cell-shaped coefficients
    //    of its 64 rows with lanes along x (coalesced; floats: the CPU widens each cell's float), then each thread walks
    //    its own row from LDS; d' (and c') go to S interleaved by line (S[i * lines + line]: contiguous across the
    //    threads), and a chunk's answers leave through LDS the same way.
    //  - a, b and u one number each (uniform): c' is the same for every line, so thread 0's is kept in LDS for the
    //    backward sweep and only d' goes through S. That kernel is for lines up to kTridiagLdsLine cells (the launch
    //    gives the LDS); <kernel>_s, the general form, takes longer ones.
    static constexpr int kTridiagRun = 16;

    void tridiag_kernels(const StageCode& s, const ir::Inst& in, int slot)
    {
        tridiag_kernel(s, in, slot, s.uniform);
        if (s.uniform) {
            tridiag_kernel(s, in, slot, false);
        }
    }

    void tridiag_kernel(const StageCode& s, const ir::Inst& in, int slot, bool u)
    {
        const char* names[4] = {"a", "b", "u", "r"};
        std::vector<int> cell; // the cell-shaped coefficients (by position)
        std::ostringstream fetch;
        for (int q = 0; q < 4; ++q) {
            const int arg = in.args[static_cast<std::size_t>(q)];
            if (scalar(arg)) {
                fetch << "    const double s_" << names[q] << " = " << d(arg) << ";\n";
            } else {
                cell.push_back(q);
            }
        }
        auto val = [&](int q, const std::string& at) {
            return scalar(in.args[static_cast<std::size_t>(q)]) ? "s_" + std::string(names[q]) : "(double)" + at;
        };
        // One forward step at i, the cell-shaped coefficients read through `at`; c' to LDS (thread 0) or S[cq].
        auto forward = [&](const std::function<std::string(int)>& at, const std::string& cq, const std::string& pad) {
            std::ostringstream f;
            f << pad << "const double a = " << val(0, at(0)) << "; const double uu = " << val(2, at(2)) << ";\n"
              << pad << "const double b = " << val(1, at(1)) << " + (i == 0 ? a : 0.0) + (i + 1 == n ? uu : 0.0);\n"
              << pad << "const double piv = i == 0 ? b : b - a * cp;\n"
              << pad << "cp = i + 1 == n ? 0.0 : uu / piv;\n"
              << pad << "dp = (" << val(3, at(3)) << " - (i == 0 ? 0.0 : a * dp)) / piv;\n"
              << pad << (u ? std::string("if (t == 0) t_cp[i] = cp;") : "k.S[" + cq + "] = cp;") << " k.S[k.cells + " << cq << "] = dp;\n";
            return f.str();
        };
        // Register arrays: `prefix`<name>[len] for each cell-shaped coefficient / for d' (and c' unless uniform).
        auto coef_arrays = [&](const std::string& prefix, const std::string& len) {
            std::string r;
            for (int q : cell) {
                r += "float " + prefix + names[q] + "[" + len + "]; ";
            }
            return r;
        };
        auto coef_load = [&](const std::string& prefix, const std::string& x, const std::string& y, const std::string& z) {
            std::string r;
            for (int q : cell) {
                r += " " + prefix + names[q] + "[j] = " + v(in.args[static_cast<std::size_t>(q)], x, y, z) + ";";
            }
            return r;
        };
        auto coef_zero = [&](const std::string& prefix) {
            std::string r;
            for (int q : cell) {
                r += " " + prefix + names[q] + "[j] = 0.0f;";
            }
            return r;
        };
        auto coef_move = [&](const std::string& to, const std::string& from) {
            std::string r;
            for (int q : cell) {
                r += " " + to + names[q] + "[j] = " + from + names[q] + "[j];";
            }
            return r;
        };
        auto sweep_arrays = [&](const std::string& prefix, const std::string& len) {
            return "double " + prefix + "d[" + len + "]; " + (u ? std::string() : "double " + prefix + "c[" + len + "]; ");
        };
        auto sweep_load = [&](const std::string& prefix, const std::string& q) {
            return prefix + "d[j] = k.S[k.cells + " + q + "];" + (u ? std::string() : " " + prefix + "c[j] = k.S[" + q + "];");
        };
        auto sweep_zero = [&](const std::string& prefix) {
            return prefix + "d[j] = 0.0;" + (u ? std::string() : " " + prefix + "c[j] = 0.0;");
        };
        auto sweep_move = [&](const std::string& to, const std::string& from) {
            return to + "d[j] = " + from + "d[j];" + (u ? std::string() : " " + to + "c[j] = " + from + "c[j];");
        };
        auto cp_of = [&](const std::string& prefix) { return u ? std::string("t_cp[i]") : prefix + "c[j]"; };

        out_ << "extern \"C\" __global__ __launch_bounds__(64) void " << s.kernel << (s.uniform && !u ? "_s" : "")
             << "(K k) {\n    const Direct m(k);\n";
        if (u) {
            out_ << "    extern __shared__ double t_cp[];\n";
        }
        out_ << fetch.str() << "    const int t = (int)threadIdx.x;\n";
        const std::string R = std::to_string(kTridiagRun);
        if (s.axis != 0) {
            const bool ya = s.axis == 1;
            const std::string yi = ya ? "i" : "y", zi = ya ? "z" : "i";
            const std::string at = "IDX(x, " + yi + ", " + zi + ")";
            out_ << "    const i64 lines = (i64)k.nx * " << (ya ? "k.nz" : "k.ny") << "; const i64 line = (i64)blockIdx.x * 64 + t;\n"
                 << "    const bool live = line < lines; const int n = " << (ya ? "k.ny" : "k.nz") << ";\n"
                 << "    const int x = (int)(line % k.nx); const int " << (ya ? "z" : "y") << " = (int)(line / k.nx);\n"
                 << "    double cp = 0.0; double dp = 0.0;\n"
                 << "    if (live) {\n"
                 << "        " << coef_arrays("f_", R) << coef_arrays("n_", R) << "\n"
                 << "#pragma unroll\n"
                 << "        for (int j = 0; j < " << R << "; ++j) { const int i = j; if (i < n) {" << coef_load("f_", "x", yi, zi)
                 << " } else {" << coef_zero("f_") << " } }\n"
                 << "        for (int i0 = 0; i0 < n; i0 += " << R << ") {\n"
                 << "#pragma unroll\n"
                 << "            for (int j = 0; j < " << R << "; ++j) { const int i = i0 + " << R << " + j; if (i < n) {"
                 << coef_load("n_", "x", yi, zi) << " } else {" << coef_zero("n_") << " } }\n"
                 << "#pragma unroll\n"
                 << "            for (int j = 0; j < " << R << "; ++j) { const int i = i0 + j; if (i < n) { const i64 q = " << at << ";\n"
                 << forward([&](int q) { return "f_" + std::string(names[q]) + "[j]"; }, "q", "                ")
                 << "            } }\n"
                 << "#pragma unroll\n"
                 << "            for (int j = 0; j < " << R << "; ++j) {" << coef_move("f_", "n_") << " }\n"
                 << "        }\n"
                 << "    }\n";
            if (u) {
                out_ << "    __syncthreads();\n";
            }
            out_ << "    if (live) {\n"
                 << "        double next = 0.0;\n"
                 << "        " << sweep_arrays("g_", R) << sweep_arrays("h_", R) << "\n"
                 << "#pragma unroll\n"
                 << "        for (int j = 0; j < " << R << "; ++j) { const int i = n - 1 - j; if (i >= 0) { const i64 q = " << at << "; "
                 << sweep_load("g_", "q") << " } else { " << sweep_zero("g_") << " } }\n"
                 << "        for (int i1 = n - 1; i1 >= 0; i1 -= " << R << ") {\n"
                 << "#pragma unroll\n"
                 << "            for (int j = 0; j < " << R << "; ++j) { const int i = i1 - " << R << " - j; if (i >= 0) { const i64 q = " << at
                 << "; " << sweep_load("h_", "q") << " } else { " << sweep_zero("h_") << " } }\n"
                 << "#pragma unroll\n"
                 << "            for (int j = 0; j < " << R << "; ++j) { const int i = i1 - j; if (i >= 0) { next = g_d[j] - " << cp_of("g_")
                 << " * next; AT(" << slot << ", x, " << yi << ", " << zi << ") = fe_rf(next); } }\n"
                 << "#pragma unroll\n"
                 << "            for (int j = 0; j < " << R << "; ++j) { " << sweep_move("g_", "h_") << " }\n"
                 << "        }\n"
                 << "    }\n}\n";
            return;
        }
        // Lines along x. A thread reads for the staging the cells (x0 + (t & 31), row (t >> 5) + 2 r), r < 32, clamped
        // into the chunk and the block's rows so every read is in bounds (what falls outside is not stored).
        for (int q : cell) {
            out_ << "    __shared__ float t_" << names[q] << "[64][33];\n";
        }
        auto stage_load = [&](const std::string& x0, const std::string& w) {
            return "            { const int l = min((t >> 5) + 2 * j, rows - 1); const int c = " + x0 + " + min(t & 31, " + w + " - 1);" +
                   coef_load("g_", "c", "t_y[l]", "t_z[l]") + " }\n";
        };
        out_ << "    __shared__ float t_o[64][33]; __shared__ int t_y[64]; __shared__ int t_z[64];\n"
             << "    const i64 lines = (i64)k.ny * k.nz; const i64 line0 = (i64)blockIdx.x * 64;\n"
             << "    const i64 line = line0 + t; const bool live = line < lines; const int n = k.nx;\n"
             << "    const int rows = lines - line0 < 64 ? (int)(lines - line0) : 64;\n"
             << "    if (live) { t_y[t] = (int)(line % k.ny); t_z[t] = (int)(line / k.ny); }\n"
             << "    __syncthreads();\n"
             << "    double cp = 0.0; double dp = 0.0;\n";
        if (!cell.empty()) {
            out_ << "    " << coef_arrays("g_", "32") << "\n"
                 << "#pragma unroll\n"
                 << "    for (int j = 0; j < 32; ++j)\n" << stage_load("0", "(n < 32 ? n : 32)");
        }
        out_ << "    for (int x0 = 0; x0 < n; x0 += 32) {\n"
             << "        const int w = n - x0 < 32 ? n - x0 : 32;\n"
             << "        __syncthreads();\n";
        if (!cell.empty()) {
            std::string st;
            for (int q : cell) {
                st += " t_" + std::string(names[q]) + "[l][c] = g_" + names[q] + "[j];";
            }
            out_ << "#pragma unroll\n"
                 << "        for (int j = 0; j < 32; ++j) { const int l = (t >> 5) + 2 * j, c = t & 31; if (l < rows && c < w) {" << st << " } }\n"
                 << "        __syncthreads();\n"
                 << "        if (x0 + 32 < n) {\n"
                 << "            const int w2 = n - x0 - 32 < 32 ? n - x0 - 32 : 32;\n"
                 << "#pragma unroll\n"
                 << "            for (int j = 0; j < 32; ++j)\n" << "    " << stage_load("x0 + 32", "w2")
                 << "        }\n";
        }
        out_ << "        if (live) {\n"
             << "            for (int j = 0; j < w; ++j) {\n"
             << "                const int i = x0 + j; const i64 q = (i64)i * lines + line;\n"
             << forward([&](int q) { return "t_" + std::string(names[q]) + "[t][j]"; }, "q", "                ")
             << "            }\n"
             << "        }\n"
             << "    }\n"
             << "    __syncthreads();\n"
             << "    double next = 0.0;\n"
             << "    const int last = ((n - 1) / 32) * 32;\n"
             << "    " << sweep_arrays("g_", "32") << (u ? sweep_arrays("h_", "32") : std::string()) << "\n"
             << "#pragma unroll\n"
             << "    for (int j = 0; j < 32; ++j) { const int i = last + j; if (live && i < n) { const i64 q = (i64)i * lines + line; "
             << sweep_load("g_", "q") << " } else { " << sweep_zero("g_") << " } }\n"
             << "    for (int x0 = last; x0 >= 0; x0 -= 32) {\n"
             << "        const int w = n - x0 < 32 ? n - x0 : 32;\n";
        const std::string load_prev = "for (int j = 0; j < 32; ++j) { const int i = x0 - 32 + j; if (live && x0 > 0) { const i64 q = (i64)i * lines + line; ";
        if (u) {
            out_ << "#pragma unroll\n"
                 << "        " << load_prev << sweep_load("h_", "q") << " } else { " << sweep_zero("h_") << " } }\n";
        }
        out_ << "        if (live) {\n"
             << "#pragma unroll\n"
             << "            for (int j = 31; j >= 0; --j) { const int i = x0 + j; if (j < w) { next = g_d[j] - " << cp_of("g_")
             << " * next; t_o[t][j] = fe_rf(next); } }\n"
             << "        }\n"
             << "        __syncthreads();\n"
             << "#pragma unroll\n"
             << "        for (int j = 0; j < 32; ++j) { const int l = (t >> 5) + 2 * j, c = t & 31; if (l < rows && c < w) AT(" << slot
             << ", x0 + c, t_y[l], t_z[l]) = t_o[l][c]; }\n"
             << "        __syncthreads();\n";
        if (u) {
            out_ << "#pragma unroll\n"
                 << "        for (int j = 0; j < 32; ++j) { " << sweep_move("g_", "h_") << " }\n";
        } else {
            out_ << "#pragma unroll\n"
                 << "        " << load_prev << sweep_load("g_", "q") << " } else { " << sweep_zero("g_") << " } }\n";
        }
        out_ << "    }\n"
             << "}\n";
    }

    // A `half` net's op in three kernels around the library's MFMA forward (see StageCode::Kind::NetHalf). The inputs,
    // the limiter, the ldiffuse weights and the divergence are the fp32 path's (src/field/Law.cpp:673-738,1298-1443);
    // only the net runs in fp16 with fp32 accumulation.
    void net_half_kernels(const StageCode& s, const ir::Inst& in, int slot)
    {
        const Runtime walk(runtime_);
        const NetDecl& net = law_.nets[static_cast<std::size_t>(in.index)];
        std::vector<std::size_t> offsets;
        (void)net_blob(law_, offsets);
        const std::size_t off = offsets[static_cast<std::size_t>(in.index)];
        const int nin = net.inputs();
        const int rpc = s.rows_per_cell, pad = s.in_pad, outs = s.outs;
        const std::string head = "    const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x; if (t >= count) return;\n"
                                 "    const i64 i = c0 + t; int x, y, z; cell_xyz(k, i, x, y, z);\n    const Direct m(k);\n";
        std::ostringstream p;
        p << "extern \"C\" __global__ void " << s.kernel << "_pack(K k, _Float16* rows, i64 c0, i64 count) {\n" << head
          << "    const float* P = k.N + " << off << ";\n"
          << "    auto put = [&](int r, const float* v) { _Float16* o = rows + (t * " << rpc << " + r) * " << pad << ";\n"
          << "        for (int q = 0; q < " << pad << "; ++q) o[q] = q < " << nin << " ? (_Float16)((v[q] - P[q]) * P[" << nin
          << " + q]) : (_Float16)0.0f; };\n";
        std::ostringstream fk;
        fk << "extern \"C\" __global__ void " << s.kernel << "_face(K k, const float* outs, float* faces, i64 c0, i64 count) {\n"
           << head;
        static const char* ax[3] = {"x", "y", "z"};
        static const char* cnt[3] = {"k.nx", "k.ny", "k.nz"};
        if (in.op == Op::Mlp) {
            p << "    float v[" << nin << "];\n";
            for (std::size_t q = 0; q < in.args.size(); ++q) {
                p << "    v[" << q << "] = (float)" << d(in.args[q]) << ";\n";
            }
            p << "    put(0, v);\n";
            fk << "    AT(" << slot << ", x, y, z) = outs[t * " << outs << "];\n";
        } else if (in.op == Op::Flux) {
            const std::size_t mm = in.args.size() - 2;
            const double lo = at(in.args[mm]).number, hi = at(in.args[mm + 1]).number;
            p << "    float xij[" << 2 * mm << "], xji[" << 2 * mm << "];\n";
            fk << "    const float per = k.dt > 0.0 ? (float)(0.999 * k.h / (6.0 * k.dt)) : 0.0f;\n";
            for (int a = 0; a < 3; ++a) {
                const std::string j[3] = {a == 0 ? "x + 1" : "x", a == 1 ? "y + 1" : "y", a == 2 ? "z + 1" : "z"};
                p << "    if (" << ax[a] << " + 1 < " << cnt[a] << ") {\n";
                for (std::size_t q = 0; q < mm; ++q) {
                    p << "        xij[" << q << "] = (float)" << d(in.args[q]) << "; xij[" << mm + q << "] = (float)"
                      << d(in.args[q], j[0], j[1], j[2]) << ";\n"
                      << "        xji[" << q << "] = xij[" << mm + q << "]; xji[" << mm + q << "] = xij[" << q << "];\n";
                }
                p << "    } else {\n        for (int q = 0; q < " << 2 * mm << "; ++q) xij[q] = xji[q] = 0.0f;\n    }\n"
                  << "    put(" << 2 * a << ", xij); put(" << 2 * a + 1 << ", xji);\n";
                fk << "    if (" << ax[a] << " + 1 < " << cnt[a] << ") {\n"
                   << "        const float gij = outs[(t * 6 + " << 2 * a << ") * " << outs << "], gji = outs[(t * 6 + " << 2 * a + 1 << ") * "
                   << outs << "];\n"
                   << "        float r = 0.5f * (gij - gji);\n"
                   << "        const float fi = (float)" << d(in.args[0]) << ", fj = (float)" << d(in.args[0], j[0], j[1], j[2]) << ";\n"
                   << "        if (k.dt > 0.0) {\n            if (r > 0.0f) {\n";
                if (lo > -1e29) {
                    fk << "                r = fe_minf(r, fe_maxf(0.0f, (fi - (float)" << literal(lo) << ") * per));\n";
                }
                if (hi < 1e29) {
                    fk << "                r = fe_minf(r, fe_maxf(0.0f, ((float)" << literal(hi) << " - fj) * per));\n";
                }
                fk << "            } else if (r < 0.0f) {\n";
                if (lo > -1e29) {
                    fk << "                r = fe_maxf(r, -fe_maxf(0.0f, (fj - (float)" << literal(lo) << ") * per));\n";
                }
                if (hi < 1e29) {
                    fk << "                r = fe_maxf(r, -fe_maxf(0.0f, ((float)" << literal(hi) << " - fi) * per));\n";
                }
                fk << "            }\n        }\n        faces[3 * i + " << a << "] = r;\n    } else {\n        faces[3 * i + " << a
                   << "] = 0.0f;\n    }\n";
            }
        } else {
            // ldiffuse(NET, k, F): the face's window F[i-1] F[i] | F[j] F[j+1] (ends repeated), both ways round
            const int ka = in.args[0], fa = in.args[1];
            auto window = [&](std::ostringstream& o, int a) {
                const int sx = a == 0, sy = a == 1, sz = a == 2;
                const std::string c = ax[a];
                o << "        const int xj = x + " << sx << ", yj = y + " << sy << ", zj = z + " << sz << ";\n"
                  << "        const int xm = " << c << " > 0 ? x - " << sx << " : x, ym = " << c << " > 0 ? y - " << sy << " : y, zm = " << c
                  << " > 0 ? z - " << sz << " : z;\n"
                  << "        const int xp = " << c << " + 2 < " << cnt[a] << " ? xj + " << sx << " : xj, yp = " << c << " + 2 < " << cnt[a]
                  << " ? yj + " << sy << " : yj, zp = " << c << " + 2 < " << cnt[a] << " ? zj + " << sz << " : zj;\n"
                  << "        const float w[4] = {(float)" << d(fa, "xm", "ym", "zm") << ", (float)" << d(fa) << ", (float)"
                  << d(fa, "xj", "yj", "zj") << ", (float)" << d(fa, "xp", "yp", "zp") << "};\n"
                  << "        const double ka = " << d(ka) << ", kb = " << d(ka, "xj", "yj", "zj") << ";\n"
                  << "        const float kf = (float)((ka + kb) > 0.0 ? 2.0 * ka * kb / (ka + kb) : 0.0);\n"
                  << "        const float mirror[4] = {w[3], w[2], w[1], w[0]};\n";
            };
            p << "    auto inputs = [&](const float* w, float kf, float* v) {\n"
              << "        const float centre = 0.5f * (w[1] + w[2]);\n        float spread = 1e-6f;\n"
              << "        for (int a = 0; a < 4; ++a) spread = fe_maxf(spread, fabsf(w[a] - centre));\n"
              << "        for (int a = 0; a < 4; ++a) v[a] = (w[a] - centre) / spread;\n"
              << "        const float dtf = (float)k.dt, hf = (float)k.h;\n        v[4] = kf * dtf / (hf * hf);\n    };\n"
              << "    float v1[5], v2[5];\n";
            fk << "    auto raw = [&](const float* w, float kf, const float* y) -> float {\n"
               << "        const float hf = (float)k.h;\n"
               << "        const float mean = 0.25f * (y[0] + y[1] + y[2] + y[3]);\n"
               << "        const float base[4] = {0.0f, -1.0f, 1.0f, 0.0f};\n        float acc = 0.0f;\n"
               << "        for (int a = 0; a < 4; ++a) acc += (base[a] + (y[a] - mean)) * w[a];\n"
               << "        return -kf / hf * acc;\n    };\n";
            for (int a = 0; a < 3; ++a) {
                p << "    if (" << ax[a] << " + 1 < " << cnt[a] << ") {\n";
                window(p, a);
                p << "        inputs(w, kf, v1); inputs(mirror, kf, v2);\n"
                  << "    } else {\n        for (int q = 0; q < 5; ++q) v1[q] = v2[q] = 0.0f;\n    }\n"
                  << "    put(" << 2 * a << ", v1); put(" << 2 * a + 1 << ", v2);\n";
                fk << "    if (" << ax[a] << " + 1 < " << cnt[a] << ") {\n";
                window(fk, a);
                fk << "        const float ij = raw(w, kf, outs + (t * 6 + " << 2 * a << ") * " << outs << "), ji = raw(mirror, kf, outs + (t * 6 + "
                   << 2 * a + 1 << ") * " << outs << ");\n"
                   << "        faces[3 * i + " << a << "] = 0.5f * (ij - ji);\n    } else {\n        faces[3 * i + " << a << "] = 0.0f;\n    }\n";
            }
        }
        p << "}\n";
        fk << "}\n";
        out_ << p.str() << fk.str();
        if (in.op != Op::Mlp) {
            // the cell: in through its -x, -y, -z faces, out through +x, +y, +z, in the CPU's order, over h
            out_ << "extern \"C\" __global__ void " << s.kernel << "(K k, const float* faces) { CELL_HEAD\n"
                 << "    const i64 i = IDX(x, y, z);\n    float out = 0.0f;\n"
                 << "    if (x > 0) out += faces[3 * (i - 1) + 0];\n"
                 << "    if (x + 1 < k.nx) out -= faces[3 * i + 0];\n"
                 << "    if (y > 0) out += faces[3 * (i - k.nx) + 1];\n"
                 << "    if (y + 1 < k.ny) out -= faces[3 * i + 1];\n"
                 << "    if (z > 0) out += faces[3 * (i - k.nxy) + 2];\n"
                 << "    if (z + 1 < k.nz) out -= faces[3 * i + 2];\n"
                 << "    AT(" << slot << ", x, y, z) = out / (float)k.h;\n}\n";
        }
    }

    // distance(a) (src/field/Law.cpp:1042-1096): the set a >= 0.5, the squared distances (in cells) from every cell to
    // the nearest cell inside and outside it, and from those the signed distance in metres, the surface half a cell
    // from the last solid cell's centre and the lattice's outside counted as outside. The squared distances are
    // whole numbers, exact in float, so the exact minimum is the CPU's value whatever the method: here each cell scans
    // outward along one axis at a time until no farther cell can beat what it holds. The last step is the CPU's
    // double arithmetic.
    void distance_kernels(const StageCode& s, const ir::Inst& in, int slot)
    {
        const std::string set = std::to_string(slot + 1);
        out_ << "extern \"C\" __global__ void " << s.kernel << "_mark(K k, unsigned* w) { CELL_HEAD const Direct m(k);\n"
             << "    const float in = " << d(in.args[0]) << " >= 0.5 ? 1.0f : 0.0f;\n"
             << "    float& held = AT(" << set << ", x, y, z);\n"
             << "    const bool changed = __float_as_uint(held) != __float_as_uint(in) || w[3] == 0u;\n"
             << "    held = in;\n"
             << "    const u64 bc = __ballot(changed), bi = __ballot(in != 0.0f), bo = __ballot(in == 0.0f);\n"
             << "    const int lane = (int)__lane_id();\n"
             << "    if (changed && lane == __ffsll((long long)bc) - 1) w[0] = 1u;\n"
             << "    if (in != 0.0f && lane == __ffsll((long long)bi) - 1) w[1] = 1u;\n"
             << "    if (in == 0.0f && lane == __ffsll((long long)bo) - 1) w[2] = 1u;\n"
             << "}\n";
        // One axis for both sets; from_in < 0: the first axis, seeded from the set (0 on it, inf off it).
        out_ << "extern \"C\" __global__ void " << s.kernel
             << "_scan(K k, const unsigned* w, int axis, int from_in, int from_out, int to_in, int to_out) {\n"
             << "    if (w[0] == 0u) return;\n    CELL_HEAD\n"
             << "    const int n = axis == 0 ? k.nx : (axis == 1 ? k.ny : k.nz);\n"
             << "    const int q = axis == 0 ? x : (axis == 1 ? y : z);\n"
             << "    const i64 step = axis == 0 ? 1 : (axis == 1 ? (i64)k.nx : k.nxy);\n"
             << "    const i64 here = IDX(x, y, z);\n"
             << "    const float inf = __builtin_inff();\n"
             << "    const float* fi = k.A + (i64)(from_in < 0 ? " << set << " : from_in) * k.cells;\n"
             << "    const float* fo = k.A + (i64)(from_in < 0 ? " << set << " : from_out) * k.cells;\n"
             << "    float bi = inf, bo = inf;\n"
             << "    for (int r = 0; r < n; ++r) {\n"
             << "        const float rr = (float)(r * r);\n"
             << "        if (rr >= bi && rr >= bo) break;\n"
             << "        const int lo = q - r, hi = q + r;\n"
             << "        if (lo < 0 && hi >= n) break;\n"
             << "        if (lo >= 0) {\n"
             << "            const i64 j = here - (i64)r * step;\n"
             << "            const float a = from_in < 0 ? (fi[j] != 0.0f ? 0.0f : inf) : fi[j];\n"
             << "            const float b = from_in < 0 ? (fo[j] == 0.0f ? 0.0f : inf) : fo[j];\n"
             << "            bi = fminf(bi, a + rr); bo = fminf(bo, b + rr);\n"
             << "        }\n"
             << "        if (hi < n) {\n"
             << "            const i64 j = here + (i64)r * step;\n"
             << "            const float a = from_in < 0 ? (fi[j] != 0.0f ? 0.0f : inf) : fi[j];\n"
             << "            const float b = from_in < 0 ? (fo[j] == 0.0f ? 0.0f : inf) : fo[j];\n"
             << "            bi = fminf(bi, a + rr); bo = fminf(bo, b + rr);\n"
             << "        }\n"
             << "    }\n"
             << "    k.A[(i64)to_in * k.cells + here] = bi; k.A[(i64)to_out * k.cells + here] = bo;\n"
             << "}\n";
        // The same axis a line a thread: the lower envelope of the parabolas f(j) + (q - j)^2 (Felzenszwalb and
        // Huttenlocher's 1-D transform) built once along the line, then read at every q, in O(n) a line instead of an
        // outward search from every cell (which walks whole lines where a set is far or absent). The comparisons are
        // exact: the f(j) are whole numbers below 2^24 and the intersections are compared cross-multiplied in double
        // (FeEnv), so each q gets the exact least value, the outward search's to the bit. The stacks
        // (indices and f) of the block's lines sit in LDS (the launch gives 12 n bytes a line, n the line's length).
        out_ << "extern \"C\" __global__ void " << s.kernel
             << "_lines(K k, const unsigned* w, int axis, int from_in, int from_out, int to_in, int to_out) {\n"
             << "    if (w[0] == 0u) return;\n"
             << "    extern __shared__ __attribute__((aligned(16))) unsigned char env_lds[];\n"
             // a block: L lines, both sets (the squared distances to the set's cells and to the others), a thread each
             << "    const int L = (int)blockDim.x / 2, t = (int)threadIdx.x, out = t >= L ? 1 : 0, lt = t - out * L;\n"
             << "    const int n = axis == 0 ? k.nx : (axis == 1 ? k.ny : k.nz);\n"
             << "    const i64 lines = k.cells / n;\n"
             << "    const i64 line = (i64)blockIdx.x * L + lt;\n"
             << "    if (line >= lines) return;\n"
             << "    int x = 0, y = 0, z = 0;\n"
             << "    if (axis == 0) { y = (int)(line % k.ny); z = (int)(line / k.ny); }\n"
             << "    else if (axis == 1) { x = (int)(line % k.nx); z = (int)(line / k.nx); }\n"
             << "    else { x = (int)(line % k.nx); y = (int)(line / k.nx); }\n"
             << "    const i64 base = IDX(x, y, z), step = axis == 0 ? 1 : (axis == 1 ? (i64)k.nx : k.nxy);\n"
             << "    FeEnv e;\n"
             << "    e.st = reinterpret_cast<unsigned short*>(env_lds) + (i64)t * n;\n"
             << "    e.fs = reinterpret_cast<float*>(env_lds + (i64)4 * L * n) + (i64)t * n;\n"
             << "    const float inf = __builtin_inff();\n"
             << "    const int from = from_in < 0 ? " << set << " : (out ? from_out : from_in);\n"
             << "    const float* f = k.A + (i64)from * k.cells + base;\n"
             // sixteen cells' loads in flight before their pushes (a push waits on the one before it)
             << "    for (int q0 = 0; q0 < n; q0 += 16) {\n"
             << "        float av[16];\n"
             << "#pragma unroll\n"
             << "        for (int u = 0; u < 16; ++u) { const i64 q = q0 + u < n ? q0 + u : n - 1; av[u] = f[q * step]; }\n"
             << "#pragma unroll\n"
             << "        for (int u = 0; u < 16; ++u) {\n"
             << "            const int q = q0 + u;\n"
             << "            if (q < n) e.push(from_in < 0 ? ((av[u] != 0.0f) != (out != 0) ? 0.0f : inf) : av[u], q);\n"
             << "        }\n"
             << "    }\n"
             << "    float* o = k.A + (i64)(out ? to_out : to_in) * k.cells + base;\n"
             << "    e.start();\n"
             << "    for (int q = 0; q < n; ++q) o[(i64)q * step] = e.at(q);\n"
             << "}\n";
        out_ << "extern \"C\" __global__ void " << s.kernel << "_signed(K k, unsigned* w) {\n"
             << "    if (w[0] == 0u) return;\n    CELL_HEAD\n"
             << "    const bool any_in = w[1] != 0u, any_out = w[2] != 0u;\n"
             << "    const double far = (double)(k.nx + k.ny + k.nz) * k.h;\n"
             << "    double dd;\n"
             << "    if (AT(" << set << ", x, y, z) != 0.0f) {\n"
             << "        double border = __builtin_inf();\n"
             << "        if (k.nx > 1) border = fe_min(fe_min(border, (double)(x + 1)), (double)(k.nx - x));\n"
             << "        if (k.ny > 1) border = fe_min(fe_min(border, (double)(y + 1)), (double)(k.ny - y));\n"
             << "        if (k.nz > 1) border = fe_min(fe_min(border, (double)(z + 1)), (double)(k.nz - z));\n"
             << "        const double cells = fe_min(any_out ? sqrt((double)AT(" << slot + 3 << ", x, y, z)) : far / k.h, border);\n"
             << "        dd = -(cells - 0.5) * k.h;\n"
             << "    } else {\n"
             << "        dd = any_in ? (sqrt((double)AT(" << slot + 2 << ", x, y, z)) - 0.5) * k.h : far;\n"
             << "    }\n"
             << "    AT(" << slot << ", x, y, z) = fe_rf(dd);\n"
             << "    if (x == 0 && y == 0 && z == 0) w[3] = 1u;\n"
             << "}\n";
    }

    // poisson(s[, n]) (src/field/Law.cpp:1139-1181): n Jacobi sweeps of lap(phi) = s with phi = 0 outside the lattice,
    // warm-started from the last answer (zero at first), each sweep's sum in double in the CPU's order; skipped when
    // the source holds the bits it held at the last solve.
    void poisson_kernels(const StageCode& s, const ir::Inst& in, int slot)
    {
        const int src = in.args[0];
        if (s.scalar_source) {
            out_ << "extern \"C\" __global__ void " << s.kernel << "_mark(K k, unsigned* w) {\n"
                 << "    const Direct m(k);\n"
                 << "    const u64 b = (u64)__double_as_longlong(" << d(src) << ");\n"
                 << "    const bool changed = w[3] == 0u || w[1] != (unsigned)b || w[2] != (unsigned)(b >> 32);\n"
                 << "    w[1] = (unsigned)b; w[2] = (unsigned)(b >> 32);\n"
                 << "    if (changed) w[0] = 1u;\n"
                 << "}\n";
        } else {
            out_ << "extern \"C\" __global__ void " << s.kernel << "_mark(K k, unsigned* w) { CELL_HEAD const Direct m(k);\n"
                 << "    const float v = " << v(src) << ";\n"
                 << "    float& held = AT(" << slot + 2 << ", x, y, z);\n"
                 << "    const bool changed = __float_as_uint(held) != __float_as_uint(v) || w[3] == 0u;\n"
                 << "    held = v;\n"
                 << "    const u64 bc = __ballot(changed);\n"
                 << "    if (changed && (int)__lane_id() == __ffsll((long long)bc) - 1) w[0] = 1u;\n"
                 << "}\n";
        }
        out_ << "extern \"C\" __global__ void " << s.kernel << "_sweep(K k, const unsigned* w, int from, int to) {\n"
             << "    if (w[0] == 0u) return;\n    CELL_HEAD\n"
             << (s.scalar_source ? "    const Direct m(k);\n" : "")
             << "    const float* phi = k.A + (i64)from * k.cells;\n"
             << "    const i64 i = IDX(x, y, z);\n"
             << "    double sum = 0.0;\n"
             << "    if (k.nx > 1) { if (x > 0) sum += phi[i - 1]; if (x + 1 < k.nx) sum += phi[i + 1]; }\n"
             << "    if (k.ny > 1) { if (y > 0) sum += phi[i - k.nx]; if (y + 1 < k.ny) sum += phi[i + k.nx]; }\n"
             << "    if (k.nz > 1) { if (z > 0) sum += phi[i - k.nxy]; if (z + 1 < k.nz) sum += phi[i + k.nxy]; }\n"
             << "    const int axes = (k.nx > 1 ? 1 : 0) + (k.ny > 1 ? 1 : 0) + (k.nz > 1 ? 1 : 0);\n"
             << "    const double denom = 2.0 * (double)(axes > 1 ? axes : 1);\n"
             << "    const double s = " << (s.scalar_source ? d(src) : "(double)AT(" + std::to_string(slot + 2) + ", x, y, z)") << ";\n"
             << "    k.A[(i64)to * k.cells + i] = fe_rf((sum - k.h * k.h * s) / denom);\n"
             << "}\n";
        out_ << "extern \"C\" __global__ void " << s.kernel << "_copy(K k, unsigned* w, int from) {\n"
             << "    if (w[0] == 0u) return;\n    CELL_HEAD\n"
             << "    if (from != " << slot << ") AT(" << slot << ", x, y, z) = k.A[(i64)from * k.cells + IDX(x, y, z)];\n"
             << "    if (x == 0 && y == 0 && z == 0) w[3] = 1u;\n"
             << "}\n";
    }

    ProgramCode program(const ir::Program& p, const std::string& tag, bool derive)
    {
        ProgramCode code;
        code.stages = p.stages;
        code.limit_stages = p.limit_stages;
        code.max_dt = p.max_dt;
        if (p.equations.empty()) {
            return code;
        }
        code.gather = "g_" + tag;
        std::ostringstream g;
        if (derive) {
            // out = float((accumulate ? double(out) : 0.0) + value) (src/field/Solver.cpp:252-268).
            const ir::Equation& eq = p.equations[0];
            const bool reads_itself = std::find(p.reads.begin(), p.reads.end(), eq.target) != p.reads.end();
            code.in_place = !reads_itself;
            const std::string t = std::to_string(eq.target);
            g << "extern \"C\" __global__ void " << code.gather << "(K k) { CELL_HEAD\n    const Direct m(k);\n"
              << "    const i64 i = IDX(x, y, z);\n"
              << "    const double val = " << d(eq.value) << ";\n    const float old = k.now[" << t << "][i];\n"
              << "    const float nv = fe_rf(" << (p.accumulate ? "(double)old" : "0.0") << " + val);\n"
              << "    const bool ch = __float_as_uint(old) != __float_as_uint(nv);\n";
            if (code.in_place) {
                g << "    if (ch) k.now[" << t << "][i] = nv;\n";
            } else {
                g << "    k.nxt[" << t << "][i] = nv;\n";
                code.swaps.push_back(eq.target);
            }
            g << "    mark(k, " << t << ", ch);\n}\n";
            out_ << g.str();
            return code;
        }
        code.in_place = p.in_place;
        // The cell function: every rate first, from the pre-state; then each target's deltas in equation order. o[q]
        // and n[q] are target q's old and new value.
        std::vector<int> targets;
        for (const ir::Equation& eq : p.equations) {
            if (std::find(targets.begin(), targets.end(), eq.target) == targets.end()) {
                targets.push_back(eq.target);
            }
        }
        const std::string cellfn = "c_" + code.gather;
        std::ostringstream c;
        c << "template <class M> __device__ __forceinline__ void " << cellfn
          << "(const K& k, const M& m, int x, int y, int z, float* o, float* n)\n{\n";
        for (std::size_t e = 0; e < p.equations.size(); ++e) {
            c << "    const double r" << e << " = " << d(p.equations[e].value) << ";\n";
        }
        for (std::size_t q = 0; q < targets.size(); ++q) {
            const int target = targets[q];
            c << "    o[" << q << "] = m.template fld<" << target << ">(k, x, y, z); n[" << q << "] = o[" << q << "];\n";
            for (std::size_t e = 0; e < p.equations.size(); ++e) {
                if (p.equations[e].target != target) {
                    continue;
                }
                const std::string upd = "n[" + std::to_string(q) + "] = fe_ftz(n[" + std::to_string(q) + "] + fe_rf(r" + std::to_string(e) +
                    " * k.dt));\n";
                if (scalar(p.equations[e].value)) {
                    // A rate that is one number and 0 is idle on the CPU: nothing is added.
                    c << "    if (r" << e << " != 0.0) " << upd;
                } else {
                    c << "    " << upd;
                }
            }
        }
        c << "}\n";
        out_ << c.str();
        if (!code.in_place) {
            code.swaps = targets;
        }
        const int nt = static_cast<int>(targets.size());
        // The direct kernel: a thread a cell; `<gather>_32` the same with 32-bit reads (lattices under 2^31 cells).
        g << "template <class M> __device__ __forceinline__ void d_" << code.gather << "(const K& k, int x, int y, int z)\n{\n"
          << "    const M m(k);\n    float o[" << nt << "], n[" << nt << "];\n    " << cellfn << "(k, m, x, y, z, o, n);\n"
          << "    const i64 i = IDX(x, y, z);\n";
        for (int q = 0; q < nt; ++q) {
            const std::string t = std::to_string(targets[static_cast<std::size_t>(q)]);
            const std::string qs = std::to_string(q);
            g << "    { const bool ch = __float_as_uint(o[" << qs << "]) != __float_as_uint(n[" << qs << "]);\n";
            if (code.in_place) {
                g << "      if (ch) k.now[" << t << "][i] = n[" << qs << "];\n";
            } else {
                g << "      k.nxt[" << t << "][i] = n[" << qs << "];\n";
            }
            g << "      mark(k, " << t << ", ch); }\n";
        }
        g << "}\n";
        for (const bool w32 : {false, true}) {
            const std::string w = w32 ? "true" : "false";
            g << "extern \"C\" __global__ void " << code.gather << (w32 ? "_32" : "") << "(K k) { CELL_HEAD\n"
              << "    if (FE_POW2(k)) d_" << code.gather << "<DirectT<" << w << ", true>>(k, x, y, z);\n"
              << "    else d_" << code.gather << "<DirectT<" << w << ", false>>(k, x, y, z);\n}\n";
        }
        out_ << g.str();
        std::map<int, int> reach;
        int cost = 0;
        const bool bounded = field_reach(p, reach, cost);
        int radius = 0;
        std::vector<int> windowed;
        for (const auto& [f, r] : reach) {
            if (r > 0) {
                windowed.push_back(f);
                radius = std::max(radius, r);
            }
        }
        if (bounded && radius >= 1 && radius <= 2 && cost <= 4000) {
            window_kernels(code, cellfn, targets, windowed, radius);
            const int kk = conduction_form(p);
            if (kk >= 0 && code.window) {
                flux_kernels(code, cellfn, p.equations[0].target, kk);
            }
        }
        return code;
    }

    // A term that is one conduction law, d F = diffuse(k, F) with k one number: its scalar k's instruction, else -1.
    // Such a term also gets the flux-carrying window (flux_kernels).
    int conduction_form(const ir::Program& p) const
    {
        if (p.in_place || p.equations.size() != 1) {
            return -1;
        }
        const ir::Equation& eq = p.equations[0];
        const ir::Inst& in = at(eq.value);
        if (in.op != Op::Diffuse || lay_.stages.count(eq.value) != 0 || !scalar(in.args[0])) {
            return -1;
        }
        const ir::Inst& fa = at(in.args[1]);
        if (fa.op != Op::Field || fa.index != eq.target || lay_.stages.count(in.args[1]) != 0) {
            return -1;
        }
        return in.args[0];
    }

    // The window of a conduction term, carrying its face fluxes. Tiles of 256 x TY cells march in z, a lane four
    // consecutive cells of a row (float4 loads and stores; nx a multiple of 4); plane z's x and y neighbours from a
    // double-buffered LDS tile (one barrier a plane), its z neighbours from a four-slot register ring loaded three
    // planes ahead. Two ways to bring the tile's halo rows: form A (`_f4_`), two more waves that only load them (the
    // compute waves' loads stay one row each: the better form once the lattice streams from HBM); form B (`_h4_`),
    // the tile's first and last row waves load them beside their own row, the others their own row twice (no idle
    // waves, so more compute waves a CU: the better form while the lattice sits in the last-level cache; measured on
    // MI355X at 256^3: 27.2 us against 30.6, at 512^3: 265 against 243).
    // The CPU's face sum, in its order (-x, +x, -y, +y, -z, +z; src/field/Law.cpp:986-1018), from fluxes computed once
    // where IEEE arithmetic makes the share exact: kf * (a - b) is -(kf * (b - a)) to the bit, so the flux a cell sends
    // through its +x face is the one its right neighbour receives, and the flux through its +z face is carried into
    // the next plane as that plane's -z flux, with the plane's values widened to double once. A face past the
    // lattice's edge reads the cell itself: kf * (c - c) = +0 for finite c and kf, and adding +0 to a sum that starts
    // at +0 changes no bit. A NaN sum (an infinity or NaN among the inputs) is summed again at once from the same face
    // fluxes, the missing faces left out, as the CPU sums them.
    // TOTAL (a step that checks the term's contracts): the tile also totals the new values and their magnitudes in a
    // fixed order (each lane its cells in order in double, at most 4 x 64 of them; the lanes' sums folded as compensated
    // pairs by a butterfly, the waves in order) for the "after" side of the conservation contract, tsum[2 tile..], and
    // counts the cells outside the field's range (rec, first): the separate totals pass is not needed.
    void flux_kernels(ProgramCode& code, const std::string& cellfn, int target, int kk)
    {
        const std::vector<int> targets{target};
        const std::string t = std::to_string(target);
        struct Form {
            bool rows; // form B: halo rows loaded by the edge row waves
            int ty;
            int zchunk;
        };
        for (const Form form : {Form{false, 4, 32}, Form{false, 6, 32}, Form{true, 8, 16}}) {
            const int TY = form.ty;
            const int WAVES = form.rows ? TY : TY + 2;
            const std::string name = code.gather + (form.rows ? "_h4_" : "_f4_") + std::to_string(TY);
            const std::string LDS = "L_" + name;
            const std::string MK = "M_" + name;
            std::ostringstream s;
            s << "template <bool P2> struct " << MK << " { static constexpr bool kPow2 = P2; };\n"
              << "__shared__ __attribute__((aligned(16))) float " << LDS << "[" << 2 * (TY + 2) * 264 << "];\n"
              << "template <bool P2, bool XH, bool TOTAL> __device__ __forceinline__ void run_" << name
              << "(const K& k, int zchunk, int tiles_x, int tiles_y, int tiles_z, const bool TRACK, const bool REMAP, const bool STREAM,\n"
              << "    double* tsum, unsigned* rec, u64* first, double lo_r, double hi_r) {\n"
              << "    const int tl = REMAP ? xcd_tile((int)blockIdx.x, tiles_x * tiles_y * tiles_z) : (int)blockIdx.x;\n"
              << "    const int lane = (int)threadIdx.x; const int r = (int)threadIdx.y;\n"
              << "    const int x = (tl % tiles_x) * 256 + 4 * lane;\n"
              << "    const int y0 = ((tl / tiles_x) % tiles_y) * " << TY << ";\n";
            if (form.rows) {
                s << "    const int y = y0 + r; const int slot = r + 1;\n"
                  << "    const bool live = x < k.nx && y < k.ny;\n";
            } else {
                s << "    const int y = y0 + r - 1; const int slot = r;\n"
                  << "    const bool live = x < k.nx && r >= 1 && r <= " << TY << " && y < k.ny;\n";
            }
            s << "    const int z0 = (tl / (tiles_x * tiles_y)) * zchunk; const int z1 = z0 + zchunk < k.nz ? z0 + zchunk : k.nz;\n"
              << "    const int xs = x < k.nx - 4 ? x : k.nx - 4;\n"
              << "    const int ys = clampi(y, 0, k.ny - 1);\n"
              << "    const float* base = k.now[" << t << "] + (i64)ys * k.nx + xs;\n"
              << "    float* outp = k.nxt[" << t << "] + (i64)ys * k.nx + xs;\n";
            if (form.rows) {
                // the row above the tile (wave 0) or below it (the last wave); the others their own row again
                s << "    const int hy = r == 0 ? clampi(y0 - 1, 0, k.ny - 1) : (r == " << TY - 1 << " ? clampi(y0 + " << TY
                  << ", 0, k.ny - 1) : ys);\n"
                  << "    const int hslot = r == 0 ? 0 : (r == " << TY - 1 << " ? " << TY + 1 << " : -1);\n"
                  << "    const float* hbase = k.now[" << t << "] + (i64)hy * k.nx + xs;\n";
            }
            s << "    const int zlim = z1 < k.nz - 1 ? z1 : k.nz - 1;\n"
              << "    const int lofs = XH && lane == 0 && xs > 0 ? -1 : 0;\n"
              << "    const int rofs = XH && lane == 63 && xs + 4 < k.nx ? 4 : 0;\n"
              << "    const Direct md(k);\n"
              << "    const double kv = s" << kk << "(k, md);\n"
              << "    const double kf = (kv + kv) > 0.0 ? 2.0 * kv * kv / (kv + kv) : 0.0;\n"
              << "    float* const tile = " << LDS << ";\n";
            // Every load unconditional, from a clamped plane (past the top, the plane itself: the zero-flux edge).
            if (form.rows) {
                s << "    struct Pl { float4 v; float4 h; float l; float rr; float hl; float hr; };\n"
                  << "    auto load = [&](Pl& p, int zp) { const i64 o = (i64)(zp < zlim ? zp : zlim) * k.nxy;\n"
                  << "        const float* q = base + o; const float* qh = hbase + o;\n"
                  << "        p.v = *reinterpret_cast<const float4*>(q); p.h = *reinterpret_cast<const float4*>(qh);\n"
                  << "        if (XH) { p.l = q[lofs]; p.rr = q[rofs]; p.hl = qh[lofs]; p.hr = qh[rofs]; } };\n";
            } else {
                s << "    struct Pl { float4 v; float l; float rr; };\n"
                  << "    auto load = [&](Pl& p, int zp) { const float* q = base + (i64)(zp < zlim ? zp : zlim) * k.nxy;\n"
                  << "        p.v = *reinterpret_cast<const float4*>(q); if (XH) { p.l = q[lofs]; p.rr = q[rofs]; } };\n";
            }
            s << "    Pl A, B, C, D;\n"
              << "    load(A, z0); __builtin_amdgcn_sched_barrier(0); load(B, z0 + 1); __builtin_amdgcn_sched_barrier(0);\n"
              << "    load(C, z0 + 2); __builtin_amdgcn_sched_barrier(0);\n"
              << "    const float4 lo = *reinterpret_cast<const float4*>(base + (i64)(z0 > 0 ? z0 - 1 : 0) * k.nxy);\n"
              << "    double c0 = (double)A.v.x, c1 = (double)A.v.y, c2 = (double)A.v.z, c3 = (double)A.v.w;\n"
              << "    double fzm0 = kf * ((double)lo.x - c0), fzm1 = kf * ((double)lo.y - c1), fzm2 = kf * ((double)lo.z - c2),\n"
              << "           fzm3 = kf * ((double)lo.w - c3);\n"
              << "    int buf = 0;\n"
              << "    bool seen[1] = {false};\n"
              << "    double ts = 0.0, tm = 0.0; unsigned nout = 0, nbad = 0; u64 firstbad = ~0ull;\n"
              << "    auto step = [&](const int z, const Pl& cur, const Pl& nxt, Pl& far) {\n"
              << "        load(far, z + 3);\n"
              << "        const float4 c = cur.v;\n"
              << "        float* const plane = tile + buf * " << (TY + 2) * 264 << ";\n"
              << "        float* const row = plane + slot * 264;\n"
              << "        *reinterpret_cast<float4*>(row + 4 + 4 * lane) = c;\n"
              << "        if (XH) { if (lane == 0) row[3] = cur.l; if (lane == 63) row[260] = cur.rr; }\n";
            if (form.rows) {
                s << "        if (hslot >= 0) { float* const hrow = plane + hslot * 264; *reinterpret_cast<float4*>(hrow + 4 + 4 * lane) = cur.h;\n"
                  << "            if (XH) { if (lane == 0) hrow[3] = cur.hl; if (lane == 63) hrow[260] = cur.hr; } }\n";
            }
            s << "        __syncthreads();\n"
              << "        const float4 hi = nxt.v;\n"
              << "        const double d0 = (double)hi.x, d1 = (double)hi.y, d2 = (double)hi.z, d3 = (double)hi.w;\n"
              << "        if (live) {\n"
              // values first, then the edge selects (a select between an LDS lvalue and a register takes addresses)
              << "            const float4 ym = *reinterpret_cast<const float4*>(row - 264 + 4 + 4 * lane);\n"
              << "            const float4 yp = *reinterpret_cast<const float4*>(row + 264 + 4 + 4 * lane);\n"
              << "            const float txm = row[3 + 4 * lane]; const float txp = row[8 + 4 * lane];\n"
              << "            const float xme = x > 0 ? txm : c.x; const float xpe = x + 4 < k.nx ? txp : c.w;\n"
              << "            const double fxm = kf * ((double)xme - c0), f01 = kf * (c1 - c0), f12 = kf * (c2 - c1), f23 = kf * (c3 - c2),\n"
              << "                         fxp = kf * ((double)xpe - c3);\n"
              << "            const double fzp0 = kf * (d0 - c0), fzp1 = kf * (d1 - c1), fzp2 = kf * (d2 - c2), fzp3 = kf * (d3 - c3);\n"
              << "            double q0 = 0.0 + fxm; q0 += f01; q0 += kf * ((double)ym.x - c0); q0 += kf * ((double)yp.x - c0); q0 += fzm0; q0 += fzp0;\n"
              << "            double q1 = 0.0 + (-f01); q1 += f12; q1 += kf * ((double)ym.y - c1); q1 += kf * ((double)yp.y - c1); q1 += fzm1; q1 += fzp1;\n"
              << "            double q2 = 0.0 + (-f12); q2 += f23; q2 += kf * ((double)ym.z - c2); q2 += kf * ((double)yp.z - c2); q2 += fzm2; q2 += fzp2;\n"
              << "            double q3 = 0.0 + (-f23); q3 += fxp; q3 += kf * ((double)ym.w - c3); q3 += kf * ((double)yp.w - c3); q3 += fzm3; q3 += fzp3;\n"
              << "            if (q0 != q0 || q1 != q1 || q2 != q2 || q3 != q3) {\n"
              // An infinity or NaN among the inputs: the faces the cells have, in the CPU's order, from the plane still
              // in LDS (nothing extra kept in registers for this rare path but the incoming -z fluxes).
              << "                const bool xl = x > 0, xr = x + 4 < k.nx, yl = y > 0, yr = y + 1 < k.ny, zl = z > 0, zr = z + 1 < k.nz;\n"
              << "                const float4 ya = *reinterpret_cast<const float4*>(row - 264 + 4 + 4 * lane);\n"
              << "                const float4 yb = *reinterpret_cast<const float4*>(row + 264 + 4 + 4 * lane);\n"
              << "                const double gm = kf * ((double)row[3 + 4 * lane] - c0), gp = kf * ((double)row[8 + 4 * lane] - c3);\n"
              << "                const double g01 = kf * (c1 - c0), g12 = kf * (c2 - c1), g23 = kf * (c3 - c2);\n"
              << "                q0 = 0.0; if (xl) q0 += gm; q0 += g01; if (yl) q0 += kf * ((double)ya.x - c0); if (yr) q0 += kf * ((double)yb.x - c0);\n"
              << "                if (zl) q0 += fzm0; if (zr) q0 += kf * (d0 - c0);\n"
              << "                q1 = 0.0; q1 += -g01; q1 += g12; if (yl) q1 += kf * ((double)ya.y - c1); if (yr) q1 += kf * ((double)yb.y - c1);\n"
              << "                if (zl) q1 += fzm1; if (zr) q1 += kf * (d1 - c1);\n"
              << "                q2 = 0.0; q2 += -g12; q2 += g23; if (yl) q2 += kf * ((double)ya.z - c2); if (yr) q2 += kf * ((double)yb.z - c2);\n"
              << "                if (zl) q2 += fzm2; if (zr) q2 += kf * (d2 - c2);\n"
              << "                q3 = 0.0; q3 += -g23; if (xr) q3 += gp; if (yl) q3 += kf * ((double)ya.w - c3); if (yr) q3 += kf * ((double)yb.w - c3);\n"
              << "                if (zl) q3 += fzm3; if (zr) q3 += kf * (d3 - c3);\n"
              << "            }\n"
              << "            q0 = div_hh<" << MK << "<P2>>(k, q0); q1 = div_hh<" << MK << "<P2>>(k, q1);\n"
              << "            q2 = div_hh<" << MK << "<P2>>(k, q2); q3 = div_hh<" << MK << "<P2>>(k, q3);\n"
              << "            const f4v nv = {fe_ftz(c.x + fe_rf((double)fe_rf(q0) * k.dt)), fe_ftz(c.y + fe_rf((double)fe_rf(q1) * k.dt)),\n"
              << "                            fe_ftz(c.z + fe_rf((double)fe_rf(q2) * k.dt)), fe_ftz(c.w + fe_rf((double)fe_rf(q3) * k.dt))};\n"
              << "            f4v* dst = reinterpret_cast<f4v*>(outp + (i64)z * k.nxy);\n"
              << "            if (STREAM) { asm volatile(\"\"); __builtin_nontemporal_store(nv, dst); asm volatile(\"\"); } else *dst = nv;\n"
              << "            if (TRACK) { asm volatile(\"\");\n"
              << "                const bool ch = __float_as_uint(nv.x) != __float_as_uint(c.x) || __float_as_uint(nv.y) != __float_as_uint(c.y) ||\n"
              << "                    __float_as_uint(nv.z) != __float_as_uint(c.z) || __float_as_uint(nv.w) != __float_as_uint(c.w);\n"
              << "                seen[0] = seen[0] || ch; }\n"
              << "            if (TOTAL) {\n"
              << "                const double v0 = (double)nv.x, v1 = (double)nv.y, v2 = (double)nv.z, v3 = (double)nv.w;\n"
              << "                ts += v0; tm += fabs(v0); ts += v1; tm += fabs(v1); ts += v2; tm += fabs(v2); ts += v3; tm += fabs(v3);\n"
              // out of range or not finite (the rare path: counted, and the least index kept)
              << "                const bool b0 = !(v0 >= lo_r && v0 <= hi_r), b1 = !(v1 >= lo_r && v1 <= hi_r), b2 = !(v2 >= lo_r && v2 <= hi_r),\n"
              << "                           b3 = !(v3 >= lo_r && v3 <= hi_r);\n"
              << "                if (b0 || b1 || b2 || b3) {\n"
              << "                    const double vs[4] = {v0, v1, v2, v3};\n"
              << "                    const bool bs[4] = {b0, b1, b2, b3};\n"
              << "#pragma unroll\n"
              << "                    for (int q = 0; q < 4; ++q) {\n"
              << "                        if (!bs[q]) continue;\n"
              << "                        if (finite_d(vs[q])) ++nout; else ++nbad;\n"
              << "                        const u64 i = (u64)z * (u64)k.nxy + (u64)y * (u64)k.nx + (u64)(x + q); firstbad = i < firstbad ? i : firstbad; } } }\n"
              << "            fzm0 = -fzp0; fzm1 = -fzp1; fzm2 = -fzp2; fzm3 = -fzp3;\n"
              << "        }\n"
              << "        buf ^= 1; c0 = d0; c1 = d1; c2 = d2; c3 = d3;\n"
              << "    };\n"
              // Whole groups of four planes, then the rest (a break inside the group would make the compiler wait for
              // nearly every load in flight at the loop's head).
              << "    int z = z0;\n"
              << "    for (; z + 4 <= z1; z += 4) { step(z, A, B, D); step(z + 1, B, C, A); step(z + 2, C, D, B); step(z + 3, D, A, C); }\n"
              << "    if (z < z1) { step(z, A, B, D); if (z + 1 < z1) { step(z + 1, B, C, A); if (z + 2 < z1) step(z + 2, C, D, B); } }\n"
              << "    if (TRACK) mark(k, " << t << ", seen[0]);\n"
              << "    if (TOTAL) {\n"
              << "        if (nout) atomicAdd(rec, nout);\n"
              << "        if (nbad) atomicAdd(rec + 1, nbad);\n"
              << "        if (firstbad != ~0ull) atomicMin(first, firstbad);\n"
              << "        __shared__ double wv[" << WAVES << "][4];\n"
              << "        Sum ps = {ts, 0.0}, pm = {tm, 0.0};\n"
              << "        wave_fold(ps); wave_fold(pm);\n"
              << "        if (lane == 0) { wv[r][0] = ps.s; wv[r][1] = ps.c; wv[r][2] = pm.s; wv[r][3] = pm.c; }\n"
              << "        __syncthreads();\n"
              << "        if (lane == 0 && r == 0) {\n"
              << "            Sum a = {wv[0][0], wv[0][1]}, b = {wv[0][2], wv[0][3]};\n"
              << "            for (int q = 1; q < " << WAVES << "; ++q) { merge(a, Sum{wv[q][0], wv[q][1]}); merge(b, Sum{wv[q][2], wv[q][3]}); }\n"
              << "            tsum[2 * tl] = a.s + a.c; tsum[2 * tl + 1] = b.s + b.c; } }\n"
              << "}\n";
            // Four kernels: rows within one tile or spanning several (the second loads halo columns, which costs
            // registers the first should not pay for: name + "x"), each as the step and as the step that also takes
            // the contract's totals (name + "_t").
            for (const bool xh : {false, true}) {
                for (const bool total : {false, true}) {
                    s << "extern \"C\" __global__ __launch_bounds__(" << 64 * WAVES << ")" << (total ? " __attribute__((amdgpu_waves_per_eu(4)))" : "") << " void " << name << (xh ? "x" : "") << (total ? "_t" : "")
                      << "(K k, int zchunk, int tiles_x, int tiles_y, int tiles_z, int mode, double* tsum, unsigned* rec, u64* first, double lo,\n"
                      << "    double hi) {\n"
                      << "    const bool tr = (mode & 1) != 0, rm = (mode & 2) != 0, st = (mode & 4) != 0;\n";
                    for (const bool p2 : {true, false}) {
                        s << "    " << (p2 ? "if ((k.flags & 3) == 3) " : "else ") << "run_" << name << "<" << (p2 ? "true" : "false") << ", "
                          << (xh ? "true" : "false") << ", " << (total ? "true" : "false")
                          << ">(k, zchunk, tiles_x, tiles_y, tiles_z, tr, rm, st, tsum, rec, first, lo, hi);\n";
                    }
                    s << "}\n";
                }
            }
            out_ << s.str();
            WindowKernel w;
            w.name = name;
            w.vx = 4;
            w.ty = TY;
            w.block_y = WAVES;
            w.flux = form.rows ? 2 : 1;
            w.zchunk = form.zchunk;
            code.windows.push_back(w);
        }
    }

    // How far from its cell the gather reads each field through what it inlines (stages are read where they lie).
    // Every op's reads lie within its radius: a transpose's runs at the lattice's edges stay within its offset
    // (shift_t) or one cell (gx_t, gy_t, gz_t) of the cell, at in-range coordinates.
    bool field_reach(const ir::Program& p, std::map<int, int>& out, int& cost) const
    {
        bool bounded = true;
        std::set<std::pair<int, int>> seen;
        std::function<void(int, int, bool)> walk = [&](int id, int off, bool root) {
            if (!seen.insert({id, off}).second) {
                return;
            }
            const ir::Inst& in = at(id);
            if (!root && lay_.stages.count(id) != 0) {
                return;
            }
            if (root && lay_.stages.count(id) != 0) {
                return; // the whole value was computed before: read at its own cell
            }
            if (in.op == Op::Field) {
                int& r = out[in.index];
                r = std::max(r, off);
                return;
            }
            for (const int a : in.args) {
                walk(a, off + in.radius, false);
            }
        };
        cost = 0;
        for (const ir::Equation& eq : p.equations) {
            walk(eq.value, 0, true);
            cost += at(eq.value).cost;
        }
        return bounded;
    }

    // Window gathers: tiles of cells marching in z, the fields read off-cell (`wf`, reach <= R) from an LDS ring of
    // planes. The ring holds planes z-R..z+R and the one being filled (RS >= 2R + 2 slots, a slot a plane, slot of
    // plane p = (p + RS) % RS, in unsigned arithmetic: p >= -R, so p + RS > 0); each plane of each field is loaded once by its tile, rows and halo columns at clamped
    // coordinates, so every coordinate within R of a cell holds the clamped cell's value (what a clamped read gets),
    // three planes ahead of use through a four-slot register ring (unrolled: register copies of loads in flight would
    // make the compiler wait for them). One barrier a plane: the slot being filled is one no lane reads that plane.
    //   x1: 64 x TY cells, a lane a cell; rows of 64 + 2R.
    //   x4: 256 x TY cells, a lane four consecutive cells (float4 loads and stores; nx a multiple of 4); rows of
    //       256 + 8 with the cells at [4, 260), so each lane's four are 16-byte aligned, the left halo before them
    //       and the right after. Lanes past the end of a row write the row's last value into their slots, which is
    //       the clamped value of every coordinate past the end.
    // The arithmetic is the cell function's, read through the window: the direct kernel's bits.
    void window_kernels(ProgramCode& code, const std::string& cellfn, const std::vector<int>& targets, const std::vector<int>& wf,
        int radius)
    {
        const int R = radius;
        const int NW = static_cast<int>(wf.size());
        const int RS = 2 * R + 2;
        // gfx950 gives a workgroup up to 160 KB of LDS, but a window that spills (a law with many fields at reach 2)
        // can reach its ring through flat addresses, and flat accesses past the first 64 KB of LDS fault: a
        // 76 KB ring of such a kernel did. Windows stay within 64 KB.
        constexpr int kLdsBudget = 64 * 1024;
        struct Shape {
            int vx;
            int ty;
        };
        std::vector<Shape> shapes;
        auto lds_bytes = [&](int vx, int ty) {
            const int tw = vx == 4 ? 264 : 64 + 2 * R;
            return NW * RS * (ty + 2 * R) * tw * 4;
        };
        for (const Shape sh : {Shape{1, R == 1 ? 6 : 4}, Shape{4, 2}, Shape{4, 4}, Shape{4, 6}}) {
            if (NW > 0 && lds_bytes(sh.vx, sh.ty) <= kLdsBudget) {
                shapes.push_back(sh);
            }
        }
        if (shapes.empty()) {
            return;
        }
        code.window = true;
        code.radius = R;
        code.zchunk = 32;
        for (const Shape sh : shapes) {
            const std::string name = code.gather + "_w" + std::to_string(sh.vx) + "_" + std::to_string(sh.ty);
            window_kernel(cellfn, name, targets, wf, R, sh.vx, sh.ty, code.in_place);
            code.windows.push_back({name, sh.vx, sh.ty, sh.ty + 2 * R});
        }
    }

    // A lane's VX cells through the cell function: the window's fast forms, where any cell flags (a NaN sum: an
    // infinity or NaN among its inputs) the lane's plane is kept in `redo` and taken again after the last plane,
    // exactly, by the direct gather's view; then the stores and change tracking. A flagged lane's ping-pong store
    // still happens (every store unconditional: one under a branch would make the compiler wait for loads in flight
    // at the next wait) and the redo overwrites it; its in-place stores and change marks wait for the redo, which
    // reads the cells as they were.
    std::string cells_body(const std::string& cellfn, int VX, const std::vector<int>& targets, bool in_place) const
    {
        const int NT = static_cast<int>(targets.size());
        std::ostringstream body;
        body << "            float o[" << VX << "][" << std::max(1, NT) << "], n[" << VX << "][" << std::max(1, NT) << "];\n";
        body << "            m.bad = false; m.zc = z;\n";
        for (int c = 0; c < VX; ++c) {
            body << "            " << cellfn << "(k, m, x + " << c << ", y, z, o[" << c << "], n[" << c << "]);\n";
        }
        body << "            const bool fine = !m.bad;\n            if (!fine) redo |= 1ull << (z - z0);\n"
             << store_body(VX, targets, in_place, "            ", "fine");
        return body.str();
    }

    // After a tile's last plane: the lane's flagged planes again, in the exact forms of the direct gather.
    std::string redo_body(const std::string& cellfn, int VX, const std::vector<int>& targets, bool in_place) const
    {
        const int NT = static_cast<int>(targets.size());
        std::ostringstream body;
        body << "    if (redo != 0) {\n        const Direct md(k);\n        do {\n"
             << "            const int z = z0 + __builtin_ctzll(redo); redo &= redo - 1;\n"
             << "            float o[" << VX << "][" << std::max(1, NT) << "], n[" << VX << "][" << std::max(1, NT) << "];\n";
        for (int c = 0; c < VX; ++c) {
            body << "            " << cellfn << "(k, md, x + " << c << ", y, z, o[" << c << "], n[" << c << "]);\n";
        }
        body << store_body(VX, targets, in_place, "            ", "true") << "        } while (redo != 0);\n    }\n";
        return body.str();
    }

    // The stores of a lane's VX cells at plane z (o: the old values, n: the new) and their change tracking: TRACK
    // only where the step reads the flags back; an in-place store still skips unchanged cells. `fine`: whether the
    // values are final (in-place stores and change marks only then).
    std::string store_body(int VX, const std::vector<int>& targets, bool in_place, const std::string& ind, const std::string& fine) const
    {
        const int NT = static_cast<int>(targets.size());
        std::ostringstream body;
        body << ind << "{ const i64 i = IDX(x, y, z);\n";
        for (int q = 0; q < NT; ++q) {
            const std::string t = std::to_string(targets[static_cast<std::size_t>(q)]);
            body << ind << "{ bool ch = false;\n";
            if (!in_place) {
                // A branch the compiler keeps (the empty asm cannot be hoisted out of it): untracked steps skip the
                // comparisons instead of computing them and masking the answer.
                body << ind << "  if (TRACK) { asm volatile(\"\");";
                for (int c = 0; c < VX; ++c) {
                    body << " ch = ch || __float_as_uint(o[" << c << "][" << q << "]) != __float_as_uint(n[" << c << "][" << q << "]);";
                }
                body << " }\n";
            }
            for (int c = 0; c < VX && in_place; ++c) {
                body << ind << "  { const bool d = __float_as_uint(o[" << c << "][" << q << "]) != __float_as_uint(n[" << c << "][" << q
                     << "]); ch = ch || d; if (d && " << fine << ") k.now[" << t << "][i + " << c << "] = n[" << c << "][" << q << "]; }\n";
            }
            if (!in_place) {
                if (VX == 4) {
                    // STREAM: a streaming (nontemporal) store, for lattices past the last-level cache (the empty asm
                    // around it keeps the compiler from hoisting or sinking the two stores into one ordinary store)
                    body << ind << "  { const f4v nv = {n[0][" << q << "], n[1][" << q << "], n[2][" << q << "], n[3][" << q << "]};\n"
                         << ind << "    if (STREAM) { asm volatile(\"\"); __builtin_nontemporal_store(nv, reinterpret_cast<f4v*>(k.nxt[" << t << "] + i)); asm volatile(\"\"); }\n"
                         << ind << "    else *reinterpret_cast<f4v*>(k.nxt[" << t << "] + i) = nv; }\n";
                } else {
                    body << ind << "  k.nxt[" << t << "][i] = n[0][" << q << "];\n";
                }
            }
            body << ind << "  seen[" << q << "] = seen[" << q << "] || (ch && " << fine << "); }\n";
        }
        body << ind << "}\n";
        return body.str();
    }

    void window_kernel(const std::string& cellfn, const std::string& name, const std::vector<int>& targets, const std::vector<int>& wf,
        int R, int VX, int TY, bool in_place)
    {
        const int NW = static_cast<int>(wf.size());
        const int RS = 2 * R + 2;
        const int TW = VX == 4 ? 264 : 64 + 2 * R; // row stride
        const int X0 = VX == 4 ? 4 : R;            // where the tile's first cell sits in its row
        const int TP = (TY + 2 * R) * TW;
        const int RF = RS * TP;
        const int NT = static_cast<int>(targets.size());
        const int SPAN = 64 * VX;
        const std::string W = "W_" + name;
        const std::string LDS = "L_" + name;
        std::ostringstream s;
        // The ring at file scope: reads through it keep its 16-byte alignment, so a lane's four neighbouring cells merge
        // into one wide LDS read.
        s << "__shared__ __attribute__((aligned(16))) float " << LDS << "[" << NW * RF << "];\n";
        // Read through the window in the fast forms (a flagged cell is taken again after the tile, in the exact forms).
        // A read at an offset from the lane's own cells that the compiler knows (a stencil's fixed neighbours, once
        // inlined) is served from registers when it lies in the lane's own column: the cells the lane loaded itself
        // for planes z-R..z+R (`own`), as the LDS ring holds them. Every other read goes through the ring.
        const int NOWN = 2 * R + 1;
        s << "template <bool P2> struct " << W << " {\n    static constexpr bool kClamps = false;\n    static constexpr bool kExact = false;\n"
          << "    static constexpr bool kPow2 = P2;\n"
          << "    int x0; int y0; int xc; int yc; int zc; mutable bool bad; double u[" << std::max(1, lay_.uniforms) << "];\n"
          << "    float own[" << NW << "][" << NOWN << "][" << VX << "];\n"
          << "    __device__ __forceinline__ void flag(bool b) const { bad = bad || b; }\n"
          << "    template <int F> __device__ __forceinline__ float fld(const K& k, int x, int y, int z) const {\n";
        for (int w = 0; w < NW; ++w) {
            // Selects, not an index: a known offset folds to one register, an unknown one costs selects, never
            // private memory.
            s << "        " << (w ? "else " : "") << "if constexpr (F == " << wf[static_cast<std::size_t>(w)] << ") {\n"
              << "            const int dx = x - xc, dy = y - yc, dz = z - zc;\n"
              << "            if (dy == 0 && (unsigned)(dz + " << R << ") <= " << 2 * R << "u && (unsigned)dx < " << VX << "u) {\n"
              << "                float v = own[" << w << "][0][0];\n"
              << "#pragma unroll\n                for (int j = 0; j < " << 2 * R + 1 << "; ++j)\n"
              << "#pragma unroll\n                    for (int c = 0; c < " << VX << "; ++c) v = (dz + " << R << " == j && dx == c) ? own[" << w
              << "][j][c] : v;\n"
              << "                return v;\n            }\n"
              << "            return " << LDS << "[" << w * RF << " + ((unsigned)(z + " << RS << ") % " << RS << "u) * " << TP << " + (y - y0 + " << R
              << ") * " << TW << " + (x - x0 + " << X0 << ")];\n        }\n";
        }
        s << "        else return k.now[F][IDX(x, y, z)];\n    }\n"
          << "    template <int F> __device__ __forceinline__ float fld_any(const K& k, int x, int y, int z) const {\n";
        for (int w = 0; w < NW; ++w) {
            s << "        " << (w ? "else " : "") << "if constexpr (F == " << wf[static_cast<std::size_t>(w)] << ") return " << LDS << "["
              << w * RF << " + ((unsigned)(z + " << RS << ") % " << RS << "u) * " << TP << " + (y - y0 + " << R << ") * " << TW << " + (x - x0 + "
              << X0 << ")];\n";
        }
        s << "        else return k.now[F][IDX(x, y, z)];\n    }\n};\n";
        // The kernel's body, instantiated for "h and h^2 are powers of two" (their divisions are multiplications,
        // with no branch between cells) and for the general case.
        s << "template <bool P2> __device__ __forceinline__ void run_" << name
          << "(const K& k, int zchunk, int tiles_x, int tiles_y, int tiles_z, const bool TRACK, const bool REMAP, const bool STREAM) {\n"
          << "    float* const lds = " << LDS << ";\n"
          << "    const int t = REMAP ? xcd_tile((int)blockIdx.x, tiles_x * tiles_y * tiles_z) : (int)blockIdx.x;\n"
          // x0 passes through an empty asm: the compiler then knows nothing of its low bits, so a neighbour's x + c
          // stays an add (not an or), and its offset from the lane's own cell, x + c - x, folds to c.
          << "    int x0 = (t % tiles_x) * " << SPAN << "; asm volatile(\"\" : \"+s\"(x0)); const int y0 = ((t / tiles_x) % tiles_y) * "
          << TY << ";\n"
          << "    const int z0 = (t / (tiles_x * tiles_y)) * zchunk; const int z1 = z0 + zchunk < k.nz ? z0 + zchunk : k.nz;\n"
          << "    const int lane = (int)threadIdx.x; const int r = (int)threadIdx.y;\n"
          << "    const int x = x0 + " << VX << " * lane; const int y = y0 + r - " << R << ";\n"
          << "    const bool live = r >= " << R << " && r < " << R + TY << " && x < k.nx && y < k.ny;\n"
          << "    const int ys = clampi(y, 0, k.ny - 1);\n"
          << "    const i64 rowo = (i64)ys * k.nx;\n"
          << "    const int zlim = z1 - 1 + " << R << " < k.nz - 1 ? z1 - 1 + " << R << " : k.nz - 1;\n"
          << "    " << W << "<P2> m; m.x0 = x0; m.y0 = y0; m.xc = x; m.yc = y;\n"
          << "    for (int q = 0; q < " << std::max(1, lay_.uniforms) << "; ++q) m.u[q] = " << (lay_.uniforms > 0 ? "k.U[q]" : "0.0")
          << ";\n"
          << "    float* row = lds + r * " << TW << ";\n";
        if (VX == 1) {
            // Lanes 0..R-1 also load the left halo columns, 64-R..63 the right; the others their own cell again (unused).
            s << "    const int xs = clampi(x, 0, k.nx - 1);\n"
              << "    const int hslot = lane < " << R << " ? lane : (lane >= " << 64 - R << " ? lane + " << 2 * R << " : -1);\n"
              << "    const int hx = clampi(lane < " << R << " ? x0 - " << R << " + lane : (lane >= " << 64 - R << " ? x0 + lane + " << R
              << " : x), 0, k.nx - 1);\n"
              << "    struct Pl { float v[" << NW << "]; float h[" << NW << "]; };\n"
              << "    auto load = [&](Pl& p, int zp) { const i64 o = (i64)clampi(zp, 0, zlim) * k.nxy + rowo;\n";
            for (int w = 0; w < NW; ++w) {
                const int f = wf[static_cast<std::size_t>(w)];
                s << "        p.v[" << w << "] = k.now[" << f << "][o + xs]; p.h[" << w << "] = k.now[" << f << "][o + hx];\n";
            }
            s << "    };\n"
              << "    auto put = [&](const Pl& p, int zp) { float* sl = row + ((unsigned)(zp + " << RS << ") % " << RS << "u) * " << TP << ";\n";
            for (int w = 0; w < NW; ++w) {
                s << "        sl[" << w * RF << " + " << R << " + lane] = p.v[" << w << "]; if (hslot >= 0) sl[" << w * RF << " + hslot] = p.h["
                  << w << "];\n";
            }
            s << "    };\n"
              << "    auto keep = [&](const Pl& p, const int slot) {\n";
            for (int w = 0; w < NW; ++w) {
                s << "        m.own[" << w << "][slot][0] = p.v[" << w << "];\n";
            }
            s << "    };\n";
        } else {
            // Lane 0 also loads the R columns left of the tile, lane 63 the R right of it; the others their own first
            // cell again (unused): every load is unconditional.
            s << "    const bool dead = x >= k.nx;\n"
              << "    const int xs = x < k.nx - 4 ? x : k.nx - 4;\n"
              << "    const bool hl = lane == 0 || lane == 63;\n"
              << "    const int hpos = lane == 0 ? " << 4 - R << " : 260;\n"
              << "    int hx[" << R << "];\n"
              << "    for (int j = 0; j < " << R << "; ++j) hx[j] = lane == 0 ? clampi(x0 - " << R << " + j, 0, k.nx - 1) : (lane == 63 ? "
              << "clampi(x0 + 256 + j, 0, k.nx - 1) : xs);\n"
              << "    struct Pl { float4 v[" << NW << "]; float h[" << NW << "][" << R << "]; };\n"
              << "    auto load = [&](Pl& p, int zp) { const i64 o = (i64)clampi(zp, 0, zlim) * k.nxy + rowo;\n";
            for (int w = 0; w < NW; ++w) {
                const int f = wf[static_cast<std::size_t>(w)];
                s << "        p.v[" << w << "] = *reinterpret_cast<const float4*>(k.now[" << f << "] + o + xs);\n"
                  << "        for (int j = 0; j < " << R << "; ++j) p.h[" << w << "][j] = k.now[" << f << "][o + hx[j]];\n";
            }
            s << "    };\n"
              << "    auto put = [&](const Pl& p, int zp) { float* sl = row + ((unsigned)(zp + " << RS << ") % " << RS << "u) * " << TP << ";\n";
            for (int w = 0; w < NW; ++w) {
                s << "        { float4 v = p.v[" << w << "]; if (dead) v = make_float4(v.w, v.w, v.w, v.w);\n"
                  << "          *reinterpret_cast<float4*>(sl + " << w * RF << " + 4 + 4 * lane) = v;\n"
                  << "          if (hl) for (int j = 0; j < " << R << "; ++j) sl[" << w * RF << " + hpos + j] = p.h[" << w << "][j]; }\n";
            }
            s << "    };\n"
              << "    auto keep = [&](const Pl& p, const int slot) {\n";
            for (int w = 0; w < NW; ++w) {
                s << "        m.own[" << w << "][slot][0] = p.v[" << w << "].x; m.own[" << w << "][slot][1] = p.v[" << w << "].y; m.own[" << w
                  << "][slot][2] = p.v[" << w << "].z; m.own[" << w << "][slot][3] = p.v[" << w << "].w;\n";
            }
            s << "    };\n";
        }
        // The cells of a lane: their new values and the stores.
        const std::string body = cells_body(cellfn, VX, targets, in_place);
        // A tile's change flags are marked once, after its last plane (a store a plane would be waited on with the loads).
        s << "    bool seen[" << std::max(1, NT) << "] = {};\n    unsigned long long redo = 0; // planes z0 + b whose cells a NaN sum flagged\n";
        s << "    for (int zp = z0 - " << R << "; zp < z0 + " << R << "; ++zp) { Pl p; load(p, zp); put(p, zp); keep(p, zp - z0 + " << R
          << "); }\n"
          << "    Pl A, B, C, D;\n"
          // In this order: the loop's first wait counts A as the oldest load in flight (were A issued last, the
          // compiler would wait for every load there).
          << "    load(A, z0 + " << R << "); __builtin_amdgcn_sched_barrier(0); load(B, z0 + " << R + 1
          << "); __builtin_amdgcn_sched_barrier(0); load(C, z0 + " << R + 2 << "); __builtin_amdgcn_sched_barrier(0);\n"
          << "    auto step = [&](const int z, const Pl& cur, Pl& far) {\n"
          << "        load(far, z + " << R + 3 << ");\n"
          << "        put(cur, z + " << R << ");\n"
          << "        keep(cur, " << 2 * R << ");\n"
          << "        __syncthreads();\n"
          << "        if (live) {\n" << body << "        }\n"
          << "        for (int w = 0; w < " << NW << "; ++w) for (int j = 0; j < " << 2 * R << "; ++j) for (int c = 0; c < " << VX
          << "; ++c) m.own[w][j][c] = m.own[w][j + 1][c];\n"
          << "    };\n"
          // Whole groups of four planes, then the rest: a break inside the group would join the loop's latch, and
          // the compiler, counting the loads in flight over every path into the loop's head, would wait for nearly all.
          << "    int z = z0;\n"
          << "    for (; z + 4 <= z1; z += 4) {\n"
          << "        step(z, A, D); step(z + 1, B, A); step(z + 2, C, B); step(z + 3, D, C);\n"
          << "    }\n"
          << "    if (z < z1) { step(z, A, D); if (z + 1 < z1) { step(z + 1, B, A); if (z + 2 < z1) step(z + 2, C, B); } }\n";
        s << redo_body(cellfn, VX, targets, in_place);
        for (int q = 0; q < NT; ++q) {
            s << "    if (TRACK) mark(k, " << targets[static_cast<std::size_t>(q)] << ", seen[" << q << "]);\n";
        }
        s << "}\n";
        s << "extern \"C\" __global__ __launch_bounds__(" << 64 * (TY + 2 * R) << ") void " << name
          << "(K k, int zchunk, int tiles_x, int tiles_y, int tiles_z, int mode) {\n"
          << "    if ((k.flags & 3) == 3) run_" << name << "<true>(k, zchunk, tiles_x, tiles_y, tiles_z, (mode & 1) != 0, (mode & 2) != 0, (mode & 4) != 0);\n"
          << "    else run_" << name << "<false>(k, zchunk, tiles_x, tiles_y, tiles_z, (mode & 1) != 0, (mode & 2) != 0, (mode & 4) != 0);\n}\n";
        out_ << s.str();
    }

    const ir::Module& m_;
    const Law& law_;
    LawLayout& lay_;
    std::ostringstream out_;
    std::vector<char> live_;
};

} // namespace

std::string generate(const ir::Module& m, const Law& law, LawLayout& layout, std::string& why)
{
    layout = LawLayout{};
    try {
        return Generator(m, law, layout).run(why);
    } catch (const std::exception& e) {
        why = e.what();
        return "";
    }
}

std::vector<float> net_blob(const Law& law, std::vector<std::size_t>& offsets)
{
    std::vector<float> blob;
    offsets.clear();
    for (const NetDecl& net : law.nets) {
        offsets.push_back(blob.size());
        const int nin = net.inputs();
        for (int i = 0; i < nin; ++i) {
            blob.push_back(i < static_cast<int>(net.in_shift.size()) ? net.in_shift[static_cast<std::size_t>(i)] : 0.0f);
        }
        for (int i = 0; i < nin; ++i) {
            blob.push_back(i < static_cast<int>(net.in_scale.size()) ? net.in_scale[static_cast<std::size_t>(i)] : 1.0f);
        }
        blob.push_back(net.out_scale);
        blob.insert(blob.end(), net.weights.begin(), net.weights.end());
        blob.resize(blob.size() + ((4 - blob.size() % 4) % 4), 0.0f); // each net 16-byte aligned
    }
    return blob;
}

std::vector<std::string> compile_options()
{
    // -O3; no fused multiply-adds (the CPU reference has none, and its roundings are what the kernels reproduce);
    // f32 denormals flushed by the hardware (the CPU's FTZ/DAZ threads); packed FP32 off (some packed forms have
    // given wrong results on this hardware).
    return {"-O3", "-std=c++17", "-ffp-contract=off", "-fgpu-flush-denormals-to-zero", "-Xclang", "-target-feature", "-Xclang",
        "-packed-fp32-ops"};
}

} // namespace fieldengine::field::hip
