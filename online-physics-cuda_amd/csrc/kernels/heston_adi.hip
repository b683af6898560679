// Batched Heston ADI on gfx950: thousands of options of uneven lattice sizes stepped together, every step four
// batched passes over all options, the whole run one captured graph.
//
// The scheme and operators are those of the engine's Heston ADI laws (one option a plane; the reference's
// bench/laws/heston_adi.law:1-51 and bench/finance.py:253-288): x = ln S along x (cells of h = 6 / nx on ln 100 +- 3),
// var = sigma (j + 1/2) h along y, F = A0 + A1 + A2 with three-point A1 (x) and A2 (y), A0 = rho var gx(gy(.)), the
// edge cell repeating past the ends. In increments (delta form), for Craig-Sneyd (scheme 0) or Hundsdorfer-Verwer (1):
//   pass 1 (rows):    T1 = (I - th dt A1)^-1 dt F(U)
//   pass 2 (columns): D2 = (I - th dt A2)^-1 T1
//   pass 3 (rows):    T1 = (I - th dt A1)^-1 (CS: dt/2 A0 D2;  HV: dt F(U) - D2 + dt/2 F(D2))
//   pass 4 (columns): U += D2 + (I - th dt A2)^-1 T1
// in fp32 (multiply-adds fused).
//
// Rows (passes 1 and 3): L lanes a row, K = nx / L cells a lane, in registers: each lane reduces its chunk of the
// tridiagonal system (a downward and an upward sweep) to one equation in the last unknowns of its own and its
// neighbours' chunks, the L equations are solved by parallel cyclic reduction across the lanes (shuffles), and each
// lane back-substitutes its chunk. K (Thomas work a lane) against L (PCR depth) is the split the tuner picks.
// Columns (passes 2 and 4): a thread a column, Thomas with the option's factorisation of (I - th dt A2) (the same for
// every column: it depends on the row only), computed once in double on the host; the forward sweep writes d' in
// place and the backward sweep reads it back. Loads are coalesced across the columns of a block.
//
// Options of one lattice shape form a class; a pass is one launch a class, each block a run of rows (columns) of one
// option from a task table.
#include "fe_adi.h"
#include "fe_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#pragma clang fp contract(fast)

namespace {

struct HOpt {
    long long off; // first cell of the option's lattice (row-major, x fastest)
    int nx, ny;
    int row;       // its first entry in the row table
    float h, dt, thdt, r;
    int part8, part16; // its partitioned column solve's coefficients in the part table, 8 or 16 rows a thread (-1: none)
};

// The part table (part_table): a row's m, a, C (downward sweep), Q, R (after the upward sweep), then a chunk's C_(M-1)
// (the last row's downward C), al, m and c' (its equation's factorisation).
constexpr int kPartRow = 5;
constexpr int kPartChunk = 4;

struct HRow {
    float lx, ux, ly, uy; // A1, A2 coefficients of the row
    float rv;             // rho * var of the row
    float ya, ycp, ym;    // (I - th dt A2) factorised: lower coefficient, c', 1 / pivot (edges folded)
    float yb, yc;         // (I - th dt A2) itself: the diagonal (edges folded) and the upper coefficient (0 on the last row)
};

enum Mode { kFirst = 0, kCsSecond = 1, kHvSecond = 2 };

template <int K, int L>
struct Row {
    float v[K];
};

// Lane p of a group of L: `send` of lane p - 1 (of p + 1), or `edge` where the row ends there (the edge cell
// repeating past the end).
template <int L>
__device__ __forceinline__ float from_left(float send, float edge, int p)
{
    const float o = __shfl_up(send, 1, L);
    return p == 0 ? edge : o;
}
template <int L>
__device__ __forceinline__ float from_right(float send, float edge, int p)
{
    const float o = __shfl_down(send, 1, L);
    return p == L - 1 ? edge : o;
}

// F restricted to what `mode` needs, at the K cells of a lane in row j: dt F(W) or dt/2 A0 W, from rows jm, j, jp of W.
template <int K, int L>
__device__ __forceinline__ void apply(const Row<K, L>& wm, const Row<K, L>& w, const Row<K, L>& wp, const HRow& cf, float h,
    float r, float span_y, int p, bool mixed_only, float scale, float* out)
{
    // gy at the lane's cells and at the cells either side (from the neighbouring lanes); gx(gy) one-sided at the ends.
    float gy[K];
    const float iy = span_y > 0.0f ? 1.0f / (h * span_y) : 0.0f;
#pragma unroll
    for (int q = 0; q < K; ++q) {
        gy[q] = (wp.v[q] - wm.v[q]) * iy;
    }
    const float gy_left = from_left<L>(gy[K - 1], gy[0], p);
    const float gy_right = from_right<L>(gy[0], gy[K - 1], p);
    const float w_left = from_left<L>(w.v[K - 1], w.v[0], p);
    const float w_right = from_right<L>(w.v[0], w.v[K - 1], p);
    const float ih = 1.0f / h, ih2 = 0.5f / h;
#pragma unroll
    for (int q = 0; q < K; ++q) {
        const bool first = p == 0 && q == 0;
        const bool last = p == L - 1 && q == K - 1;
        const float gl = q == 0 ? gy_left : gy[q - 1];
        const float gr = q == K - 1 ? gy_right : gy[q + 1];
        // central (gr - gl) / 2h; at an end the cell itself stands for the missing one and the span is h
        const float gxy = (first || last) ? (gr - gl) * ih : (gr - gl) * ih2;
        float f = cf.rv * gxy;
        if (!mixed_only) {
            const float c = w.v[q];
            const float l = q == 0 ? w_left : w.v[q - 1];
            const float rr = q == K - 1 ? w_right : w.v[q + 1];
            f += cf.lx * (l - c) + cf.ux * (rr - c) + cf.ly * (wm.v[q] - c) + cf.uy * (wp.v[q] - c) - r * c;
        }
        out[q] = scale * f;
    }
}

template <int K>
__device__ __forceinline__ void load_row(const float* p, float (&v)[K])
{
    if constexpr (K % 4 == 0) {
#pragma unroll
        for (int q = 0; q < K; q += 4) {
            const float4 t = *reinterpret_cast<const float4*>(p + q);
            v[q] = t.x;
            v[q + 1] = t.y;
            v[q + 2] = t.z;
            v[q + 3] = t.w;
        }
    } else {
#pragma unroll
        for (int q = 0; q < K; q += 2) {
            const float2 t = *reinterpret_cast<const float2*>(p + q);
            v[q] = t.x;
            v[q + 1] = t.y;
        }
    }
}

template <int K>
__device__ __forceinline__ void store_row(float* p, const float (&v)[K])
{
    if constexpr (K % 4 == 0) {
#pragma unroll
        for (int q = 0; q < K; q += 4) {
            *reinterpret_cast<float4*>(p + q) = make_float4(v[q], v[q + 1], v[q + 2], v[q + 3]);
        }
    } else {
#pragma unroll
        for (int q = 0; q < K; q += 2) {
            *reinterpret_cast<float2*>(p + q) = make_float2(v[q], v[q + 1]);
        }
    }
}

// Solves (b on the diagonal, a below, c above; the first row's a and the last row's c folded onto the diagonal) for
// the row whose right-hand side the L lanes of a group hold, K cells a lane; the answer replaces d. Quotients are taken
// through the hardware reciprocal (v_rcp_f32, 1 ulp): an IEEE division is a sequence of about ten instructions, and the
// row passes are bound by their instruction issue; the systems are diagonally dominant, so the extra half ulp a quotient
// does not grow.
template <int K, int L>
__device__ __forceinline__ void solve_row(float (&d)[K], float a, float b, float c, int p)
{
    static_assert(K >= 2, "a lane holds two cells at least");
    float C[K], A[K], D[K];
    // Downward: x_i + C_i x_{i+1} + A_i x_{s-1} = D_i, x_{s-1} the last unknown of the lane before.
    {
        const bool first = p == 0;
        const float bi = b + (first ? a : 0.0f);
        const float m = __builtin_amdgcn_rcpf(bi);
        C[0] = c * m;
        A[0] = first ? 0.0f : a * m;
        D[0] = d[0] * m;
    }
#pragma unroll
    for (int q = 1; q < K; ++q) {
        const bool last = p == L - 1 && q == K - 1;
        const float bi = b + (last ? c : 0.0f);
        const float m = __builtin_amdgcn_rcpf(bi - a * C[q - 1]);
        C[q] = last ? 0.0f : c * m;
        A[q] = -a * A[q - 1] * m;
        D[q] = (d[q] - a * D[q - 1]) * m;
    }
    // Upward: x_q = P_q - Q_q x_e - R_q x_{s-1} for q < K - 1 (x_e the lane's own last unknown).
    float P[K], Q[K], R[K];
    P[K - 2] = D[K - 2];
    Q[K - 2] = C[K - 2];
    R[K - 2] = A[K - 2];
#pragma unroll
    for (int q = K - 3; q >= 0; --q) {
        P[q] = D[q] - C[q] * P[q + 1];
        Q[q] = -C[q] * Q[q + 1];
        R[q] = A[q] - C[q] * R[q + 1];
    }
    // One equation a lane in u_p = x_e: the last row, with x_{e+1} (the next lane's first) = P' - Q' u_{p+1} - R' u_p.
    const float p1 = __shfl_down(P[0], 1, L), q1 = __shfl_down(Q[0], 1, L), r1 = __shfl_down(R[0], 1, L);
    const bool tail = p == L - 1;
    const float ce = C[K - 1];
    float al = A[K - 1];
    float be = 1.0f - (tail ? 0.0f : ce * r1);
    float ga = tail ? 0.0f : -ce * q1;
    float de = D[K - 1] - (tail ? 0.0f : ce * p1);
    // Parallel cyclic reduction over the L lanes.
#pragma unroll
    for (int s = 1; s < L; s <<= 1) {
        const float alm = __shfl_up(al, s, L), bem = __shfl_up(be, s, L), gam = __shfl_up(ga, s, L), dem = __shfl_up(de, s, L);
        const float alp = __shfl_down(al, s, L), bep = __shfl_down(be, s, L), gap = __shfl_down(ga, s, L),
                    dep = __shfl_down(de, s, L);
        const bool has_m = p >= s, has_p = p + s < L;
        const float k1 = has_m ? al * __builtin_amdgcn_rcpf(bem) : 0.0f;
        const float k2 = has_p ? ga * __builtin_amdgcn_rcpf(bep) : 0.0f;
        const float nal = has_m ? -k1 * alm : 0.0f;
        const float nga = has_p ? -k2 * gap : 0.0f;
        const float nbe = be - (has_m ? k1 * gam : 0.0f) - (has_p ? k2 * alp : 0.0f);
        const float nde = de - (has_m ? k1 * dem : 0.0f) - (has_p ? k2 * dep : 0.0f);
        al = nal;
        be = nbe;
        ga = nga;
        de = nde;
    }
    const float u = de * __builtin_amdgcn_rcpf(be);
    const float ul = from_left<L>(u, 0.0f, p); // lane 0 has no unknown before it (its R terms are 0)
    d[K - 1] = u;
#pragma unroll
    for (int q = 0; q < K - 1; ++q) {
        d[q] = P[q] - Q[q] * u - R[q] * ul;
    }
}

struct Task {
    int opt;
    int first; // first row (passes 1, 3) or first column (passes 2, 4)
};

// Passes 1 and 3: a block of WX waves, each wave 64 / L rows, one option's consecutive rows.
template <int K, int L, int MODE>
__global__ __launch_bounds__(256) void k_rows(const HOpt* __restrict__ opts, const HRow* __restrict__ rows,
    const Task* __restrict__ tasks, const float* __restrict__ U, const float* __restrict__ D2, float* __restrict__ T1)
{
    const Task t = tasks[blockIdx.x];
    const HOpt o = opts[t.opt];
    const int lane = (int)(threadIdx.x & 63u);
    const int group = (int)(threadIdx.x >> 6) * (64 / L) + lane / L;
    const int p = lane % L;
    const int j = t.first + group;
    if (j >= o.ny) {
        return; // the whole group: shuffles stay inside groups
    }
    const HRow cf = rows[o.row + j];
    const int jm = j > 0 ? j - 1 : j, jp = j + 1 < o.ny ? j + 1 : j;
    const float span_y = (float)(jp - jm);
    const long long col = (long long)p * K;
    float rhs[K];
    Row<K, L> wm, w, wp;
    const float* base = U + o.off + col; // the launch passes D2 here for the Craig-Sneyd correction (dt/2 A0 D2)
    load_row<K>(base + (long long)jm * o.nx, wm.v);
    load_row<K>(base + (long long)j * o.nx, w.v);
    load_row<K>(base + (long long)jp * o.nx, wp.v);
    if (MODE == kCsSecond) {
        apply<K, L>(wm, w, wp, cf, o.h, o.r, span_y, p, true, 0.5f * o.dt, rhs);
    } else {
        apply<K, L>(wm, w, wp, cf, o.h, o.r, span_y, p, false, o.dt, rhs);
    }
    if (MODE == kHvSecond) {
        // dt F(U) - D2 + dt/2 F(D2)
        const float* dbase = D2 + o.off + col;
        load_row<K>(dbase + (long long)jm * o.nx, wm.v);
        load_row<K>(dbase + (long long)j * o.nx, w.v);
        load_row<K>(dbase + (long long)jp * o.nx, wp.v);
        float fd[K];
        apply<K, L>(wm, w, wp, cf, o.h, o.r, span_y, p, false, 0.5f * o.dt, fd);
#pragma unroll
        for (int q = 0; q < K; ++q) {
            rhs[q] += fd[q] - w.v[q];
        }
    }
    const float cx = -cf.lx - cf.ux - 0.5f * o.r;
    solve_row<K, L>(rhs, -o.thdt * cf.lx, 1.0f - o.thdt * cx, -o.thdt * cf.ux, p);
    store_row<K>(T1 + o.off + (long long)j * o.nx + col, rhs);
}

// Passes 2 and 4: a thread a column of one option. FINAL: U += D2 + x (and D2 is only read); else D2 = x.
template <bool FINAL>
__global__ __launch_bounds__(256) void k_cols(const HOpt* __restrict__ opts, const HRow* __restrict__ rows,
    const Task* __restrict__ tasks, float* __restrict__ U, float* __restrict__ D2, float* __restrict__ T1)
{
    const Task t = tasks[blockIdx.x];
    const HOpt o = opts[t.opt];
    const int i = t.first + (int)threadIdx.x;
    if (i >= o.nx) {
        return;
    }
    const long long nx = o.nx;
    const int ny = o.ny;
    float* col = T1 + o.off + i;
    const HRow* cf = rows + o.row;
    // The sweeps walk kChunk rows at a time: the chunk's loads are issued together (they do not depend on the
    // recurrence), so a column waits for memory once a chunk instead of once a row. The arithmetic and its order are
    // the row-by-row sweep's.
    constexpr int kChunk = 8;
    float prev = 0.0f;
    int j = 0;
    for (; j + kChunk <= ny; j += kChunk) {
        float d[kChunk];
#pragma unroll
        for (int q = 0; q < kChunk; ++q) {
            d[q] = col[(j + q) * nx];
        }
#pragma unroll
        for (int q = 0; q < kChunk; ++q) {
            prev = (d[q] - (j + q == 0 ? 0.0f : cf[j + q].ya * prev)) * cf[j + q].ym;
            col[(j + q) * nx] = prev;
        }
    }
    for (; j < ny; ++j) {
        const float dj = col[j * nx];
        prev = (dj - (j == 0 ? 0.0f : cf[j].ya * prev)) * cf[j].ym;
        col[j * nx] = prev;
    }
    float next = 0.0f;
    float* d2 = D2 + o.off + i;
    float* u = U + o.off + i;
    j = ny - 1;
    for (; j - kChunk + 1 >= 0; j -= kChunk) {
        float c[kChunk], a[kChunk], b[kChunk];
#pragma unroll
        for (int q = 0; q < kChunk; ++q) {
            c[q] = col[(j - q) * nx];
            if (FINAL) {
                a[q] = u[(j - q) * nx];
                b[q] = d2[(j - q) * nx];
            }
        }
#pragma unroll
        for (int q = 0; q < kChunk; ++q) {
            next = c[q] - cf[j - q].ycp * next;
            if (FINAL) {
                u[(j - q) * nx] = a[q] + (b[q] + next);
            } else {
                d2[(j - q) * nx] = next;
            }
        }
    }
    for (; j >= 0; --j) {
        next = col[j * nx] - cf[j].ycp * next;
        if (FINAL) {
            u[j * nx] = u[j * nx] + (d2[j * nx] + next);
        } else {
            d2[j * nx] = next;
        }
    }
}

// Passes 2 and 4 with the forward sweep's d' kept in LDS (a 64-thread block, a column a thread, [ny][64] floats) instead
// of written to T1 and read back: 8 bytes a cell less HBM traffic a pass. The same arithmetic in the same order as
// k_cols; a thread touches only its own column, so no barrier.
template <bool FINAL>
__global__ __launch_bounds__(64) void k_cols_lds(const HOpt* __restrict__ opts, const HRow* __restrict__ rows,
    const Task* __restrict__ tasks, float* __restrict__ U, float* __restrict__ D2, const float* __restrict__ T1)
{
    extern __shared__ float dprime[];
    const Task t = tasks[blockIdx.x];
    const HOpt o = opts[t.opt];
    const int i = t.first + (int)threadIdx.x;
    if (i >= o.nx) {
        return;
    }
    const long long nx = o.nx;
    const int ny = o.ny;
    const float* col = T1 + o.off + i;
    float* mine = dprime + threadIdx.x;
    const HRow* cf = rows + o.row;
    constexpr int kChunk = 8;
    float prev = 0.0f;
    int j = 0;
    for (; j + kChunk <= ny; j += kChunk) {
        float d[kChunk];
#pragma unroll
        for (int q = 0; q < kChunk; ++q) {
            d[q] = col[(j + q) * nx];
        }
#pragma unroll
        for (int q = 0; q < kChunk; ++q) {
            prev = (d[q] - (j + q == 0 ? 0.0f : cf[j + q].ya * prev)) * cf[j + q].ym;
            mine[(j + q) * 64] = prev;
        }
    }
    for (; j < ny; ++j) {
        prev = (col[j * nx] - (j == 0 ? 0.0f : cf[j].ya * prev)) * cf[j].ym;
        mine[j * 64] = prev;
    }
    float next = 0.0f;
    float* d2 = D2 + o.off + i;
    float* u = U + o.off + i;
    j = ny - 1;
    for (; j - kChunk + 1 >= 0; j -= kChunk) {
        float c[kChunk], a[kChunk], b[kChunk];
#pragma unroll
        for (int q = 0; q < kChunk; ++q) {
            c[q] = mine[(j - q) * 64];
            if (FINAL) {
                a[q] = u[(j - q) * nx];
                b[q] = d2[(j - q) * nx];
            }
        }
#pragma unroll
        for (int q = 0; q < kChunk; ++q) {
            next = c[q] - cf[j - q].ycp * next;
            if (FINAL) {
                u[(j - q) * nx] = a[q] + (b[q] + next);
            } else {
                d2[(j - q) * nx] = next;
            }
        }
    }
    for (; j >= 0; --j) {
        next = mine[j * 64] - cf[j].ycp * next;
        if (FINAL) {
            u[j * nx] = u[j * nx] + (d2[j * nx] + next);
        } else {
            d2[j * nx] = next;
        }
    }
}

// Passes 2 and 4 partitioned: a block takes 64 columns of one option, P = blockDim.y waves, wave p the rows
// [p M, (p + 1) M) of every column in registers (rows past the option's ny are x = 0 equations, which couple to
// nothing: the last real row's upper coefficient is folded). Each thread reduces its chunk as the row solver does (a
// downward and an upward sweep express every unknown in the chunk's last unknown and the one before the chunk); the P
// equations of a column in those unknowns are solved down the column by one thread (Thomas through LDS; measured
// faster than parallel cyclic reduction over the waves, whose extra barriers and divisions cost more than the idle
// waves), and each thread substitutes back. (I - th dt A2) is the same matrix for every column of an option, so every
// coefficient of that elimination (the sweeps' multipliers, the chunk equations and their factorisation) is the
// option's own, computed once on the host in double (part_table) and read here as scalars; a column carries only its
// right-hand side through them: three multiply-adds a cell and no division. HBM traffic is the pass's input and output
// once: 8 bytes a cell (16 with the update of U), against 16 (24) for k_cols. The arithmetic differs from the Thomas
// sweep's, so its bits do too (both fp32).
template <int M, bool FINAL>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8))) void k_cols_part(const HOpt* __restrict__ opts, const float* __restrict__ ptab,
    const Task* __restrict__ tasks, float* __restrict__ U, float* __restrict__ D2, const float* __restrict__ T1)
{
    static_assert(M >= 2, "a chunk holds two rows at least");
    extern __shared__ float red[]; // [3][P][64]: a chunk's first unknown P0, the chunk equations' right sides, u
    const Task t = tasks[blockIdx.x];
    const HOpt o = opts[t.opt];
    // blockDim.x is the wave: p is the same across a wave, so the table's entries are scalar loads
    const int c = (int)threadIdx.x, p = __builtin_amdgcn_readfirstlane((int)threadIdx.y), P = (int)blockDim.y;
    const int i = t.first + c;
    const bool on = i < o.nx;
    const int ic = on ? i : o.nx - 1;
    const int ny = o.ny;
    const long long nx = o.nx;
    const float* rt = ptab + (M == 8 ? o.part8 : o.part16); // [P M rows][kPartRow], then [P chunks][kPartChunk]
    const float* ct = rt + P * M * kPartRow;
    const int j0 = p * M;
    const int stride = P * 64;
    float* sP = red;
    float* sDe = red + stride;
    float* sU = red + 2 * stride;
    const int me = p * 64 + c;
    // every load unconditional (clamped rows and column), padding selected afterwards
    float x[M];
#pragma unroll
    for (int q = 0; q < M; ++q) {
        const int j = min(j0 + q, ny - 1);
        x[q] = T1[o.off + j * nx + ic];
    }
    // downward: x_q = (d_q - a_q x_(q-1)) m_q
#pragma unroll
    for (int q = 0; q < M; ++q) {
        const float* e = rt + (j0 + q) * kPartRow;
        const float d = j0 + q < ny ? x[q] : 0.0f;
        x[q] = q == 0 ? d * e[0] : (d - e[1] * x[q - 1]) * e[0];
    }
    // upward: x_q -= C_q x_(q+1), C the downward sweep's
#pragma unroll
    for (int q = M - 3; q >= 0; --q) {
        x[q] = x[q] - rt[(j0 + q) * kPartRow + 2] * x[q + 1];
    }
    sP[me] = x[0];
    __syncthreads();
    // the chunk's equation in (its last unknown, the next chunk's): right side x_(M-1) - C_(M-1) P0 of the next chunk
    sDe[me] = x[M - 1] - ct[p * kPartChunk] * (p == P - 1 ? 0.0f : sP[me + 64]);
    __syncthreads();
    if (p == 0) {
        // Thomas down the column's P chunk equations, the factorisation the table's: d'_k = (de_k - al_k d'_(k-1)) m_k,
        // then u_k = d'_k - c'_k u_(k+1)
        float dp = 0.0f;
        for (int k = 0; k < P; ++k) {
            const float* e = ct + k * kPartChunk;
            const int s = k * 64 + c;
            dp = (sDe[s] - e[1] * dp) * e[2];
            sDe[s] = dp;
        }
        float u = 0.0f;
        for (int k = P - 1; k >= 0; --k) {
            const int s = k * 64 + c;
            u = sDe[s] - ct[k * kPartChunk + 3] * u;
            sU[s] = u;
        }
    }
    __syncthreads();
    const float u = sU[me];
    const float ul = p > 0 ? sU[me - 64] : 0.0f;
    x[M - 1] = u;
#pragma unroll
    for (int q = 0; q < M - 1; ++q) {
        const float* e = rt + (j0 + q) * kPartRow;
        x[q] = x[q] - e[3] * u - e[4] * ul;
    }
    if (FINAL) {
        __builtin_amdgcn_sched_barrier(0);
        float a[M], b[M];
#pragma unroll
        for (int q = 0; q < M; ++q) {
            const long long at = o.off + min(j0 + q, ny - 1) * nx + ic;
            a[q] = U[at];
            b[q] = D2[at];
        }
#pragma unroll
        for (int q = 0; q < M; ++q) {
            if (on && j0 + q < ny) {
                U[o.off + (j0 + q) * nx + i] = a[q] + (b[q] + x[q]);
            }
        }
    } else {
#pragma unroll
        for (int q = 0; q < M; ++q) {
            if (on && j0 + q < ny) {
                D2[o.off + (j0 + q) * nx + i] = x[q];
            }
        }
    }
}

// The coefficients k_cols_part needs for one option, P chunks of M rows (rows past ny the x = 0 equations), in double
// from the row table's (I - th dt A2): per row the downward sweep's multiplier m = 1 / (b - a C_(q-1)) (1 / b on a
// chunk's first row), its a and C = c m, and after the upward sweep the coefficients Q, R of the chunk's last unknown and
// of the one before the chunk; per chunk C_(M-1), and the chunk equations al u_(k-1) + be u_k + ga u_(k+1) = de
// (al = A_(M-1), be = 1 - C_(M-1) R_0', ga = -C_(M-1) Q_0', ' the next chunk's) factorised: al, m = 1 / (be - al c'),
// c' = ga m.
void part_table(const HRow* cf, int ny, int P, int M, std::vector<float>& out)
{
    const int R = P * M;
    std::vector<double> m(R), a(R), C(R), A(R), Q(R), Rr(R);
    for (int j = 0; j < R; ++j) {
        const int q = j % M;
        const bool real = j < ny;
        const HRow& r = cf[std::min(j, ny - 1)];
        const double aj = real && j > 0 ? r.ya : 0.0;
        const double bj = real ? r.yb : 1.0;
        const double cj = real ? r.yc : 0.0;
        m[j] = q == 0 ? 1.0 / bj : 1.0 / (bj - aj * C[j - 1]);
        C[j] = cj * m[j];
        A[j] = q == 0 ? aj * m[j] : -aj * A[j - 1] * m[j];
        a[j] = aj;
    }
    for (int p = 0; p < P; ++p) {
        const int j0 = p * M;
        for (int q = 0; q < M; ++q) {
            Q[j0 + q] = C[j0 + q];
            Rr[j0 + q] = A[j0 + q];
        }
        for (int q = M - 3; q >= 0; --q) {
            const int j = j0 + q;
            Rr[j] = Rr[j] - C[j] * Rr[j + 1];
            Q[j] = -C[j] * Q[j + 1];
        }
    }
    for (int j = 0; j < R; ++j) {
        for (const double v : {m[j], a[j], C[j], Q[j], Rr[j]}) {
            out.push_back(static_cast<float>(v));
        }
    }
    double cp = 0.0;
    for (int p = 0; p < P; ++p) {
        const int last = p * M + M - 1;
        const bool tail = p == P - 1;
        const double ce = C[last], al = p > 0 ? A[last] : 0.0;
        const double r1 = tail ? 0.0 : Rr[(p + 1) * M], q1 = tail ? 0.0 : Q[(p + 1) * M];
        const double be = 1.0 - ce * r1, ga = -ce * q1;
        const double mm = 1.0 / (be - al * cp);
        cp = ga * mm;
        for (const double v : {ce, al, mm, cp}) {
            out.push_back(static_cast<float>(v));
        }
    }
}

// The cell-averaged put payoff (in double, rounded once).
__global__ void k_payoff(const HOpt* __restrict__ opts, const double* __restrict__ strikes, int n, float* __restrict__ U)
{
    const int o_id = (int)blockIdx.y;
    if (o_id >= n) {
        return;
    }
    const HOpt o = opts[o_id];
    const long long cells = (long long)o.nx * o.ny;
    const double k = strikes[o_id], lk = log(k), h = (double)6.0 / o.nx;
    const double x0 = log(100.0) - 3.0 + 0.5 * h;
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < cells; c += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(c % o.nx);
        const double x = x0 + i * h;
        const double lo = fmin(x - 0.5 * h, lk), hi = fmin(x + 0.5 * h, lk);
        U[o.off + c] = (float)((k * (hi - lo) - (exp(hi) - exp(lo))) / h);
    }
}

// Prices: the 4 x 4 Lagrange stencil of each option at (ln 100, v0), weights from the host.
struct Stencil {
    int i0, j0;
    double wx[4], wy[4];
};

__global__ void k_price(const HOpt* __restrict__ opts, const Stencil* __restrict__ st, int n, const float* __restrict__ U,
    double* __restrict__ out)
{
    const int o_id = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (o_id >= n) {
        return;
    }
    const HOpt o = opts[o_id];
    const Stencil s = st[o_id];
    double p = 0.0;
    for (int b = 0; b < 4; ++b) {
        double row = 0.0;
        for (int a = 0; a < 4; ++a) {
            row += s.wx[a] * (double)U[o.off + (long long)(s.j0 + b) * o.nx + s.i0 + a];
        }
        p += s.wy[b] * row;
    }
    out[o_id] = p;
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
    Dev(Dev&& o) noexcept : p(o.p), n(o.n)
    {
        o.p = nullptr;
        o.n = 0;
    }
    ~Dev()
    {
        if (p != nullptr) {
            (void)hipFree(p);
        }
    }
    void alloc(std::size_t count)
    {
        if (p != nullptr) {
            (void)hipFree(p);
            p = nullptr;
        }
        n = count;
        hcheck(hipMalloc(reinterpret_cast<void**>(&p), std::max<std::size_t>(1, count) * sizeof(T)), "hipMalloc");
    }
    void upload(const std::vector<T>& v)
    {
        alloc(v.size());
        hcheck(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "upload");
    }
};

using RowKernel = void (*)(const HOpt*, const HRow*, const Task*, const float*, const float*, float*);

template <int K, int L>
RowKernel row_kernel(int mode)
{
    switch (mode) {
    case kFirst: return k_rows<K, L, kFirst>;
    case kCsSecond: return k_rows<K, L, kCsSecond>;
    default: return k_rows<K, L, kHvSecond>;
    }
}

RowKernel pick_row_kernel(int k, int l, int mode)
{
#define FE_ROWK(KV, LV) if (k == KV && l == LV) return row_kernel<KV, LV>(mode);
    FE_ROWK(2, 16) FE_ROWK(4, 16) FE_ROWK(8, 16) FE_ROWK(16, 16)
    FE_ROWK(2, 32) FE_ROWK(4, 32) FE_ROWK(8, 32) FE_ROWK(16, 32)
    FE_ROWK(2, 64) FE_ROWK(4, 64) FE_ROWK(8, 64) FE_ROWK(16, 64)
#undef FE_ROWK
    return nullptr;
}

} // namespace

struct fe_heston {
    struct Class {
        int nx = 0, ny_max = 0;
        std::vector<int> members;
        int lanes = 64;  // L: lanes a row
        int row_waves = 4;  // waves a block in the row passes
        int col_threads = 64; // threads a block in the column passes
        bool col_lds = false;  // the column passes keep d' in LDS (k_cols_lds; 64 threads a block, ny_max <= 256)
        int col_part = 0;      // M > 0: the column passes partitioned, M rows a thread (k_cols_part; 64 x P threads)
        Dev<Task> row_tasks, col_tasks;
        int n_row_tasks = 0, n_col_tasks = 0;
    };
    int n = 0;
    int steps = 0;
    int scheme = 0;
    double theta = 0.5;
    long long cells = 0;
    std::vector<HOpt> host_opts;
    std::vector<double> params; // 8 an option
    std::vector<Class> classes;
    Dev<HOpt> opts;
    Dev<HRow> rows;
    Dev<float> ptab; // the partitioned column solves' coefficients (part_table)
    Dev<Stencil> stencils;
    Dev<double> strikes;
    Dev<double> prices;
    Dev<float> U, D2, T1;
    fe_hip_graph graph = nullptr;
    std::string error;

    ~fe_heston()
    {
        if (graph != nullptr) {
            fe_hip_graph_destroy(graph);
        }
    }

    hipStream_t stream() const { return static_cast<hipStream_t>(fe_hip_stream()); }

    void build_tasks(Class& c)
    {
        std::vector<Task> rt, ct;
        const int rows_a_block = c.row_waves * (64 / c.lanes);
        for (const int o : c.members) {
            const int ny = host_opts[static_cast<std::size_t>(o)].ny;
            for (int j = 0; j < ny; j += rows_a_block) {
                rt.push_back({o, j});
            }
            for (int i = 0; i < c.nx; i += c.col_threads) {
                ct.push_back({o, i});
            }
        }
        c.row_tasks.upload(rt);
        c.col_tasks.upload(ct);
        c.n_row_tasks = static_cast<int>(rt.size());
        c.n_col_tasks = static_cast<int>(ct.size());
    }

    static unsigned lds_bytes(const Class& c) { return static_cast<unsigned>(c.ny_max) * 64u * 4u; }
    static bool lds_fits(const Class& c) { return c.ny_max >= 1 && c.ny_max <= 256; }
    // The partitioned column passes: P = ceil(ny_max / M) waves a block (at most 16: 1024 threads).
    static int part_waves(const Class& c, int m) { return (c.ny_max + m - 1) / m; }
    static bool part_fits(const Class& c, int m) { return (m == 8 || m == 16) && c.ny_max >= 1 && part_waves(c, m) <= 16; }
    void launch_cols_part(const Class& c, bool final_pass, hipStream_t s)
    {
        const int pw = part_waves(c, c.col_part);
        const dim3 block(64, static_cast<unsigned>(pw));
        const unsigned shm = 3u * static_cast<unsigned>(pw) * 64u * 4u;
        const dim3 grid(static_cast<unsigned>(c.n_col_tasks));
        if (c.col_part == 8) {
            if (final_pass) {
                hipLaunchKernelGGL((k_cols_part<8, true>), grid, block, shm, s, opts.p, ptab.p, c.col_tasks.p, U.p, D2.p, T1.p);
            } else {
                hipLaunchKernelGGL((k_cols_part<8, false>), grid, block, shm, s, opts.p, ptab.p, c.col_tasks.p, U.p, D2.p, T1.p);
            }
        } else if (final_pass) {
            hipLaunchKernelGGL((k_cols_part<16, true>), grid, block, shm, s, opts.p, ptab.p, c.col_tasks.p, U.p, D2.p, T1.p);
        } else {
            hipLaunchKernelGGL((k_cols_part<16, false>), grid, block, shm, s, opts.p, ptab.p, c.col_tasks.p, U.p, D2.p, T1.p);
        }
    }

    void queue_step() { queue_step_of(-1); }

    // One step of every class (cls < 0) or of one class only (the tuner's measurements).
    void queue_step_of(int cls)
    {
        const hipStream_t s = stream();
        const int second = scheme == 0 ? kCsSecond : kHvSecond;
        for (int pass = 0; pass < 4; ++pass) {
            for (std::size_t ci = 0; ci < classes.size(); ++ci) {
                if (cls >= 0 && static_cast<int>(ci) != cls) {
                    continue;
                }
                Class& c = classes[ci];
                if (pass == 0 || pass == 2) {
                    RowKernel k = pick_row_kernel(c.nx / c.lanes, c.lanes, pass == 0 ? kFirst : second);
                    // pass 3 reads D2 rows for the mixed term (CS) or U and D2 (HV)
                    const float* src = pass == 2 && scheme == 0 ? D2.p : U.p;
                    hipLaunchKernelGGL(k, dim3(static_cast<unsigned>(c.n_row_tasks)), dim3(64 * c.row_waves), 0, s, opts.p, rows.p,
                        c.row_tasks.p, src, D2.p, T1.p);
                } else if (c.col_part > 0) {
                    launch_cols_part(c, pass == 3, s);
                } else if (pass == 1) {
                    if (c.col_lds) {
                        hipLaunchKernelGGL(k_cols_lds<false>, dim3(static_cast<unsigned>(c.n_col_tasks)), dim3(64), lds_bytes(c), s, opts.p,
                            rows.p, c.col_tasks.p, U.p, D2.p, T1.p);
                    } else {
                        hipLaunchKernelGGL(k_cols<false>, dim3(static_cast<unsigned>(c.n_col_tasks)), dim3(c.col_threads), 0, s, opts.p,
                            rows.p, c.col_tasks.p, U.p, D2.p, T1.p);
                    }
                } else {
                    if (c.col_lds) {
                        hipLaunchKernelGGL(k_cols_lds<true>, dim3(static_cast<unsigned>(c.n_col_tasks)), dim3(64), lds_bytes(c), s, opts.p,
                            rows.p, c.col_tasks.p, U.p, D2.p, T1.p);
                    } else {
                        hipLaunchKernelGGL(k_cols<true>, dim3(static_cast<unsigned>(c.n_col_tasks)), dim3(c.col_threads), 0, s, opts.p,
                            rows.p, c.col_tasks.p, U.p, D2.p, T1.p);
                    }
                }
            }
        }
        hcheck(hipGetLastError(), "launch");
    }

    // Card time (ms) of `reps` steps of one class with its current shape, after one untimed step.
    float time_class(int cls, int reps)
    {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        hcheck(hipEventCreate(&e0), "event");
        hcheck(hipEventCreate(&e1), "event");
        queue_step_of(cls);
        hcheck(hipEventRecord(e0, stream()), "event");
        for (int r = 0; r < reps; ++r) {
            queue_step_of(cls);
        }
        hcheck(hipEventRecord(e1, stream()), "event");
        hcheck(hipEventSynchronize(e1), "event");
        float ms = 0.0f;
        hcheck(hipEventElapsedTime(&ms, e0, e1), "event");
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return ms;
    }

    // For each class: the row sweeps' split (lanes a row: Thomas cells a lane against PCR depth) and waves a block,
    // then the column sweeps' block size, each the fastest of its candidates on this batch; then every lattice back
    // to its payoff (the measurements stepped it).
    std::string tune(int reps)
    {
        std::ostringstream log;
        log << "[";
        for (std::size_t ci = 0; ci < classes.size(); ++ci) {
            Class& c = classes[ci];
            float best = 1e30f;
            int bl = c.lanes, bw = c.row_waves, bc = c.col_threads;
            for (const int lanes : {16, 32, 64}) {
                const int k = c.nx / lanes;
                if (c.nx % lanes != 0 || !(k == 2 || k == 4 || k == 8 || k == 16)) {
                    continue;
                }
                for (const int waves : {1, 2, 4}) {
                    c.lanes = lanes;
                    c.row_waves = waves;
                    build_tasks(c);
                    const float ms = time_class(static_cast<int>(ci), reps);
                    if (ms < best) {
                        best = ms;
                        bl = lanes;
                        bw = waves;
                    }
                }
            }
            c.lanes = bl;
            c.row_waves = bw;
            best = 1e30f;
            bool blds = false;
            int bpart = 0;
            // -64: 64 threads, d' in LDS; -8, -16: partitioned, 8 or 16 rows a thread
            for (const int threads : {64, 128, 256, -64, -8, -16}) {
                if ((threads == -64 && !lds_fits(c)) || ((threads == -8 || threads == -16) && !part_fits(c, -threads))) {
                    continue;
                }
                set_cols(c, threads);
                build_tasks(c);
                const float ms = time_class(static_cast<int>(ci), reps);
                if (ms < best) {
                    best = ms;
                    bc = c.col_threads;
                    blds = c.col_lds;
                    bpart = c.col_part;
                }
            }
            c.col_threads = bc;
            c.col_lds = blds;
            c.col_part = bpart;
            build_tasks(c);
            log << (ci ? "," : "") << "{\"nx\":" << c.nx << ",\"options\":" << c.members.size() << ",\"lanes\":" << c.lanes
                << ",\"row_waves\":" << c.row_waves << ",\"col_threads\":" << c.col_threads << ",\"col_lds\":" << (c.col_lds ? "true" : "false")
                << ",\"col_rows_a_thread\":" << c.col_part
                << ",\"ms_a_step\":" << best / reps << "}";
        }
        log << "]";
        reset();
        hcheck(hipStreamSynchronize(stream()), "sync");
        return log.str();
    }

    // The column passes' launch shape from its code: 64, 128 or 256 threads a column each; -64 d' in LDS; -8 or -16
    // partitioned with that many rows a thread.
    static void set_cols(Class& c, int code)
    {
        c.col_lds = code == -64;
        c.col_part = code == -8 || code == -16 ? -code : 0;
        c.col_threads = code > 0 ? code : 64;
    }
    static int cols_code(const Class& c)
    {
        return c.col_part > 0 ? -c.col_part : (c.col_lds ? -64 : c.col_threads);
    }

    void reset()
    {
        const unsigned gy = static_cast<unsigned>(n);
        hipLaunchKernelGGL(k_payoff, dim3(64, gy), dim3(256), 0, stream(), opts.p, strikes.p, n, U.p);
        hcheck(hipGetLastError(), "payoff");
    }
};

namespace {

int fail(fe_heston* b, const std::string& what)
{
    if (b != nullptr) {
        b->error = what;
    }
    return -1;
}

void set_why(char* why, std::size_t bytes, const std::string& text)
{
    if (why != nullptr && bytes > 0) {
        std::snprintf(why, bytes, "%s", text.c_str());
    }
}

} // namespace

extern "C" {

fe_heston* fe_heston_create(int n, const int* nx, const int* ny, const double* options, int steps, int scheme, double theta, char* why,
    size_t why_bytes)
{
    try {
        if (n <= 0 || n > 65535 || nx == nullptr || ny == nullptr || options == nullptr || steps <= 0 || (scheme != 0 && scheme != 1) ||
            !(theta > 0.0)) {
            set_why(why, why_bytes, "invalid arguments");
            return nullptr;
        }
        fe_hip_device_info info{};
        if (fe_hip_device_open(&info) != 0) {
            set_why(why, why_bytes, std::string("no HIP device: ") + fe_hip_unavailable());
            return nullptr;
        }
        auto b = std::make_unique<fe_heston>();
        b->n = n;
        b->steps = steps;
        b->scheme = scheme;
        b->theta = theta;
        b->params.assign(options, options + 8 * static_cast<std::size_t>(n));
        std::vector<HRow> rows;
        std::vector<Stencil> st(static_cast<std::size_t>(n));
        std::vector<double> strikes(static_cast<std::size_t>(n));
        std::map<int, int> class_of; // nx -> class index
        long long off = 0;
        for (int o = 0; o < n; ++o) {
            const int X = nx[o], Y = ny[o];
            const double* q = options + 8 * static_cast<std::size_t>(o);
            const double kappa = q[0], theta_v = q[1], sigma = q[2], rho = q[3], r = q[4], strike = q[5], mat = q[6], v0 = q[7];
            bool shape_ok = false;
            for (const int l : {64, 32, 16}) {
                const int k = X / l;
                shape_ok = shape_ok || (X % l == 0 && (k == 2 || k == 4 || k == 8 || k == 16));
            }
            if (!shape_ok || Y < 4) {
                set_why(why, why_bytes, "option " + std::to_string(o) + ": lattice " + std::to_string(X) + "x" + std::to_string(Y) +
                    " (nx must be 32..1024, a power of two; ny >= 4)");
                return nullptr;
            }
            if (!(sigma > 0.0) || !(mat > 0.0) || !(strike > 0.0)) {
                set_why(why, why_bytes, "option " + std::to_string(o) + ": sigma, maturity and strike must be positive");
                return nullptr;
            }
            const double h = 6.0 / X;
            const double dt = mat / steps;
            const double thdt = theta * dt;
            HOpt ho{};
            ho.off = off;
            ho.nx = X;
            ho.ny = Y;
            ho.row = static_cast<int>(rows.size());
            ho.h = static_cast<float>(h);
            ho.dt = static_cast<float>(dt);
            ho.thdt = static_cast<float>(thdt);
            ho.r = static_cast<float>(r);
            b->host_opts.push_back(ho);
            off += static_cast<long long>(X) * Y;
            // Rows: the operator coefficients (as the oracle's grid) and (I - th dt A2) factorised in double.
            std::vector<double> lyv(static_cast<std::size_t>(Y)), uyv(static_cast<std::size_t>(Y));
            double cprev = 0.0;
            for (int j = 0; j < Y; ++j) {
                const double v = sigma * (j + 0.5) * h;
                const double e0 = v < sigma * h ? 1.0 : 0.0;
                const double bb = kappa * (theta_v - v) / sigma - 0.5 * sigma;
                HRow hr{};
                const double lx = 0.5 * v / (h * h) - (r - 0.5 * v) / (2 * h);
                const double ux = 0.5 * v / (h * h) + (r - 0.5 * v) / (2 * h);
                const double ly = 0.5 * (v - 0.5 * sigma * h) / (h * h) - (1 - e0) * bb / (2 * h);
                const double uy = 0.5 * (v + 0.5 * sigma * h) / (h * h) + (1 + e0) * bb / (2 * h);
                hr.lx = static_cast<float>(lx);
                hr.ux = static_cast<float>(ux);
                hr.ly = static_cast<float>(ly);
                hr.uy = static_cast<float>(uy);
                hr.rv = static_cast<float>(rho * v);
                const double a = -thdt * ly, c = -thdt * uy;
                const double cy = -ly - uy - 0.5 * r;
                const double bd = 1.0 - thdt * cy + (j == 0 ? a : 0.0) + (j + 1 == Y ? c : 0.0);
                const double piv = j == 0 ? bd : bd - a * cprev;
                const double cp = j + 1 == Y ? 0.0 : c / piv;
                hr.ya = static_cast<float>(a);
                hr.ycp = static_cast<float>(cp);
                hr.ym = static_cast<float>(1.0 / piv);
                hr.yb = static_cast<float>(bd);
                hr.yc = static_cast<float>(j + 1 == Y ? 0.0 : c);
                cprev = cp;
                rows.push_back(hr);
            }
            // Price stencil: 4-point Lagrange at ln 100 along x and at v0 along y.
            Stencil s{};
            const double x0 = std::log(100.0) - 3.0 + 0.5 * h;
            const double tx = (std::log(100.0) - x0) / h;
            s.i0 = std::clamp(static_cast<int>(std::floor(tx)) - 1, 0, X - 4);
            const double ty = (v0 / sigma) / h - 0.5;
            s.j0 = std::clamp(static_cast<int>(std::floor(ty)) - 1, 0, Y - 4);
            auto lag = [](double t, double* w) {
                w[0] = -(t - 1) * (t - 2) * (t - 3) / 6.0;
                w[1] = t * (t - 2) * (t - 3) / 2.0;
                w[2] = -t * (t - 1) * (t - 3) / 2.0;
                w[3] = t * (t - 1) * (t - 2) / 6.0;
            };
            lag(tx - s.i0, s.wx);
            lag(ty - s.j0, s.wy);
            st[static_cast<std::size_t>(o)] = s;
            strikes[static_cast<std::size_t>(o)] = strike;
            auto it = class_of.find(X);
            if (it == class_of.end()) {
                it = class_of.emplace(X, static_cast<int>(b->classes.size())).first;
                b->classes.emplace_back();
                b->classes.back().nx = X;
            }
            fe_heston::Class& cl = b->classes[static_cast<std::size_t>(it->second)];
            cl.members.push_back(o);
            cl.ny_max = std::max(cl.ny_max, Y);
        }
        b->cells = off;
        std::vector<float> ptab;
        for (fe_heston::Class& c : b->classes) {
            for (const int o : c.members) {
                HOpt& ho = b->host_opts[static_cast<std::size_t>(o)];
                const HRow* cf = rows.data() + ho.row;
                for (const int m : {8, 16}) {
                    int& at = m == 8 ? ho.part8 : ho.part16;
                    at = -1;
                    if (fe_heston::part_fits(c, m)) {
                        at = static_cast<int>(ptab.size());
                        part_table(cf, ho.ny, fe_heston::part_waves(c, m), m, ptab);
                    }
                }
            }
        }
        b->opts.upload(b->host_opts);
        b->rows.upload(rows);
        b->ptab.upload(ptab.empty() ? std::vector<float>(1, 0.0f) : ptab);
        b->stencils.upload(st);
        b->strikes.upload(strikes);
        b->prices.alloc(static_cast<std::size_t>(n));
        b->U.alloc(static_cast<std::size_t>(off));
        b->D2.alloc(static_cast<std::size_t>(off));
        b->T1.alloc(static_cast<std::size_t>(off));
        for (fe_heston::Class& c : b->classes) {
            // default split: four cells a lane at most 64 lanes (the tuner may change it)
            c.lanes = std::min(64, std::max(16, c.nx / 4));
            while (c.nx / c.lanes > 16) {
                c.lanes *= 2;
            }
            b->build_tasks(c);
        }
        b->reset();
        hcheck(hipStreamSynchronize(b->stream()), "sync");
        return b.release();
    } catch (const std::exception& e) {
        set_why(why, why_bytes, e.what());
        return nullptr;
    }
}

void fe_heston_free(fe_heston* b) { delete b; }

const char* fe_heston_error(fe_heston* b) { return b != nullptr ? b->error.c_str() : "no batch"; }

long long fe_heston_cells(fe_heston* b) { return b != nullptr ? b->cells : 0; }

int fe_heston_reset(fe_heston* b)
{
    if (b == nullptr) {
        return -1;
    }
    try {
        b->reset();
        return 0;
    } catch (const std::exception& e) {
        return fail(b, e.what());
    }
}

int fe_heston_step(fe_heston* b, int count)
{
    if (b == nullptr || count < 0) {
        return -1;
    }
    try {
        for (int s = 0; s < count; ++s) {
            b->queue_step();
        }
        return 0;
    } catch (const std::exception& e) {
        return fail(b, e.what());
    }
}

int fe_heston_capture(fe_heston* b, int count)
{
    if (b == nullptr || count <= 0) {
        return -1;
    }
    try {
        if (b->graph != nullptr) {
            fe_hip_graph_destroy(b->graph);
            b->graph = nullptr;
        }
        hcheck(hipStreamSynchronize(b->stream()), "sync");
        if (fe_hip_graph_begin_capture() != 0) {
            return fail(b, "graph capture did not start");
        }
        try {
            for (int s = 0; s < count; ++s) {
                b->queue_step();
            }
        } catch (...) {
            fe_hip_graph graph = nullptr;
            fe_hip_graph_end_capture(&graph);
            if (graph != nullptr) {
                fe_hip_graph_destroy(graph);
            }
            throw;
        }
        const int r = fe_hip_graph_end_capture(&b->graph);
        if (r != 0) {
            b->graph = nullptr;
            return fail(b, std::string("graph capture: ") + fe_hip_error_string(r));
        }
        return 0;
    } catch (const std::exception& e) {
        return fail(b, e.what());
    }
}

int fe_heston_launch(fe_heston* b)
{
    if (b == nullptr || b->graph == nullptr) {
        return fail(b, "no captured graph");
    }
    const int r = fe_hip_graph_launch(b->graph);
    return r == 0 ? 0 : fail(b, std::string("graph launch: ") + fe_hip_error_string(r));
}

int fe_heston_sync(fe_heston* b)
{
    if (b == nullptr) {
        return -1;
    }
    const hipError_t e = hipStreamSynchronize(b->stream());
    return e == hipSuccess ? 0 : fail(b, hipGetErrorString(e));
}

int fe_heston_prices(fe_heston* b, double* out)
{
    if (b == nullptr || out == nullptr) {
        return -1;
    }
    try {
        hipLaunchKernelGGL(k_price, dim3(static_cast<unsigned>((b->n + 255) / 256)), dim3(256), 0, b->stream(), b->opts.p, b->stencils.p,
            b->n, b->U.p, b->prices.p);
        hcheck(hipGetLastError(), "price");
        hcheck(hipMemcpyAsync(out, b->prices.p, sizeof(double) * static_cast<std::size_t>(b->n), hipMemcpyDeviceToHost, b->stream()),
            "prices");
        hcheck(hipStreamSynchronize(b->stream()), "sync");
        return 0;
    } catch (const std::exception& e) {
        return fail(b, e.what());
    }
}

int fe_heston_lattice(fe_heston* b, int option, float* out)
{
    if (b == nullptr || out == nullptr || option < 0 || option >= b->n) {
        return -1;
    }
    const HOpt& o = b->host_opts[static_cast<std::size_t>(option)];
    const hipError_t e = hipMemcpyAsync(out, b->U.p + o.off, sizeof(float) * static_cast<std::size_t>(o.nx) * o.ny, hipMemcpyDeviceToHost,
        b->stream());
    if (e != hipSuccess) {
        return fail(b, hipGetErrorString(e));
    }
    return hipStreamSynchronize(b->stream()) == hipSuccess ? 0 : fail(b, "sync");
}

int fe_heston_tune(fe_heston* b, int reps, char* report, size_t report_bytes)
{
    if (b == nullptr || reps <= 0) {
        return -1;
    }
    try {
        if (b->graph != nullptr) {
            fe_hip_graph_destroy(b->graph);
            b->graph = nullptr;
        }
        const std::string r = b->tune(reps);
        set_why(report, report_bytes, r);
        return 0;
    } catch (const std::exception& e) {
        return fail(b, e.what());
    }
}

int fe_heston_classes(fe_heston* b) { return b != nullptr ? static_cast<int>(b->classes.size()) : 0; }

int fe_heston_get_shape(fe_heston* b, int cls, int* nx, int* lanes, int* row_waves, int* col_threads)
{
    if (b == nullptr || cls < 0 || cls >= static_cast<int>(b->classes.size())) {
        return -1;
    }
    const fe_heston::Class& c = b->classes[static_cast<std::size_t>(cls)];
    *nx = c.nx;
    *lanes = c.lanes;
    *row_waves = c.row_waves;
    *col_threads = fe_heston::cols_code(c);
    return 0;
}

int fe_heston_set_shape(fe_heston* b, int cls, int lanes, int row_waves, int col_threads)
{
    if (b == nullptr || cls < 0 || cls >= static_cast<int>(b->classes.size())) {
        return -1;
    }
    fe_heston::Class& c = b->classes[static_cast<std::size_t>(cls)];
    const int k = lanes > 0 ? c.nx / lanes : 0;
    if (!(lanes == 16 || lanes == 32 || lanes == 64) || c.nx % lanes != 0 || !(k == 2 || k == 4 || k == 8 || k == 16) ||
        !(row_waves == 1 || row_waves == 2 || row_waves == 4) ||
        !(col_threads == 64 || col_threads == 128 || col_threads == 256 || (col_threads == -64 && fe_heston::lds_fits(c)) ||
            ((col_threads == -8 || col_threads == -16) && fe_heston::part_fits(c, -col_threads)))) {
        return fail(b, "shape out of range");
    }
    try {
        if (b->graph != nullptr) {
            fe_hip_graph_destroy(b->graph);
            b->graph = nullptr;
        }
        hcheck(hipStreamSynchronize(b->stream()), "sync");
        c.lanes = lanes;
        c.row_waves = row_waves;
        fe_heston::set_cols(c, col_threads);
        b->build_tasks(c);
        return 0;
    } catch (const std::exception& e) {
        return fail(b, e.what());
    }
}

} // extern "C"
