// A world model's MPPI planner on gfx950: rollouts of a learned ODE on the matrix cores and the softmin plan update.
//
// The reference plans with one thread a future, the dynamics MLP evaluated in fp32 registers with the weights in shared
// memory (src/field/mind/WorldModel.cpp:192-240). Here a wave carries 16 futures at a time through the whole horizon:
// their states in LDS (fp32), each step's inputs (state and action, normalised) rounded to fp16 into an LDS tile, and
// the three layers as chains of v_mfma_f32_16x16x16_f16 (A: 16 futures x 16 inputs from the tile, lane l future l % 16
// and inputs 4 (l / 16)..+3; B: the fp16 weights W[out][in], output l % 16; D: lane l output l % 16, futures
// 4 (l / 16)..+3), fp32 accumulation, tanh on D into the other tile. The last layer's rates move the states in fp32.
// The update weighs every future by exp(-(J - J_min) / lambda), one thread an (step, axis), summing the futures in
// order: the same inputs give the same plan, bit for bit.
#include "fe_mppi.h"
#include "fe_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

typedef _Float16 half4 __attribute__((ext_vector_type(4)));
typedef float float4v __attribute__((ext_vector_type(4)));
typedef unsigned long long u64;

constexpr int kMaxWidth = 128;        // padded inputs and hidden width
constexpr int kMaxState = 64;         // features, padded

struct Desc {
    int F, in_pad, hid, f_pad, horizon, samples;
    const _Float16* W1; // [hid][in_pad]
    const _Float16* W2; // [hid][hid]
    const _Float16* W3; // [f_pad][hid]
    const float* b1;    // [hid]
    const float* b2;    // [hid]
    const float* b3;    // [f_pad]
    const float* shift; // [in_pad]
    const float* inscale; // [in_pad]
    const float* scale;   // [F]
    const float* target;  // [F]
    const float* weight;  // [F]
    float action_weight;
};

// The noise of (agent id, tick, future, step, axis): a standard normal by Box-Muller from two uniforms, each the top 24
// bits of a 64-bit mix (splitmix64's finaliser, Steele, Lea and Flood 2014) of the key and of the key's mix.
__host__ __device__ inline u64 mix64(u64 h)
{
    h ^= h >> 30;
    h *= 0xbf58476d1ce4e5b9ull;
    h ^= h >> 27;
    h *= 0x94d049bb133111ebull;
    h ^= h >> 31;
    return h;
}
// The part of a key that is the same for every future and step of one agent's plan; the key of (future k, step t,
// axis d) is noise_base ^ (k << 24) ^ (t << 8) ^ d.
__host__ __device__ inline u64 noise_base(u64 id, u64 tick)
{
    return mix64(id * 0x9e3779b97f4a7c15ull + 1) ^ mix64(tick * 0xd1b54a32d192ed03ull + 2);
}
// The hardware's logarithm, square root and cosine (v_log_f32, v_sqrt_f32, v_cos_f32, the last taking turns, so
// cos(2 pi u2) is one instruction): a few ulp, where the library's range reductions were a large share of a step.
__device__ inline float gauss(u64 key)
{
    const u64 a = mix64(key), b = mix64(a ^ 0x632be59bd9b4e019ull);
    const float u1 = ((float)(a >> 40) + 0.5f) * (1.0f / 16777216.0f);
    const float u2 = ((float)(b >> 40) + 0.5f) * (1.0f / 16777216.0f);
    return __builtin_amdgcn_sqrtf(-2.0f * 0.6931471805599453f * __builtin_amdgcn_logf(u1)) * __builtin_amdgcn_cosf(u2);
}

// tanh u = 1 - 2 / (1 + e^(2u)) through the hardware exponential and reciprocal: a few ulp of fp32, below the fp16
// rounding of the activation it becomes; the library tanhf's long polynomial sequence was most of a layer's issue slots.
// Saturates to +-1 (e^(2u) overflowing to inf, or to 0).
__device__ __forceinline__ float fast_tanh(float u)
{
    const float e = __builtin_amdgcn_exp2f((2.0f * 1.4426950408889634f) * u); // e^(2u), 2 log2(e) folded
    return 1.0f - 2.0f * __builtin_amdgcn_rcpf(1.0f + e);
}

// The block's LDS, sized to the net (a block is one wave): two fp16 activation tiles of 16 rows at a stride of ts
// halves (the wider of the padded inputs and the hidden width, plus 8 so the 16 rows a product reads start in
// different banks), the 16 futures' states at a stride of ss floats (odd), and their actions for two steps.
struct Lds {
    int ts, ss;
    __host__ __device__ static Lds of(const Desc& d)
    {
        return Lds{(d.in_pad > d.hid ? d.in_pad : d.hid) + 8, d.F | 1};
    }
    __host__ __device__ int bytes() const { return 2 * 16 * ts * 2 + 16 * ss * 4 + 2 * 16 * 2 * 4; }
};

// One layer for a wave's 16 futures: in (K halves a row) -> out (N outputs), tanh or (last) the rates into the states.
__device__ __forceinline__ void layer(const Desc& d, const _Float16* in, _Float16* out, int ts, int K, int N, const _Float16* W,
    const float* b, bool last, float* st, int ss, float dt, int lane)
{
    const int r = lane & 15, g = lane >> 4;
    for (int nb = 0; nb < N; nb += 16) {
        float4v acc = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int kc = 0; kc < K; kc += 16) {
            const half4 a = *reinterpret_cast<const half4*>(in + r * ts + kc + 4 * g);
            const half4 w = *reinterpret_cast<const half4*>(W + (long long)(nb + r) * K + kc + 4 * g);
            acc = __builtin_amdgcn_mfma_f32_16x16x16f16(a, w, acc, 0, 0, 0);
        }
        const int col = nb + r;
        const float bias = b[col];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int row = 4 * g + j;
            const float v = acc[j] + bias;
            if (!last) {
                out[row * ts + col] = (_Float16)fast_tanh(v);
            } else if (col < d.F) {
                st[row * ss + col] += dt * v * d.scale[col];
            }
        }
    }
}

// A block (agent, part), a wave a tile of 16 futures at a time: part p takes tiles 16 p, then every 16 * parts futures
// on. Actions into acts[agent][k][t][axis], each future's cost into cost[agent][k]. LDS: Lds::of(d).bytes(). The noise
// of two steps is drawn at once (lane l: future l / 2 % 16, axis l % 2, step t + l / 32), so every lane draws.
__global__ __launch_bounds__(64) void k_mppi_rollout(Desc d, const float* __restrict__ obs, const float* __restrict__ mean,
    const unsigned* __restrict__ ids, float* __restrict__ acts, float* __restrict__ cost, u64 tick, float sigma, float dt)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
    const Lds L = Lds::of(d);
    const int ts = L.ts, ss = L.ss;
    _Float16* tile0 = reinterpret_cast<_Float16*>(lds);
    _Float16* tile1 = tile0 + 16 * ts;
    float* st = reinterpret_cast<float*>(tile1 + 16 * ts);
    float* act = st + 16 * ss; // [2 steps][16 futures][2 axes]
    const int lane = (int)threadIdx.x, r16 = lane & 15, g4 = lane >> 4;
    const int agent = (int)blockIdx.x;
    const int F = d.F, H = d.horizon, K = d.samples;
    const u64 base = noise_base(ids[agent], tick);
    for (int t0 = (int)blockIdx.y * 16; t0 < K; t0 += 16 * (int)gridDim.y) {
        for (int q = lane; q < 16 * F; q += 64) {
            st[(q / F) * ss + q % F] = obs[(long long)agent * F + q % F];
        }
        float J = 0.0f; // lanes 0..15: future t0 + lane
        for (int t = 0; t < H; ++t) {
            if ((t & 1) == 0) {
                const int row = (lane >> 1) & 15, axis = lane & 1, step = t + (lane >> 5), k = t0 + row;
                if (step < H) {
                    const float e = k == 0 ? 0.0f : gauss(base ^ ((u64)k << 24) ^ ((u64)step << 8) ^ (u64)axis);
                    const float a = fminf(fmaxf(mean[((long long)agent * H + step) * 2 + axis] + sigma * e, -1.0f), 1.0f);
                    act[lane] = a;
                    acts[(((long long)agent * K + k) * H + step) * 2 + axis] = a;
                }
            }
            __syncthreads();
            const float* at = act + 32 * (t & 1);
            // the inputs (state, then action), normalised, in fp16; zero past them: lane l row l % 16, features l / 16,
            // l / 16 + 4, ...
            for (int f = g4; f < d.in_pad; f += 4) {
                const float x = f < F ? st[r16 * ss + f] : (f < F + 2 ? at[r16 * 2 + f - F] : 0.0f);
                tile0[r16 * ts + f] = (_Float16)(f < F + 2 ? (x - d.shift[f]) * d.inscale[f] : 0.0f);
            }
            __syncthreads();
            layer(d, tile0, tile1, ts, d.in_pad, d.hid, d.W1, d.b1, false, st, ss, dt, lane);
            __syncthreads();
            layer(d, tile1, tile0, ts, d.hid, d.hid, d.W2, d.b2, false, st, ss, dt, lane);
            __syncthreads();
            layer(d, tile0, tile1, ts, d.hid, d.f_pad, d.W3, d.b3, true, st, ss, dt, lane);
            __syncthreads();
            // the step's cost: lane l the features l / 16, l / 16 + 4, ... of future l % 16, then the four lane groups
            // added in a fixed order
            float c = 0.0f;
            for (int f = g4; f < F; f += 4) {
                const float e = st[r16 * ss + f] - d.target[f];
                c += d.weight[f] * e * e;
            }
            c += __shfl_xor(c, 16);
            c += __shfl_xor(c, 32);
            if (lane < 16) {
                const float ax = at[lane * 2], ay = at[lane * 2 + 1];
                J += c + d.action_weight * (ax * ax + ay * ay);
            }
            __syncthreads();
        }
        if (lane < 16) {
            cost[(long long)agent * K + t0 + lane] = J;
        }
    }
}

// The new plan: a thread an (step, axis); the futures' actions weighted by exp(-(J - J_min) / lambda), lambda a quarter
// of (mean cost - least cost), at least 1e-4, summed in the futures' order. The least cost, the sum of costs (in order)
// and each future's weight are found once a block (in LDS) instead of by every thread.
constexpr int kMaxSamples = 1024;
__global__ void k_mppi_update(const float* __restrict__ cost, const float* __restrict__ acts, float* __restrict__ mean, int K, int H)
{
    __shared__ float cs[kMaxSamples];
    __shared__ float lo_s, lam_s;
    const int agent = (int)blockIdx.x, q = (int)threadIdx.x;
    const float* c = cost + (long long)agent * K;
    for (int k = q; k < K; k += (int)blockDim.x) {
        cs[k] = c[k];
    }
    __syncthreads();
    if (q == 0) {
        float lo = cs[0], sum = 0.0f;
#pragma unroll 8
        for (int k = 0; k < K; ++k) { // eight reads in flight, added in order
            lo = fminf(lo, cs[k]);
            sum += cs[k];
        }
        lo_s = lo;
        lam_s = fmaxf(1e-4f, 0.25f * (sum / (float)K - lo));
    }
    __syncthreads();
    const float lo = lo_s, lam = lam_s;
    for (int k = q; k < K; k += (int)blockDim.x) {
        cs[k] = expf(-(cs[k] - lo) / lam);
    }
    __syncthreads();
    if (q >= 2 * H) {
        return;
    }
    float wsum = 0.0f, acc = 0.0f;
#pragma unroll 8
    for (int k = 0; k < K; ++k) { // eight loads in flight, added in order
        const float wk = cs[k];
        wsum += wk;
        acc += wk * acts[((long long)agent * K + k) * 2 * H + q];
    }
    mean[(long long)agent * 2 * H + q] = acc / wsum;
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
    void put(const T* host, std::size_t count)
    {
        alloc(count);
        hcheck(hipMemcpy(p, host, count * sizeof(T), hipMemcpyHostToDevice), "upload");
    }
};

int pad16(int v) { return (v + 15) / 16 * 16; }

} // namespace

struct fe_mppi {
    int F = 0, hid = 0, horizon = 0, samples = 0, in_pad = 0, f_pad = 0, cus = 0;
    bool dynamics = false, costed = false;
    Dev<_Float16> W1, W2, W3;
    Dev<float> b1, b2, b3, shift, inscale, scale, target, weight;
    float action_weight = 0.0f;
    Dev<float> obs, mean, acts, cost;
    Dev<unsigned> ids;
    std::string error;
};

namespace {

// Blocks an agent's futures spread over: one tile of 16 a block while every tile's block fits the card at once; past
// that each block walks a few tiles in turn, so the launch is one round (a second, part-filled round of one-tile
// blocks took longer than the first: 824 against 777 us for 256 agents x 512 futures). A CU holds as many blocks as
// its 160 KB of LDS allows, at most 28 (seven waves a SIMD at the rollout's registers).
int rollout_parts(const fe_mppi& p, const Desc& d, int agents)
{
    const int tiles = p.samples / 16;
    const int per_cu = std::min(160 * 1024 / Lds::of(d).bytes(), 28);
    const long long fit = static_cast<long long>(std::max(p.cus, 1)) * per_cu;
    const long long rounds = (static_cast<long long>(agents) * tiles + fit - 1) / fit;
    return static_cast<int>((tiles + rounds - 1) / rounds);
}

int fail(fe_mppi* p, const std::string& what)
{
    if (p != nullptr) {
        p->error = what;
    }
    return -1;
}

// W[out][in] (row-major, fp32) into a zero-padded fp16 [out_pad][in_pad].
std::vector<_Float16> pack(const float* w, int out, int in, int out_pad, int in_pad)
{
    std::vector<_Float16> h(static_cast<std::size_t>(out_pad) * in_pad, static_cast<_Float16>(0.0f));
    for (int o = 0; o < out; ++o) {
        for (int i = 0; i < in; ++i) {
            h[static_cast<std::size_t>(o) * in_pad + i] = static_cast<_Float16>(w[static_cast<std::size_t>(o) * in + i]);
        }
    }
    return h;
}

std::vector<float> padded(const float* v, int n, int to, float fill = 0.0f)
{
    std::vector<float> out(static_cast<std::size_t>(to), fill);
    std::copy(v, v + n, out.begin());
    return out;
}

} // namespace

extern "C" {

fe_mppi* fe_mppi_create(int features, int hidden, int horizon, int samples, char* why, size_t why_bytes)
{
    auto say = [&](const std::string& s) {
        if (why != nullptr && why_bytes > 0) {
            std::snprintf(why, why_bytes, "%s", s.c_str());
        }
    };
    if (features < 1 || features + 2 > kMaxState || hidden < 16 || hidden > kMaxWidth || hidden % 16 != 0 || horizon < 1 || horizon > 64 ||
        samples < 16 || samples > 1024 || samples % 16 != 0) {
        say("features 1..62, hidden 16..128 (a multiple of 16), horizon 1..64, samples 16..1024 (a multiple of 16)");
        return nullptr;
    }
    fe_hip_device_info info{};
    if (fe_hip_device_open(&info) != 0) {
        say(std::string("no HIP device: ") + fe_hip_unavailable());
        return nullptr;
    }
    auto p = std::make_unique<fe_mppi>();
    p->cus = info.compute_units;
    p->F = features;
    p->hid = hidden;
    p->horizon = horizon;
    p->samples = samples;
    p->in_pad = pad16(features + 2);
    p->f_pad = pad16(features);
    return p.release();
}

void fe_mppi_free(fe_mppi* p) { delete p; }
const char* fe_mppi_error(fe_mppi* p) { return p != nullptr ? p->error.c_str() : "no planner"; }

int fe_mppi_set_dynamics(fe_mppi* p, const float* weights, const float* in_shift, const float* in_scale, const float* scale)
{
    if (p == nullptr || weights == nullptr || in_shift == nullptr || in_scale == nullptr || scale == nullptr) {
        return fail(p, "invalid arguments");
    }
    try {
        const int F = p->F, I = F + 2, Hd = p->hid;
        const float* w1 = weights;
        const float* b1 = w1 + static_cast<std::size_t>(I) * Hd;
        const float* w2 = b1 + Hd;
        const float* b2 = w2 + static_cast<std::size_t>(Hd) * Hd;
        const float* w3 = b2 + Hd;
        const float* b3 = w3 + static_cast<std::size_t>(Hd) * F;
        const std::vector<_Float16> h1 = pack(w1, Hd, I, Hd, p->in_pad), h2 = pack(w2, Hd, Hd, Hd, Hd), h3 = pack(w3, F, Hd, p->f_pad, Hd);
        p->W1.put(h1.data(), h1.size());
        p->W2.put(h2.data(), h2.size());
        p->W3.put(h3.data(), h3.size());
        const std::vector<float> pb3 = padded(b3, F, p->f_pad), sh = padded(in_shift, I, p->in_pad), sc = padded(in_scale, I, p->in_pad);
        p->b1.put(b1, static_cast<std::size_t>(Hd));
        p->b2.put(b2, static_cast<std::size_t>(Hd));
        p->b3.put(pb3.data(), pb3.size());
        p->shift.put(sh.data(), sh.size());
        p->inscale.put(sc.data(), sc.size());
        p->scale.put(scale, static_cast<std::size_t>(F));
        p->dynamics = true;
        return 0;
    } catch (const std::exception& e) {
        return fail(p, e.what());
    }
}

int fe_mppi_set_cost(fe_mppi* p, const float* target, const float* weight, float action_weight)
{
    if (p == nullptr || target == nullptr || weight == nullptr) {
        return fail(p, "invalid arguments");
    }
    try {
        p->target.put(target, static_cast<std::size_t>(p->F));
        p->weight.put(weight, static_cast<std::size_t>(p->F));
        p->action_weight = action_weight;
        p->costed = true;
        return 0;
    } catch (const std::exception& e) {
        return fail(p, e.what());
    }
}

int fe_mppi_plan(fe_mppi* p, int agents, const float* obs, const uint32_t* ids, uint64_t tick, float sigma, float dt, float* mean,
    float* cost_out)
{
    if (p == nullptr || agents < 1 || obs == nullptr || ids == nullptr || mean == nullptr) {
        return fail(p, "invalid arguments");
    }
    if (!p->dynamics || !p->costed) {
        return fail(p, "set the dynamics and the cost first");
    }
    try {
        const int F = p->F, H = p->horizon, K = p->samples;
        const hipStream_t s = static_cast<hipStream_t>(fe_hip_stream());
        p->obs.put(obs, static_cast<std::size_t>(agents) * F);
        p->ids.put(ids, static_cast<std::size_t>(agents));
        p->mean.put(mean, static_cast<std::size_t>(agents) * H * 2);
        p->acts.alloc(static_cast<std::size_t>(agents) * K * H * 2);
        p->cost.alloc(static_cast<std::size_t>(agents) * K);
        Desc d{F, p->in_pad, p->hid, p->f_pad, H, K, p->W1.p, p->W2.p, p->W3.p, p->b1.p, p->b2.p, p->b3.p, p->shift.p, p->inscale.p,
            p->scale.p, p->target.p, p->weight.p, p->action_weight};
        // a block is one wave (its barriers wait for no other wave), an agent's futures spread over samples / 16 blocks
        hipLaunchKernelGGL(k_mppi_rollout, dim3(static_cast<unsigned>(agents), static_cast<unsigned>(rollout_parts(*p, d, agents))), dim3(64),
            static_cast<unsigned>(Lds::of(d).bytes()), s,
            d, p->obs.p, p->mean.p, p->ids.p,
            p->acts.p, p->cost.p, static_cast<unsigned long long>(tick), sigma, dt);
        hipLaunchKernelGGL(k_mppi_update, dim3(static_cast<unsigned>(agents)), dim3(static_cast<unsigned>(pad16(2 * H))), 0, s, p->cost.p,
            p->acts.p, p->mean.p, K, H);
        hcheck(hipGetLastError(), "mppi launch");
        hcheck(hipMemcpyAsync(mean, p->mean.p, sizeof(float) * static_cast<std::size_t>(agents) * H * 2, hipMemcpyDeviceToHost, s), "download");
        if (cost_out != nullptr) {
            hcheck(hipMemcpyAsync(cost_out, p->cost.p, sizeof(float) * static_cast<std::size_t>(agents) * K, hipMemcpyDeviceToHost, s), "download");
        }
        hcheck(hipStreamSynchronize(s), "sync");
        return 0;
    } catch (const std::exception& e) {
        return fail(p, e.what());
    }
}

} // extern "C"
