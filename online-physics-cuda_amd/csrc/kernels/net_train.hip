// A law net's trainer on gfx950 (include/fe_nettrain.h): the loss and gradient of a regression over rows, and a
// clipped Adam step.
//
// A block of four waves takes tiles of 16 rows. A tile's inputs are normalised into LDS; each layer's product runs on
// the matrix cores in fp32 (v_mfma_f32_16x16x4f32: A 16 x 4, lane l holds A[l % 16][l / 16]; B 4 x 16, lane l holds
// B[l / 16][l % 16]; D 16 x 16, lane l holds D[4 (l / 16) + j][l % 16], j = 0..3), its 16-column blocks spread over the
// waves, the pre-activations and activations kept in LDS for the way back. The output layer gives each row's error
// e = out_scale * y - t and delta m e out_scale; going back, each layer's weight gradient (deltas^T x inputs, the 16 rows
// as the product's depth) and bias gradient (the rows summed in order) are added into the block's own slot of a partial
// array, and the deltas move down through W^T and the activation's derivative. A block walks its tiles in a fixed order
// and the blocks' partials are summed in a fixed order (in one level up to 64 blocks, past that in chunks of 32 first),
// so the same rows give the same bits; the last level's pass also takes the Adam step.
#include "fe_nettrain.h"
#include "fe_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

typedef float float4v __attribute__((ext_vector_type(4)));

constexpr int kMaxLayers = 8;   // weight matrices
constexpr int kMaxWidth = 128;  // any layer's width
constexpr int kRows = 16;       // rows a tile (a wave)
constexpr int kChunk = 32;      // blocks' partials summed together in the first level (past kOneLevel blocks)
constexpr int kOneLevel = 64;   // up to this many blocks their partials are added in one level
constexpr int kBlocks = 1024;   // the most blocks a pass (four blocks of four waves a CU)
constexpr int kMaxLds = 65536;  // bytes of LDS a block may use
constexpr int kWaves = 4;       // waves a block: a tile's 16-column blocks spread over them

struct Shape {
    int layers = 0;
    int sizes[kMaxLayers + 1] = {};
    int act = 0;
    int w_at[kMaxLayers] = {};  // layer k's W in the image (its b follows W)
    int a_at[kMaxLayers] = {};  // LDS floats: layer k's inputs, 16 rows of a_ld(sizes[k])
    int p_at[kMaxLayers] = {};  // LDS floats: hidden layer k's pre-activations, 16 x sizes[k + 1]
    int d_at[2] = {};           // LDS floats: two delta buffers, 16 rows of d_ld each
    int d_ld = 0;               // a delta row's stride: a_ld of the widest layer output
    int loss_at = 0;            // LDS floats: the waves' losses
    int lds = 0;                // LDS floats a block
    int base = 0;               // the first weight: 2 in + 1 padded to a multiple of 4
    int image = 0;              // image floats
    int trained = 0;            // weights and biases: image - base
};

__host__ __device__ inline int pad16(int v) { return (v + 15) / 16 * 16; }

// An input row's stride in LDS: padded to 16 (a product's depth goes 16 at a time, four steps of 4 whose loads are
// issued together; the padding stays zero) plus 1, so the 16 rows a product reads at once fall in different banks.
__host__ __device__ inline int a_ld(int n) { return pad16(n) + 1; }

// The activations through the hardware base-2 exponential and reciprocal (v_exp_f32, v_rcp_f32, a few ulp of fp32;
// near zero tanh's 1 - 2 / (1 + e^(2x)) is within an ulp of 1 absolute): tanh x = 1 - 2 / (1 + e^(2x)),
// silu x = x / (1 + e^(-x)), gelu (tanh form) z / (1 + e^(-2u)), u = sqrt(2 / pi) (z + 0.044715 z^3).
constexpr float kLog2e = 1.4426950408889634f;
constexpr float kGeluK0 = 0.7978845608028654f; // sqrt(2 / pi)

__device__ __forceinline__ float logistic2(float y) { return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(y)); } // 1 / (1 + 2^y)

__device__ __forceinline__ float gelu_s(float x) // 1 / (1 + e^(-2u))
{
    return logistic2(x * ((-2.0f * kGeluK0 * kLog2e) + (-2.0f * kGeluK0 * 0.044715f * kLog2e) * (x * x)));
}

__device__ __forceinline__ float act_f(int a, float x)
{
    switch (a) {
    case 0: return 1.0f - 2.0f * logistic2((2.0f * kLog2e) * x);
    case 1: return x > 0.0f ? x : 0.0f;
    case 2: return x * logistic2(-kLog2e * x);
    case 4: return x * gelu_s(x);
    default: return x;
    }
}

__device__ __forceinline__ float act_d(int a, float x)
{
    switch (a) {
    case 0: {
        const float t = 1.0f - 2.0f * logistic2((2.0f * kLog2e) * x);
        return 1.0f - t * t;
    }
    case 1: return x > 0.0f ? 1.0f : 0.0f;
    case 2: {
        const float s = logistic2(-kLog2e * x);
        return s + x * s * (1.0f - s);
    }
    case 4: {
        const float s = gelu_s(x);
        return s + x * s * (1.0f - s) * (2.0f * kGeluK0) * (1.0f + 3.0f * 0.044715f * (x * x));
    }
    default: return 1.0f;
    }
}

// The tiles blockIdx.x, blockIdx.x + gridDim.x, ... of 16 rows each, in that order: forward, loss, backward; the
// block's gradient summed over its tiles into partial[blockIdx.x][0, trained), its loss into column trained (each
// thread adds to the same elements for every tile, so the order is fixed). The four waves share a tile: each product's
// 16-column blocks (or the weight gradient's 16 x 16 blocks) go round the waves. LDS: s.lds floats (the launch gives
// them).
__global__ __launch_bounds__(64 * kWaves) void k_nt_tile(Shape s, const float* __restrict__ w, const float* __restrict__ X,
    const float* __restrict__ T, const float* __restrict__ M, int rows, float* __restrict__ partial)
{
    extern __shared__ __attribute__((aligned(16))) float sh[];
    const int tid = (int)threadIdx.x, lane = tid & 63, wave = tid >> 6, c16 = lane & 15, g4 = lane >> 4;
    const int in = s.sizes[0], out = s.sizes[s.layers], ld0 = a_ld(in), dl = s.d_ld;
    const int tiles = (rows + kRows - 1) / kRows, base = s.base;
    const float scale = w[2 * in];
    float* P = partial + (long long)blockIdx.x * (s.trained + 1);
    // padding columns stay zero for every tile: only the first sizes[k] columns of a row are written
    for (int q = tid; q < s.lds; q += 64 * kWaves) {
        sh[q] = 0.0f;
    }
    __syncthreads();
    float loss = 0.0f; // this lane's share, in a fixed order
    for (int tile = (int)blockIdx.x; tile < tiles; tile += (int)gridDim.x) {
        const int r0 = tile * kRows;
        const bool first = tile == (int)blockIdx.x;
        float* a0 = sh + s.a_at[0];
        for (int q = tid; q < kRows * in; q += 64 * kWaves) {
            const int r = q / in, i = q % in, row = r0 + r;
            a0[r * ld0 + i] = row < rows ? (X[(long long)row * in + i] - w[i]) * w[in + i] : 0.0f;
        }
        __syncthreads();
        for (int k = 0; k < s.layers; ++k) {
            const int ni = s.sizes[k], no = s.sizes[k + 1], ni16 = pad16(ni), la = a_ld(ni), lo = a_ld(no);
            const float* A = sh + s.a_at[k];
            const float* W = w + s.w_at[k];
            const float* b = W + (long long)no * ni;
            const bool last = k + 1 == s.layers;
            for (int nb = 16 * wave; nb < no; nb += 16 * kWaves) {
                const int col = nb + c16;
                float4v acc = {0.0f, 0.0f, 0.0f, 0.0f};
                for (int kc = 0; kc < ni16; kc += 16) {
                    float av[4], bv[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int kk = kc + 4 * u + g4;
                        av[u] = A[c16 * la + kk];
                        bv[u] = col < no && kk < ni ? W[(long long)col * ni + kk] : 0.0f;
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av[u], bv[u], acc, 0, 0, 0);
                    }
                }
                if (col < no) {
                    for (int j = 0; j < 4; ++j) {
                        const int r = 4 * g4 + j;
                        const float pre = acc[j] + b[col];
                        if (!last) {
                            sh[s.p_at[k] + r * no + col] = pre;
                            sh[s.a_at[k + 1] + r * lo + col] = act_f(s.act, pre);
                        } else {
                            const int row = r0 + r;
                            float d = 0.0f;
                            if (row < rows) {
                                const float m = M != nullptr ? M[row] : 1.0f;
                                const float e = pre * scale - T[(long long)row * out + col];
                                loss += 0.5f * m * e * e;
                                d = m * e * scale;
                            }
                            sh[s.d_at[0] + r * dl + col] = d;
                        }
                    }
                }
            }
            __syncthreads();
        }
        int cur = 0;
        for (int k = s.layers - 1; k >= 0; --k) {
            const int ni = s.sizes[k], no = s.sizes[k + 1], la = a_ld(ni);
            const float* A = sh + s.a_at[k];
            const float* D = sh + s.d_at[cur];
            const int g_at = s.w_at[k] - base;
            // weights: G[o][i] = sum over the tile's rows of D[r][o] A[r][i] (A: o x rows, B: rows x i), a 16 x 16 block
            // a wave at a time
            const int nib = (ni + 15) / 16, blocks = (no + 15) / 16 * nib;
            for (int q = wave; q < blocks; q += kWaves) {
                const int ob = q / nib * 16, ib = q % nib * 16;
                float4v acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int rc = 0; rc < kRows; rc += 4) {
                    const int r = rc + g4;
                    const float av = ob + c16 < no ? D[r * dl + ob + c16] : 0.0f;
                    const float bv = ib + c16 < ni ? A[r * la + ib + c16] : 0.0f;
                    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc, 0, 0, 0);
                }
                const int i = ib + c16;
                for (int j = 0; j < 4; ++j) {
                    const int o = ob + 4 * g4 + j;
                    if (o < no && i < ni) {
                        float* g = P + g_at + o * ni + i;
                        *g = first ? acc[j] : *g + acc[j];
                    }
                }
            }
            // biases: the rows in order
            for (int o = tid; o < no; o += 64 * kWaves) {
                float g = 0.0f;
                for (int r = 0; r < kRows; ++r) {
                    g += D[r * dl + o];
                }
                float* gp = P + g_at + no * ni + o;
                *gp = first ? g : *gp + g;
            }
            if (k > 0) {
                // the deltas below: (D W)[r][i] times the activation's derivative at the pre-activation
                const float* W = w + s.w_at[k];
                float* Dn = sh + s.d_at[cur ^ 1];
                const int no16 = pad16(no);
                for (int ib = 16 * wave; ib < ni; ib += 16 * kWaves) {
                    const int i = ib + c16;
                    float4v acc = {0.0f, 0.0f, 0.0f, 0.0f};
                    for (int oc = 0; oc < no16; oc += 16) {
                        float av[4], bv[4];
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            const int o = oc + 4 * u + g4;
                            av[u] = o < no ? D[c16 * dl + o] : 0.0f;
                            bv[u] = o < no && i < ni ? W[(long long)o * ni + i] : 0.0f;
                        }
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av[u], bv[u], acc, 0, 0, 0);
                        }
                    }
                    if (i < ni) {
                        for (int j = 0; j < 4; ++j) {
                            const int r = 4 * g4 + j;
                            Dn[r * dl + i] = acc[j] * act_d(s.act, sh[s.p_at[k - 1] + r * ni + i]);
                        }
                    }
                }
                cur ^= 1;
            }
            __syncthreads();
        }
    }
    // the block's loss: each wave's lanes folded by a fixed butterfly, then the waves in order
    for (int o = 32; o > 0; o >>= 1) {
        loss += __shfl_xor(loss, o);
    }
    if (lane == 0) {
        sh[s.loss_at + wave] = loss;
    }
    __syncthreads();
    if (tid == 0) {
        float l = 0.0f;
        for (int v = 0; v < kWaves; ++v) {
            l += sh[s.loss_at + v];
        }
        P[s.trained] = l;
    }
}

// First level: column c of the blocks' partials [chunk j, chunk j + 1) summed in block order into out[j][c].
__global__ void k_nt_chunks(const float* __restrict__ parts, int blocks, int width, float* __restrict__ out)
{
    const int c = (int)(blockIdx.x * blockDim.x + threadIdx.x), j = (int)blockIdx.y;
    if (c >= width) {
        return;
    }
    const int t1 = (j + 1) * kChunk < blocks ? (j + 1) * kChunk : blocks;
    float s = 0.0f;
#pragma unroll 8
    for (int t = j * kChunk; t < t1; ++t) {
        s += parts[(long long)t * width + c];
    }
    out[(long long)j * width + c] = s;
}

// Adam on the trained part of the image (the reference's form, src/field/cuda/CudaNet.cpp:356-366): each gradient
// clipped to +-clip, then the bias-corrected step.
struct Adam {
    float* w; // the image's trained part
    float* m1;
    float* m2;
    int n; // 0: no step
    float lr, b1, b2, c1, c2, clip;
};

__device__ __forceinline__ void adam_at(const Adam& a, int p, float grad)
{
    const float g = fminf(fmaxf(grad, -a.clip), a.clip);
    const float m = a.b1 * a.m1[p] + (1.0f - a.b1) * g, v = a.b2 * a.m2[p] + (1.0f - a.b2) * g * g;
    a.m1[p] = m;
    a.m2[p] = v;
    a.w[p] -= a.lr * (m / a.c1) / (sqrtf(v / a.c2) + 1e-8f);
}

// Last level: `count` rows of partial sums (the blocks' own, or the chunks') added in order, divided by the rows (the
// mean; the last column is the loss), and, when adam.n > 0, the Adam step on each weight and bias with its mean.
__global__ void k_nt_fold(const float* __restrict__ sums, int count, int width, float rows_f, float* __restrict__ grad, Adam adam)
{
    const int c = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (c < width) {
        float s = 0.0f;
#pragma unroll 8
        for (int j = 0; j < count; ++j) {
            s += sums[(long long)j * width + c];
        }
        const float g = s / rows_f;
        grad[c] = g;
        if (c < adam.n) {
            adam_at(adam, c, g);
        }
    }
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
        if (count <= n) {
            return;
        }
        if (p != nullptr) {
            hcheck(hipFree(p), "free");
            p = nullptr;
        }
        hcheck(hipMalloc(&p, std::max<std::size_t>(1, count) * sizeof(T)), "device allocation");
        n = count;
    }
    void put(const T* src, std::size_t count, hipStream_t s)
    {
        alloc(count);
        hcheck(hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, s), "upload");
    }
};

unsigned blocks_for(long long n, int per) { return static_cast<unsigned>((n + per - 1) / per); }

} // namespace

struct fe_nettrain {
    Shape s;
    Dev<float> w, m1, m2, x, t, m, partial, chunks, grad;
    int adam_t = 0;
    std::string error;
    hipStream_t stream() const { return static_cast<hipStream_t>(fe_hip_stream()); }
    const float* loss() const { return grad.p + s.trained; }

    // The gradient and loss for the rows already uploaded; leaves the mean gradient in grad[0, trained), the mean loss
    // in grad[trained], and takes the Adam step when adam.n > 0. Up to kOneLevel blocks their partials are added in one
    // level, past it in chunks of kChunk first.
    void gradient(int rows, bool masked, const Adam& adam)
    {
        const hipStream_t st = stream();
        const int blocks = std::min((rows + kRows - 1) / kRows, kBlocks), width = s.trained + 1;
        partial.alloc(static_cast<std::size_t>(blocks) * width);
        grad.alloc(static_cast<std::size_t>(width));
        hipLaunchKernelGGL(k_nt_tile, dim3(static_cast<unsigned>(blocks)), dim3(64 * kWaves), static_cast<unsigned>(s.lds) * 4u, st, s, w.p, x.p,
            t.p, masked ? m.p : nullptr, rows, partial.p);
        const float* sums = partial.p;
        int count = blocks;
        if (blocks > kOneLevel) {
            count = (blocks + kChunk - 1) / kChunk;
            chunks.alloc(static_cast<std::size_t>(count) * width);
            hipLaunchKernelGGL(k_nt_chunks, dim3(blocks_for(width, 256), static_cast<unsigned>(count)), dim3(256), 0, st, partial.p, blocks, width,
                chunks.p);
            sums = chunks.p;
        }
        hipLaunchKernelGGL(k_nt_fold, dim3(blocks_for(width, 256)), dim3(256), 0, st, sums, count, width, static_cast<float>(rows), grad.p, adam);
        hcheck(hipGetLastError(), "gradient launch");
    }

    void upload(const float* xs, const float* ts, const float* ms, int rows)
    {
        const hipStream_t st = stream();
        x.put(xs, static_cast<std::size_t>(rows) * s.sizes[0], st);
        t.put(ts, static_cast<std::size_t>(rows) * s.sizes[s.layers], st);
        if (ms != nullptr) {
            m.put(ms, static_cast<std::size_t>(rows), st);
        }
    }
};

namespace {

int fail(fe_nettrain* t, const std::string& what)
{
    if (t != nullptr) {
        t->error = what;
    }
    return -1;
}

} // namespace

extern "C" {

fe_nettrain* fe_nettrain_create(const int* sizes, int count, int act, char* why, size_t why_bytes)
{
    auto say = [&](const std::string& m) {
        if (why != nullptr && why_bytes > 0) {
            std::snprintf(why, why_bytes, "%s", m.c_str());
        }
    };
    if (sizes == nullptr || count < 2 || count > kMaxLayers + 1 || act < 0 || act > 4) {
        say("2..9 sizes (inputs, hidden widths, outputs) and an activation 0..4 (tanh, relu, silu, linear, gelu)");
        return nullptr;
    }
    Shape s;
    s.layers = count - 1;
    s.act = act;
    for (int k = 0; k < count; ++k) {
        if (sizes[k] < 1 || sizes[k] > kMaxWidth) {
            say("every size 1..128");
            return nullptr;
        }
        s.sizes[k] = sizes[k];
    }
    s.base = (2 * sizes[0] + 1 + 3) / 4 * 4;
    int at = s.base, lds = 0, widest = 0;
    for (int k = 0; k < s.layers; ++k) {
        s.w_at[k] = at;
        at += s.sizes[k] * s.sizes[k + 1] + s.sizes[k + 1];
        s.a_at[k] = lds;
        lds += kRows * a_ld(s.sizes[k]);
        widest = std::max(widest, s.sizes[k + 1]);
    }
    for (int k = 0; k + 1 < s.layers; ++k) {
        s.p_at[k] = lds;
        lds += kRows * s.sizes[k + 1];
    }
    s.d_ld = a_ld(widest);
    s.d_at[0] = lds;
    s.d_at[1] = lds + kRows * s.d_ld;
    s.loss_at = lds + 2 * kRows * s.d_ld;
    s.lds = s.loss_at + kWaves;
    s.image = at;
    s.trained = at - s.base;
    if (s.lds * 4 > kMaxLds) {
        say("the net's tile needs " + std::to_string(s.lds * 4) + " bytes of LDS, more than 65536");
        return nullptr;
    }
    fe_hip_device_info info{};
    if (fe_hip_device_open(&info) != 0) {
        say(std::string("no HIP device: ") + fe_hip_unavailable());
        return nullptr;
    }
    try {
        auto* t = new fe_nettrain();
        t->s = s;
        std::vector<float> image(static_cast<std::size_t>(s.image), 0.0f);
        for (int i = 0; i < sizes[0]; ++i) {
            image[static_cast<std::size_t>(sizes[0] + i)] = 1.0f;
        }
        image[static_cast<std::size_t>(2 * sizes[0])] = 1.0f;
        t->w.put(image.data(), image.size(), t->stream());
        std::vector<float> zeros(static_cast<std::size_t>(s.trained), 0.0f);
        t->m1.put(zeros.data(), zeros.size(), t->stream());
        t->m2.put(zeros.data(), zeros.size(), t->stream());
        hcheck(hipStreamSynchronize(t->stream()), "sync");
        return t;
    } catch (const std::exception& e) {
        say(e.what());
        return nullptr;
    }
}

void fe_nettrain_free(fe_nettrain* t)
{
    if (t != nullptr) {
        (void)hipStreamSynchronize(t->stream());
        delete t;
    }
}

const char* fe_nettrain_error(fe_nettrain* t) { return t != nullptr ? t->error.c_str() : "no trainer"; }

int fe_nettrain_image_size(fe_nettrain* t) { return t != nullptr ? t->s.image : -1; }

int fe_nettrain_set(fe_nettrain* t, const float* image)
{
    if (t == nullptr || image == nullptr) {
        return fail(t, "no trainer or no image");
    }
    try {
        t->w.put(image, static_cast<std::size_t>(t->s.image), t->stream());
        hcheck(hipStreamSynchronize(t->stream()), "sync");
        return 0;
    } catch (const std::exception& e) {
        return fail(t, e.what());
    }
}

int fe_nettrain_get(fe_nettrain* t, float* image)
{
    if (t == nullptr || image == nullptr) {
        return fail(t, "no trainer or no image");
    }
    try {
        hcheck(hipMemcpyAsync(image, t->w.p, sizeof(float) * static_cast<std::size_t>(t->s.image), hipMemcpyDeviceToHost, t->stream()), "download");
        hcheck(hipStreamSynchronize(t->stream()), "sync");
        return 0;
    } catch (const std::exception& e) {
        return fail(t, e.what());
    }
}

int fe_nettrain_gradient(fe_nettrain* t, const float* x, const float* target, const float* m, int rows, float* loss, float* grad)
{
    if (t == nullptr || x == nullptr || target == nullptr || rows < 1) {
        return fail(t, "no trainer, no rows or no targets");
    }
    try {
        t->upload(x, target, m, rows);
        t->gradient(rows, m != nullptr, Adam{});
        const hipStream_t st = t->stream();
        if (grad != nullptr) {
            hcheck(hipMemcpyAsync(grad, t->grad.p, sizeof(float) * static_cast<std::size_t>(t->s.trained), hipMemcpyDeviceToHost, st), "download");
        }
        float l = 0.0f;
        hcheck(hipMemcpyAsync(&l, t->loss(), sizeof(float), hipMemcpyDeviceToHost, st), "download");
        hcheck(hipStreamSynchronize(st), "sync");
        if (loss != nullptr) {
            *loss = l;
        }
        return 0;
    } catch (const std::exception& e) {
        return fail(t, e.what());
    }
}

int fe_nettrain_step(fe_nettrain* t, const float* x, const float* target, const float* m, int rows, float lr, float clip, float* loss)
{
    if (t == nullptr || x == nullptr || target == nullptr || rows < 1) {
        return fail(t, "no trainer, no rows or no targets");
    }
    try {
        t->upload(x, target, m, rows);
        const float b1 = 0.9f, b2 = 0.999f;
        const int step = t->adam_t + 1;
        const float c1 = 1.0f - std::pow(b1, static_cast<float>(step));
        const float c2 = 1.0f - std::pow(b2, static_cast<float>(step));
        t->gradient(rows, m != nullptr, Adam{t->w.p + t->s.base, t->m1.p, t->m2.p, t->s.trained, lr, b1, b2, c1, c2, clip});
        t->adam_t = step;
        const hipStream_t st = t->stream();
        float l = 0.0f;
        hcheck(hipMemcpyAsync(&l, t->loss(), sizeof(float), hipMemcpyDeviceToHost, st), "download");
        hcheck(hipStreamSynchronize(st), "sync");
        if (loss != nullptr) {
            *loss = l;
        }
        return 0;
    } catch (const std::exception& e) {
        return fail(t, e.what());
    }
}

} // extern "C"
