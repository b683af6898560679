// Law nets declared `half` (net NAME ... half): the forward pass of a batch of rows on the matrix cores, fp16 operands
// and fp32 accumulation (v_mfma_f32_16x16x16_f16), the reference's tensor-core path restated for gfx950
// (src/field/cuda/CudaNet.cpp:189-277 runs 32 rows a warp through m16n8k16 fp16; the CPU, the oracle, stays fp32).
//
// A row is one evaluation of the net: its inputs already normalised ((x - shift) * scale) and rounded to fp16 by the
// generated pack kernel, zero past the net's inputs. A 64-lane block takes 16 rows at a time: the layer's activations
// sit in LDS as a 16 x width fp16 tile; each 16-wide block of a layer's outputs is one chain of MFMAs over its inputs
// in 16-wide chunks (A: 16 rows x 16 inputs from LDS, lane l row l % 16, inputs 4 (l / 16)..+3; B: 16 inputs x 16
// outputs straight from the fp16 weights W[out][in], the same four inputs of output l % 16; D: lane l output l % 16,
// rows 4 (l / 16)..+3). Bias and activation in fp32 on D, rounded to fp16 into the other LDS tile; the last layer's
// outputs go out in fp32 times out_scale. Widths are padded to 16 with zero weights, and every activation maps 0 to 0,
// so padding carries nothing.
#include "kernels/net_mfma.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <stdexcept>
#include <vector>

#include "fe_hip.h"

namespace fieldengine::field::hip {

namespace {

typedef _Float16 half4 __attribute__((ext_vector_type(4)));
typedef float float4v __attribute__((ext_vector_type(4)));

constexpr int kStride = kHalfNetMaxWidth + 8; // LDS row stride in halves (16-byte aligned rows, banks staggered)

__device__ __forceinline__ float activate(int act, float v)
{
    switch (act) {
    case 0: return tanhf(v);
    case 1: return v > 0.0f ? v : 0.0f;
    case 2: return v / (1.0f + expf(-v));
    case 3: return v;
    default: return 0.5f * v * (1.0f + tanhf(0.7978845608028654f * (v + 0.044715f * v * v * v)));
    }
}

__global__ __launch_bounds__(64) void k_net_half(HalfNetDesc d, const _Float16* __restrict__ rows, long long n_rows,
    float* __restrict__ out)
{
    __shared__ __attribute__((aligned(16))) _Float16 act[2][16][kStride];
    const int lane = (int)threadIdx.x;
    const int r = lane & 15, g = lane >> 4;
    const long long tiles = (n_rows + 15) / 16;
    const int chunks = d.in_pad / 4;
    for (long long t = blockIdx.x; t < tiles; t += gridDim.x) {
        const long long row0 = t * 16;
        for (int q = lane; q < 16 * chunks; q += 64) {
            const int rr = q / chunks, cc = (q % chunks) * 4;
            half4 v = {0, 0, 0, 0};
            if (row0 + rr < n_rows) {
                v = *reinterpret_cast<const half4*>(rows + (row0 + rr) * d.in_pad + cc);
            }
            *reinterpret_cast<half4*>(&act[0][rr][cc]) = v;
        }
        __syncthreads();
        int cur = 0, K = d.in_pad;
        for (int l = 0; l < d.layers; ++l) {
            const int N = d.width[l];
            const bool last = l + 1 == d.layers;
            const _Float16* W = static_cast<const _Float16*>(d.w[l]);
            for (int nb = 0; nb < N; nb += 16) {
                float4v acc = {0.0f, 0.0f, 0.0f, 0.0f};
                for (int kc = 0; kc < K; kc += 16) {
                    const half4 a = *reinterpret_cast<const half4*>(&act[cur][r][kc + 4 * g]);
                    const half4 b = *reinterpret_cast<const half4*>(W + (long long)(nb + r) * K + kc + 4 * g);
                    acc = __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, acc, 0, 0, 0);
                }
                const float bias = d.b[l][nb + r];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int row = 4 * g + j;
                    const float v = acc[j] + bias;
                    if (!last) {
                        act[cur ^ 1][row][nb + r] = (_Float16)activate(d.act, v);
                    } else if (nb + r < d.outs && row0 + row < n_rows) {
                        out[(row0 + row) * d.outs + nb + r] = v * d.out_scale;
                    }
                }
            }
            __syncthreads();
            cur ^= 1;
            K = N;
        }
    }
}

int pad16(int v) { return (v + 15) / 16 * 16; }

} // namespace

int half_net_in_pad(const NetDecl& net) { return pad16(net.inputs()); }

HalfNetCard::~HalfNetCard()
{
    if (blob != nullptr) {
        (void)hipFree(blob);
    }
}

namespace {

// fp16 weights (each layer W[out_pad][in_pad], zero padded) then fp32 biases (b[out_pad]); widths into d.
void pack_half(const NetDecl& net, HalfNetDesc& d, std::vector<_Float16>& wh, std::vector<float>& bf, std::size_t* woff, std::size_t* boff)
{
    const int layers = net.layers();
    int K = d.in_pad;
    for (int l = 0; l < layers; ++l) {
        const int ni = net.sizes[static_cast<std::size_t>(l)], no = net.sizes[static_cast<std::size_t>(l) + 1];
        const int N = pad16(no);
        if (K > kHalfNetMaxWidth || N > kHalfNetMaxWidth) {
            throw std::runtime_error("net '" + net.name + "': a layer wider than " + std::to_string(kHalfNetMaxWidth));
        }
        d.width[l] = N;
        const float* w = net.weights.data() + net.weight_offset(l);
        woff[l] = wh.size();
        for (int o = 0; o < N; ++o) {
            for (int i = 0; i < K; ++i) {
                wh.push_back(o < no && i < ni ? static_cast<_Float16>(w[static_cast<std::size_t>(o) * ni + i]) : static_cast<_Float16>(0.0f));
            }
        }
        boff[l] = bf.size();
        for (int o = 0; o < N; ++o) {
            bf.push_back(o < no ? w[static_cast<std::size_t>(ni) * no + o] : 0.0f);
        }
        K = N;
    }
}

} // namespace

std::unique_ptr<HalfNetCard> half_net_create(const NetDecl& net)
{
    const int layers = net.layers();
    if (layers < 1 || layers > kHalfNetMaxLayers) {
        throw std::runtime_error("net '" + net.name + "': " + std::to_string(layers) + " layers (the matrix-core path takes 1.." +
            std::to_string(kHalfNetMaxLayers) + ")");
    }
    auto card = std::make_unique<HalfNetCard>();
    HalfNetDesc& d = card->desc;
    d.layers = layers;
    d.in_pad = pad16(net.inputs());
    d.outs = net.outputs();
    d.act = static_cast<int>(net.act);
    d.out_scale = net.out_scale;
    std::vector<_Float16> wh;
    std::vector<float> bf;
    std::size_t woff[kHalfNetMaxLayers], boff[kHalfNetMaxLayers];
    pack_half(net, d, wh, bf, woff, boff);
    const std::size_t wbytes = (wh.size() * 2 + 255) / 256 * 256;
    const std::size_t bytes = wbytes + bf.size() * 4;
    if (hipMalloc(&card->blob, bytes) != hipSuccess ||
        hipMemcpy(card->blob, wh.data(), wh.size() * 2, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(static_cast<char*>(card->blob) + wbytes, bf.data(), bf.size() * 4, hipMemcpyHostToDevice) != hipSuccess) {
        throw std::runtime_error("net '" + net.name + "': device copy of its half weights failed");
    }
    for (int l = 0; l < layers; ++l) {
        d.w[l] = static_cast<const _Float16*>(card->blob) + woff[l];  // stored as const void*
        d.b[l] = reinterpret_cast<const float*>(static_cast<const char*>(card->blob) + wbytes) + boff[l];
    }
    return card;
}

void half_net_update(HalfNetCard& card, const NetDecl& net)
{
    HalfNetDesc d = card.desc;
    d.out_scale = net.out_scale;
    std::vector<_Float16> wh;
    std::vector<float> bf;
    std::size_t woff[kHalfNetMaxLayers], boff[kHalfNetMaxLayers];
    pack_half(net, d, wh, bf, woff, boff);
    const std::size_t wbytes = (wh.size() * 2 + 255) / 256 * 256;
    const hipStream_t s = static_cast<hipStream_t>(fe_hip_stream());
    if (hipMemcpyAsync(card.blob, wh.data(), wh.size() * 2, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(static_cast<char*>(card.blob) + wbytes, bf.data(), bf.size() * 4, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        throw std::runtime_error("net '" + net.name + "': device copy of its half weights failed");
    }
    card.desc.out_scale = net.out_scale;
}

int half_net_run(const HalfNetCard& net, const void* rows, long long n_rows, float* out)
{
    if (n_rows <= 0) {
        return 0;
    }
    const long long tiles = (n_rows + 15) / 16;
    const unsigned blocks = static_cast<unsigned>(tiles < 8192 ? tiles : 8192);
    hipLaunchKernelGGL(k_net_half, dim3(blocks), dim3(64), 0, static_cast<hipStream_t>(fe_hip_stream()), net.desc,
        static_cast<const _Float16*>(rows), n_rows, out);
    return static_cast<int>(hipGetLastError());
}

} // namespace fieldengine::field::hip

// The same forward pass through the C ABI, for tests: nets given by their sizes, activation and weights in the
// reference's layout (per layer W (out x in, row-major) then b), rows already normalised.
extern "C" int fe_net_half_forward(const int* sizes, int n_sizes, int act, float out_scale, const float* weights,
    const float* rows, long long n_rows, float* out)
{
    using namespace fieldengine::field;
    if (sizes == nullptr || n_sizes < 2 || weights == nullptr || rows == nullptr || out == nullptr || n_rows < 0 || act < 0 || act > 4) {
        return -1;
    }
    try {
        NetDecl net;
        net.name = "test";
        net.sizes.assign(sizes, sizes + n_sizes);
        net.act = static_cast<Activation>(act);
        net.out_scale = out_scale;
        net.weights.assign(weights, weights + net.parameter_count());
        auto card = hip::half_net_create(net);
        const int in_pad = hip::half_net_in_pad(net);
        std::vector<_Float16> h(static_cast<std::size_t>(n_rows) * in_pad, static_cast<_Float16>(0.0f));
        for (long long r = 0; r < n_rows; ++r) {
            for (int i = 0; i < net.inputs(); ++i) {
                h[static_cast<std::size_t>(r) * in_pad + i] = static_cast<_Float16>(rows[r * net.inputs() + i]);
            }
        }
        void* drows = nullptr;
        float* dout = nullptr;
        const std::size_t out_bytes = static_cast<std::size_t>(n_rows) * net.outputs() * 4;
        if (hipMalloc(&drows, h.size() * 2 + 16) != hipSuccess || hipMalloc(reinterpret_cast<void**>(&dout), out_bytes + 16) != hipSuccess) {
            return 2;
        }
        int rc = static_cast<int>(hipMemcpy(drows, h.data(), h.size() * 2, hipMemcpyHostToDevice));
        if (rc == 0) {
            rc = hip::half_net_run(*card, drows, n_rows, dout);
        }
        if (rc == 0) {
            rc = static_cast<int>(hipStreamSynchronize(static_cast<hipStream_t>(fe_hip_stream())));
        }
        if (rc == 0) {
            rc = static_cast<int>(hipMemcpy(out, dout, out_bytes, hipMemcpyDeviceToHost));
        }
        (void)hipFree(drows);
        (void)hipFree(dout);
        return rc;
    } catch (const std::exception&) {
        return -2;
    }
}
