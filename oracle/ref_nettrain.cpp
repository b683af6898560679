// Test infrastructure: the loss and gradient of the net trainer's regression by the REFERENCE's own CPU net code
// (net_forward_taped / net_backward, src/field/nn/Net.hpp:56-60, linked from its objects by oracle/Makefile.ref), the way
// its world model's host path learns (src/field/mind/WorldModel.cpp:430-460: forward with the tape, dL/dy = e / rows,
// back, the rows added in order), with a weight m a row. tests/golden/make_nettrain_golden.py runs it to write the
// vectors that pin oracle/nettrain.py and the card.
//
//   ref_nettrain IN OUT
//
// IN, little-endian: int32 count, int32 sizes[count], int32 act (0 tanh, 1 relu, 2 silu, 3 linear), int32 rows,
// float in_shift[in], in_scale[in], out_scale, weights[parameters] (per layer W[out][in], then b[out]),
// x[rows][in], t[rows][out], m[rows]. OUT: float64 loss (the rows' mean of 0.5 m |y - t|^2), float32 grad[parameters].
#include "field/nn/Net.hpp"

#include <cstdio>
#include <cstdint>
#include <vector>

using namespace fieldengine::field;

namespace {

template <typename T>
bool take(std::FILE* f, T* p, std::size_t n)
{
    return std::fread(p, sizeof(T), n, f) == n;
}

} // namespace

int main(int argc, char** argv)
{
    if (argc != 3) {
        std::fprintf(stderr, "usage: ref_nettrain IN OUT\n");
        return 2;
    }
    std::FILE* f = std::fopen(argv[1], "rb");
    if (f == nullptr) {
        std::fprintf(stderr, "cannot read %s\n", argv[1]);
        return 2;
    }
    std::int32_t count = 0, act = 0, rows = 0;
    NetDecl net;
    bool ok = take(f, &count, 1) && count >= 2 && count <= 16;
    std::vector<std::int32_t> sizes(ok ? static_cast<std::size_t>(count) : 0);
    ok = ok && take(f, sizes.data(), sizes.size()) && take(f, &act, 1) && take(f, &rows, 1) && act >= 0 && act <= 3 && rows >= 1;
    if (!ok) {
        std::fprintf(stderr, "bad header\n");
        return 2;
    }
    net.name = "ref_nettrain";
    net.sizes.assign(sizes.begin(), sizes.end());
    net.act = static_cast<Activation>(act);
    const int in = net.inputs(), out = net.outputs();
    net.in_shift.resize(static_cast<std::size_t>(in));
    net.in_scale.resize(static_cast<std::size_t>(in));
    net.weights.resize(net.parameter_count());
    std::vector<float> x(static_cast<std::size_t>(rows) * in), t(static_cast<std::size_t>(rows) * out), m(static_cast<std::size_t>(rows));
    ok = take(f, net.in_shift.data(), net.in_shift.size()) && take(f, net.in_scale.data(), net.in_scale.size()) && take(f, &net.out_scale, 1) &&
        take(f, net.weights.data(), net.weights.size()) && take(f, x.data(), x.size()) && take(f, t.data(), t.size()) && take(f, m.data(), m.size());
    std::fclose(f);
    if (!ok) {
        std::fprintf(stderr, "short input\n");
        return 2;
    }
    net.loaded = true;
    int taped = 0;
    for (std::size_t k = 1; k < net.sizes.size(); ++k) {
        taped += net.sizes[k];
    }
    std::vector<float> tape(static_cast<std::size_t>(taped)), y(static_cast<std::size_t>(out)), dy(static_cast<std::size_t>(out));
    std::vector<float> grad(net.parameter_count(), 0.0f), scratch(2 * static_cast<std::size_t>(net.widest()));
    double loss = 0.0;
    for (int r = 0; r < rows; ++r) {
        const float* xr = &x[static_cast<std::size_t>(r) * in];
        net_forward_taped(net, xr, y.data(), tape.data());
        for (int o = 0; o < out; ++o) {
            const float e = y[static_cast<std::size_t>(o)] - t[static_cast<std::size_t>(r) * out + o];
            loss += 0.5 * m[static_cast<std::size_t>(r)] * e * e;
            dy[static_cast<std::size_t>(o)] = m[static_cast<std::size_t>(r)] * e / static_cast<float>(rows);
        }
        net_backward(net, xr, tape.data(), dy.data(), grad.data(), nullptr, scratch.data());
    }
    loss /= rows;
    std::FILE* g = std::fopen(argv[2], "wb");
    if (g == nullptr || std::fwrite(&loss, sizeof loss, 1, g) != 1 || std::fwrite(grad.data(), sizeof(float), grad.size(), g) != grad.size()) {
        std::fprintf(stderr, "cannot write %s\n", argv[2]);
        return 2;
    }
    std::fclose(g);
    return 0;
}
