/* ORACLE (test infrastructure only): a plain-C restatement of the reference CPU solver for one law,
 *
 *     term conduction
 *       d T = diffuse(k, T)          k one number (a param)
 *       max_dt = c * h * h / k       (optional; c <= 0 means no limit)
 *
 * following the reference's CPU path step by step:
 *   - substeps: ceil(dt / max_dt) when dt > max_dt, capped at 4096 (src/field/Solver.cpp:356-371);
 *   - the rate: diffuse() in double per cell, faces -x +x -y +y -z +z, each kf * (double(T[nb]) - double(T[c])) with
 *     kf = 2 ka kb / (ka + kb), stored as float (src/field/Law.cpp:986-1018); a uniform T is the scalar 0 and the term
 *     is idle (src/field/Law.cpp:565-580, src/field/Solver.cpp:389-394);
 *   - the store: delta = float(double(rate) * sub_dt), then T += delta in float (src/field/Solver.cpp:398-427);
 *   - FTZ and DAZ on, as on every thread the reference computes cells on (src/field/Parallel.hpp:18-28).
 *
 * Only tests/, the smoke check and bench.py's CPU baseline call this; the product never does.
 */
#include <math.h>
#include <omp.h>
#include <pmmintrin.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <xmmintrin.h>

static void flush_denormals(void)
{
    _MM_SET_FLUSH_ZERO_MODE(_MM_FLUSH_ZERO_ON);
    _MM_SET_DENORMALS_ZERO_MODE(_MM_DENORMALS_ZERO_ON);
}

static int uniform(const float* v, size_t n)
{
    for (size_t i = 1; i < n; ++i) {
        if (v[i] != v[0]) {
            return 0;
        }
    }
    return 1;
}

/* One substep: out = T + float(double(rate) * sub_dt). Returns 0 when the term was idle (out is then T). */
int oracle_heat_substep(const float* t, float* out, int nx, int ny, int nz, double k, double h, double sub_dt)
{
    const size_t n = (size_t)nx * (size_t)ny * (size_t)nz;
    const size_t sy = (size_t)nx;
    const size_t sz = (size_t)nx * (size_t)ny;
    flush_denormals();
    if (uniform(t, n)) {
        memcpy(out, t, n * sizeof(float));
        return 0;
    }
    const double kf = (k + k) > 0.0 ? 2.0 * k * k / (k + k) : 0.0;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; ++z) {
        flush_denormals();
        for (int y = 0; y < ny; ++y) {
            for (int x = 0; x < nx; ++x) {
                const size_t i = (size_t)z * sz + (size_t)y * sy + (size_t)x;
                const double c = (double)t[i];
                double acc = 0.0;
                if (x > 0) acc += kf * ((double)t[i - 1] - c);
                if (x + 1 < nx) acc += kf * ((double)t[i + 1] - c);
                if (y > 0) acc += kf * ((double)t[i - sy] - c);
                if (y + 1 < ny) acc += kf * ((double)t[i + sy] - c);
                if (z > 0) acc += kf * ((double)t[i - sz] - c);
                if (z + 1 < nz) acc += kf * ((double)t[i + sz] - c);
                const float rate = (float)(acc / (h * h));
                const float delta = (float)((double)rate * sub_dt);
                out[i] = t[i] + delta;
            }
        }
    }
    return 1;
}

/* The substeps a step of dt takes under max_dt = coeff * h * h / k (coeff <= 0: no limit). */
int oracle_heat_substeps(double dt, double coeff, double k, double h)
{
    if (coeff <= 0.0) {
        return 1;
    }
    const double limit = coeff * h * h / k;
    if (limit > 0.0 && isfinite(limit) && dt > limit) {
        const double wanted = ceil(dt / limit);
        return wanted > 4096.0 ? 4096 : (int)wanted;
    }
    return 1;
}

/* `steps` steps of dt on T in place (scratch: n floats). Returns the substeps of the last step. */
int oracle_heat_run(float* t, float* scratch, int nx, int ny, int nz, double k, double h, double dt, double coeff, int steps)
{
    const size_t n = (size_t)nx * (size_t)ny * (size_t)nz;
    const int substeps = oracle_heat_substeps(dt, coeff, k, h);
    const double sub_dt = dt / substeps;
    for (int s = 0; s < steps; ++s) {
        for (int q = 0; q < substeps; ++q) {
            if (!oracle_heat_substep(t, scratch, nx, ny, nz, k, h, sub_dt)) {
                break;
            }
            memcpy(t, scratch, n * sizeof(float));
        }
    }
    return substeps;
}

/* The compensated fixed-order total the build reports beside the CPU's double total: sum and sum of |v| in double, per
 * 8192-cell block and then the blocks in order (src/field/Parallel.hpp:439-457, src/field/Solver.cpp:31-39). */
void oracle_total(const float* v, size_t n, double* sum, double* mag)
{
    const size_t block = 8192;
    const size_t blocks = (n + block - 1) / block;
    double s = 0.0;
    double m = 0.0;
    for (size_t b = 0; b < blocks; ++b) {
        double ps = 0.0;
        double pm = 0.0;
        const size_t end = (b + 1) * block < n ? (b + 1) * block : n;
        for (size_t i = b * block; i < end; ++i) {
            ps += (double)v[i];
            pm += fabs((double)v[i]);
        }
        s += ps;
        m += pm;
    }
    *sum = s;
    *mag = m;
}
