/* ORACLE (test infrastructure only): the plain 2-D diffusion-reaction law in FP64 on the CPU, the run the coarse
 * lattices' error is measured against. Only tests/ and bench.py call it.
 *
 *   du/dt = Du lap u + u - u^3 - k - v,   dv/dt = Dv lap v + u - v
 *
 * on `count` lattices of ny x nx cells of side h, the 5-point Laplacian with the edge cell repeating past the ends
 * (zero flux), explicit Euler steps of dt: the law the reference ships as bench/laws/diff_react.law (its x-only shift
 * form there, :19-20), here in two dimensions as the batched coarse lattice of the learned correction computes it. */
#include <omp.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

void oracle_dr_run_f64(int nx, int ny, int count, double h, double du, double dv, double k, double dt, int steps, double* u, double* v)
{
    const size_t cells = (size_t)nx * (size_t)ny;
    const size_t n = cells * (size_t)count;
    double* u2 = (double*)malloc(n * sizeof(double));
    double* v2 = (double*)malloc(n * sizeof(double));
    const double ih2 = 1.0 / (h * h);
    for (int s = 0; s < steps; ++s) {
#pragma omp parallel for schedule(static)
        for (long long row = 0; row < (long long)ny * count; ++row) {
            const size_t lat = (size_t)(row / ny);
            const int y = (int)(row % ny);
            const double* U = u + lat * cells;
            const double* V = v + lat * cells;
            const int ym = y > 0 ? y - 1 : y, yp = y + 1 < ny ? y + 1 : y;
            for (int x = 0; x < nx; ++x) {
                const int xm = x > 0 ? x - 1 : x, xp = x + 1 < nx ? x + 1 : x;
                const size_t i = (size_t)y * nx + x;
                const double uc = U[i], vc = V[i];
                const double lu = (U[(size_t)y * nx + xm] + U[(size_t)y * nx + xp] + U[(size_t)ym * nx + x] + U[(size_t)yp * nx + x] - 4.0 * uc) * ih2;
                const double lv = (V[(size_t)y * nx + xm] + V[(size_t)y * nx + xp] + V[(size_t)ym * nx + x] + V[(size_t)yp * nx + x] - 4.0 * vc) * ih2;
                u2[lat * cells + i] = uc + dt * (du * lu + uc - uc * uc * uc - k - vc);
                v2[lat * cells + i] = vc + dt * (dv * lv + uc - vc);
            }
        }
        memcpy(u, u2, n * sizeof(double));
        memcpy(v, v2, n * sizeof(double));
    }
    free(u2);
    free(v2);
}
