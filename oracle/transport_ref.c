/* ORACLE (test infrastructure only): the three-field reaction-advection-diffusion law the adjoint is checked on,
 * stepped in FP64 on the CPU, and its loss L = sum over the lattice of c^2 after the last step. Central differences of
 * L here are what the card's checkpointed adjoint gradients are held to. Only tests/ and bench.py call it.
 *
 *   dX/dt = diffuse(DX, X) - vx (X - X[x-1]) / h - vy (X[y+1] - X) / h ... (upwind for vx = 0.3 > 0, vy = -0.2 < 0,
 *           vz = 0.1 > 0) + reaction:  da = db = -a b + 0.1 c,  dc = a b - 0.1 c
 *
 * diffuse() is the flux form with missing faces skipped (src/field/Law.cpp:986-1018; one k, so every face conducts k),
 * shift() repeats the edge cell (src/field/Law.cpp:928-929): the upwind difference is 0 across the lattice's edge. */
#include <omp.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

static double diffuse_at(const double* F, int nx, int ny, int nz, int x, int y, int z, double k, double h)
{
    const size_t sy = (size_t)nx, sz = (size_t)nx * ny;
    const size_t i = (size_t)z * sz + (size_t)y * sy + x;
    const double c = F[i];
    double acc = 0.0;
    if (x > 0) acc += k * (F[i - 1] - c);
    if (x + 1 < nx) acc += k * (F[i + 1] - c);
    if (y > 0) acc += k * (F[i - sy] - c);
    if (y + 1 < ny) acc += k * (F[i + sy] - c);
    if (z > 0) acc += k * (F[i - sz] - c);
    if (z + 1 < nz) acc += k * (F[i + sz] - c);
    return acc / (h * h);
}

static double advect_at(const double* F, int nx, int ny, int nz, int x, int y, int z, double h)
{
    const size_t sy = (size_t)nx, sz = (size_t)nx * ny;
    const size_t i = (size_t)z * sz + (size_t)y * sy + x;
    const double c = F[i];
    const double xm = x > 0 ? F[i - 1] : c;
    const double yp = y + 1 < ny ? F[i + sy] : c;
    const double zm = z > 0 ? F[i - sz] : c;
    return -0.3 * (c - xm) / h + 0.2 * (yp - c) / h - 0.1 * (c - zm) / h;
}

/* `steps` steps of dt from (a, b, c) in place; returns sum c^2 at the end. D: the three diffusivities. */
double oracle_transport_run(int nx, int ny, int nz, double h, double dt, int steps, const double* D, double* a, double* b, double* c)
{
    const size_t n = (size_t)nx * ny * nz;
    double* na = (double*)malloc(n * sizeof(double));
    double* nb = (double*)malloc(n * sizeof(double));
    double* nc = (double*)malloc(n * sizeof(double));
    for (int s = 0; s < steps; ++s) {
#pragma omp parallel for schedule(static)
        for (int z = 0; z < nz; ++z) {
            for (int y = 0; y < ny; ++y) {
                for (int x = 0; x < nx; ++x) {
                    const size_t i = (size_t)z * nx * ny + (size_t)y * nx + x;
                    const double r = -a[i] * b[i] + 0.1 * c[i];
                    na[i] = a[i] + dt * (diffuse_at(a, nx, ny, nz, x, y, z, D[0], h) + advect_at(a, nx, ny, nz, x, y, z, h) + r);
                    nb[i] = b[i] + dt * (diffuse_at(b, nx, ny, nz, x, y, z, D[1], h) + advect_at(b, nx, ny, nz, x, y, z, h) + r);
                    nc[i] = c[i] + dt * (diffuse_at(c, nx, ny, nz, x, y, z, D[2], h) + advect_at(c, nx, ny, nz, x, y, z, h) - r);
                }
            }
        }
        memcpy(a, na, n * sizeof(double));
        memcpy(b, nb, n * sizeof(double));
        memcpy(c, nc, n * sizeof(double));
    }
    double loss = 0.0;
    for (size_t i = 0; i < n; ++i) {
        loss += c[i] * c[i];
    }
    free(na);
    free(nb);
    free(nc);
    return loss;
}
