/* ORACLE (test infrastructure only): Heston by ADI on the CPU, one option a lattice, for checking the card's batched
 * solver. Only tests/, the smoke check and bench.py's CPU baseline call it; the product never does.
 *
 * The PDE and its three-point operators are those of the reference's ADI law (bench/laws/heston_adi.law:1-51, run by
 * bench/finance.py:253-288): x = ln S along x with cells of h, the variance var = sigma (j + 1/2) h along y,
 *   A1 W = lx (W[i-1] - W) + ux (W[i+1] - W) - r/2 W,  lx,ux = var/(2h^2) -+ (r - var/2)/(2h)            (heston_adi.law:37-38)
 *   A2 W = ly (W[j-1] - W) + uy (W[j+1] - W) - r/2 W,  ly = (var - sigma h/2)/(2h^2) - (1 - e0) b/(2h),
 *                                                    uy = (var + sigma h/2)/(2h^2) + (1 + e0) b/(2h)      (heston_adi.law:39-40)
 *        b = kappa (theta - var)/sigma - sigma/2, e0 = 1 on the first row (var < sigma h), else 0            (heston_adi.law:35-36)
 *   A0 W = rho var gx(gy(W)), central, one-sided at an edge                                                (heston_adi.law:43)
 * the edge cell repeating past the ends (shift's and tridiag's rule), so each implicit operator is exactly its explicit
 * one. F = A0 + A1 + A2. In increments (delta form, heston_adi.law:10-16):
 *   D1 = (I - th dt A1)^-1 dt F(U),  D2 = (I - th dt A2)^-1 D1
 *   Craig-Sneyd:          E1 = (I - th dt A1)^-1 (dt/2 A0 D2)                       (In 't Hout & Foulon 2010, scheme CS)
 *   Hundsdorfer-Verwer:   E1 = (I - th dt A1)^-1 (dt F(U) - D2 + dt/2 F(D2))        (heston_adi.law:44-47)
 *   E2 = (I - th dt A2)^-1 E1,  U' = U + D2 + E2.
 * The start is the put's payoff averaged over each cell (bench/finance.py:241-251); the price is read at (ln 100, v0)
 * by 4-point Lagrange interpolation along each axis (the card's solver reads it the same way).
 */
#include <math.h>
#include <omp.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    double kappa, theta, sigma, rho, r, strike, maturity, v0;
} heston_option;

typedef struct {
    int nx, ny;
    double h, x0;            /* cell size; x of cell 0's centre */
    heston_option opt;
    double *var, *lx, *ux, *ly, *uy; /* per row */
} heston_grid;

static const double kSpot = 100.0;
static const double kHalfWidth = 3.0; /* ln 100 +- 3 along x */

static void grid_make(heston_grid* g, int nx, int ny, const heston_option* o)
{
    g->nx = nx;
    g->ny = ny;
    g->h = 2.0 * kHalfWidth / nx;
    g->x0 = log(kSpot) - kHalfWidth + 0.5 * g->h;
    g->opt = *o;
    g->var = (double*)malloc(sizeof(double) * 5 * (size_t)ny);
    g->lx = g->var + ny;
    g->ux = g->lx + ny;
    g->ly = g->ux + ny;
    g->uy = g->ly + ny;
    const double h = g->h, s = o->sigma, r = o->r;
    for (int j = 0; j < ny; ++j) {
        const double v = s * (j + 0.5) * h;
        const double e0 = v < s * h ? 1.0 : 0.0;
        const double b = o->kappa * (o->theta - v) / s - 0.5 * s;
        g->var[j] = v;
        g->lx[j] = 0.5 * v / (h * h) - (r - 0.5 * v) / (2 * h);
        g->ux[j] = 0.5 * v / (h * h) + (r - 0.5 * v) / (2 * h);
        g->ly[j] = 0.5 * (v - 0.5 * s * h) / (h * h) - (1 - e0) * b / (2 * h);
        g->uy[j] = 0.5 * (v + 0.5 * s * h) / (h * h) + (1 + e0) * b / (2 * h);
    }
}

static void grid_free(heston_grid* g) { free(g->var); }

#define REAL double
#define SFX(name) name##_f64
#include "heston_core.h"
#undef REAL
#undef SFX
#define REAL float
#define SFX(name) name##_f32
#include "heston_core.h"
#undef REAL
#undef SFX

/* The cell-averaged put payoff of one option into V (ny rows of nx). */
void oracle_heston_payoff(int nx, int ny, const double* params, double* V)
{
    heston_grid g;
    grid_make(&g, nx, ny, (const heston_option*)params);
    const double k = g.opt.strike, lk = log(k), h = g.h;
    for (int i = 0; i < nx; ++i) {
        const double x = g.x0 + i * h;
        const double lo = fmin(x - 0.5 * h, lk), hi = fmin(x + 0.5 * h, lk);
        const double p = (k * (hi - lo) - (exp(hi) - exp(lo))) / h;
        for (int j = 0; j < ny; ++j) {
            V[(size_t)j * nx + i] = p;
        }
    }
    grid_free(&g);
}

/* 4-point Lagrange weights at t on the unit grid 0..3. */
static void lagrange4(double t, double w[4])
{
    w[0] = -(t - 1) * (t - 2) * (t - 3) / 6.0;
    w[1] = t * (t - 2) * (t - 3) / 2.0;
    w[2] = -t * (t - 1) * (t - 3) / 2.0;
    w[3] = t * (t - 1) * (t - 2) / 6.0;
}

/* The first cell of the 4-point stencil and its weights along x (at ln 100) and along y (at var = v0). */
void oracle_heston_stencil(int nx, int ny, const double* params, int* i0, int* j0, double* wx, double* wy)
{
    heston_grid g;
    grid_make(&g, nx, ny, (const heston_option*)params);
    const double tx = (log(kSpot) - g.x0) / g.h;
    int a = (int)floor(tx) - 1;
    a = a < 0 ? 0 : (a > nx - 4 ? nx - 4 : a);
    const double ty = (g.opt.v0 / g.opt.sigma) / g.h - 0.5;
    int b = (int)floor(ty) - 1;
    b = b < 0 ? 0 : (b > ny - 4 ? ny - 4 : b);
    lagrange4(tx - a, wx);
    lagrange4(ty - b, wy);
    *i0 = a;
    *j0 = b;
    grid_free(&g);
}

static double price_of(int nx, int ny, const double* params, const double* V)
{
    int i0, j0;
    double wx[4], wy[4];
    oracle_heston_stencil(nx, ny, params, &i0, &j0, wx, wy);
    double p = 0.0;
    for (int b = 0; b < 4; ++b) {
        double row = 0.0;
        for (int a = 0; a < 4; ++a) {
            row += wx[a] * V[(size_t)(j0 + b) * nx + i0 + a];
        }
        p += wy[b] * row;
    }
    return p;
}

double oracle_heston_price(int nx, int ny, const double* params, const double* V) { return price_of(nx, ny, params, V); }

/* One option, FP64: V (the payoff, or any state) advanced `steps` steps of maturity/steps in place. */
void oracle_heston_run_f64(int nx, int ny, const double* params, double* V, int steps, int scheme, double th)
{
    heston_grid g;
    grid_make(&g, nx, ny, (const heston_option*)params);
    run_f64(&g, V, steps, scheme, th);
    grid_free(&g);
}

void oracle_heston_run_f32(int nx, int ny, const double* params, float* V, int steps, int scheme, double th)
{
    heston_grid g;
    grid_make(&g, nx, ny, (const heston_option*)params);
    run_f32(&g, V, steps, scheme, th);
    grid_free(&g);
}

/* A batch, options in parallel over the host's cores (OpenMP, dynamic: the lattices are uneven). V: every option's
 * lattice back to back from offsets[o]; params: 8 doubles an option (kappa theta sigma rho r strike maturity v0).
 * Prices into prices[o] when not NULL. */
void oracle_heston_batch_f64(int n, const int* nx, const int* ny, const double* params, const long long* offsets, double* V,
    int steps, int scheme, double th, double* prices)
{
#pragma omp parallel for schedule(dynamic, 1)
    for (int o = 0; o < n; ++o) {
        oracle_heston_run_f64(nx[o], ny[o], params + 8 * (size_t)o, V + offsets[o], steps, scheme, th);
        if (prices != NULL) {
            prices[o] = price_of(nx[o], ny[o], params + 8 * (size_t)o, V + offsets[o]);
        }
    }
}

void oracle_heston_batch_f32(int n, const int* nx, const int* ny, const double* params, const long long* offsets, float* V,
    int steps, int scheme, double th)
{
#pragma omp parallel for schedule(dynamic, 1)
    for (int o = 0; o < n; ++o) {
        oracle_heston_run_f32(nx[o], ny[o], params + 8 * (size_t)o, V + offsets[o], steps, scheme, th);
    }
}

/* The Thomas solve the schemes use, exposed for checking against a dense solve: (a, b, c) constant, edges folded. */
void oracle_heston_thomas(int n, double a, double b, double c, const double* d, double* x)
{
    double* cp = (double*)malloc(sizeof(double) * (size_t)n);
    double* dp = (double*)malloc(sizeof(double) * (size_t)n);
    thomas_f64(n, a, b, c, d, 1, x, 1, cp, dp);
    free(cp);
    free(dp);
}

/* One application of F (CS and HV's explicit part) to W, for checking the operators against the engine's law. */
void oracle_heston_apply_f(int nx, int ny, const double* params, const double* W, double* out)
{
    heston_grid g;
    grid_make(&g, nx, ny, (const heston_option*)params);
    for (int j = 0; j < ny; ++j) {
        for (int i = 0; i < nx; ++i) {
            out[(size_t)j * nx + i] = op_mixed_f64(&g, W, i, j) + op_x_f64(&g, W, i, j) + op_y_f64(&g, W, i, j);
        }
    }
    grid_free(&g);
}
