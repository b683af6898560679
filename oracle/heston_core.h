/* ORACLE (test infrastructure only): the body of heston_ref.c, included once with REAL = double (the FP64 run the
 * card's results are measured against) and once with REAL = float (the host-core baseline the card's speed is
 * reported beside). See heston_ref.c for the scheme and the reference lines it follows. */

/* F restricted to one operator family at (i, j) of W (row-major, x fastest), the edge cell repeating past the ends. */
static REAL SFX(op_x)(const heston_grid* g, const REAL* W, int i, int j)
{
    const int nx = g->nx;
    const REAL* row = W + (size_t)j * nx;
    const REAL w = row[i];
    const REAL wl = row[i > 0 ? i - 1 : i];
    const REAL wr = row[i + 1 < nx ? i + 1 : i];
    return (REAL)g->lx[j] * (wl - w) + (REAL)g->ux[j] * (wr - w) - (REAL)(0.5 * g->opt.r) * w;
}

static REAL SFX(op_y)(const heston_grid* g, const REAL* W, int i, int j)
{
    const int nx = g->nx, ny = g->ny;
    const REAL w = W[(size_t)j * nx + i];
    const REAL wd = W[(size_t)(j > 0 ? j - 1 : j) * nx + i];
    const REAL wu = W[(size_t)(j + 1 < ny ? j + 1 : j) * nx + i];
    return (REAL)g->ly[j] * (wd - w) + (REAL)g->uy[j] * (wu - w) - (REAL)(0.5 * g->opt.r) * w;
}

/* rho var gx(gy(W)): central differences, one-sided (span 1) at an edge, as the engine's gx/gy. */
static REAL SFX(gy_at)(const heston_grid* g, const REAL* W, int i, int j)
{
    const int nx = g->nx, ny = g->ny;
    const int lo = j > 0 ? j - 1 : j, hi = j + 1 < ny ? j + 1 : j;
    if (hi == lo) {
        return 0;
    }
    return (W[(size_t)hi * nx + i] - W[(size_t)lo * nx + i]) / (REAL)(g->h * (hi - lo));
}

static REAL SFX(op_mixed)(const heston_grid* g, const REAL* W, int i, int j)
{
    const int nx = g->nx;
    const int lo = i > 0 ? i - 1 : i, hi = i + 1 < nx ? i + 1 : i;
    if (hi == lo) {
        return 0;
    }
    const REAL d = (SFX(gy_at)(g, W, hi, j) - SFX(gy_at)(g, W, lo, j)) / (REAL)(g->h * (hi - lo));
    return (REAL)(g->opt.rho * g->var[j]) * d;
}

/* Thomas on one line of n unknowns with constant a, b, c, the edges folded onto the diagonal (the edge cell repeats
 * past the ends: src/field/Law.cpp:1098-1138). x may alias d; cp is scratch of n. */
static void SFX(thomas)(int n, REAL a, REAL b, REAL c, const REAL* d, size_t ds, REAL* x, size_t xs, REAL* cp, REAL* dp)
{
    for (int i = 0; i < n; ++i) {
        const REAL bi = b + (i == 0 ? a : 0) + (i + 1 == n ? c : 0);
        const REAL piv = i == 0 ? bi : bi - a * cp[i - 1];
        cp[i] = i + 1 == n ? 0 : c / piv;
        dp[i] = (d[(size_t)i * ds] - (i == 0 ? 0 : a * dp[i - 1])) / piv;
    }
    REAL next = 0;
    for (int i = n - 1; i >= 0; --i) {
        next = dp[i] - cp[i] * next;
        x[(size_t)i * xs] = next;
    }
}

/* (I - th dt A1)^-1 along every row of R, into X (may alias R). */
static void SFX(solve_x)(const heston_grid* g, double th_dt, const REAL* R, REAL* X, REAL* cp, REAL* dp)
{
    for (int j = 0; j < g->ny; ++j) {
        const double cx = -g->lx[j] - g->ux[j] - 0.5 * g->opt.r;
        SFX(thomas)(g->nx, (REAL)(-th_dt * g->lx[j]), (REAL)(1.0 - th_dt * cx), (REAL)(-th_dt * g->ux[j]), R + (size_t)j * g->nx, 1,
            X + (size_t)j * g->nx, 1, cp, dp);
    }
}

/* (I - th dt A2)^-1 down every column: the coefficients change with the row, so Thomas is written out here. */
static void SFX(solve_y)(const heston_grid* g, double th_dt, const REAL* R, REAL* X, REAL* cp, REAL* dp)
{
    const int nx = g->nx, ny = g->ny;
    for (int i = 0; i < nx; ++i) {
        for (int j = 0; j < ny; ++j) {
            const double cy = -g->ly[j] - g->uy[j] - 0.5 * g->opt.r;
            const REAL a = (REAL)(-th_dt * g->ly[j]), c = (REAL)(-th_dt * g->uy[j]);
            const REAL b = (REAL)(1.0 - th_dt * cy) + (j == 0 ? a : 0) + (j + 1 == ny ? c : 0);
            const REAL piv = j == 0 ? b : b - a * cp[j - 1];
            cp[j] = j + 1 == ny ? 0 : c / piv;
            dp[j] = (R[(size_t)j * nx + i] - (j == 0 ? 0 : a * dp[j - 1])) / piv;
        }
        REAL next = 0;
        for (int j = ny - 1; j >= 0; --j) {
            next = dp[j] - cp[j] * next;
            X[(size_t)j * nx + i] = next;
        }
    }
}

/* `steps` steps of one option in place. scheme 0: Craig-Sneyd, 1: Hundsdorfer-Verwer; th: the scheme's theta. */
static void SFX(run)(const heston_grid* g, REAL* U, int steps, int scheme, double th)
{
    const size_t n = (size_t)g->nx * (size_t)g->ny;
    const double dt = g->opt.maturity / steps;
    const double th_dt = th * dt;
    REAL* rhs = (REAL*)malloc(n * sizeof(REAL));
    REAL* d2 = (REAL*)malloc(n * sizeof(REAL));
    REAL* e = (REAL*)malloc(n * sizeof(REAL));
    const int line = g->nx > g->ny ? g->nx : g->ny;
    REAL* cp = (REAL*)malloc((size_t)line * sizeof(REAL));
    REAL* dp = (REAL*)malloc((size_t)line * sizeof(REAL));
    for (int s = 0; s < steps; ++s) {
        /* D1 = (I - th dt A1)^-1 dt F(U); D2 = (I - th dt A2)^-1 D1 */
        for (int j = 0; j < g->ny; ++j) {
            for (int i = 0; i < g->nx; ++i) {
                const REAL f = SFX(op_mixed)(g, U, i, j) + SFX(op_x)(g, U, i, j) + SFX(op_y)(g, U, i, j);
                rhs[(size_t)j * g->nx + i] = (REAL)dt * f;
            }
        }
        SFX(solve_x)(g, th_dt, rhs, rhs, cp, dp);
        SFX(solve_y)(g, th_dt, rhs, d2, cp, dp);
        /* CS: E1 = (I - th dt A1)^-1 (dt/2 A0 D2). HV: E1 = (I - th dt A1)^-1 (dt F(U) - D2 + dt/2 F(D2)). */
        for (int j = 0; j < g->ny; ++j) {
            for (int i = 0; i < g->nx; ++i) {
                const size_t q = (size_t)j * g->nx + i;
                if (scheme == 0) {
                    rhs[q] = (REAL)(0.5 * dt) * SFX(op_mixed)(g, d2, i, j);
                } else {
                    const REAL fu = SFX(op_mixed)(g, U, i, j) + SFX(op_x)(g, U, i, j) + SFX(op_y)(g, U, i, j);
                    const REAL fd = SFX(op_mixed)(g, d2, i, j) + SFX(op_x)(g, d2, i, j) + SFX(op_y)(g, d2, i, j);
                    rhs[q] = (REAL)dt * fu - d2[q] + (REAL)(0.5 * dt) * fd;
                }
            }
        }
        SFX(solve_x)(g, th_dt, rhs, rhs, cp, dp);
        SFX(solve_y)(g, th_dt, rhs, e, cp, dp);
        for (size_t q = 0; q < n; ++q) {
            U[q] = U[q] + (d2[q] + e[q]);
        }
    }
    free(rhs);
    free(d2);
    free(e);
    free(cp);
    free(dp);
}
