/* Batched Heston ADI on the card: C ABI.
 *
 * The reference prices a batch of Heston options by running its ADI law (bench/laws/heston_adi.law:1-51) with one
 * option a z-plane, every plane the same lattice, one kernel launch a stage (bench/finance.py:253-288,
 * bench/results/finance.md "A batch: thousands of options in one run"). This library steps options of uneven lattice
 * sizes together: four batched passes a step (two row sweeps, two column sweeps, the increments' delta form), Craig-Sneyd
 * or Hundsdorfer-Verwer, fp32 on the card, every step of a run capturable as one graph. An option is 8 doubles: kappa,
 * theta, sigma, rho, r, strike, maturity, v0; its lattice is ny rows of nx cells, x = ln S on ln 100 +- 3 (h = 6 / nx),
 * var = sigma (j + 1/2) h; its price is read at (ln 100, v0) by 4-point Lagrange interpolation along each axis.
 *
 * Entry points return 0 on success and -1 on failure with the reason in fe_heston_error(). */
#ifndef FE_ADI_H
#define FE_ADI_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fe_heston fe_heston;

/* A batch of n options (nx[o] x ny[o] each; nx a power of two in 32..1024, ny >= 4), `steps` steps to each option's
 * maturity, scheme 0 Craig-Sneyd or 1 Hundsdorfer-Verwer with its theta (1/2, or 1/2 + sqrt(3)/6); the lattices hold
 * the payoffs. NULL with the reason in `why`. No counterpart in the reference, whose batches are z-planes of one law
 * (bench/finance.py:253-288). */
fe_heston* fe_heston_create(int n, const int* nx, const int* ny, const double* options, int steps, int scheme, double theta,
    char* why, size_t why_bytes);
void fe_heston_free(fe_heston* b);
const char* fe_heston_error(fe_heston* b);
long long fe_heston_cells(fe_heston* b);
/* Every lattice back to its payoff (queued on the library's stream). */
int fe_heston_reset(fe_heston* b);
/* Queues `count` steps (four passes each, one launch a pass a lattice class). */
int fe_heston_step(fe_heston* b, int count);
/* Captures `count` steps as one graph (replacing any before); fe_heston_launch runs it. */
int fe_heston_capture(fe_heston* b, int count);
int fe_heston_launch(fe_heston* b);
int fe_heston_sync(fe_heston* b);
/* Every option's price (waits for the card). */
int fe_heston_prices(fe_heston* b, double* out);
/* One option's lattice (ny * nx floats, x fastest; waits for the card). */
int fe_heston_lattice(fe_heston* b, int option, float* out);
/* Lattice classes (options of one nx) and their launch shapes: lanes a row in the row sweeps (16, 32, 64: the
 * Thomas-against-PCR split, nx / lanes cells a lane in 2..16), waves a block there (1, 2, 4), threads a block in the
 * column sweeps (64, 128, 256; -64: 64 threads keeping the forward sweep's results in LDS instead of HBM, for classes
 * of at most 256 rows; -8, -16: the column split over waves, that many rows a thread in registers and the chunks'
 * equations solved through LDS, 64 columns by ceil(rows / 8 or 16) waves a block, for classes of at most 16 such
 * chunks). Setting a shape drops a captured graph. */
int fe_heston_classes(fe_heston* b);
int fe_heston_get_shape(fe_heston* b, int cls, int* nx, int* lanes, int* row_waves, int* col_threads);
int fe_heston_set_shape(fe_heston* b, int cls, int lanes, int row_waves, int col_threads);
/* The autotuner: for each class, times `reps` steps of that class with every row-sweep split and block, keeps the
 * fastest, then every column-sweep block; the choice as JSON in `report`. Leaves every lattice at its payoff and drops
 * a captured graph. */
int fe_heston_tune(fe_heston* b, int reps, char* report, size_t report_bytes);

#ifdef __cplusplus
}
#endif

#endif
