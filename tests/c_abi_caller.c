/* A plain C caller of libparallelkde_hip, through include/parallelkde_hip.h
 * only: what a C or C++ host (or any FFI) does to use the :gradepro path.
 *
 *   c_abi_caller OUT_FILE
 *
 * Exit 0 when every check passes, 3 when no GPU is usable (PKDE_NO_DEVICE;
 * nothing else is run), 1 on a failed check (the message names it).
 *
 * Runs, on seeded synthetic samples drawn here:
 *   - the plain form (pkde_gradepro_estimate) on a 1-D 512-point grid over
 *     [-5, 5] (n = 10000) and a 2-D 64 x 64 grid over [-6, 6]^2 (n = 20000),
 *     in-kernel Philox draws keyed by the plan's rng_seed; checks that the
 *     density is finite and non-negative with unit mass and the chosen rung
 *     in [0, S];
 *   - the batched form (pkde_gradepro_estimate_batch) and the serving form
 *     (pkde_server_*) on three requests of different dimension (1-D 256,
 *     2-D 64^2, 3-D 32^3): each result equal bit for bit to the plain form's;
 *   - the error paths: a sample outside the grid (PKDE_OUT_OF_GRID), an axis
 *     of 5000 points (PKDE_UNSUPPORTED).
 * OUT_FILE receives, for the 1-D and the 2-D plain estimates in turn:
 * int32 ndims, int64 n, uint64 seed, int64 G, the samples (ndims x n
 * doubles, coordinate fastest), the density (G doubles, axis 1 fastest) and
 * the chosen rung (G int16), for the caller's comparison with the oracle. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "parallelkde_hip.h"

static int failures = 0;

#define CHECK(cond, ...)                                     \
  do {                                                       \
    if (!(cond)) {                                           \
      fprintf(stderr, "check failed (%s:%d): ", __FILE__, __LINE__); \
      fprintf(stderr, __VA_ARGS__);                          \
      fprintf(stderr, " [%s]\n", pkde_last_error());         \
      ++failures;                                            \
    }                                                        \
  } while (0)

/* xorshift64* uniforms in (0, 1) and Box-Muller normals */
static uint64_t rng_state = 0x9E3779B97F4A7C15ull;
static double uniform01(void) {
  rng_state ^= rng_state >> 12;
  rng_state ^= rng_state << 25;
  rng_state ^= rng_state >> 27;
  const uint64_t v = rng_state * 0x2545F4914F6CDD1Dull;
  return ((double)(v >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}
static double normal01(void) {
  const double u = uniform01(), v = uniform01();
  return sqrt(-2.0 * log(u)) * cos(6.283185307179586 * v);
}

/* n draws of a d-dimensional N(mu_c, s^2 I) mixture inside |x_i| < bound */
static double* draw(int d, int64_t n, double bound, int mixture) {
  double* x = (double*)malloc(sizeof(double) * (size_t)d * (size_t)n);
  for (int64_t j = 0; j < n; ++j) {
    for (;;) {
      const double mu = mixture ? (uniform01() < 0.5 ? -1.5 : 1.5) : 0.0;
      const double s = mixture ? 0.7 : 1.0;
      int ok = 1;
      for (int i = 0; i < d; ++i) {
        x[j * d + i] = mu + s * normal01();
        if (fabs(x[j * d + i]) >= bound) ok = 0;
      }
      if (ok) break;
    }
  }
  return x;
}

typedef struct {
  int ndims;
  int64_t n_axis[3];
  double lo[3], hi[3], h[3];
  int64_t G;
} grid_t;

static grid_t make_grid(int d, int64_t pts, double lo, double hi) {
  grid_t g;
  memset(&g, 0, sizeof(g));
  g.ndims = d;
  g.G = 1;
  for (int i = 0; i < d; ++i) {
    g.n_axis[i] = pts;
    g.lo[i] = lo;
    g.hi[i] = hi;
    g.h[i] = (hi - lo) / (double)(pts - 1);
    g.G *= pts;
  }
  return g;
}

/* the plain form on one grid; returns the status */
static int plain(pkde_ctx* ctx, const grid_t* g, uint64_t seed, const double* x, int64_t n,
                 double* dens, int16_t* chosen, pkde_stats* st) {
  pkde_gradepro_params p;
  pkde_gradepro_params_default(&p);
  p.rng_seed = seed;
  pkde_plan* plan = NULL;
  int rc = pkde_gradepro_plan_create(ctx, g->ndims, g->n_axis, g->lo, g->hi, g->h, &p, &plan);
  if (rc) return rc;
  rc = pkde_gradepro_estimate(plan, x, n, NULL, dens, chosen, st);
  pkde_gradepro_plan_destroy(plan);
  return rc;
}

static void check_density(const char* name, const grid_t* g, const double* dens,
                          const int16_t* chosen, int S) {
  double mass = 0.0, cell = 1.0;
  int bad = 0;
  for (int i = 0; i < g->ndims; ++i) cell *= g->h[i];
  for (int64_t k = 0; k < g->G; ++k) {
    if (!isfinite(dens[k]) || dens[k] < 0.0) ++bad;
    if (chosen[k] < 0 || chosen[k] > S) ++bad;
    mass += dens[k] * cell;
  }
  CHECK(bad == 0, "%s: %d points non-finite, negative or with a rung outside [0, %d]", name,
        bad, S);
  CHECK(fabs(mass - 1.0) < 0.2, "%s: mass %.6f", name, mass);
  printf("%s: G=%lld mass %.6f\n", name, (long long)g->G, mass);
}

static void write_case(FILE* f, const grid_t* g, uint64_t seed, const double* x, int64_t n,
                       const double* dens, const int16_t* chosen) {
  const int32_t d = g->ndims;
  fwrite(&d, sizeof(d), 1, f);
  fwrite(&n, sizeof(n), 1, f);
  fwrite(&seed, sizeof(seed), 1, f);
  fwrite(&g->G, sizeof(g->G), 1, f);
  fwrite(x, sizeof(double), (size_t)d * (size_t)n, f);
  fwrite(dens, sizeof(double), (size_t)g->G, f);
  fwrite(chosen, sizeof(int16_t), (size_t)g->G, f);
}

int main(int argc, char** argv) {
  if (argc < 2) {
    fprintf(stderr, "usage: %s OUT_FILE\n", argv[0]);
    return 2;
  }
  CHECK(pkde_version() == PKDE_ABI_VERSION, "ABI version %d, header %d", pkde_version(),
        PKDE_ABI_VERSION);
  pkde_ctx* ctx = NULL;
  int rc = pkde_context_create(0, &ctx);
  if (rc == PKDE_NO_DEVICE) {
    printf("no usable GPU: %s\n", pkde_last_error());
    return 3;
  }
  CHECK(rc == PKDE_OK, "context: status %d", rc);
  if (rc) return 1;
  FILE* out = fopen(argv[1], "wb");
  CHECK(out != NULL, "cannot open %s", argv[1]);
  if (!out) return 1;

  /* ---- plain form: 1-D and 2-D ---- */
  const grid_t g1 = make_grid(1, 512, -5.0, 5.0), g2 = make_grid(2, 64, -6.0, 6.0);
  const int64_t n1 = 10000, n2 = 20000;
  double* x1 = draw(1, n1, 4.5, 0);
  double* x2 = draw(2, n2, 5.5, 1);
  double* d1 = (double*)malloc(sizeof(double) * (size_t)g1.G);
  double* d2 = (double*)malloc(sizeof(double) * (size_t)g2.G);
  int16_t* c1 = (int16_t*)malloc(sizeof(int16_t) * (size_t)g1.G);
  int16_t* c2 = (int16_t*)malloc(sizeof(int16_t) * (size_t)g2.G);
  pkde_stats st;
  rc = plain(ctx, &g1, 101, x1, n1, d1, c1, &st);
  CHECK(rc == PKDE_OK, "1-D plain estimate: status %d", rc);
  if (!rc) {
    check_density("1-D 512 plain", &g1, d1, c1, st.n_rungs);
    printf("  %d rungs, %.3f ms on the device (ladder %.3f ms)\n", st.n_rungs, st.ms_total,
           st.ms_ladder);
    write_case(out, &g1, 101, x1, n1, d1, c1);
  }
  rc = plain(ctx, &g2, 202, x2, n2, d2, c2, &st);
  CHECK(rc == PKDE_OK, "2-D plain estimate: status %d", rc);
  if (!rc) {
    check_density("2-D 64x64 plain", &g2, d2, c2, st.n_rungs);
    write_case(out, &g2, 202, x2, n2, d2, c2);
  }
  fclose(out);

  /* ---- batched and serving forms against the plain form ---- */
  enum { NREQ = 3 };
  grid_t gr[NREQ] = {make_grid(1, 256, -5.0, 5.0), make_grid(2, 64, -6.0, 6.0),
                     make_grid(3, 32, -4.0, 4.0)};
  const int64_t nr[NREQ] = {5000, 8000, 12000};
  const double bound[NREQ] = {4.5, 5.5, 3.5};
  double* xr[NREQ];
  double* ref[NREQ];
  double* bat[NREQ];
  double* srvd[NREQ];
  int16_t* refc[NREQ];
  int16_t* batc[NREQ];
  int16_t* srvc[NREQ];
  pkde_request req[NREQ];
  for (int q = 0; q < NREQ; ++q) {
    xr[q] = draw(gr[q].ndims, nr[q], bound[q], q == 1);
    const size_t G = (size_t)gr[q].G;
    ref[q] = (double*)malloc(sizeof(double) * G);
    bat[q] = (double*)malloc(sizeof(double) * G);
    srvd[q] = (double*)malloc(sizeof(double) * G);
    refc[q] = (int16_t*)malloc(sizeof(int16_t) * G);
    batc[q] = (int16_t*)malloc(sizeof(int16_t) * G);
    srvc[q] = (int16_t*)malloc(sizeof(int16_t) * G);
    rc = plain(ctx, &gr[q], 500 + q, xr[q], nr[q], ref[q], refc[q], NULL);
    CHECK(rc == PKDE_OK, "plain request %d: status %d", q, rc);
    memset(&req[q], 0, sizeof(req[q]));
    req[q].ndims = gr[q].ndims;
    for (int i = 0; i < gr[q].ndims; ++i) {
      req[q].n_per_axis[i] = gr[q].n_axis[i];
      req[q].lo[i] = gr[q].lo[i];
      req[q].hi[i] = gr[q].hi[i];
      req[q].h[i] = gr[q].h[i];
    }
    req[q].samples = xr[q];
    req[q].n = nr[q];
    pkde_gradepro_params_default(&req[q].params);
    req[q].params.rng_seed = 500 + q;
    req[q].density_out = bat[q];
    req[q].chosen_out = batc[q];
  }
  int32_t status[NREQ];
  rc = pkde_gradepro_estimate_batch(ctx, NREQ, req, status);
  CHECK(rc == PKDE_OK, "batch: status %d", rc);
  for (int q = 0; q < NREQ; ++q) {
    const size_t G = (size_t)gr[q].G;
    CHECK(status[q] == PKDE_OK, "batch request %d: status %d", q, status[q]);
    CHECK(memcmp(bat[q], ref[q], G * sizeof(double)) == 0 &&
              memcmp(batc[q], refc[q], G * sizeof(int16_t)) == 0,
          "batch request %d (%d-D) differs from the plain form", q, gr[q].ndims);
  }
  printf("batch of %d requests (1-D 256, 2-D 64^2, 3-D 32^3): equal to the plain form\n", NREQ);

  pkde_server_cfg cfg;
  cfg.max_inflight = 2;
  cfg.memory_budget = 0; /* default: 80 % of free device memory */
  pkde_server* srv = NULL;
  rc = pkde_server_create(ctx, &cfg, &srv);
  CHECK(rc == PKDE_OK, "server: status %d", rc);
  if (!rc) {
    uint64_t tickets[NREQ];
    for (int q = 0; q < NREQ; ++q) {
      req[q].density_out = srvd[q];
      req[q].chosen_out = srvc[q];
      rc = pkde_server_submit(srv, &req[q], &tickets[q]);
      CHECK(rc == PKDE_OK, "submit %d: status %d", q, rc);
    }
    for (int q = 0; q < NREQ; ++q) {
      const size_t G = (size_t)gr[q].G;
      rc = pkde_server_wait(srv, tickets[q]);
      CHECK(rc == PKDE_OK, "served request %d: status %d", q, rc);
      CHECK(memcmp(srvd[q], ref[q], G * sizeof(double)) == 0 &&
                memcmp(srvc[q], refc[q], G * sizeof(int16_t)) == 0,
            "served request %d (%d-D) differs from the plain form", q, gr[q].ndims);
    }
    pkde_server_stats ss;
    rc = pkde_server_stats_get(srv, &ss);
    CHECK(rc == PKDE_OK && ss.completed == NREQ, "server stats: status %d, completed %llu", rc,
          (unsigned long long)ss.completed);
    printf("served %d requests: equal to the plain form; p50 %.1f ms, p99 %.1f ms\n", NREQ,
           ss.p50_ms, ss.p99_ms);
    pkde_server_destroy(srv);
  }

  /* ---- error paths ---- */
  {
    double xo[4] = {0.0, 1.0, 5.5, -1.0}; /* 5.5 lies outside [-5, 5] */
    pkde_stats so;
    rc = plain(ctx, &g1, 1, xo, 4, d1, c1, &so);
    CHECK(rc == PKDE_OUT_OF_GRID, "out-of-grid sample: status %d", rc);
    grid_t big = make_grid(1, 5000, -5.0, 5.0);
    pkde_gradepro_params p;
    pkde_gradepro_params_default(&p);
    pkde_plan* plan = NULL;
    rc = pkde_gradepro_plan_create(ctx, 1, big.n_axis, big.lo, big.hi, big.h, &p, &plan);
    CHECK(rc == PKDE_UNSUPPORTED && plan == NULL, "5000-point axis: status %d", rc);
    printf("error paths: out of grid and unsupported axis size reported\n");
  }

  for (int q = 0; q < NREQ; ++q) {
    free(xr[q]); free(ref[q]); free(bat[q]); free(srvd[q]);
    free(refc[q]); free(batc[q]); free(srvc[q]);
  }
  free(x1); free(x2); free(d1); free(d2); free(c1); free(c2);
  pkde_context_destroy(ctx);
  if (failures) {
    fprintf(stderr, "%d check(s) failed\n", failures);
    return 1;
  }
  printf("all checks passed\n");
  return 0;
}
