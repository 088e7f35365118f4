/*
 * oracle_driver.c -- TEST INFRASTRUCTURE: runs the oracle (lbm_oracle.c,
 * linked in directly) through every code path on small lattices, for the
 * AddressSanitizer / UndefinedBehaviorSanitizer and ThreadSanitizer builds
 * of the oracle (its OpenMP list sweeps with several threads).
 *
 * Cases: lid cavity (BGK, Zou-He velocity lid), channel with a disc (TRT,
 * Zou-He velocity inlet, Zou-He pressure outlet, lift/drag, reductions),
 * open boundaries on every side with the no-op outlet, partially saturated
 * bounce-back obstacles, and the by-value copies of the timing baseline.
 * Exit status 0 when every run stays finite.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct orc_lattice orc_lattice;
orc_lattice *orc_create(int64_t nr, int64_t nc, const uint8_t *cls, int collision, double tau, double tau_even,
                        double tau_odd, int inlet_policy, int outlet_policy, int *err);
void orc_destroy(orc_lattice *L);
int64_t orc_count(const orc_lattice *L, int which);
void orc_get_list(const orc_lattice *L, int which, int64_t *ij, int32_t *tag);
void orc_set_open_values(orc_lattice *L, int which, const double *ux, const double *uy, const double *rho);
void orc_set_threads(int n);
void orc_init(orc_lattice *L);
void orc_step(orc_lattice *L, int64_t n);
void orc_get_populations(const orc_lattice *L, double *out);
void orc_get_after_collision(const orc_lattice *L, double *out);
void orc_get_macro(const orc_lattice *L, double *rho, double *u);
void orc_lift_drag(const orc_lattice *L, const int64_t *ij, const int32_t *masks, int64_t npts, double *out,
                   double *per_point);
void orc_reduce(const orc_lattice *L, double *out);
void orc_set_by_value_copies(orc_lattice *L, int on);
void orc_set_partial_bounce_back(orc_lattice *L, double tau, const double *b);

enum { FLUID = 0, BOUNDARY = 2, INLET = 3, OUTLET = 4, OBSTACLE = 5 };
enum { BGK = 0, TRT = 1 };
enum { VELOCITY = 0, PRESSURE = 1, NOOP = 2 };
enum { L_FLUID, L_WALL, L_INLET, L_OUTLET, L_OBSTACLE };

static uint8_t *box(int64_t nr, int64_t nc)
{
    uint8_t *c = calloc((size_t)(nr * nc), 1);
    for (int64_t j = 0; j < nc; ++j)
        for (int64_t i = 0; i < nr; ++i)
            if (i == 0 || j == 0 || i == nr - 1 || j == nc - 1) c[i + nr * j] = BOUNDARY;
    return c;
}

static void disc(uint8_t *c, int64_t nr, int64_t ci, int64_t cj, int64_t r)
{
    for (int64_t i = ci - r; i <= ci + r; ++i)
        for (int64_t j = cj - r; j <= cj + r; ++j)
            if ((i - ci) * (i - ci) + (j - cj) * (j - cj) <= r * r) c[i + nr * j] = OBSTACLE;
}

static int run(const char *name, int64_t nr, int64_t nc, uint8_t *cls, int coll, int pin, int pout, int psbb,
               int by_value, int steps)
{
    int err = 0;
    orc_lattice *L = orc_create(nr, nc, cls, coll, 0.8, 0.7, 1.2, pin, pout, &err);
    if (!L) { fprintf(stderr, "%s: create failed (%d)\n", name, err); return 1; }
    for (int w = 0; w < 2; ++w) {
        int64_t n = orc_count(L, L_INLET + w);
        double *ux = malloc(sizeof(double) * (size_t)(n + 1)), *uy = malloc(sizeof(double) * (size_t)(n + 1)),
               *rho = malloc(sizeof(double) * (size_t)(n + 1));
        for (int64_t q = 0; q < n; ++q) { ux[q] = 0.03 + 0.001 * (double)(q % 7); uy[q] = 0.0; rho[q] = w ? 1.0 : 1.001; }
        orc_set_open_values(L, w, ux, uy, rho);
        free(ux); free(uy); free(rho);
    }
    int64_t nob = orc_count(L, L_OBSTACLE);
    if (psbb) {
        double *b = malloc(sizeof(double) * (size_t)(nob + 1));
        for (int64_t q = 0; q < nob; ++q) b[q] = (double)(q % 5) / 4.0;
        orc_set_partial_bounce_back(L, 0.7, b);
        free(b);
    }
    orc_set_by_value_copies(L, by_value);
    orc_init(L);
    orc_step(L, steps);
    size_t n = (size_t)(nr * nc);
    double *p = malloc(9 * n * sizeof(double)), *a = malloc(9 * n * sizeof(double));
    double *rho = malloc(n * sizeof(double)), *u = malloc(2 * n * sizeof(double));
    orc_get_populations(L, p);
    orc_get_after_collision(L, a);
    orc_get_macro(L, rho, u);
    double red[5], ld[2] = {0, 0};
    orc_reduce(L, red);
    int64_t *ij = malloc(sizeof(int64_t) * (size_t)(2 * nob + 2));
    int32_t *tag = malloc(sizeof(int32_t) * (size_t)(nob + 1));
    double *pp = malloc(sizeof(double) * (size_t)(2 * nob + 2));
    orc_get_list(L, L_OBSTACLE, ij, tag);
    orc_lift_drag(L, ij, tag, nob, ld, pp);
    orc_lift_drag(L, NULL, NULL, nob, ld, NULL);
    int bad = !isfinite(red[0]) || !isfinite(ld[0]) || !isfinite(ld[1]);
    printf("%s: mass %.12f lift %.6e drag %.6e%s\n", name, red[0], ld[0], ld[1], bad ? " NOT FINITE" : "");
    free(p); free(a); free(rho); free(u); free(ij); free(tag); free(pp);
    orc_destroy(L);
    return bad;
}

int main(int argc, char **argv)
{
    int threads = argc > 1 ? atoi(argv[1]) : 4;
    orc_set_threads(threads);
    int bad = 0;
    {
        const int64_t nr = 33, nc = 37;
        uint8_t *c = box(nr, nc);
        for (int64_t j = 1; j < nc - 1; ++j) c[0 + nr * j] = INLET;
        bad |= run("cavity BGK", nr, nc, c, BGK, VELOCITY, NOOP, 0, 0, 60);
        free(c);
    }
    {
        const int64_t nr = 40, nc = 90;
        uint8_t *c = box(nr, nc);
        for (int64_t i = 1; i < nr - 1; ++i) { c[i] = INLET; c[i + nr * (nc - 1)] = OUTLET; }
        disc(c, nr, 20, 30, 6);
        bad |= run("channel TRT", nr, nc, c, TRT, VELOCITY, PRESSURE, 0, 0, 60);
        bad |= run("channel PSBB", nr, nc, c, BGK, VELOCITY, NOOP, 1, 0, 60);
        bad |= run("channel by-value copies", nr, nc, c, TRT, PRESSURE, PRESSURE, 0, 1, 20);
        free(c);
    }
    {
        const int64_t nr = 30, nc = 30;
        uint8_t *c = box(nr, nc);
        for (int64_t j = 1; j < nc - 1; ++j) { c[0 + nr * j] = INLET; c[(nr - 1) + nr * j] = OUTLET; }
        for (int64_t i = 1; i < nr - 1; ++i) { c[i] = INLET; c[i + nr * (nc - 1)] = OUTLET; }
        c[0] = INLET;
        disc(c, nr, 14, 15, 4);
        bad |= run("all sides", nr, nc, c, TRT, VELOCITY, NOOP, 0, 0, 60);
        free(c);
    }
    return bad;
}
