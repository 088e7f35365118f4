/*
 * lbm_oracle.c -- CPU restatement of the LLALBM D2Q9 time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: only the test
 * suite, the smoke check and the CPU-baseline leg of the benchmark may load
 * it, and only as the checker.  The product path (the HIP library) never
 * links or calls it.
 *
 * It restates the reference's *Serial* policy semantics in the reference's
 * own structure: list-driven sweeps over node-coordinate vectors, push
 * streaming, three full population tensors in the reference's ColMajor
 * layout (element (i,j,d) at i + NR*(j + NC*d)), and the same floating-point
 * expression forms (build with -ffp-contract=off).  The list loops carry
 * OpenMP pragmas so the same code doubles as the "reference-structured
 * OpenMP" CPU timing baseline; every loop body writes disjoint locations
 * for wet nodes, so wet-node results do not depend on the thread count.
 *
 * Reference files followed (paths relative to the reference root):
 *   step order           llalbm/src/core/lattices/Lattice.hpp:288-438
 *   equilibrium / macro  llalbm/src/core/equilibriums/DefaultEquilibrium.hpp:54-155
 *   BGK collide/stream   llalbm/src/core/collisions/BGKCollision.hpp:119-304
 *   TRT collide/stream   llalbm/src/core/collisions/TRTCollision.hpp:146-343
 *   wall / obstacle BB   llalbm/src/core/boundaries/BounceBackPolicy.hpp:68-187
 *   Zou-He velocity      llalbm/src/core/boundaries/ZouHePolicy.hpp:58-123
 *   open (no-op) outlet  llalbm/src/core/boundaries/OpenBoundaryPolicy.hpp:58-61
 *   inlet/outlet values  llalbm/src/core/initializers/InletOutletInitializer.hpp:90-129
 *   node identification  llalbm/src/utils/analysis/NodeIdentification.hpp:45-242
 *   list construction    llalbm/src/utils/readers/LatticeReader.hpp:136-205 (row-major scan)
 *   lift / drag          llalbm/src/utils/analysis/FlowAnalyzer.hpp:348-451
 *   partially saturated  llalbm/src/core/boundaries/PSBounceBackPolicy.hpp:177-297
 *   bounce-back          (weights: Lattice.hpp:245-278, ObstaclePoint b, aliases.hpp:340-357)
 *
 * Extensions with no reference counterpart (defined here, identically in
 * the HIP path):
 *   - Zou-He *pressure* open boundary: density prescribed, normal velocity
 *     from the Zou & He (1997) closure, tangential velocity zero; the node
 *     then takes its equilibrium from its own moments (like a fluid node).
 *   - mass / momentum reductions over the wet nodes.
 *
 * Deliberate deviation: a fluid node on the domain perimeter would make the
 * reference push out of bounds (undefined behaviour, BGKCollision.hpp:180-187);
 * here such pushes are dropped, like the guarded open-boundary stream.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* node classes = reference InputNodeType (aliases.hpp:41-48) */
enum { CLS_FLUID = 0, CLS_SOLID = 1, CLS_BOUNDARY = 2, CLS_INLET = 3, CLS_OUTLET = 4, CLS_OBSTACLE = 5 };
/* boundary types = reference btypes (aliases.hpp:28-38) */
enum { BT_NONE = 0, BT_TOP, BT_RIGHT, BT_BOTTOM, BT_LEFT, BT_RIGHT_TOP_C, BT_RIGHT_BOTTOM_C, BT_LEFT_BOTTOM_C, BT_LEFT_TOP_C };
enum { COLL_BGK = 0, COLL_TRT = 1 };
enum { OPEN_ZOUHE_VELOCITY = 0, OPEN_ZOUHE_PRESSURE = 1, OPEN_NOOP = 2 };
enum { LIST_FLUID = 0, LIST_WALL = 1, LIST_INLET = 2, LIST_OUTLET = 3, LIST_OBSTACLE = 4 };

typedef struct { int64_t i, j; int32_t tag; } node_t; /* tag: btype or 9-bit direction mask */

typedef struct orc_lattice {
    int64_t nr, nc;
    double *P, *E, *A;       /* populations, equilibrium, after-collision */
    double *rho, *u;          /* global_rho (NR,NC), global_u (NR,NC,2) */
    node_t *list[5];
    int64_t count[5];
    int collision;
    double tau, tau_even, tau_odd;
    int policy[2];            /* inlet, outlet */
    double *val_ux[2], *val_uy[2], *val_rho[2];
    int64_t step;
    int by_value;             /* timing baseline: make the reference's by-value field copies */
    double ps_tau;            /* > 0: obstacles are partially saturated bounce-back nodes */
    double *ps_b;             /* their weights, obstacle list order */
} orc_lattice;

#define P3(L, a, i, j, d) ((a)[(i) + (L)->nr * ((j) + (L)->nc * (int64_t)(d))])
#define R2(L, a, i, j) ((a)[(i) + (L)->nr * (j)])

/* ------------------------------------------------------------------ */
/* node identification                                                 */
/* ------------------------------------------------------------------ */

/* NodeIdentification.hpp:45-110 */
static int identify_type(int64_t i, int64_t j, int64_t nr, int64_t nc)
{
    if (i == 0) {
        if (j == 0) return BT_LEFT_TOP_C;
        if (j == nc - 1) return BT_RIGHT_TOP_C;
        return BT_TOP;
    }
    if (i == nr - 1) {
        if (j == 0) return BT_LEFT_BOTTOM_C;
        if (j == nc - 1) return BT_RIGHT_BOTTOM_C;
        return BT_BOTTOM;
    }
    if (j == 0) return BT_LEFT;
    if (j == nc - 1) return BT_RIGHT;
    return BT_NONE;
}

/* push offsets (row, col) of direction d, from the stream at BGKCollision.hpp:179-187 */
static const int OFF_I[9] = {0, 0, -1, 0, 1, -1, -1, 1, 1};
static const int OFF_J[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};

static int is_wet_class(int c) { return c == CLS_FLUID || c == CLS_INLET || c == CLS_OUTLET; }

/* Obstacle direction mask, NodeIdentification.hpp:137-211: bit d set iff the
 * neighbour at the push target of d is fluid, inlet or outlet.  The
 * reference finds neighbours by linear search in the node lists; a lookup in
 * the class plane answers the same question. */
static int32_t obstacle_mask(const uint8_t *cls, int64_t nr, int64_t nc, int64_t i, int64_t j)
{
    int32_t m = 0;
    for (int d = 1; d < 9; ++d) {
        int64_t ii = i + OFF_I[d], jj = j + OFF_J[d];
        if (ii < 0 || ii >= nr || jj < 0 || jj >= nc) continue;
        if (is_wet_class(cls[ii + nr * jj])) m |= 1 << d;
    }
    return m;
}

/* ------------------------------------------------------------------ */
/* lifetime                                                            */
/* ------------------------------------------------------------------ */

void orc_destroy(orc_lattice *L)
{
    if (!L) return;
    free(L->P); free(L->E); free(L->A); free(L->rho); free(L->u);
    for (int k = 0; k < 5; ++k) free(L->list[k]);
    for (int k = 0; k < 2; ++k) { free(L->val_ux[k]); free(L->val_uy[k]); free(L->val_rho[k]); }
    free(L->ps_b);
    free(L);
}

/* Builds the node lists from a class plane (ColMajor, cls[i + NR*j]) in the
 * order the reference's file reader produces them (row-major scan,
 * LatticeReader.hpp:136-205, which also equals the std::set order of
 * ConstructionTools.hpp:94-118), then identifies inlet/outlet types and
 * obstacle masks and drops the non-wet obstacles (NodeIdentification.hpp:229-242).
 * Returns NULL on bad input (err receives a code). */
orc_lattice *orc_create(int64_t nr, int64_t nc, const uint8_t *cls, int collision, double tau,
                        double tau_even, double tau_odd, int inlet_policy, int outlet_policy, int *err)
{
    *err = 0;
    if (nr < 3 || nc < 3 || !cls) { *err = 1; return NULL; }
    orc_lattice *L = (orc_lattice *)calloc(1, sizeof(orc_lattice));
    L->nr = nr; L->nc = nc;
    L->collision = collision; L->tau = tau; L->tau_even = tau_even; L->tau_odd = tau_odd;
    L->policy[0] = inlet_policy; L->policy[1] = outlet_policy;
    size_t n = (size_t)(nr * nc);
    L->P = (double *)calloc(9 * n, sizeof(double));
    L->E = (double *)calloc(9 * n, sizeof(double));
    L->A = (double *)calloc(9 * n, sizeof(double));
    L->rho = (double *)calloc(n, sizeof(double));
    L->u = (double *)calloc(2 * n, sizeof(double));
    if (!L->P || !L->E || !L->A || !L->rho || !L->u) { orc_destroy(L); *err = 2; return NULL; }
    for (size_t k = 0; k < n; ++k) L->rho[k] = 1.0; /* Lattice.hpp:295, also before the first orc_init */

    int64_t cnt[6] = {0};
    for (size_t k = 0; k < n; ++k) {
        if (cls[k] > CLS_OBSTACLE) { orc_destroy(L); *err = 1; return NULL; }
        cnt[cls[k]]++;
    }
    const int map_cls[6] = {LIST_FLUID, -1, LIST_WALL, LIST_INLET, LIST_OUTLET, LIST_OBSTACLE};
    for (int c = 0; c < 6; ++c)
        if (map_cls[c] >= 0) L->list[map_cls[c]] = (node_t *)malloc(sizeof(node_t) * (size_t)(cnt[c] ? cnt[c] : 1));
    for (int64_t i = 0; i < nr; ++i) {
        for (int64_t j = 0; j < nc; ++j) {
            int c = cls[i + nr * j];
            int l = map_cls[c];
            if (l < 0) continue;
            node_t nd = {i, j, 0};
            if (l == LIST_INLET || l == LIST_OUTLET) nd.tag = identify_type(i, j, nr, nc);
            if (l == LIST_OBSTACLE) {
                nd.tag = obstacle_mask(cls, nr, nc, i, j);
                if (nd.tag == 0) continue; /* eliminate_non_wet_nodes */
            }
            L->list[l][L->count[l]++] = nd;
        }
    }
    for (int k = 0; k < 2; ++k) {
        int64_t m = L->count[LIST_INLET + k] ? L->count[LIST_INLET + k] : 1;
        L->val_ux[k] = (double *)calloc((size_t)m, sizeof(double));
        L->val_uy[k] = (double *)calloc((size_t)m, sizeof(double));
        L->val_rho[k] = (double *)calloc((size_t)m, sizeof(double));
        for (int64_t q = 0; q < m; ++q) L->val_rho[k][q] = 1.0;
    }
    return L;
}

int64_t orc_count(const orc_lattice *L, int which) { return (which >= 0 && which < 5) ? L->count[which] : -1; }

void orc_get_list(const orc_lattice *L, int which, int64_t *ij, int32_t *tag)
{
    for (int64_t k = 0; k < L->count[which]; ++k) {
        ij[2 * k] = L->list[which][k].i;
        ij[2 * k + 1] = L->list[which][k].j;
        if (tag) tag[k] = L->list[which][k].tag;
    }
}

/* Per-node prescribed values for the inlet (which=0) or outlet (which=1)
 * list, in list order.  They stand for the reference's update functions
 * evaluated at the coming step (InletOutletInitializer.hpp:90-129). */
void orc_set_open_values(orc_lattice *L, int which, const double *ux, const double *uy, const double *rho)
{
    int64_t m = L->count[LIST_INLET + which];
    for (int64_t k = 0; k < m; ++k) {
        if (ux) L->val_ux[which][k] = ux[k];
        if (uy) L->val_uy[which][k] = uy[k];
        if (rho) L->val_rho[which][k] = rho[k];
    }
}

void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------ */
/* single-node kernels (also exported for closed-form tests)           */
/* ------------------------------------------------------------------ */

/* DefaultEquilibrium.hpp:62-76 (the same polynomial serves both overloads) */
void orc_feq(double rho, double ux, double uy, double *e)
{
    const double two9 = (2.0 / 9.0);
    const double one18 = (1.0 / 18.0);
    const double one36 = (1.0 / 36.0);
    double u2 = ux * ux + uy * uy;
    e[0] = two9 * rho * (2 - 3 * u2);
    e[1] = one18 * rho * (2 + 6 * ux + 9 * ux * ux - 3 * u2);
    e[2] = one18 * rho * (2 + 6 * uy + 9 * uy * uy - 3 * u2);
    e[3] = one18 * rho * (2 - 6 * ux + 9 * ux * ux - 3 * u2);
    e[4] = one18 * rho * (2 - 6 * uy + 9 * uy * uy - 3 * u2);
    e[5] = one36 * rho * (1 + 3 * (ux + uy) + 9 * ux * uy + 3 * u2);
    e[6] = one36 * rho * (1 - 3 * (ux - uy) - 9 * ux * uy + 3 * u2);
    e[7] = one36 * rho * (1 - 3 * (ux + uy) + 9 * ux * uy + 3 * u2);
    e[8] = one36 * rho * (1 + 3 * (ux - uy) - 9 * ux * uy + 3 * u2);
}

/* DefaultEquilibrium.hpp:123-155 */
void orc_moments(const double *p, double *rho_out, double *ux_out, double *uy_out)
{
    double rho = p[0] + p[1] + p[2] + p[3] + p[4] + p[5] + p[6] + p[7] + p[8];
    double rhoinv = 1.0 / rho;
    *rho_out = rho;
    *ux_out = rhoinv * (p[1] + p[5] + p[8] - (p[3] + p[6] + p[7]));
    *uy_out = rhoinv * (p[2] + p[5] + p[6] - (p[4] + p[7] + p[8]));
}

/* BGKCollision.hpp:140-153 */
void orc_collide_bgk_node(double tau, const double *p, const double *e, double *out)
{
    double t = 1.0 / tau;
    double o_m_t = 1.0 - t;
    for (int d = 0; d < 9; ++d) out[d] = o_m_t * p[d] + t * e[d];
}

/* TRTCollision.hpp:155-188; opposite pairs (1,3) (2,4) (5,7) (6,8) */
void orc_collide_trt_node(double te, double to, const double *p, const double *e, double *out)
{
    static const int OPP[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
    out[0] = p[0] - (1.0 / te) * (p[0] - e[0]);
    for (int d = 1; d < 9; ++d) {
        int o = OPP[d];
        out[d] = p[d] - 0.5 * (1.0 / te) * (p[d] + p[o] - e[d] - e[o]) -
                 0.5 * (1.0 / to) * (p[d] - p[o] - e[d] + e[o]);
    }
}

/* ZouHePolicy.hpp:68-120, one node; p is modified in place.  Corner and
 * NONE types are left untouched (ZouHePolicy.hpp:119-120).  The density the
 * reference computes is written into a by-value copy and lost, so it is
 * only returned here. */
double orc_zouhe_velocity_node(int btype, double *p, double ux, double uy)
{
    const double two_thirds = 2.0 / 3.0;
    const double one_sixth = 1.0 / 6.0;
    const double one_half = 0.5;
    double rho, ru, rv;
    switch (btype) {
    case BT_TOP:
        rho = (p[0] + p[1] + p[3] + 2.0 * (p[2] + p[5] + p[6])) / (1.0 + uy);
        ru = rho * ux; rv = rho * uy;
        p[4] = p[2] - two_thirds * rv;
        p[7] = p[5] - one_sixth * rv - one_half * ru + one_half * (p[1] - p[3]);
        p[8] = p[6] - one_sixth * rv + one_half * ru - one_half * (p[1] - p[3]);
        return rho;
    case BT_BOTTOM:
        rho = (p[0] + p[1] + p[3] + 2.0 * (p[4] + p[7] + p[8])) / (1.0 - uy);
        ru = rho * ux; rv = rho * uy;
        p[2] = p[4] + two_thirds * rv;
        p[5] = p[7] + one_sixth * rv + one_half * ru - one_half * (p[1] - p[3]);
        p[6] = p[8] + one_sixth * rv - one_half * ru + one_half * (p[1] - p[3]);
        return rho;
    case BT_LEFT:
        rho = (p[0] + p[2] + p[4] + 2.0 * (p[3] + p[7] + p[6])) / (1.0 - ux);
        ru = rho * ux; rv = rho * uy;
        p[1] = p[3] + two_thirds * ru;
        p[5] = p[7] + one_sixth * ru + one_half * rv - one_half * (p[2] - p[4]);
        p[8] = p[6] + one_sixth * ru - one_half * rv + one_half * (p[2] - p[4]);
        return rho;
    case BT_RIGHT:
        rho = (p[0] + p[2] + p[4] + 2.0 * (p[1] + p[5] + p[8])) / (1.0 + ux);
        ru = rho * ux; rv = rho * uy;
        p[3] = p[1] - two_thirds * ru;
        p[7] = p[5] - one_sixth * ru - one_half * rv + one_half * (p[2] - p[4]);
        p[6] = p[8] - one_sixth * ru + one_half * rv - one_half * (p[2] - p[4]);
        return rho;
    default:
        return 0.0;
    }
}

/* Zou-He pressure closure (no reference counterpart): the known
 * populations and the prescribed density give the normal velocity, the
 * tangential velocity is zero, and the unknowns follow from the velocity
 * closure above with that (rho, u).  Returns the normal velocity. */
double orc_zouhe_pressure_node(int btype, double *p, double rho)
{
    double ux = 0.0, uy = 0.0;
    switch (btype) {
    case BT_TOP:    uy = (p[0] + p[1] + p[3] + 2.0 * (p[2] + p[5] + p[6])) / rho - 1.0; break;
    case BT_BOTTOM: uy = 1.0 - (p[0] + p[1] + p[3] + 2.0 * (p[4] + p[7] + p[8])) / rho; break;
    case BT_LEFT:   ux = 1.0 - (p[0] + p[2] + p[4] + 2.0 * (p[3] + p[7] + p[6])) / rho; break;
    case BT_RIGHT:  ux = (p[0] + p[2] + p[4] + 2.0 * (p[1] + p[5] + p[8])) / rho - 1.0; break;
    default: return 0.0;
    }
    const double two_thirds = 2.0 / 3.0;
    const double one_sixth = 1.0 / 6.0;
    const double one_half = 0.5;
    double ru = rho * ux, rv = rho * uy;
    switch (btype) {
    case BT_TOP:
        p[4] = p[2] - two_thirds * rv;
        p[7] = p[5] - one_sixth * rv - one_half * ru + one_half * (p[1] - p[3]);
        p[8] = p[6] - one_sixth * rv + one_half * ru - one_half * (p[1] - p[3]);
        return uy;
    case BT_BOTTOM:
        p[2] = p[4] + two_thirds * rv;
        p[5] = p[7] + one_sixth * rv + one_half * ru - one_half * (p[1] - p[3]);
        p[6] = p[8] + one_sixth * rv - one_half * ru + one_half * (p[1] - p[3]);
        return uy;
    case BT_LEFT:
        p[1] = p[3] + two_thirds * ru;
        p[5] = p[7] + one_sixth * ru + one_half * rv - one_half * (p[2] - p[4]);
        p[8] = p[6] + one_sixth * ru - one_half * rv + one_half * (p[2] - p[4]);
        return ux;
    default: /* BT_RIGHT */
        p[3] = p[1] - two_thirds * ru;
        p[7] = p[5] - one_sixth * ru - one_half * rv + one_half * (p[2] - p[4]);
        p[6] = p[8] - one_sixth * ru + one_half * rv - one_half * (p[2] - p[4]);
        return ux;
    }
}

/* ------------------------------------------------------------------ */
/* list sweeps (the policy methods)                                    */
/* ------------------------------------------------------------------ */

/* InletOutletInitializer.hpp:90-129: u at inlet/outlet nodes <- prescribed
 * values.  Pressure-policy lists take their velocity from the macro update. */
static void update_nodes(orc_lattice *L)
{
    for (int k = 0; k < 2; ++k) {
        if (L->policy[k] == OPEN_ZOUHE_PRESSURE) continue;
        const node_t *lst = L->list[LIST_INLET + k];
        int64_t m = L->count[LIST_INLET + k];
        for (int64_t q = 0; q < m; ++q) {
            R2(L, L->u, lst[q].i, lst[q].j) = L->val_ux[k][q];
            L->u[lst[q].i + L->nr * (lst[q].j + L->nc)] = L->val_uy[k][q];
        }
    }
}

static void macro_list(orc_lattice *L, const node_t *lst, int64_t m)
{
#pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < m; ++q) {
        int64_t i = lst[q].i, j = lst[q].j;
        double p[9], rho, ux, uy;
        for (int d = 0; d < 9; ++d) p[d] = P3(L, L->P, i, j, d);
        orc_moments(p, &rho, &ux, &uy);
        R2(L, L->rho, i, j) = rho;
        L->u[i + L->nr * j] = ux;
        L->u[i + L->nr * (j + L->nc)] = uy;
    }
}

/* DefaultEquilibrium.hpp:123-155 (fluid list), plus pressure-policy lists */
static void update_macro(orc_lattice *L)
{
    macro_list(L, L->list[LIST_FLUID], L->count[LIST_FLUID]);
    for (int k = 0; k < 2; ++k)
        if (L->policy[k] == OPEN_ZOUHE_PRESSURE) macro_list(L, L->list[LIST_INLET + k], L->count[LIST_INLET + k]);
}

/* DefaultEquilibrium.hpp:54-112 */
static void calc_equilibrium(orc_lattice *L, double *dst, const node_t *lst, int64_t m)
{
#pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < m; ++q) {
        int64_t i = lst[q].i, j = lst[q].j;
        double e[9];
        orc_feq(R2(L, L->rho, i, j), L->u[i + L->nr * j], L->u[i + L->nr * (j + L->nc)], e);
        for (int d = 0; d < 9; ++d) P3(L, dst, i, j, d) = e[d];
    }
}

/* BGKCollision.hpp:119-155,201-237 / TRTCollision.hpp:146-191,233-278 */
static void collide(orc_lattice *L, const node_t *lst, int64_t m)
{
#pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < m; ++q) {
        int64_t i = lst[q].i, j = lst[q].j;
        double p[9], e[9], o[9];
        for (int d = 0; d < 9; ++d) { p[d] = P3(L, L->P, i, j, d); e[d] = P3(L, L->E, i, j, d); }
        if (L->collision == COLL_TRT) orc_collide_trt_node(L->tau_even, L->tau_odd, p, e, o);
        else orc_collide_bgk_node(L->tau, p, e, o);
        for (int d = 0; d < 9; ++d) P3(L, L->A, i, j, d) = o[d];
    }
}

/* BGKCollision.hpp:164-189 (fluid, unconditional) and :246-304 (open
 * boundary, guarded by the domain edges).  The fluid push is guarded too,
 * see the header note. */
static void stream(orc_lattice *L, const node_t *lst, int64_t m)
{
    const int64_t nr = L->nr, nc = L->nc;
#pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < m; ++q) {
        int64_t i = lst[q].i, j = lst[q].j;
        P3(L, L->P, i, j, 0) = P3(L, L->A, i, j, 0);
        for (int d = 1; d < 9; ++d) {
            int64_t ti = i + OFF_I[d], tj = j + OFF_J[d];
            if (ti < 0 || ti >= nr || tj < 0 || tj >= nc) continue;
            P3(L, L->P, ti, tj, d) = P3(L, L->A, i, j, d);
        }
    }
}

/* BounceBackPolicy.hpp:68-133: every wall node reflects its 8 populations
 * into its in-domain neighbours, population opp(d) of the wall going into
 * slot d of the neighbour at the push target of d.  Runs serially in list
 * order: wall->wall writes are order dependent (they only ever touch wall
 * slots, which no wet node reads). */
static void bounce_back_walls(orc_lattice *L)
{
    static const int OPP[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
    const int64_t nr = L->nr, nc = L->nc;
    const node_t *lst = L->list[LIST_WALL];
    for (int64_t q = 0; q < L->count[LIST_WALL]; ++q) {
        int64_t i = lst[q].i, j = lst[q].j;
        double p[9];
        for (int d = 0; d < 9; ++d) p[d] = P3(L, L->P, i, j, d);
        for (int d = 1; d < 9; ++d) {
            int64_t ti = i + OFF_I[d], tj = j + OFF_J[d];
            if (ti < 0 || ti >= nr || tj < 0 || tj >= nc) continue;
            P3(L, L->P, ti, tj, d) = p[OPP[d]];
        }
    }
}

/* BounceBackPolicy.hpp:143-187: masked reflection from obstacle nodes */
static void bounce_back_obstacles(orc_lattice *L)
{
    static const int OPP[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
    const node_t *lst = L->list[LIST_OBSTACLE];
    for (int64_t q = 0; q < L->count[LIST_OBSTACLE]; ++q) {
        int64_t i = lst[q].i, j = lst[q].j;
        int32_t mask = lst[q].tag;
        for (int d = 1; d < 9; ++d)
            if (mask & (1 << d)) P3(L, L->P, i + OFF_I[d], j + OFF_J[d], d) = P3(L, L->P, i, j, OPP[d]);
    }
}

/* PSBounceBackPolicy<2>::update_boundaries (PSBounceBackPolicy.hpp:177-297),
 * in list order: obstacle o reads its populations p1..p8 (pushed into it by
 * its wet neighbours), then for every direction bit d writes slot d of the
 * neighbour n = o + c_d from p_opp(d), the equilibrium of n at this step
 * (global_rho / global_u of n, DefaultEquilibrium's forms) and the value
 * slot d of n held before, with weight b of o and t = 1 / tau of the
 * policy.  global_rho of an obstacle keeps its initial 1. */
static void partially_saturated_obstacles(orc_lattice *L)
{
    const double one18 = (1.0 / 18.0);
    const double one36 = (1.0 / 36.0);
    const node_t *lst = L->list[LIST_OBSTACLE];
    for (int64_t q = 0; q < L->count[LIST_OBSTACLE]; ++q) {
        const int64_t i = lst[q].i, j = lst[q].j;
        const int32_t mask = lst[q].tag;
        const double b = L->ps_b[q];
        double p[9];
        for (int d = 1; d < 9; ++d) p[d] = P3(L, L->P, i, j, d);
        const double t = 1.0 / L->ps_tau;
        const double rho_o = R2(L, L->rho, i, j);
        for (int d = 1; d < 9; ++d) {
            if (!(mask & (1 << d))) continue;
            const int64_t ni = i + OFF_I[d], nj = j + OFF_J[d];
            double e[9];
            orc_feq(R2(L, L->rho, ni, nj), L->u[ni + L->nr * nj], L->u[ni + L->nr * (nj + L->nc)], e);
            static const int OPP[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
            const int o = OPP[d];
            const double po = p[o];
            const double old = P3(L, L->P, ni, nj, d);
            if (d < 5)
                P3(L, L->P, ni, nj, d) = po - (1.0 - b) * (po - e[o]) * t + b * ((old - e[d]) - (po - 2 * one18 * rho_o));
            else
                P3(L, L->P, ni, nj, d) = po - (1.0 - b) * (po - e[o]) * t + b * ((old - e[d]) - (po - one36 * rho_o));
        }
    }
}

void orc_set_partial_bounce_back(orc_lattice *L, double tau, const double *b)
{
    free(L->ps_b);
    L->ps_b = NULL;
    L->ps_tau = 0.0;
    if (!(tau > 0)) return;
    int64_t n = L->count[LIST_OBSTACLE];
    L->ps_b = (double *)malloc(sizeof(double) * (size_t)(n ? n : 1));
    for (int64_t q = 0; q < n; ++q) L->ps_b[q] = b ? b[q] : 1.0;
    L->ps_tau = tau;
}

/* ZouHePolicy.hpp:58-123 (velocity), pressure extension, or no-op
 * (OpenBoundaryPolicy.hpp:58-61) for one open list. */
static void update_open(orc_lattice *L, int k)
{
    const node_t *lst = L->list[LIST_INLET + k];
    int64_t m = L->count[LIST_INLET + k];
    int pol = L->policy[k];
    if (pol == OPEN_NOOP) return;
    for (int64_t q = 0; q < m; ++q) {
        int64_t i = lst[q].i, j = lst[q].j;
        double p[9];
        for (int d = 0; d < 9; ++d) p[d] = P3(L, L->P, i, j, d);
        if (pol == OPEN_ZOUHE_VELOCITY)
            orc_zouhe_velocity_node(lst[q].tag, p, L->u[i + L->nr * j], L->u[i + L->nr * (j + L->nc)]);
        else
            orc_zouhe_pressure_node(lst[q].tag, p, L->val_rho[k][q]);
        for (int d = 0; d < 9; ++d) P3(L, L->P, i, j, d) = p[d];
    }
}

/* ------------------------------------------------------------------ */
/* driver: Lattice::perform_lbm                                        */
/* ------------------------------------------------------------------ */

/* Lattice.hpp:294-328: rho = 1, u = 0, P0 = feq(1, 0) on the wet lists.
 * Wall/obstacle slots are left at zero (the reference leaves them
 * uninitialised). */
void orc_init(orc_lattice *L)
{
    size_t n = (size_t)(L->nr * L->nc);
    for (size_t k = 0; k < n; ++k) { L->rho[k] = 1.0; L->u[k] = 0.0; L->u[k + n] = 0.0; }
    memset(L->P, 0, 9 * n * sizeof(double));
    memset(L->E, 0, 9 * n * sizeof(double));
    memset(L->A, 0, 9 * n * sizeof(double));
    calc_equilibrium(L, L->P, L->list[LIST_FLUID], L->count[LIST_FLUID]);
    calc_equilibrium(L, L->P, L->list[LIST_INLET], L->count[LIST_INLET]);
    calc_equilibrium(L, L->P, L->list[LIST_OUTLET], L->count[LIST_OUTLET]);
    L->step = 0;
}

/* The reference hands global_rho / global_u to the boundary policies and
 * to update_nodes BY VALUE, at two call levels (OMPPolicy.hpp:50-72 ->
 * BounceBackPolicy.hpp:68,143, ZouHePolicy.hpp:58; OMPPolicy.hpp:139 ->
 * InletOutletInitializer.hpp:90): every such call constructs a fresh
 * Eigen tensor and copies the field into it on the calling thread.  With
 * by_value set (CPU timing baseline only; results do not change) these
 * copies are made the same way: allocate, copy serially, free. */
static void by_value_copy(const double *src, size_t n)
{
    double *t = (double *)malloc(n * sizeof(double));
    if (!t) return;
    for (size_t k = 0; k < n; ++k) t[k] = src[k];
    /* keep the copy observable so it is not optimised away */
    __asm__ volatile("" : : "r"(t) : "memory");
    free(t);
}

static void by_value_fields(const orc_lattice *L, int with_u)
{
    const size_t n = (size_t)(L->nr * L->nc);
    for (int level = 0; level < 2; ++level) {
        by_value_copy(L->rho, n);
        if (with_u) by_value_copy(L->u, 2 * n);
    }
}

void orc_set_by_value_copies(orc_lattice *L, int on) { L->by_value = on != 0; }

/* one iteration of the loop body at Lattice.hpp:338-423 (no file output) */
static void one_step(orc_lattice *L)
{
    if (L->by_value) by_value_fields(L, 0); /* update_nodes: density tensor by value */
    update_nodes(L);
    update_macro(L);
    calc_equilibrium(L, L->E, L->list[LIST_FLUID], L->count[LIST_FLUID]);
    calc_equilibrium(L, L->E, L->list[LIST_INLET], L->count[LIST_INLET]);
    calc_equilibrium(L, L->E, L->list[LIST_OUTLET], L->count[LIST_OUTLET]);
    collide(L, L->list[LIST_FLUID], L->count[LIST_FLUID]);
    collide(L, L->list[LIST_INLET], L->count[LIST_INLET]);
    collide(L, L->list[LIST_OUTLET], L->count[LIST_OUTLET]);
    stream(L, L->list[LIST_FLUID], L->count[LIST_FLUID]);
    stream(L, L->list[LIST_INLET], L->count[LIST_INLET]);
    stream(L, L->list[LIST_OUTLET], L->count[LIST_OUTLET]);
    if (L->by_value) by_value_fields(L, 1); /* update_domain_boundaries */
    bounce_back_walls(L);
    if (L->by_value) by_value_fields(L, 1); /* update_obstacle_boundaries */
    if (L->ps_tau > 0) partially_saturated_obstacles(L);
    else bounce_back_obstacles(L);
    if (L->by_value) by_value_fields(L, 1); /* update_inlet_boundaries */
    update_open(L, 0);
    if (L->by_value) by_value_fields(L, 1); /* update_outlet_boundaries */
    update_open(L, 1);
    L->step++;
}

void orc_step(orc_lattice *L, int64_t n)
{
    for (int64_t s = 0; s < n; ++s) one_step(L);
}

int64_t orc_get_step(const orc_lattice *L) { return L->step; }

void orc_get_populations(const orc_lattice *L, double *out) { memcpy(out, L->P, 9 * (size_t)(L->nr * L->nc) * sizeof(double)); }
void orc_set_populations(orc_lattice *L, const double *in) { memcpy(L->P, in, 9 * (size_t)(L->nr * L->nc) * sizeof(double)); }
void orc_get_after_collision(const orc_lattice *L, double *out) { memcpy(out, L->A, 9 * (size_t)(L->nr * L->nc) * sizeof(double)); }

void orc_get_macro(const orc_lattice *L, double *rho, double *u)
{
    size_t n = (size_t)(L->nr * L->nc);
    memcpy(rho, L->rho, n * sizeof(double));
    memcpy(u, L->u, 2 * n * sizeof(double));
}

/* FlowAnalyzer.hpp:348-451 on the after-collision tensor, over the given
 * points (coordinates + direction masks; NULL = the obstacle list). */
void orc_lift_drag(const orc_lattice *L, const int64_t *ij, const int32_t *masks, int64_t npts,
                   double *lift_drag, double *per_point)
{
    const node_t *lst = L->list[LIST_OBSTACLE];
    if (!ij) npts = L->count[LIST_OBSTACLE];
    double total_lift = 0., total_drag = 0.;
    const double *A = L->A;
    for (int64_t q = 0; q < npts; ++q) {
        int64_t i = ij ? ij[2 * q] : lst[q].i;
        int64_t j = ij ? ij[2 * q + 1] : lst[q].j;
        int32_t m = ij ? masks[q] : lst[q].tag;
        double local_lift = 0., local_drag = 0.;
        if (m & (1 << 1)) local_drag -= P3(L, A, i, j + 1, 3);
        if (m & (1 << 2)) local_lift -= P3(L, A, i - 1, j, 4);
        if (m & (1 << 3)) local_drag += P3(L, A, i, j - 1, 1);
        if (m & (1 << 4)) local_lift += P3(L, A, i + 1, j, 2);
        if (m & (1 << 5)) {
            local_drag -= (P3(L, A, i - 1, j + 1, 7)) * sqrt(2.0) / 2.0;
            local_lift -= (P3(L, A, i - 1, j + 1, 7)) * sqrt(2.0) / 2.0;
        }
        if (m & (1 << 6)) {
            local_drag += (P3(L, A, i - 1, j - 1, 8)) * sqrt(2.0) / 2.0;
            local_lift -= (P3(L, A, i - 1, j - 1, 8)) * sqrt(2.0) / 2.0;
        }
        if (m & (1 << 7)) {
            local_drag += (P3(L, A, i + 1, j - 1, 5)) * sqrt(2.0) / 2.0;
            local_lift += (P3(L, A, i + 1, j - 1, 5)) * sqrt(2.0) / 2.0;
        }
        if (m & (1 << 8)) {
            local_drag -= (P3(L, A, i + 1, j + 1, 6)) * sqrt(2.0) / 2.0;
            local_lift += (P3(L, A, i + 1, j + 1, 6)) * sqrt(2.0) / 2.0;
        }
        if (per_point) { per_point[2 * q] = local_lift; per_point[2 * q + 1] = local_drag; }
        total_drag += local_drag;
        total_lift += local_lift;
    }
    lift_drag[0] = total_lift;
    lift_drag[1] = total_drag;
}

/* Sums over the current populations P: out[0] = total mass of the wet
 * nodes, out[1..2] = momentum (x, y) of the fluid nodes, out[3] = mass of
 * the fluid nodes, out[4] = number of fluid nodes.  Sequential sums in list
 * order. */
void orc_reduce(const orc_lattice *L, double *out)
{
    double mass = 0., jx = 0., jy = 0., fmass = 0.;
    for (int l = LIST_FLUID; l <= LIST_OUTLET; ++l) {
        if (l == LIST_WALL) continue;
        for (int64_t q = 0; q < L->count[l]; ++q) {
            int64_t i = L->list[l][q].i, j = L->list[l][q].j;
            double p[9];
            for (int d = 0; d < 9; ++d) p[d] = P3(L, L->P, i, j, d);
            double rho = p[0] + p[1] + p[2] + p[3] + p[4] + p[5] + p[6] + p[7] + p[8];
            mass += rho;
            if (l == LIST_FLUID) {
                fmass += rho;
                jx += p[1] + p[5] + p[8] - (p[3] + p[6] + p[7]);
                jy += p[2] + p[5] + p[6] - (p[4] + p[7] + p[8]);
            }
        }
    }
    out[0] = mass; out[1] = jx; out[2] = jy; out[3] = fmass; out[4] = (double)L->count[LIST_FLUID];
}
