// Host-side geometry helpers of the C ABI: they write node-class planes
// (ColMajor, cls[i + NR*j]) with the node sets the reference's
// ConstructionInfo builds (llalbm/src/utils/generation/ConstructionInfo.hpp).
#include <cstdint>
#include <random>

#include "../../include/lbm_amd.h"

namespace {

// Disc bounds rule of add_obstacle_hyper_sphere (ConstructionInfo.hpp:388-402
// with check_obstacle_domain_bounds, :630-640): centre +- r must stay in
// rows/columns 1..N-2.
bool disc_fits(int64_t nr, int64_t nc, int64_t ci, int64_t cj, int64_t r)
{
    if (ci < 0 || ci >= nr || cj < 0 || cj >= nc || r < 0) return false;
    return ci - r > 0 && ci + r < nr - 1 && cj - r > 0 && cj + r < nc - 1;
}

int64_t paint_disc(int64_t nr, uint8_t *cls, int64_t ci, int64_t cj, int64_t r)
{
    int64_t added = 0;
    const int64_t r2 = r * r;
    for (int64_t i = ci - r; i <= ci + r; ++i)
        for (int64_t j = cj - r; j <= cj + r; ++j) {
            int64_t di = i - ci, dj = j - cj;
            if (di * di + dj * dj > r2) continue;  // ConstructionInfo.hpp:426-439
            uint8_t &k = cls[i + nr * j];
            if (k != LBM_OBSTACLE) {
                k = LBM_OBSTACLE;
                ++added;
            }
        }
    return added;
}

// Integer in [lo, hi] from one 64-bit draw: lo + floor(x * n / 2^64).
int64_t draw(std::mt19937_64 &g, int64_t lo, int64_t hi)
{
    unsigned __int128 x = g();
    uint64_t n = (uint64_t)(hi - lo + 1);
    return lo + (int64_t)((x * n) >> 64);
}

}  // namespace

extern "C" {

int lbm_geometry_cavity(int64_t nr, int64_t nc, uint8_t *cls)
{
    if (nr < 3 || nc < 3 || !cls) return LBM_ERR_INVALID;
    for (int64_t j = 0; j < nc; ++j)
        for (int64_t i = 0; i < nr; ++i) {
            bool perim = i == 0 || j == 0 || i == nr - 1 || j == nc - 1;
            cls[i + nr * j] = perim ? LBM_BOUNDARY : LBM_FLUID;
        }
    for (int64_t j = 1; j < nc - 1; ++j) cls[0 + nr * j] = LBM_INLET;
    return LBM_OK;
}

int lbm_geometry_channel(int64_t nr, int64_t nc, uint8_t *cls)
{
    if (nr < 3 || nc < 3 || !cls) return LBM_ERR_INVALID;
    for (int64_t j = 0; j < nc; ++j)
        for (int64_t i = 0; i < nr; ++i) {
            uint8_t k = LBM_FLUID;
            if (i == 0 || i == nr - 1) k = LBM_BOUNDARY;
            else if (j == 0) k = LBM_INLET;
            else if (j == nc - 1) k = LBM_OUTLET;
            cls[i + nr * j] = k;
        }
    return LBM_OK;
}

int lbm_geometry_add_disc(int64_t nr, int64_t nc, uint8_t *cls, int64_t ci, int64_t cj, int64_t r, int64_t *added)
{
    if (!cls) return LBM_ERR_INVALID;
    if (!disc_fits(nr, nc, ci, cj, r)) return LBM_ERR_INVALID;
    int64_t a = paint_disc(nr, cls, ci, cj, r);
    if (added) *added = a;
    return LBM_OK;
}

int lbm_geometry_porous(int64_t nr, int64_t nc, uint8_t *cls, uint64_t seed, int64_t rmin, int64_t rmax,
                        double target_solid, double *solid, int64_t *ndiscs)
{
    if (!cls || rmin < 0 || rmax < rmin || !(target_solid >= 0.0) || target_solid >= 0.9) return LBM_ERR_INVALID;
    if (2 * rmax + 3 > nr || 2 * rmax + 3 > nc) return LBM_ERR_INVALID;
    std::mt19937_64 gen(seed);
    const double total = (double)nr * (double)nc;
    int64_t filled = 0, placed = 0;
    for (int64_t q = 0; q < nr * nc; ++q) filled += cls[q] == LBM_OBSTACLE;
    // Each accepted disc covers at least one node, so the loop ends; the
    // attempt cap only guards against a target the free area cannot reach.
    const int64_t max_attempts = 64 * (int64_t)(total / 1.0) + 1000000;
    for (int64_t attempt = 0; filled < target_solid * total && attempt < max_attempts; ++attempt) {
        int64_t r = draw(gen, rmin, rmax);
        int64_t ci = draw(gen, 0, nr - 1);
        int64_t cj = draw(gen, 0, nc - 1);
        if (!disc_fits(nr, nc, ci, cj, r)) continue;
        filled += paint_disc(nr, cls, ci, cj, r);
        ++placed;
    }
    if (solid) *solid = (double)filled / total;
    if (ndiscs) *ndiscs = placed;
    return filled >= target_solid * total ? LBM_OK : LBM_ERR_INVALID;
}

}  // extern "C"
