// Host bookkeeping of a kinetic Monte Carlo step: rates, event choice,
// clock, basin comparison, event recycling and the excluded-atom history.
// Plain C++ on host arrays; the device part of the step (basin-2 kick,
// relax and basin check of every converged search) runs in the walker
// state machines.
#include <math.h>

#include "akmc.h"

namespace {

// fractional <-> Cartesian for a cell whose rows are the lattice vectors:
// r = s L, s = r L^-1 (L^-1 from the cofactors)
struct CellFrame {
    double L[9], R[9];
    explicit CellFrame(const double lat[9])
    {
        for (int k = 0; k < 9; ++k) L[k] = lat[k];
        const double c0 = L[4] * L[8] - L[5] * L[7], c1 = L[2] * L[7] - L[1] * L[8], c2 = L[1] * L[5] - L[2] * L[4];
        const double c3 = L[5] * L[6] - L[3] * L[8], c4 = L[0] * L[8] - L[2] * L[6], c5 = L[2] * L[3] - L[0] * L[5];
        const double c6 = L[3] * L[7] - L[4] * L[6], c7 = L[1] * L[6] - L[0] * L[7], c8 = L[0] * L[4] - L[1] * L[3];
        const double det = L[0] * c0 + L[1] * c3 + L[2] * c6;
        const double f = det != 0.0 ? 1.0 / det : 1.0;
        const double c[9] = {c0, c1, c2, c3, c4, c5, c6, c7, c8};
        for (int k = 0; k < 9; ++k) R[k] = c[k] * f;
    }
    // squared minimum-image length of the displacement (rx, ry, rz)
    double min_image_r2(double rx, double ry, double rz) const
    {
        double s[3];
        for (int j = 0; j < 3; ++j) {
            s[j] = rx * R[j] + ry * R[3 + j] + rz * R[6 + j];
            s[j] -= rint(s[j]);
        }
        double r2 = 0.0;
        for (int j = 0; j < 3; ++j) {
            const double d = s[0] * L[j] + s[1] * L[3 + j] + s[2] * L[6 + j];
            r2 += d * d;
        }
        return r2;
    }
};

}  // namespace

extern "C" {

double akmc_kmc_rate(double activation_energy, double prefactor, double temperature)
{
    const double beta = 1.0 / (AKMC_BOLTZMANN * temperature);
    return prefactor * exp(-activation_energy * beta);
}

double akmc_kmc_max_cap(double min_cap, double window, double temperature)
{
    return min_cap + window * temperature * AKMC_BOLTZMANN;
}

int akmc_kmc_select(int n, const double *rates, double u, double *total_rate)
{
    double total = 0.0;
    for (int i = 0; i < n; ++i) total += rates[i];
    if (total_rate) *total_rate = total;
    const double target = u * total;
    double cum = 0.0;
    for (int i = 0; i < n; ++i) {
        cum += rates[i];
        if (cum > target) return i;
    }
    return -1;
}

double akmc_kmc_time_increase(double total_rate, double u)
{
    if (fabs(total_rate) <= 1.0e-18) return 0.0;
    return -(1.0 / total_rate) * log(u);
}

int akmc_kmc_check_basin(int n, const double *basin1, const double *basin2, const double lat[9], double radius,
                         double max_radius)
{
    if (n < 0 || (n > 0 && (!basin1 || !basin2)) || !lat) return AKMC_EINVAL;
    const CellFrame cf(lat);
    const double r2min = radius * radius, r2max = max_radius * max_radius;
    int moved = 0;
    for (int i = 0; i < n; ++i) {
        const double r2 = cf.min_image_r2(basin1[i] - basin2[i], basin1[n + i] - basin2[n + i],
                                          basin1[2 * n + i] - basin2[2 * n + i]);
        if (r2 >= r2min) moved = 1;
        if (r2 > r2max) return 0;
    }
    return moved;
}

int akmc_kmc_displaced_atoms(int n, const double *pos, const double *basin, const double lat[9], double radius,
                             int *atoms)
{
    if (n < 0 || (n > 0 && (!pos || !basin || !atoms)) || !lat) return AKMC_EINVAL;
    const CellFrame cf(lat);
    const double r2min = radius * radius;
    int m = 0;
    for (int i = 0; i < n; ++i) {
        const double r2 =
            cf.min_image_r2(pos[i] - basin[i], pos[n + i] - basin[n + i], pos[2 * n + i] - basin[2 * n + i]);
        if (r2 > r2min) atoms[m++] = i;
    }
    return m;
}

int akmc_kmc_recycle(int n, const double *centres, int chosen, const double lat[9], double radius, int recycle,
                     int mode, double percent, const double *u, const int *blocked, int *keep)
{
    if (n < 0 || (n > 0 && (!centres || !keep)) || !lat || (recycle && (chosen < 0 || chosen >= n)) ||
        (recycle && mode == 2 && n > 0 && !u))
        return AKMC_EINVAL;
    for (int i = 0; i < n; ++i) keep[i] = 0;
    if (!recycle) return 0;
    const CellFrame cf(lat);
    const double ratio = mode == 2 ? percent / 100.0 : 1.2;
    const double *ac = centres + 3 * chosen;
    int kept = 0;
    for (int i = 0; i < n; ++i) {
        if (i == chosen) continue;
        const double *bc = centres + 3 * i;
        const double r = sqrt(cf.min_image_r2(ac[0] - bc[0], ac[1] - bc[1], ac[2] - bc[2]));
        if (r < radius) continue;
        const double choice = mode == 2 ? u[i] : 0.0;
        if (choice < ratio && !(blocked && blocked[i])) {
            keep[i] = 1;
            ++kept;
        }
    }
    return kept;
}

int akmc_kmc_update_excluded(int natoms, int *excluded, int nd, const int *displaced, int history)
{
    if (natoms < 0 || nd < 0 || (natoms > 0 && !excluded) || (nd > 0 && !displaced)) return AKMC_EINVAL;
    for (int i = 0; i < natoms; ++i)
        if (excluded[i] > 0) excluded[i] -= 1;
    for (int k = 0; k < nd; ++k) {
        if (displaced[k] < 0 || displaced[k] >= natoms) return AKMC_EINVAL;
        excluded[displaced[k]] += history;
    }
    return AKMC_SUCCESS;
}

}  // extern "C"
