/*
 * akmc_oracle.c -- CPU FP64 oracle (TEST INFRASTRUCTURE ONLY).
 *
 * Plain-C restatement of the reference arithmetic for the force-and-saddle
 * path, used only by tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg.  Nothing in the product library calls it.
 * Every function cites the reference file:line it follows; the checks that
 * pin each one are listed in tests/test_oracle_*.py.
 */
#include "akmc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define AKO_PI 3.14159265358979323846

/* ================================================================== */
/* cell helpers                                                        */
/* ================================================================== */

/* Cell lengths, angle cosines, perpendicular widths and volume of a cell
 * given as three row vectors (Common/Utility.cpp:23-84, the DL_POLY dcell
 * quantities: out[0..2] lengths, out[3..5] cosines, out[6..8] widths,
 * out[9] volume). */
void ako_dcell(const double a[9], double out[10])
{
    double len[3], cr[3][3], vol;
    for (int r = 0; r < 3; ++r)
        len[r] = sqrt(a[3 * r] * a[3 * r] + a[3 * r + 1] * a[3 * r + 1] + a[3 * r + 2] * a[3 * r + 2]);
    out[0] = len[0];
    out[1] = len[1];
    out[2] = len[2];
    out[3] = (a[0] * a[3] + a[1] * a[4] + a[2] * a[5]) / (len[0] * len[1]);
    out[4] = (a[0] * a[6] + a[1] * a[7] + a[2] * a[8]) / (len[0] * len[2]);
    out[5] = (a[3] * a[6] + a[4] * a[7] + a[5] * a[8]) / (len[1] * len[2]);
    /* cr[0] = b x c, cr[1] = c x a, cr[2] = a x b */
    for (int r = 0; r < 3; ++r) {
        const double *u = a + 3 * ((r + 1) % 3), *v = a + 3 * ((r + 2) % 3);
        cr[r][0] = u[1] * v[2] - u[2] * v[1];
        cr[r][1] = u[2] * v[0] - u[0] * v[2];
        cr[r][2] = u[0] * v[1] - u[1] * v[0];
    }
    vol = fabs(a[0] * cr[0][0] + a[1] * cr[0][1] + a[2] * cr[0][2]);
    out[9] = vol;
    for (int r = 0; r < 3; ++r)
        out[6 + r] = vol / sqrt(cr[r][0] * cr[r][0] + cr[r][1] * cr[r][1] + cr[r][2] * cr[r][2]);
}

/* Inverse of the row-vector cell matrix, laid out so that the fractional
 * coordinate s_a = sum_b r_b rcp[3b+a] (the convention the Field classes
 * use, potentials/metal/Field.cpp:396-398). */
void ako_inverse_cell(const double m[9], double rcp[9])
{
    double c[9];
    c[0] = m[4] * m[8] - m[5] * m[7];
    c[1] = m[2] * m[7] - m[1] * m[8];
    c[2] = m[1] * m[5] - m[2] * m[4];
    c[3] = m[5] * m[6] - m[3] * m[8];
    c[4] = m[0] * m[8] - m[2] * m[6];
    c[5] = m[2] * m[3] - m[0] * m[5];
    c[6] = m[3] * m[7] - m[4] * m[6];
    c[7] = m[1] * m[6] - m[0] * m[7];
    c[8] = m[0] * m[4] - m[1] * m[3];
    double det = m[0] * c[0] + m[1] * c[3] + m[2] * c[6];
    for (int i = 0; i < 9; ++i)
        rcp[i] = c[i] / det;
}

static inline void frac_of(double rx, double ry, double rz, const double *rcp,
                           double *xx, double *yy, double *zz)
{
    *xx = rx * rcp[0] + ry * rcp[3] + rz * rcp[6];
    *yy = rx * rcp[1] + ry * rcp[4] + rz * rcp[7];
    *zz = rx * rcp[2] + ry * rcp[5] + rz * rcp[8];
}

static inline void cart_of(double xx, double yy, double zz, const double *lat,
                           double *rx, double *ry, double *rz)
{
    *rx = xx * lat[0] + yy * lat[3] + zz * lat[6];
    *ry = xx * lat[1] + yy * lat[4] + zz * lat[7];
    *rz = xx * lat[2] + yy * lat[5] + zz * lat[8];
}

/* minimum-image displacement (potentials/metal/Field.cpp:571-583) */
static inline double min_image_rsq(double rx, double ry, double rz, const double *lat,
                                   const double *rcp, double *ox, double *oy, double *oz)
{
    double xx, yy, zz;
    frac_of(rx, ry, rz, rcp, &xx, &yy, &zz);
    xx -= rint(xx);
    yy -= rint(yy);
    zz -= rint(zz);
    cart_of(xx, yy, zz, lat, ox, oy, oz);
    return (*ox) * (*ox) + (*oy) * (*oy) + (*oz) * (*oz);
}

/* in-place wrap x <- L (s - rint(s)) (potentials/metal/Field.cpp:1894-1920) */
void ako_wrap_min_image(int n, double *x, double *y, double *z,
                        const double lat[9], const double rcp[9])
{
    for (int i = 0; i < n; ++i) {
        double xx, yy, zz;
        frac_of(x[i], y[i], z[i], rcp, &xx, &yy, &zz);
        xx -= rint(xx);
        yy -= rint(yy);
        zz -= rint(zz);
        cart_of(xx, yy, zz, lat, &x[i], &y[i], &z[i]);
    }
}

/* in-place wrap to s in [0,1) by one shift (potentials/metal/Field.cpp:1855-1892) */
void ako_wrap_no_image(int n, double *x, double *y, double *z,
                       const double lat[9], const double rcp[9])
{
    for (int i = 0; i < n; ++i) {
        double xx, yy, zz;
        frac_of(x[i], y[i], z[i], rcp, &xx, &yy, &zz);
        if (xx < 0.0) xx += 1.0;
        if (yy < 0.0) yy += 1.0;
        if (zz < 0.0) zz += 1.0;
        if (xx >= 1.0) xx -= 1.0;
        if (yy >= 1.0) yy -= 1.0;
        if (zz >= 1.0) zz -= 1.0;
        cart_of(xx, yy, zz, lat, &x[i], &y[i], &z[i]);
    }
}

/* explicit image shells c = int(rc / width) + 1 (potentials/metal/Field.cpp:70-76) */
static void image_cells(const double lat[9], double rc, int cells[3])
{
    double cp[10];
    ako_dcell(lat, cp);
    cells[0] = (int)(rc / cp[6]) + 1;
    cells[1] = (int)(rc / cp[7]) + 1;
    cells[2] = (int)(rc / cp[8]) + 1;
}

/* ================================================================== */
/* Sutton-Chen                                                         */
/* ================================================================== */

/* phi(r) = (eps c)^2 (a/r)^m  (ManyBody.cpp:497-504) */
static inline double sc_dens(const ako_sc_field *f, int ta, int tb, double r)
{
    if (r < 1.0e-8) return 0.0;
    int p = f->table[ta * f->nspec + tb];
    if (p < 0) return 0.0;
    const double *m = f->par + 9 * p;
    double arij = m[1] / r;
    double t = m[0] * m[4];
    return t * t * pow(arij, m[3]);
}

/* V = eps (a/r)^n, fs = -r V' = n eps (a/r)^n  (ManyBody.cpp:654-659) */
static inline void sc_pair(const ako_sc_field *f, int ta, int tb, double r, double *v, double *fs)
{
    *v = *fs = 0.0;
    if (r < 1.0e-8) return;
    int p = f->table[ta * f->nspec + tb];
    if (p < 0) return;
    const double *m = f->par + 9 * p;
    double arij = m[1] / r;
    *v = m[0] * pow(arij, m[2]);
    *fs = m[0] * m[2] * pow(arij, m[2]);
}

/* fm = (eps c/sqrt(rho_i) + eps c/sqrt(rho_j)) * 0.5 eps m c (a/r)^m
 * (ManyBody.cpp:711-757) */
static inline double sc_many(const ako_sc_field *f, int ta, int tb, double r, double rhoi, double rhoj)
{
    const double tiny = 1.0e-8;
    if (r < tiny || rhoi < tiny || rhoj < tiny) return 0.0;
    int p = f->table[ta * f->nspec + tb];
    if (p < 0) return 0.0;
    const double *m = f->par + 9 * p;
    double arij = m[1] / r;
    double fact = m[0] * m[4];
    double t1 = fact / sqrt(rhoi) + fact / sqrt(rhoj);
    double t2 = 0.5 * m[0] * m[3] * m[4];
    return t1 * t2 * pow(arij, m[3]);
}

/* F(rho) = -sqrt(rho) (ManyBody.cpp:554-575) */
static inline double sc_embed(double rho) { return -sqrt(rho); }

/* rho_i over the minimum image (Field.cpp:1782-1853) or explicit images
 * with rsq > 1e-3 (Field.cpp:1707-1781) */
static double sc_atom_rho(const ako_sc_field *f, int a, int n, const double *x, const double *y,
                          const double *z, const int *type, const double *lat, const double *rcp,
                          const int *cells)
{
    double rho = 0.0, radius = f->rc * f->rc;
    int ta = type[a];
    for (int b = 0; b < n; ++b) {
        if (f->min_image) {
            if (a == b) continue;
            double rx, ry, rz;
            double rsq = min_image_rsq(x[a] - x[b], y[a] - y[b], z[a] - z[b], lat, rcp, &rx, &ry, &rz);
            if (rsq <= radius) rho += sc_dens(f, ta, type[b], sqrt(rsq));
        } else {
            for (int nx = -cells[0]; nx <= cells[0]; ++nx)
                for (int ny = -cells[1]; ny <= cells[1]; ++ny)
                    for (int nz = -cells[2]; nz <= cells[2]; ++nz) {
                        double bx = x[b] + nx * lat[0] + ny * lat[3] + nz * lat[6];
                        double by = y[b] + nx * lat[1] + ny * lat[4] + nz * lat[7];
                        double bz = z[b] + nx * lat[2] + ny * lat[5] + nz * lat[8];
                        double rx = x[a] - bx, ry = y[a] - by, rz = z[a] - bz;
                        double rsq = rx * rx + ry * ry + rz * rz;
                        if (rsq <= radius && rsq > 1.0e-3) rho += sc_dens(f, ta, type[b], sqrt(rsq));
                    }
        }
    }
    return rho;
}

/* one ordered pair of pass 2 (Field.cpp:601-639): adds to f_i, subtracts
 * from f_j, accumulates stress and 1/2 V */
static inline void sc_pass2_pair(const ako_sc_field *f, int ta, int tb, double rx, double ry,
                                 double rz, double rsq, double rhoi, double rhoj, double *fi,
                                 double *fj, double *stress, double *epair)
{
    double r = sqrt(rsq), es, fs, fm;
    sc_pair(f, ta, tb, r, &es, &fs);
    fm = sc_many(f, ta, tb, r, rhoi, rhoj);
    double ft = (fs - fm) / (rsq * 2.0);
    *epair += 0.5 * es;
    fi[0] += rx * ft;
    fi[1] += ry * ft;
    fi[2] += rz * ft;
    fj[0] -= rx * ft;
    fj[1] -= ry * ft;
    fj[2] -= rz * ft;
    stress[0] -= ft * rx * rx;
    stress[1] -= ft * ry * ry;
    stress[2] -= ft * rz * rz;
    stress[3] -= ft * ry * rz;
    stress[4] -= ft * rx * rz;
    stress[5] -= ft * rx * ry;
}

/* Field::calculateForces (metal, Field.cpp:211-276) with the serial
 * realSpaceForceImage (Field.cpp:494-653) or realSpaceForceNoImage
 * (Field.cpp:817-956). */
double ako_sc_forces(const ako_sc_field *f, int n, double *x, double *y, double *z,
                     const int *type, const double lat[9], const double rcp[9],
                     double *fx, double *fy, double *fz, double stress[6],
                     double *e_many, double *e_pair, double *rho_out)
{
    int cells[3] = {0, 0, 0};
    double radius = f->rc * f->rc, emany = 0.0, epair = 0.0;
    double *rho = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    for (int i = 0; i < n; ++i) fx[i] = fy[i] = fz[i] = 0.0;
    for (int k = 0; k < 6; ++k) stress[k] = 0.0;
    if (f->min_image) {
        ako_wrap_min_image(n, x, y, z, lat, rcp);
    } else {
        ako_wrap_no_image(n, x, y, z, lat, rcp);
        image_cells(lat, f->rc, cells);
    }
    for (int a = 0; a < n; ++a) {
        rho[a] = sc_atom_rho(f, a, n, x, y, z, type, lat, rcp, cells);
        emany += sc_embed(rho[a]);
    }
    for (int a = 0; a < n; ++a) {
        int ta = type[a];
        double fi[3] = {0.0, 0.0, 0.0};
        for (int b = 0; b < n; ++b) {
            if (f->min_image) {
                if (a == b) continue;
                double rx, ry, rz, fj[3] = {0.0, 0.0, 0.0};
                double rsq = min_image_rsq(x[a] - x[b], y[a] - y[b], z[a] - z[b], lat, rcp, &rx, &ry, &rz);
                if (rsq <= radius) {
                    sc_pass2_pair(f, ta, type[b], rx, ry, rz, rsq, rho[a], rho[b], fi, fj, stress, &epair);
                    fx[b] += fj[0];
                    fy[b] += fj[1];
                    fz[b] += fj[2];
                }
            } else {
                for (int nx = -cells[0]; nx <= cells[0]; ++nx)
                    for (int ny = -cells[1]; ny <= cells[1]; ++ny)
                        for (int nz = -cells[2]; nz <= cells[2]; ++nz) {
                            double bx = x[b] + nx * lat[0] + ny * lat[3] + nz * lat[6];
                            double by = y[b] + nx * lat[1] + ny * lat[4] + nz * lat[7];
                            double bz = z[b] + nx * lat[2] + ny * lat[5] + nz * lat[8];
                            double rx = x[a] - bx, ry = y[a] - by, rz = z[a] - bz;
                            double rsq = rx * rx + ry * ry + rz * rz;
                            if (rsq <= radius && rsq > 1.0e-3) {
                                /* the no-image loop updates f_a directly
                                 * (Field.cpp:917-922) */
                                double fa[3] = {0.0, 0.0, 0.0}, fb[3] = {0.0, 0.0, 0.0};
                                sc_pass2_pair(f, ta, type[b], rx, ry, rz, rsq, rho[a], rho[b], fa, fb,
                                              stress, &epair);
                                fx[a] += fa[0];
                                fy[a] += fa[1];
                                fz[a] += fa[2];
                                fx[b] += fb[0];
                                fy[b] += fb[1];
                                fz[b] += fb[2];
                            }
                        }
            }
        }
        fx[a] += fi[0];
        fy[a] += fi[1];
        fz[a] += fi[2];
    }
    if (rho_out)
        for (int i = 0; i < n; ++i) rho_out[i] = rho[i];
    free(rho);
    if (e_many) *e_many = emany;
    if (e_pair) *e_pair = epair;
    return emany + epair;
}

/* Same per-pair arithmetic as ako_sc_forces (minimum image), threaded over
 * the outer atom loop with schedule(static,4) and thread-private force
 * copies merged at the end, the structure of Field.cpp:293-490 (the
 * reference's timed CPU build, CMakeLists.txt:46). */
double ako_sc_forces_omp(const ako_sc_field *f, int n, double *x, double *y, double *z,
                         const int *type, const double lat[9], const double rcp[9],
                         double *fx, double *fy, double *fz, double stress[6],
                         double *e_many, double *e_pair)
{
    if (!f->min_image)
        return ako_sc_forces(f, n, x, y, z, type, lat, rcp, fx, fy, fz, stress, e_many, e_pair, NULL);
    int cells[3] = {0, 0, 0};
    double radius = f->rc * f->rc, emany = 0.0, epair = 0.0;
    double *rho = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    for (int i = 0; i < n; ++i) fx[i] = fy[i] = fz[i] = 0.0;
    for (int k = 0; k < 6; ++k) stress[k] = 0.0;
    ako_wrap_min_image(n, x, y, z, lat, rcp);
#pragma omp parallel for schedule(static, 4) reduction(+ : emany)
    for (int a = 0; a < n; ++a) {
        rho[a] = sc_atom_rho(f, a, n, x, y, z, type, lat, rcp, cells);
        emany += sc_embed(rho[a]);
    }
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0, s4 = 0, s5 = 0;
#pragma omp parallel reduction(+ : epair, s0, s1, s2, s3, s4, s5)
    {
        double *tf = (double *)calloc((size_t)3 * n, sizeof(double));
        double st[6] = {0, 0, 0, 0, 0, 0};
#pragma omp for schedule(static, 4)
        for (int a = 0; a < n; ++a) {
            int ta = type[a];
            double fi[3] = {0.0, 0.0, 0.0};
            for (int b = 0; b < n; ++b) {
                if (a == b) continue;
                double rx, ry, rz, fj[3] = {0.0, 0.0, 0.0};
                double rsq = min_image_rsq(x[a] - x[b], y[a] - y[b], z[a] - z[b], lat, rcp, &rx, &ry, &rz);
                if (rsq <= radius) {
                    sc_pass2_pair(f, ta, type[b], rx, ry, rz, rsq, rho[a], rho[b], fi, fj, st, &epair);
                    tf[3 * b] += fj[0];
                    tf[3 * b + 1] += fj[1];
                    tf[3 * b + 2] += fj[2];
                }
            }
            tf[3 * a] += fi[0];
            tf[3 * a + 1] += fi[1];
            tf[3 * a + 2] += fi[2];
        }
#pragma omp critical
        for (int i = 0; i < n; ++i) {
            fx[i] += tf[3 * i];
            fy[i] += tf[3 * i + 1];
            fz[i] += tf[3 * i + 2];
        }
        s0 += st[0]; s1 += st[1]; s2 += st[2]; s3 += st[3]; s4 += st[4]; s5 += st[5];
        free(tf);
    }
    stress[0] = s0; stress[1] = s1; stress[2] = s2; stress[3] = s3; stress[4] = s4; stress[5] = s5;
    free(rho);
    if (e_many) *e_many = emany;
    if (e_pair) *e_pair = epair;
    return emany + epair;
}

/* energy only: the sums of realSpaceEnergyImage (Field.cpp:1088-1297) */
double ako_sc_energy(const ako_sc_field *f, int n, double *x, double *y, double *z,
                     const int *type, const double lat[9], const double rcp[9],
                     double *e_many, double *e_pair)
{
    int cells[3] = {0, 0, 0};
    double radius = f->rc * f->rc, emany = 0.0, epair = 0.0;
    if (f->min_image) ako_wrap_min_image(n, x, y, z, lat, rcp);
    else {
        ako_wrap_no_image(n, x, y, z, lat, rcp);
        image_cells(lat, f->rc, cells);
    }
    for (int a = 0; a < n; ++a) {
        emany += sc_embed(sc_atom_rho(f, a, n, x, y, z, type, lat, rcp, cells));
        for (int b = 0; b < n; ++b) {
            if (f->min_image) {
                if (a == b) continue;
                double rx, ry, rz, v, fs;
                double rsq = min_image_rsq(x[a] - x[b], y[a] - y[b], z[a] - z[b], lat, rcp, &rx, &ry, &rz);
                if (rsq <= radius) {
                    sc_pair(f, type[a], type[b], sqrt(rsq), &v, &fs);
                    epair += 0.5 * v;
                }
            } else {
                for (int nx = -cells[0]; nx <= cells[0]; ++nx)
                    for (int ny = -cells[1]; ny <= cells[1]; ++ny)
                        for (int nz = -cells[2]; nz <= cells[2]; ++nz) {
                            double rx = x[a] - (x[b] + nx * lat[0] + ny * lat[3] + nz * lat[6]);
                            double ry = y[a] - (y[b] + nx * lat[1] + ny * lat[4] + nz * lat[7]);
                            double rz = z[a] - (z[b] + nx * lat[2] + ny * lat[5] + nz * lat[8]);
                            double rsq = rx * rx + ry * ry + rz * rz, v, fs;
                            if (rsq <= radius && rsq > 1.0e-3) {
                                sc_pair(f, type[a], type[b], sqrt(rsq), &v, &fs);
                                epair += 0.5 * v;
                            }
                        }
            }
        }
    }
    if (e_many) *e_many = emany;
    if (e_pair) *e_pair = epair;
    return emany + epair;
}

long long ako_sc_count_pairs(const ako_sc_field *f, int n, const double *x, const double *y,
                             const double *z, const double lat[9], const double rcp[9])
{
    long long cnt = 0;
    double radius = f->rc * f->rc;
    int cells[3];
    image_cells(lat, f->rc, cells);
    for (int a = 0; a < n; ++a)
        for (int b = 0; b < n; ++b) {
            if (f->min_image) {
                if (a == b) continue;
                double rx, ry, rz;
                if (min_image_rsq(x[a] - x[b], y[a] - y[b], z[a] - z[b], lat, rcp, &rx, &ry, &rz) <= radius) ++cnt;
            } else {
                for (int nx = -cells[0]; nx <= cells[0]; ++nx)
                    for (int ny = -cells[1]; ny <= cells[1]; ++ny)
                        for (int nz = -cells[2]; nz <= cells[2]; ++nz) {
                            double rx = x[a] - (x[b] + nx * lat[0] + ny * lat[3] + nz * lat[6]);
                            double ry = y[a] - (y[b] + nx * lat[1] + ny * lat[4] + nz * lat[7]);
                            double rz = z[a] - (z[b] + nx * lat[2] + ny * lat[5] + nz * lat[8]);
                            double rsq = rx * rx + ry * ry + rz * rz;
                            if (rsq <= radius && rsq > 1.0e-3) ++cnt;
                        }
            }
        }
    return cnt;
}

/* ================================================================== */
/* Ewald + Buckingham                                                  */
/* ================================================================== */

/* eta from the Madelung accuracy (pair Field.cpp:129-132) */
void ako_ewald_eta_from_acc(double acc, double rc, double *eta, double *eps, double *tol)
{
    double e = fmin(fabs(acc), 0.5);
    double t = sqrt(fabs(log(e * rc)));
    *eps = e;
    *tol = t;
    *eta = sqrt(fabs(log(e * rc * t))) / rc;
}

/* reciprocal shell counts and cutoff (Ewald.cpp:88-120) */
void ako_ewald_gcell(const double lat[9], const double rcp[9], double eta, double rc,
                     double eps, double tol, int gcell[3], double *rcpcut)
{
    double cp[10];
    ako_dcell(lat, cp);
    double tol1 = sqrt(-log(eps * rc * pow(2.0 * tol * eta, 2)));
    for (int k = 0; k < 3; ++k) gcell[k] = (int)rint(0.25 + cp[6 + k] * eta * tol1 / AKO_PI);
    ako_dcell(rcp, cp);
    double c = gcell[0] * cp[6];
    if (gcell[1] * cp[7] < c) c = gcell[1] * cp[7];
    if (gcell[2] * cp[8] < c) c = gcell[2] * cp[8];
    *rcpcut = c;
}

/* half-space k set and prefactors (Ewald.cpp:188-319): n_x >= 0, whole
 * n_x = 0 plane, 1e-8 < |n rcp|^2 < rcpcut^2, k = 2 pi n rcp;
 * g0 = 4 pi CTOEV exp(-k^2/4eta^2)/(k^2 V), doubled when k_x >= 1e-8. */
int ako_ewald_kvectors(const double rcp[9], double volume, double eta, double cfact,
                       const int gcell[3], double rcpcut, int maxk,
                       double *gx, double *gy, double *gz, double *g0, double *g1)
{
    double etasq = eta * eta, gsqfct = 1.0 / (4.0 * etasq);
    double gfct0 = cfact * AKO_PI / (etasq * volume), gfct1 = gsqfct * gfct0;
    double cut2 = rcpcut * rcpcut;
    int nk = 0;
    for (int nx = 0; nx <= gcell[0]; ++nx)
        for (int ny = -gcell[1]; ny <= gcell[1]; ++ny)
            for (int nz = -gcell[2]; nz <= gcell[2]; ++nz) {
                double x = nx * rcp[0] + ny * rcp[1] + nz * rcp[2];
                double y = nx * rcp[3] + ny * rcp[4] + nz * rcp[5];
                double z = nx * rcp[6] + ny * rcp[7] + nz * rcp[8];
                double gsq = x * x + y * y + z * z;
                if (gsq < cut2 && gsq > 1.0e-8) {
                    if (nk >= maxk) return -1;
                    gx[nk] = x;
                    gy[nk] = y;
                    gz[nk] = z;
                    ++nk;
                }
            }
    for (int i = 0; i < nk; ++i) {
        gx[i] *= 2.0 * AKO_PI;
        gy[i] *= 2.0 * AKO_PI;
        gz[i] *= 2.0 * AKO_PI;
        double gsq = gx[i] * gx[i] + gy[i] * gy[i] + gz[i] * gz[i];
        double xg = gsq * gsqfct;
        double gzz = exp(-xg) / xg;
        g0[i] = (gx[i] < 1.0e-8) ? gzz * gfct0 : gzz * gfct0 * 2.0;
        g1[i] = -2.0 * gzz * (1.0 + 1.0 / xg) * gfct1;
    }
    return nk;
}

/* serial reciprocal-space force (Ewald.cpp:560-662).  The reference's
 * phaseTotal starts uninitialised; here it starts at 0 (stress only). */
double ako_ewald_recip_force(int n, const double *x, const double *y, const double *z,
                             const double *q, int nk, const double *gx, const double *gy,
                             const double *gz, const double *g0, const double *g1,
                             double *fx, double *fy, double *fz, double stress[6])
{
    double *cs = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    double *sn = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    double eng = 0.0, phtot = 0.0;
    for (int k = 0; k < nk; ++k) {
        double C = 0.0, S = 0.0, vx = gx[k], vy = gy[k], vz = gz[k], v0 = g0[k], v1 = g1[k];
        for (int a = 0; a < n; ++a) {
            double ph = vx * x[a] + vy * y[a] + vz * z[a];
            cs[a] = q[a] * cos(ph);
            sn[a] = q[a] * sin(ph);
            C += cs[a];
            S += sn[a];
        }
        double ph2 = 0.5 * (C * C + S * S);
        if (ph2 > 1.0e-16) {
            phtot += ph2 * v0;
            double fac = 2.0 * ph2 * v1;
            stress[0] += fac * vx * vx;
            stress[1] += fac * vy * vy;
            stress[2] += fac * vz * vz;
            stress[3] += fac * vy * vz;
            stress[4] += fac * vx * vz;
            stress[5] += fac * vx * vy;
        }
        for (int a = 0; a < n; ++a) {
            eng += v0 * (cs[a] * C + sn[a] * S);
            double pf = -v0 * (-sn[a] * C + cs[a] * S);
            fx[a] += pf * vx;
            fy[a] += pf * vy;
            fz[a] += pf * vz;
        }
    }
    stress[0] -= phtot;
    stress[1] -= phtot;
    stress[2] -= phtot;
    free(cs);
    free(sn);
    return 0.5 * eng;
}

/* reciprocal energy 1/2 sum g0 |S(k)|^2 (Ewald.cpp:368-451) */
double ako_ewald_recip_energy(int n, const double *x, const double *y, const double *z,
                              const double *q, int nk, const double *gx, const double *gy,
                              const double *gz, const double *g0)
{
    double e = 0.0;
    for (int k = 0; k < nk; ++k) {
        double C = 0.0, S = 0.0;
        for (int a = 0; a < n; ++a) {
            double ph = gx[k] * x[a] + gy[k] * y[a] + gz[k] * z[a];
            C += q[a] * cos(ph);
            S += q[a] * sin(ph);
        }
        e += g0[k] * (C * C + S * S);
    }
    return 0.5 * e;
}

/* erfc(x)/x and its radial derivative term (pair Field.cpp:2218-2236) */
static inline void hfunc1(double xv, double *h0, double *h1)
{
    if (xv == 0.0) return;
    const double fact = 1.128379167095512;
    *h0 = erfc(xv) / xv;
    double xsq = xv * xv;
    *h1 = (-*h0 - fact * exp(-xsq)) / xsq;
}

/* The short-range pair forms and their "force" (-(dV/dr)/r where the
 * reference's expression is that) as TwoBody::shortrangeone computes them
 * (TwoBody.h:173-276); with energy_path the energy is that of the energy-
 * only form (TwoBody.h:81-150), which differs for Morse by + C12/r^12.
 * The table default 0 for unlisted pairs (TwoBody.cpp:509) is the
 * caller's table. */
static inline void buck(const ako_ion_field *f, int ta, int tb, double r, double rsq, double *e,
                        double *frc)
{
    *e = *frc = 0.0;
    if (r < 1.0e-8) return;
    int p = f->table[ta * f->nspec + tb];
    if (p < 0) return;
    const double *m = f->par + 6 * p;
    switch (f->form[p]) {
    case AKO_MORSE: {
        double e1 = exp(-m[1] * (r - m[2]));
        *e = m[0] * (1.0 - e1) * (1.0 - e1) - m[0];
        if (f->energy_path) *e += m[3] / pow(rsq, 6);
        *frc = -(2.0 * m[0] * m[1] * (1.0 - e1) * e1) / r;
        break;
    }
    case AKO_LJONES: {
        double t = m[1] / r;
        double e1 = 4.0 * m[0] * pow(t, 12.0), e2 = -4.0 * m[0] * pow(t, 6.0);
        *e = e1 + e2;
        *frc = -(12.0 * e1 / rsq + 6.0 * e2 / rsq);
        break;
    }
    case AKO_BHM: {
        double r6 = rsq * rsq * rsq, r8 = r6 * rsq, ex = exp(m[1] * (m[2] - r));
        *e = m[0] * ex - m[3] / r6 - m[4] / r8;
        *frc = (r * m[0] * m[1] * ex - 6.0 * m[3] / r6 - 8.0 * m[4] / r8) / rsq;
        break;
    }
    case AKO_SEVNSIX: {
        double t = m[1] / r;
        double r7 = rsq * rsq * rsq * r, r8 = r7 * r;
        double f1 = -m[0] * 7.0 * pow(m[1], 7.0) / r8, f2 = -m[0] * 12.0 * pow(m[1], 6.0) / r7;
        *e = m[0] * pow(t, 7.0) - 2.0 * m[0] * pow(t, 6.0);
        *frc = -(f1 + f2) / rsq;
        break;
    }
    default: {
        double e1 = m[0] * exp(-r / m[1]);
        double f1 = e1 / (r * m[1]);
        double e2 = -m[2] / (rsq * rsq * rsq);
        double f2 = 6 * e2 / rsq;
        *e = e1 + e2;
        *frc = f1 + f2;
    }
    }
}

double ako_damped_mesh(double alpha, double rc, int nmesh, double *e, double *f, double *rfld0, double *rfld1)
{
    const double a1 = 0.254829592, a2 = -0.284496736, a3 = 1.421413741, a4 = -1.453152027, a5 = 1.061405429,
                 pp = 0.3275911;  /* Abramowitz & Stegun 7.1.26 */
    const double del = rc / (nmesh - 4);
    e[0] = f[0] = 0.0;
    for (int i = 1; i < nmesh; ++i) {
        double r = (double)i * del, rsq = r * r;
        double tt = 1.0 / (1.0 + pp * alpha * r);
        double ex = exp(-pow(alpha * r, 2.0));
        e[i] = tt * (a1 + tt * (a2 + tt * (a3 + tt * (a4 + tt * a5)))) * ex / r;
        f[i] = (e[i] + 2.0 * (alpha / 1.7724538509055159) * ex) / rsq;
    }
    *rfld0 = e[nmesh - 4];
    *rfld1 = rc * f[nmesh - 4];
    return del;
}

/* interpolate_shifted_damped_potential (pair Field.cpp:2049-2094): the
 * three-point interpolation of the mesh; energy and -(dV/dr)/r with the
 * shift terms */
static inline void damped(const ako_ion_field *f, double r, double chg, double *e0, double *f1)
{
    double rstep = 1.0 / f->dsf_del;
    int point = (int)(r * rstep);
    double dlt = r * rstep - (double)point;
    const double *t = f->dsf_e;
    double d1 = t[point] + (t[point + 1] - t[point]) * dlt;
    double d2 = t[point + 1] + (t[point + 2] - t[point + 1]) * (dlt - 1.0);
    double omega = d1 + (d2 - d1) * dlt * 0.5;
    *e0 = chg * (omega - f->dsf_rfld0 + f->dsf_rfld1 * (r - f->rc));
    t = f->dsf_f;
    d1 = t[point] + (t[point + 1] - t[point]) * dlt;
    d2 = t[point + 1] + (t[point + 2] - t[point + 1]) * (dlt - 1.0);
    double gamma = d1 + (d2 - d1) * dlt * 0.5;
    *f1 = -chg * (gamma - f->dsf_rfld1 / r);
}

/* one real-space pair term of realSpaceForceImage (pair Field.cpp:780-818) */
static inline void ion_pair(const ako_ion_field *f, int ta, int tb, double qa_c, double qb,
                            double rx, double ry, double rz, double rsq, double *fa, double *fb,
                            double *stress, double *ereal, double *evdw)
{
    double r = sqrt(rsq), h0 = 0.0, h1 = 0.0, eng = 0.0, forc = 0.0, es, fs;
    double hf0 = f->eta, hf1 = hf0 * hf0 * hf0;
    double cp = qa_c * qb;
    if (f->do_ewald == 1) {
        hfunc1(hf0 * r, &h0, &h1);
        eng = h0 * cp * hf0;
        forc = h1 * cp * hf1;
    } else if (f->do_ewald == 2) {
        damped(f, r, cp, &eng, &forc);
    }
    buck(f, ta, tb, r, rsq, &es, &fs);
    *ereal += eng;
    *evdw += es;
    forc -= fs;
    fa[0] -= rx * forc;
    fa[1] -= ry * forc;
    fa[2] -= rz * forc;
    fb[0] += rx * forc;
    fb[1] += ry * forc;
    fb[2] += rz * forc;
    stress[0] += forc * rx * rx;
    stress[1] += forc * ry * ry;
    stress[2] += forc * rz * rz;
    stress[3] += forc * ry * rz;
    stress[4] += forc * rx * rz;
    stress[5] += forc * rx * ry;
}

/* pair Field::calculateForces (Field.cpp:323-381): reciprocal force on the
 * unwrapped positions, wrap, then realSpaceForceImage (Field.cpp:660-842)
 * or realSpaceForceNoImage (Field.cpp:1016-1201). */
double ako_ion_forces(const ako_ion_field *f, int n, double *x, double *y, double *z,
                      const int *type, const double *q, const double lat[9], const double rcp[9],
                      double *fx, double *fy, double *fz, double stress[6],
                      double *e_rcp, double *e_real, double *e_vdw)
{
    double ercp = 0.0, ereal = 0.0, evdw = 0.0, radius = f->rc * f->rc;
    for (int i = 0; i < n; ++i) fx[i] = fy[i] = fz[i] = 0.0;
    for (int k = 0; k < 6; ++k) stress[k] = 0.0;
    if (f->do_ewald == 1)
        ercp = ako_ewald_recip_force(n, x, y, z, q, f->nk, f->gx, f->gy, f->gz, f->g0, f->g1, fx, fy, fz, stress);
    if (f->min_image) ako_wrap_min_image(n, x, y, z, lat, rcp);
    else ako_wrap_no_image(n, x, y, z, lat, rcp);
    /* self term via hfunc2(0): -eta/sqrt(pi) CTOEV q^2 (Field.cpp:704-720) */
    if (f->do_ewald == 1) {
        const double fact = 1.128379167095512;
        for (int a = 0; a < n; ++a) ereal += -fact * (0.5 * f->cfact * q[a] * q[a]) * f->eta;
    }
    for (int a = 0; a < n; ++a) {
        int ta = type[a];
        double qa = f->cfact * q[a];
        double fi[3] = {0.0, 0.0, 0.0};
        for (int b = a; b < n; ++b) {
            if (f->min_image) {
                if (a == b) continue;
                double rx, ry, rz;
                double rsq = min_image_rsq(x[a] - x[b], y[a] - y[b], z[a] - z[b], lat, rcp, &rx, &ry, &rz);
                if (rsq <= radius) {
                    double fb[3] = {0.0, 0.0, 0.0};
                    ion_pair(f, ta, type[b], qa, q[b], rx, ry, rz, rsq, fi, fb, stress, &ereal, &evdw);
                    fx[b] += fb[0];
                    fy[b] += fb[1];
                    fz[b] += fb[2];
                }
            } else {
                const int *c = f->cells;
                for (int nx = -c[0]; nx <= c[0]; ++nx)
                    for (int ny = -c[1]; ny <= c[1]; ++ny)
                        for (int nz = -c[2]; nz <= c[2]; ++nz) {
                            if (a == b && nx == 0 && ny == 0 && nz == 0) continue;
                            double rx = x[a] - (x[b] + nx * lat[0] + ny * lat[3] + nz * lat[6]);
                            double ry = y[a] - (y[b] + nx * lat[1] + ny * lat[4] + nz * lat[7]);
                            double rz = z[a] - (z[b] + nx * lat[2] + ny * lat[5] + nz * lat[8]);
                            double rsq = rx * rx + ry * ry + rz * rz;
                            if (rsq <= radius) {
                                double fa[3] = {0.0, 0.0, 0.0}, fb[3] = {0.0, 0.0, 0.0};
                                ion_pair(f, ta, type[b], qa, q[b], rx, ry, rz, rsq, fa, fb, stress, &ereal, &evdw);
                                fx[a] += fa[0];
                                fy[a] += fa[1];
                                fz[a] += fa[2];
                                fx[b] += fb[0];
                                fy[b] += fb[1];
                                fz[b] += fb[2];
                            }
                        }
            }
        }
        fx[a] += fi[0];
        fy[a] += fi[1];
        fz[a] += fi[2];
    }
    if (e_rcp) *e_rcp = ercp;
    if (e_real) *e_real = ereal;
    if (e_vdw) *e_vdw = evdw;
    return ercp + ereal + evdw;
}

/* unordered in-cutoff pairs counted as the real-space loop visits them */
long long ako_ion_count_pairs(const ako_ion_field *f, int n, const double *x, const double *y,
                              const double *z, const double lat[9], const double rcp[9])
{
    long long cnt = 0;
    double radius = f->rc * f->rc;
    for (int a = 0; a < n; ++a)
        for (int b = a; b < n; ++b) {
            if (f->min_image) {
                if (a == b) continue;
                double rx, ry, rz;
                if (min_image_rsq(x[a] - x[b], y[a] - y[b], z[a] - z[b], lat, rcp, &rx, &ry, &rz) <= radius) ++cnt;
            } else {
                const int *c = f->cells;
                for (int nx = -c[0]; nx <= c[0]; ++nx)
                    for (int ny = -c[1]; ny <= c[1]; ++ny)
                        for (int nz = -c[2]; nz <= c[2]; ++nz) {
                            if (a == b && nx == 0 && ny == 0 && nz == 0) continue;
                            double rx = x[a] - (x[b] + nx * lat[0] + ny * lat[3] + nz * lat[6]);
                            double ry = y[a] - (y[b] + nx * lat[1] + ny * lat[4] + nz * lat[7]);
                            double rz = z[a] - (z[b] + nx * lat[2] + ny * lat[5] + nz * lat[8]);
                            if (rx * rx + ry * ry + rz * rz <= radius) ++cnt;
                        }
            }
        }
    return cnt;
}

/* ================================================================== */
/* field callback                                                      */
/* ================================================================== */
double ako_field_cb(void *vctx, int n, double *x, double *y, double *z,
                    double *fx, double *fy, double *fz)
{
    ako_field_ctx *c = (ako_field_ctx *)vctx;
    double st[6];
    c->ncalls++;
    if (c->kind == 0)
        return ako_sc_forces((const ako_sc_field *)c->field, n, x, y, z, c->type, c->lat, c->rcp,
                             fx, fy, fz, st, NULL, NULL, NULL);
    if (c->kind == 2)
        return ako_sc_forces_omp((const ako_sc_field *)c->field, n, x, y, z, c->type, c->lat, c->rcp,
                                 fx, fy, fz, st, NULL, NULL);
    return ako_ion_forces((const ako_ion_field *)c->field, n, x, y, z, c->type, c->q, c->lat,
                          c->rcp, fx, fy, fz, st, NULL, NULL, NULL);
}

/* ================================================================== */
/* FIRE relax                                                          */
/* ================================================================== */

/* defaults of Common/RelaxCntrl.h:15-48; euler = true (optim/Relax.h:29) */
void ako_relax_defaults(ako_relax_cntrl *c)
{
    c->maxIter = 3000;
    c->normTol = 1.0e-5;
    c->forceTol = 1.0e-4;
    c->N_min = 5;
    c->initTimeStep = 0.0005;
    c->alphaDecrease = 0.99;
    c->stepIncrease = 1.1;
    c->stepDecrease = 0.5;
    c->alphaStart = 0.1;
    c->maxTimeStep = 0.00075;
    c->euler = 1;
    c->timeStepUnit = AKO_EVTODL;
    c->N_delaySteps = 20;
    c->N_maxUphill = 50;
    c->initialDelay = 1;
    c->minTimeStep = 0.0001;
}

/* Relax::vectorNormMax over atoms (optim/Relax.cpp:926-946) */
static void norm_max3(int n, const double *ax, const double *ay, const double *az, double *nrm, double *mx)
{
    double s = 0.0, m = 0.0;
    for (int i = 0; i < n; ++i) {
        double q = ax[i] * ax[i] + ay[i] * ay[i] + az[i] * az[i];
        if (q > m) m = q;
        s += q;
    }
    *nrm = sqrt(s);
    *mx = sqrt(m);
}

/* Relax::FIRERelax (optim/Relax.cpp:44-315) */
int ako_fire_relax(ako_force_fn fn, void *ctx, int n, double *x, double *y, double *z,
                   const int *frozen, const ako_relax_cntrl *c, double *energy, int *iters)
{
    double *vx = (double *)calloc(n, sizeof(double)), *vy = (double *)calloc(n, sizeof(double)),
           *vz = (double *)calloc(n, sizeof(double)), *fx = (double *)calloc(n, sizeof(double)),
           *fy = (double *)calloc(n, sizeof(double)), *fz = (double *)calloc(n, sizeof(double));
    double dt = c->initTimeStep, dtmax = c->maxTimeStep, alpha = c->alphaStart;
    double fnorm, fmax, vnorm, vmax, e;
    int nstep = 1, status = AKO_FAILED, it;
    e = fn(ctx, n, x, y, z, fx, fy, fz);
    for (int i = 0; i < n; ++i)
        if (frozen && frozen[i] == 1) fx[i] = fy[i] = fz[i] = 0.0;
    double half = 0.5 * dt * c->timeStepUnit;
    for (it = 1; it <= c->maxIter; ++it) {
        double P = 0.0;
        for (int i = 0; i < n; ++i) P += fx[i] * vx[i] + fy[i] * vy[i] + fz[i] * vz[i];
        double oma = 1.0 - alpha;
        norm_max3(n, fx, fy, fz, &fnorm, &fmax);
        norm_max3(n, vx, vy, vz, &vnorm, &vmax);
        double af = alpha * vnorm / fnorm;
        if (P > 0) {
            for (int i = 0; i < n; ++i)
                if (!frozen || frozen[i] == 0) {
                    vx[i] = oma * vx[i] + af * fx[i];
                    vy[i] = oma * vy[i] + af * fy[i];
                    vz[i] = oma * vz[i] + af * fz[i];
                }
            if (nstep > c->N_min) {
                dt = fmin(dt * c->stepIncrease, dtmax);
                alpha *= c->alphaDecrease;
            }
            nstep += 1;
        } else {
            nstep = 0;
            dt *= c->stepDecrease;
            alpha = c->alphaStart;
            for (int i = 0; i < n; ++i) vx[i] = vy[i] = vz[i] = 0.0;
        }
        if (c->euler) {
            double fact = dt * c->timeStepUnit;
            for (int i = 0; i < n; ++i)
                if (!frozen || frozen[i] == 0) {
                    x[i] += dt * vx[i];
                    y[i] += dt * vy[i];
                    z[i] += dt * vz[i];
                    vx[i] += fact * fx[i];
                    vy[i] += fact * fy[i];
                    vz[i] += fact * fz[i];
                }
        } else {
            half = 0.5 * dt * c->timeStepUnit;
            for (int i = 0; i < n; ++i) {
                if (!frozen || frozen[i] == 0) {
                    vx[i] += half * fx[i];
                    vy[i] += half * fy[i];
                    vz[i] += half * fz[i];
                } else {
                    vx[i] = vy[i] = vz[i] = 0.0;
                }
            }
            for (int i = 0; i < n; ++i)
                if (!frozen || frozen[i] == 0) {
                    x[i] += dt * vx[i];
                    y[i] += dt * vy[i];
                    z[i] += dt * vz[i];
                }
        }
        e = fn(ctx, n, x, y, z, fx, fy, fz);
        if (frozen)
            for (int i = 0; i < n; ++i)
                if (frozen[i] == 1) fx[i] = fy[i] = fz[i] = 0.0;
        if (!c->euler) {
            for (int i = 0; i < n; ++i) {
                if (!frozen || frozen[i] == 0) {
                    vx[i] += half * fx[i];
                    vy[i] += half * fy[i];
                    vz[i] += half * fz[i];
                } else {
                    vx[i] = vy[i] = vz[i] = 0.0;
                    fx[i] = fy[i] = fz[i] = 0.0;
                }
            }
        }
        /* Relax::centreOfMass over all atoms (optim/Relax.cpp:964-996) */
        double mx = 0.0, my = 0.0, mz = 0.0;
        for (int i = 0; i < n; ++i) {
            mx += vx[i];
            my += vy[i];
            mz += vz[i];
        }
        mx /= (double)n;
        my /= (double)n;
        mz /= (double)n;
        for (int i = 0; i < n; ++i) {
            vx[i] -= mx;
            vy[i] -= my;
            vz[i] -= mz;
        }
        norm_max3(n, fx, fy, fz, &fnorm, &fmax);
        if (fnorm < c->normTol || fmax < c->forceTol) {
            status = AKO_SUCCESS;
            break;
        }
    }
    if (iters) *iters = (it > c->maxIter) ? c->maxIter : it;
    if (energy) *energy = e;
    free(vx); free(vy); free(vz); free(fx); free(fy); free(fz);
    return status;
}

/* Relax::FIRE2Relax (optim/Relax.cpp:317-648).  Differences from FIRE:
 * the velocity mixing moves into the integration loop (:496-540), an
 * uphill step counts towards N_maxUphill (fail past it, :446-459), only
 * decreases dt and resets alpha from iteration N_delaySteps on when
 * initialDelay is set (:461-478), clamps dt at minTimeStep, and puts the
 * free atoms back by half a step of the new dt before zeroing the
 * velocities (:481-491). */
int ako_fire2_relax(ako_force_fn fn, void *ctx, int n, double *x, double *y, double *z,
                    const int *frozen, const ako_relax_cntrl *c, double *energy, int *iters)
{
    double *vx = (double *)calloc(n, sizeof(double)), *vy = (double *)calloc(n, sizeof(double)),
           *vz = (double *)calloc(n, sizeof(double)), *fx = (double *)calloc(n, sizeof(double)),
           *fy = (double *)calloc(n, sizeof(double)), *fz = (double *)calloc(n, sizeof(double));
    double dt = c->initTimeStep, dtmax = c->maxTimeStep, dtmin = c->minTimeStep, alpha = c->alphaStart;
    double fnorm, fmax, vnorm, vmax, e;
    int nstep = 1, nuphill = 0, status = AKO_FAILED, it;
    e = fn(ctx, n, x, y, z, fx, fy, fz);
    for (int i = 0; i < n; ++i)
        if (frozen && frozen[i] == 1) fx[i] = fy[i] = fz[i] = 0.0;
    double half = 0.5 * dt * c->timeStepUnit;
    for (it = 1; it <= c->maxIter; ++it) {
        double P = 0.0;
        for (int i = 0; i < n; ++i) P += fx[i] * vx[i] + fy[i] * vy[i] + fz[i] * vz[i];
        double oma = 1.0 - alpha;
        norm_max3(n, fx, fy, fz, &fnorm, &fmax);
        norm_max3(n, vx, vy, vz, &vnorm, &vmax);
        double af = alpha * vnorm / fnorm;
        if (P > 0) {
            if (nstep > c->N_min) {
                dt = fmin(dt * c->stepIncrease, dtmax);
                alpha *= c->alphaDecrease;
            }
            nstep += 1;
            nuphill = 0;
        } else {
            nstep = 0;
            nuphill += 1;
            if (nuphill > c->N_maxUphill) {
                status = AKO_FAILED;
                break;
            }
            if (c->initialDelay) {
                if (it >= c->N_delaySteps) {
                    dt *= c->stepDecrease;
                    if (dt < dtmin) dt = dtmin;
                    alpha = c->alphaStart;
                }
            } else {
                dt *= c->stepDecrease;
                alpha = c->alphaStart;
            }
            for (int i = 0; i < n; ++i)
                if (!frozen || frozen[i] == 0) {
                    x[i] -= 0.5 * dt * vx[i];
                    y[i] -= 0.5 * dt * vy[i];
                    z[i] -= 0.5 * dt * vz[i];
                }
            for (int i = 0; i < n; ++i) vx[i] = vy[i] = vz[i] = 0.0;
        }
        if (c->euler) {
            double fact = dt * c->timeStepUnit;
            for (int i = 0; i < n; ++i)
                if (!frozen || frozen[i] == 0) {
                    if (P > 0.0) {
                        vx[i] = oma * vx[i] + af * fx[i];
                        vy[i] = oma * vy[i] + af * fy[i];
                        vz[i] = oma * vz[i] + af * fz[i];
                    }
                    x[i] += dt * vx[i];
                    y[i] += dt * vy[i];
                    z[i] += dt * vz[i];
                    vx[i] += fact * fx[i];
                    vy[i] += fact * fy[i];
                    vz[i] += fact * fz[i];
                }
        } else {
            half = 0.5 * dt * c->timeStepUnit;
            for (int i = 0; i < n; ++i) {
                if (!frozen || frozen[i] == 0) {
                    if (P > 0.0) {
                        vx[i] = oma * vx[i] + af * fx[i];
                        vy[i] = oma * vy[i] + af * fy[i];
                        vz[i] = oma * vz[i] + af * fz[i];
                    }
                    vx[i] += half * fx[i];
                    vy[i] += half * fy[i];
                    vz[i] += half * fz[i];
                } else {
                    vx[i] = vy[i] = vz[i] = 0.0;
                }
            }
            for (int i = 0; i < n; ++i)
                if (!frozen || frozen[i] == 0) {
                    x[i] += dt * vx[i];
                    y[i] += dt * vy[i];
                    z[i] += dt * vz[i];
                }
        }
        e = fn(ctx, n, x, y, z, fx, fy, fz);
        if (frozen)
            for (int i = 0; i < n; ++i)
                if (frozen[i] == 1) fx[i] = fy[i] = fz[i] = 0.0;
        if (!c->euler) {
            for (int i = 0; i < n; ++i) {
                if (!frozen || frozen[i] == 0) {
                    vx[i] += half * fx[i];
                    vy[i] += half * fy[i];
                    vz[i] += half * fz[i];
                } else {
                    vx[i] = vy[i] = vz[i] = 0.0;
                }
            }
        }
        double mx = 0.0, my = 0.0, mz = 0.0;
        for (int i = 0; i < n; ++i) {
            mx += vx[i];
            my += vy[i];
            mz += vz[i];
        }
        mx /= (double)n;
        my /= (double)n;
        mz /= (double)n;
        for (int i = 0; i < n; ++i) {
            vx[i] -= mx;
            vy[i] -= my;
            vz[i] -= mz;
        }
        norm_max3(n, fx, fy, fz, &fnorm, &fmax);
        if (fnorm < c->normTol || fmax < c->forceTol) {
            status = AKO_SUCCESS;
            break;
        }
    }
    if (iters) *iters = (it > c->maxIter) ? c->maxIter : it;
    if (energy) *energy = e;
    free(vx); free(vy); free(vz); free(fx); free(fy); free(fz);
    return status;
}

/* ================================================================== */
/* saddle-search helpers                                               */
/* ================================================================== */

/* regionStyle 2: freeze atoms beyond radius of the centre atom
 * (optim/SaddleSearch.cpp:73-135); returns the number of active atoms */
int ako_active_sphere(int n, const double *x, const double *y, const double *z,
                      const double lat[9], const double rcp[9], int centre, double radius,
                      int *frozen)
{
    int act = 0;
    double r2 = radius * radius;
    for (int i = 0; i < n; ++i) {
        if (i == centre || frozen[i] == 1) {
            if (frozen[i] == 0) ++act;
            continue;
        }
        double rx, ry, rz;
        if (min_image_rsq(x[i] - x[centre], y[i] - y[centre], z[i] - z[centre], lat, rcp, &rx, &ry, &rz) > r2)
            frozen[i] = 1;
        else
            ++act;
    }
    return act;
}

/* makeRandomDisplacement (optim/SaddleSearch.cpp:356-421): u - 1/2 per
 * active component, centre-of-mass removal over the active vector, unit
 * direction in rvec, displacement = (u - 1/2 - mean) * d0 (not normalised) */
void ako_random_displacement(int n, double *x, double *y, double *z, const int *frozen,
                             const double *u, double d0, double *rvec)
{
    int m = 0;
    for (int i = 0; i < n; ++i)
        if (frozen[i] == 0) m += 3;
    double *d = (double *)malloc(sizeof(double) * (m > 0 ? m : 1));
    for (int j = 0; j < m; ++j) d[j] = u[j] - 0.5;
    int na = m / 3;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int i = 0; i < na; ++i) {
        sx += d[3 * i];
        sy += d[3 * i + 1];
        sz += d[3 * i + 2];
    }
    sx /= na;
    sy /= na;
    sz /= na;
    for (int i = 0; i < na; ++i) {
        d[3 * i] -= sx;
        d[3 * i + 1] -= sy;
        d[3 * i + 2] -= sz;
    }
    double nrm = 0.0;
    for (int j = 0; j < m; ++j) nrm += d[j] * d[j];
    double inv = 1.0 / sqrt(nrm);
    for (int j = 0; j < m; ++j) {
        rvec[j] = inv * d[j];
        d[j] *= d0;
    }
    for (int i = 0, j = 0; i < n; ++i)
        if (frozen[i] == 0) {
            x[i] += d[j];
            y[i] += d[j + 1];
            z[i] += d[j + 2];
            j += 3;
        }
    free(d);
}

/* Symmetric tridiagonal eigenproblem by the implicit QL algorithm with
 * Wilkinson-type shifts (Bowdler, Martin, Reinsch & Wilkinson 1968, the
 * tql2 procedure), accumulating the plane rotations into z.  Replaces the
 * LAPACK dstev call of optim/ART.cpp:764-794; eigenvalues are returned in
 * ascending order like dstev's. */
int ako_tridiag_eig(int n, const double *d_in, const double *e_in, double *w, double *z)
{
    double *d = (double *)malloc(sizeof(double) * n), *e = (double *)malloc(sizeof(double) * n);
    for (int i = 0; i < n; ++i) {
        d[i] = d_in[i];
        e[i] = (i < n - 1) ? e_in[i] : 0.0;
    }
    for (int i = 0; i < n * n; ++i) z[i] = 0.0;
    for (int i = 0; i < n; ++i) z[i * n + i] = 1.0;
    int rc = 0;
    for (int l = 0; l < n; ++l) {
        int iter = 0, m;
        for (;;) {
            for (m = l; m < n - 1; ++m) {
                double dd = fabs(d[m]) + fabs(d[m + 1]);
                if (fabs(e[m]) <= 2.220446049250313e-16 * dd) break;
            }
            if (m == l) break;
            if (++iter > 60) {
                rc = 1;
                goto done;
            }
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = hypot(g, 1.0);
            g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? fabs(r) : -fabs(r)));
            double s = 1.0, c = 1.0, p = 0.0;
            int i;
            for (i = m - 1; i >= l; --i) {
                double f = s * e[i], b = c * e[i];
                r = hypot(f, g);
                e[i + 1] = r;
                if (r == 0.0) {
                    d[i + 1] -= p;
                    e[m] = 0.0;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2.0 * c * b;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - b;
                /* column-major z: column k is z[k*n .. k*n+n-1] */
                for (int k = 0; k < n; ++k) {
                    double t = z[(i + 1) * n + k];
                    z[(i + 1) * n + k] = s * z[i * n + k] + c * t;
                    z[i * n + k] = c * z[i * n + k] - s * t;
                }
            }
            if (r == 0.0 && i >= l) continue;
            d[l] -= p;
            e[l] = g;
            e[m] = 0.0;
        }
    }
    /* sort ascending with vectors */
    for (int i = 0; i < n - 1; ++i) {
        int k = i;
        for (int j = i + 1; j < n; ++j)
            if (d[j] < d[k]) k = j;
        if (k != i) {
            double t = d[i];
            d[i] = d[k];
            d[k] = t;
            for (int r = 0; r < n; ++r) {
                t = z[i * n + r];
                z[i * n + r] = z[k * n + r];
                z[k * n + r] = t;
            }
        }
    }
    for (int i = 0; i < n; ++i) w[i] = d[i];
done:
    free(d);
    free(e);
    return rc;
}

/* ================================================================== */
/* ART                                                                 */
/* ================================================================== */

/* defaults of Common/ArtCntrl.h:14-74 */
void ako_art_defaults(ako_art_cntrl *c)
{
    c->numLanczosVectors = 20;
    c->deltaX = 0.01;
    c->initialDisplacement = 0.01;
    c->maxEigenvalue = 0.0;
    c->eigValTol = 0.001;
    c->parDamp = 0.5;
    c->goPerpendicular = 0;
    c->maxParIter = 1000;
    c->maxPerpIter = 20;
    c->normTol = 1.0e-4;
    c->forceTol = 1.0e-3;
    c->N_min = 5;
    c->initTimeStep = 0.005;
    c->alphaDecrease = 0.99;
    c->stepIncrease = 1.1;
    c->stepDecrease = 0.5;
    c->alphaStart = 0.1;
    c->maxTimeStep = 0.01;
    c->initTimeStepPerp = 0.005;
    c->maxTimeStepPerp = 0.01;
    c->timeStepUnit = AKO_EVTODL;
}

static void displace_active(int n, double *x, double *y, double *z, const int *frozen,
                            const double *d, double s)
{
    for (int i = 0, j = 0; i < n; ++i)
        if (frozen[i] == 0) {
            x[i] += s * d[j];
            y[i] += s * d[j + 1];
            z[i] += s * d[j + 2];
            j += 3;
        }
}

/* force on the active atoms packed as a 3*nactive vector
 * (ART::calculateForces, optim/ART.cpp:372-417) */
static double active_force(ako_force_fn fn, void *ctx, int n, double *x, double *y, double *z,
                           const int *frozen, double *fv, double *wx, double *wy, double *wz)
{
    double e = fn(ctx, n, x, y, z, wx, wy, wz);
    for (int i = 0, j = 0; i < n; ++i)
        if (frozen[i] == 0) {
            fv[j] = wx[i];
            fv[j + 1] = wy[i];
            fv[j + 2] = wz[i];
            j += 3;
        }
    return e;
}

static double vdot(const double *a, const double *b, int m)
{
    double d = 0.0;
    for (int i = 0; i < m; ++i) d += a[i] * b[i];
    return d;
}

/* ART::calculateLanczosForce (optim/ART.cpp:133-371); returns converged */
int ako_lanczos(ako_force_fn fn, void *ctx, int n, double *x, double *y, double *z,
                const int *frozen, double *rvec, int nl, double deltaX, double eigValTol,
                double *eigval, double *fpar, double *fperp, double *energy)
{
    int m = 0;
    for (int i = 0; i < n; ++i)
        if (frozen[i] == 0) m += 3;
    if (nl > m) nl = m;
    double *q = (double *)malloc(sizeof(double) * m), *u = (double *)malloc(sizeof(double) * m),
           *f0 = (double *)malloc(sizeof(double) * m), *fd = (double *)malloc(sizeof(double) * m),
           *Q = (double *)calloc((size_t)nl * m, sizeof(double)), *dT = (double *)calloc(nl, sizeof(double)),
           *sT = (double *)calloc(nl, sizeof(double)), *yv = (double *)calloc(nl, sizeof(double)),
           *wx = (double *)malloc(sizeof(double) * n), *wy = (double *)malloc(sizeof(double) * n),
           *wz = (double *)malloc(sizeof(double) * n), *w = (double *)malloc(sizeof(double) * nl),
           *zz = (double *)malloc(sizeof(double) * nl * nl);
    double lowest = 1.0e6, evnew = 0.0, evold = 0.0;
    int converged = 0;
    double beta = sqrt(vdot(rvec, rvec, m));
    *energy = active_force(fn, ctx, n, x, y, z, frozen, f0, wx, wy, wz);
    int i = 0;
    do {
        for (int j = 0; j < m; ++j) {
            q[j] = rvec[j] / beta;
            Q[(size_t)i * m + j] = q[j];
        }
        displace_active(n, x, y, z, frozen, q, deltaX);
        active_force(fn, ctx, n, x, y, z, frozen, fd, wx, wy, wz);
        displace_active(n, x, y, z, frozen, q, -deltaX);
        for (int j = 0; j < m; ++j) u[j] = -(fd[j] - f0[j]) / deltaX;
        if (i == 0)
            for (int j = 0; j < m; ++j) rvec[j] = u[j];
        else
            for (int j = 0; j < m; ++j) rvec[j] = u[j] - beta * Q[(size_t)(i - 1) * m + j];
        double alpha = vdot(q, rvec, m);
        for (int j = 0; j < m; ++j) rvec[j] -= alpha * Q[(size_t)i * m + j];
        dT[i] = alpha;
        if (i > 0) sT[i - 1] = beta;
        beta = sqrt(vdot(rvec, rvec, m));
        if (i == 0) {
            evnew = evold = alpha;
        } else {
            if (ako_tridiag_eig(i + 1, dT, sT, w, zz) != 0) {
                converged = -1;
                break;
            }
            evnew = w[0];
            for (int k = 0; k <= i; ++k) yv[k] = zz[k];
            double delta = fabs((evnew - evold) / evold);
            evold = evnew;
            lowest = evnew;
            if (delta < eigValTol) converged = 1;
        }
        ++i;
    } while (i < nl && converged == 0);
    if (converged >= 0) {
        double *ev = u;
        for (int j = 0; j < m; ++j) ev[j] = 0.0;
        for (int k = 0; k < nl; ++k)
            for (int j = 0; j < m; ++j) ev[j] += yv[k] * Q[(size_t)k * m + j];
        double nrm = sqrt(vdot(ev, ev, m));
        for (int j = 0; j < m; ++j) ev[j] /= nrm;
        double dot = vdot(f0, ev, m);
        for (int j = 0; j < m; ++j) {
            double fvv = -dot * ev[j];
            fpar[j] = f0[j] + 2.0 * fvv;
            fperp[j] = fvv;
            rvec[j] = ev[j];
        }
    }
    *eigval = lowest;
    free(q); free(u); free(f0); free(fd); free(Q); free(dT); free(sT); free(yv);
    free(wx); free(wy); free(wz); free(w); free(zz);
    return converged;
}

static void vnorm_max(const double *v, int m, double *nrm, double *mx)
{
    double s = 0.0, x = 0.0;
    for (int i = 0; i < m; ++i) {
        double q = v[i] * v[i];
        if (q > x) x = q;
        s += q;
    }
    *nrm = sqrt(s);
    *mx = sqrt(x);
}

/* ART::FIRESaddlePoint (optim/ART.cpp:421-756) for the parallel phase and,
 * when goPerpendicular is set, the perpendicular phase first
 * (optim/ART.cpp:90-110) */
static int art_fire(ako_force_fn fn, void *ctx, int n, double *x, double *y, double *z,
                    const int *frozen, double *rvec, const ako_art_cntrl *c, int goPerp,
                    int *itART, double *efinal, double *evfinal)
{
    int m = 0;
    for (int i = 0; i < n; ++i)
        if (frozen[i] == 0) m += 3;
    int nl = c->numLanczosVectors < m ? c->numLanczosVectors : m;
    double *v = (double *)calloc(m, sizeof(double)), *fo = (double *)calloc(m, sizeof(double)),
           *fpar = (double *)calloc(m, sizeof(double)), *fperp = (double *)calloc(m, sizeof(double));
    int maxIt = goPerp ? c->maxPerpIter : c->maxParIter;
    double dt = goPerp ? c->initTimeStepPerp : c->initTimeStep;
    double dtmax = goPerp ? c->maxTimeStepPerp : c->maxTimeStep;
    double alpha = c->alphaStart, ev, eold, enew;
    int nstep = 1, status = AKO_FAILED;
    if (ako_lanczos(fn, ctx, n, x, y, z, frozen, rvec, nl, c->deltaX, c->eigValTol, &ev, fpar, fperp, &eold) < 0)
        goto out;
    double half = 0.5 * dt * c->timeStepUnit;
    while (*itART <= maxIt) {
        (*itART)++;
        int perp = (ev >= c->maxEigenvalue) || goPerp;
        for (int i = 0; i < m; ++i) fo[i] = perp ? fperp[i] : fpar[i] * c->parDamp;
        for (int i = 0; i < m; ++i) v[i] += half * fo[i];
        double npar, mpar, nperp, mperp, vn, vm;
        vnorm_max(fpar, m, &npar, &mpar);
        vnorm_max(fperp, m, &nperp, &mperp);
        vnorm_max(v, m, &vn, &vm);
        double P;
        double oma = 1.0 - alpha;
        if (perp) {
            P = vdot(fperp, v, m);
            for (int i = 0; i < m; ++i) v[i] = oma * v[i] + (alpha * fperp[i] * vn / nperp);
        } else {
            P = vdot(fpar, v, m);
            for (int i = 0; i < m; ++i) v[i] = oma * v[i] + (alpha * fpar[i] * vn / npar);
        }
        if (P > 0) {
            if (nstep > c->N_min) {
                dt = fmin(dt * c->stepIncrease, dtmax);
                alpha *= c->alphaDecrease;
            }
            nstep += 1;
        } else {
            nstep = 0;
            dt *= c->stepDecrease;
            alpha = c->alphaStart;
            for (int i = 0; i < m; ++i) v[i] = 0.0;
        }
        for (int i = 0, j = 0; i < n; ++i)
            if (frozen[i] == 0) {
                x[i] += dt * v[j];
                y[i] += dt * v[j + 1];
                z[i] += dt * v[j + 2];
                j += 3;
            }
        if (ako_lanczos(fn, ctx, n, x, y, z, frozen, rvec, nl, c->deltaX, c->eigValTol, &ev, fpar, fperp, &enew) < 0)
            goto out;
        perp = (ev >= c->maxEigenvalue) || goPerp;
        for (int i = 0; i < m; ++i) v[i] += half * (perp ? fperp[i] : fpar[i]);
        /* the convergence test uses the norms taken before the move
         * (optim/ART.cpp:567-569 vs 694-697) */
        if (*itART > 5) {
            if ((npar + nperp) < c->normTol || (mpar + mperp) < c->forceTol) {
                status = (ev <= c->maxEigenvalue) ? AKO_SUCCESS : AKO_FAILED;
                *efinal = enew;
                *evfinal = ev;
                goto out;
            }
        }
        half = 0.5 * dt * c->timeStepUnit;
        eold = enew;
        if (fabs(half) < 1.0e-10) goto out;
    }
out:
    (void)eold;
    free(v); free(fo); free(fpar); free(fperp);
    return status;
}

int ako_art_search(ako_force_fn fn, void *ctx, int n, double *x, double *y, double *z,
                   const int *frozen, double *rvec, const ako_art_cntrl *c,
                   double *saddle_energy, double *eigval, int *iters)
{
    int itART = 0, st;
    double e = 0.0, ev = 0.0;
    if (c->goPerpendicular) {
        st = art_fire(fn, ctx, n, x, y, z, frozen, rvec, c, 1, &itART, &e, &ev);
        if (st != AKO_SUCCESS) {
            if (iters) *iters = itART;
            return st;
        }
    }
    st = art_fire(fn, ctx, n, x, y, z, frozen, rvec, c, 0, &itART, &e, &ev);
    if (saddle_energy) *saddle_energy = e;
    if (eigval) *eigval = ev;
    if (iters) *iters = itART;
    return st;
}

/* ================================================================== */
/* Dimer                                                               */
/* ================================================================== */

/* DimerCntrl defaults (Common/DimerCntrl.h:18-88) where the reference has
 * them; maxParIter and the translation time steps are this build's (the
 * reference's translation loop never runs, optim/Dimer.cpp:15-42). */
void ako_dimer_defaults(ako_dimer_cntrl *c)
{
    c->deltaTheta = 0.005;
    c->dimerDisplacement = 0.01;
    c->maxDeltaX = 0.5;
    c->maxRotations = 20;
    c->maxParIter = 1000;
    c->maxEigenvalue = 0.0;
    c->normTol = 0.1;
    c->forceTol = 1.0e-2;
    c->N_min = 5;
    c->alphaDecrease = 0.99;
    c->stepIncrease = 1.1;
    c->stepDecrease = 0.5;
    c->alphaStart = 0.1;
    c->initTimeStep = 0.005;
    c->maxTimeStep = 0.01;
    c->maxStep = 0.1;
    c->timeStepUnit = AKO_EVTODL;
}

/*
 * Minimum-mode following with a dimer (Henkelman & Jonsson 1999; rotation
 * by a trial angle and harmonic fit, Olsen et al. 2004 / Heyden et al.
 * 2005), translated by FIRE on the effective force.  The per-iteration
 * structure follows optim/Dimer.cpp:416-619: F2 = 2F0 - F1 (:454-456),
 * rotational force from the two ends (:463-474), curvature from the end
 * forces (:492), F_eff = -F_par if C > 0 else F - 2F_par (:527-536).
 *
 * One iteration:
 *   F0 = F(x);  F1 = F(x + D n);
 *   repeat (<= maxRotations):
 *     G = (F0 - F1)/D  (~ H n);  C = G.n;  Gp = G - C n;  g = |Gp|
 *     stop if g < normTol
 *     T = -Gp/g;  trial angle t1 = atan(g/|C|)/2 (Heyden et al. 2005); stop
 *       if t1 < deltaTheta
 *     n' = n cos t1 + T sin t1;  F1' = F(x + D n')
 *     C' = (F0 - F1').n'/D;  b1 = -g;  a1 = (C - C' + b1 sin 2t1)/(1 - cos 2t1)
 *     th = atan(b1/a1)/2; if a1 cos 2th + b1 sin 2th > 0: th += pi/2 (then
 *       folded into (-pi/2, pi/2])
 *     F1 <- F1 sin(t1-th)/sin t1 + F1' sin th/sin t1 + F0 (1 - cos th - sin th tan(t1/2))
 *     n <- normalise(n cos th + T sin th)
 *   C = (F0 - F1).n/D
 *   converged if max_i |F0_i| < forceTol and C < maxEigenvalue
 *   Fe = -(F0.n)n if C > 0 else F0 - 2(F0.n)n
 *   FIRE: P = Fe.v; P > 0: v = (1-a)v + a|v| Fe/|Fe|, grow dt after N_min;
 *         else v = 0, dt *= stepDecrease, a = alphaStart;
 *         v += dt*unit*Fe;  dx = dt*v;  if max_i |dx_i| > maxStep scale dx and v
 *         by maxStep / max_i |dx_i|;  x += dx
 */
int ako_dimer_search(ako_force_fn fn, void *ctx, int n, double *x, double *y, double *z,
                     const int *frozen, double *mode, const ako_dimer_cntrl *c,
                     double *saddle_energy, double *curvature, int *iters, int *fcalls)
{
    int m = 0;
    for (int i = 0; i < n; ++i)
        if (frozen[i] == 0) m += 3;
    double *F0 = (double *)malloc(sizeof(double) * m), *F1 = (double *)malloc(sizeof(double) * m),
           *F1t = (double *)malloc(sizeof(double) * m), *G = (double *)malloc(sizeof(double) * m),
           *T = (double *)malloc(sizeof(double) * m), *nt = (double *)malloc(sizeof(double) * m),
           *v = (double *)calloc(m, sizeof(double)), *Fe = (double *)malloc(sizeof(double) * m),
           *wx = (double *)malloc(sizeof(double) * n), *wy = (double *)malloc(sizeof(double) * n),
           *wz = (double *)malloc(sizeof(double) * n);
    const double D = c->dimerDisplacement, dth = c->deltaTheta;
    double dt = c->initTimeStep, alpha = c->alphaStart, C = 0.0, E0 = 0.0;
    int nstep = 0, status = AKO_FAILED, it, nf = 0;
    double nn = sqrt(vdot(mode, mode, m));
    for (int j = 0; j < m; ++j) mode[j] /= nn;
    for (it = 1; it <= c->maxParIter; ++it) {
        E0 = active_force(fn, ctx, n, x, y, z, frozen, F0, wx, wy, wz);
        displace_active(n, x, y, z, frozen, mode, D);
        active_force(fn, ctx, n, x, y, z, frozen, F1, wx, wy, wz);
        displace_active(n, x, y, z, frozen, mode, -D);
        nf += 2;
        for (int r = 0; r < c->maxRotations; ++r) {
            for (int j = 0; j < m; ++j) G[j] = (F0[j] - F1[j]) / D;
            C = vdot(G, mode, m);
            double g2 = 0.0;
            for (int j = 0; j < m; ++j) {
                T[j] = G[j] - C * mode[j];
                g2 += T[j] * T[j];
            }
            double g = sqrt(g2);
            if (g < c->normTol) break;
            double t1 = 0.5 * atan(g / fabs(C));
            if (t1 < dth) break;
            for (int j = 0; j < m; ++j) {
                T[j] = -T[j] / g;
                nt[j] = mode[j] * cos(t1) + T[j] * sin(t1);
            }
            displace_active(n, x, y, z, frozen, nt, D);
            active_force(fn, ctx, n, x, y, z, frozen, F1t, wx, wy, wz);
            displace_active(n, x, y, z, frozen, nt, -D);
            nf += 1;
            double Ct = 0.0;
            for (int j = 0; j < m; ++j) Ct += (F0[j] - F1t[j]) * nt[j];
            Ct /= D;
            double b1 = -g;
            double a1 = (C - Ct + b1 * sin(2.0 * t1)) / (1.0 - cos(2.0 * t1));
            double th = 0.5 * atan(b1 / a1);
            if (a1 * cos(2.0 * th) + b1 * sin(2.0 * th) > 0.0) th += 0.5 * AKO_PI;
            if (th > 0.5 * AKO_PI) th -= AKO_PI;
            double s0 = sin(t1 - th) / sin(t1), s1 = sin(th) / sin(t1);
            double s2 = 1.0 - cos(th) - sin(th) * tan(0.5 * t1);
            double ct = cos(th), st = sin(th), nn2 = 0.0;
            for (int j = 0; j < m; ++j) {
                F1[j] = F1[j] * s0 + F1t[j] * s1 + F0[j] * s2;
                mode[j] = mode[j] * ct + T[j] * st;
                nn2 += mode[j] * mode[j];
            }
            nn2 = 1.0 / sqrt(nn2);
            for (int j = 0; j < m; ++j) mode[j] *= nn2;
        }
        C = 0.0;
        for (int j = 0; j < m; ++j) C += (F0[j] - F1[j]) * mode[j];
        C /= D;
        double fmax = 0.0;
        for (int j = 0; j < m; j += 3) {
            double q = F0[j] * F0[j] + F0[j + 1] * F0[j + 1] + F0[j + 2] * F0[j + 2];
            if (q > fmax) fmax = q;
        }
        fmax = sqrt(fmax);
        if (fmax < c->forceTol && C < c->maxEigenvalue) {
            status = AKO_SUCCESS;
            break;
        }
        double fn_ = vdot(F0, mode, m);
        for (int j = 0; j < m; ++j) Fe[j] = (C > 0.0) ? -fn_ * mode[j] : F0[j] - 2.0 * fn_ * mode[j];
        double P = vdot(Fe, v, m);
        if (P > 0.0) {
            double vn = sqrt(vdot(v, v, m)), fe = sqrt(vdot(Fe, Fe, m));
            for (int j = 0; j < m; ++j) v[j] = (1.0 - alpha) * v[j] + alpha * vn * Fe[j] / fe;
            if (nstep > c->N_min) {
                dt = fmin(dt * c->stepIncrease, c->maxTimeStep);
                alpha *= c->alphaDecrease;
            }
            nstep += 1;
        } else {
            for (int j = 0; j < m; ++j) v[j] = 0.0;
            dt *= c->stepDecrease;
            alpha = c->alphaStart;
            nstep = 0;
        }
        double fact = dt * c->timeStepUnit, dmax = 0.0;
        for (int j = 0; j < m; ++j) v[j] += fact * Fe[j];
        for (int j = 0; j < m; j += 3) {
            double q = (dt * v[j]) * (dt * v[j]) + (dt * v[j + 1]) * (dt * v[j + 1]) + (dt * v[j + 2]) * (dt * v[j + 2]);
            if (q > dmax) dmax = q;
        }
        dmax = sqrt(dmax);
        double sc = (dmax > c->maxStep) ? c->maxStep / dmax : 1.0;
        for (int j = 0; j < m; ++j) {
            v[j] *= sc;
            T[j] = dt * v[j];
        }
        displace_active(n, x, y, z, frozen, T, 1.0);
    }
    if (saddle_energy) *saddle_energy = E0;
    if (curvature) *curvature = C;
    if (iters) *iters = (it > c->maxParIter) ? c->maxParIter : it;
    if (fcalls) *fcalls = nf;
    free(F0); free(F1); free(F1t); free(G); free(T); free(nt); free(v); free(Fe);
    free(wx); free(wy); free(wz);
    return status;
}
