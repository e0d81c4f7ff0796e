/*
 * ntust_markers.c -- CPU restatement of the Lagrangian-marker direct
 * forcing of the HIP library (k_marker.hip, ntust_set_markers).
 *
 * TEST INFRASTRUCTURE ONLY (the checker; see ntust_oracle.h).
 *
 * No counterpart in the reference: its immersed boundary is the Eulerian
 * volume fraction ETA (calcul_new_velocity.f90:61-100, airfoil.f90:580-892).
 * The marker form replaces that blend at the same place of the step:
 *   U = I u2 (u2 = u1 when ETA = 0), F = -U/dt, u2 += dt S F, FY, FZ = S F,
 * with I and S built on the cubic B-spline kernel (support 2 widths).  Windows, spread lists and summation orders
 * are the library's, written out again here, so the results agree bit for
 * bit (built without FMA contraction, IEEE division and sqrt).  Parity of
 * these functions is pinned by mathematics, not by the reference
 * (tests/test_marker_pins.py): I and S are transposes, the kernel's
 * weights sum to one and reproduce linear fields, the forcing stops a
 * uniform flow at the markers.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "ntust_oracle.h"

#define H(i, j, k) ((size_t)((i) + 1) + (size_t)(g->nx + 4) * ((size_t)((j) + 1) + (size_t)(g->ny + 4) * (size_t)((k) + 1)))
#define M1(a, i) ((a)[(i) + 1])
#define WIN 4          /* window points per axis in y and z */
#define RAD 2.0        /* kernel support, in widths */

typedef struct {
    int n;
    double hx, hy, hz;
    double *zm, *ym, *wm, *Xn, *Yn, *Zn;
    int *jw[3], *kw[3];
    int ne[3];
    int *sj[3], *sk[3], *so[3], *sm[3];
} or_markers;

static or_markers MK;

static void mk_free(void)
{
    free(MK.zm); free(MK.ym); free(MK.wm); free(MK.Xn); free(MK.Yn); free(MK.Zn);
    for (int a = 0; a < 3; a++) {
        free(MK.jw[a]); free(MK.kw[a]); free(MK.sj[a]); free(MK.sk[a]); free(MK.so[a]);
        free(MK.sm[a]);
    }
    memset(&MK, 0, sizeof MK);
}

/* the cubic B-spline (support |r| < 2) */
double or_mk_phi(double r)
{
    const double a = fabs(r);
    if (a <= 1.0) return (4.0 - 6.0 * (a * a) + 3.0 * (a * a * a)) / 6.0;
    if (a < 2.0) {
        const double b = 2.0 - a;
        return (b * b * b) / 6.0;
    }
    return 0.0;
}

double or_mk_dphi(double r)
{
    const double a = fabs(r);
    double d = 0.0;
    if (a <= 1.0) {
        d = (-12.0 * a + 9.0 * (a * a)) / 6.0;
    } else if (a < 2.0) {
        const double b = 2.0 - a;
        d = -(b * b) / 2.0;
    }
    return r < 0.0 ? -d : d;
}

/* component point along one axis: the upper face (node i+1) of cell i on the
 * component's own axis, else the cell centre (node arrays: element i at i+1) */
static double pt(const double *N, int i, int node)
{
    return node ? N[i + 2] : 0.5 * (N[i + 1] + N[i + 2]);
}

static double dvol(const or_grid *g, int a, int i, int j, int k)
{
    if (a == 0) return (M1(g->Dxs, i) * M1(g->iDy, j)) * M1(g->iDz, k);
    if (a == 1) return (M1(g->iDx, i) * M1(g->Dys, j)) * M1(g->iDz, k);
    return (M1(g->iDx, i) * M1(g->iDy, j)) * M1(g->Dzs, k);
}

static double *dup(const double *s, int n)
{
    double *d = malloc(sizeof(double) * (size_t)n);
    memcpy(d, s, sizeof(double) * (size_t)n);
    return d;
}

/* ntust_set_markers: 0, or -2 when the grid is finer than a kernel width at
 * a marker (more than 4 points within 2 h) */
int or_set_markers(const or_grid *g, int n, const double *zm, const double *ym,
                   const double *wm, double hx, double hy, double hz, const double *Xn,
                   const double *Yn, const double *Zn)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    mk_free();
    if (n <= 0) return 0;
    MK.hx = hx; MK.hy = hy; MK.hz = hz;
    MK.zm = dup(zm, n); MK.ym = dup(ym, n); MK.wm = dup(wm, n);
    MK.Xn = dup(Xn, nx + 5); MK.Yn = dup(Yn, ny + 5); MK.Zn = dup(Zn, nz + 5);
    for (int a = 0; a < 3; a++) {
        MK.jw[a] = malloc(sizeof(int) * (size_t)n);
        MK.kw[a] = malloc(sizeof(int) * (size_t)n);
        for (int m = 0; m < n; m++) {
            int j = 1, k = 1;
            while (j <= ny && !(pt(Yn, j, a == 1) > ym[m] - RAD * hy)) j++;
            while (k <= nz && !(pt(Zn, k, a == 2) > zm[m] - RAD * hz)) k++;
            if ((j + WIN <= ny && pt(Yn, j + WIN, a == 1) < ym[m] + RAD * hy) ||
                (k + WIN <= nz && pt(Zn, k + WIN, a == 2) < zm[m] + RAD * hz)) {
                mk_free();
                return -2;
            }
            MK.jw[a][m] = j;
            MK.kw[a][m] = k;
        }
        /* spread lists: interior points of the windows in increasing (k, j),
         * markers in increasing order */
        const size_t np = (size_t)(ny + 2) * (nz + 2);
        int *cnt = calloc(np, sizeof(int));
        for (int m = 0; m < n; m++)
            for (int dk = 0; dk < WIN; dk++)
                for (int dj = 0; dj < WIN; dj++) {
                    const int j = MK.jw[a][m] + dj, k = MK.kw[a][m] + dk;
                    if (j >= 1 && j <= ny && k >= 1 && k <= nz) cnt[(size_t)k * (ny + 2) + j]++;
                }
        int ne = 0, tot = 0;
        for (size_t q = 0; q < np; q++)
            if (cnt[q]) { ne++; tot += cnt[q]; }
        MK.ne[a] = ne;
        MK.sj[a] = malloc(sizeof(int) * (size_t)(ne > 0 ? ne : 1));
        MK.sk[a] = malloc(sizeof(int) * (size_t)(ne > 0 ? ne : 1));
        MK.so[a] = malloc(sizeof(int) * (size_t)(ne + 1));
        MK.sm[a] = malloc(sizeof(int) * (size_t)(tot > 0 ? tot : 1));
        int *start = calloc(np, sizeof(int));
        int e = 0, o = 0;
        MK.so[a][0] = 0;
        for (int k = 1; k <= nz; k++)
            for (int j = 1; j <= ny; j++) {
                const size_t q = (size_t)k * (ny + 2) + j;
                if (!cnt[q]) continue;
                MK.sj[a][e] = j;
                MK.sk[a][e] = k;
                start[q] = o;
                o += cnt[q];
                MK.so[a][++e] = o;
            }
        memset(cnt, 0, np * sizeof(int));
        for (int m = 0; m < n; m++)
            for (int dk = 0; dk < WIN; dk++)
                for (int dj = 0; dj < WIN; dj++) {
                    const int j = MK.jw[a][m] + dj, k = MK.kw[a][m] + dk;
                    if (j >= 1 && j <= ny && k >= 1 && k <= nz) {
                        const size_t q = (size_t)k * (ny + 2) + j;
                        MK.sm[a][start[q] + cnt[q]++] = m;
                    }
                }
        free(cnt);
        free(start);
    }
    MK.n = n;
    return 0;
}

int or_marker_count(void) { return MK.n; }

/* U[(a*n + m)*nx + ix-1] = sum_c w u_c dv_c over the window of (m, ix) */
void or_marker_interp(const or_grid *g, const double *u, const double *v, const double *w,
                      double *U)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz, n = MK.n;
    for (int m = 0; m < n; m++)
        for (int ix = 1; ix <= nx; ix++) {
            const double xm = 0.5 * (MK.Xn[ix + 1] + MK.Xn[ix + 2]), ym = MK.ym[m], zm = MK.zm[m];
            for (int a = 0; a < 3; a++) {
                const double *f = a == 0 ? u : (a == 1 ? v : w);
                double acc = 0.0;
                for (int dk = 0; dk < WIN; dk++) {
                    const int k = MK.kw[a][m] + dk;
                    if (k < 1 || k > nz) continue;
                    const double fz = or_mk_phi((pt(MK.Zn, k, a == 2) - zm) / MK.hz) / MK.hz;
                    for (int dj = 0; dj < WIN; dj++) {
                        const int j = MK.jw[a][m] + dj;
                        if (j < 1 || j > ny) continue;
                        const double fy = or_mk_phi((pt(MK.Yn, j, a == 1) - ym) / MK.hy) / MK.hy;
                        for (int i = ix - 2; i <= ix + 2; i++) {
                            if (i < 1 || i > nx) continue;
                            const double fx =
                                or_mk_phi((pt(MK.Xn, i, a == 0) - xm) / MK.hx) / MK.hx;
                            const double wt = (fx * fy) * fz;
                            acc = acc + (wt * f[H(i, j, k)]) * dvol(g, a, i, j, k);
                        }
                    }
                }
                U[((size_t)a * n + m) * nx + (ix - 1)] = acc;
            }
        }
}

/* f_a(c) = sum_m sum_ix w W F_a(m, ix) at the listed points of component a
 * (other entries untouched; sum order: list, then ix) */
void or_marker_spread(const or_grid *g, const double *F, double *fu, double *fv, double *fw)
{
    const int nx = g->nx, n = MK.n;
    double *out[3] = {fu, fv, fw};
    for (int a = 0; a < 3; a++)
        for (int e = 0; e < MK.ne[a]; e++)
            for (int i = 1; i <= nx; i++) {
                const int j = MK.sj[a][e], k = MK.sk[a][e];
                const double xp = pt(MK.Xn, i, a == 0), yp = pt(MK.Yn, j, a == 1),
                             zp = pt(MK.Zn, k, a == 2);
                double acc = 0.0;
                for (int q = MK.so[a][e]; q < MK.so[a][e + 1]; q++) {
                    const int m = MK.sm[a][q];
                    const double fy = or_mk_phi((yp - MK.ym[m]) / MK.hy) / MK.hy;
                    const double fz = or_mk_phi((zp - MK.zm[m]) / MK.hz) / MK.hz;
                    for (int ix = i - 2; ix <= i + 2; ix++) {
                        if (ix < 1 || ix > nx) continue;
                        const double xm = 0.5 * (MK.Xn[ix + 1] + MK.Xn[ix + 2]);
                        const double fx = or_mk_phi((xp - xm) / MK.hx) / MK.hx;
                        const double wt = (fx * fy) * fz;
                        const double W = MK.wm[m] * M1(g->iDx, ix);
                        acc = acc + (wt * W) * F[((size_t)a * n + m) * nx + (ix - 1)];
                    }
                }
                out[a][H(i, j, k)] = acc;
            }
}

/* The step's forcing on the blended velocity u2: U = I u2, F = -U/dt,
 * u2 += dt S F, FY, FZ = S F on the listed points; then uc, vc, wc
 * re-centred from u2 (calcul_new_velocity.f90:171-181).  Sum orders as
 * k_marker_spread<0>. */
void or_marker_forcing(const or_grid *g, double dt, double *u2, double *v2, double *w2,
                       double *FY, double *FZ, double *uc, double *vc, double *wc)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz, n = MK.n;
    if (n <= 0) return;
    double *U = malloc(sizeof(double) * 3 * (size_t)n * nx);
    or_marker_interp(g, u2, v2, w2, U);
    double *vel[3] = {u2, v2, w2}, *fo[3] = {NULL, FY, FZ};
    for (int a = 0; a < 3; a++)
        for (int e = 0; e < MK.ne[a]; e++)
            for (int i = 1; i <= nx; i++) {
                const int j = MK.sj[a][e], k = MK.sk[a][e];
                const double xp = pt(MK.Xn, i, a == 0), yp = pt(MK.Yn, j, a == 1),
                             zp = pt(MK.Zn, k, a == 2);
                double acc = 0.0;
                for (int q = MK.so[a][e]; q < MK.so[a][e + 1]; q++) {
                    const int m = MK.sm[a][q];
                    const double fy = or_mk_phi((yp - MK.ym[m]) / MK.hy) / MK.hy;
                    const double fz = or_mk_phi((zp - MK.zm[m]) / MK.hz) / MK.hz;
                    for (int ix = i - 2; ix <= i + 2; ix++) {
                        if (ix < 1 || ix > nx) continue;
                        const double xm = 0.5 * (MK.Xn[ix + 1] + MK.Xn[ix + 2]);
                        const double fx = or_mk_phi((xp - xm) / MK.hx) / MK.hx;
                        const double wt = (fx * fy) * fz;
                        const double W = MK.wm[m] * M1(g->iDx, ix);
                        const double val = U[((size_t)a * n + m) * nx + (ix - 1)];
                        acc = acc + (wt * W) * (-val / dt);
                    }
                }
                const size_t o = H(i, j, k);
                vel[a][o] = vel[a][o] + dt * acc;
                if (fo[a]) fo[a][o] = acc;
            }
    free(U);
    if (uc)
        for (int k = 1; k <= nz; k++)
            for (int j = 1; j <= ny; j++)
                for (int i = 1; i <= nx; i++) {
                    const size_t c = (size_t)(i - 1) + (size_t)nx * ((size_t)(j - 1) + (size_t)ny * (k - 1));
                    uc[c] = 0.5 * (u2[H(i, j, k)] + u2[H(i - 1, j, k)]);
                    vc[c] = 0.5 * (v2[H(i, j, k)] + v2[H(i, j - 1, k)]);
                    wc[c] = 0.5 * (w2[H(i, j, k)] + w2[H(i, j, k - 1)]);
                }
}
