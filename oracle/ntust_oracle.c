/*
 * ntust_oracle.c -- CPU restatement of the projection-method time step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library, never part of
 * the product path.  Built with -O2 -ffp-contract=off (no FMA contraction,
 * IEEE division, correctly rounded sqrt); glibc pow for the LES filter.
 *
 * Each function cites the reference subroutine (file:line) it restates.
 * Expressions keep the reference's textual operation order: C and Fortran
 * both associate * / + - left to right, and a leading unary minus that C
 * binds tighter than * yields the same bits because IEEE rounding is sign
 * symmetric.
 *
 * Loops whose cells are independent (the predictors, the LES gradients,
 * the right-hand side, one colour of the red-black sweep, the projection)
 * run over k with OpenMP, each cell computed exactly as in the serial loop;
 * the only reduction is a maximum, so the results do not depend on the
 * thread count.  GS.f90's lexicographic sweep and every floating-point sum
 * stay serial.  GS.f90:36 itself parallelises its sweep with PRIVATE(i)
 * only (pNew, pChange, mChange and both maxima shared, a race on the
 * order of updates); the red-black restatement here is the race-free form.
 */
#include "ntust_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define HX(g) ((g)->nx + 4)
#define HY(g) ((g)->ny + 4)
#define HZ(g) ((g)->nz + 4)
/* halo element (i,j,k), bounds -1..n+2 */
#define H(i, j, k) ((size_t)((i) + 1) + (size_t)sx * ((size_t)((j) + 1) + (size_t)sy * (size_t)((k) + 1)))
/* interior element (i,j,k), bounds 1..n */
#define I3(i, j, k) ((size_t)((i) - 1) + (size_t)nx * ((size_t)((j) - 1) + (size_t)ny * (size_t)((k) - 1)))
/* metric element m, bounds -1..n+2 */
#define M(a, m) ((a)[(m) + 1])

#define GRID_LOCALS(g)                    \
    const int nx = (g)->nx, ny = (g)->ny, nz = (g)->nz; \
    const int sx = nx + 4, sy = ny + 4;   \
    (void)nz; (void)sx; (void)sy

/* threads of the OpenMP loops (n <= 0: leave as is); returns the count */
int or_set_threads(int n)
{
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
}

/* ------------------------------------------------------------------ */
/* boundary_conditions.f90                                             */
/* ------------------------------------------------------------------ */

/* tangential component ghost (the mirror rules, e.g. :31-37) */
static double bc_tangent(int code, double interior, double old)
{
    if (code == OR_NOSLIP) return 0.0 - interior;   /* 2*0-v(1)   */
    if (code == OR_NEUMANN) return interior;        /* v(1)       */
    if (code == OR_DIRICHLET) return 2.0 - interior; /* 2*1-v(1)  */
    return old;                                     /* no match   */
}

/* wall-normal component on the wall face itself (e.g. :18-24) */
static double bc_normal(int code, double interior, double old)
{
    if (code == OR_NOSLIP) return 0.0;
    if (code == OR_NEUMANN) return interior;
    if (code == OR_DIRICHLET) return 1.0;
    return old;
}

void or_boundary_conditions(const or_grid *g, const int bc[18],
                            double *u, double *v, double *w, double *p,
                            double *u1, double *v1, double *w1,
                            double *u2, double *v2, double *w2,
                            double *us, double *vs, double *ws)
{
    GRID_LOCALS(g);
    int i, j, k;

    /* West, boundary_conditions.f90:14-57 */
    for (k = -1; k <= nz + 2; k++)
        for (j = -1; j <= ny + 2; j++) {
            u[H(0, j, k)] = bc_normal(bc[0], u[H(1, j, k)], u[H(0, j, k)]);
            u[H(-1, j, k)] = u[H(0, j, k)];
            v[H(0, j, k)] = bc_tangent(bc[1], v[H(1, j, k)], v[H(0, j, k)]);
            v[H(-1, j, k)] = v[H(0, j, k)];
            w[H(0, j, k)] = bc_tangent(bc[2], w[H(1, j, k)], w[H(0, j, k)]);
            w[H(-1, j, k)] = w[H(0, j, k)];
        }
    /* East, :60-105 */
    for (k = -1; k <= nz + 2; k++)
        for (j = -1; j <= ny + 2; j++) {
            u[H(nx, j, k)] = bc_normal(bc[3], u[H(nx - 1, j, k)], u[H(nx, j, k)]);
            u[H(nx + 1, j, k)] = u[H(nx, j, k)];
            u[H(nx + 2, j, k)] = u[H(nx + 1, j, k)];
            v[H(nx + 1, j, k)] = bc_tangent(bc[4], v[H(nx, j, k)], v[H(nx + 1, j, k)]);
            v[H(nx + 2, j, k)] = v[H(nx + 1, j, k)];
            w[H(nx + 1, j, k)] = bc_tangent(bc[5], w[H(nx, j, k)], w[H(nx + 1, j, k)]);
            w[H(nx + 2, j, k)] = w[H(nx + 1, j, k)];
        }
    /* South, :108-153 */
    for (k = -1; k <= nz + 2; k++)
        for (i = -1; i <= nx + 2; i++) {
            u[H(i, 0, k)] = bc_tangent(bc[6], u[H(i, 1, k)], u[H(i, 0, k)]);
            u[H(i, -1, k)] = u[H(i, 0, k)];
            v[H(i, 0, k)] = bc_normal(bc[7], v[H(i, 1, k)], v[H(i, 0, k)]);
            v[H(i, -1, k)] = v[H(i, 0, k)];
            w[H(i, 0, k)] = bc_tangent(bc[8], w[H(i, 1, k)], w[H(i, 0, k)]);
            w[H(i, -1, k)] = w[H(i, 0, k)];
        }
    /* North, :156-202 */
    for (k = -1; k <= nz + 2; k++)
        for (i = -1; i <= nx + 2; i++) {
            u[H(i, ny + 1, k)] = bc_tangent(bc[9], u[H(i, ny, k)], u[H(i, ny + 1, k)]);
            u[H(i, ny + 2, k)] = u[H(i, ny + 1, k)];
            v[H(i, ny, k)] = bc_normal(bc[10], v[H(i, ny - 1, k)], v[H(i, ny, k)]);
            v[H(i, ny + 1, k)] = v[H(i, ny, k)];
            v[H(i, ny + 2, k)] = v[H(i, ny + 1, k)];
            w[H(i, ny + 1, k)] = bc_tangent(bc[11], w[H(i, ny, k)], w[H(i, ny + 1, k)]);
            w[H(i, ny + 2, k)] = w[H(i, ny + 1, k)];
        }
    /* Back (inlet), :207-252.  The Neumann rule for v reads v(i,j,0)
     * itself (:228), so it leaves the ghost unchanged. */
    for (j = -1; j <= ny + 2; j++)
        for (i = -1; i <= nx + 2; i++) {
            u[H(i, j, 0)] = bc_tangent(bc[12], u[H(i, j, 1)], u[H(i, j, 0)]);
            u[H(i, j, -1)] = u[H(i, j, 0)];
            if (bc[13] == OR_NEUMANN)
                v[H(i, j, 0)] = v[H(i, j, 0)];
            else
                v[H(i, j, 0)] = bc_tangent(bc[13], v[H(i, j, 1)], v[H(i, j, 0)]);
            v[H(i, j, -1)] = v[H(i, j, 0)];
            w[H(i, j, 0)] = bc_normal(bc[14], w[H(i, j, 1)], w[H(i, j, 0)]);
            w[H(i, j, -1)] = w[H(i, j, 0)];
        }
    /* Front (outlet), :257-301 */
    for (j = -1; j <= ny + 2; j++)
        for (i = -1; i <= nx + 2; i++) {
            u[H(i, j, nz + 1)] = bc_tangent(bc[15], u[H(i, j, nz)], u[H(i, j, nz + 1)]);
            u[H(i, j, nz + 2)] = u[H(i, j, nz + 1)];
            v[H(i, j, nz + 1)] = bc_tangent(bc[16], v[H(i, j, nz)], v[H(i, j, nz + 1)]);
            v[H(i, j, nz + 2)] = v[H(i, j, nz + 1)];
            w[H(i, j, nz)] = bc_normal(bc[17], w[H(i, j, nz - 1)], w[H(i, j, nz)]);
            w[H(i, j, nz + 1)] = w[H(i, j, nz)];
            w[H(i, j, nz + 2)] = w[H(i, j, nz + 1)];
        }

    /* whole-array copies, :305-324 */
    {
        size_t n = (size_t)HX(g) * HY(g) * HZ(g);
        size_t b = n * sizeof(double);
        if (u1) memcpy(u1, u, b);
        if (v1) memcpy(v1, v, b);
        if (w1) memcpy(w1, w, b);
        if (u2) memcpy(u2, u, b);
        if (v2) memcpy(v2, v, b);
        if (w2) memcpy(w2, w, b);
        if (us) memcpy(us, u, b);
        if (vs) memcpy(vs, v, b);
        if (ws) memcpy(ws, w, b);
    }

    /* Neumann pressure ghosts, :329-374 */
    if (p) {
        for (k = -1; k <= nz + 2; k++)
            for (j = -1; j <= ny + 2; j++) {
                p[H(0, j, k)] = p[H(1, j, k)];
                p[H(-1, j, k)] = p[H(0, j, k)];
            }
        for (k = -1; k <= nz + 2; k++)
            for (j = -1; j <= ny + 2; j++) {
                p[H(nx + 1, j, k)] = p[H(nx, j, k)];
                p[H(nx + 2, j, k)] = p[H(nx + 1, j, k)];
            }
        for (k = -1; k <= nz + 2; k++)
            for (i = -1; i <= nx + 2; i++) {
                p[H(i, 0, k)] = p[H(i, 1, k)];
                p[H(i, -1, k)] = p[H(i, 0, k)];
            }
        for (k = -1; k <= nz + 2; k++)
            for (i = -1; i <= nx + 2; i++) {
                p[H(i, ny + 1, k)] = p[H(i, ny, k)];
                p[H(i, ny + 2, k)] = p[H(i, ny + 1, k)];
            }
        for (j = -1; j <= ny + 2; j++)
            for (i = -1; i <= nx + 2; i++) {
                p[H(i, j, 0)] = p[H(i, j, 1)];
                p[H(i, j, -1)] = p[H(i, j, 0)];
            }
        for (j = -1; j <= ny + 2; j++)
            for (i = -1; i <= nx + 2; i++) {
                p[H(i, j, nz + 1)] = p[H(i, j, nz)];
                p[H(i, j, nz + 2)] = p[H(i, j, nz + 1)];
            }
    }
}

/* ------------------------------------------------------------------ */
/* smagorinsky.f90                                                      */
/* ------------------------------------------------------------------ */

/* Phi(i,j,k) of the (nx,ny,nz) view = flat element (i-1)+nx(j-1)+nx*ny(k-1)
 * counted from the base address */
static double flat_read(const double *Phi, const double *prebase, long nxy,
                        long off)
{
    if (off >= 0) return Phi[off];
    if (prebase) return prebase[nxy + off];
    return 0.0;
}

void or_gradient_phi_gauss(const or_grid *g, const double *Phi,
                           const double *prebase, double *dPhidX,
                           const double *Dxm, const double *Dym,
                           const double *Dzm, int kstart, int kend)
{
    GRID_LOCALS(g);
    const long nxy = (long)nx * ny;
    const size_t n3 = (size_t)nx * ny * nz;
    #pragma omp parallel for schedule(static)
    for (int k = kstart; k <= kend; k++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) {
                long f = (long)(i - 1) + (long)nx * (j - 1) + nxy * (k - 1);
                double c = flat_read(Phi, prebase, nxy, f);
                /* smagorinsky.f90:60-67 */
                double pe = 0.5 * (flat_read(Phi, prebase, nxy, f + 1) + c);
                double pw = 0.5 * (flat_read(Phi, prebase, nxy, f - 1) + c);
                double pn = 0.5 * (flat_read(Phi, prebase, nxy, f + nx) + c);
                double ps = 0.5 * (flat_read(Phi, prebase, nxy, f - nx) + c);
                double pf = 0.5 * (flat_read(Phi, prebase, nxy, f + nxy) + c);
                double pb = 0.5 * (flat_read(Phi, prebase, nxy, f - nxy) + c);
                double dxm = M(Dxm, i), dym = M(Dym, j), dzm = M(Dzm, k);
                /* :69-73 */
                double gx = pw * (-1.0) * (dym * dzm) + pe * (1.0) * (dym * dzm);
                double gy = ps * (-1.0) * (dxm * dzm) + pn * (1.0) * (dxm * dzm);
                double gz = pb * (-1.0) * (dxm * dym) + pf * (1.0) * (dxm * dym);
                /* :75-79 */
                double fact = 1.0 / (dxm * dym * dzm);
                size_t o = I3(i, j, k);
                dPhidX[o] = fact * gx;
                dPhidX[o + n3] = fact * gy;
                dPhidX[o + 2 * n3] = fact * gz;
            }
}

/* delta = (iDx*iDy*iDz)**(1./3.) with the single-precision exponent
 * (smagorinsky.f90:20); (Cs*delta)**2 is (Cs*delta)*(Cs*delta) */
void or_smag_filter_table(const or_grid *g, double Cs, double *tab)
{
    GRID_LOCALS(g);
    const double third = (double)(1.0f / 3.0f);
    #pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) {
                double delta = pow(M(g->iDx, i) * M(g->iDy, j) * M(g->iDz, k), third);
                double cd = Cs * delta;
                tab[I3(i, j, k)] = cd * cd;
            }
}

void or_smagorinsky(const or_grid *g, const double *u, const double *v,
                    const double *w, const double *pre_u, const double *pre_v,
                    const double *pre_w, double Cs, int kstart, int kend,
                    double *dudx, double *dvdx, double *dwdx, double *Viseff)
{
    GRID_LOCALS(g);
    const size_t n3 = (size_t)nx * ny * nz;
    double *tu = dudx, *tv = dvdx, *tw = dwdx, *tab;
    if (!tu) tu = malloc(3 * n3 * sizeof(double));
    if (!tv) tv = malloc(3 * n3 * sizeof(double));
    if (!tw) tw = malloc(3 * n3 * sizeof(double));
    tab = malloc(n3 * sizeof(double));
    /* smagorinsky.f90:6-8 */
    or_gradient_phi_gauss(g, u, pre_u, tu, g->Dxs, g->iDy, g->iDz, kstart, kend);
    or_gradient_phi_gauss(g, v, pre_v, tv, g->iDx, g->Dys, g->iDz, kstart, kend);
    or_gradient_phi_gauss(g, w, pre_w, tw, g->iDx, g->iDy, g->Dzs, kstart, kend);
    or_smag_filter_table(g, Cs, tab);
    #pragma omp parallel for schedule(static)
    for (int k = kstart; k <= kend; k++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) {
                size_t o = I3(i, j, k);
                double ux = tu[o], uy = tu[o + n3], uz = tu[o + 2 * n3];
                double vx = tv[o], vy = tv[o + n3], vz = tv[o + 2 * n3];
                double wx = tw[o], wy = tw[o + n3], wz = tw[o + 2 * n3];
                /* :14-18 */
                double m = fabs(2.0 * (ux * ux) + 2.0 * (vy * vy) + 2.0 * (wz * wz) +
                                (uy + vx) * (uy + vx) + (uz + wx) * (uz + wx) +
                                (vz + wy) * (vz + wy));
                /* :20-24 */
                Viseff[o] = tab[o] * sqrt(m);
            }
    if (!dudx) free(tu);
    if (!dvdx) free(tv);
    if (!dwdx) free(tw);
    free(tab);
}

/* ------------------------------------------------------------------ */
/* discretisation_QUICK_centre.f90                                      */
/* ------------------------------------------------------------------ */

/* one QUICK face value in the reference's form
 *   0.5*(a+b) - 0.125*h*h/den*( (x1-x2)/d1 - (x3-x4)/d2 ) */
static inline double qface(double a, double b, double h, double den,
                           double x1, double x2, double d1,
                           double x3, double x4, double d2)
{
    return 0.5 * (a + b) - 0.125 * h * h / den * ((x1 - x2) / d1 - (x3 - x4) / d2);
}

void or_predictor_quick(const or_grid *g, const double *u, const double *v,
                        const double *w, const double *Viseff, double dt,
                        double nu, int LES, int kstart, int kend,
                        double *us, double *vs, double *ws)
{
    GRID_LOCALS(g);
    const double *iDx = g->iDx, *Dxs = g->Dxs, *iDy = g->iDy, *Dys = g->Dys;
    const double *iDz = g->iDz, *Dzs = g->Dzs;
    const double les = (double)LES;
    int i, j, k;
#define U(a, b, c) u[H(a, b, c)]
#define V(a, b, c) v[H(a, b, c)]
#define W(a, b, c) w[H(a, b, c)]

    /* u_star, :7-117 */
    #pragma omp parallel for private(i, j) schedule(static)
    for (k = kstart; k <= kend; k++)
        for (j = 1; j <= ny; j++)
            for (i = 1; i <= nx; i++) {
                double utx1 = 0.5 * (U(i + 1, j, k) + U(i, j, k));
                double utx2 = 0.5 * (U(i - 1, j, k) + U(i, j, k));
                double vtx1 = 0.5 * (V(i, j, k) + V(i + 1, j, k));
                double vtx2 = 0.5 * (V(i, j - 1, k) + V(i + 1, j - 1, k));
                double wtx1 = 0.5 * (W(i + 1, j, k) + W(i, j, k));
                double wtx2 = 0.5 * (W(i, j, k - 1) + W(i + 1, j, k - 1));
                double ue, uw, un, usx, uf, ub;
                if (utx1 > 0)
                    ue = qface(U(i, j, k), U(i + 1, j, k), M(iDx, i + 1), M(Dxs, i),
                               U(i + 1, j, k), U(i, j, k), M(iDx, i + 1),
                               U(i, j, k), U(i - 1, j, k), M(iDx, i));
                else
                    ue = qface(U(i, j, k), U(i + 1, j, k), M(iDx, i + 1), M(Dxs, i + 1),
                               U(i + 2, j, k), U(i + 1, j, k), M(iDx, i + 2),
                               U(i + 1, j, k), U(i, j, k), M(iDx, i + 1));
                if (utx2 > 0)
                    uw = qface(U(i - 1, j, k), U(i, j, k), M(iDx, i), M(Dxs, i - 1),
                               U(i, j, k), U(i - 1, j, k), M(iDx, i),
                               U(i - 1, j, k), U(i - 2, j, k), M(iDx, i - 1));
                else
                    uw = qface(U(i - 1, j, k), U(i, j, k), M(iDx, i), M(Dxs, i),
                               U(i + 1, j, k), U(i, j, k), M(iDx, i + 1),
                               U(i, j, k), U(i - 1, j, k), M(iDx, i));
                if (vtx1 > 0)
                    un = qface(U(i, j, k), U(i, j + 1, k), M(Dys, j), M(iDy, j),
                               U(i, j + 1, k), U(i, j, k), M(Dys, j),
                               U(i, j, k), U(i, j - 1, k), M(Dys, j - 1));
                else
                    un = qface(U(i, j, k), U(i, j + 1, k), M(Dys, j), M(iDy, j + 1),
                               U(i, j + 2, k), U(i, j + 1, k), M(Dys, j + 1),
                               U(i, j + 1, k), U(i, j, k), M(Dys, j));
                if (vtx2 > 0)
                    usx = qface(U(i, j - 1, k), U(i, j, k), M(Dys, j - 1), M(iDy, j - 1),
                                U(i, j, k), U(i, j - 1, k), M(Dys, j - 1),
                                U(i, j - 1, k), U(i, j - 2, k), M(Dys, j - 2));
                else
                    usx = qface(U(i, j - 1, k), U(i, j, k), M(Dys, j - 1), M(iDy, j),
                                U(i, j + 1, k), U(i, j, k), M(Dys, j),
                                U(i, j, k), U(i, j - 1, k), M(Dys, j - 1));
                if (wtx1 > 0)
                    uf = qface(U(i, j, k), U(i, j, k + 1), M(Dzs, k), M(iDz, k),
                               U(i, j, k + 1), U(i, j, k), M(Dzs, k),
                               U(i, j, k), U(i, j, k - 1), M(Dzs, k - 1));
                else
                    uf = qface(U(i, j, k), U(i, j, k + 1), M(Dzs, k), M(iDz, k + 1),
                               U(i, j, k + 2), U(i, j, k + 1), M(Dzs, k + 1),
                               U(i, j, k + 1), U(i, j, k), M(Dzs, k));
                if (wtx2 > 0)
                    ub = qface(U(i, j, k - 1), U(i, j, k), M(Dzs, k - 1), M(iDz, k - 1),
                               U(i, j, k), U(i, j, k - 1), M(Dzs, k - 1),
                               U(i, j, k - 1), U(i, j, k - 2), M(Dzs, k - 2));
                else
                    ub = qface(U(i, j, k - 1), U(i, j, k), M(Dzs, k - 1), M(iDz, k),
                               U(i, j, k + 1), U(i, j, k), M(Dzs, k),
                               U(i, j, k), U(i, j, k - 1), M(Dzs, k - 1));
                double nut = Viseff[I3(i, j, k)];
                double c = nu + nut * les;
                double u0 = U(i, j, k);
                /* :106-112 */
                us[H(i, j, k)] = u0 - dt * (utx1 * ue - utx2 * uw) / M(Dxs, i)
                                 - dt * (vtx1 * un - vtx2 * usx) / M(iDy, j)
                                 - dt * (wtx1 * uf - wtx2 * ub) / M(iDz, k)
                                 + c * dt * ((U(i + 1, j, k) - u0) / M(iDx, i + 1) - (u0 - U(i - 1, j, k)) / M(iDx, i)) / M(Dxs, i)
                                 + c * dt * ((U(i, j + 1, k) - u0) / M(Dys, j) - (u0 - U(i, j - 1, k)) / M(Dys, j - 1)) / M(iDy, j)
                                 + c * dt * ((U(i, j, k + 1) - u0) / M(Dzs, k) - (u0 - U(i, j, k - 1)) / M(Dzs, k - 1)) / M(iDz, k);
            }

    /* v_star, :123-230 */
    #pragma omp parallel for private(i, j) schedule(static)
    for (k = kstart; k <= kend; k++)
        for (j = 1; j <= ny; j++)
            for (i = 1; i <= nx; i++) {
                double uty1 = 0.5 * (U(i, j, k) + U(i, j + 1, k));
                double uty2 = 0.5 * (U(i - 1, j, k) + U(i - 1, j + 1, k));
                double vty1 = 0.5 * (V(i, j, k) + V(i, j + 1, k));
                double vty2 = 0.5 * (V(i, j - 1, k) + V(i, j, k));
                double wty1 = 0.5 * (W(i, j, k) + W(i, j + 1, k));
                double wty2 = 0.5 * (W(i, j, k - 1) + W(i, j + 1, k - 1));
                double ve, vw, vn, vsx, vf, vb;
                if (uty1 > 0)
                    ve = qface(V(i, j, k), V(i + 1, j, k), M(Dxs, i), M(iDx, i),
                               V(i + 1, j, k), V(i, j, k), M(Dxs, i),
                               V(i, j, k), V(i - 1, j, k), M(Dxs, i - 1));
                else
                    ve = qface(V(i, j, k), V(i + 1, j, k), M(Dxs, i), M(iDx, i + 1),
                               V(i + 2, j, k), V(i + 1, j, k), M(Dxs, i + 1),
                               V(i + 1, j, k), V(i, j, k), M(Dxs, i));
                if (uty2 > 0)
                    vw = qface(V(i - 1, j, k), V(i, j, k), M(Dxs, i - 1), M(iDx, i - 1),
                               V(i, j, k), V(i - 1, j, k), M(Dxs, i - 1),
                               V(i - 1, j, k), V(i - 2, j, k), M(Dxs, i - 2));
                else
                    vw = qface(V(i - 1, j, k), V(i, j, k), M(Dxs, i - 1), M(iDx, i),
                               V(i + 1, j, k), V(i, j, k), M(Dxs, i),
                               V(i, j, k), V(i - 1, j, k), M(Dxs, i - 1));
                if (vty1 > 0)
                    vn = qface(V(i, j, k), V(i, j + 1, k), M(iDy, j + 1), M(Dys, j),
                               V(i, j + 1, k), V(i, j, k), M(iDy, j + 1),
                               V(i, j, k), V(i, j - 1, k), M(iDy, j));
                else
                    vn = qface(V(i, j, k), V(i, j + 1, k), M(iDy, j + 1), M(Dys, j + 1),
                               V(i, j + 2, k), V(i, j + 1, k), M(iDy, j + 2),
                               V(i, j + 1, k), V(i, j, k), M(iDy, j + 1));
                if (vty2 > 0)
                    vsx = qface(V(i, j - 1, k), V(i, j, k), M(iDy, j), M(Dys, j - 1),
                                V(i, j, k), V(i, j - 1, k), M(iDy, j),
                                V(i, j - 1, k), V(i, j - 2, k), M(iDy, j - 1));
                else
                    vsx = qface(V(i, j - 1, k), V(i, j, k), M(iDy, j), M(Dys, j),
                                V(i, j + 1, k), V(i, j, k), M(iDy, j + 1),
                                V(i, j, k), V(i, j - 1, k), M(iDy, j));
                if (wty1 > 0)
                    vf = qface(V(i, j, k), V(i, j, k + 1), M(Dzs, k), M(iDz, k),
                               V(i, j, k + 1), V(i, j, k), M(Dzs, k),
                               V(i, j, k), V(i, j, k - 1), M(Dzs, k - 1));
                else
                    vf = qface(V(i, j, k), V(i, j, k + 1), M(Dzs, k), M(iDz, k + 1),
                               V(i, j, k + 2), V(i, j, k + 1), M(Dzs, k + 1),
                               V(i, j, k + 1), V(i, j, k), M(Dzs, k));
                if (wty2 > 0)
                    vb = qface(V(i, j, k - 1), V(i, j, k), M(Dzs, k - 1), M(iDz, k - 1),
                               V(i, j, k), V(i, j, k - 1), M(Dzs, k - 1),
                               V(i, j, k - 1), V(i, j, k - 2), M(Dzs, k - 2));
                else
                    vb = qface(V(i, j, k - 1), V(i, j, k), M(Dzs, k - 1), M(iDz, k),
                               V(i, j, k + 1), V(i, j, k), M(Dzs, k),
                               V(i, j, k), V(i, j, k - 1), M(Dzs, k - 1));
                double nut = Viseff[I3(i, j, k)];
                double c = nu + nut * les;
                double v0 = V(i, j, k);
                /* :220-226 */
                vs[H(i, j, k)] = v0 - dt * (uty1 * ve - uty2 * vw) / M(iDx, i)
                                 - dt * (vty1 * vn - vty2 * vsx) / M(Dys, j)
                                 - dt * (wty1 * vf - wty2 * vb) / M(iDz, k)
                                 + c * dt * ((V(i + 1, j, k) - v0) / M(Dxs, i) - (v0 - V(i - 1, j, k)) / M(Dxs, i - 1)) / M(iDx, i)
                                 + c * dt * ((V(i, j + 1, k) - v0) / M(iDy, j + 1) - (v0 - V(i, j - 1, k)) / M(iDy, j)) / M(Dys, j)
                                 + c * dt * ((V(i, j, k + 1) - v0) / M(Dzs, k) - (v0 - V(i, j, k - 1)) / M(Dzs, k - 1)) / M(iDz, k);
            }

    /* w_star, :236-342 */
    #pragma omp parallel for private(i, j) schedule(static)
    for (k = kstart; k <= kend; k++)
        for (j = 1; j <= ny; j++)
            for (i = 1; i <= nx; i++) {
                double utz1 = 0.5 * (U(i, j, k + 1) + U(i, j, k));
                double utz2 = 0.5 * (U(i - 1, j, k + 1) + U(i - 1, j, k));
                double vtz1 = 0.5 * (V(i, j, k) + V(i, j, k + 1));
                double vtz2 = 0.5 * (V(i, j - 1, k + 1) + V(i, j - 1, k));
                double wtz1 = 0.5 * (W(i, j, k + 1) + W(i, j, k));
                double wtz2 = 0.5 * (W(i, j, k) + W(i, j, k - 1));
                double we, ww, wn, wsx, wf, wb;
                if (utz1 > 0)
                    we = qface(W(i, j, k), W(i + 1, j, k), M(Dxs, i), M(iDx, i),
                               W(i + 1, j, k), W(i, j, k), M(Dxs, i),
                               W(i, j, k), W(i - 1, j, k), M(Dxs, i - 1));
                else
                    we = qface(W(i, j, k), W(i + 1, j, k), M(Dxs, i), M(iDx, i + 1),
                               W(i + 2, j, k), W(i + 1, j, k), M(Dxs, i + 1),
                               W(i + 1, j, k), W(i, j, k), M(Dxs, i));
                if (utz2 > 0)
                    ww = qface(W(i - 1, j, k), W(i, j, k), M(Dxs, i - 1), M(iDx, i - 1),
                               W(i, j, k), W(i - 1, j, k), M(Dxs, i - 1),
                               W(i - 1, j, k), W(i - 2, j, k), M(Dxs, i - 2));
                else
                    ww = qface(W(i - 1, j, k), W(i, j, k), M(Dxs, i - 1), M(iDx, i),
                               W(i + 1, j, k), W(i, j, k), M(Dxs, i),
                               W(i, j, k), W(i - 1, j, k), M(Dxs, i - 1));
                if (vtz1 > 0)
                    wn = qface(W(i, j, k), W(i, j + 1, k), M(Dys, j), M(iDy, j),
                               W(i, j + 1, k), W(i, j, k), M(Dys, j),
                               W(i, j, k), W(i, j - 1, k), M(Dys, j - 1));
                else
                    wn = qface(W(i, j, k), W(i, j + 1, k), M(Dys, j), M(iDy, j + 1),
                               W(i, j + 2, k), W(i, j + 1, k), M(Dys, j + 1),
                               W(i, j + 1, k), W(i, j, k), M(Dys, j));
                if (vtz2 > 0)
                    wsx = qface(W(i, j - 1, k), W(i, j, k), M(Dys, j - 1), M(iDy, j - 1),
                                W(i, j, k), W(i, j - 1, k), M(Dys, j - 1),
                                W(i, j - 1, k), W(i, j - 2, k), M(Dys, j - 2));
                else
                    wsx = qface(W(i, j - 1, k), W(i, j, k), M(Dys, j - 1), M(iDy, j),
                                W(i, j + 1, k), W(i, j, k), M(Dys, j),
                                W(i, j, k), W(i, j - 1, k), M(Dys, j - 1));
                if (wtz1 > 0)
                    wf = qface(W(i, j, k), W(i, j, k + 1), M(iDz, k + 1), M(Dzs, k),
                               W(i, j, k + 1), W(i, j, k), M(iDz, k + 1),
                               W(i, j, k), W(i, j, k - 1), M(iDz, k));
                else
                    wf = qface(W(i, j, k), W(i, j, k + 1), M(iDz, k + 1), M(Dzs, k + 1),
                               W(i, j, k + 2), W(i, j, k + 1), M(iDz, k + 2),
                               W(i, j, k + 1), W(i, j, k), M(iDz, k + 1));
                if (wtz2 > 0)
                    wb = qface(W(i, j, k - 1), W(i, j, k), M(iDz, k), M(Dzs, k - 1),
                               W(i, j, k), W(i, j, k - 1), M(iDz, k),
                               W(i, j, k - 1), W(i, j, k - 2), M(iDz, k - 1));
                else
                    wb = qface(W(i, j, k - 1), W(i, j, k), M(iDz, k), M(Dzs, k),
                               W(i, j, k + 1), W(i, j, k), M(iDz, k + 1),
                               W(i, j, k), W(i, j, k - 1), M(iDz, k));
                double nut = Viseff[I3(i, j, k)];
                double c = nu + nut * les;
                double w0 = W(i, j, k);
                /* :333-339 */
                ws[H(i, j, k)] = w0 - dt * (utz1 * we - utz2 * ww) / M(iDx, i)
                                 - dt * (vtz1 * wn - vtz2 * wsx) / M(iDy, j)
                                 - dt * (wtz1 * wf - wtz2 * wb) / M(Dzs, k)
                                 + c * dt * ((W(i + 1, j, k) - w0) / M(Dxs, i) - (w0 - W(i - 1, j, k)) / M(Dxs, i - 1)) / M(iDx, i)
                                 + c * dt * ((W(i, j + 1, k) - w0) / M(Dys, j) - (w0 - W(i, j - 1, k)) / M(Dys, j - 1)) / M(iDy, j)
                                 + c * dt * ((W(i, j, k + 1) - w0) / M(iDz, k + 1) - (w0 - W(i, j, k - 1)) / M(iDz, k)) / M(Dzs, k);
            }
}

/* ------------------------------------------------------------------ */
/* discretisation_upwind.f90 (Q1 = 0)                                   */
/* ------------------------------------------------------------------ */

void or_predictor_upwind(const or_grid *g, const double *u, const double *v,
                         const double *w, double dt, double nu,
                         double *us, double *vs, double *ws)
{
    GRID_LOCALS(g);
    const double *iDx = g->iDx, *Dxs = g->Dxs, *iDy = g->iDy, *Dys = g->Dys;
    const double *iDz = g->iDz, *Dzs = g->Dzs;
    /* (nu+nut*Q1) with Q1 = 0 and nut = 0: nu + 0.0 */
    const double c = nu + 0.0 * 0.0;
    int i, j, k;

    /* u_star, :8-93 */
    #pragma omp parallel for private(i, j) schedule(static)
    for (k = 1; k <= nz; k++)
        for (j = 1; j <= ny; j++)
            for (i = 1; i <= nx - 1; i++) {
                double utx1 = 0.5 * (U(i, j, k) + U(i + 1, j, k));
                double utx2 = 0.5 * (U(i - 1, j, k) + U(i, j, k));
                double vtx1 = 0.5 * (V(i, j, k) + V(i + 1, j, k));
                double vtx2 = 0.5 * (V(i, j - 1, k) + V(i + 1, j - 1, k));
                double wtx1 = 0.5 * (W(i + 1, j, k) + W(i, j, k));
                double wtx2 = 0.5 * (W(i, j, k - 1) + W(i + 1, j, k - 1));
                double ue = (utx1 >= 0) ? U(i, j, k) : U(i + 1, j, k);
                double uw = (utx2 >= 0) ? U(i - 1, j, k) : U(i, j, k);
                double un = (vtx1 >= 0) ? U(i, j, k) : U(i, j + 1, k);
                double usx = (vtx2 >= 0) ? U(i, j - 1, k) : U(i, j, k);
                double uf = (wtx1 >= 0) ? U(i, j, k) : U(i, j, k + 1);
                double ub = (wtx2 >= 0) ? U(i, j, k - 1) : U(i, j, k);
                double u0 = U(i, j, k);
                /* :83-89 */
                us[H(i, j, k)] = u0 - dt * (utx1 * ue - utx2 * uw) / M(Dxs, i)
                                 - dt * (vtx1 * un - vtx2 * usx) / M(iDy, j)
                                 - dt * (wtx1 * uf - wtx2 * ub) / M(iDz, k)
                                 + c * dt * ((U(i + 1, j, k) - u0) / M(iDx, i + 1) - (u0 - U(i - 1, j, k)) / M(iDx, i)) / M(Dxs, i)
                                 + c * dt * ((U(i, j + 1, k) - u0) / M(Dys, j) - (u0 - U(i, j - 1, k)) / M(Dys, j - 1)) / M(iDy, j)
                                 + c * dt * ((U(i, j, k + 1) - u0) / M(Dzs, k) - (u0 - U(i, j, k - 1)) / M(Dzs, k - 1)) / M(iDz, k);
            }

    /* v_star, :99-184 */
    #pragma omp parallel for private(i, j) schedule(static)
    for (k = 1; k <= nz; k++)
        for (j = 1; j <= ny - 1; j++)
            for (i = 1; i <= nx; i++) {
                double uty1 = 0.5 * (U(i, j, k) + U(i, j + 1, k));
                double uty2 = 0.5 * (U(i - 1, j, k) + U(i - 1, j + 1, k));
                double vty1 = 0.5 * (V(i, j, k) + V(i, j + 1, k));
                double vty2 = 0.5 * (V(i, j - 1, k) + V(i, j, k));
                double wty1 = 0.5 * (W(i, j, k) + W(i, j + 1, k));
                double wty2 = 0.5 * (W(i, j, k - 1) + W(i, j + 1, k - 1));
                double ve = (uty1 >= 0) ? V(i, j, k) : V(i + 1, j, k);
                double vw = (uty2 >= 0) ? V(i - 1, j, k) : V(i, j, k);
                double vn = (vty1 >= 0) ? V(i, j, k) : V(i, j + 1, k);
                double vsx = (vty2 >= 0) ? V(i, j - 1, k) : V(i, j, k);
                double vf = (wty1 >= 0) ? V(i, j, k) : V(i, j, k + 1);
                double vb = (wty2 >= 0) ? V(i, j, k - 1) : V(i, j, k);
                double v0 = V(i, j, k);
                /* :167-173 */
                vs[H(i, j, k)] = v0 - dt * (uty1 * ve - uty2 * vw) / M(iDx, i)
                                 - dt * (vty1 * vn - vty2 * vsx) / M(Dys, j)
                                 - dt * (wty1 * vf - wty2 * vb) / M(iDz, k)
                                 + c * dt * ((V(i + 1, j, k) - v0) / M(Dxs, i) - (v0 - V(i - 1, j, k)) / M(Dxs, i - 1)) / M(iDx, i)
                                 + c * dt * ((V(i, j + 1, k) - v0) / M(iDy, j + 1) - (v0 - V(i, j - 1, k)) / M(iDy, j)) / M(Dys, j)
                                 + c * dt * ((V(i, j, k + 1) - v0) / M(Dzs, k) - (v0 - V(i, j, k - 1)) / M(Dzs, k - 1)) / M(iDz, k);
            }

    /* w_star, :188-273.  The x-face advecting v of :194 averages
     * v(i,j,k) and v(i+1,j,k+1), as the reference writes it. */
    #pragma omp parallel for private(i, j) schedule(static)
    for (k = 1; k <= nz - 1; k++)
        for (j = 1; j <= ny; j++)
            for (i = 1; i <= nx; i++) {
                double utz1 = 0.5 * (U(i, j, k + 1) + U(i, j, k));
                double utz2 = 0.5 * (U(i - 1, j, k + 1) + U(i - 1, j, k));
                double vtz1 = 0.5 * (V(i, j, k) + V(i + 1, j, k + 1));
                double vtz2 = 0.5 * (V(i, j - 1, k + 1) + V(i, j - 1, k));
                double wtz1 = 0.5 * (W(i, j, k + 1) + W(i, j, k));
                double wtz2 = 0.5 * (W(i, j, k) + W(i, j, k - 1));
                double we = (utz1 >= 0) ? W(i, j, k) : W(i + 1, j, k);
                double ww = (utz2 >= 0) ? W(i - 1, j, k) : W(i, j, k);
                double wn = (vtz1 >= 0) ? W(i, j, k) : W(i, j + 1, k);
                double wsx = (vtz2 >= 0) ? W(i, j - 1, k) : W(i, j, k);
                double wf = (wtz1 >= 0) ? W(i, j, k) : W(i, j, k + 1);
                double wb = (wtz2 >= 0) ? W(i, j, k - 1) : W(i, j, k);
                double w0 = W(i, j, k);
                /* :255-261 */
                ws[H(i, j, k)] = w0 - dt * (utz1 * we - utz2 * ww) / M(iDx, i)
                                 - dt * (vtz1 * wn - vtz2 * wsx) / M(iDy, j)
                                 - dt * (wtz1 * wf - wtz2 * wb) / M(Dzs, k)
                                 + c * dt * ((W(i + 1, j, k) - w0) / M(Dxs, i) - (w0 - W(i - 1, j, k)) / M(Dxs, i - 1)) / M(iDx, i)
                                 + c * dt * ((W(i, j + 1, k) - w0) / M(Dys, j) - (w0 - W(i, j - 1, k)) / M(Dys, j - 1)) / M(iDy, j)
                                 + c * dt * ((W(i, j, k + 1) - w0) / M(iDz, k + 1) - (w0 - W(i, j, k - 1)) / M(iDz, k)) / M(Dzs, k);
            }
#undef U
#undef V
#undef W
}

/* ------------------------------------------------------------------ */
/* GS.f90                                                               */
/* ------------------------------------------------------------------ */

int or_gauss_seidel(const or_grid *g, const double *us, const double *vs,
                    const double *ws, double *p, double *pre, double dt,
                    double omega, double zeta, double itmax, double *pmax)
{
    GRID_LOCALS(g);
    const double *iDx = g->iDx, *Dxs = g->Dxs, *iDy = g->iDy, *Dys = g->Dys;
    const double *iDz = g->iDz, *Dzs = g->Dzs;
    int ik = 0;
    double pChangeMax_ = 1.0;
    /* GS.f90:11 */
    while (pChangeMax_ > zeta && (double)ik < itmax) {
        ik++;
        double pChangeMax = 0.0;
        for (int k = 1; k <= nz; k++)
            for (int j = 1; j <= ny; j++)
                for (int i = 1; i <= nx; i++) {
                    double ay = M(iDy, j), az = M(iDz, k), ax = M(iDx, i);
                    /* :39-41 */
                    double mChange = (us[H(i, j, k)] - us[H(i - 1, j, k)]) * ay * az
                                   + (vs[H(i, j, k)] - vs[H(i, j - 1, k)]) * ax * az
                                   + (ws[H(i, j, k)] - ws[H(i, j, k - 1)]) * ax * ay;
                    /* :43-51 */
                    double pNew = (-p[H(i + 1, j, k)] * ay * az / M(Dxs, i)
                                   - p[H(i - 1, j, k)] * ay * az / M(Dxs, i - 1)
                                   - p[H(i, j + 1, k)] * ax * az / M(Dys, j)
                                   - p[H(i, j - 1, k)] * ax * az / M(Dys, j - 1)
                                   - p[H(i, j, k + 1)] * ax * ay / M(Dzs, k)
                                   - p[H(i, j, k - 1)] * ax * ay / M(Dzs, k - 1)
                                   + mChange / dt)
                                / (-ay * az / M(Dxs, i) - ay * az / M(Dxs, i - 1)
                                   - ax * az / M(Dys, j) - ax * az / M(Dys, j - 1)
                                   - ax * ay / M(Dzs, k) - ax * ay / M(Dzs, k - 1));
                    size_t o = H(i, j, k);
                    /* :53-63 */
                    double pChange = fabs(pNew - p[o]);
                    p[o] = p[o] + (omega * (pNew - p[o]));
                    if (pChange > pChangeMax) pChangeMax = pChange;
                }
        pChangeMax_ = pChangeMax; /* MPI_ALLREDUCE MAX over one rank, :70 */
    }
    /* :86-94 */
    if (pre)
        for (int k = 1; k <= nz; k++)
            for (int j = 1; j <= ny; j++)
                for (int i = 1; i <= nx; i++) pre[I3(i, j, k)] = p[H(i, j, k)];
    if (pmax) *pmax = pChangeMax_;
    return ik;
}

void or_poisson_b(const or_grid *g, const double *us, const double *vs,
                  const double *ws, double dt, double *b)
{
    GRID_LOCALS(g);
    #pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) {
                double ay = M(g->iDy, j), az = M(g->iDz, k), ax = M(g->iDx, i);
                double mChange = (us[H(i, j, k)] - us[H(i - 1, j, k)]) * ay * az
                               + (vs[H(i, j, k)] - vs[H(i, j - 1, k)]) * ax * az
                               + (ws[H(i, j, k)] - ws[H(i, j, k - 1)]) * ax * ay;
                b[I3(i, j, k)] = mChange / dt;
            }
}

int or_sor_redblack(const or_grid *g, const double *b, double *p,
                    double *pre, double omega, double zeta, double itmax,
                    double *pmax)
{
    GRID_LOCALS(g);
    double *rx = malloc(sizeof(double) * (nx + 4));
    double *ry = malloc(sizeof(double) * (ny + 4));
    double *rz = malloc(sizeof(double) * (nz + 4));
    for (int m = 0; m < nx + 4; m++) rx[m] = 1.0 / g->Dxs[m];
    for (int m = 0; m < ny + 4; m++) ry[m] = 1.0 / g->Dys[m];
    for (int m = 0; m < nz + 4; m++) rz[m] = 1.0 / g->Dzs[m];
    int ik = 0;
    double pChangeMax_ = 1.0;
    while (pChangeMax_ > zeta && (double)ik < itmax) {
        ik++;
        double pChangeMax = 0.0;
        for (int colour = 0; colour < 2; colour++)
            #pragma omp parallel for reduction(max : pChangeMax) schedule(static)
            for (int k = 1; k <= nz; k++)
                for (int j = 1; j <= ny; j++)
                    for (int i = 1; i <= nx; i++) {
                        if (((i + j + k) & 1) != colour) continue;
                        double ax = M(g->iDx, i), ay = M(g->iDy, j), az = M(g->iDz, k);
                        double ayz = ay * az, axz = ax * az, axy = ax * ay;
                        double rE = M(rx, i), rW = M(rx, i - 1), rN = M(ry, j), rS = M(ry, j - 1);
                        double rF = M(rz, k), rB = M(rz, k - 1);
                        double t1 = fma(rE, p[H(i + 1, j, k)], rW * p[H(i - 1, j, k)]);
                        double t2 = fma(rN, p[H(i, j + 1, k)], rS * p[H(i, j - 1, k)]);
                        double t3 = fma(rF, p[H(i, j, k + 1)], rB * p[H(i, j, k - 1)]);
                        double s = fma(ayz, t1, fma(axz, t2, axy * t3));
                        double aP = fma(ayz, rE + rW, fma(axz, rN + rS, axy * (rF + rB)));
                        double pNew = (s - b[I3(i, j, k)]) / aP;
                        size_t o = H(i, j, k);
                        double pChange = fabs(pNew - p[o]);
                        p[o] = fma(omega, pNew - p[o], p[o]);
                        if (pChange > pChangeMax) pChangeMax = pChange;
                    }
        pChangeMax_ = pChangeMax;
    }
    if (pre)
        #pragma omp parallel for schedule(static)
        for (int k = 1; k <= nz; k++)
            for (int j = 1; j <= ny; j++)
                for (int i = 1; i <= nx; i++) pre[I3(i, j, k)] = p[H(i, j, k)];
    if (pmax) *pmax = pChangeMax_;
    free(rx);
    free(ry);
    free(rz);
    return ik;
}

/* ------------------------------------------------------------------ */
/* calcul_new_velocity.f90                                              */
/* ------------------------------------------------------------------ */

void or_calcul_new_velocity(const or_grid *g, const double *us,
                            const double *vs, const double *ws,
                            const double *p, const double *ETA,
                            double u_solid, double v_solid, double w_solid,
                            double dt, int kstart, int kend,
                            double *u1, double *v1, double *w1,
                            double *u2, double *v2, double *w2,
                            double *FX, double *FY, double *FZ,
                            double *uc, double *vc, double *wc)
{
    GRID_LOCALS(g);
    int i, j, k;
    /* :13-49 */
    #pragma omp parallel for private(i, j) schedule(static)
    for (k = kstart; k <= kend; k++)
        for (j = 1; j <= ny; j++)
            for (i = 1; i <= nx; i++) {
                size_t o = H(i, j, k);
                u1[o] = us[o] - dt * (p[H(i + 1, j, k)] - p[o]) / M(g->Dxs, i);
                v1[o] = vs[o] - dt * (p[H(i, j + 1, k)] - p[o]) / M(g->Dys, j);
                w1[o] = ws[o] - dt * (p[H(i, j, k + 1)] - p[o]) / M(g->Dzs, k);
            }
    /* :61-100 */
    #pragma omp parallel for private(i, j) schedule(static)
    for (k = kstart; k <= kend; k++)
        for (j = 1; j <= ny; j++)
            for (i = 1; i <= nx; i++) {
                size_t o = H(i, j, k);
                double e = ETA[o];
                u2[o] = e * u_solid + (1.0 - e) * u1[o];
                FX[o] = (u2[o] - u1[o]) / dt;
                v2[o] = e * v_solid + (1.0 - e) * v1[o];
                FY[o] = (v2[o] - v1[o]) / dt;
                w2[o] = e * w_solid + (1.0 - e) * w1[o];
                FZ[o] = (w2[o] - w1[o]) / dt;
            }
    /* :171-181 */
    #pragma omp parallel for private(i, j) schedule(static)
    for (k = kstart; k <= kend; k++)
        for (j = 1; j <= ny; j++)
            for (i = 1; i <= nx; i++) {
                size_t o = I3(i, j, k);
                uc[o] = 0.5 * (u2[H(i, j, k)] + u2[H(i - 1, j, k)]);
                vc[o] = 0.5 * (v2[H(i, j, k)] + v2[H(i, j - 1, k)]);
                wc[o] = 0.5 * (w2[H(i, j, k)] + w2[H(i, j, k - 1)]);
            }
}

void or_updating_velocity(const or_grid *g, const double *u2,
                          const double *v2, const double *w2, int kstart,
                          int kend, double *u, double *v, double *w)
{
    GRID_LOCALS(g);
    #pragma omp parallel for schedule(static)
    for (int k = kstart; k <= kend; k++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) {
                size_t o = H(i, j, k);
                u[o] = u2[o];
                v[o] = v2[o];
                w[o] = w2[o];
            }
}

/* ------------------------------------------------------------------ */
/* virtualForceIntegrator.f90                                           */
/* ------------------------------------------------------------------ */

void or_virtual_force_integrator(const or_grid *g, const double *FY,
                                 const double *FZ, double lx, double *cDrag,
                                 double *cLift)
{
    GRID_LOCALS(g);
    /* FXz/FYz with one zero row on either side in j; FXy/FYy with one zero
     * entry on either side in i (the out-of-bounds reads of :49,53,68,72) */
    const int pj = ny + 2;
    double *FXz = calloc((size_t)(nx + 2) * pj, sizeof(double));
    double *FYz = calloc((size_t)(nx + 2) * pj, sizeof(double));
    double *FXy = calloc((size_t)nx + 2, sizeof(double));
    double *FYy = calloc((size_t)nx + 2, sizeof(double));
#define ZI(i, j) ((size_t)(i) + (size_t)(nx + 2) * (size_t)(j))
    const double *Dzs = g->Dzs, *Dys = g->Dys, *Dxs = g->Dxs;
    /* z pass, :22-36 */
    for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) {
            double fx = 0.0, fy = 0.0;
            for (int k = 1; k <= nz; k++) {
                fx = fx + (FZ[H(i, j, k - 1)] * M(Dzs, k - 1) + 4.0 * FZ[H(i, j, k)] * M(Dzs, k)
                           + FZ[H(i, j, k + 1)] * M(Dzs, k + 1)) / 6.0;
                fy = fy + (FY[H(i, j, k - 1)] * M(Dzs, k - 1) + 4.0 * FY[H(i, j, k)] * M(Dzs, k)
                           + FY[H(i, j, k + 1)] * M(Dzs, k + 1)) / 6.0;
            }
            FXz[ZI(i, j)] = fx;
            FYz[ZI(i, j)] = fy;
        }
    /* y pass, :41-58 */
    for (int i = 1; i <= nx; i++) {
        double fx = 0.0, fy = 0.0;
        for (int j = 1; j <= ny; j++) {
            if (FXz[ZI(i, j)] != 0.0)
                fx = fx + (FXz[ZI(i, j - 1)] * M(Dys, j - 1) + 4.0 * FXz[ZI(i, j)] * M(Dys, j)
                           + FXz[ZI(i, j + 1)] * M(Dys, j + 1)) / 6.0;
            if (FYz[ZI(i, j)] != 0.0)
                fy = fy + (FYz[ZI(i, j - 1)] * M(Dys, j - 1) + 4.0 * FYz[ZI(i, j)] * M(Dys, j)
                           + FYz[ZI(i, j + 1)] * M(Dys, j + 1)) / 6.0;
        }
        FXy[i] = fx;
        FYy[i] = fy;
    }
    /* x pass, :63-75 */
    double totalFX = 0.0, totalFY = 0.0;
    for (int i = 1; i <= nx; i++) {
        if (FXy[i] != 0.0)
            totalFX = totalFX + (FXy[i - 1] * M(Dxs, i - 1) + 4.0 * FXy[i] * M(Dxs, i)
                                 + FXy[i + 1] * M(Dxs, i + 1)) / 6.0;
        if (FYy[i] != 0.0)
            totalFY = totalFY + (FYy[i - 1] * M(Dxs, i - 1) + 4.0 * FYy[i] * M(Dxs, i)
                                 + FYy[i + 1] * M(Dxs, i + 1)) / 6.0;
    }
    /* :80-81 */
    *cDrag = (-2.0) * totalFX / lx;
    *cLift = (-2.0) * totalFY / lx;
#undef ZI
    free(FXz);
    free(FYz);
    free(FXy);
    free(FYy);
}

/* ------------------------------------------------------------------ */
/* solver_1st_member.f90                                                */
/* ------------------------------------------------------------------ */

void or_poisson_coeffs(const or_grid *g, double *coef, int *jcoef)
{
    GRID_LOCALS(g);
    const size_t nd = (size_t)nx * ny * nz;
    const double *iDx = g->iDx, *Dxs = g->Dxs, *iDy = g->iDy, *Dys = g->Dys;
    const double *iDz = g->iDz, *Dzs = g->Dzs;
    size_t ik = 0;
    /* :18-29; 1/a/b is (1/a)/b */
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++, ik++) {
                double cE = 1.0 / M(iDx, i) / M(Dxs, i), cW = 1.0 / M(iDx, i) / M(Dxs, i - 1);
                double cN = 1.0 / M(iDy, j) / M(Dys, j), cS = 1.0 / M(iDy, j) / M(Dys, j - 1);
                double cF = 1.0 / M(iDz, k) / M(Dzs, k), cB = 1.0 / M(iDz, k) / M(Dzs, k - 1);
                coef[ik] = cE + cW + cN + cS + cF + cB;
                coef[ik + nd] = -1.0 / M(iDx, i) / M(Dxs, i);
                coef[ik + 2 * nd] = -1.0 / M(iDx, i) / M(Dxs, i - 1);
                coef[ik + 3 * nd] = -1.0 / M(iDy, j) / M(Dys, j);
                coef[ik + 4 * nd] = -1.0 / M(iDy, j) / M(Dys, j - 1);
                coef[ik + 5 * nd] = -1.0 / M(iDz, k) / M(Dzs, k);
                coef[ik + 6 * nd] = -1.0 / M(iDz, k) / M(Dzs, k - 1);
            }
    /* :33-69 boundary rows */
    ik = 0;
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++, ik++) {
                double aa = 0.0, bb = 0.0, cc = 0.0;
                if (i == 1) coef[ik + 2 * nd] = 0.0;
                else if (i == nx) { aa = -1.0 / M(iDx, i) / M(Dxs, i); coef[ik + nd] = 0.0; }
                if (j == 1) coef[ik + 4 * nd] = 0.0;
                else if (j == ny) coef[ik + 3 * nd] = 0.0;
                if (k == 1) coef[ik + 6 * nd] = 0.0;
                else if (k == nz) coef[ik + 5 * nd] = 0.0;
                coef[ik] = coef[ik] + aa + bb + cc;
            }
    /* :71-74 */
    jcoef[0] = 0;
    jcoef[1] = 1;
    jcoef[2] = nx;
    jcoef[3] = nx * ny;
}

/* ------------------------------------------------------------------ */
/* main.f90:139-259, one time step                                      */
/* ------------------------------------------------------------------ */

int or_step(const or_grid *g, const or_params *prm, or_fields *f,
            double *cd, double *cl, double *pmax)
{
    const int nz = g->nz;
    /* main.f90:209 */
    or_smagorinsky(g, f->u, f->v, f->w, NULL, NULL, NULL, prm->Cs, 1, nz,
                   NULL, NULL, NULL, f->Viseff);
    /* main.f90:213 */
    if (prm->scheme == 0)
        or_predictor_quick(g, f->u, f->v, f->w, f->Viseff, prm->dt, prm->nu,
                           prm->LES, 1, nz, f->us, f->vs, f->ws);
    else
        or_predictor_upwind(g, f->u, f->v, f->w, prm->dt, prm->nu,
                            f->us, f->vs, f->ws);
    if (f->fu) {
        const long sx = g->nx + 4, sxy = sx * (g->ny + 4);
        #pragma omp parallel for schedule(static)
        for (int k = 1; k <= nz; k++)
            for (int j = 1; j <= g->ny; j++)
                for (int i = 1; i <= g->nx; i++) {
                    const long o = (i + 1) + sx * (j + 1) + sxy * (k + 1);
                    f->us[o] = f->us[o] + prm->dt * f->fu[o];
                    f->vs[o] = f->vs[o] + prm->dt * f->fv[o];
                    f->ws[o] = f->ws[o] + prm->dt * f->fw[o];
                }
    }
    /* main.f90:231 */
    int ik;
    if (prm->sor == 0) {
        ik = or_gauss_seidel(g, f->us, f->vs, f->ws, f->p, f->pre, prm->dt,
                             prm->omega, prm->zeta, prm->itmax, pmax);
    } else {
        size_t n3 = (size_t)g->nx * g->ny * g->nz;
        double *b = malloc(n3 * sizeof(double));
        or_poisson_b(g, f->us, f->vs, f->ws, prm->dt, b);
        ik = or_sor_redblack(g, b, f->p, f->pre, prm->omega, prm->zeta,
                             prm->itmax, pmax);
        free(b);
    }
    /* main.f90:247 */
    or_calcul_new_velocity(g, f->us, f->vs, f->ws, f->p, f->ETA, 0.0, 0.0, 0.0,
                           prm->dt, 1, nz, f->u1, f->v1, f->w1, f->u2, f->v2,
                           f->w2, f->FX, f->FY, f->FZ, f->uc, f->vc, f->wc);
    /* Lagrangian-marker forcing of the library, when markers are set (no
     * counterpart in the reference; ntust_markers.c) */
    if (prm->markers && or_marker_count() > 0)
        or_marker_forcing(g, prm->dt, f->u2, f->v2, f->w2, f->FY, f->FZ, f->uc, f->vc, f->wc);
    /* main.f90:251 */
    or_virtual_force_integrator(g, f->FY, f->FZ, prm->lx, cd, cl);
    /* main.f90:255 */
    or_updating_velocity(g, f->u2, f->v2, f->w2, 1, nz, f->u, f->v, f->w);
    /* main.f90:259 */
    or_boundary_conditions(g, prm->bc, f->u, f->v, f->w, f->p, f->u1, f->v1,
                           f->w1, f->u2, f->v2, f->w2, f->us, f->vs, f->ws);
    return ik;
}

/* ------------------------------------------------------------------ */
/* output, steady check and the .q restart reader                      */
/* ------------------------------------------------------------------ */

/* filer.f90:163-171: Qout(:,:,:,1..5) = pre, uc, vc, wc, ETA, each value
 * assigned to a default real (round to nearest even). */
void or_pack_q(const or_grid *g, const double *pre, const double *uc,
               const double *vc, const double *wc, const double *ETA,
               float *qout)
{
    GRID_LOCALS(g);
    const size_t ni = (size_t)nx * ny * nz;
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) {
                const size_t c = I3(i, j, k);
                qout[c] = (float)pre[c];
                qout[c + ni] = (float)uc[c];
                qout[c + 2 * ni] = (float)vc[c];
                qout[c + 3 * ni] = (float)wc[c];
                qout[c + 4 * ni] = (float)ETA[H(i, j, k)];
            }
}

/* One unformatted sequential record: 4-byte length, data, 4-byte length. */
static int put_record(FILE *f, const void *data, size_t len)
{
    if (len > 0x7fffffffu) return -1;
    const int32_t n = (int32_t)len;
    if (fwrite(&n, 4, 1, f) != 1) return -1;
    if (len && fwrite(data, 1, len, f) != len) return -1;
    if (fwrite(&n, 4, 1, f) != 1) return -1;
    return 0;
}

/* filer.f90:173-182: appended records nblocks; nx ny nz; temp x4; Qout */
int or_write_q(const char *path, int nx, int ny, int nz, const float *qout)
{
    FILE *f = fopen(path, "ab");
    if (!f) return -1;
    const int32_t nb = 1, dims[3] = {nx, ny, nz};
    const float temp[4] = {1.0f, 1.0f, 1.0f, 1.0f};
    int rc = put_record(f, &nb, 4);
    if (!rc) rc = put_record(f, dims, 12);
    if (!rc) rc = put_record(f, temp, 16);
    if (!rc) rc = put_record(f, qout, (size_t)5 * nx * ny * nz * sizeof(float));
    if (fclose(f)) rc = -1;
    return rc;
}

/* gridder.f90:355-378: Xout(i,j,k) = Xs(i) (Y, Z alike) as default reals,
 * written as nblocks; nx ny nz; one record with X, Y, Z.  Xs(1:nx). */
int or_write_mesh_x(const char *path, int nx, int ny, int nz, const double *Xs,
                    const double *Ys, const double *Zs)
{
    const size_t ni = (size_t)nx * ny * nz;
    float *xyz = malloc(3 * ni * sizeof(float));
    if (!xyz) return -1;
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) {
                const size_t c = I3(i, j, k);
                xyz[c] = (float)Xs[i - 1];
                xyz[c + ni] = (float)Ys[j - 1];
                xyz[c + 2 * ni] = (float)Zs[k - 1];
            }
    FILE *f = fopen(path, "wb");
    int rc = f ? 0 : -1;
    const int32_t nb = 1, dims[3] = {nx, ny, nz};
    if (!rc) rc = put_record(f, &nb, 4);
    if (!rc) rc = put_record(f, dims, 12);
    if (!rc) rc = put_record(f, xyz, 3 * ni * sizeof(float));
    if (f && fclose(f)) rc = -1;
    free(xyz);
    return rc;
}

/* check_steady.f90:1-30.  When istep /= isto, *vdiff = max |wc - last|
 * (a NaN difference never compares greater, so it never wins); otherwise
 * *vdiff is left as it is.  Then last = wc.  last: interior array. */
void or_check_steady(const or_grid *g, int istep, int isto, const double *wc,
                     double *last, double *vdiff)
{
    GRID_LOCALS(g);
    const size_t ni = (size_t)nx * ny * nz;
    if (istep != isto) {
        double m = 0.0;
        for (size_t c = 0; c < ni; c++)
            if (fabs(wc[c] - last[c]) > m) m = fabs(wc[c] - last[c]);
        *vdiff = m;
    }
    memcpy(last, wc, ni * sizeof(double));
}

/* reading_data.f90:47-83 reading_variables, from the Qout array of a .q
 * file.  u, v, w take 0.5*(Q(i)+Q(i+1)) evaluated in single precision over
 * i, j, k = 1..n-1 for all three components; p (interior) and pre take Q1;
 * uc, vc, wc = 0.5*(u(i)+u(i-1)) in double precision.  Entries the loops
 * do not reach keep their values. */
void or_reading_variables(const or_grid *g, const float *qout, double *u,
                          double *v, double *w, double *p, double *pre,
                          double *uc, double *vc, double *wc)
{
    GRID_LOCALS(g);
    const size_t ni = (size_t)nx * ny * nz;
    const float *q1 = qout, *q2 = qout + ni, *q3 = qout + 2 * ni, *q4 = qout + 3 * ni;
    for (int k = 1; k <= nz - 1; k++)
        for (int j = 1; j <= ny - 1; j++)
            for (int i = 1; i <= nx - 1; i++) {
                const float a = 0.5f * (q2[I3(i, j, k)] + q2[I3(i + 1, j, k)]);
                const float b = 0.5f * (q3[I3(i, j, k)] + q3[I3(i, j + 1, k)]);
                const float c = 0.5f * (q4[I3(i, j, k)] + q4[I3(i, j, k + 1)]);
                u[H(i, j, k)] = (double)a;
                v[H(i, j, k)] = (double)b;
                w[H(i, j, k)] = (double)c;
            }
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) {
                p[H(i, j, k)] = (double)q1[I3(i, j, k)];
                pre[I3(i, j, k)] = p[H(i, j, k)];
            }
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) {
                uc[I3(i, j, k)] = 0.5 * (u[H(i, j, k)] + u[H(i - 1, j, k)]);
                vc[I3(i, j, k)] = 0.5 * (v[H(i, j, k)] + v[H(i, j - 1, k)]);
                wc[I3(i, j, k)] = 0.5 * (w[H(i, j, k)] + w[H(i, j, k - 1)]);
            }
}

/* ------------------------------------------------------------------ */
/* airfoil.f90: RayCasting                                             */
/* ------------------------------------------------------------------ */

/* airfoil.f90:597-611.  The rotation is not rigid: ay is rotated with the
 * already rotated az (:604-605). */
void or_airfoil_transform(int poly, const double *az0, const double *ay0,
                          double AOA, double zc, double yc, double *az,
                          double *ay)
{
    const double PI = 4.0 * atan(1.0);   /* :594 */
    for (int i = 0; i < poly; i++) {
        double z = az0[i], y = ay0[i];
        z = z - 0.25;
        z = cos((-AOA) * PI / 180) * z - sin((-AOA) * PI / 180) * y;
        y = sin((-AOA) * PI / 180) * z + cos((-AOA) * PI / 180) * y;
        z = z + zc;
        y = y + yc;
        z = z + 0.25;
        az[i] = z;
        ay[i] = y;
    }
}

/* One edge of the crossing-number test (:695-760 for cells, :788-860 for
 * sub-cells): the edge from vertex m to m+1 (0-based here) against the
 * row through yrow and the point at zcen, with (Yn, Zn) the node the
 * slope test uses (the cell's or sub-cell's lower corner, not its
 * centre).  strict: the sub-cell form (zcen > max) instead of >=.  The
 * "nudge" of :701-702 adds 0.d01 = 0 and is left out. */
static int ray_edge(const double *az, const double *ay, int m, double yrow,
                    double zcen, double Yn, double Zn, int strict)
{
    int c = 0;
    for (int blk = 0; blk < 2; blk++) {
        int P, Q;
        if (blk == 0) {
            if (!(ay[m] >= yrow && ay[m + 1] <= yrow)) continue;
            P = m + 1;
            Q = m;
        } else {
            if (!(ay[m] <= yrow && ay[m + 1] >= yrow)) continue;
            P = m;
            Q = m + 1;
        }
        const double zmax = az[m] > az[m + 1] ? az[m] : az[m + 1];
        const double zmin = az[m] < az[m + 1] ? az[m] : az[m + 1];
        if (strict ? zcen > zmax : zcen >= zmax) {
            /* + 0 */
        } else if (zcen < zmin) {
            c += 1;
        } else {
            const double m_pa = (Yn - ay[P]) / (Zn - az[P]);
            const double m_ab = (ay[Q] - ay[P]) / (az[Q] - az[P]);
            if (m_pa > m_ab) c += 1;
        }
    }
    return c;
}

int or_ray_casting(const or_grid *g, int poly, const double *az,
                   const double *ay, const double *Yn, const double *Zn,
                   int nsub, double *ETA, double *eta1_out, int *points_out)
{
    GRID_LOCALS(g);
#define YN(j) Yn[(j) + 1]
#define ZN(k) Zn[(k) + 1]
    double ymin = ay[0], ymax = ay[0], zmin = az[0], zmax = az[0];
    for (int i = 1; i < poly; i++) {           /* minval / maxval */
        if (ay[i] < ymin) ymin = ay[i];
        if (ay[i] > ymax) ymax = ay[i];
        if (az[i] < zmin) zmin = az[i];
        if (az[i] > zmax) zmax = az[i];
    }
    /* :614-654; a bound the search does not find keeps the module
     * variable's value, 0 on the first call */
    int A = 0, B = 0, C = 0, E = 0;
    for (int j = 1; j <= ny; j++) if (YN(j) > ymin) { A = j - 1; break; }
    for (int j = 1; j <= ny; j++) if (YN(j) > ymax) { B = j; break; }
    for (int k = 1; k <= nz; k++) if (ZN(k) > zmin) { C = k - 1; break; }
    for (int k = 1; k <= nz; k++) if (ZN(k) > zmax) { E = k; break; }
    /* with both ranges non-empty the loops below touch points(C..E-1,
     * A..B-1), outside the array when A or C is 0 */
    const int boxed = E - 1 >= C && B - 1 >= A;
    if (boxed && (A < 1 || C < 1)) return -1;

    const size_t n2 = (size_t)nz * ny;
    int *points = calloc(n2, sizeof(int)), *inter = calloc(n2, sizeof(int));
    double *eta1 = malloc(n2 * sizeof(double));
#define P2(k, j) ((size_t)((k) - 1) + (size_t)nz * (size_t)((j) - 1))
    for (int m = 0; m < poly && boxed; m++)    /* :659-673 */
        for (int k = C; k <= E - 1; k++)
            if (az[m] >= ZN(k) && az[m] < ZN(k + 1))
                for (int j = A; j <= B - 1; j++)
                    if (ay[m] >= YN(j) && ay[m] < YN(j + 1)) points[P2(k, j)] = 1;

    for (int m = 0; m < poly - 1; m++)         /* :684-766 */
        for (int j = 1; j <= ny; j++) {
            const double yrow = (YN(j) + YN(j + 1)) / 2.0;
            if (!((ay[m] >= yrow && ay[m + 1] <= yrow) || (ay[m] <= yrow && ay[m + 1] >= yrow)))
                continue;
            for (int k = 1; k <= nz; k++)
                inter[P2(k, j)] += ray_edge(az, ay, m, yrow, (ZN(k) + ZN(k + 1)) / 2.0,
                                            YN(j), ZN(k), 0);
        }
    for (size_t c = 0; c < n2; c++) eta1[c] = (inter[c] % 2 == 0) ? 0.0 : 1.0;

    double *ZZ = malloc((nsub + 1) * sizeof(double)), *YY = malloc((nsub + 1) * sizeof(double));
    for (int k = C; k <= E - 1 && boxed; k++)  /* :776-875 */
        for (int j = A; j <= B - 1; j++) {
            if (points[P2(k, j)] <= 0) continue;
            for (int l = 1; l <= nsub + 1; l++) {
                ZZ[l - 1] = ZN(k) + (double)(l - 1) * (ZN(k + 1) - ZN(k)) / (double)nsub;
                YY[l - 1] = YN(j) + (double)(l - 1) * (YN(j + 1) - YN(j)) / (double)nsub;
            }
            long total = 0;
            for (int l = 0; l < nsub; l++) {
                const double yrow = (YY[l] + YY[l + 1]) / 2.0;
                for (int n = 0; n < nsub; n++) {
                    const double zcen = (ZZ[n] + ZZ[n + 1]) / 2.0;
                    int cnt = 0;
                    for (int m = 0; m < poly - 1; m++)
                        cnt += ray_edge(az, ay, m, yrow, zcen, YY[l], ZZ[n], 1);
                    total += cnt % 2;
                }
            }
            eta1[P2(k, j)] = (double)total / (double)nsub / (double)nsub;
        }
    for (int k = 1; k <= nz; k++)              /* :879-883 */
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) ETA[H(i, j, k)] = eta1[P2(k, j)];
    if (eta1_out) memcpy(eta1_out, eta1, n2 * sizeof(double));
    if (points_out) memcpy(points_out, points, n2 * sizeof(int));
    free(ZZ);
    free(YY);
    free(points);
    free(inter);
    free(eta1);
    return 0;
#undef P2
#undef YN
#undef ZN
}

/* airfoil.f90:311-321 dynamic_AOA; PI is the default-real literal widened */
double or_dynamic_aoa(double AOA1, double time, double dt, double StartDynamic)
{
    const double PI = (double)3.14159265359f, reduce_frequency = 2.0;
    const double frequency = reduce_frequency / PI;
    return AOA1 + 10 * fabs(sin(2 * PI * frequency * (time - dt * (StartDynamic - 1))));
}

/* ------------------------------------------------------------------ */
/* solver.f90: BiCGSTAB on the DIA matrix                              */
/* ------------------------------------------------------------------ */

/* solver.f90:146-190 matmul_ell: x = coef(:,1) y, then for j = 2..m the
 * upper diagonal coef(:,j+a) y(i+col) and the lower coef(:,j+b) y(i-col),
 * col = jcoef(j), (a,b) = (0,1), (1,2), (2,3) for j = 2, 3, 4: per row the
 * terms are added in the order diag, +1, -1, +nx, -nx, +nxny, -nxny. */
void or_matmul_ell(double *x, const double *coef, const int *jcoef,
                   const double *y, int n, int m)
{
    for (int i = 0; i < n; i++) x[i] = coef[i] * y[i];
    int a = 0, b = 0;
    for (int j = 2; j <= m; j++) {
        const int col = jcoef[j - 1];
        if (j == 2) { a = 0; b = 1; }
        else if (j == 3) { a = 1; b = 2; }
        else if (j == 4) { a = 2; b = 3; }
        const double *up = coef + (size_t)(j + a - 1) * n, *lo = coef + (size_t)(j + b - 1) * n;
        for (int i = 0; i < n - col; i++) x[i] = x[i] + up[i] * y[i + col];
        for (int i = col; i < n; i++) x[i] = x[i] + lo[i] * y[i - col];
    }
}

static double or_scal(const double *x, const double *y, int n)
{
    double s = 0.0;
    for (int i = 0; i < n; i++) s = s + x[i] * y[i];
    return s;
}

/* solver.f90:1-144 solver_BICG3 (BiCGSTAB, no preconditioner), serial:
 * the reductions sum left to right.  x: initial guess in, solution out.
 * Returns the iteration count j; *norm gets ||r|| / ||b||. */
int or_solver_bicg3(const double *coef, const int *jcoef, const double *b,
                    double *x, int n, int m, double eps, int itmax, double *normout)
{
    double *p = calloc(n, sizeof(double)), *r = malloc(n * sizeof(double));
    double *r2 = malloc(n * sizeof(double)), *v = calloc(n, sizeof(double));
    double *ss = malloc(n * sizeof(double)), *t = malloc(n * sizeof(double));
    const double norm0 = sqrt(or_scal(b, b, n));
    or_matmul_ell(p, coef, jcoef, x, n, m);
    for (int i = 0; i < n; i++) r[i] = b[i] - p[i];
    for (int i = 0; i < n; i++) r2[i] = r[i];
    double rho1 = 1, alpha = 1, omega = 1, rho2, beta;
    for (int i = 0; i < n; i++) { v[i] = 0; p[i] = 0; }
    double norm = sqrt(or_scal(r, r, n)) / norm0;
    int j = 0;
    while (norm > eps && j < itmax) {
        j = j + 1;
        rho2 = or_scal(r2, r, n);
        beta = (rho2 / rho1) * (alpha / omega);
        for (int i = 0; i < n; i++) p[i] = r[i] + beta * (p[i] - omega * v[i]);
        or_matmul_ell(v, coef, jcoef, p, n, m);
        alpha = rho2 / or_scal(r2, v, n);
        for (int i = 0; i < n; i++) ss[i] = r[i] - alpha * v[i];
        or_matmul_ell(t, coef, jcoef, ss, n, m);
        omega = or_scal(t, ss, n) / or_scal(t, t, n);
        for (int i = 0; i < n; i++) x[i] = x[i] + alpha * p[i] + omega * ss[i];
        for (int i = 0; i < n; i++) r[i] = ss[i] - omega * t[i];
        rho1 = rho2;
        norm = sqrt(or_scal(r, r, n)) / norm0;
    }
    if (normout) *normout = norm;
    free(p); free(r); free(r2); free(v); free(ss); free(t);
    return j;
}
