// k_adjoint.hip -- discrete adjoint of the device-resident time step (gfx950).
//
// The step (main.f90:209-259, as the resident driver runs it) maps the
// post-boundary-condition state (u, v, w, p) to the next one:
//   nu_t = Smag(u)                          smagorinsky.f90:1-87
//   u*   = Pred(u, nu_t) + dt f             discretisation_QUICK_centre.f90,
//                                           discretisation_upwind.f90; f is a
//                                           body force (the control), 0 in the
//                                           reference
//   p    = k red-black SOR sweeps from p    GS.f90:39-63 reordered; b = D u*/dt,
//          (ghosts fixed at their values)   the ghosts are the previous step's
//   u    = (1-ETA) (u* - dt G p)            calcul_new_velocity.f90:13-100,195-243
//   BC(u, v, w, p)                          boundary_conditions.f90:1-376
// The transpose of each stage is applied in reverse order.  Every stage is
// linear except the predictor (bilinear in u once the upwind choices of
// QUICK / upwind are frozen at the forward state) and nu_t (differentiated
// where the strain rate is non-zero).  The pressure stage is transposed
// sweep by sweep with the forward's sweep count, so the result is the exact
// gradient of the truncated solve the forward ran, converged or not.
//
// Adjoint arrays have the layout of the arrays they belong to.  The
// transposed stencils of the predictor and of the eddy viscosity are
// scatters with FP64 atomic adds (each output cell adds to every input it
// read), so the adjoint is reproducible only to rounding (summation order).
#include "ntust_kernels.h"

namespace ntust {

__device__ __forceinline__ void aadd(double *a, double v)
{
    if (v != 0.0) unsafeAtomicAdd(a, v);
}

// ---------------------------------------------------------------------
// boundary conditions, transposed (faces in reverse order F, B, N, S, E, W)
// ---------------------------------------------------------------------

// derivative of the ghost rules with respect to the interior value they read
__device__ __forceinline__ double dtangent(int code)
{
    if (code == BC_NOSLIP) return -1.0;
    if (code == BC_NEUMANN) return 1.0;
    if (code == BC_DIRICHLET) return -1.0;
    return 0.0;
}
__device__ __forceinline__ double dnormal(int code)
{
    return code == BC_NEUMANN ? 1.0 : 0.0;
}
__device__ __forceinline__ bool has_rule(int code)
{
    return code == BC_NOSLIP || code == BC_NEUMANN || code == BC_DIRICHLET;
}

// x_dst = alpha * x_src (+ const), transposed: l_src += alpha l_dst, l_dst = 0
__device__ __forceinline__ void tassign(double *l, long dst, long src, double alpha)
{
    const double ld = l[dst];
    l[dst] = 0.0;
    l[src] += alpha * ld;
}

template <int FACE>
__global__ void __launch_bounds__(256)
k_adj_bc_face(KGrid g, BcTable bc, double *lu, double *lv, double *lw, double *lp)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x - 1;
    const int b = blockIdx.y * blockDim.y + threadIdx.y - 1;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const int *c = bc.c + 3 * FACE;
    // the forward face, written as assignments in order; here reversed
    if (FACE == 0 || FACE == 1) {
        const int j = a, k = b;
        if (j > ny + 2 || k > nz + 2) return;
        if (FACE == 0) {
            const long o = hix(g, 0, j, k);
            if (lp) { tassign(lp, o - 1, o, 1.0); tassign(lp, o, o + 1, 1.0); }
            tassign(lw, o - 1, o, 1.0);
            if (has_rule(c[2])) tassign(lw, o, o + 1, dtangent(c[2]));
            tassign(lv, o - 1, o, 1.0);
            if (has_rule(c[1])) tassign(lv, o, o + 1, dtangent(c[1]));
            tassign(lu, o - 1, o, 1.0);
            if (has_rule(c[0])) tassign(lu, o, o + 1, dnormal(c[0]));
        } else {
            const long o = hix(g, nx, j, k), t = o + 1;
            if (lp) { tassign(lp, t + 1, t, 1.0); tassign(lp, t, t - 1, 1.0); }
            tassign(lw, t + 1, t, 1.0);
            if (has_rule(c[2])) tassign(lw, t, t - 1, dtangent(c[2]));
            tassign(lv, t + 1, t, 1.0);
            if (has_rule(c[1])) tassign(lv, t, t - 1, dtangent(c[1]));
            tassign(lu, o + 2, o + 1, 1.0);
            tassign(lu, o + 1, o, 1.0);
            if (has_rule(c[0])) tassign(lu, o, o - 1, dnormal(c[0]));
        }
    } else if (FACE == 2 || FACE == 3) {
        const int i = a, k = b;
        if (i > nx + 2 || k > nz + 2) return;
        const long s = g.sx;
        if (FACE == 2) {
            const long o = hix(g, i, 0, k);
            if (lp) { tassign(lp, o - s, o, 1.0); tassign(lp, o, o + s, 1.0); }
            tassign(lw, o - s, o, 1.0);
            if (has_rule(c[2])) tassign(lw, o, o + s, dtangent(c[2]));
            tassign(lv, o - s, o, 1.0);
            if (has_rule(c[1])) tassign(lv, o, o + s, dnormal(c[1]));
            tassign(lu, o - s, o, 1.0);
            if (has_rule(c[0])) tassign(lu, o, o + s, dtangent(c[0]));
        } else {
            const long t = hix(g, i, ny + 1, k), o = t - s;
            if (lp) { tassign(lp, t + s, t, 1.0); tassign(lp, t, t - s, 1.0); }
            tassign(lw, t + s, t, 1.0);
            if (has_rule(c[2])) tassign(lw, t, t - s, dtangent(c[2]));
            tassign(lv, o + 2 * s, o + s, 1.0);
            tassign(lv, o + s, o, 1.0);
            if (has_rule(c[1])) tassign(lv, o, o - s, dnormal(c[1]));
            tassign(lu, t + s, t, 1.0);
            if (has_rule(c[0])) tassign(lu, t, t - s, dtangent(c[0]));
        }
    } else {
        const int i = a, j = b;
        if (i > nx + 2 || j > ny + 2) return;
        const long s = g.sxy;
        if (FACE == 4) {
            const long o = hix(g, i, j, 0);
            if (lp) { tassign(lp, o - s, o, 1.0); tassign(lp, o, o + s, 1.0); }
            tassign(lw, o - s, o, 1.0);
            if (has_rule(c[2])) tassign(lw, o, o + s, dnormal(c[2]));
            tassign(lv, o - s, o, 1.0);
            // Back-face Neumann rule for v copies v(i,j,0) onto itself (identity)
            if (has_rule(c[1]) && c[1] != BC_NEUMANN) tassign(lv, o, o + s, dtangent(c[1]));
            tassign(lu, o - s, o, 1.0);
            if (has_rule(c[0])) tassign(lu, o, o + s, dtangent(c[0]));
        } else {
            const long t = hix(g, i, j, nz + 1), o = t - s;
            if (lp) { tassign(lp, t + s, t, 1.0); tassign(lp, t, t - s, 1.0); }
            tassign(lw, o + 2 * s, o + s, 1.0);
            tassign(lw, o + s, o, 1.0);
            if (has_rule(c[2])) tassign(lw, o, o - s, dnormal(c[2]));
            tassign(lv, t + s, t, 1.0);
            if (has_rule(c[1])) tassign(lv, t, t - s, dtangent(c[1]));
            tassign(lu, t + s, t, 1.0);
            if (has_rule(c[0])) tassign(lu, t, t - s, dtangent(c[0]));
        }
    }
}

void launch_adj_boundary_conditions(const KGrid &g, const BcTable &bc, double *lu, double *lv,
                                    double *lw, double *lp, hipStream_t st)
{
    const dim3 blk(64, 4);
    const int ex = g.nx + 4, ey = g.ny + 4, ez = g.nz + 4;
    const dim3 gx((ey + 63) / 64, (ez + 3) / 4), gy((ex + 63) / 64, (ez + 3) / 4),
        gz((ex + 63) / 64, (ey + 3) / 4);
    hipLaunchKernelGGL(k_adj_bc_face<5>, gz, blk, 0, st, g, bc, lu, lv, lw, lp);
    hipLaunchKernelGGL(k_adj_bc_face<4>, gz, blk, 0, st, g, bc, lu, lv, lw, lp);
    hipLaunchKernelGGL(k_adj_bc_face<3>, gy, blk, 0, st, g, bc, lu, lv, lw, lp);
    hipLaunchKernelGGL(k_adj_bc_face<2>, gy, blk, 0, st, g, bc, lu, lv, lw, lp);
    hipLaunchKernelGGL(k_adj_bc_face<1>, gx, blk, 0, st, g, bc, lu, lv, lw, lp);
    hipLaunchKernelGGL(k_adj_bc_face<0>, gx, blk, 0, st, g, bc, lu, lv, lw, lp);
}

// ---------------------------------------------------------------------
// projection and direct forcing, transposed
// ---------------------------------------------------------------------

// Drag and lift of one step (virtualForceIntegrator.f90:22-81) as functions
// of the projection's u1.  With the step's skip tests frozen (a y or x line
// sum skips a sample that is exactly zero), C_D is linear in FZ:
//   C_D = -2/lx sum_ijk Ax(i) Ay(i,j) Az(k) FZ(i,j,k)
//   Ax(i)   = Dxs(i)/6 (4 mx(i) + mx(i-1) [i>1] + mx(i+1) [i<nx]),   mx(i)   = [FXy(i) != 0]
//   Ay(i,j) = Dys(j)/6 (4 my(i,j) + my(i,j-1) [j>1] + my(i,j+1) [j<ny]), my(i,j) = [FXz(i,j) != 0]
//   Az(k)   = Dzs(k)/6 (4 + [k>1] + [k<nz])
// (the halo planes of FZ and the padding of the line sums are constants),
// and C_L alike in FY with FYz, FYy.  FZ = (w2 - w1)/dt with
// w2 = ETA w_solid + (1-ETA) w1 and w_solid = 0 in the resident step, so
// dFZ/dw1 = ((1-ETA) - 1)/dt and dFZ/dETA = -w1/dt.
// u = (1-ETA) u1 + ETA u_solid,  u1 = u* - dt (p(i+1) - p(i)) / Dxs(i):
// lus = (1-ETA) lu on the interior, plus the drag / lift terms of the
// objective (F.sd = dJ/dC_D, F.sl = dJ/dC_L of this step, with the step's
// line sums).  u1(nx), v1(ny), w1(nz) read the p ghost planes nx+1, ny+1,
// nz+1, which hold the previous step's values during the step; their
// adjoint goes to the ghost entries of lp (zero here: the transposed
// boundary conditions cleared them), each written by exactly one thread.
// F.geta (with the step's u1, v1, w1): dJ/dETA += lu du2/dETA + ... =
// -(lu u1 + lv v1 + lw w1) - (dJ/dFY v1 + dJ/dFZ w1)/dt.
__global__ void __launch_bounds__(256)
k_adj_project_a(KGrid g, const double *__restrict__ lu, const double *__restrict__ lv,
                const double *__restrict__ lw, const double *__restrict__ eta, double dt,
                AdjForce F, double *__restrict__ lus, double *__restrict__ lvs,
                double *__restrict__ lws, double *__restrict__ lp)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.nx || j > g.ny) return;
    const long o = hix(g, i, j, k);
    const double e = eta[o];
    const double s = 1.0 - e;
    double a = s * lu[o], b = s * lv[o], c = s * lw[o];
    const double wd = F.sd != 0.0 ? F.sd * force_weight(g, F.fxz, F.fxy, i, j, k, F.lx) : 0.0;
    const double wl = F.sl != 0.0 ? F.sl * force_weight(g, F.fyz, F.fyy, i, j, k, F.lx) : 0.0;
    const double dfd = ((1.0 - e) - 1.0) / dt;
    b += wl * dfd;
    c += wd * dfd;
    if (F.geta) {
        const double u1 = F.u1[o], v1 = F.v1[o], w1 = F.w1[o];
        F.geta[o] += -(lu[o] * u1 + lv[o] * v1 + lw[o] * w1) - (wl * v1 + wd * w1) / dt;
    }
    lus[o] = a;
    lvs[o] = b;
    lws[o] = c;
    if (i == g.nx) lp[o + 1] -= dt * a / g.Dxs[i];
    if (j == g.ny) lp[o + g.sx] -= dt * b / g.Dys[j];
    if (k == g.nz) lp[o + g.sxy] -= dt * c / g.Dzs[k];
}

// lp(i,j,k) += dt (lus(i)/Dxs(i) - lus(i-1)/Dxs(i-1)) + (y, z alike), the
// lower terms only for a lower neighbour inside the domain; and the
// interior of lu, lv, lw is cleared: the step overwrites the interior of
// the velocity, so its old interior values only enter through the
// predictor, whose transpose adds to them next.
__global__ void __launch_bounds__(256)
k_adj_project_b(KGrid g, const double *__restrict__ lus, const double *__restrict__ lvs,
                const double *__restrict__ lws, double dt, double *__restrict__ lp,
                double *__restrict__ lu, double *__restrict__ lv, double *__restrict__ lw)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.nx || j > g.ny) return;
    const long o = hix(g, i, j, k);
    double acc = dt * lus[o] / g.Dxs[i] + dt * lvs[o] / g.Dys[j] + dt * lws[o] / g.Dzs[k];
    if (i > 1) acc -= dt * lus[o - 1] / g.Dxs[i - 1];
    if (j > 1) acc -= dt * lvs[o - g.sx] / g.Dys[j - 1];
    if (k > 1) acc -= dt * lws[o - g.sxy] / g.Dzs[k - 1];
    lp[o] += acc;
    lu[o] = 0.0;
    lv[o] = 0.0;
    lw[o] = 0.0;
}

// ---------------------------------------------------------------------
// red-black SOR, transposed
// ---------------------------------------------------------------------
//
// A half-sweep of colour C updates every interior cell P of that colour,
//   p_P <- (1-omega) p_P + omega (sum_nb c_nb p_nb - b_P) / aP_P,
// with c_nb the face coefficients (ayz/Dxs, ...; the same for both cells of
// a face, the operator is symmetric) and aP_P their sum; the neighbours are
// of the other colour, interior or ghost.  Its transpose, in the scaled
// adjoint z_P = omega lambda_P / aP_P, is
//   z_Q <- z_Q + omega (sum_nb c_nb z_nb) / aP_Q      (Q of the other colour)
//   z_P <- (1-omega) z_P,   lambda_b_P -= z_P,   lambda_ghost += c z_P,
// so the reverse of the forward's k sweeps (red then black each) is a
// sequence of homogeneous SOR half-sweeps on z with the colours in the
// opposite order -- the forward kernel's update with b = 0 -- in which
// each colour's (1-omega) factor is applied by its next update:
//   for t = k..1:  red   (self factor 1 at t = k, else 1-omega)
//                  black (self factor 1-omega)
//   then red *= 1-omega.
// Z accumulates, per cell, the z values that the transposed half-sweeps of
// its own colour read: lambda_b = -Z, and a ghost adjacent to the boundary
// cell P collects c_g Z_P.  No convergence test: the sweep count is the
// forward's.
//
// Only the red half-sweeps accumulate.  The black sum follows from the red
// one: the black update is z1 <- (1-omega) z1 + omega S(z0)/aP with S
// linear, so summed over its k updates
//   sum_{m<k} z1_m = S(sum_{m<=k} z0_m)/aP + (z1_0 - z1_k)/omega,
// with z0_m the red values they read (k_adj_black_sum).  This saves the
// read and write of a black sum plane per black half-sweep.

__device__ __forceinline__ double cell_aP(const KGrid &g, const double *rx, const double *ry,
                                          const double *rz, int i, int j, int k)
{
    const double ax = g.iDx[i], ay = g.iDy[j], az = g.iDz[k];
    return ay * az * (rx[i] + rx[i - 1]) + ax * az * (ry[j] + ry[j - 1]) +
           ax * ay * (rz[k] + rz[k - 1]);
}

// lp (interior, halo layout) -> z in the colour planes; Z: the black
// cells start with their z (read by the first transposed black half-sweep),
// the red ones with 0.
__global__ void __launch_bounds__(256)
k_adj_to_split(KGrid g, SplitGrid s, const double *__restrict__ rx,
               const double *__restrict__ ry, const double *__restrict__ rz,
               const double *__restrict__ lp, double omega, double *__restrict__ z0,
               double *__restrict__ z1, double *__restrict__ Z0, double *__restrict__ Z1)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.nx || j > g.ny) return;
    const double z = omega * lp[hix(g, i, j, k)] / cell_aP(g, rx, ry, rz, i, j, k);
    const long q = six(s, i, j, k);
    if ((i + j + k) & 1) {
        z1[q] = z;
        Z1[q] = z;
    } else {
        z0[q] = z;
        Z0[q] = 0.0;
    }
}

// one transposed half-sweep: colour C cells gather the other colour's z,
//   z <- self z + omega S(z_other)/aP,   and (ACC) Z += the new z,
// in the layout of the forward FP64 sweep k_rb_sweep: RB_KPT64 cells along
// k per thread, the other colour's W/E entries of rows k0-1 .. k0+KPT
// loaded once and shared as the +-k neighbours, XCD-slab block order,
// loads first.  The arithmetic is the plain one-cell form's, term for term.
template <int C, bool ACC>
__global__ void __launch_bounds__(256)
k_adj_sweep(KGrid g, SplitGrid s, const double *__restrict__ rx,
            const double *__restrict__ ry, const double *__restrict__ rz,
            double *__restrict__ z_own, const double *__restrict__ z_oth,
            double *__restrict__ Z_own, double self, double omega)
{
    constexpr int KPT = RB_KPT64;
    int bx, by, bz;
    xcd_block(bx, by, bz);
    const int q = bx * blockDim.x + threadIdx.x;
    const int jr = by * blockDim.y + threadIdx.y + 1;
    const int j = min(jr, g.ny);
    const int k0 = bz * KPT + 1;
    const int sh0 = (C + 1 + j + k0) & 1;
    const int ia = min(2 * q + 1 + sh0, g.nx + 1), ib = min(2 * q + 2 - sh0, g.nx + 1);
    const double axa = g.iDx[ia], rEa = rx[ia], rWa = rx[ia - 1];
    const double axb = g.iDx[ib], rEb = rx[ib], rWb = rx[ib - 1];
    const double ay = g.iDy[j], rN = ry[j], rS = ry[j - 1];
    const unsigned mp = (unsigned)s.mp;
    const unsigned q1 = (unsigned)min(q, (g.nx + 1) / 2);
    double vW[KPT + 2], vE[KPT + 2], vN[KPT], vS[KPT], vo[KPT], vZ[KPT];
    unsigned at[KPT];
#pragma unroll
    for (int r = -1; r <= KPT; r++) {
        const int k = min(k0 + r, g.nz + 1);
        const unsigned row = mp * ((unsigned)j + (unsigned)s.ry * (unsigned)k);
        vW[r + 1] = z_oth[row + q1];
        vE[r + 1] = z_oth[row + q1 + 1];
    }
#pragma unroll
    for (int r = 0; r < KPT; r++) {
        const int k = min(k0 + r, g.nz);
        const unsigned row = mp * ((unsigned)j + (unsigned)s.ry * (unsigned)k);
        const unsigned o = row + q1 + (unsigned)(sh0 ^ (r & 1));
        at[r] = o;
        vN[r] = z_oth[o + mp];
        vS[r] = z_oth[o - mp];
        vo[r] = z_own[o];
        if (ACC) vZ[r] = Z_own[o];
    }
#pragma unroll
    for (int r = 0; r < KPT; r++) {
        const int k = min(k0 + r, g.nz);
        const bool odd = r & 1;
        const bool sh = sh0 ^ (int)odd;
        const double ax = odd ? axb : axa, rE = odd ? rEb : rEa, rW = odd ? rWb : rWa;
        const double az = g.iDz[k], rF = rz[k], rB = rz[k - 1];
        const double zF = sh ? vE[r + 2] : vW[r + 2], zB = sh ? vE[r] : vW[r];
        const double sum = ay * az * (rE * vE[r + 1] + rW * vW[r + 1]) +
                           ax * az * (rN * vN[r] + rS * vS[r]) + ax * ay * (rF * zF + rB * zB);
        const double aP = ay * az * (rE + rW) + ax * az * (rN + rS) + ax * ay * (rF + rB);
        const double zn = self * vo[r] + omega * sum / aP;
        const int i = 2 * q + 1 + (int)sh;
        if (jr <= g.ny && k0 + r <= g.nz && i <= g.nx) {
            z_own[at[r]] = zn;
            if (ACC) Z_own[at[r]] = vZ[r] + zn;
        }
    }
}

// after the FP64 transposed sweeps (the black ones not accumulating): the
// black sum from the red one, Z1 = S(Z0)/aP + (z1_0 - z1)/omega, with z1_0
// the black cells' initial z, which k_adj_to_split stores in Z1
__global__ void __launch_bounds__(256)
k_adj_black_sum(KGrid g, SplitGrid s, const double *__restrict__ rx,
                const double *__restrict__ ry, const double *__restrict__ rz,
                const double *__restrict__ Z0, const double *__restrict__ z1,
                double *__restrict__ Z1, double omega)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    const int i = 2 * q + ((1 + j + k) & 1);
    if (i < 1 || i > g.nx || j > g.ny) return;
    const long row = (long)s.mp * ((long)j + (long)s.ry * k);
    const long o = row + q;
    const long pl = (long)s.mp * s.ry;
    const double ZW = Z0[row + ((i - 1) >> 1)], ZE = Z0[row + ((i + 1) >> 1)];
    const double ZS = Z0[o - s.mp], ZN = Z0[o + s.mp];
    const double ZB = Z0[o - pl], ZF = Z0[o + pl];
    const double ax = g.iDx[i], ay = g.iDy[j], az = g.iDz[k];
    const double sum = ay * az * (rx[i] * ZE + rx[i - 1] * ZW) +
                       ax * az * (ry[j] * ZN + ry[j - 1] * ZS) +
                       ax * ay * (rz[k] * ZF + rz[k - 1] * ZB);
    const double aP = ay * az * (rx[i] + rx[i - 1]) + ax * az * (ry[j] + ry[j - 1]) +
                      ax * ay * (rz[k] + rz[k - 1]);
    Z1[o] = sum / aP + (Z1[o] - z1[o]) / omega;
}

// The transposed half-sweeps with FP32 arithmetic, for the adjoint of the
// FP32-sweep solver.  The transposed half-sweep (1-omega) z + omega S(z)/aP
// is affine in z like the forward one, so within a segment of half-sweeps
// z = z0 + e with z0 FP64 and fixed, and the FP32 correction e follows the
// forward FP32 sweep's update with the segment's residual
//   r = S(z0) - aP z0            (the forward residual with b = 0)
//   e <- e + omega ((S(e) + r)/aP - e)
// and (ACC, the red half-sweeps) Ef sums e after every half-sweep, so that
// at the end of the segment Z += n z0 + Ef and z0 += e (k_adj_seg_red); the
// black sum is derived there from Ef (k_adj_seg_black).  The black
// correction runs shifted as in the forward, e1' = e1 - s1 with s1 = r1/aP
// in the r1 plane and the red residual r0 + S(s1) (launch_rb_residual with
// the shift), so the black half-sweeps read no residual.  The FP32
// rounding is then relative to one segment's change of z, as in the
// forward.  Thread layout of the forward FP32 sweep: RB_KPT cells along k,
// XCD-slab block order, loads first; no convergence maxima (the count is
// fixed).
template <int C, bool ACC>
__global__ void __launch_bounds__(256)
k_adj_sweep32(KGrid g, SplitGrid s, FMetric fm, float *__restrict__ e_own,
              const float *__restrict__ e_oth, const float *__restrict__ r_own,
              float *__restrict__ Ef_own, float omega)
{
    int bx, by, bz;
    xcd_block(bx, by, bz);
    const int q = bx * blockDim.x + threadIdx.x;
    const int jr = by * blockDim.y + threadIdx.y + 1;
    const int j = min(jr, g.ny);
    const int k0 = bz * RB_KPT + 1;
    const int sh0 = (C + 1 + j + k0) & 1;
    const int ia = min(2 * q + 1 + sh0, g.nx + 1), ib = min(2 * q + 2 - sh0, g.nx + 1);
    const float axa = fm.iDx[ia], rEa = fm.rx[ia], rWa = fm.rx[ia - 1];
    const float axb = fm.iDx[ib], rEb = fm.rx[ib], rWb = fm.rx[ib - 1];
    const float ay = fm.iDy[j], rN = fm.ry[j], rS = fm.ry[j - 1];
    const unsigned mp = (unsigned)s.mp, pl = (unsigned)s.mp * (unsigned)s.ry;
    const unsigned q1 = (unsigned)min(q, (g.nx + 1) / 2);
    float vW[RB_KPT], vE[RB_KPT], vN[RB_KPT], vS[RB_KPT], vF[RB_KPT], vB[RB_KPT], vr[RB_KPT],
        vo[RB_KPT], vf[RB_KPT];
    unsigned at[RB_KPT];
#pragma unroll
    for (int r = 0; r < RB_KPT; r++) {
        const int k = min(k0 + r, g.nz);
        const unsigned row = mp * ((unsigned)j + (unsigned)s.ry * (unsigned)k);
        const unsigned o = row + q1 + (unsigned)(sh0 ^ (r & 1));
        at[r] = o;
        vW[r] = e_oth[row + q1];
        vE[r] = e_oth[row + q1 + 1];
        vN[r] = e_oth[o + mp];
        vS[r] = e_oth[o - mp];
        vF[r] = e_oth[o + pl];
        vB[r] = e_oth[o - pl];
        vr[r] = C == 0 ? r_own[o] : 0.0f;   // colour 1: shifted, no right-hand side
        vo[r] = e_own[o];
        if (ACC) vf[r] = Ef_own[o];
    }
#pragma unroll
    for (int r = 0; r < RB_KPT; r++) {
        const int k = min(k0 + r, g.nz);
        const bool odd = r & 1;
        const float ax = odd ? axb : axa, rE = odd ? rEb : rEa, rW = odd ? rWb : rWa;
        const float az = fm.iDz[k], rF = fm.rz[k], rB = fm.rz[k - 1];
        const float ayz = ay * az, axz = ax * az, axy = ax * ay;
        const float t1 = fmaf(rE, vE[r], rW * vW[r]);
        const float t2 = fmaf(rN, vN[r], rS * vS[r]);
        const float t3 = fmaf(rF, vF[r], rB * vB[r]);
        const float sum = fmaf(ayz, t1, fmaf(axz, t2, axy * t3));
        const float aP = fmaf(ayz, rE + rW, fmaf(axz, rN + rS, axy * (rF + rB)));
        const float nv = fmaf(omega, (sum + vr[r]) / aP - vo[r], vo[r]);
        const int i = 2 * q + 1 + (sh0 ^ (int)odd);
        if (jr <= g.ny && k0 + r <= g.nz && i <= g.nx) {
            e_own[at[r]] = nv;
            if (ACC) Ef_own[at[r]] = vf[r] + nv;
        }
    }
}

// A segment's end, fused with the next segment's start, in two colour
// launches (k_adj_seg_black, then k_adj_seg_red); with next = 0 (the last
// segment) only the folds.
// Colour 1, the black fold (before the red fold clears e0 and Ef0).
// The black half-sweeps do not accumulate: summing the black update, in
// the shifted correction e1' = e1 - s1 (s1 in the r1 plane),
//   e1' <- (1-omega) e1' + omega S(e0)/aP,   e1' = -s1 at the start,
// over the segment's nb black updates (e0 = 0 at its start) gives
//   sum e1 = S(E)/aP + nb s1 - ((1-omega)/omega) e1_end,
// E the sum of the red corrections they read: Ef0, less the last red
// correction when the segment ends with a red update (ends_red).  tail is
// (1-omega)/omega, plus 1 when the last black update does not count.
// Then Z1 += nacc z1 + sum and z1 += e1_end = e1' + s1 on the interior
// black cells (the ghost and padding entries are zero and stay so), in
// FP64 arithmetic on the FP32 metric the sweeps use.  Then e1' = 0 (last
// segment), or the next segment's shifted residual
// r1 = S(z0 + e0) - aP z1 (b = 0; the red fold is not done yet, so the red
// values are formed on the fly) with the FP64 metric as launch_rb_residual:
// s1 = r1/aP into the r1 plane and e1' = -s1.
__global__ void __launch_bounds__(256)
k_adj_seg_black(KGrid g, SplitGrid s, FMetric fm, const double *__restrict__ rx,
                const double *__restrict__ ry, const double *__restrict__ rz, double nb,
                double nacc, double tail, int ends_red, int next,
                const float *__restrict__ Ef0, const float *__restrict__ e0,
                const double *__restrict__ z0, float *__restrict__ s1, float *__restrict__ e1,
                double *__restrict__ z1, double *__restrict__ Z1)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    const int i = 2 * q + ((1 + j + k) & 1);
    if (i < 1 || i > g.nx || j > g.ny) return;
    const long row = (long)s.mp * ((long)j + (long)s.ry * k);
    const long o = row + q;
    const long pl = (long)s.mp * s.ry;
    // W, E, S, N, B, F
    const long at[6] = {row + ((i - 1) >> 1), row + ((i + 1) >> 1), o - s.mp, o + s.mp,
                        o - pl, o + pl};
    double E[6], zr[6];
#pragma unroll
    for (int n = 0; n < 6; n++) {
        const double ee = (double)e0[at[n]];
        E[n] = (double)Ef0[at[n]] - (ends_red ? ee : 0.0);
        zr[n] = z0[at[n]] + ee;   // ghost entries: 0 + 0
    }
    {
        const double ax = fm.iDx[i], ay = fm.iDy[j], az = fm.iDz[k];
        const double rW = fm.rx[i - 1], rE = fm.rx[i], rS = fm.ry[j - 1], rN = fm.ry[j];
        const double rB = fm.rz[k - 1], rF = fm.rz[k];
        const double sum = ay * az * (rE * E[1] + rW * E[0]) +
                           ax * az * (rN * E[3] + rS * E[2]) + ax * ay * (rF * E[5] + rB * E[4]);
        const double aP = ay * az * (rE + rW) + ax * az * (rN + rS) + ax * ay * (rF + rB);
        const double sh = (double)s1[o];
        const double e = (double)e1[o] + sh;
        const double z = z1[o];
        Z1[o] = Z1[o] + (nacc * z + ((sum / aP + nb * sh) - tail * e));
        z1[o] = z + e;
    }
    if (!next) {
        e1[o] = 0.0f;
        return;
    }
    const double ax = g.iDx[i], ay = g.iDy[j], az = g.iDz[k];
    const double rW = rx[i - 1], rE = rx[i], rS = ry[j - 1], rN = ry[j];
    const double rB = rz[k - 1], rF = rz[k];
    const double ayz = ay * az, axz = ax * az, axy = ax * ay;
    const double t1 = fma(rE, zr[1], rW * zr[0]);
    const double t2 = fma(rN, zr[3], rS * zr[2]);
    const double t3 = fma(rF, zr[5], rB * zr[4]);
    const double sum = fma(ayz, t1, fma(axz, t2, axy * t3));
    const double aP = fma(ayz, rE + rW, fma(axz, rN + rS, axy * (rF + rB)));
    const double res = (sum - 0.0) - aP * z1[o];
    const float sv = (float)(res / aP);
    s1[o] = sv;
    e1[o] = -sv;
}

// Colour 0: the red fold (Z0 += nr z0 + Ef0, z0 += e0, e0 = Ef0 = 0), then
// the next segment's residual r0 = S(z1) - aP z0 + S(s1) (b = 0; the black
// values already folded, s1 read as -e1'), FP64 metric.
__global__ void __launch_bounds__(256)
k_adj_seg_red(KGrid g, SplitGrid s, const double *__restrict__ rx,
              const double *__restrict__ ry, const double *__restrict__ rz, double nr, int next,
              float *__restrict__ e0, float *__restrict__ Ef0, double *__restrict__ z0,
              double *__restrict__ Z0, const double *__restrict__ z1,
              const float *__restrict__ e1, float *__restrict__ r0)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    const int i = 2 * q + ((j + k) & 1);
    if (i < 1 || i > g.nx || j > g.ny) return;
    const long row = (long)s.mp * ((long)j + (long)s.ry * k);
    const long o = row + q;
    const long pl = (long)s.mp * s.ry;
    const double z = z0[o];
    Z0[o] = Z0[o] + (nr * z + (double)Ef0[o]);
    const double zn = z + (double)e0[o];
    z0[o] = zn;
    e0[o] = 0.0f;
    Ef0[o] = 0.0f;
    if (!next) return;
    const long at[6] = {row + ((i - 1) >> 1), row + ((i + 1) >> 1), o - s.mp, o + s.mp,
                        o - pl, o + pl};
    double zb[6], eb[6];
#pragma unroll
    for (int n = 0; n < 6; n++) {
        zb[n] = z1[at[n]];
        eb[n] = (double)e1[at[n]];
    }
    const double ax = g.iDx[i], ay = g.iDy[j], az = g.iDz[k];
    const double rW = rx[i - 1], rE = rx[i], rS = ry[j - 1], rN = ry[j];
    const double rB = rz[k - 1], rF = rz[k];
    const double ayz = ay * az, axz = ax * az, axy = ax * ay;
    const double t1 = fma(rE, zb[1], rW * zb[0]);
    const double t2 = fma(rN, zb[3], rS * zb[2]);
    const double t3 = fma(rF, zb[5], rB * zb[4]);
    const double sum = fma(ayz, t1, fma(axz, t2, axy * t3));
    const double aP = fma(ayz, rE + rW, fma(axz, rN + rS, axy * (rF + rB)));
    double res = (sum - 0.0) - aP * zn;
    const double u1 = fma(rE, eb[1], rW * eb[0]);
    const double u2 = fma(rN, eb[3], rS * eb[2]);
    const double u3 = fma(rF, eb[5], rB * eb[4]);
    res -= fma(ayz, u1, fma(axz, u2, axy * u3));
    r0[o] = (float)res;
}

// colour planes -> lp (interior: adjoint of the initial guess, the previous
// step's pressure) and lb = -Z (adjoint of the right-hand side b)
__global__ void __launch_bounds__(256)
k_adj_from_split(KGrid g, SplitGrid s, const double *__restrict__ rx,
                 const double *__restrict__ ry, const double *__restrict__ rz,
                 const double *__restrict__ z0, const double *__restrict__ z1,
                 const double *__restrict__ Z0, const double *__restrict__ Z1, double omega,
                 double red_factor, double *__restrict__ lp, double *__restrict__ lb)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.nx || j > g.ny) return;
    const long q = six(s, i, j, k), o = hix(g, i, j, k);
    const bool black = (i + j + k) & 1;
    const double aP = cell_aP(g, rx, ry, rz, i, j, k);
    const double z = black ? z1[q] : z0[q];
    lp[o] = aP * z / omega * (black ? 1.0 : red_factor);
    lb[o] = -(black ? Z1[q] : Z0[q]);
}

// The ghost neighbours of the boundary cells enter every half-sweep with
// their face coefficient: lambda_ghost += c_g Z_P = -c_g lb_P.  Each face
// ghost has one interior neighbour, so each is written by one thread.
__global__ void __launch_bounds__(256)
k_adj_gamma(KGrid g, const double *__restrict__ lb, double *__restrict__ lp)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int b = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int face = blockIdx.z;   // 0 W, 1 E, 2 S, 3 N, 4 B, 5 F
    int i, j, k;
    long d;
    double coef;
    if (face < 2) {
        if (a > g.ny || b > g.nz) return;
        j = a; k = b; i = face == 0 ? 1 : g.nx;
        const double ayz = g.iDy[j] * g.iDz[k];
        coef = face == 0 ? ayz / g.Dxs[0] : ayz / g.Dxs[g.nx];
        d = face == 0 ? -1 : 1;
    } else if (face < 4) {
        if (a > g.nx || b > g.nz) return;
        i = a; k = b; j = face == 2 ? 1 : g.ny;
        const double axz = g.iDx[i] * g.iDz[k];
        coef = face == 2 ? axz / g.Dys[0] : axz / g.Dys[g.ny];
        d = face == 2 ? -(long)g.sx : (long)g.sx;
    } else {
        if (a > g.nx || b > g.ny) return;
        i = a; j = b; k = face == 4 ? 1 : g.nz;
        const double axy = g.iDx[i] * g.iDy[j];
        coef = face == 4 ? axy / g.Dzs[0] : axy / g.Dzs[g.nz];
        d = face == 4 ? -g.sxy : g.sxy;
    }
    const long o = hix(g, i, j, k);
    lp[o + d] -= coef * lb[o];
}

// ---------------------------------------------------------------------
// divergence (b = D u* / dt), transposed
// ---------------------------------------------------------------------

// lus(i) += (lb(cell i) - lb(cell i+1)) ay az / dt on the interior faces;
// the lowest faces u*(0), v*(0), w*(0) are the state's ghosts (the copy of
// boundary_conditions.f90:305-324) and feed lu, lv, lw there.
__global__ void __launch_bounds__(256)
k_adj_divergence(KGrid g, const double *__restrict__ lb, double dt, double *__restrict__ lus,
                 double *__restrict__ lvs, double *__restrict__ lws, double *__restrict__ lu,
                 double *__restrict__ lv, double *__restrict__ lw)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.nx || j > g.ny) return;
    const long o = hix(g, i, j, k);
    const double ax = g.iDx[i], ay = g.iDy[j], az = g.iDz[k];
    const double l0 = lb[o] / dt;
    const double le = i < g.nx ? lb[o + 1] / dt : 0.0;
    const double ln = j < g.ny ? lb[o + g.sx] / dt : 0.0;
    const double lf = k < g.nz ? lb[o + g.sxy] / dt : 0.0;
    // face i of u* is read by cell i (+) and cell i+1 (-); the areas of the
    // two cells are the same ay*az
    lus[o] += (l0 - le) * ay * az;
    lvs[o] += (l0 - ln) * ax * az;
    lws[o] += (l0 - lf) * ax * ay;
    if (i == 1) lu[o - 1] -= l0 * ay * az;
    if (j == 1) lv[o - g.sx] -= l0 * ax * az;
    if (k == 1) lw[o - g.sxy] -= l0 * ax * ay;
}

// ---------------------------------------------------------------------
// predictor, transposed (scatter)
// ---------------------------------------------------------------------

template <int C, int A>
struct Ax {
    __device__ static const double *iD(const KGrid &g) { return A == 0 ? g.iDx : (A == 1 ? g.iDy : g.iDz); }
    __device__ static const double *Ds(const KGrid &g) { return A == 0 ? g.Dxs : (A == 1 ? g.Dys : g.Dzs); }
    __device__ static double L(const KGrid &g, int m) { return C == A ? iD(g)[m + 1] : Ds(g)[m]; }
    __device__ static double Cw(const KGrid &g, int m) { return C == A ? Ds(g)[m] : iD(g)[m]; }
    __device__ static long stride(const KGrid &g) { return A == 0 ? 1L : (A == 1 ? (long)g.sx : g.sxy); }
};

// QUICK face value on the face between nodes m and m+1 as a linear
// combination of three node values: value = w0*phi(n0) + w1*phi(n1) + w2*phi(n2)
struct Face3 {
    long n0, n1, n2;
    double w0, w1, w2;
};

// forward: 0.5*(p0+p1) - c*((p1-p0)/L0 - (p0-pm1)/Lm1),  c = 0.125 L0 L0 / C0
__device__ __forceinline__ Face3 face_pos(long o0, long o1, long om1, double L0, double Lm1,
                                          double C0)
{
    const double c = 0.125 * L0 * L0 / C0;
    Face3 f;
    f.n0 = o0; f.n1 = o1; f.n2 = om1;
    f.w0 = 0.5 + c / L0 + c / Lm1;
    f.w1 = 0.5 - c / L0;
    f.w2 = -c / Lm1;
    return f;
}
// forward: 0.5*(p0+p1) - c*((p2-p1)/L1 - (p1-p0)/L0),  c = 0.125 L0 L0 / C1
__device__ __forceinline__ Face3 face_neg(long o0, long o1, long o2, double L0, double L1,
                                          double C1)
{
    const double c = 0.125 * L0 * L0 / C1;
    Face3 f;
    f.n0 = o0; f.n1 = o1; f.n2 = o2;
    f.w0 = 0.5 - c / L0;
    f.w1 = 0.5 + c / L1 + c / L0;
    f.w2 = -c / L1;
    return f;
}
__device__ __forceinline__ double face_val(const Face3 &f, const double *phi)
{
    return f.w0 * phi[f.n0] + f.w1 * phi[f.n1] + f.w2 * phi[f.n2];
}

// axis A of component C: flux difference (ap f+ - am f-)/C_A and diffusion,
// transposed for the adjoint lam of phi*(P); upwind==true uses first-order
// face values (node values) with the >= 0 tests
template <int C, int A, bool UPWIND>
__device__ __forceinline__ void adj_axis(const KGrid &g, const double *vel_a, double *lvel_a,
                                         const double *phi, double *lphi, long o, int m,
                                         double dt, double cnu, double lam, double &c0,
                                         double &lap_sum)
{
    typedef Ax<C, A> X;
    const long s = X::stride(g);
    // advecting velocities (linear in vel_a)
    long ap0, ap1, am0, am1;
    if (C == A) {
        ap0 = o + s; ap1 = o; am0 = o - s; am1 = o;
    } else {
        const long sc = Ax<C, C>::stride(g);
        ap0 = o; ap1 = o + sc; am0 = o - s; am1 = o - s + sc;
    }
    const double ap = 0.5 * (vel_a[ap0] + vel_a[ap1]);
    const double am = 0.5 * (vel_a[am0] + vel_a[am1]);
    Face3 fp, fm;
    if (UPWIND) {
        fp.n0 = ap >= 0 ? o : o + s; fp.n1 = o; fp.n2 = o; fp.w0 = 1.0; fp.w1 = 0.0; fp.w2 = 0.0;
        fm.n0 = am >= 0 ? o - s : o; fm.n1 = o; fm.n2 = o; fm.w0 = 1.0; fm.w1 = 0.0; fm.w2 = 0.0;
    } else {
        if (ap > 0) fp = face_pos(o, o + s, o - s, X::L(g, m), X::L(g, m - 1), X::Cw(g, m));
        else fp = face_neg(o, o + s, o + 2 * s, X::L(g, m), X::L(g, m + 1), X::Cw(g, m + 1));
        if (am > 0) fm = face_pos(o - s, o, o - 2 * s, X::L(g, m - 1), X::L(g, m - 2), X::Cw(g, m - 1));
        else fm = face_neg(o - s, o, o + s, X::L(g, m - 1), X::L(g, m), X::Cw(g, m));
    }
    const double Cw = X::Cw(g, m);
    const double wv = -dt * lam / Cw;   // d phi*/d (ap f+ - am f-)
    const double vp = face_val(fp, phi), vm = face_val(fm, phi);
    // This axis's contributions to lphi at o + d s (d = -2 .. 2) are summed
    // in registers and added with one atomic per node; the centre one, c0,
    // is shared by the three axes and added by the caller.
    double cm2 = 0.0, cm1 = 0.0, cp1 = 0.0, cp2 = 0.0;
    auto add = [&](long n, double v) {   // n is one of o + d s
        if (n == o) c0 += v;
        else if (n == o + s) cp1 += v;
        else if (n == o - s) cm1 += v;
        else if (n == o + 2 * s) cp2 += v;
        else cm2 += v;
    };
    const double lp = wv * ap, lm = -wv * am;
    add(fp.n0, fp.w0 * lp);
    add(fp.n1, fp.w1 * lp);
    add(fp.n2, fp.w2 * lp);
    add(fm.n0, fm.w0 * lm);
    add(fm.n1, fm.w1 * lm);
    add(fm.n2, fm.w2 * lm);
    if (C == A) {
        // the advecting velocity is phi itself (lvel_a is lphi):
        // ap0 = o + s, ap1 = am1 = o, am0 = o - s
        cp1 += 0.5 * wv * vp;
        c0 += 0.5 * wv * vp;
        cm1 += -0.5 * wv * vm;
        c0 += -0.5 * wv * vm;
    } else {
        aadd(&lvel_a[ap0], 0.5 * wv * vp);
        aadd(&lvel_a[ap1], 0.5 * wv * vp);
        aadd(&lvel_a[am0], -0.5 * wv * vm);
        aadd(&lvel_a[am1], -0.5 * wv * vm);
    }
    // diffusion: cnu*dt*((phi+ - phi)/L(m) - (phi - phi-)/L(m-1))/C
    const double Lp = X::L(g, m), Lm = X::L(g, m - 1);
    const double wd = cnu * dt * lam / Cw;
    cp1 += wd / Lp;
    c0 += -wd / Lp - wd / Lm;
    cm1 += wd / Lm;
    aadd(&lphi[o + s], cp1);
    aadd(&lphi[o - s], cm1);
    aadd(&lphi[o + 2 * s], cp2);
    aadd(&lphi[o - 2 * s], cm2);
    lap_sum += ((phi[o + s] - phi[o]) / Lp - (phi[o] - phi[o - s]) / Lm) / Cw;
}

template <int C, bool UPWIND>
__device__ __forceinline__ void adj_component(const KGrid &g, const double *u, const double *v,
                                              const double *w, double *lu, double *lv,
                                              double *lw, int i, int j, int k, double dt,
                                              double cnu, double les, double lam,
                                              double *lvisc_cell)
{
    const double *phi = C == 0 ? u : (C == 1 ? v : w);
    double *lphi = C == 0 ? lu : (C == 1 ? lv : lw);
    const long o = hix(g, i, j, k);
    double c0 = lam;   // phi* = phi + ...; the axes add their centre terms
    double lap = 0.0;
    adj_axis<C, 0, UPWIND>(g, u, lu, phi, lphi, o, i, dt, cnu, lam, c0, lap);
    if (C == 2 && UPWIND) {
        // discretisation_upwind.f90:194: the y-advecting v of the w equation
        // averages v(i,j,k) and v(i+1,j,k+1)
        typedef Ax<2, 1> Y;
        const long s = Y::stride(g);
        const long ap0 = o, ap1 = o + 1 + g.sxy, am0 = o - s, am1 = o - s + g.sxy;
        const double ap = 0.5 * (v[ap0] + v[ap1]), am = 0.5 * (v[am0] + v[am1]);
        const long np = ap >= 0 ? o : o + s, nm = am >= 0 ? o - s : o;
        const double Cw = Y::Cw(g, j);
        const double wv = -dt * lam / Cw;
        aadd(&lphi[np], wv * ap);
        aadd(&lphi[nm], -wv * am);
        aadd(&lv[ap0], 0.5 * wv * phi[np]);
        aadd(&lv[ap1], 0.5 * wv * phi[np]);
        aadd(&lv[am0], -0.5 * wv * phi[nm]);
        aadd(&lv[am1], -0.5 * wv * phi[nm]);
        const double Lp = Y::L(g, j), Lm = Y::L(g, j - 1);
        const double wd = cnu * dt * lam / Cw;
        aadd(&lphi[o + s], wd / Lp);
        aadd(&lphi[o], -wd / Lp - wd / Lm);
        aadd(&lphi[o - s], wd / Lm);
        lap += ((phi[o + s] - phi[o]) / Lp - (phi[o] - phi[o - s]) / Lm) / Cw;
    } else {
        adj_axis<C, 1, UPWIND>(g, v, lv, phi, lphi, o, j, dt, cnu, lam, c0, lap);
    }
    adj_axis<C, 2, UPWIND>(g, w, lw, phi, lphi, o, k, dt, cnu, lam, c0, lap);
    aadd(&lphi[o], c0);
    if (lvisc_cell) *lvisc_cell += les * dt * lap * lam;
}

// QUICK: every interior cell and component.  lvisc is interior-shaped.
template <bool UPWIND>
__global__ void __launch_bounds__(256)
k_adj_predict(KGrid g, const double *__restrict__ u, const double *__restrict__ v,
              const double *__restrict__ w, const double *__restrict__ visc, double dt,
              double nu, double les, const double *__restrict__ lus,
              const double *__restrict__ lvs, const double *__restrict__ lws, double *lu,
              double *lv, double *lw, double *lvisc)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int comp = blockIdx.z / g.nz;
    const int k = blockIdx.z - comp * g.nz + 1;
    if (i > g.nx || j > g.ny) return;
    const long o = hix(g, i, j, k);
    const long ci = iix(g, i, j, k);
    const double nut = visc ? visc[ci] : 0.0;
    const double cnu = nu + nut * les;
    double lnu = 0.0;
    if (comp == 0) {
        if (UPWIND && i == g.nx) { aadd(&lu[o], lus[o]); return; }   // wall face copies u
        adj_component<0, UPWIND>(g, u, v, w, lu, lv, lw, i, j, k, dt, cnu, les, lus[o], &lnu);
    } else if (comp == 1) {
        if (UPWIND && j == g.ny) { aadd(&lv[o], lvs[o]); return; }
        adj_component<1, UPWIND>(g, u, v, w, lu, lv, lw, i, j, k, dt, cnu, les, lvs[o], &lnu);
    } else {
        if (UPWIND && k == g.nz) { aadd(&lw[o], lws[o]); return; }
        adj_component<2, UPWIND>(g, u, v, w, lu, lv, lw, i, j, k, dt, cnu, les, lws[o], &lnu);
    }
    if (lvisc && !UPWIND) aadd(&lvisc[ci], lnu);
}

void launch_adj_predict(const KGrid &g, int upwind, const double *u, const double *v,
                        const double *w, const double *visc, double dt, double nu, int les,
                        const double *lus, const double *lvs, const double *lws, double *lu,
                        double *lv, double *lw, double *lvisc, hipStream_t st)
{
    dim3 blk(64, 4), grd((g.nx + 63) / 64, (g.ny + 3) / 4, 3 * g.nz);
    // the upwind predictor has no eddy viscosity (its LES branch is off)
    if (upwind)
        hipLaunchKernelGGL(k_adj_predict<true>, grd, blk, 0, st, g, u, v, w, nullptr, dt, nu,
                           0.0, lus, lvs, lws, lu, lv, lw, nullptr);
    else
        hipLaunchKernelGGL(k_adj_predict<false>, grd, blk, 0, st, g, u, v, w, visc, dt, nu,
                           (double)les, lus, lvs, lws, lu, lv, lw, lvisc);
}

// ---------------------------------------------------------------------
// Smagorinsky, linearised and transposed
// ---------------------------------------------------------------------

// nu_t = filt*sqrt(m), m = 2ux^2+2vy^2+2wz^2+(uy+vx)^2+(uz+wx)^2+(vz+wy)^2,
// gradients of the flat (nx,ny,nz) views: d/dx = (Phi(f+1)-Phi(f-1))/(2 Dxm)
// etc.; reads below the base are constants (zero), so they get no adjoint.
__device__ __forceinline__ void flat_adj(double *lphi, long f, double val)
{
    if (f >= 0) aadd(&lphi[f], val);
}

__global__ void __launch_bounds__(256)
k_adj_smagorinsky(KGrid g, const double *__restrict__ u, const double *__restrict__ v,
                  const double *__restrict__ w, const double *__restrict__ filt,
                  const double *__restrict__ lvisc, double *lu, double *lv, double *lw)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.nx || j > g.ny) return;
    const long nxy = (long)g.nx * g.ny;
    const long f = iix(g, i, j, k);
    const double lnu = lvisc[f];
    if (lnu == 0.0) return;
    auto rd = [&](const double *p, long q) { return q >= 0 ? p[q] : 0.0; };
    // half-differences over the centre: d = 0.5*(Phi(f+s) - Phi(f-s)) / Dm
    const double ux = 0.5 * (rd(u, f + 1) - rd(u, f - 1)) / g.Dxs[i];
    const double uy = 0.5 * (rd(u, f + g.nx) - rd(u, f - g.nx)) / g.iDy[j];
    const double uz = 0.5 * (rd(u, f + nxy) - rd(u, f - nxy)) / g.iDz[k];
    const double vx = 0.5 * (rd(v, f + 1) - rd(v, f - 1)) / g.iDx[i];
    const double vy = 0.5 * (rd(v, f + g.nx) - rd(v, f - g.nx)) / g.Dys[j];
    const double vz = 0.5 * (rd(v, f + nxy) - rd(v, f - nxy)) / g.iDz[k];
    const double wx = 0.5 * (rd(w, f + 1) - rd(w, f - 1)) / g.iDx[i];
    const double wy = 0.5 * (rd(w, f + g.nx) - rd(w, f - g.nx)) / g.iDy[j];
    const double wz = 0.5 * (rd(w, f + nxy) - rd(w, f - nxy)) / g.Dzs[k];
    const double m = 2.0 * ux * ux + 2.0 * vy * vy + 2.0 * wz * wz + (uy + vx) * (uy + vx) +
                     (uz + wx) * (uz + wx) + (vz + wy) * (vz + wy);
    if (!(m > 0.0)) return;
    const double lm = lnu * filt[f] * 0.5 / sqrt(m);
    const double gux = lm * 4.0 * ux, gvy = lm * 4.0 * vy, gwz = lm * 4.0 * wz;
    const double gxy = lm * 2.0 * (uy + vx), gxz = lm * 2.0 * (uz + wx), gyz = lm * 2.0 * (vz + wy);
    // u: ux (Dxs), uy (iDy), uz (iDz)
    flat_adj(lu, f + 1, 0.5 * gux / g.Dxs[i]);   flat_adj(lu, f - 1, -0.5 * gux / g.Dxs[i]);
    flat_adj(lu, f + g.nx, 0.5 * gxy / g.iDy[j]); flat_adj(lu, f - g.nx, -0.5 * gxy / g.iDy[j]);
    flat_adj(lu, f + nxy, 0.5 * gxz / g.iDz[k]);  flat_adj(lu, f - nxy, -0.5 * gxz / g.iDz[k]);
    // v: vx (iDx), vy (Dys), vz (iDz)
    flat_adj(lv, f + 1, 0.5 * gxy / g.iDx[i]);    flat_adj(lv, f - 1, -0.5 * gxy / g.iDx[i]);
    flat_adj(lv, f + g.nx, 0.5 * gvy / g.Dys[j]); flat_adj(lv, f - g.nx, -0.5 * gvy / g.Dys[j]);
    flat_adj(lv, f + nxy, 0.5 * gyz / g.iDz[k]);  flat_adj(lv, f - nxy, -0.5 * gyz / g.iDz[k]);
    // w: wx (iDx), wy (iDy), wz (Dzs)
    flat_adj(lw, f + 1, 0.5 * gxz / g.iDx[i]);    flat_adj(lw, f - 1, -0.5 * gxz / g.iDx[i]);
    flat_adj(lw, f + g.nx, 0.5 * gyz / g.iDy[j]); flat_adj(lw, f - g.nx, -0.5 * gyz / g.iDy[j]);
    flat_adj(lw, f + nxy, 0.5 * gwz / g.Dzs[k]);  flat_adj(lw, f - nxy, -0.5 * gwz / g.Dzs[k]);
}

// ---------------------------------------------------------------------
// objective and control
// ---------------------------------------------------------------------

// kinetic energy 0.5*sum (u^2+v^2+w^2) iDx iDy iDz over the interior, block
// partial sums in part[] (reduced on the host, FP64)
__global__ void __launch_bounds__(256)
k_kinetic_energy(KGrid g, const double *__restrict__ u, const double *__restrict__ v,
                 const double *__restrict__ w, double *part)
{
    __shared__ double sh[256];
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    const int tid = threadIdx.x + threadIdx.y * blockDim.x;
    double e = 0.0;
    if (i <= g.nx && j <= g.ny) {
        const long o = hix(g, i, j, k);
        const double vol = g.iDx[i] * g.iDy[j] * g.iDz[k];
        e = 0.5 * (u[o] * u[o] + v[o] * v[o] + w[o] * w[o]) * vol;
    }
    sh[tid] = e;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) sh[tid] += sh[tid + s];
        __syncthreads();
    }
    if (tid == 0) part[blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z)] = sh[0];
}

// lu += scale * u * vol on the interior (d KE/d u)
__global__ void __launch_bounds__(256)
k_ke_gradient(KGrid g, const double *__restrict__ u, const double *__restrict__ v,
              const double *__restrict__ w, double scale, double *lu, double *lv, double *lw)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.nx || j > g.ny) return;
    const long o = hix(g, i, j, k);
    const double s = scale * g.iDx[i] * g.iDy[j] * g.iDz[k];
    lu[o] += s * u[o];
    lv[o] += s * v[o];
    lw[o] += s * w[o];
}

// u* += dt f (interior), and the control gradient gf += dt lus
__global__ void __launch_bounds__(256)
k_axpy_interior(KGrid g, double a, const double *__restrict__ x, const double *__restrict__ y,
                const double *__restrict__ z, double *X, double *Y, double *Z)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.nx || j > g.ny) return;
    const long o = hix(g, i, j, k);
    X[o] += a * x[o];
    Y[o] += a * y[o];
    Z[o] += a * z[o];
}

// ---------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------

static dim3 cells_grid(const KGrid &g) { return dim3((g.nx + 63) / 64, (g.ny + 3) / 4, g.nz); }

void launch_adj_project(const KGrid &g, const double *lu, const double *lv, const double *lw,
                        const double *eta, double dt, const AdjForce &F, double *lus,
                        double *lvs, double *lws, double *lp, double *lu_clear,
                        double *lv_clear, double *lw_clear, hipStream_t st)
{
    hipLaunchKernelGGL(k_adj_project_a, cells_grid(g), dim3(64, 4), 0, st, g, lu, lv, lw, eta,
                       dt, F, lus, lvs, lws, lp);
    hipLaunchKernelGGL(k_adj_project_b, cells_grid(g), dim3(64, 4), 0, st, g, lus, lvs, lws, dt,
                       lp, lu_clear, lv_clear, lw_clear);
}

void launch_adj_pressure(const KGrid &g, const SplitGrid &s, const double *rx, const double *ry,
                         const double *rz, double omega, int sweeps, double *lp, double *lb,
                         double *z0, double *z1, double *Z0, double *Z1,
                         const AdjMixed *mx, hipStream_t st)
{
    const dim3 blk(64, 4);
    const int nq = (g.nx + 2) / 2;
    const dim3 grs((nq + 63) / 64, (g.ny + 3) / 4, g.nz);
    // the transposed sweeps: the forward FP64 sweep's grid (cell pairs,
    // RB_KPT64 planes per thread)
    const int nq64 = (g.nx + 1) / 2;
    const dim3 g64((nq64 + 63) / 64, (g.ny + 3) / 4, (g.nz + RB_KPT64 - 1) / RB_KPT64);
    hipLaunchKernelGGL(k_adj_to_split, cells_grid(g), blk, 0, st, g, s, rx, ry, rz, lp, omega,
                       z0, z1, Z0, Z1);
    if (!mx) {
        for (int t = sweeps; t >= 1; t--) {
            hipLaunchKernelGGL((k_adj_sweep<0, true>), g64, blk, 0, st, g, s, rx, ry, rz, z0, z1,
                               Z0, t == sweeps ? 1.0 : 1.0 - omega, omega);
            hipLaunchKernelGGL((k_adj_sweep<1, false>), g64, blk, 0, st, g, s, rx, ry, rz, z1,
                               z0, nullptr, 1.0 - omega, omega);
        }
        if (sweeps > 0)
            hipLaunchKernelGGL(k_adj_black_sum, grs, blk, 0, st, g, s, rx, ry, rz, Z0, z1, Z1,
                               omega);
    } else if (sweeps > 0) {
        // the first red half-sweep (self factor 1) in FP64, then the
        // sequence black, (red, black) x (sweeps-1) -- the last black not
        // accumulating -- in FP32 segments of mx->seg1 half-sweeps, then of
        // mx->seg (both even: every segment starts with a black update, and
        // only the last can end with one)
        hipLaunchKernelGGL((k_adj_sweep<0, true>), g64, blk, 0, st, g, s, rx, ry, rz, z0, z1, Z0,
                           1.0, omega);
        const int nq32 = (g.nx + 1) / 2;
        const dim3 g32((nq32 + 63) / 64, (g.ny + 3) / 4, (g.nz + RB_KPT - 1) / RB_KPT);
        const long nh = 2L * sweeps - 1;   // half-sweeps after the first
        // the first segment's residual; each segment's end computes the next
        // one's (k_adj_seg_black, k_adj_seg_red)
        launch_rb_residual(g, s, rx, ry, rz, z0, z1, mx->bzero, mx->bzero, mx->r0, mx->r1,
                           mx->e1, nullptr, st);
        for (long h0 = 0, h1; h0 < nh; h0 = h1) {
            const long len = h0 == 0 ? mx->seg1 : mx->seg;
            h1 = h0 + len < nh ? h0 + len : nh;
            for (long h = h0; h < h1; h++) {
                if (h & 1)
                    hipLaunchKernelGGL((k_adj_sweep32<0, true>), g32, blk, 0, st, g, s, mx->fm,
                                       mx->e0, mx->e1, mx->r0, mx->f0, (float)omega);
                else
                    hipLaunchKernelGGL((k_adj_sweep32<1, false>), g32, blk, 0, st, g, s, mx->fm,
                                       mx->e1, mx->e0, mx->r1, nullptr, (float)omega);
            }
            const long nb = (h1 - h0 + 1) / 2, nr = (h1 - h0) / 2;
            const bool last = h1 == nh;   // ends with the black update that does not count
            hipLaunchKernelGGL(k_adj_seg_black, grs, blk, 0, st, g, s, mx->fm, rx, ry, rz,
                               (double)nb, (double)(last ? nb - 1 : nb),
                               (1.0 - omega) / omega + (last ? 1.0 : 0.0), last ? 0 : 1,
                               last ? 0 : 1, mx->f0, mx->e0, z0, mx->r1, mx->e1, z1, Z1);
            hipLaunchKernelGGL(k_adj_seg_red, grs, blk, 0, st, g, s, rx, ry, rz, (double)nr,
                               last ? 0 : 1, mx->e0, mx->f0, z0, Z0, z1, mx->e1, mx->r0);
        }
    }
    if (sweeps > 0) {
        hipLaunchKernelGGL(k_adj_from_split, cells_grid(g), blk, 0, st, g, s, rx, ry, rz, z0, z1,
                           Z0, Z1, omega, 1.0 - omega, lp, lb);
        const int m = g.nx > g.ny ? g.nx : g.ny;
        const int mm = m > g.nz ? m : g.nz;
        hipLaunchKernelGGL(k_adj_gamma, dim3((mm + 63) / 64, (mm + 3) / 4, 6), blk, 0, st, g, lb,
                           lp);
    }
}

void launch_adj_divergence(const KGrid &g, const double *lb, double dt, double *lus,
                           double *lvs, double *lws, double *lu, double *lv, double *lw,
                           hipStream_t st)
{
    hipLaunchKernelGGL(k_adj_divergence, cells_grid(g), dim3(64, 4), 0, st, g, lb, dt, lus, lvs,
                       lws, lu, lv, lw);
}

void launch_adj_smagorinsky(const KGrid &g, const double *u, const double *v, const double *w,
                            const double *filt, const double *lvisc, double *lu, double *lv,
                            double *lw, hipStream_t st)
{
    hipLaunchKernelGGL(k_adj_smagorinsky, cells_grid(g), dim3(64, 4), 0, st, g, u, v, w, filt,
                       lvisc, lu, lv, lw);
}

int launch_kinetic_energy(const KGrid &g, const double *u, const double *v, const double *w,
                          double *part, hipStream_t st)
{
    const dim3 grd = cells_grid(g);
    hipLaunchKernelGGL(k_kinetic_energy, grd, dim3(64, 4), 0, st, g, u, v, w, part);
    return (int)(grd.x * grd.y * grd.z);
}

void launch_ke_gradient(const KGrid &g, const double *u, const double *v, const double *w,
                        double scale, double *lu, double *lv, double *lw, hipStream_t st)
{
    hipLaunchKernelGGL(k_ke_gradient, cells_grid(g), dim3(64, 4), 0, st, g, u, v, w, scale, lu,
                       lv, lw);
}

void launch_axpy_interior(const KGrid &g, double a, const double *x, const double *y,
                          const double *z, double *X, double *Y, double *Z, hipStream_t st)
{
    hipLaunchKernelGGL(k_axpy_interior, cells_grid(g), dim3(64, 4), 0, st, g, a, x, y, z, X, Y,
                       Z);
}

}  // namespace ntust
