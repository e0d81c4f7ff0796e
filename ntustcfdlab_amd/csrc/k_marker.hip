// k_marker.hip -- Lagrangian-marker direct forcing and its adjoint (gfx950).
//
// No counterpart in the reference, whose immersed boundary is the Eulerian
// volume fraction ETA blended into the velocity after the projection
// (calcul_new_velocity.f90:61-100).  The marker form replaces that blend at
// the same place in the step: with the post-projection velocity u1,
//   U = I u1                   (the velocity at the markers)
//   F = (U_solid - U)/dt       (U_solid = 0, the resting body)
//   u2 = u1 + dt S F,  FX, FY, FZ = S F,
// and the force integral (virtualForceIntegrator.f90:22-81) sums FY, FZ as
// before.  I and S use the discrete delta of ntust_kernels.h (MarkerSet,
// a cubic B-spline) and are transposes of each other.
//
// Every sum runs in a fixed order -- window z, then y, then x for the
// interpolation; list order, then x for the spreading -- with one thread per
// output, no atomics, so results are reproducible bit for bit (and equal
// the CPU restatement's, which sums in the same order without FMA).
#include "ntust_kernels.h"

namespace ntust {

__device__ __forceinline__ double node_or_centre(const double *N, int i, bool node)
{
    // component point along one axis: the upper face (node i+1) of cell i
    // for the component's own axis, else the cell centre
    return node ? N[i + 2] : 0.5 * (N[i + 1] + N[i + 2]);
}

__device__ __forceinline__ double dvol(const KGrid &g, int a, int i, int j, int k)
{
    // control volume of component a's point (i, j, k)
    if (a == 0) return (g.Dxs[i] * g.iDy[j]) * g.iDz[k];
    if (a == 1) return (g.iDx[i] * g.Dys[j]) * g.iDz[k];
    return (g.iDx[i] * g.iDy[j]) * g.Dzs[k];
}

__device__ __forceinline__ long mk_index(const KGrid &g, const MarkerSet &M, int a, int m, int ix)
{
    return ((long)a * M.n + m) * g.nx + (ix - 1);
}

// U at marker (m, ix) for the three components; thread per (ix, m)
__global__ void __launch_bounds__(64)
k_marker_interp(KGrid g, MarkerSet M, const double *__restrict__ u,
                const double *__restrict__ v, const double *__restrict__ w,
                double *__restrict__ U)
{
    const int ix = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int m = blockIdx.y;
    if (ix > g.nx) return;
    const double xm = 0.5 * (M.Xn[ix + 1] + M.Xn[ix + 2]), ym = M.ym[m], zm = M.zm[m];
    for (int a = 0; a < 3; a++) {
        const double *f = a == 0 ? u : (a == 1 ? v : w);
        const int j0 = M.jw[a][m], k0 = M.kw[a][m];
        double acc = 0.0;
        for (int dk = 0; dk < MK_WIN; dk++) {
            const int k = k0 + dk;
            if (k < 1 || k > g.nz) continue;
            const double fz = mk_phi((node_or_centre(M.Zn, k, a == 2) - zm) / M.hz) / M.hz;
            for (int dj = 0; dj < MK_WIN; dj++) {
                const int j = j0 + dj;
                if (j < 1 || j > g.ny) continue;
                const double fy = mk_phi((node_or_centre(M.Yn, j, a == 1) - ym) / M.hy) / M.hy;
                for (int i = ix - 2; i <= ix + 2; i++) {
                    if (i < 1 || i > g.nx) continue;
                    const double fx =
                        mk_phi((node_or_centre(M.Xn, i, a == 0) - xm) / M.hx) / M.hx;
                    const double wt = (fx * fy) * fz;
                    acc = acc + (wt * f[hix(g, i, j, k)]) * dvol(g, a, i, j, k);
                }
            }
        }
        U[mk_index(g, M, a, m, ix)] = acc;
    }
}

// MODE 0 (forward): F = -U/dt spread to component a's listed points, u_a +=
// dt S F, Fout = S F.  MODE 1 (adjoint): lambda_a -= dv (S L), with L the
// transposed interpolation of the adjoint.  Thread per (i, list entry).
template <int MODE>
__global__ void __launch_bounds__(64)
k_marker_spread(KGrid g, MarkerSet M, int a, const double *__restrict__ U, double dt,
                double *__restrict__ out, double *__restrict__ Fout)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int e = blockIdx.y;
    if (i > g.nx) return;
    const int j = M.sj[a][e], k = M.sk[a][e];
    const double xp = node_or_centre(M.Xn, i, a == 0);
    const double yp = node_or_centre(M.Yn, j, a == 1);
    const double zp = node_or_centre(M.Zn, k, a == 2);
    double acc = 0.0;
    for (int q = M.so[a][e]; q < M.so[a][e + 1]; q++) {
        const int m = M.sm[a][q];
        const double fy = mk_phi((yp - M.ym[m]) / M.hy) / M.hy;
        const double fz = mk_phi((zp - M.zm[m]) / M.hz) / M.hz;
        for (int ix = i - 2; ix <= i + 2; ix++) {
            if (ix < 1 || ix > g.nx) continue;
            const double xm = 0.5 * (M.Xn[ix + 1] + M.Xn[ix + 2]);
            const double fx = mk_phi((xp - xm) / M.hx) / M.hx;
            const double wt = (fx * fy) * fz;
            const double W = M.wm[m] * g.iDx[ix];
            const double val = U[mk_index(g, M, a, m, ix)];
            acc = acc + (wt * W) * (MODE == 0 ? -val / dt : val);
        }
    }
    const long o = hix(g, i, j, k);
    if (MODE == 0) {
        out[o] = out[o] + dt * acc;
        if (Fout) Fout[o] = acc;
    } else {
        out[o] = out[o] - dvol(g, a, i, j, k) * acc;
    }
}

// Adjoint of one step's forcing, first half, thread per (ix, m): with
// q = lambda + dJ/dF/dt (dJ/dFY, dJ/dFZ the force weights of the step),
//   L_a(m, ix) = sum_c w q_c                 (the transpose of S, over W)
// and, with the step's u1 (may be null),
//   dJ/dX_m = -W sum_a (dL_a U_a + L_a dU_a),  U_a = sum w u1 dv,
//   dU_a = sum (dw/dX) u1 dv,  dL_a = sum (dw/dX) q,
// for X = Y, Z (the x position of a marker is fixed): the derivative of
// -<q, S I u1> with respect to the marker's position.  GP holds the
// per-(m, ix) parts, summed over ix in order by k_marker_grad.
__global__ void __launch_bounds__(64)
k_marker_adj_interp(KGrid g, MarkerSet M, AdjForce F, double dt, const double *__restrict__ lu,
                    const double *__restrict__ lv, const double *__restrict__ lw,
                    const double *__restrict__ u1, const double *__restrict__ v1,
                    const double *__restrict__ w1, double *__restrict__ L,
                    double *__restrict__ GP)
{
    const int ix = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int m = blockIdx.y;
    if (ix > g.nx) return;
    const double xm = 0.5 * (M.Xn[ix + 1] + M.Xn[ix + 2]), ym = M.ym[m], zm = M.zm[m];
    const bool grad = u1 != nullptr;
    double gz = 0.0, gy = 0.0;
    for (int a = 0; a < 3; a++) {
        const double *lam = a == 0 ? lu : (a == 1 ? lv : lw);
        const double *vel = a == 0 ? u1 : (a == 1 ? v1 : w1);
        const int j0 = M.jw[a][m], k0 = M.kw[a][m];
        double La = 0.0, Ly = 0.0, Lz = 0.0, Ua = 0.0, Uy = 0.0, Uz = 0.0;
        for (int dk = 0; dk < MK_WIN; dk++) {
            const int k = k0 + dk;
            if (k < 1 || k > g.nz) continue;
            const double rz = (node_or_centre(M.Zn, k, a == 2) - zm) / M.hz;
            const double fz = mk_phi(rz) / M.hz, dfz = -(mk_dphi(rz) / M.hz) / M.hz;
            for (int dj = 0; dj < MK_WIN; dj++) {
                const int j = j0 + dj;
                if (j < 1 || j > g.ny) continue;
                const double ry = (node_or_centre(M.Yn, j, a == 1) - ym) / M.hy;
                const double fy = mk_phi(ry) / M.hy, dfy = -(mk_dphi(ry) / M.hy) / M.hy;
                for (int i = ix - 2; i <= ix + 2; i++) {
                    if (i < 1 || i > g.nx) continue;
                    const double fx =
                        mk_phi((node_or_centre(M.Xn, i, a == 0) - xm) / M.hx) / M.hx;
                    const long o = hix(g, i, j, k);
                    double q = lam[o];
                    if (a == 1 && F.sl != 0.0)
                        q = q + F.sl * force_weight(g, F.fyz, F.fyy, i, j, k, F.lx) / dt;
                    if (a == 2 && F.sd != 0.0)
                        q = q + F.sd * force_weight(g, F.fxz, F.fxy, i, j, k, F.lx) / dt;
                    const double wt = (fx * fy) * fz;
                    La = La + wt * q;
                    if (grad) {
                        const double wy = (fx * dfy) * fz, wz = (fx * fy) * dfz;
                        const double ud = vel[o] * dvol(g, a, i, j, k);
                        Ly = Ly + wy * q;
                        Lz = Lz + wz * q;
                        Ua = Ua + wt * ud;
                        Uy = Uy + wy * ud;
                        Uz = Uz + wz * ud;
                    }
                }
            }
        }
        L[mk_index(g, M, a, m, ix)] = La;
        gz = gz + (Lz * Ua + La * Uz);
        gy = gy + (Ly * Ua + La * Uy);
    }
    if (grad) {
        const double W = M.wm[m] * g.iDx[ix];
        GP[2 * ((long)m * g.nx + (ix - 1))] = -W * gz;
        GP[2 * ((long)m * g.nx + (ix - 1)) + 1] = -W * gy;
    }
}

// gz[m], gy[m] += the per-(m, ix) parts, summed over ix in order
__global__ void __launch_bounds__(64)
k_marker_grad(int n, int nx, const double *__restrict__ GP, double *__restrict__ gz,
              double *__restrict__ gy)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    double sz = 0.0, sy = 0.0;
    for (int ix = 0; ix < nx; ix++) {
        sz = sz + GP[2 * ((long)m * nx + ix)];
        sy = sy + GP[2 * ((long)m * nx + ix) + 1];
    }
    gz[m] = gz[m] + sz;
    gy[m] = gy[m] + sy;
}

void launch_marker_interp(const KGrid &g, const MarkerSet &M, const double *u, const double *v,
                          const double *w, double *U, hipStream_t st)
{
    if (M.n <= 0) return;
    hipLaunchKernelGGL(k_marker_interp, dim3((g.nx + 63) / 64, M.n), dim3(64), 0, st, g, M, u,
                       v, w, U);
}

void launch_marker_force(const KGrid &g, const MarkerSet &M, const double *U, double dt,
                         double *u, double *v, double *w, double *FY, double *FZ,
                         hipStream_t st)
{
    double *vel[3] = {u, v, w}, *fo[3] = {nullptr, FY, FZ};
    for (int a = 0; a < 3; a++)
        if (M.ne[a] > 0)
            hipLaunchKernelGGL(k_marker_spread<0>, dim3((g.nx + 63) / 64, M.ne[a]), dim3(64), 0,
                               st, g, M, a, U, dt, vel[a], fo[a]);
}

void launch_marker_adjoint(const KGrid &g, const MarkerSet &M, const AdjForce &F, double dt,
                           double *lu, double *lv, double *lw, const double *u1,
                           const double *v1, const double *w1, double *L, double *GP,
                           double *gz, double *gy, hipStream_t st)
{
    if (M.n <= 0) return;
    hipLaunchKernelGGL(k_marker_adj_interp, dim3((g.nx + 63) / 64, M.n), dim3(64), 0, st, g, M,
                       F, dt, lu, lv, lw, u1, v1, w1, L, GP);
    double *lam[3] = {lu, lv, lw};
    for (int a = 0; a < 3; a++)
        if (M.ne[a] > 0)
            hipLaunchKernelGGL(k_marker_spread<1>, dim3((g.nx + 63) / 64, M.ne[a]), dim3(64), 0,
                               st, g, M, a, L, dt, lam[a], nullptr);
    if (u1)
        hipLaunchKernelGGL(k_marker_grad, dim3((M.n + 63) / 64), dim3(64), 0, st, M.n, g.nx,
                           GP, gz, gy);
}

}  // namespace ntust
