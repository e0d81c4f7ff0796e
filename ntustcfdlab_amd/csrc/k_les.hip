// k_les.hip -- Gauss cell gradient and Smagorinsky eddy viscosity (gfx950).
//
// Restates smagorinsky.f90:1-87.  The gradient routine of the reference
// (GradientPhiGauss, :42-87) declares its argument Phi(nx,ny,nz) and is
// handed the halo arrays u, v, w, so Phi(i,j,k) is the flat element
// f = (i-1) + nx(j-1) + nx*ny(k-1) counted from element (-1,-1,-1); the
// neighbours are f+-1, f+-nx, f+-nx*ny.  The kernels read exactly those
// elements.  Offsets below zero (the k=1 plane) read memory that precedes
// the array in the reference; here they read prebase[nx*ny+f] when a
// pre-base window is supplied, else 0.0.
//
// One thread per interior cell, x fastest across the wavefront, so every
// read of the flat view is a coalesced run of 64 consecutive words.
// Traffic (fused viscosity): read u, v, w once (L2 serves the +-nx and
// +-nx*ny neighbours), read the filter table, write nu_t: 40 B/cell.
#include "ntust_kernels.h"

namespace ntust {

// Branch-free: the in-array candidate is loaded from a clamped index and
// selected, so all loads of a cell issue together (f >= -nxy always: the
// most negative offset is the -z neighbour on the k = 1 plane).
__device__ __forceinline__ double flat_at(const double *phi, const double *pre, long nxy,
                                          long f)
{
    const double a = phi[f >= 0 ? f : 0];
    if (!pre) return f >= 0 ? a : 0.0;
    const double b = pre[nxy + (f < 0 ? f : -1)];
    return f >= 0 ? a : b;
}

// Gauss gradient of the flat view at cell (i,j,k); reference order of
// operations (smagorinsky.f90:60-79).
__device__ __forceinline__ void gauss_grad(const double *phi, const double *pre, int nx,
                                           long nxy, long f, double dxm, double dym,
                                           double dzm, double &gx, double &gy, double &gz)
{
    const double c = flat_at(phi, pre, nxy, f);
    const double pe = 0.5 * (flat_at(phi, pre, nxy, f + 1) + c);
    const double pw = 0.5 * (flat_at(phi, pre, nxy, f - 1) + c);
    const double pn = 0.5 * (flat_at(phi, pre, nxy, f + nx) + c);
    const double ps = 0.5 * (flat_at(phi, pre, nxy, f - nx) + c);
    const double pf = 0.5 * (flat_at(phi, pre, nxy, f + nxy) + c);
    const double pb = 0.5 * (flat_at(phi, pre, nxy, f - nxy) + c);
    const double ax = pw * (-1.0) * (dym * dzm) + pe * (1.0) * (dym * dzm);
    const double ay = ps * (-1.0) * (dxm * dzm) + pn * (1.0) * (dxm * dzm);
    const double az = pb * (-1.0) * (dxm * dym) + pf * (1.0) * (dxm * dym);
    const double fact = 1.0 / (dxm * dym * dzm);
    gx = fact * ax;
    gy = fact * ay;
    gz = fact * az;
}

__global__ void __launch_bounds__(256)
k_gradient_flat(KGrid g, const double *__restrict__ phi, const double *__restrict__ pre,
                double *__restrict__ dphi, const double *__restrict__ Dxm,
                const double *__restrict__ Dym, const double *__restrict__ Dzm, int k0,
                int k1)
{
    int bx, by, bz;
    xcd_band(bx, by, bz);
    const int i = bx * blockDim.x + threadIdx.x + 1;
    const int j = by * blockDim.y + threadIdx.y + 1;
    const int k = bz + k0;
    if (i > g.nx || j > g.ny || k > k1) return;
    const long nxy = (long)g.nx * g.ny;
    const long f = iix(g, i, j, k);
    double gx, gy, gz;
    gauss_grad(phi, pre, g.nx, nxy, f, Dxm[i], Dym[j], Dzm[k], gx, gy, gz);
    const long n3 = nxy * g.nz;
    dphi[f] = gx;
    dphi[f + n3] = gy;
    dphi[f + 2 * n3] = gz;
}

// CalculateSmagorinskyViscosity (smagorinsky.f90:1-37) fused with the three
// gradient calls (:6-8).  filt holds (Cs*delta)**2 per cell, computed on the
// host with the same libm pow the reference calls (:20,22).
__global__ void __launch_bounds__(256)
k_smagorinsky(KGrid g, const double *__restrict__ u, const double *__restrict__ v,
              const double *__restrict__ w, const double *__restrict__ pu,
              const double *__restrict__ pv, const double *__restrict__ pw,
              const double *__restrict__ filt, double *__restrict__ visc,
              double *__restrict__ dudx, double *__restrict__ dvdx,
              double *__restrict__ dwdx, int k0, int k1)
{
    int bx, by, bz;
    xcd_band(bx, by, bz);
    const int i = bx * blockDim.x + threadIdx.x + 1;
    const int j = by * blockDim.y + threadIdx.y + 1;
    const int k = bz + k0;
    if (i > g.nx || j > g.ny || k > k1) return;
    const long nxy = (long)g.nx * g.ny;
    const long f = iix(g, i, j, k);
    const double fl = filt[f];   // issued with the gradient loads
    double ux, uy, uz, vx, vy, vz, wx, wy, wz;
    gauss_grad(u, pu, g.nx, nxy, f, g.Dxs[i], g.iDy[j], g.iDz[k], ux, uy, uz);
    gauss_grad(v, pv, g.nx, nxy, f, g.iDx[i], g.Dys[j], g.iDz[k], vx, vy, vz);
    gauss_grad(w, pw, g.nx, nxy, f, g.iDx[i], g.iDy[j], g.Dzs[k], wx, wy, wz);
    if (dudx) {
        const long n3 = nxy * g.nz;
        dudx[f] = ux; dudx[f + n3] = uy; dudx[f + 2 * n3] = uz;
        dvdx[f] = vx; dvdx[f + n3] = vy; dvdx[f + 2 * n3] = vz;
        dwdx[f] = wx; dwdx[f + n3] = wy; dwdx[f + 2 * n3] = wz;
    }
    const double m = fabs(2.0 * (ux * ux) + 2.0 * (vy * vy) + 2.0 * (wz * wz) +
                          (uy + vx) * (uy + vx) + (uz + wx) * (uz + wx) +
                          (vz + wy) * (vz + wy));
    visc[f] = fl * sqrt(m);
}

void launch_gradient_flat(const KGrid &g, const double *phi, const double *prebase,
                          double *dphi, const double *Dxm, const double *Dym,
                          const double *Dzm, int k0, int k1, hipStream_t st)
{
    if (k1 < k0) return;
    dim3 blk(64, 4), grd((g.nx + 63) / 64, (g.ny + 3) / 4, k1 - k0 + 1);
    hipLaunchKernelGGL(k_gradient_flat, grd, blk, 0, st, g, phi, prebase, dphi, Dxm, Dym,
                       Dzm, k0, k1);
}

void launch_smagorinsky(const KGrid &g, const double *u, const double *v, const double *w,
                        const double *pre_u, const double *pre_v, const double *pre_w,
                        const double *filt, double *visc, double *dudx, double *dvdx,
                        double *dwdx, int k0, int k1, hipStream_t st)
{
    if (k1 < k0) return;
    dim3 blk(64, 4), grd((g.nx + 63) / 64, (g.ny + 3) / 4, k1 - k0 + 1);
    hipLaunchKernelGGL(k_smagorinsky, grd, blk, 0, st, g, u, v, w, pre_u, pre_v, pre_w, filt,
                       visc, dudx, dvdx, dwdx, k0, k1);
}

}  // namespace ntust
