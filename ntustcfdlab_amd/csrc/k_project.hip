// k_project.hip -- projection, direct forcing, centring, velocity update and
// the lift/drag integral (gfx950).
//
// Restates calcul_new_velocity.f90:1-243 and virtualForceIntegrator.f90:3-85.
//
// Projection (calcul_new_velocity.f90:13-49, 61-100, 171-181):
//   u1 = u* - dt*(p(i+1)-p(i))/Dxs(i)              (v, w alike)
//   u2 = ETA*u_solid + (1-ETA)*u1,  FX = (u2-u1)/dt (v, w alike)
//   uc = 0.5*(u2(i)+u2(i-1))                        (v, w alike)
// One thread per cell.  The centred value needs u2 of the lower neighbour;
// the thread recomputes it from u*, p and ETA there (a pure function of
// them, so the bits are the neighbour's), and reads the array's ghost on the
// lowest layer.  In the device-resident step the kernel writes u directly
// (Updating_velocity, :195-243, fused) and skips u1/u2 and whatever outputs
// the step does not consume.  Traffic: read u*,v*,w*,p,ETA; write u,v,w
// (+FY,FZ for the force integral): 64 (80) B/cell.
//
// Force integral (virtualForceIntegrator.f90:22-81): the reference sums the
// weighted triples sequentially, z, then y, then x, and skips a sample
// whose own value is exactly zero.  To keep the bits, each z line is summed
// in order by one thread (coalesced across x), each y line by one thread,
// and the x line by one thread.  Samples outside the declared bounds of the
// reference's scratch arrays (FXz(i,0), FXz(i,ny+1), FXy(0), FXy(nx+1)) read
// the zero padding of the device scratch.
#include "ntust_kernels.h"

namespace ntust {

struct Proj {
    const double *us, *vs, *ws, *p, *eta;
    double usol, vsol, wsol, dt;
};

__device__ __forceinline__ double blend(double e, double sol, double v1)
{
    return e * sol + (1.0 - e) * v1;
}

__global__ void __launch_bounds__(256)
k_project(KGrid g, Proj q, double *u1, double *v1, double *w1, double *u2, double *v2,
          double *w2, double *FX, double *FY, double *FZ, double *uc, double *vc, double *wc,
          double *uo, double *vo, double *wo, const double *u2g, const double *v2g,
          const double *w2g, int k0, int k1)
{
    int bx, by, bz;
    xcd_band(bx, by, bz);
    const int i = bx * blockDim.x + threadIdx.x + 1;
    const int j = by * blockDim.y + threadIdx.y + 1;
    const int k = bz + k0;
    if (i > g.nx || j > g.ny || k > k1) return;
    const long o = hix(g, i, j, k);
    const double dt = q.dt;
    const double pc = q.p[o];
    const double a1 = q.us[o] - dt * (q.p[o + 1] - pc) / g.Dxs[i];
    const double b1 = q.vs[o] - dt * (q.p[o + g.sx] - pc) / g.Dys[j];
    const double c1 = q.ws[o] - dt * (q.p[o + g.sxy] - pc) / g.Dzs[k];
    const double e = q.eta[o];
    const double a2 = blend(e, q.usol, a1);
    const double b2 = blend(e, q.vsol, b1);
    const double c2 = blend(e, q.wsol, c1);
    if (u1) { u1[o] = a1; v1[o] = b1; w1[o] = c1; }
    if (u2) { u2[o] = a2; v2[o] = b2; w2[o] = c2; }
    if (FX) FX[o] = (a2 - a1) / dt;
    if (FY) { FY[o] = (b2 - b1) / dt; FZ[o] = (c2 - c1) / dt; }
    if (uo) { uo[o] = a2; vo[o] = b2; wo[o] = c2; }
    if (uc) {
        double am, bm, cm;
        if (i > 1) {
            const long n = o - 1;
            am = blend(q.eta[n], q.usol, q.us[n] - dt * (q.p[o] - q.p[n]) / g.Dxs[i - 1]);
        } else {
            am = u2g[o - 1];
        }
        if (j > 1) {
            const long n = o - g.sx;
            bm = blend(q.eta[n], q.vsol, q.vs[n] - dt * (q.p[o] - q.p[n]) / g.Dys[j - 1]);
        } else {
            bm = v2g[o - g.sx];
        }
        if (k > 1) {
            const long n = o - g.sxy;
            cm = blend(q.eta[n], q.wsol, q.ws[n] - dt * (q.p[o] - q.p[n]) / g.Dzs[k - 1]);
        } else {
            cm = w2g[o - g.sxy];
        }
        const long ci = iix(g, i, j, k);
        uc[ci] = 0.5 * (a2 + am);
        vc[ci] = 0.5 * (b2 + bm);
        wc[ci] = 0.5 * (c2 + cm);
    }
}

__global__ void __launch_bounds__(256)
k_copy_interior(KGrid g, const double *__restrict__ a, const double *__restrict__ b,
                const double *__restrict__ c, double *__restrict__ x, double *__restrict__ y,
                double *__restrict__ z, int k0, int k1)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + k0;
    if (i > g.nx || j > g.ny || k > k1) return;
    const long o = hix(g, i, j, k);
    x[o] = a[o];
    y[o] = b[o];
    z[o] = c[o];
}

// halo-array interior -> interior array (pre = p, GS.f90:86-94)
__global__ void __launch_bounds__(256)
k_pack_interior(KGrid g, const double *__restrict__ a, double *__restrict__ x)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.nx || j > g.ny) return;
    x[iix(g, i, j, k)] = a[hix(g, i, j, k)];
}

void launch_pack_interior(const KGrid &g, const double *a, double *x, hipStream_t st)
{
    dim3 blk(64, 4), grd((g.nx + 63) / 64, (g.ny + 3) / 4, g.nz);
    hipLaunchKernelGGL(k_pack_interior, grd, blk, 0, st, g, a, x);
}

// ---- virtualForceIntegrator ------------------------------------------

// scratch layouts: FXz/FYz (nx+2)*(ny+2) indexed (i, j) for i,j in 0..n+1,
// FXy/FYy (nx+2); the border entries stay zero.  Each sum keeps the
// reference's order; the loads of the next FORCE_CHUNK samples are issued
// together ahead of the sequential additions, so a line's sum pays one
// memory latency per chunk instead of one per sample.
constexpr int FORCE_CHUNK = 16;

__global__ void __launch_bounds__(256)
k_force_z(KGrid g, const double *__restrict__ FY, const double *__restrict__ FZ,
          double *__restrict__ FXz, double *__restrict__ FYz)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    if (i > g.nx || j > g.ny) return;
    const double *Dzs = g.Dzs;
    const long o = hix(g, i, j, 0);
    double zm = FZ[o], z0 = FZ[o + g.sxy];
    double ym = FY[o], y0 = FY[o + g.sxy];
    double fx = 0.0, fy = 0.0;
    for (int k0 = 1; k0 <= g.nz; k0 += FORCE_CHUNK) {
        // samples k0+1 .. k0+FORCE_CHUNK (plane nz+1 is the last one read)
        double zc[FORCE_CHUNK], yc[FORCE_CHUNK];
#pragma unroll
        for (int c = 0; c < FORCE_CHUNK; c++) {
            const long oc = o + g.sxy * (long)min(k0 + 1 + c, g.nz + 1);
            zc[c] = FZ[oc];
            yc[c] = FY[oc];
        }
#pragma unroll
        for (int c = 0; c < FORCE_CHUNK; c++) {
            const int k = k0 + c;
            if (k <= g.nz) {
                const double zp = zc[c], yp = yc[c];
                fx = fx + (zm * Dzs[k - 1] + 4.0 * z0 * Dzs[k] + zp * Dzs[k + 1]) / 6.0;
                fy = fy + (ym * Dzs[k - 1] + 4.0 * y0 * Dzs[k] + yp * Dzs[k + 1]) / 6.0;
                zm = z0; z0 = zp;
                ym = y0; y0 = yp;
            }
        }
    }
    const long s = (long)i + (long)(g.nx + 2) * j;
    FXz[s] = fx;
    FYz[s] = fy;
}

__global__ void __launch_bounds__(64)
k_force_y(KGrid g, const double *__restrict__ FXz, const double *__restrict__ FYz,
          double *__restrict__ FXy, double *__restrict__ FYy)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (i > g.nx) return;
    const long px = g.nx + 2;
    const double *Dys = g.Dys;
    double fx = 0.0, fy = 0.0;
    double xm = FXz[i], x0 = FXz[i + px], ym = FYz[i], y0 = FYz[i + px];
    for (int j0 = 1; j0 <= g.ny; j0 += FORCE_CHUNK) {
        double xc[FORCE_CHUNK], yc[FORCE_CHUNK];
#pragma unroll
        for (int c = 0; c < FORCE_CHUNK; c++) {
            const long q = i + px * (long)min(j0 + 1 + c, g.ny + 1);
            xc[c] = FXz[q];
            yc[c] = FYz[q];
        }
#pragma unroll
        for (int c = 0; c < FORCE_CHUNK; c++) {
            const int j = j0 + c;
            if (j <= g.ny) {
                const double xp = xc[c], yp = yc[c];
                if (x0 != 0.0) fx = fx + (xm * Dys[j - 1] + 4.0 * x0 * Dys[j] + xp * Dys[j + 1]) / 6.0;
                if (y0 != 0.0) fy = fy + (ym * Dys[j - 1] + 4.0 * y0 * Dys[j] + yp * Dys[j + 1]) / 6.0;
                xm = x0; x0 = xp;
                ym = y0; y0 = yp;
            }
        }
    }
    FXy[i] = fx;
    FYy[i] = fy;
}

__global__ void k_force_x(KGrid g, const double *__restrict__ FXy,
                          const double *__restrict__ FYy, double lx, double *coef)
{
    if (threadIdx.x != 0) return;
    const double *Dxs = g.Dxs;
    double tx = 0.0, ty = 0.0;
    double xm = FXy[0], x0 = FXy[1], ym = FYy[0], y0 = FYy[1];
    for (int i0 = 1; i0 <= g.nx; i0 += FORCE_CHUNK) {
        double xc[FORCE_CHUNK], yc[FORCE_CHUNK];
#pragma unroll
        for (int c = 0; c < FORCE_CHUNK; c++) {
            const int q = min(i0 + 1 + c, g.nx + 1);
            xc[c] = FXy[q];
            yc[c] = FYy[q];
        }
#pragma unroll
        for (int c = 0; c < FORCE_CHUNK; c++) {
            const int i = i0 + c;
            if (i <= g.nx) {
                const double xp = xc[c], yp = yc[c];
                if (x0 != 0.0) tx = tx + (xm * Dxs[i - 1] + 4.0 * x0 * Dxs[i] + xp * Dxs[i + 1]) / 6.0;
                if (y0 != 0.0) ty = ty + (ym * Dxs[i - 1] + 4.0 * y0 * Dxs[i] + yp * Dxs[i + 1]) / 6.0;
                xm = x0; x0 = xp;
                ym = y0; y0 = yp;
            }
        }
    }
    coef[0] = (-2.0) * tx / lx;   // cDrag, virtualForceIntegrator.f90:80
    coef[1] = (-2.0) * ty / lx;   // cLift, :81
}

void launch_project(const KGrid &g, const double *us, const double *vs, const double *ws,
                    const double *p, const double *eta, double usol, double vsol,
                    double wsol, double dt, double *u1, double *v1, double *w1,
                    double *u2, double *v2, double *w2, double *FX, double *FY,
                    double *FZ, double *uc, double *vc, double *wc, double *uout,
                    double *vout, double *wout, const double *u2g, const double *v2g,
                    const double *w2g, int k0, int k1, hipStream_t st)
{
    if (k1 < k0) return;
    Proj q{us, vs, ws, p, eta, usol, vsol, wsol, dt};
    dim3 blk(64, 4), grd((g.nx + 63) / 64, (g.ny + 3) / 4, k1 - k0 + 1);
    hipLaunchKernelGGL(k_project, grd, blk, 0, st, g, q, u1, v1, w1, u2, v2, w2, FX, FY, FZ,
                       uc, vc, wc, uout, vout, wout, u2g, v2g, w2g, k0, k1);
}

void launch_copy_interior(const KGrid &g, const double *a, const double *b, const double *c,
                          double *x, double *y, double *z, int k0, int k1, hipStream_t st)
{
    if (k1 < k0) return;
    dim3 blk(64, 4), grd((g.nx + 63) / 64, (g.ny + 3) / 4, k1 - k0 + 1);
    hipLaunchKernelGGL(k_copy_interior, grd, blk, 0, st, g, a, b, c, x, y, z, k0, k1);
}

// The entries of the line sums that k_force_y and k_force_x read but no
// kernel writes -- rows j = 0 and ny+1 of FXz, FYz and entries 0 and nx+1 of
// FXy, FYy, the reference's out-of-bounds reads (virtualForceIntegrator.f90:
// 49,53,68,72) -- set to 0 for scratch memory that a previous call, possibly
// of another grid, has left holding other values.
__global__ void __launch_bounds__(64)
k_force_clear(int nx, int ny, double *__restrict__ FXz, double *__restrict__ FYz,
              double *__restrict__ FXy, double *__restrict__ FYy)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > nx + 1) return;
    const long top = (long)(nx + 2) * (ny + 1);
    FXz[i] = 0.0;
    FYz[i] = 0.0;
    FXz[i + top] = 0.0;
    FYz[i + top] = 0.0;
    if (i == 0 || i == nx + 1) {
        FXy[i] = 0.0;
        FYy[i] = 0.0;
    }
}

void launch_force_clear(const KGrid &g, double *FXz, double *FYz, double *FXy, double *FYy,
                        hipStream_t st)
{
    hipLaunchKernelGGL(k_force_clear, dim3((g.nx + 2 + 63) / 64), dim3(64), 0, st, g.nx, g.ny,
                       FXz, FYz, FXy, FYy);
}

void launch_force_integral(const KGrid &g, const double *FY, const double *FZ, double lx,
                           double *FXz, double *FYz, double *FXy, double *FYy,
                           double *coef_out, hipStream_t st)
{
    dim3 blk(64, 4), grd((g.nx + 63) / 64, (g.ny + 3) / 4, 1);
    hipLaunchKernelGGL(k_force_z, grd, blk, 0, st, g, FY, FZ, FXz, FYz);
    hipLaunchKernelGGL(k_force_y, dim3((g.nx + 63) / 64), dim3(64), 0, st, g, FXz, FYz, FXy,
                       FYy);
    hipLaunchKernelGGL(k_force_x, dim3(1), dim3(64), 0, st, g, FXy, FYy, lx, coef_out);
}

}  // namespace ntust
