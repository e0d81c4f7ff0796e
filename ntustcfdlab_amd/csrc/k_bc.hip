// k_bc.hip -- boundary conditions (gfx950).
//
// Restates boundary_conditions.f90:1-376.  Each face sweeps the full
// -1..n+2 range of the other two indices, and faces are applied in the order
// West, East, South, North, Back, Front, so on the edges a later face
// overwrites an earlier one (e.g. the Front face resets the lid ghost
// w(i,ny+1,nz)).  One launch per face keeps that order; within a face the
// rules of u, v and w read only interior-side values the face does not
// write, so the cells of a face are independent.  The Neumann pressure
// ghosts of each face (:329-374) only touch p and ride in the same launch.
//
// Velocity rules: the tangential components get the mirror ghost
//   no-slip 0-v(1), Neumann v(1), Dirichlet 2-v(1)
// and the wall-normal component is set on the wall face itself
//   no-slip 0, Neumann u(1), Dirichlet 1
// with the second ghost layer copying the first.  The reference's Back-face
// Neumann rule for v (:228) copies v(i,j,0) onto itself, a no-op, kept here.
// An unknown code applies no rule, as a string that matches none.
#include "ntust_kernels.h"

namespace ntust {

__device__ __forceinline__ double tangent(int code, double in, double old)
{
    if (code == BC_NOSLIP) return 0.0 - in;
    if (code == BC_NEUMANN) return in;
    if (code == BC_DIRICHLET) return 2.0 - in;
    return old;
}

__device__ __forceinline__ double normal(int code, double in, double old)
{
    if (code == BC_NOSLIP) return 0.0;
    if (code == BC_NEUMANN) return in;
    if (code == BC_DIRICHLET) return 1.0;
    return old;
}

// FACE: 0 W, 1 E, 2 S, 3 N, 4 B, 5 F.  (a, b) run over -1..n+2 of the two
// tangential axes.
template <int FACE>
__global__ void __launch_bounds__(256)
k_bc_face(KGrid g, BcTable bc, double *u, double *v, double *w, double *p)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x - 1;
    const int b = blockIdx.y * blockDim.y + threadIdx.y - 1;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const int *c = bc.c + 3 * FACE;
    if (FACE == 0 || FACE == 1) {
        const int j = a, k = b;
        if (j > ny + 2 || k > nz + 2) return;
        const long s = 1;
        if (FACE == 0) {
            const long o = hix(g, 0, j, k);
            u[o] = normal(c[0], u[o + s], u[o]);
            u[o - s] = u[o];
            v[o] = tangent(c[1], v[o + s], v[o]);
            v[o - s] = v[o];
            w[o] = tangent(c[2], w[o + s], w[o]);
            w[o - s] = w[o];
            if (p) { p[o] = p[o + s]; p[o - s] = p[o]; }
        } else {
            const long o = hix(g, nx, j, k);
            u[o] = normal(c[0], u[o - s], u[o]);
            u[o + s] = u[o];
            u[o + 2 * s] = u[o + s];
            const long t = o + s;   // (nx+1, j, k)
            v[t] = tangent(c[1], v[t - s], v[t]);
            v[t + s] = v[t];
            w[t] = tangent(c[2], w[t - s], w[t]);
            w[t + s] = w[t];
            if (p) { p[t] = p[t - s]; p[t + s] = p[t]; }
        }
    } else if (FACE == 2 || FACE == 3) {
        const int i = a, k = b;
        if (i > nx + 2 || k > nz + 2) return;
        const long s = g.sx;
        if (FACE == 2) {
            const long o = hix(g, i, 0, k);
            u[o] = tangent(c[0], u[o + s], u[o]);
            u[o - s] = u[o];
            v[o] = normal(c[1], v[o + s], v[o]);
            v[o - s] = v[o];
            w[o] = tangent(c[2], w[o + s], w[o]);
            w[o - s] = w[o];
            if (p) { p[o] = p[o + s]; p[o - s] = p[o]; }
        } else {
            const long t = hix(g, i, ny + 1, k);
            u[t] = tangent(c[0], u[t - s], u[t]);
            u[t + s] = u[t];
            const long o = t - s;   // (i, ny, k)
            v[o] = normal(c[1], v[o - s], v[o]);
            v[o + s] = v[o];
            v[o + 2 * s] = v[o + s];
            w[t] = tangent(c[2], w[t - s], w[t]);
            w[t + s] = w[t];
            if (p) { p[t] = p[t - s]; p[t + s] = p[t]; }
        }
    } else {
        const int i = a, j = b;
        if (i > nx + 2 || j > ny + 2) return;
        const long s = g.sxy;
        if (FACE == 4) {
            const long o = hix(g, i, j, 0);
            u[o] = tangent(c[0], u[o + s], u[o]);
            u[o - s] = u[o];
            if (c[1] != BC_NEUMANN) v[o] = tangent(c[1], v[o + s], v[o]);
            v[o - s] = v[o];
            w[o] = normal(c[2], w[o + s], w[o]);
            w[o - s] = w[o];
            if (p) { p[o] = p[o + s]; p[o - s] = p[o]; }
        } else {
            const long t = hix(g, i, j, nz + 1);
            u[t] = tangent(c[0], u[t - s], u[t]);
            u[t + s] = u[t];
            v[t] = tangent(c[1], v[t - s], v[t]);
            v[t + s] = v[t];
            const long o = t - s;   // (i, j, nz)
            w[o] = normal(c[2], w[o - s], w[o]);
            w[o + s] = w[o];
            w[o + 2 * s] = w[o + s];
            if (p) { p[t] = p[t - s]; p[t + s] = p[t]; }
        }
    }
}

void launch_boundary_conditions(const KGrid &g, const BcTable &bc, double *u, double *v,
                                double *w, double *p, hipStream_t st)
{
    const dim3 blk(64, 4);
    const int ex = g.nx + 4, ey = g.ny + 4, ez = g.nz + 4;
    const dim3 gx((ey + 63) / 64, (ez + 3) / 4), gy((ex + 63) / 64, (ez + 3) / 4),
        gz((ex + 63) / 64, (ey + 3) / 4);
    hipLaunchKernelGGL(k_bc_face<0>, gx, blk, 0, st, g, bc, u, v, w, p);
    hipLaunchKernelGGL(k_bc_face<1>, gx, blk, 0, st, g, bc, u, v, w, p);
    hipLaunchKernelGGL(k_bc_face<2>, gy, blk, 0, st, g, bc, u, v, w, p);
    hipLaunchKernelGGL(k_bc_face<3>, gy, blk, 0, st, g, bc, u, v, w, p);
    hipLaunchKernelGGL(k_bc_face<4>, gz, blk, 0, st, g, bc, u, v, w, p);
    hipLaunchKernelGGL(k_bc_face<5>, gz, blk, 0, st, g, bc, u, v, w, p);
}

// boundary_conditions.f90:305-324: u1 = u2 = u* = u over the whole array
__global__ void __launch_bounds__(256)
k_copy_full(long n, const double *__restrict__ s, double *__restrict__ a,
            double *__restrict__ b, double *__restrict__ c)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const double x = s[t];
    if (a) a[t] = x;
    if (b) b[t] = x;
    if (c) c[t] = x;
}

void launch_copy_full(long n, const double *src, double *dst0, double *dst1, double *dst2,
                      hipStream_t st)
{
    hipLaunchKernelGGL(k_copy_full, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, src,
                       dst0, dst1, dst2);
}

// Device-resident step: the whole-array copies only matter outside the
// interior box 1..n of each axis (u* and u2 are rewritten inside it before
// anything reads them), so copy that shell of u, v, w into the u* and u2
// arrays: the four end planes k = -1, 0, nz+1, nz+2 whole, and in every
// plane k = 1..nz the frame of columns outside 1..nx x 1..ny (rows j = -1,
// 0, ny+1, ny+2 whole, then columns i = -1, 0, nx+1, nx+2 of rows 1..ny).
// One thread per shell cell.
__device__ __forceinline__ void copy6(const KGrid &g, int i, int j, int k, const double *u,
                                      const double *v, const double *w, double *a0, double *a1,
                                      double *a2, double *b0, double *b1, double *b2)
{
    const long o = hix(g, i, j, k);
    const double x = u[o], y = v[o], z = w[o];
    a0[o] = x; a1[o] = y; a2[o] = z;
    b0[o] = x; b1[o] = y; b2[o] = z;
}

__global__ void __launch_bounds__(256)
k_copy_caps(KGrid g, const double *__restrict__ u, const double *__restrict__ v,
            const double *__restrict__ w, double *a0, double *a1, double *a2, double *b0,
            double *b1, double *b2)
{
    const int ex = g.nx + 4;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)ex * (g.ny + 4)) return;
    const int i = (int)(t % ex) - 1, j = (int)(t / ex) - 1;
    const int c = blockIdx.y;
    const int k = c < 2 ? c - 1 : g.nz + c - 1;
    copy6(g, i, j, k, u, v, w, a0, a1, a2, b0, b1, b2);
}

__global__ void __launch_bounds__(256)
k_copy_frame(KGrid g, const double *__restrict__ u, const double *__restrict__ v,
             const double *__restrict__ w, double *a0, double *a1, double *a2, double *b0,
             double *b1, double *b2)
{
    const int ex = g.nx + 4;
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= 4 * ex + 4 * g.ny) return;
    int i, j;
    if (f < 4 * ex) {
        const int r = f / ex;
        j = r < 2 ? r - 1 : g.ny + r - 1;
        i = f % ex - 1;
    } else {
        const int h = f - 4 * ex, c = h % 4;
        j = h / 4 + 1;
        i = c < 2 ? c - 1 : g.nx + c - 1;
    }
    copy6(g, i, j, (int)blockIdx.y + 1, u, v, w, a0, a1, a2, b0, b1, b2);
}

void launch_copy_shell(const KGrid &g, const double *u, const double *v, const double *w,
                       double *a0, double *a1, double *a2, double *b0, double *b1,
                       double *b2, hipStream_t st)
{
    const long plane = (long)(g.nx + 4) * (g.ny + 4);
    const int frame = 4 * (g.nx + 4) + 4 * g.ny;
    hipLaunchKernelGGL(k_copy_caps, dim3((unsigned)((plane + 255) / 256), 4), dim3(256), 0, st,
                       g, u, v, w, a0, a1, a2, b0, b1, b2);
    hipLaunchKernelGGL(k_copy_frame, dim3((unsigned)((frame + 255) / 256), g.nz), dim3(256), 0,
                       st, g, u, v, w, a0, a1, a2, b0, b1, b2);
}

}  // namespace ntust
