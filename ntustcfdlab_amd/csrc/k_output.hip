// k_output.hip -- device side of the solver's output and restart paths:
// the PLOT3D solution array of filereachtime (filer.f90:161-187), the
// steady-state check (check_steady.f90:1-30) and the state rebuilt from a
// solution file by reading_variables (reading_data.f90:47-83).
//
// All of it is HBM-bound elementwise work on the resident state; the point
// of doing it here is that an output step moves 20 B per cell (the FP32
// solution array) over PCIe instead of the 40 B of the five FP64 fields it
// is made from, and the steady check moves 8 B in total.
#include "ntust_kernels.h"

namespace ntust {

static dim3 interior_grid(const KGrid &g) { return dim3((g.nx + 63) / 64, (g.ny + 3) / 4, g.nz); }

// Qout(i,j,k,h), h = 1..5: pre, uc, vc, wc, ETA, each rounded to a default
// real (round to nearest even, as the Fortran assignment); Qout is
// (nx,ny,nz,5) column-major.
__global__ void __launch_bounds__(256)
k_pack_q(KGrid g, const double *__restrict__ pre, const double *__restrict__ uc,
         const double *__restrict__ vc, const double *__restrict__ wc,
         const double *__restrict__ eta, float *__restrict__ q)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.nx || j > g.ny) return;
    const long c = iix(g, i, j, k), ni = (long)g.nx * g.ny * g.nz;
    q[c] = (float)pre[c];
    q[c + ni] = (float)uc[c];
    q[c + 2 * ni] = (float)vc[c];
    q[c + 3 * ni] = (float)wc[c];
    q[c + 4 * ni] = (float)eta[hix(g, i, j, k)];
}

// max |wc - last| (compare != 0) and last = wc.  As in check_steady.f90:
// a difference replaces the running maximum only when it compares greater,
// so a NaN never wins.  Non-negative doubles order like their bit patterns,
// so the block maxima meet in one 64-bit atomic max (vmax starts at 0).
__global__ void __launch_bounds__(256)
k_steady(KGrid g, const double *__restrict__ wc, double *__restrict__ last, int compare,
         unsigned long long *vmax)
{
    __shared__ double red[4];
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    const int tid = threadIdx.x + threadIdx.y * blockDim.x;
    double m = 0.0;
    if (i <= g.nx && j <= g.ny) {
        const long c = iix(g, i, j, k);
        const double x = wc[c];
        const double d = fabs(x - last[c]);
        if (d > m) m = d;
        last[c] = x;
    }
    if (!compare) return;   // uniform over the grid: no thread waits below
    m = wave_max(m);
    if ((tid & 63) == 0) red[tid >> 6] = m;
    __syncthreads();
    if (tid == 0) {
        double b = red[0];
        for (int w = 1; w < 4; w++) b = red[w] > b ? red[w] : b;
        if (b > 0.0) atomicMax(vmax, (unsigned long long)__double_as_longlong(b));
    }
}

// reading_variables: u(i,j,k) = 0.5*(Q2(i)+Q2(i+1)) in single precision,
// then widened, for i, j, k = 1..n-1 (all three components over the same
// range; other entries keep their values); p = Q1 widened and pre = p on
// the interior.
__global__ void __launch_bounds__(256)
k_unpack_q(KGrid g, const float *__restrict__ q, double *__restrict__ u,
           double *__restrict__ v, double *__restrict__ w, double *__restrict__ p,
           double *__restrict__ pre)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.nx || j > g.ny) return;
    const long c = iix(g, i, j, k), o = hix(g, i, j, k);
    const long ni = (long)g.nx * g.ny * g.nz, sj = g.nx, sk = (long)g.nx * g.ny;
    const double pv = (double)q[c];
    p[o] = pv;
    pre[c] = pv;
    if (i < g.nx && j < g.ny && k < g.nz) {
        const float *q2 = q + ni, *q3 = q + 2 * ni, *q4 = q + 3 * ni;
        const float a = 0.5f * (q2[c] + q2[c + 1]);
        const float b = 0.5f * (q3[c] + q3[c + sj]);
        const float d = 0.5f * (q4[c] + q4[c + sk]);
        u[o] = (double)a;
        v[o] = (double)b;
        w[o] = (double)d;
    }
}

// uc = 0.5*(u(i)+u(i-1)), vc, wc alike, in double precision (the literal
// 0.5 widens exactly)
__global__ void __launch_bounds__(256)
k_centre(KGrid g, const double *__restrict__ u, const double *__restrict__ v,
         const double *__restrict__ w, double *__restrict__ uc, double *__restrict__ vc,
         double *__restrict__ wc)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.nx || j > g.ny) return;
    const long c = iix(g, i, j, k), o = hix(g, i, j, k);
    uc[c] = 0.5 * (u[o] + u[o - 1]);
    vc[c] = 0.5 * (v[o] + v[o - g.sx]);
    wc[c] = 0.5 * (w[o] + w[o - g.sxy]);
}

void launch_centre(const KGrid &g, const double *u, const double *v, const double *w, double *uc,
                   double *vc, double *wc, hipStream_t st)
{
    hipLaunchKernelGGL(k_centre, interior_grid(g), dim3(64, 4), 0, st, g, u, v, w, uc, vc, wc);
}

void launch_pack_q(const KGrid &g, const double *pre, const double *uc, const double *vc,
                   const double *wc, const double *eta, float *q, hipStream_t st)
{
    hipLaunchKernelGGL(k_pack_q, interior_grid(g), dim3(64, 4), 0, st, g, pre, uc, vc, wc, eta,
                       q);
}

void launch_steady(const KGrid &g, const double *wc, double *last, int compare,
                   unsigned long long *vmax, hipStream_t st)
{
    hipLaunchKernelGGL(k_steady, interior_grid(g), dim3(64, 4), 0, st, g, wc, last, compare,
                       vmax);
}

void launch_unpack_q(const KGrid &g, const float *q, double *u, double *v, double *w, double *p,
                     double *pre, double *uc, double *vc, double *wc, hipStream_t st)
{
    hipLaunchKernelGGL(k_unpack_q, interior_grid(g), dim3(64, 4), 0, st, g, q, u, v, w, p, pre);
    hipLaunchKernelGGL(k_centre, interior_grid(g), dim3(64, 4), 0, st, g, u, v, w, uc, vc, wc);
}

// A plain streaming copy of n 16-byte words (each read once and written
// once): the known traffic the memory counters of a profile are checked
// against.  Grid-stride, 4 words in flight per thread.
__global__ void __launch_bounds__(256)
k_stream_copy(const double2 *__restrict__ src, double2 *__restrict__ dst, long n)
{
    const long stride = (long)gridDim.x * blockDim.x;
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; q + 3 * stride < n; q += 4 * stride) {
        const double2 a = src[q], b = src[q + stride], c = src[q + 2 * stride],
                      d = src[q + 3 * stride];
        dst[q] = a;
        dst[q + stride] = b;
        dst[q + 2 * stride] = c;
        dst[q + 3 * stride] = d;
    }
    for (; q < n; q += stride) dst[q] = src[q];
}

void launch_stream_copy(const void *src, void *dst, long words16, hipStream_t st)
{
    hipLaunchKernelGGL(k_stream_copy, dim3(2048), dim3(256), 0, st, (const double2 *)src,
                       (double2 *)dst, words16);
}

}  // namespace ntust
