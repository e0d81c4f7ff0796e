// k_fast.hip -- predictor and eddy viscosity in the fast arithmetic modes
// (gfx950): the same discretisation as k_predict.hip / k_les.hip
// (discretisation_QUICK_centre.f90:1-345, smagorinsky.f90:1-87), with every
// division replaced by a product with a precomputed per-axis table and
// products fused into multiply-adds.  Not bit-identical to the reference;
// two arithmetic types:
//   T = double  (NTUST_ARITH_FAST): FP64 throughout;
//   T = float   (NTUST_ARITH_FP32): the increment u* - u and nu_t in FP32,
//               formed from first differences of the FP64 fields taken in
//               FP64, so the rounding scales with the increment and the
//               differences, not with |u|; u* = u + increment in FP64.
// The QUICK upwind choice is taken on the FP64 advecting velocities in both,
// so it falls as in the exact kernel.
//
// QUICK face values as linear combinations (the published form, restated):
// on the face between nodes m and m+1, with L the node spacing and C the
// control width of the axis (k_predict.hip's Axis), a = L(m)^2/(8 C(m)) and
// b = L(m)^2/(8 C(m+1)),
//   positive velocity: phi(m-1) (-a/L(m-1)) + phi(m) (1/2 + a/L(m) + a/L(m-1))
//                      + phi(m+1) (1/2 - a/L(m))
//   otherwise:         phi(m) (1/2 - b/L(m)) + phi(m+1) (1/2 + b/L(m+1) + b/L(m))
//                      + phi(m+2) (-b/L(m+1))
// which are discretisation_QUICK_centre.f90:18-102's expressions multiplied
// out.  The coefficients of each branch sum to 1, so relative to the node's
// own value phi(m) a face value is phi(m) plus a combination of differences.
#include "ntust_kernels.h"

namespace ntust {

template <typename T>
__device__ __forceinline__ T fmaT(T a, T b, T c)
{
    return fma(a, b, c);
}

// Advecting velocities of component C on the +/- faces along axis A, their
// difference taken from FP64 differences (exact up to one rounding).
template <int C, int A>
__device__ __forceinline__ void adv_fast(const KGrid &g, const double *vel, long o, double &plus,
                                         double &minus, double &diff)
{
    const long sa = A == 0 ? 1L : (A == 1 ? (long)g.sx : g.sxy);
    if (C == A) {
        const double a = vel[o + sa], b = vel[o], c = vel[o - sa];
        plus = 0.5 * (a + b);
        minus = 0.5 * (c + b);
        diff = 0.5 * (a - c);
    } else {
        const long sc = C == 0 ? 1L : (C == 1 ? (long)g.sx : g.sxy);
        const double a = vel[o], b = vel[o + sc], c = vel[o - sa], d = vel[o - sa + sc];
        plus = 0.5 * (a + b);
        minus = 0.5 * (c + d);
        diff = 0.5 * ((a - c) + (b - d));
    }
}

// flux difference (a+ f+ - a- f-) and second difference along axis A of
// component C at node m, in type T.  FP64: face values from the node values;
// FP32: every term relative to phi(m), from FP64 differences.
template <typename T, int C, int A>
__device__ __forceinline__ void axis_fast(const KGrid &g, const QCoef<T> *tab, const double *phi,
                                          long o, int m, double ap, double am, double dadv,
                                          T &flux, T &lap, T &rc)
{
    const long s = A == 0 ? 1L : (A == 1 ? (long)g.sx : g.sxy);
    const QCoef<T> F = tab[m + 1], Fm = tab[m];   // faces m+1/2 and m-1/2
    const double p0 = phi[o], pm1 = phi[o - s], pp1 = phi[o + s];
    const double pm2 = phi[o - 2 * s], pp2 = phi[o + 2 * s];
    rc = F.rC;
    if (sizeof(T) == 8) {
        const T fp = ap > 0 ? fmaT<T>(F.am1, pm1, fmaT<T>(F.a0, p0, F.ap1 * pp1))
                            : fmaT<T>(F.b0, p0, fmaT<T>(F.bp1, pp1, F.bp2 * pp2));
        const T fm = am > 0 ? fmaT<T>(Fm.am1, pm2, fmaT<T>(Fm.a0, pm1, Fm.ap1 * p0))
                            : fmaT<T>(Fm.b0, pm1, fmaT<T>(Fm.bp1, p0, Fm.bp2 * pp1));
        flux = fmaT<T>((T)ap, fp, -(T)am * fm);
        lap = fmaT<T>(pp1 - p0, F.rL, -(p0 - pm1) * Fm.rL);
    } else {
        const T dm2 = (T)(pm2 - p0), dm1 = (T)(pm1 - p0), dp1 = (T)(pp1 - p0),
                dp2 = (T)(pp2 - p0);
        const T ep = ap > 0 ? fmaT<T>(F.am1, dm1, F.ap1 * dp1) : fmaT<T>(F.bp1, dp1, F.bp2 * dp2);
        const T em = am > 0 ? fmaT<T>(Fm.am1, dm2, Fm.a0 * dm1) : fmaT<T>(Fm.b0, dm1, Fm.bp2 * dp1);
        // a+ (p0 + ep) - a- (p0 + em) = (a+ - a-) p0 + a+ ep - a- em
        flux = fmaT<T>((T)dadv, (T)p0, fmaT<T>((T)ap, ep, -(T)am * em));
        lap = fmaT<T>(dp1, F.rL, dm1 * Fm.rL);
    }
}

template <typename T, int C>
__device__ __forceinline__ double component_fast(const KGrid &g, const QTables<T> &q,
                                                 const double *u, const double *v,
                                                 const double *w, const double *phi, int i,
                                                 int j, int k, double dt, double cnu)
{
    const long o = hix(g, i, j, k);
    double ap0, am0, d0, ap1, am1, d1, ap2, am2, d2;
    adv_fast<C, 0>(g, u, o, ap0, am0, d0);
    adv_fast<C, 1>(g, v, o, ap1, am1, d1);
    adv_fast<C, 2>(g, w, o, ap2, am2, d2);
    T f0, l0, r0, f1, l1, r1, f2, l2, r2;
    axis_fast<T, C, 0>(g, q.t[0][C == 0 ? 0 : 1], phi, o, i, ap0, am0, d0, f0, l0, r0);
    axis_fast<T, C, 1>(g, q.t[1][C == 1 ? 0 : 1], phi, o, j, ap1, am1, d1, f1, l1, r1);
    axis_fast<T, C, 2>(g, q.t[2][C == 2 ? 0 : 1], phi, o, k, ap2, am2, d2, f2, l2, r2);
    const T adv = fmaT<T>(r0, f0, fmaT<T>(r1, f1, r2 * f2));
    const T dif = fmaT<T>(r0, l0, fmaT<T>(r1, l1, r2 * l2));
    const T inc = fmaT<T>((T)(cnu * dt), dif, -(T)dt * adv);
    return phi[o] + (double)inc;
}

template <typename T>
__global__ void __launch_bounds__(256)
k_predict_quick_fast(KGrid g, QTables<T> q, const double *__restrict__ u,
                     const double *__restrict__ v, const double *__restrict__ w,
                     const double *__restrict__ visc, double dt, double nu, double les,
                     double *__restrict__ us, double *__restrict__ vs, double *__restrict__ ws,
                     int k0)
{
    int bx, by, bz;
    xcd_band(bx, by, bz);
    const int i = bx * blockDim.x + threadIdx.x + 1;
    const int j = by * blockDim.y + threadIdx.y + 1;
    const int comp = bz % 3;
    const int k = bz / 3 + k0;
    if (i > g.nx || j > g.ny) return;
    const double nut = visc ? visc[iix(g, i, j, k)] : 0.0;
    const double cnu = nu + nut * les;
    const long o = hix(g, i, j, k);
    if (comp == 0)
        us[o] = component_fast<T, 0>(g, q, u, v, w, u, i, j, k, dt, cnu);
    else if (comp == 1)
        vs[o] = component_fast<T, 1>(g, q, u, v, w, v, i, j, k, dt, cnu);
    else
        ws[o] = component_fast<T, 2>(g, q, u, v, w, w, i, j, k, dt, cnu);
}

// ---- eddy viscosity ------------------------------------------------------

__device__ __forceinline__ double flat_fast(const double *phi, const double *pre, long nxy,
                                            long f)
{
    const double a = phi[f >= 0 ? f : 0];
    if (!pre) return f >= 0 ? a : 0.0;
    const double b = pre[nxy + (f < 0 ? f : -1)];
    return f >= 0 ? a : b;
}

// The Gauss gradient of smagorinsky.f90:60-79 on the flat view is, in real
// arithmetic, 0.5 (Phi(f+s) - Phi(f-s)) / D for each axis: the half sum of
// a face pair times the face area over the volume.
template <typename T>
__device__ __forceinline__ void grad_fast(const double *phi, const double *pre, int nx, long nxy,
                                          long f, T rx, T ry, T rz, T &gx, T &gy, T &gz)
{
    gx = (T)(0.5 * (flat_fast(phi, pre, nxy, f + 1) - flat_fast(phi, pre, nxy, f - 1))) * rx;
    gy = (T)(0.5 * (flat_fast(phi, pre, nxy, f + nx) - flat_fast(phi, pre, nxy, f - nx))) * ry;
    gz = (T)(0.5 * (flat_fast(phi, pre, nxy, f + nxy) - flat_fast(phi, pre, nxy, f - nxy))) * rz;
}

template <typename T>
__global__ void __launch_bounds__(256)
k_smagorinsky_fast(KGrid g, RMetric<T> r, const double *__restrict__ u,
                   const double *__restrict__ v, const double *__restrict__ w,
                   const double *__restrict__ filt, double *__restrict__ visc, int k0, int k1)
{
    int bx, by, bz;
    xcd_band(bx, by, bz);
    const int i = bx * blockDim.x + threadIdx.x + 1;
    const int j = by * blockDim.y + threadIdx.y + 1;
    const int k = bz + k0;
    if (i > g.nx || j > g.ny || k > k1) return;
    const long nxy = (long)g.nx * g.ny;
    const long f = iix(g, i, j, k);
    const double fl = filt[f];
    T ux, uy, uz, vx, vy, vz, wx, wy, wz;
    grad_fast<T>(u, nullptr, g.nx, nxy, f, r.rDxs[i], r.riDy[j], r.riDz[k], ux, uy, uz);
    grad_fast<T>(v, nullptr, g.nx, nxy, f, r.riDx[i], r.rDys[j], r.riDz[k], vx, vy, vz);
    grad_fast<T>(w, nullptr, g.nx, nxy, f, r.riDx[i], r.riDy[j], r.rDzs[k], wx, wy, wz);
    const T a = uy + vx, b = uz + wx, c = vz + wy;
    const T m = fmaT<T>((T)2, fmaT<T>(ux, ux, fmaT<T>(vy, vy, wz * wz)),
                        fmaT<T>(a, a, fmaT<T>(b, b, c * c)));
    visc[f] = (double)((T)fl * sqrt(m));
}

// ---- launchers -------------------------------------------------------------

template <typename T>
void launch_predict_quick_fast(const KGrid &g, const QTables<T> &q, const double *u,
                               const double *v, const double *w, const double *visc, double dt,
                               double nu, int les, double *us, double *vs, double *ws, int k0,
                               int k1, hipStream_t st)
{
    if (k1 < k0) return;
    dim3 blk(64, 4), grd((g.nx + 63) / 64, (g.ny + 3) / 4, 3 * (k1 - k0 + 1));
    hipLaunchKernelGGL(k_predict_quick_fast<T>, grd, blk, 0, st, g, q, u, v, w, visc, dt, nu,
                       (double)les, us, vs, ws, k0);
}

template <typename T>
void launch_smagorinsky_fast(const KGrid &g, const RMetric<T> &r, const double *u,
                             const double *v, const double *w, const double *filt,
                             double *visc, int k0, int k1, hipStream_t st)
{
    if (k1 < k0) return;
    dim3 blk(64, 4), grd((g.nx + 63) / 64, (g.ny + 3) / 4, k1 - k0 + 1);
    hipLaunchKernelGGL(k_smagorinsky_fast<T>, grd, blk, 0, st, g, r, u, v, w, filt, visc, k0,
                       k1);
}

template void launch_predict_quick_fast<double>(const KGrid &, const QTables<double> &,
                                                const double *, const double *, const double *,
                                                const double *, double, double, int, double *,
                                                double *, double *, int, int, hipStream_t);
template void launch_predict_quick_fast<float>(const KGrid &, const QTables<float> &,
                                               const double *, const double *, const double *,
                                               const double *, double, double, int, double *,
                                               double *, double *, int, int, hipStream_t);
template void launch_smagorinsky_fast<double>(const KGrid &, const RMetric<double> &,
                                              const double *, const double *, const double *,
                                              const double *, double *, int, int, hipStream_t);
template void launch_smagorinsky_fast<float>(const KGrid &, const RMetric<float> &,
                                             const double *, const double *, const double *,
                                             const double *, double *, int, int, hipStream_t);

}  // namespace ntust
