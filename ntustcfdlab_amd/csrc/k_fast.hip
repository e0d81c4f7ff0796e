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

// Advecting velocities on the +/- faces: along the component's own axis
// from the component itself (a = phi(m+1), b = phi(m), c = phi(m-1)); along
// another axis A from that axis' velocity at o, o + s_C, o - s_A, o - s_A +
// s_C (a, b, c, d).  Their difference is taken from FP64 differences (exact
// up to one rounding).
__device__ __forceinline__ void adv_same(double a, double b, double c, double &plus,
                                         double &minus, double &diff)
{
    plus = 0.5 * (a + b);
    minus = 0.5 * (c + b);
    diff = 0.5 * (a - c);
}

__device__ __forceinline__ void adv_cross(double a, double b, double c, double d, double &plus,
                                          double &minus, double &diff)
{
    plus = 0.5 * (a + b);
    minus = 0.5 * (c + d);
    diff = 0.5 * ((a - c) + (b - d));
}

// flux difference (a+ f+ - a- f-) and second difference along one axis of
// a component at node m, in type T, from the node values p(m-2) .. p(m+2) and
// the faces m+1/2 (F) and m-1/2 (Fm).  FP64: face values from the node
// values; FP32: every term relative to p(m), from FP64 differences.
template <typename T>
__device__ __forceinline__ void axis_vals(const QCoef<T> &F, const QCoef<T> &Fm, double p0,
                                          double pm1, double pp1, double pm2, double pp2,
                                          double ap, double am, double dadv, T &flux, T &lap,
                                          T &rc)
{
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

// u* = p0 + (cnu dt (diffusion) - dt (advection)), the increment in T
template <typename T>
__device__ __forceinline__ double increment_fast(double p0, T f0, T l0, T r0, T f1, T l1, T r1,
                                                 T f2, T l2, T r2, double dt, double cnu)
{
    const T adv = fmaT<T>(r0, f0, fmaT<T>(r1, f1, r2 * f2));
    const T dif = fmaT<T>(r0, l0, fmaT<T>(r1, l1, r2 * l2));
    const T inc = fmaT<T>((T)(cnu * dt), dif, -(T)dt * adv);
    return p0 + (double)inc;
}

// The predictor in the fast modes, marching along k: a thread owns the
// column (i, j) of a chunk of PRED_KC planes and computes all three
// components of each cell, keeping u, v, w of planes k-2 .. k+2 of its
// column in registers (the z stars and the z neighbours of the advecting
// velocities), loading the x and y neighbours and the six cross-plane
// terms.  Each wavefront is one row j (blocks of 64 x 4), so j is
// wave-uniform and the y and z coefficient tables are read with scalar
// loads.  At 256^3 (FP32) it takes 425 us against 856 us for a thread per
// cell and component reading every star from the caches (the same
// arithmetic, bit for bit); forcing 3 or 4 waves per SIMD spills and takes
// 775 / 1154 us.
constexpr int PRED_KC = 8;

template <typename T>
__global__ void __launch_bounds__(256)
k_predict_quick_march(KGrid g, QTables<T> q, const double *__restrict__ u,
                      const double *__restrict__ v, const double *__restrict__ w,
                      const double *__restrict__ visc, double dt, double nu, double les,
                      double *__restrict__ us, double *__restrict__ vs,
                      double *__restrict__ ws, int k0, int k1)
{
    int bx, by, bz;
    xcd_block(bx, by, bz);
    const int il = bx * blockDim.x + threadIdx.x + 1;
    const int i = min(il, g.nx);   // lanes past nx load cell nx and store nothing
    const int j = __builtin_amdgcn_readfirstlane(min(by * (int)blockDim.y + (int)threadIdx.y + 1,
                                                     g.ny));
    const bool row = by * (int)blockDim.y + (int)threadIdx.y + 1 <= g.ny;
    const int kb = k0 + bz * PRED_KC, ke = min(kb + PRED_KC - 1, k1);
    const long sx = g.sx, sxy = g.sxy;
    // x and y faces of this column: constant along the march
    const QCoef<T> Xo = q.t[0][0][i + 1], Xom = q.t[0][0][i];   // own axis (u)
    const QCoef<T> Xt = q.t[0][1][i + 1], Xtm = q.t[0][1][i];   // transverse (v, w)
    const QCoef<T> Yo = q.t[1][0][j + 1], Yom = q.t[1][0][j];   // own axis (v)
    const QCoef<T> Yt = q.t[1][1][j + 1], Ytm = q.t[1][1][j];   // transverse (u, w)
    // planes kb-2 .. kb+2 of the column (halo planes -1 .. nz+2)
    long o = hix(g, i, j, kb);
    double um2 = u[o - 2 * sxy], um1 = u[o - sxy], u0 = u[o], up1 = u[o + sxy], up2 = u[o + 2 * sxy];
    double vm2 = v[o - 2 * sxy], vm1 = v[o - sxy], v0 = v[o], vp1 = v[o + sxy], vp2 = v[o + 2 * sxy];
    double wm2 = w[o - 2 * sxy], wm1 = w[o - sxy], w0 = w[o], wp1 = w[o + sxy], wp2 = w[o + 2 * sxy];
#pragma unroll 1
    for (int k = kb; k <= ke; k++) {
        // x and y neighbours in plane k, and the cross-plane terms
        const double uxm1 = u[o - 1], uxp1 = u[o + 1], uxm2 = u[o - 2], uxp2 = u[o + 2];
        const double vxm1 = v[o - 1], vxp1 = v[o + 1], vxm2 = v[o - 2], vxp2 = v[o + 2];
        const double wxm1 = w[o - 1], wxp1 = w[o + 1], wxm2 = w[o - 2], wxp2 = w[o + 2];
        const double uym1 = u[o - sx], uyp1 = u[o + sx], uym2 = u[o - 2 * sx], uyp2 = u[o + 2 * sx];
        const double vym1 = v[o - sx], vyp1 = v[o + sx], vym2 = v[o - 2 * sx], vyp2 = v[o + 2 * sx];
        const double wym1 = w[o - sx], wyp1 = w[o + sx], wym2 = w[o - 2 * sx], wyp2 = w[o + 2 * sx];
        const double u_xm_yp = u[o - 1 + sx], u_xm_zp = u[o - 1 + sxy];
        const double v_xp_ym = v[o + 1 - sx], v_ym_zp = v[o - sx + sxy];
        const double w_xp_zm = w[o + 1 - sxy], w_yp_zm = w[o + sx - sxy];
        const double nut = visc ? visc[iix(g, i, j, k)] : 0.0;
        const double cnu = nu + nut * les;
        const QCoef<T> Zo = q.t[2][0][k + 1], Zom = q.t[2][0][k];   // own axis (w)
        const QCoef<T> Zt = q.t[2][1][k + 1], Ztm = q.t[2][1][k];   // transverse (u, v)
        double ap0, am0, d0, ap1, am1, d1, ap2, am2, d2;
        T f0, l0, r0, f1, l1, r1, f2, l2, r2;
        // u (discretisation_QUICK_centre.f90:7-117)
        adv_same(uxp1, u0, uxm1, ap0, am0, d0);
        adv_cross(v0, vxp1, vym1, v_xp_ym, ap1, am1, d1);
        adv_cross(w0, wxp1, wm1, w_xp_zm, ap2, am2, d2);
        axis_vals<T>(Xo, Xom, u0, uxm1, uxp1, uxm2, uxp2, ap0, am0, d0, f0, l0, r0);
        axis_vals<T>(Yt, Ytm, u0, uym1, uyp1, uym2, uyp2, ap1, am1, d1, f1, l1, r1);
        axis_vals<T>(Zt, Ztm, u0, um1, up1, um2, up2, ap2, am2, d2, f2, l2, r2);
        const double nu_ = increment_fast<T>(u0, f0, l0, r0, f1, l1, r1, f2, l2, r2, dt, cnu);
        // v (:121-231)
        adv_cross(u0, uyp1, uxm1, u_xm_yp, ap0, am0, d0);
        adv_same(vyp1, v0, vym1, ap1, am1, d1);
        adv_cross(w0, wyp1, wm1, w_yp_zm, ap2, am2, d2);
        axis_vals<T>(Xt, Xtm, v0, vxm1, vxp1, vxm2, vxp2, ap0, am0, d0, f0, l0, r0);
        axis_vals<T>(Yo, Yom, v0, vym1, vyp1, vym2, vyp2, ap1, am1, d1, f1, l1, r1);
        axis_vals<T>(Zt, Ztm, v0, vm1, vp1, vm2, vp2, ap2, am2, d2, f2, l2, r2);
        const double nv_ = increment_fast<T>(v0, f0, l0, r0, f1, l1, r1, f2, l2, r2, dt, cnu);
        // w (:235-345)
        adv_cross(u0, up1, uxm1, u_xm_zp, ap0, am0, d0);
        adv_cross(v0, vp1, vym1, v_ym_zp, ap1, am1, d1);
        adv_same(wp1, w0, wm1, ap2, am2, d2);
        axis_vals<T>(Xt, Xtm, w0, wxm1, wxp1, wxm2, wxp2, ap0, am0, d0, f0, l0, r0);
        axis_vals<T>(Yt, Ytm, w0, wym1, wyp1, wym2, wyp2, ap1, am1, d1, f1, l1, r1);
        axis_vals<T>(Zo, Zom, w0, wm1, wp1, wm2, wp2, ap2, am2, d2, f2, l2, r2);
        const double nw_ = increment_fast<T>(w0, f0, l0, r0, f1, l1, r1, f2, l2, r2, dt, cnu);
        if (il <= g.nx && row) {
            us[o] = nu_;
            vs[o] = nv_;
            ws[o] = nw_;
        }
        // next plane: shift the window, load plane k+3 (clamped into the halo)
        const long on = o + sxy;
        const long o3 = hix(g, i, j, min(k + 3, g.nz + 2));
        um2 = um1; um1 = u0; u0 = up1; up1 = up2; up2 = u[o3];
        vm2 = vm1; vm1 = v0; v0 = vp1; vp1 = vp2; vp2 = v[o3];
        wm2 = wm1; wm1 = w0; w0 = wp1; wp1 = wp2; wp2 = w[o3];
        o = on;
    }
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
// The eddy viscosity marching along k: a thread owns column i of SMAG_R
// consecutive rows and SMAG_KC planes, keeping u, v, w of planes k-1, k,
// k+1 of its rows in registers (the z neighbours) and taking the y
// neighbours from its own rows but at the ends of its segment; the x
// neighbours are loaded (the row's own cache lines).  At 256^3 (FP32) it
// takes 174 us against 230 us for a thread per cell loading all 18
// neighbours, with the same bits; reading only the centres of u, v, w and
// filt takes 132 us.
constexpr int SMAG_R = 4, SMAG_KC = 32;

template <typename T>
__global__ void __launch_bounds__(256)
k_smagorinsky_fast(KGrid g, RMetric<T> r, const double *__restrict__ u,
                   const double *__restrict__ v, const double *__restrict__ w,
                   const double *__restrict__ filt, double *__restrict__ visc, int k0, int k1)
{
    constexpr int R = SMAG_R;
    int bx, by, bz;
    xcd_block(bx, by, bz);
    const int i = bx * blockDim.x + threadIdx.x + 1;
    const int j0 = (by * blockDim.y + threadIdx.y) * R + 1;
    if (i > g.nx || j0 > g.ny) return;
    const int kb = k0 + bz * SMAG_KC, ke = min(kb + SMAG_KC - 1, k1);
    const long nx = g.nx, nxy = (long)g.nx * g.ny;
    long f = iix(g, i, j0, kb);
    const T rx0 = r.rDxs[i], rx1 = r.riDx[i];   // u: own axis; v, w: transverse
    T ry0[R], ry1[R];
    bool rin[R];
#pragma unroll
    for (int q = 0; q < R; q++) {
        const int jj = min(j0 + q, g.ny);
        ry0[q] = r.riDy[jj];
        ry1[q] = r.rDys[jj];
        rin[q] = j0 + q <= g.ny;
    }
    // [q][0..2]: the flat view at planes k-1 (0 below the view), k, k+1 of
    // row j0+q (rows past ny read memory the view does not own, unused)
    double U[R][3], V[R][3], W[R][3];
#pragma unroll
    for (int q = 0; q < R; q++) {
        const long fq = f + q * nx;
        U[q][0] = flat_fast(u, nullptr, nxy, fq - nxy);
        V[q][0] = flat_fast(v, nullptr, nxy, fq - nxy);
        W[q][0] = flat_fast(w, nullptr, nxy, fq - nxy);
        U[q][1] = u[fq]; V[q][1] = v[fq]; W[q][1] = w[fq];
        U[q][2] = u[fq + nxy]; V[q][2] = v[fq + nxy]; W[q][2] = w[fq + nxy];
    }
    for (int k = kb; k <= ke; k++) {
        const T rzo = r.rDzs[k], rzt = r.riDz[k];
        // rows j0-1 and j0+R of plane k
        const long flo = f - nx, fhi = f + R * nx;
        const double ulo = flat_fast(u, nullptr, nxy, flo), vlo = flat_fast(v, nullptr, nxy, flo),
                     wlo = flat_fast(w, nullptr, nxy, flo);
        const double uhi = u[fhi], vhi = v[fhi], whi = w[fhi];
#pragma unroll
        for (int q = 0; q < R; q++) {
            const long fq = f + q * nx;
            const int qm = q > 0 ? q - 1 : 0, qp = q < R - 1 ? q + 1 : R - 1;
            const double uym = q == 0 ? ulo : U[qm][1], uyp = q == R - 1 ? uhi : U[qp][1];
            const double vym = q == 0 ? vlo : V[qm][1], vyp = q == R - 1 ? vhi : V[qp][1];
            const double wym = q == 0 ? wlo : W[qm][1], wyp = q == R - 1 ? whi : W[qp][1];
            // 0.5 (Phi(f+s) - Phi(f-s)) / D per axis, as grad_fast
            const T ux = (T)(0.5 * (u[fq + 1] - flat_fast(u, nullptr, nxy, fq - 1))) * rx0;
            const T vx = (T)(0.5 * (v[fq + 1] - flat_fast(v, nullptr, nxy, fq - 1))) * rx1;
            const T wx = (T)(0.5 * (w[fq + 1] - flat_fast(w, nullptr, nxy, fq - 1))) * rx1;
            const T uy = (T)(0.5 * (uyp - uym)) * ry0[q];
            const T vy = (T)(0.5 * (vyp - vym)) * ry1[q];
            const T wy = (T)(0.5 * (wyp - wym)) * ry0[q];
            const T uz = (T)(0.5 * (U[q][2] - U[q][0])) * rzt;
            const T vz = (T)(0.5 * (V[q][2] - V[q][0])) * rzt;
            const T wz = (T)(0.5 * (W[q][2] - W[q][0])) * rzo;
            const T a = uy + vx, b = uz + wx, c = vz + wy;
            const T m = fmaT<T>((T)2, fmaT<T>(ux, ux, fmaT<T>(vy, vy, wz * wz)),
                                fmaT<T>(a, a, fmaT<T>(b, b, c * c)));
            if (rin[q]) visc[fq] = (double)((T)filt[fq] * sqrt(m));
        }
        const long fn = f + nxy;
#pragma unroll
        for (int q = 0; q < R; q++) {
            U[q][0] = U[q][1]; U[q][1] = U[q][2];
            V[q][0] = V[q][1]; V[q][1] = V[q][2];
            W[q][0] = W[q][1]; W[q][1] = W[q][2];
            if (k < ke) {
                const long fq = fn + q * nx + nxy;
                U[q][2] = u[fq]; V[q][2] = v[fq]; W[q][2] = w[fq];
            }
        }
        f = fn;
    }
}

// ---- launchers -------------------------------------------------------------

template <typename T>
void launch_predict_quick_march(const KGrid &g, const QTables<T> &q, const double *u,
                                const double *v, const double *w, const double *visc, double dt,
                                double nu, int les, double *us, double *vs, double *ws, int k0,
                                int k1, hipStream_t st)
{
    if (k1 < k0) return;
    // blocks of 64 x 4: every wavefront one row j (the kernel reads j as
    // wave-uniform)
    dim3 blk(64, 4), grd((g.nx + 63) / 64, (g.ny + 3) / 4, (k1 - k0 + PRED_KC) / PRED_KC);
    hipLaunchKernelGGL(k_predict_quick_march<T>, grd, blk, 0, st, g, q, u, v, w, visc, dt, nu,
                       (double)les, us, vs, ws, k0, k1);
}

template <typename T>
void launch_smagorinsky_fast(const KGrid &g, const RMetric<T> &r, const double *u,
                             const double *v, const double *w, const double *filt,
                             double *visc, int k0, int k1, hipStream_t st)
{
    if (k1 < k0) return;
    dim3 blk(64, 4), grd((g.nx + 63) / 64, (g.ny + 4 * SMAG_R - 1) / (4 * SMAG_R),
                         (k1 - k0 + SMAG_KC) / SMAG_KC);
    hipLaunchKernelGGL(k_smagorinsky_fast<T>, grd, blk, 0, st, g, r, u, v, w, filt, visc, k0,
                       k1);
}

template void launch_predict_quick_march<double>(const KGrid &, const QTables<double> &,
                                                 const double *, const double *, const double *,
                                                 const double *, double, double, int, double *,
                                                 double *, double *, int, int, hipStream_t);
template void launch_predict_quick_march<float>(const KGrid &, const QTables<float> &,
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
