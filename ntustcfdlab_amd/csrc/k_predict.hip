// k_predict.hip -- momentum predictor u*, v*, w* (gfx950).
//
// Restates discretisation_QUICK_centre.f90:1-345 (QUICK advection of the
// face-averaged advecting velocity plus central diffusion with nu + nu_t*LES)
// and discretisation_upwind.f90:1-275 (first-order upwind, the undeclared Q1
// taken as 0 so its inline LES branch is off).
//
// The reference writes nine near-copies of one formula.  Here one templated
// routine covers all three components: on a staggered grid the component's
// own axis ("normal") and the two transverse axes differ only in which metric
// plays the node spacing L(m) and which the control width C(m):
//   normal axis a == c:      L(m) = iD_a(m+1)  C(m) = D_as(m)
//   transverse axis a != c:  L(m) = D_as(m)    C(m) = iD_a(m)
// With those, every QUICK face value of the reference is
//   0.5*(phi(m)+phi(m+1)) - 0.125*L(m)*L(m)/C(m)  *((phi(m+1)-phi(m))/L(m)  -(phi(m)-phi(m-1))/L(m-1))
//   0.5*(phi(m)+phi(m+1)) - 0.125*L(m)*L(m)/C(m+1)*((phi(m+2)-phi(m+1))/L(m+1)-(phi(m+1)-phi(m))/L(m))
// for a positive / non-positive advecting velocity on the face between nodes
// m and m+1, and the update divides the x, y, z flux differences by C of each
// axis.  The operation order of each expression is the reference's, so the
// results are bit-identical to it (compiled without FMA contraction).
//
// One thread per (cell, component); x fastest.  Reads u, v, w (radius-2 star
// per axis plus the four cross points of the advecting velocities) and nu_t;
// writes one component.  Exact arithmetic keeps ~30 FP64 divisions per
// component, so this kernel is FP64-issue bound, not HBM bound.
#include "ntust_kernels.h"

namespace ntust {

template <int C, int A>
struct Axis {
    // metric pointers of axis A
    __device__ static const double *iD(const KGrid &g) { return A == 0 ? g.iDx : (A == 1 ? g.iDy : g.iDz); }
    __device__ static const double *Ds(const KGrid &g) { return A == 0 ? g.Dxs : (A == 1 ? g.Dys : g.Dzs); }
    __device__ static double L(const KGrid &g, int m) { return C == A ? iD(g)[m + 1] : Ds(g)[m]; }
    __device__ static double Cw(const KGrid &g, int m) { return C == A ? Ds(g)[m] : iD(g)[m]; }
    __device__ static long stride(const KGrid &g) { return A == 0 ? 1L : (A == 1 ? (long)g.sx : g.sxy); }
};

__device__ __forceinline__ double quick_pos(double p0, double p1, double pm1, double L0,
                                            double Lm1, double C0)
{
    // 0.5*(phi(m)+phi(m+1)) - 0.125*L(m)*L(m)/C(m)*((phi(m+1)-phi(m))/L(m) - (phi(m)-phi(m-1))/L(m-1))
    return 0.5 * (p0 + p1) - 0.125 * L0 * L0 / C0 * ((p1 - p0) / L0 - (p0 - pm1) / Lm1);
}

__device__ __forceinline__ double quick_neg(double p0, double p1, double p2, double L0,
                                            double L1, double C1)
{
    // 0.5*(phi(m)+phi(m+1)) - 0.125*L(m)*L(m)/C(m+1)*((phi(m+2)-phi(m+1))/L(m+1) - (phi(m+1)-phi(m))/L(m))
    return 0.5 * (p0 + p1) - 0.125 * L0 * L0 / C1 * ((p2 - p1) / L1 - (p1 - p0) / L0);
}

// Advecting velocities of component C on the +/- faces along axis A.
template <int C, int A>
__device__ __forceinline__ void advecting(const KGrid &g, const double *vel_a, long o,
                                          double &plus, double &minus)
{
    const long sa = Axis<C, A>::stride(g);
    if (C == A) {
        plus = 0.5 * (vel_a[o + sa] + vel_a[o]);
        minus = 0.5 * (vel_a[o - sa] + vel_a[o]);
    } else {
        const long sc = Axis<C, C>::stride(g);
        plus = 0.5 * (vel_a[o] + vel_a[o + sc]);
        minus = 0.5 * (vel_a[o - sa] + vel_a[o - sa + sc]);
    }
}

// Flux difference and diffusion along axis A for component C at node m.
template <int C, int A>
__device__ __forceinline__ void quick_axis(const KGrid &g, const double *phi, long o, int m,
                                           double adv_p, double adv_m, double &fplus,
                                           double &fminus, double &lap)
{
    typedef Axis<C, A> X;
    const long s = X::stride(g);
    const double p0 = phi[o];
    const double pp1 = phi[o + s], pm1 = phi[o - s];
    // face between m and m+1
    if (adv_p > 0)
        fplus = quick_pos(p0, pp1, pm1, X::L(g, m), X::L(g, m - 1), X::Cw(g, m));
    else
        fplus = quick_neg(p0, pp1, phi[o + 2 * s], X::L(g, m), X::L(g, m + 1), X::Cw(g, m + 1));
    // face between m-1 and m
    if (adv_m > 0)
        fminus = quick_pos(pm1, p0, phi[o - 2 * s], X::L(g, m - 1), X::L(g, m - 2),
                           X::Cw(g, m - 1));
    else
        fminus = quick_neg(pm1, p0, pp1, X::L(g, m - 1), X::L(g, m), X::Cw(g, m));
    lap = (pp1 - p0) / X::L(g, m) - (p0 - pm1) / X::L(g, m - 1);
}

template <int C>
__device__ __forceinline__ double quick_component(const KGrid &g, const double *u,
                                                  const double *v, const double *w,
                                                  const double *phi, int i, int j, int k,
                                                  double dt, double cnu)
{
    const long o = hix(g, i, j, k);
    double ap0, am0, ap1, am1, ap2, am2;
    advecting<C, 0>(g, u, o, ap0, am0);
    advecting<C, 1>(g, v, o, ap1, am1);
    advecting<C, 2>(g, w, o, ap2, am2);
    double fp0, fm0, l0, fp1, fm1, l1, fp2, fm2, l2;
    quick_axis<C, 0>(g, phi, o, i, ap0, am0, fp0, fm0, l0);
    quick_axis<C, 1>(g, phi, o, j, ap1, am1, fp1, fm1, l1);
    quick_axis<C, 2>(g, phi, o, k, ap2, am2, fp2, fm2, l2);
    const double c0 = Axis<C, 0>::Cw(g, i), c1 = Axis<C, 1>::Cw(g, j), c2 = Axis<C, 2>::Cw(g, k);
    const double p = phi[o];
    return p - dt * (ap0 * fp0 - am0 * fm0) / c0
             - dt * (ap1 * fp1 - am1 * fm1) / c1
             - dt * (ap2 * fp2 - am2 * fm2) / c2
             + cnu * dt * (l0) / c0
             + cnu * dt * (l1) / c1
             + cnu * dt * (l2) / c2;
}

// blockIdx.z enumerates (k, component), the component fastest, so the three
// components of a plane run together and share its u, v, w in L2
__global__ void __launch_bounds__(256)
k_predict_quick(KGrid g, const double *__restrict__ u, const double *__restrict__ v,
                const double *__restrict__ w, const double *__restrict__ visc, double dt,
                double nu, double les, double *__restrict__ us, double *__restrict__ vs,
                double *__restrict__ ws, int k0, int nk)
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
        us[o] = quick_component<0>(g, u, v, w, u, i, j, k, dt, cnu);
    else if (comp == 1)
        vs[o] = quick_component<1>(g, u, v, w, v, i, j, k, dt, cnu);
    else
        ws[o] = quick_component<2>(g, u, v, w, w, i, j, k, dt, cnu);
}

// ---- first-order upwind ------------------------------------------------

template <int C, int A>
__device__ __forceinline__ void upwind_axis(const KGrid &g, const double *phi, long o, int m,
                                            double adv_p, double adv_m, double &fplus,
                                            double &fminus, double &lap)
{
    typedef Axis<C, A> X;
    const long s = X::stride(g);
    const double p0 = phi[o], pp1 = phi[o + s], pm1 = phi[o - s];
    fplus = (adv_p >= 0) ? p0 : pp1;
    fminus = (adv_m >= 0) ? pm1 : p0;
    lap = (pp1 - p0) / X::L(g, m) - (p0 - pm1) / X::L(g, m - 1);
}

template <int C>
__device__ __forceinline__ double upwind_component(const KGrid &g, const double *u,
                                                   const double *v, const double *w,
                                                   const double *phi, int i, int j, int k,
                                                   double dt, double cnu)
{
    const long o = hix(g, i, j, k);
    double ap0, am0, ap1, am1, ap2, am2;
    advecting<C, 0>(g, u, o, ap0, am0);
    advecting<C, 1>(g, v, o, ap1, am1);
    advecting<C, 2>(g, w, o, ap2, am2);
    if (C == 2) {
        // discretisation_upwind.f90:194 averages v(i,j,k) and v(i+1,j,k+1)
        ap1 = 0.5 * (v[o] + v[o + 1 + g.sxy]);
    }
    double fp0, fm0, l0, fp1, fm1, l1, fp2, fm2, l2;
    upwind_axis<C, 0>(g, phi, o, i, ap0, am0, fp0, fm0, l0);
    upwind_axis<C, 1>(g, phi, o, j, ap1, am1, fp1, fm1, l1);
    upwind_axis<C, 2>(g, phi, o, k, ap2, am2, fp2, fm2, l2);
    const double c0 = Axis<C, 0>::Cw(g, i), c1 = Axis<C, 1>::Cw(g, j), c2 = Axis<C, 2>::Cw(g, k);
    const double p = phi[o];
    return p - dt * (ap0 * fp0 - am0 * fm0) / c0
             - dt * (ap1 * fp1 - am1 * fm1) / c1
             - dt * (ap2 * fp2 - am2 * fm2) / c2
             + cnu * dt * (l0) / c0
             + cnu * dt * (l1) / c1
             + cnu * dt * (l2) / c2;
}

// The reference's upwind loops stop one face short of the east (u), north
// (v) and front (w) wall, leaving the value the boundary routine copied
// there.  wall_copy=1 (device-resident step, where the whole-array copies of
// boundary_conditions.f90:305-324 are not made) writes that value, u(nx,j,k)
// etc., explicitly.
__global__ void __launch_bounds__(256)
k_predict_upwind(KGrid g, const double *__restrict__ u, const double *__restrict__ v,
                 const double *__restrict__ w, double dt, double cnu,
                 double *__restrict__ us, double *__restrict__ vs, double *__restrict__ ws,
                 int wall_copy)
{
    int bx, by, bz;
    xcd_band(bx, by, bz);
    const int i = bx * blockDim.x + threadIdx.x + 1;
    const int j = by * blockDim.y + threadIdx.y + 1;
    const int comp = bz % 3;
    const int k = bz / 3 + 1;
    if (i > g.nx || j > g.ny) return;
    const long o = hix(g, i, j, k);
    if (comp == 0) {
        if (i < g.nx) us[o] = upwind_component<0>(g, u, v, w, u, i, j, k, dt, cnu);
        else if (wall_copy) us[o] = u[o];
    } else if (comp == 1) {
        if (j < g.ny) vs[o] = upwind_component<1>(g, u, v, w, v, i, j, k, dt, cnu);
        else if (wall_copy) vs[o] = v[o];
    } else {
        if (k < g.nz) ws[o] = upwind_component<2>(g, u, v, w, w, i, j, k, dt, cnu);
        else if (wall_copy) ws[o] = w[o];
    }
}

void launch_predict_quick(const KGrid &g, const double *u, const double *v, const double *w,
                          const double *visc, double dt, double nu, int les, double *us,
                          double *vs, double *ws, int k0, int k1, hipStream_t st)
{
    if (k1 < k0) return;
    const int nk = k1 - k0 + 1;
    dim3 blk(64, 4), grd((g.nx + 63) / 64, (g.ny + 3) / 4, 3 * nk);
    hipLaunchKernelGGL(k_predict_quick, grd, blk, 0, st, g, u, v, w, visc, dt, nu,
                       (double)les, us, vs, ws, k0, nk);
}

void launch_predict_upwind(const KGrid &g, const double *u, const double *v, const double *w,
                           double dt, double nu, double *us, double *vs, double *ws,
                           int wall_copy, hipStream_t st)
{
    // (nu + nut*Q1) with Q1 = 0 and the module scalar nut = 0
    const double cnu = nu + 0.0 * 0.0;
    dim3 blk(64, 4), grd((g.nx + 63) / 64, (g.ny + 3) / 4, 3 * g.nz);
    hipLaunchKernelGGL(k_predict_upwind, grd, blk, 0, st, g, u, v, w, dt, cnu, us, vs, ws,
                       wall_copy);
}

}  // namespace ntust
