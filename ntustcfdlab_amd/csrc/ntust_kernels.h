// ntust_kernels.h -- shared definitions of the gfx950 kernels.
//
// Storage follows the Fortran module arrays of the solver (column-major,
// x fastest, one 64-bit word per value):
//   halo arrays (-1:nx+2,-1:ny+2,-1:nz+2): (i,j,k) at (i+1)+sx*((j+1)+sy*(k+1))
//   interior    (1:nx,1:ny,1:nz):          (i,j,k) at (i-1)+nx*((j-1)+ny*(k-1))
//   metrics     (-1:n+2): the device pointers below are offset by one
//               element, so m[i] addresses element i for i = -1 .. n+2.
//
// Every exact-mode kernel is compiled with -ffp-contract=off and writes its
// expressions in the reference's operation order, so its results equal the
// CPU restatement bit for bit.  Fast-mode kernels spell out every fused
// multiply-add with fma().
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace ntust {

struct KGrid {
    int nx, ny, nz;
    int sx, sy;          // halo strides: sx = nx+4, sy = ny+4
    long sxy;            // sx*sy
    const double *iDx, *Dxs, *iDy, *Dys, *iDz, *Dzs;   // offset by +1
};

__host__ __device__ inline long hix(const KGrid &g, int i, int j, int k)
{
    return (long)(i + 1) + (long)g.sx * (long)(j + 1) + g.sxy * (long)(k + 1);
}

__host__ __device__ inline long iix(const KGrid &g, int i, int j, int k)
{
    return (long)(i - 1) + (long)g.nx * ((long)(j - 1) + (long)g.ny * (long)(k - 1));
}

// Colour-split pressure storage of the red-black solver.  Cell (i,j,k),
// i in 0..nx+1, j in 0..ny+1, k in 0..nz+1, has colour (i+j+k)&1 and lives
// in plane[colour] at (i>>1) + mp*(j + (ny+2)*k).
struct SplitGrid {
    int mp;              // padded half-row length >= (nx+3)/2
    int ry;              // ny+2
    long plane;          // mp*(ny+2)*(nz+2)
};

__host__ __device__ inline long six(const SplitGrid &s, int i, int j, int k)
{
    return (long)(i >> 1) + (long)s.mp * ((long)j + (long)s.ry * (long)k);
}

// Loads of values that earlier launches wrote with atomics or as flags (the
// per-sweep maxima, the stop and dirty flags).  System scope: the load
// bypasses the XCD's L2, so it does not depend on that L2 having been
// invalidated since a stale copy of the line was cached there (a memset
// before the solve, or the filter read of a bucket).
__device__ __forceinline__ double load_flag(const double *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ int load_flag(const int *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// max over a 64-wide wavefront (non-negative, non-NaN inputs)
__device__ __forceinline__ double wave_max(double v)
{
    for (int o = 32; o > 0; o >>= 1) {
        double t = __shfl_xor(v, o, 64);
        v = t > v ? t : v;
    }
    return v;
}

// XCD-aware block order for the stencil kernels.  The hardware hands the
// workgroups of a launch to the 8 XCDs round-robin (workgroup n to XCD
// n mod 8, in order), and each XCD has its own L2.  Renumbering so that the
// workgroups of XCD x take a contiguous range of logical blocks (x fastest,
// then y, then z) gives each XCD a slab of the domain, whose halo rows its
// L2 then serves.  A bijection for any block count; only the speed depends
// on the dispatch order.
__device__ __forceinline__ void xcd_block(int &bx, int &by, int &bz)
{
    const int nb = gridDim.x * gridDim.y * gridDim.z;
    const int n = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    const int x = n & 7, idx = n >> 3, per = nb >> 3, rem = nb & 7;
    const int logical = x * per + (x < rem ? x : rem) + idx;
    bx = logical % gridDim.x;
    const int r = logical / gridDim.x;
    by = r % gridDim.y;
    bz = r / gridDim.y;
}

// XCD y-band order for kernels whose blocks each cover one plane of one or
// more full-size arrays (a slab of planes per XCD would not fit its 4 MB
// L2).  Within each z-layer of the launch the blocks are renumbered so that
// XCD x takes a contiguous band of them (x fastest, then y); all XCDs
// advance through z together, so an XCD's +-y, +-x and +-z neighbours stay
// in its own L2 while its working set is a band of a few planes.  The
// z-layer keeps its index.  A bijection for any grid; the band lands on one
// XCD in every layer when gridDim.x*gridDim.y is a multiple of 8.
__device__ __forceinline__ void xcd_band(int &bx, int &by, int &bz)
{
    const int nb = gridDim.x * gridDim.y;
    const int m = blockIdx.x + gridDim.x * blockIdx.y;
    const int x = m & 7, idx = m >> 3, per = nb >> 3, rem = nb & 7;
    const int logical = x * per + (x < rem ? x : rem) + idx;
    bx = logical % gridDim.x;
    by = logical / gridDim.x;
    bz = blockIdx.z;
}

// Boundary-condition codes, one per (face, component) in the order
//   W_u W_v W_w  E_u E_v E_w  S_u S_v S_w  N_u N_v N_w  B_u B_v B_w  F_u F_v F_w
enum BcCode { BC_NOSLIP = 0, BC_NEUMANN = 1, BC_DIRICHLET = 2 };
struct BcTable { int c[18]; };

// ---- launchers (one translation unit per kernel family) ----------------

// LES: gradient of a flat (nx,ny,nz) view and the fused eddy viscosity
void launch_gradient_flat(const KGrid &g, const double *phi, const double *prebase,
                          double *dphi, const double *Dxm, const double *Dym,
                          const double *Dzm, int k0, int k1, hipStream_t st);
void launch_smagorinsky(const KGrid &g, const double *u, const double *v, const double *w,
                        const double *pre_u, const double *pre_v, const double *pre_w,
                        const double *filt, double *visc, double *dudx, double *dvdx,
                        double *dwdx, int k0, int k1, hipStream_t st);

// fast arithmetic modes of the predictor and the eddy viscosity (k_fast.hip).
// QCoef: the QUICK face between nodes m and m+1 of one axis -- coefficients
// of phi(m-1), phi(m), phi(m+1) for a positive advecting velocity (am1, a0,
// ap1), of phi(m), phi(m+1), phi(m+2) otherwise (b0, bp1, bp2) -- and the
// reciprocal node spacing 1/L(m) and control width 1/C(m).  QTables: per
// axis, kind 0 the component's own axis (L(m) = iD(m+1), C(m) = D_s(m)),
// kind 1 a transverse one (L(m) = D_s(m), C(m) = iD(m)); entry m+1 holds
// face m.  RMetric: reciprocal metrics, offset by +1 like KGrid.
template <typename T>
struct QCoef { T am1, a0, ap1, b0, bp1, bp2, rL, rC; };
template <typename T>
struct QTables { const QCoef<T> *t[3][2]; };
template <typename T>
struct RMetric { const T *riDx, *rDxs, *riDy, *rDys, *riDz, *rDzs; };
template <typename T>
void launch_predict_quick_march(const KGrid &g, const QTables<T> &q, const double *u,
                                const double *v, const double *w, const double *visc, double dt,
                                double nu, int les, double *us, double *vs, double *ws, int k0,
                                int k1, hipStream_t st);
template <typename T>
void launch_smagorinsky_fast(const KGrid &g, const RMetric<T> &r, const double *u,
                             const double *v, const double *w, const double *filt,
                             double *visc, int k0, int k1, hipStream_t st);

// predictor
void launch_predict_quick(const KGrid &g, const double *u, const double *v, const double *w,
                          const double *visc, double dt, double nu, int les, double *us,
                          double *vs, double *ws, int k0, int k1, hipStream_t st);
void launch_predict_upwind(const KGrid &g, const double *u, const double *v, const double *w,
                           double dt, double nu, double *us, double *vs, double *ws,
                           int wall_copy, hipStream_t st);

// pressure: lexicographic SOR (exact) and red-black SOR (fast)
//
// Per-sweep maxima of |pNew - p|: sweep t owns SLOT_BUCKETS buckets, each
// on its own 128-byte line; the sweep's maximum is the largest bucket.
// Sweep 0 is the state before the first sweep: bucket 0 holds 1.0 (GS.f90:9).
constexpr int SLOT_BUCKETS = 32;
constexpr int SLOT_STRIDE = 16;   // doubles
__host__ __device__ inline long slot_index(long t, int b)
{
    return (t * SLOT_BUCKETS + b) * SLOT_STRIDE;
}
// cells per thread along k in the red-black sweep kernels: FP32 and the
// mixed solver's kernels, and the FP64 sweep
constexpr int RB_KPT = 4;
constexpr int RB_KPT64 = 8;
// FP32 copies of the metrics used by the FP32 sweeps (offset by +1 like KGrid)
struct FMetric {
    const float *iDx, *iDy, *iDz, *rx, *ry, *rz;
};
struct SorControl {
    int stop;            // set by the first sweep that does not run
    int ik;              // sweeps run
};
// buckets of sweep 0 (1.0 in bucket 0, GS.f90:9) and sweep 1 (0), control
// cleared; every running sweep t clears the buckets of sweep t+1
void launch_sor_init(double *slots, SorControl *ctl, hipStream_t st);
__device__ __forceinline__ void clear_next_buckets(double *slots, int t, int tid)
{
    if (tid < SLOT_BUCKETS) slots[slot_index(t + 1, tid)] = 0.0;
}
// GS.f90's lexicographic SOR pipelined over sweeps (k_sor_lex.hip): sweeps
// t0..t1 on the colour planes, the launches tau = 3+2*t0 .. nx+ny+nz+2*t1;
// flags[t] = epoch for every sweep t < t1 that changed some cell by more
// than zeta, the maximum change of sweep t1 in its buckets (zeroed by the
// caller)
void launch_lex_sweeps(const KGrid &g, const SplitGrid &s, double *p0, double *p1,
                       const double *b0, const double *b1, double omega, double zeta, int t0,
                       int t1, int *flags, int epoch, double *slots, hipStream_t st);
long lex_launches(const KGrid &g, long t0, long t1);
// members of the Krylov pressure path
void launch_poisson_coeffs(const KGrid &g, double *coef, hipStream_t st);
void launch_poisson_rhs(const KGrid &g, const double *us, const double *vs, const double *ws,
                        double dt, double *b, hipStream_t st);
void launch_pressure_field(const KGrid &g, const double *x, double *p, double *pre,
                           hipStream_t st);
void launch_rb_rhs(const KGrid &g, const SplitGrid &s, const double *us, const double *vs,
                   const double *ws, double dt, double *b0, double *b1, hipStream_t st);
void launch_rb_gather(const KGrid &g, const SplitGrid &s, const double *p, double *p0,
                      double *p1, hipStream_t st);
void launch_rb_scatter(const KGrid &g, const SplitGrid &s, const double *p0, const double *p1,
                       double *p, double *pre, hipStream_t st);
void launch_rb_sweep(const KGrid &g, const SplitGrid &s, const double *rx, const double *ry,
                     const double *rz, double *p0, double *p1, const double *b0,
                     const double *b1, double omega, double zeta, double itmax, int t,
                     double *slots, SorControl *ctl, hipStream_t st);
// sweep t (from 1); first: the first sweep of a segment
// (clears the segment's dirty flag when it does not run)
void launch_rb_sweep32(const KGrid &g, const SplitGrid &s, const FMetric &fm, float *e0,
                       float *e1, const float *r0, const float *r1, float omega, double zeta,
                       double itmax, int t, double *slots, SorControl *ctl, int *dirty,
                       int first, hipStream_t st);
// speculative sweeps (no gate): sweep t runs and records its maximum in
// the buckets of slot t, which launch_slots_clear has zeroed -- or, with
// flags non-null, only whether it exceeds zeta (flags[t] = 1, flags zeroed
// by the caller); the host reads them and replays from a checkpoint if the
// reference's loop would have stopped earlier
void launch_rb_sweep_spec(const KGrid &g, const SplitGrid &s, const double *rx, const double *ry,
                          const double *rz, double *p0, double *p1, const double *b0,
                          const double *b1, double omega, int t, double *slots, double zeta,
                          int *flags, hipStream_t st);
void launch_rb_sweep32_spec(const KGrid &g, const SplitGrid &s, const FMetric &fm, float *e0,
                            float *e1, const float *r0, const float *r1, float omega, int t,
                            double *slots, double zeta, int *flags, hipStream_t st);
void launch_slots_clear(double *slots, int t0, int n, hipStream_t st);
// the residual for the shifted colour-1 sweeps: r1 holds s1 = r1/aP1,
// e1_shift = -s1, r0 gains S(s1)
void launch_rb_residual(const KGrid &g, const SplitGrid &s, const double *rx, const double *ry,
                        const double *rz, const double *p0, const double *p1, const double *b0,
                        const double *b1, float *r0, float *r1, float *e1_shift,
                        const int *dirty, hipStream_t st);
// absorb followed by the residual, in two colour launches (same bits),
// for the shifted colour-1 sweeps
void launch_rb_correct(const KGrid &g, const SplitGrid &s, const double *rx, const double *ry,
                       const double *rz, double *p0, double *p1, float *e0, float *e1,
                       const double *b0, const double *b1, float *r0, float *r1,
                       const int *dirty, hipStream_t st);

// projection, direct forcing, centring, update
void launch_project(const KGrid &g, const double *us, const double *vs, const double *ws,
                    const double *p, const double *eta, double usol, double vsol,
                    double wsol, double dt, double *u1, double *v1, double *w1,
                    double *u2, double *v2, double *w2, double *FX, double *FY,
                    double *FZ, double *uc, double *vc, double *wc, double *uout,
                    double *vout, double *wout, const double *u2g, const double *v2g,
                    const double *w2g, int k0, int k1, hipStream_t st);
void launch_copy_interior(const KGrid &g, const double *a, const double *b, const double *c,
                          double *x, double *y, double *z, int k0, int k1, hipStream_t st);
void launch_pack_interior(const KGrid &g, const double *a, double *x, hipStream_t st);
// zeroes the entries of the line sums that the force kernels read and do not
// write (rows 0 and ny+1 of FXz, FYz; entries 0 and nx+1 of FXy, FYy)
void launch_force_clear(const KGrid &g, double *FXz, double *FYz, double *FXy, double *FYy,
                        hipStream_t st);
void launch_force_integral(const KGrid &g, const double *FY, const double *FZ, double lx,
                           double *FXz, double *FYz, double *FXy, double *FYy,
                           double *coef_out, hipStream_t st);

// boundary conditions
void launch_boundary_conditions(const KGrid &g, const BcTable &bc, double *u, double *v,
                                double *w, double *p, hipStream_t st);
void launch_copy_full(long n, const double *src, double *dst0, double *dst1, double *dst2,
                      hipStream_t st);
void launch_copy_shell(const KGrid &g, const double *u, const double *v, const double *w,
                       double *a0, double *a1, double *a2, double *b0, double *b1,
                       double *b2, hipStream_t st);

// weight of a cell's FZ (FY) in C_D (C_L) with the step's skip pattern:
// virtualForceIntegrator.f90:22-81 with the line sums fz = FXz (FYz), fy =
// FXy (FYy); a line sample that is exactly zero is skipped, so only the
// masks of the line sums enter
__device__ __forceinline__ double nz_mask(const double *a, long s)
{
    return a[s] != 0.0 ? 1.0 : 0.0;
}

__device__ __forceinline__ double force_weight(const KGrid &g, const double *fz, const double *fy,
                                               int i, int j, int k, double lx)
{
    const long px = g.nx + 2;
    double ax = 4.0 * nz_mask(fy, i);
    if (i > 1) ax += nz_mask(fy, i - 1);
    if (i < g.nx) ax += nz_mask(fy, i + 1);
    double ay = 4.0 * nz_mask(fz, i + px * j);
    if (j > 1) ay += nz_mask(fz, i + px * (j - 1));
    if (j < g.ny) ay += nz_mask(fz, i + px * (j + 1));
    const double az = 4.0 + (k > 1 ? 1.0 : 0.0) + (k < g.nz ? 1.0 : 0.0);
    return (-2.0 / lx) * (g.Dxs[i] / 6.0 * ax) * (g.Dys[j] / 6.0 * ay) * (g.Dzs[k] / 6.0 * az);
}

// ---- Lagrangian markers (k_marker.hip) -----------------------------------
//
// Markers on a curve in the (z, y) plane, extruded along x: marker (m, ix)
// sits at (xc(ix), Y_m, Z_m), ix = 1..nx, with weight W = wm[m] * iDx(ix).
// The discrete delta of component a at its staggered point c is
//   w = (phi(dx/hx)/hx * phi(dy/hy)/hy) * phi(dz/hz)/hz,  d = x_c - X_m,
// phi the cubic B-spline (support |r| < 2; weights over unit-spaced points
// sum to 1, first moment 0, twice continuously differentiable, so the
// forcing is twice differentiable in the marker positions).  Each marker
// reads the points of a 5 (x) x 4 (y) x 4 (z) window, the y and z windows
// starting at jw[a][m], kw[a][m]; the points outside the interior
// are left out.  Spreading gathers, for each interior point of the
// spread lists, the markers whose windows hold it (CSR per component), so
// interpolation I u = sum_c w u_c dv_c and spreading S F = sum_m w W F_m
// are exact transposes in the dv- and W-weighted inner products.
constexpr int MK_WIN = 4;          // window points per axis in y and z
constexpr double MK_RADIUS = 2.0;  // support of the kernel, in kernel widths

__host__ __device__ inline double mk_phi(double r)
{
    const double a = fabs(r);
    if (a <= 1.0) return (4.0 - 6.0 * (a * a) + 3.0 * (a * a * a)) / 6.0;
    if (a < 2.0) {
        const double b = 2.0 - a;
        return (b * b * b) / 6.0;
    }
    return 0.0;
}

// d phi / d r
__host__ __device__ inline double mk_dphi(double r)
{
    const double a = fabs(r);
    double d = 0.0;
    if (a <= 1.0) {
        d = (-12.0 * a + 9.0 * (a * a)) / 6.0;
    } else if (a < 2.0) {
        const double b = 2.0 - a;
        d = -(b * b) / 2.0;
    }
    return r < 0.0 ? -d : d;
}

struct AdjForce;

struct MarkerSet {
    int n;                                  // markers in the (z, y) plane
    double hx, hy, hz;                      // kernel widths
    const double *zm, *ym, *wm;             // positions, weights per unit span
    const int *jw[3], *kw[3];               // window starts per component
    const double *Xn, *Yn, *Zn;             // node coordinates, element i at [i+1]
    // spread lists per component: entry e covers the points (i, sj[e], sk[e])
    // for all i, with markers sm[so[e] .. so[e+1]-1] in increasing order
    int ne[3];
    const int *sj[3], *sk[3], *so[3], *sm[3];
};

// U[(a*n + m)*nx + ix-1] = (I u)_a at marker (m, ix) from the velocity u, v, w
void launch_marker_interp(const KGrid &g, const MarkerSet &M, const double *u, const double *v,
                          const double *w, double *U, hipStream_t st);
// forcing F = -U/dt spread and applied: u += dt S F; FY, FZ (may be null)
// get the spread force of v and w on the listed points
void launch_marker_force(const KGrid &g, const MarkerSet &M, const double *U, double dt,
                         double *u, double *v, double *w, double *FY, double *FZ,
                         hipStream_t st);
// adjoint of the forcing of one step: q = lambda + dJ/dF / dt on the listed
// points (the force weights of F, as in the transposed projection);
// lambda -= (S I)^T q; with u1 (the step's post-projection velocity, may be
// null: no marker gradient) gz, gy += dJ/dZ_m, dJ/dY_m
void launch_marker_adjoint(const KGrid &g, const MarkerSet &M, const AdjForce &F, double dt,
                           double *lu, double *lv, double *lw, const double *u1,
                           const double *v1, const double *w1, double *L, double *GP,
                           double *gz, double *gy, hipStream_t st);

// discrete adjoint of the resident step (transposed stages, reverse order)
void launch_adj_boundary_conditions(const KGrid &g, const BcTable &bc, double *lu, double *lv,
                                    double *lw, double *lp, hipStream_t st);
// objective terms of the transposed projection of one step: dJ/dC_D and
// dJ/dC_L of the step with its force-integral line sums (FXz, FXy, FYz,
// FYy; unused when the scale is 0), and the ETA gradient accumulator with
// the step's u1, v1, w1 (geta null: none)
struct AdjForce {
    double sd = 0.0, sl = 0.0;
    const double *fxz = nullptr, *fxy = nullptr, *fyz = nullptr, *fyy = nullptr;
    double lx = 1.0;
    const double *u1 = nullptr, *v1 = nullptr, *w1 = nullptr;
    double *geta = nullptr;
};
void launch_adj_project(const KGrid &g, const double *lu, const double *lv, const double *lw,
                        const double *eta, double dt, const AdjForce &F, double *lus,
                        double *lvs, double *lws, double *lp, double *lu_clear,
                        double *lv_clear, double *lw_clear, hipStream_t st);
// FP32 transposed sweeps for the adjoint of the FP32-sweep solver: the
// correction e, its residual r and running sum Ef per colour (FP32 colour
// planes, ghost and padding entries zero), a zero FP64 plane for b = 0, and
// the segment length in half-sweeps between FP64 folds
struct AdjMixed {
    FMetric fm;
    float *e0, *e1, *r0, *r1, *f0;   // f0: the red running sum
    const double *bzero;
    long seg1, seg;   // half-sweeps of the first segment and of the later ones (even)
};
// mx null: the FP64 transposed sweeps
void launch_adj_pressure(const KGrid &g, const SplitGrid &s, const double *rx, const double *ry,
                         const double *rz, double omega, int sweeps, double *lp, double *lb,
                         double *z0, double *z1, double *Z0, double *Z1, const AdjMixed *mx,
                         hipStream_t st);
void launch_adj_divergence(const KGrid &g, const double *lb, double dt, double *lus,
                           double *lvs, double *lws, double *lu, double *lv, double *lw,
                           hipStream_t st);
void launch_adj_predict(const KGrid &g, int upwind, const double *u, const double *v,
                        const double *w, const double *visc, double dt, double nu, int les,
                        const double *lus, const double *lvs, const double *lws, double *lu,
                        double *lv, double *lw, double *lvisc, hipStream_t st);
void launch_adj_smagorinsky(const KGrid &g, const double *u, const double *v, const double *w,
                            const double *filt, const double *lvisc, double *lu, double *lv,
                            double *lw, hipStream_t st);
// airfoil volume fraction (k_ray.hip, airfoil.f90:580-892).  Yn, Zn: node
// coordinates with element j at Yn[j+1] (j = -1..n+3); az, ay: the rotated
// and placed polygon; A..B-1 (y) and C..E-1 (z): the cells of its bounding
// box, as the reference's bound search finds them.
constexpr int RAY_NSUB_MAX = 256;
struct RayGeom {
    int nx, ny, nz, poly;
    int A, B, C, E;
    int nsub;
    const double *Yn, *Zn, *az, *ay;
};
void launch_ray_casting(const KGrid &g, const RayGeom &r, int *points, int *counts,
                        double *eta1, double *eta, hipStream_t st);

// Krylov path (k_krylov.hip, solver.f90).  coef: the n x 7 DIA columns;
// col: jcoef(2..4); m: jcoef entries used (1..4).
constexpr int KRY_MAX_BLOCKS = 1024;
struct BicgState {
    double rho1, rho2, alpha, omega, norm0, norm, eps;
    int j, itmax, stop;
};
struct BicgVectors {
    double *p, *r, *r2, *v, *ss, *t;
    double *part;   // 2 * KRY_MAX_BLOCKS partial sums
};
void launch_dia_matvec(const double *coef, const long *col, int m, long n, const double *y,
                       double *x, hipStream_t st);
void launch_bicg_setup(const double *coef, const long *col, int m, long n, const double *b,
                       const double *x, BicgVectors &w, BicgState *st, hipStream_t s);
void launch_bicg_iteration(const double *coef, const long *col, int m, long n, double *x,
                           BicgVectors &w, BicgState *st, hipStream_t s);

// output and restart (k_output.hip)
void launch_pack_q(const KGrid &g, const double *pre, const double *uc, const double *vc,
                   const double *wc, const double *eta, float *q, hipStream_t st);
void launch_steady(const KGrid &g, const double *wc, double *last, int compare,
                   unsigned long long *vmax, hipStream_t st);
// plain copy of words16 16-byte words (counter calibration)
void launch_stream_copy(const void *src, void *dst, long words16, hipStream_t st);
// uc, vc, wc = the cell-centred velocity of u, v, w (calcul_new_velocity.f90:171-181)
void launch_centre(const KGrid &g, const double *u, const double *v, const double *w, double *uc,
                   double *vc, double *wc, hipStream_t st);
void launch_unpack_q(const KGrid &g, const float *q, double *u, double *v, double *w, double *p,
                     double *pre, double *uc, double *vc, double *wc, hipStream_t st);

// kinetic energy 0.5 sum (u^2+v^2+w^2) iDx iDy iDz over the interior: block
// partial sums into part[], returns their count
int launch_kinetic_energy(const KGrid &g, const double *u, const double *v, const double *w,
                          double *part, hipStream_t st);
void launch_ke_gradient(const KGrid &g, const double *u, const double *v, const double *w,
                        double scale, double *lu, double *lv, double *lw, hipStream_t st);
// X += a x on the interior of three halo arrays
void launch_axpy_interior(const KGrid &g, double a, const double *x, const double *y,
                          const double *z, double *X, double *Y, double *Z, hipStream_t st);

}  // namespace ntust
