// k_sor_fused.hip -- one red-black SOR sweep per launch (gfx950).
//
// The two-pass red-black solver (k_sor.hip, k_sor_mixed.hip) launches one
// kernel per colour: two launches per sweep, and the other colour's plane
// is read once per launch.  Here a workgroup owns a tile of TX x TY x TZ
// cells and does the whole sweep on it:
//   1. load both colour planes over the tile plus two cells on every side
//      into LDS (the values of the previous sweep);
//   2. update the red cells of the tile plus one cell on every side --
//      the outer layer belongs to the neighbouring tiles, which compute the
//      same values from the same inputs, so the copies agree bit for bit;
//   3. update the black cells of the tile from the red values in LDS;
//   4. write the tile's red and black cells.
// Neighbouring tiles read the previous sweep's values of the cells they do
// not own while their owners write the new ones, so the sweep reads one copy
// of the planes and writes the other (sweep t reads copy (t-1)&1, writes
// copy t&1); the result of ik sweeps is copy ik&1.  Compulsory traffic per
// cell and sweep: read p (both colours) and b, write p -- 24 B in FP64
// (two-pass: 32), 12 B in FP32.
//
// The point update, the gate on the previous sweep's maximum change
// (GS.f90:11) and the per-sweep maxima in buckets are those of the two-pass
// kernels, so the iterates, the stopping sweep and the reported change are
// bit-identical to them.
#include "ntust_kernels.h"

namespace ntust {

// tile of owned cells; LDS holds (TX/2+3) x (TY+4) x (TZ+4) entries per colour
constexpr int FX = 32, FY = 8, FZ = 8;
constexpr int FW = FX / 2 + 3, FHY = FY + 4, FHZ = FZ + 4;
constexpr int FPL = FW * FHY * FHZ;        // entries per colour
constexpr int FTHREADS = 256;
constexpr int FNL = (2 * FPL + FTHREADS - 1) / FTHREADS;              // loads per thread
constexpr int FNR = ((FX / 2 + 1) * (FY + 2) * (FZ + 2) + FTHREADS - 1) / FTHREADS;
constexpr int FNB = ((FX / 2) * FY * FZ + FTHREADS - 1) / FTHREADS;
static_assert(2 * (FX + 4) + 2 * (FY + 4) + 2 * (FZ + 4) <= FTHREADS, "one metric per thread");

template <typename T>
struct SweepMetric {
    const T *iDx, *iDy, *iDz, *rx, *ry, *rz;   // offset by +1
};

__device__ __forceinline__ double dfma(double a, double b, double c) { return fma(a, b, c); }
__device__ __forceinline__ float dfma(float a, float b, float c) { return fmaf(a, b, c); }
__device__ __forceinline__ double dabs(double a) { return fabs(a); }
__device__ __forceinline__ float dabs(float a) { return fabsf(a); }

// MIXED = false: p and b, d = (sum - b)/aP - p   (k_rb_sweep)
// MIXED = true:  e and r, d = (sum + r)/aP - e   (k_rb_sweep32)
template <typename T, bool MIXED>
__global__ void __launch_bounds__(FTHREADS)
k_rb_sweep_fused(KGrid g, SplitGrid s, SweepMetric<T> m, const T *__restrict__ in0,
                 const T *__restrict__ in1, T *__restrict__ out0, T *__restrict__ out1,
                 const T *__restrict__ rhs0, const T *__restrict__ rhs1, T omega, double zeta,
                 double itmax, int t, double *slots, SorControl *ctl, int *dirty)
{
    __shared__ T lp[2 * FPL];
    // metrics of the tile: x from x0-1, y from j0-2, z from k0-2
    __shared__ T mxD[FX + 4], mxR[FX + 4], myD[FY + 4], myR[FY + 4], mzD[FZ + 4], mzR[FZ + 4];
    __shared__ double red[FTHREADS / 64];
    const int tid = threadIdx.x;
    const int blin = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    const int lane = tid & 63;
    double gate = lane < SLOT_BUCKETS ? slots[slot_index(t - 1, lane)] : 0.0;
    const int stopped = ctl->stop;

    const int x0 = blockIdx.x * FX, j0 = blockIdx.y * FY + 1, k0 = blockIdx.z * FZ + 1;
    const int mb = (x0 >> 1) - 1;   // first split index held
    const long pl = s.plane / (g.nz + 2);   // mp*(ny+2)

    // 1. previous sweep's values of both colours into LDS (cells that do not
    //    exist read as 0 and are never used), the metrics, the right-hand sides
    T v[FNL];
#pragma unroll
    for (int r = 0; r < FNL; r++) {
        const int e = tid + r * FTHREADS;
        v[r] = (T)0;
        if (e < 2 * FPL) {
            const int c = e >= FPL;
            const int q = e - c * FPL;
            const int mm = q % FW, row = q / FW;
            const int jj = row % FHY, kk = row / FHY;
            const int j = j0 - 2 + jj, k = k0 - 2 + kk, mi = mb + mm;
            const int i = 2 * mi + ((c + j + k) & 1);
            if (j >= 0 && j <= g.ny + 1 && k >= 0 && k <= g.nz + 1 && i >= 0 && i <= g.nx + 1) {
                const long o = (long)mi + (long)s.mp * j + pl * k;
                v[r] = c ? in1[o] : in0[o];
            }
        }
    }
    T mv = (T)0;
    {
        // thread tid loads one metric value: 2*(FX+4) + 2*(FY+4) + 2*(FZ+4) <= FTHREADS
        int q = tid;
        if (q < 2 * (FX + 4)) {
            const int i = x0 - 1 + (q >> 1);
            if (i <= g.nx + 2) mv = (q & 1) ? m.rx[i] : m.iDx[i];
        } else if ((q -= 2 * (FX + 4)) < 2 * (FY + 4)) {
            const int j = j0 - 2 + (q >> 1);
            if (j <= g.ny + 2) mv = (q & 1) ? m.ry[j] : m.iDy[j];
        } else if ((q -= 2 * (FY + 4)) < 2 * (FZ + 4)) {
            const int k = k0 - 2 + (q >> 1);
            if (k <= g.nz + 2) mv = (q & 1) ? m.rz[k] : m.iDz[k];
        }
    }
    T rb_r[FNR], rb_b[FNB];
#pragma unroll
    for (int r = 0; r < FNR; r++) {
        const int e = tid + r * FTHREADS;
        const int n = e % (FX / 2 + 1), row = e / (FX / 2 + 1);
        const int j = j0 - 1 + row % (FY + 2), k = k0 - 1 + row / (FY + 2);
        const int i = x0 + ((x0 + j + k) & 1) + 2 * n;
        rb_r[r] = (T)0;
        if (row < (FY + 2) * (FZ + 2) && i >= 1 && i <= g.nx && j >= 1 && j <= g.ny && k >= 1 &&
            k <= g.nz)
            rb_r[r] = rhs0[(long)(i >> 1) + (long)s.mp * j + pl * k];
    }
#pragma unroll
    for (int r = 0; r < FNB; r++) {
        const int e = tid + r * FTHREADS;
        const int n = e % (FX / 2), row = e / (FX / 2);
        const int j = j0 + row % FY, k = k0 + row / FY;
        const int i = x0 + 1 + ((x0 + j + k) & 1) + 2 * n;
        rb_b[r] = (T)0;
        if (row < FY * FZ && i <= g.nx && j <= g.ny && k <= g.nz)
            rb_b[r] = rhs1[(long)(i >> 1) + (long)s.mp * j + pl * k];
    }
#pragma unroll
    for (int r = 0; r < FNL; r++) {
        const int e = tid + r * FTHREADS;
        if (e < 2 * FPL) lp[e] = v[r];
    }
    {
        int q = tid;
        if (q < 2 * (FX + 4)) (q & 1 ? mxR : mxD)[q >> 1] = mv;
        else if ((q -= 2 * (FX + 4)) < 2 * (FY + 4)) (q & 1 ? myR : myD)[q >> 1] = mv;
        else if ((q -= 2 * (FY + 4)) < 2 * (FZ + 4)) (q & 1 ? mzR : mzD)[q >> 1] = mv;
    }
    // every wavefront reads the same buckets: the decision is uniform over
    // the block, so returning here cannot split a barrier
    gate = wave_max(gate);
    const bool active = !stopped && ((double)(t - 1) < itmax) && (gate > zeta);
    if (!active) {
        if (blin == 0 && tid == 0) ctl->stop = 1;
        return;
    }
    __syncthreads();

    auto at = [&](int c, int i, int j, int k) {
        return c * FPL + ((i >> 1) - mb) + FW * ((j - j0 + 2) + FHY * (k - k0 + 2));
    };
    // the point update of colour c at (i,j,k); returns the new value
    auto update = [&](int c, int i, int j, int k, T rhs, T &change) {
        const int oc = c ^ 1;
        const T pW = lp[at(oc, i - 1, j, k)], pE = lp[at(oc, i + 1, j, k)];
        const T pS = lp[at(oc, i, j - 1, k)], pN = lp[at(oc, i, j + 1, k)];
        const T pB = lp[at(oc, i, j, k - 1)], pF = lp[at(oc, i, j, k + 1)];
        const T po = lp[at(c, i, j, k)];
        const int xi = i - x0 + 1, yj = j - j0 + 2, zk = k - k0 + 2;
        const T ax = mxD[xi], ay = myD[yj], az = mzD[zk];
        const T rE = mxR[xi], rW = mxR[xi - 1], rN = myR[yj], rS = myR[yj - 1];
        const T rF = mzR[zk], rB = mzR[zk - 1];
        const T ayz = ay * az, axz = ax * az, axy = ax * ay;
        const T t1 = dfma(rE, pE, rW * pW);
        const T t2 = dfma(rN, pN, rS * pS);
        const T t3 = dfma(rF, pF, rB * pB);
        const T sum = dfma(ayz, t1, dfma(axz, t2, axy * t3));
        const T aP = dfma(ayz, rE + rW, dfma(axz, rN + rS, axy * (rF + rB)));
        const T d = (MIXED ? (sum + rhs) : (sum - rhs)) / aP - po;
        change = dabs(d);
        return dfma(omega, d, po);
    };

    // 2. red: tile plus one layer.  Only black values are read here, so all
    //    red values are computed before any is written back.
    T pcm = (T)0;
    T nr[FNR];
#pragma unroll
    for (int r = 0; r < FNR; r++) {
        const int e = tid + r * FTHREADS;
        const int n = e % (FX / 2 + 1), row = e / (FX / 2 + 1);
        const int j = j0 - 1 + row % (FY + 2), k = k0 - 1 + row / (FY + 2);
        const int i = x0 + ((x0 + j + k) & 1) + 2 * n;
        nr[r] = (T)0;
        if (row < (FY + 2) * (FZ + 2) && i >= 1 && i <= g.nx && j >= 1 && j <= g.ny && k >= 1 &&
            k <= g.nz) {
            T ch = (T)0;
            nr[r] = update(0, i, j, k, rb_r[r], ch);
            if (ch > pcm) pcm = ch;   // a NaN change never raises the max (GS.f90:61)
        }
    }
#pragma unroll
    for (int r = 0; r < FNR; r++) {
        const int e = tid + r * FTHREADS;
        const int n = e % (FX / 2 + 1), row = e / (FX / 2 + 1);
        const int j = j0 - 1 + row % (FY + 2), k = k0 - 1 + row / (FY + 2);
        const int i = x0 + ((x0 + j + k) & 1) + 2 * n;
        if (row < (FY + 2) * (FZ + 2) && i >= 1 && i <= g.nx && j >= 1 && j <= g.ny && k >= 1 &&
            k <= g.nz) {
            lp[at(0, i, j, k)] = nr[r];
            if (i > x0 && i <= x0 + FX && j >= j0 && j < j0 + FY && k >= k0 && k < k0 + FZ)
                out0[(long)(i >> 1) + (long)s.mp * j + pl * k] = nr[r];
        }
    }
    __syncthreads();
    // 3. black: the tile
#pragma unroll
    for (int r = 0; r < FNB; r++) {
        const int e = tid + r * FTHREADS;
        const int n = e % (FX / 2), row = e / (FX / 2);
        const int j = j0 + row % FY, k = k0 + row / FY;
        const int i = x0 + 1 + ((x0 + j + k) & 1) + 2 * n;
        if (row < FY * FZ && i <= g.nx && j <= g.ny && k <= g.nz) {
            T ch = (T)0;
            const T nv = update(1, i, j, k, rb_b[r], ch);
            if (ch > pcm) pcm = ch;
            out1[(long)(i >> 1) + (long)s.mp * j + pl * k] = nv;
        }
    }
    // 4. the sweep's maximum change
    double pm = wave_max((double)pcm);
    if (lane == 0) red[tid >> 6] = pm;
    __syncthreads();
    if (tid == 0) {
        double mm = red[0];
        for (int w = 1; w < FTHREADS / 64; w++) mm = red[w] > mm ? red[w] : mm;
        double *bk = &slots[slot_index(t, blin % SLOT_BUCKETS)];
        if (mm > 0.0 && mm > *(volatile double *)bk)
            atomicMax((unsigned long long *)bk, (unsigned long long)__double_as_longlong(mm));
        if (blin == 0) {
            ctl->ik = t;
            if (MIXED) *dirty = 1;
        }
    }
}

static dim3 fused_grid(const KGrid &g)
{
    return dim3((g.nx + FX - 1) / FX, (g.ny + FY - 1) / FY, (g.nz + FZ - 1) / FZ);
}

void launch_rb_sweep_fused(const KGrid &g, const SplitGrid &s, const double *rx,
                           const double *ry, const double *rz, const double *in0,
                           const double *in1, double *out0, double *out1, const double *b0,
                           const double *b1, double omega, double zeta, double itmax, int t,
                           double *slots, SorControl *ctl, hipStream_t st)
{
    SweepMetric<double> m{g.iDx, g.iDy, g.iDz, rx, ry, rz};
    hipLaunchKernelGGL((k_rb_sweep_fused<double, false>), fused_grid(g), dim3(FTHREADS), 0, st,
                       g, s, m, in0, in1, out0, out1, b0, b1, omega, zeta, itmax, t, slots, ctl,
                       (int *)nullptr);
}

void launch_rb_sweep32_fused(const KGrid &g, const SplitGrid &s, const FMetric &fm,
                             const float *in0, const float *in1, float *out0, float *out1,
                             const float *r0, const float *r1, float omega, double zeta,
                             double itmax, int t, double *slots, SorControl *ctl, int *dirty,
                             hipStream_t st)
{
    SweepMetric<float> m{fm.iDx, fm.iDy, fm.iDz, fm.rx, fm.ry, fm.rz};
    hipLaunchKernelGGL((k_rb_sweep_fused<float, true>), fused_grid(g), dim3(FTHREADS), 0, st, g,
                       s, m, in0, in1, out0, out1, r0, r1, omega, zeta, itmax, t, slots, ctl,
                       dirty);
}

// p0 += e on the interior and e = 0, for the copy of e the last sweep wrote
// (copy ik&1); skipped when no sweep of the segment ran
__global__ void __launch_bounds__(256)
k_rb_absorb_pp(KGrid g, SplitGrid s, double *__restrict__ p0, double *__restrict__ p1,
               float *ea0, float *ea1, float *eb0, float *eb1, const int *dirty,
               const SorControl *ctl)
{
    if (!*dirty) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.nx || j > g.ny) return;
    const bool second = ctl->ik & 1;
    float *e0 = second ? eb0 : ea0, *e1 = second ? eb1 : ea1;
    const long o = six(s, i, j, k);
    if ((i + j + k) & 1) {
        p1[o] = p1[o] + (double)e1[o];
        e1[o] = 0.0f;
    } else {
        p0[o] = p0[o] + (double)e0[o];
        e0[o] = 0.0f;
    }
}

void launch_rb_absorb_pp(const KGrid &g, const SplitGrid &s, double *p0, double *p1,
                         float *ea0, float *ea1, float *eb0, float *eb1, const int *dirty,
                         const SorControl *ctl, hipStream_t st)
{
    dim3 blk(64, 4), grd((g.nx + 63) / 64, (g.ny + 3) / 4, g.nz);
    hipLaunchKernelGGL(k_rb_absorb_pp, grd, blk, 0, st, g, s, p0, p1, ea0, ea1, eb0, eb1, dirty,
                       ctl);
}

}  // namespace ntust
