// k_ray.hip -- the airfoil's volume fraction ETA on the GPU: RayCasting
// (airfoil.f90:580-892).
//
// The reference decides, for each cell of the (z, y) plane, whether its
// centre lies inside the airfoil polygon by counting the polygon edges that
// cross the horizontal row through the centre on one side of it (even-odd
// rule, with a slope test taken at the cell's lower corner when the edge's
// z-range straddles the centre), refines the cells that hold a polygon
// vertex into nsub x nsub sub-cells whose inside-fraction becomes ETA there,
// and extrudes the plane along x.  Every decision is a comparison of doubles
// computed with the same operations as the reference and every count is an
// integer, so the result is bit-identical to the CPU restatement.
//
// GPU mapping:
//   k_ray_rows      one workgroup per row j: the edges crossing the row are
//                   gathered in LDS (a few per row), then one thread per
//                   cell k counts them;
//   k_ray_points    one thread per vertex marks the cell that holds it;
//   k_ray_subcells  workgroups over (cell of the polygon's bounding box,
//                   band of 16 sub-rows); cells without a vertex return at
//                   once; the end points of the edges near the band are
//                   staged in LDS and each wave walks them once for its 4
//                   sub-rows, lanes over sub-cells; counts add up per cell;
//   k_ray_fraction  ETA_1 = inside sub-cells / nsub^2 in the vertex cells;
//   k_ray_extrude   ETA(i,j,k) = ETA_1(k,j).
// A gather that overflows its LDS list falls back to testing every edge.
#include "ntust_kernels.h"

namespace ntust {

constexpr int RAY_CAP = 2048;

// The edge from vertex m to m+1 against the row through yrow and the point
// zcen; (yn, zn) is the node the slope test uses.  strict: the sub-cell
// form of the first test (zcen > max instead of >=).  Both orientation
// blocks are tried, as in the reference (both apply to an edge lying in
// the row).
__device__ __forceinline__ int ray_edge_v(double z0, double a0, double z1, double a1,
                                          double yrow, double zcen, double yn, double zn,
                                          bool strict)
{
    const double zmax = z0 > z1 ? z0 : z1, zmin = z0 < z1 ? z0 : z1;
    int c = 0;
#pragma unroll
    for (int blk = 0; blk < 2; blk++) {
        const bool hit = blk == 0 ? (a0 >= yrow && a1 <= yrow) : (a0 <= yrow && a1 >= yrow);
        if (!hit) continue;
        // blk 0: slope from vertex m+1 (P) towards m (Q); blk 1: from m to m+1
        const double yP = blk == 0 ? a1 : a0, zP = blk == 0 ? z1 : z0;
        const double yQ = blk == 0 ? a0 : a1, zQ = blk == 0 ? z0 : z1;
        if (strict ? zcen > zmax : zcen >= zmax) {
        } else if (zcen < zmin) {
            c += 1;
        } else {
            const double m_pa = (yn - yP) / (zn - zP);
            const double m_ab = (yQ - yP) / (zQ - zP);
            if (m_pa > m_ab) c += 1;
        }
    }
    return c;
}

__device__ __forceinline__ int ray_edge(const double *__restrict__ az,
                                        const double *__restrict__ ay, int m, double yrow,
                                        double zcen, double yn, double zn, bool strict)
{
    return ray_edge_v(az[m], ay[m], az[m + 1], ay[m + 1], yrow, zcen, yn, zn, strict);
}

__global__ void __launch_bounds__(256)
k_ray_rows(RayGeom r, double *__restrict__ eta1)
{
    __shared__ int list[RAY_CAP];
    __shared__ int count;
    const int j = blockIdx.x + 1;
    const double yrow = (r.Yn[j + 1] + r.Yn[j + 2]) / 2.0;
    if (threadIdx.x == 0) count = 0;
    __syncthreads();
    for (int m = threadIdx.x; m < r.poly - 1; m += blockDim.x) {
        const double a0 = r.ay[m], a1 = r.ay[m + 1];
        if ((a0 >= yrow && a1 <= yrow) || (a0 <= yrow && a1 >= yrow)) {
            const int s = atomicAdd(&count, 1);
            if (s < RAY_CAP) list[s] = m;
        }
    }
    __syncthreads();
    const int n = count;
    const double yn = r.Yn[j + 1];
    for (int k = threadIdx.x + 1; k <= r.nz; k += blockDim.x) {
        const double zcen = (r.Zn[k + 1] + r.Zn[k + 2]) / 2.0, zn = r.Zn[k + 1];
        int c = 0;
        if (n <= RAY_CAP) {
            for (int q = 0; q < n; q++)
                c += ray_edge(r.az, r.ay, list[q], yrow, zcen, yn, zn, false);
        } else {
            for (int m = 0; m < r.poly - 1; m++)
                c += ray_edge(r.az, r.ay, m, yrow, zcen, yn, zn, false);
        }
        eta1[(k - 1) + (long)r.nz * (j - 1)] = (c % 2 == 0) ? 0.0 : 1.0;
    }
}

__global__ void __launch_bounds__(256)
k_ray_points(RayGeom r, int *__restrict__ points)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= r.poly) return;
    const double z = r.az[m], y = r.ay[m];
    for (int k = r.C; k <= r.E - 1; k++) {
        if (!(z >= r.Zn[k + 1] && z < r.Zn[k + 2])) continue;
        for (int j = r.A; j <= r.B - 1; j++)
            if (y >= r.Yn[j + 1] && y < r.Yn[j + 2])
                points[(k - 1) + (long)r.nz * (j - 1)] = 1;
    }
}

// Sub-cells of the vertex cells.  A workgroup takes RAY_ROWS_WG sub-rows
// of one cell (each wave RAY_ROWS_WAVE of them), stages in LDS the end
// points of the edges whose y-range meets those rows, and walks the edges
// once: for each edge, each of the wave's rows it crosses toggles one bit
// of a per-lane mask (only the parity of the count matters), so the LDS
// reads of an edge serve all the wave's rows.  The inside sub-cells of the
// workgroup are added to the cell's integer count.
constexpr int RAY_ROWS_WAVE = 1, RAY_ROWS_WG = 4 * RAY_ROWS_WAVE;
constexpr int RAY_SUB_CAP = 1536;   // 4 doubles per edge: 48 KB

__global__ void __launch_bounds__(256)
k_ray_subcells(RayGeom r, const int *__restrict__ points, int *__restrict__ counts)
{
    __shared__ double ZZ[RAY_NSUB_MAX + 1], YY[RAY_NSUB_MAX + 1];
    __shared__ double ez0[RAY_SUB_CAP], ey0[RAY_SUB_CAP], ez1[RAY_SUB_CAP], ey1[RAY_SUB_CAP];
    __shared__ int count;
    __shared__ int part[4];
    const int ns = r.nsub, tid = threadIdx.x;
    const int nrb = (ns + RAY_ROWS_WG - 1) / RAY_ROWS_WG;
    const int span = r.E - r.C;
    const int cellbox = (int)(blockIdx.x / nrb), rb = (int)(blockIdx.x % nrb);
    const int k = r.C + cellbox % span;
    const int j = r.A + cellbox / span;
    const long cell = (k - 1) + (long)r.nz * (j - 1);
    if (points[cell] <= 0) return;   // the whole workgroup, before any barrier
    const double zk = r.Zn[k + 1], zk1 = r.Zn[k + 2], yj = r.Yn[j + 1], yj1 = r.Yn[j + 2];
    for (int l = tid; l <= ns; l += blockDim.x) {
        ZZ[l] = zk + (double)l * (zk1 - zk) / (double)ns;
        YY[l] = yj + (double)l * (yj1 - yj) / (double)ns;
    }
    if (tid == 0) count = 0;
    __syncthreads();
    const int l0 = rb * RAY_ROWS_WG, l1 = min(l0 + RAY_ROWS_WG, ns) - 1;
    // an edge can count only where its y-range holds one of the rows'
    // centres; the centres grow with l, so the first and last bound them
    const double ylo = (YY[l0] + YY[l0 + 1]) / 2.0, yhi = (YY[l1] + YY[l1 + 1]) / 2.0;
    for (int m = tid; m < r.poly - 1; m += blockDim.x) {
        const double a0 = r.ay[m], a1 = r.ay[m + 1];
        const double lo = a0 < a1 ? a0 : a1, hi = a0 > a1 ? a0 : a1;
        if (hi >= ylo && lo <= yhi) {
            const int s = atomicAdd(&count, 1);
            if (s < RAY_SUB_CAP) {
                ez0[s] = r.az[m];
                ey0[s] = a0;
                ez1[s] = r.az[m + 1];
                ey1[s] = a1;
            }
        }
    }
    __syncthreads();
    const int n = count;
    const int wave = tid >> 6, lane = tid & 63;
    double yrow[RAY_ROWS_WAVE], yn[RAY_ROWS_WAVE];
    bool live[RAY_ROWS_WAVE];
#pragma unroll
    for (int q = 0; q < RAY_ROWS_WAVE; q++) {
        const int l = l0 + wave * RAY_ROWS_WAVE + q;
        live[q] = l <= l1;
        const int lc = live[q] ? l : l1;
        yrow[q] = (YY[lc] + YY[lc + 1]) / 2.0;
        yn[q] = YY[lc];
    }
    int inside = 0;
    for (int c0 = 0; c0 < ns; c0 += 64) {
        const int col = c0 + lane;
        const bool on = col < ns;
        const int cc = on ? col : ns - 1;
        const double zcen = (ZZ[cc] + ZZ[cc + 1]) / 2.0, zn = ZZ[cc];
        unsigned mask = 0;
        if (n <= RAY_SUB_CAP) {
            for (int e = 0; e < n; e++) {
                const double z0 = ez0[e], a0 = ey0[e], z1 = ez1[e], a1 = ey1[e];
                const double lo = a0 < a1 ? a0 : a1, hi = a0 > a1 ? a0 : a1;
#pragma unroll
                for (int q = 0; q < RAY_ROWS_WAVE; q++)
                    if (live[q] && yrow[q] >= lo && yrow[q] <= hi)
                        mask ^= (unsigned)(ray_edge_v(z0, a0, z1, a1, yrow[q], zcen, yn[q], zn,
                                                      true) & 1) << q;
            }
        } else {
            for (int m = 0; m < r.poly - 1; m++) {
#pragma unroll
                for (int q = 0; q < RAY_ROWS_WAVE; q++)
                    if (live[q])
                        mask ^= (unsigned)(ray_edge(r.az, r.ay, m, yrow[q], zcen, yn[q], zn,
                                                    true) & 1) << q;
            }
        }
        if (on) inside += __popc(mask);
    }
    for (int o = 32; o > 0; o >>= 1) inside += __shfl_xor(inside, o, 64);
    if (lane == 0) part[wave] = inside;
    __syncthreads();
    if (tid == 0) atomicAdd(&counts[cell], part[0] + part[1] + part[2] + part[3]);
}

// ETA_1 of the vertex cells: the inside fraction of their sub-cells
__global__ void __launch_bounds__(256)
k_ray_fraction(RayGeom r, const int *__restrict__ points, const int *__restrict__ counts,
               double *__restrict__ eta1)
{
    const int span = r.E - r.C;
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= span * (r.B - r.A)) return;
    const int k = r.C + b % span, j = r.A + b / span;
    const long cell = (k - 1) + (long)r.nz * (j - 1);
    if (points[cell] > 0) eta1[cell] = (double)counts[cell] / (double)r.nsub / (double)r.nsub;
}

__global__ void __launch_bounds__(256)
k_ray_extrude(KGrid g, const double *__restrict__ eta1, double *__restrict__ eta)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.nx || j > g.ny) return;
    eta[hix(g, i, j, k)] = eta1[(k - 1) + (long)g.nz * (j - 1)];
}

void launch_ray_casting(const KGrid &g, const RayGeom &r, int *points, int *counts,
                        double *eta1, double *eta, hipStream_t st)
{
    hipLaunchKernelGGL(k_ray_rows, dim3(g.ny), dim3(256), 0, st, r, eta1);
    // the box loops run only for a non-empty box (then A, C >= 1)
    if (r.E > r.C && r.B > r.A) {
        hipLaunchKernelGGL(k_ray_points, dim3((r.poly + 255) / 256), dim3(256), 0, st, r,
                           points);
        const int box = (r.E - r.C) * (r.B - r.A);
        const int nrb = (r.nsub + RAY_ROWS_WG - 1) / RAY_ROWS_WG;
        hipLaunchKernelGGL(k_ray_subcells, dim3(box * nrb), dim3(256), 0, st, r, points,
                           counts);
        hipLaunchKernelGGL(k_ray_fraction, dim3((box + 255) / 256), dim3(256), 0, st, r, points,
                           counts, eta1);
    }
    hipLaunchKernelGGL(k_ray_extrude, dim3((g.nx + 63) / 64, (g.ny + 3) / 4, g.nz), dim3(64, 4), 0,
                       st, g, eta1, eta);
}

}  // namespace ntust
