// gfx950 kernels for compact derivatives and filters along one direction.
//
// One workgroup owns a tile of L lines.  Phase 1 stages the tile (plus the
// periodic / transformed-periodic halo) into LDS with coalesced loads: for
// x-lines consecutive lanes walk along the line, for y/z-lines consecutive
// lanes walk across neighbouring lines (consecutive x), so every global access
// is unit-stride whatever the direction.  Phase 2 gives each line P lanes
// (one per chunk of <= MM rows, rows held in registers): explicit stencil,
// banded LU solve of the chunk block, exchange of 2W interface values per
// chunk through LDS, dense interface correction.  Phase 3 writes the tile back
// through LDS with the same coalesced pattern.
#include <hip/hip_runtime.h>

#include "lineop.h"
#include "lineop_launch.h"

namespace mc3d {

// first point of line q (z-lines: within the member of grid z)
template <int DIR>
__device__ __forceinline__ long line_base(int q, const LaunchArgs &a) {
    if (DIR == 0) return (long)q * a.Nx;
    if (DIR == 1) return (long)(q / a.Nx) * a.Nx * a.Ny + (q % a.Nx);
    return (long)q;
}

template <int DIR>
__device__ __forceinline__ long point_stride(int Nx, int Ny) {
    if (DIR == 0) return 1;
    if (DIR == 1) return Nx;
    return (long)Nx * Ny;
}

// Explicit stage of one row, evaluated exactly like the reference: taps summed
// in the reference's order with separately rounded products and sums (no FMA
// contraction), then divided by the grid spacing.  Rows at the Dirichlet ends
// cancel an O(1) base value, so this keeps the stage bitwise reproducible.
__device__ __forceinline__ double rhs_row(const OpDev &op, const double *trow, int p) {
#pragma clang fp contract(off)
    double r = 0.0;
    const int N = op.N;
    if (p < op.nbl || p >= N - op.nbr) {
        const int ridx = (p < op.nbl) ? p : op.nbl + (p - (N - op.nbr));
        const double *rec = op.rows + ridx * kRowRec;
        const int off = (int)rec[0], step = (int)rec[1], wd = (int)rec[2];
        for (int k = 0; k < wd; ++k) r = r + rec[3 + k] * trow[p + off + k * step];
    } else {
#pragma unroll
        for (int k = 0; k < 2 * kHMax + 1; ++k)
            if (k < op.int_w) r = r + op.int_c[k] * trow[p + op.int_off + k];
    }
    if (op.divide) r = r / op.dd;
    return r;
}

template <int W, int DIR, int MODE, int MM>
__global__ __launch_bounds__(256) void lineop_kernel(OpDev op, LaunchArgs a) {
    extern __shared__ double smem[];
    const int L = op.L, P = op.P, N = op.N, H = op.H, EXT = op.ext;
    double *tile = smem;
    double *red = smem + (size_t)L * EXT;

    const int f = blockIdx.y;
    const long member = DIR == 2 ? (long)blockIdx.z * a.Nx * a.Ny * a.Nz : 0;
    const double *__restrict__ in = a.in[f] + member;
    double *__restrict__ out = a.out[f] + member;
    const int q0 = blockIdx.x * L;
    const int nthr = blockDim.x;
    const int tid = threadIdx.x;
    const long ps = point_stride<DIR>(a.Nx, a.Ny);

    // ---- phase 1: stage the tile ----
    const bool halo = (MODE != kModePlainDirichlet);
    const int e0 = halo ? 0 : H;
    const int span = halo ? N + 2 * H : N;
    const int total = L * span;
    for (int idx = tid; idx < total; idx += nthr) {
        int l, e;
        if (DIR == 0) { l = idx / span; e = e0 + idx % span; }
        else { l = idx % L; e = e0 + idx / L; }
        const int q = q0 + l;
        if (q >= a.nlines) continue;
        int p = e - H;
        double v;
        if (p >= 0 && p < N) {
            v = in[line_base<DIR>(q, a) + p * ps];
        } else if (MODE == kModeTPB) {
            const int slot = (p < 0) ? 3 + p : 3 + (p - N);  // -1 -> Nm1, N -> Np1
            const double *w = a.wrap[slot];
            v = w ? w[q] : 0.0;
        } else {
            p = (p < 0) ? p + N : p - N;
            v = in[line_base<DIR>(q, a) + p * ps];
        }
        tile[l * EXT + e] = v;
    }
    __syncthreads();

    // ---- phase 2: per-chunk compute ----
    const int l = tid % L;
    const int c = tid / L;
    const int q = q0 + l;
    const bool live = (c < P) && (q < a.nlines);
    const int cc = (c < P) ? c : 0;
    const int cs = op.cstart[cc];
    const int cm = op.clen[cc];
    const double *trow = tile + l * EXT + H;

    double y[MM];
#pragma unroll
    for (int i = 0; i < MM; ++i) y[i] = (live && i < cm) ? rhs_row(op, trow, cs + i) : 0.0;

    if (W > 0) {
        const double *lu = op.lu + (size_t)cc * MM * kLuRec;
#pragma unroll
        for (int i = 1; i < MM; ++i) {
            double z = y[i] - lu[i * kLuRec + 0] * y[i - 1];
            if (W == 2 && i >= 2) z -= lu[i * kLuRec + 1] * y[i - 2];
            y[i] = z;
        }
#pragma unroll
        for (int i = MM - 1; i >= 0; --i) {
            double x = y[i];
            if (i + 1 < MM) x -= lu[i * kLuRec + 2] * y[i + 1];
            if (W == 2 && i + 2 < MM) x -= lu[i * kLuRec + 3] * y[i + 2];
            y[i] = x * lu[i * kLuRec + 4];
        }
        // interface values
        const int NI = op.NI;
        double *myred = red + l * NI;
        if (c < P) {
#pragma unroll
            for (int k = 0; k < W; ++k) {
                myred[c * 2 * W + k] = y[k];
                double v = 0.0;
                const int want = cm - W + k;
#pragma unroll
                for (int i = 0; i < MM; ++i) v = (i == want) ? y[i] : v;
                myred[c * 2 * W + W + k] = v;
            }
        }
        __syncthreads();
        double s[2 * W > 0 ? 2 * W : 1];
        const double *G = op.G + (size_t)cc * 2 * W * NI;
#pragma unroll
        for (int j = 0; j < 2 * W; ++j) {
            double acc = 0.0;
            for (int u = 0; u < NI; ++u) acc += G[j * NI + u] * myred[u];
            s[j] = acc;
        }
        const double *sp = op.sp + (size_t)cc * MM * 2 * W;
#pragma unroll
        for (int i = 0; i < MM; ++i) {
            double x = y[i];
#pragma unroll
            for (int j = 0; j < 2 * W; ++j) x -= sp[i * 2 * W + j] * s[j];
            y[i] = x;
        }
    } else {
        __syncthreads();  // every lane has finished reading the tile
    }

    // pass-through lines on Dirichlet faces of the other directions
    bool pass = false;
    if (op.passA | op.passB) {
        int ia, ib;
        if (DIR == 0) { ia = q % a.Ny; ib = q / a.Ny; }
        else if (DIR == 1) { ia = q % a.Nx; ib = q / a.Nx; }
        else { ia = q % a.Nx; ib = q / a.Nx; }
        pass = (op.passA && (ia == 0 || ia == op.NA - 1)) || (op.passB && (ib == 0 || ib == op.NB - 1));
    }
    if (live && !pass) {
#pragma unroll
        for (int i = 0; i < MM; ++i)
            if (i < cm) tile[l * EXT + H + cs + i] = y[i];
    }
    __syncthreads();

    // ---- phase 3: write back ----
    const int tot_out = L * N;
    for (int idx = tid; idx < tot_out; idx += nthr) {
        int ll, p;
        if (DIR == 0) { ll = idx / N; p = idx % N; }
        else { ll = idx % L; p = idx / L; }
        const int qq = q0 + ll;
        if (qq >= a.nlines) continue;
        const long g = line_base<DIR>(qq, a) + p * ps;
        const double v = tile[ll * EXT + H + p];
        if (a.accumulate) out[g] += v;
        else out[g] = v;
    }
}

template <int W, int DIR, int MM>
static hipError_t launch_mode(const OpDev &op, const LaunchArgs &a, int mode, dim3 grid, dim3 block,
                              size_t lds, hipStream_t st) {
    switch (mode) {
        case kModePlainPeriodic:
            hipLaunchKernelGGL((lineop_kernel<W, DIR, kModePlainPeriodic, MM>), grid, block, lds, st, op, a);
            break;
        case kModePlainDirichlet:
            hipLaunchKernelGGL((lineop_kernel<W, DIR, kModePlainDirichlet, MM>), grid, block, lds, st, op, a);
            break;
        default:
            hipLaunchKernelGGL((lineop_kernel<W, DIR, kModeTPB, MM>), grid, block, lds, st, op, a);
            break;
    }
    return hipGetLastError();
}

template <int W, int DIR>
static hipError_t launch_mm(const OpDev &op, const LaunchArgs &a, int mode, dim3 grid, dim3 block,
                            size_t lds, hipStream_t st) {
    if (op.MM == 32) return launch_mode<W, DIR, 32>(op, a, mode, grid, block, lds, st);
    return launch_mode<W, DIR, 16>(op, a, mode, grid, block, lds, st);
}

template <int W>
static hipError_t launch_dir(const OpDev &op, const LaunchArgs &a, int mode, dim3 grid, dim3 block,
                             size_t lds, hipStream_t st) {
    if (op.dir == 0) return launch_mm<W, 0>(op, a, mode, grid, block, lds, st);
    if (op.dir == 1) return launch_mm<W, 1>(op, a, mode, grid, block, lds, st);
    return launch_mm<W, 2>(op, a, mode, grid, block, lds, st);
}

size_t lineop_lds_bytes(const OpDev &op) {
    return sizeof(double) * ((size_t)op.L * op.ext + (size_t)op.L * (op.NI > 0 ? op.NI : 1));
}

hipError_t lineop_launch(const OpDev &op, const LaunchArgs &a, int nfields, int mode, hipStream_t st) {
    const int nblk = (a.nlines + op.L - 1) / op.L;
    dim3 grid(nblk, nfields, op.dir == 2 ? a.nbatch : 1);
    dim3 block(op.L * op.P);
    const size_t lds = lineop_lds_bytes(op);
    if (op.W == 0) return launch_dir<0>(op, a, mode, grid, block, lds, st);
    if (op.W == 1) return launch_dir<1>(op, a, mode, grid, block, lds, st);
    return launch_dir<2>(op, a, mode, grid, block, lds, st);
}

}  // namespace mc3d
