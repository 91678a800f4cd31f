// Fast gfx950 line-operator kernels for evenly chunked lines (N = P * MM) and
// ragged ones (one short chunk).  Header of templates: each (W, H) pair is
// instantiated in a translation unit of its own (lineop_fast_w*h*.hip), so
// the library builds in parallel; lineop_fast.hip holds the host-side checks
// and the dispatch.
//
// Strided directions (y, z): a wave is 64 consecutive lines (consecutive x)
// and one chunk of MM rows, so every global load and store is a 512-byte
// contiguous run and the chunk index is wave-uniform: the factor tables are
// read with scalar loads.  Each lane keeps its chunk plus the stencil halo in
// registers; LDS is only used for the interface exchange.  The window loads
// and the stores go through a buffer descriptor (one 32-bit lane offset, the
// row offset in a scalar register), and all window loads are issued before
// any other work of the lane.  Lines of more than 16 chunks: evenly chunked
// ones two chunks to a wave (lineop_strided_pair; up to 32 chunks, 30 with
// special rows), the others in rounds of 16 chunk waves with a second,
// recomputing pass (lineop_strided_multi).
//
// Contiguous direction (x): the workgroup stages 8, 16 (periodic, explicit)
// or 64 lines (banded Dirichlet or ragged: a wave per chunk; 32 lines and two
// chunks a wave for Dirichlet lines of 17 to 30 chunks of 32 rows,
// lineop_contig_pair) into LDS in a
// transposed [point][line] layout (row stride lines + 1) with halo rows, so
// the coalesced fill and the per-lane chunk reads are both free of bank
// conflicts; each thread issues batches of loads before its LDS writes.
// 64-line tiles too large for LDS stage their chunks in groups
// (lineop_contig_ph).
//
// Per chunk: explicit stencil (taps in the reference's summation order, no
// FMA contraction, so the explicit stage is bitwise that of the reference),
// banded LU solve of the chunk block, then the interface correction
// x = y - S * (G * yhat) with the 2W-per-chunk interface values exchanged
// through LDS.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>

#include "lineop.h"
#include "lineop_launch.h"
#include "lineop_plan.h"

namespace mc3d {

namespace {

// one dense special row: taps d[1..K] over the window points v[0..K), summed
// in increasing (d[0] > 0) or decreasing position
template <int K>
__device__ __forceinline__ double dense_row(const double *d, const double *v) {
#pragma clang fp contract(off)
    double r = 0.0;
    if (d[0] > 0.0) {
#pragma unroll
        for (int t = 0; t < K; ++t) r = r + d[1 + t] * v[t];
    } else {
#pragma unroll
        for (int t = K - 1; t >= 0; --t) r = r + d[1 + t] * v[t];
    }
    return r;
}

// explicit stage of the MM rows of one chunk from the window w[0..MM+2H)
// (w[j] = phi[cs - H + j]).  With ROWS, the first chunk's first H rows and
// the last chunk's last H rows are special rows of K = 2H + 1 dense taps:
// Dirichlet left rows over phi[0..K) = w[H..H+K), right rows over
// phi[N-K..N) = w[MM+H-K..MM+H), periodic rows (CD2's wrap rows) over their
// own stencil w[i..i+K).  Each end of the chunk is one uniform branch between
// straight-line blocks of H rows, so the window values either block reads
// are live only where the interior rows read them anyway.  (Per-row branches
// between the two stencils, or special rows computed over the interior ones
// afterwards, cost 8-30 vector registers and a wave per SIMD.)
template <int H, int MM, bool PER, bool ROWS>
__device__ __forceinline__ void chunk_rhs(const OpDev &op, const double *w, double *y, bool first,
                                          bool last) {
#pragma clang fp contract(off)
    constexpr int K = 2 * H + 1;
    auto scale = [&](double r) {
        if (op.inv_dd != 0.0) return r * op.inv_dd;
        if (op.divide) return r / op.dd;
        return r;
    };
    auto interior = [&](int i) {
        double r = 0.0;
#pragma unroll
        for (int k = 0; k < K; ++k) r = r + op.int_c[k] * w[i + k];
        return scale(r);
    };
    if (ROWS && first) {
#pragma unroll
        for (int i = 0; i < H; ++i) y[i] = scale(dense_row<K>(op.drows + i * kDRowRec, w + (PER ? i : H)));
    } else {
#pragma unroll
        for (int i = 0; i < H; ++i) y[i] = interior(i);
    }
#pragma unroll
    for (int i = H; i < MM - H; ++i) y[i] = interior(i);
    if (ROWS && last) {
#pragma unroll
        for (int i = MM - H; i < MM; ++i)
            y[i] = scale(dense_row<K>(op.drows + (i - MM + 2 * H) * kDRowRec, w + (PER ? i : MM + H - K)));
    } else {
#pragma unroll
        for (int i = MM - H; i < MM; ++i) y[i] = interior(i);
    }
}

// banded LU solve of one chunk block with the factor rows lu[MM][kLuRec]
// (lu(e): entry e of the rows).  The factor rows are read with scalar loads; a
// scheduling fence every eight rows keeps the scheduler from hoisting every
// row's loads to the top, where they outgrow the scalar registers and are
// spilled to vector-register lanes (one v_writelane per dword, one v_readlane
// per dword at each use).
template <int W, int MM, class LU>
__device__ __forceinline__ void chunk_solve_by(LU lu, double *y) {
#pragma unroll
    for (int i = 1; i < MM; ++i) {
        if (i % 8 == 0) __builtin_amdgcn_sched_barrier(0);
        double z = y[i] - lu(i * kLuRec + 0) * y[i - 1];
        if (W == 2 && i >= 2) z -= lu(i * kLuRec + 1) * y[i - 2];
        y[i] = z;
    }
#pragma unroll
    for (int i = MM - 1; i >= 0; --i) {
        if (i % 8 == 7) __builtin_amdgcn_sched_barrier(0);
        double x = y[i];
        if (i + 1 < MM) x -= lu(i * kLuRec + 2) * y[i + 1];
        if (W == 2 && i + 2 < MM) x -= lu(i * kLuRec + 3) * y[i + 2];
        y[i] = x * lu(i * kLuRec + 4);
    }
}
template <int W, int MM>
__device__ __forceinline__ void chunk_solve(const double *lu, double *y) {
    chunk_solve_by<W, MM>([lu](int e) { return lu[e]; }, y);
}

// interface correction y = y - S s of one chunk with the spike rows
// sp[MM][2W] (sp(e): entry e)
template <int W, int MM, class SP>
__device__ __forceinline__ void chunk_correct_by(SP sp, const double *s, double *y) {
#pragma unroll
    for (int i = 0; i < MM; ++i) {
        if (i % 8 == 0) __builtin_amdgcn_sched_barrier(0);
        double x = y[i];
#pragma unroll
        for (int j = 0; j < 2 * W; ++j) x -= sp(i * 2 * W + j) * s[j];
        y[i] = x;
    }
}

// s[j] = sum over the chunks u of the band around chunk c of G_c[j][u's 2W
// columns] . yhat[u's 2W values].  Evenly chunked periodic lines read chunk
// 0's rows G_0 rotated by c chunks (G_c[j][u] = G_0[j][u - 2Wc]), ragged ones
// (RG) their own rows; chunk offsets -gb..gb, or every chunk once when the
// band covers the line.  Dirichlet lines read G_c over chunks
// max(0, c - gb) .. min(P - 1, c + gb).
template <int W, bool PER, bool RG>
__device__ __forceinline__ void interface_sums(const OpDev &op, int c, const double *yhat, double *s) {
    constexpr int W2 = 2 * W;
    const int P = op.P, NI = op.NI, gb = op.gb;
#pragma unroll
    for (int j = 0; j < W2; ++j) s[j] = 0.0;
    if (PER) {
        const bool band = 2 * gb + 1 < P;
        const int nb = band ? 2 * gb + 1 : P, d0 = band ? -gb : 0;
        for (int t = 0; t < nb; ++t) {
            const int d = d0 + t;
            const int cb = d < 0 ? d + P : d;  // column block of G_0
            int ub = c + d;                    // chunk whose interface values it multiplies
            ub = ub < 0 ? ub + P : (ub >= P ? ub - P : ub);
            const double *g = RG ? op.G + (size_t)c * W2 * NI + ub * W2 : op.G + cb * W2;
            const double *yv = yhat + ub * W2;
#pragma unroll
            for (int k = 0; k < W2; ++k) {
                const double v = yv[k];
#pragma unroll
                for (int j = 0; j < W2; ++j) s[j] += g[j * NI + k] * v;
            }
        }
    } else {
        const double *G = op.G + (size_t)c * W2 * NI;
        const int lo = c - gb < 0 ? 0 : c - gb, hi = c + gb > P - 1 ? P - 1 : c + gb;
        for (int ub = lo; ub <= hi; ++ub) {
            const double *g = G + ub * W2;
            const double *yv = yhat + ub * W2;
#pragma unroll
            for (int k = 0; k < W2; ++k) {
                const double v = yv[k];
#pragma unroll
                for (int j = 0; j < W2; ++j) s[j] += g[j * NI + k] * v;
            }
        }
    }
}

// interface_sums of Dirichlet lines for a wave of two chunks, cl in the lower
// half-wave and ch in the upper (hi): every lane forms both chunks' sums with
// the rows G_cl and G_ch as scalar operands, over its own line's interface
// values, and keeps its half's (rows selected per lane instead took 28 more
// registers).  The terms come in the order of interface_sums (increasing
// chunk); a chunk offset d = -gb .. gb that falls off the line for one chunk
// adds zero terms there.
template <int W>
__device__ __forceinline__ void interface_sums_pair(const OpDev &op, int cl, int ch, bool hi, const double *yhat,
                                                    double *s) {
    constexpr int W2 = 2 * W;
    const int P = op.P, NI = op.NI, gb = op.gb;
    const double *Gl = op.G + (size_t)cl * W2 * NI, *Gh = op.G + (size_t)ch * W2 * NI;
    double sl[W2], sh[W2];
#pragma unroll
    for (int j = 0; j < W2; ++j) sl[j] = sh[j] = 0.0;
    for (int d = -gb; d <= gb; ++d) {
        const int ul = cl + d, uh = ch + d;
        const bool vl = ul >= 0 && ul < P, vh = uh >= 0 && uh < P;
        if (vl) {
            const double *g = Gl + ul * W2, *yv = yhat + ul * W2;
#pragma unroll
            for (int k = 0; k < W2; ++k) {
                const double v = yv[k];
#pragma unroll
                for (int j = 0; j < W2; ++j) sl[j] += g[j * NI + k] * v;
            }
        }
        __builtin_amdgcn_sched_barrier(0);
        if (vh) {
            const double *g = Gh + uh * W2, *yv = yhat + uh * W2;
#pragma unroll
            for (int k = 0; k < W2; ++k) {
                const double v = yv[k];
#pragma unroll
                for (int j = 0; j < W2; ++j) sh[j] += g[j * NI + k] * v;
            }
        }
    }
#pragma unroll
    for (int j = 0; j < W2; ++j) s[j] = hi ? sh[j] : sl[j];
}

// raw buffer access: descriptor in SGPRs, one 32-bit lane offset, row offset in an SGPR
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
constexpr int kRawBufferWord3 = 0x00020000;  // gfx9-family raw buffer descriptor, dword 3
__device__ __forceinline__ double bload(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
// rows no other chunk reads: non-temporal loads (cache-policy bit 1), which keep
// the halo rows the neighbouring chunks re-read in the cache
__device__ __forceinline__ double bload_nt(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 2));
}

// results are written once and read by a later launch (from HBM at these sizes):
// non-temporal stores (cache-policy bit 1)
__device__ __forceinline__ void bstore(double v, __amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, voff, soff, 2);
}

__device__ __forceinline__ bool pass_line(const OpDev &op, int dir, int q, int Nx, int Ny) {
    if (!(op.passA | op.passB)) return false;
    int ia, ib;
    if (dir == 0) { ia = q % Ny; ib = q / Ny; }
    else { ia = q % Nx; ib = q / Nx; }
    return (op.passA && (ia == 0 || ia == op.NA - 1)) || (op.passB && (ib == 0 || ib == op.NB - 1));
}

// A y- or z-line qq of this launch row (field blockIdx.y): y-lines run x
// fastest within a z-plane over every stacked member, z-lines belong to member
// blockIdx.z.  The field pointers (member included), the line's first point,
// the point stride, and the bytes a buffer descriptor over the field covers.
struct StridedLine {
    const double *in;
    double *out;
    long base, ps;
    unsigned fbytes;
};
__device__ __forceinline__ StridedLine strided_line(const OpDev &op, const LaunchArgs &a, int qq) {
    StridedLine r{a.in[blockIdx.y], a.out[blockIdx.y], 0, 0,
                  (unsigned)a.Nx * (unsigned)a.Ny * (unsigned)a.Nz * 8u};  // one member
    if (op.dir == 1) {
        r.base = (long)(qq / a.Nx) * a.Nx * a.Ny + qq % a.Nx;
        r.ps = a.Nx;
        r.fbytes *= (unsigned)a.nbatch;
    } else {
        const long member = (long)blockIdx.z * a.Nx * a.Ny * a.Nz;
        r.in += member;
        r.out += member;
        r.base = qq;
        r.ps = (long)a.Nx * a.Ny;
    }
    return r;
}

// Ragged lines (RG): chunk cshort (1) has lshort < MM rows, so later chunks
// start MM - lshort rows earlier.  The chunk's rows beyond its length are
// computed and never stored; its factor rows there are identity rows.
template <int MM, bool RG>
__device__ __forceinline__ int chunk_start(const OpDev &op, int c) {
    return RG && c > op.cshort ? c * MM - (MM - op.lshort) : c * MM;
}

// publishes a chunk's first and last W values; the last W of the short chunk
// of a ragged line sit at rows len - W .. len - 1, picked by a select over the
// rows (registers cannot be indexed at run time)
template <int W, int MM, bool RG>
__device__ __forceinline__ void publish(double *myred, int c, bool short_chunk, int len, const double *y) {
#pragma unroll
    for (int k = 0; k < W; ++k) myred[c * 2 * W + k] = y[k];
    if (RG && short_chunk) {
#pragma unroll
        for (int k = 0; k < W; ++k) {
            double v = 0.0;
#pragma unroll
            for (int i = 0; i < MM; ++i) v = (i == len - W + k) ? y[i] : v;
            myred[c * 2 * W + W + k] = v;
        }
    } else {
#pragma unroll
        for (int k = 0; k < W; ++k) myred[c * 2 * W + W + k] = y[MM - W + k];
    }
}

// special end rows: every Dirichlet line; of the periodic operators only CD2
// (W = 0, H = 1), whose wrap rows carry the opposite sign of its interior
template <int W, int H, bool PER>
constexpr bool has_rows() { return !PER || (W == 0 && H == 1); }

template <int W, int H, int MM, bool PER, bool RG>
__global__ __launch_bounds__(1024) void lineop_strided(OpDev op, LaunchArgs a) {
    extern __shared__ double red[];
    const int l = threadIdx.x;
    const int c = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const int P = op.P, N = op.N, NI = op.NI;
    const int q = blockIdx.x * 64 + l;
    const bool live = q < a.nlines;
    const int qq = live ? q : 0;
    const StridedLine ln = strided_line(op, a, qq);
    const double *__restrict__ in = ln.in;
    double *__restrict__ out = ln.out;
    const long base = ln.base, ps = ln.ps;
    const unsigned fbytes = ln.fbytes;
    const int cs = chunk_start<MM, RG>(op, c);
    const bool short_chunk = RG && c == op.cshort;
    const int len = short_chunk ? op.lshort : MM;
    // window loads through a buffer descriptor (fields below 2^28 points)
    const __amdgpu_buffer_rsrc_t rin =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<double *>(in), (short)0, (int)fbytes, kRawBufferWord3);
    const unsigned boff = (unsigned)base * 8u, rowb = (unsigned)ps * 8u;
    double w[MM + 2 * H];
    // row offsets: consecutive rows from (cs - H); only the halo rows of the
    // first and last chunk wrap (periodic) or fall outside the line (zero)
    const int s0 = cs - H;
    const bool first = c == 0, last = c == P - 1;
#pragma unroll
    for (int j = 0; j < MM + 2 * H; ++j) {
        if (j < H || j >= MM + H) {
            int p = s0 + j;
            const bool outside = (j < H && first) || (j >= MM + H && last);
            if (PER) {
                if (outside) p = j < H ? p + N : p - N;
                w[j] = bload(rin, boff, (unsigned)p * rowb);
            } else {
                w[j] = outside ? 0.0 : bload(rin, boff, (unsigned)p * rowb);
            }
        } else if (j >= 2 * H && j < MM) {
            w[j] = bload_nt(rin, boff, (unsigned)(s0 + j) * rowb);
        } else {
            w[j] = bload(rin, boff, (unsigned)(s0 + j) * rowb);
        }
    }
    // issue the window loads before any table load or arithmetic
    __builtin_amdgcn_sched_barrier(0);
    double y[MM];
    chunk_rhs<H, MM, PER, has_rows<W, H, PER>()>(op, w, y, c == 0, c == P - 1);
    if (W > 0) {
        chunk_solve<W, MM>(op.lu + (PER && !RG ? 0 : (size_t)c * MM * kLuRec), y);
        publish<W, MM, RG>(red + l * (NI + 1), c, short_chunk, len, y);
    }
    // every lane has read its window (in may equal out) and published its interface
    __syncthreads();
    if (W > 0) {
        const double *myred = red + l * (NI + 1);
        double s[2 * W > 0 ? 2 * W : 1];
        interface_sums<W, PER, RG>(op, c, myred, s);
        const double *sp = op.sp + (PER && !RG ? 0 : (size_t)c * MM * 2 * W);
        chunk_correct_by<W, MM>([sp](int e) { return sp[e]; }, s, y);
    }
    if (!live) return;
    const __amdgpu_buffer_rsrc_t rout =
        __builtin_amdgcn_make_buffer_rsrc(out, (short)0, (int)fbytes, kRawBufferWord3);
    if (pass_line(op, op.dir, q, a.Nx, a.Ny)) {
#pragma unroll
        for (int i = 0; i < MM; ++i)
            if (!RG || i < len) y[i] = bload(rin, boff, (unsigned)(cs + i) * rowb);
    }
    if (a.accumulate) {
#pragma unroll
        for (int i = 0; i < MM; ++i)
            if (!RG || i < len) y[i] = bload_nt(rout, boff, (unsigned)(cs + i) * rowb) + y[i];
    }
#pragma unroll
    for (int i = 0; i < MM; ++i)
        if (!RG || i < len) bstore(y[i], rout, boff, (unsigned)(cs + i) * rowb);
}

// y- and z-lines of more than 16 chunks of 16 rows (640 or 1024 points, say): a
// workgroup of 64 lines x 16 chunk waves, each wave taking chunks ty, ty + 16,
// ... in rounds.  Pass 1 computes every chunk's explicit stage and block solve
// and publishes its interface values; pass 2 recomputes each chunk (its window
// read again) and applies the interface correction.  Nothing is held across
// chunks but the interface values, so the registers stay those of one chunk;
// the price is a second read of the input and a second explicit stage and
// solve.  Within a round every window is read before any of the round's
// stores; the rows a window needs from a chunk of an earlier round (the H rows
// below each round's first chunk, and the line's first H rows for the
// periodic wrap of the last chunk) are stashed in LDS in pass 1, so that the
// operator may run in place.
template <int W, int H, int MM, bool PER, bool RG>
__global__ __launch_bounds__(1024) void lineop_strided_multi(OpDev op, LaunchArgs a) {
    extern __shared__ double smem[];
    constexpr int CW = 16;  // chunk waves per workgroup
    const int l = threadIdx.x;
    const int ty = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const int P = op.P, NI = op.NI, N = op.N;
    const int rounds = (P + CW - 1) / CW;
    double *red = smem;                               // [64][NI + 1]
    double *stash = smem + (size_t)64 * (NI + 1);     // [64][rounds][H]
    const int q = blockIdx.x * 64 + l;
    const bool live = q < a.nlines;
    const int qq = live ? q : 0;
    const StridedLine ln = strided_line(op, a, qq);
    const double *__restrict__ in = ln.in;
    double *__restrict__ out = ln.out;
    const long base = ln.base, ps = ln.ps;
    const unsigned fbytes = ln.fbytes;
    const __amdgpu_buffer_rsrc_t rin =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<double *>(in), (short)0, (int)fbytes, kRawBufferWord3);
    const __amdgpu_buffer_rsrc_t rout =
        __builtin_amdgcn_make_buffer_rsrc(out, (short)0, (int)fbytes, kRawBufferWord3);
    const unsigned boff = (unsigned)base * 8u, rowb = (unsigned)ps * 8u;
    double *mystash = stash + (size_t)l * rounds * H;
    // the window of chunk c: rows cs - H .. cs + MM + H - 1 (pass 2 takes the
    // rows stored in an earlier round from the stash)
    auto window = [&](int c, bool pass2, double *w) {
        const int s0 = chunk_start<MM, RG>(op, c) - H;
        const bool first = c == 0, last = c == P - 1;
        const bool below = pass2 && c > 0 && c % CW == 0;        // rows of chunk c - 1, an earlier round
        const bool wrap = pass2 && PER && last && rounds > 1;    // rows 0 .. H - 1, stored in round 0
#pragma unroll
        for (int j = 0; j < MM + 2 * H; ++j) {
            if (j < H) {
                if (below) {
                    w[j] = mystash[(c / CW) * H + j];
                } else if (first) {
                    w[j] = PER ? bload(rin, boff, (unsigned)(s0 + j + N) * rowb) : 0.0;
                } else {
                    w[j] = bload(rin, boff, (unsigned)(s0 + j) * rowb);
                }
            } else if (j >= MM + H) {
                if (wrap) {
                    w[j] = mystash[j - MM - H];
                } else if (last) {
                    w[j] = PER ? bload(rin, boff, (unsigned)(s0 + j - N) * rowb) : 0.0;
                } else {
                    w[j] = bload(rin, boff, (unsigned)(s0 + j) * rowb);
                }
            } else {
                w[j] = bload(rin, boff, (unsigned)(s0 + j) * rowb);
            }
        }
    };
    // pass 1: interface values of every chunk, and the stash
#pragma unroll 1
    for (int r = 0; r < rounds; ++r) {
        const int c = ty + CW * r;
        // (explicit operators read a window here only for the stash)
        if (c < P && (W > 0 || c % CW == CW - 1 || (PER && c == 0))) {
            double w[MM + 2 * H];
            window(c, false, w);
            if (c % CW == CW - 1 && c + 1 < P) {  // the H rows below chunk c + 1 (c is a full chunk)
#pragma unroll
                for (int j = 0; j < H; ++j) mystash[((c + 1) / CW) * H + j] = w[MM + j];
            }
            if (PER && c == 0) {  // the line's first H rows
#pragma unroll
                for (int j = 0; j < H; ++j) mystash[j] = w[H + j];
            }
            if (W > 0) {
                double y[MM];
                chunk_rhs<H, MM, PER, has_rows<W, H, PER>()>(op, w, y, c == 0, c == P - 1);
                chunk_solve<W, MM>(op.lu + (PER && !RG ? 0 : (size_t)c * MM * kLuRec), y);
                publish<W, MM, RG>(red + l * (NI + 1), c, RG && c == op.cshort, RG && c == op.cshort ? op.lshort : MM,
                                   y);
            }
        }
    }
    __syncthreads();
    // pass 2: a round at a time, every window read before the round's stores.
    // The interface sums come first: computed after the block solve, they made
    // the compiler spill the chunk's registers to scratch.
#pragma unroll 1
    for (int r = 0; r < rounds; ++r) {
        const int c = ty + CW * r;
        const bool has = c < P;
        const int cs = has ? chunk_start<MM, RG>(op, c) : 0;
        const int len = RG && c == op.cshort ? op.lshort : MM;
        double y[MM];
        double s[2 * W > 0 ? 2 * W : 1];
        if (has && W > 0) interface_sums<W, PER, RG>(op, c, red + l * (NI + 1), s);
        if (has) {
            double w[MM + 2 * H];
            window(c, true, w);
            __builtin_amdgcn_sched_barrier(0);
            chunk_rhs<H, MM, PER, has_rows<W, H, PER>()>(op, w, y, c == 0, c == P - 1);
            if (W > 0) {
                chunk_solve<W, MM>(op.lu + (PER && !RG ? 0 : (size_t)c * MM * kLuRec), y);
                const double *sp = op.sp + (PER && !RG ? 0 : (size_t)c * MM * 2 * W);
                chunk_correct_by<W, MM>([sp](int e) { return sp[e]; }, s, y);
            }
        }
        __syncthreads();  // the round's windows are read (in may equal out)
        if (has && live) {
            if (pass_line(op, op.dir, q, a.Nx, a.Ny)) {
#pragma unroll
                for (int i = 0; i < MM; ++i)
                    if (!RG || i < len) y[i] = bload(rin, boff, (unsigned)(cs + i) * rowb);
            }
            if (a.accumulate) {
#pragma unroll
                for (int i = 0; i < MM; ++i)
                    if (!RG || i < len) y[i] = bload_nt(rout, boff, (unsigned)(cs + i) * rowb) + y[i];
            }
#pragma unroll
            for (int i = 0; i < MM; ++i)
                if (!RG || i < len) bstore(y[i], rout, boff, (unsigned)(cs + i) * rowb);
        }
    }
}

// Evenly chunked y- and z-lines of 17 to 32 chunks (576 or 1024 points, say),
// two chunks to a wave, one per half-wave, each over 32 consecutive lines
// (256-byte runs).  The window's row offset within the chunk stays a scalar
// register; the chunk's start moves into the lane's offset.
//  - periodic lines without special rows: lanes 0-31 chunk ty, lanes 32-63
//    chunk ty + P2 (P2 = P/2 rounded up; when P is odd the last wave's upper
//    half is idle); every chunk reads chunk 0's factor rows and the rotation
//    of its interface rows;
//  - lines with special rows (Dirichlet, CD2's periodic wrap rows), up to 30
//    chunks: chunk 0 and chunk P - 1 take a wave each (the upper half redoes
//    the lower half's chunk and stores nothing), the interior chunks 1 .. P - 2
//    two to a wave (1 + k and 1 + k + Q, Q = (P - 1) / 2); ragged Dirichlet
//    lines (RG) up to 29 chunks give the short chunk 1 a wave of its own too
//    and pair chunks 2 .. P - 2 (2 + k and 2 + k + Q, Q = (P - 2) / 2).  The end-row flags
//    and the factor and spike tables are then wave-uniform (the interior
//    chunks' tables are equal bitwise; the host checks it); the interface
//    rows G_c differ for every chunk of a Dirichlet line, so each lane forms
//    both of its wave's sums and keeps its own (interface_sums_pair).
//    (Tables and end rows selected per lane in the waves holding an end
//    chunk took the kernel past 128 registers.)
template <int W, int H, int MM, bool PER, bool RG>
__global__ __launch_bounds__(1024) void lineop_strided_pair(OpDev op, LaunchArgs a) {
    static_assert(!(PER && RG), "ragged lines in pairs: Dirichlet only");
    constexpr bool ROWS = has_rows<W, H, PER>();
    constexpr bool OWN = !PER && W > 0;  // tables per chunk
    extern __shared__ double red[];  // [32 lines][NI + 1]
    const int l = threadIdx.x, ll = l & 31, half = l >> 5;
    const int P = op.P, N = op.N, NI = op.NI;
    const int P2 = (P + 1) / 2;
    // the wave's chunks (cl, ch), wave-uniform with special rows
    const int ty = ROWS ? __builtin_amdgcn_readfirstlane(threadIdx.y) : 0;
    constexpr int F = RG ? 2 : 1;  // first paired chunk (waves 0 .. F hold one chunk each)
    const int Q = (P - F) / 2;
    const int cl = ty == 0 ? 0 : ty == 1 ? P - 1 : ty - 1;
    const bool hi_has = ty >= F + 1 && ty - 1 + Q <= P - 2;
    const int ch = hi_has ? ty - 1 + Q : cl;
    int c;     // this lane's chunk
    bool has;  // whether the lane stores it
    if constexpr (ROWS) {
        c = half ? ch : cl;
        has = !half || hi_has;
    } else {
        c = (int)threadIdx.y + half * P2;
        has = c < P;  // (the last wave's upper half when P is odd)
    }
    const int cc = has || ROWS ? c : 0;
    const int q = blockIdx.x * 32 + ll;
    const bool live = q < a.nlines;
    const int qq = live ? q : 0;
    // (the addressing of strided_line, written out: through the helper this
    // kernel took two more registers and a wave fewer per SIMD)
    const double *__restrict__ in = a.in[blockIdx.y];
    double *__restrict__ out = a.out[blockIdx.y];
    long base, ps;
    unsigned fbytes = (unsigned)a.Nx * (unsigned)a.Ny * (unsigned)a.Nz * 8u;  // one member
    if (op.dir == 1) {
        base = (long)(qq / a.Nx) * a.Nx * a.Ny + qq % a.Nx;
        ps = a.Nx;
        fbytes *= (unsigned)a.nbatch;
    } else {
        const long member = (long)blockIdx.z * a.Nx * a.Ny * a.Nz;
        in += member;
        out += member;
        base = qq;
        ps = (long)a.Nx * a.Ny;
    }
    const int cs = chunk_start<MM, RG>(op, cc);
    const bool short_chunk = RG && cl == op.cshort;  // (wave-uniform: the short chunk's wave holds it alone)
    const int len = short_chunk ? op.lshort : MM;
    const __amdgpu_buffer_rsrc_t rin =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<double *>(in), (short)0, (int)fbytes, kRawBufferWord3);
    const unsigned rowb = (unsigned)ps * 8u;
    // The lane's offset is the first row of each of the window's three parts
    // (the H rows below the chunk, wrapped to N - H for chunk 0; the chunk's
    // rows; the H rows above it, wrapped to 0 for chunk P - 1), the row within
    // the part the scalar offset; every lane offset is a row of the field.
    // (Dirichlet lines read the wrapped rows too; their end rows never use them.)
    const unsigned b8 = (unsigned)base * 8u;
    const unsigned vlo = b8 + (unsigned)(cc == 0 ? N - H : cs - H) * rowb;
    const unsigned vmid = b8 + (unsigned)cs * rowb;
    const unsigned vhi = b8 + (unsigned)(cc == P - 1 ? 0 : cs + MM) * rowb;
    double w[MM + 2 * H];
#pragma unroll
    for (int j = 0; j < MM + 2 * H; ++j) {
        if (j < H) {
            w[j] = bload(rin, vlo, (unsigned)j * rowb);
        } else if (j >= MM + H) {
            w[j] = bload(rin, vhi, (unsigned)(j - MM - H) * rowb);
        } else if (j >= 2 * H && j < MM) {
            w[j] = bload_nt(rin, vmid, (unsigned)(j - H) * rowb);
        } else {
            w[j] = bload(rin, vmid, (unsigned)(j - H) * rowb);
        }
    }
    __builtin_amdgcn_sched_barrier(0);
    double y[MM];
    chunk_rhs<H, MM, PER, ROWS>(op, w, y, cl == 0, cl == P - 1);
    double *myred = red + ll * (NI + 1);
    const double *lu = op.lu + (OWN ? (size_t)cl * MM * kLuRec : 0);
    const double *sp = op.sp + (OWN ? (size_t)cl * MM * 2 * W : 0);
    if (W > 0) {
        chunk_solve<W, MM>(lu, y);
        if (has) publish<W, MM, RG>(myred, c, short_chunk, len, y);
    }
    __syncthreads();
    if (W > 0) {
        double s[2 * W > 0 ? 2 * W : 1];
        if constexpr (OWN) interface_sums_pair<W>(op, cl, ch, half, myred, s);
        else interface_sums<W, true, false>(op, cc, myred, s);
        chunk_correct_by<W, MM>([=](int e) { return sp[e]; }, s, y);
    }
    if (!live || !has) return;
    const __amdgpu_buffer_rsrc_t rout =
        __builtin_amdgcn_make_buffer_rsrc(out, (short)0, (int)fbytes, kRawBufferWord3);
    if (pass_line(op, op.dir, q, a.Nx, a.Ny)) {
#pragma unroll
        for (int i = 0; i < MM; ++i)
            if (!RG || i < len) y[i] = bload(rin, vmid, (unsigned)i * rowb);
    }
    if (a.accumulate) {
#pragma unroll
        for (int i = 0; i < MM; ++i)
            if (!RG || i < len) y[i] = bload_nt(rout, vmid, (unsigned)i * rowb) + y[i];
    }
#pragma unroll
    for (int i = 0; i < MM; ++i)
        if (!RG || i < len) bstore(y[i], rout, vmid, (unsigned)i * rowb);
}

// Dirichlet x-lines of 17 to 30 chunks of 32 rows of banded operators (544 to
// 960 points), which the 64-line tiles cannot hold a wave per chunk:
// 32-line tiles, two chunks to a wave, one per half-wave, laid out as in
// lineop_strided_pair (chunk 0 and chunk P - 1 a wave each, the interior
// chunks two to a wave, so that tables and end rows are wave-uniform and only
// the interface sums are formed for both chunks of a wave), here
// with neighbouring chunks paired: wave t = 1 .. Q holds chunks 2t - 1 and 2t
// (Q = (P - 1) / 2; for odd P the upper half of wave Q redoes the lower
// half's chunk and stores nothing), wave Q + 1 chunk P - 1.  The waves then
// hold the chunks in line order, and the staging of lineop_contig_ph applies
// to groups of waves: op.phases groups of consecutive waves, each group's rows
// loaded into one tile [row][line] (row stride 33), its windows read, the next
// group reusing the tile; the results drain the same way.
template <int W, int H, int MM>
__global__ __launch_bounds__(1024) void lineop_contig_pair(OpDev op, LaunchArgs a) {
    extern __shared__ double smem[];
    constexpr int L = 32, LS = L + 1, kB = 8;
    const int N = op.N, P = op.P, NI = op.NI, K = op.phases;
    const int Q = (P - 1) / 2, NW = Q + 2;
    const int wpp = (NW + K - 1) / K;  // waves per group
    // first and last chunk of wave t
    auto wfirst = [=](int t) { return t == 0 ? 0 : t == NW - 1 ? P - 1 : 2 * t - 1; };
    auto wlast = [=](int t) { return t == 0 ? 0 : t == NW - 1 ? P - 1 : 2 * t <= P - 2 ? 2 * t : 2 * t - 1; };
    double *tile = smem;
    constexpr bool kShare = W == 2;  // interface values inside the tile (dead by then)
    // (a group spans at most 2 wpp chunks)
    double *red = kShare ? smem : smem + (size_t)(2 * wpp * MM + 2 * H) * LS;
    const int l = threadIdx.x, ll = l & 31, half = l >> 5;
    const int ty = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const int cl = wfirst(ty);
    const bool hi_has = ty >= 1 && ty <= Q && 2 * ty <= P - 2;
    const int ch = hi_has ? 2 * ty : cl;
    const int c = half ? ch : cl;  // this lane's chunk
    const bool has = !half || hi_has;
    const int tid = threadIdx.y * 64 + l;
    const int nthr = 64 * NW;
    const int q0 = blockIdx.x * L;
    const int rem = a.nlines - q0;
    const unsigned nrec = (unsigned)(rem < L ? rem : L) * (unsigned)N * 8u;
    const __amdgpu_buffer_rsrc_t rin = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<double *>(a.in[blockIdx.y]) + (long)q0 * N, (short)0, (int)nrec, kRawBufferWord3);
    const int q = q0 + ll;
    const bool live = q < a.nlines;
    const int cs = c * MM;
    const int myk = ty / wpp;
    double w[MM + 2 * H];
    for (int k = 0; k < K; ++k) {
        const int t0 = k * wpp;
        if (t0 >= NW) break;
        const int t1 = t0 + wpp < NW ? t0 + wpp : NW;
        const int r0 = wfirst(t0) * MM - H;  // line row held by tile row 0
        const int R = (wlast(t1 - 1) + 1) * MM + H - r0;
        if (k > 0) __syncthreads();  // the previous group has read its windows
        {
            int lq = tid / R, rr = tid - lq * R;
            const int qd = nthr / R, rd = nthr - qd * R;
            for (int e0 = tid; e0 < L * R; e0 += kB * nthr) {
                double v[kB];
                int slot[kB];
#pragma unroll
                for (int b = 0; b < kB; ++b) {
                    const int p = r0 + rr;
                    v[b] = (lq < L && p >= 0 && p < N) ? bload_nt(rin, (unsigned)(lq * N + p) * 8u, 0) : 0.0;
                    slot[b] = lq < L ? rr * LS + lq : -1;
                    lq += qd;
                    rr += rd;
                    if (rr >= R) { rr -= R; ++lq; }
                }
#pragma unroll
                for (int b = 0; b < kB; ++b)
                    if (slot[b] >= 0) tile[slot[b]] = v[b];
            }
        }
        __syncthreads();
        if (myk == k) {
            const double *trow = tile + (size_t)(cs - H - r0) * LS + ll;
#pragma unroll
            for (int j = 0; j < MM + 2 * H; ++j) w[j] = trow[j * LS];
        }
    }
    if (kShare) __syncthreads();  // every window is read before the interface values overwrite the tile
    double y[MM];
    chunk_rhs<H, MM, false, true>(op, w, y, cl == 0, cl == P - 1);
    double *myred = red + ll * (NI + 1);
    chunk_solve<W, MM>(op.lu + (size_t)cl * MM * kLuRec, y);
    if (has) publish<W, MM, false>(myred, c, false, MM, y);
    __syncthreads();
    {
        double s[2 * W];
        interface_sums_pair<W>(op, cl, ch, half, myred, s);
        const double *sp = op.sp + (size_t)cl * MM * 2 * W;
        chunk_correct_by<W, MM>([=](int e) { return sp[e]; }, s, y);
    }
    // pass-through lines keep their input (nothing is stored yet)
    if (live && pass_line(op, 0, q, a.Nx, a.Ny)) {
#pragma unroll
        for (int i = 0; i < MM; ++i) y[i] = bload(rin, (unsigned)(ll * N + cs + i) * 8u, 0);
    }
    const __amdgpu_buffer_rsrc_t rout = __builtin_amdgcn_make_buffer_rsrc(
        a.out[blockIdx.y] + (long)q0 * N, (short)0, (int)nrec, kRawBufferWord3);
    for (int k = 0; k < K; ++k) {
        const int t0 = k * wpp;
        if (t0 >= NW) break;
        const int t1 = t0 + wpp < NW ? t0 + wpp : NW;
        const int s0 = wfirst(t0) * MM;  // line row held by tile row 0
        const int R = (wlast(t1 - 1) + 1) * MM - s0;
        __syncthreads();  // interface values read (first group) or the previous group stored
        if (myk == k && has && live) {
#pragma unroll
            for (int i = 0; i < MM; ++i) tile[(cs - s0 + i) * LS + ll] = y[i];
        }
        __syncthreads();
        int lq = tid / R, rr = tid - lq * R;
        const int qd = nthr / R, rd = nthr - qd * R;
        for (int e0 = tid; e0 < L * R; e0 += kB * nthr) {
            double v[kB];
            unsigned off[kB];
            bool ok[kB];
#pragma unroll
            for (int b = 0; b < kB; ++b) {
                ok[b] = lq < L;
                v[b] = ok[b] ? tile[rr * LS + lq] : 0.0;
                off[b] = (unsigned)(lq * N + s0 + rr) * 8u;
                lq += qd;
                rr += rd;
                if (rr >= R) { rr -= R; ++lq; }
            }
            if (a.accumulate) {
#pragma unroll
                for (int b = 0; b < kB; ++b)
                    if (ok[b]) v[b] = bload_nt(rout, off[b], 0) + v[b];
            }
#pragma unroll
            for (int b = 0; b < kB; ++b)
                if (ok[b]) bstore(v[b], rout, off[b], 0);
        }
    }
}

// lines per workgroup of the contiguous kernel: 16 for periodic lines (every
// table read is chunk independent, so a wave may span chunks; the smaller tile
// lets four workgroups of 256-point lines share a CU, which overlaps one
// workgroup's staging with another's solve: 32 lines measured 4-14% slower,
// 8 lines mixed) and for explicit operators (no tables: Dirichlet CD2 and
// CommN5Exp x 4.7 -> 5.7-5.8 TB/s with 16 lines), 64 for banded Dirichlet
// ones (a wave is then 64 lines x one chunk, so the chunk's factor tables are
// scalar loads and the end-row branches uniform; with 16 lines the per-lane
// table loads made Penta10 x twice slower).  Banded operators on ragged
// lines (RG) read chunk-own tables, periodic ones included: 64 lines.
// Banded periodic lines longer than 256 points: 8 lines (the 16-line tile of
// 576-point lines, 79 KB, let two workgroups share a CU; 8 lines measured
// 13-49% faster there, and mixed, -5% to +8%, at 256 points).
__host__ __device__ constexpr int contig_lines(bool per, int w, bool rg, int n = 0) {
    return w == 0 ? 16 : (per && !rg) ? (n > 256 ? 8 : 16) : 64;
}

// loads in flight per thread while staging.  The tile caps a CU at one to four
// workgroups, so the fill is latency bound: 16 at a time (8 at a time left
// Penta10 x at 3.4 TB/s; the 16-row Dirichlet kernels and the H = 5 periodic
// ones now fit 16 in 78-104 VGPRs, 1-3% faster than 8)
template <int MM>
constexpr int contig_batch() {
    return MM % 16 ? (MM <= 12 ? MM : MM / 2) : 16;
}

template <int W, int H, int MM, bool PER, bool RG, int L>
__global__ __launch_bounds__(1024) void lineop_contig(OpDev op, LaunchArgs a) {
    extern __shared__ double smem[];
    constexpr int LS = L + 1;  // row stride of the transposed tile
    const int N = op.N, P = op.P, NI = op.NI;
    double *tile = smem;  // [N + 2H][LS], row e holds phi[e - H]
    // [L][NI+1] interface values.  Pentadiagonal operators keep them inside the
    // tile, which is dead from the window reads to the write-back: their
    // interface is twice as large, and sharing lets two workgroups fit a CU
    // (two more barriers); tridiagonal ones fit anyway and append it.
    constexpr bool kShare = W == 2;
    double *red = kShare ? smem : smem + (size_t)(N + 2 * H) * LS;
    const int tid = threadIdx.y * L + threadIdx.x;
    const int nthr = L * P;  // L * N = nthr * MM elements (ragged lines: fewer)
    const int LN = L * N;
    const int q0 = blockIdx.x * L;
    // The workgroup's lines are consecutive rows of the field: one contiguous
    // block of L * N doubles, element e at point e % N of line e / N.  Buffer
    // descriptors over that block end at the last line of the field, so the
    // loads of a partial last workgroup past it return 0 and its stores there
    // are dropped.
    const int rem = a.nlines - q0;
    const unsigned nrec = (unsigned)(rem < L ? rem : L) * (unsigned)N * 8u;
    const __amdgpu_buffer_rsrc_t rin = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<double *>(a.in[blockIdx.y]) + (long)q0 * N, (short)0, (int)nrec, kRawBufferWord3);
    // tile slot (p + H) * LS + line of element tid + k * nthr, advanced k by
    // k: by (nthr / N lines, nthr % N points), wrapping the point.  When the
    // chunk rows divide the lines per workgroup, nthr is a whole number of
    // lines and the slot advances by a constant.
    constexpr bool kWhole = !RG && L % MM == 0;
    const int ll0 = tid / N, p0 = tid - ll0 * N;
    const int slot0 = (p0 + H) * LS + ll0;
    const int qd = nthr / N, rd = nthr - qd * N;
    const int step = rd * LS + qd, wrapfix = 1 - N * LS;
    constexpr int kB = contig_batch<MM>();
    // fill: batches of coalesced loads, then their LDS writes
    {
        // periodic halo rows: phi[N - H .. N) above row H, phi[0 .. H) below row N + H
        double hv = 0.0;
        int hslot = -1;
        if (PER && tid < 2 * H * L) {
            const int r = tid / L, ll = tid - r * L;
            hv = bload(rin, (unsigned)(ll * N + (r < H ? N - H + r : r - H)) * 8u, 0);
            hslot = (r < H ? r : N + r) * LS + ll;
        }
        int slot = slot0, pp = p0;
#pragma unroll 1
        for (int k0 = 0; k0 < MM; k0 += kB) {
            double v[kB];
#pragma unroll
            for (int k = 0; k < kB; ++k) v[k] = bload_nt(rin, (unsigned)(tid + (k0 + k) * nthr) * 8u, 0);
#pragma unroll
            for (int k = 0; k < kB; ++k) {
                if (!RG || tid + (k0 + k) * nthr < LN) tile[slot] = v[k];
                if (kWhole) {
                    slot += L / MM;
                } else {
                    pp += rd;
                    slot += step;
                    if (pp >= N) { pp -= N; slot += wrapfix; }
                }
            }
        }
        if (PER) {
            if (hslot >= 0) tile[hslot] = hv;
            for (int e = tid + nthr; e < 2 * H * L; e += nthr) {  // lines shorter than 2H chunks
                const int r = e / L, ll = e - r * L;
                tile[(r < H ? r : N + r) * LS + ll] = bload(rin, (unsigned)(ll * N + (r < H ? N - H + r : r - H)) * 8u, 0);
            }
        } else {  // halo rows: zeros (only end rows reach past the ends, with zero weight)
            for (int e = tid; e < 2 * H * L; e += nthr) {
                const int r = e / L, ll = e - r * L;
                tile[(r < H ? r : N + r) * LS + ll] = 0.0;
            }
        }
    }
    __syncthreads();
    const int l = threadIdx.x;
    const int c = L < 64 ? (int)threadIdx.y : __builtin_amdgcn_readfirstlane(threadIdx.y);
    const int q = q0 + l;
    const bool live = q < a.nlines;
    const int cs = chunk_start<MM, RG>(op, c);
    const bool short_chunk = RG && c == op.cshort;
    const int len = short_chunk ? op.lshort : MM;
    double w[MM + 2 * H];
    const double *trow = tile + (size_t)cs * LS + l;  // row cs holds phi[cs - H]
#pragma unroll
    for (int j = 0; j < MM + 2 * H; ++j) w[j] = trow[j * LS];
    if (kShare) __syncthreads();  // every window is read before the interface values overwrite the tile
    double y[MM];
    chunk_rhs<H, MM, PER, has_rows<W, H, PER>()>(op, w, y, c == 0, c == P - 1);
    if (W > 0) {
        chunk_solve<W, MM>(op.lu + (PER && !RG ? 0 : (size_t)c * MM * kLuRec), y);
        publish<W, MM, RG>(red + l * (NI + 1), c, short_chunk, len, y);
    }
    // every lane has read its window and published its interface values
    __syncthreads();
    if (W > 0) {
        const double *myred = red + l * (NI + 1);
        double s[2 * W > 0 ? 2 * W : 1];
        interface_sums<W, PER, RG>(op, c, myred, s);
        const double *sp = op.sp + (PER && !RG ? 0 : (size_t)c * MM * 2 * W);
#pragma unroll
        for (int i = 0; i < MM; ++i) {
            if (i % 8 == 0) __builtin_amdgcn_sched_barrier(0);
            double x = y[i];
#pragma unroll
            for (int j = 0; j < 2 * W; ++j) x -= sp[i * 2 * W + j] * s[j];
            y[i] = x;
        }
        if (kShare) __syncthreads();  // every interface value is read before the results overwrite the tile
    }
    const bool pass = live && pass_line(op, 0, q, a.Nx, a.Ny);
    if (kShare && pass) {  // the tile no longer holds the input: reload it (nothing is stored yet)
#pragma unroll
        for (int i = 0; i < MM; ++i) y[i] = bload(rin, (unsigned)(l * N + cs + i) * 8u, 0);
    }
    if (live && (kShare || !pass)) {
#pragma unroll
        for (int i = 0; i < MM; ++i)
            if (!RG || i < len) tile[(cs + H + i) * LS + l] = y[i];
    }
    __syncthreads();
    // store: batches of coalesced stores (and loads when accumulating)
    {
        const __amdgpu_buffer_rsrc_t rout = __builtin_amdgcn_make_buffer_rsrc(
            a.out[blockIdx.y] + (long)q0 * N, (short)0, (int)nrec, kRawBufferWord3);
        int slot = slot0, pp = p0;
#pragma unroll 1
        for (int k0 = 0; k0 < MM; k0 += kB) {
            double v[kB];
#pragma unroll
            for (int k = 0; k < kB; ++k) {
                v[k] = !RG || tid + (k0 + k) * nthr < LN ? tile[slot] : 0.0;
                if (kWhole) {
                    slot += L / MM;
                } else {
                    pp += rd;
                    slot += step;
                    if (pp >= N) { pp -= N; slot += wrapfix; }
                }
            }
            if (a.accumulate) {
                double o[kB];
#pragma unroll
                for (int k = 0; k < kB; ++k) o[k] = bload_nt(rout, (unsigned)(tid + (k0 + k) * nthr) * 8u, 0);
#pragma unroll
                for (int k = 0; k < kB; ++k) v[k] = o[k] + v[k];
            }
#pragma unroll
            for (int k = 0; k < kB; ++k) bstore(v[k], rout, (unsigned)(tid + (k0 + k) * nthr) * 8u, 0);
        }
    }
}

// x-lines in 64-line tiles (banded Dirichlet or ragged operators, a wave per
// chunk) whose whole-line tile exceeds LDS: the chunks are staged in
// op.phases groups of cpp chunks, one after another through one tile of
// cpp * MM + 2H rows.  Phase k loads the rows its chunks' windows span (zero or
// wrapped outside the line), its waves read their windows, and the next phase
// reuses the tile; the results go back the same way, a group at a time.  The
// tile holds the rows of a group of lines [row][line] (row stride 65), filled
// and drained as 64 contiguous segments of the workgroup's block with a
// running (line, row) walk.
template <int W, int H, int MM, bool PER, bool RG>
__global__ __launch_bounds__(1024) void lineop_contig_ph(OpDev op, LaunchArgs a) {
    extern __shared__ double smem[];
    constexpr int L = 64, LS = L + 1, kB = 8;
    const int N = op.N, P = op.P, NI = op.NI, K = op.phases;
    const int cpp = (P + K - 1) / K;
    double *tile = smem;
    constexpr bool kShare = W == 2;  // interface values inside the tile (dead by then)
    double *red = kShare ? smem : smem + (size_t)(cpp * MM + 2 * H) * LS;
    const int l = threadIdx.x;
    const int c = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const int tid = threadIdx.y * L + l;
    const int nthr = L * P;
    const int q0 = blockIdx.x * L;
    const int rem = a.nlines - q0;
    const unsigned nrec = (unsigned)(rem < L ? rem : L) * (unsigned)N * 8u;
    const __amdgpu_buffer_rsrc_t rin = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<double *>(a.in[blockIdx.y]) + (long)q0 * N, (short)0, (int)nrec, kRawBufferWord3);
    const int q = q0 + l;
    const bool live = q < a.nlines;
    const int cs = chunk_start<MM, RG>(op, c);
    const bool short_chunk = RG && c == op.cshort;
    const int len = short_chunk ? op.lshort : MM;
    const int myk = c / cpp;
    double w[MM + 2 * H];
    for (int k = 0; k < K; ++k) {
        const int c0 = k * cpp;
        if (c0 >= P) break;
        const int c1 = c0 + cpp < P ? c0 + cpp : P;
        const int r0 = chunk_start<MM, RG>(op, c0) - H;  // line row held by tile row 0
        const int R = chunk_start<MM, RG>(op, c1 - 1) + MM + H - r0;
        if (k > 0) __syncthreads();  // the previous group has read its windows
        {
            int ll = tid / R, rr = tid - ll * R;
            const int qd = nthr / R, rd = nthr - qd * R;
            for (int e0 = tid; e0 < L * R; e0 += kB * nthr) {
                double v[kB];
                int slot[kB];
#pragma unroll
                for (int b = 0; b < kB; ++b) {
                    const int p = r0 + rr;
                    const int pg = PER ? (p < 0 ? p + N : (p >= N ? p - N : p)) : p;
                    const bool in_line = PER || (p >= 0 && p < N);
                    v[b] = (ll < L && in_line) ? bload_nt(rin, (unsigned)(ll * N + pg) * 8u, 0) : 0.0;
                    slot[b] = ll < L ? rr * LS + ll : -1;
                    ll += qd;
                    rr += rd;
                    if (rr >= R) { rr -= R; ++ll; }
                }
#pragma unroll
                for (int b = 0; b < kB; ++b)
                    if (slot[b] >= 0) tile[slot[b]] = v[b];
            }
        }
        __syncthreads();
        if (myk == k) {
            const double *trow = tile + (size_t)(cs - H - r0) * LS + l;
#pragma unroll
            for (int j = 0; j < MM + 2 * H; ++j) w[j] = trow[j * LS];
        }
    }
    if (kShare) __syncthreads();  // every window is read before the interface values overwrite the tile
    double y[MM];
    chunk_rhs<H, MM, PER, has_rows<W, H, PER>()>(op, w, y, c == 0, c == P - 1);
    if (W > 0) {
        chunk_solve<W, MM>(op.lu + (PER && !RG ? 0 : (size_t)c * MM * kLuRec), y);
        publish<W, MM, RG>(red + l * (NI + 1), c, short_chunk, len, y);
    }
    __syncthreads();
    if (W > 0) {
        const double *myred = red + l * (NI + 1);
        double s[2 * W > 0 ? 2 * W : 1];
        interface_sums<W, PER, RG>(op, c, myred, s);
        const double *sp = op.sp + (PER && !RG ? 0 : (size_t)c * MM * 2 * W);
#pragma unroll
        for (int i = 0; i < MM; ++i) {
            if (i % 8 == 0) __builtin_amdgcn_sched_barrier(0);
            double x = y[i];
#pragma unroll
            for (int j = 0; j < 2 * W; ++j) x -= sp[i * 2 * W + j] * s[j];
            y[i] = x;
        }
    }
    // pass-through lines keep their input (nothing is stored yet)
    if (live && pass_line(op, 0, q, a.Nx, a.Ny)) {
#pragma unroll
        for (int i = 0; i < MM; ++i) y[i] = bload(rin, (unsigned)(l * N + cs + i) * 8u, 0);
    }
    const __amdgpu_buffer_rsrc_t rout = __builtin_amdgcn_make_buffer_rsrc(
        a.out[blockIdx.y] + (long)q0 * N, (short)0, (int)nrec, kRawBufferWord3);
    for (int k = 0; k < K; ++k) {
        const int c0 = k * cpp;
        if (c0 >= P) break;
        const int c1 = c0 + cpp < P ? c0 + cpp : P;
        const int s0 = chunk_start<MM, RG>(op, c0);  // line row held by tile row 0
        const int R = chunk_start<MM, RG>(op, c1 - 1) + (RG && c1 - 1 == op.cshort ? op.lshort : MM) - s0;
        __syncthreads();  // interface values read (first group) or the previous group stored
        if (myk == k && live) {
#pragma unroll
            for (int i = 0; i < MM; ++i)
                if (!RG || i < len) tile[(cs - s0 + i) * LS + l] = y[i];
        }
        __syncthreads();
        int ll = tid / R, rr = tid - ll * R;
        const int qd = nthr / R, rd = nthr - qd * R;
        for (int e0 = tid; e0 < L * R; e0 += kB * nthr) {
            double v[kB];
            unsigned off[kB];
            bool ok[kB];
#pragma unroll
            for (int b = 0; b < kB; ++b) {
                ok[b] = ll < L;
                v[b] = ok[b] ? tile[rr * LS + ll] : 0.0;
                off[b] = (unsigned)(ll * N + s0 + rr) * 8u;
                ll += qd;
                rr += rd;
                if (rr >= R) { rr -= R; ++ll; }
            }
            if (a.accumulate) {
#pragma unroll
                for (int b = 0; b < kB; ++b)
                    if (ok[b]) v[b] = bload_nt(rout, off[b], 0) + v[b];
            }
#pragma unroll
            for (int b = 0; b < kB; ++b)
                if (ok[b]) bstore(v[b], rout, off[b], 0);
        }
    }
}

template <int W, int H, int MM, bool PER, bool RG>
hipError_t launch_fast(const OpDev &op, const LaunchArgs &a, int nfields, hipStream_t st) {
    if (op.dir == 0) {
        if constexpr (!PER && !RG && W > 0 && MM == 32) {
            if (lineop_fast_pair(op)) {
                dim3 grid((a.nlines + 31) / 32, nfields), block(64, lineop_pair_waves(op));
                hipLaunchKernelGGL((lineop_contig_pair<W, H, MM>), grid, block, lineop_fast_lds_bytes(op), st, op, a);
                return hipGetLastError();
            }
        }
        constexpr int L = contig_lines(PER, W, RG);
        if constexpr (PER && !RG && W > 0) {
            if (contig_lines(PER, W, RG, op.N) == 8) {
                dim3 grid((a.nlines + 7) / 8, nfields), block(8, op.P);
                hipLaunchKernelGGL((lineop_contig<W, H, MM, PER, RG, 8>), grid, block, lineop_fast_lds_bytes(op), st,
                                   op, a);
                return hipGetLastError();
            }
        }
        dim3 grid((a.nlines + L - 1) / L, nfields), block(L, op.P);
        if constexpr (L == 64) {
            if (op.phases > 1) {
                hipLaunchKernelGGL((lineop_contig_ph<W, H, MM, PER, RG>), grid, block, lineop_fast_lds_bytes(op), st,
                                   op, a);
                return hipGetLastError();
            }
        }
        hipLaunchKernelGGL((lineop_contig<W, H, MM, PER, RG, L>), grid, block, lineop_fast_lds_bytes(op), st, op, a);
    } else {
        if constexpr ((!RG && ((PER && !has_rows<W, H, PER>()) || MM == 16 || MM == 32)) || (RG && !PER)) {
            if (lineop_fast_pair(op)) {
                dim3 grid((a.nlines + 31) / 32, nfields, op.dir == 2 ? a.nbatch : 1), block(64, lineop_pair_waves(op));
                hipLaunchKernelGGL((lineop_strided_pair<W, H, MM, PER, RG>), grid, block, lineop_fast_lds_bytes(op),
                                   st, op, a);
                return hipGetLastError();
            }
        }
        if constexpr (MM == 16) {
            if (op.P > 16) {
                dim3 grid((a.nlines + 63) / 64, nfields, op.dir == 2 ? a.nbatch : 1), block(64, 16);
                hipLaunchKernelGGL((lineop_strided_multi<W, H, MM, PER, RG>), grid, block, lineop_fast_lds_bytes(op),
                                   st, op, a);
                return hipGetLastError();
            }
        }
        dim3 grid((a.nlines + 63) / 64, nfields, op.dir == 2 ? a.nbatch : 1), block(64, op.P);
        hipLaunchKernelGGL((lineop_strided<W, H, MM, PER, RG>), grid, block, lineop_fast_lds_bytes(op), st, op, a);
    }
    return hipGetLastError();
}

template <int W, int H, int MM, bool RG = false>
hipError_t launch_fast_bc(const OpDev &op, const LaunchArgs &a, int nfields, hipStream_t st) {
    return op.periodic ? launch_fast<W, H, MM, true, RG>(op, a, nfields, st)
                       : launch_fast<W, H, MM, false, RG>(op, a, nfields, st);
}

template <int W, int H>
hipError_t launch_fast_mm(const OpDev &op, const LaunchArgs &a, int nfields, hipStream_t st) {
    if (op.cshort >= 0) {  // ragged lines: 16- or 32-row chunks and one short chunk
        if (op.MM == 16) return launch_fast_bc<W, H, 16, true>(op, a, nfields, st);
        if (op.MM == 32) return launch_fast_bc<W, H, 32, true>(op, a, nfields, st);
        return hipErrorInvalidValue;
    }
    switch (op.MM) {
        case 32: return launch_fast_bc<W, H, 32>(op, a, nfields, st);
        case 28: return launch_fast_bc<W, H, 28>(op, a, nfields, st);
        case 24: return launch_fast_bc<W, H, 24>(op, a, nfields, st);
        case 20: return launch_fast_bc<W, H, 20>(op, a, nfields, st);
        case 16: return launch_fast_bc<W, H, 16>(op, a, nfields, st);
        // 8- and 12-row chunks: periodic lines only (not instantiated for
        // Dirichlet lines, which the plan never splits that way)
        case 12: return launch_fast<W, H, 12, true, false>(op, a, nfields, st);
        case 8: return launch_fast<W, H, 8, true, false>(op, a, nfields, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

// launches of one (W, H) pair (lineop_fast_w*h*.hip)
hipError_t lineop_launch_fast_w0h1(const OpDev &op, const LaunchArgs &a, int nfields, hipStream_t st);
hipError_t lineop_launch_fast_w0h3(const OpDev &op, const LaunchArgs &a, int nfields, hipStream_t st);
hipError_t lineop_launch_fast_w1h2(const OpDev &op, const LaunchArgs &a, int nfields, hipStream_t st);
hipError_t lineop_launch_fast_w1h4(const OpDev &op, const LaunchArgs &a, int nfields, hipStream_t st);
hipError_t lineop_launch_fast_w1h5(const OpDev &op, const LaunchArgs &a, int nfields, hipStream_t st);
hipError_t lineop_launch_fast_w2h3(const OpDev &op, const LaunchArgs &a, int nfields, hipStream_t st);
hipError_t lineop_launch_fast_w2h5(const OpDev &op, const LaunchArgs &a, int nfields, hipStream_t st);

}  // namespace mc3d
