// Partitioned line operator: shared host/device description.
//
// An operator maps each line phi[0..N) of a field to out = A^{-1} (B phi),
// where B is an explicit stencil (interior stencil plus special end rows) and
// A is banded with half-bandwidth W (0: explicit, 1: tridiagonal,
// 2: pentadiagonal), cyclic for periodic lines.
//
// The line is split into P chunks of at most MMAX rows, one lane per chunk.
// Each lane solves its diagonal block D_c with a precomputed LU, then the
// chunks are coupled through 2W interface unknowns per chunk:
//     x_c = y_c - S_c s_c,     s_c = G_c yhat
// where y_c = D_c^{-1} r_c, S_c = D_c^{-1} [C_left | C_right] (the "spikes"),
// yhat gathers the first and last W rows of every y_c, and G_c holds the rows
// of the inverse interface matrix that yield chunk c's neighbour values.
// All factors depend only on (scheme, N, boundary type) and are built once on
// the host in long double.
#pragma once

#include <cstdint>

namespace mc3d {

constexpr int kMaxFields = 16;
constexpr int kMMax = 32;      // largest chunk (rows held in registers by one lane)
constexpr int kHMax = 5;       // max interior stencil half width
constexpr int kSpecRows = 5;   // max special rows at each end
constexpr int kSpecW = 16;     // max taps of a special row
constexpr int kRowRec = 3 + kSpecW;  // {offset, step, width, w[16]}
constexpr int kLuRec = 5;      // {l1, l2, u1, u2, invd}
constexpr int kDRowRec = 17;   // dense special row {order, w[2H+1], zeros}

enum OpClass { kDeriv = 0, kFilter = 1 };

// Passed by value to the kernels.
struct OpDev {
    int N;          // line length
    int dir;        // 0,1,2
    int W;          // half bandwidth of A
    int H;          // interior stencil half width (halo)
    int P;          // chunks per line
    int L;          // lines per workgroup
    int NI;         // interface unknowns per line = 2*W*P
    int ext;        // LDS row stride in doubles (>= N + 2H, odd)
    int MM;         // table row stride: rows per chunk capacity (16 or 32)
    int periodic;   // halo by wrap
    int nbl, nbr;   // special rows at the left / right end
    int int_off, int_w;        // interior stencil: first offset, taps
    double int_c[2 * kHMax + 1];
    double dd;      // divisor applied to the stencil sum (1 for filters)
    int divide;     // 1: result /= dd
    // pass-through of lines on Dirichlet faces of the other directions
    int passA, NA, passB, NB;
    const double *rows;   // (nbl+nbr) * kRowRec, taps in summation order
    const double *drows;  // (nbl+nbr) * kDRowRec dense special rows (fast kernels)
    double inv_dd;        // 1/dd when that reciprocal is exact (dd a power of two), else 0
    const double *lu;     // P * MM * kLuRec
    const double *sp;     // P * MM * 2W
    const double *G;      // P * 2W * NI
    int gb;               // interface band in chunks (Plan::gb)
    int cshort, lshort;   // ragged lines: chunk cshort has lshort rows (else cshort = -1)
    int phases;           // x-lines in 64-line tiles: groups of chunks staged one after another (1: all)
    int same_interior;    // interior chunks share bitwise their lu and sp tables (Plan::same_interior)
    const int *cstart;    // P
    const int *clen;      // P
};

}  // namespace mc3d
