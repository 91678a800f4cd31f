// Host-side plan of a partitioned line operator (tables later copied to the GPU).
#pragma once

#include <string>
#include <vector>

#include "lineop.h"

namespace mc3d {

// rows per chunk the plans and the fast kernels are built for (the LDS-tiled
// kernel takes 16 and 32 only)
inline bool chunk_rows_ok(int mm) {
    return mm == 8 || mm == 12 || mm == 16 || mm == 20 || mm == 24 || mm == 28 || mm == 32;
}

struct PlanRequest {
    int cls;        // kDeriv or kFilter
    int kind;       // scheme / filter id
    int bc;         // 0 periodic, 1 Dirichlet
    int N;
    double dd;      // derivative spacing
    double alphaF;  // filter coefficient
    int mm = 0;     // rows per chunk (16 or 32); 0 chooses
};

struct Plan {
    int N = 0, W = 0, H = 0, P = 0, L = 0, NI = 0, ext = 0;
    int MM = 16;         // table row stride (rows per chunk capacity)
    bool uniform = false;  // every chunk has MM rows
    // ragged lines: every chunk has MM rows but chunk cshort (1), which has
    // lshort < MM; cshort = -1 otherwise
    int cshort = -1, lshort = 0;
    bool dense_ok = false;  // special rows fit the (2H+1)-wide dense form
    // chunks 1 .. P - 2 (ragged lines: 2 .. P - 2) hold bitwise the same factor
    // and spike tables (their blocks are interior rows of the matrix); false
    // when there are no tables
    bool same_interior = false;
    bool periodic = false, divide = false;
    double dd = 1.0;
    int int_off = 0;
    std::vector<double> int_c;
    int nbl = 0, nbr = 0;
    std::vector<double> rows, lu, sp, G;
    // interface band: every entry of G_c outside the column blocks of chunks
    // c - gb .. c + gb (cyclically on periodic lines) is below 1e-17 of the
    // largest entry of G_c, so the fast kernels skip those blocks
    int gb = 0;
    // special rows as dense rows of K = 2H + 1 taps (the fast kernels' window
    // holds every point they read): {order (+1 increasing, -1 decreasing),
    // w[K], zeros}.  Dirichlet left rows read phi[0..K), right rows
    // phi[N-K..N); periodic rows (the wrap rows of CD2) phi[p-H..p+H] around
    // their own point p
    std::vector<double> drows;
    std::vector<int> cstart, clen;
};

// Returns 0 or a negative MC3D_E* code with a message in err.
int build_plan(const PlanRequest &rq, Plan &plan, std::string &err);

}  // namespace mc3d
