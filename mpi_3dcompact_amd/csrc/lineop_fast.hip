// Host side of the fast line kernels: LDS size, staging groups, the checks that
// decide whether an operator takes them, and the dispatch to the (W, H) pair's
// translation unit.
#include "lineop_fast_kernels.h"

namespace mc3d {

bool lineop_fast_pair(const OpDev &op) {
    if (op.cshort >= 0)  // ragged Dirichlet y/z-lines, the short chunk a wave of its own (29 chunks in 16 waves)
        return op.dir != 0 && !op.periodic && (op.MM == 16 || op.MM == 32) && op.cshort == 1 && op.P > 16 &&
               op.P <= 29 && (op.W == 0 || op.same_interior);
    if (op.N != op.P * op.MM || op.P <= 16 || op.P > 32) return false;
    if (op.dir == 0)  // banded Dirichlet x-lines of 32-row chunks (lineop_contig_pair; shorter
                      // lines keep the 64-line tiles, in 32-row chunks if need be)
        return !op.periodic && op.W > 0 && op.MM == 32 && op.P <= 30 && op.same_interior;
    if (op.periodic && op.nbl + op.nbr == 0) return true;  // chunk 0's tables serve every chunk
    // special rows: 16- or 32-row chunks, a wave for each end chunk (30 chunks
    // in 16 waves), with no tables or interior ones shared
    return (op.MM == 16 || op.MM == 32) && op.P <= 30 && (op.W == 0 || (!op.periodic && op.same_interior));
}

int lineop_pair_waves(const OpDev &op) {
    if (op.periodic && op.nbl + op.nbr == 0 && op.cshort < 0) return (op.P + 1) / 2;
    const int f = op.cshort >= 0 ? 2 : 1;  // chunks 0 .. f - 1 and P - 1 alone, the rest in pairs
    return f + 1 + (op.P - f) / 2;
}

size_t lineop_fast_lds_bytes(const OpDev &op) {
    if (op.dir == 0 && lineop_fast_pair(op)) {  // 32-line tile of a group of waves, the interface values
        const int nw = lineop_pair_waves(op), K = std::max(op.phases, 1);
        const int wpp = (nw + K - 1) / K;
        const size_t tile = (size_t)(2 * wpp * op.MM + 2 * op.H) * 33, red = (size_t)32 * (op.NI + 1);
        return sizeof(double) * (op.W == 2 ? std::max(tile, red) : tile + red);
    }
    if (op.dir == 0) {  // the tile, with the interface values inside it (W = 2) or after it
        const size_t L = contig_lines(op.periodic, op.W, op.cshort >= 0, op.N);
        const int K = L == 64 && op.phases > 1 ? op.phases : 1;
        const size_t rows = K > 1 ? (size_t)((op.P + K - 1) / K) * op.MM + 2 * op.H : (size_t)op.N + 2 * op.H;
        const size_t tile = rows * (L + 1), red = L * (op.NI + 1);
        return sizeof(double) * (op.W == 2 ? std::max(tile, red) : tile + red);
    }
    // the interface values (32 lines when two chunks share a wave), and beyond
    // 16 chunks in rounds the stash of rows between rounds
    if (lineop_fast_pair(op)) return sizeof(double) * (size_t)32 * (op.NI + 1);
    const size_t stash = op.P > 16 ? (size_t)64 * ((op.P + 15) / 16) * op.H : 0;
    return sizeof(double) * ((size_t)64 * (op.NI + 1) + stash);
}

int lineop_fast_phases(const OpDev &op, size_t lds_limit) {
    if (op.dir == 0 && lineop_fast_pair(op)) {  // groups of waves
        for (int k = 1; k <= lineop_pair_waves(op); ++k) {
            OpDev t = op;
            t.phases = k;
            if (lineop_fast_lds_bytes(t) <= lds_limit) return k;
        }
        return 1;
    }
    if (op.dir != 0 || contig_lines(op.periodic, op.W, op.cshort >= 0) != 64) return 1;
    for (int k = 1; k <= op.P; ++k) {
        OpDev t = op;
        t.phases = k;
        if (lineop_fast_lds_bytes(t) <= lds_limit) return k;
    }
    return 1;
}

bool lineop_fast_supported(const OpDev &op, size_t lds_limit) {
    const int w = op.W, h = op.H;
    const bool pair = (w == 0 && (h == 1 || h == 3)) || (w == 1 && (h == 2 || h == 4 || h == 5)) ||
                      (w == 2 && (h == 3 || h == 5));
    // workgroup size as launched: (lines per workgroup) x P
    const int lines = op.dir != 0 ? 64 : contig_lines(op.periodic, op.W, op.cshort >= 0, op.N);
    // y- and z-lines of more than 16 chunks: evenly chunked ones up to 32
    // chunks two to a wave; else 16-row chunks, 16 chunk waves in rounds (up to
    // 64 chunks; the 32-row chunk state spills in that kernel)
    const bool multi = op.dir != 0 && op.P > 16;
    const int threads = lineop_fast_pair(op) ? 64 * lineop_pair_waves(op)
                        : multi          ? (op.MM == 16 && op.P <= 64 ? 1024 : 2048)
                                         : lines * op.P;
    // special rows: H at each end of a Dirichlet line (every scheme has that
    // many); on periodic lines only where the kernels compile them in (CD2)
    const bool ends = op.periodic ? (op.nbl + op.nbr == 0 || (w == 0 && h == 1 && op.nbl == h && op.nbr == h))
                                  : (op.nbl == h && op.nbr == h);
    const bool rows = chunk_rows_ok(op.MM) && (op.MM >= 16 || op.periodic) && op.MM >= 2 * h;
    // evenly chunked, or ragged: 16- or 32-row chunks and one short second chunk
    const bool chunks = op.cshort < 0 ? op.N == op.P * op.MM
                                      : (op.MM == 16 || op.MM == 32) && op.cshort == 1 && op.P >= 3 &&
                                            op.lshort >= std::max(2 * w, 1) && op.lshort < op.MM &&
                                            op.N == (op.P - 1) * op.MM + op.lshort;
    return pair && ends && rows && chunks && threads <= 1024 && lineop_fast_lds_bytes(op) <= lds_limit;
}

hipError_t lineop_launch_fast(const OpDev &op, const LaunchArgs &a, int nfields, hipStream_t st) {
    switch (op.W * 8 + op.H) {
        case 0 * 8 + 1: return lineop_launch_fast_w0h1(op, a, nfields, st);
        case 0 * 8 + 3: return lineop_launch_fast_w0h3(op, a, nfields, st);
        case 1 * 8 + 2: return lineop_launch_fast_w1h2(op, a, nfields, st);
        case 1 * 8 + 4: return lineop_launch_fast_w1h4(op, a, nfields, st);
        case 1 * 8 + 5: return lineop_launch_fast_w1h5(op, a, nfields, st);
        case 2 * 8 + 3: return lineop_launch_fast_w2h3(op, a, nfields, st);
        case 2 * 8 + 5: return lineop_launch_fast_w2h5(op, a, nfields, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace mc3d
