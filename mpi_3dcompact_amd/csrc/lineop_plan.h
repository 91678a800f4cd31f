// Host-side plan of a partitioned line operator (tables later copied to the GPU).
#pragma once

#include <string>
#include <vector>

#include "lineop.h"

namespace mc3d {

struct PlanRequest {
    int cls;        // kDeriv or kFilter
    int kind;       // scheme / filter id
    int bc;         // 0 periodic, 1 Dirichlet
    int N;
    double dd;      // derivative spacing
    double alphaF;  // filter coefficient
};

struct Plan {
    int N = 0, W = 0, H = 0, P = 0, L = 0, NI = 0, ext = 0;
    bool periodic = false, divide = false;
    double dd = 1.0;
    int int_off = 0;
    std::vector<double> int_c;
    int nbl = 0, nbr = 0;
    std::vector<double> rows, lu, sp, G;
    std::vector<int> cstart, clen;
};

// Returns 0 or a negative MC3D_E* code with a message in err.
int build_plan(const PlanRequest &rq, Plan &plan, std::string &err);

}  // namespace mc3d
