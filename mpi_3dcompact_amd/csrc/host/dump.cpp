// SolutionDump reader / writer (layout in dump.hpp).
#include "dump.hpp"

#include <cmath>
#include <cstdio>
#include <memory>
#include <vector>

#include "mc3d.h"
#include "solver.hpp"

namespace mc3d {

namespace {

struct File {
    FILE *f = nullptr;
    File(const std::string &p, const char *mode) : f(std::fopen(p.c_str(), mode)) {}
    ~File() {
        if (f) std::fclose(f);
    }
};

long points(const DumpHeader &h) { return (long)h.Nx * h.Ny * h.Nz; }

}  // namespace

std::string dump_name(int timeStep) { return "SolutionDump." + std::to_string(timeStep); }

DumpHeader read_dump_header(const std::string &path) {
    File f(path, "rb");
    if (!f.f) throw Error(MC3D_EINVAL, "cannot open " + path);
    double c[5];
    if (std::fread(c, sizeof(double), 5, f.f) != 5) throw Error(MC3D_EINVAL, path + ": short header");
    for (double v : c)
        if (!std::isfinite(v)) throw Error(MC3D_EINVAL, path + ": header is not a SolutionDump header");
    DumpHeader h;
    h.timeStep = (int)c[0];
    h.time = c[1];
    h.Nx = (int)c[2];
    h.Ny = (int)c[3];
    h.Nz = (int)c[4];
    if (h.Nx <= 0 || h.Ny <= 0 || h.Nz <= 0 || (double)h.Nx != c[2] || (double)h.Ny != c[3] ||
        (double)h.Nz != c[4])
        throw Error(MC3D_EINVAL, path + ": header dimensions are not positive integers");
    std::fseek(f.f, 0, SEEK_END);
    const long want = (long)sizeof(double) * (5 + (long)kDumpFields * points(h));
    if (std::ftell(f.f) < want)
        throw Error(MC3D_EINVAL, path + ": file shorter than its header's " + std::to_string(kDumpFields) +
                                     " fields of " + std::to_string(h.Nx) + "x" + std::to_string(h.Ny) +
                                     "x" + std::to_string(h.Nz));
    return h;
}

void read_dump_fields(const std::string &path, const DumpHeader &h, double *const fields[kDumpFields]) {
    File f(path, "rb");
    if (!f.f) throw Error(MC3D_EINVAL, "cannot open " + path);
    const long n = points(h);
    for (int k = 0; k < kDumpFields; ++k) {
        if (!fields[k]) continue;
        if (std::fseek(f.f, (long)sizeof(double) * (5 + (long)k * n), SEEK_SET) != 0 ||
            std::fread(fields[k], sizeof(double), (size_t)n, f.f) != (size_t)n)
            throw Error(MC3D_EINVAL, path + ": read of field " + std::to_string(k) + " failed");
    }
}

void write_dump(const std::string &path, const DumpHeader &h, const double *const fields[kDumpFields]) {
    File f(path, "wb");
    if (!f.f) throw Error(MC3D_EINVAL, "cannot create " + path);
    const double c[5] = {(double)h.timeStep, h.time, (double)h.Nx, (double)h.Ny, (double)h.Nz};
    if (std::fwrite(c, sizeof(double), 5, f.f) != 5) throw Error(MC3D_EINVAL, path + ": write failed");
    const long n = points(h);
    std::vector<double> zeros;
    for (int k = 0; k < kDumpFields; ++k) {
        const double *p = fields[k];
        if (!p) {
            zeros.assign((size_t)n, 0.0);
            p = zeros.data();
        }
        if (std::fwrite(p, sizeof(double), (size_t)n, f.f) != (size_t)n)
            throw Error(MC3D_EINVAL, path + ": write failed");
    }
}

}  // namespace mc3d
