// SolutionDump files: the reference's restart / output format.
//
// Layout (src/CurvilinearCSolver_Core.cpp:1333-1409, read back by
// src/MPI_3DCompact.cpp:80-126,214-271): five doubles {timeStep, time, Nx, Ny,
// Nz}, then eight Nx*Ny*Nz arrays of doubles, x fastest: x, y, z (mesh
// coordinates), rho, rhoU, rhoV, rhoW, rhoE.  Native byte order (the
// reference writes it with MPI-IO in native representation).
#pragma once

#include <string>

namespace mc3d {

struct DumpHeader {
    int timeStep = 0;
    double time = 0.0;
    int Nx = 0, Ny = 0, Nz = 0;
};

constexpr int kDumpFields = 8;  // x y z rho rhoU rhoV rhoW rhoE

// "SolutionDump.<timeStep>", the reference's file name
std::string dump_name(int timeStep);

// All three throw mc3d::Error (MC3D_EINVAL) on I/O or format errors.
DumpHeader read_dump_header(const std::string &path);
// fields[k] (k < 8, null entries skipped) receive array k of the file
void read_dump_fields(const std::string &path, const DumpHeader &h, double *const fields[kDumpFields]);
// null entries are written as zeros
void write_dump(const std::string &path, const DumpHeader &h, const double *const fields[kDumpFields]);

}  // namespace mc3d
