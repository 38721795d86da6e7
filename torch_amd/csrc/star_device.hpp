// star_device.hpp -- per-cell geometry of a star on the grid (device code).
//
//   star_neighbours   Grid::calculateNearestNeighbours  (src/Fluid/Grid.cpp:922-951)
//                     with Grid::getRayPlane (:909-920) and traverse3D (:167-178)
//   star_path_length  Radiation::cellPathLength         (src/Integrators/Radiation.cpp:102-117)
//   star_shell_volume Radiation::shellVolume            (Radiation.cpp:119-138)
//   star_wind_cell    the CausalWind test of Fluid::initialiseGrid (src/Fluid/Fluid.cpp:41-54)
//   star_fuv          F_FUV of Thermodynamics::preTimeStepCalculations
//                     (src/Integrators/Thermodynamics.cpp:237-243, fluxFUV :113-118)
//
// Every value is computed from the cell's indices when needed (no per-cell
// geometry arrays in HBM).  The arithmetic follows the reference's
// expression order, so that with -ffp-contract=off the results are the
// reference's bit for bit.
#pragma once

#include "torchgpu_internal.hpp"

namespace tg {

// offset of cell (i, j, k) from the star along axis a, cell units (exact)
__device__ __forceinline__ double star_offset(const StarGeom& s, const int ijk[3], int a) {
    return ((double)ijk[a] + 0.5) - ((double)s.s[a] + 0.5);
}

__device__ __forceinline__ long long star_flat(const StarGeom& s, int i, int j, int k) {
    return i + (long long)s.n[0] * (j + (long long)s.n[1] * k);
}

// The four upstream neighbours of a cell (flat indices, -1 off the grid) and
// their interpolation weights.  The ray plane is the axis of largest offset
// (first on ties); the neighbours step one cell towards the star along it
// and, in the other two axes, zero or one cell towards the star.
__device__ __forceinline__ void star_neighbours(const StarGeom& s, const int ijk[3], long long nb[4], double w[4]) {
    int plane = 0;
    double m = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double e = fabs(star_offset(s, ijk, a));
        if (e > m) {
            m = e;
            plane = a;
        }
    }
    const int ax[3] = {(plane + 1) % 3, (plane + 2) % 3, plane};
    double d[3];
    int sg[3], step[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        d[a] = star_offset(s, ijk, ax[a]);
        sg[a] = d[a] < -1.0 / 10.0 ? -1 : 1;
        step[a] = fabs(d[a]) < 1.0 / 10.0 ? 0 : -sg[a];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        // q = 0: plane axis only; 1: + second axis; 2: + first axis; 3: all three
        int c[3] = {ijk[0], ijk[1], ijk[2]};
        c[ax[2]] += step[2];
        if (q == 1 || q == 3) c[ax[1]] += step[1];
        if (q == 2 || q == 3) c[ax[0]] += step[0];
        const bool in = c[0] >= 0 && c[0] < s.n[0] && c[1] >= 0 && c[1] < s.n[1] && c[2] >= 0 && c[2] < s.n[2];
        nb[q] = in ? star_flat(s, c[0], c[1], c[2]) : -1;
    }
    // the ray through the cell centre crosses the upstream face at
    // (ic0, ic1); the weights split it bilinearly between the neighbours
    const double i0 = (double)ijk[ax[0]], i1 = (double)ijk[ax[1]];
    const double ic0 = i0 - 0.5 * ((double)sg[2] * d[0] / d[2]);
    const double ic1 = i1 - 0.5 * ((double)sg[2] * d[1] / d[2]);
    const double del0 = fabs(2.0 * ic0 - 2.0 * i0 + (double)sg[0]);
    const double del1 = fabs(2.0 * ic1 - 2.0 * i1 + (double)sg[1]);
    const bool f0 = fabs(d[0]) > 0.9, f1 = fabs(d[1]) > 0.9, f2 = fabs(d[2]) > 0.9;
    w[0] = f2 ? del0 * del1 : 0.0;
    w[1] = (f1 && f2) ? del0 * (1.0 - del1) : 0.0;
    w[2] = (f0 && f2) ? (1.0 - del0) * del1 : 0.0;
    w[3] = (f0 && f1 && f2) ? (1.0 - del0) * (1.0 - del1) : 0.0;
}

// The same neighbours as coordinates (valid[q] false off the grid).
__device__ __forceinline__ void star_neighbour_cells(const StarGeom& s, const int ijk[3], int nbc[4][3], bool valid[4],
                                                     double w[4]) {
    int plane = 0;
    double m = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double e = fabs(star_offset(s, ijk, a));
        if (e > m) {
            m = e;
            plane = a;
        }
    }
    const int ax[3] = {(plane + 1) % 3, (plane + 2) % 3, plane};
    double d[3];
    int sg[3], step[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        d[a] = star_offset(s, ijk, ax[a]);
        sg[a] = d[a] < -1.0 / 10.0 ? -1 : 1;
        step[a] = fabs(d[a]) < 1.0 / 10.0 ? 0 : -sg[a];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        nbc[q][0] = ijk[0];
        nbc[q][1] = ijk[1];
        nbc[q][2] = ijk[2];
        nbc[q][ax[2]] += step[2];
        if (q == 1 || q == 3) nbc[q][ax[1]] += step[1];
        if (q == 2 || q == 3) nbc[q][ax[0]] += step[0];
        valid[q] = nbc[q][0] >= 0 && nbc[q][0] < s.n[0] && nbc[q][1] >= 0 && nbc[q][1] < s.n[1] && nbc[q][2] >= 0 &&
                   nbc[q][2] < s.n[2];
    }
    const double i0 = (double)ijk[ax[0]], i1 = (double)ijk[ax[1]];
    const double ic0 = i0 - 0.5 * ((double)sg[2] * d[0] / d[2]);
    const double ic1 = i1 - 0.5 * ((double)sg[2] * d[1] / d[2]);
    const double del0 = fabs(2.0 * ic0 - 2.0 * i0 + (double)sg[0]);
    const double del1 = fabs(2.0 * ic1 - 2.0 * i1 + (double)sg[1]);
    const bool f0 = fabs(d[0]) > 0.9, f1 = fabs(d[1]) > 0.9, f2 = fabs(d[2]) > 0.9;
    w[0] = f2 ? del0 * del1 : 0.0;
    w[1] = (f1 && f2) ? del0 * (1.0 - del1) : 0.0;
    w[2] = (f0 && f2) ? (1.0 - del0) * del1 : 0.0;
    w[3] = (f0 && f1 && f2) ? (1.0 - del0) * (1.0 - del1) : 0.0;
}

// Path length of the ray through the cell: distance / largest |offset|
// (ties go to z, then y); 0 at the star.
__device__ __forceinline__ double star_path_length(const StarGeom& s, const int ijk[3]) {
    const double e0 = fabs(star_offset(s, ijk, 0)), e1 = fabs(star_offset(s, ijk, 1)),
                 e2 = fabs(star_offset(s, ijk, 2));
    double denom;
    if (e2 >= e1 && e2 >= e0) denom = e2;
    else if (e1 > e2 && e1 >= e0) denom = e1;
    else denom = e0;
    if (denom == 0) return 0.0;
    const double dx = s.dx;
    return sqrt(e0 * e0 * dx * dx + e1 * e1 * dx * dx + e2 * e2 * dx * dx) / denom;
}

// squared distance to the star in code units
__device__ __forceinline__ double star_r2(const StarGeom& s, const int ijk[3]) {
    double r2 = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double e = star_offset(s, ijk, a);
        r2 += e * e * s.dx * s.dx;
    }
    return r2;
}

__device__ __forceinline__ double star_shell_volume(const StarGeom& s, const int ijk[3], double ds) {
    const double r2 = star_r2(s, ijk);
    if (r2 <= 100.0 * ds * ds) {
        const double r = sqrt(r2), rmin = r - 0.5 * ds, rmax = r + 0.5 * ds;
        return 4.0 * s.pi * (rmax * rmax * rmax - rmin * rmin * rmin) / 3.0;
    }
    return 4.0 * s.pi * r2 * ds;
}

__device__ __forceinline__ bool star_wind_cell(const StarGeom& s, const int ijk[3]) {
    double d2 = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) d2 += (double)(ijk[a] - s.s[a]) * (((double)ijk[a] + 0.5) - (double)s.s[a]);
    const bool here = ijk[0] == s.s[0] && ijk[1] == s.s[1] && ijk[2] == s.s[2];
    return d2 <= (double)(s.wind_r * s.wind_r) || here;
}

// Cell t of a shell's enumeration (face by face, the face's first axis
// fastest).
// A face holds at most n^2 < 2^32 cells, so the position within it is split
// with 32-bit unsigned division (the 64-bit one is a long call chain).
__device__ __forceinline__ void shell_cell(const ShellFaces& f, long long t, int ijk[3]) {
    int q = 0;
    while (q + 1 < f.nf && t >= f.start[q + 1]) ++q;
    const unsigned int l = (unsigned int)(t - f.start[q]);
    const unsigned int w = (unsigned int)f.w0[q];
    const unsigned int l1 = l / w;
    const int a = f.axis[q];
    ijk[a] = f.fixed[q];
    ijk[a == 0 ? 1 : 0] = f.lo0[q] + (int)(l - l1 * w);
    ijk[a == 2 ? 1 : 2] = f.lo1[q] + (int)l1;
}

// Indices of flat cell c of a grid with fewer than 2^32 cells (32-bit
// unsigned division; the handles that call it check the grid size).
__device__ __forceinline__ void cell_ijk32(const StarGeom& s, long long c, int ijk[3]) {
    const unsigned int u = (unsigned int)c, n0 = (unsigned int)s.n[0], n1 = (unsigned int)s.n[1];
    const unsigned int r = u / n0, k = r / n1;
    ijk[0] = (int)(u - r * n0);
    ijk[1] = (int)(r - k * n1);
    ijk[2] = (int)k;
}

__device__ __forceinline__ double star_fuv(const StarGeom& s, const int ijk[3]) {
    const double r2 = star_r2(s, ijk);
    return r2 != 0 ? (0.5 * s.photon_rate) / (1.2e7 * 4 * s.pi * r2) : 0.0;
}

}  // namespace tg
