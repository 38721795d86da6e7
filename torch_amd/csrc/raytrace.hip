// raytrace.hip -- the causal column-density ray trace of the star-on thermo
// operator (Thermodynamics::rayTrace + updateColDen,
// src/Integrators/Thermodynamics.cpp:339-417) and the star geometry kernel.
//
// The reference walks the cells in a causal order (Grid::buildCausal) so
// that a cell's four upstream neighbours (Grid::calculateNearestNeighbours)
// are done before it.  Those neighbours always lie exactly one Chebyshev
// shell (max_d |i_d - star_d|) closer to the star: one step along the ray
// plane and at most one along the other axes, whose offsets are no larger.
// So the GPU sweeps the shells outwards, one launch per shell, every cell
// of a shell in parallel: shell k reads only shell k - 1, written by the
// previous launch on the same stream.  Neighbours of zero weight may be in
// the same shell (ties between axes); their values are multiplied by zero
// and drop out, as in the reference, so a concurrent write there cannot
// change a result (the buffers are kept finite).
//
// Per cell: 8 B density read, 2 x 8 B written, and 4 x 16 B neighbour reads
// that hit in L2 (the previous shell was just written).
#include <hip/hip_runtime.h>

#include "star_device.hpp"
#include "torchgpu_internal.hpp"

namespace tg {

// updateColDen (Thermodynamics.cpp:339-363) for the cells of one shell.
//
// The x-fixed faces of a shell have no contiguous direction in the x-fastest
// grid, so their cells also keep COL_DEN and DCOL_DEN in face-contiguous
// buffers xf (per side, y fastest, in a frame centred on the star, two
// shells ping-ponged): every upstream neighbour of nonzero weight of an
// x-face cell is on the same side's x-face of the previous shell.
// Neighbours of zero weight are not read (their term is 0 x a finite
// value, as the reference adds it).
struct XFaces {
    double* col;   // [2 parities][2 sides][w * w]
    double* dcol;
    int K, w;      // frame half-width (largest shell) and w = 2K + 1
};

__device__ __forceinline__ long long xface_index(const StarGeom& s, const XFaces& xf, int parity, int side, int y,
                                                 int z) {
    return ((long long)(parity * 2 + side) * xf.w + (z - s.s[2] + xf.K)) * xf.w + (y - s.s[1] + xf.K);
}

__global__ __launch_bounds__(256) void coldens_shell_kernel(StarGeom s, ShellFaces f, int k, XFaces xf,
                                                            const double* __restrict__ den, double den_floor,
                                                            double* col, double* dcol) {
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    if (t >= f.start[f.nf]) return;
    int ijk[3];
    shell_cell(f, t, ijk);
    const long long c = star_flat(s, ijk[0], ijk[1], ijk[2]);
    const bool xface = ijk[0] - s.s[0] == k || s.s[0] - ijk[0] == k;
    const int side = ijk[0] > s.s[0] ? 1 : 0;
    // the cell's density first: for x-face cells a strided load, whose
    // latency then overlaps the neighbour reads; floored as fixPrimitives
    // floors it (den_floor -HUGE_VAL: as given)
    const double den_raw = den[c];
    const double den_c = den_raw < den_floor ? den_floor : den_raw;
    // distance to the star in cell units
    double dist2 = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double e = star_offset(s, ijk, a);
        dist2 += e * e;
    }
    double v = 0;
    if (dist2 > 0.95 * 0.95) {
        int nbc[4][3];
        bool valid[4];
        double w[4], cd[4], wr[4];
        star_neighbour_cells(s, ijk, nbc, valid, w);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            cd[q] = 0.0;
            if (valid[q] && w[q] != 0.0) {
                if (xface) {
                    const long long i = xface_index(s, xf, (k - 1) & 1, side, nbc[q][1], nbc[q][2]);
                    cd[q] = xf.col[i] + xf.dcol[i];
                } else {
                    const long long i = star_flat(s, nbc[q][0], nbc[q][1], nbc[q][2]);
                    cd[q] = col[i] + dcol[i];
                }
            }
            wr[q] = cd[q] == 0 ? 0.0 : w[q] / cd[q];
        }
        const double sum_w = wr[0] + wr[1] + wr[2] + wr[3];
        if (sum_w != 0) {
#pragma unroll
            for (int q = 0; q < 4; ++q) v += (wr[q] / sum_w) * cd[q];
        }
    }
    const double dc = (den_c / s.hydrogen_mass) * star_path_length(s, ijk);
    col[c] = v;
    dcol[c] = dc;
    if (xface) {
        // the star's cell (k = 0) feeds both sides of shell 1
        for (int sd = (k == 0 ? 0 : side); sd <= (k == 0 ? 1 : side); ++sd) {
            const long long i = xface_index(s, xf, k & 1, sd, ijk[1], ijk[2]);
            xf.col[i] = v;
            xf.dcol[i] = dc;
        }
    }
}

__global__ void star_geometry_kernel(StarGeom s, long long* __restrict__ nb, double* __restrict__ w,
                                     double* __restrict__ ds, double* __restrict__ shell_vol) {
    const long long n = (long long)s.n[0] * s.n[1] * s.n[2];
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (long long)gridDim.x * blockDim.x) {
        const int ijk[3] = {(int)(c % s.n[0]), (int)((c / s.n[0]) % s.n[1]), (int)(c / ((long long)s.n[0] * s.n[1]))};
        long long b[4];
        double wq[4];
        star_neighbours(s, ijk, b, wq);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            nb[4 * c + q] = b[q];
            w[4 * c + q] = wq[q];
        }
        const double d = star_path_length(s, ijk);
        ds[c] = d;
        shell_vol[c] = star_shell_volume(s, ijk, d);
    }
}

// Faces of shell k clipped to the grid (see ShellFaces).
void shell_faces(const StarGeom& s, int k, ShellFaces& f) {
    f.nf = 0;
    f.start[0] = 0;
    auto add = [&](int axis, int fixed, int lo0, int hi0, int lo1, int hi1) {
        const int b0 = axis == 0 ? 1 : 0, b1 = axis == 2 ? 1 : 2;
        if (fixed < 0 || fixed >= s.n[axis]) return;
        lo0 = lo0 < 0 ? 0 : lo0;
        lo1 = lo1 < 0 ? 0 : lo1;
        hi0 = hi0 > s.n[b0] - 1 ? s.n[b0] - 1 : hi0;
        hi1 = hi1 > s.n[b1] - 1 ? s.n[b1] - 1 : hi1;
        if (hi0 < lo0 || hi1 < lo1) return;
        const int q = f.nf++;
        f.axis[q] = axis;
        f.fixed[q] = fixed;
        f.lo0[q] = lo0;
        f.lo1[q] = lo1;
        f.w0[q] = hi0 - lo0 + 1;
        f.start[q + 1] = f.start[q] + (long long)f.w0[q] * (hi1 - lo1 + 1);
    };
    const int x = s.s[0], y = s.s[1], z = s.s[2];
    if (k == 0) {
        add(0, x, y, y, z, z);
        return;
    }
    for (int sgn = -1; sgn <= 1; sgn += 2) add(0, x + sgn * k, y - k, y + k, z - k, z + k);
    for (int sgn = -1; sgn <= 1; sgn += 2) add(1, y + sgn * k, x - k + 1, x + k - 1, z - k, z + k);
    for (int sgn = -1; sgn <= 1; sgn += 2) add(2, z + sgn * k, x - k + 1, x + k - 1, y - k + 1, y + k - 1);
}

int shell_max(const StarGeom& s) {
    int kmax = 0;
    for (int a = 0; a < 3; ++a) {
        const int m = s.s[a] > s.n[a] - 1 - s.s[a] ? s.s[a] : s.n[a] - 1 - s.s[a];
        kmax = m > kmax ? m : kmax;
    }
    return kmax;
}

long long xface_doubles(const StarGeom& s) {
    const long long w = 2LL * shell_max(s) + 1;
    return 2 * 2 * 2 * w * w;  // col and dcol, 2 parities, 2 sides
}

hipError_t launch_thermo_ray_trace(const StarGeom& s, const double* den, double den_floor, double* col, double* dcol,
                                   double* xface, hipStream_t stream) {
    const int kmax = shell_max(s);
    XFaces xf;
    xf.K = kmax;
    xf.w = 2 * kmax + 1;
    xf.col = xface;
    xf.dcol = xface + 2LL * 2 * xf.w * xf.w;
    ShellFaces f;
    for (int k = 0; k <= kmax; ++k) {
        shell_faces(s, k, f);
        const long long cells = f.start[f.nf];
        if (cells == 0) continue;
        const long long nb = (cells + 255) / 256;
        hipLaunchKernelGGL(coldens_shell_kernel, dim3((unsigned)nb), dim3(256), 0, stream, s, f, k, xf, den,
                           den_floor, col, dcol);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_star_geometry(const StarGeom& s, long long* nb, double* w, double* ds, double* shell_vol,
                                hipStream_t stream) {
    const long long n = (long long)s.n[0] * s.n[1] * s.n[2];
    long long b = (n + 255) / 256;
    if (b > 65536) b = 65536;
    hipLaunchKernelGGL(star_geometry_kernel, dim3((unsigned)b), dim3(256), 0, stream, s, nb, w, ds, shell_vol);
    return hipGetLastError();
}

}  // namespace tg
