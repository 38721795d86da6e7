// snapshot.cpp -- the reference's data2D snapshot text format (host code).
//
//   torchgpu_snapshot_write   DataPrinter::print2D   (src/IO/DataPrinter.cpp:139-174)
//   torchgpu_snapshot_header  DataReader::readDataParameters (src/IO/DataReader.cpp:17-68)
//   torchgpu_snapshot_read    DataReader::readGrid   (src/IO/DataReader.cpp:70-125)
//   torchgpu_snapshot_patch   DataReader::patchGrid  (src/IO/DataReader.cpp:126-231)
//   torchgpu_heating_write    DataPrinter::printHeating (src/IO/DataPrinter.cpp:325-354)
//
// Format: the time (s), then nx, ny and nz, one per line; then one line per
// cell, x fastest, with the nd cell-centre coordinates (cm), density
// (g/cm^3), pressure (dyn/cm^2), HII fraction and the nd velocity components
// (cm/s), tab separated, each double as std::scientific with precision 10
// (printf "%.10e").  nd = 1 when ny == 1, 2 when nz == 1, else 3.  The writer
// gzips like the reference's OutputStreamGZ when asked; the readers take
// plain or gzip text (zlib reads both).
#include <zlib.h>

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <string>
#include <vector>

#include "../../include/torchgpu.h"

namespace {

struct TextOut {
    FILE* f = nullptr;
    gzFile z = nullptr;
    bool open(const char* path, bool gz) {
        if (gz) z = gzopen(path, "wb6");
        else f = std::fopen(path, "wb");
        return z || f;
    }
    bool write(const char* s, size_t n) {
        if (z) return gzwrite(z, s, (unsigned)n) == (int)n;
        return std::fwrite(s, 1, n, f) == n;
    }
    bool close() {
        bool ok = true;
        if (z) ok = gzclose(z) == Z_OK;
        if (f) ok = std::fclose(f) == 0;
        z = nullptr;
        f = nullptr;
        return ok;
    }
};

// Whitespace-separated tokens of a plain or gzip text file.
struct TokenIn {
    gzFile z = nullptr;
    std::string buf;
    size_t pos = 0;
    bool eof = false;
    bool open(const char* path) {
        z = gzopen(path, "rb");
        return z != nullptr;
    }
    ~TokenIn() {
        if (z) gzclose(z);
    }
    bool fill() {
        if (eof) return false;
        char tmp[1 << 16];
        const int n = gzread(z, tmp, sizeof tmp);
        if (n <= 0) {
            eof = true;
            return false;
        }
        buf.erase(0, pos);
        pos = 0;
        buf.append(tmp, (size_t)n);
        return true;
    }
    bool next(std::string& tok) {
        tok.clear();
        for (;;) {
            while (pos < buf.size() && std::isspace((unsigned char)buf[pos])) ++pos;
            if (pos < buf.size()) break;
            if (!fill()) return false;
        }
        for (;;) {
            while (pos < buf.size() && !std::isspace((unsigned char)buf[pos])) tok.push_back(buf[pos++]);
            if (pos < buf.size() || !fill()) return !tok.empty();
        }
    }
    bool number(double& v) {
        std::string t;
        if (!next(t)) return false;
        char* end = nullptr;
        v = std::strtod(t.c_str(), &end);
        return end && *end == '\0';
    }
};

int dims_of(int ny, int nz) { return ny == 1 ? 1 : (nz == 1 ? 2 : 3); }

double from_code(double scales[3], double v, double m, double l, double t) {
    double out = 0.0;
    torchgpu_units(scales[0], scales[1], scales[2], v, m, l, t, 0, &out);
    return out;
}

}  // namespace

extern "C" int torchgpu_snapshot_write(const char* path, int gz, double t, const double* Q, int nx, int ny, int nz,
                                       double side_length, double density_scale, double pressure_scale,
                                       double time_scale) {
    if (!path || !Q || nx < 1 || ny < 1 || nz < 1 || !(side_length > 0.0)) return TORCHGPU_ERR_INVALID_ARG;
    double sc[3] = {density_scale, pressure_scale, time_scale};
    const int nd = dims_of(ny, nz);
    const long long n = (long long)nx * ny * nz;
    // Grid::initialise: dx[i] = sideLength / ncells[0] (Grid.cpp:471-472);
    // cell centres xc = index + 0.5 in the nd active dimensions (Grid.cpp:537-541)
    const double dx = side_length / (double)nx;
    // unit factors as Converter::fromCodeUnits applies them (one conversion
    // per value, as the reference prints)
    TextOut out;
    if (!out.open(path, gz != 0)) return TORCHGPU_ERR_INVALID_ARG;
    char line[512];
    int k = std::snprintf(line, sizeof line, "%.10e\n%d\n%d\n%d\n", from_code(sc, t, 0, 0, 1), nx, ny, nz);
    bool ok = out.write(line, (size_t)k);
    for (long long c = 0; c < n && ok; ++c) {
        const long long idx[3] = {c % nx, (c / nx) % ny, c / ((long long)nx * ny)};
        k = 0;
        for (int d = 0; d < nd; ++d)
            k += std::snprintf(line + k, sizeof line - k, "%.10e\t",
                               from_code(sc, ((double)idx[d] + 0.5) * dx, 0, 1, 0));
        k += std::snprintf(line + k, sizeof line - k, "%.10e", from_code(sc, Q[0 * n + c], 1, -3, 0));
        k += std::snprintf(line + k, sizeof line - k, "\t%.10e", from_code(sc, Q[1 * n + c], 1, -1, -2));
        k += std::snprintf(line + k, sizeof line - k, "\t%.10e", Q[2 * n + c]);
        for (int d = 0; d < nd; ++d)
            k += std::snprintf(line + k, sizeof line - k, "\t%.10e", from_code(sc, Q[(4 + d) * n + c], 0, 1, -1));
        line[k++] = '\n';
        ok = out.write(line, (size_t)k);
    }
    ok = out.close() && ok;
    return ok ? TORCHGPU_OK : TORCHGPU_ERR_INVALID_ARG;
}

extern "C" int torchgpu_heating_write(const char* path, int gz, double t, const double* H, int nx, int ny, int nz,
                                      double density_scale, double pressure_scale, double time_scale) {
    if (!path || !H || nx < 1 || ny < 1 || nz < 1) return TORCHGPU_ERR_INVALID_ARG;
    double sc[3] = {density_scale, pressure_scale, time_scale};
    const int nd = dims_of(ny, nz);
    const long long n = (long long)nx * ny * nz;
    TextOut out;
    if (!out.open(path, gz != 0)) return TORCHGPU_ERR_INVALID_ARG;
    char line[640];
    int k = std::snprintf(line, sizeof line, "%.10e\n%d\n%d\n%d\n", from_code(sc, t, 0, 0, 1), nx, ny, nz);
    bool ok = out.write(line, (size_t)k);
    for (long long c = 0; c < n && ok; ++c) {
        const long long idx[3] = {c % nx, (c / nx) % ny, c / ((long long)nx * ny)};
        k = 0;
        // GridCell::xc as it is, in cell units (the reference does not scale it here)
        for (int d = 0; d < nd; ++d) k += std::snprintf(line + k, sizeof line - k, "%.10e\t", (double)idx[d] + 0.5);
        k += std::snprintf(line + k, sizeof line - k, "%.10e", from_code(sc, H[c], 1, -1, -3));
        for (int i = 1; i < TORCHGPU_HEAT_N; ++i)
            k += std::snprintf(line + k, sizeof line - k, "\t%.10e", from_code(sc, H[i * n + c], 1, -1, -3));
        line[k++] = '\n';
        ok = out.write(line, (size_t)k);
    }
    ok = out.close() && ok;
    return ok ? TORCHGPU_OK : TORCHGPU_ERR_INVALID_ARG;
}

namespace {

// DataReader::readDataParameters (DataReader.cpp:17-68): time, cell counts,
// nd and dx, the smallest nonzero coordinate difference of the first two
// cells listed.
int read_header(const char* path, double* t, int n[3], int* nd, double* dx) {
    TokenIn in;
    if (!in.open(path)) return TORCHGPU_ERR_INVALID_ARG;
    double v[4];
    for (double& x : v)
        if (!in.number(x)) return TORCHGPU_ERR_INVALID_ARG;
    *t = v[0];
    for (int a = 0; a < 3; ++a) n[a] = (int)v[1 + a];
    *nd = dims_of(n[1], n[2]);
    double x0[3] = {0, 0, 0}, x1[3] = {0, 0, 0}, skip;
    for (int d = 0; d < *nd; ++d)
        if (!in.number(x0[d])) return TORCHGPU_ERR_INVALID_ARG;
    for (int i = 0; i < 3 + *nd; ++i)
        if (!in.number(skip)) return TORCHGPU_ERR_INVALID_ARG;
    for (int d = 0; d < *nd; ++d)
        if (!in.number(x1[d])) return TORCHGPU_ERR_INVALID_ARG;
    bool any = false;
    double m = 0.0;
    for (int d = 0; d < *nd; ++d) {
        const double e = std::abs(x1[d] - x0[d]);
        if (e > std::numeric_limits<double>::min() && (!any || e < m)) {
            m = e;
            any = true;
        }
    }
    if (!any) return TORCHGPU_ERR_INVALID_ARG;
    *dx = m;
    return TORCHGPU_OK;
}

}  // namespace

extern "C" int torchgpu_snapshot_header(const char* path, double* t, int* nx, int* ny, int* nz, int* nd,
                                        double* side_length) {
    if (!path || !t || !nx || !ny || !nz || !nd || !side_length) return TORCHGPU_ERR_INVALID_ARG;
    int n[3];
    double dx;
    const int rc = read_header(path, t, n, nd, &dx);
    if (rc) return rc;
    *nx = n[0];
    *ny = n[1];
    *nz = n[2];
    *side_length = dx * n[0];
    return TORCHGPU_OK;
}

extern "C" int torchgpu_snapshot_patch(const char* path, double* Q, int nx, int ny, int nz, double grid_dx, int ox,
                                       int oy, int oz) {
    // DataReader::patchGrid (DataReader.cpp:126-231): the patch's cells
    // (coordinates / grid_dx + offset, truncated) first zero the grid cells
    // they land in, then add their values / r, r = (grid_dx / patch dx)^nd
    if (!path || !Q || nx < 1 || ny < 1 || nz < 1 || !(grid_dx > 0.0)) return TORCHGPU_ERR_INVALID_ARG;
    double t, pdx;
    int pn[3], nd;
    int rc = read_header(path, &t, pn, &nd, &pdx);
    if (rc) return rc;
    if (nd != dims_of(ny, nz)) return TORCHGPU_ERR_INVALID_ARG;  // "patch ndims != grid ndims"
    const double ratio = grid_dx / pdx;
    if (std::abs(ratio - (int)ratio) > std::numeric_limits<double>::min()) return TORCHGPU_ERR_INVALID_ARG;
    for (int a = 0; a < 3; ++a) {
        const double covered = pn[a] / ratio;
        if (std::abs(covered - (int)covered) > std::numeric_limits<double>::min()) return TORCHGPU_ERR_INVALID_ARG;
    }
    const int r = (int)std::pow((int)(ratio + 0.5), nd);
    const int off[3] = {ox, oy, oz};
    const long long n = (long long)nx * ny * nz;
    const long long npatch = (long long)pn[0] * pn[1] * pn[2];
    for (int pass = 0; pass < 2; ++pass) {
        TokenIn in;
        if (!in.open(path)) return TORCHGPU_ERR_INVALID_ARG;
        double skip;
        for (int i = 0; i < 4; ++i)
            if (!in.number(skip)) return TORCHGPU_ERR_INVALID_ARG;
        for (long long p = 0; p < npatch; ++p) {
            int xc[3] = {0, 0, 0};
            for (int d = 0; d < nd; ++d) {
                double x;
                if (!in.number(x)) return TORCHGPU_ERR_INVALID_ARG;
                xc[d] = (int)((x / grid_dx) + off[d]);
            }
            double val[6] = {0, 0, 0, 0, 0, 0};
            for (int q = 0; q < 3 + nd; ++q)
                if (!in.number(val[q])) return TORCHGPU_ERR_INVALID_ARG;
            if (xc[0] < 0 || xc[0] > nx) continue;
            if (xc[0] >= nx || xc[1] < 0 || xc[1] >= ny || xc[2] < 0 || xc[2] >= nz) continue;  // Grid::locate
            const long long c = xc[0] + (long long)nx * (xc[1] + (long long)ny * xc[2]);
            if (pass == 0) {
                Q[0 * n + c] = 0;
                Q[1 * n + c] = 0;
                Q[2 * n + c] = 0;
                for (int d = 0; d < nd; ++d) Q[(4 + d) * n + c] = 0;
            } else {
                Q[0 * n + c] += val[0] / r;
                Q[1 * n + c] += val[1] / r;
                Q[2 * n + c] += val[2] / r;
                for (int d = 0; d < nd; ++d) Q[(4 + d) * n + c] += val[3 + d] / r;
            }
        }
    }
    return TORCHGPU_OK;
}

extern "C" int torchgpu_snapshot_read(const char* path, double* Q, int nx, int ny, int nz) {
    if (!path || !Q || nx < 1 || ny < 1 || nz < 1) return TORCHGPU_ERR_INVALID_ARG;
    double t, side;
    int hx, hy, hz, nd;
    int rc = torchgpu_snapshot_header(path, &t, &hx, &hy, &hz, &nd, &side);
    if (rc) return rc;
    const double dx = side / hx;
    TokenIn in;
    if (!in.open(path)) return TORCHGPU_ERR_INVALID_ARG;
    double skip;
    for (int i = 0; i < 4; ++i)
        if (!in.number(skip)) return TORCHGPU_ERR_INVALID_ARG;
    // one line per cell of this grid, as DataReader::readGrid reads them
    const long long n = (long long)nx * ny * nz;
    for (long long c = 0; c < n; ++c) {
        int xc[3] = {0, 0, 0};
        for (int d = 0; d < nd; ++d) {
            double x;
            if (!in.number(x)) return TORCHGPU_ERR_INVALID_ARG;
            xc[d] = (int)(x / dx);
        }
        double rho, pre, hii, vel[3] = {0, 0, 0};
        if (!in.number(rho) || !in.number(pre) || !in.number(hii)) return TORCHGPU_ERR_INVALID_ARG;
        for (int d = 0; d < nd; ++d)
            if (!in.number(vel[d])) return TORCHGPU_ERR_INVALID_ARG;
        // Grid::locate: cells outside this grid are skipped
        if (xc[0] < 0 || xc[0] >= nx || xc[1] < 0 || xc[1] >= ny || xc[2] < 0 || xc[2] >= nz) continue;
        const long long i = xc[0] + (long long)nx * (xc[1] + (long long)ny * xc[2]);
        Q[0 * n + i] = rho;
        Q[1 * n + i] = pre;
        Q[2 * n + i] = hii;
        for (int d = 0; d < nd; ++d) Q[(4 + d) * n + i] = vel[d];
    }
    return TORCHGPU_OK;
}
