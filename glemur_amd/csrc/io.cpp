// Run files in the reference's formats (src/io.cu): legacy VTK, BINARY,
// big-endian, and the ASCII time series, so the reference's python/glemur
// readers read them unchanged.  Host code only.
#include "gm_internal.hpp"

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

using gm::fail;

namespace {

// Big-endian encoder for one run of values (the reference swaps per value,
// src/io.cu:86-136).
struct BE {
    std::vector<unsigned char> buf;
    void f64(double v)
    {
        uint64_t u;
        memcpy(&u, &v, 8);
        for (int s = 56; s >= 0; s -= 8) buf.push_back((unsigned char)(u >> s));
    }
    void f32(float v)
    {
        uint32_t u;
        memcpy(&u, &v, 4);
        for (int s = 24; s >= 0; s -= 8) buf.push_back((unsigned char)(u >> s));
    }
    void i32(int v)
    {
        uint32_t u = (uint32_t)v;
        for (int s = 24; s >= 0; s -= 8) buf.push_back((unsigned char)(u >> s));
    }
    void real(double v, bool dbl) { dbl ? f64(v) : f32((float)v); }
    bool flush(FILE *fd)
    {
        bool ok = fwrite(buf.data(), 1, buf.size(), fd) == buf.size();
        buf.clear();
        return ok;
    }
};

double rd_be(const unsigned char *b, bool dbl)
{
    if (dbl) {
        uint64_t u = 0;
        for (int i = 0; i < 8; ++i) u = (u << 8) | b[i];
        double v;
        memcpy(&v, &u, 8);
        return v;
    }
    uint32_t u = 0;
    for (int i = 0; i < 4; ++i) u = (u << 8) | b[i];
    float v;
    memcpy(&v, &u, 4);
    return v;
}

std::string path_in(const char *dir, const char *name)
{
    std::string d = dir && *dir ? dir : ".";
    if (d.back() != '/') d += '/';
    return d + name;
}

const char *real_str(bool dbl) { return dbl ? "double" : "float"; }

}  // namespace

// writeB0, src/io.cu:167-205
extern "C" int gm_write_b0(const char *dir, const gm_params *p, const double *B0, int precision)
{
    if (!p || !B0) return fail(GM_ERR_ARG, "gm_write_b0: null pointer");
    bool dbl = precision != 32;
    std::string fn = path_in(dir, "data/B0.vtk");
    FILE *fd = fopen(fn.c_str(), "w");
    if (!fd) return fail(GM_ERR_IO, "could not open file '" + fn + "'");
    fprintf(fd, "# vtk DataFile Version 2.0\n");
    fprintf(fd, "GLEMuR B0 dump\n");
    fprintf(fd, "BINARY\n");
    fprintf(fd, "DATASET STRUCTURED_POINTS\n");
    fprintf(fd, "DIMENSIONS %9i %9i %9i\n", p->nx + 2, p->ny + 2, p->nz + 2);
    fprintf(fd, "ORIGIN %8.12f %8.12f %8.12f\n", p->Ox - p->dx, p->Oy - p->dy, p->Oz - p->dz);
    fprintf(fd, "SPACING %8.12f %8.12f %8.12f\n", p->Lx / (p->nx - 1), p->Ly / (p->ny - 1), p->Lz / (p->nz - 1));
    fprintf(fd, "POINT_DATA %9i\n", (p->nx + 2) * (p->ny + 2) * (p->nz + 2));
    fprintf(fd, "VECTORS bfield %s\n", real_str(dbl));
    BE be;
    long long n3 = 3 * gm::npad(*p);
    bool ok = true;
    for (long long a = 0; a < n3; ++a) {
        be.real(B0[a], dbl);
        if (be.buf.size() > (1 << 20)) ok &= be.flush(fd);
    }
    ok &= be.flush(fd);
    fclose(fd);
    return ok ? GM_OK : fail(GM_ERR_IO, "short write to '" + fn + "'");
}

// dumpState + writeParams, src/io.cu:209-386, 503-538
extern "C" int gm_dump_state(const char *dir, const gm_params *p, double t, double dt, int n,
                             const double *xb, const double *B, const double *uu, const double *J,
                             const double *detJac, const double *cellVol, const double *convexity,
                             const double *wedgeMin, int precision)
{
    if (!p || !xb || !B) return fail(GM_ERR_ARG, "gm_dump_state: null pointer");
    bool dbl = precision != 32;
    char name[64];
    if (n == -1)
        snprintf(name, sizeof name, "data/save.vtk");
    else
        snprintf(name, sizeof name, "data/dump%d.vtk", n);
    std::string fn = path_in(dir, name);
    FILE *fd = fopen(fn.c_str(), "w");
    if (!fd) return fail(GM_ERR_IO, "could not open file '" + fn + "'");
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    const long long P = gm::npad(*p);
    BE be;
    bool ok = true;
    fprintf(fd, "# vtk DataFile Version 2.0\n");
    fprintf(fd, "GLEMuR data dump\n");
    fprintf(fd, "BINARY\n");
    fprintf(fd, "DATASET STRUCTURED_GRID\n");
    fprintf(fd, "DIMENSIONS %9i %9i %9i\n", nx + 2, ny + 2, nz + 2);
    // FIELD parameters 9 (src/io.cu:503-538)
    fprintf(fd, "FIELD parameters %i\n", 9);
    fprintf(fd, "t 1 1 %s\n", real_str(dbl));
    be.real(t, dbl); ok &= be.flush(fd);
    fprintf(fd, "dt 1 1 %s\n", real_str(dbl));
    be.real(dt, dbl); ok &= be.flush(fd);
    fprintf(fd, "nx_ny_nz 1 3 int\n");
    be.i32(nx); be.i32(ny); be.i32(nz); ok &= be.flush(fd);
    fprintf(fd, "Lx_Ly_Lz 1 3 %s\n", real_str(dbl));
    be.real(p->Lx, dbl); be.real(p->Ly, dbl); be.real(p->Lz, dbl); ok &= be.flush(fd);
    fprintf(fd, "Ox_Oy_Oz 1 3 %s\n", real_str(dbl));
    be.real(p->Ox, dbl); be.real(p->Oy, dbl); be.real(p->Oz, dbl); ok &= be.flush(fd);
    fprintf(fd, "dx_dy_dz 1 3 %s\n", real_str(dbl));
    be.real(p->dx, dbl); be.real(p->dy, dbl); be.real(p->dz, dbl); ok &= be.flush(fd);
    fprintf(fd, "ampl 1 1 %s\n", real_str(dbl));
    be.real(p->ampl, dbl); ok &= be.flush(fd);
    fprintf(fd, "phi1_phi2 1 2 %s\n", real_str(dbl));
    be.real(p->phi1, dbl); be.real(p->phi2, dbl); ok &= be.flush(fd);
    fprintf(fd, "maxError 1 1 %s\n", real_str(dbl));
    be.real(p->maxError, dbl); ok &= be.flush(fd);

    auto vec_block = [&](const double *v) {
        for (long long a = 0; a < 3 * P; ++a) {
            be.real(v[a], dbl);
            if (be.buf.size() > (1 << 20)) ok &= be.flush(fd);
        }
        ok &= be.flush(fd);
    };
    fprintf(fd, "POINTS %9lld %s\n", P, real_str(dbl));
    vec_block(xb);
    fprintf(fd, "POINT_DATA %9lld\n", P);
    fprintf(fd, "VECTORS bfield %s\n", real_str(dbl));
    vec_block(B);
    // the velocity lives on inner points; ghosts are written as zeros
    // (src/io.cu:269-290)
    auto inner_vec = [&](const double *v) {
        for (int k = 0; k < nz + 2; ++k)
            for (int j = 0; j < ny + 2; ++j)
                for (int i = 0; i < nx + 2; ++i) {
                    bool ghost = i == 0 || i == nx + 1 || j == 0 || j == ny + 1 || k == 0 || k == nz + 1;
                    for (int l = 0; l < 3; ++l)
                        be.real(ghost ? 0. : v[l + 3 * ((i - 1) + (long long)nx * ((j - 1) + (long long)ny * (k - 1)))], dbl);
                    if (be.buf.size() > (1 << 20)) ok &= be.flush(fd);
                }
        ok &= be.flush(fd);
    };
    if (p->inertia && uu) {
        fprintf(fd, "VECTORS ufield %s\n", real_str(dbl));
        inner_vec(uu);
    }
    if (p->dumpJ && J) {
        // J lives on inner points; ghosts are written as zeros (src/io.cu:295-317)
        fprintf(fd, "VECTORS jfield %s\n", real_str(dbl));
        inner_vec(J);
    }
    if (p->dumpDetJac && detJac) {
        fprintf(fd, "SCALARS detJac %s\n", real_str(dbl));
        fprintf(fd, "LOOKUP_TABLE default\n");
        for (long long a = 0; a < P; ++a) {
            be.real(detJac[a], dbl);
            if (be.buf.size() > (1 << 20)) ok &= be.flush(fd);
        }
        ok &= be.flush(fd);
    }
    // inner-point scalars, ghosts written as zeros (src/io.cu:334-385)
    auto inner_scalar = [&](const char *nm, const double *v) {
        fprintf(fd, "SCALARS %s %s\n", nm, real_str(dbl));
        fprintf(fd, "LOOKUP_TABLE default\n");
        for (int k = 0; k < nz + 2; ++k)
            for (int j = 0; j < ny + 2; ++j)
                for (int i = 0; i < nx + 2; ++i) {
                    bool ghost = i == 0 || i == nx + 1 || j == 0 || j == ny + 1 || k == 0 || k == nz + 1;
                    be.real(ghost ? 0. : v[(i - 1) + (long long)nx * ((j - 1) + (long long)ny * (k - 1))], dbl);
                    if (be.buf.size() > (1 << 20)) ok &= be.flush(fd);
                }
        ok &= be.flush(fd);
    };
    if (p->dumpCellVol && cellVol) inner_scalar("cellVol", cellVol);
    if (p->dumpConvexity && convexity) inner_scalar("convexity", convexity);
    if (p->dumpWedgeMin && wedgeMin) inner_scalar("wedgeMin", wedgeMin);
    fclose(fd);
    return ok ? GM_OK : fail(GM_ERR_IO, "short write to '" + fn + "'");
}

// readState, src/io.cu:390-499.  FIELD entries are skipped by their declared
// size, so every entry, maxError included, is consumed (the reference stops
// reading inside the maxError bytes, src/io.cu:415-460).
extern "C" int gm_read_state(const char *dir, const gm_params *p, double *xb, double *B0, double *uu, double *t,
                             double *dt, int precision)
{
    (void)precision;
    if (!p || !xb || !B0 || !t || !dt) return fail(GM_ERR_ARG, "gm_read_state: null pointer");
    if (p->inertia && !uu) return fail(GM_ERR_ARG, "gm_read_state: inertia = t needs the velocity array");
    const long long P = gm::npad(*p);
    std::string fn = path_in(dir, "data/save.vtk");
    FILE *fd = fopen(fn.c_str(), "rb");
    if (!fd) return fail(GM_ERR_IO, "could not open file '" + fn + "'");
    char line[512];
    auto getline = [&]() { return fgets(line, sizeof line, fd) != nullptr; };
    for (int i = 0; i < 5; ++i)
        if (!getline()) { fclose(fd); return fail(GM_ERR_IO, fn + ": truncated header"); }
    if (!getline()) { fclose(fd); return fail(GM_ERR_IO, fn + ": no FIELD line"); }
    int nparams = 0;
    if (sscanf(line, "FIELD parameters %d", &nparams) != 1) { fclose(fd); return fail(GM_ERR_IO, fn + ": bad FIELD line"); }
    for (int q = 0; q < nparams; ++q) {
        if (!getline()) { fclose(fd); return fail(GM_ERR_IO, fn + ": truncated FIELD"); }
        char nm[128], ty[32];
        int nc = 0, nt = 0;
        if (sscanf(line, "%127s %d %d %31s", nm, &nc, &nt, ty) != 4) { fclose(fd); return fail(GM_ERR_IO, fn + ": bad FIELD entry"); }
        int sz = strcmp(ty, "double") == 0 ? 8 : 4;
        std::vector<unsigned char> b((size_t)nc * nt * sz);
        if (fread(b.data(), 1, b.size(), fd) != b.size()) { fclose(fd); return fail(GM_ERR_IO, fn + ": truncated FIELD data"); }
        if (!strcmp(nm, "t")) *t = rd_be(b.data(), sz == 8);
        if (!strcmp(nm, "dt")) *dt = rd_be(b.data(), sz == 8);
        if (!strcmp(nm, "nx_ny_nz")) {
            int dims[3];
            for (int i = 0; i < 3; ++i) dims[i] = (int)(((uint32_t)b[4 * i] << 24) | ((uint32_t)b[4 * i + 1] << 16) | ((uint32_t)b[4 * i + 2] << 8) | b[4 * i + 3]);
            if (dims[0] != p->nx || dims[1] != p->ny || dims[2] != p->nz) {
                fclose(fd);
                return fail(GM_ERR_IO, fn + ": grid size differs from params.in");
            }
        }
        if (fgetc(fd) != '\n') fseek(fd, -1, SEEK_CUR);
    }
    if (!getline()) { fclose(fd); return fail(GM_ERR_IO, fn + ": no POINTS"); }
    char ty[32];
    long long np = 0;
    if (sscanf(line, "POINTS %lld %31s", &np, ty) != 2 || np != P) { fclose(fd); return fail(GM_ERR_IO, fn + ": bad POINTS line"); }
    int sz = strcmp(ty, "double") == 0 ? 8 : 4;
    std::vector<unsigned char> b((size_t)3 * P * sz);
    if (fread(b.data(), 1, b.size(), fd) != b.size()) { fclose(fd); return fail(GM_ERR_IO, fn + ": truncated POINTS"); }
    for (long long a = 0; a < 3 * P; ++a) xb[a] = rd_be(b.data() + a * sz, sz == 8);
    if (p->inertia) {
        // the velocity follows B (src/io.cu:466-479).  The blocks are found
        // by their headers: the reference reads B from the start of the
        // "VECTORS bfield" line, one line early, and the velocity shifted by
        // as much; that misalignment is not reproduced
        if (!getline() || strncmp(line, "POINT_DATA", 10) != 0) { fclose(fd); return fail(GM_ERR_IO, fn + ": no POINT_DATA"); }
        if (!getline() || strncmp(line, "VECTORS bfield", 14) != 0) { fclose(fd); return fail(GM_ERR_IO, fn + ": no bfield"); }
        if (fread(b.data(), 1, b.size(), fd) != b.size()) { fclose(fd); return fail(GM_ERR_IO, fn + ": truncated bfield"); }
        if (!getline() || strncmp(line, "VECTORS ufield", 14) != 0) { fclose(fd); return fail(GM_ERR_IO, fn + ": no ufield"); }
        if (fread(b.data(), 1, b.size(), fd) != b.size()) { fclose(fd); return fail(GM_ERR_IO, fn + ": truncated ufield"); }
        const int nx = p->nx, ny = p->ny, nz = p->nz;
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    const long long a = (i + 1) + (long long)(nx + 2) * ((j + 1) + (long long)(ny + 2) * (k + 1));
                    const long long q = i + (long long)nx * (j + (long long)ny * k);
                    for (int l = 0; l < 3; ++l) uu[3 * q + l] = rd_be(b.data() + (3 * a + l) * sz, sz == 8);
                }
    }
    fclose(fd);

    std::string fb = path_in(dir, "data/B0.vtk");
    fd = fopen(fb.c_str(), "rb");
    if (!fd) return fail(GM_ERR_IO, "could not open file '" + fb + "'");
    for (int i = 0; i < 9; ++i)
        if (!getline()) { fclose(fd); return fail(GM_ERR_IO, fb + ": truncated header"); }
    sz = strstr(line, "double") ? 8 : 4;
    if (fread(b.data(), 1, (size_t)3 * P * sz, fd) != (size_t)3 * P * sz) { fclose(fd); return fail(GM_ERR_IO, fb + ": truncated data"); }
    for (long long a = 0; a < 3 * P; ++a) B0[a] = rd_be(b.data() + a * sz, sz == 8);
    fclose(fd);
    return GM_OK;
}

// writeTs, src/io.cu:542-678.  jxb_jb_col adds two columns after the
// reference's: JxB_JBMax and JxBMax_JBMax (python/glemur/ts.py takes the
// names from the header).
extern "C" int gm_write_ts(const char *dir, const gm_params *p, const gm_reduction *red, double t, double dt,
                           int it, double maxDelta, int echo, int jxb_jb_col)
{
    if (!p) return fail(GM_ERR_ARG, "gm_write_ts: null pointer");
    std::string fn = path_in(dir, "data/time_series.dat");
    struct Col {
        bool on;
        const char *name;
        double v;
    };
    gm_reduction z{};
    const gm_reduction &r = red ? *red : z;
    const Col cols[] = {
        {p->redJMax, "JMax", r.JMax},
        {p->redJxB_B2Max, "JxB_B2Max", r.JxB_B2Max},
        {p->redEpsilonStar, "epsilonStar", r.epsilonStar},
        {p->redErrB_1ez, "errB_1ez", r.errB_1ez},
        {p->redErrXb_XbAn, "errXb_XbAn", r.errXb_XbAn},
        {p->redB2, "B2", r.B2},
        {p->redB2f, "B2f", r.B2 - r.B2res},
        {p->redConvex, "convex", r.convex},
        {p->redWedgeMin, "wedgeMin", r.wedgeMin},
        {p->redU2 && p->inertia, "U2", r.U2},
        {p->redUMax && p->inertia, "UMax", r.UMax},
        {jxb_jb_col != 0, "JxB_JBMax", r.JxB_JBMax},
        {jxb_jb_col != 0, "JxBMax_JBMax", r.JxBMax_JBMax},
    };
    FILE *fd = fopen(fn.c_str(), t == 0 ? "w+" : "a");
    if (!fd) return fail(GM_ERR_IO, "could not open file '" + fn + "'");
    if (t == 0) {
        if (echo) {
            printf("#%7s", "it");
            printf("%13s", "t");
            printf("%13s", "dt");
            printf("%13s", "maxError");
            for (const Col &c : cols)
                if (c.on) printf("%13s", c.name);
            printf("\n");
        }
        fprintf(fd, "#%7s", "it");
        fprintf(fd, "%13s", "t");
        fprintf(fd, "%13s", "dt");
        fprintf(fd, "%13s", "maxDelta");
        for (const Col &c : cols)
            if (c.on) fprintf(fd, "%13s", c.name);
        fprintf(fd, "\n");
    } else {
        if (echo) {
            printf("%8i ", it);
            printf("%12.5e ", t);
            printf("%12.5e ", dt);
            printf("%12.5e ", maxDelta);
            for (const Col &c : cols)
                if (c.on) printf("%12.5e ", c.v);
            printf("\n");
        }
        fprintf(fd, "%8i ", it);
        fprintf(fd, "%12.5e ", t);
        fprintf(fd, "%12.5e ", dt);
        fprintf(fd, "%12.5e ", maxDelta);
        for (const Col &c : cols)
            if (c.on) fprintf(fd, "%12.5e ", c.v);
        fprintf(fd, "\n");
    }
    fclose(fd);
    return GM_OK;
}
