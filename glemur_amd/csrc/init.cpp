// Host initial state: B0 on the padded grid and the undistorted (or
// initially distorted) positions, in the reference's AoS padded layout.
#include "gm_internal.hpp"

#include <cmath>
#include <cstring>

using gm::fail;

namespace {

constexpr double kPi = 3.141592653589793;  // the reference's PI (src/global.h:17)

struct Grid {
    const gm_params &p;
    long long sx, sxy;
    explicit Grid(const gm_params &q) : p(q), sx(q.nx + 2), sxy((long long)(q.nx + 2) * (q.ny + 2)) {}
    long long at(int i, int j, int k) const { return i + sx * j + sxy * k; }
    // coordinate of padded index i: h*(i-1) + O  (src/init.cu:50-55)
    static double coord(double h, int i, double o) { return h * (i - 1) + o; }
};

// Twisted flux blob of the "blobs" field (src/init.cu:109-124): adds
// c (-(y-yc), x-xc, 0) exp(...) with c = 2 ampl twist/scale to b.
inline void add_blob(const gm_params &p, int m, double x, double y, double z, double *b)
{
    double rx = x - p.blobXc[m], ry = y - p.blobYc[m], rz = z - p.blobZc[m];
    double arg = (-rx * rx - ry * ry) / (p.blobScale[m] * p.blobScale[m]) - rz * rz / (p.blobZl[m] * p.blobZl[m]);
    double g = 2 * p.ampl * p.blobTwist[m] / p.blobScale[m] * std::exp(arg);
    b[0] += g * (-ry);
    b[1] += g * rx;
}

bool field_at(const gm_params &p, double x, double y, double z, double *b)
{
    const char *f = p.bInit;
    if (gm::starts(f, "blobs ")) {
        b[0] = 0.; b[1] = 0.; b[2] = 0.;
        for (int m = 0; m < p.nBlobs; ++m) add_blob(p, m, x, y, z, b);
        b[2] = p.ampl;
        return true;
    }
    if (gm::starts(f, "homZ ")) {
        b[0] = 0.; b[1] = 0.; b[2] = p.ampl;
        return true;
    }
    if (gm::starts(f, "analytic ")) {   // closed-form relaxed state known (src/init.cu:71-78)
        double e = std::exp(-(x * x + y * y) / (p.ar * p.ar) - z * z / (p.az * p.az));
        b[0] = 2 * p.ampl * z * e * p.phi1 / (p.az * p.az) * y;
        b[1] = -2 * p.ampl * z * e * p.phi1 / (p.az * p.az) * x;
        b[2] = p.ampl;
        return true;
    }
    if (gm::starts(f, "analyticX ")) {
        double e = std::exp(-(z * z + y * y) / (p.ar * p.ar) - x * x / (p.az * p.az));
        b[0] = p.ampl;
        b[1] = -2 * p.ampl * x * e * p.phi1 / (p.az * p.az) * z;
        b[2] = 2 * p.ampl * x * e * p.phi1 / (p.az * p.az) * y;
        return true;
    }
    if (gm::starts(f, "analyticY ")) {
        double e = std::exp(-(x * x + z * z) / (p.ar * p.ar) - y * y / (p.az * p.az));
        b[0] = 2 * p.ampl * y * e * p.phi1 / (p.az * p.az) * z;
        b[1] = p.ampl;
        b[2] = -2 * p.ampl * y * e * p.phi1 / (p.az * p.az) * x;
        return true;
    }
    if (gm::starts(f, "Pontin09 ")) {   // src/init.cu:57-68
        double r2 = x * x + y * y;
        double e1 = std::exp(-r2 / (p.ar * p.ar) - (z - p.L1) * (z - p.L1) / (p.az * p.az));
        double e2 = std::exp(-r2 / (p.ar * p.ar) - (z - p.L2) * (z - p.L2) / (p.az * p.az));
        double s = p.phi1 * e1 + p.phi2 * e2;
        b[0] = -2 * p.ampl / (kPi * p.ar) * y * s;
        b[1] = 2 * p.ampl / (kPi * p.ar) * x * s;
        b[2] = p.ampl;
        return true;
    }
    return false;
}

void periodic_field(const gm_params &p, double *f)
{
    Grid g(p);
    auto cp = [&](long long dst, long long src) {
        for (int l = 0; l < 3; ++l) f[3 * dst + l] = f[3 * src + l];
    };
    if (p.zPeri)
        for (int j = 0; j < p.ny + 2; ++j)
            for (int i = 0; i < p.nx + 2; ++i) {
                cp(g.at(i, j, 0), g.at(i, j, p.nz));
                cp(g.at(i, j, p.nz + 1), g.at(i, j, 1));
            }
    if (p.yPeri)
        for (int k = 0; k < p.nz + 2; ++k)
            for (int i = 0; i < p.nx + 2; ++i) {
                cp(g.at(i, 0, k), g.at(i, p.ny, k));
                cp(g.at(i, p.ny + 1, k), g.at(i, 1, k));
            }
    if (p.xPeri)
        for (int k = 0; k < p.nz + 2; ++k)
            for (int j = 0; j < p.ny + 2; ++j) {
                cp(g.at(0, j, k), g.at(p.nx, j, k));
                cp(g.at(p.nx + 1, j, k), g.at(1, j, k));
            }
}

// Ghost planes of a periodic axis hold the opposite inner plane shifted by
// -+(L + h) along that axis (src/boundaries.cu:532-582).
void periodic_grid(const gm_params &p, double *x)
{
    Grid g(p);
    auto lo =[&](long long dst, long long src, int axis, double L, double h) {
        for (int l = 0; l < 3; ++l) x[3 * dst + l] = x[3 * src + l];
        x[3 * dst + axis] = x[3 * src + axis] - L - h;
    };
    auto hi = [&](long long dst, long long src, int axis, double L, double h) {
        for (int l = 0; l < 3; ++l) x[3 * dst + l] = x[3 * src + l];
        x[3 * dst + axis] = x[3 * src + axis] + L + h;
    };
    if (p.zPeri)
        for (int j = 0; j < p.ny + 2; ++j)
            for (int i = 0; i < p.nx + 2; ++i) {
                lo(g.at(i, j, 0), g.at(i, j, p.nz), 2, p.Lz, p.dz);
                hi(g.at(i, j, p.nz + 1), g.at(i, j, 1), 2, p.Lz, p.dz);
            }
    if (p.yPeri)
        for (int k = 0; k < p.nz + 2; ++k)
            for (int i = 0; i < p.nx + 2; ++i) {
                lo(g.at(i, 0, k), g.at(i, p.ny, k), 1, p.Ly, p.dy);
                hi(g.at(i, p.ny + 1, k), g.at(i, 1, k), 1, p.Ly, p.dy);
            }
    if (p.xPeri)
        for (int k = 0; k < p.nz + 2; ++k)
            for (int j = 0; j < p.ny + 2; ++j) {
                lo(g.at(0, j, k), g.at(p.nx, j, k), 0, p.Lx, p.dx);
                hi(g.at(p.nx + 1, j, k), g.at(1, j, k), 0, p.Lx, p.dx);
            }
}

}  // namespace

extern "C" double gm_init_residual(const gm_params *p)
{
    if (!p) return 0;
    const char *f = p->bInit;
    if (gm::starts(f, "Pontin09 ") || gm::starts(f, "analytic ") || gm::starts(f, "homZ ") ||
        gm::starts(f, "blobs "))
        return 1 * p->Lx * p->Ly * p->Lz * p->ampl;
    if (gm::starts(f, "Borromean ")) return 1 * p->Lx * p->Ly * p->Lz * p->bGround;
    return 0;
}

extern "C" int gm_init_state(const gm_params *pp, double *B0, double *xb)
{
    if (!pp || !B0 || !xb) return fail(GM_ERR_ARG, "gm_init_state: null pointer");
    const gm_params &p = *pp;
    Grid g(p);
    double probe[3];
    if (!field_at(p, 0, 0, 0, probe))
        return fail(GM_ERR_UNSUPPORTED, std::string("initial field '") + std::string(p.bInit, 29) +
                                            "' is not one of blobs, homZ, analytic, analyticX, analyticY, Pontin09");
    for (int k = 0; k < p.nz + 2; ++k) {
        double z = Grid::coord(p.dz, k, p.Oz);
        for (int j = 0; j < p.ny + 2; ++j) {
            double y = Grid::coord(p.dy, j, p.Oy);
            for (int i = 0; i < p.nx + 2; ++i) {
                double x = Grid::coord(p.dx, i, p.Ox);
                long long a = g.at(i, j, k);
                field_at(p, x, y, z, B0 + 3 * a);
                xb[3 * a + 0] = x;
                xb[3 * a + 1] = y;
                xb[3 * a + 2] = z;
            }
        }
    }
    // initDistortion (src/init.cu:388-419)
    bool shearX = gm::starts(p.initDist, "initShearX "), centre = gm::starts(p.initDist, "centerShift ");
    if (shearX || centre)
        for (int k = 0; k < p.nz + 2; ++k) {
            double z = Grid::coord(p.dz, k, p.Oz);
            for (int j = 0; j < p.ny + 2; ++j) {
                double y = Grid::coord(p.dy, j, p.Oy);
                for (int i = 0; i < p.nx + 2; ++i) {
                    double x = Grid::coord(p.dx, i, p.Ox);
                    double *q = xb + 3 * g.at(i, j, k);
                    if (shearX) {
                        q[1] -= p.initShearA * std::sin(p.initShearK * 2 * kPi * (x + p.Ox - p.dx / 2) / (p.Lx + p.dx)) * z;
                        double yy = q[1];
                        q[0] -= p.initShearB * std::sin(p.initShearK * 2 * kPi * (yy + p.Oy - p.dy / 2) / (p.Ly + p.dy)) * z;
                    }
                    if (centre)
                        q[1] += p.initShearA * std::exp(-p.initShearK * x * x) * (1 - y * y) *
                                (std::exp(-1 * z + 1) - 5.0804508111226507);
                }
            }
        }
    periodic_field(p, B0);
    periodic_grid(p, xb);
    return GM_OK;
}
