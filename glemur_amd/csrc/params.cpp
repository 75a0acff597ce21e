// Parameter set of the reference (src/global.h:44-108) and a reader for its
// params.in namelist file (src/readNamelist.f90).  Host code only.
#include "gm_internal.hpp"

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <vector>

namespace gm {

static thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}
void clear_error() { g_err.clear(); }
const char *last_error() { return g_err.c_str(); }

}  // namespace gm

using gm::fail;

extern "C" const char *gm_last_error(void) { return gm::last_error(); }
extern "C" int gm_version(void) { return 1; }

// Fortran character(len=30) value copied by copyString
// (src/readNamelist.f90:267-283): blank padded, 29 characters and a NUL.
static void set_fstring(char dst[30], const std::string &v)
{
    for (int i = 0; i < 29; ++i)
        dst[i] = i < (int)v.size() ? v[i] : ' ';
    dst[29] = '\0';
}

static std::string ftrim(const char s[30])
{
    std::string r(s, strnlen(s, 30));
    while (!r.empty() && r.back() == ' ') r.pop_back();
    return r;
}

// src/readNamelist.f90:92-150.  In the FP64 build the variables are c_double
// but the literals are default (single-precision) reals, so a literal such as
// 3e-4 arrives as the double nearest to the float nearest to 3e-4.
static double freal(double literal) { return (double)(float)literal; }

extern "C" int gm_params_default(gm_params *p)
{
    if (!p) return fail(GM_ERR_ARG, "gm_params_default: null pointer");
    memset(p, 0, sizeof *p);
    p->nx = 16; p->ny = 16; p->nz = 16;
    p->Lx = freal(8.); p->Ly = freal(8.); p->Lz = freal(20.);
    p->Ox = 0; p->Oy = 0; p->Oz = 0;
    set_fstring(p->bInit, "Pontin09");
    set_fstring(p->uInit, "nil");
    p->ampl = freal(1.);
    p->phi1 = freal(2.); p->phi2 = freal(-2.);
    p->L1 = freal(-4.); p->L2 = freal(4.);
    p->ar = freal(1.414213562373095); p->az = freal(2.);
    p->nBlobs = 0;
    p->width = freal(0.6);
    p->minor = freal(1.0); p->major = freal(2.5);
    p->stretch = freal(1.0);
    p->bGround = 0.;
    set_fstring(p->initDist, "none");
    p->initDistCode = 0;
    p->initShearA = 0.; p->initShearB = 0.; p->initShearK = freal(1.);
    p->pert = freal(1.); p->twist = freal(1.);
    p->fRestart = false;
    p->nCycle = 10;
    p->dt0 = freal(3e-4);
    p->dtMin = freal(1e-9);
    p->maxError = freal(5e-5);
    p->zUpdate = true;
    p->inertia = false;
    p->pressure = false;
    p->nu = 1; p->beta = 1;
    p->jxbAver = false; p->jxbAverWeight = 0.;
    p->xPeri = p->yPeri = p->zPeri = false;
    p->epsilonProf = true;
    set_fstring(p->jMethod, "Stokes");
    set_fstring(p->pMethod, "Classic");
    p->nTs = 10;
    p->dtDump = freal(10.);
    p->dtSave = freal(1.);
    return gm_params_derive(p);
}

extern "C" int gm_params_derive(gm_params *p)
{
    if (!p) return fail(GM_ERR_ARG, "gm_params_derive: null pointer");
    if (p->nx < 2 || p->ny < 2 || p->nz < 2)
        return fail(GM_ERR_ARG, "grid needs at least 2 inner points per axis");
    // src/readNamelist.f90:185-186
    p->dx = p->Lx / (p->nx - 1); p->dy = p->Ly / (p->ny - 1); p->dz = p->Lz / (p->nz - 1);
    p->dx1 = 1 / p->dx; p->dy1 = 1 / p->dy; p->dz1 = 1 / p->dz;
    // src/readNamelist.f90:257-262 (no assignment for other strings)
    std::string d = ftrim(p->initDist);
    if (d == "initShearX") p->initDistCode = 0;
    if (d == "centerShift") p->initDistCode = 1;
    return GM_OK;
}

// ---------------------------------------------------------------------------
// Namelist input.  Groups &comp, &start, &run, &io are read in that order
// (src/readNamelist.f90:173-178).  Accepted syntax: name = value list with
// comma, blank, newline or ';' separators (';' as gfortran accepts it),
// '!' comments, quoted strings, r*value repeats, name(i) = ... element
// assignment, logicals T/F/.true./.false. in any case, d or e exponents.
// ---------------------------------------------------------------------------
namespace {

enum class Ty { Int, Real, Bool, Str, RealArr };
struct Field {
    const char *name;
    Ty ty;
    size_t off;
};

#define F(n, t) {#n, Ty::t, offsetof(gm_params, n)}
const Field kFields[] = {
    F(nx, Int), F(ny, Int), F(nz, Int),
    F(Lx, Real), F(Ly, Real), F(Lz, Real), F(Ox, Real), F(Oy, Real), F(Oz, Real),
    F(bInit, Str), F(uInit, Str), F(ampl, Real), F(phi1, Real), F(phi2, Real),
    F(L1, Real), F(L2, Real), F(ar, Real), F(az, Real), F(nBlobs, Int),
    F(blobXc, RealArr), F(blobYc, RealArr), F(blobZc, RealArr), F(blobZl, RealArr),
    F(blobTwist, RealArr), F(blobScale, RealArr),
    F(width, Real), F(minor, Real), F(major, Real), F(stretch, Real), F(bGround, Real),
    F(initDist, Str), F(initShearA, Real), F(initShearB, Real), F(initShearK, Real),
    F(pert, Real), F(twist, Real), F(fRestart, Bool),
    F(nCycle, Int), F(dt0, Real), F(dtMin, Real), F(maxError, Real), F(zUpdate, Bool),
    F(inertia, Bool), F(pressure, Bool), F(nu, Real), F(beta, Real), F(jxbAver, Bool),
    F(jxbAverWeight, Real), F(xPeri, Bool), F(yPeri, Bool), F(zPeri, Bool),
    F(epsilonProf, Bool), F(jMethod, Str), F(pMethod, Str),
    F(nTs, Int), F(dtDump, Real), F(dtSave, Real), F(dumpJ, Bool), F(dumpDetJac, Bool),
    F(dumpCellVol, Bool), F(dumpConvexity, Bool), F(dumpWedgeMin, Bool),
    F(redB2, Bool), F(redB2f, Bool), F(redBMax, Bool), F(redJMax, Bool),
    F(redJxB_B2Max, Bool), F(redEpsilonStar, Bool), F(redErrB_1ez, Bool),
    F(redErrXb_XbAn, Bool), F(redConvex, Bool), F(redWedgeMin, Bool), F(redU2, Bool),
    F(redUMax, Bool),
};
#undef F

// Members of each group (src/readNamelist.f90:152-171).
const char *kComp[] = {"nx", "ny", "nz", nullptr};
const char *kStart[] = {"Lx", "Ly", "Lz", "Ox", "Oy", "Oz", "bInit", "uInit", "ampl", "phi1", "phi2",
                        "L1", "L2", "ar", "az", "nBlobs", "blobXc", "blobYc", "blobZc", "blobZl",
                        "blobTwist", "blobScale", "width", "minor", "major", "stretch", "bGround",
                        "initDist", "initShearA", "initShearB", "initShearK", "pert", "twist",
                        "fRestart", nullptr};
const char *kRun[] = {"nCycle", "dt0", "dtMin", "maxError", "zUpdate", "inertia", "pressure", "nu",
                      "beta", "jxbAver", "jxbAverWeight", "xPeri", "yPeri", "zPeri", "epsilonProf",
                      "jMethod", "pMethod", nullptr};
const char *kIo[] = {"nTs", "dtDump", "dtSave", "dumpJ", "dumpDetJac", "dumpCellVol", "dumpConvexity",
                     "dumpWedgeMin", "redB2", "redB2f", "redBMax", "redJMax", "redJxB_B2Max",
                     "redEpsilonStar", "redErrB_1ez", "redErrXb_XbAn", "redConvex", "redWedgeMin",
                     "redU2", "redUMax", nullptr};

std::string lower(std::string s)
{
    for (auto &c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}

const Field *find_field(const std::string &name, const char *const *group)
{
    std::string ln = lower(name);
    for (int g = 0; group[g]; ++g)
        if (lower(group[g]) == ln)
            for (const Field &f : kFields)
                if (lower(f.name) == ln) return &f;
    return nullptr;
}

struct Tok {
    enum K { Word, Str, Eq, Slash, End } k;
    std::string s;
};

// Tokens of one group body, from just after "&name" to the closing '/'.
bool lex_group(const std::string &t, size_t &pos, std::vector<Tok> &out, std::string &err)
{
    while (true) {
        if (pos >= t.size()) {
            err = "namelist group not closed with '/'";
            return false;
        }
        char c = t[pos];
        if (c == '!') {
            while (pos < t.size() && t[pos] != '\n') ++pos;
        } else if (std::isspace((unsigned char)c) || c == ',' || c == ';') {
            ++pos;
        } else if (c == '/') {
            ++pos;
            out.push_back({Tok::Slash, ""});
            return true;
        } else if (c == '=') {
            ++pos;
            out.push_back({Tok::Eq, ""});
        } else if (c == '"' || c == '\'') {
            char q = c;
            std::string s;
            ++pos;
            while (pos < t.size()) {
                if (t[pos] == q) {
                    if (pos + 1 < t.size() && t[pos + 1] == q) {  // doubled quote
                        s += q;
                        pos += 2;
                        continue;
                    }
                    break;
                }
                s += t[pos++];
            }
            if (pos >= t.size()) {
                err = "unterminated string";
                return false;
            }
            ++pos;
            out.push_back({Tok::Str, s});
        } else {
            std::string s;
            while (pos < t.size()) {
                char d = t[pos];
                if (std::isspace((unsigned char)d) || d == ',' || d == ';' || d == '/' || d == '=' ||
                    d == '!' || d == '"' || d == '\'')
                    break;
                s += d;
                ++pos;
            }
            out.push_back({Tok::Word, s});
        }
    }
}

bool parse_real(const std::string &w, double &v)
{
    std::string s = w;
    for (auto &c : s)
        if (c == 'd' || c == 'D' || c == 'q' || c == 'Q') c = 'e';
    char *end = nullptr;
    v = strtod(s.c_str(), &end);
    return end && *end == '\0' && !s.empty();
}

bool parse_bool(const std::string &w, bool &v)
{
    size_t i = 0;
    if (i < w.size() && w[i] == '.') ++i;
    if (i >= w.size()) return false;
    char c = (char)std::tolower((unsigned char)w[i]);
    if (c == 't') { v = true; return true; }
    if (c == 'f') { v = false; return true; }
    return false;
}

bool assign(gm_params *p, const Field &f, int index, const std::vector<Tok> &vals, std::string &err)
{
    char *base = reinterpret_cast<char *>(p) + f.off;
    int slot = index;
    for (const Tok &tk : vals) {
        int rep = 1;
        std::string v = tk.s;
        if (tk.k == Tok::Word) {
            size_t star = v.find('*');
            if (star != std::string::npos) {
                rep = atoi(v.substr(0, star).c_str());
                v = v.substr(star + 1);
                if (rep < 1) {
                    err = "bad repeat count in '" + tk.s + "'";
                    return false;
                }
            }
        }
        for (int r = 0; r < rep; ++r, ++slot) {
            switch (f.ty) {
            case Ty::Int: {
                if (slot > 0) { err = std::string(f.name) + ": too many values"; return false; }
                char *end = nullptr;
                long x = strtol(v.c_str(), &end, 10);
                if (!end || *end || v.empty()) { err = std::string(f.name) + ": bad integer '" + v + "'"; return false; }
                *reinterpret_cast<int *>(base) = (int)x;
                break;
            }
            case Ty::Real: {
                if (slot > 0) { err = std::string(f.name) + ": too many values"; return false; }
                double x;
                if (!parse_real(v, x)) { err = std::string(f.name) + ": bad real '" + v + "'"; return false; }
                *reinterpret_cast<double *>(base) = x;
                break;
            }
            case Ty::Bool: {
                if (slot > 0) { err = std::string(f.name) + ": too many values"; return false; }
                bool x;
                if (!parse_bool(v, x)) { err = std::string(f.name) + ": bad logical '" + v + "'"; return false; }
                *reinterpret_cast<bool *>(base) = x;
                break;
            }
            case Ty::Str: {
                if (slot > 0) { err = std::string(f.name) + ": too many values"; return false; }
                set_fstring(reinterpret_cast<char *>(base), v);
                break;
            }
            case Ty::RealArr: {
                if (slot >= 10) { err = std::string(f.name) + ": more than 10 values"; return false; }
                double x;
                if (!parse_real(v, x)) { err = std::string(f.name) + ": bad real '" + v + "'"; return false; }
                reinterpret_cast<double *>(base)[slot] = x;
                break;
            }
            }
        }
    }
    return true;
}

bool read_group(const std::string &text, size_t &pos, const char *gname, const char *const *members,
                gm_params *p, std::string &err)
{
    // find "&gname" (case-insensitive) followed by a separator
    std::string lt = lower(text), key = std::string("&") + lower(gname);
    size_t at = pos;
    while (true) {
        at = lt.find(key, at);
        if (at == std::string::npos) {
            err = std::string("namelist group &") + gname + " not found";
            return false;
        }
        size_t e = at + key.size();
        if (e >= lt.size() || !(std::isalnum((unsigned char)lt[e]) || lt[e] == '_')) break;
        at = e;
    }
    pos = at + key.size();
    std::vector<Tok> toks;
    if (!lex_group(text, pos, toks, err)) return false;
    size_t i = 0;
    while (i < toks.size() && toks[i].k != Tok::Slash) {
        if (toks[i].k != Tok::Word || i + 1 >= toks.size() || toks[i + 1].k != Tok::Eq) {
            err = std::string("&") + gname + ": expected 'name =' near '" + toks[i].s + "'";
            return false;
        }
        std::string name = toks[i].s;
        int index = 0;
        size_t lp = name.find('(');
        if (lp != std::string::npos) {
            size_t rp = name.find(')', lp);
            index = atoi(name.substr(lp + 1, rp - lp - 1).c_str()) - 1;
            name = name.substr(0, lp);
            if (index < 0) { err = name + ": bad index"; return false; }
        }
        const Field *f = find_field(name, members);
        if (!f) {
            err = std::string("&") + gname + ": unknown variable '" + name + "'";
            return false;
        }
        i += 2;
        std::vector<Tok> vals;
        while (i < toks.size() && toks[i].k != Tok::Slash) {
            if (toks[i].k == Tok::Word && i + 1 < toks.size() && toks[i + 1].k == Tok::Eq) break;
            if (toks[i].k == Tok::Eq) { err = name + ": unexpected '='"; return false; }
            vals.push_back(toks[i]);
            ++i;
        }
        if (!assign(p, *f, index, vals, err)) return false;
    }
    return true;
}

}  // namespace

extern "C" int gm_params_read(const char *path, gm_params *p)
{
    if (!path || !p) return fail(GM_ERR_ARG, "gm_params_read: null pointer");
    std::ifstream in(path, std::ios::binary);
    if (!in) return fail(GM_ERR_IO, std::string("could not open '") + path + "'");
    std::stringstream ss;
    ss << in.rdbuf();
    std::string text = ss.str();
    size_t pos = 0;
    std::string err;
    if (!read_group(text, pos, "comp", kComp, p, err) || !read_group(text, pos, "start", kStart, p, err) ||
        !read_group(text, pos, "run", kRun, p, err) || !read_group(text, pos, "io", kIo, p, err))
        return fail(GM_ERR_PARSE, std::string(path) + ": " + err);
    return gm_params_derive(p);
}

// Seeded braid of the workloads (no reference counterpart).  The field is
// the reference's "blobs" form (src/init.cu:109-124):
//   B0 = ampl z^ + sum_k c_k (-(y-y_k), x-x_k, 0) exp(-rho^2/a_k^2 - (z-z_k)^2/l_k^2)
// with c_k = 2 ampl twist_k / scale_k.
//  random_centres = 0: six blobs at x_k = +1,-1,..., y_k = 0,
//      z_k = -20,-12,...,20, a^2 = 2, l^2 = 4, c_k = (-1)^k U(0.5,1.5).
//  random_centres = 1: n in [nmin, nmax], x_k, y_k ~ U(-1.5,1.5), z_k evenly
//      spaced on [-20,20], c_k = +-1, a^2 ~ U(1.5,2.5), l^2 ~ U(3,5).
extern "C" int gm_braid_seeded(gm_params *p, uint64_t seed, int nmin, int nmax, int random_centres)
{
    if (!p) return fail(GM_ERR_ARG, "gm_braid_seeded: null pointer");
    gm::SplitMix64 rng(seed);
    set_fstring(p->bInit, "blobs");
    if (!random_centres) {
        p->nBlobs = 6;
        for (int k = 0; k < 6; ++k) {
            double c = (k % 2 == 0 ? 1.0 : -1.0) * rng.uniform(0.5, 1.5);
            double scale = std::sqrt(2.0);
            p->blobXc[k] = (k % 2 == 0) ? 1.0 : -1.0;
            p->blobYc[k] = 0.0;
            p->blobZc[k] = -20.0 + 8.0 * k;
            p->blobScale[k] = scale;
            p->blobZl[k] = 2.0;
            p->blobTwist[k] = c * scale / (2 * p->ampl);
        }
    } else {
        if (nmin < 2 || nmax > 10 || nmin > nmax) return fail(GM_ERR_ARG, "blob count range must lie in [2,10]");
        int n = nmin + (int)(rng.next() % (uint64_t)(nmax - nmin + 1));
        p->nBlobs = n;
        for (int k = 0; k < n; ++k) {
            double sgn = (rng.next() & 1) ? 1.0 : -1.0;
            double scale = std::sqrt(rng.uniform(1.5, 2.5));
            p->blobXc[k] = rng.uniform(-1.5, 1.5);
            p->blobYc[k] = rng.uniform(-1.5, 1.5);
            p->blobZc[k] = -20.0 + 40.0 * k / (n - 1);
            p->blobScale[k] = scale;
            p->blobZl[k] = std::sqrt(rng.uniform(3.0, 5.0));
            p->blobTwist[k] = sgn * scale / (2 * p->ampl);
        }
    }
    return GM_OK;
}
