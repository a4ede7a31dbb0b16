// ECLIPSE (GRDECL) grid and rock input, host side (include/eclipse_grid.h).
//
// The reader keeps the keyword arrays of the GRID and PROPS sections that the
// grid and the rock need; the geometry is the exact corner-point geometry of
// vertical pillars (bilinear top and bottom surfaces, planar lateral faces cut
// at faults into the overlaps of the two columns' layers); the rock follows
// the permeability keyword rules of ReservoirPropertyCommon_impl.hpp.
#include "../../include/eclipse_grid.h"

#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <map>
#include <new>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

constexpr double kMilliDarcy = 1e-3 * 9.869233e-13;   // opm-core units: milli * darcy
constexpr double kFoot = 0.3048;

struct Face {
    int cell, key, nb, bid;
    double area, cen[3], nrm[3];
};

bool is_keyword(const std::string &t) {
    if (t.empty() || t.size() > 8 || !std::isupper((unsigned char)t[0])) return false;
    for (char ch : t)
        if (!(std::isupper((unsigned char)ch) || std::isdigit((unsigned char)ch) || ch == '_')) return false;
    return true;
}

const std::set<std::string> &section_keywords() {
    static const std::set<std::string> s = {"RUNSPEC", "GRID", "EDIT", "PROPS", "REGIONS", "SOLUTION",
                                            "SUMMARY", "SCHEDULE"};
    return s;
}

const std::set<std::string> &array_keywords() {
    static const std::set<std::string> s = {"SPECGRID", "DIMENS", "COORD", "ZCORN", "ACTNUM", "PORO",
                                            "PERMX", "PERMY", "PERMZ", "PERMXY", "PERMXZ", "PERMYX",
                                            "PERMYZ", "PERMZX", "PERMZY"};
    return s;
}

std::vector<std::string> split_tokens(const std::string &line) {
    std::vector<std::string> out;
    std::istringstream is(line);
    std::string t;
    while (is >> t) out.push_back(t);
    return out;
}

}  // namespace

struct mhs_deck {
    std::string err;
    std::map<std::string, std::vector<double>> kw;
    std::string units = "METRIC";
    int nx = 0, ny = 0, nz = 0;
    // grid
    bool built = false;
    std::vector<int> facepos, nb, bid, ijk;
    std::vector<double> cvol, ccen, farea, fcen, fnrm, corners;

    void parse(const std::string &text);
    void build_grid();
};

void mhs_deck::parse(const std::string &text) {
    std::istringstream in(text);
    std::string raw;
    std::vector<std::string> lines;
    while (std::getline(in, raw)) {
        const auto c = raw.find("--");
        if (c != std::string::npos) raw.erase(c);
        lines.push_back(raw);
    }
    size_t li = 0;
    // a keyword stands alone on its line; a section keyword may carry a
    // banner after it (RUNSPEC ====) and an array keyword the first of its values
    auto keyword_line = [&](const std::vector<std::string> &toks) {
        if (toks.empty() || !is_keyword(toks[0])) return false;
        return toks.size() == 1 || section_keywords().count(toks[0]) > 0 || array_keywords().count(toks[0]) > 0;
    };
    while (li < lines.size()) {
        const std::vector<std::string> toks = split_tokens(lines[li]);
        if (toks.empty()) { ++li; continue; }
        if (!keyword_line(toks)) {
            throw std::runtime_error("read_grdecl: expected a keyword, got '" + toks[0] + "' (line " +
                                     std::to_string(li + 1) + ")");
        }
        const std::string k = toks[0];
        ++li;
        if (k == "SOLUTION" || k == "SUMMARY" || k == "SCHEDULE" || k == "REGIONS") break;
        if (k == "METRIC" || k == "FIELD" || k == "LAB") {
            if (k == "LAB") throw std::runtime_error("read_grdecl: only METRIC and FIELD units are supported");
            units = k;
            continue;
        }
        if (k == "TITLE") { ++li; continue; }
        if (array_keywords().count(k)) {
            std::vector<double> vals;
            bool done = false;
            // the values may start on the keyword's own line
            std::vector<std::string> rest(toks.begin() + 1, toks.end());
            bool own_line = true;
            while (!done && (own_line || li < lines.size())) {
                const std::vector<std::string> cur = own_line ? rest : split_tokens(lines[li]);
                for (std::string t : cur) {
                    if (done) break;
                    const auto slash = t.find('/');
                    if (slash != std::string::npos) {
                        done = true;
                        t.erase(slash);
                        if (t.empty()) break;
                    }
                    if (k == "SPECGRID" && vals.size() >= 3) continue;   // NUMRES, radial flag
                    const auto star = t.find('*');
                    double v = 0.0;
                    long rep = 1;
                    std::string num = t;
                    if (star != std::string::npos) {
                        num = t.substr(star + 1);
                        if (num.empty())
                            throw std::runtime_error(k + ": default values ('" + t + "') are not supported");
                        rep = std::strtol(t.substr(0, star).c_str(), nullptr, 10);
                    }
                    char *end = nullptr;
                    v = std::strtod(num.c_str(), &end);
                    if (end == num.c_str() || *end != '\0')
                        throw std::runtime_error("read_grdecl: unreadable value in " + k + ": '" + t + "'");
                    for (long r = 0; r < rep; ++r) vals.push_back(v);
                }
                if (own_line) own_line = false;
                else ++li;
            }
            if (!done) throw std::runtime_error("read_grdecl: keyword " + k + " is not terminated by '/'");
            kw[k] = std::move(vals);
            continue;
        }
        // any other keyword: its data runs to the next keyword line
        while (li < lines.size() && !keyword_line(split_tokens(lines[li]))) ++li;
    }
    for (const char *k : {"SPECGRID", "DIMENS"}) {
        auto it = kw.find(k);
        if (it != kw.end() && it->second.size() >= 3) {
            nx = (int)it->second[0];
            ny = (int)it->second[1];
            nz = (int)it->second[2];
            break;
        }
    }
}

namespace {

// For two bands {aL(s) <= z <= aH(s)} and {bL(s) <= z <= bH(s)}, s in [0, 1],
// with linear bounds (value 1 at s = 0, value 2 at s = 1): the integrals of 1,
// s and z over their intersection.  The height is piecewise linear with kinks
// where two bounds cross, so Simpson's rule between the sorted crossings is
// exact.
void overlap_integrals(double aL1, double aL2, double aH1, double aH2, double bL1, double bL2, double bH1,
                       double bH2, double *I0, double *Is, double *Iz) {
    const double L1[4] = {aL1, bL1, aH1, bH1}, L2[4] = {aL2, bL2, aH2, bH2};
    double pts[8];
    int np = 0;
    pts[np++] = 0.0;
    pts[np++] = 1.0;
    for (int p = 0; p < 4; ++p)
        for (int q = p + 1; q < 4; ++q) {
            const double d1 = L1[p] - L1[q], d2 = L2[p] - L2[q], den = d1 - d2;
            const double s = den != 0.0 ? d1 / den : 0.0;
            pts[np++] = (s > 0.0 && s < 1.0) ? s : 0.0;
        }
    std::sort(pts, pts + np);
    auto band = [&](double s, double *h, double *q) {
        const double lo = std::max(aL1 + s * (aL2 - aL1), bL1 + s * (bL2 - bL1));
        const double hi = std::min(aH1 + s * (aH2 - aH1), bH1 + s * (bH2 - bH1));
        *h = hi - lo;
        *q = 0.5 * (hi * hi - lo * lo);
    };
    *I0 = *Is = *Iz = 0.0;
    for (int t = 0; t + 1 < np; ++t) {
        const double a = pts[t], b = pts[t + 1], m = 0.5 * (a + b), w = (b - a) / 6.0;
        double ha, qa, hm, qm, hb, qb;
        band(a, &ha, &qa);
        band(m, &hm, &qm);
        band(b, &hb, &qb);
        if (!(hm > 0.0 && b > a)) continue;
        *I0 += w * (ha + 4 * hm + hb);
        *Is += w * (a * ha + 4 * m * hm + b * hb);
        *Iz += w * (qa + 4 * qm + qb);
    }
}

}  // namespace

void mhs_deck::build_grid() {
    if (nx <= 0 || ny <= 0 || nz <= 0) throw std::runtime_error("deck has neither SPECGRID nor DIMENS");
    auto ci = kw.find("COORD"), zi = kw.find("ZCORN");
    if (ci == kw.end() || zi == kw.end()) throw std::runtime_error("deck needs COORD and ZCORN");
    const std::vector<double> &coord = ci->second, &zcorn = zi->second;
    const size_t np = (size_t)(nx + 1) * (ny + 1);
    if (coord.size() != 6 * np)
        throw std::runtime_error("COORD has " + std::to_string(coord.size()) + " values, expected " +
                                 std::to_string(6 * np));
    if (zcorn.size() != 8 * (size_t)nx * ny * nz)
        throw std::runtime_error("ZCORN has " + std::to_string(zcorn.size()) + " values, expected " +
                                 std::to_string(8 * (size_t)nx * ny * nz));
    auto ai = kw.find("ACTNUM");
    if (ai != kw.end())
        for (double a : ai->second)
            if (a == 0.0) throw std::runtime_error("inactive cells (ACTNUM 0) are not supported");
    const double lf = units == "FIELD" ? kFoot : 1.0;
    auto px = [&](int i, int j, int e) { return lf * coord[6 * ((size_t)j * (nx + 1) + i) + e]; };
    double xmin = 1e300, xmax = -1e300, ymin = 1e300, ymax = -1e300;
    for (int j = 0; j <= ny; ++j)
        for (int i = 0; i <= nx; ++i) {
            xmin = std::min(xmin, px(i, j, 0)); xmax = std::max(xmax, px(i, j, 0));
            ymin = std::min(ymin, px(i, j, 1)); ymax = std::max(ymax, px(i, j, 1));
        }
    const double rtol = 1e-9, scale = std::max(std::max(xmax - xmin, ymax - ymin), 1.0);
    for (int j = 0; j <= ny; ++j)
        for (int i = 0; i <= nx; ++i)
            if (std::fabs(px(i, j, 3) - px(i, j, 0)) > rtol * scale || std::fabs(px(i, j, 4) - px(i, j, 1)) > rtol * scale)
                throw std::runtime_error("only vertical pillars are supported");
    const double x0 = px(0, 0, 0), y0 = px(0, 0, 1);
    const double dx = (px(nx, 0, 0) - x0) / nx, dy = (px(0, ny, 1) - y0) / ny;
    bool lattice = dx > 0 && dy > 0;
    for (int j = 0; j <= ny && lattice; ++j)
        for (int i = 0; i <= nx; ++i)
            if (std::fabs(px(i, j, 0) - (x0 + i * dx)) > rtol * scale ||
                std::fabs(px(i, j, 1) - (y0 + j * dy)) > rtol * scale)
                lattice = false;
    if (!lattice) throw std::runtime_error("pillars must lie on a regular lattice with increasing x and y");
    // ZCORN: per layer k, top then bottom surface; per surface, per row j the
    // near (y-) and far (y+) corner rows; per row, per cell i its two corners
    auto z = [&](int k, int t, int j, int cy, int i, int cx) {
        return lf * zcorn[(((((size_t)k * 2 + t) * ny + j) * 2 + cy) * nx + i) * 2 + cx];
    };
    for (int k = 0; k + 1 < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int cy = 0; cy < 2; ++cy)
                for (int i = 0; i < nx; ++i)
                    for (int cx = 0; cx < 2; ++cx)
                        if (z(k, 1, j, cy, i, cx) != z(k + 1, 0, j, cy, i, cx))
                            throw std::runtime_error("gaps between layers are not supported");
    // zc(j, i, c, l): depth of corner c (0: (i, j), 1: (i+1, j), 2: (i, j+1),
    // 3: (i+1, j+1)) of column (i, j) at level l (top of layer l; l = nz the
    // bottom of the last layer)
    const int nl = nz + 1;
    std::vector<double> zc((size_t)ny * nx * 4 * nl);
    auto ZC = [&](int j, int i, int c, int l) -> double & { return zc[(((size_t)j * nx + i) * 4 + c) * nl + l]; };
    const int cxs[4] = {0, 1, 0, 1}, cys[4] = {0, 0, 1, 1};
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i)
            for (int c = 0; c < 4; ++c)
                for (int l = 0; l < nl; ++l)
                    ZC(j, i, c, l) = l < nz ? z(l, 0, j, cys[c], i, cxs[c]) : z(nz - 1, 1, j, cys[c], i, cxs[c]);
    // --- the geometry of corner_point_from_depths ------------------------------
    double thick = 0.0, zlo = 1e300, zhi = -1e300;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i)
            for (int c = 0; c < 4; ++c) {
                thick += ZC(j, i, c, nz) - ZC(j, i, c, 0);
                for (int l = 0; l < nl; ++l) {
                    zlo = std::min(zlo, ZC(j, i, c, l));
                    zhi = std::max(zhi, ZC(j, i, c, l));
                }
            }
    const double dz = thick / (4.0 * nx * ny) / nz;
    const double zmin = zlo - 10.0 * dz, zmax = zhi + 10.0 * dz;
    const int nc = nx * ny * nz;
    auto cidx = [&](int i, int j, int k) { return i + nx * (j + ny * k); };
    cvol.assign(nc, 0.0);
    ccen.assign(3 * (size_t)nc, 0.0);
    const double g1 = 0.5 - 0.5 / std::sqrt(3.0), g2 = 0.5 + 0.5 / std::sqrt(3.0);
    const double gauss[2] = {g1, g2};
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                double V = 0, MX = 0, MY = 0, MZ = 0;
                for (double xi : gauss)
                    for (double eta : gauss) {
                        const double w[4] = {(1 - xi) * (1 - eta), xi * (1 - eta), (1 - xi) * eta, xi * eta};
                        double zt = 0, zb = 0;
                        for (int c = 0; c < 4; ++c) {
                            zt += w[c] * ZC(j, i, c, k);
                            zb += w[c] * ZC(j, i, c, k + 1);
                        }
                        const double h = 0.25 * dx * dy * (zb - zt);
                        V += h;
                        MX += h * ((i + xi) * dx);
                        MY += h * ((j + eta) * dy);
                        MZ += 0.25 * dx * dy * 0.5 * (zb * zb - zt * zt);
                    }
                const int c = cidx(i, j, k);
                if (!(V > 0.0)) throw std::runtime_error("corner_point: non-positive cell volume");
                cvol[c] = V;
                ccen[3 * (size_t)c] = MX / V + x0;
                ccen[3 * (size_t)c + 1] = MY / V + y0;
                ccen[3 * (size_t)c + 2] = MZ / V;
            }
    std::vector<Face> F;
    F.reserve((size_t)nc * 7);
    auto add = [&](int cell, int key, int nbr, double area, const double cen[3], const double nrm[3], int b) {
        Face f;
        f.cell = cell; f.key = key; f.nb = nbr; f.bid = b; f.area = area;
        for (int e = 0; e < 3; ++e) { f.cen[e] = cen[e]; f.nrm[e] = nrm[e]; }
        F.push_back(f);
    };
    // top / bottom faces: bilinear surfaces given by their vector area
    for (int l = 0; l <= nz; ++l)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const double z0 = ZC(j, i, 0, l), z1 = ZC(j, i, 1, l), z2 = ZC(j, i, 2, l), z3 = ZC(j, i, 3, l);
                const double gx = dy * 0.5 * ((z1 - z0) + (z3 - z2)), gy = dx * 0.5 * ((z2 - z0) + (z3 - z1));
                const double N[3] = {-gx, -gy, dx * dy};
                const double A = std::sqrt(N[0] * N[0] + N[1] * N[1] + N[2] * N[2]);
                const double nrm[3] = {N[0] / A, N[1] / A, N[2] / A}, mnrm[3] = {-nrm[0], -nrm[1], -nrm[2]};
                const double cen[3] = {(i + 0.5) * dx, (j + 0.5) * dy, 0.25 * (z0 + z1 + z2 + z3)};
                if (l > 0) add(cidx(i, j, l - 1), 5000, l < nz ? cidx(i, j, l) : -1, A, cen, nrm, l < nz ? 0 : 6);
                if (l < nz) add(cidx(i, j, l), 4000, l > 0 ? cidx(i, j, l - 1) : -1, A, cen, mnrm, l > 0 ? 0 : 5);
            }
    // lateral faces between column A = (iA, jA) and B across the plane through
    // A's corners cA; B absent (iB < 0): the domain side of A
    auto lateral = [&](int iA, int jA, int iB, int jB, const int cA[2], const int cB[2], int dir) {
        double nrm[3] = {0, 0, 0}, mnrm[3] = {0, 0, 0};
        nrm[dir] = 1.0;
        mnrm[dir] = -1.0;
        const int sideA = 2 * dir + 1, sideB = 2 * dir;
        const double L = dir == 0 ? dy : dx;
        auto point = [&](double sb, double zb, double out[3]) {
            if (dir == 0) { out[0] = (iA + 1) * dx; out[1] = jA * dy + sb * dy; }
            else { out[0] = iA * dx + sb * dx; out[1] = (jA + 1) * dy; }
            out[2] = zb;
        };
        auto zA1 = [&](int l) { return ZC(jA, iA, cA[0], l); };
        auto zA2 = [&](int l) { return ZC(jA, iA, cA[1], l); };
        if (iB < 0) {
            for (int k = 0; k < nz; ++k) {
                double I0, Is, Iz, cen[3];
                overlap_integrals(zA1(k), zA2(k), zA1(k + 1), zA2(k + 1), zA1(k), zA2(k), zA1(k + 1), zA2(k + 1),
                                  &I0, &Is, &Iz);
                point(Is / I0, Iz / I0, cen);
                add(cidx(iA, jA, k), sideA * 1000 + k + 1, -1, L * I0, cen, nrm, sideA + 1);
            }
            return;
        }
        auto zB1 = [&](int l) { return ZC(jB, iB, cB[0], l); };
        auto zB2 = [&](int l) { return ZC(jB, iB, cB[1], l); };
        bool same = true;
        for (int l = 0; l < nl && same; ++l) same = zA1(l) == zB1(l) && zA2(l) == zB2(l);
        if (same) {
            for (int k = 0; k < nz; ++k) {
                double I0, Is, Iz, cen[3];
                overlap_integrals(zA1(k), zA2(k), zA1(k + 1), zA2(k + 1), zA1(k), zA2(k), zA1(k + 1), zA2(k + 1),
                                  &I0, &Is, &Iz);
                point(Is / I0, Iz / I0, cen);
                const int a = cidx(iA, jA, k), b = cidx(iB, jB, k);
                add(a, sideA * 1000 + k + 1, b, L * I0, cen, nrm, 0);
                add(b, sideB * 1000 + k + 1, a, L * I0, cen, mnrm, 0);
            }
            return;
        }
        // a fault: every overlap of A's layers with B's, and with the virtual
        // layers above (-1) and below (nz) each column
        auto bounds = [&](bool isA, int k, double b4[4]) {
            auto z1 = [&](int l) { return isA ? zA1(l) : zB1(l); };
            auto z2 = [&](int l) { return isA ? zA2(l) : zB2(l); };
            if (k == -1) { b4[0] = zmin; b4[1] = zmin; b4[2] = z1(0); b4[3] = z2(0); }
            else if (k == nz) { b4[0] = z1(nz); b4[1] = z2(nz); b4[2] = zmax; b4[3] = zmax; }
            else { b4[0] = z1(k); b4[1] = z2(k); b4[2] = z1(k + 1); b4[3] = z2(k + 1); }
        };
        for (int ka = -1; ka <= nz; ++ka)
            for (int kb = -1; kb <= nz; ++kb) {
                const bool real_a = ka >= 0 && ka < nz, real_b = kb >= 0 && kb < nz;
                if (!real_a && !real_b) continue;
                if (real_a && real_b && std::abs(ka - kb) > 3) continue;
                double a4[4], b4[4], I0, Is, Iz, cen[3];
                bounds(true, ka, a4);
                bounds(false, kb, b4);
                overlap_integrals(a4[0], a4[1], a4[2], a4[3], b4[0], b4[1], b4[2], b4[3], &I0, &Is, &Iz);
                if (!(I0 > 1e-12 * dz)) continue;
                point(Is / I0, Iz / I0, cen);
                const double area = L * I0;
                if (real_a)
                    add(cidx(iA, jA, ka), sideA * 1000 + kb + 1, real_b ? cidx(iB, jB, kb) : -1, area, cen, nrm,
                        real_b ? 0 : 7);
                if (real_b)
                    add(cidx(iB, jB, kb), sideB * 1000 + ka + 1, real_a ? cidx(iA, jA, ka) : -1, area, cen, mnrm,
                        real_a ? 0 : 7);
            }
    };
    const int cx13[2] = {1, 3}, cx02[2] = {0, 2}, cy23[2] = {2, 3}, cy01[2] = {0, 1};
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i + 1 < nx; ++i) lateral(i, j, i + 1, j, cx13, cx02, 0);
    for (int j = 0; j + 1 < ny; ++j)
        for (int i = 0; i < nx; ++i) lateral(i, j, i, j + 1, cy23, cy01, 1);
    for (int j = 0; j < ny; ++j) lateral(nx - 1, j, -1, -1, cx13, nullptr, 0);
    for (int i = 0; i < nx; ++i) lateral(i, ny - 1, -1, -1, cy23, nullptr, 1);
    // x- and y- sides of the domain: the mirrored planes through corners (0, 2)
    // resp. (0, 1), normals reversed
    for (int dir = 0; dir < 2; ++dir) {
        const int *corners = dir == 0 ? cx02 : cy01;
        const int side = 2 * dir;
        double nrm[3] = {0, 0, 0};
        nrm[dir] = -1.0;
        const int n = dir == 0 ? ny : nx;
        for (int m = 0; m < n; ++m) {
            const int iA = dir == 0 ? 0 : m, jA = dir == 0 ? m : 0;
            for (int k = 0; k < nz; ++k) {
                const double a1 = ZC(jA, iA, corners[0], k), a2 = ZC(jA, iA, corners[1], k);
                const double b1 = ZC(jA, iA, corners[0], k + 1), b2 = ZC(jA, iA, corners[1], k + 1);
                double I0, Is, Iz;
                overlap_integrals(a1, a2, b1, b2, a1, a2, b1, b2, &I0, &Is, &Iz);
                const double sb = Is / I0, zb = Iz / I0;
                double cen[3];
                if (dir == 0) { cen[0] = 0.0; cen[1] = (jA + sb) * dy; }
                else { cen[0] = (iA + sb) * dx; cen[1] = 0.0; }
                cen[2] = zb;
                add(cidx(iA, jA, k), side * 1000 + k + 1, -1, (dir == 0 ? dy : dx) * I0, cen, nrm, side + 1);
            }
        }
    }
    std::stable_sort(F.begin(), F.end(), [](const Face &a, const Face &b) {
        return a.cell != b.cell ? a.cell < b.cell : a.key < b.key;
    });
    const size_t nhf = F.size();
    facepos.assign(nc + 1, 0);
    nb.resize(nhf);
    bid.resize(nhf);
    farea.resize(nhf);
    fcen.resize(3 * nhf);
    fnrm.resize(3 * nhf);
    for (size_t h = 0; h < nhf; ++h) {
        const Face &f = F[h];
        ++facepos[f.cell + 1];
        nb[h] = f.nb;
        bid[h] = f.bid;
        farea[h] = f.area;
        fcen[3 * h] = f.cen[0] + x0;
        fcen[3 * h + 1] = f.cen[1] + y0;
        fcen[3 * h + 2] = f.cen[2];
        for (int e = 0; e < 3; ++e) fnrm[3 * h + e] = f.nrm[e];
    }
    for (int c = 0; c < nc; ++c) facepos[c + 1] += facepos[c];
    ijk.resize(3 * (size_t)nc);
    corners.resize(24 * (size_t)nc);
    // VTK hexahedron: top ring (i, j), (i+1, j), (i+1, j+1), (i, j+1), then the bottom ring
    const int ring[4][3] = {{0, 0, 0}, {1, 1, 0}, {3, 1, 1}, {2, 0, 1}};
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const int c = cidx(i, j, k);
                ijk[3 * (size_t)c] = i;
                ijk[3 * (size_t)c + 1] = j;
                ijk[3 * (size_t)c + 2] = k;
                for (int lev = 0; lev < 2; ++lev)
                    for (int r = 0; r < 4; ++r) {
                        double *p = &corners[24 * (size_t)c + 3 * (4 * lev + r)];
                        p[0] = x0 + (i + ring[r][1]) * dx;
                        p[1] = y0 + (j + ring[r][2]) * dy;
                        p[2] = ZC(j, i, ring[r][0], k + lev);
                    }
            }
    built = true;
}

namespace {

int deck_fail(mhs_deck *d, const std::string &msg) {
    if (d) d->err = msg;
    return 1;
}

int read_common(mhs_deck **out, const std::string &text) {
    if (!out) return 1;
    mhs_deck *d = new (std::nothrow) mhs_deck();
    *out = d;
    if (!d) return 5;
    try {
        d->parse(text);
    } catch (const std::exception &e) {
        return deck_fail(d, e.what());
    }
    return 0;
}

// classifyPermeability (ReservoirPropertyCommon_impl.hpp:70-112): 0 none,
// 1 scalar, 2 diagonal, 3 tensor, -1 invalid.
int classify(const mhs_deck *d) {
    auto has = [&](const char *k) { return d->kw.count(k) > 0; };
    const bool xx = has("PERMX"), xy = has("PERMXY"), xz = has("PERMXZ");
    const bool yx = has("PERMYX"), yy = has("PERMY"), yz = has("PERMYZ");
    const bool zx = has("PERMZX"), zy = has("PERMZY"), zz = has("PERMZ");
    const int ncross = xy + xz + yx + yz + zx + zy, ncomp = xx + yy + zz + ncross;
    int kind = 0;
    if (ncross > 0) kind = 3;
    else if (ncomp == 1) kind = 1;
    else if (ncomp >= 2) kind = 2;
    bool ok = true;
    if (ncomp > 0) {
        ok = xx || !(xy || xz || yx || zx);
        ok = ok && (yy || !(yx || yz || xy || zy));
        ok = ok && (zz || !(zx || zy || xz || yz));
    }
    return ok ? kind : -1;
}

}  // namespace

extern "C" {

int mhs_deck_read_file(mhs_deck **out, const char *path) {
    if (!out || !path) return 1;
    std::ifstream f(path);
    if (!f) {
        *out = new (std::nothrow) mhs_deck();
        return deck_fail(*out, std::string("Could not open file ") + path);
    }
    std::stringstream ss;
    ss << f.rdbuf();
    return read_common(out, ss.str());
}

int mhs_deck_read_text(mhs_deck **out, const char *text) {
    if (!text) return 1;
    return read_common(out, std::string(text));
}

int mhs_deck_destroy(mhs_deck *d) {
    delete d;
    return 0;
}

const char *mhs_deck_last_error(const mhs_deck *d) { return d ? d->err.c_str() : "null deck"; }

int mhs_deck_dims(mhs_deck *d, int dims[3]) {
    if (!d || !dims) return 1;
    if (d->nx <= 0) return deck_fail(d, "deck has neither SPECGRID nor DIMENS");
    dims[0] = d->nx;
    dims[1] = d->ny;
    dims[2] = d->nz;
    return 0;
}

int mhs_deck_build_grid(mhs_deck *d, int *num_cells, int *num_half_faces) {
    if (!d) return 1;
    try {
        d->build_grid();
    } catch (const std::exception &e) {
        return deck_fail(d, e.what());
    }
    if (num_cells) *num_cells = (int)d->cvol.size();
    if (num_half_faces) *num_half_faces = (int)d->nb.size();
    return 0;
}

int mhs_deck_grid_arrays(mhs_deck *d, int *facepos, int *neighbour, int *boundary_id, double *cell_volume,
                         double *cell_centroid, double *face_area, double *face_centroid, double *face_normal,
                         int *ijk, double *corners) {
    if (!d) return 1;
    if (!d->built) return deck_fail(d, "grid_arrays: no grid (call mhs_deck_build_grid)");
    auto put = [](auto *dst, const auto &src) {
        if (dst) std::copy(src.begin(), src.end(), dst);
    };
    put(facepos, d->facepos);
    put(neighbour, d->nb);
    put(boundary_id, d->bid);
    put(cell_volume, d->cvol);
    put(cell_centroid, d->ccen);
    put(face_area, d->farea);
    put(face_centroid, d->fcen);
    put(face_normal, d->fnrm);
    put(ijk, d->ijk);
    put(corners, d->corners);
    return 0;
}

int mhs_deck_rock(mhs_deck *d, int num_cells, const int *global_cell, double perm_threshold, double *perm,
                  double *porosity, int *kind) {
    if (!d || num_cells < 0) return 1;
    if (d->nx <= 0) return deck_fail(d, "deck has neither SPECGRID nor DIMENS");
    const long long nglob = (long long)d->nx * d->ny * d->nz;
    for (int c = 0; c < num_cells; ++c) {
        const long long g = global_cell ? global_cell[c] : c;
        if (g < 0 || g >= nglob) return deck_fail(d, "rock: global cell index out of range");
    }
    const int k = classify(d);
    if (k < 0) return deck_fail(d, "Invalid set of permeability fields given.");
    if (kind) *kind = k;
    // fillTensor (:178-244): kmap[e] picks the component array of tensor entry e
    // (row-major; 0: none); cross terms are symmetric and a diagonal keyword
    // fills the other diagonal entries not yet set
    enum { XX, XY, XZ, YX, YY, YZ, ZX, ZY, ZZ };
    int kmap[9] = {0};
    std::vector<const std::vector<double> *> comps = {nullptr};
    auto push = [&](const char *name) {
        comps.push_back(&d->kw[name]);
        return (int)comps.size() - 1;
    };
    auto scalar_if_needed = [&](int i, int j, int l) {
        if (kmap[j] == 0) kmap[j] = kmap[i];
        if (kmap[l] == 0) kmap[l] = kmap[i];
    };
    auto has = [&](const char *name) { return d->kw.count(name) > 0; };
    if (has("PERMX")) { kmap[XX] = push("PERMX"); scalar_if_needed(XX, YY, ZZ); }
    if (has("PERMXY")) { kmap[XY] = kmap[YX] = push("PERMXY"); }
    if (has("PERMXZ")) { kmap[XZ] = kmap[ZX] = push("PERMXZ"); }
    if (has("PERMYX")) { kmap[YX] = kmap[XY] = push("PERMYX"); }
    if (has("PERMY")) { kmap[YY] = push("PERMY"); scalar_if_needed(YY, ZZ, XX); }
    if (has("PERMYZ")) { kmap[YZ] = kmap[ZY] = push("PERMYZ"); }
    if (has("PERMZX")) { kmap[ZX] = kmap[XZ] = push("PERMZX"); }
    if (has("PERMZY")) { kmap[ZY] = kmap[YZ] = push("PERMZY"); }
    if (has("PERMZ")) { kmap[ZZ] = push("PERMZ"); scalar_if_needed(ZZ, XX, YY); }
    for (size_t q = 1; q < comps.size(); ++q)
        if ((long long)comps[q]->size() != nglob)
            return deck_fail(d, "All permeability fields must have the same size as the logical cartesian size "
                                "of the grid: " + std::to_string(comps[q]->size()) + " != " + std::to_string(nglob));
    auto pi = d->kw.find("PORO");
    if (pi != d->kw.end() && (long long)pi->second.size() != nglob)
        return deck_fail(d, "PORO field must have the same size as the logical cartesian size of the grid: " +
                                std::to_string(pi->second.size()) + " != " + std::to_string(nglob));
    for (int c = 0; c < num_cells; ++c) {
        const long long g = global_cell ? global_cell[c] : c;
        if (perm) {
            for (int e = 0; e < 9; ++e)
                perm[9 * (size_t)c + e] = kmap[e] ? (*comps[kmap[e]])[g] * kMilliDarcy : 0.0;
            if (comps.size() > 1)
                for (int e : {0, 4, 8}) perm[9 * (size_t)c + e] = std::max(perm[9 * (size_t)c + e], perm_threshold);
        }
        if (porosity) porosity[c] = pi != d->kw.end() ? pi->second[g] : 1.0;
    }
    return 0;
}

}  // extern "C"
