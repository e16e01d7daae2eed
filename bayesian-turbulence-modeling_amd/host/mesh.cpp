// mesh.cpp -- mesh factories, MSH1 I/O, wall distance (host side).
//
// The factories follow the reference's construction order and formulas
// (core/src/Mesh.cpp) so that face indices, owner/neighbour and every
// geometric value agree bit for bit; tests/test_mesh_factories.py checks each
// factory against the reference's own output.
#include "mesh.hpp"

#include <algorithm>
#include <cctype>
#include <cstdint>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <limits>
#include <sstream>
#include <stdexcept>

namespace rsh {

namespace {

constexpr uint32_t kMagic = 0x4D534831;   // "MSH1", Mesh.cpp:602

// tanh clustering toward both walls of [0, L] (Mesh.cpp:727-730)
double tanhNode(double L, double s, int j, int n) {
    double eta = static_cast<double>(j) / n;
    return 0.5 * L * (1.0 + std::tanh(s * (2.0 * eta - 1.0)) / std::tanh(s));
}

// stretch s giving first-cell height y1t (Mesh.cpp:1087-1113, :409-429)
double bisectStretch(double L, int n, double y1t) {
    auto fch = [&](double s) { return 0.5 * L * (1.0 + std::tanh(s * (2.0 / n - 1.0)) / std::tanh(s)); };
    double sLo = 0.1, sHi = 20.0;
    if (y1t < fch(sHi)) return sHi;
    for (int it = 0; it < 100; ++it) {
        double sm = 0.5 * (sLo + sHi);
        if (fch(sm) > y1t) sLo = sm; else sHi = sm;
        if (sHi - sLo < 1e-10) break;
    }
    return 0.5 * (sLo + sHi);
}

// first-cell target height from the flat-plate friction estimate (Mesh.cpp:1067-1077)
double firstCellTarget(double Lref, double Re, double yPlus) {
    double Cf = 0.058 * std::pow(Re, -0.2);
    double uTau = std::sqrt(Cf / 2.0);
    double nu = Lref / Re;
    return yPlus * nu / uTau;
}

}  // namespace

void Mesh::bump() { ++version_; flat_.builtFor = -1; }

// Mesh.cpp:301-316
void Mesh::computeFaceGeometry() {
    for (Face& f : faces_) {
        if (!f.isBoundary()) f.d = cells_[f.neighbor].center - cells_[f.owner].center;
        else f.d = f.center - cells_[f.owner].center;
        f.delta = f.d.norm();
        if (f.delta < 1e-30) f.delta = 1e-20;
    }
}

// Mesh.cpp:318-328
void Mesh::computeInterpolationWeights() {
    for (int fi = 0; fi < nInternal_; ++fi) {
        Face& f = faces_[fi];
        double dP = (f.center - cells_[f.owner].center).norm();
        double dN = (f.center - cells_[f.neighbor].center).norm();
        double sum = dP + dN;
        f.weight = (sum > 1e-30) ? dN / sum : 0.5;
    }
}

// Cartesian channel on given node lines; face order: horizontal internal,
// vertical internal, bottom, top, inlet, outlet (Mesh.cpp:719-865, :1116-1277).
Mesh Mesh::cartesian(int nx, int ny, const std::vector<double>& xc, const std::vector<double>& yc,
                     const char* topType) {
    Mesh m;
    const double dz = 1.0;
    const int pp = (nx + 1) * (ny + 1);
    m.nodes_.resize(2 * pp);
    for (int j = 0; j <= ny; ++j)
        for (int i = 0; i <= nx; ++i) {
            m.nodes_[j * (nx + 1) + i] = Vec3(xc[i], yc[j], 0.0);
            m.nodes_[pp + j * (nx + 1) + i] = Vec3(xc[i], yc[j], dz);
        }
    m.cells_.resize(nx * ny);
    auto cid = [&](int i, int j) { return j * nx + i; };
    m.nInternal_ = nx * (ny - 1) + (nx - 1) * ny;
    m.faces_.resize(m.nInternal_ + 2 * nx + 2 * ny);
    int fi = 0;
    auto link = [&](int f) {
        m.cells_[m.faces_[f].owner].faces.push_back(f);
        if (m.faces_[f].neighbor >= 0) m.cells_[m.faces_[f].neighbor].faces.push_back(f);
    };
    for (int j = 0; j < ny - 1; ++j)
        for (int i = 0; i < nx; ++i) {
            Face& f = m.faces_[fi];
            f.owner = cid(i, j); f.neighbor = cid(i, j + 1);
            f.center = Vec3(0.5 * (xc[i] + xc[i + 1]), yc[j + 1], 0.5 * dz);
            f.area = (xc[i + 1] - xc[i]) * dz;
            f.normal = Vec3(0, 1, 0);
            link(fi++);
        }
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx - 1; ++i) {
            Face& f = m.faces_[fi];
            f.owner = cid(i, j); f.neighbor = cid(i + 1, j);
            f.center = Vec3(xc[i + 1], 0.5 * (yc[j] + yc[j + 1]), 0.5 * dz);
            f.area = (yc[j + 1] - yc[j]) * dz;
            f.normal = Vec3(1, 0, 0);
            link(fi++);
        }
    auto bpatch = [&](const char* name, const char* type, int count, auto&& fill) {
        Patch p; p.name = name; p.type = type;
        for (int k = 0; k < count; ++k) {
            Face& f = m.faces_[fi];
            f.neighbor = -1;
            f.patchID = (int)m.patches_.size();
            fill(f, k);
            link(fi);
            p.faces.push_back(fi++);
        }
        m.patches_.push_back(std::move(p));
    };
    const double Lx = xc[nx], Ly = yc[ny];
    bpatch("bottom", "wall", nx, [&](Face& f, int i) {
        f.owner = cid(i, 0); f.center = Vec3(0.5 * (xc[i] + xc[i + 1]), 0.0, 0.5 * dz);
        f.area = (xc[i + 1] - xc[i]) * dz; f.normal = Vec3(0, -1, 0); });
    bpatch("top", topType, nx, [&](Face& f, int i) {
        f.owner = cid(i, ny - 1); f.center = Vec3(0.5 * (xc[i] + xc[i + 1]), Ly, 0.5 * dz);
        f.area = (xc[i + 1] - xc[i]) * dz; f.normal = Vec3(0, 1, 0); });
    bpatch("inlet", "inlet", ny, [&](Face& f, int j) {
        f.owner = cid(0, j); f.center = Vec3(0.0, 0.5 * (yc[j] + yc[j + 1]), 0.5 * dz);
        f.area = (yc[j + 1] - yc[j]) * dz; f.normal = Vec3(-1, 0, 0); });
    bpatch("outlet", "outlet", ny, [&](Face& f, int j) {
        f.owner = cid(nx - 1, j); f.center = Vec3(Lx, 0.5 * (yc[j] + yc[j + 1]), 0.5 * dz);
        f.area = (yc[j + 1] - yc[j]) * dz; f.normal = Vec3(1, 0, 0); });
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            Cell& c = m.cells_[cid(i, j)];
            c.center = Vec3(0.5 * (xc[i] + xc[i + 1]), 0.5 * (yc[j] + yc[j + 1]), 0.5 * dz);
            c.volume = (xc[i + 1] - xc[i]) * (yc[j + 1] - yc[j]) * dz;
        }
    m.computeFaceGeometry();
    m.computeInterpolationWeights();
    return m;
}

Mesh Mesh::makeChannel2D(int nx, int ny, double Lx, double Ly) {
    if (nx < 1 || ny < 1) throw std::invalid_argument("makeChannel2D: nx, ny must be >= 1");
    std::vector<double> yc(ny + 1), xc(nx + 1);
    for (int j = 0; j <= ny; ++j) yc[j] = tanhNode(Ly, 2.0, j, ny);
    for (int i = 0; i <= nx; ++i) xc[i] = Lx * static_cast<double>(i) / nx;
    return cartesian(nx, ny, xc, yc, "wall");
}

Mesh Mesh::makeChannel2D(int nx, int ny, double Lx, double Ly, double Re, double yPlusTarget) {
    if (nx < 1 || ny < 1) throw std::invalid_argument("makeChannel2D: nx, ny must be >= 1");
    double y1t = firstCellTarget(Ly / 2.0, Re, yPlusTarget);
    double stretch = 0.0;
    if (y1t < Ly / ny) stretch = bisectStretch(Ly, ny, y1t);
    std::vector<double> yc(ny + 1), xc(nx + 1);
    for (int j = 0; j <= ny; ++j)
        yc[j] = stretch > 1e-12 ? tanhNode(Ly, stretch, j, ny) : Ly * static_cast<double>(j) / ny;
    for (int i = 0; i <= nx; ++i) xc[i] = Lx * static_cast<double>(i) / nx;
    return cartesian(nx, ny, xc, yc, "wall");
}

// Mesh.cpp:880-1059: one-sided geometric growth from the bottom wall
Mesh Mesh::makePlate2D(int nx, int ny, double Lx, double Ly, double Re, double yPlusTarget) {
    if (nx < 1 || ny < 1) throw std::invalid_argument("makePlate2D: nx, ny must be >= 1");
    double y1t = firstCellTarget(Ly / 2.0, Re, yPlusTarget);
    double ratio = 1.0;
    if (y1t < Ly / ny) {
        auto total = [&](double r) { return y1t * (std::pow(r, ny) - 1.0) / (r - 1.0); };
        double rLo = 1.0 + 1e-9, rHi = 1.5;
        if (total(rHi) < Ly) {
            ratio = rHi;
        } else {
            for (int it = 0; it < 200; ++it) {
                double rMid = 0.5 * (rLo + rHi);
                if (total(rMid) < Ly) rLo = rMid; else rHi = rMid;
                if (rHi - rLo < 1e-12) break;
            }
            ratio = 0.5 * (rLo + rHi);
        }
    }
    std::vector<double> yc(ny + 1), xc(nx + 1);
    if (ratio > 1.0 + 1e-9) {
        yc[0] = 0.0;
        double h = y1t;
        for (int j = 1; j <= ny; ++j) { yc[j] = yc[j - 1] + h; h *= ratio; }
        double scale = Ly / yc[ny];
        for (int j = 1; j <= ny; ++j) yc[j] *= scale;
    } else {
        for (int j = 0; j <= ny; ++j) yc[j] = Ly * static_cast<double>(j) / ny;
    }
    for (int i = 0; i <= nx; ++i) xc[i] = Lx * static_cast<double>(i) / nx;
    return cartesian(nx, ny, xc, yc, "patch");
}

// Mesh.cpp:1280-1546: upstream channel over an expanded duct
Mesh Mesh::makeBackwardFacingStep2D(int nxUp, int nxDown, int nyUp, int nyDown, double Lu, double Ld,
                                    double hs, double H, double Re, double yPlusTarget) {
    if (nxUp < 1 || nxDown < 1 || nyUp < 1 || nyDown < 1)
        throw std::invalid_argument("makeBackwardFacingStep2D: block sizes must be >= 1");
    const int nxT = nxUp + nxDown;
    const double dz = 1.0, Hup = H - hs;
    std::vector<double> xc(nxT + 1);
    for (int i = 0; i <= nxUp; ++i) xc[i] = -Lu + Lu * i / nxUp;
    for (int i = 1; i <= nxDown; ++i) xc[nxUp + i] = Ld * i / nxDown;
    auto blockStretch = [&](double L, int n) {
        double s = 2.0;
        if (Re > 0) {
            double y1t = firstCellTarget(L, Re, yPlusTarget);
            if (y1t < L / n) s = bisectStretch(L, n, y1t);
        }
        return s;
    };
    const double sUp = blockStretch(Hup, nyUp), sLow = blockStretch(hs, nyDown);
    std::vector<double> yu(nyUp + 1), yl(nyDown + 1);
    for (int j = 0; j <= nyUp; ++j) {
        double eta = static_cast<double>(j) / nyUp;
        yu[j] = hs + 0.5 * Hup * (1.0 + std::tanh(sUp * (2.0 * eta - 1.0)) / std::tanh(sUp));
    }
    for (int j = 0; j <= nyDown; ++j) {
        double eta = static_cast<double>(j) / nyDown;
        yl[j] = 0.5 * hs * (1.0 + std::tanh(sLow * (2.0 * eta - 1.0)) / std::tanh(sLow));
    }
    auto cu = [&](int i, int j) { return j * nxT + i; };
    auto cl = [&](int i, int j) { return nxT * nyUp + j * nxDown + i; };
    Mesh m;
    const int nC = nxT * nyUp + nxDown * nyDown;
    m.cells_.resize(nC);
    m.nInternal_ = nxT * (nyUp - 1) + nxDown * (nyDown - 1) + nxDown + (nxT - 1) * nyUp + (nxDown - 1) * nyDown;
    const int nBnd = nxT + nxUp + nyDown + nxDown + nyUp + (nyUp + nyDown);
    m.faces_.resize(m.nInternal_ + nBnd);
    int fi = 0;
    auto add = [&](int o, int n, Vec3 c, double A, Vec3 nrm) {
        Face& f = m.faces_[fi];
        f.owner = o; f.neighbor = n; f.center = c; f.area = A; f.normal = nrm;
        m.cells_[o].faces.push_back(fi);
        if (n >= 0) m.cells_[n].faces.push_back(fi);
        return fi++;
    };
    for (int j = 0; j < nyUp - 1; ++j)
        for (int i = 0; i < nxT; ++i)
            add(cu(i, j), cu(i, j + 1), Vec3(0.5 * (xc[i] + xc[i + 1]), yu[j + 1], 0.5 * dz), (xc[i + 1] - xc[i]) * dz, Vec3(0, 1, 0));
    for (int j = 0; j < nyDown - 1; ++j)
        for (int i = 0; i < nxDown; ++i)
            add(cl(i, j), cl(i, j + 1), Vec3(0.5 * (xc[nxUp + i] + xc[nxUp + i + 1]), yl[j + 1], 0.5 * dz),
                (xc[nxUp + i + 1] - xc[nxUp + i]) * dz, Vec3(0, 1, 0));
    for (int i = 0; i < nxDown; ++i)
        add(cu(nxUp + i, 0), cl(i, nyDown - 1), Vec3(0.5 * (xc[nxUp + i] + xc[nxUp + i + 1]), hs, 0.5 * dz),
            (xc[nxUp + i + 1] - xc[nxUp + i]) * dz, Vec3(0, -1, 0));
    for (int j = 0; j < nyUp; ++j)
        for (int i = 0; i < nxT - 1; ++i)
            add(cu(i, j), cu(i + 1, j), Vec3(xc[i + 1], 0.5 * (yu[j] + yu[j + 1]), 0.5 * dz), (yu[j + 1] - yu[j]) * dz, Vec3(1, 0, 0));
    for (int j = 0; j < nyDown; ++j)
        for (int i = 0; i < nxDown - 1; ++i)
            add(cl(i, j), cl(i + 1, j), Vec3(xc[nxUp + i + 1], 0.5 * (yl[j] + yl[j + 1]), 0.5 * dz),
                (yl[j + 1] - yl[j]) * dz, Vec3(1, 0, 0));
    auto patch = [&](const char* name, const char* type, auto&& body) {
        Patch p; p.name = name; p.type = type;
        const int pid = (int)m.patches_.size();
        body([&](int o, Vec3 c, double A, Vec3 nrm) {
            int f = add(o, -1, c, A, nrm);
            m.faces_[f].patchID = pid;
            p.faces.push_back(f);
        });
        m.patches_.push_back(std::move(p));
    };
    patch("top_wall", "wall", [&](auto&& put) {
        for (int i = 0; i < nxT; ++i) put(cu(i, nyUp - 1), Vec3(0.5 * (xc[i] + xc[i + 1]), H, 0.5 * dz), (xc[i + 1] - xc[i]) * dz, Vec3(0, 1, 0));
    });
    patch("bottom_wall_up", "wall", [&](auto&& put) {
        for (int i = 0; i < nxUp; ++i) put(cu(i, 0), Vec3(0.5 * (xc[i] + xc[i + 1]), hs, 0.5 * dz), (xc[i + 1] - xc[i]) * dz, Vec3(0, -1, 0));
    });
    patch("step_face", "wall", [&](auto&& put) {
        for (int j = 0; j < nyDown; ++j) put(cl(0, j), Vec3(0.0, 0.5 * (yl[j] + yl[j + 1]), 0.5 * dz), (yl[j + 1] - yl[j]) * dz, Vec3(-1, 0, 0));
    });
    patch("bottom_wall_down", "wall", [&](auto&& put) {
        for (int i = 0; i < nxDown; ++i)
            put(cl(i, 0), Vec3(0.5 * (xc[nxUp + i] + xc[nxUp + i + 1]), 0.0, 0.5 * dz), (xc[nxUp + i + 1] - xc[nxUp + i]) * dz, Vec3(0, -1, 0));
    });
    patch("inlet", "inlet", [&](auto&& put) {
        for (int j = 0; j < nyUp; ++j) put(cu(0, j), Vec3(-Lu, 0.5 * (yu[j] + yu[j + 1]), 0.5 * dz), (yu[j + 1] - yu[j]) * dz, Vec3(-1, 0, 0));
    });
    patch("outlet", "outlet", [&](auto&& put) {
        for (int j = 0; j < nyUp; ++j) put(cu(nxT - 1, j), Vec3(Ld, 0.5 * (yu[j] + yu[j + 1]), 0.5 * dz), (yu[j + 1] - yu[j]) * dz, Vec3(1, 0, 0));
        for (int j = 0; j < nyDown; ++j) put(cl(nxDown - 1, j), Vec3(Ld, 0.5 * (yl[j] + yl[j + 1]), 0.5 * dz), (yl[j + 1] - yl[j]) * dz, Vec3(1, 0, 0));
    });
    for (int j = 0; j < nyUp; ++j)
        for (int i = 0; i < nxT; ++i) {
            Cell& c = m.cells_[cu(i, j)];
            c.center = Vec3(0.5 * (xc[i] + xc[i + 1]), 0.5 * (yu[j] + yu[j + 1]), 0.5 * dz);
            c.volume = (xc[i + 1] - xc[i]) * (yu[j + 1] - yu[j]) * dz;
        }
    for (int j = 0; j < nyDown; ++j)
        for (int i = 0; i < nxDown; ++i) {
            Cell& c = m.cells_[cl(i, j)];
            c.center = Vec3(0.5 * (xc[nxUp + i] + xc[nxUp + i + 1]), 0.5 * (yl[j] + yl[j + 1]), 0.5 * dz);
            c.volume = (xc[nxUp + i + 1] - xc[nxUp + i]) * (yl[j + 1] - yl[j]) * dz;
        }
    m.computeFaceGeometry();
    m.computeInterpolationWeights();
    return m;
}

// Periodic channel from node coordinates Y[i][j] (Mesh.cpp:462-585):
// exact quad centroids, wrap faces measured against the neighbour's periodic
// image.  Face order: vertical internal, wrap, horizontal internal, bottom
// wall, top wall.
Mesh Mesh::periodicFromNodes(const std::vector<double>& xN, const std::vector<std::vector<double>>& Y) {
    const int nx = (int)xN.size() - 1;
    const int ny = (int)Y[0].size() - 1;
    const double Lx = xN.back() - xN.front();
    const double dz = 1.0;
    Mesh m;
    const int pp = (nx + 1) * (ny + 1);
    m.nodes_.resize(2 * pp);
    for (int j = 0; j <= ny; ++j)
        for (int i = 0; i <= nx; ++i) {
            m.nodes_[j * (nx + 1) + i] = Vec3(xN[i], Y[i][j], 0.0);
            m.nodes_[pp + j * (nx + 1) + i] = Vec3(xN[i], Y[i][j], dz);
        }
    m.cells_.resize(nx * ny);
    auto cid = [&](int i, int j) { return j * nx + i; };
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            const double px[4] = {xN[i], xN[i + 1], xN[i + 1], xN[i]};
            const double py[4] = {Y[i][j], Y[i + 1][j], Y[i + 1][j + 1], Y[i][j + 1]};
            double A2 = 0.0, cxA = 0.0, cyA = 0.0;
            for (int k = 0; k < 4; ++k) {
                int kn = (k + 1) % 4;
                double cross = px[k] * py[kn] - px[kn] * py[k];
                A2 += cross;
                cxA += (px[k] + px[kn]) * cross;
                cyA += (py[k] + py[kn]) * cross;
            }
            double A = 0.5 * A2;
            Cell& c = m.cells_[cid(i, j)];
            c.center = Vec3(cxA / (6.0 * A), cyA / (6.0 * A), 0.5 * dz);
            c.volume = std::abs(A) * dz;
        }
    m.nInternal_ = (nx - 1) * ny + ny + nx * (ny - 1);
    m.faces_.resize(m.nInternal_ + 2 * nx);
    int fi = 0;
    auto edge = [&](Face& f, double x0, double y0, double x1, double y1, bool plusX) {
        double ex = x1 - x0, ey = y1 - y0;
        double len = std::sqrt(ex * ex + ey * ey);
        f.center = Vec3(0.5 * (x0 + x1), 0.5 * (y0 + y1), 0.5 * dz);
        f.area = len * dz;
        f.normal = plusX ? Vec3(ey / len, -ex / len, 0.0) : Vec3(-ey / len, ex / len, 0.0);
    };
    auto setDW = [&](Face& f, const Vec3& cO, const Vec3& cN) {
        f.d = cN - cO;
        f.delta = std::max(f.d.norm(), 1e-20);
        double dP = (f.center - cO).norm(), dN = (f.center - cN).norm();
        double sum = dP + dN;
        f.weight = (sum > 1e-30) ? dN / sum : 0.5;
    };
    auto link = [&](int f) {
        m.cells_[m.faces_[f].owner].faces.push_back(f);
        if (m.faces_[f].neighbor >= 0) m.cells_[m.faces_[f].neighbor].faces.push_back(f);
    };
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx - 1; ++i) {
            Face& f = m.faces_[fi];
            f.owner = cid(i, j); f.neighbor = cid(i + 1, j);
            edge(f, xN[i + 1], Y[i + 1][j], xN[i + 1], Y[i + 1][j + 1], true);
            setDW(f, m.cells_[f.owner].center, m.cells_[f.neighbor].center);
            link(fi++);
        }
    for (int j = 0; j < ny; ++j) {
        Face& f = m.faces_[fi];
        f.owner = cid(nx - 1, j); f.neighbor = cid(0, j);
        edge(f, xN[nx], Y[nx][j], xN[nx], Y[nx][j + 1], true);
        setDW(f, m.cells_[f.owner].center, m.cells_[f.neighbor].center + Vec3(Lx, 0.0, 0.0));
        link(fi++);
    }
    for (int j = 0; j < ny - 1; ++j)
        for (int i = 0; i < nx; ++i) {
            Face& f = m.faces_[fi];
            f.owner = cid(i, j); f.neighbor = cid(i, j + 1);
            edge(f, xN[i], Y[i][j + 1], xN[i + 1], Y[i + 1][j + 1], false);
            setDW(f, m.cells_[f.owner].center, m.cells_[f.neighbor].center);
            link(fi++);
        }
    for (int side = 0; side < 2; ++side) {
        Patch p; p.name = side == 0 ? "bottom_wall" : "top_wall"; p.type = "wall";
        for (int i = 0; i < nx; ++i) {
            Face& f = m.faces_[fi];
            const int j = side == 0 ? 0 : ny;
            f.owner = cid(i, side == 0 ? 0 : ny - 1); f.neighbor = -1;
            f.patchID = (int)m.patches_.size();
            edge(f, xN[i], Y[i][j], xN[i + 1], Y[i + 1][j], side == 0);
            f.d = f.center - m.cells_[f.owner].center;
            f.delta = std::max(f.d.norm(), 1e-20);
            link(fi);
            p.faces.push_back(fi++);
        }
        m.patches_.push_back(std::move(p));
    }
    return m;
}

Mesh Mesh::makeCurvedChannelPeriodic2D(const std::vector<double>& xNodes, const std::vector<double>& yBottom,
                                       double yTop, int ny, double Re, double yPlusTarget) {
    const int nx = (int)xNodes.size() - 1;
    if (nx < 3 || (int)yBottom.size() != nx + 1)
        throw std::runtime_error("makeCurvedChannelPeriodic2D: need nx+1 x nodes and matching yBottom samples");
    if (std::abs(yBottom.front() - yBottom.back()) > 1e-12)
        throw std::runtime_error("makeCurvedChannelPeriodic2D: yBottom must be periodic (first == last sample)");
    if (ny < 1) throw std::invalid_argument("makeCurvedChannelPeriodic2D: ny must be >= 1");
    double meanH = 0.0;
    for (int i = 0; i <= nx; ++i) meanH += (yTop - yBottom[i]);
    meanH /= (nx + 1);
    double stretch = 2.0;
    if (Re > 0) {
        double y1t = firstCellTarget(meanH, Re, yPlusTarget);
        if (y1t < meanH / ny) {
            auto fc = [&](double s) { return 0.5 * meanH * (1.0 + std::tanh(s * (2.0 / ny - 1.0)) / std::tanh(s)); };
            double sLo = 0.1, sHi = 20.0;
            if (y1t < fc(sHi)) {
                stretch = sHi;
            } else {
                for (int it = 0; it < 100; ++it) {
                    double sm = 0.5 * (sLo + sHi);
                    if (fc(sm) > y1t) sLo = sm; else sHi = sm;
                    if (sHi - sLo < 1e-10) break;
                }
                stretch = 0.5 * (sLo + sHi);
            }
        }
    }
    std::vector<double> eta(ny + 1);
    for (int j = 0; j <= ny; ++j) {
        double e = static_cast<double>(j) / ny;
        eta[j] = 0.5 * (1.0 + std::tanh(stretch * (2.0 * e - 1.0)) / std::tanh(stretch));
    }
    std::vector<std::vector<double>> Y(nx + 1, std::vector<double>(ny + 1));
    for (int i = 0; i <= nx; ++i)
        for (int j = 0; j <= ny; ++j) Y[i][j] = yBottom[i] + (yTop - yBottom[i]) * eta[j];
    return periodicFromNodes(xNodes, Y);
}

Mesh Mesh::makeChannelPeriodic2D(const std::vector<double>& xNodes, const std::vector<double>& yNodes) {
    if (xNodes.size() < 2 || yNodes.size() < 2)
        throw std::invalid_argument("makeChannelPeriodic2D: need >= 2 x and y nodes");
    std::vector<std::vector<double>> Y(xNodes.size(), yNodes);
    return periodicFromNodes(xNodes, Y);
}

std::vector<double> Mesh::geometricChannelNodes(int nHalf, double ratio, double H) {
    if (nHalf < 1 || !(ratio > 0)) throw std::invalid_argument("geometricChannelNodes: nHalf >= 1, ratio > 0");
    std::vector<double> h(nHalf);
    double s = 0.0;
    for (int j = 0; j < nHalf; ++j) { h[j] = std::pow(ratio, j); s += h[j]; }
    std::vector<double> y(2 * nHalf + 1);
    y[0] = 0.0;
    for (int j = 0; j < nHalf; ++j) y[j + 1] = y[j] + H * h[j] / s;
    y[nHalf] = H;
    for (int j = 1; j <= nHalf; ++j) y[nHalf + j] = 2.0 * H - y[nHalf - j];
    return y;
}

Mesh Mesh::makeChannelColumn1D(const std::vector<double>& yNodes, double dx) {
    const int ny = (int)yNodes.size() - 1;
    if (ny < 2) throw std::invalid_argument("makeChannelColumn1D: need >= 3 y nodes");
    const std::vector<double> xN = {0.0, dx};
    std::vector<std::vector<double>> Y(2, yNodes);
    Mesh full = periodicFromNodes(xN, Y);   // nx = 1: one wrap face per row
    // drop the wrap faces (owner == neighbour, zero net flux)
    Mesh m;
    m.nodes_ = full.nodes_;
    m.cells_.resize(full.cells_.size());
    for (size_t c = 0; c < full.cells_.size(); ++c) {
        m.cells_[c].center = full.cells_[c].center;
        m.cells_[c].volume = full.cells_[c].volume;
    }
    std::vector<int> newId(full.faces_.size(), -1);
    for (int f = 0; f < full.nFaces(); ++f) {
        const Face& F = full.faces_[f];
        if (!F.isBoundary() && F.owner == F.neighbor) continue;
        newId[f] = (int)m.faces_.size();
        m.faces_.push_back(F);
        if (!F.isBoundary()) m.nInternal_++;
    }
    for (int f = 0; f < m.nFaces(); ++f) {
        const Face& F = m.faces_[f];
        m.cells_[F.owner].faces.push_back(f);
        if (!F.isBoundary()) m.cells_[F.neighbor].faces.push_back(f);
    }
    for (const Patch& P : full.patches_) {
        Patch q; q.name = P.name; q.type = P.type;
        for (int f : P.faces) q.faces.push_back(newId[f]);
        m.patches_.push_back(std::move(q));
    }
    return m;
}

// Mesh.cpp:345-373 (brute force, minimum over wall-face centres)
void Mesh::computeWallDistance() {
    std::vector<Vec3> wp;
    for (const Patch& p : patches_)
        if (p.type == "wall")
            for (int f : p.faces) wp.push_back(faces_[f].center);
    wallDist_.assign(nCells(), std::numeric_limits<double>::max());
    if (wp.empty()) { std::fill(wallDist_.begin(), wallDist_.end(), 1e10); bump(); return; }
    for (int c = 0; c < nCells(); ++c) {
        double minD = std::numeric_limits<double>::max();
        const Vec3& cc = cells_[c].center;
        for (const Vec3& w : wp) {
            double d = (cc - w).norm();
            if (d < minD) minD = d;
        }
        wallDist_[c] = minD;
    }
    bump();
}

std::vector<double> Mesh::wallFaceCenters() const {
    std::vector<double> w;
    for (const Patch& p : patches_)
        if (p.type == "wall")
            for (int f : p.faces) {
                const Vec3& c = faces_[f].center;
                w.insert(w.end(), {c.x, c.y, c.z});
            }
    return w;
}

std::vector<double> Mesh::cellCenters() const {
    std::vector<double> c(3 * (size_t)nCells());
    for (int i = 0; i < nCells(); ++i) {
        c[3 * (size_t)i] = cells_[i].center.x;
        c[3 * (size_t)i + 1] = cells_[i].center.y;
        c[3 * (size_t)i + 2] = cells_[i].center.z;
    }
    return c;
}

void Mesh::setWallDistance(std::vector<double> d) {
    if ((int)d.size() != nCells()) throw std::invalid_argument("wall distance needs one value per cell");
    wallDist_ = std::move(d);
    bump();
}

int Mesh::patchByName(const std::string& name) const {
    for (int i = 0; i < nPatches(); ++i)
        if (patches_[i].name == name) return i;
    throw std::runtime_error("Patch not found: " + name);
}

void Mesh::setPatchType(const std::string& name, const std::string& type) {
    patches_[patchByName(name)].type = type;
    bump();
}

// MSH1 writer / reader (Mesh.cpp:602-713): same field order and widths
void Mesh::saveBinary(const std::string& path) const {
    std::ofstream o(path, std::ios::binary);
    if (!o) throw std::runtime_error("Cannot open for writing: " + path);
    auto wi = [&](int v) { o.write(reinterpret_cast<const char*>(&v), 4); };
    auto wd = [&](double v) { o.write(reinterpret_cast<const char*>(&v), 8); };
    auto wv = [&](const Vec3& v) { wd(v.x); wd(v.y); wd(v.z); };
    auto ws = [&](const std::string& s) { wi((int)s.size()); o.write(s.data(), s.size()); };
    o.write(reinterpret_cast<const char*>(&kMagic), 4);
    wi(nCells()); wi(nFaces()); wi(nNodes()); wi(nPatches()); wi(nInternal_);
    for (const Vec3& n : nodes_) wv(n);
    for (const Face& f : faces_) {
        wi(f.owner); wi(f.neighbor); wi(f.patchID);
        wv(f.center); wv(f.normal); wd(f.area); wd(f.delta); wv(f.d); wd(f.weight);
    }
    for (const Cell& c : cells_) {
        wv(c.center); wd(c.volume); wi((int)c.faces.size());
        for (int f : c.faces) wi(f);
    }
    for (const Patch& p : patches_) {
        ws(p.name); ws(p.type); wi((int)p.faces.size());
        for (int f : p.faces) wi(f);
    }
    if (!o) throw std::runtime_error("write failed: " + path);
}

// ------------------------------------------------------------ OpenFOAM polyMesh
namespace {

// An ASCII OpenFOAM file as tokens: comments removed, the FoamFile header
// dictionary skipped, '(' and ')' as tokens of their own.
struct FoamTokens {
    std::vector<std::string> t;
    size_t i = 0;
    explicit FoamTokens(const std::string& path) {
        std::ifstream in(path);
        if (!in) throw std::runtime_error("Cannot open " + path);
        std::stringstream ss;
        ss << in.rdbuf();
        const std::string s = ss.str();
        std::string cur;
        auto flush = [&]() { if (!cur.empty()) { t.push_back(cur); cur.clear(); } };
        for (size_t k = 0; k < s.size(); ++k) {
            const char c = s[k];
            if (c == '/' && k + 1 < s.size() && s[k + 1] == '/') {
                flush();
                while (k < s.size() && s[k] != '\n') ++k;
            } else if (c == '/' && k + 1 < s.size() && s[k + 1] == '*') {
                flush();
                k += 2;
                while (k + 1 < s.size() && !(s[k] == '*' && s[k + 1] == '/')) ++k;
                ++k;
            } else if (c == '(' || c == ')' || c == '{' || c == '}' || c == ';') {
                flush();
                t.push_back(std::string(1, c));
            } else if (std::isspace(static_cast<unsigned char>(c))) {
                flush();
            } else {
                cur.push_back(c);
            }
        }
        flush();
        // skip "FoamFile { ... }"
        if (i < t.size() && t[i] == "FoamFile") {
            while (i < t.size() && t[i] != "}") ++i;
            ++i;
        }
    }
    bool done() const { return i >= t.size(); }
    const std::string& next() {
        if (i >= t.size()) throw std::runtime_error("unexpected end of OpenFOAM file");
        return t[i++];
    }
    void expect(const char* s) {
        if (next() != s) throw std::runtime_error(std::string("malformed OpenFOAM file: expected ") + s);
    }
    int nextInt() { return std::stoi(next()); }
    double nextDouble() { return std::stod(next()); }
    // the list size that opens the data block
    int count() {
        while (!done()) {
            const std::string& s = t[i];
            if (!s.empty() && std::isdigit(static_cast<unsigned char>(s[0]))) return nextInt();
            ++i;
        }
        throw std::runtime_error("Could not read count");
    }
};

}  // namespace

// Mesh::loadOpenFOAM (Mesh.cpp:88-265): points, faces, owner, neighbour and
// boundary of an ASCII polyMesh; the cell count from the largest cell index;
// each cell's faces in face order (owner side, then neighbour side); face
// centre = node average, area vector = sum of the node-fan triangles about
// that centre; then cell centres and volumes, face deltas and weights.
Mesh Mesh::loadOpenFOAM(const std::string& polyMeshDir) {
    namespace fs = std::filesystem;
    const fs::path dir(polyMeshDir);
    Mesh m;
    {
        FoamTokens T((dir / "points").string());
        const int n = T.count();
        T.expect("(");
        m.nodes_.resize(n);
        for (int k = 0; k < n; ++k) {
            T.expect("(");
            const double x = T.nextDouble(), y = T.nextDouble(), z = T.nextDouble();
            m.nodes_[k] = Vec3(x, y, z);
            T.expect(")");
        }
    }
    std::vector<std::vector<int>> faceNodes;
    {
        FoamTokens T((dir / "faces").string());
        const int n = T.count();
        T.expect("(");
        faceNodes.resize(n);
        for (int k = 0; k < n; ++k) {
            const int nv = T.nextInt();
            T.expect("(");
            faceNodes[k].resize(nv);
            for (int& v : faceNodes[k]) {
                v = T.nextInt();
                if (v < 0 || v >= (int)m.nodes_.size()) throw std::runtime_error("face node index out of range");
            }
            T.expect(")");
        }
        m.faces_.resize(n);
    }
    {
        FoamTokens T((dir / "owner").string());
        const int n = T.count();
        if (n > (int)m.faces_.size()) throw std::runtime_error("owner list longer than the face list");
        T.expect("(");
        for (int k = 0; k < n; ++k) m.faces_[k].owner = T.nextInt();
    }
    {
        FoamTokens T((dir / "neighbour").string());
        const int n = T.count();
        if (n > (int)m.faces_.size()) throw std::runtime_error("neighbour list longer than the face list");
        m.nInternal_ = n;
        T.expect("(");
        for (int k = 0; k < n; ++k) m.faces_[k].neighbor = T.nextInt();
    }
    int nCells = 0;
    for (const Face& f : m.faces_) {
        if (f.owner < 0) throw std::runtime_error("face without an owner cell");
        nCells = std::max(nCells, f.owner + 1);
        if (f.neighbor >= 0) nCells = std::max(nCells, f.neighbor + 1);
    }
    m.cells_.resize(nCells);
    for (int fi = 0; fi < (int)m.faces_.size(); ++fi) {
        m.cells_[m.faces_[fi].owner].faces.push_back(fi);
        if (m.faces_[fi].neighbor >= 0) m.cells_[m.faces_[fi].neighbor].faces.push_back(fi);
    }
    if (fs::exists(dir / "boundary")) {
        FoamTokens T((dir / "boundary").string());
        const int np = T.count();
        T.expect("(");
        m.patches_.resize(np);
        for (int p = 0; p < np; ++p) {
            Patch& P = m.patches_[p];
            P.name = T.next();
            T.expect("{");
            int nF = 0, start = 0;
            int depth = 1;
            while (depth > 0) {
                const std::string key = T.next();
                if (key == "{") { ++depth; continue; }
                if (key == "}") { --depth; continue; }
                if (key == ";" || depth > 1) continue;
                if (key == "type") P.type = T.next();
                else if (key == "nFaces") nF = T.nextInt();
                else if (key == "startFace") start = T.nextInt();
            }
            if (nF < 0 || start < 0 || start + nF > (int)m.faces_.size())
                throw std::runtime_error("patch face range out of bounds: " + P.name);
            P.faces.resize(nF);
            for (int k = 0; k < nF; ++k) {
                P.faces[k] = start + k;
                m.faces_[start + k].patchID = p;
            }
        }
    }
    for (int fi = 0; fi < (int)m.faces_.size(); ++fi) {
        const std::vector<int>& fn = faceNodes[fi];
        const int nv = (int)fn.size();
        Vec3 ctr;
        for (int j = 0; j < nv; ++j) ctr = ctr + m.nodes_[fn[j]];
        ctr = ctr / static_cast<double>(nv);
        Vec3 sa;
        for (int j = 0; j < nv; ++j) {
            const Vec3 a = m.nodes_[fn[j]] - ctr, b = m.nodes_[fn[(j + 1) % nv]] - ctr;
            const Vec3 c(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
            sa = sa + c * 0.5;
        }
        Face& f = m.faces_[fi];
        f.center = ctr;
        f.area = sa.norm();
        f.normal = (f.area > 1e-300) ? sa / f.area : Vec3();
    }
    m.computeCellCentersAndVolumes();
    m.computeFaceGeometry();
    m.computeInterpolationWeights();
    return m;
}

// Mesh.cpp:268-297: centre = mean of the face centres; volume from the
// divergence theorem about that centre, |sum_f ±S_f . (c_f - c_C)| / 3.
void Mesh::computeCellCentersAndVolumes() {
    for (Cell& c : cells_) {
        Vec3 ctr;
        for (int fi : c.faces) ctr = ctr + faces_[fi].center;
        c.center = ctr / static_cast<double>(c.faces.size());
    }
    for (int ci = 0; ci < (int)cells_.size(); ++ci) {
        Cell& c = cells_[ci];
        double vol = 0.0;
        for (int fi : c.faces) {
            const Face& f = faces_[fi];
            const Vec3 Sf = f.normal * f.area;
            const double sign = (f.owner == ci) ? 1.0 : -1.0;
            vol += sign * Sf.dot(f.center - c.center);
        }
        c.volume = std::abs(vol) / 3.0;
        if (c.volume < 1e-30) c.volume = 1e-20;
    }
}

// Mesh::loadFromFile (Mesh.cpp:21-47): a directory holding points and faces,
// or a case directory with constant/polyMesh, is read as OpenFOAM; a file
// ending in .msh as MSH1; a .foam file or a path without extension as a
// polyMesh directory.
Mesh Mesh::loadFromFile(const std::string& path) {
    namespace fs = std::filesystem;
    if (fs::is_directory(path)) {
        const fs::path p(path);
        if (fs::exists(p / "points") && fs::exists(p / "faces")) return loadOpenFOAM(path);
        const fs::path poly = p / "constant" / "polyMesh";
        if (fs::exists(poly / "points")) return loadOpenFOAM(poly.string());
        throw std::runtime_error("Cannot find polyMesh in: " + path);
    }
    const std::string ext = fs::path(path).extension().string();
    if (ext == ".foam" || ext.empty()) return loadOpenFOAM(path);
    if (ext != ".msh") throw std::runtime_error("Unknown mesh format: " + ext);
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("Cannot open binary mesh: " + path);
    auto ri = [&]() { int v = 0; in.read(reinterpret_cast<char*>(&v), 4); return v; };
    auto rd = [&]() { double v = 0; in.read(reinterpret_cast<char*>(&v), 8); return v; };
    auto rv = [&]() { double a = rd(), b = rd(), c = rd(); return Vec3(a, b, c); };
    auto rs = [&]() {
        int n = ri();
        if (n < 0 || n > (1 << 20)) throw std::runtime_error("corrupt MSH1 string");
        std::string s(n, '\0'); in.read(&s[0], n); return s;
    };
    uint32_t magic = 0;
    in.read(reinterpret_cast<char*>(&magic), 4);
    if (magic != kMagic) throw std::runtime_error("Invalid binary mesh magic number");
    Mesh m;
    int nc = ri(), nf = ri(), nn = ri(), np = ri();
    m.nInternal_ = ri();
    if (nc < 0 || nf < 0 || nn < 0 || np < 0 || m.nInternal_ < 0 || m.nInternal_ > nf)
        throw std::runtime_error("corrupt MSH1 header");
    m.nodes_.resize(nn);
    for (auto& v : m.nodes_) v = rv();
    m.faces_.resize(nf);
    for (Face& f : m.faces_) {
        f.owner = ri(); f.neighbor = ri(); f.patchID = ri();
        f.center = rv(); f.normal = rv(); f.area = rd(); f.delta = rd(); f.d = rv(); f.weight = rd();
    }
    m.cells_.resize(nc);
    for (Cell& c : m.cells_) {
        c.center = rv(); c.volume = rd();
        int k = ri();
        if (k < 0 || k > nf) throw std::runtime_error("corrupt MSH1 cell");
        c.faces.resize(k);
        for (int& f : c.faces) f = ri();
    }
    m.patches_.resize(np);
    for (Patch& p : m.patches_) {
        p.name = rs(); p.type = rs();
        int k = ri();
        if (k < 0 || k > nf) throw std::runtime_error("corrupt MSH1 patch");
        p.faces.resize(k);
        for (int& f : p.faces) f = ri();
    }
    if (!in) throw std::runtime_error("truncated MSH1 file: " + path);
    return m;
}

const rsg_mesh_desc& Mesh::desc() const {
    Flat& F = flat_;
    if (F.builtFor == version_ && F.builtFor >= 0) return F.desc;
    const int nf = nFaces(), nif = nInternal_, nc = nCells();
    F.owner.resize(nf); F.neighbor.resize(nif);
    F.normal.resize(3 * (size_t)nf); F.area.resize(nf); F.delta.resize(nf);
    F.d.resize(3 * (size_t)nf); F.weight.resize(nf);
    for (int f = 0; f < nf; ++f) {
        const Face& x = faces_[f];
        F.owner[f] = x.owner;
        if (f < nif) F.neighbor[f] = x.neighbor;
        F.normal[3 * f] = x.normal.x; F.normal[3 * f + 1] = x.normal.y; F.normal[3 * f + 2] = x.normal.z;
        F.area[f] = x.area; F.delta[f] = x.delta; F.weight[f] = x.weight;
        F.d[3 * f] = x.d.x; F.d[3 * f + 1] = x.d.y; F.d[3 * f + 2] = x.d.z;
    }
    F.vol.resize(nc);
    for (int c = 0; c < nc; ++c) F.vol[c] = cells_[c].volume;
    F.wd = wallDist_;
    F.ptype.resize(nPatches()); F.pptr.assign(nPatches() + 1, 0); F.pfaces.clear();
    for (int p = 0; p < nPatches(); ++p) {
        const std::string& t = patches_[p].type;
        F.ptype[p] = t == "wall" ? RSG_PATCH_WALL : t == "inlet" ? RSG_PATCH_INLET : t == "outlet" ? RSG_PATCH_OUTLET
                   : t == "symmetry" ? RSG_PATCH_SYMMETRY : RSG_PATCH_OTHER;
        for (int f : patches_[p].faces) F.pfaces.push_back(f);
        F.pptr[p + 1] = (int)F.pfaces.size();
    }
    rsg_mesh_desc& d = F.desc;
    d.n_cells = nc; d.n_faces = nf; d.n_internal_faces = nif; d.n_patches = nPatches();
    d.owner = F.owner.data(); d.neighbor = F.neighbor.data(); d.face_normal = F.normal.data();
    d.face_area = F.area.data(); d.face_delta = F.delta.data(); d.face_d = F.d.data();
    d.face_weight = F.weight.data(); d.cell_volume = F.vol.data(); d.wall_distance = F.wd.data();
    d.patch_type = F.ptype.data(); d.patch_face_ptr = F.pptr.data(); d.patch_faces = F.pfaces.data();
    F.builtFor = version_;
    return d;
}

}  // namespace rsh
