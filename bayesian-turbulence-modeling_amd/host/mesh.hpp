// mesh.hpp -- unstructured finite-volume mesh of the forward model (host side).
//
// Same data model as the reference (core/include/Mesh.hpp:11-197): faces with
// owner/neighbour, centre, unit normal, area, delta = |d|, d and interpolation
// weight; internal faces first, boundary faces after; named, typed patches;
// 2D meshes extruded by dz = 1 with no front/back faces.  The factories
// reproduce the reference's face and cell ordering and geometry arithmetic so
// a mesh built here is index- and bit-identical to the reference's.
#pragma once
#include <cmath>
#include <string>
#include <vector>

#include "rsg.h"

namespace rsh {

struct Vec3 {
    double x = 0.0, y = 0.0, z = 0.0;
    Vec3() = default;
    Vec3(double a, double b, double c) : x(a), y(b), z(c) {}
    Vec3 operator+(const Vec3& o) const { return {x + o.x, y + o.y, z + o.z}; }
    Vec3 operator-(const Vec3& o) const { return {x - o.x, y - o.y, z - o.z}; }
    Vec3 operator*(double s) const { return {x * s, y * s, z * s}; }
    Vec3 operator/(double s) const { return {x / s, y / s, z / s}; }
    double dot(const Vec3& o) const { return x * o.x + y * o.y + z * o.z; }
    double norm() const { return std::sqrt(x * x + y * y + z * z); }
    double norm2() const { return x * x + y * y + z * z; }
};

struct Face {
    int owner = -1, neighbor = -1;
    Vec3 center, normal;
    double area = 0.0, delta = 0.0;
    Vec3 d;
    double weight = 0.5;
    int patchID = -1;
    bool isBoundary() const { return neighbor < 0; }
};

struct Cell {
    Vec3 center;
    double volume = 0.0;
    std::vector<int> faces;
};

struct Patch {
    std::string name, type;
    std::vector<int> faces;
};

class Mesh {
public:
    // factories (reference Mesh.cpp:719-1546, :395-598)
    static Mesh makeChannel2D(int nx, int ny, double Lx, double Ly);
    static Mesh makeChannel2D(int nx, int ny, double Lx, double Ly, double Re, double yPlusTarget);
    static Mesh makePlate2D(int nx, int ny, double Lx, double Ly, double Re, double yPlusTarget);
    static Mesh makeBackwardFacingStep2D(int nxUp, int nxDown, int nyUp, int nyDown, double Lu,
                                         double Ld, double hs, double H, double Re = 0.0,
                                         double yPlusTarget = 1.0);
    static Mesh makeCurvedChannelPeriodic2D(const std::vector<double>& xNodes,
                                            const std::vector<double>& yBottom, double yTop, int ny,
                                            double Re, double yPlusTarget);
    // Streamwise-periodic flat channel with explicit wall-normal node
    // coordinates (same topology and face order as the curved factory).
    static Mesh makeChannelPeriodic2D(const std::vector<double>& xNodes, const std::vector<double>& yNodes);
    // Node distribution of a channel of height 2H with symmetric geometric
    // stretching: nHalf cells per half, growth ratio r from each wall.
    static std::vector<double> geometricChannelNodes(int nHalf, double ratio, double H);
    // Streamwise-invariant single-column channel: the periodic column with its
    // streamwise faces dropped (they carry zero net flux in x-invariant flow).
    static Mesh makeChannelColumn1D(const std::vector<double>& yNodes, double dx);

    static Mesh loadFromFile(const std::string& path);      // MSH1 file or OpenFOAM polyMesh
    static Mesh loadOpenFOAM(const std::string& polyMeshDir);
    void saveBinary(const std::string& path) const;

    void computeWallDistance();
    // Wall-face centres in patch order (the points computeWallDistance searches)
    // and the cell centres, as x,y,z triples; and installing a wall distance
    // computed elsewhere (the device kernel).
    std::vector<double> wallFaceCenters() const;
    std::vector<double> cellCenters() const;
    void setWallDistance(std::vector<double> d);
    void setPatchType(const std::string& name, const std::string& type);
    int patchByName(const std::string& name) const;

    int nCells() const { return (int)cells_.size(); }
    int nFaces() const { return (int)faces_.size(); }
    int nNodes() const { return (int)nodes_.size(); }
    int nPatches() const { return (int)patches_.size(); }
    int nInternalFaces() const { return nInternal_; }
    const Cell& cell(int i) const { return cells_[i]; }
    const Face& face(int i) const { return faces_[i]; }
    const Vec3& node(int i) const { return nodes_[i]; }
    const Patch& patch(int i) const { return patches_[i]; }
    const std::vector<double>& wallDistance() const { return wallDist_; }
    bool hasWallDistance() const { return (int)wallDist_.size() == nCells(); }

    // Flattened view for the C-ABI (valid until the mesh changes).
    const rsg_mesh_desc& desc() const;
    // identity of the geometry for device-side caching
    long version() const { return version_; }

private:
    std::vector<Cell> cells_;
    std::vector<Face> faces_;
    std::vector<Vec3> nodes_;
    std::vector<Patch> patches_;
    std::vector<double> wallDist_;
    int nInternal_ = 0;
    long version_ = 0;

    struct Flat {
        std::vector<int32_t> owner, neighbor, ptype, pptr, pfaces;
        std::vector<double> normal, area, delta, d, weight, vol, wd;
        rsg_mesh_desc desc{};
        long builtFor = -1;
    };
    mutable Flat flat_;

    void computeFaceGeometry();
    void computeInterpolationWeights();
    void computeCellCentersAndVolumes();
    void bump();
    static Mesh periodicFromNodes(const std::vector<double>& xNodes,
                                  const std::vector<std::vector<double>>& Y);   // Y[i][j]
    static Mesh cartesian(int nx, int ny, const std::vector<double>& xc, const std::vector<double>& yc,
                          const char* topType);
};

}  // namespace rsh
