// rsg_prep.h -- host-side preparation of the device data layout.
//
// Turns an rsg_mesh_desc (the reference's Mesh, flattened) into the
// structure-of-arrays image of RsgMeshDev, and computes the byte layout of a
// member workspace.  Plain C++: used by the HIP library (rsg_capi.hip) and by
// the sequential emulation build of the CPU tests.
#pragma once
#include <stdint.h>
#include <string.h>
#include <string>
#include <vector>
#include <algorithm>

#include "rsg.h"
#include "rsg_layout.h"

namespace rsg {

// Bump allocator over one block; with base == nullptr it only measures.
struct Carver {
    char* base = nullptr;
    size_t off = 0;
    template <class T> T* take(size_t n) {
        off = (off + 255) & ~size_t(255);
        T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;
        off += n * sizeof(T);
        return p;
    }
};

// Host image of one mesh, in the order RsgMeshDev expects.
struct PrepMesh {
    int nC = 0, nIF = 0, nF = 0, nB = 0, hasOutlet = 0, nnz0 = 0, nP = 0;
    std::vector<int> own, nbr, cfPtr, cfIdx, bFlags, rp0, ci0, src0, ellCnt, ellNb, ellFace, ovPtr, ovFace, ovNb,
        bPtr, bFace;
    int ellK = 0;
    int planar = 0;   // every face normal and area vector has a zero z component
    std::vector<double> w, dlt, area, nx, ny, nz, sfx, sfy, sfz, ex, ey, ez, vol, wd, pinY;
    std::vector<int> patchType, patchPtr, patchFaces;
};

inline bool prep_mesh(const rsg_mesh_desc* d, PrepMesh& P, std::string& err) {
    if (!d || d->n_cells <= 0 || d->n_faces < d->n_internal_faces || d->n_internal_faces < 0 ||
        d->n_patches < 0 || !d->owner || (d->n_internal_faces > 0 && !d->neighbor) ||
        !d->face_normal || !d->face_area || !d->face_delta || !d->face_d || !d->face_weight ||
        !d->cell_volume || !d->wall_distance ||
        (d->n_patches > 0 && (!d->patch_type || !d->patch_face_ptr || !d->patch_faces))) {
        err = "mesh descriptor: missing array or inconsistent sizes";
        return false;
    }
    const int nC = d->n_cells, nF = d->n_faces, nIF = d->n_internal_faces, nB = nF - nIF;
    P.nC = nC; P.nF = nF; P.nIF = nIF; P.nB = nB; P.nP = d->n_patches;
    for (int f = 0; f < nF; ++f) {
        const int o = d->owner[f];
        if (o < 0 || o >= nC) { err = "mesh descriptor: owner out of range"; return false; }
        if (f < nIF && (d->neighbor[f] < 0 || d->neighbor[f] >= nC)) {
            err = "mesh descriptor: neighbour out of range"; return false;
        }
    }
    P.own.assign(d->owner, d->owner + nF);
    P.nbr.assign(d->neighbor, d->neighbor + nIF);
    P.w.resize(nIF); P.ex.resize(nIF); P.ey.resize(nIF); P.ez.resize(nIF);
    P.dlt.resize(nF); P.area.resize(nF); P.nx.resize(nF); P.ny.resize(nF); P.nz.resize(nF);
    P.sfx.resize(nF); P.sfy.resize(nF); P.sfz.resize(nF);
    for (int f = 0; f < nF; ++f) {
        const double dl = d->face_delta[f], A = d->face_area[f];
        const double* n = d->face_normal + 3 * f;
        P.dlt[f] = std::max(dl, 1e-20);       // SIMPLESolver.cpp:66 et al.
        P.area[f] = A;
        P.nx[f] = n[0]; P.ny[f] = n[1]; P.nz[f] = n[2];
        P.sfx[f] = n[0] * A; P.sfy[f] = n[1] * A; P.sfz[f] = n[2] * A;   // Field.hpp:132
        if (f < nIF) {
            P.w[f] = d->face_weight[f];
            const double dd = std::max(dl, 1e-30);   // SIMPLESolver.cpp:401
            P.ex[f] = d->face_d[3 * f] / dd; P.ey[f] = d->face_d[3 * f + 1] / dd;
            P.ez[f] = d->face_d[3 * f + 2] / dd;
        }
    }
    P.vol.assign(d->cell_volume, d->cell_volume + nC);
    P.wd.assign(d->wall_distance, d->wall_distance + nC);
    // cell -> faces in ascending face index (the reference's scatter order)
    std::vector<int> cnt(nC + 1, 0);
    for (int f = 0; f < nF; ++f) {
        cnt[P.own[f] + 1]++;
        if (f < nIF) cnt[P.nbr[f] + 1]++;
    }
    P.cfPtr.assign(nC + 1, 0);
    for (int c = 0; c < nC; ++c) P.cfPtr[c + 1] = P.cfPtr[c] + cnt[c + 1];
    P.cfIdx.assign(P.cfPtr[nC], 0);
    std::vector<int> fill(P.cfPtr.begin(), P.cfPtr.end() - 1);
    for (int f = 0; f < nF; ++f) {
        P.cfIdx[fill[P.own[f]]++] = f << 1;
        if (f < nIF) P.cfIdx[fill[P.nbr[f]]++] = (f << 1) | 1;
    }
    // patches: flags, wall pin distance (last wall face of the cell in patch
    // order, SIMPLESolver.cpp:884-908), outlet presence (:337-340)
    P.bFlags.assign(nB, 0);
    P.pinY.assign(nC, 0.0);
    P.hasOutlet = 0;
    if (d->n_patches > 0) {
        P.patchType.assign(d->patch_type, d->patch_type + d->n_patches);
        P.patchPtr.assign(d->patch_face_ptr, d->patch_face_ptr + d->n_patches + 1);
    } else {
        P.patchType.clear();
        P.patchPtr.assign(1, 0);   // no patches: empty face list
    }
    if (P.patchPtr[0] != 0) { err = "mesh descriptor: patch_face_ptr[0] != 0"; return false; }
    for (int p = 0; p < d->n_patches; ++p) {
        const int a = P.patchPtr[p], b = P.patchPtr[p + 1];
        if (b < a || b > nB) { err = "mesh descriptor: patch_face_ptr not monotone"; return false; }
        for (int j = a; j < b; ++j) {
            const int f = d->patch_faces[j];
            if (f < nIF || f >= nF) { err = "mesh descriptor: patch face is not a boundary face"; return false; }
            if (P.patchType[p] == RSG_PATCH_WALL) {
                P.bFlags[f - nIF] |= RSG_BF_WALL;
                P.pinY[P.own[f]] = std::max(d->face_delta[f], 1e-20);
            }
            if (P.patchType[p] == RSG_PATCH_OUTLET) P.bFlags[f - nIF] |= RSG_BF_OUTLET;
        }
        if (P.patchType[p] == RSG_PATCH_OUTLET && b > a) P.hasOutlet = 1;
    }
    P.patchFaces.assign(d->patch_faces, d->patch_faces + (d->n_patches > 0 ? P.patchPtr[d->n_patches] : 0));
    // ELL image of the internal faces of each cell, in ascending face order
    P.ellCnt.assign(nC, 0);
    P.ellNb.assign((size_t)RSG_ELL_K * nC, 0);
    P.ellFace.assign((size_t)RSG_ELL_K * nC, -1);
    P.ovPtr.assign(nC + 1, 0); P.ovFace.clear(); P.ovNb.clear();
    P.bPtr.assign(nC + 1, 0); P.bFace.clear();
    for (int c = 0; c < nC; ++c) {
        int k = 0;
        for (int e = P.cfPtr[c]; e < P.cfPtr[c + 1]; ++e) {
            const int code = P.cfIdx[e], f = code >> 1;
            if (f >= nIF) { P.bFace.push_back(f); continue; }
            const int other = (code & 1) ? P.own[f] : P.nbr[f];
            if (k < RSG_ELL_K) {
                P.ellNb[(size_t)k * nC + c] = other;
                P.ellFace[(size_t)k * nC + c] = code;
            } else {
                P.ovFace.push_back(code); P.ovNb.push_back(other);
            }
            ++k;
        }
        P.ellCnt[c] = std::min(k, RSG_ELL_K);
        for (int j = k; j < RSG_ELL_K; ++j) P.ellNb[(size_t)j * nC + c] = c;
        P.ovPtr[c + 1] = (int)P.ovFace.size();
        P.bPtr[c + 1] = (int)P.bFace.size();
    }
    P.ellK = P.ovFace.empty() ? 1 : 0;
    P.planar = 1;
    for (int f = 0; f < nF && P.planar; ++f)
        if (d->face_normal[3 * f + 2] != 0.0) P.planar = 0;
    if (P.ovFace.empty()) { P.ovFace.push_back(0); P.ovNb.push_back(0); }   // keep every array non-empty
    if (P.bFace.empty()) P.bFace.push_back(0);
    // level-0 CSR pattern, built as lduToCSR does (LinearSolver.hpp:344-404):
    // diagonal first, then faces in order, then a stable insertion sort per row
    P.nnz0 = nC + 2 * nIF;
    P.rp0.assign(nC + 1, 0);
    std::vector<int> rc(nC, 1);
    for (int f = 0; f < nIF; ++f) { rc[P.own[f]]++; rc[P.nbr[f]]++; }
    for (int i = 0; i < nC; ++i) P.rp0[i + 1] = P.rp0[i] + rc[i];
    P.ci0.assign(P.nnz0, 0);
    P.src0.assign(P.nnz0, 0);
    std::vector<int> pos(P.rp0.begin(), P.rp0.end() - 1);
    for (int i = 0; i < nC; ++i) { P.ci0[pos[i]] = i; P.src0[pos[i]] = -1 - i; pos[i]++; }
    for (int f = 0; f < nIF; ++f) {
        const int o = P.own[f], nb = P.nbr[f];
        P.ci0[pos[o]] = nb; P.src0[pos[o]] = f; pos[o]++;
        P.ci0[pos[nb]] = o; P.src0[pos[nb]] = nIF + f; pos[nb]++;
    }
    for (int i = 0; i < nC; ++i) {
        const int start = P.rp0[i], end = P.rp0[i + 1];
        for (int a = start + 1; a < end; ++a) {
            const int kc = P.ci0[a], ks = P.src0[a];
            int b = a - 1;
            while (b >= start && P.ci0[b] > kc) { P.ci0[b + 1] = P.ci0[b]; P.src0[b + 1] = P.src0[b]; --b; }
            P.ci0[b + 1] = kc; P.src0[b + 1] = ks;
        }
    }
    return true;
}

// Topology-only mesh of an LDU system (every face internal, no geometry),
// used by the stand-alone linear-solve entry point.
inline bool prep_topology(int n, int nif, const int* own, const int* nbr, PrepMesh& P, std::string& err) {
    if (n <= 0 || nif < 0 || !own || (nif > 0 && !nbr)) { err = "LDU system: bad sizes or null topology"; return false; }
    std::vector<double> z3(3 * (size_t)std::max(nif, 1), 0.0), z(std::max(nif, 1), 0.0), one(n, 1.0);
    rsg_mesh_desc d;
    d.n_cells = n; d.n_faces = nif; d.n_internal_faces = nif; d.n_patches = 0;
    d.owner = own; d.neighbor = nbr; d.face_normal = z3.data(); d.face_area = z.data();
    d.face_delta = z.data(); d.face_d = z3.data(); d.face_weight = z.data();
    d.cell_volume = one.data(); d.wall_distance = one.data();
    d.patch_type = nullptr; d.patch_face_ptr = nullptr; d.patch_faces = nullptr;
    return prep_mesh(&d, P, err);
}

// Byte offsets of every RsgMeshDev array inside one contiguous block.
struct MeshOffsets {
    size_t own, nbr, w, dlt, area, nx, ny, nz, sfx, sfy, sfz, ex, ey, ez, vol, wd, pinY,
        cfPtr, cfIdx, bFlags, rp0, ci0, src0, ellCnt, ellNb, ellFace, ovPtr, ovFace, ovNb, bPtr, bFace, total;
};

inline MeshOffsets mesh_offsets(const PrepMesh& P) {
    Carver cv;
    MeshOffsets o;
    auto at = [&](size_t bytes) { cv.off = (cv.off + 255) & ~size_t(255); size_t r = cv.off; cv.off += bytes; return r; };
    o.own = at(4 * P.own.size()); o.nbr = at(4 * P.nbr.size()); o.w = at(8 * P.w.size());
    o.dlt = at(8 * P.dlt.size()); o.area = at(8 * P.area.size());
    o.nx = at(8 * P.nx.size()); o.ny = at(8 * P.ny.size()); o.nz = at(8 * P.nz.size());
    o.sfx = at(8 * P.sfx.size()); o.sfy = at(8 * P.sfy.size()); o.sfz = at(8 * P.sfz.size());
    o.ex = at(8 * P.ex.size()); o.ey = at(8 * P.ey.size()); o.ez = at(8 * P.ez.size());
    o.vol = at(8 * P.vol.size()); o.wd = at(8 * P.wd.size()); o.pinY = at(8 * P.pinY.size());
    o.cfPtr = at(4 * P.cfPtr.size()); o.cfIdx = at(4 * P.cfIdx.size()); o.bFlags = at(4 * P.bFlags.size());
    o.rp0 = at(4 * P.rp0.size()); o.ci0 = at(4 * P.ci0.size()); o.src0 = at(4 * P.src0.size());
    o.ellCnt = at(4 * P.ellCnt.size()); o.ellNb = at(4 * P.ellNb.size()); o.ellFace = at(4 * P.ellFace.size());
    o.ovPtr = at(4 * P.ovPtr.size()); o.ovFace = at(4 * P.ovFace.size()); o.ovNb = at(4 * P.ovNb.size());
    o.bPtr = at(4 * P.bPtr.size()); o.bFace = at(4 * P.bFace.size());
    o.total = (cv.off + 255) & ~size_t(255);
    return o;
}

// Writes the arrays into a host staging block laid out by mesh_offsets.
inline void mesh_stage(const PrepMesh& P, const MeshOffsets& o, char* base) {
    auto put = [&](size_t off, const void* src, size_t bytes) { if (bytes) memcpy(base + off, src, bytes); };
    put(o.own, P.own.data(), 4 * P.own.size()); put(o.nbr, P.nbr.data(), 4 * P.nbr.size());
    put(o.w, P.w.data(), 8 * P.w.size()); put(o.dlt, P.dlt.data(), 8 * P.dlt.size());
    put(o.area, P.area.data(), 8 * P.area.size());
    put(o.nx, P.nx.data(), 8 * P.nx.size()); put(o.ny, P.ny.data(), 8 * P.ny.size());
    put(o.nz, P.nz.data(), 8 * P.nz.size());
    put(o.sfx, P.sfx.data(), 8 * P.sfx.size()); put(o.sfy, P.sfy.data(), 8 * P.sfy.size());
    put(o.sfz, P.sfz.data(), 8 * P.sfz.size());
    put(o.ex, P.ex.data(), 8 * P.ex.size()); put(o.ey, P.ey.data(), 8 * P.ey.size());
    put(o.ez, P.ez.data(), 8 * P.ez.size());
    put(o.vol, P.vol.data(), 8 * P.vol.size()); put(o.wd, P.wd.data(), 8 * P.wd.size());
    put(o.pinY, P.pinY.data(), 8 * P.pinY.size());
    put(o.cfPtr, P.cfPtr.data(), 4 * P.cfPtr.size()); put(o.cfIdx, P.cfIdx.data(), 4 * P.cfIdx.size());
    put(o.bFlags, P.bFlags.data(), 4 * P.bFlags.size());
    put(o.rp0, P.rp0.data(), 4 * P.rp0.size()); put(o.ci0, P.ci0.data(), 4 * P.ci0.size());
    put(o.src0, P.src0.data(), 4 * P.src0.size());
    put(o.ellCnt, P.ellCnt.data(), 4 * P.ellCnt.size()); put(o.ellNb, P.ellNb.data(), 4 * P.ellNb.size());
    put(o.ellFace, P.ellFace.data(), 4 * P.ellFace.size());
    put(o.ovPtr, P.ovPtr.data(), 4 * P.ovPtr.size()); put(o.ovFace, P.ovFace.data(), 4 * P.ovFace.size());
    put(o.ovNb, P.ovNb.data(), 4 * P.ovNb.size());
    put(o.bPtr, P.bPtr.data(), 4 * P.bPtr.size()); put(o.bFace, P.bFace.data(), 4 * P.bFace.size());
}

// RsgMeshDev whose pointers address `base` (host staging or device copy).
inline RsgMeshDev mesh_view(const PrepMesh& P, const MeshOffsets& o, char* base) {
    RsgMeshDev m;
    m.nC = P.nC; m.nIF = P.nIF; m.nF = P.nF; m.nB = P.nB; m.hasOutlet = P.hasOutlet; m.nnz0 = P.nnz0;
    m.own = (const int*)(base + o.own); m.nbr = (const int*)(base + o.nbr);
    m.w = (const double*)(base + o.w); m.dlt = (const double*)(base + o.dlt);
    m.area = (const double*)(base + o.area);
    m.nx = (const double*)(base + o.nx); m.ny = (const double*)(base + o.ny); m.nz = (const double*)(base + o.nz);
    m.sfx = (const double*)(base + o.sfx); m.sfy = (const double*)(base + o.sfy); m.sfz = (const double*)(base + o.sfz);
    m.ex = (const double*)(base + o.ex); m.ey = (const double*)(base + o.ey); m.ez = (const double*)(base + o.ez);
    m.vol = (const double*)(base + o.vol); m.wd = (const double*)(base + o.wd);
    m.pinY = (const double*)(base + o.pinY);
    m.cfPtr = (const int*)(base + o.cfPtr); m.cfIdx = (const int*)(base + o.cfIdx);
    m.bFlags = (const int*)(base + o.bFlags);
    m.rp0 = (const int*)(base + o.rp0); m.ci0 = (const int*)(base + o.ci0); m.src0 = (const int*)(base + o.src0);
    m.ellK = P.ellK;
    m.ellCnt = (const int*)(base + o.ellCnt); m.ellNb = (const int*)(base + o.ellNb);
    m.ellFace = (const int*)(base + o.ellFace);
    m.ovPtr = (const int*)(base + o.ovPtr); m.ovFace = (const int*)(base + o.ovFace);
    m.ovNb = (const int*)(base + o.ovNb);
    m.bPtr = (const int*)(base + o.bPtr); m.bFace = (const int*)(base + o.bFace);
    return m;
}

// Per-problem BC block: [velT velV pT pV kT kV omT omV]
struct BcOffsets { size_t velT, velV, pT, pV, kT, kV, omT, omV, total; };
inline BcOffsets bc_offsets(int nB) {
    Carver cv; BcOffsets o;
    auto at = [&](size_t bytes) { cv.off = (cv.off + 255) & ~size_t(255); size_t r = cv.off; cv.off += bytes; return r; };
    o.velT = at(4 * nB); o.velV = at(24 * nB); o.pT = at(4 * nB); o.pV = at(8 * nB);
    o.kT = at(4 * nB); o.kV = at(8 * nB); o.omT = at(4 * nB); o.omV = at(8 * nB);
    o.total = (cv.off + 255) & ~size_t(255);
    return o;
}
inline void bc_stage(const rsg_bc_desc* d, int nB, const BcOffsets& o, char* base) {
    memcpy(base + o.velT, d->vel_type, 4 * nB); memcpy(base + o.velV, d->vel_value, 24 * nB);
    memcpy(base + o.pT, d->p_type, 4 * nB); memcpy(base + o.pV, d->p_value, 8 * nB);
    memcpy(base + o.kT, d->k_type, 4 * nB); memcpy(base + o.kV, d->k_value, 8 * nB);
    memcpy(base + o.omT, d->om_type, 4 * nB); memcpy(base + o.omV, d->om_value, 8 * nB);
}
inline RsgBcDev bc_view(const BcOffsets& o, char* base) {
    RsgBcDev b;
    b.velT = (const int*)(base + o.velT); b.velV = (const double*)(base + o.velV);
    b.pT = (const int*)(base + o.pT); b.pV = (const double*)(base + o.pV);
    b.kT = (const int*)(base + o.kT); b.kV = (const double*)(base + o.kV);
    b.omT = (const int*)(base + o.omT); b.omV = (const double*)(base + o.omV);
    return b;
}

// Member workspace layout.  With cv.base == nullptr only measures.
// lean: a member solved by the kernel-per-phase path (rsg_big.hip), which
// keeps its AMG hierarchy outside the workspace: no device-AMG arrays.
inline void carve_member(Carver& cv, int nC, int nIF, int nB, int nnz0, int histCap, RsgMemberDev& M,
                         bool lean = false) {
    double** cellArrays[] = {&M.Ux, &M.Uy, &M.Uz, &M.p, &M.k, &M.om, &M.nuT, &M.F1, &M.F2, &M.Pk,
                             &M.CDkw, &M.Smag, &M.aPx, &M.aPc, &M.pp, &M.kOld, &M.omOld,
                             &M.diag, &M.src};
    for (double** a : cellArrays) *a = cv.take<double>(nC);
    for (int i = 0; i < 15; ++i) M.g[i] = cv.take<double>(nC);
    for (int i = 0; i < 3; ++i) M.gp[i] = cv.take<double>(nC);
    for (int i = 0; i < 8; ++i) M.kv[i] = cv.take<double>(nC);
    double** bArrays[] = {&M.Ubx, &M.Uby, &M.Ubz, &M.pb, &M.kb, &M.omb, &M.ppb};
    for (double** a : bArrays) *a = cv.take<double>(nB > 0 ? nB : 1);
    M.upper = cv.take<double>(nIF > 0 ? nIF : 1);
    M.lower = cv.take<double>(nIF > 0 ? nIF : 1);
    RsgAmgDev& A = M.amg;
    if (lean) {
        A = RsgAmgDev{};
        M.histCap = histCap;
        M.hist = histCap > 0 ? cv.take<double>(6 * (size_t)histCap) : nullptr;
        M.sc = cv.take<RsgScalars>(1);
        cv.off = (cv.off + 255) & ~size_t(255);
        return;
    }
    A.rowCap = 3 * nC + 64;
    A.nnzCap = 3 * nnz0 + 64;
    const int LV = RSG_MAX_LEVELS + 1;
    A.lvN = cv.take<int>(LV); A.lvRow = cv.take<int>(LV); A.lvRp = cv.take<int>(LV); A.lvNz = cv.take<int>(LV);
    A.rp = cv.take<int>(A.rowCap + LV);
    A.ci = cv.take<int>(A.nnzCap);
    A.v = cv.take<double>(A.nnzCap);
    A.dg = cv.take<double>(A.rowCap);
    A.map = cv.take<int>(A.rowCap);
    A.aggPtr = cv.take<int>(A.rowCap + LV);
    A.aggIdx = cv.take<int>(A.rowCap);
    A.gPtr = cv.take<int>(A.nnzCap + LV);
    A.gIdx = cv.take<int>(A.nnzCap);
    A.x = cv.take<double>(A.rowCap); A.b = cv.take<double>(A.rowCap); A.r = cv.take<double>(A.rowCap);
    A.cinv = cv.take<double>(RSG_COARSE_DENSE * RSG_COARSE_DENSE);
    A.scr = cv.take<int>(3 * (size_t)A.rowCap + 2 * (size_t)A.nnzCap);
    A.eCap = A.nnzCap; A.aCap = A.rowCap;
    A.lvK = cv.take<int>(LV); A.lvE = cv.take<int>(LV); A.lvAK = cv.take<int>(LV); A.lvAE = cv.take<int>(LV);
    A.used = cv.take<int>(2);
    A.eCol = cv.take<int>(A.eCap);
    A.eVal = cv.take<double>(A.eCap);
    A.aCol = cv.take<int>(A.aCap);
    M.histCap = histCap;
    M.hist = histCap > 0 ? cv.take<double>(6 * (size_t)histCap) : nullptr;
    M.sc = cv.take<RsgScalars>(1);
    cv.off = (cv.off + 255) & ~size_t(255);
}

inline size_t member_bytes(int nC, int nIF, int nB, int nnz0, int histCap) {
    Carver cv; RsgMemberDev M;
    carve_member(cv, nC, nIF, nB, nnz0, histCap, M);
    return cv.off;
}

inline RsgSettingsDev settings_dev(const rsg_settings* s) {
    RsgSettingsDev d;
    d.maxIter = s->max_iterations; d.tol = s->convergence_tol; d.divLimit = s->divergence_limit;
    d.aU = s->alpha_u; d.aP = s->alpha_p; d.aK = s->alpha_k; d.aOm = s->alpha_omega;
    d.innerIter = s->inner_iterations; d.innerTol = s->inner_tolerance;
    d.turbStart = s->turb_start_iter; d.turbInterval = s->turb_update_interval;
    d.nutFloorIters = s->nut_floor_iters; d.rcAll = s->rhie_chow_all_meshes;
    d.bf[0] = s->body_force[0]; d.bf[1] = s->body_force[1]; d.bf[2] = s->body_force[2];
    d.kMin = s->k_min; d.kMax = s->k_max; d.omMin = s->omega_min;
    d.wallFn = s->use_wall_functions; d.vonK = s->von_karman; d.variant = s->sst_variant;
    d.pSolver = s->pressure_solver; d.mSolver = s->momentum_solver; d.tSolver = s->turbulence_solver;
    d.amgRefresh = s->amg_refresh > 0 ? s->amg_refresh : 1;
    d.coarseRefresh = s->coarse_refresh > 0 ? s->coarse_refresh : 1;
    d.recordHistory = s->record_history;
    return d;
}

inline bool check_settings(const rsg_settings* s, std::string& err) {
    if (s->max_iterations < 0 || s->inner_iterations < 0) { err = "settings: negative iteration count"; return false; }
    if (s->turb_update_interval <= 0) { err = "settings: turb_update_interval must be >= 1"; return false; }
    if (!(s->alpha_u > 0) || !(s->alpha_k > 0) || !(s->alpha_omega > 0)) { err = "settings: relaxation factors must be > 0"; return false; }
    if (s->momentum_solver == RSG_SOLVER_AMG || s->turbulence_solver == RSG_SOLVER_AMG) {
        err = "settings: AMG is supported for the pressure solver only"; return false;
    }
    auto ok = [](int c) {
        return c == RSG_SOLVER_AMG || c == RSG_SOLVER_PCG || c == RSG_SOLVER_BICGSTAB || c == RSG_SOLVER_GAUSS_SEIDEL;
    };
    if (!ok(s->pressure_solver) || !ok(s->momentum_solver) || !ok(s->turbulence_solver)) {
        err = "settings: unknown linear solver code"; return false;
    }
    return true;
}


// Whether a member's Uz stays exactly zero, so that skipping the z-momentum
// assembly and solve leaves every result bit-identical: a planar mesh (zero z
// normals, hence zero z area vectors and z gradients), no z body force, zero
// z velocity on every boundary face and in the initial state.  Under these
// the reference's z system has source and residual exactly 0, its solve
// returns at once with initialRes 0, and Uz, Ubz stay 0.
inline bool bc_zero_z(const rsg_bc_desc* bc, int nB) {
    for (int b = 0; b < nB; ++b)
        if (bc->vel_value[3 * b + 2] != 0.0) return false;
    return true;
}

inline bool planar_z(const PrepMesh& P, bool bcZeroZ, const rsg_settings* s,
                     double initUz, const double* warmU, const double* warmUb) {
    if (!P.planar || !bcZeroZ || s->body_force[2] != 0.0 || initUz != 0.0) return false;
    if (warmU)
        for (int c = 0; c < P.nC; ++c) if (warmU[3 * c + 2] != 0.0) return false;
    if (warmUb)
        for (int b = 0; b < P.nB; ++b) if (warmUb[3 * b + 2] != 0.0) return false;
    return true;
}

}  // namespace rsg
