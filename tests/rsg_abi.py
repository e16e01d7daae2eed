"""ctypes mirror of include/rsg.h for the tests (struct layouts and helpers).

Test infrastructure: builds rsg_mesh_desc / rsg_bc_desc / rsg_settings values
from numpy arrays and MSH1 files so the same inputs can be fed to the C-ABI
library, the CPU emulation build and the oracle.
"""
import ctypes as C
import struct

import numpy as np

PD = C.POINTER(C.c_double)
PI = C.POINTER(C.c_int32)

PATCH = {"wall": 0, "inlet": 1, "outlet": 2, "symmetry": 3}
BC = dict(DIRICHLET=0, NEUMANN=1, INLET_VELOCITY=2, OUTLET_PRESSURE=3, WALL_NOSLIP=4,
          WALL_KOMEGA=5, WALL_MOVING=6, SYMMETRY=7, CYCLIC=8, KEEP=9)
DEFAULT_COEFFS = [0.85, 0.5, 0.075, 5.0 / 9.0, 1.0, 0.856, 0.0828, 0.44, 0.09, 0.31, 0.41]


class MeshDesc(C.Structure):
    _fields_ = [("n_cells", C.c_int32), ("n_faces", C.c_int32), ("n_internal_faces", C.c_int32),
                ("n_patches", C.c_int32), ("owner", PI), ("neighbor", PI), ("face_normal", PD),
                ("face_area", PD), ("face_delta", PD), ("face_d", PD), ("face_weight", PD),
                ("cell_volume", PD), ("wall_distance", PD), ("patch_type", PI),
                ("patch_face_ptr", PI), ("patch_faces", PI)]


class BcDesc(C.Structure):
    _fields_ = [("vel_type", PI), ("vel_value", PD), ("p_type", PI), ("p_value", PD),
                ("k_type", PI), ("k_value", PD), ("om_type", PI), ("om_value", PD)]


class Settings(C.Structure):
    _fields_ = [("max_iterations", C.c_int32), ("convergence_tol", C.c_double),
                ("divergence_limit", C.c_double), ("alpha_u", C.c_double), ("alpha_p", C.c_double),
                ("alpha_k", C.c_double), ("alpha_omega", C.c_double), ("inner_iterations", C.c_int32),
                ("inner_tolerance", C.c_double), ("turb_start_iter", C.c_int32),
                ("turb_update_interval", C.c_int32), ("nut_floor_iters", C.c_int32),
                ("rhie_chow_all_meshes", C.c_int32), ("body_force", C.c_double * 3),
                ("k_min", C.c_double), ("k_max", C.c_double), ("omega_min", C.c_double),
                ("use_wall_functions", C.c_int32), ("von_karman", C.c_double),
                ("sst_variant", C.c_int32), ("pressure_solver", C.c_int32),
                ("momentum_solver", C.c_int32), ("turbulence_solver", C.c_int32),
                ("amg_refresh", C.c_int32), ("coarse_refresh", C.c_int32),
                ("record_history", C.c_int32)]


class Init(C.Structure):
    _fields_ = [("kind", C.c_int32), ("U", C.c_double * 3), ("p", C.c_double), ("k", C.c_double),
                ("omega", C.c_double), ("U_cells", PD), ("p_cells", PD), ("k_cells", PD),
                ("om_cells", PD), ("nuT", PD), ("F1", PD), ("F2", PD), ("Pk", PD), ("CDkw", PD),
                ("U_bnd", PD), ("p_bnd", PD), ("k_bnd", PD), ("om_bnd", PD),
                ("turb_established", C.c_int32)]


class Result(C.Structure):
    _fields_ = [("status", C.c_int32), ("final_iter", C.c_int32), ("converged", C.c_int32),
                ("diverged", C.c_int32), ("turb_established", C.c_int32),
                ("U", PD), ("p", PD), ("k", PD), ("omega", PD), ("nuT", PD), ("F1", PD),
                ("F2", PD), ("Pk", PD), ("CDkw", PD), ("U_bnd", PD), ("p_bnd", PD),
                ("k_bnd", PD), ("om_bnd", PD), ("history", PD), ("history_cap", C.c_int32),
                ("history_len", C.c_int32), ("inner_iters", C.c_double * 6),
                ("t_submit", C.c_double), ("t_start", C.c_double), ("t_end", C.c_double),
                ("predictions", PD), ("predictions_cap", C.c_int32), ("n_predictions", C.c_int32),
                ("log_lik", C.c_double)]


class CompDesc(C.Structure):
    _fields_ = [("t_type", PI), ("t_value", PD), ("gamma", C.c_double), ("R", C.c_double), ("Pr", C.c_double),
                ("Pr_T", C.c_double), ("mu_ref", C.c_double), ("T_ref", C.c_double), ("S_suth", C.c_double),
                ("alpha_t", C.c_double)]


class CompInit(C.Structure):
    _fields_ = [("U", C.c_double * 3), ("p", C.c_double), ("T", C.c_double), ("k", C.c_double),
                ("omega", C.c_double)]


TBC = dict(DIRICHLET=0, ZERO_GRADIENT=1, KEEP=2)
EOS_AIR = (1.4, 287.0, 0.72, 0.9, 1.716e-5, 273.15, 110.4)


class Observation(C.Structure):
    _fields_ = [("type", C.c_int32), ("index", C.c_int32), ("component", C.c_int32),
                ("face_begin", C.c_int32), ("face_end", C.c_int32), ("ref_velocity", C.c_double),
                ("ref_area", C.c_double), ("ref_density", C.c_double), ("ref_pressure", C.c_double),
                ("observed", C.c_double), ("sigma", C.c_double), ("sigma_model", C.c_double)]


def dptr(a):
    return a.ctypes.data_as(PD)


def iptr(a):
    return a.ctypes.data_as(PI)


# ORS_* packed settings layout of the oracle / reference driver
def packed_settings(s):
    return [s.max_iterations, s.convergence_tol, s.divergence_limit, s.alpha_u, s.alpha_p,
            s.alpha_k, s.alpha_omega, s.inner_iterations, s.inner_tolerance, s.turb_start_iter,
            s.turb_update_interval, s.nut_floor_iters, s.rhie_chow_all_meshes,
            s.body_force[0], s.body_force[1], s.body_force[2], s.k_min, s.k_max, s.omega_min,
            s.use_wall_functions, s.von_karman, s.sst_variant, s.pressure_solver,
            s.momentum_solver, s.turbulence_solver]


def default_settings(**kw):
    s = Settings()
    s.max_iterations = 500; s.convergence_tol = 1e-5; s.divergence_limit = 1e6
    s.alpha_u = 0.7; s.alpha_p = 0.3; s.alpha_k = 0.5; s.alpha_omega = 0.5
    s.inner_iterations = 200; s.inner_tolerance = 1e-3
    s.turb_start_iter = 5; s.turb_update_interval = 1; s.nut_floor_iters = 500
    s.rhie_chow_all_meshes = 0
    s.k_min = 1e-10; s.k_max = 1e10; s.omega_min = 1e-6
    s.use_wall_functions = 0; s.von_karman = 0.41; s.sst_variant = 0
    s.pressure_solver = 0; s.momentum_solver = 2; s.turbulence_solver = 2
    s.amg_refresh = 1; s.coarse_refresh = 1; s.record_history = 0
    for k, v in kw.items():
        if k == "body_force":
            for i in range(3):
                s.body_force[i] = v[i]
        else:
            setattr(s, k, v)
    return s


class Msh:
    """MSH1 file (Mesh::saveBinary layout, core/src/Mesh.cpp:602-648)."""

    def __init__(self, path):
        b = open(path, "rb").read()
        off = 0

        def take(fmt):
            nonlocal off
            v = struct.unpack_from("<" + fmt, b, off)
            off += struct.calcsize("<" + fmt)
            return v
        assert take("I")[0] == 0x4D534831
        nc, nf, nn, npch, nif = take("5i")
        off += 24 * nn
        fr = np.frombuffer(b, dtype=np.dtype([("i", "<i4", 3), ("d", "<f8", 12)]), count=nf, offset=off)
        off += fr.nbytes
        self.nC, self.nF, self.nIF, self.nP = nc, nf, nif, npch
        self.owner = np.ascontiguousarray(fr["i"][:, 0]).astype(np.int32)
        self.neighbor = np.ascontiguousarray(fr["i"][:nif, 1]).astype(np.int32)
        d = fr["d"]
        self.fc = np.ascontiguousarray(d[:, 0:3]); self.fn = np.ascontiguousarray(d[:, 3:6])
        self.area = np.ascontiguousarray(d[:, 6]); self.delta = np.ascontiguousarray(d[:, 7])
        self.fd = np.ascontiguousarray(d[:, 8:11]); self.weight = np.ascontiguousarray(d[:, 11])
        cc = np.zeros((nc, 3)); vol = np.zeros(nc)
        for c in range(nc):
            x = take("4d"); cc[c] = x[:3]; vol[c] = x[3]
            k = take("i")[0]; off += 4 * k
        self.cc, self.vol = cc, vol
        self.ptypes, self.pnames, pf = [], [], []
        for p in range(npch):
            ln = take("i")[0]; self.pnames.append(b[off:off + ln].decode()); off += ln
            ln = take("i")[0]; self.ptypes.append(b[off:off + ln].decode()); off += ln
            k = take("i")[0]; pf.append(np.frombuffer(b, "<i4", k, off).copy()); off += 4 * k
        self.pfaces = pf
        self.ptr = np.zeros(npch + 1, np.int32)
        for p in range(npch):
            self.ptr[p + 1] = self.ptr[p] + len(pf[p])
        self.pf = np.concatenate(pf).astype(np.int32) if pf else np.zeros(0, np.int32)
        self.ptype_codes = np.array([PATCH.get(t, 4) for t in self.ptypes], np.int32)
        self.wd = self.wall_distance()

    def wall_distance(self):
        # Mesh::computeWallDistance (core/src/Mesh.cpp:345-373): min over wall
        # face centres of the Euclidean distance, evaluated elementwise in the
        # reference's operation order
        wp = [self.fc[self.pfaces[p]] for p in range(self.nP) if self.ptypes[p] == "wall"]
        if not wp:
            return np.full(self.nC, 1e10)
        wp = np.concatenate(wp)
        out = np.empty(self.nC)
        for c in range(self.nC):
            d = self.cc[c][None, :] - wp
            out[c] = np.min(np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]))
        return out

    def desc(self):
        keep = [np.ascontiguousarray(a) for a in (
            self.owner, self.neighbor, self.fn.ravel(), self.area, self.delta, self.fd.ravel(),
            self.weight, self.vol, self.wd, self.ptype_codes, self.ptr, self.pf)]
        d = MeshDesc(self.nC, self.nF, self.nIF, self.nP, iptr(keep[0]), iptr(keep[1]),
                     dptr(keep[2]), dptr(keep[3]), dptr(keep[4]), dptr(keep[5]), dptr(keep[6]),
                     dptr(keep[7]), dptr(keep[8]), iptr(keep[9]), iptr(keep[10]), iptr(keep[11]))
        d._keep = keep
        return d

    def bc(self, kind=0, vals=(1.0, 1e-4, 1.0)):
        """Per-face BC arrays of channelDefaults (0) / couetteDefaults (1) /
        flatPlateDefaults (2), BoundaryCondition.hpp:44-152."""
        nB = self.nF - self.nIF
        vt = np.full(nB, BC["NEUMANN"], np.int32); vv = np.zeros(3 * nB)
        pt = np.full(nB, BC["NEUMANN"], np.int32); pv = np.zeros(nB)
        kt = np.full(nB, BC["NEUMANN"], np.int32); kv = np.zeros(nB)
        ot = np.full(nB, BC["NEUMANN"], np.int32); ov = np.zeros(nB)
        Uin = 0.5 * vals[0] if kind == 1 else vals[0]
        for p in range(self.nP):
            b = self.pfaces[p] - self.nIF
            t = self.ptypes[p]
            if t == "wall":
                vt[b] = BC["WALL_NOSLIP"]; kt[b] = BC["WALL_KOMEGA"]; ot[b] = BC["WALL_KOMEGA"]
            elif t == "inlet":
                vt[b] = BC["INLET_VELOCITY"]
                for i in b:
                    vv[3 * i:3 * i + 3] = (Uin, 0, 0)
                kt[b] = BC["DIRICHLET"]; kv[b] = vals[1]; ot[b] = BC["DIRICHLET"]; ov[b] = vals[2]
            elif t == "outlet":
                pt[b] = BC["OUTLET_PRESSURE"]
            elif t == "symmetry":
                vt[b] = BC["SYMMETRY"]
            name = self.pnames[p]
            if kind == 1 and name == "top":
                vt[b] = BC["WALL_MOVING"]
                for i in b:
                    vv[3 * i:3 * i + 3] = (vals[0], 0, 0)
            if kind == 1 and t == "inlet":
                vt[b] = BC["NEUMANN"]; pt[b] = BC["NEUMANN"]; kt[b] = BC["NEUMANN"]; ot[b] = BC["NEUMANN"]
            if kind == 2 and name == "top":
                vt[b] = BC["DIRICHLET"]
                for i in b:
                    vv[3 * i:3 * i + 3] = (vals[0], 0, 0)
                pt[b] = BC["NEUMANN"]; kt[b] = BC["DIRICHLET"]; kv[b] = vals[1]
                ot[b] = BC["DIRICHLET"]; ov[b] = vals[2]
        keep = [vt, vv, pt, pv, kt, kv, ot, ov]
        d = BcDesc(iptr(vt), dptr(vv), iptr(pt), dptr(pv), iptr(kt), dptr(kv), iptr(ot), dptr(ov))
        d._keep = keep
        return d

    def comp_bc(self, Uin, T_in, p_out, kIn, omIn, eos=EOS_AIR, alpha_t=0.7):
        """CompressibleBoundaryConditions::channelDefaults (CompressibleBCs.hpp:27-72)
        as per-face arrays: (rsg_bc_desc, rsg_comp_desc)."""
        nB = self.nF - self.nIF
        vt = np.full(nB, BC["NEUMANN"], np.int32); vv = np.zeros(3 * nB)
        pt = np.full(nB, BC["NEUMANN"], np.int32); pv = np.zeros(nB)
        kt = np.full(nB, BC["NEUMANN"], np.int32); kv = np.zeros(nB)
        ot = np.full(nB, BC["NEUMANN"], np.int32); ov = np.zeros(nB)
        tt = np.full(nB, TBC["KEEP"], np.int32); tv = np.zeros(nB)
        for p in range(self.nP):
            b = self.pfaces[p] - self.nIF
            t = self.ptypes[p]
            tt[b] = TBC["ZERO_GRADIENT"]; tv[b] = 300.0
            if t == "wall":
                vt[b] = BC["WALL_NOSLIP"]; kt[b] = BC["WALL_KOMEGA"]; ot[b] = BC["WALL_KOMEGA"]
            elif t == "inlet":
                vt[b] = BC["INLET_VELOCITY"]
                for i in b:
                    vv[3 * i:3 * i + 3] = (Uin, 0, 0)
                tt[b] = TBC["DIRICHLET"]; tv[b] = T_in
                kt[b] = BC["DIRICHLET"]; kv[b] = kIn; ot[b] = BC["DIRICHLET"]; ov[b] = omIn
            elif t == "outlet":
                pt[b] = BC["OUTLET_PRESSURE"]; pv[b] = p_out
        d = BcDesc(iptr(vt), dptr(vv), iptr(pt), dptr(pv), iptr(kt), dptr(kv), iptr(ot), dptr(ov))
        d._keep = [vt, vv, pt, pv, kt, kv, ot, ov]
        c = CompDesc(iptr(tt), dptr(tv), *eos, alpha_t)
        c._keep = [tt, tv]
        return d, c


def new_result(n, nB, hist_cap=0):
    arrs = {k: np.zeros(n * (3 if k == "U" else 1)) for k in
            ("U", "p", "k", "omega", "nuT", "F1", "F2", "Pk", "CDkw")}
    arrs.update({k: np.zeros(nB * (3 if k == "U_bnd" else 1)) for k in ("U_bnd", "p_bnd", "k_bnd", "om_bnd")})
    arrs["history"] = np.zeros(6 * max(hist_cap, 1))
    r = Result()
    for k, a in arrs.items():
        setattr(r, k, dptr(a))
    r.history_cap = hist_cap
    return r, arrs


OBS_ARRAYS = ("type", "patch", "loc", "comp", "refArea", "refVel", "refP", "refRho", "observed", "sigma", "sigmaModel")


def observation_set(msh, rng):
    """Observables of every type on a mesh (ObservationOperator.hpp:20-32): on
    every wall patch a drag, a separation point, a reattachment length and
    skin friction at three points; pressure taps and velocity probes (all
    three components) at random points in the bounding box; measurement
    values and sigmas random, one sigma pair zero (the 1e-30 variance floor)."""
    lo, hi = msh.cc.min(axis=0), msh.cc.max(axis=0)
    rows = []
    walls = [q for q in range(msh.nP) if msh.ptypes[q] == "wall"]
    for q in walls:
        rows.append((0, q, (0.0, 0.0, 0.0), 0, 0.7, 1.3, 0.0, 1.0))
        rows.append((4, q, (0.0, 0.0, 0.0), 0, 1.0, 1.0, 0.0, 1.0))
        rows.append((5, q, (0.0, 0.0, 0.0), 0, 1.0, 1.0, 0.0, 1.0))
        for _ in range(3):
            rows.append((3, q, tuple(rng.uniform(lo, hi)), 0, 1.0, 0.9, 0.0, 1.0))
    for _ in range(3):
        rows.append((1, -1, tuple(rng.uniform(lo, hi)), 0, 1.0, 1.1, 0.2, 1.2))
    for comp in (0, 1, 2):
        rows.append((2, -1, tuple(rng.uniform(lo, hi)), comp, 1.0, 1.0, 0.0, 1.0))
    n = len(rows)
    o = {"type": np.array([r[0] for r in rows], np.int32), "patch": np.array([r[1] for r in rows], np.int32),
         "loc": np.array([r[2] for r in rows], float).ravel(), "comp": np.array([r[3] for r in rows], np.int32),
         "refArea": np.array([r[4] for r in rows]), "refVel": np.array([r[5] for r in rows]),
         "refP": np.array([r[6] for r in rows]), "refRho": np.array([r[7] for r in rows]),
         "observed": rng.normal(size=n), "sigma": rng.uniform(0.05, 0.5, n), "sigmaModel": rng.uniform(0.0, 0.2, n)}
    o["sigma"][-1] = 0.0; o["sigmaModel"][-1] = 0.0
    return o


def wavy_state(msh, rng):
    """U, U boundary values and p whose wall shear changes sign along x (so
    that separation and reattachment crossings exist)."""
    L = msh.cc[:, 0].max() - msh.cc[:, 0].min() + 1e-12
    U = np.zeros((msh.nC, 3))
    U[:, 0] = np.sin(6.0 * np.pi * msh.cc[:, 0] / L) + 0.1 * rng.normal(size=msh.nC)
    U[:, 1] = 0.1 * rng.normal(size=msh.nC)
    nB = msh.nF - msh.nIF
    Ub = 0.01 * rng.normal(size=(nB, 3))
    p = rng.normal(size=msh.nC)
    return U.ravel(), Ub.ravel(), p


def resolve_observations(msh, o):
    """rsg_observation records for observation_set(): the reference's nearest
    cell / nearest patch face searches (first minimum of the squared distance,
    ObservationOperator.hpp:160-183) and the patch face lists."""
    n = len(o["type"])
    recs = (Observation * n)()
    faces, fx = [], []
    loc = o["loc"].reshape(-1, 3)
    for i in range(n):
        r = recs[i]
        r.type = int(o["type"][i]); r.component = int(o["comp"][i]); r.index = -1
        r.ref_velocity = o["refVel"][i]; r.ref_area = o["refArea"][i]; r.ref_density = o["refRho"][i]
        r.ref_pressure = o["refP"][i]; r.observed = o["observed"][i]; r.sigma = o["sigma"][i]
        r.sigma_model = o["sigmaModel"][i]
        q = int(o["patch"][i])
        if r.type in (1, 2):
            d = msh.cc - loc[i]
            r.index = int(np.argmin(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]))
        elif r.type == 3:
            pf = msh.pfaces[q]
            if len(pf):
                d = msh.fc[pf] - loc[i]
                r.index = int(pf[np.argmin(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2])])
        else:
            r.face_begin = len(faces)
            faces += [int(f) for f in msh.pfaces[q]]
            fx += [float(msh.fc[f][0]) for f in msh.pfaces[q]]
            r.face_end = len(faces)
    return recs, np.array(faces, np.int32), np.array(fx)
