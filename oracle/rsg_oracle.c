/* rsg_oracle.c -- CPU restatement of the reference forward model.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Loaded solely by tests/, the smoke
 * check and bench.py's cpu_baseline leg; never by the product path.
 *
 * Each routine restates the reference algorithm in the reference's
 * floating-point operation order (citations are to files of the reference
 * repository).  Built with -O2 -ffp-contract=off it reproduces the reference
 * solver bit-for-bit; tests/test_oracle_vs_ref.py pins that against the
 * reference compiled from its own sources (oracle/_ref) and against the
 * closed-form / known-answer values of the reference's own C++ tests.
 */
#include "rsg_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

/* std::max / std::min semantics (first argument wins ties / NaN handling) */
static inline double smax(double a, double b) { return (a < b) ? b : a; }
static inline double smin(double a, double b) { return (b < a) ? b : a; }

typedef struct { double x, y, z; } v3;
static inline v3 v3make(double x, double y, double z) { v3 r = {x, y, z}; return r; }
static inline v3 v3add(v3 a, v3 b) { return v3make(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 v3sub(v3 a, v3 b) { return v3make(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 v3mul(v3 a, double s) { return v3make(a.x * s, a.y * s, a.z * s); }
static inline v3 v3div(v3 a, double s) { return v3make(a.x / s, a.y / s, a.z / s); }
static inline double v3dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline double v3norm(v3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }

/* ------------------------------------------------------------------ mesh */
/* Mesh.hpp:44-197; MSH1 reader Mesh.cpp:651-713; wall distance :345-373 */
struct om_mesh {
    int nC, nF, nIF, nP, nN;
    int *own, *nbr, *patchID;
    v3 *fc, *fn, *fd;
    double *area, *delta, *weight;
    v3 *cc;
    double *vol, *wd;
    int *cfPtr, *cfIdx;
    int *ptype, *pfPtr, *pfIdx;
};

static int rd(FILE* f, void* p, size_t n) { return fread(p, 1, n, f) == n ? 0 : -1; }

static int ptype_code(const char* s) {
    if (!strcmp(s, "wall")) return OPT_WALL;
    if (!strcmp(s, "inlet")) return OPT_INLET;
    if (!strcmp(s, "outlet")) return OPT_OUTLET;
    if (!strcmp(s, "symmetry")) return OPT_SYMMETRY;
    return OPT_OTHER;
}

void om_mesh_free(om_mesh* m) {
    if (!m) return;
    free(m->own); free(m->nbr); free(m->patchID); free(m->fc); free(m->fn); free(m->fd);
    free(m->area); free(m->delta); free(m->weight); free(m->cc); free(m->vol); free(m->wd);
    free(m->cfPtr); free(m->cfIdx); free(m->ptype); free(m->pfPtr); free(m->pfIdx);
    free(m);
}

static void wall_distance(om_mesh* m) {
    /* Mesh.cpp:345-373: min over wall-face centres of |cc - wp| */
    int nw = 0;
    for (int p = 0; p < m->nP; ++p)
        if (m->ptype[p] == OPT_WALL) nw += m->pfPtr[p + 1] - m->pfPtr[p];
    if (nw == 0) { for (int c = 0; c < m->nC; ++c) m->wd[c] = 1e10; return; }
    v3* wp = (v3*)malloc(sizeof(v3) * nw);
    int k = 0;
    for (int p = 0; p < m->nP; ++p)
        if (m->ptype[p] == OPT_WALL)
            for (int j = m->pfPtr[p]; j < m->pfPtr[p + 1]; ++j) wp[k++] = m->fc[m->pfIdx[j]];
    for (int c = 0; c < m->nC; ++c) {
        double minD = DBL_MAX;
        for (int j = 0; j < nw; ++j) {
            double d = v3norm(v3sub(m->cc[c], wp[j]));
            if (d < minD) minD = d;
        }
        m->wd[c] = minD;
    }
    free(wp);
}

om_mesh* om_mesh_load(const char* path) {
    FILE* f = fopen(path, "rb");
    if (!f) return NULL;
    uint32_t magic = 0;
    int hdr[5];
    om_mesh* m = NULL;
    if (rd(f, &magic, 4) || magic != 0x4D534831u || rd(f, hdr, sizeof hdr)) goto fail;
    m = (om_mesh*)calloc(1, sizeof(om_mesh));
    m->nC = hdr[0]; m->nF = hdr[1]; m->nN = hdr[2]; m->nP = hdr[3]; m->nIF = hdr[4];
    {
        /* nodes are not used by the solver: skip */
        if (fseek(f, (long)m->nN * 24, SEEK_CUR)) goto fail;
        int nF = m->nF;
        m->own = malloc(sizeof(int) * nF); m->nbr = malloc(sizeof(int) * nF);
        m->patchID = malloc(sizeof(int) * nF);
        m->fc = malloc(sizeof(v3) * nF); m->fn = malloc(sizeof(v3) * nF); m->fd = malloc(sizeof(v3) * nF);
        m->area = malloc(sizeof(double) * nF); m->delta = malloc(sizeof(double) * nF);
        m->weight = malloc(sizeof(double) * nF);
        for (int i = 0; i < nF; ++i) {
            int ii[3]; double dd[12];
            if (rd(f, ii, sizeof ii) || rd(f, dd, sizeof dd)) goto fail;
            m->own[i] = ii[0]; m->nbr[i] = ii[1]; m->patchID[i] = ii[2];
            m->fc[i] = v3make(dd[0], dd[1], dd[2]);
            m->fn[i] = v3make(dd[3], dd[4], dd[5]);
            m->area[i] = dd[6]; m->delta[i] = dd[7];
            m->fd[i] = v3make(dd[8], dd[9], dd[10]);
            m->weight[i] = dd[11];
        }
        int nC = m->nC;
        m->cc = malloc(sizeof(v3) * nC); m->vol = malloc(sizeof(double) * nC);
        m->wd = malloc(sizeof(double) * nC);
        m->cfPtr = malloc(sizeof(int) * (nC + 1));
        int cap = 8 * nC + 8, used = 0;
        m->cfIdx = malloc(sizeof(int) * cap);
        m->cfPtr[0] = 0;
        for (int c = 0; c < nC; ++c) {
            double dd[4]; int nf;
            if (rd(f, dd, sizeof dd) || rd(f, &nf, 4)) goto fail;
            m->cc[c] = v3make(dd[0], dd[1], dd[2]); m->vol[c] = dd[3];
            if (used + nf > cap) { cap = 2 * (used + nf); m->cfIdx = realloc(m->cfIdx, sizeof(int) * cap); }
            if (rd(f, m->cfIdx + used, sizeof(int) * nf)) goto fail;
            used += nf; m->cfPtr[c + 1] = used;
        }
        m->ptype = malloc(sizeof(int) * (m->nP + 1));
        m->pfPtr = malloc(sizeof(int) * (m->nP + 1));
        int pcap = (m->nF - m->nIF) + 8, pused = 0;
        m->pfIdx = malloc(sizeof(int) * pcap);
        m->pfPtr[0] = 0;
        for (int p = 0; p < m->nP; ++p) {
            int len; char buf[256];
            if (rd(f, &len, 4) || len < 0 || len > 255 || rd(f, buf, len)) goto fail;
            if (rd(f, &len, 4) || len < 0 || len > 255 || rd(f, buf, len)) goto fail;
            buf[len] = 0; m->ptype[p] = ptype_code(buf);
            int nf;
            if (rd(f, &nf, 4)) goto fail;
            if (pused + nf > pcap) { pcap = 2 * (pused + nf); m->pfIdx = realloc(m->pfIdx, sizeof(int) * pcap); }
            if (rd(f, m->pfIdx + pused, sizeof(int) * nf)) goto fail;
            pused += nf; m->pfPtr[p + 1] = pused;
        }
    }
    fclose(f);
    wall_distance(m);
    return m;
fail:
    fclose(f);
    om_mesh_free(m);
    return NULL;
}

void om_mesh_info(const om_mesh* m, int* out) {
    out[0] = m->nC; out[1] = m->nF; out[2] = m->nIF; out[3] = m->nP; out[4] = m->nN;
}
int om_wall_distance(const om_mesh* m, double* wd) {
    memcpy(wd, m->wd, sizeof(double) * m->nC); return 0;
}
int om_patch_type(const om_mesh* m, int p) { return m->ptype[p]; }

/* ------------------------------------------------------------ operators */
/* greenGaussGrad, Field.hpp:122-152 */
static void gg(const om_mesh* m, const double* phi, const double* phib, v3* g) {
    for (int c = 0; c < m->nC; ++c) g[c] = v3make(0, 0, 0);
    for (int f = 0; f < m->nIF; ++f) {
        int o = m->own[f], n = m->nbr[f];
        double w = m->weight[f];
        double phiF = phi[o] * w + phi[n] * (1.0 - w);
        v3 Sf = v3mul(m->fn[f], m->area[f]);
        g[o] = v3add(g[o], v3mul(Sf, phiF));
        g[n] = v3sub(g[n], v3mul(Sf, phiF));
    }
    for (int f = m->nIF; f < m->nF; ++f) {
        v3 Sf = v3mul(m->fn[f], m->area[f]);
        int o = m->own[f];
        g[o] = v3add(g[o], v3mul(Sf, phib[f - m->nIF]));
    }
    for (int c = 0; c < m->nC; ++c) {
        double V = m->vol[c];
        if (V > 1e-30) g[c] = v3div(g[c], V);
    }
}

int om_green_gauss(const om_mesh* m, const double* phi, const double* phib, double* grad) {
    gg(m, phi, phib, (v3*)grad); return 0;
}

/* computeVelocityGradients Field.hpp:232-246 -- vg[0..2] = grad Ux, Uy, Uz */
typedef struct { v3 *dudx, *dvdx, *dwdx; } velgrad;

static void vel_grads(const om_mesh* m, const v3* U, const v3* Ub, velgrad* vg,
                      double* scratchC, double* scratchB) {
    int nB = m->nF - m->nIF;
    for (int comp = 0; comp < 3; ++comp) {
        for (int c = 0; c < m->nC; ++c)
            scratchC[c] = comp == 0 ? U[c].x : comp == 1 ? U[c].y : U[c].z;
        for (int b = 0; b < nB; ++b)
            scratchB[b] = comp == 0 ? Ub[b].x : comp == 1 ? Ub[b].y : Ub[b].z;
        gg(m, scratchC, scratchB, comp == 0 ? vg->dudx : comp == 1 ? vg->dvdx : vg->dwdx);
    }
}

/* strainRateMagnitude Field.hpp:254-267 */
static void strain_mag(const om_mesh* m, const velgrad* vg, double* S) {
    for (int c = 0; c < m->nC; ++c) {
        double S11 = vg->dudx[c].x, S22 = vg->dvdx[c].y, S33 = vg->dwdx[c].z;
        double S12 = 0.5 * (vg->dudx[c].y + vg->dvdx[c].x);
        double S13 = 0.5 * (vg->dudx[c].z + vg->dwdx[c].x);
        double S23 = 0.5 * (vg->dvdx[c].z + vg->dwdx[c].y);
        S[c] = sqrt(2.0 * (S11 * S11 + S22 * S22 + S33 * S33
                           + 2.0 * (S12 * S12 + S13 * S13 + S23 * S23)));
    }
}

/* ------------------------------------------------------------------ SST */
/* SSTCoefficients SSTModel.hpp:9-40; index order InferenceParameters.hpp:12-14 */
typedef struct {
    double sk1, sw1, b1, a1_, sk2, sw2, b2, a2_, bs, a1, kap;
    int variant;
} sstc;

static sstc mk_sst(const double* c, int variant) {
    sstc s = {c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7], c[8], c[9], c[10], variant};
    return s;
}
static inline double blend(double p1, double p2, double F1) { return F1 * p1 + (1.0 - F1) * p2; }

/* SSTModel.cpp:9-26 */
static double sst_F1(const sstc* s, double k, double om, double y, double nu, double CDpos_in) {
    double sqrtK = sqrt(smax(k, 0.0));
    double omSafe = smax(om, 1e-20);
    double ySafe = smax(y, 1e-20);
    double y2s = ySafe * ySafe;
    double term1 = sqrtK / (s->bs * omSafe * ySafe);
    double term2 = 500.0 * nu / (y2s * omSafe);
    double CDpos = smax(CDpos_in, 1e-20);
    double term3 = 4.0 * s->sw2 * smax(k, 0.0) / (CDpos * y2s);
    double arg1 = smin(smax(term1, term2), term3);
    double a4 = arg1 * arg1 * arg1 * arg1;
    return tanh(a4);
}
/* SSTModel.cpp:30-44 */
static double sst_F2(const sstc* s, double k, double om, double y, double nu) {
    double sqrtK = sqrt(smax(k, 0.0));
    double omSafe = smax(om, 1e-20);
    double ySafe = smax(y, 1e-20);
    double y2 = ySafe * ySafe;
    double term1 = 2.0 * sqrtK / (s->bs * omSafe * ySafe);
    double term2 = 500.0 * nu / (y2 * omSafe);
    double arg2 = smax(term1, term2);
    return tanh(arg2 * arg2);
}
/* SSTModel.cpp:57-65 */
static double sst_nut(const sstc* s, double k, double om, double S, double F2) {
    double kSafe = smax(k, 0.0);
    if (s->variant == 1) return kSafe / smax(om, 1e-20);
    double denom = smax(s->a1 * om, S * F2);
    denom = smax(denom, 1e-20);
    return s->a1 * kSafe / denom;
}
/* SSTModel.cpp:70-74 */
static double sst_prod(const sstc* s, double nuT, double S, double k, double om) {
    double raw = nuT * S * S;
    double lim = 10.0 * s->bs * smax(k, 0.0) * smax(om, 1e-20);
    return smin(raw, lim);
}
/* SSTModel.cpp:91-96 */
static double sst_prod_om(const sstc* s, double nuT, double S, double k, double om) {
    double S2 = S * S;
    double lim = 10.0 * s->bs * smax(k, 0.0) * smax(om, 1e-20) / smax(nuT, 1e-30);
    return smin(S2, lim);
}

double om_sst_point(int which, const double* coeffs, int variant, const double* a) {
    sstc s = mk_sst(coeffs, variant);
    switch (which) {
        case 0: return sst_F1(&s, a[0], a[1], a[2], a[3], a[4]);
        case 1: return sst_F2(&s, a[0], a[1], a[2], a[3]);
        case 2: return sst_nut(&s, a[0], a[1], a[2], a[3]);
        case 3: return sst_prod(&s, a[0], a[1], a[2], a[3]);
        case 4: return sst_prod_om(&s, a[0], a[1], a[2], a[3]);
        default: return 0.0;
    }
}

/* ------------------------------------------------------------ BCs/fields */
/* BoundaryCondition.hpp:20-32 */
typedef struct { int type; double value; v3 vec; } pbc;
typedef struct { pbc *vel, *prs, *k, *om; } flowbc;

/* channelDefaults :44-98, couetteDefaults :134-152, flatPlateDefaults :108-122.
 * The "top" patch is identified by name in the reference; MSH1 meshes from the
 * factories keep the patch order bottom, top, inlet, outlet, so the oracle
 * takes it as patch index 1 when kind != 0 (documented restriction). */
static flowbc mk_bc(const om_mesh* m, int kind, const double* v) {
    flowbc bc;
    int np = m->nP;
    bc.vel = calloc(np + 1, sizeof(pbc)); bc.prs = calloc(np + 1, sizeof(pbc));
    bc.k = calloc(np + 1, sizeof(pbc)); bc.om = calloc(np + 1, sizeof(pbc));
    double Uin = kind == 1 ? 0.5 * v[0] : v[0], kIn = v[1], omIn = v[2];
    for (int p = 0; p < np; ++p) {
        pbc neu = {OBC_NEUMANN, 0.0, {0, 0, 0}};
        bc.vel[p] = bc.prs[p] = bc.k[p] = bc.om[p] = neu;
        switch (m->ptype[p]) {
            case OPT_WALL:
                bc.vel[p].type = OBC_WALL_NOSLIP; bc.prs[p].type = OBC_NEUMANN;
                bc.k[p].type = OBC_WALL_KOMEGA; bc.om[p].type = OBC_WALL_KOMEGA; break;
            case OPT_INLET:
                bc.vel[p].type = OBC_INLET_VELOCITY; bc.vel[p].vec = v3make(Uin, 0, 0);
                bc.k[p].type = OBC_DIRICHLET; bc.k[p].value = kIn;
                bc.om[p].type = OBC_DIRICHLET; bc.om[p].value = omIn; break;
            case OPT_OUTLET:
                bc.prs[p].type = OBC_OUTLET_PRESSURE; break;
            case OPT_SYMMETRY:
                bc.vel[p].type = OBC_SYMMETRY; break;
            default: break;
        }
    }
    if (kind == 1 && np > 1) {
        /* couette: top wall moves at Uwall; inlet zero-gradient on every field */
        bc.vel[1].type = OBC_WALL_MOVING; bc.vel[1].vec = v3make(v[0], 0, 0);
        for (int p = 0; p < np; ++p)
            if (m->ptype[p] == OPT_INLET) {
                pbc neu = {OBC_NEUMANN, 0.0, {0, 0, 0}};
                bc.vel[p] = bc.prs[p] = bc.k[p] = bc.om[p] = neu;
            }
    }
    if (kind == 2 && np > 1) {
        /* flat plate: top is free stream */
        bc.vel[1].type = OBC_DIRICHLET; bc.vel[1].vec = v3make(v[0], 0, 0);
        bc.prs[1].type = OBC_NEUMANN;
        bc.k[1].type = OBC_DIRICHLET; bc.k[1].value = kIn;
        bc.om[1].type = OBC_DIRICHLET; bc.om[1].value = omIn;
    }
    return bc;
}
static void free_bc(flowbc* bc) { free(bc->vel); free(bc->prs); free(bc->k); free(bc->om); }

/* applyVelocityBC :196-225 */
static void apply_vel_bc(const om_mesh* m, const flowbc* bc, const v3* U, v3* Ub) {
    for (int p = 0; p < m->nP; ++p)
        for (int j = m->pfPtr[p]; j < m->pfPtr[p + 1]; ++j) {
            int f = m->pfIdx[j], b = f - m->nIF;
            switch (bc->vel[p].type) {
                case OBC_WALL_NOSLIP: Ub[b] = v3make(0, 0, 0); break;
                case OBC_WALL_MOVING: Ub[b] = bc->vel[p].vec; break;
                case OBC_INLET_VELOCITY:
                case OBC_DIRICHLET: Ub[b] = bc->vel[p].vec; break;
                case OBC_SYMMETRY: {
                    v3 Uo = U[m->own[f]];
                    Ub[b] = v3sub(Uo, v3mul(m->fn[f], v3dot(Uo, m->fn[f])));
                    break;
                }
                default: Ub[b] = U[m->own[f]]; break;
            }
        }
}
/* applyPressureBC :229-245 */
static void apply_p_bc(const om_mesh* m, const flowbc* bc, const double* p, double* pb) {
    for (int q = 0; q < m->nP; ++q)
        for (int j = m->pfPtr[q]; j < m->pfPtr[q + 1]; ++j) {
            int f = m->pfIdx[j], b = f - m->nIF;
            int t = bc->prs[q].type;
            pb[b] = (t == OBC_OUTLET_PRESSURE || t == OBC_DIRICHLET) ? bc->prs[q].value : p[m->own[f]];
        }
}
/* applyKBC :249-269 */
static void apply_k_bc(const om_mesh* m, const flowbc* bc, const double* k, double* kb) {
    for (int q = 0; q < m->nP; ++q)
        for (int j = m->pfPtr[q]; j < m->pfPtr[q + 1]; ++j) {
            int f = m->pfIdx[j], b = f - m->nIF;
            int t = bc->k[q].type;
            kb[b] = t == OBC_WALL_KOMEGA ? 0.0 : t == OBC_DIRICHLET ? bc->k[q].value : k[m->own[f]];
        }
}
/* applyOmegaBC :275-300 */
static void apply_om_bc(const om_mesh* m, const flowbc* bc, const double* om, double* omb,
                        double nu, double beta1) {
    for (int q = 0; q < m->nP; ++q)
        for (int j = m->pfPtr[q]; j < m->pfPtr[q + 1]; ++j) {
            int f = m->pfIdx[j], b = f - m->nIF;
            int t = bc->om[q].type;
            if (t == OBC_WALL_KOMEGA) {
                double y1 = smax(m->delta[f], 1e-20);
                omb[b] = 60.0 * nu / (beta1 * y1 * y1);
            } else if (t == OBC_DIRICHLET) {
                omb[b] = bc->om[q].value;
            } else {
                omb[b] = om[m->own[f]];
            }
        }
}

/* ------------------------------------------------------------ LDU system */
/* LinearSystem LinearSolver.hpp:18-85 */
typedef struct {
    int n, nif;
    const int *own, *nbr;
    double *diag, *upper, *lower, *src;
} ldu;

static void ldu_zero(ldu* A) {
    memset(A->diag, 0, sizeof(double) * A->n); memset(A->src, 0, sizeof(double) * A->n);
    memset(A->upper, 0, sizeof(double) * A->nif); memset(A->lower, 0, sizeof(double) * A->nif);
}
static void ldu_matvec(const ldu* A, const double* x, double* y) {
    for (int i = 0; i < A->n; ++i) y[i] = A->diag[i] * x[i];
    for (int f = 0; f < A->nif; ++f) {
        y[A->own[f]] += A->upper[f] * x[A->nbr[f]];
        y[A->nbr[f]] += A->lower[f] * x[A->own[f]];
    }
}
static void ldu_residual(const ldu* A, const double* x, double* r) {
    ldu_matvec(A, x, r);
    for (int i = 0; i < A->n; ++i) r[i] = A->src[i] - r[i];
}
static double dot(const double* a, const double* b, int n) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}
static double vnorm(const double* v, int n) { return sqrt(dot(v, v, n)); }
static void jacobi(const ldu* A, const double* r, double* z) {
    for (int i = 0; i < A->n; ++i) z[i] = (fabs(A->diag[i]) > 1e-30) ? r[i] / A->diag[i] : r[i];
}

typedef struct { int iterations; double finalRes, initialRes; int converged; } sres;

/* PCGSolver :147-186 */
static sres solve_pcg(const ldu* A, double* x, int maxIter, double tol) {
    int n = A->n;
    double *r = calloc(n, 8), *z = calloc(n, 8), *p = calloc(n, 8), *Ap = calloc(n, 8);
    sres res = {0, 0.0, 0.0, 0};
    ldu_residual(A, x, r);
    double r0 = vnorm(r, n);
    res.initialRes = r0;
    if (r0 < 1e-30) { res.converged = 1; goto out; }
    jacobi(A, r, z);
    memcpy(p, z, 8 * n);
    double rz = dot(r, z, n);
    for (int it = 0; it < maxIter; ++it) {
        ldu_matvec(A, p, Ap);
        double pAp = dot(p, Ap, n);
        if (fabs(pAp) < 1e-30) break;
        double alpha = rz / pAp;
        for (int i = 0; i < n; ++i) { x[i] += alpha * p[i]; r[i] -= alpha * Ap[i]; }
        double rn = vnorm(r, n);
        res.iterations = it + 1;
        res.finalRes = rn / r0;
        if (res.finalRes < tol) { res.converged = 1; goto out; }
        jacobi(A, r, z);
        double rzNew = dot(r, z, n);
        double beta = rzNew / (rz + 1e-30);
        for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
        rz = rzNew;
    }
out:
    free(r); free(z); free(p); free(Ap);
    return res;
}

/* BiCGSTABSolver :191-247 */
static sres solve_bicgstab(const ldu* A, double* x, int maxIter, double tol) {
    int n = A->n;
    double *r = calloc(n, 8), *rh = calloc(n, 8), *p = calloc(n, 8), *v = calloc(n, 8);
    double *s = calloc(n, 8), *t = calloc(n, 8), *ph = calloc(n, 8), *sh = calloc(n, 8);
    sres res = {0, 0.0, 0.0, 0};
    ldu_residual(A, x, r);
    double r0 = vnorm(r, n);
    res.initialRes = r0;
    if (r0 < 1e-30) { res.converged = 1; goto out; }
    memcpy(rh, r, 8 * n);
    double rho = 1, alpha = 1, omega = 1;
    for (int it = 0; it < maxIter; ++it) {
        double rhoN = dot(rh, r, n);
        if (fabs(rhoN) < 1e-30) break;
        double beta = (rhoN / (rho + 1e-30)) * (alpha / (omega + 1e-30));
        for (int i = 0; i < n; ++i) p[i] = r[i] + beta * (p[i] - omega * v[i]);
        jacobi(A, p, ph);
        ldu_matvec(A, ph, v);
        double rv = dot(rh, v, n);
        if (fabs(rv) < 1e-30) break;
        alpha = rhoN / rv;
        for (int i = 0; i < n; ++i) s[i] = r[i] - alpha * v[i];
        double sn = vnorm(s, n);
        if (sn / r0 < tol) {
            for (int i = 0; i < n; ++i) x[i] += alpha * ph[i];
            res.iterations = it + 1; res.finalRes = sn / r0; res.initialRes = r0; res.converged = 1;
            goto out;
        }
        jacobi(A, s, sh);
        ldu_matvec(A, sh, t);
        double tt = dot(t, t, n);
        omega = (tt > 1e-30) ? dot(t, s, n) / tt : 0.0;
        for (int i = 0; i < n; ++i) x[i] += alpha * ph[i] + omega * sh[i];
        for (int i = 0; i < n; ++i) r[i] = s[i] - omega * t[i];
        double rn = vnorm(r, n);
        res.iterations = it + 1;
        res.finalRes = rn / r0;
        if (res.finalRes < tol) { res.converged = 1; goto out; }
        if (fabs(omega) < 1e-30) break;
        rho = rhoN;
    }
out:
    free(r); free(rh); free(p); free(v); free(s); free(t); free(ph); free(sh);
    return res;
}

/* GaussSeidelSolver :258-299 */
static sres solve_gs(const ldu* A, double* x, int maxIter, double tol) {
    int n = A->n;
    sres res = {0, 0.0, 0.0, 0};
    double* r = calloc(n, 8);
    double b0 = vnorm(A->src, n);
    res.initialRes = b0;
    if (b0 < 1e-30) b0 = 1.0;
    int* deg = calloc(n, sizeof(int));
    for (int f = 0; f < A->nif; ++f) { ++deg[A->own[f]]; ++deg[A->nbr[f]]; }
    int* off = calloc(n + 1, sizeof(int));
    for (int i = 0; i < n; ++i) off[i + 1] = off[i] + deg[i];
    int* adj = malloc(sizeof(int) * (off[n] + 1));
    double* coef = malloc(sizeof(double) * (off[n] + 1));
    int* cur = malloc(sizeof(int) * (n + 1));
    for (int i = 0; i < n; ++i) cur[i] = off[i];
    for (int f = 0; f < A->nif; ++f) {
        int o = A->own[f], nb = A->nbr[f];
        adj[cur[o]] = nb; coef[cur[o]++] = A->upper[f];
        adj[cur[nb]] = o; coef[cur[nb]++] = A->lower[f];
    }
    for (int it = 0; it < maxIter; ++it) {
        for (int i = 0; i < n; ++i) {
            if (fabs(A->diag[i]) <= 1e-30) continue;
            double s = A->src[i];
            for (int a = off[i]; a < off[i + 1]; ++a) s -= coef[a] * x[adj[a]];
            x[i] = s / A->diag[i];
        }
        ldu_residual(A, x, r);
        res.iterations = it + 1;
        res.finalRes = vnorm(r, n) / b0;
        if (res.finalRes < tol) { res.converged = 1; break; }
    }
    free(r); free(deg); free(off); free(adj); free(coef); free(cur);
    return res;
}

/* ------------------------------------------------------------------ AMG */
/* CSRMatrix :304-341 */
typedef struct { int n; int *rp, *ci; double *v, *dg; } csr;
static void csr_free(csr* A) { free(A->rp); free(A->ci); free(A->v); free(A->dg); }

/* lduToCSR :344-404 (stable insertion sort per row) */
static csr ldu_to_csr(const ldu* A) {
    int n = A->n;
    csr C;
    C.n = n;
    C.rp = calloc(n + 1, sizeof(int));
    int* cnt = malloc(sizeof(int) * (n + 1));
    for (int i = 0; i < n; ++i) cnt[i] = 1;
    for (int f = 0; f < A->nif; ++f) { cnt[A->own[f]]++; cnt[A->nbr[f]]++; }
    for (int i = 0; i < n; ++i) C.rp[i + 1] = C.rp[i] + cnt[i];
    int nnz = C.rp[n];
    C.ci = malloc(sizeof(int) * (nnz + 1)); C.v = malloc(sizeof(double) * (nnz + 1));
    C.dg = malloc(sizeof(double) * (n + 1));
    int* pos = malloc(sizeof(int) * (n + 1));
    for (int i = 0; i < n; ++i) pos[i] = C.rp[i];
    for (int i = 0; i < n; ++i) {
        C.ci[pos[i]] = i; C.v[pos[i]] = A->diag[i]; C.dg[i] = A->diag[i]; pos[i]++;
    }
    for (int f = 0; f < A->nif; ++f) {
        int o = A->own[f], nb = A->nbr[f];
        C.ci[pos[o]] = nb; C.v[pos[o]] = A->upper[f]; pos[o]++;
        C.ci[pos[nb]] = o; C.v[pos[nb]] = A->lower[f]; pos[nb]++;
    }
    for (int i = 0; i < n; ++i) {
        int start = C.rp[i], end = C.rp[i + 1];
        for (int a = start + 1; a < end; ++a) {
            int kc = C.ci[a]; double kv = C.v[a];
            int b = a - 1;
            while (b >= start && C.ci[b] > kc) { C.ci[b + 1] = C.ci[b]; C.v[b + 1] = C.v[b]; b--; }
            C.ci[b + 1] = kc; C.v[b + 1] = kv;
        }
    }
    free(cnt); free(pos);
    return C;
}

/* buildAggregates :407-443 */
static int aggregate(int n, const int* rp, const int* ci, const double* v, const double* dg, int* map) {
    int nc = 0;
    for (int i = 0; i < n; ++i) map[i] = -1;
    for (int i = 0; i < n; ++i) {
        if (map[i] >= 0) continue;
        int agg = nc++;
        map[i] = agg;
        double aii = fabs(dg[i]);
        for (int j = rp[i]; j < rp[i + 1]; ++j) {
            int c = ci[j];
            if (c == i || map[c] >= 0) continue;
            double ajj = fabs(dg[c]);
            double strength = fabs(v[j]) / sqrt(aii * ajj + 1e-30);
            if (strength > 0.25) map[c] = agg;
        }
    }
    for (int i = 0; i < n; ++i) {
        if (map[i] >= 0) continue;
        double best = -1.0;
        int bestAgg = 0;
        for (int j = rp[i]; j < rp[i + 1]; ++j) {
            int c = ci[j];
            if (c == i && map[c] < 0) continue;
            if (map[c] >= 0 && fabs(v[j]) > best) { best = fabs(v[j]); bestAgg = map[c]; }
        }
        map[i] = bestAgg;
    }
    return nc;
}

int om_build_aggregates(int n, const int* rowPtr, const int* colIdx,
                        const double* values, const double* diag, int* map) {
    return aggregate(n, rowPtr, colIdx, values, diag, map);
}

/* galerkinProduct :446-485.  Entries of coarse row I accumulate in the order
 * (fine row i ascending, fine entry j ascending); the reference's hash map only
 * stores them, so a per-row scatter into a dense accumulator with first-touch
 * initialisation (0.0 + v) reproduces the sums exactly. */
static csr galerkin(const csr* A, const int* map, int nc) {
    csr C;
    C.n = nc;
    /* bucket fine rows by aggregate, preserving ascending fine-row order */
    int* cnt = calloc(nc + 1, sizeof(int));
    for (int i = 0; i < A->n; ++i) cnt[map[i] + 1]++;
    for (int I = 0; I < nc; ++I) cnt[I + 1] += cnt[I];
    int* rows = malloc(sizeof(int) * (A->n + 1));
    int* fill = malloc(sizeof(int) * (nc + 1));
    memcpy(fill, cnt, sizeof(int) * nc);
    for (int i = 0; i < A->n; ++i) rows[fill[map[i]]++] = i;
    double* acc = calloc(nc, sizeof(double));
    int* mark = malloc(sizeof(int) * nc);
    for (int J = 0; J < nc; ++J) mark[J] = -1;
    int* cols = malloc(sizeof(int) * (nc + 1));
    int cap = A->rp[A->n] + 16, nnz = 0;
    C.rp = calloc(nc + 1, sizeof(int));
    C.ci = malloc(sizeof(int) * cap); C.v = malloc(sizeof(double) * cap);
    C.dg = calloc(nc + 1, sizeof(double));
    for (int I = 0; I < nc; ++I) {
        int ncol = 0;
        for (int q = cnt[I]; q < cnt[I + 1]; ++q) {
            int i = rows[q];
            for (int j = A->rp[i]; j < A->rp[i + 1]; ++j) {
                int J = map[A->ci[j]];
                if (mark[J] != I) { mark[J] = I; acc[J] = 0.0; cols[ncol++] = J; }
                acc[J] += A->v[j];
            }
        }
        /* sort the row's columns ascending */
        for (int a = 1; a < ncol; ++a) {
            int key = cols[a], b = a - 1;
            while (b >= 0 && cols[b] > key) { cols[b + 1] = cols[b]; b--; }
            cols[b + 1] = key;
        }
        C.rp[I] = nnz;
        for (int a = 0; a < ncol; ++a) {
            if (nnz >= cap) { cap *= 2; C.ci = realloc(C.ci, sizeof(int) * cap); C.v = realloc(C.v, sizeof(double) * cap); }
            C.ci[nnz] = cols[a]; C.v[nnz] = acc[cols[a]];
            if (cols[a] == I) C.dg[I] = acc[cols[a]];
            nnz++;
        }
    }
    C.rp[nc] = nnz;
    free(cnt); free(rows); free(fill); free(acc); free(mark); free(cols);
    return C;
}

/* AMGSolver :497-656 */
typedef struct {
    csr A; int* map; int nc;
    double *x, *b, *r;
} amglev;
typedef struct { amglev lev[21]; int nl; } amg;

static void amg_free(amg* H) {
    for (int l = 0; l < H->nl; ++l) {
        csr_free(&H->lev[l].A); free(H->lev[l].map);
        free(H->lev[l].x); free(H->lev[l].b); free(H->lev[l].r);
    }
    H->nl = 0;
}

static void amg_setup(amg* H, const ldu* A) {
    H->nl = 1;
    memset(&H->lev[0], 0, sizeof(amglev));
    H->lev[0].A = ldu_to_csr(A);
    for (int l = 0; l < 20; ++l) {
        amglev* fine = &H->lev[l];
        if (fine->A.n <= 50) break;
        int* map = malloc(sizeof(int) * (fine->A.n + 1));
        int nc = aggregate(fine->A.n, fine->A.rp, fine->A.ci, fine->A.v, fine->A.dg, map);
        if (nc == 0 || nc >= fine->A.n) { free(map); break; }
        fine->map = map; fine->nc = nc;
        amglev* co = &H->lev[l + 1];
        memset(co, 0, sizeof(amglev));
        co->A = galerkin(&fine->A, map, nc);
        H->nl++;
    }
    for (int l = 0; l < H->nl; ++l) {
        int n = H->lev[l].A.n;
        H->lev[l].x = calloc(n + 1, 8); H->lev[l].b = calloc(n + 1, 8); H->lev[l].r = calloc(n + 1, 8);
    }
}

static void csr_gs(const csr* A, double* x, const double* b) {
    for (int i = 0; i < A->n; ++i) {
        double s = b[i];
        for (int j = A->rp[i]; j < A->rp[i + 1]; ++j)
            if (A->ci[j] != i) s -= A->v[j] * x[A->ci[j]];
        if (fabs(A->dg[i]) > 1e-30) x[i] = s / A->dg[i];
    }
}
static void csr_res(const csr* A, const double* x, const double* b, double* r) {
    for (int i = 0; i < A->n; ++i) {
        double s = 0.0;
        for (int j = A->rp[i]; j < A->rp[i + 1]; ++j) s += A->v[j] * x[A->ci[j]];
        r[i] = b[i] - s;
    }
}

static void vcycle(amg* H, int l) {
    amglev* L = &H->lev[l];
    int n = L->A.n;
    if (l == H->nl - 1) { for (int s = 0; s < 50; ++s) csr_gs(&L->A, L->x, L->b); return; }
    for (int s = 0; s < 2; ++s) csr_gs(&L->A, L->x, L->b);
    csr_res(&L->A, L->x, L->b, L->r);
    amglev* C = &H->lev[l + 1];
    memset(C->b, 0, 8 * C->A.n); memset(C->x, 0, 8 * C->A.n);
    for (int i = 0; i < n; ++i) C->b[L->map[i]] += L->r[i];
    vcycle(H, l + 1);
    for (int i = 0; i < n; ++i) L->x[i] += C->x[L->map[i]];
    for (int s = 0; s < 2; ++s) csr_gs(&L->A, L->x, L->b);
}

static void amg_precond(amg* H, const double* r, double* z) {
    int n = H->lev[0].A.n;
    memcpy(H->lev[0].b, r, 8 * n);
    memset(H->lev[0].x, 0, 8 * n);
    vcycle(H, 0);
    memcpy(z, H->lev[0].x, 8 * n);
}

static sres solve_amg(const ldu* A, double* x, int maxIter, double tol) {
    amg H; memset(&H, 0, sizeof H);
    amg_setup(&H, A);
    int n = A->n;
    double *r = calloc(n, 8), *z = calloc(n, 8), *p = calloc(n, 8), *Ap = calloc(n, 8);
    sres res = {0, 0.0, 0.0, 0};
    ldu_residual(A, x, r);
    double r0 = vnorm(r, n);
    res.initialRes = r0;
    if (r0 < 1e-30) { res.converged = 1; goto out; }
    amg_precond(&H, r, z);
    memcpy(p, z, 8 * n);
    double rz = dot(r, z, n);
    for (int it = 0; it < maxIter; ++it) {
        ldu_matvec(A, p, Ap);
        double pAp = dot(p, Ap, n);
        if (fabs(pAp) < 1e-30) break;
        double alpha = rz / pAp;
        for (int i = 0; i < n; ++i) { x[i] += alpha * p[i]; r[i] -= alpha * Ap[i]; }
        double rn = vnorm(r, n);
        res.iterations = it + 1;
        res.finalRes = rn / r0;
        if (res.finalRes < tol) { res.converged = 1; goto out; }
        amg_precond(&H, r, z);
        double rzNew = dot(r, z, n);
        double beta = rzNew / (rz + 1e-30);
        for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
        rz = rzNew;
    }
out:
    free(r); free(z); free(p); free(Ap);
    amg_free(&H);
    return res;
}

static sres solve_any(int code, const ldu* A, double* x, int maxIter, double tol) {
    switch (code) {
        case ORS_AMG: return solve_amg(A, x, maxIter, tol);
        case ORS_PCG: return solve_pcg(A, x, maxIter, tol);
        case ORS_GS: return solve_gs(A, x, maxIter, tol);
        default: return solve_bicgstab(A, x, maxIter, tol);
    }
}

int om_ldu_solve(int solver, int n, int nif, const int* own, const int* nbr,
                 const double* diag, const double* upper, const double* lower,
                 const double* src, double* x, int maxIter, double tol, double* out) {
    ldu A = {n, nif, own, nbr, (double*)diag, (double*)upper, (double*)lower, (double*)src};
    sres r = solve_any(solver, &A, x, maxIter, tol);
    out[0] = r.iterations; out[1] = r.finalRes; out[2] = r.initialRes; out[3] = r.converged;
    return 0;
}

int om_amg_hierarchy(int n, int nif, const int* own, const int* nbr,
                     const double* diag, const double* upper, const double* lower,
                     int* levelSizes, int* levelNnz, int maxLev, int* maps, int mapsCap) {
    double* zero = calloc(n + 1, 8);
    ldu A = {n, nif, own, nbr, (double*)diag, (double*)upper, (double*)lower, zero};
    amg H; memset(&H, 0, sizeof H);
    amg_setup(&H, &A);
    int off = 0;
    for (int l = 0; l < H.nl; ++l) {
        if (l < maxLev) { levelSizes[l] = H.lev[l].A.n; levelNnz[l] = H.lev[l].A.rp[H.lev[l].A.n]; }
        if (l < H.nl - 1)
            for (int i = 0; i < H.lev[l].A.n && off < mapsCap; ++i) maps[off++] = H.lev[l].map[i];
    }
    int nl = H.nl;
    amg_free(&H); free(zero);
    return nl;
}

/* ---------------------------------------------------------------- SIMPLE */
typedef struct {
    int maxIter; double tol, divLimit, aU, aP, aK, aOm;
    int innerIter; double innerTol;
    int turbStart, turbInterval, nutFloorIters, rcAll;
    v3 bf; double kMin, kMax, omMin; int wallFn; double vonK; int variant;
    int pSolver, mSolver, tSolver;
} settings_t;

static settings_t mk_settings(const double* s) {
    settings_t o;
    o.maxIter = (int)s[ORS_MAX_ITER]; o.tol = s[ORS_CONV_TOL]; o.divLimit = s[ORS_DIV_LIMIT];
    o.aU = s[ORS_ALPHA_U]; o.aP = s[ORS_ALPHA_P]; o.aK = s[ORS_ALPHA_K]; o.aOm = s[ORS_ALPHA_OM];
    o.innerIter = (int)s[ORS_INNER_ITER]; o.innerTol = s[ORS_INNER_TOL];
    o.turbStart = (int)s[ORS_TURB_START]; o.turbInterval = (int)s[ORS_TURB_INTERVAL];
    o.nutFloorIters = (int)s[ORS_NUT_FLOOR_ITERS]; o.rcAll = s[ORS_RC_ALL] != 0.0;
    o.bf = v3make(s[ORS_BF_X], s[ORS_BF_Y], s[ORS_BF_Z]);
    o.kMin = s[ORS_K_MIN]; o.kMax = s[ORS_K_MAX]; o.omMin = s[ORS_OM_MIN];
    o.wallFn = s[ORS_WALL_FN] != 0.0; o.vonK = s[ORS_VON_KARMAN];
    o.variant = (int)s[ORS_SST_VARIANT];
    o.pSolver = (int)s[ORS_P_SOLVER]; o.mSolver = (int)s[ORS_M_SOLVER]; o.tSolver = (int)s[ORS_T_SOLVER];
    return o;
}

typedef struct {
    v3 *U, *Ub;
    double *p, *pb, *k, *kb, *om, *omb;
    double *nuT, *F1, *F2, *Pk, *CDkw;
    int turbEstablished;
} fields_t;

typedef struct {
    const om_mesh* m;
    flowbc bc;
    sstc sst;
    settings_t s;
    double nu;
    double *aP, *aPun;
    double normUx0, normUy0, normP0;
    /* scratch */
    velgrad vg;
    v3 *gradP;
    double *scC, *scB;
    char* isWallFace;
} solver_t;

static void solver_init(solver_t* S, const om_mesh* m, int bcKind, const double* bcVals,
                        double nu, const double* coeffs, const double* settings) {
    memset(S, 0, sizeof *S);
    S->m = m; S->bc = mk_bc(m, bcKind, bcVals); S->s = mk_settings(settings);
    S->sst = mk_sst(coeffs, S->s.variant); S->nu = nu;
    int n = m->nC, nB = m->nF - m->nIF;
    S->aP = calloc(n + 1, 8); S->aPun = calloc(n + 1, 8);
    S->normUx0 = S->normUy0 = S->normP0 = 1.0;
    S->vg.dudx = calloc(n + 1, sizeof(v3)); S->vg.dvdx = calloc(n + 1, sizeof(v3));
    S->vg.dwdx = calloc(n + 1, sizeof(v3));
    S->gradP = calloc(n + 1, sizeof(v3));
    S->scC = calloc(n + 1, 8); S->scB = calloc(nB + 1, 8);
    S->isWallFace = calloc(m->nF + 1, 1);
    for (int p = 0; p < m->nP; ++p)
        if (m->ptype[p] == OPT_WALL)
            for (int j = m->pfPtr[p]; j < m->pfPtr[p + 1]; ++j) S->isWallFace[m->pfIdx[j]] = 1;
}
static void solver_free(solver_t* S) {
    free_bc(&S->bc); free(S->aP); free(S->aPun); free(S->vg.dudx); free(S->vg.dvdx);
    free(S->vg.dwdx); free(S->gradP); free(S->scC); free(S->scB); free(S->isWallFace);
}

static fields_t fields_alloc(const om_mesh* m) {
    fields_t f;
    int n = m->nC, nB = m->nF - m->nIF;
    f.U = calloc(n + 1, sizeof(v3)); f.Ub = calloc(nB + 1, sizeof(v3));
    f.p = calloc(n + 1, 8); f.pb = calloc(nB + 1, 8);
    f.k = calloc(n + 1, 8); f.kb = calloc(nB + 1, 8);
    f.om = calloc(n + 1, 8); f.omb = calloc(nB + 1, 8);
    f.nuT = calloc(n + 1, 8); f.F1 = calloc(n + 1, 8); f.F2 = calloc(n + 1, 8);
    f.Pk = calloc(n + 1, 8); f.CDkw = calloc(n + 1, 8);
    f.turbEstablished = 0;
    return f;
}
static void fields_free(fields_t* f) {
    free(f->U); free(f->Ub); free(f->p); free(f->pb); free(f->k); free(f->kb);
    free(f->om); free(f->omb); free(f->nuT); free(f->F1); free(f->F2); free(f->Pk); free(f->CDkw);
}

/* SSTModel::computeFields SSTModel.cpp:122-172 (Smag output is the same |S|) */
static void compute_fields(solver_t* S, fields_t* f, double* Smag) {
    const om_mesh* m = S->m;
    int n = m->nC;
    vel_grads(m, f->U, f->Ub, &S->vg, S->scC, S->scB);
    strain_mag(m, &S->vg, Smag);
    v3* gk = calloc(n + 1, sizeof(v3));
    v3* gw = calloc(n + 1, sizeof(v3));
    gg(m, f->k, f->kb, gk);
    gg(m, f->om, f->omb, gw);
    for (int c = 0; c < n; ++c) {
        double kc = smax(f->k[c], 0.0);
        double wc = smax(f->om[c], 1e-20);
        double y = smax(m->wd[c], 1e-20);
        double Sc = Smag[c];
        double CDkw = 2.0 * S->sst.sw2 / wc * v3dot(gk[c], gw[c]);
        f->CDkw[c] = CDkw;
        double CDpos = smax(CDkw, 1e-20);
        f->F1[c] = (S->sst.variant == 2) ? 1.0 : sst_F1(&S->sst, kc, wc, y, S->nu, CDpos);
        f->F2[c] = sst_F2(&S->sst, kc, wc, y, S->nu);
        f->nuT[c] = sst_nut(&S->sst, kc, wc, Sc, f->F2[c]);
        f->Pk[c] = sst_prod(&S->sst, f->nuT[c], Sc, kc, wc);
    }
    free(gk); free(gw);
}

static inline double comp_of(v3 a, int c) { return c == 0 ? a.x : c == 1 ? a.y : a.z; }

/* transposeStressSource StressOperators.hpp:31-84 with the wall-zero coefficient */
static void transpose_src(solver_t* S, const double* coeff, const fields_t* f, int comp, double* src) {
    const om_mesh* m = S->m;
    vel_grads(m, f->U, f->Ub, &S->vg, S->scC, S->scB);
    velgrad* vg = &S->vg;
#define GCOL(ci) (comp == 0 ? v3make(vg->dudx[ci].x, vg->dvdx[ci].x, vg->dwdx[ci].x) : \
                  comp == 1 ? v3make(vg->dudx[ci].y, vg->dvdx[ci].y, vg->dwdx[ci].y) : \
                              v3make(vg->dudx[ci].z, vg->dvdx[ci].z, vg->dwdx[ci].z))
    for (int c = 0; c < m->nC; ++c) src[c] = 0.0;
    for (int fi = 0; fi < m->nIF; ++fi) {
        int o = m->own[fi], n = m->nbr[fi];
        double w = m->weight[fi];
        double cf = w * coeff[o] + (1.0 - w) * coeff[n];
        v3 g = v3add(v3mul(GCOL(o), w), v3mul(GCOL(n), 1.0 - w));
        double flux = cf * v3dot(g, m->fn[fi]) * m->area[fi];
        src[o] += flux;
        src[n] -= flux;
    }
    for (int p = 0; p < m->nP; ++p) {
        int isWall = m->ptype[p] == OPT_WALL;
        for (int j = m->pfPtr[p]; j < m->pfPtr[p + 1]; ++j) {
            int fi = m->pfIdx[j], o = m->own[fi];
            double cf = isWall ? 0.0 : coeff[o];
            double flux = cf * v3dot(GCOL(o), m->fn[fi]) * m->area[fi];
            src[o] += flux;
        }
    }
#undef GCOL
}

int om_transpose_stress(const om_mesh* m, const double* coeff, const double* U, const double* Ub,
                        int comp, int wallZero, double* src) {
    /* transposeStressSource with (wallZero) or without the wall coefficient */
    double zeroBc[3] = {0.0, 0.0, 0.0};
    (void)zeroBc;
    solver_t S;
    memset(&S, 0, sizeof S);
    S.m = m;
    int n = m->nC, nB = m->nF - m->nIF;
    S.vg.dudx = calloc(n + 1, sizeof(v3)); S.vg.dvdx = calloc(n + 1, sizeof(v3)); S.vg.dwdx = calloc(n + 1, sizeof(v3));
    S.scC = calloc(n + 1, 8); S.scB = calloc(nB + 1, 8);
    fields_t f; memset(&f, 0, sizeof f);
    f.U = (v3*)U; f.Ub = (v3*)Ub;
    if (wallZero) {
        transpose_src(&S, coeff, &f, comp, src);
    } else {
        /* owner extrapolation on every boundary face (StressOperators.hpp:60-67) */
        vel_grads(m, f.U, f.Ub, &S.vg, S.scC, S.scB);
        velgrad* vg = &S.vg;
#define GC(ci) (comp == 0 ? v3make(vg->dudx[ci].x, vg->dvdx[ci].x, vg->dwdx[ci].x) : \
                comp == 1 ? v3make(vg->dudx[ci].y, vg->dvdx[ci].y, vg->dwdx[ci].y) : \
                            v3make(vg->dudx[ci].z, vg->dvdx[ci].z, vg->dwdx[ci].z))
        for (int c = 0; c < n; ++c) src[c] = 0.0;
        for (int fi = 0; fi < m->nIF; ++fi) {
            int o = m->own[fi], nb = m->nbr[fi];
            double w = m->weight[fi];
            double cf = w * coeff[o] + (1.0 - w) * coeff[nb];
            v3 g = v3add(v3mul(GC(o), w), v3mul(GC(nb), 1.0 - w));
            double flux = cf * v3dot(g, m->fn[fi]) * m->area[fi];
            src[o] += flux; src[nb] -= flux;
        }
        for (int fi = m->nIF; fi < m->nF; ++fi) {
            int o = m->own[fi];
            double flux = coeff[o] * v3dot(GC(o), m->fn[fi]) * m->area[fi];
            src[o] += flux;
        }
#undef GC
    }
    free(S.vg.dudx); free(S.vg.dvdx); free(S.vg.dwdx); free(S.scC); free(S.scB);
    return 0;
}

/* assembleMomentum SIMPLESolver.cpp:56-203 (injection source excluded) */
static void assemble_momentum(solver_t* S, ldu* A, const fields_t* f, int comp, double* aP) {
    const om_mesh* m = S->m;
    double nu = S->nu;
    ldu_zero(A);
    for (int fi = 0; fi < m->nIF; ++fi) {
        int o = m->own[fi], n = m->nbr[fi];
        double Sf = m->area[fi], delta = smax(m->delta[fi], 1e-20), w = m->weight[fi];
        double nuEff_o = nu + f->nuT[o];
        double nuEff_n = nu + f->nuT[n];
        double nuEff_f = w * nuEff_o + (1.0 - w) * nuEff_n;
        double Df = nuEff_f * Sf / delta;
        v3 Uf = v3add(v3mul(f->U[o], w), v3mul(f->U[n], 1.0 - w));
        v3 nn = m->fn[fi];
        double mFlux = (Uf.x * nn.x + Uf.y * nn.y + Uf.z * nn.z) * Sf;
        double cPos = smax(mFlux, 0.0), cNeg = smax(-mFlux, 0.0);
        A->diag[o] += Df + cPos;
        A->diag[n] += Df + cNeg;
        A->upper[fi] = -(Df + cNeg);
        A->lower[fi] = -(Df + cPos);
    }
    for (int fi = m->nIF; fi < m->nF; ++fi) {
        int o = m->own[fi], b = fi - m->nIF;
        double Sf = m->area[fi], delta = smax(m->delta[fi], 1e-20);
        double nuEff = (S->isWallFace[fi] && !S->s.wallFn) ? nu : nu + f->nuT[o];
        double Db = nuEff * Sf / delta;
        double Ub = comp_of(f->Ub[b], comp);
        v3 Ubv = f->Ub[b], nn = m->fn[fi];
        double mFlux = (Ubv.x * nn.x + Ubv.y * nn.y + Ubv.z * nn.z) * Sf;
        if (mFlux >= 0) { A->diag[o] += mFlux + Db; A->src[o] += Db * Ub; }
        else { A->diag[o] += Db; A->src[o] += (Db - mFlux) * Ub; }
    }
    gg(m, f->p, f->pb, S->gradP);
    for (int c = 0; c < m->nC; ++c) A->src[c] -= comp_of(S->gradP[c], comp) * m->vol[c];
    double fb = comp_of(S->s.bf, comp);
    if (fb != 0.0)
        for (int c = 0; c < m->nC; ++c) A->src[c] += fb * m->vol[c];
    double* ts = calloc(m->nC + 1, 8);
    transpose_src(S, f->nuT, f, comp, ts);
    for (int c = 0; c < m->nC; ++c) A->src[c] += ts[c];
    free(ts);
    double a = S->s.aU;
    for (int c = 0; c < m->nC; ++c) {
        double phi_old = comp_of(f->U[c], comp);
        S->aPun[c] = A->diag[c];
        A->src[c] += (1.0 - a) / a * A->diag[c] * phi_old;
        A->diag[c] /= a;
        aP[c] = A->diag[c];
    }
}

static int has_outlet(const om_mesh* m) {
    for (int p = 0; p < m->nP; ++p)
        if (m->ptype[p] == OPT_OUTLET && m->pfPtr[p + 1] > m->pfPtr[p]) return 1;
    return 0;
}

/* assemblePressureCorrection SIMPLESolver.cpp:322-450 */
static void assemble_pcorr(solver_t* S, ldu* A, const fields_t* f, const double* aP) {
    const om_mesh* m = S->m;
    ldu_zero(A);
    int hasOut = has_outlet(m);
    int rc = !hasOut || S->s.rcAll;
    if (rc) gg(m, f->p, f->pb, S->gradP);
    for (int fi = 0; fi < m->nIF; ++fi) {
        int o = m->own[fi], n = m->nbr[fi];
        double Sf = m->area[fi], delta = smax(m->delta[fi], 1e-20), w = m->weight[fi];
        double Vo = m->vol[o], Vn = m->vol[n];
        double dP_o = (fabs(aP[o]) > 1e-30) ? Vo / aP[o] : 0.0;
        double dP_n = (fabs(aP[n]) > 1e-30) ? Vn / aP[n] : 0.0;
        double dP_f = w * dP_o + (1.0 - w) * dP_n;
        double coeff = dP_f * Sf / delta;
        A->diag[o] += coeff; A->diag[n] += coeff;
        A->upper[fi] = -coeff; A->lower[fi] = -coeff;
        v3 Uf = v3add(v3mul(f->U[o], w), v3mul(f->U[n], 1.0 - w));
        v3 nn = m->fn[fi];
        double massFlux = (Uf.x * nn.x + Uf.y * nn.y + Uf.z * nn.z) * Sf;
        if (rc) {
            v3 ehat = v3div(m->fd[fi], smax(m->delta[fi], 1e-30));
            v3 gbar = v3add(v3mul(S->gradP[o], w), v3mul(S->gradP[n], 1.0 - w));
            massFlux += -dP_f * Sf * ((f->p[n] - f->p[o]) / delta - v3dot(gbar, ehat));
        }
        A->src[o] -= massFlux;
        A->src[n] += massFlux;
    }
    for (int p = 0; p < m->nP; ++p) {
        int isOut = m->ptype[p] == OPT_OUTLET;
        for (int j = m->pfPtr[p]; j < m->pfPtr[p + 1]; ++j) {
            int fi = m->pfIdx[j], o = m->own[fi], b = fi - m->nIF;
            double Sf = m->area[fi], delta = smax(m->delta[fi], 1e-20);
            v3 Ub = f->Ub[b], nn = m->fn[fi];
            double massFlux = (Ub.x * nn.x + Ub.y * nn.y + Ub.z * nn.z) * Sf;
            A->src[o] -= massFlux;
            if (isOut) {
                double Vo = m->vol[o];
                double dP_o = (fabs(aP[o]) > 1e-30) ? Vo / aP[o] : 0.0;
                double coeff = dP_o * Sf / delta;
                A->diag[o] += coeff;
            }
        }
    }
    if (!hasOut && m->nC > 0) A->diag[0] += fabs(A->diag[0]) > 1e-30 ? A->diag[0] : 1.0;
}

/* assembleKEquation SIMPLESolver.cpp:483-560 */
static void assemble_k(solver_t* S, ldu* A, const fields_t* f) {
    const om_mesh* m = S->m;
    double nu = S->nu;
    ldu_zero(A);
    for (int fi = 0; fi < m->nIF; ++fi) {
        int o = m->own[fi], n = m->nbr[fi];
        double Sf = m->area[fi], delta = smax(m->delta[fi], 1e-20), w = m->weight[fi];
        double F1_f = w * f->F1[o] + (1.0 - w) * f->F1[n];
        double sk = blend(S->sst.sk1, S->sst.sk2, F1_f);
        double nuEff_o = nu + sk * f->nuT[o];
        double nuEff_n = nu + sk * f->nuT[n];
        double nuEff_f = w * nuEff_o + (1.0 - w) * nuEff_n;
        double Df = nuEff_f * Sf / delta;
        v3 Uf = v3add(v3mul(f->U[o], w), v3mul(f->U[n], 1.0 - w));
        v3 nn = m->fn[fi];
        double mFlux = (Uf.x * nn.x + Uf.y * nn.y + Uf.z * nn.z) * Sf;
        double cPos = smax(mFlux, 0.0), cNeg = smax(-mFlux, 0.0);
        A->diag[o] += Df + cPos; A->diag[n] += Df + cNeg;
        A->upper[fi] = -(Df + cNeg); A->lower[fi] = -(Df + cPos);
    }
    for (int fi = m->nIF; fi < m->nF; ++fi) {
        int o = m->own[fi], b = fi - m->nIF;
        double Sf = m->area[fi], delta = smax(m->delta[fi], 1e-20);
        double sk = blend(S->sst.sk1, S->sst.sk2, f->F1[o]);
        double nuEff = nu + sk * f->nuT[o];
        double Db = nuEff * Sf / delta;
        double kb = f->kb[b];
        v3 Ub = f->Ub[b], nn = m->fn[fi];
        double mFlux = (Ub.x * nn.x + Ub.y * nn.y + Ub.z * nn.z) * Sf;
        if (mFlux >= 0) { A->diag[o] += mFlux + Db; A->src[o] += Db * kb; }
        else { A->diag[o] += Db; A->src[o] += (Db - mFlux) * kb; }
    }
    for (int c = 0; c < m->nC; ++c) {
        double vol = m->vol[c];
        A->src[c] += f->Pk[c] * vol;
        double destruction = S->sst.bs * smax(f->om[c], 1e-20);
        A->diag[c] += destruction * vol;
    }
    double a = S->s.aK;
    for (int c = 0; c < m->nC; ++c) {
        A->src[c] += (1.0 - a) / a * A->diag[c] * f->k[c];
        A->diag[c] /= a;
    }
}

/* assembleOmegaEquation SIMPLESolver.cpp:567-669 */
static void assemble_om(solver_t* S, ldu* A, const fields_t* f, const double* Smag) {
    const om_mesh* m = S->m;
    double nu = S->nu;
    ldu_zero(A);
    for (int fi = 0; fi < m->nIF; ++fi) {
        int o = m->own[fi], n = m->nbr[fi];
        double Sf = m->area[fi], delta = smax(m->delta[fi], 1e-20), w = m->weight[fi];
        double F1_f = w * f->F1[o] + (1.0 - w) * f->F1[n];
        double sw = blend(S->sst.sw1, S->sst.sw2, F1_f);
        double nuEff_o = nu + sw * f->nuT[o];
        double nuEff_n = nu + sw * f->nuT[n];
        double nuEff_f = w * nuEff_o + (1.0 - w) * nuEff_n;
        double Df = nuEff_f * Sf / delta;
        v3 Uf = v3add(v3mul(f->U[o], w), v3mul(f->U[n], 1.0 - w));
        v3 nn = m->fn[fi];
        double mFlux = (Uf.x * nn.x + Uf.y * nn.y + Uf.z * nn.z) * Sf;
        double cPos = smax(mFlux, 0.0), cNeg = smax(-mFlux, 0.0);
        A->diag[o] += Df + cPos; A->diag[n] += Df + cNeg;
        A->upper[fi] = -(Df + cNeg); A->lower[fi] = -(Df + cPos);
    }
    for (int fi = m->nIF; fi < m->nF; ++fi) {
        int o = m->own[fi], b = fi - m->nIF;
        double Sf = m->area[fi], delta = smax(m->delta[fi], 1e-20);
        double sw = blend(S->sst.sw1, S->sst.sw2, f->F1[o]);
        double nuEff = nu + sw * f->nuT[o];
        double Db = nuEff * Sf / delta;
        double wb = f->omb[b];
        v3 Ub = f->Ub[b], nn = m->fn[fi];
        double mFlux = (Ub.x * nn.x + Ub.y * nn.y + Ub.z * nn.z) * Sf;
        if (mFlux >= 0) { A->diag[o] += mFlux + Db; A->src[o] += Db * wb; }
        else { A->diag[o] += Db; A->src[o] += (Db - mFlux) * wb; }
    }
    for (int c = 0; c < m->nC; ++c) {
        double vol = m->vol[c];
        double F1 = f->F1[c];
        double omC = smax(f->om[c], 1e-20);
        double alphaB = blend(S->sst.a1_, S->sst.a2_, F1);
        double Sm = Smag[c];
        A->src[c] += alphaB * sst_prod_om(&S->sst, f->nuT[c], Sm, f->k[c], f->om[c]) * vol;
        double betaB = blend(S->sst.b1, S->sst.b2, F1);
        A->diag[c] += betaB * omC * vol;
        A->src[c] += (1.0 - F1) * f->CDkw[c] * vol;
    }
    double a = S->s.aOm;
    for (int c = 0; c < m->nC; ++c) {
        A->src[c] += (1.0 - a) / a * A->diag[c] * f->om[c];
        A->diag[c] /= a;
    }
}

/* initUniform SIMPLESolver.cpp:23-51 */
static void init_uniform(solver_t* S, fields_t* f, const double* init) {
    const om_mesh* m = S->m;
    int n = m->nC, nB = m->nF - m->nIF;
    v3 U0 = v3make(init[0], init[1], init[2]);
    for (int c = 0; c < n; ++c) {
        f->U[c] = U0; f->p[c] = init[3]; f->k[c] = init[4]; f->om[c] = init[5];
        f->F1[c] = 1.0; f->F2[c] = 1.0; f->Pk[c] = 0.0; f->CDkw[c] = 0.0;
    }
    for (int b = 0; b < nB; ++b) { f->Ub[b] = U0; f->pb[b] = init[3]; f->kb[b] = init[4]; f->omb[b] = init[5]; }
    f->turbEstablished = 0;
    for (int c = 0; c < n; ++c) {
        double y = smax(m->wd[c], 1e-20);
        double omWall = 60.0 * S->nu / (S->sst.b1 * y * y);
        f->om[c] = smax(omWall, init[5]);
    }
    for (int c = 0; c < n; ++c) f->nuT[c] = S->sst.a1 * init[4] / smax(S->sst.a1 * f->om[c], 1e-20);
    apply_vel_bc(m, &S->bc, f->U, f->Ub);
    apply_p_bc(m, &S->bc, f->p, f->pb);
    apply_k_bc(m, &S->bc, f->k, f->kb);
    apply_om_bc(m, &S->bc, f->om, f->omb, S->nu, 0.075);  /* applyAllBCs default beta1 */
}

static double clampd(double v, double lo, double hi) { return smax(lo, smin(hi, v)); }

/* SIMPLESolver::solve SIMPLESolver.cpp:696-1043 */
typedef struct { int finalIter, converged, diverged, nEntries; } hist_t;

static hist_t simple_solve(solver_t* S, fields_t* f, double* hist, int histCap) {
    const om_mesh* m = S->m;
    int n = m->nC, nIF = m->nIF, nB = m->nF - m->nIF;
    settings_t* st = &S->s;
    hist_t H = {0, 0, 0, 0};
    double *buf = calloc((size_t)(4 * n + 2 * nIF) * 4 + 8, 8);
    ldu mom = {n, nIF, m->own, m->nbr, buf, buf + n, buf + n + nIF, buf + n + 2 * nIF};
    double* b2 = buf + 2 * n + 2 * nIF;
    ldu ps = {n, nIF, m->own, m->nbr, b2, b2 + n, b2 + n + nIF, b2 + n + 2 * nIF};
    double* b3 = b2 + 2 * n + 2 * nIF;
    ldu ks = {n, nIF, m->own, m->nbr, b3, b3 + n, b3 + n + nIF, b3 + n + 2 * nIF};
    double* b4 = b3 + 2 * n + 2 * nIF;
    ldu os = {n, nIF, m->own, m->nbr, b4, b4 + n, b4 + n + nIF, b4 + n + 2 * nIF};
    double *pp = calloc(n + 1, 8), *ppb = calloc(nB + 1, 8);
    double *Smag = calloc(n + 1, 8), *vec = calloc(n + 1, 8), *aPstore = calloc(n + 1, 8);
    double *kOld = calloc(n + 1, 8), *omOld = calloc(n + 1, 8);
    v3* gpp = calloc(n + 1, sizeof(v3));
    double lastK = 1.0, lastOm = 1.0;
    int turbScheduled = st->turbStart < st->maxIter;

    for (int iter = 0; iter < st->maxIter; ++iter) {
        int turbActive = iter >= st->turbStart;
        int turbUpdate = turbActive && ((iter - st->turbStart) % st->turbInterval == 0);
        if (turbUpdate) {
            compute_fields(S, f, Smag);
            if (!f->turbEstablished && iter >= st->turbStart + st->nutFloorIters) f->turbEstablished = 1;
            double nuTMin = f->turbEstablished ? 0.0 : 0.1 * S->nu;
            for (int c = 0; c < n; ++c) f->nuT[c] = smax(f->nuT[c], nuTMin);
        }
        sres rU[3];
        for (int comp = 0; comp < 3; ++comp) {
            assemble_momentum(S, &mom, f, comp, S->aP);
            for (int c = 0; c < n; ++c) vec[c] = comp_of(f->U[c], comp);
            rU[comp] = solve_any(st->mSolver, &mom, vec, st->innerIter, st->innerTol);
            for (int c = 0; c < n; ++c) {
                if (comp == 0) f->U[c].x = vec[c];
                else if (comp == 1) f->U[c].y = vec[c];
                else f->U[c].z = vec[c];
            }
            if (comp == 0) memcpy(aPstore, S->aP, 8 * n);
        }
        apply_vel_bc(m, &S->bc, f->U, f->Ub);

        assemble_pcorr(S, &ps, f, aPstore);
        for (int c = 0; c < n; ++c) pp[c] = 0.0;
        for (int b = 0; b < nB; ++b) ppb[b] = 0.0;
        sres rP = solve_any(st->pSolver, &ps, pp, st->innerIter, st->innerTol);
        for (int q = 0; q < m->nP; ++q) {
            if (m->ptype[q] == OPT_OUTLET) continue;
            for (int j = m->pfPtr[q]; j < m->pfPtr[q + 1]; ++j) {
                int fi = m->pfIdx[j];
                ppb[fi - nIF] = pp[m->own[fi]];
            }
        }
        /* correctVelocity :456-466, correctPressure :470-475 */
        gg(m, pp, ppb, gpp);
        for (int c = 0; c < n; ++c) {
            double rAP = (fabs(aPstore[c]) > 1e-30) ? 1.0 / aPstore[c] : 0.0;
            double vol = m->vol[c];
            f->U[c].x -= rAP * gpp[c].x * vol;
            f->U[c].y -= rAP * gpp[c].y * vol;
            f->U[c].z -= rAP * gpp[c].z * vol;
        }
        apply_vel_bc(m, &S->bc, f->U, f->Ub);
        for (int c = 0; c < n; ++c) f->p[c] += st->aP * pp[c];
        apply_p_bc(m, &S->bc, f->p, f->pb);

        if (turbUpdate) {
            memcpy(kOld, f->k, 8 * n); memcpy(omOld, f->om, 8 * n);
            assemble_k(S, &ks, f);
            memcpy(vec, f->k, 8 * n);
            solve_any(st->tSolver, &ks, vec, st->innerIter, st->innerTol);
            for (int c = 0; c < n; ++c) f->k[c] = clampd(vec[c], st->kMin, st->kMax);
            for (int b = 0; b < nB; ++b) f->kb[b] = clampd(f->kb[b], st->kMin, st->kMax);
            apply_k_bc(m, &S->bc, f->k, f->kb);

            assemble_om(S, &os, f, Smag);
            memcpy(vec, f->om, 8 * n);
            solve_any(st->tSolver, &os, vec, st->innerIter, st->innerTol);
            for (int c = 0; c < n; ++c) f->om[c] = clampd(vec[c], st->omMin, 1e15);
            for (int b = 0; b < nB; ++b) f->omb[b] = clampd(f->omb[b], st->omMin, 1e15);
            apply_om_bc(m, &S->bc, f->om, f->omb, S->nu, S->sst.b1);

            /* wall-cell omega pin :884-908 */
            double betaStar = S->sst.bs, kappa = st->vonK;
            for (int q = 0; q < m->nP; ++q) {
                if (m->ptype[q] != OPT_WALL) continue;
                for (int j = m->pfPtr[q]; j < m->pfPtr[q + 1]; ++j) {
                    int fi = m->pfIdx[j], o = m->own[fi];
                    double y1 = smax(m->delta[fi], 1e-20);
                    double omRes = 60.0 * S->nu / (S->sst.b1 * y1 * y1);
                    if (st->wallFn) {
                        double k_p = smax(f->k[o], 1e-30);
                        double uTau = sqrt(sqrt(betaStar) * k_p);
                        double omLog = uTau / (sqrt(betaStar) * kappa * y1);
                        f->om[o] = sqrt(omRes * omRes + omLog * omLog);
                    } else {
                        f->om[o] = omRes;
                    }
                }
            }
            double kMaxVal = 1e-30, omMaxVal = 1e-30, kMaxDiff = 0.0, omMaxDiff = 0.0;
            for (int c = 0; c < n; ++c) {
                kMaxVal = smax(kMaxVal, fabs(f->k[c]));
                omMaxVal = smax(omMaxVal, fabs(f->om[c]));
                kMaxDiff = smax(kMaxDiff, fabs(f->k[c] - kOld[c]));
                omMaxDiff = smax(omMaxDiff, fabs(f->om[c] - omOld[c]));
            }
            lastK = kMaxDiff / kMaxVal;
            lastOm = omMaxDiff / omMaxVal;
        }

        double eUx = rU[0].initialRes, eUy = rU[1].initialRes, eUz = rU[2].initialRes, eP = rP.initialRes;
        double eK = turbActive ? lastK : 0.0, eOm = turbActive ? lastOm : 0.0;
        if (hist && iter < histCap) {
            double* h = hist + 6 * (size_t)iter;
            h[0] = eUx; h[1] = eUy; h[2] = eUz; h[3] = eP; h[4] = eK; h[5] = eOm;
        }
        H.nEntries = iter + 1;
        if (iter == 0) {
            S->normUx0 = smax(eUx, 1e-30); S->normUy0 = smax(eUy, 1e-30); S->normP0 = smax(eP, 1e-30);
        } else if (iter < 50) {
            S->normP0 = smax(S->normP0, eP);
        }
        double normMom0 = smax(S->normUx0, S->normUy0);
        double nUx = normMom0 < 1e-20 ? 0.0 : eUx / normMom0;
        double nUy = normMom0 < 1e-20 ? 0.0 : eUy / normMom0;
        double nP = S->normP0 < 1e-20 ? 0.0 : eP / S->normP0;
        double maxRes = smax(smax(nUx, nUy), nP);
        if (turbActive) maxRes = smax(smax(maxRes, lastK), lastOm);
        if (!isfinite(eUx) || !isfinite(eUy) || !isfinite(eUz) || !isfinite(eP) || !isfinite(eK)
            || !isfinite(eOm) || !isfinite(lastK) || !isfinite(lastOm))
            maxRes = INFINITY;
        if (maxRes < st->tol && iter > 0 && (!turbScheduled || (turbActive && f->turbEstablished))) {
            H.converged = 1; H.finalIter = iter; goto done;
        }
        if (isnan(maxRes) || isinf(maxRes) || maxRes > st->divLimit) {
            H.diverged = 1; H.finalIter = iter; goto done;
        }
    }
    H.finalIter = st->maxIter;
done:
    free(buf); free(pp); free(ppb); free(Smag); free(vec); free(aPstore);
    free(kOld); free(omOld); free(gpp);
    return H;
}

static void export_fields(const om_mesh* m, const fields_t* f, double* U, double* p, double* k,
                          double* om, double* nuT, double* F1, double* F2, double* Pk, double* CDkw,
                          double* Ub, double* pb, double* kb, double* omb) {
    int n = m->nC, nB = m->nF - m->nIF;
    if (U) memcpy(U, f->U, sizeof(v3) * n);
    if (Ub) memcpy(Ub, f->Ub, sizeof(v3) * nB);
    if (p) memcpy(p, f->p, 8 * n);
    if (k) memcpy(k, f->k, 8 * n);
    if (om) memcpy(om, f->om, 8 * n);
    if (nuT) memcpy(nuT, f->nuT, 8 * n);
    if (F1) memcpy(F1, f->F1, 8 * n);
    if (F2) memcpy(F2, f->F2, 8 * n);
    if (Pk) memcpy(Pk, f->Pk, 8 * n);
    if (CDkw) memcpy(CDkw, f->CDkw, 8 * n);
    if (pb) memcpy(pb, f->pb, 8 * nB);
    if (kb) memcpy(kb, f->kb, 8 * nB);
    if (omb) memcpy(omb, f->omb, 8 * nB);
}

int om_simple_solve(const om_mesh* m, int bcKind, const double* bcVals, double nu,
                    const double* coeffs, const double* settings, const double* init,
                    double* U, double* p, double* k, double* om, double* nuT,
                    double* F1, double* F2, double* Pk, double* CDkw,
                    double* Ub, double* pb, double* kb, double* omb,
                    double* hist, int histCap, int* info) {
    solver_t S;
    solver_init(&S, m, bcKind, bcVals, nu, coeffs, settings);
    fields_t f = fields_alloc(m);
    init_uniform(&S, &f, init);
    hist_t H = simple_solve(&S, &f, hist, histCap);
    export_fields(m, &f, U, p, k, om, nuT, F1, F2, Pk, CDkw, Ub, pb, kb, omb);
    info[0] = H.finalIter; info[1] = H.converged; info[2] = H.diverged; info[3] = H.nEntries;
    info[4] = f.turbEstablished;
    fields_free(&f);
    solver_free(&S);
    return 0;
}

static void load_state(const om_mesh* m, fields_t* f, const double* U, const double* Ub,
                       const double* p, const double* pb, const double* k, const double* kb,
                       const double* om, const double* omb) {
    int n = m->nC, nB = m->nF - m->nIF;
    memcpy(f->U, U, sizeof(v3) * n); memcpy(f->Ub, Ub, sizeof(v3) * nB);
    memcpy(f->p, p, 8 * n); memcpy(f->pb, pb, 8 * nB);
    memcpy(f->k, k, 8 * n); memcpy(f->kb, kb, 8 * nB);
    memcpy(f->om, om, 8 * n); memcpy(f->omb, omb, 8 * nB);
}

/* Warm start (ForwardModel.cpp:56-60: fields = cached.fields; solve(fields)).
 * state: U[3n] Ub[3nB] p pb k kb omega omegab nuT F1 F2 Pk CDkw (13 arrays). */
int om_simple_solve_warm(const om_mesh* m, int bcKind, const double* bcVals, double nu,
                         const double* coeffs, const double* settings,
                         const double* const* state, int turbEstablished,
                         double* U, double* p, double* k, double* om, double* nuT,
                         double* F1, double* F2, double* Pk, double* CDkw,
                         double* Ub, double* pb, double* kb, double* omb,
                         double* hist, int histCap, int* info) {
    solver_t S;
    solver_init(&S, m, bcKind, bcVals, nu, coeffs, settings);
    fields_t f = fields_alloc(m);
    int n = m->nC;
    load_state(m, &f, state[0], state[1], state[2], state[3], state[4], state[5], state[6], state[7]);
    memcpy(f.nuT, state[8], 8 * n); memcpy(f.F1, state[9], 8 * n); memcpy(f.F2, state[10], 8 * n);
    memcpy(f.Pk, state[11], 8 * n); memcpy(f.CDkw, state[12], 8 * n);
    f.turbEstablished = turbEstablished != 0;
    hist_t H = simple_solve(&S, &f, hist, histCap);
    export_fields(m, &f, U, p, k, om, nuT, F1, F2, Pk, CDkw, Ub, pb, kb, omb);
    info[0] = H.finalIter; info[1] = H.converged; info[2] = H.diverged; info[3] = H.nEntries;
    info[4] = f.turbEstablished;
    fields_free(&f);
    solver_free(&S);
    return 0;
}

/* SIMPLESolver::recomputeClosure :1054-1064 */
static void recompute_closure(solver_t* S, fields_t* f, double* Smag) {
    compute_fields(S, f, Smag);
    for (int c = 0; c < S->m->nC; ++c) f->nuT[c] = smax(f->nuT[c], 0.0);
}

static void dump_ldu(const ldu* A, double* o) {
    memcpy(o, A->diag, 8 * A->n);
    memcpy(o + A->n, A->upper, 8 * A->nif);
    memcpy(o + A->n + A->nif, A->lower, 8 * A->nif);
    memcpy(o + A->n + 2 * A->nif, A->src, 8 * A->n);
}

int om_precon_blocks(const om_mesh* m, int bcKind, const double* bcVals, double nu,
                     const double* coeffs, const double* settings,
                     const double* U, const double* Ub, const double* p, const double* pb,
                     const double* k, const double* kb, const double* om, const double* omb,
                     double* mom, double* kout, double* omout, double* aP, double* closure) {
    solver_t S;
    solver_init(&S, m, bcKind, bcVals, nu, coeffs, settings);
    fields_t f = fields_alloc(m);
    load_state(m, &f, U, Ub, p, pb, k, kb, om, omb);
    int n = m->nC, nIF = m->nIF;
    double* Smag = calloc(n + 1, 8);
    recompute_closure(&S, &f, Smag);
    double* buf = calloc(2 * n + 2 * nIF + 8, 8);
    ldu A = {n, nIF, m->own, m->nbr, buf, buf + n, buf + n + nIF, buf + n + 2 * nIF};
    assemble_momentum(&S, &A, &f, 0, aP); dump_ldu(&A, mom);
    assemble_k(&S, &A, &f); dump_ldu(&A, kout);
    assemble_om(&S, &A, &f, Smag); dump_ldu(&A, omout);
    if (closure) {
        memcpy(closure, f.nuT, 8 * n); memcpy(closure + n, f.F1, 8 * n);
        memcpy(closure + 2 * n, f.F2, 8 * n); memcpy(closure + 3 * n, f.Pk, 8 * n);
        memcpy(closure + 4 * n, f.CDkw, 8 * n);
    }
    free(buf); free(Smag);
    fields_free(&f); solver_free(&S);
    return 0;
}

int om_residual(const om_mesh* m, int bcKind, const double* bcVals, double nu,
                const double* coeffs, const double* settings,
                const double* U, const double* Ub, const double* p, const double* pb,
                const double* k, const double* kb, const double* om, const double* omb,
                double* R) {
    solver_t S;
    solver_init(&S, m, bcKind, bcVals, nu, coeffs, settings);
    fields_t f = fields_alloc(m);
    load_state(m, &f, U, Ub, p, pb, k, kb, om, omb);
    int n = m->nC, nIF = m->nIF;
    double* Smag = calloc(n + 1, 8);
    recompute_closure(&S, &f, Smag);
    double* buf = calloc(2 * n + 2 * nIF + 8, 8);
    ldu A = {n, nIF, m->own, m->nbr, buf, buf + n, buf + n + nIF, buf + n + 2 * nIF};
    double* phi = calloc(n + 1, 8);
    double* aP = calloc(n + 1, 8);
    for (int blk = 0; blk < 4; ++blk) {
        if (blk < 2) {
            assemble_momentum(&S, &A, &f, blk, aP);
            for (int c = 0; c < n; ++c) phi[c] = comp_of(f.U[c], blk);
        } else if (blk == 2) {
            assemble_k(&S, &A, &f); memcpy(phi, f.k, 8 * n);
        } else {
            assemble_om(&S, &A, &f, Smag); memcpy(phi, f.om, 8 * n);
        }
        ldu_residual(&A, phi, R + (size_t)blk * n);
    }
    free(buf); free(Smag); free(phi); free(aP);
    fields_free(&f); solver_free(&S);
    return 0;
}

/* -------------------------------------------------------- observations */
/* ObservationOperator.hpp:159-364 */
static int nearest_cell(const om_mesh* m, v3 loc) {
    int best = 0; double bd = 1e30;
    for (int c = 0; c < m->nC; ++c) {
        v3 d = v3sub(m->cc[c], loc);
        double d2 = d.x * d.x + d.y * d.y + d.z * d.z;
        if (d2 < bd) { bd = d2; best = c; }
    }
    return best;
}

int om_observe(const om_mesh* m, int nObs, const int* type, const int* patch,
               const double* loc, const int* comp, const double* refArea,
               const double* refVel, const double* refP, const double* refRho,
               const double* Uin, const double* Ubin, const double* p, double nu,
               double* pred) {
    const v3* U = (const v3*)Uin;
    const v3* Ub = (const v3*)Ubin;
    for (int i = 0; i < nObs; ++i) {
        v3 L = v3make(loc[3 * i], loc[3 * i + 1], loc[3 * i + 2]);
        int q = patch[i];
        double out = 0.0;
        switch (type[i]) {
            case 0: { /* computeDrag :186-223 */
                double force = 0.0;
                for (int j = m->pfPtr[q]; j < m->pfPtr[q + 1]; ++j) {
                    int fi = m->pfIdx[j], ow = m->own[fi];
                    double Sf = m->area[fi];
                    double pf = p[ow];
                    force += pf * m->fn[fi].x * Sf;
                    double delta = smax(m->delta[fi], 1e-20);
                    double Uw_x = Ub[fi - m->nIF].x;
                    double tau_x = nu * (U[ow].x - Uw_x) / delta;
                    force += tau_x * Sf;
                }
                double dyn = 0.5 * refVel[i] * refVel[i];
                out = force / (dyn * refArea[i]);
                break;
            }
            case 1: { /* computePressureAt :261-271 */
                int c = nearest_cell(m, L);
                double dyn = 0.5 * refRho[i] * refVel[i] * refVel[i];
                out = (p[c] - refP[i]) / smax(dyn, 1e-20);
                break;
            }
            case 2: { /* computeVelocityAt :251-258 */
                int c = nearest_cell(m, L);
                out = comp[i] == 0 ? U[c].x : comp[i] == 1 ? U[c].y : U[c].z;
                break;
            }
            case 3: { /* computeSkinFrictionAt :226-248 */
                int best = m->pfPtr[q + 1] > m->pfPtr[q] ? m->pfIdx[m->pfPtr[q]] : -1;
                double bd = 1e30;
                for (int j = m->pfPtr[q]; j < m->pfPtr[q + 1]; ++j) {
                    int fi = m->pfIdx[j];
                    v3 d = v3sub(m->fc[fi], L);
                    double d2 = d.x * d.x + d.y * d.y + d.z * d.z;
                    if (d2 < bd) { bd = d2; best = fi; }
                }
                if (best < 0) { out = 0.0; break; }
                int ow = m->own[best];
                double delta = smax(m->delta[best], 1e-20);
                v3 Uc = U[ow];
                v3 Un = v3mul(m->fn[best], v3dot(Uc, m->fn[best]));
                v3 Ut = v3sub(Uc, Un);
                double tau = nu * v3norm(Ut) / delta;
                double dynP = 0.5 * refVel[i] * refVel[i];
                out = tau / smax(dynP, 1e-20);
                break;
            }
            case 4: /* computeSeparationPoint :277-307 */
            case 5: { /* computeReattachmentLength :311-340 */
                int sep = type[i] == 4;
                int havePrev = 0;
                double prevTau = sep ? 1.0 : -1.0, prevX = 0.0;
                out = sep ? 0.0 : (m->pfPtr[q + 1] > m->pfPtr[q] ? m->fc[m->pfIdx[m->pfPtr[q + 1] - 1]].x : 0.0);
                for (int j = m->pfPtr[q]; j < m->pfPtr[q + 1]; ++j) {
                    int fi = m->pfIdx[j], ow = m->own[fi];
                    double delta = smax(m->delta[fi], 1e-20);
                    v3 Uc = U[ow];
                    v3 Ut = v3sub(Uc, v3mul(m->fn[fi], v3dot(Uc, m->fn[fi])));
                    double tau = nu * Ut.x / delta;
                    double x = m->fc[fi].x;
                    if (sep && havePrev && prevTau > 0 && tau <= 0) {
                        double denom = prevTau - tau;
                        double frac = (fabs(denom) > 1e-30) ? prevTau / denom : 0.0;
                        out = prevX + frac * (x - prevX); break;
                    }
                    if (!sep && havePrev && prevTau < 0 && tau >= 0) {
                        double denom = tau - prevTau;
                        double frac = (fabs(denom) > 1e-30) ? (-prevTau) / denom : 0.0;
                        out = prevX + frac * (x - prevX); break;
                    }
                    prevTau = tau; prevX = x; havePrev = 1;
                }
                break;
            }
            default: out = 0.0;
        }
        pred[i] = out;
    }
    return 0;
}

/* ObservationOperator.hpp:364-378 (Measurement::totalVariance, :16) */
double om_loglik(int nObs, const double* pred, const double* observed, const double* sigma,
                 const double* sigmaModel) {
    double ll = 0.0;
    for (int i = 0; i < nObs; ++i) {
        double var = sigma[i] * sigma[i] + sigmaModel[i] * sigmaModel[i];
        if (var < 1e-30) var = 1e-30;
        double diff = pred[i] - observed[i];
        ll += -0.5 * diff * diff / var;
    }
    return ll;
}
