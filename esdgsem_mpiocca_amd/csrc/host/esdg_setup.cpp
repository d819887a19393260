// esdg_setup.cpp -- host-side FP64 problem setup (basis, meshes, geometry, bathymetry, states).
//
// The operator and geometry arrays here are the constant inputs of the per-stage kernels.  They
// follow the reference's own setup arithmetic operation by operation, so that the same mesh and
// degree give the same bits as the reference host code:
//   LGL/Gauss nodes, weights, Legendre recursion   src/basis.cpp:566-776
//   barycentric weights, derivative matrix         src/basis.cpp:826-868
//   flux-differencing D, Dstrong, Dhat             src/basis.cpp:108-133
//   modal transform VdmInv                         src/basis.cpp:495-564, 778-824
//   element sizes                                  src/basis.cpp:939-966
//   Cartesian mesh and connectivity                src/Mesh.cpp:316-696, 92-113, 282-307
//   mapped geometry (metrics, normals, scal)       src/Mesh.cpp:1667-1808
//   single-rank EdgeData layout                    src/MeshPartitioning.cpp:65-132, main.cpp:650-716
//   split-form bathymetry gradient Bx, By          src/SW2D.cpp:1048-1105
//   RK coefficients                                src/RungeKutta.cpp:18-90
// The initial states are new (they define the benchmark configurations) and are documented in
// esdg_case_state.h, which the device case builder shares.
#include "esdg_setup.h"
#include "esdg_output.h"
#include "../esdg_case_state.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <exception>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(const std::string& msg, int code = -1) {
    g_err = msg;
    return code;
}

// ------------------------------------------------------------------------------------------
// Legendre polynomials and the LGL / Gauss point sets
// ------------------------------------------------------------------------------------------

// L_n(x) and L_n'(x) by the three-term recursion (basis.cpp:566-602).
void legendre(int n, double x, double* L, double* dL) {
    if (n == 0) { *L = 1.0; *dL = 0.0; return; }
    if (n == 1) { *L = x; *dL = 1.0; return; }
    double Lm2 = 1.0, Lm1 = x, dLm2 = 0.0, dLm1 = 1.0, Ln = 0.0, dLn = 0.0;
    for (int k = 2; k <= n; k++) {
        Ln = ((2.0 * k - 1.0) / k) * x * Lm1 - (k - 1.0) / k * Lm2;
        dLn = dLm2 + (2.0 * k - 1.0) * Lm1;
        Lm2 = Lm1; Lm1 = Ln; dLm2 = dLm1; dLm1 = dLn;
    }
    *L = Ln; *dL = dLn;
}

// q = L_{N+1} - L_{N-1}, q' and L_N at x: the LGL interior nodes are the roots of q
// (basis.cpp:623-653).
void q_and_L(int N, double x, double* q, double* dq, double* LN) {
    double Lm2 = 1.0, Lm1 = x, dLm2 = 0.0, dLm1 = 1.0, dLN = 0.0, Ln = 0.0;
    for (int k = 2; k <= N; k++) {
        Ln = (2.0 * k - 1.0) / k * x * Lm1 - (k - 1.0) / k * Lm2;
        dLN = dLm2 + (2.0 * k - 1.0) * Lm1;
        if (k < N) { Lm2 = Lm1; Lm1 = Ln; dLm2 = dLm1; dLm1 = dLN; }
    }
    const int k = N + 1;
    const double Lp1 = (2.0 * k - 1.0) / k * x * Ln - (k - 1.0) / k * Lm1;
    const double dLp1 = dLm1 + (2.0 * k - 1.0) * Ln;
    *q = Lp1 - Lm1;
    *dq = dLp1 - dLm1;
    *LN = Ln;
    (void)dLN;
}

struct Basis {
    int N = 0, ngl = 0;
    std::vector<double> xGL, wGL, xG, wG, wbary, D, D0, Dstrong, Dhat, Vdm, VdmInv;
};

void lgl_nodes(Basis& B) {
    const int N = B.N, ngl = B.ngl;
    const int newton_it = 4;
    const double TOL = 4. * std::pow(10., -16);
    double q = 0, dq = 0, LN = 0;
    if (N == 1) {
        B.xGL[0] = -1.0; B.wGL[0] = 1.0; B.xGL[1] = 1.0; B.wGL[1] = B.wGL[0];
    } else if (N == 2) {
        B.xGL[0] = -1.0; B.wGL[0] = 2.0 / (N * (N + 1.0));
        B.xGL[ngl - 1] = 1.0; B.wGL[ngl - 1] = B.wGL[0];
        q_and_L(N, 0.0, &q, &dq, &LN);
        B.xGL[1] = 0.0; B.wGL[1] = 2.0 / (N * (N + 1) * std::pow(LN, 2));
    } else {
        B.xGL[0] = -1.0; B.wGL[0] = 2.0 / (N * (N + 1.0));
        B.xGL[ngl - 1] = 1.0; B.wGL[ngl - 1] = B.wGL[0];
        for (int j = 1; j <= (int)std::floor(N / 2.0); j++) {
            double x = -std::cos((j + 1.0 / 4.0) * M_PI / N - 3.0 / (8.0 * N * M_PI) * 1.0 / (j + 1.0 / 4.0));
            for (int k = 0; k <= newton_it; k++) {  // Newton on q
                q_and_L(N, x, &q, &dq, &LN);
                const double delta = -q / dq;
                x = x + delta;
                if (std::fabs(delta) <= TOL * std::fabs(x)) break;
            }
            q_and_L(N, x, &q, &dq, &LN);
            B.xGL[j] = x;
            B.xGL[ngl - j - 1] = -x;
            B.wGL[j] = 2.0 / (N * (N + 1) * std::pow(LN, 2));
            B.wGL[ngl - j - 1] = B.wGL[j];
        }
        if (N % 2 == 0) {
            q_and_L(N, 0.0, &q, &dq, &LN);
            B.xGL[N / 2] = 0.0;
            B.wGL[N / 2] = 2.0 / (N * (N + 1) * std::pow(LN, 2));
        }
    }
}

void gauss_nodes(Basis& B) {
    const int N = B.N, ngl = B.ngl;
    const int newton_it = 4;
    const double TOL = 4. * std::pow(10., -16);
    double L1 = 0, L2 = 0;
    if (N == 0) { B.xG[0] = 0.0; B.wG[0] = 2.0; return; }
    if (N == 1) {
        B.xG[0] = -std::sqrt(1.0 / 3.0); B.wG[0] = 1.0; B.xG[1] = -B.xG[0]; B.wG[1] = B.wG[0];
        return;
    }
    for (int j = 0; j < (int)std::floor((N + 1.0) / 2.0); j++) {
        double x = -std::cos((2.0 * j + 1.0) / (2.0 * N + 2.0) * M_PI);
        for (int k = 0; k <= newton_it; k++) {
            legendre(N + 1, x, &L1, &L2);
            const double delta = -L1 / L2;
            x = x + delta;
            if (std::fabs(delta) <= TOL * std::fabs(x)) break;
        }
        legendre(N + 1, x, &L1, &L2);
        B.xG[j] = x;
        B.xG[ngl - j - 1] = -x;
        B.wG[j] = 2.0 / ((1.0 - std::pow(x, 2)) * std::pow(L2, 2));
        B.wG[ngl - j - 1] = B.wG[j];
    }
    if (N % 2 == 0) {
        legendre(N + 1, 0.0, &L1, &L2);
        B.xG[N / 2] = 0.0;
        B.wG[N / 2] = 2.0 / std::pow(L2, 2);
    }
}

// Lagrange basis through nodes x (barycentric weights w) evaluated at x0 (basis.cpp:778-809).
void lagrange_at(double x0, int N, const double* x, const double* w, double* P) {
    const double TOL = 4. * std::pow(10., -16);
    bool match = false;
    for (int j = 0; j < N + 1; j++) {
        P[j] = 0.0;
        if (std::fabs(x0 - x[j]) < TOL) { match = true; P[j] = 1; }
    }
    if (!match) {
        double s = 0.0;
        for (int j = 0; j < N + 1; j++) {
            const double t = w[j] / (x0 - x[j]);
            P[j] = t;
            s = s + t;
        }
        for (int j = 0; j < N + 1; j++) P[j] = P[j] / s;
    }
}

Basis make_basis(int N, int flux_differencing) {
    if (N < 1 || N > 15) throw std::invalid_argument("polynomial degree N must be in [1,15]");
    Basis B;
    B.N = N; B.ngl = N + 1;
    const int ngl = B.ngl, ngl2 = ngl * ngl;
    B.xGL.assign(ngl, 0.0); B.wGL.assign(ngl, 0.0); B.xG.assign(ngl, 0.0); B.wG.assign(ngl, 0.0);
    B.wbary.assign(ngl, 1.0);
    B.D.assign(ngl2, 0.0); B.D0.assign(ngl2, 0.0); B.Dstrong.assign(ngl2, 0.0); B.Dhat.assign(ngl2, 0.0);
    B.Vdm.assign(ngl2, 0.0); B.VdmInv.assign(ngl2, 0.0);
    lgl_nodes(B);
    gauss_nodes(B);
    // barycentric weights (basis.cpp:826-849)
    for (int j = 1; j < ngl; j++)
        for (int k = 0; k < j; k++) {
            B.wbary[k] = B.wbary[k] * (B.xGL[k] - B.xGL[j]);
            B.wbary[j] = B.wbary[j] * (B.xGL[j] - B.xGL[k]);
        }
    for (int j = 0; j < ngl; j++) B.wbary[j] = 1.0 / B.wbary[j];
    // derivative matrix, diagonal by the negative-sum trick in column order (basis.cpp:851-868)
    for (int i = 0; i < ngl; i++)
        for (int j = 0; j < ngl; j++)
            if (i != j) {
                const double dij = (B.wbary[j] / B.wbary[i]) * (1.0 / (B.xGL[i] - B.xGL[j]));
                B.D[i * ngl + j] = dij;
                B.D[i * ngl + i] = B.D[i * ngl + i] - dij;
            }
    for (int i = 0; i < ngl2; i++) {
        B.D0[i] = B.D[i];
        B.Dstrong[i] = B.D[i];
        if (flux_differencing) B.D[i] = 2.0 * B.D[i];
    }
    for (int i = 0; i < ngl; i++)
        for (int j = 0; j < ngl; j++)
            B.Dhat[i * ngl + j] = -B.wGL[j] * B.D0[j * ngl + i] / B.wGL[i];
    B.D[0] = B.D[0] + 1.0 / B.wGL[0];
    B.D[ngl2 - 1] = B.D[ngl2 - 1] - 1.0 / B.wGL[ngl - 1];
    B.Dstrong[0] = B.D0[0] + 1.0 / B.wGL[0];
    B.Dstrong[ngl2 - 1] = B.D0[ngl2 - 1] - 1.0 / B.wGL[ngl - 1];
    // modal transform: L2 projection onto orthonormal Legendre modes by Gauss quadrature
    std::vector<double> vt(ngl2, 0.0), interp(ngl2, 0.0), row(ngl, 0.0);
    for (int i = 0; i < ngl; i++)
        for (int j = 0; j < ngl; j++) {
            double L, dL;
            legendre(j, B.xG[i], &L, &dL);
            vt[j * ngl + i] = L * std::sqrt(j + 0.5) * B.wG[i];
            legendre(j, B.xGL[i], &L, &dL);
            B.Vdm[i * ngl + j] = L * std::sqrt(j + 0.5);
        }
    for (int i = 0; i <= N; i++) {
        lagrange_at(B.xG[i], N, B.xGL.data(), B.wbary.data(), row.data());
        for (int j = 0; j < ngl; j++) interp[i * ngl + j] = row[j];
    }
    for (int i = 0; i < ngl; i++)
        for (int j = 0; j < ngl; j++)
            for (int l = 0; l < ngl; l++)
                B.VdmInv[i * ngl + j] += vt[i * ngl + l] * interp[l * ngl + j];
    return B;
}

// ------------------------------------------------------------------------------------------
// Meshes
// ------------------------------------------------------------------------------------------

struct Geometry {
    int ngl = 0, Nelem = 0, Nfaces = 0;
    std::vector<double> x, y, xXi, xEta, yXi, yEta, J;       // per node; J is the Jacobian (not inverted)
    std::vector<double> enx, eny, escal;                     // per element side node [e][side][k]
    std::vector<int> edge_info;                              // reference Mesh::EdgeInfo [f][7]
    std::vector<int> elem_to_edge;                           // [e][4], 1-based
    std::vector<double> b;                                   // bathymetry from the mesh file (read_bottom)
};

// Metric terms by the nodal derivative matrix and J (Mesh.cpp:1722-1742) of one element's nodes.
void element_metrics(int ngl, const double* D, const double* xe, const double* ye, double* xxi, double* xeta,
                     double* yxi, double* yeta, double* Je) {
    for (int j = 0; j < ngl; j++)
        for (int i = 0; i < ngl; i++) {
            const int ij = j * ngl + i;
            double ax = 0.0, bx = 0.0, ay = 0.0, by = 0.0;
            for (int k = 0; k < ngl; k++) {
                ax += D[i * ngl + k] * xe[j * ngl + k];
                bx += D[j * ngl + k] * xe[k * ngl + i];
                ay += D[i * ngl + k] * ye[j * ngl + k];
                by += D[j * ngl + k] * ye[k * ngl + i];
            }
            xxi[ij] = ax; xeta[ij] = bx; yxi[ij] = ay; yeta[ij] = by;
            Je[ij] = xxi[ij] * yeta[ij] - xeta[ij] * yxi[ij];
        }
}

// Outward unit normals and surface Jacobians of the four sides (Mesh.cpp:1744-1806).
void side_normals(int ngl, const double* xxi, const double* xeta, const double* yxi, const double* yeta,
                  const double* Je, double* nxs, double* nys, double* scs) {
    for (int j = 0; j < ngl; j++) {
        const int s1 = ngl + j, s3 = 3 * ngl + j;
        const int Nj = j * ngl + ngl - 1, Zj = j * ngl;
        double ye_ = yeta[Nj], xe_ = xeta[Nj], Jt = Je[Nj];
        scs[s1] = std::sqrt(ye_ * ye_ + xe_ * xe_);
        nxs[s1] = std::copysign(1.0, Jt) * (ye_ / scs[s1]);
        nys[s1] = std::copysign(1.0, Jt) * (-xe_ / scs[s1]);
        ye_ = yeta[Zj]; xe_ = xeta[Zj]; Jt = Je[Zj];
        scs[s3] = std::sqrt(ye_ * ye_ + xe_ * xe_);
        nxs[s3] = -std::copysign(1.0, Jt) * (ye_ / scs[s3]);
        nys[s3] = -std::copysign(1.0, Jt) * (-xe_ / scs[s3]);
    }
    for (int i = 0; i < ngl; i++) {
        const int s0 = i, s2 = 2 * ngl + i;
        const int iZ = i, iN = (ngl - 1) * ngl + i;
        double yx = yxi[iZ], xx = xxi[iZ], Jt = Je[iZ];
        scs[s0] = std::sqrt(yx * yx + xx * xx);
        nxs[s0] = -std::copysign(1.0, Jt) * (-yx / scs[s0]);
        nys[s0] = -std::copysign(1.0, Jt) * (xx / scs[s0]);
        yx = yxi[iN]; xx = xxi[iN]; Jt = Je[iN];
        scs[s2] = std::sqrt(yx * yx + xx * xx);
        nxs[s2] = std::copysign(1.0, Jt) * (-yx / scs[s2]);
        nys[s2] = std::copysign(1.0, Jt) * (xx / scs[s2]);
    }
}

// ---- ISM-V2 meshes (Mesh::ReadMesh, src/Mesh.cpp:698-1503) --------------------------------
// Boundary curves are given, or built from their end corners, at the Chebyshev-Lobatto points
// -cos(k pi / M) of the file's curve order M and evaluated by barycentric Lagrange interpolation
// (Mesh.cpp:1832-1858, 1946-1977); an element with a given curve is mapped by the transfinite
// (Coons) map of its four curves, a straight-sided one bilinearly (Mesh.cpp:1506-1665).
struct CurveBasis {
    int M = 0;
    std::vector<double> xc, wb;
    explicit CurveBasis(int order) : M(order), xc(order + 1), wb(order + 1, 1.0) {
        for (int k = 0; k <= M; ++k) xc[k] = -std::cos(k * M_PI / M);
        for (int j = 1; j <= M; j++)
            for (int k = 0; k < j; k++) {
                wb[k] = wb[k] * (xc[k] - xc[j]);
                wb[j] = wb[j] * (xc[j] - xc[k]);
            }
        for (int j = 0; j <= M; j++) wb[j] = 1.0 / wb[j];
    }
    // barycentric form 2; a point within 1e-12 of a node takes that node's value
    double at(const double* f, double xpt) const {
        bool hit = false;
        double out = 0.0;
        for (int j = 0; j <= M; j++)
            if (std::fabs(xpt - xc[j]) < std::pow(10.0, -12)) { hit = true; out = f[j]; }
        if (hit) return out;
        double num = 0.0, den = 0.0;
        for (int j = 0; j <= M; j++) {
            const double t = wb[j] / (xpt - xc[j]);
            num += t * f[j];
            den += t;
        }
        return num / den;
    }
    // Coons map of the four curves (sides 1, 2: xi = -1..1 along eta = -1 / xi = 1; 3, 4 opposite)
    double transfinite(const double* g1, const double* g2, const double* g3, const double* g4, double psi,
                       double eta) const {
        const double x1_ref = at(g1, -1.0), x2_ref = at(g1, 1.0), x3_ref = at(g3, 1.0), x4_ref = at(g3, -1.0);
        const double x1_cp = at(g1, psi), x2_cp = at(g2, eta), x3_cp = at(g3, psi), x4_cp = at(g4, eta);
        return 0.5 * ((1.0 - psi) * x4_cp + (1.0 + psi) * x2_cp + (1.0 - eta) * x1_cp + (1.0 + eta) * x3_cp) -
               0.25 * ((1.0 - psi) * ((1.0 - eta) * x1_ref + (1.0 + eta) * x4_ref) +
                       (1.0 + psi) * ((1.0 - eta) * x2_ref + (1.0 + eta) * x3_ref));
    }
};

inline double bilinear(const double* c, double psi, double eta) {
    return 0.25 * (c[0] * (1.0 - psi) * (1.0 - eta) + c[1] * (1.0 + psi) * (1.0 - eta) +
                   c[2] * (1.0 + psi) * (1.0 + eta) + c[3] * (1.0 - psi) * (1.0 + eta));
}

Geometry read_ism(const esdg_case_spec& s, const Basis& B) {
    if (!s.mesh_file) throw std::invalid_argument("ISM mesh kind without a mesh file");
    std::ifstream in(s.mesh_file);
    if (!in) throw std::invalid_argument(std::string("mesh file not found: ") + s.mesh_file);
    const bool rb = s.read_bottom != 0;
    // ConvertToKM (Mesh.cpp:770-776): longitude scaled, latitude by the Mercator projection
    const double earth_radius = 6.378e3;
    auto to_km = [&](double& x, double& y) {
        if (!s.convert_to_km) return;
        x = x * earth_radius * M_PI / 180.0;
        y = earth_radius * std::asinh(std::tan(y / 180.0 * M_PI));
    };
    std::string line;
    auto next = [&](const char* what) -> std::istringstream {
        if (!std::getline(in, line)) throw std::invalid_argument(std::string("ISM mesh: unexpected end of file in ") + what);
        return std::istringstream(line);
    };
    next("format");
    int nn = 0, ne = 0, nel = 0, order = 0;
    {
        auto ls = next("header");
        if (!(ls >> nn >> ne >> nel >> order) || nn < 1 || ne < 1 || nel < 1 || order < 1)
            throw std::invalid_argument("ISM mesh: bad header");
    }
    const int ngl = B.ngl, ngl2 = ngl * ngl;
    if ((int64_t)nel * ngl2 * 3 > (int64_t)INT32_MAX) throw std::invalid_argument("mesh too large for 32-bit node indices");
    const CurveBasis CB(order);
    std::vector<double> xn(nn), yn(nn), bn(nn, 0.0);
    for (int i = 0; i < nn; ++i) {
        auto ls = next("nodes");
        if (!(ls >> xn[i] >> yn[i]) || (rb && !(ls >> bn[i]))) throw std::invalid_argument("ISM mesh: bad node line");
        to_km(xn[i], yn[i]);
    }
    Geometry G;
    G.ngl = ngl; G.Nelem = nel; G.Nfaces = ne;
    G.edge_info.assign((size_t)ne * 7, 0);
    G.elem_to_edge.assign((size_t)nel * 4, 0);
    for (int is = 0; is < ne; ++is) {
        auto ls = next("edges");
        int n0, n1, eL, eR, sL, sR;
        if (!(ls >> n0 >> n1 >> eL >> eR >> sL >> sR)) throw std::invalid_argument("ISM mesh: bad edge line");
        if (eL < 1 || eL > nel || eR < 0 || eR > nel || sL < 1 || sL > 4 || std::abs(sR) > 4)
            throw std::invalid_argument("ISM mesh: edge references a missing element or side");
        int* E = &G.edge_info[(size_t)is * 7];
        E[0] = n0 - 1; E[1] = n1 - 1; E[2] = eL - 1; E[3] = eR - 1; E[4] = sL - 1;
        G.elem_to_edge[(size_t)(eL - 1) * 4 + sL - 1] = is + 1;
        // a negative right side marks a reversed node order on the right element
        E[5] = (sR < 0) ? -sR - 1 : sR - 1;
        E[6] = (sR < 0) ? 0 : 1;
        if (E[3] != -1) G.elem_to_edge[(size_t)(eR - 1) * 4 + E[5]] = is + 1;
    }
    const size_t nnodes = (size_t)nel * ngl2;
    G.x.assign(nnodes, 0.0); G.y.assign(nnodes, 0.0);
    G.xXi.assign(nnodes, 0.0); G.xEta.assign(nnodes, 0.0); G.yXi.assign(nnodes, 0.0); G.yEta.assign(nnodes, 0.0);
    G.J.assign(nnodes, 0.0);
    if (rb) G.b.assign(nnodes, 0.0);
    G.enx.assign((size_t)nel * 4 * ngl, 0.0);
    G.eny.assign((size_t)nel * 4 * ngl, 0.0);
    G.escal.assign((size_t)nel * 4 * ngl, 0.0);
    const int M = order;
    std::vector<double> gx(4 * (M + 1)), gy(4 * (M + 1)), gb(4 * (M + 1));
    std::vector<double> xe(ngl2), ye(ngl2), be(ngl2), xxi(ngl2), xeta(ngl2), yxi(ngl2), yeta(ngl2), Je(ngl2);
    const double* D = B.D0.data();
    // side j runs from corner first[j] to corner last[j] (Mesh.cpp:1014-1031)
    const int first[4] = {0, 1, 3, 0}, last[4] = {1, 2, 2, 3};
    for (int ie = 0; ie < nel; ++ie) {
        int c[4], given[4];
        {
            auto ls = next("element corners");
            if (!(ls >> c[0] >> c[1] >> c[2] >> c[3])) throw std::invalid_argument("ISM mesh: bad element corners");
            for (int j = 0; j < 4; ++j) {
                c[j] -= 1;
                if (c[j] < 0 || c[j] >= nn) throw std::invalid_argument("ISM mesh: corner references a missing node");
            }
        }
        {
            auto ls = next("curve flags");
            if (!(ls >> given[0] >> given[1] >> given[2] >> given[3])) throw std::invalid_argument("ISM mesh: bad curve flags");
        }
        const bool curved = given[0] || given[1] || given[2] || given[3];
        double cx[4], cy[4], cb[4];
        for (int j = 0; j < 4; ++j) {
            cx[j] = xn[c[j]]; cy[j] = yn[c[j]]; cb[j] = bn[c[j]];
            double* Gx = &gx[j * (M + 1)];
            double* Gy = &gy[j * (M + 1)];
            double* Gb = &gb[j * (M + 1)];
            if (!given[j]) {
                const int a = c[first[j]], z = c[last[j]];
                for (int k = 0; k <= M; ++k) {
                    Gx[k] = ((xn[z] - xn[a]) * CB.xc[k] + xn[a] + xn[z]) / 2.0;
                    Gy[k] = ((yn[z] - yn[a]) * CB.xc[k] + yn[a] + yn[z]) / 2.0;
                    if (rb) Gb[k] = ((bn[z] - bn[a]) * CB.xc[k] + bn[a] + bn[z]) / 2.0;
                }
            } else {
                for (int k = 0; k <= M; ++k) {
                    auto ls = next("boundary curve");
                    if (!(ls >> Gx[k] >> Gy[k]) || (rb && !(ls >> Gb[k])))
                        throw std::invalid_argument("ISM mesh: bad boundary curve point");
                    to_km(Gx[k], Gy[k]);
                }
            }
        }
        next("boundary names");
        const double *g1x = &gx[0], *g2x = &gx[M + 1], *g3x = &gx[2 * (M + 1)], *g4x = &gx[3 * (M + 1)];
        const double *g1y = &gy[0], *g2y = &gy[M + 1], *g3y = &gy[2 * (M + 1)], *g4y = &gy[3 * (M + 1)];
        const double *g1b = &gb[0], *g2b = &gb[M + 1], *g3b = &gb[2 * (M + 1)], *g4b = &gb[3 * (M + 1)];
        for (int j = 0; j < ngl; j++)
            for (int i = 0; i < ngl; i++) {
                const int ij = j * ngl + i;
                const double psi = B.xGL[i], eta = B.xGL[j];
                if (curved) {
                    xe[ij] = CB.transfinite(g1x, g2x, g3x, g4x, psi, eta);
                    ye[ij] = CB.transfinite(g1y, g2y, g3y, g4y, psi, eta);
                    if (rb) be[ij] = CB.transfinite(g1b, g2b, g3b, g4b, psi, eta);
                } else {
                    xe[ij] = bilinear(cx, psi, eta);
                    ye[ij] = bilinear(cy, psi, eta);
                    if (rb) be[ij] = bilinear(cb, psi, eta);
                }
            }
        element_metrics(ngl, D, xe.data(), ye.data(), xxi.data(), xeta.data(), yxi.data(), yeta.data(), Je.data());
        const size_t base = (size_t)ie * ngl2;
        for (int k = 0; k < ngl2; ++k) {
            G.x[base + k] = xe[k]; G.y[base + k] = ye[k];
            G.xXi[base + k] = xxi[k]; G.xEta[base + k] = xeta[k];
            G.yXi[base + k] = yxi[k]; G.yEta[base + k] = yeta[k];
            G.J[base + k] = Je[k];
            if (rb) G.b[base + k] = be[k];
        }
        side_normals(ngl, xxi.data(), xeta.data(), yxi.data(), yeta.data(), Je.data(), &G.enx[(size_t)ie * 4 * ngl],
                     &G.eny[(size_t)ie * 4 * ngl], &G.escal[(size_t)ie * 4 * ngl]);
    }
    for (int ie = 0; ie < nel; ++ie)
        for (int j = 0; j < 4; ++j)
            if (G.elem_to_edge[(size_t)ie * 4 + j] < 1) throw std::invalid_argument("ISM mesh: element side without an edge");
    return G;
}

// Structured quad topology and node placement of Mesh::GenerateMesh (Mesh.cpp:316-696).  For
// the Cartesian kind the metrics, J and side normals are the generator's constants; for the
// curved kind the nodes are displaced and the metrics/normals are recomputed the way
// Mesh::ConstructMappedGeometry does.
Geometry make_mesh(const esdg_case_spec& s, const Basis& B) {
    if (s.mesh_kind == ESDG_MESH_ISM) return read_ism(s, B);
    Geometry G;
    const int ngl = B.ngl, ngl2 = ngl * ngl;
    const int NelemX = s.nelx, NelemY = s.nely;
    if (NelemX < 1 || NelemY < 1) throw std::invalid_argument("nelx and nely must be >= 1");
    if ((int64_t)NelemX * NelemY * ngl2 * 3 > (int64_t)INT32_MAX)
        throw std::invalid_argument("mesh too large for 32-bit node indices");
    const bool PX = s.periodic_x != 0, PY = s.periodic_y != 0;
    G.ngl = ngl;
    G.Nelem = NelemX * NelemY;
    int nedges = (NelemX + 1) * NelemY + (NelemY + 1) * NelemX;
    if (PX) nedges -= NelemX;
    if (PY) nedges -= NelemY;
    G.Nfaces = nedges;
    const size_t nn = (size_t)G.Nelem * ngl2;
    G.x.assign(nn, 0.0); G.y.assign(nn, 0.0);
    G.xXi.assign(nn, 0.0); G.xEta.assign(nn, 0.0); G.yXi.assign(nn, 0.0); G.yEta.assign(nn, 0.0);
    G.J.assign(nn, 0.0);
    G.enx.assign((size_t)G.Nelem * 4 * ngl, 0.0);
    G.eny.assign((size_t)G.Nelem * 4 * ngl, 0.0);
    G.escal.assign((size_t)G.Nelem * 4 * ngl, 0.0);
    G.edge_info.assign((size_t)nedges * 7, 0);
    G.elem_to_edge.assign((size_t)G.Nelem * 4, 0);

    const double deltaX = (s.xR - s.xL) / NelemX;
    const double deltaY = (s.yR - s.yL) / NelemY;
    const double x_xi = 0.5 * deltaX;
    const double y_eta = 0.5 * deltaY;
    const double Jac = x_xi * y_eta;
    for (int ieY = 0; ieY < NelemY; ++ieY)
        for (int ieX = 0; ieX < NelemX; ++ieX)
            for (int j = 0; j < ngl; ++j)
                for (int i = 0; i < ngl; ++i) {
                    const size_t id = ((size_t)ieY * NelemX + ieX) * ngl2 + j * ngl + i;
                    G.x[id] = s.xL + (B.xGL[i] + 1.0) / 2.0 * deltaX + ieX * deltaX;
                    G.y[id] = s.yL + (B.xGL[j] + 1.0) / 2.0 * deltaY + ieY * deltaY;
                }
    for (size_t id = 0; id < nn; ++id) {
        G.xXi[id] = x_xi; G.xEta[id] = 0.0; G.yXi[id] = 0.0; G.yEta[id] = y_eta; G.J[id] = Jac;
    }

    // Face numbering: horizontal faces first, then vertical ones (Mesh.cpp:412-655).
    auto side_set = [&](int e, int side, double nxv, double nyv, double sc) {
        for (int k = 0; k < ngl; ++k) {
            const size_t id = (size_t)e * 4 * ngl + side * ngl + k;
            G.enx[id] = nxv; G.eny[id] = nyv; G.escal[id] = sc;
        }
    };
    auto finish_edge = [&](int is) {
        int* E = &G.edge_info[(size_t)(is - 1) * 7];
        E[0] = 0; E[1] = 0; E[6] = 1;
        G.elem_to_edge[(size_t)E[2] * 4 + E[4]] = is;
        if (E[3] != -1) G.elem_to_edge[(size_t)E[3] * 4 + E[5]] = is;
    };
    int jLoopBound = NelemY;
    if (PX) jLoopBound -= 1;
    for (int j = 0; j <= jLoopBound; j++)
        for (int i = 0; i < NelemX; i++) {
            const int is = (j * NelemX + i) + 1;
            int* E = &G.edge_info[(size_t)(is - 1) * 7];
            if (j == 0) {
                E[2] = i; E[4] = 0; E[3] = -1; E[5] = -1;
                side_set(E[2], E[4], 0.0, -1.0, x_xi);
                if (PX) {
                    E[3] = (NelemY - 1) * NelemX + i; E[5] = 2;
                    side_set(E[3], E[5], 0.0, 1.0, x_xi);
                }
            } else if (j == NelemY) {
                if (!PX) {
                    E[2] = (NelemY - 1) * NelemX + i; E[4] = 2; E[3] = -1; E[5] = -1;
                    side_set(E[2], E[4], 0.0, 1.0, x_xi);
                }
            } else {
                E[2] = (j - 1) * NelemX + i; E[3] = j * NelemX + i; E[4] = 2; E[5] = 0;
                side_set(E[2], E[4], 0.0, 1.0, x_xi);
                side_set(E[3], E[5], 0.0, -1.0, x_xi);
            }
            finish_edge(is);
        }
    int iLoopBound = NelemX;
    if (PY) iLoopBound -= 1;
    for (int j = 0; j < NelemY; j++)
        for (int i = 0; i <= iLoopBound; i++) {
            const int is = ((jLoopBound + 1) * NelemX) + (j * (iLoopBound + 1) + i) + 1;
            int* E = &G.edge_info[(size_t)(is - 1) * 7];
            if (i == 0) {
                E[2] = j * NelemX; E[4] = 3; E[3] = -1; E[5] = -1;
                side_set(E[2], E[4], -1.0, 0.0, y_eta);
                if (PY) {
                    E[3] = (j + 1) * NelemX - 1; E[5] = 1;
                    side_set(E[3], E[5], 1.0, 0.0, y_eta);
                }
            } else if (i == NelemX) {
                if (!PY) {
                    E[2] = (j + 1) * NelemX - 1; E[4] = 1; E[3] = -1; E[5] = -1;
                    side_set(E[2], E[4], 1.0, 0.0, y_eta);
                }
            } else {
                E[2] = j * NelemX + i - 1; E[3] = j * NelemX + i; E[4] = 1; E[5] = 3;
                side_set(E[2], E[4], 1.0, 0.0, y_eta);
                side_set(E[3], E[5], -1.0, 0.0, y_eta);
            }
            finish_edge(is);
        }

    if (s.mesh_kind == ESDG_MESH_CURVED_SINE) {
        if (PX || PY) throw std::invalid_argument("the curved mesh kind is not periodic");
        const double a = s.perturb_amp, w = s.perturb_len;
        if (!(w > 0.0)) throw std::invalid_argument("perturb_len must be > 0");
        std::vector<double> xe(ngl2), ye(ngl2), xxi(ngl2), xeta(ngl2), yxi(ngl2), yeta(ngl2), Je(ngl2);
        const double* D = B.D0.data();
        for (int e = 0; e < G.Nelem; ++e) {
            const size_t base = (size_t)e * ngl2;
            for (int k = 0; k < ngl2; ++k) {
                const double xc = G.x[base + k], yc = G.y[base + k];
                const double d = a * std::sin(M_PI * xc / w) * std::sin(M_PI * yc / w);
                xe[k] = xc + d;
                ye[k] = yc + d;
            }
            element_metrics(ngl, D, xe.data(), ye.data(), xxi.data(), xeta.data(), yxi.data(), yeta.data(), Je.data());
            for (int k = 0; k < ngl2; ++k) {
                G.x[base + k] = xe[k]; G.y[base + k] = ye[k];
                G.xXi[base + k] = xxi[k]; G.xEta[base + k] = xeta[k];
                G.yXi[base + k] = yxi[k]; G.yEta[base + k] = yeta[k];
                G.J[base + k] = Je[k];
            }
            side_normals(ngl, xxi.data(), xeta.data(), yxi.data(), yeta.data(), Je.data(), &G.enx[(size_t)e * 4 * ngl],
                         &G.eny[(size_t)e * 4 * ngl], &G.escal[(size_t)e * 4 * ngl]);
        }
    } else if (s.mesh_kind != ESDG_MESH_CARTESIAN) {
        throw std::invalid_argument("unknown mesh kind");
    }
    return G;
}

}  // namespace

// ------------------------------------------------------------------------------------------
// Case object
// ------------------------------------------------------------------------------------------
struct esdg_case {
    esdg_case_spec spec{};
    Basis B;
    int Nelem = 0, Nfaces = 0;
    double min_ele_size = 0.0;
    std::map<std::string, std::vector<double>> d;
    std::map<std::string, std::vector<int>> iv;
};

namespace {

// Bathymetry and initial state at every node: the benchmark configurations, whose formulas
// (shared with the device case builder) are documented in esdg_case_state.h.
void make_state(esdg_case& C, const std::vector<double>& x, const std::vector<double>& y,
                const std::vector<double>* bmesh) {
    const esdg_case_spec& s = C.spec;
    if (!esdg_case_state::valid_case(s)) {
        if (s.case_kind == ESDG_CASE_CONVERGENCE)
            throw std::invalid_argument("convergence case: p0 must be Testcase 1, 8, 88 or 89");
        throw std::invalid_argument("unknown case kind");
    }
    const int ngl2 = C.B.ngl * C.B.ngl;
    const size_t nn = (size_t)C.Nelem * ngl2;
    std::vector<double>& b = C.d["b"];
    std::vector<double>& q = C.d["q"];
    b.assign(nn, 0.0);
    q.assign(nn * 3, 0.0);
    const esdg_case_state::FloodModes F = esdg_case_state::flood_modes(s.seed);
    for (int e = 0; e < C.Nelem; ++e)
        for (int k = 0; k < ngl2; ++k) {
            const size_t id = (size_t)e * ngl2 + k;
            const size_t qid = (size_t)e * ngl2 * 3 + k;
            esdg_case_state::node_state(s, F, x[id], y[id], qid, bmesh != nullptr, bmesh ? (*bmesh)[id] : 0.0,
                                        &b[id], &q[qid], &q[qid + ngl2], &q[qid + 2 * ngl2]);
        }
}

void build_case(esdg_case& C) {
    const esdg_case_spec& s = C.spec;
    C.B = make_basis(s.N, 1);
    const Basis& B = C.B;
    const int ngl = B.ngl, ngl2 = ngl * ngl;
    Geometry G = make_mesh(s, B);
    C.Nelem = G.Nelem;
    C.Nfaces = G.Nfaces;
    const size_t nn = (size_t)G.Nelem * ngl2;

    C.d["x_GL"] = B.xGL; C.d["w_GL"] = B.wGL;
    C.d["D"] = B.D; C.d["D0"] = B.D0; C.d["Dstrong"] = B.Dstrong; C.d["Dhat"] = B.Dhat;
    C.d["VdmInv"] = B.VdmInv;

    // per-face normals and scal from the left element's side (Mesh.cpp:282-307)
    std::vector<double> fnx((size_t)G.Nfaces * ngl), fny((size_t)G.Nfaces * ngl), fsc((size_t)G.Nfaces * ngl);
    for (int is = 0; is < G.Nfaces; is++)
        for (int i = 0; i < ngl; i++) {
            const size_t id = (size_t)is * ngl + i;
            const size_t idLoc = (size_t)G.edge_info[(size_t)is * 7 + 2] * 4 * ngl + (size_t)G.edge_info[(size_t)is * 7 + 4] * ngl + i;
            fnx[id] = G.enx[idLoc]; fny[id] = G.eny[idLoc]; fsc[id] = G.escal[idLoc];
        }
    // EdgeData (MeshPartitioning.cpp:104-131 then main.cpp:653-663) and per-element face maps
    std::vector<int> ED((size_t)G.Nfaces * 8), ER(G.Nfaces), E2E((size_t)G.Nelem * 4), EMS((size_t)G.Nelem * 4), EOR((size_t)G.Nelem * 4);
    for (int is = 0; is < G.Nfaces; is++) {
        const int* E = &G.edge_info[(size_t)is * 7];
        int* o = &ED[(size_t)is * 8];
        o[0] = E[2]; o[1] = E[3]; o[2] = E[4]; o[3] = E[5]; o[4] = E[6];
        o[5] = 0; o[6] = (E[3] == -1) ? -1 : 0; o[7] = is;
        ER[is] = o[4];
    }
    for (int ie = 0; ie < G.Nelem; ++ie)
        for (int is = 0; is < 4; is++) {
            const size_t id = (size_t)ie * 4 + is;
            const int ifa = G.elem_to_edge[id];
            if (ifa < 1) throw std::runtime_error("mesh element side without a face");
            E2E[id] = ifa;
            const int* E = &G.edge_info[(size_t)(ifa - 1) * 7];
            if (E[2] == ie) { EMS[id] = +1; EOR[id] = 1; }
            else { EMS[id] = -1; EOR[id] = E[6]; }
        }
    // 1/J and element sizes (main.cpp:700, basis.cpp:939-966)
    std::vector<double> Jinv(nn);
    for (size_t i = 0; i < nn; ++i) Jinv[i] = 1.0 / G.J[i];
    std::vector<double> sizes(G.Nelem, 0.0);
    double tmin = 0.0;
    for (int ie = 0; ie < G.Nelem; ie++) {
        for (int j = 0; j < ngl; ++j)
            for (int i = 0; i < ngl; ++i) {
                const size_t xid = (size_t)ie * ngl2 + j * ngl + i;
                sizes[ie] += 1.0 / Jinv[xid] * B.wGL[i] * B.wGL[j];
            }
        tmin = (ie == 0) ? sizes[ie] : std::min(tmin, sizes[ie]);
    }
    C.min_ele_size = std::sqrt(tmin);

    make_state(C, G.x, G.y, G.b.empty() ? nullptr : &G.b);

    // split-form bathymetry gradient (SW2D.cpp:1048-1105)
    const std::vector<double>& b = C.d["b"];
    std::vector<double> Bx(nn, 0.0), By(nn, 0.0);
    const double* D = B.D0.data();
    for (int ie = 0; ie < G.Nelem; ie++)
        for (int j = 0; j < ngl; ++j)
            for (int i = 0; i < ngl; ++i) {
                const size_t id = (size_t)ie * ngl2;
                double B_xi = 0.0, BYeta_xi = 0.0, BXxi_eta = 0.0;
                double B_eta = 0.0, BYxi_eta = 0.0, BXeta_xi = 0.0;
                for (int l = 0; l < ngl; ++l) {
                    const int il = i * ngl + l, jl = j * ngl + l, li = l * ngl + i;
                    B_xi += D[il] * b[id + jl];
                    BYeta_xi += D[il] * b[id + jl] * G.yEta[id + jl];
                    BXeta_xi += D[il] * b[id + jl] * G.xEta[id + jl];
                    B_eta += D[jl] * b[id + li];
                    BYxi_eta += D[jl] * b[id + li] * G.yXi[id + li];
                    BXxi_eta += D[jl] * b[id + li] * G.xXi[id + li];
                }
                const size_t xid = id + j * ngl + i;
                Bx[xid] = 0.5 * Jinv[xid] * (G.yEta[xid] * B_xi + BYeta_xi - G.yXi[xid] * B_eta - BYxi_eta);
                By[xid] = 0.5 * Jinv[xid] * (-G.xEta[xid] * B_xi - BXeta_xi + G.xXi[xid] * B_eta + BXxi_eta);
            }

    C.d["x"] = std::move(G.x); C.d["y"] = std::move(G.y);
    C.d["xXi"] = std::move(G.xXi); C.d["xEta"] = std::move(G.xEta);
    C.d["yXi"] = std::move(G.yXi); C.d["yEta"] = std::move(G.yEta);
    C.d["Jac"] = std::move(Jinv);
    C.d["Bx"] = std::move(Bx); C.d["By"] = std::move(By);
    C.d["nx"] = std::move(fnx); C.d["ny"] = std::move(fny); C.d["scal"] = std::move(fsc);
    C.d["EleSizes"] = std::move(sizes);
    C.iv["EdgeData"] = std::move(ED); C.iv["EdgeReversed"] = std::move(ER);
    C.iv["ElemToEdge"] = std::move(E2E); C.iv["ElemEdgeMasterSlave"] = std::move(EMS);
    C.iv["ElemEdgeOrientation"] = std::move(EOR);
}

}  // namespace

// ------------------------------------------------------------------------------------------
// C ABI
// ------------------------------------------------------------------------------------------
extern "C" {

const char* esdg_setup_last_error(void) { return g_err.c_str(); }

int esdg_basis_build(int N, int flux_differencing, double* x_GL, double* w_GL, double* D, double* D0,
                     double* Dstrong, double* Dhat, double* VdmInv) {
    try {
        Basis B = make_basis(N, flux_differencing);
        const size_t ngl = B.ngl, ngl2 = ngl * ngl;
        auto cp = [](double* dst, const std::vector<double>& v, size_t n) {
            if (dst) std::memcpy(dst, v.data(), n * sizeof(double));
        };
        cp(x_GL, B.xGL, ngl); cp(w_GL, B.wGL, ngl);
        cp(D, B.D, ngl2); cp(D0, B.D0, ngl2); cp(Dstrong, B.Dstrong, ngl2); cp(Dhat, B.Dhat, ngl2);
        cp(VdmInv, B.VdmInv, ngl2);
        return 0;
    } catch (const std::exception& e) {
        return fail(e.what());
    }
}

int esdg_rk_coefficients(int order, int ssp, int* stages, double* A, double* B, double* C) {
    if (!stages || !A || !B || !C) return fail("null argument");
    for (int i = 0; i < 5; ++i) { A[i] = 0.0; B[i] = 0.0; C[i] = 0.0; }
    if (order == 3) {
        *stages = 3;
        if (ssp) {  // Shu-Osher SSP-RK3
            A[0] = 1.0; A[1] = 3.0 / 4.0; A[2] = 1.0 / 3.0;
            B[0] = 0.0; B[1] = 1.0 / 4.0; B[2] = 2.0 / 3.0;
            C[0] = 1.0; C[1] = 1.0 / 4.0; C[2] = 2.0 / 3.0;
        } else {    // Williamson low-storage RK3
            A[0] = 0.0; A[1] = -5.0 / 9.0; A[2] = -153.0 / 128.0;
            B[0] = 0.0; B[1] = 1.0 / 3.0; B[2] = 3.0 / 4.0;
            C[0] = 1.0 / 3.0; C[1] = 15.0 / 16.0; C[2] = 8.0 / 15.0;
        }
        return 0;
    }
    if (order == 4) {  // Carpenter-Kennedy 5-stage 4th-order 2N-storage
        *stages = 5;
        A[0] = 0.0;
        A[1] = -(567301805773.0 / 1357537059087.0);
        A[2] = -(2404267990393.0 / 2016746695238.0);
        A[3] = -(3550918686646.0 / 2091501179385.0);
        A[4] = -(1275806237668.0 / 842570457699.0);
        B[0] = 0.0;
        B[1] = (1432997174477.0 / 9575080441755.0);
        B[2] = (2526269341429.0 / 6820363962896.0);
        B[3] = (2006345519317.0 / 3224310063776.0);
        B[4] = (2802321613138.0 / 2924317926251.0);
        C[0] = (1432997174477.0 / 9575080441755.0);
        C[1] = (5161836677717.0 / 13612068292357.0);
        C[2] = (1720146321549.0 / 2090206949498.0);
        C[3] = (3134564353537.0 / 4481467310338.0);
        C[4] = (2277821191437.0 / 14882151754819.0);
        return 0;
    }
    return fail("unsupported RK order (3 or 4)");
}

int esdg_case_create(const esdg_case_spec* spec, esdg_case** out) {
    if (!spec || !out) return fail("null argument");
    *out = nullptr;
    esdg_case* C = new (std::nothrow) esdg_case();
    if (!C) return fail("out of host memory");
    try {
        C->spec = *spec;
        build_case(*C);
    } catch (const std::exception& e) {
        delete C;
        return fail(e.what());
    }
    *out = C;
    return 0;
}

int esdg_case_get_info(const esdg_case* c, esdg_case_info* info) {
    if (!c || !info) return fail("null argument");
    info->N = c->B.N; info->ngl = c->B.ngl; info->Nelem = c->Nelem; info->Nfaces = c->Nfaces;
    info->min_ele_size = c->min_ele_size;
    return 0;
}

int esdg_case_array(const esdg_case* c, const char* name, void** data, int64_t* count, int* is_int) {
    if (!c || !name || !data || !count || !is_int) return fail("null argument");
    auto it = c->d.find(name);
    if (it != c->d.end()) {
        *data = (void*)it->second.data(); *count = (int64_t)it->second.size(); *is_int = 0;
        return 0;
    }
    auto jt = c->iv.find(name);
    if (jt != c->iv.end()) {
        *data = (void*)jt->second.data(); *count = (int64_t)jt->second.size(); *is_int = 1;
        return 0;
    }
    return fail(std::string("no case array named ") + name);
}

void esdg_case_destroy(esdg_case* c) { delete c; }

int esdg_total_entropy(int ngl, int Nelem, double g, const double* w_GL, const double* q, const double* b,
                       const double* Jac, double* out) {
    if (!w_GL || !q || !b || !Jac || !out) return fail("null argument");
    const int ngl2 = ngl * ngl;
    double acc = 0.0;
    for (int ie = 0; ie < Nelem; ie++)
        for (int j = 0; j < ngl; ++j)
            for (int i = 0; i < ngl; ++i) {
                const size_t id = (size_t)ie * ngl2 * 3 + j * ngl + i;
                const size_t xid = (size_t)ie * ngl2 + j * ngl + i;
                const double h = q[id], hu = q[id + ngl2], hv = q[id + 2 * ngl2];
                double u = 0.0, v = 0.0;
                if (h > std::pow(10, -12)) { u = hu / h; v = hv / h; }
                const double E = 0.5 * h * (u * u + v * v) + 0.5 * g * h * h + g * h * b[xid];
                acc += E / Jac[xid] * w_GL[i] * w_GL[j];
            }
    *out = acc;
    return 0;
}

int esdg_total_mass(int ngl, int Nelem, const double* w_GL, const double* q, const double* Jac, double* out) {
    if (!w_GL || !q || !Jac || !out) return fail("null argument");
    const int ngl2 = ngl * ngl;
    double acc = 0.0;
    for (int ie = 0; ie < Nelem; ie++)
        for (int j = 0; j < ngl; ++j)
            for (int i = 0; i < ngl; ++i) {
                const size_t id = (size_t)ie * ngl2 * 3 + j * ngl + i;
                const size_t xid = (size_t)ie * ngl2 + j * ngl + i;
                acc += q[id] / Jac[xid] * w_GL[i] * w_GL[j];
            }
    *out = acc;
    return 0;
}

// ---- Tecplot ASCII output (the reference's plot files, src/Plots.cpp) ------------------------
// One POINT zone of ngl x ngl nodes per element, nodes eta outer / xi inner, numbers in the
// stream's default format (6 significant digits), as the reference writes them.

}  // extern "C"

namespace {
template <class Row>
int write_zones(const char* path, const char* title, const char* vars, int Nelem, int ngl, Row row,
                const char* trailer = "") {
    if (!path || Nelem < 0 || ngl < 1) return fail("esdg_plot: bad argument");
    std::ofstream f(path);
    if (!f) return fail(std::string("esdg_plot: cannot open ") + path);
    f << "TITLE = " << title << "\n";
    f << "VARIABLES = " << vars << "\n";
    const int ngl2 = ngl * ngl;
    for (int ie = 0; ie < Nelem; ie++) {
        f << "ZONE I =" << ngl << ",J=" << ngl << ",F=POINT\n";
        for (int j = 0; j < ngl; ++j)
            for (int i = 0; i < ngl; ++i) row(f, ie, (size_t)ie * ngl2 + j * ngl + i, (size_t)ie * ngl2 * 3 + j * ngl + i);
    }
    f << trailer;
    f.close();
    return f ? 0 : fail(std::string("esdg_plot: write failed: ") + path);
}
}  // namespace

extern "C" {

int esdg_plot_solution(const char* path, int Nelem, int ngl, int plot_var, const double* x, const double* y,
                       const double* q, const double* b, double h_0) {
    if (!x || !y || !q || !b) return fail("esdg_plot_solution: null argument");
    const int ngl2 = ngl * ngl;
    if (plot_var == 1)
        return write_zones(path, "H_solution.tec", "\"x\",\"y\",\"H\",\"hu\",\"hv\",\"bottom\"", Nelem, ngl,
                           [&](std::ofstream& f, int, size_t xid, size_t id) {
                               const double H = q[id] + b[xid];
                               f << x[xid] << " " << y[xid] << " " << H << " " << q[id + ngl2] << " "
                                 << q[id + 2 * ngl2] << " " << b[xid] << " \n";
                           });
    if (plot_var == 2)
        return write_zones(path, "H_solution.tec", "\"x\",\"y\",\"FreeSurface\",\"hu\",\"hv\",\"bottom\"", Nelem, ngl,
                           [&](std::ofstream& f, int, size_t xid, size_t id) {
                               const double H = q[id] + b[xid] - h_0;
                               const double bottom = b[xid] - h_0;
                               f << x[xid] << " " << y[xid] << " " << H << " " << q[id + ngl2] << " "
                                 << q[id + 2 * ngl2] << " " << bottom << " \n";
                           });
    return fail("esdg_plot_solution: plot_var must be 1 or 2");
}

int esdg_plot_arrival_times(const char* path, int Nelem, int ngl, const double* x, const double* y, const double* arr) {
    if (!x || !y || !arr) return fail("esdg_plot_arrival_times: null argument");
    return write_zones(path, "Arrivals.tec", "\"x\",\"y\",\"Arrivaltime\",", Nelem, ngl,
                       [&](std::ofstream& f, int, size_t xid, size_t) {
                           f << x[xid] << " " << y[xid] << " " << arr[xid] << " \n";
                       });
}

int esdg_plot_maximum_elevation(const char* path, int Nelem, int ngl, const double* x, const double* y,
                                const double* mx) {
    if (!x || !y || !mx) return fail("esdg_plot_maximum_elevation: null argument");
    return write_zones(path, "MaximumElevations.tec", "\"x\",\"y\",\"MaximumElevation\",", Nelem, ngl,
                       [&](std::ofstream& f, int, size_t xid, size_t) {
                           f << x[xid] << " " << y[xid] << " " << mx[xid] << " \n";
                       });
}

// time series of the solution integrals (one line per recorded time)
static int write_series(const char* path, const char* title, const char* vars, int n, const double* t, const double* v,
                        bool relative) {
    if (!path || n < 0 || (n > 0 && (!t || !v))) return fail("esdg_plot: bad argument");
    std::ofstream f(path);
    if (!f) return fail(std::string("esdg_plot: cannot open ") + path);
    f << "TITLE = " << title << "\n";
    f << "VARIABLES = " << vars << "\n";
    for (int i = 0; i < n; ++i) {
        if (relative) f << t[i] << " " << (v[i] - v[0]) / v[0] << " \n";
        else f << t[i] << " " << v[i] << " \n";
    }
    f.close();
    return f ? 0 : fail(std::string("esdg_plot: write failed: ") + path);
}

int esdg_plot_entropy(const char* path, int n, const double* times, const double* total_entropy) {
    return write_series(path, "Entropyplot.tec", "\"t\",\"Entropy\"", n, times, total_entropy, false);
}

int esdg_plot_mass(const char* path, int n, const double* times, const double* total_mass) {
    return write_series(path, "Massplot.tec", "\"t\",\"Massdifference\"", n, times, total_mass, true);
}

int esdg_plot_time_series(const char* path, int n, const double* times, int ngauges, const char* const* names,
                          const double* series) {
    if (!path || n < 0 || ngauges < 0 || (ngauges > 0 && !names) || (n > 0 && (!times || (ngauges > 0 && !series))))
        return fail("esdg_plot_time_series: bad argument");
    for (int k = 0; k < ngauges; ++k)
        if (!names[k]) return fail("esdg_plot_time_series: null gauge name");
    std::ofstream f(path);
    if (!f) return fail(std::string("esdg_plot_time_series: cannot open ") + path);
    f << "TITLE = Timeseries.tec\n";
    f << "VARIABLES = \"t\"";
    for (int k = 0; k < ngauges; ++k) f << ",\"" << names[k] << "\"";
    f << "\n";
    for (int i = 0; i < n; ++i) {
        f << times[i];
        for (int k = 0; k < ngauges; ++k) f << " " << series[(size_t)k * n + i];
        f << " \n";
    }
    f.close();
    return f ? 0 : fail(std::string("esdg_plot_time_series: write failed: ") + path);
}

int esdg_plot_viscosity(const char* path, int Nelem, int ngl, const double* x, const double* y, const double* eps) {
    if (!x || !y || !eps) return fail("esdg_plot_viscosity: null argument");
    return write_zones(path, "ArtVisc.tec", "\"x\",\"y\",\"Eps\"", Nelem, ngl,
                       [&](std::ofstream& f, int ie, size_t xid, size_t) {
                           f << x[xid] << " " << y[xid] << " " << eps[ie] << " \n";
                       },
                       "Test");
}

}  // extern "C"
