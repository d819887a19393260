// ref_setup_driver.cpp -- test-only driver around the reference's own host setup sources.
//
// This file is part of the test oracle, not of the product.  It is compiled together with the
// reference's src/basis.cpp, src/RungeKutta.cpp, src/Mesh.cpp, src/SW2D.cpp and
// src/ReadInput.cpp and src/Plots.cpp (compiled where they lie, never copied) into oracle/_ref/, and exposes their
// results through a small C interface so the tests can pin the product's setup arrays
// (esdg_setup.cpp) bit for bit against the reference's own code.
#include <unistd.h>

#include <cstdio>
#include <cstring>
#include <string>

#include <sys/stat.h>

#include "Mesh.h"
#include "RungeKutta.h"
#include "SW2D.h"
#include "basis.h"
#include "plots.h"

extern "C" {

int ref_basis(int N, int FD, double* xGL, double* wGL, double* D, double* D0, double* Dstrong,
              double* Dhat, double* VdmInv) {
    basis B(N, FD);  // src/basis.cpp:28-135
    const int ngl = N + 1, ngl2 = ngl * ngl;
    std::memcpy(xGL, B.x_GL, ngl * sizeof(double));
    std::memcpy(wGL, B.w_GL, ngl * sizeof(double));
    std::memcpy(D, B.D, ngl2 * sizeof(double));
    std::memcpy(D0, B.D0, ngl2 * sizeof(double));
    std::memcpy(Dstrong, B.Dstrong, ngl2 * sizeof(double));
    std::memcpy(Dhat, B.Dhat, ngl2 * sizeof(double));
    std::memcpy(VdmInv, B.VdmInv, ngl2 * sizeof(double));
    return 0;
}

int ref_rk(int order, int ssp, int* stages, double* A, double* B, double* C) {
    RungeKutta RK(order, ssp);  // src/RungeKutta.cpp:18-90
    *stages = RK.rkStages;
    for (int i = 0; i < RK.rkStages; ++i) {
        A[i] = RK.CoeffsA[i]; B[i] = RK.CoeffsB[i]; C[i] = RK.CoeffsC[i];
    }
    return 0;
}

// Cartesian mesh through Mesh::InitMesh (Cartesian branch, default test case, so that the domain
// comes from a Cartesian.dat written into workdir): Mesh.cpp:60-307, 316-696.
int ref_cartesian_mesh(const char* workdir, int N, int nelx, int nely, double xL, double xR,
                       double yL, double yR, int px, int py, double* x, double* y, double* xXi,
                       double* xEta, double* yXi, double* yEta, double* J, double* nx, double* ny,
                       double* scal, int* EdgeInfo, int* ElemToEdge, int* MasterSlave,
                       int* Orientation, int* Nfaces_out) {
    char cwd[4096];
    if (!getcwd(cwd, sizeof cwd)) return -1;
    if (chdir(workdir) != 0) return -2;
    FILE* f = std::fopen("Cartesian.dat", "w");
    if (!f) { (void)chdir(cwd); return -3; }
    std::fprintf(f, "xL=%.17g\nxR=%.17g\nyL=%.17g\nyR=%.17g\nNelemX=%d\nNelemY=%d\nPeriodicBD_X=%d\nPeriodicBD_Y=%d\n",
                 xL, xR, yL, yR, nelx, nely, px, py);
    std::fclose(f);
    basis B(N, 1);
    const int ngl = N + 1, ngl2 = ngl * ngl;
    Mesh M(B.x_GL, ngl, 0, 0, 0);
    M.InitMesh(std::string(""), true, 999);
    (void)chdir(cwd);
    const int Ne = M.m_num_elements, Nf = M.m_num_edges;
    *Nfaces_out = Nf;
    std::memcpy(x, M.x_global, sizeof(double) * Ne * ngl2);
    std::memcpy(y, M.y_global, sizeof(double) * Ne * ngl2);
    std::memcpy(xXi, M.xXi_global, sizeof(double) * Ne * ngl2);
    std::memcpy(xEta, M.xEta_global, sizeof(double) * Ne * ngl2);
    std::memcpy(yXi, M.yXi_global, sizeof(double) * Ne * ngl2);
    std::memcpy(yEta, M.yEta_global, sizeof(double) * Ne * ngl2);
    std::memcpy(J, M.J_global, sizeof(double) * Ne * ngl2);
    std::memcpy(nx, M.NormalsX, sizeof(double) * Nf * ngl);
    std::memcpy(ny, M.NormalsY, sizeof(double) * Nf * ngl);
    std::memcpy(scal, M.Scal, sizeof(double) * Nf * ngl);
    std::memcpy(EdgeInfo, M.EdgeInfo, sizeof(int) * Nf * 7);
    std::memcpy(ElemToEdge, M.ElementToEdge, sizeof(int) * Ne * 4);
    std::memcpy(MasterSlave, M.ElemEdgeMasterSlave, sizeof(int) * Ne * 4);
    std::memcpy(Orientation, M.ElemEdgeOrientation, sizeof(int) * Ne * 4);
    return 0;
}

// ISM-V2 mesh through Mesh::InitMesh (file branch): Mesh.cpp:62-307, 698-1808.  Sizes first
// (arrays null), then the arrays.
int ref_ism_mesh(const char* path, int N, int read_bottom, int to_km, int* Nelem_out, int* Nfaces_out, double* x,
                 double* y, double* xXi, double* xEta, double* yXi, double* yEta, double* J, double* b,
                 double* nx, double* ny, double* scal, int* EdgeInfo, int* ElemToEdge, int* MasterSlave,
                 int* Orientation) {
    basis B(N, 1);
    const int ngl = N + 1, ngl2 = ngl * ngl;
    Mesh M(B.x_GL, ngl, 0, read_bottom, to_km);
    M.InitMesh(std::string(path), false, 0);
    const int Ne = M.m_num_elements, Nf = M.m_num_edges;
    *Nelem_out = Ne;
    *Nfaces_out = Nf;
    if (!x) return 0;
    std::memcpy(x, M.x_global, sizeof(double) * Ne * ngl2);
    std::memcpy(y, M.y_global, sizeof(double) * Ne * ngl2);
    std::memcpy(xXi, M.xXi_global, sizeof(double) * Ne * ngl2);
    std::memcpy(xEta, M.xEta_global, sizeof(double) * Ne * ngl2);
    std::memcpy(yXi, M.yXi_global, sizeof(double) * Ne * ngl2);
    std::memcpy(yEta, M.yEta_global, sizeof(double) * Ne * ngl2);
    std::memcpy(J, M.J_global, sizeof(double) * Ne * ngl2);
    if (read_bottom) std::memcpy(b, M.b_global, sizeof(double) * Ne * ngl2);
    std::memcpy(nx, M.NormalsX, sizeof(double) * Nf * ngl);
    std::memcpy(ny, M.NormalsY, sizeof(double) * Nf * ngl);
    std::memcpy(scal, M.Scal, sizeof(double) * Nf * ngl);
    std::memcpy(EdgeInfo, M.EdgeInfo, sizeof(int) * Nf * 7);
    std::memcpy(ElemToEdge, M.ElementToEdge, sizeof(int) * Ne * 4);
    std::memcpy(MasterSlave, M.ElemEdgeMasterSlave, sizeof(int) * Ne * 4);
    std::memcpy(Orientation, M.ElemEdgeOrientation, sizeof(int) * Ne * 4);
    return 0;
}

int ref_bathymetry_gradient(int Nelem, int ngl, double g, const double* x, const double* y,
                            const double* b, const double* D0, const double* yEta, const double* yXi,
                            const double* xEta, const double* xXi, double* Bx, double* By,
                            const double* Jinv) {
    SW2D P(0, 0.0);  // src/SW2D.cpp:1048-1105
    P.CalcBDerivatives(Nelem, ngl, ngl * ngl, g, x, y, b, D0, yEta, yXi, xEta, xXi, Bx, By, Jinv);
    return 0;
}

int ref_element_sizes(int N, int Nelem, const double* Jinv, double* sizes, double* min_size) {
    basis B(N, 1);  // src/basis.cpp:939-966
    B.setNelemLocal(Nelem);
    for (int i = 0; i < Nelem; ++i) sizes[i] = 0.0;
    B.calcElementSizes(Jinv, sizes, min_size);
    return 0;
}

int ref_total_entropy_mass(int N, int Nelem, double g, const double* q, const double* b,
                           const double* Jinv, double* entropy, double* mass) {
    basis B(N, 1);  // src/basis.cpp:1017-1111
    B.setNelemGlobal(Nelem);
    B.calcTotalEntropy(g, q, b, Jinv, entropy);
    B.calcTotalMass(q, Jinv, mass);
    return 0;
}

// The reference's plot writers (src/Plots.cpp), run in workdir: they write under movie/.
// which: 0 PlotSolution, 1 PlotArrivalTimings, 2 PlotMaximumElevation, 3 PlotViscoseParameter,
// 4 PlotEntropy, 5 PlotMass (count records: times in x, totals in values), 6 PlotTimeSeries
// (values: q for 0, the per-node buffer for 1 and 2, the element viscosity for 3).
int ref_plot(const char* workdir, int which, int Nelem, int ngl, int plot_var, const double* x, const double* y,
             const double* values, const double* b, int count, double h0) {
    char cwd[4096];
    if (!getcwd(cwd, sizeof cwd)) return -1;
    if (chdir(workdir) != 0) return -2;
    mkdir("movie", 0755);
    switch (which) {
        case 0: PlotSolution(Nelem, ngl, plot_var, x, y, values, b, count, h0); break;
        case 1: PlotArrivalTimings(Nelem, ngl, x, y, values); break;
        case 2: PlotMaximumElevation(Nelem, ngl, x, y, values); break;
        case 3: PlotViscoseParameter(Nelem, ngl, x, y, values, count); break;
        case 4: PlotEntropy(count, x, values); break;     // times in x, totals in values
        case 5: PlotMass(count, x, values); break;
        case 6:   // times in x, the seven gauge series [7][count] in values
            PlotTimeSeries(count, x, values, values + count, values + 2 * count, values + 3 * count,
                           values + 4 * count, values + 5 * count, values + 6 * count);
            break;
        default: break;
    }
    return chdir(cwd) == 0 ? 0 : -3;
}

}  // extern "C"
