/*
 * esdg_setup.h -- host-side problem setup for the ESDGSEM shallow-water hot path.
 *
 * Builds, on the host and in FP64, every constant array the per-stage kernels read:
 * the LGL basis and its operators, quad meshes (Cartesian and structured curved) in the
 * reference's face/element connectivity layout, metric terms, normals, bathymetry and its
 * split-form gradient, element sizes and initial states.  Pure C++ (no HIP); the arrays are
 * handed to the device layer (esdg.h) exactly as the reference hands them to
 * deviceclass::copyDeviceVariables (src/deviceclass.cpp:577-626).
 *
 * Layouts (identical to the reference's kernel contract):
 *   node (i = xi index, j = eta index) of element e:   e*ngl^2 + j*ngl + i
 *   state q / Qt / gRK:                                 e*ngl^2*3 + v*ngl^2 + j*ngl + i, v in {h,hu,hv}
 *   face arrays nx, ny, scal, bL, bR:                   f*ngl + k   (k in the left element's side order)
 *   face state qL, qR, SurfaceParts:                    f*ngl*3 + v*ngl + k
 *   EdgeData[8*f + ...] = {elemL, elemR(-1 boundary), sideL, sideR, orient(1 same, 0 reversed),
 *                          procL(=0), procR(-1 boundary, else 0), globalFaceId}
 *   ElemToEdge[4*e + s] is 1-based; sides 0 bottom, 1 right, 2 top, 3 left.
 *   Jac (device) = 1/J.
 *
 * All functions return 0 on success and a negative code on error; esdg_setup_last_error()
 * gives the message.  No C++ exception crosses this ABI.
 */
#ifndef ESDG_SETUP_H
#define ESDG_SETUP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces basis::basis (src/basis.cpp:28-135): LGL nodes/weights, derivative matrix D0,
 * flux-differencing D (2*D0 with SBP corner terms), Dstrong, Dhat = -W^-1 D0^T W and the
 * modal transform VdmInv (src/basis.cpp:495-564).  Every matrix is ngl x ngl row-major. */
int esdg_basis_build(int N, int flux_differencing,
                     double* x_GL, double* w_GL,
                     double* D, double* D0, double* Dstrong, double* Dhat, double* VdmInv);

/* Replaces RungeKutta::RungeKutta (src/RungeKutta.cpp:18-90).  A/B/C must hold 5 entries.
 * Returns the number of stages in *stages. */
int esdg_rk_coefficients(int order, int ssp, int* stages, double* A, double* B, double* C);

/* ---- test/benchmark cases --------------------------------------------------------- */

enum esdg_mesh_kind {
    ESDG_MESH_CARTESIAN = 0,      /* Mesh::GenerateMesh, src/Mesh.cpp:316-696 */
    ESDG_MESH_CURVED_SINE = 1,    /* same topology, nodes x += a*sin(pi*x/w)*sin(pi*y/w), same in y,
                                     metrics and normals as Mesh::ConstructMappedGeometry
                                     (src/Mesh.cpp:1667-1808) */
    ESDG_MESH_ISM = 2             /* ISM-V2 mesh file (Mesh::ReadMesh, src/Mesh.cpp:698-1503): corner
                                     nodes, edges, Chebyshev-Lobatto boundary curves, transfinite map
                                     (:1506-1808); spec.mesh_file, spec.read_bottom */
};

enum esdg_case_kind {
    ESDG_CASE_GAUSSIAN_BUMP = 0,  /* b = 0.8 e^{-5 r1^2}, h = 1 - b + amp e^{-20 r2^2}, u = v = 0 */
    ESDG_CASE_DAM_BREAK = 1,      /* b = 0.1 sin x cos y, h = 2 (x<5) else 1, + seeded U(-noise,noise);
                                     p1 = height ratio, p2 = step angle (radians) when set */
    ESDG_CASE_FLOOD = 2,          /* 16-mode random-phase Fourier bathymetry, reservoir over dry land */
    ESDG_CASE_OCEAN_UPLIFT = 3,   /* 4000 m ocean with linear shelf to 50 m, elliptic Gaussian uplift */
    ESDG_CASE_CONSTANT = 4,       /* constant state (h0, hu0, hv0) over flat b=0: free-stream test */
    ESDG_CASE_CONVERGENCE = 5     /* manufactured solution of Testcase p0 (SW2D.cpp:312-318, 829-833 for
                                     p0 = 1 or 0): b = 2 + sin(2 pi y)/2 + cos(2 pi x), h = 8 + cos x sin y - b;
                                     p0 = 8: b = 2 + sin y + cos x, h = 20 + cos x sin y - b;
                                     p0 = 88 / 89: b = 0, h = 20 + cos(x/1e3) sin(y/1e3) - b /
                                     20000 + cos(x/1e6) sin(y/1e6) - b; hu = h/2, hv = 3h/2 at t = 0 */
};

typedef struct esdg_case_spec {
    int case_kind;        /* esdg_case_kind */
    int mesh_kind;        /* esdg_mesh_kind */
    int N;                /* polynomial degree */
    int nelx, nely;       /* elements in x and y */
    double xL, xR, yL, yR;
    int periodic_x;       /* reference flag PeriodicBD_X: wraps the horizontal faces (y direction) */
    int periodic_y;       /* reference flag PeriodicBD_Y: wraps the vertical faces (x direction) */
    double perturb_amp;   /* curved mesh amplitude a */
    double perturb_len;   /* curved mesh half-wavelength w */
    double g;             /* gravity, used by nothing but the stored b gradients (kept for symmetry) */
    double amp;           /* case amplitude (bump height, noise amplitude, uplift amplitude ...) */
    double p0, p1, p2, p3;/* case parameters, meaning per case kind (see esdg_setup.cpp) */
    uint64_t seed;        /* seeds the counter-based generator (SplitMix64) */
    const char* mesh_file;/* ESDG_MESH_ISM: path of the ISM-V2 file (nelx, nely, x/y bounds unused) */
    int read_bottom;      /* ESDG_MESH_ISM: b from the file's third node/curve column (ReadBottom)
                             instead of the case's bathymetry */
    int convert_to_km;    /* ESDG_MESH_ISM: coordinates in degrees, mapped to km by the Mercator
                             projection of ReadMesh (ConvertToKM, src/Mesh.cpp:770-776) */
} esdg_case_spec;

typedef struct esdg_case esdg_case;

typedef struct esdg_case_info {
    int N, ngl, Nelem, Nfaces;
    double min_ele_size;  /* sqrt(min element area), basis::calcElementSizes (src/basis.cpp:939-966) */
} esdg_case_info;

int  esdg_case_create(const esdg_case_spec* spec, esdg_case** out);
int  esdg_case_get_info(const esdg_case* c, esdg_case_info* info);
/* Named host arrays of a case: "x","y","xXi","xEta","yXi","yEta","Jac","b","Bx","By","q",
 * "nx","ny","scal","EleSizes","x_GL","w_GL","D","D0","Dstrong","Dhat","VdmInv" (double) and
 * "EdgeData","EdgeReversed","ElemToEdge","ElemEdgeMasterSlave","ElemEdgeOrientation" (int32).
 * *is_int is set to 1 for int32 arrays.  The pointer stays owned by the case. */
int  esdg_case_array(const esdg_case* c, const char* name, void** data, int64_t* count, int* is_int);
void esdg_case_destroy(esdg_case* c);

/* Diagnostics on host arrays (basis.cpp:1017-1111): total entropy and mass with GL weights. */
int esdg_total_entropy(int ngl, int Nelem, double g, const double* w_GL, const double* q,
                       const double* b, const double* Jac, double* out);
int esdg_total_mass(int ngl, int Nelem, const double* w_GL, const double* q, const double* Jac,
                    double* out);

const char* esdg_setup_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
