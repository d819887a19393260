// esdg_geometry.hip -- mesh geometry and the benchmark cases built on the device
// (include/esdg_device_setup.h).
//
// Each kernel restates one host setup step of the reference with its operation order, one
// thread per output value, so that the arrays equal the host setup's bit for bit (the library is
// compiled without FMA contraction):
//   structured face numbering, EdgeInfo, ElementToEdge   src/Mesh.cpp:412-655
//   EdgeData, ElemEdgeMasterSlave / Orientation          src/MeshPartitioning.cpp:104-132, main.cpp:650-716
//   node placement                                       src/Mesh.cpp:361-375
//   metric terms and J                                   src/Mesh.cpp:1722-1742
//   side normals and scal, gathered per face             src/Mesh.cpp:1744-1806, 282-307
//   1/J                                                  main.cpp:700
//   element sizes and their minimum                      src/basis.cpp:939-966
//   split-form bathymetry gradient                       src/SW2D.cpp:1048-1105
#include <cmath>
#include <map>
#include <string>
#include <vector>

#include "esdg_case_state.h"
#include "esdg_device_setup.h"
#include "esdg_internal.h"

struct esdg_device_case {
    esdg_ctx* ctx = nullptr;
    int N = 0, ngl = 0, Nelem = 0, Nfaces = 0;
    double min_ele_size = 0.0;
    struct Arr { void* p; int64_t count; int is_int; };
    std::map<std::string, Arr> arrays;
    ~esdg_device_case() {
        for (auto& kv : arrays) (void)esdg_free(ctx, kv.second.p);
    }
};

namespace esdg {

namespace {

constexpr int TPB = 256;
inline unsigned grid_of(int64_t n) { return (unsigned)std::min<int64_t>(std::max<int64_t>((n + TPB - 1) / TPB, 1), 65536); }

// ---- structured face numbering (Mesh::GenerateMesh, src/Mesh.cpp:412-655) ----------------------
// Horizontal faces first: face is = j*NelemX + i + 1 for j = 0..jb (jb = NelemY, or NelemY-1 when the
// horizontal faces wrap), then vertical ones: is = (jb+1)*NelemX + j*(ib+1) + i + 1 for i = 0..ib.
// Writes EdgeData [f][8], EdgeReversed [f] and ElemToEdge [e][4] (1-based).
struct Structured { int nx, ny, px, py; };

__global__ void __launch_bounds__(TPB) k_structured_faces(Structured m, int Nfaces, int* __restrict__ ED,
                                                          int* __restrict__ ER, int* __restrict__ E2E) {
    const int jb = m.ny - (m.px ? 1 : 0), ib = m.nx - (m.py ? 1 : 0);
    const int nh = (jb + 1) * m.nx;
    for (int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; f < Nfaces; f += (int64_t)gridDim.x * blockDim.x) {
        int eL, eR, sL, sR;
        if (f < nh) {
            const int j = (int)(f / m.nx), i = (int)(f % m.nx);
            if (j == 0) {
                eL = i; sL = 0; eR = -1; sR = -1;
                if (m.px) { eR = (m.ny - 1) * m.nx + i; sR = 2; }
            } else if (j == m.ny) {
                eL = (m.ny - 1) * m.nx + i; sL = 2; eR = -1; sR = -1;
            } else {
                eL = (j - 1) * m.nx + i; eR = j * m.nx + i; sL = 2; sR = 0;
            }
        } else {
            const int64_t g = f - nh;
            const int j = (int)(g / (ib + 1)), i = (int)(g % (ib + 1));
            if (i == 0) {
                eL = j * m.nx; sL = 3; eR = -1; sR = -1;
                if (m.py) { eR = (j + 1) * m.nx - 1; sR = 1; }
            } else if (i == m.nx) {
                eL = (j + 1) * m.nx - 1; sL = 1; eR = -1; sR = -1;
            } else {
                eL = j * m.nx + i - 1; eR = j * m.nx + i; sL = 1; sR = 3;
            }
        }
        int* o = ED + f * 8;
        o[0] = eL; o[1] = eR; o[2] = sL; o[3] = sR; o[4] = 1;
        o[5] = 0; o[6] = (eR == -1) ? -1 : 0; o[7] = (int)f;
        ER[f] = 1;
        E2E[(int64_t)eL * 4 + sL] = (int)f + 1;
        if (eR != -1) E2E[(int64_t)eR * 4 + sR] = (int)f + 1;
    }
}

// ElemEdgeMasterSlave / ElemEdgeOrientation of every element side (main.cpp:650-716): +1 and 1
// where the element is the face's left element, else -1 and the face's orientation flag.
__global__ void __launch_bounds__(TPB) k_element_face_maps(int Nelem, const int* __restrict__ E2E,
                                                           const int* __restrict__ ED, int* __restrict__ EMS,
                                                           int* __restrict__ EOR, int* __restrict__ bad) {
    for (int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; id < (int64_t)Nelem * 4;
         id += (int64_t)gridDim.x * blockDim.x) {
        const int ifa = E2E[id];
        if (ifa < 1) { atomicOr(bad, 1); continue; }
        const int* E = ED + (int64_t)(ifa - 1) * 8;
        if (E[0] == (int)(id / 4)) { EMS[id] = 1; EOR[id] = 1; }
        else { EMS[id] = -1; EOR[id] = E[4]; }
    }
}

// ---- node placement (src/Mesh.cpp:361-375) with the curved displacement ----------------------------
// xc[ieX*ngl + i] = xL + (x_GL[i]+1)/2 dX + ieX dX and likewise yc; sx / sy = sin(pi xc / w) and
// sin(pi yc / w), evaluated on the host.  x = xc + (a sx) sy, y = yc + (a sx) sy (curved), or xc, yc.
__global__ void __launch_bounds__(TPB) k_structured_nodes(int nx, int ny, int ngl, int curved, double a,
                                                          const double* __restrict__ xc, const double* __restrict__ yc,
                                                          const double* __restrict__ sx, const double* __restrict__ sy,
                                                          double* __restrict__ x, double* __restrict__ y) {
    const int ngl2 = ngl * ngl;
    const int64_t n = (int64_t)nx * ny * ngl2;
    for (int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; id < n; id += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = id / ngl2;
        const int k = (int)(id - e * ngl2), j = k / ngl, i = k - j * ngl;
        const int ieY = (int)(e / nx), ieX = (int)(e - (int64_t)ieY * nx);
        const int ix = ieX * ngl + i, iy = ieY * ngl + j;
        double X = xc[ix], Y = yc[iy];
        if (curved) {
            const double d = a * sx[ix] * sy[iy];
            X = X + d;
            Y = Y + d;
        }
        x[id] = X;
        y[id] = Y;
    }
}

// ---- metric terms and J (src/Mesh.cpp:1722-1742), 1/J (main.cpp:700) -------------------------------
__global__ void __launch_bounds__(TPB) k_mapped_metrics(int Nelem, int ngl, const double* __restrict__ D,
                                                        const double* __restrict__ x, const double* __restrict__ y,
                                                        double* __restrict__ xXi, double* __restrict__ xEta,
                                                        double* __restrict__ yXi, double* __restrict__ yEta,
                                                        double* __restrict__ Jac) {
    const int ngl2 = ngl * ngl;
    const int64_t n = (int64_t)Nelem * ngl2;
    for (int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; id < n; id += (int64_t)gridDim.x * blockDim.x) {
        const int64_t base = (id / ngl2) * ngl2;
        const int ij = (int)(id - base), j = ij / ngl, i = ij - j * ngl;
        double ax = 0.0, bx = 0.0, ay = 0.0, by = 0.0;
        for (int k = 0; k < ngl; k++) {
            ax += D[i * ngl + k] * x[base + j * ngl + k];
            bx += D[j * ngl + k] * x[base + k * ngl + i];
            ay += D[i * ngl + k] * y[base + j * ngl + k];
            by += D[j * ngl + k] * y[base + k * ngl + i];
        }
        xXi[id] = ax; xEta[id] = bx; yXi[id] = ay; yEta[id] = by;
        const double J = ax * by - bx * ay;
        Jac[id] = 1.0 / J;
    }
}

// Constant metrics of the Cartesian generator (src/Mesh.cpp:378-398).
__global__ void __launch_bounds__(TPB) k_cartesian_metrics(int64_t n, double x_xi, double y_eta, double Jc,
                                                           double* __restrict__ xXi, double* __restrict__ xEta,
                                                           double* __restrict__ yXi, double* __restrict__ yEta,
                                                           double* __restrict__ Jac) {
    const double jinv = 1.0 / Jc;
    for (int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; id < n; id += (int64_t)gridDim.x * blockDim.x) {
        xXi[id] = x_xi; xEta[id] = 0.0; yXi[id] = 0.0; yEta[id] = y_eta; Jac[id] = jinv;
    }
}

// ---- face normals (src/Mesh.cpp:1744-1806 for the left element's side, gathered as :282-307) ------
// The sign of J is read from 1/J (the same sign, also for J = +-0).
__global__ void __launch_bounds__(TPB) k_face_normals(int Nfaces, int ngl, const int* __restrict__ ED,
                                                      const double* __restrict__ xXi, const double* __restrict__ xEta,
                                                      const double* __restrict__ yXi, const double* __restrict__ yEta,
                                                      const double* __restrict__ Jac, double* __restrict__ nx,
                                                      double* __restrict__ ny, double* __restrict__ scal) {
    const int64_t n = (int64_t)Nfaces * ngl;
    for (int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; id < n; id += (int64_t)gridDim.x * blockDim.x) {
        const int64_t f = id / ngl;
        const int k = (int)(id - f * ngl);
        const int e = ED[f * 8 + 0], side = ED[f * 8 + 2];
        int node;
        switch (side) {
            case 0: node = k; break;                             // eta = -1, node (k, 0)
            case 1: node = k * ngl + ngl - 1; break;             // xi = 1, node (N, k)
            case 2: node = (ngl - 1) * ngl + k; break;           // eta = 1, node (k, N)
            default: node = k * ngl; break;                      // xi = -1, node (0, k)
        }
        const int64_t xid = (int64_t)e * ngl * ngl + node;
        const double sg = copysign(1.0, Jac[xid]);
        double sc, vx, vy;
        if (side == 0 || side == 2) {
            const double yx = yXi[xid], xx = xXi[xid];
            sc = sqrt(yx * yx + xx * xx);
            if (side == 0) { vx = -sg * (-yx / sc); vy = -sg * (xx / sc); }
            else { vx = sg * (-yx / sc); vy = sg * (xx / sc); }
        } else {
            const double ye = yEta[xid], xe = xEta[xid];
            sc = sqrt(ye * ye + xe * xe);
            if (side == 1) { vx = sg * (ye / sc); vy = sg * (-xe / sc); }
            else { vx = -sg * (ye / sc); vy = -sg * (-xe / sc); }
        }
        nx[id] = vx; ny[id] = vy; scal[id] = sc;
    }
}

// Face normals of the Cartesian generator (constant per side: Mesh.cpp:412-655's side_set values).
__global__ void __launch_bounds__(TPB) k_cartesian_face_normals(int Nfaces, int ngl, const int* __restrict__ ED,
                                                                double x_xi, double y_eta, double* __restrict__ nx,
                                                                double* __restrict__ ny, double* __restrict__ scal) {
    const int64_t n = (int64_t)Nfaces * ngl;
    for (int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; id < n; id += (int64_t)gridDim.x * blockDim.x) {
        const int side = ED[(id / ngl) * 8 + 2];
        double vx = 0.0, vy = 0.0, sc;
        switch (side) {
            case 0: vy = -1.0; sc = x_xi; break;
            case 2: vy = 1.0; sc = x_xi; break;
            case 3: vx = -1.0; sc = y_eta; break;
            default: vx = 1.0; sc = y_eta; break;
        }
        nx[id] = vx; ny[id] = vy; scal[id] = sc;
    }
}

// ---- element sizes (src/basis.cpp:939-966) -------------------------------------------------------
// One thread per element sums in the reference's (j, i) order; block minima go to part[], and a
// second launch folds them (min is exact, so the order of the fold does not matter).
__global__ void __launch_bounds__(TPB) k_element_sizes(int Nelem, int ngl, const double* __restrict__ Jac,
                                                       const double* __restrict__ w, double* __restrict__ sizes,
                                                       double* __restrict__ part) {
    __shared__ double s[TPB];
    double m = INFINITY;
    const int ngl2 = ngl * ngl;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < Nelem; e += (int64_t)gridDim.x * blockDim.x) {
        double acc = 0.0;
        for (int j = 0; j < ngl; ++j)
            for (int i = 0; i < ngl; ++i) acc += 1.0 / Jac[e * ngl2 + j * ngl + i] * w[i] * w[j];
        sizes[e] = acc;
        m = fmin(m, acc);
    }
    s[threadIdx.x] = m;
    __syncthreads();
    for (int k = TPB / 2; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) s[threadIdx.x] = fmin(s[threadIdx.x], s[threadIdx.x + k]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ void __launch_bounds__(TPB) k_min_fold(const double* __restrict__ part, int n, double* __restrict__ out) {
    __shared__ double s[TPB];
    double m = INFINITY;
    for (int k = threadIdx.x; k < n; k += TPB) m = fmin(m, part[k]);
    s[threadIdx.x] = m;
    __syncthreads();
    for (int k = TPB / 2; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) s[threadIdx.x] = fmin(s[threadIdx.x], s[threadIdx.x + k]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = s[0];
}

// ---- split-form bathymetry gradient (src/SW2D.cpp:1048-1105) --------------------------------------
__global__ void __launch_bounds__(TPB) k_bathymetry_gradient(int Nelem, int ngl, const double* __restrict__ D,
                                                             const double* __restrict__ b,
                                                             const double* __restrict__ yEta,
                                                             const double* __restrict__ yXi,
                                                             const double* __restrict__ xEta,
                                                             const double* __restrict__ xXi,
                                                             const double* __restrict__ Jac, double* __restrict__ Bx,
                                                             double* __restrict__ By) {
    const int ngl2 = ngl * ngl;
    const int64_t n = (int64_t)Nelem * ngl2;
    for (int64_t xid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; xid < n; xid += (int64_t)gridDim.x * blockDim.x) {
        const int64_t id = (xid / ngl2) * ngl2;
        const int ji = (int)(xid - id), j = ji / ngl, i = ji - j * ngl;
        double B_xi = 0.0, BYeta_xi = 0.0, BXxi_eta = 0.0;
        double B_eta = 0.0, BYxi_eta = 0.0, BXeta_xi = 0.0;
        for (int l = 0; l < ngl; ++l) {
            const int il = i * ngl + l, jl = j * ngl + l, li = l * ngl + i;
            B_xi += D[il] * b[id + jl];
            BYeta_xi += D[il] * b[id + jl] * yEta[id + jl];
            BXeta_xi += D[il] * b[id + jl] * xEta[id + jl];
            B_eta += D[jl] * b[id + li];
            BYxi_eta += D[jl] * b[id + li] * yXi[id + li];
            BXxi_eta += D[jl] * b[id + li] * xXi[id + li];
        }
        Bx[xid] = 0.5 * Jac[xid] * (yEta[xid] * B_xi + BYeta_xi - yXi[xid] * B_eta - BYxi_eta);
        By[xid] = 0.5 * Jac[xid] * (-xEta[xid] * B_xi - BXeta_xi + xXi[xid] * B_eta + BXxi_eta);
    }
}

// ---- bathymetry and initial state (esdg_case_state.h, shared with the host setup) ---------------
__global__ void __launch_bounds__(TPB) k_case_state(esdg_case_spec s, esdg_case_state::FloodModes F, int Nelem,
                                                    int ngl2, const double* __restrict__ x,
                                                    const double* __restrict__ y, double* __restrict__ b,
                                                    double* __restrict__ q) {
    const int64_t n = (int64_t)Nelem * ngl2;
    for (int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; id < n; id += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = id / ngl2;
        const int k = (int)(id - e * ngl2);
        const int64_t qid = e * ngl2 * 3 + k;
        esdg_case_state::node_state(s, F, x[id], y[id], (uint64_t)qid, false, 0.0, &b[id], &q[qid], &q[qid + ngl2],
                                    &q[qid + 2 * ngl2]);
    }
}

int size_check(esdg_ctx* ctx, int Nelem, int ngl) {
    ESDG_REQUIRE(ctx, Nelem >= 0, "bad element count");
    ESDG_REQUIRE(ctx, ngl == ctx->p.ngl, "ngl differs from the context's");
    return ESDG_OK;
}

}  // namespace

}  // namespace esdg

using namespace esdg;

extern "C" {

int esdg_mapped_metrics(esdg_ctx* ctx, int Nelem, const double* D0, const double* x, const double* y, double* xXi,
                        double* xEta, double* yXi, double* yEta, double* Jac) {
    ESDG_REQUIRE(ctx, ctx, "null context");
    ESDG_REQUIRE(ctx, Nelem >= 0 && D0 && x && y && xXi && xEta && yXi && yEta && Jac, "esdg_mapped_metrics: bad argument");
    if (Nelem == 0) return ESDG_OK;
    KernelScope ks(ctx, "mapped_metrics");
    const int ngl = ctx->p.ngl;
    k_mapped_metrics<<<grid_of((int64_t)Nelem * ngl * ngl), TPB, 0, ctx->stream>>>(Nelem, ngl, D0, x, y, xXi, xEta, yXi,
                                                                                  yEta, Jac);
    return launch_check(ctx, "mapped_metrics");
}

int esdg_face_normals(esdg_ctx* ctx, int Nfaces, const int* EdgeData, const double* xXi, const double* xEta,
                      const double* yXi, const double* yEta, const double* Jac, double* nx, double* ny, double* scal) {
    ESDG_REQUIRE(ctx, ctx, "null context");
    ESDG_REQUIRE(ctx, Nfaces >= 0 && EdgeData && xXi && xEta && yXi && yEta && Jac && nx && ny && scal,
                 "esdg_face_normals: bad argument");
    if (Nfaces == 0) return ESDG_OK;
    KernelScope ks(ctx, "face_normals");
    const int ngl = ctx->p.ngl;
    k_face_normals<<<grid_of((int64_t)Nfaces * ngl), TPB, 0, ctx->stream>>>(Nfaces, ngl, EdgeData, xXi, xEta, yXi, yEta,
                                                                           Jac, nx, ny, scal);
    return launch_check(ctx, "face_normals");
}

int esdg_element_sizes(esdg_ctx* ctx, int Nelem, const double* Jac, const double* w_GL, double* EleSizes,
                       double* host_min_ele_size) {
    ESDG_REQUIRE(ctx, ctx, "null context");
    ESDG_REQUIRE(ctx, Nelem > 0 && Jac && w_GL && EleSizes && host_min_ele_size, "esdg_element_sizes: bad argument");
    KernelScope ks(ctx, "element_sizes");
    const int nb = (int)std::min<int64_t>(((int64_t)Nelem + TPB - 1) / TPB, 2048);
    static_assert(2048 + 1 <= 4096, "the partial minima use the context scratch");
    double* part = ctx->d_scratch;
    k_element_sizes<<<nb, TPB, 0, ctx->stream>>>(Nelem, ctx->p.ngl, Jac, w_GL, EleSizes, part);
    k_min_fold<<<1, TPB, 0, ctx->stream>>>(part, nb, part + 2048);
    int rc = launch_check(ctx, "element_sizes");
    if (rc) return rc;
    double tmin = 0.0;
    rc = esdg_memcpy(ctx, &tmin, part + 2048, sizeof(double), ESDG_D2H);
    if (rc) return rc;
    *host_min_ele_size = std::sqrt(tmin);
    return ESDG_OK;
}

int esdg_bathymetry_gradient(esdg_ctx* ctx, int Nelem, const double* D0, const double* B, const double* Yeta,
                             const double* Yxi, const double* Xeta, const double* Xxi, const double* Jac, double* Bx,
                             double* By) {
    ESDG_REQUIRE(ctx, ctx, "null context");
    ESDG_REQUIRE(ctx, Nelem >= 0 && D0 && B && Yeta && Yxi && Xeta && Xxi && Jac && Bx && By,
                 "esdg_bathymetry_gradient: bad argument");
    if (Nelem == 0) return ESDG_OK;
    KernelScope ks(ctx, "bathymetry_gradient");
    const int ngl = ctx->p.ngl;
    k_bathymetry_gradient<<<grid_of((int64_t)Nelem * ngl * ngl), TPB, 0, ctx->stream>>>(Nelem, ngl, D0, B, Yeta, Yxi,
                                                                                       Xeta, Xxi, Jac, Bx, By);
    return launch_check(ctx, "bathymetry_gradient");
}

int esdg_device_case_create(esdg_ctx* ctx, const esdg_case_spec* spec, esdg_device_case** out) {
    ESDG_REQUIRE(ctx, ctx, "null context");
    ESDG_REQUIRE(ctx, spec && out, "esdg_device_case_create: null argument");
    *out = nullptr;
    const esdg_case_spec& s = *spec;
    ESDG_REQUIRE(ctx, s.mesh_kind == ESDG_MESH_CARTESIAN || s.mesh_kind == ESDG_MESH_CURVED_SINE,
                 "esdg_device_case_create: structured mesh kinds only (ISM files are read by esdg_case_create)");
    ESDG_REQUIRE(ctx, s.N + 1 == ctx->p.ngl, "esdg_device_case_create: spec N + 1 differs from the context's ngl");
    ESDG_REQUIRE(ctx, s.nelx >= 1 && s.nely >= 1, "esdg_device_case_create: nelx and nely must be >= 1");
    const int ngl = s.N + 1, ngl2 = ngl * ngl;
    ESDG_REQUIRE(ctx, (int64_t)s.nelx * s.nely * ngl2 * 3 <= (int64_t)INT32_MAX,
                 "esdg_device_case_create: mesh too large for 32-bit node indices");
    const bool PX = s.periodic_x != 0, PY = s.periodic_y != 0;
    const bool curved = s.mesh_kind == ESDG_MESH_CURVED_SINE;
    ESDG_REQUIRE(ctx, !curved || (!PX && !PY), "esdg_device_case_create: the curved mesh kind is not periodic");
    ESDG_REQUIRE(ctx, !curved || s.perturb_len > 0.0, "esdg_device_case_create: perturb_len must be > 0");
    ESDG_REQUIRE(ctx, esdg_case_state::valid_case(s), "esdg_device_case_create: unknown case kind or Testcase");

    // basis on the host (ngl^2 values)
    std::vector<double> xGL(ngl), wGL(ngl), D(ngl2), D0(ngl2), Dst(ngl2), Dhat(ngl2), Vi(ngl2);
    if (esdg_basis_build(s.N, 1, xGL.data(), wGL.data(), D.data(), D0.data(), Dst.data(), Dhat.data(), Vi.data()) != 0)
        return set_error(ctx, ESDG_ERR_ARG, std::string("esdg_device_case_create: ") + esdg_setup_last_error());

    esdg_device_case* C = new (std::nothrow) esdg_device_case();
    if (!C) return set_error(ctx, ESDG_ERR_NOMEM, "esdg_device_case_create: out of host memory");
    C->ctx = ctx;
    C->N = s.N;
    C->ngl = ngl;
    C->Nelem = s.nelx * s.nely;
    int nedges = (s.nelx + 1) * s.nely + (s.nely + 1) * s.nelx;
    if (PX) nedges -= s.nelx;
    if (PY) nedges -= s.nely;
    C->Nfaces = nedges;
    const int64_t nn = (int64_t)C->Nelem * ngl2, fn = (int64_t)nedges * ngl;
    int rc = ESDG_OK;
    auto A = [&](const char* name, int64_t count, int is_int) -> void* {
        if (rc) return nullptr;
        void* p = nullptr;
        rc = esdg_malloc(ctx, &p, (size_t)std::max<int64_t>(count, 1) * (is_int ? sizeof(int) : sizeof(double)));
        if (rc) return nullptr;
        C->arrays[name] = {p, count, is_int};
        return p;
    };
    auto H = [&](const char* name, const std::vector<double>& h) {
        void* p = A(name, (int64_t)h.size(), 0);
        if (!rc) rc = esdg_memcpy(ctx, p, h.data(), h.size() * sizeof(double), ESDG_H2D);
    };
    H("x_GL", xGL); H("w_GL", wGL); H("D", D); H("D0", D0); H("Dstrong", Dst); H("Dhat", Dhat); H("VdmInv", Vi);
    int* ED = (int*)A("EdgeData", (int64_t)nedges * 8, 1);
    int* ER = (int*)A("EdgeReversed", nedges, 1);
    int* E2E = (int*)A("ElemToEdge", (int64_t)C->Nelem * 4, 1);
    int* EMS = (int*)A("ElemEdgeMasterSlave", (int64_t)C->Nelem * 4, 1);
    int* EOR = (int*)A("ElemEdgeOrientation", (int64_t)C->Nelem * 4, 1);
    double* x = (double*)A("x", nn, 0);
    double* y = (double*)A("y", nn, 0);
    double* xXi = (double*)A("xXi", nn, 0);
    double* xEta = (double*)A("xEta", nn, 0);
    double* yXi = (double*)A("yXi", nn, 0);
    double* yEta = (double*)A("yEta", nn, 0);
    double* Jac = (double*)A("Jac", nn, 0);
    double* nxv = (double*)A("nx", fn, 0);
    double* nyv = (double*)A("ny", fn, 0);
    double* scal = (double*)A("scal", fn, 0);
    double* sizes = (double*)A("EleSizes", C->Nelem, 0);
    double* b = (double*)A("b", nn, 0);
    double* q = (double*)A("q", nn * 3, 0);
    double* Bx = (double*)A("Bx", nn, 0);
    double* By = (double*)A("By", nn, 0);
    if (rc) { delete C; return rc; }

    // one-dimensional node tables (Mesh.cpp:361-375) and the curved displacement's sine factors
    const double deltaX = (s.xR - s.xL) / s.nelx, deltaY = (s.yR - s.yL) / s.nely;
    std::vector<double> xc((size_t)s.nelx * ngl), yc((size_t)s.nely * ngl), sx(xc.size(), 0.0), sy(yc.size(), 0.0);
    for (int ieX = 0; ieX < s.nelx; ++ieX)
        for (int i = 0; i < ngl; ++i) {
            const double v = s.xL + (xGL[i] + 1.0) / 2.0 * deltaX + ieX * deltaX;
            xc[(size_t)ieX * ngl + i] = v;
            if (curved) sx[(size_t)ieX * ngl + i] = std::sin(M_PI * v / s.perturb_len);
        }
    for (int ieY = 0; ieY < s.nely; ++ieY)
        for (int j = 0; j < ngl; ++j) {
            const double v = s.yL + (xGL[j] + 1.0) / 2.0 * deltaY + ieY * deltaY;
            yc[(size_t)ieY * ngl + j] = v;
            if (curved) sy[(size_t)ieY * ngl + j] = std::sin(M_PI * v / s.perturb_len);
        }
    double *dxc = nullptr, *dyc = nullptr, *dsx = nullptr, *dsy = nullptr;
    int* dbad = nullptr;
    auto tmp = [&](double** p, const std::vector<double>& h) {
        if (!rc) rc = esdg_malloc(ctx, (void**)p, h.size() * sizeof(double));
        if (!rc) rc = esdg_memcpy(ctx, *p, h.data(), h.size() * sizeof(double), ESDG_H2D);
    };
    tmp(&dxc, xc); tmp(&dyc, yc); tmp(&dsx, sx); tmp(&dsy, sy);
    if (!rc) rc = esdg_malloc(ctx, (void**)&dbad, sizeof(int));
    if (!rc) {
        KernelScope ks(ctx, "device_case_setup");
        const Structured m{s.nelx, s.nely, PX ? 1 : 0, PY ? 1 : 0};
        k_structured_faces<<<grid_of(nedges), TPB, 0, ctx->stream>>>(m, nedges, ED, ER, E2E);
        k_element_face_maps<<<grid_of((int64_t)C->Nelem * 4), TPB, 0, ctx->stream>>>(C->Nelem, E2E, ED, EMS, EOR, dbad);
        k_structured_nodes<<<grid_of(nn), TPB, 0, ctx->stream>>>(s.nelx, s.nely, ngl, curved ? 1 : 0, s.perturb_amp, dxc,
                                                                dyc, dsx, dsy, x, y);
        double* dD0 = (double*)C->arrays["D0"].p;
        if (curved) {
            k_mapped_metrics<<<grid_of(nn), TPB, 0, ctx->stream>>>(C->Nelem, ngl, dD0, x, y, xXi, xEta, yXi, yEta, Jac);
            k_face_normals<<<grid_of(fn), TPB, 0, ctx->stream>>>(nedges, ngl, ED, xXi, xEta, yXi, yEta, Jac, nxv, nyv,
                                                                scal);
        } else {
            const double x_xi = 0.5 * deltaX, y_eta = 0.5 * deltaY;
            k_cartesian_metrics<<<grid_of(nn), TPB, 0, ctx->stream>>>(nn, x_xi, y_eta, x_xi * y_eta, xXi, xEta, yXi,
                                                                     yEta, Jac);
            k_cartesian_face_normals<<<grid_of(fn), TPB, 0, ctx->stream>>>(nedges, ngl, ED, x_xi, y_eta, nxv, nyv, scal);
        }
        k_case_state<<<grid_of(nn), TPB, 0, ctx->stream>>>(s, esdg_case_state::flood_modes(s.seed), C->Nelem, ngl2, x, y,
                                                          b, q);
        k_bathymetry_gradient<<<grid_of(nn), TPB, 0, ctx->stream>>>(C->Nelem, ngl, dD0, b, yEta, yXi, xEta, xXi, Jac,
                                                                   Bx, By);
        rc = launch_check(ctx, "device_case_setup");
    }
    if (!rc) rc = esdg_element_sizes(ctx, C->Nelem, Jac, (const double*)C->arrays["w_GL"].p, sizes, &C->min_ele_size);
    int bad = 0;
    if (!rc) rc = esdg_memcpy(ctx, &bad, dbad, sizeof(int), ESDG_D2H);
    if (!rc && bad) rc = set_error(ctx, ESDG_ERR_STATE, "esdg_device_case_create: mesh element side without a face");
    for (void* p : {(void*)dxc, (void*)dyc, (void*)dsx, (void*)dsy, (void*)dbad})
        if (p) (void)esdg_free(ctx, p);
    if (rc) { delete C; return rc; }
    *out = C;
    return ESDG_OK;
}

int esdg_device_case_get_info(const esdg_device_case* c, esdg_case_info* info) {
    if (!c || !info) return set_error(nullptr, ESDG_ERR_ARG, "esdg_device_case_get_info: null argument");
    info->N = c->N; info->ngl = c->ngl; info->Nelem = c->Nelem; info->Nfaces = c->Nfaces;
    info->min_ele_size = c->min_ele_size;
    return ESDG_OK;
}

int esdg_device_case_array(const esdg_device_case* c, const char* name, void** dptr, int64_t* count, int* is_int) {
    if (!c || !name || !dptr || !count || !is_int)
        return set_error(nullptr, ESDG_ERR_ARG, "esdg_device_case_array: null argument");
    auto it = c->arrays.find(name);
    if (it == c->arrays.end()) return set_error(c->ctx, ESDG_ERR_ARG, std::string("no device case array named ") + name);
    *dptr = it->second.p; *count = it->second.count; *is_int = it->second.is_int;
    return ESDG_OK;
}

int esdg_device_case_inputs(const esdg_device_case* c, esdg_inputs* in) {
    if (!c || !in) return set_error(nullptr, ESDG_ERR_ARG, "esdg_device_case_inputs: null argument");
    auto d = [&](const char* n) { return (const double*)c->arrays.at(n).p; };
    auto i = [&](const char* n) { return (const int*)c->arrays.at(n).p; };
    in->Nelem = c->Nelem; in->Nfaces = c->Nfaces;
    in->GLw = d("w_GL"); in->nx = d("nx"); in->ny = d("ny"); in->scal = d("scal");
    in->Yxi = d("yXi"); in->Yeta = d("yEta"); in->Xxi = d("xXi"); in->Xeta = d("xEta");
    in->B = d("b"); in->Bx = d("Bx"); in->By = d("By"); in->D = d("D"); in->Dstrong = d("Dstrong");
    in->Dhat = d("Dhat"); in->Jac = d("Jac"); in->x = d("x"); in->y = d("y"); in->q = d("q");
    in->EleSizes = d("EleSizes"); in->VdmInv = d("VdmInv");
    in->ElemEdgeMasterSlave = i("ElemEdgeMasterSlave"); in->ElemEdgeOrientation = i("ElemEdgeOrientation");
    in->ElemToEdge = i("ElemToEdge"); in->EdgeData = i("EdgeData"); in->EdgeReversed = i("EdgeReversed");
    return ESDG_OK;
}

void esdg_device_case_destroy(esdg_device_case* c) { delete c; }

}  // extern "C"
