/*
 * esdg_device_setup.h -- mesh geometry and benchmark cases built on the device (gfx950).
 *
 * The reference builds its geometry on the host, once per run: node placement and face
 * numbering (Mesh::GenerateMesh, src/Mesh.cpp:316-696), metric terms, J and side normals
 * (Mesh::ConstructMappedGeometry, src/Mesh.cpp:1667-1808), per-face normals (src/Mesh.cpp:282-307),
 * 1/J (main.cpp:700), element sizes (basis::calcElementSizes, src/basis.cpp:939-966), the
 * single-rank EdgeData / ElemToEdge maps (src/MeshPartitioning.cpp:104-132, main.cpp:650-716) and
 * the split-form bathymetry gradient (SW2D::CalcBDerivatives, src/SW2D.cpp:1048-1105).  Here the
 * same arithmetic runs on the device, so a 2048^2 mesh at N=7 is ready in well under a second,
 * and the arrays never cross PCIe.
 *
 * Given the same node coordinates and bathymetry, every geometry array equals the host setup's
 * (esdg_setup.h) bit for bit: the kernels perform the host's IEEE operations in its order (no
 * FMA contraction).  The structured mesh kinds' node coordinates are bit-equal too: the curved
 * displacement a*sin(pi x/w)*sin(pi y/w) factors into one-dimensional tables that the host
 * evaluates.  The initial states (esdg_case_state.h, shared with the host) call sin/cos/exp per
 * node; those can differ from the host's in the last bit.
 *
 * All pointers are device pointers unless named host_*.  Launches run on the context's stream;
 * functions that return a host value synchronise.  Return 0 or a negative esdg_status, message in
 * esdg_last_error(ctx).
 */
#ifndef ESDG_DEVICE_SETUP_H
#define ESDG_DEVICE_SETUP_H

#include "esdg.h"
#include "esdg_setup.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- geometry kernels (each also usable on arrays built elsewhere, e.g. an ISM mesh's nodes) ---- */

/* Metric terms of ConstructMappedGeometry (src/Mesh.cpp:1722-1742) from the nodal coordinates
 * x, y [Nelem][ngl][ngl] and the nodal derivative matrix D0 (ngl x ngl), and Jac = 1/J
 * (main.cpp:700). */
int esdg_mapped_metrics(esdg_ctx* ctx, int Nelem, const double* D0, const double* x, const double* y,
                        double* xXi, double* xEta, double* yXi, double* yEta, double* Jac);
/* Outward unit normals and surface Jacobians of each face, evaluated on the side of its left
 * element (EdgeData[8f], EdgeData[8f+2]) as ConstructMappedGeometry does for every side
 * (src/Mesh.cpp:1744-1806) and Mesh.cpp:282-307 then gathers: nx, ny, scal [Nfaces][ngl]. */
int esdg_face_normals(esdg_ctx* ctx, int Nfaces, const int* EdgeData, const double* xXi, const double* xEta,
                      const double* yXi, const double* yEta, const double* Jac, double* nx, double* ny,
                      double* scal);
/* basis::calcElementSizes (src/basis.cpp:939-966): EleSizes[e] = sum_ji (1/Jac) w_i w_j and
 * *host_min_ele_size = sqrt(min_e EleSizes[e]).  Blocking. */
int esdg_element_sizes(esdg_ctx* ctx, int Nelem, const double* Jac, const double* w_GL, double* EleSizes,
                       double* host_min_ele_size);
/* SW2D::CalcBDerivatives (src/SW2D.cpp:1048-1105): the split-form bathymetry gradient. */
int esdg_bathymetry_gradient(esdg_ctx* ctx, int Nelem, const double* D0, const double* B, const double* Yeta,
                             const double* Yxi, const double* Xeta, const double* Xxi, const double* Jac,
                             double* Bx, double* By);

/* ---- a whole case on the device ------------------------------------------------------------- */

typedef struct esdg_device_case esdg_device_case;

/* The case esdg_case_create would build, for the structured mesh kinds (ESDG_MESH_CARTESIAN and
 * ESDG_MESH_CURVED_SINE; ISM files are read on the host): basis (host, ngl^2 values), face
 * numbering and connectivity, nodes, metrics, normals, sizes, bathymetry, initial state and Bx/By,
 * all on the device of ctx (whose ngl must be spec->N + 1). */
int  esdg_device_case_create(esdg_ctx* ctx, const esdg_case_spec* spec, esdg_device_case** out);
int  esdg_device_case_get_info(const esdg_device_case* c, esdg_case_info* info);
/* Named device arrays, the names of esdg_case_array.  The pointer stays owned by the case. */
int  esdg_device_case_array(const esdg_device_case* c, const char* name, void** dptr, int64_t* count,
                            int* is_int);
/* The solver inputs of the case, as device pointers (for esdg_solver_create_dev). */
int  esdg_device_case_inputs(const esdg_device_case* c, esdg_inputs* in);
void esdg_device_case_destroy(esdg_device_case* c);

/* esdg_solver_create with inputs that are device pointers (copied device to device). */
int esdg_solver_create_dev(esdg_ctx* ctx, const esdg_inputs* in, const esdg_run_params* rp,
                           esdg_solver** out);

#ifdef __cplusplus
}
#endif
#endif
