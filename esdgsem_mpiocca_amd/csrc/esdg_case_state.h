// esdg_case_state.h -- the benchmark configurations' bathymetry and initial state at one node,
// shared by the host setup (esdg_setup.cpp, g++) and the device case builder
// (esdg_geometry.hip, hipcc), so both evaluate the same formula with the same operation order.
// Host and device results can differ only where the formula calls sin, cos or exp (the two math
// libraries round those differently in the last bit); everything else is IEEE arithmetic.
//
// The states are new (they define the benchmark configurations, not a reference input):
//   GAUSSIAN_BUMP: b = 0.8 exp(-5((x-1)^2+(y-1)^2)); h = 1 - b + amp*exp(-20((x-0.7)^2+(y-1)^2)).
//   DAM_BREAK:     b = 0.1 sin(x) cos(y); h = 2 for x < p0 (default 5) else 1, plus amp*(2U-1);
//                  with p1 (height ratio) or p2 (angle) set: h = p1 on the side (x - p0, y - yc).n < 0
//                  of the line through (p0, yc) with normal n = (cos p2, sin p2), 1 elsewhere.
//   FLOOD:         b = (0.3/4) sum_{m<16} cos(2 pi (k_m x + l_m y)/Lx + phi_m), integers
//                  k_m, l_m in [0,3] (not both 0), phi_m in [0, 2 pi), all seeded; water level
//                  p0 for x < p1, level p2 elsewhere; h = max(0, level - b).  Wet/dry fronts.
//   OCEAN_UPLIFT:  depth 4000 m over the first 80% of x, linear shelf to 50 m over the last 20%;
//                  b = 4000 - depth, h = depth + eta, eta = amp*exp(-(xr/a)^2-(yr/c)^2) with
//                  rotation p2 (radians), axes a = p0, c = p1 (metres), centre at 40% / 50%.
//   CONSTANT:      h = p0, hu = p1, hv = p2, b = 0.
//   CONVERGENCE:   p0 = the reference's Testcase number.  1 (or 0): b = 2 + 0.5 sin(2 pi y) + cos(2 pi x),
//                  h = 8 + cos x sin y - b; 8: b = 2 + sin y + cos x, h = 20 + cos x sin y - b;
//                  88: b = 0, h = 20 + cos(x/1000) sin(y/1000) - b; 89: b = 0,
//                  h = 20000 + cos(1e-6 x) sin(1e-6 y) - b.  hu = h/2, hv = 3h/2 (t = 0)
//                  (SW2D::InitQNodal cases 1, 8, 88, 89, src/SW2D.cpp:312-318, 388-393, 760-775,
//                  with the bathymetry of the same Testcase, :829-833, 858-862).
#pragma once

#include <cmath>
#include <cstdint>

#include "esdg_setup.h"

#if defined(__HIPCC__)
#define ESDG_HD __host__ __device__
#else
#define ESDG_HD
#endif

namespace esdg_case_state {

// Counter-based random numbers: SplitMix64 of (seed, index), uniform in [0,1).
ESDG_HD inline uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
ESDG_HD inline double uniform01(uint64_t seed, uint64_t idx) {
    return (double)(splitmix64(seed ^ splitmix64(idx)) >> 11) * (1.0 / 9007199254740992.0);
}

// The 16 seeded Fourier modes of the FLOOD bathymetry.
struct FloodModes { double km[16], lm[16], ph[16]; };

inline FloodModes flood_modes(uint64_t seed) {
    FloodModes F{};
    for (int m = 0; m < 16; ++m) {
        int k, l;
        uint64_t t = 0;
        do {
            k = (int)(uniform01(seed + 7, 3 * m + 0 + 1000 * t) * 4.0);
            l = (int)(uniform01(seed + 7, 3 * m + 1 + 1000 * t) * 4.0);
            ++t;
        } while (k == 0 && l == 0);
        F.km[m] = k; F.lm[m] = l;
        F.ph[m] = 2.0 * M_PI * uniform01(seed + 7, 3 * m + 2);
    }
    return F;
}

// True when the spec names a case kind (and, for CONVERGENCE, a Testcase) that node_state knows.
inline bool valid_case(const esdg_case_spec& s) {
    switch (s.case_kind) {
        case ESDG_CASE_GAUSSIAN_BUMP: case ESDG_CASE_DAM_BREAK: case ESDG_CASE_FLOOD:
        case ESDG_CASE_OCEAN_UPLIFT: case ESDG_CASE_CONSTANT:
            return true;
        case ESDG_CASE_CONVERGENCE: {
            const int tc = (int)s.p0;
            return tc == 0 || tc == 1 || tc == 8 || tc == 88 || tc == 89;
        }
        default:
            return false;
    }
}

// Bathymetry b and state (h, hu, hv) at the node (X, Y) whose h has the flat state index qid
// (seeds the dam break's noise).  bmesh: the mesh file's bathymetry at the node, used in place of
// the case's formula when has_bmesh.  The spec must pass valid_case.
ESDG_HD inline void node_state(const esdg_case_spec& s, const FloodModes& F, double X, double Y, uint64_t qid,
                               bool has_bmesh, double bmesh, double* b_out, double* h_out, double* hu_out,
                               double* hv_out) {
    const double Lx = s.xR - s.xL, Ly = s.yR - s.yL;
    double bb = 0.0, h = 0.0, hu = 0.0, hv = 0.0;
    auto pick = [&](double formula) { return has_bmesh ? bmesh : formula; };
    switch (s.case_kind) {
        case ESDG_CASE_GAUSSIAN_BUMP:
            bb = pick(0.8 * exp(-5.0 * ((X - 1.0) * (X - 1.0) + (Y - 1.0) * (Y - 1.0))));
            h = 1.0 - bb + s.amp * exp(-20.0 * ((X - 0.7) * (X - 0.7) + (Y - 1.0) * (Y - 1.0)));
            break;
        case ESDG_CASE_DAM_BREAK: {
            const double xd = (s.p0 != 0.0) ? s.p0 : 5.0;
            bb = pick(0.1 * sin(X) * cos(Y));
            if (s.p1 == 0.0 && s.p2 == 0.0) {
                h = (X < xd) ? 2.0 : 1.0;
            } else {
                // step of height ratio p1 (h = p1 | 1) across the line through (xd, centre y)
                // with normal angle p2
                const double ratio = (s.p1 != 0.0) ? s.p1 : 2.0;
                const double side = (X - xd) * cos(s.p2) + (Y - 0.5 * (s.yL + s.yR)) * sin(s.p2);
                h = (side < 0.0) ? ratio : 1.0;
            }
            h += s.amp * (2.0 * uniform01(s.seed, qid) - 1.0);
            break;
        }
        case ESDG_CASE_FLOOD: {
            double acc = 0.0;
            for (int m = 0; m < 16; ++m)
                acc += cos(2.0 * M_PI * (F.km[m] * (X - s.xL) + F.lm[m] * (Y - s.yL)) / Lx + F.ph[m]);
            bb = pick(0.3 / 4.0 * acc);
            const double level = (X < s.p1) ? s.p0 : s.p2;
            h = level - bb;
            if (h < 0.0) h = 0.0;
            break;
        }
        case ESDG_CASE_OCEAN_UPLIFT: {
            const double xs = s.xL + 0.8 * Lx;
            double depth = 4000.0;
            if (X > xs) depth = 4000.0 + (50.0 - 4000.0) * (X - xs) / (0.2 * Lx);
            bb = pick(4000.0 - depth);
            const double cx = s.xL + 0.4 * Lx, cy = s.yL + 0.5 * Ly;
            const double c = cos(s.p2), sn = sin(s.p2);
            const double xr = c * (X - cx) + sn * (Y - cy);
            const double yr = -sn * (X - cx) + c * (Y - cy);
            h = depth + s.amp * exp(-(xr / s.p0) * (xr / s.p0) - (yr / s.p1) * (yr / s.p1));
            break;
        }
        case ESDG_CASE_CONVERGENCE: {
            const int tc = (int)s.p0;
            if (tc == 8) {
                bb = pick(2.0 + sin(Y) + cos(X));
                h = 20.0 + cos(X) * sin(Y) * cos(0.0) - bb;
            } else if (tc == 88) {
                bb = pick(0.0);
                h = 20.0 + cos(0.001 * X) * sin(0.001 * Y) * cos(0.0) - bb;
            } else if (tc == 89) {
                bb = pick(0.0);
                h = 20000.0 + cos(0.000001 * X) * sin(0.000001 * Y) * cos(0.0) - bb;
            } else {
                bb = pick(2.0 + 0.5 * sin(2.0 * M_PI * Y) + cos(2.0 * M_PI * X));
                h = 8.0 + cos(X) * sin(Y) * cos(0.0) - bb;
            }
            hu = h * 0.5;
            hv = h * 1.5;
            break;
        }
        case ESDG_CASE_CONSTANT:
            bb = pick(0.0); h = s.p0; hu = s.p1; hv = s.p2;
            break;
        default:
            break;
    }
    *b_out = bb;
    *h_out = h;
    *hu_out = hu;
    *hv_out = hv;
}

}  // namespace esdg_case_state
