// Shared pieces of the temporally blocked 2D loops (tl_fused2d.hip: the
// register-window kernels, the Chebyshev loop and the host entry points;
// tl_ppcg_ring.hip: the PPCG loop with its coefficient rows in LDS).
#pragma once

#include "tl_lanes.h"
#include "tl_ops.h"

namespace tl {
namespace fused {

constexpr int kMaxFused = 10;   // inner steps / iterations per launch

// Raw buffer descriptors (gfx9 family): word 3 of a descriptor with no
// format conversion, the largest range a descriptor is given here, and the
// offset that marks a lane's load as outside it (the hardware returns zeros).
constexpr int kBufferWord3 = 0x00020000;
constexpr unsigned kBufferRange = 0xFFFFFF00u;
constexpr unsigned kBufferOutside = 0xFFFFFFF0u;

template <class T, int S>
struct PpcgCoefs {
    T alpha[S], beta[S];
};

// One launch of the fused PPCG inner loop: S = count steps; MODE 0 reads sd,
// 1 forms sd = r/theta (ppcg_init), 2 first applies cg_calc_ur to r and u.
template <class T>
struct FusedArgs {
    PitchGrid g;
    const double* al;
    const double* be;
    double theta, alpha_ur;
    const double* alpha_dev;   // null: alpha_ur
    const T* sd_in;
    T* sd_out;
    const T* r_in;
    T* r_out;
    const T* u_in;
    T* u_out;
    const T* kx;
    const T* ky;
    const T* p;
    const T* w;
    const char* what;
    // Non-null: the launch also forms the squared 2-norm of the r it writes
    // (the interior): fetched to *rrn, or (norm_nb) left as per-block
    // partials in the reduction slots for the caller's finisher, their count
    // in *norm_nb.  Ring kernel only.
    double* rrn = nullptr;
    unsigned* norm_nb = nullptr;
};

// Compute units of the current device (cached; the runtime uses one device).
int device_cus();

// Tile height and prefetch overrides (tlhip_set_fused_tuning); < 0: default.
extern int g_fused_ty;
extern int g_fused_pf;
// Which PPCG kernel the fused loop launches (tlhip_set_fused_variant).
extern int g_fused_variant;
// Fraction of the device's resident waves the LDS-ring kernel's tiles fill
// (tlhip_set_ring_fill; <= 0: default).
extern double g_ring_fill;

// The PPCG loop with the coefficient rows in per-wave LDS rings
// (tl_ppcg_ring.hip); count 1..kMaxFused, mode 0..2 as FusedArgs.  Mode 3
// runs the Chebyshev loop the same way (2d/drivers/cheby_driver.c turns:
// mirror of u, r = u0 - A u, p = alpha_k p + beta_k r, u += p) with the
// fields in the PPCG roles: sd_in/sd_out = u, r_in = u0, u_in/u_out = p,
// r_out = r.
template <class T>
int launch_ppcg_ring(int count, int mode, const FusedArgs<T>& a);

}  // namespace fused
}  // namespace tl
