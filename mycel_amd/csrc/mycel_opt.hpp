// Device filter, OC bisection and SIMP helpers (opt.hip).
#pragma once

#include "mycel_internal.hpp"

namespace mycel {

// DensityFilter on the device: tap table + per-element 1/Hs.
struct DeviceFilter {
    int nx, ny, nz, n;
    double rmin;
    int n_taps = 0;
    int64_t nnz = 0;
    hipStream_t stream;
    DevBuf<int4> d_taps;
    DevBuf<double> d_tap_h;
    DevBuf<int32_t> d_to_active;  // box element -> active index or -1
    DevBuf<int32_t> d_active;
    DevBuf<double> d_inv_hs;
    DevBuf<double> d_s;           // S_i
    DevBuf<double> d_in, d_out;   // staging for host-pointer calls
    DeviceFilter(int nx, int ny, int nz, const uint8_t *domain, const int32_t *active, int n_active, double rmin,
                 hipStream_t s);
    // d_in == nullptr applies the operator to a vector of ones.
    void apply(const double *d_in, double *d_out, bool adjoint, hipStream_t s);
};

struct OcState {
    double target, l1, l2, v1, v2, lmid;
    int phase, steps, bis, pending, evals, pad;
};

struct OcResult {
    double lambda;
    int done, evals;
};

struct OcWork {
    DevBuf<OcState> states;
    DevBuf<double> parts, spart;
    DevBuf<OcResult> res;
    OcResult *h_res = nullptr;
    OcWork();
    ~OcWork();
    OcWork(const OcWork &) = delete;
    OcWork &operator=(const OcWork &) = delete;
};

// Largest design vector of oc_update_noalloc: its partial sums (5 doubles per 1,024 elements) fill a fixed
// device array of 640 KiB + 256 B.
constexpr int kOcNoAllocMaxElements = 1 << 24;
constexpr int kOcNoAllocMaxBlocks = kOcNoAllocMaxElements / 1024;

double oc_update_device(OcWork &w, const double *d_rho, const double *d_dc, const double *d_s,
                        const uint8_t *d_passive, double *d_rho_new, int n, double vf, double move, double rho_min,
                        double lambda_guess, int max_bisect, double vol_tol, hipStream_t s, int *evals);

// OcOptimizer.Update on device pointers with its scratch in a fixed device array: no allocation.
double oc_update_noalloc(const double *d_rho, const double *d_dc, const double *d_s, const uint8_t *d_passive,
                         double *d_rho_new, int n, double vf, double move, double rho_min, double lambda_guess,
                         int max_bisect, double vol_tol, hipStream_t s);

void launch_scatter_density(int na, const int32_t *d_active, const double *d_rho_bar, double *d_density,
                            hipStream_t s);
void launch_force_passive(int na, const uint8_t *d_passive, double *d_rho, hipStream_t s);
void launch_project(int n, const uint8_t *d_flags, const double *d_f, double *d_b, hipStream_t s);
void launch_add_into(int n, const double *d_src, double *d_dst, hipStream_t s);
void simp_metrics(int na, const double *d_rho, const double *d_rho_old, const double *d_rho_bar,
                  const double *d_grey_src, DevBuf<double> &part, double *h_out3, hipStream_t s);
void launch_heaviside_project(int n, const double *d_rt, double *d_rb, double beta, hipStream_t s);
void launch_heaviside_chain(int n, const double *d_rt, double *d_dc, double beta, hipStream_t s);
void compliance_maxdisp(int n_nodes, const uint8_t *d_active_nodes, const double *d_f, const double *d_u,
                        DevBuf<double> &part, double *h_out2, hipStream_t s);

}  // namespace mycel
