// C ABI of libmycel_hip.so (include/mycel_hip.h): handle lifetime, argument validation with the
// reference's exception semantics, host<->HBM staging, StaticAnalysis and the SIMP driver.
// All arithmetic on the path runs in the kernels of fem.hip / pcg.hip / opt.hip; the host code
// here validates, stages and sequences.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <vector>

#include "mycel_internal.hpp"
#include "mycel_opt.hpp"

namespace mycel {

namespace {
thread_local std::string g_last_error;
thread_local int g_device = 0;
std::mutex g_init_mutex;
inline int cdiv(int a, int b) { return (a + b - 1) / b; }

__global__ void k_select_fixed(int n, const uint8_t *__restrict__ flags, const double *__restrict__ ku,
                               double *__restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] = (flags[i] & kFlagFixed) ? ku[i] : 0.0;
}

template <typename F>
int32_t guarded(F &&f) {
    try {
        f();
        return MYCEL_OK;
    } catch (const Error &e) {
        g_last_error = e.what();
        return e.status;
    } catch (const std::bad_alloc &) {
        g_last_error = "host allocation failed";
        return MYCEL_E_OUT_OF_MEMORY;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return MYCEL_E_DEVICE;
    }
}

void require(bool cond, int32_t status, const char *msg) {
    if (!cond) throw Error(status, msg);
}

hipStream_t pick(mycel_op *op, void *stream) { return stream ? static_cast<hipStream_t>(stream) : op->stream; }

struct DeviceGuard {
    int prev = 0;
    explicit DeviceGuard(int dev) {
        MYCEL_HIP(hipGetDevice(&prev));
        if (prev != dev) MYCEL_HIP(hipSetDevice(dev));
    }
    ~DeviceGuard() { (void)hipSetDevice(prev); }
};

// A double in .NET's "E3" format with the invariant culture: d.dddE+ddd (at least three exponent digits).
std::string net_e3(double v) {
    if (std::isnan(v)) return "NaN";
    if (std::isinf(v)) return v > 0 ? "Infinity" : "-Infinity";
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.3E", v);
    std::string t(buf);
    const size_t e = t.find('E');
    if (e == std::string::npos || e + 2 > t.size()) return t;
    std::string digits = t.substr(e + 2);
    while (digits.size() < 3) digits.insert(digits.begin(), '0');
    return t.substr(0, e + 2) + digits;
}

}  // namespace

// PcgResult.Message (Fem/PcgSolver.cs:149,207,255,270,306,344,367); empty for a converged solve (Message null).
std::string pcg_message_text(int code, double value) {
    switch (code) {
        case MYCEL_PCG_RHS_ZERO: return "RHS is zero. Solution is zero.";
        case MYCEL_PCG_INITIAL_CONVERGED: return "Initial guess already converged.";
        case MYCEL_PCG_CURVATURE_BREAKDOWN:
            return "PCG curvature breakdown (p\u1D40Ap = " + net_e3(value) +
                   " <= 0). Is the model not sufficiently constrained?";
        case MYCEL_PCG_STEP_BREAKDOWN:
            return "PCG step breakdown (alpha = " + net_e3(value) + "). Is the model not sufficiently constrained?";
        case MYCEL_PCG_STAGNATION:
            return "PCG stagnated: residual stopped decreasing. Is the model not sufficiently constrained?";
        case MYCEL_PCG_DIRECTION_BREAKDOWN:
            return "PCG search direction breakdown (beta = " + net_e3(value) +
                   "). Is the model not sufficiently constrained?";
        case MYCEL_PCG_MAX_ITERATIONS:
            return "PCG reached maximum iterations without converging. Is the model not sufficiently constrained?";
        case MYCEL_PCG_CANCELED: return "The operation was canceled.";
        default: return "";
    }
}

// InvalidOperationException text of a non-converged solve (TopologyOptimizer.cs:249-253,
// StaticAnalysis.cs:88-92).
static std::string not_converged_text(const mycel_pcg_result &r) {
    return "FEM solver did not converge in " + std::to_string(r.iterations) + " iterations (" +
           pcg_message_text(r.message, r.message_value) + "). Is the model not sufficiently constrained?";
}

void set_last_error(const std::string &msg) { g_last_error = msg; }

// ------------------------------------------------------------------------------------------------
// FeModel + FemOperator construction (Fem/FeModel.cs:43-184, Fem/FemOperator.cs:33-108)
// ------------------------------------------------------------------------------------------------
mycel_op *op_create(const mycel_grid &g, const uint8_t *domain, const int32_t *fixed, int n_fixed, double e0,
                           double nu, double emin_ratio, double penal, const double *density) {
    require(g.nx > 0 && g.ny > 0 && g.nz > 0, MYCEL_E_OUT_OF_RANGE, "Grid dimensions must be strictly positive.");
    require(g.h > 0, MYCEL_E_OUT_OF_RANGE, "Element edge length h must be positive.");
    const long long total = static_cast<long long>(g.nx) * g.ny * g.nz;
    const long long nodes = static_cast<long long>(g.nx + 1) * (g.ny + 1) * (g.nz + 1);
    require(3 * nodes < (1LL << 31) - 64, MYCEL_E_OUT_OF_RANGE, "grid too large for 32-bit DOF indices");
    // The K*u gather addresses a few node layers with 32-bit byte offsets: at most 2^25 nodes per layer.
    require(static_cast<long long>(g.nx + 1) * (g.ny + 1) <= (1LL << 25), MYCEL_E_OUT_OF_RANGE,
            "grid cross-section (nx+1)(ny+1) beyond 2^25 nodes");
    (void)total;
    auto op = std::make_unique<mycel_op>();
    op->device = g_device;
    op->d = make_dims(g.nx, g.ny, g.nz);
    op->h = g.h;
    op->e0 = e0;
    op->nu = nu;
    op->emin = e0 * emin_ratio;  // FeMaterial.Emin (Fem/FeMaterial.cs:16)
    op->penal = penal;
    const Dims &d = op->d;

    op->domain.assign(domain, domain + d.n_elem);
    op->active_nodes.assign(d.n_nodes, 0);
    op->active.reserve(d.n_elem);
    for (int iz = 0; iz < d.nz; iz++)
        for (int iy = 0; iy < d.ny; iy++)
            for (int ix = 0; ix < d.nx; ix++) {
                const int e = ix + d.nx * (iy + d.ny * iz);
                if (!domain[e]) continue;
                op->active.push_back(e);
                const int n0 = ix + d.nx1 * (iy + d.ny1 * iz);
                for (int oz = 0; oz < 2; oz++)
                    for (int oy = 0; oy < 2; oy++)
                        for (int ox = 0; ox < 2; ox++) op->active_nodes[n0 + ox + d.nx1 * (oy + d.ny1 * oz)] = 1;
            }
    op->n_active_nodes = static_cast<int>(std::count(op->active_nodes.begin(), op->active_nodes.end(), 1));
    op->dof_flags.assign(d.n_dofs, 0);
    for (int n = 0; n < d.n_nodes; n++)
        if (!op->active_nodes[n]) op->dof_flags[3 * n] = op->dof_flags[3 * n + 1] = op->dof_flags[3 * n + 2] = kFlagConstrained;
    std::set<int> act_fixed;
    op->fixed.assign(fixed, fixed + n_fixed);
    for (int i = 0; i < n_fixed; i++) {
        const int dof = fixed[i];
        if (dof < 0 || dof >= d.n_dofs)
            throw Error(MYCEL_E_OUT_OF_RANGE, "Fixed DOF " + std::to_string(dof) + " is out of bounds [0, " +
                                                  std::to_string(d.n_dofs - 1) + "].");
        op->dof_flags[dof] |= kFlagConstrained | kFlagFixed;
        if (op->active_nodes[dof / 3]) act_fixed.insert(dof);
    }
    op->active_fixed = static_cast<int>(act_fixed.size());

    hex8_ke(g.h, nu, op->ke);
    for (int k = 0; k < 24; k++) op->diag_ke[k] = op->ke[k * 24 + k];
    // Ke as symmetry-class values; the identity Ke = sign * v[class] is checked on this Ke.
    double kmax = 0.0;
    for (int i = 0; i < 576; i++) kmax = std::max(kmax, std::fabs(op->ke[i]));
    for (int c = 0; c < kKeClasses; c++) {
        op->ks64.v[c] = op->ke[kKeRep[c][0] * 24 + kKeRep[c][1]];
        op->ks32.v[c] = static_cast<float>(op->ks64.v[c]);
    }
    for (int r = 0; r < 24; r++)
        for (int c = 0; c < 24; c++)
            if (std::fabs(op->ke[r * 24 + c] - kKeSign[r][c] * op->ks64.v[kKeClass[r][c]]) > 1e-14 * kmax)
                throw Error(MYCEL_E_ARGUMENT, "element stiffness does not have the isotropic cube symmetry");
    if (ke_mono_values(op->ke, op->km64) > 1e-12)
        throw Error(MYCEL_E_ARGUMENT, "element stiffness does not have the monomial-basis pattern of the cube");
    op->km32 = {static_cast<float>(op->km64.a), static_cast<float>(op->km64.b), static_cast<float>(op->km64.c),
                static_cast<float>(op->km64.d), static_cast<float>(op->km64.e), static_cast<float>(op->km64.f),
                static_cast<float>(op->km64.h), static_cast<float>(op->km64.q)};
    double cst[780];
    std::copy(op->ke, op->ke + 576, cst);
    std::copy(op->diag_ke, op->diag_ke + 24, cst + 576);
    hex8_bc(1.0, cst + 600);
    constitutive(1.0, nu, cst + 744);
    float ke32[576];
    for (int i = 0; i < 576; i++) ke32[i] = static_cast<float>(op->ke[i]);

    MYCEL_HIP(hipSetDevice(op->device));
    MYCEL_HIP(hipStreamCreateWithFlags(&op->stream, hipStreamNonBlocking));
    hipStream_t s = op->stream;
    op->d_ke.resize(780);
    op->d_ke.upload(cst, 780, s);
    op->d_ke32.resize(576);
    op->d_ke32.upload(ke32, 576, s);
    op->d_ep.resize(d.n_pad);
    op->d_ep32.resize(d.n_pad);
    op->d_flags.resize(d.n_dofs);
    op->d_flags.upload(op->dof_flags.data(), d.n_dofs, s);
    op->d_domain.resize(d.n_elem);
    op->d_domain.upload(domain, d.n_elem, s);
    op->d_active_nodes.resize(d.n_nodes);
    op->d_active_nodes.upload(op->active_nodes.data(), d.n_nodes, s);
    op->d_active.resize(std::max<size_t>(1, op->active.size()));
    if (!op->active.empty()) op->d_active.upload(op->active.data(), op->active.size(), s);
    op->d_density.resize(d.n_elem);
    if (density) {
        op->d_density.upload(density, d.n_elem, s);
    } else {
        std::vector<double> ones(d.n_elem, 1.0);  // FeModel.cs:74-78
        op->d_density.upload(ones.data(), d.n_elem, s);
    }
    op->d_diag.resize(d.n_dofs);
    op->d_inv_diag.resize(d.n_dofs);
    op->d_inv_diag32.resize(d.n_dofs);
    op->d_x.resize(d.n_dofs);
    op->d_y.resize(d.n_dofs);
    op_update_e(op.get(), op->d_density.get(), penal, s);
    MYCEL_HIP(hipStreamSynchronize(s));
    return op.release();
}

void op_destroy(mycel_op *op) {
    if (!op) return;
    (void)hipSetDevice(op->device);
    if (op->stream) (void)hipStreamSynchronize(op->stream);
    pcg_free(op->pcg);
    op->pcg = nullptr;
    ord_free(op->ord);
    op->ord = nullptr;
    gmg_free(op->gmg);
    op->gmg = nullptr;
    if (op->stream) (void)hipStreamDestroy(op->stream);
    delete op;
}

// ------------------------------------------------------------------------------------------------
// SIMP driver (Optimization/TopologyOptimizer.cs:16-407), with optional p-continuation and Heaviside projection.
// ------------------------------------------------------------------------------------------------
void simp_run(mycel_op *op, const double *f, int n_cases, const mycel_simp_options &o, const uint8_t *passive_mask,
                     const int32_t *patch_nodes, int n_patch, const float *init_density, int init_len,
                     const double *warm_u, mycel_snapshot_cb cb, void *user, mycel_simp_result *result,
                     mycel_snapshot *snaps, int snaps_cap, double *rho_out, double *rho_bar_out, double *u_out) {
    using clock = std::chrono::steady_clock;
    auto ms_since = [](clock::time_point t) {
        return std::chrono::duration<double, std::milli>(clock::now() - t).count();
    };
    const Dims &d = op->d;
    const int na = static_cast<int>(op->active.size());
    hipStream_t s = op->stream;
    *result = mycel_simp_result{};
    if (n_cases < 1) throw Error(MYCEL_E_ARGUMENT, "at least one load case is required");
    const size_t nd = static_cast<size_t>(d.n_dofs), ndc = nd * static_cast<size_t>(n_cases);

    double rmin = o.filter_radius_mm / op->h;  // TopologyOptimizer.cs:30-39
    if (rmin < 1.5 && o.enforce_min_filter_radius) rmin = 1.5;

    // Passive elements (:46-84).
    std::vector<uint8_t> passive(std::max(1, na), 0);
    if (passive_mask)
        for (int i = 0; i < na; i++)
            if (passive_mask[op->active[i]]) passive[i] = 1;
    if (o.protect_boundary_hinges) {
        std::vector<int32_t> pn(patch_nodes, patch_nodes + n_patch);
        for (int dof : op->fixed) pn.push_back(dof / 3);
        std::vector<uint8_t> shell(d.n_elem, 0);
        mycel_grid g{d.nx, d.ny, d.nz, op->h};
        if (mycel_passive_region(&g, pn.data(), static_cast<int>(pn.size()), 1, shell.data()) != MYCEL_OK)
            throw Error(MYCEL_E_ARGUMENT, g_last_error);
        for (int i = 0; i < na; i++)
            if (shell[op->active[i]]) passive[i] = 1;
    }

    // Initial design (:91-122).
    std::vector<double> rho(std::max(1, na));
    if (o.initial_design) {  // FP64 design (exact resume); same length rules as InitialDensity
        if (o.initial_design_len == d.n_elem)
            for (int i = 0; i < na; i++) rho[i] = o.initial_design[op->active[i]];
        else if (o.initial_design_len == na)
            for (int i = 0; i < na; i++) rho[i] = o.initial_design[i];
        else
            throw Error(MYCEL_E_ARGUMENT, "initial_design length (" + std::to_string(o.initial_design_len) +
                                              ") must match either TotalElements (" + std::to_string(d.n_elem) +
                                              ") or ActiveElementCount (" + std::to_string(na) + ").");
    } else if (init_density) {
        if (init_len == d.n_elem)
            for (int i = 0; i < na; i++) rho[i] = init_density[op->active[i]];
        else if (init_len == na)
            for (int i = 0; i < na; i++) rho[i] = init_density[i];
        else
            throw Error(MYCEL_E_ARGUMENT, "InitialDensity length (" + std::to_string(init_len) +
                                              ") must match either TotalElements (" + std::to_string(d.n_elem) +
                                              ") or ActiveElementCount (" + std::to_string(na) + ").");
    } else {
        for (int i = 0; i < na; i++) rho[i] = passive[i] ? 1.0 : o.volume_fraction;
    }
    for (int i = 0; i < na; i++)
        if (passive[i]) rho[i] = 1.0;

    DeviceFilter flt(d.nx, d.ny, d.nz, op->domain.data(), op->active.data(), na, rmin, s);
    const int nab = std::max(1, na);
    DevBuf<double> d_rho(nab), d_rho_bar(nab), d_rho_old(nab), d_dc_bar(nab), d_dc(nab), d_part;
    DevBuf<double> d_dc_k(n_cases > 1 ? nab : 0);
    std::vector<mycel_pcg_result> case_res(n_cases);
    const bool heav = o.enable_heaviside != 0;
    // Island deflation for this run's multigrid solves; the operator's own setting comes back afterwards.
    struct RatioGuard {
        mycel_op *op;
        double saved;
        ~RatioGuard() {
            if (op->island_ratio != saved) {
                op->island_ratio = saved;
                op->epoch++;  // the next solve sets the hierarchy up again without this run's islands
            }
        }
    } ratio_guard{op, op->island_ratio};
    if (o.island_deflation > 0.0) {
        if (!(o.island_deflation <= 1.0)) throw Error(MYCEL_E_OUT_OF_RANGE, "island_deflation must be in [0, 1]");
        op->island_ratio = o.island_deflation;
    }
    // The reference's evaluation order throughout (PCG, compliance, OC volumes, snapshot sums).
    const bool ordered = o.precision == MYCEL_PRECISION_FP64_ORDERED;
    DevBuf<double> d_rho_tilde(heav ? nab : 0);
    double beta = o.resume_state && o.resume_beta > 0.0 ? o.resume_beta : 1.0;  // :156, or the resumed run's
    DevBuf<uint8_t> d_passive(nab), d_active_nodes(d.n_nodes);
    DevBuf<double> d_u(ndc), d_f(ndc), d_b(ndc);
    d_rho.upload(rho.data(), nab, s);
    d_passive.upload(passive.data(), nab, s);
    d_active_nodes.upload(op->active_nodes.data(), d.n_nodes, s);
    d_f.upload(f, ndc, s);
    if (warm_u) d_u.upload(warm_u, ndc, s);
    else MYCEL_HIP(hipMemsetAsync(d_u.get(), 0, sizeof(double) * ndc, s));
    for (int k = 0; k < n_cases; k++)  // b = P F (:146-152)
        launch_project(d.n_dofs, op->d_flags.get(), d_f.get() + k * nd, d_b.get() + k * nd, s);
    OcWork oc;
    std::vector<double> h_rho_bar, h_rho;
    if (cb) {
        h_rho_bar.resize(nab);
        h_rho.resize(nab);
    }

    double lambda = o.initial_lambda > 0.0 && std::isfinite(o.initial_lambda) ? o.initial_lambda : 1.0;
    // A resumed run continues the tolerance schedule and the stop test from the saved state.
    const bool resumed = o.resume_state != 0;
    double max_change = resumed ? o.resume_max_change : 1.0;
    double hist[6];
    int hist_n = 0;
    if (resumed) {
        if (o.resume_history_len < 0 || o.resume_history_len > 6)
            throw Error(MYCEL_E_ARGUMENT, "resume_history_len must be in [0, 6]");
        for (int k = 0; k < o.resume_history_len; k++) hist[hist_n++] = o.resume_history[k];
    }
    int n_snaps = 0;
    const int start = o.start_iteration, max_it = o.max_iterations;
    for (int iter = start; iter <= max_it; iter++) {
        const auto t_total = clock::now();
        double penal = o.penal;
        if (o.continuation) penal = iter < 20 ? 1.0 : (iter < 40 ? 2.0 : o.penal);
        launch_force_passive(na, d_passive.get(), d_rho.get(), s);
        const auto t_filter = clock::now();
        if (heav) {  // :193-197
            flt.apply(d_rho.get(), d_rho_tilde.get(), false, s);
            launch_heaviside_project(na, d_rho_tilde.get(), d_rho_bar.get(), beta, s);
        } else {
            flt.apply(d_rho.get(), d_rho_bar.get(), false, s);
        }
        MYCEL_HIP(hipStreamSynchronize(s));
        const double filter_ms = ms_since(t_filter);
        launch_scatter_density(na, op->d_active.get(), d_rho_bar.get(), op->d_density.get(), s);
        op_update_e(op, op->d_density.get(), penal, s);
        double tol;
        if (iter == start && !resumed) {
            tol = o.solver_tol_initial;
        } else {  // :219-223
            const double a = std::clamp((max_change - 0.01) / 0.09, 0.0, 1.0);
            const double li = std::log10(std::max(o.solver_tol_initial, 1e-15));
            const double lf = std::log10(std::max(o.solver_tol_final, 1e-15));
            tol = std::pow(10.0, lf + a * (li - lf));
        }
        // Load cases (each warm-started from its own previous solution): with the multigrid solver all of them in
        // one lockstep solve, otherwise one after another. Compliance and solver iterations summed, residual and
        // max|u| the largest; raw sensitivities summed in case order.
        double solve_ms = 0.0, sens_ms = 0.0, cm[2] = {0.0, 0.0}, res_max = 0.0;
        int solve_iters = 0;
        const bool batched = n_cases > 1 && o.precision == MYCEL_PRECISION_MIXED_GMG;
        if (batched) {
            const auto t_solve = clock::now();
            pcg_solve_batch(op, d_b.get(), d_u.get(), n_cases, tol, o.solver_max_iterations, o.precision,
                            case_res.data());
            solve_ms += ms_since(t_solve);
        }
        for (int k = 0; k < n_cases; k++) {
            double *uk = d_u.get() + k * nd;
            mycel_pcg_result pr;
            if (batched) {
                pr = case_res[k];
            } else {
                const auto t_solve = clock::now();
                pcg_solve_device(op, d_b.get() + k * nd, uk, tol, o.solver_max_iterations, o.precision, nullptr,
                                 nullptr, nullptr, &pr);
                solve_ms += ms_since(t_solve);
            }
            if (!pr.converged) throw Error(MYCEL_E_NOT_CONVERGED, not_converged_text(pr));
            solve_iters += pr.iterations;
            res_max = std::max(res_max, pr.final_residual);
            double cmk[2];
            compliance_maxdisp(d.n_nodes, d_active_nodes.get(), d_f.get() + k * nd, uk, d_part, cmk, s);
            if (ordered) cmk[0] = seq_sum(op, d_f.get() + k * nd, uk, nullptr, d.n_dofs, s);  // :256-260
            cm[0] += cmk[0];
            cm[1] = std::max(cm[1], cmk[1]);
            const auto t_sens = clock::now();
            if (k == 0) {
                launch_sensitivities(op, uk, d_rho_bar.get(), d_dc_bar.get(), penal, s);
            } else {
                launch_sensitivities(op, uk, d_rho_bar.get(), d_dc_k.get(), penal, s);
                launch_add_into(na, d_dc_k.get(), d_dc_bar.get(), s);
            }
            sens_ms += ms_since(t_sens);
        }
        const auto t_sens = clock::now();
        if (heav) launch_heaviside_chain(na, d_rho_tilde.get(), d_dc_bar.get(), beta, s);  // :281-289
        flt.apply(d_dc_bar.get(), d_dc.get(), true, s);
        MYCEL_HIP(hipStreamSynchronize(s));
        sens_ms += ms_since(t_sens);
        const auto t_upd = clock::now();
        MYCEL_HIP(hipMemcpyAsync(d_rho_old.get(), d_rho.get(), sizeof(double) * nab, hipMemcpyDeviceToDevice, s));
        if (ordered)
            lambda = oc_update_ordered(op, d_rho.get(), d_dc.get(), flt.d_s.get(), d_passive.get(), d_rho.get(), na,
                                       o.volume_fraction, o.move_limit, o.rho_min, lambda, 60, 1e-8, s, nullptr);
        else
            lambda = oc_update_device(oc, d_rho.get(), d_dc.get(), flt.d_s.get(), d_passive.get(), d_rho.get(), na,
                                      o.volume_fraction, o.move_limit, o.rho_min, lambda, 60, 1e-8, s, nullptr);
        const double upd_ms = ms_since(t_upd);
        if (heav && iter % 25 == 0 && beta < 8.0) beta = std::min(8.0, beta * 2.0);  // :315-318
        double met[3];
        simp_metrics(na, d_rho.get(), d_rho_old.get(), d_rho_bar.get(), heav ? d_rho_tilde.get() : nullptr, d_part,
                     met, s);
        if (ordered) simp_sums_ordered(op, na, d_rho_bar.get(), heav ? d_rho_tilde.get() : d_rho_bar.get(), met, s);
        max_change = met[0];
        mycel_snapshot sn{};
        sn.iteration = iter;
        sn.compliance = cm[0];
        sn.volume_fraction = met[1] / na;
        sn.max_change = max_change;
        sn.grey_level = (4.0 * met[2] / na) * 100.0;
        sn.solver_iterations = solve_iters;
        sn.solver_residual = res_max;
        sn.max_displacement = std::sqrt(cm[1]);
        sn.lambda = lambda;
        sn.solve_ms = solve_ms;
        sn.sensitivity_ms = sens_ms;
        sn.filter_ms = filter_ms;
        sn.update_ms = upd_ms;
        sn.total_ms = ms_since(t_total);
        sn.beta_next = beta;
        if (snaps && n_snaps < snaps_cap) snaps[n_snaps] = sn;
        n_snaps++;
        result->total_iterations = n_snaps;
        result->final_compliance = sn.compliance;
        result->final_volume_fraction = sn.volume_fraction;
        bool stop = false;
        if (cb) {
            d_rho_bar.download(h_rho_bar.data(), nab, s);
            d_rho.download(h_rho.data(), nab, s);
            MYCEL_HIP(hipStreamSynchronize(s));
            stop = cb(&sn, h_rho_bar.data(), h_rho.data(), user) != 0;
        }
        // Convergence over the last six compliances (:374-390).
        if (hist_n == 6) {
            std::copy(hist + 1, hist + 6, hist);
            hist[5] = sn.compliance;
        } else {
            hist[hist_n++] = sn.compliance;
        }
        if (hist_n == 6) {
            const double rc = std::fabs(sn.compliance - hist[0]) / sn.compliance;
            if (max_change < 0.01 && rc < 1e-3) {
                result->converged = 1;
                break;
            }
        }
        if (stop) {
            result->canceled = 1;
            break;
        }
    }
    if (rho_out && na) d_rho.download(rho_out, na, s);
    if (rho_bar_out && na) d_rho_bar.download(rho_bar_out, na, s);
    if (u_out) d_u.download(u_out, ndc, s);
    MYCEL_HIP(hipStreamSynchronize(s));
}

// StaticAnalysis.Run body (Fem/StaticAnalysis.cs:35-142): b = P F, PCG, C = F.u, R = (K u)|fixed,
// max |u| over active nodes, element von Mises.
void static_analysis_run(mycel_op *op, const double *f, double *u, double tol, int max_iter, bool throw_on_nonconv,
                         int precision, double *compliance, double *max_disp, double *reactions, double *von_mises,
                         mycel_pcg_result *result) {
    if (op->active_fixed < 6)
        throw Error(MYCEL_E_INVALID_OPERATION,
                    "Model has fewer than 6 unique fixed DOFs on active nodes (found " +
                        std::to_string(op->active_fixed) + ").");
    hipStream_t s = op->stream;
    const Dims &d = op->d;
    DevBuf<double> d_f(d.n_dofs), d_b(d.n_dofs), d_u(d.n_dofs), d_part;
    d_f.upload(f, d.n_dofs, s);
    d_u.upload(u, d.n_dofs, s);
    launch_project(d.n_dofs, op->d_flags.get(), d_f.get(), d_b.get(), s);
    pcg_solve_device(op, d_b.get(), d_u.get(), tol, max_iter, precision, nullptr, nullptr, nullptr, result);
    if (throw_on_nonconv && !result->converged) throw Error(MYCEL_E_NOT_CONVERGED, not_converged_text(*result));
    DevBuf<uint8_t> d_an(d.n_nodes);
    d_an.upload(op->active_nodes.data(), d.n_nodes, s);
    double cm[2];
    compliance_maxdisp(d.n_nodes, d_an.get(), d_f.get(), d_u.get(), d_part, cm, s);
    // C = sum F_i u_i in index order under the reference's evaluation order (StaticAnalysis.cs:95-99).
    if (precision == MYCEL_PRECISION_FP64_ORDERED) cm[0] = seq_sum(op, d_f.get(), d_u.get(), nullptr, d.n_dofs, s);
    *compliance = cm[0];
    *max_disp = std::sqrt(cm[1]);
    if (reactions) {  // R = K u sampled at FixedDofs (:103-110)
        if (precision == MYCEL_PRECISION_FP64_ORDERED)
            launch_apply_k_ordered(op, d_u.get(), op->d_y.get(), kApplyK, s);
        else
            launch_apply_k(op, d_u.get(), op->d_y.get(), kApplyK, s);
        hipLaunchKernelGGL(k_select_fixed, dim3(cdiv(d.n_dofs, 256)), dim3(256), 0, s, d.n_dofs, op->d_flags.get(),
                           op->d_y.get(), d_b.get());
        MYCEL_LAUNCH_CHECK();
        d_b.download(reactions, d.n_dofs, s);
    }
    if (von_mises) {
        DevBuf<double> d_vm(d.n_elem);
        launch_von_mises(op, d_u.get(), d_vm.get(), s);
        d_vm.download(von_mises, d.n_elem, s);
        MYCEL_HIP(hipStreamSynchronize(s));
    }
    d_u.download(u, d.n_dofs, s);
    MYCEL_HIP(hipStreamSynchronize(s));
}

}  // namespace mycel

using namespace mycel;

// ================================================================================================
// C ABI
// ================================================================================================
extern "C" {

const char *mycel_last_error(void) { return g_last_error.c_str(); }

int32_t mycel_version(void) { return 10000; }

int32_t mycel_pcg_message(const mycel_pcg_result *result, char *buf, int32_t capacity) {
    return guarded([&] {
        require(result && buf, MYCEL_E_ARGUMENT_NULL, "result/buf");
        require(capacity > 0, MYCEL_E_ARGUMENT, "capacity");
        const std::string t = pcg_message_text(result->message, result->message_value);
        const size_t n = std::min(t.size(), static_cast<size_t>(capacity - 1));
        std::memcpy(buf, t.data(), n);
        buf[n] = '\0';
    });
}

int32_t mycel_device_count(int32_t *count) {
    return guarded([&] {
        require(count, MYCEL_E_ARGUMENT_NULL, "count");
        int c = 0;
        MYCEL_HIP(hipGetDeviceCount(&c));
        *count = c;
    });
}

int32_t mycel_set_device(int32_t ordinal) {
    return guarded([&] {
        int c = 0;
        MYCEL_HIP(hipGetDeviceCount(&c));
        require(ordinal >= 0 && ordinal < c, MYCEL_E_OUT_OF_RANGE, "device ordinal out of range");
        g_device = ordinal;
    });
}

int32_t mycel_hex8_ke(double h, double nu, double *ke576) {
    return guarded([&] {
        require(ke576, MYCEL_E_ARGUMENT_NULL, "ke");
        require(h > 0, MYCEL_E_OUT_OF_RANGE, "Element edge length h must be positive.");
        hex8_ke(h, nu, ke576);
    });
}

int32_t mycel_hex8_bc(double h, double *bc144) {
    return guarded([&] {
        require(bc144, MYCEL_E_ARGUMENT_NULL, "bc");
        require(h > 0, MYCEL_E_OUT_OF_RANGE, "Element edge length h must be positive.");
        hex8_bc(h, bc144);
    });
}

// FeModel.AssembleForces (Fem/FeModel.cs:263-321).
int32_t mycel_assemble_forces(int32_t total_nodes, int32_t n_cases, const int32_t *case_offsets,
                              const int32_t *nodes, const double *areas, const float *force_xyz, double *f) {
    return guarded([&] {
        require(f && (n_cases == 0 || (case_offsets && force_xyz)), MYCEL_E_ARGUMENT_NULL, "null argument");
        std::fill(f, f + 3 * static_cast<size_t>(total_nodes), 0.0);
        for (int lc = 0; lc < n_cases; lc++) {
            const int b = case_offsets[lc], e = case_offsets[lc + 1];
            const float fx = force_xyz[3 * lc], fy = force_xyz[3 * lc + 1], fz = force_xyz[3 * lc + 2];
            const int count = e - b;
            if (count == 0) {
                if (fx != 0.0f || fy != 0.0f || fz != 0.0f)
                    throw Error(MYCEL_E_INVALID_OPERATION,
                                "Cannot assemble load case with nonzero force: surface patch contains 0 nodes.");
                continue;
            }
            require(nodes && areas, MYCEL_E_ARGUMENT_NULL, "nodes/areas");
            double total_area = 0.0;
            for (int i = b; i < e; i++) total_area += areas[i];
            const bool equal = total_area <= 0.0;
            const double inv = equal ? 1.0 / count : 1.0 / total_area;
            for (int i = b; i < e; i++) {
                const int n = nodes[i];
                if (n < 0 || n >= total_nodes)
                    throw Error(MYCEL_E_INVALID_OPERATION, "Patch node index " + std::to_string(n) + " is out of bounds.");
                if (equal) {
                    const double ffx = fx * inv, ffy = fy * inv, ffz = fz * inv;
                    f[3 * n] += ffx;
                    f[3 * n + 1] += ffy;
                    f[3 * n + 2] += ffz;
                } else {
                    const double w = areas[i] * inv;
                    f[3 * n] += fx * w;
                    f[3 * n + 1] += fy * w;
                    f[3 * n + 2] += fz * w;
                }
            }
        }
    });
}

// PassiveRegion.Build (Voxelization/PassiveRegion.cs:13-62).
int32_t mycel_passive_region(const mycel_grid *grid, const int32_t *nodes, int32_t n_nodes, int32_t radius,
                             uint8_t *mask) {
    return guarded([&] {
        require(grid && mask && (n_nodes == 0 || nodes), MYCEL_E_ARGUMENT_NULL, "null argument");
        const int nx = grid->nx, ny = grid->ny, nz = grid->nz;
        std::fill(mask, mask + static_cast<size_t>(nx) * ny * nz, 0);
        const int nx1 = nx + 1, nx1ny1 = (nx + 1) * (ny + 1);
        auto cl = [](int v, int lo, int hi) { return std::min(std::max(v, lo), hi); };
        for (int q = 0; q < n_nodes; q++) {
            const int node = nodes[q];
            const int nk = node / nx1ny1, rem = node % nx1ny1, nj = rem / nx1, ni = rem % nx1;
            const int x0 = cl(ni - radius, 0, nx - 1), x1 = cl(ni + radius - 1, 0, nx - 1);
            const int y0 = cl(nj - radius, 0, ny - 1), y1 = cl(nj + radius - 1, 0, ny - 1);
            const int z0 = cl(nk - radius, 0, nz - 1), z1 = cl(nk + radius - 1, 0, nz - 1);
            for (int iz = z0; iz <= z1; iz++)
                for (int iy = y0; iy <= y1; iy++)
                    for (int ix = x0; ix <= x1; ix++) mask[nx * (iy + ny * iz) + ix] = 1;
        }
    });
}

int32_t mycel_op_create(const mycel_grid *grid, const uint8_t *domain, const int32_t *fixed_dofs, int32_t n_fixed,
                        double e0, double nu, double emin_ratio, double penal, const double *density_box,
                        mycel_op **out) {
    return guarded([&] {
        require(grid && domain && out, MYCEL_E_ARGUMENT_NULL, "grid, domain and out must not be null");
        require(n_fixed == 0 || fixed_dofs, MYCEL_E_ARGUMENT_NULL, "fixed_dofs");
        require(n_fixed >= 0, MYCEL_E_ARGUMENT, "n_fixed");
        std::lock_guard<std::mutex> lk(g_init_mutex);
        *out = op_create(*grid, domain, fixed_dofs, n_fixed, e0, nu, emin_ratio, penal, density_box);
    });
}

int32_t mycel_op_destroy(mycel_op *op) {
    return guarded([&] { op_destroy(op); });
}

int32_t mycel_op_info_get(const mycel_op *op, mycel_op_info *info) {
    return guarded([&] {
        require(op && info, MYCEL_E_ARGUMENT_NULL, "op/info");
        info->total_elements = op->d.n_elem;
        info->total_nodes = op->d.n_nodes;
        info->total_dofs = op->d.n_dofs;
        info->active_elements = static_cast<int32_t>(op->active.size());
        info->active_nodes = op->n_active_nodes;
        info->active_fixed_dofs = op->active_fixed;
    });
}

int32_t mycel_op_update_densities(mycel_op *op, const double *density_box, double penal) {
    return guarded([&] {
        require(op && density_box, MYCEL_E_ARGUMENT_NULL, "op/densities");
        DeviceGuard dg(op->device);
        op->d_density.upload(density_box, op->d.n_elem, op->stream);
        op_update_e(op, op->d_density.get(), penal, op->stream);
        MYCEL_HIP(hipStreamSynchronize(op->stream));
    });
}

int32_t mycel_op_update_densities_d(mycel_op *op, const double *d_density_box, double penal, void *stream) {
    return guarded([&] {
        require(op && d_density_box, MYCEL_E_ARGUMENT_NULL, "op/densities");
        DeviceGuard dg(op->device);
        hipStream_t s = pick(op, stream);
        MYCEL_HIP(hipMemcpyAsync(op->d_density.get(), d_density_box, sizeof(double) * op->d.n_elem,
                                 hipMemcpyDeviceToDevice, s));
        op_update_e(op, op->d_density.get(), penal, s);
    });
}

static void apply_host(mycel_op *op, const double *x, double *y, int mode) {
    DeviceGuard dg(op->device);
    hipStream_t s = op->stream;
    op->d_x.upload(x, op->d.n_dofs, s);
    launch_apply_k(op, op->d_x.get(), op->d_y.get(), mode, s);
    op->d_y.download(y, op->d.n_dofs, s);
    MYCEL_HIP(hipStreamSynchronize(s));
}

int32_t mycel_op_apply_k(mycel_op *op, const double *x, double *y) {
    return guarded([&] {
        require(op && x && y, MYCEL_E_ARGUMENT_NULL, "op/x/y");
        apply_host(op, x, y, kApplyK);
    });
}

int32_t mycel_op_apply_a(mycel_op *op, const double *x, double *y, int32_t projected) {
    return guarded([&] {
        require(op && x && y, MYCEL_E_ARGUMENT_NULL, "op/x/y");
        apply_host(op, x, y, projected ? kApplyAProjected : kApplyAGeneral);
    });
}

int32_t mycel_op_apply_k_d(mycel_op *op, const double *d_x, double *d_y, void *stream) {
    return guarded([&] {
        require(op && d_x && d_y, MYCEL_E_ARGUMENT_NULL, "op/x/y");
        launch_apply_k(op, d_x, d_y, kApplyK, pick(op, stream));
    });
}

int32_t mycel_op_apply_a_d(mycel_op *op, const double *d_x, double *d_y, int32_t projected, void *stream) {
    return guarded([&] {
        require(op && d_x && d_y, MYCEL_E_ARGUMENT_NULL, "op/x/y");
        launch_apply_k(op, d_x, d_y, projected ? kApplyAProjected : kApplyAGeneral, pick(op, stream));
    });
}

int32_t mycel_op_apply_k_f32_d(mycel_op *op, const float *d_x, float *d_y, void *stream) {
    return guarded([&] {
        require(op && d_x && d_y, MYCEL_E_ARGUMENT_NULL, "op/x/y");
        launch_apply_k_f32(op, d_x, d_y, kApplyK, pick(op, stream));
    });
}

int32_t mycel_op_get_diag_a(mycel_op *op, double *diag) {
    return guarded([&] {
        require(op && diag, MYCEL_E_ARGUMENT_NULL, "op/diag");
        DeviceGuard dg(op->device);
        op->d_diag.download(diag, op->d.n_dofs, op->stream);
        MYCEL_HIP(hipStreamSynchronize(op->stream));
    });
}

int32_t mycel_op_set_matvec_form(mycel_op *op, int32_t form, int32_t columns) {
    return guarded([&] {
        require(op, MYCEL_E_ARGUMENT_NULL, "op");
        if (form == MYCEL_MATVEC_NODE_GATHER)
            require(columns == 0 || columns == 2 || columns == 4 || columns == 8, MYCEL_E_ARGUMENT,
                    "gather columns must be 0, 2, 4 or 8");
        else if (form == MYCEL_MATVEC_ELEMENT)
            require(columns == 2 || columns == 3 || columns == 4 || columns == 8, MYCEL_E_ARGUMENT,
                    "element-centric columns must be 2, 3, 4 or 8");
        else
            require(form == MYCEL_MATVEC_AUTO || form == MYCEL_MATVEC_REFERENCE_ORDER, MYCEL_E_ARGUMENT,
                    "unknown matvec form");
        op->ku_form = form;
        op->ku_kz = form == MYCEL_MATVEC_AUTO || form == MYCEL_MATVEC_REFERENCE_ORDER ? 0 : columns;
    });
}

int32_t mycel_op_set_island_deflation(mycel_op *op, double solid_ratio) {
    return guarded([&] {
        require(op, MYCEL_E_ARGUMENT_NULL, "op");
        require(std::isfinite(solid_ratio) && solid_ratio >= 0.0 && solid_ratio <= 1.0, MYCEL_E_OUT_OF_RANGE,
                "solid_ratio must be in [0, 1]");
        if (op->island_ratio != solid_ratio) {
            op->island_ratio = solid_ratio;
            op->epoch++;  // the multigrid set-up runs again before the next solve
        }
    });
}

int32_t mycel_op_set_matvec_columns(mycel_op *op, int32_t columns) {
    if (columns == -1) return mycel_op_set_matvec_form(op, MYCEL_MATVEC_AUTO, 0);
    return mycel_op_set_matvec_form(op, MYCEL_MATVEC_NODE_GATHER, columns);
}

int32_t mycel_op_get_ke(mycel_op *op, double *ke) {
    return guarded([&] {
        require(op && ke, MYCEL_E_ARGUMENT_NULL, "op/ke");
        std::copy(op->ke, op->ke + 576, ke);
    });
}

int32_t mycel_op_get_constrained(mycel_op *op, uint8_t *mask) {
    return guarded([&] {
        require(op && mask, MYCEL_E_ARGUMENT_NULL, "op/mask");
        for (int i = 0; i < op->d.n_dofs; i++) mask[i] = (op->dof_flags[i] & kFlagConstrained) ? 1 : 0;
    });
}

int32_t mycel_op_get_active_elements(mycel_op *op, int32_t *active) {
    return guarded([&] {
        require(op && active, MYCEL_E_ARGUMENT_NULL, "op/active");
        std::copy(op->active.begin(), op->active.end(), active);
    });
}

void *mycel_op_stream(mycel_op *op) { return op ? static_cast<void *>(op->stream) : nullptr; }

int32_t mycel_pcg_solve(mycel_op *op, const double *b, double *x, double tol, int32_t max_iter, int32_t precision,
                        const volatile int32_t *cancel, mycel_pcg_progress_cb progress, void *user,
                        mycel_pcg_result *result) {
    return guarded([&] {
        require(op && b && x && result, MYCEL_E_ARGUMENT_NULL, "op/b/x/result");
        DeviceGuard dg(op->device);
        DevBuf<double> d_b(op->d.n_dofs);
        d_b.upload(b, op->d.n_dofs, op->stream);
        op->d_x.upload(x, op->d.n_dofs, op->stream);
        pcg_solve_device(op, d_b.get(), op->d_x.get(), tol, max_iter, precision, cancel, progress, user, result);
        op->d_x.download(x, op->d.n_dofs, op->stream);
        MYCEL_HIP(hipStreamSynchronize(op->stream));
    });
}

int32_t mycel_pcg_solve_d(mycel_op *op, const double *d_b, double *d_x, double tol, int32_t max_iter,
                          int32_t precision, mycel_pcg_result *result) {
    return guarded([&] {
        require(op && d_b && d_x && result, MYCEL_E_ARGUMENT_NULL, "op/b/x/result");
        DeviceGuard dg(op->device);
        pcg_solve_device(op, d_b, d_x, tol, max_iter, precision, nullptr, nullptr, nullptr, result);
    });
}

// IPreconditioner.Setup + Apply (Fem/IPreconditioner.cs:9-20): Jacobi (JacobiPreconditioner.cs:15-35) or one
// multigrid V-cycle.
__global__ void k_jacobi_apply(int n, const double *__restrict__ r, const double *__restrict__ inv,
                               double *__restrict__ z) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) z[i] = r[i] * inv[i];
}

int32_t mycel_precond_apply_d(mycel_op *op, int32_t precision, const double *d_r, double *d_z, void *stream) {
    return guarded([&] {
        require(op && d_r && d_z, MYCEL_E_ARGUMENT_NULL, "op/r/z");
        DeviceGuard dg(op->device);
        hipStream_t s = pick(op, stream);
        const int n = op->d.n_dofs;
        if (precision == MYCEL_PRECISION_FP64 || precision == MYCEL_PRECISION_FP64_ORDERED) {
            hipLaunchKernelGGL(k_jacobi_apply, dim3(cdiv(n, 256)), dim3(256), 0, s, n, d_r, op->d_inv_diag.get(), d_z);
            MYCEL_LAUNCH_CHECK();
            return;
        }
        require(precision == MYCEL_PRECISION_MIXED_GMG, MYCEL_E_ARGUMENT, "preconditioner: FP64 (Jacobi) or MIXED_GMG");
        require(gmg_supported(op), MYCEL_E_ARGUMENT, "the grid has no multigrid hierarchy");
        gmg_prepare(op, s);
        const size_t np = static_cast<size_t>(gmg_partials(op));
        if (op->d_part.size() < 2 * np) op->d_part.resize(2 * np);
        gmg_enqueue_vcycle(op, d_r, d_r, d_z, op->d_part.get(), op->d_part.get() + np, s);
    });
}

// StaticAnalysis.Run (Fem/StaticAnalysis.cs:35-142).
int32_t mycel_static_analysis(mycel_op *op, const double *f, double *u, double tol, int32_t max_iter,
                              int32_t throw_on_nonconv, int32_t precision, double *compliance, double *max_disp,
                              double *reactions, double *von_mises, mycel_pcg_result *result) {
    return guarded([&] {
        require(op && f && u && compliance && max_disp && result, MYCEL_E_ARGUMENT_NULL, "null argument");
        DeviceGuard dg(op->device);
        static_analysis_run(op, f, u, tol, max_iter, throw_on_nonconv != 0, precision, compliance, max_disp,
                            reactions, von_mises, result);
    });
}

int32_t mycel_von_mises(mycel_op *op, const double *u, double *vm) {
    return guarded([&] {
        require(op && u && vm, MYCEL_E_ARGUMENT_NULL, "op/u/vm");
        DeviceGuard dg(op->device);
        hipStream_t s = op->stream;
        DevBuf<double> d_vm(op->d.n_elem);
        op->d_x.upload(u, op->d.n_dofs, s);
        launch_von_mises(op, op->d_x.get(), d_vm.get(), s);
        d_vm.download(vm, op->d.n_elem, s);
        MYCEL_HIP(hipStreamSynchronize(s));
    });
}

int32_t mycel_average_element_to_node(mycel_op *op, const double *elem, double *nodal) {
    return guarded([&] {
        require(op && elem && nodal, MYCEL_E_ARGUMENT_NULL, "op/elem/nodal");
        DeviceGuard dg(op->device);
        hipStream_t s = op->stream;
        DevBuf<double> d_e(op->d.n_elem), d_n(op->d.n_nodes);
        d_e.upload(elem, op->d.n_elem, s);
        launch_avg_e2n(op, d_e.get(), d_n.get(), s);
        d_n.download(nodal, op->d.n_nodes, s);
        MYCEL_HIP(hipStreamSynchronize(s));
    });
}

int32_t mycel_sample_nodal_scalar(mycel_op *op, const double *nodal, int32_t n_points, const float *points_xyz,
                                  const float *origin_xyz, double *out) {
    return guarded([&] {
        require(op && nodal && out && origin_xyz && (points_xyz || n_points == 0), MYCEL_E_ARGUMENT_NULL,
                "op/nodal/points/origin/out");
        require(n_points >= 0, MYCEL_E_ARGUMENT, "n_points");
        if (n_points == 0) return;
        DeviceGuard dg(op->device);
        hipStream_t s = op->stream;
        DevBuf<double> d_n(op->d.n_nodes), d_out(n_points);
        DevBuf<float> d_p(3 * static_cast<size_t>(n_points));
        DevBuf<int> d_fb(n_points);
        d_n.upload(nodal, op->d.n_nodes, s);
        d_p.upload(points_xyz, 3 * static_cast<size_t>(n_points), s);
        launch_sample_nodal(op, d_n.get(), n_points, d_p.get(), origin_xyz, d_out.get(), d_fb.get(), s);
        d_out.download(out, n_points, s);
        MYCEL_HIP(hipStreamSynchronize(s));
    });
}

struct mycel_filter {
    int device;
    hipStream_t stream;
    DeviceFilter *f;
};

int32_t mycel_filter_create(const mycel_grid *grid, const uint8_t *domain, const int32_t *active, int32_t n_active,
                            double rmin, mycel_filter **out) {
    return guarded([&] {
        require(grid && domain && active && out, MYCEL_E_ARGUMENT_NULL, "null argument");
        auto h = std::make_unique<mycel_filter>();
        h->device = g_device;
        MYCEL_HIP(hipSetDevice(h->device));
        MYCEL_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
        h->f = new DeviceFilter(grid->nx, grid->ny, grid->nz, domain, active, n_active, rmin, h->stream);
        *out = h.release();
    });
}

int32_t mycel_filter_destroy(mycel_filter *f) {
    return guarded([&] {
        if (!f) return;
        (void)hipSetDevice(f->device);
        delete f->f;
        (void)hipStreamDestroy(f->stream);
        delete f;
    });
}

int32_t mycel_filter_get_s(mycel_filter *f, double *s) {
    return guarded([&] {
        require(f && s, MYCEL_E_ARGUMENT_NULL, "filter/s");
        DeviceGuard dg(f->device);
        if (f->f->n) f->f->d_s.download(s, f->f->n, f->stream);
        MYCEL_HIP(hipStreamSynchronize(f->stream));
    });
}

int32_t mycel_filter_nnz(mycel_filter *f, int64_t *nnz) {
    return guarded([&] {
        require(f && nnz, MYCEL_E_ARGUMENT_NULL, "filter/nnz");
        *nnz = f->f->nnz;
    });
}

static void filter_host(mycel_filter *f, const double *in, double *out, bool adjoint) {
    require(f && in && out, MYCEL_E_ARGUMENT_NULL, "filter/in/out");
    DeviceGuard dg(f->device);
    DeviceFilter &F = *f->f;
    if (!F.n) return;
    F.d_in.upload(in, F.n, f->stream);
    F.apply(F.d_in.get(), F.d_out.get(), adjoint, f->stream);
    F.d_out.download(out, F.n, f->stream);
    MYCEL_HIP(hipStreamSynchronize(f->stream));
}

int32_t mycel_filter_apply(mycel_filter *f, const double *in, double *out) {
    return guarded([&] { filter_host(f, in, out, false); });
}

int32_t mycel_filter_apply_adjoint(mycel_filter *f, const double *in, double *out) {
    return guarded([&] { filter_host(f, in, out, true); });
}

int32_t mycel_sensitivities(mycel_op *op, const double *u, const double *rho, double *dc, double penal) {
    return guarded([&] {
        require(op && u && rho && dc, MYCEL_E_ARGUMENT_NULL, "op/u/rho/dc");
        DeviceGuard dg(op->device);
        hipStream_t s = op->stream;
        const int na = static_cast<int>(op->active.size());
        if (!na) return;
        DevBuf<double> d_rho(na), d_dc(na);
        op->d_x.upload(u, op->d.n_dofs, s);
        d_rho.upload(rho, na, s);
        launch_sensitivities(op, op->d_x.get(), d_rho.get(), d_dc.get(), penal, s);
        d_dc.download(dc, na, s);
        MYCEL_HIP(hipStreamSynchronize(s));
    });
}

int32_t mycel_oc_update(const double *rho, const double *dc, const double *sv, const uint8_t *passive,
                        double *rho_new, int32_t n, double vf, double move, double rho_min, double lambda_guess,
                        int32_t max_bisect, double vol_tol, double *lambda_out) {
    return guarded([&] {
        require(rho && dc && sv && rho_new && lambda_out, MYCEL_E_ARGUMENT_NULL, "null argument");
        require(n >= 0, MYCEL_E_ARGUMENT, "n");
        MYCEL_HIP(hipSetDevice(g_device));
        hipStream_t s;
        MYCEL_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        try {
            const size_t m = std::max(1, n);
            DevBuf<double> d_rho(m), d_dc(m), d_s(m), d_out(m);
            DevBuf<uint8_t> d_p(m);
            if (n) {
                d_rho.upload(rho, n, s);
                d_dc.upload(dc, n, s);
                d_s.upload(sv, n, s);
                if (passive) d_p.upload(passive, n, s);
            }
            OcWork w;
            *lambda_out = oc_update_device(w, d_rho.get(), d_dc.get(), d_s.get(), passive ? d_p.get() : nullptr,
                                           d_out.get(), n, vf, move, rho_min, lambda_guess, max_bisect, vol_tol, s,
                                           nullptr);
            if (n) d_out.download(rho_new, n, s);
            MYCEL_HIP(hipStreamSynchronize(s));
        } catch (...) {
            (void)hipStreamDestroy(s);
            throw;
        }
        MYCEL_HIP(hipStreamDestroy(s));
    });
}

int32_t mycel_oc_update_noalloc_d(const double *d_rho, const double *d_dc, const double *d_s, const uint8_t *d_passive,
                                  double *d_rho_new, int32_t n, double vf, double move, double rho_min,
                                  double lambda_guess, int32_t max_bisect, double vol_tol, void *stream,
                                  double *lambda_out) {
    return guarded([&] {
        require(d_rho && d_dc && d_s && d_rho_new && lambda_out, MYCEL_E_ARGUMENT_NULL, "null argument");
        require(n >= 0, MYCEL_E_ARGUMENT, "n");
        require(n <= MYCEL_OC_NOALLOC_MAX_ELEMENTS, MYCEL_E_OUT_OF_RANGE,
                "n exceeds MYCEL_OC_NOALLOC_MAX_ELEMENTS, the no-allocation OC update's limit");
        MYCEL_HIP(hipSetDevice(g_device));
        *lambda_out = oc_update_noalloc(d_rho, d_dc, d_s, d_passive, d_rho_new, n, vf, move, rho_min, lambda_guess,
                                        max_bisect, vol_tol, static_cast<hipStream_t>(stream));
    });
}

void mycel_simp_options_default(mycel_simp_options *o) {
    if (!o) return;
    *o = mycel_simp_options{};
    o->volume_fraction = 0.4;
    o->filter_radius_mm = 2.0;
    o->penal = 3.0;
    o->move_limit = 0.2;
    o->max_iterations = 200;
    o->solver_tol_initial = 1e-4;
    o->solver_tol_final = 1e-6;
    o->solver_max_iterations = 5000;
    o->rho_min = 1e-3;
    o->protect_boundary_hinges = 1;
    o->start_iteration = 1;
    o->enforce_min_filter_radius = 1;
    o->continuation = 0;
    o->initial_lambda = 1.0;
    o->precision = MYCEL_PRECISION_FP64;
}

// Persistence/RunWriter.cs:215-229 (8-bit frame) and :376-388 (FNV-1a over the frame bytes). MathF.Round
// with MidpointRounding.AwayFromZero of v >= 0 is floor(v + 0.5), formed exactly in double.
int32_t mycel_run_quantize(const float *phys, int64_t n, uint8_t *q, uint32_t *checksum) {
    return guarded([&] {
        require(n >= 0, MYCEL_E_ARGUMENT, "negative length");
        require((phys || n == 0) && checksum, MYCEL_E_ARGUMENT_NULL, "null argument");
        uint32_t h = 2166136261u;
        for (int64_t i = 0; i < n; i++) {
            const float c = std::min(std::max(phys[i], 0.0f), 1.0f);
            const float v = c * 255.0f;
            const uint8_t b = (v == v) ? static_cast<uint8_t>(std::floor(static_cast<double>(v) + 0.5)) : 0;
            if (q) q[i] = b;
            h = (h ^ b) * 16777619u;
        }
        *checksum = h;
    });
}

// Persistence/RunReader.cs:213-223: the checksum a reader recomputes over a frame's bytes.
int32_t mycel_run_checksum(const uint8_t *q, int64_t n, uint32_t *checksum) {
    return guarded([&] {
        require(n >= 0, MYCEL_E_ARGUMENT, "negative length");
        require((q || n == 0) && checksum, MYCEL_E_ARGUMENT_NULL, "null argument");
        uint32_t h = 2166136261u;
        for (int64_t i = 0; i < n; i++) h = (h ^ q[i]) * 16777619u;
        *checksum = h;
    });
}

int32_t mycel_simp_run(mycel_op *op, const double *f, const mycel_simp_options *opt, const uint8_t *passive_mask_box,
                       const int32_t *patch_nodes, int32_t n_patch_nodes, const float *initial_density,
                       int32_t initial_density_len, const double *warm_u, mycel_snapshot_cb cb, void *user,
                       mycel_simp_result *result, mycel_snapshot *snapshots, int32_t snapshots_capacity,
                       double *rho_out, double *rho_bar_out, double *u_out) {
    return guarded([&] {
        require(op && f && opt && result, MYCEL_E_ARGUMENT_NULL, "op/f/options/result");
        require(n_patch_nodes == 0 || patch_nodes, MYCEL_E_ARGUMENT_NULL, "patch_nodes");
        DeviceGuard dg(op->device);
        // WarmStartDisplacements of the wrong length are ignored by the reference (:134-140); the C ABI
        // carries no length for warm_u, so a non-null pointer must hold TotalDofs values.
        simp_run(op, f, 1, *opt, passive_mask_box, patch_nodes, n_patch_nodes, initial_density,
                 initial_density_len, warm_u, cb, user, result, snapshots, snapshots_capacity, rho_out, rho_bar_out,
                 u_out);
    });
}

int32_t mycel_pcg_solve_batch(mycel_op *op, const double *b, double *x, int32_t n_rhs, double tol, int32_t max_iter,
                              int32_t precision, mycel_pcg_result *results) {
    return guarded([&] {
        require(op && b && x && results, MYCEL_E_ARGUMENT_NULL, "op/b/x/results");
        require(n_rhs >= 1, MYCEL_E_ARGUMENT, "n_rhs must be >= 1");
        DeviceGuard dg(op->device);
        const size_t n = static_cast<size_t>(op->d.n_dofs) * n_rhs;
        DevBuf<double> d_b(n), d_x(n);
        d_b.upload(b, n, op->stream);
        d_x.upload(x, n, op->stream);
        pcg_solve_batch(op, d_b.get(), d_x.get(), n_rhs, tol, max_iter, precision, results);
        d_x.download(x, n, op->stream);
        MYCEL_HIP(hipStreamSynchronize(op->stream));
    });
}

int32_t mycel_pcg_solve_batch_d(mycel_op *op, const double *d_b, double *d_x, int32_t n_rhs, double tol,
                                int32_t max_iter, int32_t precision, mycel_pcg_result *results) {
    return guarded([&] {
        require(op && d_b && d_x && results, MYCEL_E_ARGUMENT_NULL, "op/b/x/results");
        require(n_rhs >= 1, MYCEL_E_ARGUMENT, "n_rhs must be >= 1");
        DeviceGuard dg(op->device);
        pcg_solve_batch(op, d_b, d_x, n_rhs, tol, max_iter, precision, results);
    });
}

int32_t mycel_pcg_solve_packed_d(mycel_op *const *ops, const double *const *d_b, double *const *d_x, int32_t count,
                                 double tol, int32_t max_iter, mycel_pcg_result *results) {
    return guarded([&] {
        require(ops && d_b && d_x && results, MYCEL_E_ARGUMENT_NULL, "ops/b/x/results");
        require(count >= 1 && count <= 16, MYCEL_E_OUT_OF_RANGE, "count must be in [1, 16]");
        for (int i = 0; i < count; i++) {
            require(ops[i] && d_b[i] && d_x[i], MYCEL_E_ARGUMENT_NULL, "ops/b/x entry");
            require(ops[i]->device == ops[0]->device, MYCEL_E_ARGUMENT, "every operator must be on one device");
        }
        DeviceGuard dg(ops[0]->device);
        pcg_solve_packed(ops, d_b, d_x, count, tol, max_iter, results);
    });
}

int32_t mycel_simp_run_cases(mycel_op *op, const double *f, int32_t n_cases, const mycel_simp_options *opt,
                             const uint8_t *passive_mask_box, const int32_t *patch_nodes, int32_t n_patch_nodes,
                             const float *initial_density, int32_t initial_density_len, const double *warm_u,
                             mycel_snapshot_cb cb, void *user, mycel_simp_result *result, mycel_snapshot *snapshots,
                             int32_t snapshots_capacity, double *rho_out, double *rho_bar_out, double *u_out) {
    return guarded([&] {
        require(op && f && opt && result, MYCEL_E_ARGUMENT_NULL, "op/f/options/result");
        require(n_patch_nodes == 0 || patch_nodes, MYCEL_E_ARGUMENT_NULL, "patch_nodes");
        require(n_cases >= 1, MYCEL_E_ARGUMENT, "n_cases must be >= 1");
        DeviceGuard dg(op->device);
        simp_run(op, f, n_cases, *opt, passive_mask_box, patch_nodes, n_patch_nodes, initial_density,
                 initial_density_len, warm_u, cb, user, result, snapshots, snapshots_capacity, rho_out, rho_bar_out,
                 u_out);
    });
}

}  // extern "C"
