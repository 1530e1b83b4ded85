/*
 * mycel_oracle.c -- FP64 CPU restatement of Mycel.Core's matrix-free Hex8 elasticity path.
 *
 * TEST INFRASTRUCTURE ONLY: the correctness oracle for libmycel_hip.so and the CPU baseline that
 * bench.py times. The product library never links or loads this file.
 *
 * Build: gcc -O2 -ffp-contract=off -fopenmp -fPIC -shared (oracle/Makefile). The reference runs on
 * .NET, which does not contract a*b+c into an FMA; -ffp-contract=off keeps the same rounding.
 * Parallel loops mirror the reference's partitioning (8 colours x 256-element chunks, 8,192-DOF
 * reduction blocks); results do not depend on the thread count.
 *
 * Summation orders that the reference leaves to the .NET BCL:
 *   - Vector256<double>.Sum(v) is taken as (v0 + v1) + (v2 + v3), the per-128-bit-lane order the
 *     RyuJIT x64 lowering uses for floating point. This is not pinned by any reference test
 *     (the reference's own V9 test compares orders at 1e-12 only).
 *   - Math.Sqrt is the C runtime's sqrt (correctly rounded). Math.Pow is the C runtime's pow, whose last
 *     bit differs between runtimes; the SIMP stiffness takes the correctly rounded rho^3 (pow_penal).
 */
#include "mycel_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PCG_BLOCK 8192  /* Fem/PcgSolver.cs:18 */
#define OP_CHUNK 256    /* Fem/FemOperator.cs:25 */

static double now_ms(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

static int resolve_threads(int n_threads) {
#ifdef _OPENMP
    return n_threads > 0 ? n_threads : omp_get_max_threads();
#else
    (void)n_threads;
    return 1;
#endif
}

/* ------------------------------------------------------------------------------------------ */
/* Hex8 element (Fem/Hex8Element.cs)                                                           */
/* ------------------------------------------------------------------------------------------ */

static const double XI[8] = {-1, 1, 1, -1, -1, 1, 1, -1};   /* Hex8Element.cs:52 */
static const double ETA[8] = {-1, -1, 1, 1, -1, -1, 1, 1};  /* :53 */
static const double ZETA[8] = {-1, -1, -1, -1, 1, 1, 1, 1}; /* :54 */

/* CreateConstitutiveMatrix (Hex8Element.cs:77-92). */
void orc_constitutive(double e, double nu, double *d) {
    double factor = e / ((1.0 + nu) * (1.0 - 2.0 * nu));
    double c1 = factor * (1.0 - nu);
    double c2 = factor * nu;
    double c3 = e / (2.0 * (1.0 + nu));
    memset(d, 0, 36 * sizeof(double));
    d[0] = c1; d[1] = c2; d[2] = c2;
    d[6] = c2; d[7] = c1; d[8] = c2;
    d[12] = c2; d[13] = c2; d[14] = c1;
    d[21] = c3; d[28] = c3; d[35] = c3;
}

/* EvaluateB (Hex8Element.cs:183-215). */
static void evaluate_b(double xi, double eta, double zeta, double h, double *b) {
    memset(b, 0, 144 * sizeof(double));
    double inv_h = 1.0 / h;
    for (int a = 0; a < 8; a++) {
        double dndx = (0.25 * inv_h) * XI[a] * (1.0 + ETA[a] * eta) * (1.0 + ZETA[a] * zeta);
        double dndy = (0.25 * inv_h) * ETA[a] * (1.0 + XI[a] * xi) * (1.0 + ZETA[a] * zeta);
        double dndz = (0.25 * inv_h) * ZETA[a] * (1.0 + XI[a] * xi) * (1.0 + ETA[a] * eta);
        int cu = 3 * a, cv = 3 * a + 1, cw = 3 * a + 2;
        b[0 * 24 + cu] = dndx;
        b[1 * 24 + cv] = dndy;
        b[2 * 24 + cw] = dndz;
        b[3 * 24 + cv] = dndz;
        b[3 * 24 + cw] = dndy;
        b[4 * 24 + cu] = dndz;
        b[4 * 24 + cw] = dndx;
        b[5 * 24 + cu] = dndy;
        b[5 * 24 + cv] = dndx;
    }
}

/* 2x2x2 Gauss integration of B^T D B for h = 1 (Hex8Element.cs:114-166). */
static void ke_unit(double nu, double *ke) {
    double g[2];
    g[0] = -1.0 / sqrt(3.0);
    g[1] = 1.0 / sqrt(3.0);
    double d[36], b[144], db[144];
    orc_constitutive(1.0, nu, d);
    memset(ke, 0, 576 * sizeof(double));
    const double w = 1.0 / 8.0;
    for (int ix = 0; ix < 2; ix++)
        for (int iy = 0; iy < 2; iy++)
            for (int iz = 0; iz < 2; iz++) {
                evaluate_b(g[ix], g[iy], g[iz], 1.0, b);
                for (int r = 0; r < 6; r++)
                    for (int c = 0; c < 24; c++) {
                        double s = 0.0;
                        for (int k = 0; k < 6; k++) s += d[r * 6 + k] * b[k * 24 + c];
                        db[r * 24 + c] = s;
                    }
                for (int i = 0; i < 24; i++)
                    for (int j = 0; j < 24; j++) {
                        double s = 0.0;
                        for (int k = 0; k < 6; k++) s += b[k * 24 + i] * db[k * 24 + j];
                        ke[i * 24 + j] += s * w;
                    }
            }
}

/* CreateKe (Hex8Element.cs:98-178): the nu = 0.3 fast path is BaseKe * h, otherwise the unit
 * quadrature scaled by h when h != 1. Both reduce to quad(nu) * h. */
int orc_hex8_ke(double h, double nu, double *ke) {
    if (!(h > 0)) return ORC_E_OUT_OF_RANGE;
    ke_unit(nu, ke);
    if (fabs(nu - 0.3) < 1e-15) {
        for (int i = 0; i < 576; i++) ke[i] = ke[i] * h;
    } else if (fabs(h - 1.0) > 1e-15) {
        for (int i = 0; i < 576; i++) ke[i] *= h;
    }
    return ORC_OK;
}

/* CreateBc (Hex8Element.cs:220-241): BaseBc = EvaluateB(0,0,0,1), scaled by 1/h. */
int orc_hex8_bc(double h, double *bc) {
    if (!(h > 0)) return ORC_E_OUT_OF_RANGE;
    double base[144];
    evaluate_b(0.0, 0.0, 0.0, 1.0, base);
    double inv_h = 1.0 / h;
    for (int i = 0; i < 144; i++) bc[i] = base[i] * inv_h;
    return ORC_OK;
}

/* ------------------------------------------------------------------------------------------ */
/* FeModel + FemOperator state                                                                 */
/* ------------------------------------------------------------------------------------------ */

struct orc_model {
    int nx, ny, nz;
    double h;
    double e0, nu, emin;
    double penal;
    int n_elem, n_nodes, n_dofs;
    int n_active;
    int32_t *active;          /* ascending box element indices */
    int color_n[8];
    int32_t *color_elems[8];  /* ElementsByColor */
    int32_t *color_dofs[8];   /* ColorElemDofs: 24 per element */
    double *color_e[8];       /* ColorElemE */
    uint8_t *active_nodes;
    uint8_t *constrained;
    int n_fixed;
    int32_t *fixed;
    int active_fixed;
    double *density;          /* box-sized ElementDensity */
    double ke[576];
    double diag_ke[24];
    double *diag_a;
};

static void elem_nodes(const orc_model *m, int e, int *nodes) {
    int nxny = m->nx * m->ny;
    int iz = e / nxny, rem = e % nxny;
    int iy = rem / m->nx, ix = rem % m->nx;
    int nx1 = m->nx + 1, ny1 = m->ny + 1;
    int n0 = ix + nx1 * (iy + ny1 * iz);
    nodes[0] = n0;
    nodes[1] = n0 + 1;
    nodes[3] = n0 + nx1;
    nodes[2] = nodes[3] + 1;
    nodes[4] = n0 + nx1 * ny1;
    nodes[5] = nodes[4] + 1;
    nodes[7] = nodes[4] + nx1;
    nodes[6] = nodes[7] + 1;
}

/* AssembleDiagonal (Fem/FemOperator.cs:143-172). */
static void assemble_diagonal(orc_model *m) {
    memset(m->diag_a, 0, (size_t)m->n_dofs * sizeof(double));
    for (int c = 0; c < 8; c++) {
        const int32_t *dofs = m->color_dofs[c];
        const double *ev = m->color_e[c];
        for (int i = 0; i < m->color_n[c]; i++) {
            double ee = ev[i];
            for (int k = 0; k < 24; k++) m->diag_a[dofs[i * 24 + k]] += ee * m->diag_ke[k];
        }
    }
    for (int i = 0; i < m->n_dofs; i++)
        if (m->constrained[i]) m->diag_a[i] = 1.0;
}

/* Math.Pow(rho, p) for the SIMP stiffness (FemOperator.cs:100,133). Math.Pow's last bit is the platform C
 * runtime's: glibc's pow(x, 3.0) is one ulp off the correctly rounded cube for about 0.08% of x in [0, 1]
 * (measured on 10^8 samples), and pow(x, 2.0) one ulp off x * x for about 0.08% of x (18 of 24,000 random
 * densities), so no CRT pins the bits. This restatement takes the correctly rounded power for the exponents in
 * use: p = 3 as rho^2 split exactly into hi + lo by an FMA, then (hi + lo) rho formed with one more error-free
 * product (exact for every input checked against rational arithmetic); p = 2, 1, 0 as rho * rho, rho and 1;
 * other p go to pow. The sensitivities' rho^(p-1) (SensitivityAnalysis.cs:99) uses the same rule. */
static double pow_penal(double r, double p) {
    if (p == 0.0) return 1.0;
    if (p == 3.0) {
        double hi = r * r, lo = fma(r, r, -hi);
        double p1 = hi * r, e1 = fma(hi, r, -p1);
        return p1 + fma(lo, r, e1);
    }
    if (p == 2.0) return r * r;
    if (p == 1.0) return r;
    return pow(r, p);
}

void orc_pow_penal_n(const double *r, int n, double p, double *out) {
    for (int i = 0; i < n; i++) out[i] = pow_penal(r[i], p);
}

static void fill_color_e(orc_model *m, const double *density, double penal) {
    double e0 = m->e0, emin = m->emin;
    for (int c = 0; c < 8; c++)
        for (int i = 0; i < m->color_n[c]; i++) {
            double rho = density[m->color_elems[c][i]];
            m->color_e[c][i] = emin + pow_penal(rho, penal) * (e0 - emin); /* FemOperator.cs:100,133 */
        }
}

int orc_model_create(int nx, int ny, int nz, double h, double e0, double nu, double emin_ratio,
                     double penal, const uint8_t *domain, const int32_t *fixed_dofs, int n_fixed,
                     const double *density, orc_model **out) {
    if (!domain || !out) return ORC_E_ARGUMENT_NULL;
    if (nx <= 0 || ny <= 0 || nz <= 0 || !(h > 0)) return ORC_E_OUT_OF_RANGE;
    orc_model *m = (orc_model *)calloc(1, sizeof(orc_model));
    m->nx = nx; m->ny = ny; m->nz = nz; m->h = h;
    m->e0 = e0; m->nu = nu; m->emin = e0 * emin_ratio; /* FeMaterial.cs:16 */
    m->penal = penal;
    m->n_elem = nx * ny * nz;
    m->n_nodes = (nx + 1) * (ny + 1) * (nz + 1);
    m->n_dofs = 3 * m->n_nodes;
    m->density = (double *)malloc((size_t)m->n_elem * sizeof(double));
    for (int e = 0; e < m->n_elem; e++) m->density[e] = density ? density[e] : 1.0;

    /* Active elements and 8-colour lists (FeModel.cs:80-147). */
    m->active = (int32_t *)malloc((size_t)m->n_elem * sizeof(int32_t));
    m->active_nodes = (uint8_t *)calloc((size_t)m->n_nodes, 1);
    int cap[8] = {0};
    for (int c = 0; c < 8; c++) m->color_elems[c] = (int32_t *)malloc((size_t)m->n_elem * sizeof(int32_t));
    int na = 0;
    for (int iz = 0; iz < nz; iz++)
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < nx; ix++) {
                int e = ix + nx * (iy + ny * iz);
                if (!domain[e]) continue;
                m->active[na++] = e;
                int color = ((iz & 1) << 2) | ((iy & 1) << 1) | (ix & 1);
                m->color_elems[color][cap[color]++] = e;
                int nodes[8];
                elem_nodes(m, e, nodes);
                for (int a = 0; a < 8; a++) m->active_nodes[nodes[a]] = 1;
            }
    m->n_active = na;
    for (int c = 0; c < 8; c++) m->color_n[c] = cap[c];

    /* Constrained mask: inactive nodes + fixed DOFs (FeModel.cs:149-183). */
    m->constrained = (uint8_t *)calloc((size_t)m->n_dofs, 1);
    for (int n = 0; n < m->n_nodes; n++)
        if (!m->active_nodes[n]) m->constrained[3 * n] = m->constrained[3 * n + 1] = m->constrained[3 * n + 2] = 1;
    m->n_fixed = n_fixed;
    m->fixed = (int32_t *)malloc((size_t)(n_fixed > 0 ? n_fixed : 1) * sizeof(int32_t));
    uint8_t *seen = (uint8_t *)calloc((size_t)m->n_dofs, 1);
    int act_fixed = 0;
    for (int i = 0; i < n_fixed; i++) {
        int dof = fixed_dofs[i];
        if (dof < 0 || dof >= m->n_dofs) {
            free(seen);
            orc_model_destroy(m);
            return ORC_E_OUT_OF_RANGE;
        }
        m->fixed[i] = dof;
        m->constrained[dof] = 1;
        if (m->active_nodes[dof / 3] && !seen[dof]) {
            seen[dof] = 1;
            act_fixed++;
        }
    }
    free(seen);
    m->active_fixed = act_fixed;

    /* FemOperator ctor (FemOperator.cs:33-108). */
    orc_hex8_ke(h, nu, m->ke);
    for (int k = 0; k < 24; k++) m->diag_ke[k] = m->ke[k * 24 + k];
    for (int c = 0; c < 8; c++) {
        m->color_dofs[c] = (int32_t *)malloc((size_t)(m->color_n[c] > 0 ? m->color_n[c] : 1) * 24 * sizeof(int32_t));
        m->color_e[c] = (double *)malloc((size_t)(m->color_n[c] > 0 ? m->color_n[c] : 1) * sizeof(double));
        for (int i = 0; i < m->color_n[c]; i++) {
            int nodes[8];
            elem_nodes(m, m->color_elems[c][i], nodes);
            for (int a = 0; a < 8; a++)
                for (int d = 0; d < 3; d++) m->color_dofs[c][i * 24 + 3 * a + d] = 3 * nodes[a] + d;
        }
    }
    fill_color_e(m, m->density, penal);
    m->diag_a = (double *)malloc((size_t)m->n_dofs * sizeof(double));
    assemble_diagonal(m);
    *out = m;
    return ORC_OK;
}

void orc_model_destroy(orc_model *m) {
    if (!m) return;
    free(m->active);
    for (int c = 0; c < 8; c++) {
        free(m->color_elems[c]);
        free(m->color_dofs[c]);
        free(m->color_e[c]);
    }
    free(m->active_nodes);
    free(m->constrained);
    free(m->fixed);
    free(m->density);
    free(m->diag_a);
    free(m);
}

int orc_model_active_count(const orc_model *m) { return m->n_active; }
int orc_model_active_fixed_count(const orc_model *m) { return m->active_fixed; }
void orc_model_active_elements(const orc_model *m, int32_t *out) { memcpy(out, m->active, (size_t)m->n_active * sizeof(int32_t)); }
void orc_model_constrained(const orc_model *m, uint8_t *out) { memcpy(out, m->constrained, (size_t)m->n_dofs); }
void orc_model_active_nodes(const orc_model *m, uint8_t *out) { memcpy(out, m->active_nodes, (size_t)m->n_nodes); }
const double *orc_model_ke(const orc_model *m) { return m->ke; }
void orc_model_diag_a(const orc_model *m, double *out) { memcpy(out, m->diag_a, (size_t)m->n_dofs * sizeof(double)); }

/* FeModel.AssembleForces (FeModel.cs:263-321). */
int orc_assemble_forces(int n_nodes_total, int n_cases, const int32_t *case_offsets,
                        const int32_t *nodes, const double *areas, const float *force_xyz,
                        double *f) {
    memset(f, 0, (size_t)3 * n_nodes_total * sizeof(double));
    for (int lc = 0; lc < n_cases; lc++) {
        int s = case_offsets[lc], e = case_offsets[lc + 1];
        float fx = force_xyz[3 * lc], fy = force_xyz[3 * lc + 1], fz = force_xyz[3 * lc + 2];
        int count = e - s;
        if (count == 0) {
            if (fx != 0.0f || fy != 0.0f || fz != 0.0f) return ORC_E_INVALID_OPERATION;
            continue;
        }
        double total_area = 0.0;
        for (int i = s; i < e; i++) total_area += areas[i];
        if (total_area <= 0.0) {
            double inv = 1.0 / count;
            double ffx = fx * inv, ffy = fy * inv, ffz = fz * inv;
            for (int i = s; i < e; i++) {
                int n = nodes[i];
                if (n < 0 || n >= n_nodes_total) return ORC_E_INVALID_OPERATION;
                f[3 * n] += ffx;
                f[3 * n + 1] += ffy;
                f[3 * n + 2] += ffz;
            }
        } else {
            double inv_area = 1.0 / total_area;
            for (int i = s; i < e; i++) {
                int n = nodes[i];
                if (n < 0 || n >= n_nodes_total) return ORC_E_INVALID_OPERATION;
                double w = areas[i] * inv_area;
                f[3 * n] += fx * w;
                f[3 * n + 1] += fy * w;
                f[3 * n + 2] += fz * w;
            }
        }
    }
    return ORC_OK;
}

void orc_update_densities(orc_model *m, const double *density_box, double penal) {
    if (density_box != m->density) memcpy(m->density, density_box, (size_t)m->n_elem * sizeof(double));
    m->penal = penal;
    fill_color_e(m, m->density, penal);
    assemble_diagonal(m);
}

/* ------------------------------------------------------------------------------------------ */
/* K·u (FemOperator.cs:196-439)                                                                */
/* ------------------------------------------------------------------------------------------ */

static inline double vsum4(double a0, double a1, double a2, double a3) { return (a0 + a1) + (a2 + a3); }

/* MultiplyElementUnsafe, Vector256 branch (FemOperator.cs:379-422). */
static inline void multiply_element(const double *ke, const int32_t *dofs, double ee, const double *x, double *y) {
    double ue[24];
    for (int k = 0; k < 24; k++) ue[k] = x[dofs[k]] * ee;
    for (int r = 0; r < 24; r++) {
        const double *row = ke + r * 24;
        double l[4];
        for (int j = 0; j < 4; j++) {
            double acc = row[j] * ue[j] + row[4 + j] * ue[4 + j];
            acc = acc + row[8 + j] * ue[8 + j];
            acc = acc + row[12 + j] * ue[12 + j];
            acc = acc + row[16 + j] * ue[16 + j];
            acc = acc + row[20 + j] * ue[20 + j];
            l[j] = acc;
        }
        y[dofs[r]] += vsum4(l[0], l[1], l[2], l[3]);
    }
}

void orc_apply_k(const orc_model *m, const double *x, double *y, int n_threads) {
    int nt = resolve_threads(n_threads);
    memset(y, 0, (size_t)m->n_dofs * sizeof(double));
    for (int c = 0; c < 8; c++) {
        int count = m->color_n[c];
        if (count == 0) continue;
        int n_chunks = (count + OP_CHUNK - 1) / OP_CHUNK;
        const int32_t *dofs = m->color_dofs[c];
        const double *ev = m->color_e[c];
        (void)nt;
#pragma omp parallel for num_threads(nt) schedule(static)
        for (int ch = 0; ch < n_chunks; ch++) {
            int s = ch * OP_CHUNK, e = s + OP_CHUNK < count ? s + OP_CHUNK : count;
            for (int i = s; i < e; i++) multiply_element(m->ke, dofs + (size_t)i * 24, ev[i], x, y);
        }
    }
}

void orc_apply_a(const orc_model *m, const double *x, double *y, int projected, int n_threads) {
    if (projected) {
        orc_apply_k(m, x, y, n_threads);
        for (int i = 0; i < m->n_fixed; i++) y[m->fixed[i]] = x[m->fixed[i]];
    } else {
        double *v = (double *)malloc((size_t)m->n_dofs * sizeof(double));
        for (int i = 0; i < m->n_dofs; i++) v[i] = m->constrained[i] ? 0.0 : x[i];
        orc_apply_k(m, v, y, n_threads);
        for (int i = 0; i < m->n_dofs; i++)
            if (m->constrained[i]) y[i] = x[i];
        free(v);
    }
}

void orc_apply_k_scalar(const orc_model *m, const double *x, double *y) {
    memset(y, 0, (size_t)m->n_dofs * sizeof(double));
    for (int idx = 0; idx < m->n_active; idx++) {
        int e = m->active[idx];
        int nodes[8], dofs[24];
        elem_nodes(m, e, nodes);
        for (int a = 0; a < 8; a++)
            for (int d = 0; d < 3; d++) dofs[3 * a + d] = 3 * nodes[a] + d;
        double ee = m->emin + pow_penal(m->density[e], m->penal) * (m->e0 - m->emin);
        double ue[24], ye[24];
        for (int k = 0; k < 24; k++) ue[k] = x[dofs[k]];
        for (int r = 0; r < 24; r++) {
            double s = 0.0;
            for (int c = 0; c < 24; c++) s += m->ke[r * 24 + c] * ue[c];
            ye[r] = s * ee;
        }
        for (int k = 0; k < 24; k++) y[dofs[k]] += ye[k];
    }
}

/* ------------------------------------------------------------------------------------------ */
/* PCG (Fem/PcgSolver.cs)                                                                      */
/* ------------------------------------------------------------------------------------------ */

typedef struct {
    int n, n_blocks, nt;
    double *partial;
} pcg_ctx;

/* ProcessDotChunk (PcgSolver.cs:424-450). */
static double dot_blocks(pcg_ctx *c, const double *a, const double *b) {
#pragma omp parallel for num_threads(c->nt) schedule(static)
    for (int blk = 0; blk < c->n_blocks; blk++) {
        int s = blk * PCG_BLOCK, cnt = c->n - s < PCG_BLOCK ? c->n - s : PCG_BLOCK;
        const double *pa = a + s, *pb = b + s;
        int vc = cnt / 4;
        double acc[4] = {0, 0, 0, 0};
        for (int i = 0; i < vc; i++)
            for (int j = 0; j < 4; j++) acc[j] = acc[j] + pa[4 * i + j] * pb[4 * i + j];
        double sum = vsum4(acc[0], acc[1], acc[2], acc[3]);
        for (int i = vc * 4; i < cnt; i++) sum += pa[i] * pb[i];
        c->partial[blk] = sum;
    }
    double sum = 0.0; /* ReducePartialSums (:414-422) */
    for (int k = 0; k < c->n_blocks; k++) sum += c->partial[k];
    return sum;
}

/* ProcessUpdateXrNormChunk (PcgSolver.cs:452-494). */
static double update_xr_norm(pcg_ctx *c, double alpha, double *x, double *r, const double *p, const double *ap) {
#pragma omp parallel for num_threads(c->nt) schedule(static)
    for (int blk = 0; blk < c->n_blocks; blk++) {
        int s = blk * PCG_BLOCK, cnt = c->n - s < PCG_BLOCK ? c->n - s : PCG_BLOCK;
        double *px = x + s, *pr = r + s;
        const double *pp = p + s, *pap = ap + s;
        int vc = cnt / 4;
        double acc[4] = {0, 0, 0, 0};
        for (int i = 0; i < vc; i++)
            for (int j = 0; j < 4; j++) {
                int k = 4 * i + j;
                px[k] = px[k] + alpha * pp[k];
                pr[k] = pr[k] - alpha * pap[k];
                acc[j] = acc[j] + pr[k] * pr[k];
            }
        double sum = vsum4(acc[0], acc[1], acc[2], acc[3]);
        for (int k = vc * 4; k < cnt; k++) {
            double xi = px[k] + alpha * pp[k];
            double ri = pr[k] - alpha * pap[k];
            px[k] = xi;
            pr[k] = ri;
            sum += ri * ri;
        }
        c->partial[blk] = sum;
    }
    double sum = 0.0;
    for (int k = 0; k < c->n_blocks; k++) sum += c->partial[k];
    return sum;
}

/* ProcessApplyPrecDotChunk (PcgSolver.cs:496-527). */
static double apply_prec_dot(pcg_ctx *c, const double *r, double *z, const double *inv) {
#pragma omp parallel for num_threads(c->nt) schedule(static)
    for (int blk = 0; blk < c->n_blocks; blk++) {
        int s = blk * PCG_BLOCK, cnt = c->n - s < PCG_BLOCK ? c->n - s : PCG_BLOCK;
        const double *pr = r + s, *pi = inv + s;
        double *pz = z + s;
        int vc = cnt / 4;
        double acc[4] = {0, 0, 0, 0};
        for (int i = 0; i < vc; i++)
            for (int j = 0; j < 4; j++) {
                int k = 4 * i + j;
                double zz = pr[k] * pi[k];
                pz[k] = zz;
                acc[j] = acc[j] + pr[k] * zz;
            }
        double sum = vsum4(acc[0], acc[1], acc[2], acc[3]);
        for (int k = vc * 4; k < cnt; k++) {
            double zi = pr[k] * pi[k];
            pz[k] = zi;
            sum += pr[k] * zi;
        }
        c->partial[blk] = sum;
    }
    double sum = 0.0;
    for (int k = 0; k < c->n_blocks; k++) sum += c->partial[k];
    return sum;
}

/* ProcessUpdatePChunk (PcgSolver.cs:529-553). */
static void update_p(pcg_ctx *c, double beta, double *p, const double *z) {
#pragma omp parallel for num_threads(c->nt) schedule(static)
    for (int i = 0; i < c->n; i++) p[i] = z[i] + beta * p[i];
}

int orc_pcg_solve(const orc_model *m, const double *b, double *x, double tol, int max_iter,
                  int n_threads, orc_pcg_result *res) {
    double t0 = now_ms();
    memset(res, 0, sizeof(*res));
    if (m->active_fixed < 6) return ORC_E_INVALID_OPERATION; /* PcgSolver.cs:95-100 */
    int n = m->n_dofs;
    pcg_ctx c;
    c.n = n;
    c.n_blocks = (n + PCG_BLOCK - 1) / PCG_BLOCK;
    c.nt = resolve_threads(n_threads);
    c.partial = (double *)calloc((size_t)c.n_blocks, sizeof(double));

    /* Project x, note whether it is non-zero (:113-125). */
    int nonzero_x = 0;
    for (int i = 0; i < n; i++) {
        if (m->constrained[i]) x[i] = 0.0;
        else if (x[i] != 0.0) nonzero_x = 1;
    }
    /* ||P b|| sequential (:128-137). */
    double b_norm_sq = 0.0;
    for (int i = 0; i < n; i++)
        if (!m->constrained[i]) b_norm_sq += b[i] * b[i];
    double b_norm = sqrt(b_norm_sq);
    if (b_norm <= 1e-30) {
        memset(x, 0, (size_t)n * sizeof(double));
        free(c.partial);
        res->iterations = 0;
        res->final_residual = 0.0;
        res->converged = 1;
        res->message = ORC_PCG_RHS_ZERO;
        res->elapsed_ms = now_ms() - t0;
        return ORC_OK;
    }
    /* JacobiPreconditioner.Setup (JacobiPreconditioner.cs:15-27). */
    double *inv = (double *)malloc((size_t)n * sizeof(double));
    for (int i = 0; i < n; i++) {
        double d = m->diag_a[i];
        inv[i] = fabs(d) > 1e-30 ? 1.0 / d : 1.0;
    }
    double *r = (double *)malloc((size_t)n * sizeof(double));
    double *z = (double *)malloc((size_t)n * sizeof(double));
    double *p = (double *)malloc((size_t)n * sizeof(double));
    double *ap = (double *)malloc((size_t)n * sizeof(double));

    double r_norm_sq;
    if (nonzero_x) {
        orc_apply_a(m, x, ap, 1, c.nt);
        for (int i = 0; i < n; i++) r[i] = m->constrained[i] ? 0.0 : (b[i] - ap[i]);
        r_norm_sq = dot_blocks(&c, r, r);
    } else {
        for (int i = 0; i < n; i++) r[i] = m->constrained[i] ? 0.0 : b[i];
        r_norm_sq = b_norm_sq;
    }
    double rel = sqrt(r_norm_sq) / b_norm;
    int status = ORC_OK;
    if (rel <= tol) {
        res->iterations = 0;
        res->final_residual = rel;
        res->converged = 1;
        res->message = ORC_PCG_INITIAL_CONVERGED;
        goto done;
    }
    double rho0 = apply_prec_dot(&c, r, z, inv);
    memcpy(p, z, (size_t)n * sizeof(double));
    int iter = 0, stag = 0;
    double prev_rel = rel;
    while (iter < max_iter) {
        iter++;
        orc_apply_a(m, p, ap, 1, c.nt);
        double pap = dot_blocks(&c, p, ap);
        if (!(pap > 1e-30) || !isfinite(pap)) {
            res->iterations = iter; res->final_residual = rel; res->converged = 0;
            res->message = ORC_PCG_CURVATURE_BREAKDOWN;
            goto done;
        }
        double alpha = rho0 / pap;
        if (!isfinite(alpha) || alpha <= 0.0) {
            res->iterations = iter; res->final_residual = rel; res->converged = 0;
            res->message = ORC_PCG_STEP_BREAKDOWN;
            goto done;
        }
        r_norm_sq = update_xr_norm(&c, alpha, x, r, p, ap);
        rel = sqrt(r_norm_sq) / b_norm;
        if (rel <= tol) {
            res->iterations = iter; res->final_residual = rel; res->converged = 1;
            res->message = ORC_PCG_CONVERGED;
            goto done;
        }
        if (fabs(rel - prev_rel) < 1e-15 * rel) {
            stag++;
            if (stag > 50) {
                res->iterations = iter; res->final_residual = rel; res->converged = 0;
                res->message = ORC_PCG_STAGNATION;
                goto done;
            }
        } else {
            stag = 0;
        }
        prev_rel = rel;
        double rho1 = apply_prec_dot(&c, r, z, inv);
        double beta = rho1 / rho0;
        if (!isfinite(beta) || beta < 0.0) {
            res->iterations = iter; res->final_residual = rel; res->converged = 0;
            res->message = ORC_PCG_DIRECTION_BREAKDOWN;
            goto done;
        }
        update_p(&c, beta, p, z);
        rho0 = rho1;
    }
    res->iterations = iter;
    res->final_residual = rel;
    res->converged = 0;
    res->message = ORC_PCG_MAX_ITERATIONS;
done:
    res->elapsed_ms = now_ms() - t0;
    free(inv); free(r); free(z); free(p); free(ap); free(c.partial);
    return status;
}

/* ------------------------------------------------------------------------------------------ */
/* StaticAnalysis (Fem/StaticAnalysis.cs)                                                      */
/* ------------------------------------------------------------------------------------------ */

void orc_von_mises(const orc_model *m, const double *u, double *vm) {
    memset(vm, 0, (size_t)m->n_elem * sizeof(double));
    double inv_h = 1.0 / m->h;
    double base_bc[144], d[36];
    evaluate_b(0.0, 0.0, 0.0, 1.0, base_bc);
    orc_constitutive(1.0, m->nu, d);
    for (int c = 0; c < 8; c++)
        for (int i = 0; i < m->color_n[c]; i++) {
            int e = m->color_elems[c][i];
            double ee = m->color_e[c][i];
            const int32_t *dofs = m->color_dofs[c] + (size_t)i * 24;
            double ue[24], eps[6], sig[6];
            for (int k = 0; k < 24; k++) ue[k] = u[dofs[k]];
            for (int r = 0; r < 6; r++) {
                double s = 0.0;
                for (int k = 0; k < 24; k++) s += base_bc[r * 24 + k] * ue[k];
                eps[r] = s * inv_h;
            }
            for (int r = 0; r < 6; r++) {
                double s = 0.0;
                for (int k = 0; k < 6; k++) s += d[r * 6 + k] * eps[k];
                sig[r] = s * ee;
            }
            double d1 = sig[0] - sig[1], d2 = sig[1] - sig[2], d3 = sig[2] - sig[0];
            vm[e] = sqrt(0.5 * (d1 * d1 + d2 * d2 + d3 * d3) + 3.0 * (sig[3] * sig[3] + sig[4] * sig[4] + sig[5] * sig[5]));
        }
}

void orc_average_element_to_node(const orc_model *m, const double *elem, double *nodal) {
    double *sum = (double *)calloc((size_t)m->n_nodes, sizeof(double));
    int *cnt = (int *)calloc((size_t)m->n_nodes, sizeof(int));
    for (int idx = 0; idx < m->n_active; idx++) {
        int e = m->active[idx], nodes[8];
        elem_nodes(m, e, nodes);
        for (int a = 0; a < 8; a++) {
            sum[nodes[a]] += elem[e];
            cnt[nodes[a]]++;
        }
    }
    for (int n = 0; n < m->n_nodes; n++) nodal[n] = cnt[n] > 0 ? sum[n] / cnt[n] : 0.0;
    free(sum);
    free(cnt);
}

/* SampleNodalScalar (Fem/StaticAnalysis.cs:278-366): local coordinates in float32 (Vector3 point, origin and the
 * float pitch), trilinear weights in double over the cell's active nodes in corner order, else the nearest active
 * node by float32 squared distance (first in index order on a tie). */
double orc_sample_nodal_scalar(const orc_model *m, const double *nodal, float px, float py, float pz, float ox,
                               float oy, float oz) {
    const float h = (float)m->h;
    const float rx = (px - ox) / h, ry = (py - oy) / h, rz = (pz - oz) / h;
    const double fx = floor((double)rx), fy = floor((double)ry), fz = floor((double)rz);
    const int nx1 = m->nx + 1, ny1 = m->ny + 1;
    if (fx >= 0.0 && fx < m->nx && fy >= 0.0 && fy < m->ny && fz >= 0.0 && fz < m->nz) {
        const int ix = (int)fx, iy = (int)fy, iz = (int)fz;
        const int n0 = ix + nx1 * (iy + ny1 * iz), n1 = n0 + 1, n3 = n0 + nx1, n2 = n3 + 1;
        const int n4 = n0 + nx1 * ny1, n5 = n4 + 1, n7 = n4 + nx1, n6 = n7 + 1;
        const double xi = (double)rx - ix, eta = (double)ry - iy, zeta = (double)rz - iz;
        const double w[8] = {(1.0 - xi) * (1.0 - eta) * (1.0 - zeta), xi * (1.0 - eta) * (1.0 - zeta),
                             xi * eta * (1.0 - zeta),                 (1.0 - xi) * eta * (1.0 - zeta),
                             (1.0 - xi) * (1.0 - eta) * zeta,         xi * (1.0 - eta) * zeta,
                             xi * eta * zeta,                         (1.0 - xi) * eta * zeta};
        const int nd[8] = {n0, n1, n2, n3, n4, n5, n6, n7};
        double ws = 0.0, vs = 0.0;
        for (int a = 0; a < 8; a++)
            if (m->active_nodes[nd[a]]) {
                ws += w[a];
                vs += w[a] * nodal[nd[a]];
            }
        if (ws > 1e-12) return vs / ws;
    }
    float best = HUGE_VALF;
    double val = 0.0;
    for (int n = 0; n < m->n_nodes; n++) {
        if (!m->active_nodes[n]) continue;
        const int i = n % nx1, t = n / nx1, j = t % ny1, k = t / ny1;
        const float dx = px - (ox + (float)i * h), dy = py - (oy + (float)j * h), dz = pz - (oz + (float)k * h);
        const float d2 = dx * dx + dy * dy + dz * dz;
        if (d2 < best) {
            best = d2;
            val = nodal[n];
        }
    }
    return val;
}

int orc_static_analysis(orc_model *m, const double *f, double *u, double tol, int max_iter,
                        int throw_on_nonconv, int n_threads, double *compliance, double *max_disp,
                        double *reactions, double *von_mises, orc_pcg_result *res) {
    if (m->active_fixed < 6) return ORC_E_INVALID_OPERATION;
    int n = m->n_dofs;
    double *b = (double *)malloc((size_t)n * sizeof(double));
    for (int i = 0; i < n; i++) b[i] = m->constrained[i] ? 0.0 : f[i];
    /* new FemOperator(model): E and DiagA from the model's densities (StaticAnalysis.cs:73). */
    orc_update_densities(m, m->density, m->penal);
    int st = orc_pcg_solve(m, b, u, tol, max_iter, n_threads, res);
    free(b);
    if (st != ORC_OK) return st;
    if (throw_on_nonconv && !res->converged) return ORC_E_NOT_CONVERGED;
    double c = 0.0;
    for (int i = 0; i < n; i++) c += f[i] * u[i];
    *compliance = c;
    if (reactions) {
        double *ku = (double *)malloc((size_t)n * sizeof(double));
        orc_apply_k(m, u, ku, n_threads);
        memset(reactions, 0, (size_t)n * sizeof(double));
        for (int i = 0; i < m->n_fixed; i++) reactions[m->fixed[i]] = ku[m->fixed[i]];
        free(ku);
    }
    double mx = 0.0;
    for (int nd = 0; nd < m->n_nodes; nd++) {
        if (!m->active_nodes[nd]) continue;
        double ux = u[3 * nd], uy = u[3 * nd + 1], uz = u[3 * nd + 2];
        double s = ux * ux + uy * uy + uz * uz;
        if (s > mx) mx = s;
    }
    *max_disp = sqrt(mx);
    if (von_mises) orc_von_mises(m, u, von_mises);
    return ORC_OK;
}

/* ------------------------------------------------------------------------------------------ */
/* DensityFilter (Optimization/DensityFilter.cs:34-303)                                        */
/* ------------------------------------------------------------------------------------------ */

struct orc_filter {
    int n;
    int32_t *fo, *fc;
    double *fw;
    int32_t *ao, *ac;
    double *aw;
    double *s;
};

int orc_filter_create(int nx, int ny, int nz, const uint8_t *domain, const int32_t *active,
                      int n_active, double rmin, orc_filter **out) {
    if (!domain || !active || !out) return ORC_E_ARGUMENT_NULL;
    int total = nx * ny * nz;
    int32_t *to_active = (int32_t *)malloc((size_t)total * sizeof(int32_t));
    for (int e = 0; e < total; e++) to_active[e] = -1;
    for (int i = 0; i < n_active; i++) to_active[active[i]] = i;
    int rc = (int)ceil(rmin);
    double rsq = rmin * rmin;
    int box = (2 * rc + 1) * (2 * rc + 1) * (2 * rc + 1);
    orc_filter *f = (orc_filter *)calloc(1, sizeof(orc_filter));
    f->n = n_active;
    f->fo = (int32_t *)malloc((size_t)(n_active + 1) * sizeof(int32_t));
    f->fc = (int32_t *)malloc((size_t)n_active * box * sizeof(int32_t) + 4);
    f->fw = (double *)malloc((size_t)n_active * box * sizeof(double) + 8);
    int nnz = 0;
    for (int ei = 0; ei < n_active; ei++) {
        int e = active[ei];
        int ez = e / (nx * ny), rem = e % (nx * ny), ey = rem / nx, ex = rem % nx;
        f->fo[ei] = nnz;
        double hs = 0.0;
        int row0 = nnz;
        int x0 = ex - rc > 0 ? ex - rc : 0, x1 = ex + rc < nx - 1 ? ex + rc : nx - 1;
        int y0 = ey - rc > 0 ? ey - rc : 0, y1 = ey + rc < ny - 1 ? ey + rc : ny - 1;
        int z0 = ez - rc > 0 ? ez - rc : 0, z1 = ez + rc < nz - 1 ? ez + rc : nz - 1;
        for (int jz = z0; jz <= z1; jz++) {
            int dz = jz - ez;
            for (int jy = y0; jy <= y1; jy++) {
                int dy = jy - ey, dyz = dz * dz + dy * dy;
                for (int jx = x0; jx <= x1; jx++) {
                    int dx = jx - ex, dsq = dyz + dx * dx;
                    if (!(dsq < rsq)) continue;
                    int nb = jx + nx * (jy + ny * jz);
                    if (!domain[nb]) continue;
                    int nbi = to_active[nb];
                    if (nbi < 0) continue;
                    double dist = sqrt((double)dsq);
                    if (dist < rmin) {
                        double w = rmin - dist;
                        f->fc[nnz] = nbi;
                        f->fw[nnz] = w;
                        nnz++;
                        hs += w;
                    }
                }
            }
        }
        double inv_hs = hs > 0.0 ? 1.0 / hs : 1.0;
        for (int k = row0; k < nnz; k++) f->fw[k] *= inv_hs;
    }
    f->fo[n_active] = nnz;
    /* Transpose (DensityFilter.cs:152-185). */
    f->ao = (int32_t *)calloc((size_t)(n_active + 1), sizeof(int32_t));
    f->ac = (int32_t *)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(int32_t));
    f->aw = (double *)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(double));
    int32_t *cnt = (int32_t *)calloc((size_t)(n_active > 0 ? n_active : 1), sizeof(int32_t));
    for (int i = 0; i < nnz; i++) cnt[f->fc[i]]++;
    int off = 0;
    for (int i = 0; i < n_active; i++) {
        f->ao[i] = off;
        off += cnt[i];
    }
    f->ao[n_active] = off;
    int32_t *cur = (int32_t *)malloc((size_t)(n_active + 1) * sizeof(int32_t));
    memcpy(cur, f->ao, (size_t)(n_active + 1) * sizeof(int32_t));
    for (int r = 0; r < n_active; r++)
        for (int k = f->fo[r]; k < f->fo[r + 1]; k++) {
            int t = cur[f->fc[k]]++;
            f->ac[t] = r;
            f->aw[t] = f->fw[k];
        }
    f->s = (double *)malloc((size_t)(n_active > 0 ? n_active : 1) * sizeof(double));
    for (int i = 0; i < n_active; i++) {
        double s = 0.0;
        for (int k = f->ao[i]; k < f->ao[i + 1]; k++) s += f->aw[k];
        f->s[i] = s;
    }
    free(cur);
    free(cnt);
    free(to_active);
    *out = f;
    return ORC_OK;
}

void orc_filter_destroy(orc_filter *f) {
    if (!f) return;
    free(f->fo); free(f->fc); free(f->fw);
    free(f->ao); free(f->ac); free(f->aw);
    free(f->s);
    free(f);
}

int orc_filter_nnz(const orc_filter *f) { return f->fo[f->n]; }
void orc_filter_s(const orc_filter *f, double *s) { memcpy(s, f->s, (size_t)f->n * sizeof(double)); }

static void csr_apply(int n, const int32_t *o, const int32_t *c, const double *w, const double *in, double *out) {
#pragma omp parallel for schedule(static)
    for (int r = 0; r < n; r++) {
        double s = 0.0;
        for (int k = o[r]; k < o[r + 1]; k++) s += w[k] * in[c[k]];
        out[r] = s;
    }
}

void orc_filter_apply(const orc_filter *f, const double *in, double *out) { csr_apply(f->n, f->fo, f->fc, f->fw, in, out); }
void orc_filter_apply_adjoint(const orc_filter *f, const double *in, double *out) { csr_apply(f->n, f->ao, f->ac, f->aw, in, out); }

/* ------------------------------------------------------------------------------------------ */
/* Sensitivities (Optimization/SensitivityAnalysis.cs:13-146)                                  */
/* ------------------------------------------------------------------------------------------ */

void orc_sensitivities(const orc_model *m, const double *u, const double *rho, double *dc, double penal) {
    double de = m->e0 - m->emin;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < m->n_active; k++) {
        int nodes[8];
        elem_nodes(m, m->active[k], nodes);
        double ue[24];
        for (int a = 0; a < 8; a++)
            for (int d = 0; d < 3; d++) ue[3 * a + d] = u[3 * nodes[a] + d];
        double se = 0.0;
        for (int r = 0; r < 24; r++) {
            double s = 0.0;
            for (int c = 0; c < 24; c++) s += m->ke[r * 24 + c] * ue[c];
            se += ue[r] * s;
        }
        double d_e = penal * pow_penal(rho[k], penal - 1.0) * de;
        dc[k] = -d_e * se;
    }
}

/* ------------------------------------------------------------------------------------------ */
/* OC (Optimization/OcOptimizer.cs:12-175)                                                     */
/* ------------------------------------------------------------------------------------------ */

static inline double clampd(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }

static double evaluate_volume(const double *rho, const double *dc, const double *s, const uint8_t *passive,
                              int n, double lambda, double move, double rho_min) {
    double vol = 0.0;
    for (int i = 0; i < n; i++) {
        if (passive && passive[i]) {
            vol += s[i] * 1.0;
            continue;
        }
        double r = rho[i], si = s[i];
        double lo = fmax(rho_min, r - move), hi = fmin(1.0, r + move);
        double neg = -dc[i];
        if (neg <= 0.0 || si <= 0.0) vol += si * lo;
        else {
            double b = neg / (lambda * si);
            vol += si * clampd(r * sqrt(b), lo, hi);
        }
    }
    return vol;
}

double orc_oc_update(const double *rho, const double *dc, const double *s, const uint8_t *passive,
                     double *rho_new, int n, double vf, double move, double rho_min,
                     double lambda_guess, int max_bisect, double vol_tol, int *n_evals) {
    int evals = 0;
    double total_s = 0.0;
    for (int i = 0; i < n; i++) total_s += s[i];
    double target = vf * total_s;
    int good = lambda_guess > 0.0 && !isinf(lambda_guess) && !isnan(lambda_guess);
    double l1 = good ? lambda_guess * 0.5 : 1e-12;
    double l2 = good ? lambda_guess * 2.0 : 1e12;
    if (l1 <= 1e-50) l1 = 1e-12;
    if (l2 <= l1) l2 = l1 * 10.0;
    double v1 = evaluate_volume(rho, dc, s, passive, n, l1, move, rho_min);
    double v2 = evaluate_volume(rho, dc, s, passive, n, l2, move, rho_min);
    evals += 2;
    int steps = 0;
    while (v1 < target && l1 > 1e-50 && steps++ < 50) {
        l2 = l1; v2 = v1; l1 *= 0.1;
        v1 = evaluate_volume(rho, dc, s, passive, n, l1, move, rho_min);
        evals++;
    }
    steps = 0;
    while (v2 > target && l2 < 1e50 && steps++ < 50) {
        l1 = l2; v1 = v2; l2 *= 10.0;
        v2 = evaluate_volume(rho, dc, s, passive, n, l2, move, rho_min);
        evals++;
    }
    (void)v1;
    double lmid = sqrt(l1 * l2);
    for (int step = 0; step < max_bisect; step++) {
        lmid = sqrt(l1 * l2);
        double vmid = evaluate_volume(rho, dc, s, passive, n, lmid, move, rho_min);
        evals++;
        double rel = fabs(vmid - target) / fmax(target, 1e-12);
        if (rel <= vol_tol) break;
        if (vmid > target) l1 = lmid;
        else l2 = lmid;
        if (fabs(l2 - l1) <= 1e-15 * lmid) break;
    }
    for (int i = 0; i < n; i++) {
        if (passive && passive[i]) {
            rho_new[i] = 1.0;
            continue;
        }
        double r = rho[i], si = s[i];
        double lo = fmax(rho_min, r - move), hi = fmin(1.0, r + move);
        double neg = -dc[i];
        if (neg <= 0.0 || si <= 0.0) rho_new[i] = lo;
        else {
            double b = neg / (lmid * si);
            rho_new[i] = clampd(r * sqrt(b), lo, hi);
        }
    }
    if (n_evals) *n_evals = evals;
    return lmid;
}

/* ------------------------------------------------------------------------------------------ */
/* PassiveRegion.Build (Voxelization/PassiveRegion.cs:13-62)                                   */
/* ------------------------------------------------------------------------------------------ */

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

void orc_passive_region(int nx, int ny, int nz, const int32_t *nodes, int n_nodes, int radius, uint8_t *mask) {
    memset(mask, 0, (size_t)nx * ny * nz);
    int nx1 = nx + 1, nx1ny1 = (nx + 1) * (ny + 1);
    for (int q = 0; q < n_nodes; q++) {
        int node = nodes[q];
        int nk = node / nx1ny1, rem = node % nx1ny1, nj = rem / nx1, ni = rem % nx1;
        int x0 = clampi(ni - radius, 0, nx - 1), x1 = clampi(ni + radius - 1, 0, nx - 1);
        int y0 = clampi(nj - radius, 0, ny - 1), y1 = clampi(nj + radius - 1, 0, ny - 1);
        int z0 = clampi(nk - radius, 0, nz - 1), z1 = clampi(nk + radius - 1, 0, nz - 1);
        for (int iz = z0; iz <= z1; iz++)
            for (int iy = y0; iy <= y1; iy++)
                for (int ix = x0; ix <= x1; ix++) mask[nx * (iy + ny * iz) + ix] = 1;
    }
}

/* ------------------------------------------------------------------------------------------ */
/* HeavisideProjection (Optimization/HeavisideProjection.cs:12-37): smooth threshold projection */
/* ------------------------------------------------------------------------------------------ */

/* rho_bar = clamp((tanh(beta eta) + tanh(beta (rt - eta))) * (1 / denom), 0, 1),
 * denom = tanh(beta eta) + tanh(beta (1 - eta)) (:14-23). */
void orc_heaviside_project(const double *rho_tilde, double *rho_bar, int n, double beta, double eta) {
    const double denom = tanh(beta * eta) + tanh(beta * (1.0 - eta));
    const double inv_denom = 1.0 / denom;
    const double tbe = tanh(beta * eta);
    for (int i = 0; i < n; i++) {
        const double num = tbe + tanh(beta * (rho_tilde[i] - eta));
        rho_bar[i] = clampd(num * inv_denom, 0.0, 1.0);
    }
}

/* d rho_bar / d rho_tilde = max(beta (1 - t^2) / denom, 1e-12), t = tanh(beta (rt - eta)) (:26-37). */
void orc_heaviside_derivative(const double *rho_tilde, double *deriv, int n, double beta, double eta) {
    const double denom = tanh(beta * eta) + tanh(beta * (1.0 - eta));
    const double inv_denom = 1.0 / denom;
    for (int i = 0; i < n; i++) {
        const double t = tanh(beta * (rho_tilde[i] - eta));
        const double d = beta * (1.0 - t * t) * inv_denom;
        deriv[i] = fmax(d, 1e-12);
    }
}

/* ------------------------------------------------------------------------------------------ */
/* TopologyOptimizer.Run (Optimization/TopologyOptimizer.cs:16-407)                              */
/* ------------------------------------------------------------------------------------------ */

/* Several load cases (f, warm_u, u_out hold n_cases box-DOF vectors back to back): the objective is the
 * summed compliance sum_k f_k.u_k. Each case is solved on its own (warm-started from its previous
 * solution) and the raw sensitivities are summed in case order before the one adjoint filter; solver
 * iterations are summed, the residual and max|u| are the largest over the cases. n_cases = 1 is
 * TopologyOptimizer.Run itself. (The reference sums load cases into one RHS, Fem/FeModel.cs:267-318;
 * independent cases are this restatement's extension for the multi-RHS configuration.) */
int orc_simp_run_cases(orc_model *m, const double *f, int n_cases, const orc_simp_options *opt,
                       const uint8_t *passive_mask_box, const int32_t *patch_nodes, int n_patch_nodes,
                       const float *initial_density, int initial_density_len, const double *warm_u,
                       int n_threads, orc_snapshot *snaps, int *n_snaps, int *converged_out,
                       double *rho_out, double *rho_bar_out, double *u_out) {
    int na = m->n_active, ne = m->n_elem, nd = m->n_dofs;
    if (n_cases < 1) return ORC_E_ARGUMENT;
    double h = m->h;
    double vf = opt->volume_fraction;
    double rmin = opt->filter_radius_mm / h; /* :30 */
    if (rmin < 1.5 && opt->enforce_min_filter_radius) rmin = 1.5;
    *n_snaps = 0;
    *converged_out = 0;

    uint8_t *passive = (uint8_t *)calloc((size_t)(na > 0 ? na : 1), 1);
    if (passive_mask_box)
        for (int i = 0; i < na; i++)
            if (passive_mask_box[m->active[i]]) passive[i] = 1;
    if (opt->protect_boundary_hinges) { /* :59-84 */
        int np = n_patch_nodes + m->n_fixed;
        int32_t *pn = (int32_t *)malloc((size_t)(np > 0 ? np : 1) * sizeof(int32_t));
        for (int i = 0; i < n_patch_nodes; i++) pn[i] = patch_nodes[i];
        for (int i = 0; i < m->n_fixed; i++) pn[n_patch_nodes + i] = m->fixed[i] / 3;
        uint8_t *shell = (uint8_t *)malloc((size_t)ne);
        orc_passive_region(m->nx, m->ny, m->nz, pn, np, 1, shell);
        for (int i = 0; i < na; i++)
            if (shell[m->active[i]]) passive[i] = 1;
        free(shell);
        free(pn);
    }
    orc_filter *flt;
    /* model.Domain != 0 is exactly the active set: rebuild it box-sized for the filter (:87). */
    uint8_t *domain = (uint8_t *)calloc((size_t)ne, 1);
    for (int i = 0; i < na; i++) domain[m->active[i]] = 1;
    orc_filter_create(m->nx, m->ny, m->nz, domain, m->active, na, rmin, &flt);
    free(domain);
    const double *s = flt->s;

    double *rho = (double *)malloc((size_t)na * sizeof(double));
    if (initial_density) {
        if (initial_density_len == ne) for (int i = 0; i < na; i++) rho[i] = initial_density[m->active[i]];
        else if (initial_density_len == na) for (int i = 0; i < na; i++) rho[i] = initial_density[i];
        else {
            free(rho); free(passive); orc_filter_destroy(flt);
            return ORC_E_ARGUMENT;
        }
    } else {
        for (int i = 0; i < na; i++) rho[i] = passive[i] ? 1.0 : vf;
    }
    for (int i = 0; i < na; i++)
        if (passive[i]) rho[i] = 1.0;

    double *rho_bar = (double *)calloc((size_t)na, sizeof(double));
    const int heav = opt->enable_heaviside;
    double *rho_tilde = (double *)calloc((size_t)na, sizeof(double));
    double *dproj = (double *)calloc((size_t)na, sizeof(double));
    double beta = 1.0; /* :156 */
    double *rho_old = (double *)calloc((size_t)na, sizeof(double));
    double *dc_bar = (double *)calloc((size_t)na, sizeof(double));
    double *dc = (double *)calloc((size_t)na, sizeof(double));
    double *dc_k = (double *)calloc((size_t)na, sizeof(double));
    const size_t ndc = (size_t)nd * (size_t)n_cases;
    double *u = (double *)calloc(ndc, sizeof(double));
    if (warm_u) memcpy(u, warm_u, ndc * sizeof(double));
    /* new FemOperator(model): E from the model's current densities (:142). */
    orc_update_densities(m, m->density, m->penal);
    double *b = (double *)malloc(ndc * sizeof(double));
    for (size_t i = 0; i < ndc; i++) b[i] = m->constrained[i % (size_t)nd] ? 0.0 : f[i];

    double lambda = 1.0, max_change = 1.0;
    double hist[6];
    int hist_n = 0;
    int status = ORC_OK;
    int start = opt->start_iteration, max_it = opt->max_iterations;
    for (int iter = start; iter <= max_it; iter++) {
        double penal = opt->penal;
        if (opt->continuation) penal = iter < 20 ? 1.0 : (iter < 40 ? 2.0 : opt->penal);
        for (int i = 0; i < na; i++)
            if (passive[i]) rho[i] = 1.0;
        if (heav) { /* :193-197 */
            orc_filter_apply(flt, rho, rho_tilde);
            orc_heaviside_project(rho_tilde, rho_bar, na, beta, 0.5);
        } else {
            orc_filter_apply(flt, rho, rho_bar);
        }
        for (int i = 0; i < na; i++) m->density[m->active[i]] = rho_bar[i];
        orc_update_densities(m, m->density, penal);
        double tol;
        if (iter == start) tol = opt->solver_tol_initial;
        else { /* :219-223 */
            double a = clampd((max_change - 0.01) / 0.09, 0.0, 1.0);
            double li = log10(fmax(opt->solver_tol_initial, 1e-15));
            double lf = log10(fmax(opt->solver_tol_final, 1e-15));
            tol = pow(10.0, lf + a * (li - lf));
        }
        double c = 0.0, mds = 0.0, res_max = 0.0;
        int iters = 0;
        for (int k = 0; k < n_cases; k++) {
            const double *fk = f + (size_t)k * nd;
            double *uk = u + (size_t)k * nd;
            orc_pcg_result pr;
            int st = orc_pcg_solve(m, b + (size_t)k * nd, uk, tol, opt->solver_max_iterations, n_threads, &pr);
            if (st != ORC_OK) { status = st; break; }
            if (!pr.converged) { status = ORC_E_NOT_CONVERGED; break; }
            iters += pr.iterations;
            if (pr.final_residual > res_max) res_max = pr.final_residual;
            double ck = 0.0;
            for (int i = 0; i < nd; i++) ck += fk[i] * uk[i];
            c += ck;
            for (int q = 0; q < m->n_nodes; q++) {
                if (!m->active_nodes[q]) continue;
                double ux = uk[3 * q], uy = uk[3 * q + 1], uz = uk[3 * q + 2];
                double ss = ux * ux + uy * uy + uz * uz;
                if (ss > mds) mds = ss;
            }
            if (k == 0) {
                orc_sensitivities(m, uk, rho_bar, dc_bar, penal);
            } else {
                orc_sensitivities(m, uk, rho_bar, dc_k, penal);
                for (int i = 0; i < na; i++) dc_bar[i] += dc_k[i];
            }
        }
        if (status != ORC_OK) break;
        if (heav) { /* :281-289: chain rule through the projection, at the current beta */
            orc_heaviside_derivative(rho_tilde, dproj, na, beta, 0.5);
            for (int i = 0; i < na; i++) dc_bar[i] = dc_bar[i] * dproj[i];
        }
        orc_filter_apply_adjoint(flt, dc_bar, dc);
        memcpy(rho_old, rho, (size_t)na * sizeof(double));
        lambda = orc_oc_update(rho, dc, s, passive, rho, na, vf, opt->move_limit, opt->rho_min, lambda, 60, 1e-8, NULL);
        if (heav && iter % 25 == 0 && beta < 8.0) beta = fmin(8.0, beta * 2.0); /* :315-318 */
        max_change = 0.0;
        double vol = 0.0, grey = 0.0;
        for (int i = 0; i < na; i++) {
            double d = fabs(rho[i] - rho_old[i]);
            if (d > max_change) max_change = d;
            vol += rho_bar[i];
            const double rb = heav ? rho_tilde[i] : rho_bar[i]; /* :333 */
            grey += rb * (1.0 - rb);
        }
        orc_snapshot *sn = &snaps[(*n_snaps)++];
        sn->iteration = iter;
        sn->compliance = c;
        sn->volume_fraction = vol / na;
        sn->max_change = max_change;
        sn->grey_level = (4.0 * grey / na) * 100.0;
        sn->solver_iterations = iters;
        sn->solver_residual = res_max;
        sn->max_displacement = sqrt(mds);
        sn->lambda = lambda;
        sn->solver_tolerance = tol;
        if (hist_n == 6) {
            memmove(hist, hist + 1, 5 * sizeof(double));
            hist[5] = c;
        } else {
            hist[hist_n++] = c;
        }
        if (hist_n == 6) {
            double rc = fabs(c - hist[0]) / c;
            if (max_change < 0.01 && rc < 1e-3) {
                *converged_out = 1;
                break;
            }
        }
    }
    if (rho_out) memcpy(rho_out, rho, (size_t)na * sizeof(double));
    if (rho_bar_out) memcpy(rho_bar_out, rho_bar, (size_t)na * sizeof(double));
    if (u_out) memcpy(u_out, u, ndc * sizeof(double));
    free(passive); free(rho); free(rho_bar); free(rho_old); free(dc_bar); free(dc); free(dc_k); free(u); free(b);
    free(rho_tilde); free(dproj);
    orc_filter_destroy(flt);
    return status;
}

int orc_simp_run(orc_model *m, const double *f, const orc_simp_options *opt,
                 const uint8_t *passive_mask_box, const int32_t *patch_nodes, int n_patch_nodes,
                 const float *initial_density, int initial_density_len, const double *warm_u,
                 int n_threads, orc_snapshot *snaps, int *n_snaps, int *converged_out,
                 double *rho_out, double *rho_bar_out, double *u_out) {
    return orc_simp_run_cases(m, f, 1, opt, passive_mask_box, patch_nodes, n_patch_nodes, initial_density,
                              initial_density_len, warm_u, n_threads, snaps, n_snaps, converged_out, rho_out,
                              rho_bar_out, u_out);
}
