/*
 * Oracle (test infrastructure): finite-volume monodomain operators on a uniform hex
 * block, restating monodomainSolver.C:103-150 (fvc::laplacian / fvm::laplacian with
 * a tensor conductivity, "Gauss linear corrected") and myocardiumDomain.C:307-374,
 * 487-550.  The OpenFOAM operators (ESI v2312..v2512 gaussLaplacianScheme with tensor
 * gamma) are restated from their documented discretisation:
 *   SfGamma     = Sf . sigma_f            (sigma_f linear-interpolated, w = 1/2)
 *   SfGammaSn   = SfGamma . n             (orthogonal coefficient)
 *   SfGammaCorr = SfGamma - SfGammaSn n   (tangential part, explicit)
 *   flux_f      = SfGammaSn * (V_N - V_P)/h + SfGammaCorr . interp(grad V)
 * grad V is the Gauss-linear cell gradient with zeroGradient walls (boundary face
 * value = cell value); its wall value is the cell gradient with the normal
 * component replaced by snGrad = 0.  Wall faces carry only the correction flux.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"
#ifdef _OPENMP
#include <omp.h>
#endif

int orc_mesh_build(orc_mesh_t* m, int nx, int ny, int nz, double hx, double hy, double hz)
{
    memset(m, 0, sizeof *m);
    m->nx = nx; m->ny = ny; m->nz = nz; m->hx = hx; m->hy = hy; m->hz = hz;
    m->n_cells = (long)nx * ny * nz;
    m->n_faces = (long)(nx - 1) * ny * nz + (long)nx * (ny - 1) * nz + (long)nx * ny * (nz - 1);
    m->owner = malloc(sizeof(int32_t) * m->n_faces);
    m->neighbour = malloc(sizeof(int32_t) * m->n_faces);
    m->face_dir = malloc(m->n_faces);
    if (!m->owner || !m->neighbour || !m->face_dir) return -1;
    long f = 0;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                int32_t c = i + nx * (j + ny * k);
                if (i < nx - 1) { m->owner[f] = c; m->neighbour[f] = c + 1; m->face_dir[f++] = 0; }
                if (j < ny - 1) { m->owner[f] = c; m->neighbour[f] = c + nx; m->face_dir[f++] = 1; }
                if (k < nz - 1) { m->owner[f] = c; m->neighbour[f] = c + nx * ny; m->face_dir[f++] = 2; }
            }
    return 0;
}

void orc_mesh_free(orc_mesh_t* m)
{
    free(m->owner); free(m->neighbour); free(m->face_dir);
    memset(m, 0, sizeof *m);
}

static inline double face_area(const orc_mesh_t* m, int d)
{
    return d == 0 ? m->hy * m->hz : d == 1 ? m->hx * m->hz : m->hx * m->hy;
}
static inline double cell_h(const orc_mesh_t* m, int d)
{
    return d == 0 ? m->hx : d == 1 ? m->hy : m->hz;
}

/* Gauss linear gradient, zeroGradient walls: grad[3n] */
static void gauss_grad(const orc_mesh_t* m, const double* V, double* grad, int nthreads)
{
    const int nx = m->nx, ny = m->ny, nz = m->nz;
    const double vol = m->hx * m->hy * m->hz;
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
    for (long c = 0; c < m->n_cells; ++c) {
        int i = (int)(c % nx), j = (int)((c / nx) % ny), k = (int)(c / ((long)nx * ny));
        const long st[3] = {1, nx, (long)nx * ny};
        const int idx[3] = {i, j, k}, nn[3] = {nx, ny, nz};
        for (int d = 0; d < 3; ++d) {
            double A = face_area(m, d);
            double vp = V[c];
            double fplus = idx[d] < nn[d] - 1 ? 0.5 * (vp + V[c + st[d]]) : vp;
            double fminus = idx[d] > 0 ? 0.5 * (vp + V[c - st[d]]) : vp;
            grad[3 * c + d] = A * (fplus - fminus) / vol;
        }
    }
}

/* correction flux through the face with normal d between P (owner side) and the
 * interpolated gradient gf, with face tensor s (row-major 3x3): A * sum_{e!=d} s[d][e] gf[e] */
static inline double corr_flux(double A, int d, const double* s, const double* gf)
{
    double q = 0.0;
    for (int e = 0; e < 3; ++e)
        if (e != d) q += s[3 * d + e] * gf[e];
    return A * q;
}

static void div_flux(const orc_mesh_t* m, const double* sigma, const double* V,
                     const double* grad, int with_orth, int with_corr, double* out,
                     int nthreads)
{
    const int nx = m->nx, ny = m->ny, nz = m->nz;
    const double vol = m->hx * m->hy * m->hz;
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
    for (long c = 0; c < m->n_cells; ++c) {
        int i = (int)(c % nx), j = (int)((c / nx) % ny), k = (int)(c / ((long)nx * ny));
        const long st[3] = {1, nx, (long)nx * ny};
        const int idx[3] = {i, j, k}, nn[3] = {nx, ny, nz};
        double acc = 0.0;
        for (int d = 0; d < 3; ++d) {
            const double A = face_area(m, d), h = cell_h(m, d);
            for (int side = 0; side < 2; ++side) {
                const int interior = side == 0 ? idx[d] < nn[d] - 1 : idx[d] > 0;
                const double sgn = side == 0 ? 1.0 : -1.0;
                if (interior) {
                    long nb = side == 0 ? c + st[d] : c - st[d];
                    double sf[9], gf[3];
                    for (int q = 0; q < 9; ++q) sf[q] = 0.5 * (sigma[9 * c + q] + sigma[9 * nb + q]);
                    for (int q = 0; q < 3; ++q) gf[q] = 0.5 * (grad ? grad[3 * c + q] + grad[3 * nb + q] : 0.0);
                    /* flux out of c through this face; Sf = sgn*A*e_d */
                    double fl = 0.0;
                    if (with_orth) fl += A * sf[4 * d] * (V[nb] - V[c]) / h;
                    if (with_corr) fl += sgn * corr_flux(A, d, sf, gf);
                    acc += fl;
                } else if (with_corr) {
                    /* wall: sigma_b = sigma_P; grad_b = grad_P with normal part = snGrad = 0 */
                    double gb[3] = {grad[3 * c], grad[3 * c + 1], grad[3 * c + 2]};
                    gb[d] = 0.0;
                    acc += sgn * corr_flux(A, d, sigma + 9 * c, gb);
                }
            }
        }
        out[c] = acc / vol;
    }
}

void orc_laplacian(const orc_mesh_t* m, const double* sigma, const double* V, double* out,
                   int nthreads)
{
    double* grad = malloc(sizeof(double) * 3 * m->n_cells);
    gauss_grad(m, V, grad, nthreads);
    div_flux(m, sigma, V, grad, 1, 1, out, nthreads);
    free(grad);
}

void orc_correction_div(const orc_mesh_t* m, const double* sigma, const double* V,
                        double* out, int nthreads)
{
    double* grad = malloc(sizeof(double) * 3 * m->n_cells);
    gauss_grad(m, V, grad, nthreads);
    div_flux(m, sigma, V, grad, 0, 1, out, nthreads);
    free(grad);
}

void orc_assemble_implicit(const orc_mesh_t* m, const double* sigma, double chiCm,
                           double dt, double coefft, double* diag, double* upper)
{
    /* per unit volume: (chiCm*coefft/dt + sum a_f) V_P - sum a_f V_N,
     * a_f = SfGammaSn_f * deltaCoeff_f / vol; upper[f] holds -a_f */
    const double vol = m->hx * m->hy * m->hz;
    for (long c = 0; c < m->n_cells; ++c) diag[c] = chiCm * coefft / dt;
    for (long f = 0; f < m->n_faces; ++f) {
        int d = m->face_dir[f];
        long P = m->owner[f], N = m->neighbour[f];
        double snn = 0.5 * (sigma[9 * P + 4 * d] + sigma[9 * N + 4 * d]);
        double a = face_area(m, d) * snn / cell_h(m, d) / vol;
        upper[f] = -a;
        diag[P] += a;
        diag[N] += a;
    }
}

void orc_ldu_amul(const orc_mesh_t* m, const double* diag, const double* upper,
                  const double* x, double* y)
{
    for (long c = 0; c < m->n_cells; ++c) y[c] = diag[c] * x[c];
    for (long f = 0; f < m->n_faces; ++f) {
        long P = m->owner[f], N = m->neighbour[f];
        y[P] += upper[f] * x[N];
        y[N] += upper[f] * x[P];
    }
}

/* the hex block's faces in cell-gather form: for cell c, the index of its face to
 * +x/+y/+z (or -1) in the owner-ordered face list, so y = A x runs in parallel */
static void hex_face_index(const orc_mesh_t* m, long* fidx)
{
    for (long c = 0; c < 3 * m->n_cells; ++c) fidx[c] = -1;
    for (long f = 0; f < m->n_faces; ++f) fidx[3 * (long)m->owner[f] + m->face_dir[f]] = f;
}

static void hex_amul(const orc_mesh_t* m, const long* fidx, const double* diag,
                     const double* upper, const double* x, double* y, int nthreads)
{
    const long st[3] = {1, m->nx, (long)m->nx * m->ny};
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (long c = 0; c < m->n_cells; ++c) {
        double acc = diag[c] * x[c];
        for (int d = 0; d < 3; ++d) {
            long f = fidx[3 * c + d];
            if (f >= 0) acc += upper[f] * x[c + st[d]];
            long b = c - st[d];
            if (b >= 0) {
                long g = fidx[3 * b + d];
                if (g >= 0 && m->neighbour[g] == c) acc += upper[g] * x[b];
            }
        }
        y[c] = acc;
    }
}

int orc_pcg(const orc_mesh_t* m, const double* diag, const double* upper, const double* b,
            double* x, double tol, int max_iter, double* res)
{
    const long n = m->n_cells;
    int nt = 1;
#ifdef _OPENMP
    nt = omp_get_max_threads();
#endif
    double* r = malloc(sizeof(double) * n);
    double* z = malloc(sizeof(double) * n);
    double* p = malloc(sizeof(double) * n);
    double* q = malloc(sizeof(double) * n);
    long* fidx = malloc(sizeof(long) * 3 * n);
    hex_face_index(m, fidx);
    hex_amul(m, fidx, diag, upper, x, q, nt);
    double bn = 0.0, rr = 0.0, rz = 0.0;
#pragma omp parallel for schedule(static) num_threads(nt) reduction(+ : bn, rr, rz)
    for (long i = 0; i < n; ++i) {
        r[i] = b[i] - q[i];
        z[i] = r[i] / diag[i];
        p[i] = z[i];
        bn += b[i] * b[i]; rr += r[i] * r[i]; rz += r[i] * z[i];
    }
    bn = sqrt(bn);
    if (bn == 0.0) bn = 1.0;
    int it = 0;
    while (sqrt(rr) > tol * bn && it < max_iter) {
        hex_amul(m, fidx, diag, upper, p, q, nt);
        double pq = 0.0;
#pragma omp parallel for schedule(static) num_threads(nt) reduction(+ : pq)
        for (long i = 0; i < n; ++i) pq += p[i] * q[i];
        const double alpha = rz / pq;
        double rz_new = 0.0;
        rr = 0.0;
#pragma omp parallel for schedule(static) num_threads(nt) reduction(+ : rz_new, rr)
        for (long i = 0; i < n; ++i) {
            x[i] += alpha * p[i];
            r[i] -= alpha * q[i];
            z[i] = r[i] / diag[i];
            rz_new += r[i] * z[i];
            rr += r[i] * r[i];
        }
        const double beta = rz_new / rz;
        rz = rz_new;
#pragma omp parallel for schedule(static) num_threads(nt)
        for (long i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
        ++it;
    }
    /* true residual */
    hex_amul(m, fidx, diag, upper, x, q, nt);
    rr = 0.0;
#pragma omp parallel for schedule(static) num_threads(nt) reduction(+ : rr)
    for (long i = 0; i < n; ++i) rr += (b[i] - q[i]) * (b[i] - q[i]);
    if (res) *res = sqrt(rr) / bn;
    free(r); free(z); free(p); free(q); free(fidx);
    return it;
}

/* ---------------------------------------------- OpenFOAM lduMatrix PCG semantics
 * Restated from OpenFOAM (ESI v2312..v2512; third party, not in the reference tree),
 * the solver the reference's fvSolution selects for Vm (PCG, tolerance, relTol, maxIter;
 * tutorials/regressionTests/NiedererEtAl2012/system/fvSolution):
 *  - normalisation factor of lduMatrix::solver::normFactor:
 *      sum_i |(A psi)_i - sumA_i xRef| + |b_i - sumA_i xRef| + 1e-20,
 *    xRef = mean(psi), sumA = row sums of A;
 *  - residual = sum_i |b - A psi|_i / normFactor (initial and final);
 *  - converged when residual < tolerance, or relTol > 1e-20 and residual < relTol x initial
 *    (SolverPerformance::checkConvergence), tested before the first iteration;
 *  - PCG::solve iteration: w = M^-1 r, (w,r), p = w (+ beta p), w = A p, (w,p), singular
 *    when |(w,p)|/normFactor < 1e-300, psi += alpha p, r -= alpha w;
 *  - preconditioner 0 = diagonal (Jacobi), 1 = DIC (incomplete Cholesky of the lduMatrix
 *    in face order: rD[u] -= upper^2 / rD[l], then the forward and backward face sweeps). */
static double of_norm_factor(const orc_mesh_t* m, const double* diag, const double* upper,
                             const double* psi, const double* b, const double* Apsi,
                             double* sumA)
{
    const long n = m->n_cells;
    double xs = 0.0;
    for (long i = 0; i < n; ++i) xs += psi[i];
    const double xRef = xs / (double)n;
    for (long i = 0; i < n; ++i) sumA[i] = diag[i];
    for (long f = 0; f < m->n_faces; ++f) {
        sumA[m->owner[f]] += upper[f];
        sumA[m->neighbour[f]] += upper[f];
    }
    double nf = 0.0;
    for (long i = 0; i < n; ++i)
        nf += fabs(Apsi[i] - sumA[i] * xRef) + fabs(b[i] - sumA[i] * xRef);
    return nf + 1e-20;
}

static void dic_rd(const orc_mesh_t* m, const double* diag, const double* upper, double* rD)
{
    const long n = m->n_cells;
    for (long i = 0; i < n; ++i) rD[i] = diag[i];
    for (long f = 0; f < m->n_faces; ++f)
        rD[m->neighbour[f]] -= upper[f] * upper[f] / rD[m->owner[f]];
    for (long i = 0; i < n; ++i) rD[i] = 1.0 / rD[i];
}

static void precondition(const orc_mesh_t* m, int kind, const double* diag,
                         const double* upper, const double* rD, const double* r, double* w)
{
    const long n = m->n_cells;
    if (kind == 0) {
        for (long i = 0; i < n; ++i) w[i] = r[i] / diag[i];
        return;
    }
    for (long i = 0; i < n; ++i) w[i] = rD[i] * r[i];
    for (long f = 0; f < m->n_faces; ++f)
        w[m->neighbour[f]] -= rD[m->neighbour[f]] * upper[f] * w[m->owner[f]];
    for (long f = m->n_faces - 1; f >= 0; --f)
        w[m->owner[f]] -= rD[m->owner[f]] * upper[f] * w[m->neighbour[f]];
}

int orc_pcg_of(const orc_mesh_t* m, const double* diag, const double* upper, const double* b,
               double* x, double tol, double rel_tol, int max_iter, int precond,
               double* init_res, double* final_res)
{
    const long n = m->n_cells;
    double* wA = malloc(sizeof(double) * n);
    double* rA = malloc(sizeof(double) * n);
    double* pA = malloc(sizeof(double) * n);
    double* rD = malloc(sizeof(double) * n);
    orc_ldu_amul(m, diag, upper, x, wA);
    for (long i = 0; i < n; ++i) rA[i] = b[i] - wA[i];
    const double nf = of_norm_factor(m, diag, upper, x, b, wA, pA);
    double l1 = 0.0;
    for (long i = 0; i < n; ++i) l1 += fabs(rA[i]);
    const double init = l1 / nf;
    double fin = init;
#define OF_CONVERGED(r) ((r) < tol || (rel_tol > 1e-20 && (r) < rel_tol * init))
    int it = 0;
    if (!OF_CONVERGED(fin)) {
        if (precond == 1) dic_rd(m, diag, upper, rD);
        double wArA = 1e300, wArAold;
        do {
            wArAold = wArA;
            precondition(m, precond, diag, upper, rD, rA, wA);
            wArA = 0.0;
            for (long i = 0; i < n; ++i) wArA += wA[i] * rA[i];
            if (it == 0) {
                for (long i = 0; i < n; ++i) pA[i] = wA[i];
            } else {
                const double beta = wArA / wArAold;
                for (long i = 0; i < n; ++i) pA[i] = wA[i] + beta * pA[i];
            }
            orc_ldu_amul(m, diag, upper, pA, wA);
            double wApA = 0.0;
            for (long i = 0; i < n; ++i) wApA += wA[i] * pA[i];
            if (fabs(wApA) / nf < 1e-300) break;
            const double alpha = wArA / wApA;
            l1 = 0.0;
            for (long i = 0; i < n; ++i) {
                x[i] += alpha * pA[i];
                rA[i] -= alpha * wA[i];
                l1 += fabs(rA[i]);
            }
            fin = l1 / nf;
        } while (++it < max_iter && !OF_CONVERGED(fin));
    }
#undef OF_CONVERGED
    if (init_res) *init_res = init;
    if (final_res) *final_res = fin;
    free(wA); free(rA); free(pA); free(rD);
    return it;
}

void orc_box_stimulus(const orc_mesh_t* m, const orc_boxes_t* b, double t0, double* src)
{
    const int nx = m->nx, ny = m->ny;
    for (long c = 0; c < m->n_cells; ++c) src[c] = 0.0;
    if (!b) return;
    for (int s = 0; s < b->n_boxes; ++s) {
        const double tS = b->t_start[s], tE = tS + b->duration[s];
        if (t0 < tS || t0 > tE) continue;
        for (long c = 0; c < m->n_cells; ++c) {
            int i = (int)(c % nx), j = (int)((c / nx) % ny), k = (int)(c / ((long)nx * ny));
            double p[3] = {(i + 0.5) * m->hx, (j + 0.5) * m->hy, (k + 0.5) * m->hz};
            if (b->centres)
                for (int d = 0; d < 3; ++d) p[d] = b->centres[3 * c + d];
            int in = 1;
            for (int d = 0; d < 3; ++d)
                if (p[d] < b->box_min[3 * s + d] || p[d] > b->box_max[3 * s + d]) in = 0;
            if (in) src[c] += b->intensity[s];
        }
    }
}

void orc_activation(long n, const double* V, const double* Vold, uint8_t* flag, double* tact,
                    double t_old, double dt)
{
    const double SMALL = 1e-15;
    for (long c = 0; c < n; ++c) {
        if (flag[c] && V[c] > SMALL) {
            flag[c] = 0;
            double w = (0.0 - Vold[c]) / (V[c] - Vold[c]);
            tact[c] = t_old + w * dt;
        }
    }
}

double orc_explicit_dt(const orc_mesh_t* m, const double* sigma, double chiCm, double maxCo)
{
    double best = INFINITY;
    for (long f = 0; f < m->n_faces; ++f) {
        int d = m->face_dir[f];
        long P = m->owner[f], N = m->neighbour[f];
        double snn = 0.5 * (sigma[9 * P + 4 * d] + sigma[9 * N + 4 * d]);
        double Df = snn / chiCm;
        double h = cell_h(m, d);
        double q = h * h / Df;
        if (q < best) best = q;
    }
    return maxCo * best;
}

/* ----------------------------------------------------------- orchestration */
int orc_tissue_create(orc_tissue_t* T, int nx, int ny, int nz, double hx, double hy,
                      double hz, const orc_tissue_cfg_t* cfg, const double* sigma9)
{
    memset(T, 0, sizeof *T);
    if (orc_mesh_build(&T->mesh, nx, ny, nz, hx, hy, hz)) return -1;
    T->cfg = *cfg;
    const long n = T->mesh.n_cells;
    T->nstates = orc_model_nstates(cfg->model);
    double y[ORC_MAX_STATES];
    if (orc_model_init(cfg->model, cfg->tissue, T->consts, y)) return -2;
    T->states = malloc(sizeof(double) * T->nstates * n);
    for (int i = 0; i < T->nstates; ++i)
        for (long c = 0; c < n; ++c) T->states[(long)i * n + c] = y[i];
    T->step = malloc(sizeof(double) * n);
    for (long c = 0; c < n; ++c) T->step[c] = cfg->dt; /* seconds used as ms, ionicModel.C:38 */
    T->sigma = malloc(sizeof(double) * 9 * n);
    memcpy(T->sigma, sigma9, sizeof(double) * 9 * n);
    T->V = malloc(sizeof(double) * n); T->V0 = malloc(sizeof(double) * n);
    T->V00 = malloc(sizeof(double) * n); T->Iion = calloc(n, sizeof(double));
    T->src = calloc(n, sizeof(double)); T->tact = calloc(n, sizeof(double));
    T->flag = malloc(n); T->work = malloc(sizeof(double) * 4 * n);
    for (long c = 0; c < n; ++c) {
        T->V[c] = T->V0[c] = T->V00[c] = -0.084; /* myocardiumDomain.C:207 */
        T->flag[c] = 1;
    }
    return 0;
}

void orc_tissue_free(orc_tissue_t* T)
{
    orc_mesh_free(&T->mesh);
    free(T->states); free(T->step); free(T->sigma); free(T->V); free(T->V0); free(T->V00);
    free(T->Iion); free(T->src); free(T->tact); free(T->flag); free(T->work);
    memset(T, 0, sizeof *T);
}

int orc_tissue_step(orc_tissue_t* T, const orc_boxes_t* boxes)
{
    const orc_tissue_cfg_t* k = &T->cfg;
    const long n = T->mesh.n_cells;
    const double dt = k->dt, chiCm = k->chi * k->cm;
    T->t = T->t + dt;
    const double t0 = T->t - dt;
    /* BDF2 with Euler start-up: deltaT0 = GREAT while fewer than 2 old levels exist */
    double coefft = 1.0, coefft0 = 1.0, coefft00 = 0.0;
    if (k->ddt == 1 && T->step_index >= 1) { coefft = 1.5; coefft0 = 2.0; coefft00 = 0.5; }
    /* store old times: V00 <- V0 <- V */
    memcpy(T->V00, T->V0, sizeof(double) * n);
    memcpy(T->V0, T->V, sizeof(double) * n);

    orc_box_stimulus(&T->mesh, boxes, t0, T->src);
    orc_ionic_tissue_step(k->model, k->tissue, T->consts, 0, T->states, n, t0, dt, T->V,
                          T->Iion, k->integrator, k->n_sub, T->step, k->abs_tol, k->rel_tol,
                          k->nthreads);
    double* lap = T->work;
    if (!k->implicit) {
        orc_laplacian(&T->mesh, T->sigma, T->V0, lap, k->nthreads);
#pragma omp parallel for schedule(static) num_threads(k->nthreads > 0 ? k->nthreads : 1)
        for (long c = 0; c < n; ++c)
            T->V[c] = (coefft0 * T->V0[c] - coefft00 * T->V00[c] +
                       dt / chiCm * (lap[c] - chiCm * T->Iion[c] + T->src[c])) / coefft;
    } else {
        /* solveDiffusionImplicit inside `while (pimple.loop())` (monodomainSolver.C:139-150):
         * every outer corrector re-assembles fvm::laplacian, whose explicit tensor
         * correction comes from the latest V, and solves from the latest V.  The loop
         * restates OpenFOAM's pimpleLoop: after the first corrector, the initial residual
         * of the previous solve below residualControl makes the current corrector the
         * final one; nOuterCorrectors caps the count (n_outer = 1: one solve). */
        double* diag = T->work + n;
        double* rhs0 = T->work + 2 * n;
        double* rhs = T->work + 3 * n;
        double* upper = malloc(sizeof(double) * T->mesh.n_faces);
        orc_assemble_implicit(&T->mesh, T->sigma, chiCm, dt, coefft, diag, upper);
        for (long c = 0; c < n; ++c)
            rhs0[c] = chiCm / dt * (coefft0 * T->V0[c] - coefft00 * T->V00[c]) -
                      chiCm * T->Iion[c] + T->src[c];
        const int n_outer = k->n_outer > 0 ? k->n_outer : 1;
        double prev_init = 0.0;
        for (int outer = 1;; ++outer) {
            const int final_corr = n_outer == 1 || outer >= n_outer ||
                                   (outer > 1 && prev_init < k->residual_control);
            orc_correction_div(&T->mesh, T->sigma, T->V, lap, k->nthreads);
            for (long c = 0; c < n; ++c) rhs[c] = rhs0[c] + lap[c];
            double init, res;
            int it;
            if (k->norm == 1) {
                it = orc_pcg_of(&T->mesh, diag, upper, rhs, T->V, k->pcg_tol, k->pcg_rel_tol,
                                k->pcg_max_iter, k->precond, &init, &res);
            } else {
                /* relative L2 initial residual, then the plain PCG */
                orc_ldu_amul(&T->mesh, diag, upper, T->V, lap);
                double rr = 0.0, bb = 0.0;
                for (long c = 0; c < n; ++c) {
                    rr += (rhs[c] - lap[c]) * (rhs[c] - lap[c]);
                    bb += rhs[c] * rhs[c];
                }
                init = sqrt(rr / (bb > 0.0 ? bb : 1.0));
                it = orc_pcg(&T->mesh, diag, upper, rhs, T->V, k->pcg_tol, k->pcg_max_iter, &res);
            }
            T->pcg_iters_total += it;
            T->outer_total++;
            T->last_init_res = init;
            prev_init = init;
            if (final_corr) break;
        }
        free(upper);
    }
    orc_activation(n, T->V, T->V0, T->flag, T->tact, T->t - dt, dt);
    T->step_index++;
    return 0;
}

/* Pseudo-ECG (PseudoECGSolver::compute, pseudoECGSolver.C:36-83): for each electrode P,
 *   phi = sum_c ((sigma_c . grad V_c) vol_c) . r / |r|^3,  r = C_c - P,  |r| > VSMALL,
 * grad V the Gauss-linear gradient with zeroGradient walls.  out[e] the sum, absout[e]
 * the sum of the magnitudes of its terms (the scale of its rounding error). */
void orc_pseudo_ecg(const orc_mesh_t* m, const double* sigma, const double* V, int ne,
                    const double* pos, double* out, double* absout, int nthreads)
{
    const long n = m->n_cells;
    const int nx = m->nx, ny = m->ny;
    const double vol = m->hx * m->hy * m->hz;
    double* grad = malloc(sizeof(double) * 3 * n);
    gauss_grad(m, V, grad, nthreads);
    for (int e = 0; e < ne; ++e) {
        double acc = 0.0, aacc = 0.0;
        const double P0 = pos[3 * e], P1 = pos[3 * e + 1], P2 = pos[3 * e + 2];
#pragma omp parallel for schedule(static) reduction(+ : acc, aacc) \
    num_threads(nthreads > 0 ? nthreads : 1)
        for (long c = 0; c < n; ++c) {
            const int i = (int)(c % nx), j = (int)((c / nx) % ny), k = (int)(c / ((long)nx * ny));
            const double* s = sigma + 9 * c;
            const double* g = grad + 3 * c;
            double d[3];
            for (int a = 0; a < 3; ++a) d[a] = (s[3 * a] * g[0] + s[3 * a + 1] * g[1] + s[3 * a + 2] * g[2]) * vol;
            const double r0 = (i + 0.5) * m->hx - P0, r1 = (j + 0.5) * m->hy - P1,
                         r2 = (k + 0.5) * m->hz - P2;
            const double r = sqrt(r0 * r0 + r1 * r1 + r2 * r2);
            if (r > 1e-300) {
                const double term = (d[0] * r0 + d[1] * r1 + d[2] * r2) / (r * r * r);
                acc += term;
                aacc += fabs(term);
            }
        }
        out[e] = acc;
        if (absout) absout[e] = aacc;
    }
    free(grad);
}
