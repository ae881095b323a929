/*
 * mn_oracle.c -- CPU FP64 restatement of the closures-2d Mn ("momopt") path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): loaded by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 *
 * Restates, in the reference's evaluation order:
 *   src/utils.cpp:36-60                 norm1, Gauss-Legendre rule
 *   src/momopt/momopt_init.cpp:51-272   quadrature, real SH table, Gaunt map, initial u/alpha
 *   src/momopt/opt/fobj.cpp:37-129,255-335   solveHMn / fobjMn
 *   src/momopt/opt/opt.cpp:27-46,82-343      isotropicAlpha / opt / optScaled
 *   src/momopt/opt/linesearch.cpp:21-99      Armijo backtracking
 *   src/momopt/momopt_update.cpp:38-382      computeAnsatz, minmod, slopefit, solveFlux,
 *                                            solveOptimization, update (Heun)
 *   src/momopt/momopt_boundaries.cpp:190-223 duplicateBoundaries (serial periodic ghosts)
 *   src/main.cpp:204-230,257-372             time loop, initializeGrid
 * LAPACK/BLAS roles (dpotrf, dpocon, dpotrs, dnrm2; utils.h:30-38) are written
 * here from the published algorithms: unblocked column Cholesky, Higham's
 * 1-norm estimator (ACM TOMS 14(4) 381-396, 1988), forward/back substitution.
 */
#include "mn_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_NUM_REG 5

/* ORC_EXP_LONG builds a sensitivity probe of the oracle: every exp() of the
 * closure and the ansatz is evaluated in long double and rounded once.  It is
 * not a restatement of the reference; tests use it to measure how far the
 * reference algorithm's own output moves under a last-bit change of exp. */
#ifdef ORC_EXP_LONG
#define ORC_EXP(x) ((double)expl((long double)(x)))
#else
#define ORC_EXP(x) exp(x)
#endif
static const double ORC_REGULARIZATIONS[ORC_NUM_REG] = {0.0, 1e-3, 1e-2, 1e-1, 5e-1}; /* opt.h:39-40 */

/* ------------------------------------------------------------------ */
/* Gauss-Legendre (utils.cpp:55-60: GSL glfixed, ascending order).      */
/* Newton iteration on P_n from the Tricomi initial guess.              */
/* ------------------------------------------------------------------ */
int orc_gauss_legendre(int n, double *x, double *w)
{
    if (n < 1)
        return ORC_ERR_ARG;
    int half = (n + 1) / 2;
    for (int i = 0; i < half; i++) {
        double z = cos(M_PI * (i + 0.75) / (n + 0.5));
        double pp = 0.0;
        for (int it = 0; it < 100; it++) {
            double p1 = 1.0, p2 = 0.0;
            for (int j = 1; j <= n; j++) {
                double p3 = p2;
                p2 = p1;
                p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
            }
            pp = n * (z * p1 - p2) / (z * z - 1.0);
            double dz = p1 / pp;
            z -= dz;
            if (fabs(dz) <= 1e-17 * (1.0 + fabs(z)) && it > 2)
                break;
        }
        /* final derivative at the converged node */
        {
            double p1 = 1.0, p2 = 0.0;
            for (int j = 1; j <= n; j++) {
                double p3 = p2;
                p2 = p1;
                p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
            }
            pp = n * (z * p1 - p2) / (z * z - 1.0);
        }
        double wt = 2.0 / ((1.0 - z * z) * pp * pp);
        /* z is the i-th largest root; ascending order puts -z first */
        x[i] = -z;
        x[n - 1 - i] = z;
        w[i] = wt;
        w[n - 1 - i] = wt;
    }
    if (n % 2 == 1)
        x[n / 2] = 0.0;
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* Normalised associated Legendre functions with Condon-Shortley phase. */
/* Pbar_m^m = (-1)^m sqrt((2m+1)/4pi) prod_{k<=m} sqrt((2k-1)/(2k)) s^m  */
/* Pbar_{m+1}^m = x sqrt(2m+3) Pbar_m^m                                  */
/* Pbar_l^m = a (x Pbar_{l-1}^m - b Pbar_{l-2}^m),                        */
/*   a = sqrt((4l^2-1)/(l^2-m^2)), b = sqrt(((l-1)^2-m^2)/(4(l-1)^2-1))   */
/* ------------------------------------------------------------------ */
double orc_sph_plm(int l, int m, double x)
{
    if (m < 0 || m > l)
        return 0.0;
    double s = sqrt((1.0 - x) * (1.0 + x));
    double pmm = sqrt((2.0 * m + 1.0) / (4.0 * M_PI));
    for (int k = 1; k <= m; k++)
        pmm *= -sqrt((2.0 * k - 1.0) / (2.0 * k)) * s;
    if (l == m)
        return pmm;
    double pm1 = x * sqrt(2.0 * m + 3.0) * pmm;
    if (l == m + 1)
        return pm1;
    double p0 = pmm, p1 = pm1, pl = 0.0;
    for (int ll = m + 2; ll <= l; ll++) {
        double a = sqrt((4.0 * ll * ll - 1.0) / ((double)ll * ll - (double)m * m));
        double b = sqrt((((double)ll - 1.0) * (ll - 1.0) - (double)m * m) /
                        (4.0 * (ll - 1.0) * (ll - 1.0) - 1.0));
        pl = a * (x * p1 - b * p0);
        p0 = p1;
        p1 = pl;
    }
    return pl;
}

/* ------------------------------------------------------------------ */
/* Tables (momopt_init.cpp:88-93,164-259).                              */
/* ------------------------------------------------------------------ */
struct orc_tables {
    int N, Q, nm, nmm, np2, nq, use_gaunt;
    double *w, *p, *xi, *eta;
    int *gptr, *gk;
    double *gv;
    int gnnz;
};

int orc_tables_nm(const orc_tables *t) { return t->nm; }
int orc_tables_np2(const orc_tables *t) { return t->np2; }
int orc_tables_nq(const orc_tables *t) { return t->nq; }
int orc_tables_gaunt_nnz(const orc_tables *t) { return t->gnnz; }

void orc_tables_get(const orc_tables *t, double *w, double *p, double *xi, double *eta)
{
    if (w) memcpy(w, t->w, sizeof(double) * t->nq);
    if (p) memcpy(p, t->p, sizeof(double) * t->nq * t->np2);
    if (xi) memcpy(xi, t->xi, sizeof(double) * t->nq);
    if (eta) memcpy(eta, t->eta, sizeof(double) * t->nq);
}

void orc_tables_gaunt_get(const orc_tables *t, int *ptr, int *k, double *v)
{
    if (!t->use_gaunt)
        return;
    if (ptr) memcpy(ptr, t->gptr, sizeof(int) * (t->nm * t->nm + 1));
    if (k) memcpy(k, t->gk, sizeof(int) * t->gnnz);
    if (v) memcpy(v, t->gv, sizeof(double) * t->gnnz);
}

void orc_tables_destroy(orc_tables *t)
{
    if (!t) return;
    free(t->w); free(t->p); free(t->xi); free(t->eta);
    free(t->gptr); free(t->gk); free(t->gv);
    free(t);
}

/* Householder QR of A (m x n, row-major, m >= n), in place: v-vectors on and
 * below the diagonal, R above it, beta[0..n) reflector scalars and
 * beta[n..2n) the diagonal of R.  Least squares for the Gaunt
 * coefficients follows momopt_init.cpp:227-258 (QR_decomp + QR_lssolve). */
static void qr_house(int m, int n, double *A, double *beta)
{
    for (int k = 0; k < n; k++) {
        double norm = 0.0;
        for (int i = k; i < m; i++)
            norm += A[i * n + k] * A[i * n + k];
        norm = sqrt(norm);
        double akk = A[k * n + k];
        double alpha = (akk > 0) ? -norm : norm;
        if (norm == 0.0) { beta[k] = 0.0; continue; }
        /* v = x - alpha e1, stored with v_k in A[k][k] */
        double vk = akk - alpha;
        A[k * n + k] = vk;
        double vtv = vk * vk;
        for (int i = k + 1; i < m; i++)
            vtv += A[i * n + k] * A[i * n + k];
        beta[k] = (vtv > 0) ? 2.0 / vtv : 0.0;
        for (int j = k + 1; j < n; j++) {
            double s = 0.0;
            for (int i = k; i < m; i++)
                s += A[i * n + k] * A[i * n + j];
            s *= beta[k];
            for (int i = k; i < m; i++)
                A[i * n + j] -= s * A[i * n + k];
        }
        /* A[k][k] keeps v_k; R_kk = alpha is kept in beta[n + k] */
        beta[n + k] = alpha;
    }
}

static void qr_lssolve(int m, int n, const double *A, const double *beta, const double *b,
                       double *x, double *work)
{
    memcpy(work, b, sizeof(double) * m);
    for (int k = 0; k < n; k++) {
        if (beta[k] == 0.0) continue;
        double s = 0.0;
        for (int i = k; i < m; i++)
            s += A[i * n + k] * work[i];
        s *= beta[k];
        for (int i = k; i < m; i++)
            work[i] -= s * A[i * n + k];
    }
    for (int i = n - 1; i >= 0; i--) {
        double s = work[i];
        for (int j = i + 1; j < n; j++)
            s -= A[i * n + j] * x[j];
        x[i] = s / beta[n + i];
    }
}

int orc_tables_create(int moment_order, int quad_order, int use_gaunt, orc_tables **out)
{
    *out = NULL;
    if (moment_order < 0 || quad_order < 2 || (quad_order % 2) != 0)
        return ORC_ERR_ARG;
    orc_tables *t = (orc_tables *)calloc(1, sizeof(orc_tables));
    if (!t) return ORC_ERR_ALLOC;
    int N = moment_order, Q = quad_order;
    t->N = N; t->Q = Q; t->use_gaunt = use_gaunt ? 1 : 0;
    t->nq = Q * Q;                                   /* momopt_init.cpp:88 */
    t->nm = (N + 1) * (N + 2) / 2;                   /* :89 */
    t->nmm = t->use_gaunt ? (2 * N + 1) * (2 * N + 2) / 2 : t->nm;  /* :90-93 */
    t->np2 = t->nmm;
    int nq = t->nq, np2 = t->np2;
    t->w = (double *)malloc(sizeof(double) * nq);
    t->p = (double *)malloc(sizeof(double) * nq * np2);
    t->xi = (double *)malloc(sizeof(double) * nq);
    t->eta = (double *)malloc(sizeof(double) * nq);
    double *glx = (double *)malloc(sizeof(double) * Q);
    double *glw = (double *)malloc(sizeof(double) * Q);
    if (!t->w || !t->p || !t->xi || !t->eta || !glx || !glw) {
        free(glx); free(glw); orc_tables_destroy(t); return ORC_ERR_ALLOC;
    }
    orc_gauss_legendre(Q, glx, glw);                 /* :164 */
    int Nsh = t->use_gaunt ? 2 * N + 1 : N + 1;      /* :193-195 */
    for (int q1 = 0; q1 < Q / 2; q1++) {             /* :176 */
        double wTemp = 2.0 * M_PI / Q * glw[q1];     /* :178 */
        for (int q2 = 0; q2 < 2 * Q; q2++) {
            double muTemp = glx[q1];
            double phiTemp = (q2 + 0.5) * M_PI / Q;  /* :170 */
            double xiTemp = sqrt(1.0 - muTemp * muTemp) * cos(phiTemp);
            double etaTemp = sqrt(1.0 - muTemp * muTemp) * sin(phiTemp);
            int q = 2 * q1 * Q + q2;                 /* :187 */
            int k = 0;
            t->w[q] = wTemp;
            t->xi[q] = xiTemp;
            t->eta[q] = etaTemp;
            for (int n = 0; n < Nsh; n++) {
                for (int m = -n; m <= n; m++) {
                    if ((m + n) % 2 != 0) continue;  /* :200 */
                    double sh;
                    if (m < 0)
                        sh = M_SQRT2 * sin(m * phiTemp) * orc_sph_plm(n, -m, muTemp);
                    else if (m > 0)
                        sh = M_SQRT2 * cos(m * phiTemp) * orc_sph_plm(n, m, muTemp);
                    else
                        sh = orc_sph_plm(n, 0, muTemp);
                    t->p[q * np2 + k] = sh;
                    k++;
                }
            }
        }
    }
    free(glx); free(glw);

    if (t->use_gaunt) {                              /* :221-259 */
        int nm = t->nm, nmm = t->nmm;
        double *A = (double *)malloc(sizeof(double) * nq * nmm);
        double *beta = (double *)malloc(sizeof(double) * 2 * nmm);
        double *b = (double *)malloc(sizeof(double) * nq);
        double *work = (double *)malloc(sizeof(double) * nq);
        double *x = (double *)malloc(sizeof(double) * nmm);
        int cap = nm * nm * nmm;
        int *cnt = (int *)calloc(nm * nm, sizeof(int));
        int *tk = (int *)malloc(sizeof(int) * cap);
        double *tv = (double *)malloc(sizeof(double) * cap);
        if (!A || !beta || !b || !work || !x || !cnt || !tk || !tv) {
            free(A); free(beta); free(b); free(work); free(x); free(cnt); free(tk); free(tv);
            orc_tables_destroy(t); return ORC_ERR_ALLOC;
        }
        memcpy(A, t->p, sizeof(double) * nq * nmm);
        qr_house(nq, nmm, A, beta);
        for (int i = 0; i < nm; i++) {
            for (int j = i; j < nm; j++) {
                for (int q = 0; q < nq; q++)
                    b[q] = t->p[q * nmm + i] * t->p[q * nmm + j];
                qr_lssolve(nq, nmm, A, beta, b, x, work);
                int c = 0;
                for (int k = 0; k < nmm; k++) {
                    if (fabs(x[k]) > 10e-10) {        /* :249 */
                        tk[(i * nm + j) * nmm + c] = k;
                        tv[(i * nm + j) * nmm + c] = x[k];
                        c++;
                    }
                }
                cnt[i * nm + j] = c;
                if (j != i) {   /* (j,i) has the identical right-hand side */
                    memcpy(&tk[(j * nm + i) * nmm], &tk[(i * nm + j) * nmm], sizeof(int) * c);
                    memcpy(&tv[(j * nm + i) * nmm], &tv[(i * nm + j) * nmm], sizeof(double) * c);
                    cnt[j * nm + i] = c;
                }
            }
        }
        int nnz = 0;
        for (int pidx = 0; pidx < nm * nm; pidx++) nnz += cnt[pidx];
        t->gnnz = nnz;
        t->gptr = (int *)malloc(sizeof(int) * (nm * nm + 1));
        t->gk = (int *)malloc(sizeof(int) * (nnz > 0 ? nnz : 1));
        t->gv = (double *)malloc(sizeof(double) * (nnz > 0 ? nnz : 1));
        int off = 0;
        for (int pidx = 0; pidx < nm * nm; pidx++) {
            t->gptr[pidx] = off;
            for (int c = 0; c < cnt[pidx]; c++) {
                t->gk[off] = tk[pidx * nmm + c];
                t->gv[off] = tv[pidx * nmm + c];
                off++;
            }
        }
        t->gptr[nm * nm] = off;
        free(A); free(beta); free(b); free(work); free(x); free(cnt); free(tk); free(tv);
    }
    *out = t;
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* fobjMn / solveHMn (fobj.cpp:37-129,255-335).                          */
/* ------------------------------------------------------------------ */
static void solve_h(const orc_tables *t, const double *alpha, double *h, int use_gaunt,
                    double *gext)
{
    int nm = t->nm, np2 = t->np2, nq = t->nq;
    const double *w = t->w, *p = t->p;
    for (int k = 0; k < np2; k++) gext[k] = 0.0;          /* :56-57 */
    for (int i = 0; i < nm * nm; i++) h[i] = 0.0;          /* :59-60 */
    for (int q = 0; q < nq; q++) {                         /* :64 */
        double temp = 0;
        for (int i = 0; i < nm; i++)
            temp = temp + alpha[i] * p[q * np2 + i];
        temp = ORC_EXP(temp);
        if (use_gaunt) {
            for (int k = 0; k < np2; k++)
                gext[k] += w[q] * p[q * np2 + k] * temp;   /* :80 */
        } else {
            for (int i = 0; i < nm; i++) {
                double pi = p[q * np2 + i];
                for (int j = 0; j <= i; j++) {
                    double pj = p[q * np2 + j];
                    h[i * nm + j] += w[q] * pi * pj * temp; /* :91 */
                }
            }
        }
    }
    if (use_gaunt) {                                       /* :99-113 */
        for (int i = 0; i < nm; i++) {
            for (int j = i; j < nm; j++) {
                double hij = 0.0;
                int idx = i * nm + j;
                for (int c = t->gptr[idx]; c < t->gptr[idx + 1]; c++)
                    hij += t->gv[c] * gext[t->gk[c]];
                h[i * nm + j] = h[j * nm + i] = hij;
            }
        }
    } else {                                               /* :117-123 */
        for (int i = 0; i < nm; i++)
            for (int j = i + 1; j < nm; j++)
                h[i * nm + j] = h[j * nm + i];
    }
}

double orc_fobj(const orc_tables *t, const double *alpha, const double *u,
                double *g, double *h, int use_gaunt)
{
    int nm = t->nm, np2 = t->np2, nq = t->nq;
    const double *w = t->w, *p = t->p;
    double f = 0;
    if (h != NULL) {                                       /* :266-287 */
        double gext_stack[256];
        double *gext = (np2 <= 256) ? gext_stack : (double *)malloc(sizeof(double) * np2);
        solve_h(t, alpha, h, use_gaunt, gext);
        if (gext != gext_stack) free(gext);
        if (g != NULL)
            for (int i = 0; i < nm; i++)
                g[i] = h[i] / p[0] - u[i];
        f = h[0] / (p[0] * p[0]);
        for (int i = 0; i < nm; i++)
            f = f - alpha[i] * u[i];
        return f;
    }
    if (g != NULL)                                         /* :291-295 */
        for (int i = 0; i < nm; i++) g[i] = 0.0;
    for (int q = 0; q < nq; q++) {                         /* :297-317 */
        double temp = 0;
        for (int i = 0; i < nm; i++)
            temp = temp + alpha[i] * p[q * np2 + i];
        temp = ORC_EXP(temp);
        if (g != NULL)
            for (int i = 0; i < nm; i++)
                g[i] += w[q] * p[q * np2 + i] * temp;
        f = f + w[q] * temp;
    }
    if (g != NULL)
        for (int i = 0; i < nm; i++) g[i] = g[i] - u[i];
    for (int i = 0; i < nm; i++)
        f = f - alpha[i] * u[i];
    return f;
}

/* ------------------------------------------------------------------ */
/* fobjPPn / solveHPPn (fobj.cpp:137-236,355-428).                       */
/* ------------------------------------------------------------------ */
/* The PPn ansatz at t = (alpha.m)/2: G = t + sqrt(t^2 + delta), taken as
 * -delta / (t - sqrt(t^2 + delta)) for t < 0 (no cancellation); s is the
 * square root. */
static double ppn_ansatz(double t, double delta, double *s)
{
    *s = sqrt(t * t + delta);
    double G = t + *s;
    if (t < 0) G = -delta / (t - sqrt(t * t + delta));
    return G;
}

double orc_fobj_ppn(const orc_tables *t, const double *alpha, const double *u, double delta,
                    double *g, double *h, int use_gaunt)
{
    int nm = t->nm, np2 = t->np2, nq = t->nq;
    const double *w = t->w, *p = t->p;
    double f = 0;
    if (g != NULL)
        for (int i = 0; i < nm; i++) g[i] = 0.0;
    for (int q = 0; q < nq; q++) {                         /* :370-396 */
        double tq = 0;
        for (int i = 0; i < nm; i++)
            tq = tq + alpha[i] * p[q * np2 + i];
        tq = 0.5 * tq;
        double sq;
        double G = ppn_ansatz(tq, delta, &sq);
        f = f + w[q] * (tq * G - 0.5 * delta + delta * log(G));
        if (g != NULL) {
            double wg = w[q] * G;
            for (int i = 0; i < nm; i++)
                g[i] += p[q * np2 + i] * wg;
        }
    }
    if (h != NULL) {                                       /* solveHPPn, :137-236 */
        double gext_stack[256];
        double *gext = (np2 <= 256) ? gext_stack : (double *)malloc(sizeof(double) * np2);
        for (int k = 0; k < np2; k++) gext[k] = 0.0;
        for (int i = 0; i < nm * nm; i++) h[i] = 0.0;
        for (int q = 0; q < nq; q++) {
            double tq = 0;
            for (int i = 0; i < nm; i++)
                tq = tq + alpha[i] * p[q * np2 + i];
            tq = 0.5 * tq;
            double sq;
            double G = ppn_ansatz(tq, delta, &sq);
            double wh = 0.5 * w[q] * (G / sq);
            if (use_gaunt) {
                for (int k = 0; k < np2; k++)
                    gext[k] += p[q * np2 + k] * wh;
            } else {
                for (int i = 0; i < nm; i++) {
                    double pi = p[q * np2 + i];
                    for (int j = 0; j <= i; j++)
                        h[i * nm + j] += pi * p[q * np2 + j] * wh;
                }
            }
        }
        if (use_gaunt) {
            for (int i = 0; i < nm; i++)
                for (int j = i; j < nm; j++) {
                    double hij = 0.0;
                    int idx = i * nm + j;
                    for (int c = t->gptr[idx]; c < t->gptr[idx + 1]; c++)
                        hij += t->gv[c] * gext[t->gk[c]];
                    h[i * nm + j] = h[j * nm + i] = hij;
                }
        } else {
            for (int i = 0; i < nm; i++)
                for (int j = i + 1; j < nm; j++)
                    h[i * nm + j] = h[j * nm + i];
        }
        if (gext != gext_stack) free(gext);
    }
    for (int i = 0; i < nm; i++) {                         /* :414-421 */
        f = f - alpha[i] * u[i];
        if (g != NULL) g[i] = g[i] - u[i];
    }
    return f;
}

/* f (and g, H) of either moment type */
static double fobj_any(const orc_tables *t, const double *alpha, const double *u, double *g, double *h,
                       const orc_options *o)
{
    if (o->moment_type == ORC_MOMENT_PPN)
        return orc_fobj_ppn(t, alpha, u, o->delta_ppn, g, h, o->use_gaunt);
    return orc_fobj(t, alpha, u, g, h, o->use_gaunt);
}

/* ------------------------------------------------------------------ */
/* Dense kernels in the roles of dpotrf / dpocon / dpotrs / dnrm2.       */
/* Matrices are row-major n x n; the factor L is stored lower (a[i*n+j], */
/* i >= j).  H is exactly symmetric, so the row/column-major view of the */
/* reference's Fortran call (opt.cpp:231) does not matter.               */
/* ------------------------------------------------------------------ */
int orc_cholesky(int n, double *a)
{
    for (int j = 0; j < n; j++) {
        double dot = 0.0;
        for (int k = 0; k < j; k++)
            dot += a[j * n + k] * a[j * n + k];
        double ajj = a[j * n + j] - dot;
        if (ajj <= 0.0 || isnan(ajj)) {
            a[j * n + j] = ajj;
            return j + 1;
        }
        ajj = sqrt(ajj);
        a[j * n + j] = ajj;
        for (int i = j + 1; i < n; i++) {
            double y = a[i * n + j];
            for (int k = 0; k < j; k++)
                y += (-a[j * n + k]) * a[i * n + k];
            a[i * n + j] = y;
        }
        double r = 1.0 / ajj;
        for (int i = j + 1; i < n; i++)
            a[i * n + j] *= r;
    }
    return 0;
}

/* x <- L^{-1} x */
static void trsv_lower_n(int n, const double *l, double *x)
{
    for (int j = 0; j < n; j++) {
        if (x[j] != 0.0) {
            x[j] = x[j] / l[j * n + j];
            double temp = x[j];
            for (int i = j + 1; i < n; i++)
                x[i] = x[i] - temp * l[i * n + j];
        }
    }
}

/* x <- L^{-T} x */
static void trsv_lower_t(int n, const double *l, double *x)
{
    for (int j = n - 1; j >= 0; j--) {
        double temp = x[j];
        for (int i = n - 1; i > j; i--)
            temp = temp - l[i * n + j] * x[i];
        temp = temp / l[j * n + j];
        x[j] = temp;
    }
}

double orc_norm1(int n, const double *a)
{
    double one = 0.0;                                      /* utils.cpp:36-52 */
    for (int i = 0; i < n; i++) {
        double s = 0;
        for (int j = 0; j < n; j++)
            s += fabs(a[i * n + j]);
        if (s > one) one = s;
    }
    return one;
}

static double dasum(int n, const double *x)
{
    double s = 0.0;
    for (int i = 0; i < n; i++) s += fabs(x[i]);
    return s;
}

static int idamax(int n, const double *x)
{
    int idx = 0;
    double m = fabs(x[0]);
    for (int i = 1; i < n; i++)
        if (fabs(x[i]) > m) { m = fabs(x[i]); idx = i; }
    return idx;
}

static void apply_inv(int n, const double *l, double *x)
{
    trsv_lower_n(n, l, x);
    trsv_lower_t(n, l, x);
}

/* Higham's estimate of ||A^{-1}||_1 for A = L L^T (A symmetric, so the
 * transpose products coincide with the plain ones). */
static double est_inv_norm1(int n, const double *l, double *x, double *v, int *isgn)
{
    const int itmax = 5;
    double est;
    for (int i = 0; i < n; i++) x[i] = 1.0 / n;
    apply_inv(n, l, x);
    if (n == 1) {
        v[0] = x[0];
        return fabs(v[0]);
    }
    est = dasum(n, x);
    for (int i = 0; i < n; i++) {
        x[i] = (x[i] >= 0.0) ? 1.0 : -1.0;
        isgn[i] = (int)x[i];
    }
    apply_inv(n, l, x);
    int j = idamax(n, x);
    int iter = 2;
    for (;;) {
        for (int i = 0; i < n; i++) x[i] = 0.0;
        x[j] = 1.0;
        apply_inv(n, l, x);
        memcpy(v, x, sizeof(double) * n);
        double estold = est;
        est = dasum(n, v);
        int repeated = 1;
        for (int i = 0; i < n; i++) {
            int xs = (x[i] >= 0.0) ? 1 : -1;
            if (xs != isgn[i]) { repeated = 0; break; }
        }
        if (repeated) break;
        if (est <= estold) break;
        for (int i = 0; i < n; i++) {
            x[i] = (x[i] >= 0.0) ? 1.0 : -1.0;
            isgn[i] = (int)x[i];
        }
        apply_inv(n, l, x);
        int jlast = j;
        j = idamax(n, x);
        if (x[jlast] != fabs(x[j]) && iter < itmax) {
            iter++;
            continue;
        }
        break;
    }
    double altsgn = 1.0;
    for (int i = 0; i < n; i++) {
        x[i] = altsgn * (1.0 + (double)i / (double)(n - 1));
        altsgn = -altsgn;
    }
    apply_inv(n, l, x);
    double temp = 2.0 * (dasum(n, x) / (double)(3 * n));
    if (temp > est) {
        memcpy(v, x, sizeof(double) * n);
        est = temp;
    }
    return est;
}

double orc_rcond(int n, const double *l, double anorm)
{
    if (n == 0) return 1.0;
    if (anorm == 0.0) return 0.0;
    double xs[256], vs[256];
    int is[256];
    double *x = xs, *v = vs;
    int *isgn = is;
    if (n > 256) {
        x = (double *)malloc(sizeof(double) * n);
        v = (double *)malloc(sizeof(double) * n);
        isgn = (int *)malloc(sizeof(int) * n);
    }
    double ainvnm = est_inv_norm1(n, l, x, v, isgn);
    if (x != xs) { free(x); free(v); free(isgn); }
    if (!isfinite(ainvnm)) return 0.0;
    if (ainvnm == 0.0) return 0.0;
    return (1.0 / ainvnm) / anorm;
}

static double nrm2(int n, const double *x)
{
    double s = 0.0;
    for (int i = 0; i < n; i++) s += x[i] * x[i];
    return sqrt(s);
}

/* ------------------------------------------------------------------ */
/* isotropicAlpha / linesearch / optScaled / opt                        */
/* ------------------------------------------------------------------ */
static void isotropic_alpha(int n, double *alpha, const orc_options *o)   /* opt.cpp:27-46 */
{
    if (o->moment_type == ORC_MOMENT_PPN)
        alpha[0] = 1.0 - 4.0 * o->delta_ppn * M_PI;
    else
        alpha[0] = 2.0 * sqrt(M_PI) * log(0.5 / sqrt(M_PI));
    for (int i = 1; i < n; i++) alpha[i] = 0.0;
}

/* The supplementary stop test's error bound (opt.cpp:52-76):
 * gamma = exp(2 |d|_1 m_inf) for Mn, 1 + 2 |d|_1 m_inf for PPn, with
 * m_inf = sqrt((2 P + 1) / (2 pi)) and P = floor(sqrt(n)) + 1. */
static double modcfl(int n, const double *da, int moment_type)
{
    double da_1 = 0;
    for (int i = 0; i < n; i++)
        da_1 = da_1 + fabs(da[i]);
    int pnOrder = (int)floor(sqrt((double)n)) + 1;
    double m_inf = sqrt((2 * pnOrder + 1) / (2 * M_PI));
    if (moment_type == ORC_MOMENT_PPN)
        return 1.0 + 2.0 * da_1 * m_inf;
    return exp(2.0 * da_1 * m_inf);
}

static double linesearch(const orc_tables *t, double *alpha, const double *d, double f,
                         const double *g, const double *u, double *alpha_old, const orc_options *o,
                         double fscale)
{                                                          /* linesearch.cpp:21-99 */
    const int robust = o->line_search;
    int n1 = t->nm;
    double tt = 1.0;
    for (int i = 0; i < n1; i++) alpha_old[i] = alpha[i];
    const double delta = 1e-3;
    const double beta = .5;
    double maxFactor = 0;
    for (int i = 0; i < n1; i++) {
        double temp = fabs(d[i]) / (fabs(alpha_old[i]) + 1e-100);
        if (temp > maxFactor) maxFactor = temp;
    }
    while (tt * maxFactor > DBL_EPSILON / 2) {
        for (int i = 0; i < n1; i++)
            alpha[i] = alpha_old[i] + tt * d[i];
        double fNew = (o->moment_type == ORC_MOMENT_PPN) ? orc_fobj_ppn(t, alpha, u, o->delta_ppn, NULL, NULL, 0)
                                                          : orc_fobj(t, alpha, u, NULL, NULL, 0);
        double temp = 0.0;
        for (int i = 0; i < n1; i++)
            temp += g[i] * d[i];
        if (fNew <= f + delta * tt * temp)
            return tt;
        if (robust && tt == 1.0 && -temp <= 1e3 * DBL_EPSILON * fscale)
            return tt;                 /* extension: decrease below f's resolution */
        tt = beta * tt;
    }
    for (int i = 0; i < n1; i++) alpha[i] = alpha_old[i];
    return 0.0;
}

static int opt_scaled(const orc_tables *t, double *u, double *alpha, const orc_options *o,
                      orc_outs *outs)
{                                                          /* opt.cpp:154-343 */
    int nm = t->nm;
    double dstk[6 * 64 + 2 * 64 * 64];
    double *buf = (nm <= 64) ? dstk : (double *)malloc(sizeof(double) * (6 * nm + 2 * nm * nm));
    double *d = buf, *g = d + nm, *uOriginal = g + nm, *alpha0 = uOriginal + nm;
    double *alphaOld = alpha0 + nm, *spare = alphaOld + nm;
    double *chol = spare + nm, *h = chol + nm * nm;
    (void)spare;
    double f = 0;
    int maxIter = o->max_iter;
    double tolRel = o->tol_rel;
    double r = 0.0;
    double condHMax = o->cond_h_max;
    int regIndex = 0;
    int status = ORC_OK;
    for (int i = 0; i < nm; i++) uOriginal[i] = u[i];
    double rtol = nrm2(nm, u) * tolRel;
    int iter = 0, totalIter = 0;
    double tt = 1;
    double err = 0.0;
    for (int i = 0; i < nm; i++) alpha0[i] = alpha[i];
    for (;;) {
        iter = iter + 1;
        totalIter = totalIter + 1;
        f = fobj_any(t, alpha, u, g, h, o);
        /* the rounding scale of f: |its integral term| + sum |alpha_i u_i| */
        double fscale = 0.0, au = 0.0;
        for (int i = 0; i < nm; i++) {
            fscale += fabs(alpha[i] * u[i]);
            au += alpha[i] * u[i];
        }
        if (o->moment_type == ORC_MOMENT_PPN) fscale += fabs(f + au);
        else fscale += fabs(h[0] / (t->p[0] * t->p[0]));
        memcpy(chol, h, sizeof(double) * nm * nm);
        int error = orc_cholesky(nm, chol);
        double norm = orc_norm1(nm, h);
        double condest = 0.0;
        if (error == 0) {
            double rc = orc_rcond(nm, chol, norm);
            condest = 1.0 / rc;
        }
        int why = 0;
        if (o->trace) why = iter >= maxIter ? 1 : error != 0 ? 2 : tt == 0 ? 3 : condest > condHMax ? 4 : 0;
        if (iter >= maxIter || error != 0 || tt == 0 || condest > condHMax) {
            if (o->trace && *o->trace_len < o->trace_cap) {
                double *rec = o->trace + (size_t)(*o->trace_len)++ * ORC_TRACE_WORDS;
                rec[0] = totalIter; rec[1] = regIndex; rec[2] = why; rec[3] = condest;
                rec[4] = nrm2(nm, g); rec[5] = rtol; rec[6] = 0.0;
            }
            if (regIndex < ORC_NUM_REG - 1) {
                regIndex++;
                r = ORC_REGULARIZATIONS[regIndex];
                u[0] = 1.0;
                for (int i = 1; i < nm; i++)
                    u[i] = uOriginal[i] * (1.0 - r);
            } else {
                status = ORC_ERR_MAXREG;
                break;
            }
            isotropic_alpha(nm, alpha0, o);
            rtol = nrm2(nm, u) * tolRel;
            memcpy(alpha, alpha0, sizeof(double) * nm);
            iter = 0;
            tt = 1;
            if (regIndex == ORC_NUM_REG - 1)
                maxIter = 10 * maxIter;
            continue;
        }
        for (int i = 0; i < nm; i++) d[i] = g[i];
        trsv_lower_n(nm, chol, d);
        trsv_lower_t(nm, chol, d);
        for (int i = 0; i < nm; i++) d[i] = -d[i];
        err = nrm2(nm, g);
        /* the supplementary criterion gamma <= tolGamma (opt.cpp:288-293) */
        int gamma_ok = 1;
        double gamma = 0.0;
        if (o->tol_gamma > 0.0) {
            gamma = modcfl(nm, d, o->moment_type);
            gamma_ok = gamma <= o->tol_gamma;
        }
        if (o->trace && *o->trace_len < o->trace_cap) {
            double *rec = o->trace + (size_t)(*o->trace_len)++ * ORC_TRACE_WORDS;
            rec[0] = totalIter; rec[1] = regIndex; rec[2] = (err <= rtol && gamma_ok) ? 5 : 0; rec[3] = condest;
            rec[4] = err; rec[5] = rtol; rec[6] = gamma;
        }
        if (err <= rtol && gamma_ok) {
            if (o->moment_type == ORC_MOMENT_MN) {             /* :296-303 */
                double u0 = u[0] + g[0];
                for (int i = 0; i < nm; i++) {
                    u[i] += g[i];
                    u[i] /= u0;
                }
                alpha[0] -= log(u0) * 2.0 * sqrt(M_PI);
            }
            break;
        }
        tt = linesearch(t, alpha, d, f, g, u, alphaOld, o, fscale);
    }
    if (outs != NULL) {
        outs->iter = totalIter;
        outs->norm_g = err;
        outs->r = r;
    }
    if (buf != dstk) free(buf);
    return status;
}

int orc_opt(const orc_tables *t, double *u, double *alpha, const orc_options *o, orc_outs *outs)
{                                                          /* opt.cpp:82-112 */
    int nm = t->nm;
    if (o->moment_type == ORC_MOMENT_PPN) {                /* :113-141 */
        double u0 = u[0];
        for (int i = 0; i < nm; i++) {
            alpha[i] = alpha[i] / u0;
            u[i] = u[i] / u0;
        }
        orc_options os = *o;
        os.delta_ppn = o->delta_ppn / (u0 * u0);
        if (u0 <= 0)
            return ORC_ERR_U0;
        int st = opt_scaled(t, u, alpha, &os, outs);
        for (int i = 0; i < nm; i++) {
            alpha[i] = alpha[i] * u0;
            u[i] = u[i] * u0;
        }
        return st;
    }
    double u0 = u[0];
    double alphaScale = log(u0) * 2.0 * sqrt(M_PI);
    alpha[0] = alpha[0] - alphaScale;
    for (int i = 0; i < nm; i++) u[i] = u[i] / u0;
    if (u0 <= 0)
        return ORC_ERR_U0;
    int st = opt_scaled(t, u, alpha, o, outs);
    for (int i = 0; i < nm; i++) u[i] = u[i] * u0;
    alpha[0] = alpha[0] + alphaScale;
    return st;
}

int orc_opt_batch(const orc_tables *t, int ncells, double *u, double *alpha,
                  const orc_options *o, orc_outs *outs, int nthreads)
{
    int nm = t->nm;
    int status = ORC_OK;
#ifdef _OPENMP
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
#endif
    for (int c = 0; c < ncells; c++) {
        orc_outs tmp;
        int st = orc_opt(t, &u[(size_t)c * nm], &alpha[(size_t)c * nm], o, &tmp);
        if (outs) outs[c] = tmp;
        if (st != ORC_OK) {
#ifdef _OPENMP
#pragma omp critical
#endif
            { if (status == ORC_OK) status = st; }
        }
    }
    (void)nthreads;
    return status;
}

/* ------------------------------------------------------------------ */
/* Change-of-basis optimiser optcb (optcb.cpp:90-415).                  */
/* P2M and the factor are column-major (element (i,j) at a[i + j*n]),   */
/* the layout of the reference's BLAS calls; the triangular kernels     */
/* follow the Netlib reference BLAS loops.                              */
/* ------------------------------------------------------------------ */
/* dtrsv('L','N','N'): x <- A^{-1} x */
static void cb_trsv_ln(int n, const double *a, double *x)
{
    for (int j = 0; j < n; j++) {
        if (x[j] != 0.0) {
            x[j] = x[j] / a[j + j * n];
            double temp = x[j];
            for (int i = j + 1; i < n; i++)
                x[i] = x[i] - temp * a[i + j * n];
        }
    }
}

/* dtrsv('L','T','N'): x <- A^{-T} x */
static void cb_trsv_lt(int n, const double *a, double *x)
{
    for (int j = n - 1; j >= 0; j--) {
        double temp = x[j];
        for (int i = n - 1; i > j; i--)
            temp = temp - a[i + j * n] * x[i];
        temp = temp / a[j + j * n];
        x[j] = temp;
    }
}

/* dtrmv('L','N','N'): x <- A x */
static void cb_trmv_ln(int n, const double *a, double *x)
{
    for (int j = n - 1; j >= 0; j--) {
        if (x[j] != 0.0) {
            double temp = x[j];
            for (int i = n - 1; i > j; i--)
                x[i] = x[i] + temp * a[i + j * n];
            x[j] = x[j] * a[j + j * n];
        }
    }
}

/* dtrmv('L','T','N'): x <- A^T x */
static void cb_trmv_lt(int n, const double *a, double *x)
{
    for (int j = 0; j < n; j++) {
        double temp = x[j];
        temp = temp * a[j + j * n];
        for (int i = j + 1; i < n; i++)
            temp = temp + a[i + j * n] * x[i];
        x[j] = temp;
    }
}

/* dtrsm('L','L','N','N', n, nq, 1, A, B): every column b_q <- A^{-1} b_q;
 * B is the table [nq][n], column q at b[q*n] */
static void cb_trsm_lln(int n, int nq, const double *a, double *b)
{
    for (int q = 0; q < nq; q++) cb_trsv_ln(n, a, b + (size_t)q * n);
}

/* dtrmm('R','L','N','N', n, n, 1, A, B): B <- B A */
static void cb_trmm_rln(int n, const double *a, double *b)
{
    for (int j = 0; j < n; j++) {
        double temp = a[j + j * n];
        for (int i = 0; i < n; i++)
            b[i + j * n] = temp * b[i + j * n];
        for (int k = j + 1; k < n; k++) {
            if (a[k + j * n] != 0.0) {
                temp = a[k + j * n];
                for (int i = 0; i < n; i++)
                    b[i + j * n] = b[i + j * n] + temp * b[i + k * n];
            }
        }
    }
}

static void cb_identity(int n, double *a)
{
    for (int i = 0; i < n * n; i++) a[i] = 0.0;
    for (int i = 0; i < n; i++) a[i + i * n] = 1.0;
}

/* Netlib dgemv('N') with beta = 0: y <- alpha A x (A column-major) */
static void bf_gemv_n(int n, double alpha, const double *a, const double *x, double *y)
{
    for (int i = 0; i < n; i++) y[i] = 0.0;
    for (int j = 0; j < n; j++) {
        const double temp = alpha * x[j];
        for (int i = 0; i < n; i++) y[i] = y[i] + temp * a[i + j * n];
    }
}
/* Netlib dger: A <- A + alpha x y^T (columns with y_j = 0 skipped) */
static void bf_ger(int n, double alpha, const double *x, const double *y, double *a)
{
    for (int j = 0; j < n; j++) {
        if (y[j] != 0.0) {
            const double temp = alpha * y[j];
            for (int i = 0; i < n; i++) a[i + j * n] = a[i + j * n] + x[i] * temp;
        }
    }
}
static double bf_dot(int n, const double *x, const double *y)
{
    double s = 0.0;
    for (int i = 0; i < n; i++) s = s + x[i] * y[i];
    return s;
}
/* cond (optbfgs.cpp:517-545): the largest |.| sum over h[i*n+j], j inner, of each */
static double bf_cond(int n, const double *h, const double *hinv)
{
    double norm1 = 0.0, norm2 = 0.0;
    for (int i = 0; i < n; i++) {
        double s = 0.0;
        for (int j = 0; j < n; j++) s += fabs(h[i * n + j]);
        if (s > norm1) norm1 = s;
    }
    for (int i = 0; i < n; i++) {
        double s = 0.0;
        for (int j = 0; j < n; j++) s += fabs(hinv[i * n + j]);
        if (s > norm2) norm2 = s;
    }
    return norm1 * norm2;
}

void orc_bfgs_update(int n, double *H, double *Hinv, const double *s, const double *y)
{                                                          /* optbfgs.cpp:440-477 */
    double tmp[64];
    double dot = bf_dot(n, s, y);                          /* H += y y^T / (s.y) */
    double coef = 1.0 / dot;
    bf_ger(n, coef, y, y, H);
    bf_gemv_n(n, 1.0, H, s, tmp);                          /* H -= (H s)(H s)^T / (s.H s) */
    dot = bf_dot(n, s, tmp);
    coef = -1.0 / dot;
    bf_ger(n, coef, tmp, tmp, H);
    bf_gemv_n(n, 1.0, Hinv, y, tmp);                       /* H^-1 += (s.y + y.H^-1 y)/(s.y)^2 s s^T */
    coef = bf_dot(n, y, tmp);
    dot = bf_dot(n, s, y);
    coef = (coef + dot) / (dot * dot);
    bf_ger(n, coef, s, s, Hinv);
    coef = -1.0 / dot;                                     /* H^-1 -= ((H^-1 y) s^T + s (H^-1 y)^T)/(s.y) */
    bf_ger(n, coef, tmp, s, Hinv);
    bf_ger(n, coef, s, tmp, Hinv);
}

/* f's resolution for the robust line search: |mass| + sum |alpha_i u_i|, with
 * the mass f + alpha.u */
static double bf_fscale(int nm, double f, const double *a, const double *u)
{
    double s = 0.0, au = 0.0;
    for (int i = 0; i < nm; i++) {
        s += fabs(a[i] * u[i]);
        au += a[i] * u[i];
    }
    return s + fabs(f + au);
}

static int optcb_scaled(const orc_tables *t, double *u, double *P2M, double *alphaP, double *alphaM,
                        const orc_options *o, orc_outs *outs, int bfgs)
{                                                          /* optcb.cpp:159-415, optbfgs.cpp:169-513 */
    const int nm = t->nm, nq = t->nq;
    int status = ORC_OK;
    double *mem = (double *)malloc(sizeof(double) * ((size_t)12 * nm + 6 * (size_t)nm * nm + (size_t)nq * nm));
    if (!mem) return ORC_ERR_ALLOC;
    double *alphaP0 = mem, *dP = alphaP0 + nm, *dM = dP + nm, *gP = dM + nm, *uP = gP + nm, *gM = uP + nm,
           *alphaOld = gM + nm;
    double *chol_rm = alphaOld + 2 * nm, *chol = chol_rm + nm * nm, *hP = chol + nm * nm, *P2Min = hP + nm * nm;
    double *hB = P2Min + nm * nm, *hInv = hB + nm * nm;  /* BFGS's H and H^-1 (column-major) */
    double *gOld = hInv + nm * nm, *aOld = gOld + nm, *sD = aOld + nm, *yD = sD + nm;
    double *pP = yD + nm;
    /* the tables seen through the current basis: p -> pP (stride nm) */
    orc_tables tv = *t;
    tv.p = pP;
    tv.np2 = nm;
    tv.use_gaunt = 0;
    orc_options od = *o;
    od.use_gaunt = 0;
    double f = 0.0;
    memcpy(alphaP0, alphaP, sizeof(double) * nm);
    memcpy(P2Min, P2M, sizeof(double) * nm * nm);
    for (int q = 0; q < nq; q++)                            /* pP <- P2M \ p (:185-187) */
        memcpy(pP + (size_t)q * nm, t->p + (size_t)q * t->np2, sizeof(double) * nm);
    cb_trsm_lln(nm, nq, P2Min, pP);
    int maxIter = o->max_iter;
    const double tolRel = o->tol_rel, condHMax = o->cond_h_max;
    double r = 0.0;
    int regIndex = 0;
    memcpy(uP, u, sizeof(double) * nm);                    /* :198-200 */
    cb_trsv_ln(nm, P2M, uP);
    double rtol = nrm2(nm, u) * tolRel;
    int iter = 0, totalIter = 0;
    double tt = 1;
    double err = 0.0;
    int hitMaxRegularization = 0;
    for (;;) {
        iter = iter + 1;
        totalIter = totalIter + 1;
        f = fobj_any(&tv, alphaP, uP, gP, hP, &od);         /* :234-242 */
        double fscale = 0.0, au = 0.0;                      /* the robust mode's resolution of f */
        for (int i = 0; i < nm; i++) {
            fscale += fabs(alphaP[i] * uP[i]);
            au += alphaP[i] * uP[i];
        }
        if (o->moment_type == ORC_MOMENT_PPN) fscale += fabs(f + au);
        else fscale += fabs(hP[0] / (pP[0] * pP[0]));
        memcpy(chol_rm, hP, sizeof(double) * nm * nm);      /* :247-259 */
        int error = orc_cholesky(nm, chol_rm);
        double norm = orc_norm1(nm, hP);
        double condest = 0.0;
        if (error == 0) condest = 1.0 / orc_rcond(nm, chol_rm, norm);
        for (int i = 0; i < nm; i++)
            for (int j = 0; j < nm; j++)
                chol[i + j * nm] = (i >= j) ? chol_rm[i * nm + j] : 0.0;
        if (totalIter == 1 && (error != 0 || condest > condHMax)) {   /* :264-281 */
            isotropic_alpha(nm, alphaP0, o);
            cb_identity(nm, P2M);
            cb_identity(nm, P2Min);
            memcpy(alphaP, alphaP0, sizeof(double) * nm);
            memcpy(uP, u, sizeof(double) * nm);
            for (int q = 0; q < nq; q++)
                memcpy(pP + (size_t)q * nm, t->p + (size_t)q * t->np2, sizeof(double) * nm);
            rtol = nrm2(nm, u) * tolRel;
            iter = 0;
            tt = 1;
            continue;
        }
        if (iter >= maxIter || error != 0 || tt == 0 || condest > condHMax) {   /* :285-329 */
            if (regIndex < ORC_NUM_REG - 1) {
                regIndex++;
                r = ORC_REGULARIZATIONS[regIndex];
                uP[0] = 1.0;
                for (int i = 1; i < nm; i++)
                    uP[i] = u[i] * (1.0 - r);
            } else {
                if (hitMaxRegularization) {                 /* "Max regularization achieved." */
                    status = ORC_ERR_MAXREG;
                    break;
                }
                isotropic_alpha(nm, alphaP0, o);
                cb_identity(nm, P2Min);
                memcpy(uP, u, sizeof(double) * nm);
                regIndex = 0;
                hitMaxRegularization = 1;
                maxIter = o->max_iter;
            }
            rtol = nrm2(nm, uP) * tolRel;
            memcpy(alphaP, alphaP0, sizeof(double) * nm);
            memcpy(P2M, P2Min, sizeof(double) * nm * nm);
            cb_trsv_ln(nm, P2M, uP);
            for (int q = 0; q < nq; q++)
                memcpy(pP + (size_t)q * nm, t->p + (size_t)q * t->np2, sizeof(double) * nm);
            cb_trsm_lln(nm, nq, P2Min, pP);
            iter = 0;
            tt = 1;
            if (regIndex == ORC_NUM_REG - 1)
                maxIter = 10 * maxIter;
            continue;
        }
        cb_trsv_ln(nm, chol, gP);                           /* :333-343: the basis moves by L */
        cb_trsm_lln(nm, nq, chol, pP);
        cb_trmm_rln(nm, chol, P2M);
        cb_trmv_lt(nm, chol, alphaP);
        cb_trsv_ln(nm, chol, uP);
        if (!bfgs) {
            for (int i = 0; i < nm; i++) dP[i] = -gP[i];    /* :347-348 */
            memcpy(gM, gP, sizeof(double) * nm);            /* :352-354 */
            cb_trmv_ln(nm, P2M, gM);
            err = nrm2(nm, gM);
            if (err <= rtol) {                              /* :358-375 */
                memcpy(dM, dP, sizeof(double) * nm);
                cb_trsv_lt(nm, P2M, dM);
                memcpy(alphaM, alphaP, sizeof(double) * nm);
                cb_trsv_lt(nm, P2M, alphaM);
                u[0] = 1.0;
                for (int i = 1; i < nm; i++)
                    u[i] = u[i] * (1.0 - r);
                break;
            }
            tt = linesearch(&tv, alphaP, dP, f, gP, uP, alphaOld, &od, fscale);   /* :387-388 */
            continue;
        }
        cb_identity(nm, hB);                                /* optbfgs.cpp:371-373 */
        cb_identity(nm, hInv);
        int solved = 0;
        for (int ib = 0; bf_cond(nm, hB, hInv) < o->cond_h_max_bfgs && ib < o->max_bfgs_iter; ib++) {
            bf_gemv_n(nm, -1.0, hInv, gP, dP);              /* :383: d = -H^-1 g */
            memcpy(gM, gP, sizeof(double) * nm);            /* :387-389 */
            cb_trmv_ln(nm, P2M, gM);
            err = nrm2(nm, gM);
            if (err <= rtol) {                              /* :393-411 */
                memcpy(alphaM, alphaP, sizeof(double) * nm);
                cb_trsv_lt(nm, P2M, alphaM);
                u[0] = 1.0;
                for (int i = 1; i < nm; i++)
                    u[i] = u[i] * (1.0 - r);
                solved = 1;
                break;
            }
            memcpy(aOld, alphaP, sizeof(double) * nm);      /* :423-425 */
            memcpy(gOld, gP, sizeof(double) * nm);
            tt = linesearch(&tv, alphaP, dP, f, gP, uP, alphaOld, &od, fscale);
            f = fobj_any(&tv, alphaP, uP, gP, NULL, &od);   /* :429-436: f and g alone */
            fscale = bf_fscale(nm, f, alphaP, uP);
            for (int i = 0; i < nm; i++) {                  /* :440-445 */
                yD[i] = gP[i] - gOld[i];
                sD[i] = alphaP[i] - aOld[i];
            }
            orc_bfgs_update(nm, hB, hInv, sD, yD);
        }
        if (solved) break;
    }
    if (outs != NULL) {
        outs->iter = totalIter;
        outs->norm_g = err;
        outs->r = r;
    }
    free(mem);
    return status;
}

static int optcb_any(const orc_tables *t, double *u, double *P2M, double *alphaP, double *alphaM,
                     const orc_options *o, orc_outs *outs, int bfgs)
{                                                          /* optcb.cpp:90-153, optbfgs.cpp:92-163 */
    const int nm = t->nm;
    if (o->use_gaunt || t->np2 != nm) return ORC_ERR_ARG;  /* pP is copied with stride nm (:183) */
    if (bfgs && (o->max_bfgs_iter < 1 || !(o->cond_h_max_bfgs > 1.0) || nm > 64)) return ORC_ERR_ARG;
    const double u0 = u[0];
    if (o->moment_type == ORC_MOMENT_PPN) {
        orc_options os = *o;
        os.delta_ppn = o->delta_ppn / (u0 * u0);
        for (int i = 0; i < nm; i++) u[i] = u[i] / u0;
        if (u0 <= 0) return ORC_ERR_U0;
        int st = optcb_scaled(t, u, P2M, alphaP, alphaM, &os, outs, bfgs);
        for (int i = 0; i < nm; i++) {
            alphaM[i] = alphaM[i] * u0;
            u[i] = u[i] * u0;
        }
        return st;
    }
    const double alphaScale = log(u0) * 2.0 * sqrt(M_PI);
    for (int i = 0; i < nm; i++) u[i] = u[i] / u0;
    if (u0 <= 0) return ORC_ERR_U0;
    int st = optcb_scaled(t, u, P2M, alphaP, alphaM, o, outs, bfgs);
    for (int i = 0; i < nm; i++) u[i] = u[i] * u0;
    alphaM[0] = alphaM[0] + alphaScale;
    return st;
}

int orc_optcb(const orc_tables *t, double *u, double *P2M, double *alphaP, double *alphaM,
              const orc_options *o, orc_outs *outs)
{
    return optcb_any(t, u, P2M, alphaP, alphaM, o, outs, 0);
}

int orc_optbfgs(const orc_tables *t, double *u, double *P2M, double *alphaP, double *alphaM,
                const orc_options *o, orc_outs *outs)
{
    return optcb_any(t, u, P2M, alphaP, alphaM, o, outs, 1);
}

static int optcb_batch_any(const orc_tables *t, int ncells, double *u, double *P2M, double *alphaP, double *alphaM,
                           const orc_options *o, orc_outs *outs, int nthreads, int bfgs)
{
    const int nm = t->nm;
    int status = ORC_OK;
#ifdef _OPENMP
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
#endif
    for (int c = 0; c < ncells; c++) {
        orc_outs tmp;
        int st = optcb_any(t, &u[(size_t)c * nm], &P2M[(size_t)c * nm * nm], &alphaP[(size_t)c * nm],
                           &alphaM[(size_t)c * nm], o, &tmp, bfgs);
        if (outs) outs[c] = tmp;
        if (st != ORC_OK) {
#ifdef _OPENMP
#pragma omp critical
#endif
            { if (status == ORC_OK) status = st; }
        }
    }
    (void)nthreads;
    return status;
}

int orc_optcb_batch(const orc_tables *t, int ncells, double *u, double *P2M, double *alphaP, double *alphaM,
                    const orc_options *o, orc_outs *outs, int nthreads)
{
    return optcb_batch_any(t, ncells, u, P2M, alphaP, alphaM, o, outs, nthreads, 0);
}

int orc_optbfgs_batch(const orc_tables *t, int ncells, double *u, double *P2M, double *alphaP, double *alphaM,
                      const orc_options *o, orc_outs *outs, int nthreads)
{
    return optcb_batch_any(t, ncells, u, P2M, alphaP, alphaM, o, outs, nthreads, 1);
}

/* ------------------------------------------------------------------ */
/* Grid solver                                                          */
/* ------------------------------------------------------------------ */
struct orc_solver {
    orc_deck deck;
    orc_tables *t;
    int nthreads;
    int nx, ny, NX, NY, nm, nq;       /* NX = nx+4 etc. */
    double dx, dy, maxdt, init_x, init_y;
    double *moments, *alpha, *flux, *moments_old, *sigma_t, *sigma_s;
    double *ansatz;                   /* [QCHUNK][NX*NY] */
    orc_stats stats;
};

#define QCHUNK 32
#define IDX2(s, i, j) ((size_t)(i) * (s)->NY + (j))
#define IDX3(s, i, j, k) (((size_t)(i) * (s)->NY + (j)) * (s)->nm + (k))

void orc_solver_destroy(orc_solver *s)
{
    if (!s) return;
    orc_tables_destroy(s->t);
    free(s->moments); free(s->alpha); free(s->flux); free(s->moments_old);
    free(s->sigma_t); free(s->sigma_s); free(s->ansatz);
    free(s);
}

int orc_solver_nm(const orc_solver *s) { return s->nm; }
double orc_solver_max_dt(const orc_solver *s) { return s->maxdt; }

int orc_solver_create(const orc_deck *deck, int nthreads, orc_solver **out)
{
    *out = NULL;
    if (deck->num_cells_x < 1 || deck->num_cells_y < 1)
        return ORC_ERR_ARG;
    orc_solver *s = (orc_solver *)calloc(1, sizeof(orc_solver));
    if (!s) return ORC_ERR_ALLOC;
    s->deck = *deck;
    s->nthreads = nthreads;
    int st = orc_tables_create(deck->moment_order, deck->quad_order, deck->use_gaunt, &s->t);
    if (st != ORC_OK) { free(s); return st; }
    s->nm = s->t->nm;
    s->nq = s->t->nq;
    s->nx = deck->num_cells_x;
    s->ny = deck->num_cells_y;
    s->NX = s->nx + 4;
    s->NY = s->ny + 4;
    s->dx = (deck->b_x - deck->a_x) / deck->num_cells_x;   /* main.cpp:140-141 */
    s->dy = (deck->b_y - deck->a_y) / deck->num_cells_y;
    s->init_x = deck->a_x;                                 /* main.cpp:154-155 */
    s->init_y = deck->a_y;
    size_t ncell = (size_t)s->NX * s->NY;
    s->moments = (double *)malloc(sizeof(double) * ncell * s->nm);
    s->alpha = (double *)malloc(sizeof(double) * ncell * s->nm);
    s->flux = (double *)calloc(ncell * s->nm, sizeof(double));
    s->moments_old = (double *)calloc(ncell * s->nm, sizeof(double));
    s->sigma_t = (double *)malloc(sizeof(double) * ncell);
    s->sigma_s = (double *)malloc(sizeof(double) * ncell);
    s->ansatz = (double *)malloc(sizeof(double) * ncell * QCHUNK);
    double *init = (double *)malloc(sizeof(double) * ncell);
    if (!s->moments || !s->alpha || !s->flux || !s->moments_old || !s->sigma_t ||
        !s->sigma_s || !s->ansatz || !init) {
        free(init); orc_solver_destroy(s); return ORC_ERR_ALLOC;
    }
    /* initializeGrid (main.cpp:257-372) */
    int g0 = 2, gx2 = 2 + s->nx - 1, gy2 = 2 + s->ny - 1;
    double dx = s->dx, dy = s->dy;
    for (int i = 0; i < s->NX; i++) {
        for (int j = 0; j < s->NY; j++) {
            double x_i = s->init_x + (i - g0) * dx + dx / 2.0;
            double y_j = s->init_y + (j - g0) * dy + dy / 2.0;
            size_t c = IDX2(s, i, j);
            if (i < g0 || i > gx2 || j < g0 || j > gy2) {
                init[c] = deck->floor_value;
                s->sigma_t[c] = deck->sigma;
                s->sigma_s[c] = deck->sigma;
            } else if (deck->init_cond == 0 && deck->gaussian_sigma == 0.0) {
                init[c] = deck->floor_value;
                /* single partition: numNodes odd -> centre cell (main.cpp:304-316) */
                if (i == (s->nx + 2) / 2 && j == (s->ny + 2) / 2)
                    init[c] = 1.0 / (dx * dy);
                s->sigma_t[c] = deck->sigma;
                s->sigma_s[c] = deck->sigma;
            } else if (deck->init_cond == 0) {
                double s2 = deck->gaussian_sigma * deck->gaussian_sigma;
                double gf = 1.0 / (2.0 * M_PI * s2) * exp(-(x_i * x_i + y_j * y_j) / (2.0 * s2));
                init[c] = (gf > deck->floor_value) ? gf : deck->floor_value;
                s->sigma_t[c] = deck->sigma;
                s->sigma_s[c] = deck->sigma;
            } else if (deck->init_cond == 1) {
                init[c] = deck->floor_value;
                s->sigma_t[c] = 1.0;
                s->sigma_s[c] = 1.0;
                if ((fabs(x_i - 2.0) < 0.5 && fabs(y_j - 2.0) < 0.5) ||
                    (fabs(x_i - 2.0) < 0.5 && fabs(y_j - 0.0) < 0.5) ||
                    (fabs(x_i - 2.0) < 0.5 && fabs(y_j + 2.0) < 0.5) ||
                    (fabs(x_i - 1.0) < 0.5 && fabs(y_j - 1.0) < 0.5) ||
                    (fabs(x_i - 1.0) < 0.5 && fabs(y_j + 1.0) < 0.5) ||
                    (fabs(x_i - 0.0) < 0.5 && fabs(y_j + 2.0) < 0.5) ||
                    (fabs(x_i + 1.0) < 0.5 && fabs(y_j - 1.0) < 0.5) ||
                    (fabs(x_i + 1.0) < 0.5 && fabs(y_j + 1.0) < 0.5) ||
                    (fabs(x_i + 2.0) < 0.5 && fabs(y_j - 2.0) < 0.5) ||
                    (fabs(x_i + 2.0) < 0.5 && fabs(y_j - 0.0) < 0.5) ||
                    (fabs(x_i + 2.0) < 0.5 && fabs(y_j + 2.0) < 0.5)) {
                    s->sigma_t[c] = 10.0;
                    s->sigma_s[c] = 0.0;
                }
            } else if (deck->init_cond == 2) {
                s->sigma_t[c] = deck->sigma;
                s->sigma_s[c] = deck->sigma;
                init[c] = 1 + sin(2 * M_PI * x_i) * cos(2 * M_PI * y_j);
            } else {
                init[c] = deck->floor_value;   /* unknown IC: reference leaves it unset */
                s->sigma_t[c] = deck->sigma;
                s->sigma_s[c] = deck->sigma;
            }
        }
    }
    /* MomOptSolver::init (momopt_init.cpp:72,120-143) */
    s->maxdt = deck->cfl_factor * (2.0 / (2.0 + deck->theta)) * (dx * dy) / (dx + dy);
    for (int i = 0; i < s->NX; i++) {
        for (int j = 0; j < s->NY; j++) {
            s->moments[IDX3(s, i, j, 0)] = init[IDX2(s, i, j)] * 2.0 * sqrt(M_PI);
            for (int k = 1; k < s->nm; k++) s->moments[IDX3(s, i, j, k)] = 0.0;
            for (int k = 1; k < s->nm; k++) s->alpha[IDX3(s, i, j, k)] = 0.0;
            const double m0 = s->moments[IDX3(s, i, j, 0)];
            if (deck->moment_type == ORC_MOMENT_PPN)           /* :138-141 */
                s->alpha[IDX3(s, i, j, 0)] = m0 - 4.0 * deck->delta_ppn * M_PI / m0;
            else
                s->alpha[IDX3(s, i, j, 0)] = 2.0 * sqrt(M_PI) * log(m0 / (2.0 * sqrt(M_PI)));
        }
    }
    free(init);
    *out = s;
    return ORC_OK;
}

void orc_solver_get_state(const orc_solver *s, double *moments, double *alpha)
{
    size_t n = (size_t)s->NX * s->NY * s->nm;
    if (moments) memcpy(moments, s->moments, sizeof(double) * n);
    if (alpha) memcpy(alpha, s->alpha, sizeof(double) * n);
}

void orc_solver_set_state(orc_solver *s, const double *moments, const double *alpha)
{
    size_t n = (size_t)s->NX * s->NY * s->nm;
    if (moments) memcpy(s->moments, moments, sizeof(double) * n);
    if (alpha) memcpy(s->alpha, alpha, sizeof(double) * n);
}

void orc_solver_get_moments(const orc_solver *s, double *out)
{
    size_t o = 0;
    for (int i = 2; i < 2 + s->nx; i++)
        for (int j = 2; j < 2 + s->ny; j++)
            for (int k = 0; k < s->nm; k++)
                out[o++] = s->moments[IDX3(s, i, j, k)];
}

void orc_solver_get_stats(const orc_solver *s, orc_stats *st) { *st = s->stats; }

void orc_solver_ghost_fill(orc_solver *s)
{   /* comm.cpp:148-155 -> momopt_boundaries.cpp:190-223 */
    if (s->deck.init_cond != 2) return;
    int x0 = 0, x1 = 2, x2 = 2 + s->nx - 1 + 1 - 2, x3 = 2 * 2 + s->nx - 1 + 1 - 2;
    int y0 = 0, y1 = 2, y2 = 2 + s->ny - 1 + 1 - 2, y3 = 2 * 2 + s->ny - 1 + 1 - 2;
    int nm = s->nm;
    for (int i = 0; i < 2; i++)
        for (int j = 2; j <= 2 + s->ny - 1; j++)
            for (int k = 0; k < nm; k++) {
                s->moments[IDX3(s, x0 + i, j, k)] = s->moments[IDX3(s, x2 + i, j, k)];
                s->moments[IDX3(s, x3 + i, j, k)] = s->moments[IDX3(s, x1 + i, j, k)];
            }
    for (int i = 2; i <= 2 + s->nx - 1; i++)
        for (int j = 0; j < 2; j++)
            for (int k = 0; k < nm; k++) {
                s->moments[IDX3(s, i, y0 + j, k)] = s->moments[IDX3(s, i, y2 + j, k)];
                s->moments[IDX3(s, i, y3 + j, k)] = s->moments[IDX3(s, i, y1 + j, k)];
            }
}

int orc_solver_solve_all(orc_solver *s)
{   /* momopt_update.cpp:208-277 */
    orc_options o;
    o.max_iter = s->deck.max_iter;
    o.tol_rel = s->deck.tol;
    o.cond_h_max = s->deck.cond_h_max;
    o.use_gaunt = s->deck.use_gaunt;
    o.line_search = s->deck.line_search;
    o.moment_type = s->deck.moment_type;
    o.delta_ppn = s->deck.delta_ppn;
    o.max_bfgs_iter = 0;
    o.cond_h_max_bfgs = 0.0;
    o.tol_gamma = s->deck.tol_gamma;
    o.trace = NULL;
    o.trace_cap = 0;
    o.trace_len = NULL;
    int ncell = s->NX * s->NY, nm = s->nm;
    int status = ORC_OK;
    int maxit = s->stats.max_iter;
    long long itsum = 0, hist[ORC_NUM_REG] = {0, 0, 0, 0, 0};
#ifdef _OPENMP
    int nt = s->nthreads > 0 ? s->nthreads : omp_get_max_threads();
#pragma omp parallel num_threads(nt)
#endif
    {
        int lmax = 0;
        long long lsum = 0, lhist[ORC_NUM_REG] = {0, 0, 0, 0, 0};
        int lstatus = ORC_OK;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
        for (int c = 0; c < ncell; c++) {
            orc_outs outs;
            int st = orc_opt(s->t, &s->moments[(size_t)c * nm], &s->alpha[(size_t)c * nm], &o, &outs);
            if (st != ORC_OK && lstatus == ORC_OK) lstatus = st;
            if (outs.iter > lmax) lmax = outs.iter;
            lsum += outs.iter;
            for (int r = 0; r < ORC_NUM_REG; r++)
                if (fabs(outs.r - ORC_REGULARIZATIONS[r]) < 1e-10) lhist[r]++;
        }
#ifdef _OPENMP
#pragma omp critical
#endif
        {
            if (lmax > maxit) maxit = lmax;
            itsum += lsum;
            for (int r = 0; r < ORC_NUM_REG; r++) hist[r] += lhist[r];
            if (lstatus != ORC_OK && status == ORC_OK) status = lstatus;
        }
    }
    s->stats.max_iter = maxit;
    s->stats.iter_sum += itsum;
    s->stats.num_dual_solves += ncell;
    for (int r = 0; r < ORC_NUM_REG; r++) s->stats.hist_reg[r] += hist[r];
    return status;
}

/* The coordinate x0 + k h of a cell edge for the lattice source test
 * |x| < 0.5 (momopt_update.cpp:328-334, kinetic_update.cpp:173-175), rounded
 * as a product and a sum as the reference's build (no FMA) does, also in the
 * FMA-contracted sensitivity build: where edges fall on +-0.5 exactly (280
 * cells on [-3.5, 3.5]) a contraction moves a row and a column of cells into
 * the source, which is a different problem, not a rounding difference. */
__attribute__((optimize("fp-contract=off"), noinline))
static double edge_coord(double x0, int k, double h)
{
    return x0 + k * h;
}

static double sign1(double x) { return (x >= 0.0) ? 1.0 : -1.0; }       /* utils.h:30 */
static double dmin(double a, double b) { return (a < b) ? a : b; }       /* utils.h:28 */
static double dmax(double a, double b) { return (a > b) ? a : b; }       /* utils.h:29 */

static double minmod(double x, double y)                                   /* :77-80 */
{
    return sign1(x) * dmax(0.0, dmin(fabs(x), y * sign1(x)));
}

static double slopefit(double left, double center, double right, double theta) /* :86-90 */
{
    return minmod(theta * (right - center), minmod(0.5 * (right - left), theta * (center - left)));
}

void orc_solver_flux(orc_solver *s, double *flux_out)
{   /* momopt_update.cpp:96-202 */
    int NX = s->NX, NY = s->NY, nm = s->nm, nq = s->nq, np2 = s->t->np2;
    const double *p = s->t->p, *w = s->t->w, *xi_t = s->t->xi, *eta_t = s->t->eta;
    double dx = s->dx, dy = s->dy, theta = s->deck.theta;
    int ncell = NX * NY;
    double *flux = s->flux;
    for (size_t c = 0; c < (size_t)ncell * nm; c++) flux[c] = 0.0;
    for (int q0 = 0; q0 < nq; q0 += QCHUNK) {
        int qn = (nq - q0 < QCHUNK) ? nq - q0 : QCHUNK;
#ifdef _OPENMP
        int nt = s->nthreads > 0 ? s->nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(nt)
#endif
        for (int c = 0; c < ncell; c++) {
            const double *a = &s->alpha[(size_t)c * nm];
            for (int qq = 0; qq < qn; qq++) {
                int q = q0 + qq;
                double kinetic = 0.0;                           /* computeAnsatz :44-58 */
                for (int k = 0; k < nm; k++)
                    kinetic += a[k] * p[q * np2 + k];
                double E;
                if (s->deck.moment_type == ORC_MOMENT_PPN) {
                    double dl = s->deck.delta_ppn;
                    if (kinetic > 0)
                        E = 0.5 * kinetic + 0.5 * sqrt(kinetic * kinetic + 4 * dl);
                    else
                        E = -dl / (0.5 * kinetic - 0.5 * sqrt(kinetic * kinetic + 4 * dl));
                } else {
                    E = ORC_EXP(kinetic);
                }
                s->ansatz[(size_t)qq * ncell + c] = E;
            }
        }
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nt)
#endif
        for (int c = 0; c < ncell; c++) {
            int i = c / NY, j = c % NY;
            if (i < 2 || i > NX - 3 || j < 2 || j > NY - 3) continue;
            for (int qq = 0; qq < qn; qq++) {
                int q = q0 + qq;
                const double *E = &s->ansatz[(size_t)qq * ncell];
                double xi = xi_t[q], eta = eta_t[q], wq = w[q];
                double flux1 = 0, flux2 = 0, flux3 = 0, flux4 = 0;
                if (xi > 0) {
                    double k1 = E[(i - 2) * NY + j], k2 = E[(i - 1) * NY + j];
                    double k3 = E[i * NY + j], k4 = E[(i + 1) * NY + j];
                    flux1 = k3 + 0.5 * slopefit(k2, k3, k4, theta) - k2 - 0.5 * slopefit(k1, k2, k3, theta);
                    flux1 = flux1 * xi * wq;
                }
                if (xi < 0) {
                    double k1 = E[(i - 1) * NY + j], k2 = E[i * NY + j];
                    double k3 = E[(i + 1) * NY + j], k4 = E[(i + 2) * NY + j];
                    flux2 = k3 - 0.5 * slopefit(k2, k3, k4, theta) - k2 + 0.5 * slopefit(k1, k2, k3, theta);
                    flux2 = flux2 * xi * wq;
                }
                if (eta > 0) {
                    double k1 = E[i * NY + (j - 2)], k2 = E[i * NY + (j - 1)];
                    double k3 = E[i * NY + j], k4 = E[i * NY + (j + 1)];
                    flux3 = k3 + 0.5 * slopefit(k2, k3, k4, theta) - k2 - 0.5 * slopefit(k1, k2, k3, theta);
                    flux3 = flux3 * eta * wq;
                }
                if (eta < 0) {
                    double k1 = E[i * NY + (j - 1)], k2 = E[i * NY + j];
                    double k3 = E[i * NY + (j + 1)], k4 = E[i * NY + (j + 2)];
                    flux4 = k3 - 0.5 * slopefit(k2, k3, k4, theta) - k2 + 0.5 * slopefit(k1, k2, k3, theta);
                    flux4 = flux4 * eta * wq;
                }
                double comb = ((flux1 + flux2) / dx + (flux3 + flux4) / dy);
                for (int k = 0; k < nm; k++)
                    flux[(size_t)c * nm + k] += p[q * np2 + k] * comb;
            }
        }
    }
    if (flux_out) memcpy(flux_out, flux, sizeof(double) * (size_t)ncell * nm);
}

static void euler_stage(orc_solver *s, double dt)
{   /* momopt_update.cpp:316-336 / 344-364 */
    int nm = s->nm;
    for (int i = 2; i <= 2 + s->nx - 1; i++) {
        for (int j = 2; j <= 2 + s->ny - 1; j++) {
            size_t c = IDX2(s, i, j);
            double *u = &s->moments[c * nm];
            const double *F = &s->flux[c * nm];
            u[0] = u[0] * (1.0 + dt * s->sigma_s[c] - dt * s->sigma_t[c]) - dt * F[0];
            for (int k = 1; k < nm; k++)
                u[k] = u[k] * (1.0 - dt * s->sigma_t[c]) - dt * F[k];
            if (s->deck.init_cond == 1) {
                double x_i = edge_coord(s->init_x, i - 2, s->dx);
                double y_j = edge_coord(s->init_y, j - 2, s->dy);
                if (fabs(x_i) < 0.5 && fabs(y_j) < 0.5)
                    u[0] = u[0] + dt * 2.0 * sqrt(M_PI);
            }
        }
    }
}

int orc_solver_update(orc_solver *s, double dt)
{   /* momopt_update.cpp:283-382 */
    int nm = s->nm;
    for (int i = 2; i <= 2 + s->nx - 1; i++)
        for (int j = 2; j <= 2 + s->ny - 1; j++)
            for (int k = 0; k < nm; k++)
                s->moments_old[IDX3(s, i, j, k)] = s->moments[IDX3(s, i, j, k)];
    for (int stage = 0; stage < 2; stage++) {
        orc_solver_ghost_fill(s);
        int st = orc_solver_solve_all(s);
        if (st != ORC_OK) return st;
        orc_solver_flux(s, NULL);
        euler_stage(s, dt);
    }
    for (int i = 2; i <= 2 + s->nx - 1; i++)
        for (int j = 2; j <= 2 + s->ny - 1; j++)
            for (int k = 0; k < nm; k++)
                s->moments[IDX3(s, i, j, k)] =
                    0.5 * (s->moments[IDX3(s, i, j, k)] + s->moments_old[IDX3(s, i, j, k)]);
    return ORC_OK;
}

int orc_solver_run(orc_solver *s, double t_final, double out_delta_t, int *steps)
{   /* main.cpp:204-230 */
    double t = 0;
    double dt = s->maxdt;
    int n = 0;
    for (double tOut = dmin(t + out_delta_t, t_final); tOut <= t_final;
         tOut = dmin(tOut + out_delta_t, t_final)) {
        dt = s->maxdt;
        while (t < tOut) {
            t += dt;
            if (t > tOut) {
                dt = tOut - (t - dt);
                t = tOut;
            }
            int st = orc_solver_update(s, dt);
            n++;
            if (st != ORC_OK) { if (steps) *steps = n; return st; }
        }
        if (tOut > t_final - 1e-12 * t_final)
            break;
    }
    if (steps) *steps = n;
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* Sn discrete ordinates (src/kinetic/), one rank: an independent truth */
/* for the moment closures.  State f[((i*NY)+j)*nq + q], NX = nx + 4,   */
/* NY = ny + 4, two ghost layers; sigma arrays [(i*NY)+j].              */
/* ------------------------------------------------------------------ */
static void sn_ghost_fill(int nx, int ny, int nq, double *f)
{   /* kinetic_boundaries.cpp:190-223 */
    const int NY = ny + 4;
    const int x0 = 0, x1 = 2, x2 = nx, x3 = nx + 2, y0 = 0, y1 = 2, y2 = ny, y3 = ny + 2;
    for (int i = 0; i < 2; i++)
        for (int j = 2; j <= ny + 1; j++)
            for (int q = 0; q < nq; q++) {
                f[((size_t)(x0 + i) * NY + j) * nq + q] = f[((size_t)(x2 + i) * NY + j) * nq + q];
                f[((size_t)(x3 + i) * NY + j) * nq + q] = f[((size_t)(x1 + i) * NY + j) * nq + q];
            }
    for (int i = 2; i <= nx + 1; i++)
        for (int j = 0; j < 2; j++)
            for (int q = 0; q < nq; q++) {
                f[((size_t)i * NY + y0 + j) * nq + q] = f[((size_t)i * NY + y2 + j) * nq + q];
                f[((size_t)i * NY + y3 + j) * nq + q] = f[((size_t)i * NY + y1 + j) * nq + q];
            }
}

/* one forward-Euler stage in place: the flux of every interior cell from the
 * stage's input, then scattering, the init_cond 1 source and the flux
 * (kinetic_update.cpp:41-114 and :149-181) */
static void sn_stage(int nx, int ny, int nq, const double *w, const double *xi, const double *eta,
                     const double *sigma_t, const double *sigma_s, int init_cond, double init_x, double init_y,
                     double dx, double dy, double dt, double *f, double *flux)
{
    const int NY = ny + 4;
#define SNI(i, j, q) (((size_t)(i) * NY + (j)) * nq + (q))
    if (init_cond == 2) sn_ghost_fill(nx, ny, nq, f);       /* comm.cpp:149-155 */
    for (int i = 2; i <= nx + 1; i++)
        for (int j = 2; j <= ny + 1; j++)
            for (int q = 0; q < nq; q++) {
                double nF, sF, eF, wF;
                if (eta[q] > 0) {
                    double s1 = f[SNI(i, j - 2, q)], s2 = f[SNI(i, j - 1, q)], s3 = f[SNI(i, j, q)], s4 = f[SNI(i, j + 1, q)];
                    nF = s3 + 0.5 * slopefit(s2, s3, s4, 2.0);
                    sF = s2 + 0.5 * slopefit(s1, s2, s3, 2.0);
                } else {
                    double s1 = f[SNI(i, j - 1, q)], s2 = f[SNI(i, j, q)], s3 = f[SNI(i, j + 1, q)], s4 = f[SNI(i, j + 2, q)];
                    nF = s3 - 0.5 * slopefit(s2, s3, s4, 2.0);
                    sF = s2 - 0.5 * slopefit(s1, s2, s3, 2.0);
                }
                if (xi[q] > 0) {
                    double s1 = f[SNI(i - 2, j, q)], s2 = f[SNI(i - 1, j, q)], s3 = f[SNI(i, j, q)], s4 = f[SNI(i + 1, j, q)];
                    eF = s3 + 0.5 * slopefit(s2, s3, s4, 2.0);
                    wF = s2 + 0.5 * slopefit(s1, s2, s3, 2.0);
                } else {
                    double s1 = f[SNI(i - 1, j, q)], s2 = f[SNI(i, j, q)], s3 = f[SNI(i + 1, j, q)], s4 = f[SNI(i + 2, j, q)];
                    eF = s3 - 0.5 * slopefit(s2, s3, s4, 2.0);
                    wF = s2 - 0.5 * slopefit(s1, s2, s3, 2.0);
                }
                flux[SNI(i, j, q)] = (eF - wF) * xi[q] / dx + (nF - sF) * eta[q] / dy;
            }
    for (int i = 2; i <= nx + 1; i++)
        for (int j = 2; j <= ny + 1; j++) {
            const size_t c = (size_t)i * NY + j;
            double integral = 0;
            for (int q = 0; q < nq; q++) integral += w[q] * f[c * nq + q];
            integral = integral * sigma_s[c] / (4 * M_PI);
            for (int q = 0; q < nq; q++) {
                double v = f[c * nq + q];
                v = v - dt * sigma_t[c] * v + dt * integral;
                if (init_cond == 1) {
                    const double x_i = edge_coord(init_x, i, dx), y_j = edge_coord(init_y, j, dy);
                    if (fabs(x_i) < 0.5 && fabs(y_j) < 0.5) v = v + dt;
                }
                f[c * nq + q] = v - dt * flux[c * nq + q];
            }
        }
#undef SNI
}

int orc_sn_step(int nx, int ny, int nq, const double *w, const double *xi, const double *eta,
                const double *sigma_t, const double *sigma_s, int init_cond, double init_x, double init_y,
                double dx, double dy, double dt, double *f)
{   /* KineticSolver::update, kinetic_update.cpp:120-235 */
    if (nx < 2 || ny < 2 || nq < 1) return ORC_ERR_ARG;
    const size_t n = (size_t)(nx + 4) * (ny + 4) * nq;
    double *old = (double *)malloc(sizeof(double) * n), *flux = (double *)calloc(n, sizeof(double));
    if (!old || !flux) { free(old); free(flux); return ORC_ERR_ALLOC; }
    memcpy(old, f, sizeof(double) * n);
    sn_stage(nx, ny, nq, w, xi, eta, sigma_t, sigma_s, init_cond, init_x, init_y, dx, dy, dt, f, flux);
    sn_stage(nx, ny, nq, w, xi, eta, sigma_t, sigma_s, init_cond, init_x, init_y, dx, dy, dt, f, flux);
    const int NY = ny + 4;
    for (int i = 2; i <= nx + 1; i++)                        /* :221-230 */
        for (int j = 2; j <= ny + 1; j++)
            for (int q = 0; q < nq; q++) {
                const size_t k = ((size_t)i * NY + j) * nq + q;
                f[k] = 0.5 * (f[k] + old[k]);
            }
    free(old);
    free(flux);
    return ORC_OK;
}
