/*
 * vbd_oracle.c -- FP64 CPU restatement of WarpVBD's Vertex Block Descent time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP solver in
 * warpvbd_amd/ is compared against, and the CPU baseline that bench.py times.
 * Nothing in the product links, loads or calls it.
 *
 * Every function cites the reference routine it restates (paths relative to the
 * WarpVBD repository).  The restatement is op-for-op where the reference's
 * numerics matter for parity:
 *   - Lame parameters                 src/vbd_solver.py:345-362
 *   - rest shape (invDm, V, m)        src/vbd_solver.py:295-334
 *   - dDs/dx constants                src/vbd_solver.py:449-467
 *   - inertia gradient/Hessian        src/vbd_solver.py:15-47
 *   - elastic gradient/Hessian (SNH)  src/vbd_solver.py:50-124   (expanded form)
 *   - elastic gradient/Hessian (StVK) src/vbd_solver.py:127-144  (commented out in
 *                                      the reference; restated from the comment)
 *   - per-slot accumulation + damping src/vbd_solver.py:149-232
 *   - 32-row sum + Cholesky solve     src/vbd_solver.py:236-274
 *   - initial guess                   src/vbd_solver.py:277-292
 *   - graph coloring                  src/coloring.py:24-72
 *   - vertex->element adjacency       src/vbd_solver.py:485-498
 *   - time step / convergence / vel.  src/vbd_solver.py:509-612
 *
 * A second, algebraically identical "compact" evaluation of the per-slot SNH and
 * StVK blocks (H = V[s2 (Fb)(Fb)^T + s1 |b|^2 I + lam' (cof F b)(cof F b)^T] for
 * SNH) is provided for speed on large meshes; tests pin it against the expanded
 * form.  Deliberate, documented deviations from the reference:
 *   - a vertex with no incident tet is left in place (the reference's Cholesky of
 *     a zero block yields NaN);
 *   - a block that is not positive definite (possible for StVK) is solved by the
 *     adjugate (the reference's Cholesky would take a square root of a negative
 *     pivot); a block whose determinant is tiny relative to its diagonal product
 *     is skipped (dx = 0).  NaN/Inf inputs still propagate.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_MAX_COLORS 256           /* src/coloring.py:50 */
#define ORC_MAX_ELEMENTS_PER_VERTEX 32 /* src/vbd_solver.py:12 */

enum { ORC_SNH = 0, ORC_STVK = 1 };

/* ------------------------------------------------------------------ 3x3 helpers */
typedef double m33[3][3];

static void mm(const m33 A, const m33 B, m33 C)
{
    m33 T;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            T[i][j] = A[i][0] * B[0][j] + A[i][1] * B[1][j] + A[i][2] * B[2][j];
    memcpy(C, T, sizeof(m33));
}
static void mt(const m33 A, m33 T)
{
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) T[i][j] = A[j][i];
}
static double tr(const m33 A) { return A[0][0] + A[1][1] + A[2][2]; }
static double det3(const m33 A)
{
    return A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1])
         - A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0])
         + A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
}
/* adjugate inverse; a singular matrix gives a zero inverse */
static void inv3(const m33 A, m33 R)
{
    double d = det3(A);
    m33 C;
    C[0][0] = A[1][1] * A[2][2] - A[1][2] * A[2][1];
    C[0][1] = A[0][2] * A[2][1] - A[0][1] * A[2][2];
    C[0][2] = A[0][1] * A[1][2] - A[0][2] * A[1][1];
    C[1][0] = A[1][2] * A[2][0] - A[1][0] * A[2][2];
    C[1][1] = A[0][0] * A[2][2] - A[0][2] * A[2][0];
    C[1][2] = A[0][2] * A[1][0] - A[0][0] * A[1][2];
    C[2][0] = A[1][0] * A[2][1] - A[1][1] * A[2][0];
    C[2][1] = A[0][1] * A[2][0] - A[0][0] * A[2][1];
    C[2][2] = A[0][0] * A[1][1] - A[0][1] * A[1][0];
    double s = (d != 0.0) ? 1.0 / d : 0.0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) R[i][j] = C[i][j] * s;
}

/* ------------------------------------------------------------------ setup */

/* src/vbd_solver.py:345-362 (convert_youngs_poisson_to_lame) */
void orc_lame(int m, const double *E, const double *nu, double *mu, double *lam)
{
    for (int i = 0; i < m; ++i) {
        mu[i] = E[i] / (2.0 * (1.0 + nu[i]));
        lam[i] = (E[i] * nu[i]) / ((1.0 + nu[i]) * (1.0 - 2.0 * nu[i]));
    }
}

/* src/vbd_solver.py:295-334 (compute_element_invDm_masses_volume).
 * invDm is stored row-major, 9 doubles per tet.  Returns the number of tets
 * whose rest shape matrix is singular (the reference only asserts on them in
 * Warp debug builds, :336-343). */
int orc_rest_shape(const double *X, int m, const int32_t *tets, const double *rho,
                   double *invDm, double *mass, double *vol)
{
    int bad = 0;
    for (int e = 0; e < m; ++e) {
        const int32_t *t = tets + 4 * e;
        const double *x3 = X + 3 * t[3];
        m33 Dm, iD;
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r) Dm[r][c] = X[3 * t[c] + r] - x3[r]; /* matrix_from_cols */
        double d = det3(Dm);
        if (d == 0.0 || !isfinite(d)) ++bad;
        inv3(Dm, iD);
        memcpy(invDm + 9 * e, iD, sizeof(m33));
        vol[e] = fabs(d) / 6.0;
        mass[e] = rho[e] * vol[e];
    }
    return bad;
}

/* src/vbd_solver.py:450-465: dDs_dx[i][k] = e_k e_i^T for i<3, -e_k 1^T for i=3. */
static void dDs_dx(int i, int k, m33 D)
{
    memset(D, 0, sizeof(m33));
    if (i < 3) D[k][i] = 1.0;
    else D[k][0] = D[k][1] = D[k][2] = -1.0;
}

/* ------------------------------------------------------------------ per-slot physics */

/* src/vbd_solver.py:15-47 (inertial_gradient_hessian): m = m_e/4,
 * g = m/dt^2 (x - (x_prev + dt v_prev)) - m gravity, H = m/dt^2 I. */
static void inertia_slot(const double x[3], const double xp[3], const double vp[3], double me,
                         const double grav[3], double dt, double g[3], m33 H)
{
    double m = me / 4.0;
    memset(H, 0, sizeof(m33));
    for (int k = 0; k < 3; ++k) {
        double y = xp[k] + dt * vp[k];
        g[k] = m / (dt * dt) * (x[k] - y);
        H[k][k] = m / (dt * dt);
    }
    for (int k = 0; k < 3; ++k) g[k] += -m * grav[k];
}

/* Expanded restatement of src/vbd_solver.py:50-124 (SNH) and :127-144 (StVK).
 * xs: the 4 current vertex positions of the tet (12 doubles); vloc: the local
 * index (0..3) of the vertex being differentiated (the reference's mask trick,
 * :88-94, selects exactly this row for a non-degenerate tet). */
static void elastic_slot_expanded(int material, int vloc, const double *xs, const m33 invD,
                                  double V, double lmu, double llam, double g[3], m33 H)
{
    m33 Ds, F, Ft, Finv, FinvT, dF[3];
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) Ds[r][c] = xs[3 * c + r] - xs[9 + r];
    mm(Ds, invD, F);
    for (int k = 0; k < 3; ++k) {
        m33 D;
        dDs_dx(vloc, k, D);
        mm(D, invD, dF[k]);
    }
    double J = det3(F);
    mt(F, Ft);
    m33 FtF;
    mm(Ft, F, FtF);
    double Ic = tr(FtF);
    inv3(F, Finv);
    mt(Finv, FinvT);
    memset(H, 0, sizeof(m33));
    g[0] = g[1] = g[2] = 0.0;

    if (material == ORC_SNH) {
        double mu = 4.0 * lmu / 3.0;           /* :108 */
        double lmbda = llam + 5.0 * lmu / 6.0; /* :109 */
        double alpha = 1.0 + 0.75 * mu / lmbda; /* :110 */
        m33 P;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c)
                P[r][c] = mu * F[r][c] * (1.0 - 1.0 / (Ic + 1.0)) + lmbda * (J - alpha) * J * FinvT[r][c];
        m33 Pt;
        mt(P, Pt);
        for (int i = 0; i < 3; ++i) {
            m33 T;
            mm(Pt, dF[i], T);
            g[i] = V * tr(T); /* :117 */
            for (int j = 0; j < 3; ++j) {
                m33 A, B, Ci, Cj, dFit, Q, R;
                mm(Ft, dF[i], A);
                mm(Ft, dF[j], B);
                mt(dF[i], dFit);
                mm(dFit, dF[j], Q);
                mm(Finv, dF[i], Ci);
                mm(Finv, dF[j], Cj);
                mm(Cj, Ci, R); /* Finv dF_j Finv dF_i */
                H[i][j] = V * (2.0 * mu / ((Ic + 1.0) * (Ic + 1.0)) * tr(A) * tr(B)
                               + mu * (1.0 - 1.0 / (Ic + 1.0)) * tr(Q)
                               + lmbda * J * (2.0 * J - alpha) * tr(Ci) * tr(Cj)
                               - lmbda * J * (J - alpha) * tr(R)); /* :119-124 */
            }
        }
    } else {
        /* StVK, unmodified mu and lambda (:128-144) */
        m33 E;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) E[r][c] = 0.5 * (FtF[r][c] - (r == c ? 1.0 : 0.0));
        m33 dE[3];
        for (int i = 0; i < 3; ++i) {
            m33 A, B, dFt;
            mm(Ft, dF[i], A);
            mt(dF[i], dFt);
            mm(dFt, F, B);
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) dE[i][r][c] = 0.5 * (A[r][c] + B[r][c]);
        }
        for (int i = 0; i < 3; ++i) {
            m33 dEt, T;
            mt(dE[i], dEt);
            mm(dEt, E, T);
            g[i] = V * (llam * tr(E) * tr(dE[i]) + 2.0 * lmu * tr(T));
            for (int j = 0; j < 3; ++j) {
                m33 dFit, dFjt, P1, P2, d2E, d2Et, T1, T2;
                mt(dF[i], dFit);
                mt(dF[j], dFjt);
                mm(dFit, dF[j], P1);
                mm(dFjt, dF[i], P2);
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c) d2E[r][c] = 0.5 * (P1[r][c] + P2[r][c]);
                mm(dEt, dE[j], T1);
                mt(d2E, d2Et);
                mm(d2Et, E, T2);
                H[i][j] += V * (llam * (tr(dE[i]) * tr(dE[j]) + tr(E) * tr(d2E))
                                + 2.0 * lmu * (tr(T1) + tr(T2)));
            }
        }
    }
}

/* Compact, algebraically identical evaluation (dF_i = e_i b^T with b the row of
 * invDm belonging to the vertex's local slot, or minus the row sum for slot 3). */
static void elastic_slot_compact(int material, int vloc, const double *xs, const m33 invD,
                                 double V, double lmu, double llam, double g[3], m33 H)
{
    m33 Ds, F;
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) Ds[r][c] = xs[3 * c + r] - xs[9 + r];
    mm(Ds, invD, F);
    double b[3];
    for (int k = 0; k < 3; ++k)
        b[k] = (vloc < 3) ? invD[vloc][k] : -(invD[0][k] + invD[1][k] + invD[2][k]);
    double Fb[3], bb = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
    for (int r = 0; r < 3; ++r) Fb[r] = F[r][0] * b[0] + F[r][1] * b[1] + F[r][2] * b[2];
    double Ic = 0.0;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) Ic += F[r][c] * F[r][c];
    if (material == ORC_SNH) {
        m33 cof;
        cof[0][0] = F[1][1] * F[2][2] - F[1][2] * F[2][1];
        cof[0][1] = F[1][2] * F[2][0] - F[1][0] * F[2][2];
        cof[0][2] = F[1][0] * F[2][1] - F[1][1] * F[2][0];
        cof[1][0] = F[0][2] * F[2][1] - F[0][1] * F[2][2];
        cof[1][1] = F[0][0] * F[2][2] - F[0][2] * F[2][0];
        cof[1][2] = F[0][1] * F[2][0] - F[0][0] * F[2][1];
        cof[2][0] = F[0][1] * F[1][2] - F[0][2] * F[1][1];
        cof[2][1] = F[0][2] * F[1][0] - F[0][0] * F[1][2];
        cof[2][2] = F[0][0] * F[1][1] - F[0][1] * F[1][0];
        double J = F[0][0] * cof[0][0] + F[0][1] * cof[0][1] + F[0][2] * cof[0][2];
        double mu = 4.0 * lmu / 3.0, lmbda = llam + 5.0 * lmu / 6.0;
        double alpha = 1.0 + 0.75 * mu / lmbda;
        double s1 = mu * (1.0 - 1.0 / (Ic + 1.0)), s2 = 2.0 * mu / ((Ic + 1.0) * (Ic + 1.0));
        double cb[3];
        for (int r = 0; r < 3; ++r) cb[r] = cof[r][0] * b[0] + cof[r][1] * b[1] + cof[r][2] * b[2];
        for (int r = 0; r < 3; ++r) {
            g[r] = V * (s1 * Fb[r] + lmbda * (J - alpha) * cb[r]);
            for (int c = 0; c < 3; ++c)
                H[r][c] = V * (s2 * Fb[r] * Fb[c] + (r == c ? s1 * bb : 0.0) + lmbda * cb[r] * cb[c]);
        }
    } else {
        m33 E;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                double s = F[0][r] * F[0][c] + F[1][r] * F[1][c] + F[2][r] * F[2][c];
                E[r][c] = 0.5 * (s - (r == c ? 1.0 : 0.0));
            }
        double trE = tr(E), Eb[3], bEb = 0.0;
        for (int r = 0; r < 3; ++r) {
            Eb[r] = E[r][0] * b[0] + E[r][1] * b[1] + E[r][2] * b[2];
            bEb += b[r] * Eb[r];
        }
        for (int r = 0; r < 3; ++r) {
            double FEb = F[r][0] * Eb[0] + F[r][1] * Eb[1] + F[r][2] * Eb[2];
            g[r] = V * (llam * trE * Fb[r] + 2.0 * lmu * FEb);
            for (int c = 0; c < 3; ++c) {
                double FFt = F[r][0] * F[c][0] + F[r][1] * F[c][1] + F[r][2] * F[c][2];
                H[r][c] = V * ((llam + lmu) * Fb[r] * Fb[c] + lmu * bb * FFt
                               + (r == c ? (llam * trE * bb + 2.0 * lmu * bEb) : 0.0));
            }
        }
    }
}

/* Public per-slot evaluation for tests: g[3], H[9] row-major. */
void orc_slot_grad_hess(int material, int expanded, int vloc, const double *xs, const double *invDm9,
                        double V, double lmu, double llam, double *g, double *H9)
{
    m33 iD, H;
    memcpy(iD, invDm9, sizeof(m33));
    if (expanded) elastic_slot_expanded(material, vloc, xs, iD, V, lmu, llam, g, H);
    else elastic_slot_compact(material, vloc, xs, iD, V, lmu, llam, g, H);
    memcpy(H9, H, sizeof(m33));
}

/* The elastic energy whose derivatives the slots are (for finite-difference tests).
 * SNH: V[mu'/2 (Ic-3) - mu'/2 ln(Ic+1) + lam'/2 (J-alpha)^2];  StVK: V[mu E:E + lam/2 tr(E)^2]. */
double orc_energy(int material, const double *xs, const double *invDm9, double V, double lmu, double llam)
{
    m33 iD, Ds, F, Ft, FtF;
    memcpy(iD, invDm9, sizeof(m33));
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) Ds[r][c] = xs[3 * c + r] - xs[9 + r];
    mm(Ds, iD, F);
    mt(F, Ft);
    mm(Ft, F, FtF);
    double Ic = tr(FtF), J = det3(F);
    if (material == ORC_SNH) {
        double mu = 4.0 * lmu / 3.0, lmbda = llam + 5.0 * lmu / 6.0;
        double alpha = 1.0 + 0.75 * mu / lmbda;
        return V * (0.5 * mu * (Ic - 3.0) - 0.5 * mu * log(Ic + 1.0) + 0.5 * lmbda * (J - alpha) * (J - alpha));
    }
    double EE = 0.0, trE = 0.0;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            double e = 0.5 * (FtF[r][c] - (r == c ? 1.0 : 0.0));
            EE += e * e;
            if (r == c) trE += e;
        }
    return V * (lmu * EE + 0.5 * llam * trE * trE);
}

/* Cholesky-Banachiewicz solve of H dx = -G (the reference's tile_cholesky /
 * tile_cholesky_solve, src/vbd_solver.py:267-268).  Falls back to the adjugate
 * when a pivot is not positive; skips (dx = 0) a block whose determinant is tiny
 * relative to its diagonal product (|det| <= 1e-6 |h00 h11 h22|).
 * Returns 0 = Cholesky, 1 = adjugate, 2 = skipped. */
int orc_solve3(const double *H9, const double *G, double *dx)
{
    const double(*H)[3] = (const double(*)[3])H9;
    m33 A;
    memcpy(A, H9, sizeof(m33));
    double d = det3(A);
    double scale = fabs(A[0][0] * A[1][1] * A[2][2]);
    if (isfinite(d) && fabs(d) <= 1e-6 * scale && G[0] == G[0] && G[1] == G[1] && G[2] == G[2]) {
        dx[0] = dx[1] = dx[2] = 0.0; /* guard: no incident tet, or (near) singular block */
        return 2;
    }
    double L00 = H[0][0];
    int spd = L00 > 0.0;
    double L10 = 0, L20 = 0, L11 = 0, L21 = 0, L22 = 0;
    if (spd) {
        L00 = sqrt(L00);
        L10 = H[1][0] / L00;
        L20 = H[2][0] / L00;
        double p = H[1][1] - L10 * L10;
        spd = p > 0.0;
        if (spd) {
            L11 = sqrt(p);
            L21 = (H[2][1] - L20 * L10) / L11;
            double q = H[2][2] - L20 * L20 - L21 * L21;
            spd = q > 0.0;
            if (spd) L22 = sqrt(q);
        }
    }
    if (spd) {
        double r0 = -G[0], r1 = -G[1], r2 = -G[2];
        double y0 = r0 / L00;
        double y1 = (r1 - L10 * y0) / L11;
        double y2 = (r2 - L20 * y0 - L21 * y1) / L22;
        dx[2] = y2 / L22;
        dx[1] = (y1 - L21 * dx[2]) / L11;
        dx[0] = (y0 - L10 * dx[1] - L20 * dx[2]) / L00;
        return 0;
    }
    if (!(d == d) || !isfinite(d)) { /* NaN/Inf propagate */
        dx[0] = dx[1] = dx[2] = NAN;
        return 1;
    }
    m33 R;
    inv3(A, R);
    for (int r = 0; r < 3; ++r) dx[r] = -(R[r][0] * G[0] + R[r][1] * G[1] + R[r][2] * G[2]);
    return 1;
}

/* ------------------------------------------------------------------ coloring */

/* src/coloring.py:24-72 (compute_adjacency_dict + graph_coloring): sequential
 * first-fit in vertex-index order over the vertex graph of the elements.  The
 * vertex count is max(element index)+1 (coloring.py:28).  colors must hold n
 * entries; vertices above the largest referenced index stay -1.  Returns the
 * number of colors, or -1 if a vertex would need more than 256 (the reference
 * leaves it uncolored, coloring.py:62-68).  *max_degree receives the value
 * compute_adjacency_dict returns (degree of each vertex sampled after each of
 * its element visits, maximised). */
int orc_color_mesh(int n, int m, const int32_t *tets, int32_t *colors, int *max_degree)
{
    int nv = 0;
    for (int i = 0; i < 4 * m; ++i)
        if (tets[i] + 1 > nv) nv = tets[i] + 1;
    for (int v = 0; v < n; ++v) colors[v] = -1;
    /* vertex -> element CSR */
    int *cnt = calloc((size_t)nv + 1, sizeof(int));
    for (int i = 0; i < 4 * m; ++i) cnt[tets[i] + 1]++;
    for (int v = 0; v < nv; ++v) cnt[v + 1] += cnt[v];
    int *cur = malloc(sizeof(int) * ((size_t)nv + 1));
    memcpy(cur, cnt, sizeof(int) * ((size_t)nv + 1));
    int *inc = malloc(sizeof(int) * (size_t)(4 * (size_t)m + 1));
    for (int e = 0; e < m; ++e)
        for (int k = 0; k < 4; ++k) inc[cur[tets[4 * e + k]]++] = e;
    /* max degree as compute_adjacency_dict measures it: the size of the
     * neighbour set of ele_i after processing element e, for each slot i. */
    if (max_degree) {
        int *deg = calloc((size_t)nv, sizeof(int));
        int *seen = malloc(sizeof(int) * (size_t)nv);
        for (int v = 0; v < nv; ++v) seen[v] = -1;
        /* deg grows monotonically; the final neighbour-set size is reached at the
         * vertex's last incidence, so max over vertices of |N(v)| is exact. */
        int md = 0;
        for (int v = 0; v < nv; ++v) {
            int d = 0;
            for (int s = cnt[v]; s < cnt[v + 1]; ++s) {
                const int32_t *t = tets + 4 * inc[s];
                for (int k = 0; k < 4; ++k) {
                    int u = t[k];
                    if (u != v && seen[u] != v) { seen[u] = v; ++d; }
                }
            }
            /* a self-loop (degenerate element) adds v to its own set */
            for (int s = cnt[v]; s < cnt[v + 1]; ++s) {
                const int32_t *t = tets + 4 * inc[s];
                int c = 0;
                for (int k = 0; k < 4; ++k) c += (t[k] == v);
                if (c > 1) { ++d; break; }
            }
            deg[v] = d;
            if (d > md) md = d;
        }
        *max_degree = md;
        free(deg);
        free(seen);
    }
    int ncol = 0, fail = 0;
    uint64_t used[ORC_MAX_COLORS / 64];
    for (int v = 0; v < nv && v < n; ++v) {
        memset(used, 0, sizeof(used));
        for (int s = cnt[v]; s < cnt[v + 1]; ++s) {
            const int32_t *t = tets + 4 * inc[s];
            for (int k = 0; k < 4; ++k) {
                int u = t[k];
                if (u == v || u >= n) continue;
                int c = colors[u];
                if (c >= 0) used[c >> 6] |= 1ull << (c & 63);
            }
        }
        int c = 0;
        while (c < ORC_MAX_COLORS && (used[c >> 6] >> (c & 63) & 1)) ++c;
        if (c == ORC_MAX_COLORS) { fail = 1; continue; }
        colors[v] = c;
        if (c + 1 > ncol) ncol = c + 1;
    }
    free(cnt);
    free(cur);
    free(inc);
    return fail ? -1 : ncol;
}

/* ------------------------------------------------------------------ solver */

typedef struct orc_solver {
    int n, m, material, expanded, ncolors, W;
    double grav[3], damping;
    double *X, *invDm, *mass, *vol, *mu, *lam;
    int32_t *tets;
    uint8_t *active;
    int32_t *colors;
    int *col_off, *col_v; /* color groups: vertices ascending within a color */
    int *adj_off, *adj_e; /* vertex -> elements, ascending element index (:485-498) */
    double *xprev, *vprev, *x, *dxs;
    int nthreads;
} orc_solver;

void orc_destroy(orc_solver *s)
{
    if (!s) return;
    free(s->X); free(s->invDm); free(s->mass); free(s->vol); free(s->mu); free(s->lam);
    free(s->tets); free(s->active); free(s->colors); free(s->col_off); free(s->col_v);
    free(s->adj_off); free(s->adj_e); free(s->xprev); free(s->vprev); free(s->x); free(s->dxs);
    free(s);
}

/* src/vbd_solver.py:367-499.  err: 0 ok, -1 bad mesh index, -2 >256 colors,
 * -3 more than 32 elements at a vertex (the assert at :529), -4 singular rest tet. */
orc_solver *orc_create(int n, const double *X, int m, const int32_t *tets, const double *rho,
                       const double *E, const double *nu, const uint8_t *active, const double *grav,
                       double damping, int material, int expanded, int *err)
{
    *err = 0;
    for (int i = 0; i < 4 * m; ++i)
        if (tets[i] < 0 || tets[i] >= n) { *err = -1; return NULL; }
    orc_solver *s = calloc(1, sizeof(orc_solver));
    s->n = n; s->m = m; s->material = material; s->expanded = expanded; s->damping = damping;
    s->nthreads = 1;
    memcpy(s->grav, grav, 3 * sizeof(double));
    s->X = malloc(sizeof(double) * 3 * (size_t)n);
    memcpy(s->X, X, sizeof(double) * 3 * (size_t)n);
    s->tets = malloc(sizeof(int32_t) * 4 * (size_t)m);
    memcpy(s->tets, tets, sizeof(int32_t) * 4 * (size_t)m);
    s->active = malloc((size_t)n);
    if (active) memcpy(s->active, active, (size_t)n);
    else memset(s->active, 1, (size_t)n); /* :414-417 */
    s->mu = malloc(sizeof(double) * (size_t)m);
    s->lam = malloc(sizeof(double) * (size_t)m);
    orc_lame(m, E, nu, s->mu, s->lam);
    s->invDm = malloc(sizeof(double) * 9 * (size_t)m);
    s->mass = malloc(sizeof(double) * (size_t)m);
    s->vol = malloc(sizeof(double) * (size_t)m);
    if (orc_rest_shape(X, m, tets, rho, s->invDm, s->mass, s->vol)) { *err = -4; orc_destroy(s); return NULL; }
    /* coloring (:470-483) */
    s->colors = malloc(sizeof(int32_t) * (size_t)n);
    int nc = orc_color_mesh(n, m, tets, s->colors, NULL);
    if (nc < 0) { *err = -2; orc_destroy(s); return NULL; }
    s->ncolors = nc;
    s->col_off = calloc((size_t)nc + 1, sizeof(int));
    for (int v = 0; v < n; ++v)
        if (s->colors[v] >= 0) s->col_off[s->colors[v] + 1]++;
    for (int c = 0; c < nc; ++c) s->col_off[c + 1] += s->col_off[c];
    s->col_v = malloc(sizeof(int) * ((size_t)s->col_off[nc] + 1));
    {
        int *cur = malloc(sizeof(int) * ((size_t)nc + 1));
        memcpy(cur, s->col_off, sizeof(int) * ((size_t)nc + 1));
        for (int v = 0; v < n; ++v)
            if (s->colors[v] >= 0) s->col_v[cur[s->colors[v]]++] = v;
        free(cur);
    }
    /* adjacency vertex -> elements, ascending element index (:485-498) */
    s->adj_off = calloc((size_t)n + 1, sizeof(int));
    for (int i = 0; i < 4 * m; ++i) s->adj_off[tets[i] + 1]++;
    int W = 0;
    for (int v = 0; v < n; ++v) {
        if (s->adj_off[v + 1] > W) W = s->adj_off[v + 1];
        s->adj_off[v + 1] += s->adj_off[v];
    }
    s->W = W;
    if (W > ORC_MAX_ELEMENTS_PER_VERTEX) { *err = -3; orc_destroy(s); return NULL; }
    s->adj_e = malloc(sizeof(int) * ((size_t)4 * m + 1));
    {
        int *cur = malloc(sizeof(int) * ((size_t)n + 1));
        memcpy(cur, s->adj_off, sizeof(int) * ((size_t)n + 1));
        for (int e = 0; e < m; ++e)
            for (int k = 0; k < 4; ++k) s->adj_e[cur[tets[4 * e + k]]++] = e;
        free(cur);
    }
    s->xprev = malloc(sizeof(double) * 3 * (size_t)n);
    s->vprev = calloc(3 * (size_t)n, sizeof(double));
    s->x = malloc(sizeof(double) * 3 * (size_t)n);
    s->dxs = calloc(3 * (size_t)n, sizeof(double));
    memcpy(s->xprev, X, sizeof(double) * 3 * (size_t)n); /* :403-404 */
    return s;
}

void orc_set_threads(orc_solver *s, int nthreads) { s->nthreads = nthreads > 0 ? nthreads : 1; }
int orc_num_colors(const orc_solver *s) { return s->ncolors; }
int orc_max_incident(const orc_solver *s) { return s->W; }
void orc_get_colors(const orc_solver *s, int32_t *colors) { memcpy(colors, s->colors, sizeof(int32_t) * (size_t)s->n); }
void orc_get_state(const orc_solver *s, double *x, double *v)
{
    if (x) memcpy(x, s->xprev, sizeof(double) * 3 * (size_t)s->n);
    if (v) memcpy(v, s->vprev, sizeof(double) * 3 * (size_t)s->n);
}
void orc_set_state(orc_solver *s, const double *x, const double *v)
{
    if (x) memcpy(s->xprev, x, sizeof(double) * 3 * (size_t)s->n);
    if (v) memcpy(s->vprev, v, sizeof(double) * 3 * (size_t)s->n);
}
/* src/vbd_solver.py:501-506 */
void orc_reset(orc_solver *s)
{
    memcpy(s->xprev, s->X, sizeof(double) * 3 * (size_t)s->n);
    memset(s->vprev, 0, sizeof(double) * 3 * (size_t)s->n);
}

/* One vertex of one color: accumulate_grad_hess (:204-232) over its slots, then
 * solve_grad_hess (:250-274). */
static void update_vertex(orc_solver *s, int v, double dt)
{
    const int e0 = s->adj_off[v], e1 = s->adj_off[v + 1];
    double G[3] = {0, 0, 0};
    double H[9] = {0};
    for (int j = e0; j < e1; ++j) {
        int e = s->adj_e[j];
        const int32_t *t = s->tets + 4 * e;
        int vloc = 0;
        for (int k = 0; k < 4; ++k)
            if (t[k] == v) { vloc = k; break; }
        double xs[12];
        for (int k = 0; k < 4; ++k)
            for (int r = 0; r < 3; ++r) xs[3 * k + r] = s->x[3 * t[k] + r];
        double gi[3], ge[3];
        m33 Hi, He, iD;
        inertia_slot(s->x + 3 * v, s->xprev + 3 * v, s->vprev + 3 * v, s->mass[e], s->grav, dt, gi, Hi);
        memcpy(iD, s->invDm + 9 * e, sizeof(m33));
        if (s->expanded) elastic_slot_expanded(s->material, vloc, xs, iD, s->vol[e], s->mu[e], s->lam[e], ge, He);
        else elastic_slot_compact(s->material, vloc, xs, iD, s->vol[e], s->mu[e], s->lam[e], ge, He);
        double kd = s->damping;
        double dxv[3];
        for (int r = 0; r < 3; ++r) dxv[r] = s->x[3 * v + r] - s->xprev[3 * v + r];
        for (int r = 0; r < 3; ++r) {
            double gd = 0.0;
            for (int c = 0; c < 3; ++c) gd += kd * He[r][c] * dxv[c];
            gd /= dt;                                                 /* :222 */
            G[r] += gi[r] + ge[r] + gd;                                /* :226, summed in slot order */
            for (int c = 0; c < 3; ++c) H[3 * r + c] += Hi[r][c] + He[r][c] + kd * He[r][c] / dt; /* :223,227 */
        }
    }
    double dx[3] = {0, 0, 0};
    orc_solve3(H, G, dx);
    for (int r = 0; r < 3; ++r) {
        s->dxs[3 * v + r] = dx[r];       /* :271 */
        s->x[3 * v + r] += dx[r];        /* :274 */
    }
}

int orc_step_guess(orc_solver *s, const double *positions, const double *guess, double dt, double tol, int max_iter,
                   double *out_x, int *iters, double *metric);

/* src/vbd_solver.py:509-612 (VBDSolver.step).  positions: the x_n the caller
 * passes (NULL = the solver's own old_positions).  Writes the new positions to
 * out_x (may be NULL), the number of iterations executed (loop index + 1) to
 * *iters and the last iteration's max|dx_x+dx_y+dx_z| to *metric.
 * A fixed iteration count is tol = 0 (the strict '<' never holds). */
int orc_step(orc_solver *s, const double *positions, double dt, double tol, int max_iter,
             double *out_x, int *iters, double *metric)
{
    return orc_step_guess(s, positions, NULL, dt, tol, max_iter, out_x, iters, metric);
}

/* orc_step with the iterations started from `guess` (positions [n,3]) at the
 * active vertices instead of the inertia prediction (NULL: the inertia
 * prediction, as the reference).  The inertia term of the energy still uses
 * x_prev and v_prev.  No counterpart in the reference: it checks a solver
 * started from a warm guess. */
int orc_step_guess(orc_solver *s, const double *positions, const double *guess, double dt, double tol, int max_iter,
                   double *out_x, int *iters, double *metric)
{
    const int n = s->n;
    if (max_iter < 1 || !(dt > 0.0)) return -1;
    if (positions) memcpy(s->xprev, positions, sizeof(double) * 3 * (size_t)n); /* :535 */
    memcpy(s->x, s->xprev, sizeof(double) * 3 * (size_t)n);                      /* :536 */
    for (int v = 0; v < n; ++v) {                                                /* :277-292 */
        if (!s->active[v]) continue;
        for (int r = 0; r < 3; ++r)
            s->x[3 * v + r] = guess ? guess[3 * v + r]
                                    : s->xprev[3 * v + r] + s->vprev[3 * v + r] * dt + s->grav[r] * dt * dt;
    }
    memset(s->dxs, 0, sizeof(double) * 3 * (size_t)n); /* :545 */
    int it = 0;
    double mval = 0.0;
#ifdef _OPENMP
    omp_set_num_threads(s->nthreads);
#endif
    for (it = 0; it < max_iter; ++it) {                  /* :550 */
        for (int c = 0; c < s->ncolors; ++c) {          /* :552 */
            const int c0 = s->col_off[c], c1 = s->col_off[c + 1];
#ifdef _OPENMP
#pragma omp parallel for schedule(static) if (s->nthreads > 1)
#endif
            for (int i = c0; i < c1; ++i) {
                int v = s->col_v[i];
                if (!s->active[v]) continue;
                update_vertex(s, v, dt);
            }
        }
        /* :578 abs(dxs.sum(1)).max() < dx_tol */
        mval = 0.0;
        int isnan_seen = 0;
        for (int v = 0; v < n; ++v) {
            double a = fabs(s->dxs[3 * v] + s->dxs[3 * v + 1] + s->dxs[3 * v + 2]);
            if (a != a) isnan_seen = 1;
            else if (a > mval) mval = a;
        }
        if (isnan_seen) mval = NAN;
        if (mval < tol) break;
    }
    if (it == max_iter) it = max_iter - 1;
    if (iters) *iters = it + 1;
    if (metric) *metric = mval;
    for (int i = 0; i < 3 * n; ++i) {                   /* :609-610 */
        s->vprev[i] = (s->x[i] - s->xprev[i]) / dt;
        s->xprev[i] = s->x[i];
    }
    if (out_x) memcpy(out_x, s->x, sizeof(double) * 3 * (size_t)n);
    return 0;
}
