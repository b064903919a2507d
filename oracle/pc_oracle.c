/*
 * pc_oracle.c -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
 *
 * A plain-C FP64 restatement of gpucsl's PC-stable skeleton discovery.  It is
 * the checker for the HIP library: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product never calls it and has
 * no CPU fallback.
 *
 * Built with -ffp-contract=off so that every a*b+c is two IEEE roundings, the
 * same as the reference helpers compiled for the host and as the HIP FP64
 * recheck (which uses `#pragma clang fp contract(off)`).
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the gpucsl repository root).  Where this oracle deliberately deviates from the
 * reference the comment says so and why:
 *   - sepset choice: the reference keeps whichever thread wins a CAS
 *     (gpucsl/pc/cuda/helpers/graph_helpers.cu:81-101), which is racy.  Here the
 *     winner is the first passing test in pcalg order: ordered pair (a,b) with
 *     a<b first, ranks in lexicographic order, then (b,a).
 *   - Gaussian zmin counts exactly the tests with key <= winner key (the tests a
 *     sequential pcalg loop runs).  The reference also counts racy extra tests
 *     and tests of already removed pairs (gaussian_ci.cu:70-119).
 *   - discrete chi^2 is summed in one fixed order (64 strided partial sums and a
 *     fixed tree) instead of the reference's atomicAdd(double)
 *     (discrete_helpers.cu:181), whose order is not defined.
 *   - binomials use 64-bit integers; the reference's int overflows once
 *     C(n,k) * k exceeds 2^31.
 *
 * Parity pins (tests/): binomial, unranking, poz, pchisq and the FP64 Gaussian
 * CI test are compared bit-for-bit with the reference helpers compiled for the
 * host into oracle/_ref; compaction, unranking, binomial, pchisq, contingency,
 * marginal and chi^2 against the reference's own known-answer vectors; the
 * skeleton against a brute-force numpy restatement and scipy closed forms.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#pragma GCC diagnostic ignored "-Wmaybe-uninitialized"

#define OR_EXPORT __attribute__((visibility("default")))
#define OR_MAXL 32

/* ------------------------------------------------------------------------- */
/* combinatorics                                                              */
/* ------------------------------------------------------------------------- */

/* C(n, k) -- gpucsl/pc/cuda/helpers/math_helpers.cu:20-33, with 64-bit
 * intermediates (multiply then divide keeps every partial product integral). */
OR_EXPORT int64_t or_binomial(int64_t n, int64_t k) {
    if (n < k || k < 0) return 0;
    if (k > n - k) k = n - k;
    __int128 acc = 1;
    for (int64_t t = 0; t < k; ++t) {
        acc *= (n - t);
        acc /= (t + 1);
    }
    return (int64_t)acc;
}

/* Rank -> 1-based lexicographic l-combination of {1..N}.
 * gpucsl/pc/cuda/helpers/set_helpers.cu:25-49: each position walks upward from
 * the previous one while the count of combinations it skips stays <= rank. */
OR_EXPORT void or_unrank(int N, int l, int64_t rank, int *out) {
    int64_t skipped = 0;
    int prev = 0;
    for (int c = 0; c < l; ++c) {
        int v = prev;
        int64_t block = 0;
        while (skipped <= rank) {
            ++v;
            block = or_binomial(N - v, l - (c + 1));
            skipped += block;
        }
        skipped -= block;
        out[c] = v;
        prev = v;
    }
}

/* 1-based positions in i's neighbour list with v_j skipped -> variable ids.
 * gpucsl/pc/cuda/helpers/set_helpers.cu:51-57. */
static void or_resolve(const int *nbrs, const int *pos, int l, int vj, int *S) {
    for (int c = 0; c < l; ++c) {
        int idx = pos[c] - 1;
        if (nbrs[idx] >= vj) idx += 1;
        S[c] = nbrs[idx];
    }
}

/* ------------------------------------------------------------------------- */
/* normal / chi-square tail probabilities                                     */
/* ------------------------------------------------------------------------- */

/* Normal CDF after Ibbetson's polynomial as published by G. Perlman (z.c).
 * gpucsl/pc/cuda/helpers/math_helpers.cu:54-105.  The coefficients are those
 * of the published method and have to be the same for parity. */
OR_EXPORT double or_poz(double z) {
    double x;
    if (z == 0.0) {
        x = 0.0;
    } else {
        double y = 0.5 * fabs(z);
        if (y >= 3.0) { /* Z_MAX * 0.5 */
            x = 1.0;
        } else if (y < 1.0) {
            double w = y * y;
            x = 0.000124818987;
            x = x * w - 0.001075204047;
            x = x * w + 0.005198775019;
            x = x * w - 0.019198292004;
            x = x * w + 0.059054035642;
            x = x * w - 0.151968751364;
            x = x * w + 0.319152932694;
            x = x * w - 0.531923007300;
            x = x * w + 0.797884560593;
            x = x * y * 2.0;
        } else {
            y -= 2.0;
            x = -0.000045255659;
            x = x * y + 0.000152529290;
            x = x * y - 0.000019538132;
            x = x * y - 0.000676904986;
            x = x * y + 0.001390604284;
            x = x * y - 0.000794620820;
            x = x * y - 0.002034254874;
            x = x * y + 0.006549791214;
            x = x * y - 0.010557625006;
            x = x * y + 0.011630447319;
            x = x * y - 0.009279453341;
            x = x * y + 0.005353579108;
            x = x * y - 0.002141268741;
            x = x * y + 0.000535310849;
            x = x * y + 0.999936657524;
        }
    }
    return z > 0.0 ? ((x + 1.0) * 0.5) : ((1.0 - x) * 0.5);
}

static double or_ex(double v) { return v < -20.0 ? 0.0 : exp(v); } /* BIGX = 20 */

/* Upper tail of chi^2(df): Hill & Pike, CACM Algorithm 299, with the 1985
 * rounding remark, as published by G. Perlman (chisq.c).
 * gpucsl/pc/cuda/helpers/math_helpers.cu:128-175. */
OR_EXPORT double or_pchisq(double x, int df) {
    if (x <= 0.0 || df < 1) return 1.0;
    const double a = 0.5 * x;
    const int even = (2 * (df / 2)) == df;
    double y = 0.0;
    if (df > 1) y = or_ex(-a);
    double s = even ? y : (2.0 * or_poz(-sqrt(x)));
    if (df <= 2) return s;
    const double half_dfm1 = 0.5 * (df - 1.0);
    double z = even ? 1.0 : 0.5;
    if (a > 20.0) {
        double e = even ? 0.0 : 0.5723649429247000870717135; /* log(sqrt(pi)) */
        const double c = log(a);
        while (z <= half_dfm1) {
            e = log(z) + e;
            s += or_ex(c * z - a - e);
            z += 1.0;
        }
        return s;
    }
    double e = even ? 1.0 : (0.5641895835477562869480795 / sqrt(a)); /* 1/sqrt(pi) */
    double c = 0.0;
    while (z <= half_dfm1) {
        e = e * (a / z);
        c = c + e;
        z += 1.0;
    }
    return c * y + s;
}

/* ------------------------------------------------------------------------- */
/* Gaussian CI test, FP64, in the reference's operation order                 */
/* ------------------------------------------------------------------------- */

/* R[d1][d3] = A[d1][d2] * B[d2][d3]      (gaussian_helpers.cu:16-28) */
static void mm_nn(int d1, int d2, int d3, const double *A, const double *B, double *R) {
    for (int r = 0; r < d1; ++r)
        for (int c = 0; c < d3; ++c) {
            double acc = 0.0;
            for (int k = 0; k < d2; ++k) acc += A[r * d2 + k] * B[k * d3 + c];
            R[r * d3 + c] = acc;
        }
}
/* R[d1][d3] = A^T * B with A stored [d2][d1]   (gaussian_helpers.cu:30-43) */
static void mm_tn(int d1, int d2, int d3, const double *A, const double *B, double *R) {
    for (int r = 0; r < d1; ++r)
        for (int c = 0; c < d3; ++c) {
            double acc = 0.0;
            for (int k = 0; k < d2; ++k) acc += A[k * d1 + r] * B[k * d3 + c];
            R[r * d3 + c] = acc;
        }
}
/* R[d1][d3] = A * B^T with B stored [d3][d2]   (gaussian_helpers.cu:45-58) */
static void mm_nt(int d1, int d2, int d3, const double *A, const double *B, double *R) {
    for (int r = 0; r < d1; ++r)
        for (int c = 0; c < d3; ++c) {
            double acc = 0.0;
            for (int k = 0; k < d2; ++k) acc += A[r * d2 + k] * B[c * d2 + k];
            R[r * d3 + c] = acc;
        }
}

/* Gauss-Jordan inverse without pivoting; M is overwritten.
 * gaussian_helpers.cu:123-151. */
static void gj_inverse(int d, double *M, double *Inv) {
    for (int r = 0; r < d; ++r)
        for (int c = 0; c < d; ++c) Inv[r * d + c] = (r == c) ? 1.0 : 0.0;
    for (int piv = 0; piv < d; ++piv) {
        double f = M[piv * d + piv];
        for (int c = 0; c < d; ++c) {
            Inv[piv * d + c] = Inv[piv * d + c] / f;
            M[piv * d + c] = M[piv * d + c] / f;
        }
        for (int r = 0; r < d; ++r) {
            double g = M[r * d + piv];
            for (int c = 0; c < d; ++c) {
                if (r != piv) {
                    Inv[r * d + c] -= Inv[piv * d + c] * g;
                    M[r * d + c] -= M[piv * d + c] * g;
                }
            }
        }
    }
}

/* Lower Cholesky factor.  gaussian_helpers.cu:183-199. */
static void cholesky(int d, const double *A, double *L) {
    memset(L, 0, sizeof(double) * (size_t)d * (size_t)d);
    for (int r = 0; r < d; ++r)
        for (int c = 0; c <= r; ++c) {
            double acc = 0.0;
            for (int k = 0; k < c; ++k) acc += L[r * d + k] * L[c * d + k];
            if (r == c)
                L[r * d + c] = sqrt(A[r * d + r] - acc);
            else
                L[r * d + c] = (1.0 / L[c * d + c] * (A[r * d + c] - acc));
        }
}

/* Pseudo-inverse of M2 through the Cholesky factor of M2^T M2.
 * gaussian_helpers.cu:201-218 (cuPC Algorithm 7). */
static void m2_pinv(int l, const double *M2, double *Out) {
    double L[OR_MAXL * OR_MAXL], R[OR_MAXL * OR_MAXL], T[OR_MAXL * OR_MAXL];
    mm_tn(l, l, l, M2, M2, T);
    cholesky(l, T, L);
    mm_tn(l, l, l, L, L, T);
    gj_inverse(l, T, R);
    mm_nn(l, l, l, L, R, T);
    mm_nn(l, l, l, T, R, Out);
    mm_nt(l, l, l, Out, L, T);
    mm_nt(l, l, l, T, M2, Out);
}

/* Fisher-z of the partial correlation of (vi, vj) given S on a row-major
 * correlation matrix with leading dimension ld.
 * gaussian_helpers.cu:56-76 (fill_m0/m1/m2), :78-89 (compute_h),
 * :220-227 (Fisher transform) and :240-262 (gaussian_ci_test). */
OR_EXPORT double or_gauss_z(const double *C, int ld, int vi, int vj, const int *S, int l) {
    double M0[4], M1[2 * OR_MAXL], M2[OR_MAXL * OR_MAXL], M2i[OR_MAXL * OR_MAXL];
    double T[2 * OR_MAXL], H[4];
    M0[0] = 1.0;
    M0[1] = C[(size_t)vi * ld + vj];
    M0[2] = C[(size_t)vj * ld + vi];
    M0[3] = 1.0;
    for (int k = 0; k < l; ++k) {
        M1[k] = C[(size_t)vi * ld + S[k]];
        M1[l + k] = C[(size_t)vj * ld + S[k]];
    }
    for (int r = 0; r < l; ++r)
        for (int c = 0; c < l; ++c) M2[r * l + c] = C[(size_t)S[r] * ld + S[c]];
    m2_pinv(l, M2, M2i);
    mm_nt(l, l, 2, M2i, M1, T);
    mm_nn(2, l, 2, M1, T, H);
    for (int t = 0; t < 4; ++t) H[t] = M0[t] - H[t];
    double rho = H[1] / (sqrt(fabs(H[0])) * sqrt(fabs(H[3])));
    return fabs(0.5 * (log(fabs(1 + rho)) - log(fabs(1 - rho))));
}

/* Level-0 Fisher-z of one correlation.  gpucsl/pc/cuda/gaussian_ci.cu:36-37. */
OR_EXPORT double or_gauss_z0(double r) { return fabs(0.5 * log(fabs((1 + r) / (1 - r)))); }

/* ------------------------------------------------------------------------- */
/* correlation matrix (np.corrcoef restated), gpucsl/pc/helpers.py:58-59      */
/* ------------------------------------------------------------------------- */

/* X is n x p row-major (observations x variables), as the reference's data. */
OR_EXPORT void or_corrcoef(const double *X, int64_t n, int p, double *C) {
    double *mean = (double *)calloc((size_t)p, sizeof(double));
    double *Xc = (double *)malloc(sizeof(double) * (size_t)n * (size_t)p);
    for (int64_t k = 0; k < n; ++k)
        for (int v = 0; v < p; ++v) mean[v] += X[k * p + v];
    for (int v = 0; v < p; ++v) mean[v] /= (double)n;
    for (int64_t k = 0; k < n; ++k)
        for (int v = 0; v < p; ++v) Xc[(size_t)v * n + k] = X[k * p + v] - mean[v];
#pragma omp parallel for schedule(dynamic, 4)
    for (int a = 0; a < p; ++a)
        for (int b = a; b < p; ++b) {
            const double *xa = Xc + (size_t)a * n, *xb = Xc + (size_t)b * n;
            double acc = 0.0;
            for (int64_t k = 0; k < n; ++k) acc += xa[k] * xb[k];
            acc /= (double)(n - 1);
            C[(size_t)a * p + b] = acc;
            C[(size_t)b * p + a] = acc;
        }
    double *sd = (double *)malloc(sizeof(double) * (size_t)p);
    for (int v = 0; v < p; ++v) sd[v] = sqrt(C[(size_t)v * p + v]);
    for (int a = 0; a < p; ++a)
        for (int b = 0; b < p; ++b) {
            double c = C[(size_t)a * p + b] / sd[a];
            c = c / sd[b];
            if (c > 1.0) c = 1.0;
            if (c < -1.0) c = -1.0;
            C[(size_t)a * p + b] = c;
        }
    free(sd);
    free(Xc);
    free(mean);
}

/* ------------------------------------------------------------------------- */
/* graph helpers                                                              */
/* ------------------------------------------------------------------------- */

/* Row r of Gc becomes [deg, nbr0, nbr1, ...] with ascending neighbour ids and
 * the diagonal excluded.  gpucsl/pc/cuda/helpers/graph_helpers.cu:35-56. */
OR_EXPORT void or_compact(const uint16_t *G, int p, int32_t *Gc) {
    for (int r = 0; r < p; ++r) {
        int cnt = 0;
        for (int c = 0; c < p; ++c) Gc[(size_t)r * p + c] = 0;
        for (int c = 0; c < p; ++c)
            if (G[(size_t)r * p + c] && c != r) Gc[(size_t)r * p + 1 + cnt++] = c;
        Gc[(size_t)r * p] = cnt;
    }
}

static void nbr_lists(const uint8_t *G, int p, int *deg, int *nbr) {
    for (int r = 0; r < p; ++r) {
        int cnt = 0;
        for (int c = 0; c < p; ++c)
            if (G[(size_t)r * p + c] && c != r) nbr[(size_t)r * p + cnt++] = c;
        deg[r] = cnt;
    }
}

static void fmin_into(float *dst, double z) {
    float zf = (float)z;
    /* fminf semantics of write_atomic_minimum (graph_helpers.cu:123-133):
     * a NaN candidate never replaces the stored value. */
    if (!(zf != zf) && zf < *dst) *dst = zf;
}

/* ------------------------------------------------------------------------- */
/* Gaussian PC-stable skeleton                                                */
/* ------------------------------------------------------------------------- */

/* sqrt(n - l - 3), the scale of pcalg's z statistic for |S| = l */
static double level_scale(int64_t n, int l) {
    const double m = (double)(n - l - 3);
    return m > 0.0 ? sqrt(m) : 0.0;
}

/* Level 0: gpucsl/pc/cuda/gaussian_ci.cu:18-51 (upper-triangle pairs, zmin at
 * [min*p+max], removal on Z < thr, diagonal cleared).
 * Levels >= 1: gpucsl/pc/discover_skeleton_gaussian.py:30-113 and
 * gpucsl/pc/cuda/gaussian_ci.cu:66-120 with the canonical sepset rule above.
 *
 * G: p*p uint8 (out), sep: p*p*max_level int32 (out, -1 filled here),
 * zmin: p*p float (out), executed/total: per-level counters (out, length
 * max_level+1).  Returns the last level that ran. */
/* flags & 1 (pcalg conventions): zmin holds min over tests of
 * sqrt(n - |S| - 3) |z| (pcalg's pMax = 2 (1 - Phi(zmin))), and the sepset is
 * stored only for the ordered pair whose test removed the edge.  The caller
 * passes pcalg's thresholds |Phi^-1(alpha/2)| / sqrt(n - l - 3). */
OR_EXPORT int or_gauss_skeleton_ex(const double *C, int p, const double *thr, int max_level, int64_t n, int flags,
                                   uint8_t *G, int32_t *sep, float *zmin, int64_t *executed, int64_t *total) {
    const int L = max_level > 0 ? max_level : 1;
    const int pcalg = flags & 1;
    for (size_t t = 0; t < (size_t)p * p; ++t) zmin[t] = INFINITY;
    for (size_t t = 0; t < (size_t)p * p * (size_t)L; ++t) sep[t] = -1;
    for (int l = 0; l <= max_level; ++l) executed[l] = total[l] = 0;

    for (int a = 0; a < p; ++a) {
        G[(size_t)a * p + a] = 0;
        for (int b = a + 1; b < p; ++b) {
            double z = or_gauss_z0(C[(size_t)a * p + b]);
            fmin_into(&zmin[(size_t)a * p + b], pcalg ? z * level_scale(n, 0) : z);
            uint8_t keep = (z < thr[0]) ? 0 : 1;
            G[(size_t)a * p + b] = G[(size_t)b * p + a] = keep;
        }
    }
    executed[0] = total[0] = (int64_t)p * (p - 1) / 2;

    int last = 0;
    int *deg = (int *)malloc(sizeof(int) * (size_t)p);
    int *nbr = (int *)malloc(sizeof(int) * (size_t)p * p);
    for (int l = 1; l <= max_level; ++l) {
        nbr_lists(G, p, deg, nbr);
        int maxdeg = 0;
        int64_t tot = 0;
        for (int r = 0; r < p; ++r) {
            if (deg[r] > maxdeg) maxdeg = deg[r];
            tot += (int64_t)deg[r] * or_binomial(deg[r] - 1, l);
        }
        total[l] = tot;
        if (maxdeg - 1 < l) break;
        last = l;
        int64_t ex = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : ex)
        for (int a = 0; a < p; ++a) {
            int pos[OR_MAXL], S[OR_MAXL];
            for (int bi = 0; bi < deg[a]; ++bi) {
                const int b = nbr[(size_t)a * p + bi];
                if (b <= a) continue;
                float zl = INFINITY;
                int found = 0;
                for (int side = 0; side < 2 && !found; ++side) {
                    const int vi = side ? b : a, vj = side ? a : b;
                    const int N = deg[vi] - 1;
                    if (N < l) continue;
                    const int64_t cnt = or_binomial(N, l);
                    for (int64_t r = 0; r < cnt; ++r) {
                        or_unrank(N, l, r, pos);
                        or_resolve(nbr + (size_t)vi * p, pos, l, vj, S);
                        double z = or_gauss_z(C, p, vi, vj, S, l);
                        ++ex;
                        fmin_into(&zl, pcalg ? z * level_scale(n, l) : z);
                        if (z < thr[l]) {
                            found = 1;
                            for (int c = 0; c < l; ++c) {
                                sep[((size_t)vi * p + vj) * L + c] = S[c];
                                if (!pcalg) sep[((size_t)vj * p + vi) * L + c] = S[c];
                            }
                            break;
                        }
                    }
                }
                if (zl < zmin[(size_t)a * p + b]) zmin[(size_t)a * p + b] = zl;
                if (found) G[(size_t)a * p + b] = G[(size_t)b * p + a] = 2; /* removed this level */
            }
        }
        for (size_t t = 0; t < (size_t)p * p; ++t)
            if (G[t] == 2) G[t] = 0;
        executed[l] = ex;
    }
    free(nbr);
    free(deg);
    return last;
}

OR_EXPORT int or_gauss_skeleton(const double *C, int p, const double *thr, int max_level,
                                uint8_t *G, int32_t *sep, float *zmin, int64_t *executed,
                                int64_t *total) {
    return or_gauss_skeleton_ex(C, p, thr, max_level, 4, 0, G, sep, zmin, executed, total);
}

/* pmax = 2 (1 - Phi(sqrt(n-3) |zmin|)) = erfc(sqrt(n-3)|zmin| / sqrt 2), then
 * -1 on and below the diagonal.  gpucsl/pc/helpers.py:62-67 and :77-82. */
OR_EXPORT void or_pmax_from_zmin(const float *zmin, int p, int64_t n, float *pmax) {
    const double s = sqrt((double)(n - 3));
    for (int a = 0; a < p; ++a)
        for (int b = 0; b < p; ++b) {
            size_t t = (size_t)a * p + b;
            if (b <= a) {
                pmax[t] = -1.0f;
                continue;
            }
            double x = fabs(s * (double)zmin[t]);
            pmax[t] = (float)erfc(x / sqrt(2.0));
        }
}

/* ------------------------------------------------------------------------- */
/* discrete (Pearson chi^2) CI test                                           */
/* ------------------------------------------------------------------------- */

/* chi^2 over a table N[g][a][b] with its marginals.  Per-cell term:
 * gpucsl/pc/cuda/helpers/discrete_helpers.cu:164-186 (level n) and :140-155
 * (level 0, where N_++ is n).  Level 0 sums the cells sequentially in (a, b)
 * order, exactly as the reference's single-thread loop does.  Level n has no
 * defined order in the reference (atomicAdd(double)), so it is fixed here: 64
 * partial sums over cells c = lane, lane+64, ... in ascending order, then a
 * halving tree (lane + 32, 16, 8, 4, 2, 1). */
OR_EXPORT double or_chi2_table(const uint32_t *N, const uint32_t *Nis, const uint32_t *Njs,
                               const uint32_t *Ns, int dim_s, int di, int dj, int sequential) {
    double part[64];
    const int cells = dim_s * di * dj;
    if (sequential) {
        double acc = 0.0;
        for (int c = 0; c < cells; ++c) {
            const int g = c / (di * dj), a = (c / dj) % di, b = c % dj;
            const double expct = Nis[g * di + a] * (double)Njs[g * dj + b] / Ns[g];
            if (Ns[g] == 0 || expct == 0.0) continue;
            const double d = N[c] - expct;
            acc += d * d / expct;
        }
        return acc;
    }
    for (int lane = 0; lane < 64; ++lane) {
        double acc = 0.0;
        for (int c = lane; c < cells; c += 64) {
            const int g = c / (di * dj), a = (c / dj) % di, b = c % dj;
            const uint32_t obs = N[c];
            const double expct = Nis[g * di + a] * (double)Njs[g * dj + b] / Ns[g];
            if (Ns[g] == 0 || expct == 0.0) continue;
            const double d = obs - expct;
            acc += d * d / expct;
        }
        part[lane] = acc;
    }
    for (int off = 32; off >= 1; off >>= 1)
        for (int lane = 0; lane < off; ++lane) part[lane] = part[lane] + part[lane + off];
    return part[0];
}

/* Contingency table of (vi, vj | S) over column-major uint8 data:
 * gpucsl/pc/cuda/helpers/discrete_helpers.cu:40-85 (stratum index in mixed
 * radix, first sepset variable least significant), marginals :120-133. */
static double disc_test(const uint8_t *data, int64_t n, const uint8_t *dims, int vi, int vj,
                        const int *S, int l, uint32_t *work, int *df_out) {
    const int di = dims[vi], dj = dims[vj];
    int dim_s = 1;
    for (int c = 0; c < l; ++c) dim_s *= dims[S[c]];
    uint32_t *N = work;
    uint32_t *Nis = N + (size_t)dim_s * di * dj;
    uint32_t *Njs = Nis + (size_t)dim_s * di;
    uint32_t *Ns = Njs + (size_t)dim_s * dj;
    memset(N, 0, sizeof(uint32_t) * ((size_t)dim_s * di * dj + (size_t)dim_s * (di + dj + 1)));
    const uint8_t *ci = data + (size_t)vi * n, *cj = data + (size_t)vj * n;
    for (int64_t k = 0; k < n; ++k) {
        int g = 0, radix = 1;
        for (int c = 0; c < l; ++c) {
            g += data[(size_t)S[c] * n + k] * radix;
            radix *= dims[S[c]];
        }
        N[((size_t)g * di + ci[k]) * dj + cj[k]] += 1;
    }
    for (int g = 0; g < dim_s; ++g)
        for (int a = 0; a < di; ++a)
            for (int b = 0; b < dj; ++b) {
                uint32_t o = N[((size_t)g * di + a) * dj + b];
                Nis[g * di + a] += o;
                Njs[g * dj + b] += o;
                Ns[g] += o;
            }
    *df_out = (di - 1) * (dj - 1) * dim_s;
    return or_chi2_table(N, Nis, Njs, Ns, dim_s, di, dj, l == 0);
}

/* Exposed for the known-answer tests: table, marginals, chi^2 and df. */
OR_EXPORT double or_disc_chi2(const uint8_t *data, int64_t n, const uint8_t *dims, int vi, int vj,
                              const int *S, int l, uint32_t *work, int *df_out) {
    return disc_test(data, n, dims, vi, vj, S, l, work, df_out);
}

static void fmax_into(float *dst, float v) {
    /* fmaxf semantics of write_atomic_maximum (graph_helpers.cu:111-121). */
    if (!(v != v) && v > *dst) *dst = v;
}

/* Discrete PC-stable skeleton.
 * Level 0: gpucsl/pc/cuda/discrete_ci.cu:37-97 (pairs v_j < v_i, table of
 * (v_i, v_j), removal on p >= alpha with the double p value).
 * Levels >= 1: discrete_ci.cu:185-246.  The reference's early exit is a
 * `continue` (discrete_ci.cu:240-243), so every sepset of every pair adjacent
 * at level start is tested: pmax is the max over that full enumeration.  The
 * p value is stored to a float before the alpha comparison (discrete_ci.cu:213,
 * 219, 236).  The sepset is the first passing test in pcalg order.
 * data: column-major p x n uint8 codes; dims: p uint8.  Outputs as in
 * or_gauss_skeleton, with pmax (p*p float, 0 filled here) instead of zmin. */
OR_EXPORT int or_disc_skeleton(const uint8_t *data, int64_t n, int p, const uint8_t *dims,
                               double alpha, int max_level, uint8_t *G, int32_t *sep, float *pmax,
                               int64_t *executed, int64_t *total) {
    const int L = max_level > 0 ? max_level : 1;
    int maxdim = 1;
    for (int v = 0; v < p; ++v)
        if (dims[v] > maxdim) maxdim = dims[v];
    for (size_t t = 0; t < (size_t)p * p; ++t) pmax[t] = 0.0f;
    for (size_t t = 0; t < (size_t)p * p * (size_t)L; ++t) sep[t] = -1;
    for (int l = 0; l <= max_level; ++l) executed[l] = total[l] = 0;

#pragma omp parallel
    {
        uint32_t *work = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(maxdim * maxdim + 3 * maxdim + 1));
#pragma omp for schedule(dynamic, 1)
        for (int vi = 0; vi < p; ++vi) {
            G[(size_t)vi * p + vi] = 0;
            for (int vj = 0; vj < vi; ++vj) {
                int df;
                double chi2 = disc_test(data, n, dims, vi, vj, NULL, 0, work, &df);
                double pv = or_pchisq(chi2, df);
                fmax_into(&pmax[(size_t)vj * p + vi], (float)pv);
                uint8_t keep = (pv >= alpha) ? 0 : 1;
                G[(size_t)vi * p + vj] = G[(size_t)vj * p + vi] = keep;
            }
        }
        free(work);
    }
    executed[0] = total[0] = (int64_t)p * (p - 1) / 2;

    int last = 0;
    int *deg = (int *)malloc(sizeof(int) * (size_t)p);
    int *nbr = (int *)malloc(sizeof(int) * (size_t)p * p);
    for (int l = 1; l <= max_level; ++l) {
        nbr_lists(G, p, deg, nbr);
        int maxdeg = 0;
        int64_t tot = 0;
        for (int r = 0; r < p; ++r) {
            if (deg[r] > maxdeg) maxdeg = deg[r];
            tot += (int64_t)deg[r] * or_binomial(deg[r] - 1, l);
        }
        total[l] = tot;
        if (maxdeg - 1 < l) break;
        last = l;
        size_t cells = 1;
        for (int t = 0; t < l + 2; ++t) cells *= (size_t)maxdim;
        size_t wsz = cells + (size_t)3 * cells / (maxdim > 1 ? maxdim : 1) + 8;
#pragma omp parallel
        {
            uint32_t *work = (uint32_t *)malloc(sizeof(uint32_t) * wsz);
            int pos[OR_MAXL], S[OR_MAXL];
#pragma omp for schedule(dynamic, 1)
            for (int a = 0; a < p; ++a) {
                for (int bi = 0; bi < deg[a]; ++bi) {
                    const int b = nbr[(size_t)a * p + bi];
                    if (b <= a) continue;
                    float pm = pmax[(size_t)a * p + b];
                    int found = 0;
                    for (int side = 0; side < 2; ++side) {
                        const int vi = side ? b : a, vj = side ? a : b;
                        const int Nn = deg[vi] - 1;
                        if (Nn < l) continue;
                        const int64_t cnt = or_binomial(Nn, l);
                        for (int64_t r = 0; r < cnt; ++r) {
                            or_unrank(Nn, l, r, pos);
                            or_resolve(nbr + (size_t)vi * p, pos, l, vj, S);
                            int df;
                            double chi2 = disc_test(data, n, dims, vi, vj, S, l, work, &df);
                            float pf = (float)or_pchisq(chi2, df);
                            fmax_into(&pm, pf);
                            if (!found && (double)pf >= alpha) {
                                found = 1;
                                for (int c = 0; c < l; ++c) {
                                    sep[((size_t)a * p + b) * L + c] = S[c];
                                    sep[((size_t)b * p + a) * L + c] = S[c];
                                }
                            }
                        }
                    }
                    pmax[(size_t)a * p + b] = pm;
                    if (found) G[(size_t)a * p + b] = G[(size_t)b * p + a] = 2;
                }
            }
            free(work);
        }
        for (size_t t = 0; t < (size_t)p * p; ++t)
            if (G[t] == 2) G[t] = 0;
        executed[l] = tot;
    }
    free(nbr);
    free(deg);
    return last;
}

OR_EXPORT int or_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* threads of the following calls (timing runs); 0 restores OpenMP's default */
OR_EXPORT void or_set_threads(int k) {
#ifdef _OPENMP
    static int dflt = 0;
    if (!dflt) dflt = omp_get_max_threads();
    omp_set_num_threads(k > 0 ? k : dflt);
#else
    (void)k;
#endif
}
