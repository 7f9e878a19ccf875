/* C restatement of the reference's CPU oracle -- TEST INFRASTRUCTURE (checker and CPU
 * baseline only; never linked into the product library).
 *
 * Algorithm: the reference's src/gpusw/reference.py:38-125 (_score_codes), int64
 * arithmetic, sentinel -(1<<30) (reference.py:27).  Rows = reference residues, columns =
 * query residues, sub[a*ns+b] scores reference code a against query code b.
 * Batch entry points score pair lists / cross products with OpenMP over pairs.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORACLE_NEG (-(INT64_C(1) << 30))

static inline int64_t max64(int64_t a, int64_t b) { return a > b ? a : b; }

int64_t oracle_score(const uint8_t *q, int n, const uint8_t *r, int m,
                     const int64_t *sub, int ns, int mode, int64_t go, int64_t ge,
                     int free_q, int free_r, int64_t *work /* 2*(n+1) or NULL */)
{
    const int local = mode == 0, global = mode == 1, semi = mode == 2;
    if (n == 0 || m == 0) {                               /* reference.py:49-54 */
        if (local || (n == 0 && m == 0)) return 0;
        if (n == 0) return (semi && free_q) ? 0 : go + (int64_t)(m - 1) * ge;
        return (semi && free_r) ? 0 : go + (int64_t)(n - 1) * ge;
    }
    int64_t *own = NULL;
    if (!work) { own = (int64_t *)malloc(sizeof(int64_t) * 2 * (size_t)(n + 1)); work = own; }
    int64_t *hp = work, *f = work + (n + 1);
    hp[0] = 0;
    for (int j = 1; j <= n; ++j) {                        /* reference.py:56-65 */
        int64_t lad = go + (int64_t)(j - 1) * ge;
        hp[j] = local ? 0 : (global ? lad : (free_r ? 0 : lad));
        f[j] = ORACLE_NEG;
    }
    f[0] = ORACLE_NEG;
    int64_t best = 0, corner = ORACLE_NEG, endcol = ORACLE_NEG, endrow = ORACLE_NEG;
    for (int i = 1; i <= m; ++i) {                        /* reference.py:70-112 */
        const int64_t *srow = sub + (size_t)r[i - 1] * ns;
        int64_t lad = go + (int64_t)(i - 1) * ge;
        int64_t hleft = local ? 0 : (global ? lad : (free_q ? 0 : lad));
        int64_t diag = hp[0];
        hp[0] = hleft;
        int64_t eleft = ORACLE_NEG;
        for (int j = 1; j <= n; ++j) {
            int64_t fj = max64(hp[j] + go, f[j] + ge);
            int64_t ej = max64(hleft + go, eleft + ge);
            int64_t h = max64(max64(diag + srow[q[j - 1]], ej), fj);
            if (local && h < 0) h = 0;
            if (h > best) best = h;
            if (!local) {
                if (i == m && j == n) corner = h;
                if (free_q && j == n && h > endcol) endcol = h;
                if (free_r && i == m && h > endrow) endrow = h;
            }
            diag = hp[j];
            hp[j] = h;
            f[j] = fj;
            hleft = h;
            eleft = ej;
        }
    }
    free(own);
    if (local) return best;                               /* reference.py:114-125 */
    if (global) return corner;
    int64_t res = corner;
    if (free_q) res = max64(res, endcol);
    if (free_r) res = max64(res, endrow);
    if (free_q && free_r) res = max64(res, 0);
    return res;
}

/* Score explicit pairs (qi[k], rj[k]) of two encoded sets. */
void oracle_score_pairs(const uint8_t *qcodes, const int32_t *qoff,
                        const uint8_t *rcodes, const int32_t *roff,
                        const int32_t *qi, const int32_t *rj, int64_t npairs,
                        const int64_t *sub, int ns, int mode, int64_t go, int64_t ge,
                        int free_q, int free_r, int64_t *out, int nthreads)
{
    (void)nthreads;
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1)
    {
        int64_t *work = NULL;
        int cap = -1;
#pragma omp for schedule(dynamic, 16)
        for (int64_t k = 0; k < npairs; ++k) {
            int a = qi[k], b = rj[k];
            int n = qoff[a + 1] - qoff[a];
            if (n > cap) { free(work); cap = n; work = (int64_t *)malloc(sizeof(int64_t) * 2 * (size_t)(cap + 1)); }
            out[k] = oracle_score(qcodes + qoff[a], n, rcodes + roff[b], roff[b + 1] - roff[b],
                                  sub, ns, mode, go, ge, free_q, free_r, work);
        }
        free(work);
    }
}

/* Score qlist[0..nq) x all nr references into out[nq*nr] (row-major). */
void oracle_score_cross(const uint8_t *qcodes, const int32_t *qoff,
                        const uint8_t *rcodes, const int32_t *roff,
                        const int32_t *qlist, int nq, int nr,
                        const int64_t *sub, int ns, int mode, int64_t go, int64_t ge,
                        int free_q, int free_r, int64_t *out, int nthreads)
{
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1)
    {
        int64_t *work = NULL;
        int cap = -1;
#pragma omp for schedule(dynamic, 16) collapse(2)
        for (int a = 0; a < nq; ++a) {
            for (int b = 0; b < nr; ++b) {
                int qa = qlist[a];
                int n = qoff[qa + 1] - qoff[qa];
                if (n > cap) { free(work); cap = n; work = (int64_t *)malloc(sizeof(int64_t) * 2 * (size_t)(cap + 1)); }
                out[(int64_t)a * nr + b] = oracle_score(qcodes + qoff[qa], n, rcodes + roff[b],
                                                        roff[b + 1] - roff[b], sub, ns, mode, go, ge,
                                                        free_q, free_r, work);
            }
        }
        free(work);
    }
}
