/*
 * oracle_batch.c -- TEST INFRASTRUCTURE ONLY.
 *
 * OpenMP driver over column blocks for the CPU restatement, used as the
 * same-box CPU baseline.  Columns are independent in the reference
 * (uwshcu.F90:1218-4551 is a loop over i with no inter-column coupling), so
 * each block is an independent call with base pointers offset by the block's
 * first column and the chunk's leading dimension kept.
 */
#include <stddef.h>
#include "uwshcu_oracle.h"
#ifdef _OPENMP
#include <omp.h>
#endif

int orc_compute_uwshcu_inv_omp(const orc_consts *c, int nthreads, int mix, int mkx, int iend, int ncnst,
    double dt, const double *ps0, const double *zs0, const double *p0, const double *z0, const double *dp0,
    const double *u0, const double *v0, const double *qv0, const double *ql0, const double *qi0,
    const double *t0, const double *s0, const double *tr0, const double *tke, const double *cldfrct,
    const double *concldfrct, const double *pblh, double *cush, double *umf, double *slflx, double *qtflx,
    double *flxprc1, double *flxsnow1, double *qvten, double *qlten, double *qiten, double *sten,
    double *uten, double *vten, double *trten, double *qrten, double *qsten, double *precip, double *snow,
    double *evapc, double *cufrc, double *qcu, double *qlu, double *qiu, double *cbmf, double *qc,
    double *rliq, double *cnt, double *cnb, const double *dpdry0, double *wu, int *diag, int *bad_col)
{
    const int blk = 32;
    const int nblk = (iend + blk - 1) / blk;
    int rc = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < nblk; ++b) {
        const size_t o = (size_t)b * blk;
        const int n = (iend - (int)o) < blk ? (iend - (int)o) : blk;
        int bc = 0;
        int r = orc_compute_uwshcu_inv(c, mix, mkx, n, ncnst, dt, ps0 + o, zs0 + o, p0 + o, z0 + o, dp0 + o,
            u0 + o, v0 + o, qv0 + o, ql0 + o, qi0 + o, t0 + o, s0 + o, tr0 + o, tke + o, cldfrct + o,
            concldfrct + o, pblh + o, cush + o, umf + o, slflx + o, qtflx + o, flxprc1 + o, flxsnow1 + o,
            qvten + o, qlten + o, qiten + o, sten + o, uten + o, vten + o, trten + o, qrten + o, qsten + o,
            precip + o, snow + o, evapc + o, cufrc + o, qcu + o, qlu + o, qiu + o, cbmf + o, qc + o, rliq + o,
            cnt + o, cnb + o, dpdry0 + o, wu + o, diag ? diag + o * ORC_NDIAG : NULL, &bc);
        if (r != 0) {
#pragma omp critical
            {
                if (rc == 0) { rc = r; if (bad_col) *bad_col = (int)o + bc; }
            }
        }
    }
    return rc;
}
