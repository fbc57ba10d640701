// Wavefront-level building blocks of the Cholesky kernels: a lane broadcast
// and the factorisation of a small diagonal block held in LDS by one
// wavefront.
#pragma once

#include <hip/hip_runtime.h>

namespace rmh {

// value of v in lane `lane` (wave-uniform lane index)
__device__ __forceinline__ double lane_bcast(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Right-looking Cholesky of the NB x NB block L (row-major in LDS, ld LD),
// run by the lanes 0..NB-1 of one wavefront (the caller guards lane < NB):
// lane i keeps row i in registers and pivot-column entries move lane to lane
// by readlane.  The block must be padded beyond its true order with the
// identity (the factor of diag(A, I) is diag(L, I)), so all NB steps run
// unconditionally.  On return L holds the factor in its lower triangle
// (garbage above), rdiag[i] = 1 / L_ii, and the result is the first
// non-positive (or NaN) pivot, -1 if none; after such a pivot the remaining
// steps compute garbage.  Pivots use the hardware 1/sqrt estimate with two
// Newton steps, y <- y (3 - d y^2) / 2, and sqrt(d) = d y corrected once by
// s <- s + (d - s^2) y / 2.
template <int NB, int LD>
__device__ __forceinline__ int wave_potrf(double* L, double* rdiag, int lane) {
    int bad = -1;  // wave-uniform
    double r[NB], piv = 1.0;
#pragma unroll
    for (int c = 0; c < NB; ++c) r[c] = L[lane * LD + c];
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        const double d = lane_bcast(r[j], j);
        bad = (bad < 0 && !(d > 0.0)) ? j : bad;
        double y = __builtin_amdgcn_rsq(d);
        const double hd = 0.5 * d;
        y = fma(y, fma(-hd * y, y, 0.5), y);
        y = fma(y, fma(-hd * y, y, 0.5), y);
        double sj = d * y;
        sj = fma(fma(-sj, sj, d), 0.5 * y, sj);
        r[j] = (lane == j) ? sj : r[j] * y;
        piv = (lane == j) ? sj : piv;
#pragma unroll
        for (int c = j + 1; c < NB; ++c) {
            r[c] = fma(-r[j], lane_bcast(r[j], c), r[c]);
            // keep the scheduler from hoisting many broadcasts at once: they
            // would pile up in scalar registers and spill
            if ((c & 7) == 7) __builtin_amdgcn_sched_barrier(0);
        }
        __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int c = 0; c < NB; ++c) L[lane * LD + c] = r[c];
    rdiag[lane] = 1.0 / piv;
    return bad;
}

// Right-looking Cholesky of an NB x NB diagonal block (NB <= 32) that also
// solves the rows below it: lanes 0..NB-1 hold the rows of the block (padded
// with the identity, as for wave_potrf), every other lane holds one row a of
// the block column below it.  The elimination applied to the block rows is
// applied to those rows too, so on return lanes i < NB hold row i of the
// factor L and the other lanes hold x with x L^T = a.  Returns the first
// non-positive (or NaN) pivot, -1 if none.
//
// Step j: the pivot (from lane j) and its 1/sqrt, the scaled column j, the
// next column j+1 updated at once (its entries move by readlane), and the
// other columns updated one step late: step j-1's pivot column is read back
// from LDS (col: 2 x NB doubles, one buffer per parity of j) at the start of
// step j, so those loads and FMAs fill the latency of step j's pivot chain.
template <int NB>
__device__ __forceinline__ int wave_chol_rows(double (&r)[NB], double* col, int lane) {
    static_assert(NB <= 32, "block rows and solved rows share the wavefront");
    int bad = -1;    // wave-uniform
    double pj = 0.0;  // this lane's entry of the previous pivot column
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        // step j-1's updates of columns j+1.. (column j took it at step j-1)
        if (j >= 1) {
            const double* cprev = col + ((j - 1) & 1) * NB;
#pragma unroll
            for (int c = j + 1; c < NB; ++c) r[c] = fma(-pj, cprev[c], r[c]);
        }
        const double d = lane_bcast(r[j], j);
        bad = (bad < 0 && !(d > 0.0)) ? j : bad;
        double y = __builtin_amdgcn_rsq(d);
        const double hd = 0.5 * d;
        y = fma(y, fma(-hd * y, y, 0.5), y);
        y = fma(y, fma(-hd * y, y, 0.5), y);
        double sj = d * y;
        sj = fma(fma(-sj, sj, d), 0.5 * y, sj);
        r[j] = (lane == j) ? sj : r[j] * y;
        pj = r[j];
        if (j + 1 < NB) {
            if (j + 2 < NB && lane < NB) col[(j & 1) * NB + lane] = r[j];
            r[j + 1] = fma(-r[j], lane_bcast(r[j], j + 1), r[j + 1]);
            // the next step reads what the other lanes wrote (one
            // wavefront, LDS operations in program order)
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
    return bad;
}

}  // namespace rmh
