// Self-consistent charge iterations of GFN2-xTB, one workgroup per system,
// all iterations inside one launch (data-dependent iteration counts never
// leave the device).
//
// Per iteration: potentials from the mixed moments -> Fock matrix built from
// row-local potentials G (F = H0 - (G + G^T)/2, coalesced reads of the
// transposed moment integrals) -> A = C^T F C in the current orbital basis C
// (S-orthonormal) -> parallel cyclic threshold Jacobi on A with the rotations
// applied straight to the columns of C -> Fermi occupations (one wavefront,
// safeguarded Newton from the previous chemical potential) ->
// P = C diag(2f) C^T -> Mulliken shell charges, atomic dipole/quadrupole
// moments and the band energy as column sweeps -> energy -> modified Broyden
// mixing (Johnson, PRB 38, 12807 (1988)).
// The Jacobi tolerance follows the SCC: 1e-3 of the previous iteration's RMS
// moment change, never looser than 1e-6 and never tighter than the requested
// floor, so the converged iterate is diagonalised to the floor.
// Systems with at most opt.lds_max_m (<= kFastM) basis functions keep A, C,
// F and P in LDS; larger ones use per-system global scratch.  C starts as X U with X = L^-T
// and U the previous call's orthogonal basis (identity on a cold start).
#pragma once
#include "k_setup.h"

namespace qx {

constexpr int kNT = QX_SCC_BLOCK;  // threads of the SCC kernel
constexpr int kMaxNx = 128 + 9 * QX_MAX_NAT;
constexpr int kMaxRef = QX_MAX_REF * QX_MAX_NAT;
constexpr int kFastM = 56;
constexpr int kRs = 11;  // row sums per AO: S, D (3), Q (6), H0

// Opt-in per-phase cycle counters of the SCC kernel (build with
// -DQX_PHASE_TIMING); compiled out otherwise.
#ifdef QX_PHASE_TIMING
constexpr int kPhaseSys = 1024, kPhaseSlots = 32;
__device__ unsigned long long g_phase[kPhaseSys * kPhaseSlots];
#define QX_PH_INIT unsigned long long ph_t = clock64();
#define QX_PH(k)                                                                   \
    do {                                                                           \
        __syncthreads();                                                           \
        if (threadIdx.x == 0 && blockIdx.x < kPhaseSys) {                          \
            const unsigned long long t_ = clock64();                               \
            g_phase[blockIdx.x * kPhaseSlots + (k)] += t_ - ph_t;                  \
            ph_t = t_;                                                             \
        }                                                                          \
    } while (0)
#define QX_PH_ADD(k, v)                                                            \
    do {                                                                           \
        if (threadIdx.x == 0 && blockIdx.x < kPhaseSys) g_phase[blockIdx.x * kPhaseSlots + (k)] += (v); \
    } while (0)
#define QX_PT_INIT(tv) unsigned long long tv = clock64(), tv##_acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#define QX_PT(k, tv)                                                               \
    do {                                                                           \
        const unsigned long long t_ = clock64();                                   \
        tv##_acc[(k) & 7] += t_ - tv;                                              \
        tv = t_;                                                                   \
    } while (0)
#define QX_PT_COUNT(tv, k) (tv##_acc[(k)&7] += 1)
#define QX_PT_FLUSH(tv)                                                            \
    do {                                                                           \
        if (threadIdx.x == 0 && blockIdx.x < kPhaseSys)                            \
            for (int k_ = 0; k_ < 8; ++k_) g_phase[blockIdx.x * kPhaseSlots + 16 + k_] += tv##_acc[k_]; \
    } while (0)
#else
#define QX_PT_INIT(tv)
#define QX_PT(k, tv) \
    do {             \
    } while (0)
#define QX_PT_FLUSH(tv) \
    do {                \
    } while (0)
#define QX_PT_COUNT(tv, k) \
    do {                   \
    } while (0)
#define QX_PH_INIT
#define QX_PH(k) \
    do {         \
    } while (0)
#define QX_PH_ADD(k, v) \
    do {                \
    } while (0)
#endif

// C[i][j] = sum_k opA[i][k] * d[k] * opB[k][j] on the FP64 matrix cores
// (v_mfma_f64_16x16x4f64).  A wavefront computes a strip of two 16 x 16
// tiles sharing the A operand; operand lanes: A[i0 + l%16][k0 + l/16],
// B[k0 + l/16][j0 + l%16]; accumulator element r of lane l is
// C[i0 + 4r + l/16][j0 + l%16].  Out-of-range rows, columns and k are zero.
typedef double qx_d4 __attribute__((ext_vector_type(4)));
template <bool TA, bool TB>
QX_DEV void mfma_gemm(int M, int N, int K, const double* A, int lda, const double* Bm, int ldb,
                      double* Cm, int ldc, const double* d)
{
    constexpr int NU = 2;  // 16 x 16 tiles per strip (more tiles cost the SCC body registers)
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int tm = (M + 15) >> 4, tc = (N + 16 * NU - 1) / (16 * NU);
    for (int t = wave; t < tm * tc; t += kNT / 64) {
        const int i0 = (t / tc) * 16, jc = (t % tc) * 16 * NU;
        const int i = i0 + li;
        const bool iin = i < M;
        const int ic = iin ? i : M - 1;
        int jcl[NU];
        bool jin[NU];
#pragma unroll
        for (int u = 0; u < NU; ++u) {
            const int j = jc + 16 * u + li;
            jin[u] = j < N;
            jcl[u] = jin[u] ? j : N - 1;
        }
        qx_d4 acc[NU];
#pragma unroll
        for (int u = 0; u < NU; ++u) acc[u] = qx_d4{0.0, 0.0, 0.0, 0.0};
        // four k-steps per pass: all operand loads first, then the MFMAs
        // (k-steps past K multiply zeros).  Loads are clamped into the
        // matrices and out-of-range lanes zeroed afterwards, so no branch
        // surrounds a load.
        constexpr int KS = 4;
        for (int k0 = 0; k0 < K; k0 += 4 * KS) {
            double a[KS], b[KS][NU];
#pragma unroll
            for (int q = 0; q < KS; ++q) {
                const int k = k0 + 4 * q + lk;
                const bool kin = k < K;
                const int kc = kin ? k : K - 1;
                double av = TA ? A[(size_t)kc * lda + ic] : A[(size_t)ic * lda + kc];
                if (d) av *= d[kc];
                a[q] = (iin && kin) ? av : 0.0;
#pragma unroll
                for (int u = 0; u < NU; ++u) {
                    const double bv = TB ? Bm[(size_t)jcl[u] * ldb + kc] : Bm[(size_t)kc * ldb + jcl[u]];
                    b[q][u] = (jin[u] && kin) ? bv : 0.0;
                }
            }
#pragma unroll
            for (int q = 0; q < KS; ++q)
#pragma unroll
                for (int u = 0; u < NU; ++u)
                    acc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[q], b[q][u], acc[u], 0, 0, 0);
        }
#pragma unroll
        for (int u = 0; u < NU; ++u)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = i0 + 4 * r + lk, col = jc + 16 * u + li;
                if (row < M && col < N) Cm[(size_t)row * ldc + col] = acc[u][r];
            }
    }
}

QX_DEV double block_max(double v, double* smem)
{
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    __syncthreads();
    if (l == 0) smem[w] = v;
    __syncthreads();
    double s = smem[0];
    for (int i = 1; i < kNT / 64; ++i) s = fmax(s, smem[i]);
    __syncthreads();
    return s;
}

// block-wide sums of K values at once (all threads call)
template <int K>
QX_DEV void block_sums(double* v, double* smem)
{
#pragma unroll
    for (int k = 0; k < K; ++k)
        for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_xor(v[k], o, 64);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    __syncthreads();
    if (l == 0)
        for (int k = 0; k < K; ++k) smem[w * K + k] = v[k];
    __syncthreads();
    for (int k = 0; k < K; ++k) {
        double s = 0.0;
        for (int i = 0; i < kNT / 64; ++i) s += smem[i * K + k];
        v[k] = s;
    }
    __syncthreads();
}

// round-robin pairing: round r, slot k of m players (m even, r < m - 1)
QX_DEV void rr_pair(int m, int r, int k, int& p, int& q)
{
    if (k == 0) {
        p = r;
        q = m - 1;
    } else {
        p = r + k;
        if (p >= m - 1) p -= m - 1;
        q = r - k;
        if (q < 0) q += m - 1;
    }
}

// largest off-diagonal magnitude of the leading m x m block of A
QX_DEV double offdiag_max(const double* A, int m, int lda, double* red)
{
    double mx = 0.0;
    for (int i = threadIdx.x >> 6; i < m; i += kNT / 64)
        for (int j = threadIdx.x & 63; j < m; j += 64)
            if (i != j) mx = fmax(mx, fabs(A[(size_t)i * lda + j]));
    return block_max(mx, red);
}

// pair-ordered 2 x 2 block index b -> (P, R) with R <= P
QX_DEV void tri_index(int b, int& P, int& R)
{
    P = (int)((sqrtf(8.0f * (float)b + 1.0f) - 1.0f) * 0.5f);
    while (P * (P + 1) / 2 > b) --P;
    while ((P + 1) * (P + 2) / 2 <= b) ++P;
    R = b - P * (P + 1) / 2;
}

// Jacobi rotation (c, s) that zeroes a_pq:
//   r = sqrt(d^2 + 4 a^2), d = a_qq - a_pp, u = r + |d|,
//   c = sqrt(u / 2r), s = 2 a sign(d) c / u   (c^2 + s^2 = 1 exactly)
// i.e. t = s / c = 2 a sign(d) / (|d| + r), the smaller root of
// t^2 + t d / a - 1 = 0.  Three dependent long-latency operations.
QX_DEV double2 jacobi_rotation(double app, double aqq, double apq)
{
    const double d = aqq - app, ad = fabs(d);
    const double r = sqrt(d * d + 4.0 * apq * apq);
    const double u = r + ad;
    const double ir = 1.0 / r, iu = 1.0 / u;
    const double c = sqrt(0.5 * u * ir);
    const double s = 2.0 * (d >= 0.0 ? apq : -apq) * c * iu;
    return make_double2(c, s);
}

// Rotated entries of the 2 x 2 blocks (P, R) and (R, P): input b00 = A[p][r],
// b01 = A[p][s], b10 = A[q][r], b11 = A[q][s] with (p, q) the pair of P and
// (r, s) the pair of R; on the diagonal block (P == R) the off-diagonal
// entries become exact zeros.
QX_DEV void rotate_block(double2 rp, double2 rr, bool diag, const double b[4], double n[4])
{
    const double cp = rp.x, sp = rp.y, cr = rr.x, sr = rr.y;
    const double m00 = b[0] * cr - b[1] * sr, m01 = b[0] * sr + b[1] * cr;
    const double m10 = b[2] * cr - b[3] * sr, m11 = b[2] * sr + b[3] * cr;
    n[0] = cp * m00 - sp * m10;
    n[1] = diag ? 0.0 : cp * m01 - sp * m11;
    n[2] = diag ? 0.0 : sp * m00 + cp * m10;
    n[3] = sp * m01 + cp * m11;
}

QX_DEV void store_block(double* A, int lda, int2 ip, int2 ir, bool diag, const double n[4])
{
    A[(size_t)ip.x * lda + ir.x] = n[0];
    A[(size_t)ip.x * lda + ir.y] = n[1];
    A[(size_t)ip.y * lda + ir.x] = n[2];
    A[(size_t)ip.y * lda + ir.y] = n[3];
    if (!diag) {
        A[(size_t)ir.x * lda + ip.x] = n[0];
        A[(size_t)ir.y * lda + ip.x] = n[1];
        A[(size_t)ir.x * lda + ip.y] = n[2];
        A[(size_t)ir.y * lda + ip.y] = n[3];
    }
}

QX_DEV void load_block(const double* A, int lda, int2 ip, int2 ir, double b[4])
{
    b[0] = A[(size_t)ip.x * lda + ir.x];
    b[1] = A[(size_t)ip.x * lda + ir.y];
    b[2] = A[(size_t)ip.y * lda + ir.x];
    b[3] = A[(size_t)ip.y * lda + ir.y];
}

// Parallel cyclic threshold Jacobi on symmetric A (m x m, m even, leading
// dimension lda).  Pairs with |a_pq| < tol, or whose rotation has |sin| <=
// min_s, are not rotated; a sweep ends the
// loop when no off-diagonal element of the result reaches tol.  Each rotation
// is also applied to columns p, q of the nrow x m matrix Cv (leading
// dimension lda), so Cv ends as Cv V with A = V^T A_in V diagonal.
// One round = rotation parameters (one thread per pair) | barrier | block and
// column updates | barrier; rounds without any rotation skip the updates.
// A thread's first two blocks (P, R) and its column pair are the same in
// every round, so their parameters and entries are fetched in two LDS round
// trips (parameters, then data) before any store.
QX_DEV int jacobi(double* A, double* Cv, int m, int nrow, int lda, double tol, int maxsweep,
                  double2* rot, int2* rpq, double* red, int* s_any, double min_s = 0.0)
{
    const int half = m / 2;
    if (offdiag_max(A, m, lda, red) < tol) return 0;
    if (threadIdx.x == 0) {
        s_any[0] = 0;
        s_any[1] = 0;
    }
    __syncthreads();
    const int nblk = half * (half + 1) / 2;
    int bP0 = 0, bR0 = 0, bP1 = 0, bR1 = 0;
    const bool has0 = threadIdx.x < nblk, has1 = threadIdx.x + kNT < nblk;
    if (has0) tri_index(threadIdx.x, bP0, bR0);
    if (has1) tri_index(threadIdx.x + kNT, bP1, bR1);
    // column pair slot from the low 5 thread bits, rows ic + kRG j
    constexpr int kRG = kNT / 32, kRU = 64 / kRG;  // row groups, rows per thread in the first 64
    const int kc = threadIdx.x & 31, ic = threadIdx.x >> 5;
    const bool hasC = kc < half;
    int sweep = 0;
    QX_PT_INIT(jt)
    while (sweep < maxsweep) {
        ++sweep;
        for (int r = 0; r < m - 1; ++r) {
            const int f = r & 1;
            QX_PT(20, jt);
            for (int k = threadIdx.x; k < half; k += kNT) {
                int p, q;
                rr_pair(m, r, k, p, q);
                const double apq = A[(size_t)p * lda + q];
                const double app = A[(size_t)p * lda + p], aqq = A[(size_t)q * lda + q];
                double2 cs = make_double2(1.0, 0.0);
                if (fabs(apq) >= tol) {
                    cs = jacobi_rotation(app, aqq, apq);
                    if (fabs(cs.y) > min_s) s_any[f] = 1;
                    else cs = make_double2(1.0, 0.0);
                }
                rot[k] = cs;
                rpq[k] = make_int2(p, q);
            }
            __syncthreads();
            QX_PT(16, jt);
            QX_PT_COUNT(jt, 3);
#ifdef QX_PHASE_TIMING
            if (threadIdx.x == 0)
                for (int k = 0; k < half; ++k) jt_acc[5] += rot[k].y != 0.0 ? 1 : 0;
#endif
            if (s_any[f]) {
                QX_PT_COUNT(jt, 2);
                // parameters
                const double2 rp0 = rot[bP0], rr0 = rot[bR0], rp1 = rot[bP1], rr1 = rot[bR1];
                const int2 ip0 = rpq[bP0], ir0 = rpq[bR0], ip1 = rpq[bP1], ir1 = rpq[bR1];
                const double2 rc = rot[hasC ? kc : 0];
                const int2 pc = rpq[hasC ? kc : 0];
                const bool do0 = has0 && (rp0.y != 0.0 || rr0.y != 0.0);
                const bool do1 = has1 && (rp1.y != 0.0 || rr1.y != 0.0);
                const bool doC = hasC && rc.y != 0.0;
                // data
                double b0[4] = {0, 0, 0, 0}, b1[4] = {0, 0, 0, 0}, vp[kRU], vq[kRU];
                if (do0) load_block(A, lda, ip0, ir0, b0);
                if (do1) load_block(A, lda, ip1, ir1, b1);
#pragma unroll
                for (int u = 0; u < kRU; ++u) {
                    const int i = ic + kRG * u;
                    const bool ok = doC && i < nrow;
                    vp[u] = ok ? Cv[(size_t)i * lda + pc.x] : 0.0;
                    vq[u] = ok ? Cv[(size_t)i * lda + pc.y] : 0.0;
                }
                // rotate and store
                if (do0) {
                    double n0[4];
                    const bool dg = bP0 == bR0;
                    rotate_block(rp0, rr0, dg, b0, n0);
                    store_block(A, lda, ip0, ir0, dg, n0);
                }
                if (do1) {
                    double n1[4];
                    const bool dg = bP1 == bR1;
                    rotate_block(rp1, rr1, dg, b1, n1);
                    store_block(A, lda, ip1, ir1, dg, n1);
                }
                if (doC) {
                    const double c = rc.x, s = rc.y;
#pragma unroll
                    for (int u = 0; u < kRU; ++u) {
                        const int i = ic + kRG * u;
                        if (i < nrow) {
                            Cv[(size_t)i * lda + pc.x] = c * vp[u] - s * vq[u];
                            Cv[(size_t)i * lda + pc.y] = s * vp[u] + c * vq[u];
                        }
                    }
                }
                // larger systems: remaining blocks, rows and column pairs
                for (int b = threadIdx.x + 2 * kNT; b < nblk; b += kNT) {
                    int P, R;
                    tri_index(b, P, R);
                    const double2 xp = rot[P], xr = rot[R];
                    if (xp.y == 0.0 && xr.y == 0.0) continue;
                    const int2 jp = rpq[P], jr = rpq[R];
                    double bb[4], nn[4];
                    load_block(A, lda, jp, jr, bb);
                    rotate_block(xp, xr, P == R, bb, nn);
                    store_block(A, lda, jp, jr, P == R, nn);
                }
                for (int k = kc; k < half; k += 32) {
                    const double2 cs = rot[k];
                    if (cs.y == 0.0) continue;
                    const int2 pq = rpq[k];
                    for (int i = (k == kc ? ic + 64 : ic); i < nrow; i += kRG) {
                        const double xv = Cv[(size_t)i * lda + pq.x], yv = Cv[(size_t)i * lda + pq.y];
                        Cv[(size_t)i * lda + pq.x] = cs.x * xv - cs.y * yv;
                        Cv[(size_t)i * lda + pq.y] = cs.y * xv + cs.x * yv;
                    }
                }
            }
            if (threadIdx.x == 0) s_any[f ^ 1] = 0;
            __syncthreads();
            QX_PT(17, jt);
        }
        if (threadIdx.x == 0) s_any[0] = 0;  // set by the sweep's last round
        if (offdiag_max(A, m, lda, red) < tol) break;
    }
    QX_PT_FLUSH(jt);
    return sweep;
}

constexpr double kCoarseSin = 0.02;  // |sin| of the rotations a coarse sweep applies
constexpr int kCoarseSweeps = 3;      // coarse sweeps before the refinement gives up

// Simultaneous-rotation refinement of a nearly diagonal symmetric A (m x m,
// leading dimension lda) by FP64 MFMA products.  Every pair (i, j) gets the
// angle of the Jacobi rotation that would zero a_ij alone; with K the
// antisymmetric matrix of the angles, Y = I + K (+ K^2 / 2 when max |K| >
// 1e-4 or <= 1e-5) is orthogonal up to O(K^2) (O(K^4)), and none, one or two
// Newton-Schulz steps R <- R (3 I - R^T R) / 2 make it orthogonal to
// rounding.  Then A <- R^T A R and Cv <- Cv R (nrow x m).  The off-diagonal
// decreases quadratically (errors of the simultaneous angles are second
// order in the off-diagonal over the eigenvalue gaps).  An angle above 0.1
// (near-degenerate pairs with a relatively large coupling, typically at the
// first SCC iteration after the geometry moved: the series would need more
// terms) calls for a coarse Jacobi sweep that rotates only the pairs whose
// rotation has |sin| > kCoarseSin, up to kCoarseSweeps times; from the first
// one on, pairs with a coupling below tol are left alone, as Jacobi leaves
// them.  Returns true when the largest off-diagonal magnitude is below tol
// after at most maxit rounds (besides the coarse sweeps); false, possibly
// after some rounds, when the angles stay above 0.1 or the rounds did not
// suffice, and the caller finishes with Jacobi.  Y, Z, T: m x m scratch
// (leading dimension lda); rot, rpq, s_any: the Jacobi buffers.  Called by
// all threads.
QX_DEV bool eig_rotate(double* A, double* Cv, int m, int nrow, int lda, double tol, int maxit, double* Y,
                       double* Z, double* T, double* red, double2* rot, int2* rpq, int* s_any)
{
    const int tid = threadIdx.x;
    if (offdiag_max(A, m, lda, red) < tol) return true;
    int coarse = 0;
    for (int it = 0; it < maxit + coarse; ++it) {
        // one angle per pair i < j (K_ji = -K_ij): rows r and m - 1 - r of
        // the upper triangle together hold m - 1 pairs, so pair b maps to
        // (r, c) = (b / (m - 1), b % (m - 1)) without a square root (m even)
        double kmax = 0.0;
        for (int b = tid; b < (m / 2) * (m - 1); b += kNT) {
            const int r = b / (m - 1), c = b - r * (m - 1);
            const bool first = c < m - 1 - r;
            const int i = first ? r : m - 1 - r;
            const int j = first ? r + 1 + c : m - r + (c - (m - 1 - r));
            const double a = A[(size_t)i * lda + j];
            const double d = A[(size_t)j * lda + j] - A[(size_t)i * lda + i];
            const double rr = sqrt(d * d + 4.0 * a * a);
            const double t = rr > 0.0 ? 2.0 * (d >= 0.0 ? a : -a) / (fabs(d) + rr) : 0.0;
            // after a coarse sweep, pairs below tol are left alone as Jacobi leaves them
            const double th = (coarse && fabs(a) < tol) ? 0.0 : atan(t);
            Y[(size_t)i * lda + j] = th;
            Y[(size_t)j * lda + i] = -th;
            kmax = fmax(kmax, fabs(th));
        }
        for (int i = tid; i < m; i += kNT) Y[(size_t)i * lda + i] = 0.0;
        kmax = block_max(kmax, red);
        if (kmax > 0.1) {
            // strongly coupled pairs (near-degenerate orbitals after the
            // geometry moved): one Jacobi sweep over the pairs whose rotation
            // exceeds kCoarseSin, up to kCoarseSweeps times, then the
            // simultaneous rotations again
            if (coarse == kCoarseSweeps) return false;
            ++coarse;
            jacobi(A, Cv, m, nrow, lda, tol, 1, rot, rpq, red, s_any, kCoarseSin);
            QX_PH_ADD(29, 1);
            continue;
        }
        // Y = I + K + K^2/2 has Y^T Y = I + K^4/4: with max |K| <= 1e-5 that
        // is below rounding and no Newton-Schulz step is needed
        const bool quad = kmax > 1e-4 || kmax <= 1e-5;
        if (quad) {
            mfma_gemm<false, false>(m, m, m, Y, lda, Y, lda, Z, lda, nullptr);
            __syncthreads();
        }
        for (int p = tid; p < m * m; p += kNT) {
            const int i = p / m, j = p - i * m;
            const size_t o = (size_t)i * lda + j;
            Y[o] = (i == j ? 1.0 : 0.0) + Y[o] + (quad ? 0.5 * Z[o] : 0.0);
        }
        __syncthreads();
        double* R = Y;
        const int nns = kmax > 0.005 ? 2 : (kmax > 1e-5 ? 1 : 0);
        for (int s = 0; s < nns; ++s) {
            mfma_gemm<true, false>(m, m, m, R, lda, R, lda, Z, lda, nullptr);
            __syncthreads();
            for (int p = tid; p < m * m; p += kNT) {
                const int i = p / m, j = p - i * m;
                const size_t o = (size_t)i * lda + j;
                Z[o] = (i == j ? 1.5 : 0.0) - 0.5 * Z[o];
            }
            __syncthreads();
            double* out = R == Y ? T : Y;
            mfma_gemm<false, false>(m, m, m, R, lda, Z, lda, out, lda, nullptr);
            __syncthreads();
            R = out;
        }
        double* W = R == Y ? T : Y;
        mfma_gemm<false, false>(m, m, m, A, lda, R, lda, W, lda, nullptr);
        __syncthreads();
        mfma_gemm<true, false>(m, m, m, R, lda, W, lda, A, lda, nullptr);
        mfma_gemm<false, false>(nrow, m, m, Cv, lda, R, lda, Z, lda, nullptr);
        __syncthreads();
        for (int p = tid; p < nrow * m; p += kNT) {
            const int i = p / m, j = p - i * m;
            Cv[(size_t)i * lda + j] = Z[(size_t)i * lda + j];
        }
        __syncthreads();
        if (offdiag_max(A, m, lda, red) < tol) return true;
    }
    return false;
}

// The Broyden weights of up to kWaveSolve secant pairs on wavefront 0: solve
// (B) x = diag(om) gm for the n x n symmetric positive definite B (leading
// dimension lb) and return gm <- diag(om) x.  Lane i keeps row i of B in
// registers; a right-looking Cholesky factor B = L L^T (column j: lane j's
// pivot, the other lanes' multipliers, rank-1 update of the trailing rows by
// lane broadcasts), then L y = diag(om) gm forward and L^T x = y backward.
// Called by the 64 threads of wavefront 0.
constexpr int kWaveSolve = 16;
QX_DEV void broyden_solve_wave(const double* B, int lb, double* gm, const double* om, int n)
{
    const int i = threadIdx.x & 63;
    double r[kWaveSolve];
#pragma unroll
    for (int k = 0; k < kWaveSolve; ++k)
        r[k] = (i < n && k < n) ? B[(size_t)i * lb + k] : (i == k ? 1.0 : 0.0);
    double g = i < n ? gm[i] * om[i] : 0.0;
#pragma unroll
    for (int j = 0; j < kWaveSolve; ++j) {
        if (j < n) {
            const double d = sqrt(__shfl(r[j], j, 64));
            if (i == j) r[j] = d;
            else if (i > j) r[j] = r[j] / d;
#pragma unroll
            for (int k = j + 1; k < kWaveSolve; ++k) {
                const double lkj = __shfl(r[j], k, 64);
                if (i >= k) r[k] = r[k] - r[j] * lkj;
            }
        }
    }
    // forward: L y = g
#pragma unroll
    for (int k = 0; k < kWaveSolve; ++k) {
        if (k < n) {
            const double yk = __shfl(g, k, 64) / __shfl(r[k], k, 64);
            if (i == k) g = yk;
            else if (i > k) g = g - r[k] * yk;
        }
    }
    // backward: L^T x = y (lane i takes L[k][i] from lane k)
#pragma unroll
    for (int k = kWaveSolve - 1; k >= 0; --k) {
        if (k < n) {
            const double xk = __shfl(g, k, 64) / __shfl(r[k], k, 64);
            double lki = 0.0;
#pragma unroll
            for (int c = 0; c < kWaveSolve; ++c) {
                const double v = __shfl(r[c], k, 64);
                if (c == i) lki = v;
            }
            if (i == k) g = xk;
            else if (i < k) g = g - lki * xk;
        }
    }
    if (i < n) gm[i] = g * om[i];
}

// Fermi occupations per spin on wavefront 0 for nalpha electrons; returns -TS.
// The chemical potential is found by Newton steps safeguarded by a bracket
// (bisection whenever a step leaves it), starting from mu_io when it lies
// inside the bracket, until the electron count is met to 1e-13 relative, and
// polished by one more Newton step; the root is written back to mu_io.
QX_DEV double fermi_wave(int n, const double* e, double kt, double nalpha, double* f, double& mu_io,
                         int& iters)
{
    const int l = threadIdx.x & 63;
    double lo = 1e300, hi = -1e300;
    for (int i = l; i < n; i += 64) {
        lo = fmin(lo, e[i]);
        hi = fmax(hi, e[i]);
    }
    for (int o = 32; o > 0; o >>= 1) {
        lo = fmin(lo, __shfl_xor(lo, o, 64));
        hi = fmax(hi, __shfl_xor(hi, o, 64));
    }
    lo = lo - 50.0 * kt - 1.0;
    hi = hi + 50.0 * kt + 1.0;
    double mu = (mu_io > lo && mu_io < hi) ? mu_io : 0.5 * (lo + hi);
    iters = 0;
    for (int it = 0; it < 300; ++it) {
        ++iters;
        double s = 0.0, ds = 0.0;
        for (int i = l; i < n; i += 64) {
            const double fi = 1.0 / (1.0 + exp((e[i] - mu) / kt));
            s += fi;
            ds += fi * (1.0 - fi) / kt;
        }
        s = wave_sum(s);
        ds = wave_sum(ds);
        const double diff = nalpha - s;
        if (fabs(diff) <= 1e-13 * fmax(nalpha, 1.0)) {
            if (ds > 0.0) mu += diff / ds;  // final Newton polish
            break;
        }
        if (s > nalpha) hi = mu;
        else lo = mu;
        double mun = (ds > 0.0) ? mu + diff / ds : 0.5 * (lo + hi);
        if (!(mun > lo && mun < hi)) mun = 0.5 * (lo + hi);
        mu = mun;
        if (hi - lo <= 1e-15 * (fabs(mu) + 1.0)) break;
    }
    double ts = 0.0;
    for (int i = l; i < n; i += 64) {
        const double fi = 1.0 / (1.0 + exp((e[i] - mu) / kt));
        f[i] = fi;
        if (fi > 0.0 && fi < 1.0) ts += fi * log(fi) + (1.0 - fi) * log(1.0 - fi);
    }
    mu_io = mu;
    return 2.0 * kt * wave_sum(ts);
}

struct SysShared {
    ShellDesc sh[QX_MAX_NSH];
    int ao_at[QX_MAX_NAO], ao_sh[QX_MAX_NAO];
    int shoff[QX_MAX_NAT + 1], aooff[QX_MAX_NAT + 1];  // shells/AOs of atom a: [off[a], off[a+1])
    int sel[QX_MAX_NAT], refoff[QX_MAX_NAT + 1];
    double sx[3 * QX_MAX_NAT];
    double mrad[QX_MAX_NAT];
    double qat[QX_MAX_NAT], vat[QX_MAX_NAT];
    double zr[kMaxRef], dzr[kMaxRef], yr[kMaxRef];
    double red[8 * (kNT / 64)];
    double eparts[6];  // es2, es3, aes, axc, d4, unused
};

// Potentials vs (per shell), vd (3 per atom), vq (6 per atom) and interaction
// energies (ES2, ES3, AES, AXC, D4 two-body) at moments x = [qsh, dp, qp].
// Wavefront 0 evaluates the multipole terms (lane = atom, loop over partners
// through the precomputed kernel planes); wavefronts 1-3 evaluate the D4
// response y = W z and the shell Coulomb potential gamma q as column sweeps
// of the symmetric matrices.  The AES energy is x . v_aes / 2 (the
// multipole energy is a symmetric bilinear form of the moments).
QX_DEV void potentials(const SysInfo& si, const BatchPtrs& B, SysShared& ss, const double* x,
                       double* vs, double* vd, double* vq)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nat = si.nat, nsh = si.nsh, nd = si.ndim;
    const double* qsh = x;
    const double* dp = x + nsh;
    const double* qp = x + nsh + 3 * nat;
    QX_PT_INIT(pt)
    for (int a = tid; a < nat; a += kNT) {
        double q = 0.0;
        for (int s = ss.shoff[a]; s < ss.shoff[a + 1]; ++s) q += qsh[s];
        ss.qat[a] = q;
    }
    __syncthreads();
    for (int k = tid; k < nd; k += kNT) {
        int a = 0;
        while (ss.refoff[a + 1] <= k) ++a;
        const ElemParams& e = g_tab.e[ss.sel[a]];
        const int r = k - ss.refoff[a];
        const double c = e.gam * gfn2::kD4Gc;
        ss.zr[k] = d4_zeta(gfn2::kD4Ga, c, e.refq[r] + e.zeff, ss.qat[a] + e.zeff);
        ss.dzr[k] = d4_dzeta(gfn2::kD4Ga, c, e.refq[r] + e.zeff, ss.qat[a] + e.zeff);
    }
    __syncthreads();
    QX_PT(22, pt);
    double eaes = 0.0, eaxc = 0.0, es2 = 0.0, es3 = 0.0, ed4 = 0.0;
    if (wave == 0) {
        // up to 32 atoms: lanes i and i + 32 take the two halves of the
        // partners of atom i and add their sums; more atoms: one lane per atom
        const bool split = nat <= 32;
        const int i = split ? (lane & 31) : lane, h = split ? (lane >> 5) : 0;
        const int jm = split ? (nat + 1) / 2 : nat;
        const int j0 = h ? jm : 0, j1 = (split && !h) ? jm : nat;
        double v0 = 0.0, d0 = 0.0, d1 = 0.0, d2 = 0.0;
        double w0 = 0.0, w1 = 0.0, w2 = 0.0, w3 = 0.0, w4 = 0.0, w5 = 0.0;
        if (i < nat) {
            const int nn = nat * nat;
            const double* ag = B.aesg + (size_t)si.pair_off / QX_PAIR_STRIDE * QX_AES_COMP + i;
#pragma unroll 4
            for (int j = j0; j < j1; ++j) {
                const double* g = ag + (size_t)j * nat;
                const double g0 = g[0], g1 = g[nn], g2 = g[2 * nn], g3 = g[3 * nn], g4 = g[4 * nn],
                             g5 = g[5 * nn], g6 = g[6 * nn], g7 = g[7 * nn], g8 = g[8 * nn],
                             g9 = g[9 * nn];
                const double qj = ss.qat[j];
                const double m0 = dp[3 * j], m1 = dp[3 * j + 1], m2 = dp[3 * j + 2];
                const double* t = qp + 6 * j;
                v0 += m0 * g0 + m1 * g1 + m2 * g2 + t[0] * g3 + 2.0 * t[1] * g4 + t[2] * g5 +
                      2.0 * t[3] * g6 + 2.0 * t[4] * g7 + t[5] * g8;
                d0 += -qj * g0 + g9 * m0 - 3.0 * (g3 * m0 + g4 * m1 + g6 * m2);
                d1 += -qj * g1 + g9 * m1 - 3.0 * (g4 * m0 + g5 * m1 + g7 * m2);
                d2 += -qj * g2 + g9 * m2 - 3.0 * (g6 * m0 + g7 * m1 + g8 * m2);
                w0 += qj * g3;
                w1 += qj * g4;
                w2 += qj * g5;
                w3 += qj * g6;
                w4 += qj * g7;
                w5 += qj * g8;
            }
        }
        if (split) {
            v0 += __shfl_xor(v0, 32, 64);
            d0 += __shfl_xor(d0, 32, 64);
            d1 += __shfl_xor(d1, 32, 64);
            d2 += __shfl_xor(d2, 32, 64);
            w0 += __shfl_xor(w0, 32, 64);
            w1 += __shfl_xor(w1, 32, 64);
            w2 += __shfl_xor(w2, 32, 64);
            w3 += __shfl_xor(w3, 32, 64);
            w4 += __shfl_xor(w4, 32, 64);
            w5 += __shfl_xor(w5, 32, 64);
        }
        if (i < nat && h == 0) {
            w1 *= 2.0;
            w3 *= 2.0;
            w4 *= 2.0;
            const ElemParams& ei = g_tab.e[ss.sel[i]];
            const double qi = ss.qat[i];
            const double mi0 = dp[3 * i], mi1 = dp[3 * i + 1], mi2 = dp[3 * i + 2];
            const double* ti = qp + 6 * i;
            eaes = 0.5 * (qi * v0 + mi0 * d0 + mi1 * d1 + mi2 * d2 + ti[0] * w0 + ti[1] * w1 +
                          ti[2] * w2 + ti[3] * w3 + ti[4] * w4 + ti[5] * w5);
            const double dk = ei.dkernel, qk = ei.qkernel;
            eaxc = dk * (mi0 * mi0 + mi1 * mi1 + mi2 * mi2) +
                   qk * (ti[0] * ti[0] + ti[2] * ti[2] + ti[5] * ti[5] +
                         2.0 * (ti[1] * ti[1] + ti[3] * ti[3] + ti[4] * ti[4]));
            vd[3 * i + 0] = d0 + 2.0 * dk * mi0;
            vd[3 * i + 1] = d1 + 2.0 * dk * mi1;
            vd[3 * i + 2] = d2 + 2.0 * dk * mi2;
            vq[6 * i + 0] = w0 + 2.0 * qk * ti[0];
            vq[6 * i + 1] = w1 + 4.0 * qk * ti[1];
            vq[6 * i + 2] = w2 + 2.0 * qk * ti[2];
            vq[6 * i + 3] = w3 + 4.0 * qk * ti[3];
            vq[6 * i + 4] = w4 + 4.0 * qk * ti[4];
            vq[6 * i + 5] = w5 + 2.0 * qk * ti[5];
            ss.vat[i] = v0;
        }
    } else {
        const int t0 = tid - 64, nt = kNT - 64;
        const double* WD = B.wd4 + si.d4_off;
        for (int k = t0; k < nd; k += nt) {
            double y = 0.0;
#pragma unroll 16
            for (int l = 0; l < nd; ++l) y += WD[(size_t)l * nd + k] * ss.zr[l];
            ss.yr[k] = y;
            ed4 += -0.5 * ss.zr[k] * y;
        }
        // the shell Coulomb rows on the threads the D4 rows leave free
        const double* G = B.gam + si.gam_off;
        const int ga = nd + nsh <= nt ? nd : 0;
        for (int a = t0 - ga; a < nsh; a += nt) {
            if (a < 0) break;
            double v = 0.0;
#pragma unroll 8
            for (int b = 0; b < nsh; ++b) v += G[(size_t)b * nsh + a] * qsh[b];
            const double q = qsh[a];
            const double g3 = g_tab.e[ss.sh[a].el].third[ss.sh[a].k];
            es2 += 0.5 * q * v;
            es3 += g3 * q * q * q / 3.0;
            vs[a] = v + g3 * q * q;
        }
    }
    __syncthreads();
    QX_PT(22, pt);
    for (int a = tid; a < nsh; a += kNT) {
        const int at = ss.sh[a].at;
        double vdisp = 0.0;
        for (int k = ss.refoff[at]; k < ss.refoff[at + 1]; ++k) vdisp += -ss.dzr[k] * ss.yr[k];
        vs[a] += ss.vat[at] + vdisp;
    }
    double e5[5] = {es2, es3, eaes, eaxc, ed4};
    block_sums<5>(e5, ss.red);
    if (tid == 0)
        for (int k = 0; k < 5; ++k) ss.eparts[k] = e5[k];
    __syncthreads();
    QX_PT(23, pt);
    QX_PT_FLUSH(pt);
}

QX_DEV void load_system(const SysInfo& si, const BatchPtrs& B, SysShared& ss)
{
    const int tid = threadIdx.x, nt = blockDim.x, nat = si.nat;
    for (int i = tid; i < 3 * nat; i += nt) ss.sx[i] = B.xyz[3 * (size_t)si.atom_off + i];
    for (int i = tid; i < nat; i += nt) {
        ss.sel[i] = B.elem[si.atom_off + i];
        ss.mrad[i] = B.mrad[si.atom_off + i];
    }
    build_shell_tables(si, B.elem, ss.sh, ss.ao_at, ss.ao_sh);
    if (tid == 0) {
        int o = 0, so = 0, ao = 0;
        for (int a = 0; a < nat; ++a) {
            const ElemParams& e = g_tab.e[ss.sel[a]];
            ss.refoff[a] = o;
            ss.shoff[a] = so;
            ss.aooff[a] = ao;
            o += e.nref;
            so += e.nsh;
            ao += (e.nsh == 2) ? 4 : 1;
        }
        ss.refoff[nat] = o;
        ss.shoff[nat] = so;
        ss.aooff[nat] = ao;
    }
    __syncthreads();
}

struct SccBuffers {
    double *A, *C, *F, *P;  // m x m working matrices (LDS or global), leading dimension ld
    double* rs;             // kRs row sums per AO
    int ld;
    double* E;              // fifth m x m matrix for the MFMA eigen refinement, or null
};

template <bool FAST, bool EIG>
QX_DEV void scc_body(int sys, const SysInfo& si, const BatchPtrs& B, const SccOpts& opt, SysShared& ss,
                     SccBuffers bf, double* xin, double* xout, double* vs, double* vd, double* vq,
                     double* eps, double* focc, double* f2, double2* rot, int2* rpq, double* s_e,
                     int* s_any)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nat = si.nat, nsh = si.nsh, n = si.nao, ld = si.ld, nx = si.nx;
    const int m = (n + 1) & ~1;
    const int lb = bf.ld;
    const size_t ms = (size_t)ld * ld;
    const double* S = B.S + si.mat_off;
    const double* H = B.H0 + si.mat_off;
    const double* Dm = B.D + 3 * si.mat_off;  // transposed: (mu, nu) at nu * ld + mu
    const double* Qm = B.Q + 6 * si.mat_off;
    double* xg = B.x + si.x_off;
    double* xp = B.xprev + si.x_off;    // converged moments of the step before
    double* xpp = B.xprev2 + si.x_off;  // and of the one before that
    double* xp3 = B.xprev3 + si.x_off;  // and of the one before that
    double* xp4 = B.xprev4 + si.x_off;  // and of the one before that
    const int warm = B.warm[sys];
    double* A = bf.A;
    double* Cc = bf.C;
    double* Ft = bf.F;
    double* P = bf.P;
    double* rs = bf.rs;

    QX_PH_INIT
    // initial moments (polynomial extrapolation of order 1-4 over the
    // previous steps' converged moments in MD) and basis C = X U (padding
    // column zero)
    for (int i = tid; i < nx; i += kNT) {
        double v = 0.0;
        if (warm >= 5 && opt.extrapolate >= 4)
            v = 5.0 * (xg[i] - xp3[i]) - 10.0 * (xp[i] - xpp[i]) + xp4[i];
        else if (warm >= 4 && opt.extrapolate >= 3) v = 4.0 * (xg[i] + xpp[i]) - 6.0 * xp[i] - xp3[i];
        else if (warm >= 3 && opt.extrapolate >= 2) v = 3.0 * (xg[i] - xp[i]) + xpp[i];
        else if (warm >= 2 && opt.extrapolate >= 1) v = 2.0 * xg[i] - xp[i];
        else if (warm >= 1) v = xg[i];
        xin[i] = v;
        if (warm >= 4) xp4[i] = xp3[i];
        if (warm >= 3) xp3[i] = xpp[i];
        if (warm >= 2) xpp[i] = xp[i];
        if (warm >= 1) xp[i] = xg[i];
    }
    if (warm) {
        mfma_gemm<false, false>(n, n, n, B.X + si.mat_off, ld, B.U + si.mat_off, ld, Cc, lb, nullptr);
    } else {
        const double* X = B.X + si.mat_off;
        for (int p = tid; p < n * n; p += kNT) Cc[(size_t)(p / n) * lb + p % n] = X[(size_t)(p / n) * ld + p % n];
    }
    if (m > n)
        for (int i = tid; i < n; i += kNT) Cc[(size_t)i * lb + n] = 0.0;
    if (tid == 0) s_e[1] = -1e300;
    __syncthreads();
    QX_PH(11);

    double* hist = B.hist + si.hist_off;
    const int mem = opt.mem;
    double* hdf = hist;
    double* hu = hist + (size_t)mem * nx;
    double* ha = hist + 2 * (size_t)mem * nx;
    double* hom = ha + (size_t)mem * mem;
    double* hql = hom + mem;
    double* hdql = hql + nx;
    double* hbeta = hdql + nx;
    double* hgam = hbeta + (size_t)mem * mem;
    // secant pairs form a ring of `ring` slots; with opt.keep > 0 a warm
    // start continues from the pairs of the previous call (count in hcnt)
    const int ring = (opt.keep > 0 && opt.keep < mem) ? opt.keep : mem;
    double* hcnt = hgam + mem;
    int hcount = (opt.keep > 0 && warm >= 1) ? (int)hcnt[0] : 0;

    const double nalpha = 0.5 * si.nel;
    double eold = 0.0, etot = 0.0, ts = 0.0, eband = 0.0, errp = 1e-3;
    int it = 1, conv = 0;
    for (; it <= opt.maxiter; ++it) {
        potentials(si, B, ss, xin, vs, vd, vq);
        QX_PH(0);
        // G^T (row nu holds G_{mu nu} over mu), then F = H0 - (G + G^T) / 2
        for (int mu = lane; mu < n; mu += 64) {
            const int am = ss.ao_at[mu];
            const double w0 = vs[ss.ao_sh[mu]], w1 = vd[3 * am], w2 = vd[3 * am + 1],
                         w3 = vd[3 * am + 2];
            const double* q = vq + 6 * am;
            const double w4 = q[0], w5 = q[1], w6 = q[2], w7 = q[3], w8 = q[4], w9 = q[5];
            for (int nu0 = wave; nu0 < n; nu0 += 4 * (kNT / 64)) {
                double g[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int nu = nu0 + u * (kNT / 64);
                    const size_t o = (size_t)(nu < n ? nu : 0) * ld + mu;
                    g[u] = S[o] * w0 + Dm[o] * w1 + Dm[ms + o] * w2 + Dm[2 * ms + o] * w3 +
                           Qm[o] * w4 + Qm[ms + o] * w5 + Qm[2 * ms + o] * w6 +
                           Qm[3 * ms + o] * w7 + Qm[4 * ms + o] * w8 + Qm[5 * ms + o] * w9;
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int nu = nu0 + u * (kNT / 64);
                    if (nu < n) Ft[(size_t)nu * lb + mu] = g[u];
                }
            }
        }
        __syncthreads();
        for (int p = tid; p < n * n; p += kNT) {
            const int mu = p / n, nu = p % n;
            if (nu < mu) continue;
            const double gmn = Ft[(size_t)mu * lb + nu], gnm = Ft[(size_t)nu * lb + mu];
            const double f = H[(size_t)mu * ld + nu] - 0.5 * (gmn + gnm);
            Ft[(size_t)mu * lb + nu] = f;
            Ft[(size_t)nu * lb + mu] = f;
        }
        __syncthreads();
        QX_PH(1);
        // A = C^T F C (F C staged in the density buffer)
        mfma_gemm<false, false>(n, n, n, Ft, lb, Cc, lb, P, lb, nullptr);
        __syncthreads();
        mfma_gemm<true, false>(n, n, n, Cc, lb, P, lb, A, lb, nullptr);
        if (m > n)
            for (int j = tid; j < m; j += kNT) {
                A[(size_t)n * lb + j] = 0.0;
                A[(size_t)j * lb + n] = 0.0;
            }
        __syncthreads();
        QX_PH(2);
        double jtol = fmin(1e-6, 1e-3 * errp);
        if (jtol < opt.jacobi_tol) jtol = opt.jacobi_tol;
        int nsweep = 0;
        if (EIG && bf.E && eig_rotate(A, Cc, m, n, lb, jtol, 3, Ft, P, bf.E, ss.red, rot, rpq, s_any)) {
            QX_PH_ADD(15, 1);
        } else {
            nsweep = jacobi(A, Cc, m, n, lb, jtol, 50, rot, rpq, ss.red, s_any);
        }
        QX_PH(3);
        QX_PH_ADD(12, nsweep);
        (void)nsweep;
        for (int i = tid; i < n; i += kNT) eps[i] = A[(size_t)i * lb + i];
        __syncthreads();
        if (tid < 64) {
            double mu = s_e[1];
            int fit = 0;
            const double t = fermi_wave(n, eps, si.kt, nalpha, focc, mu, fit);
            if (tid == 0) {
                s_e[0] = t;
                s_e[1] = mu;
            }
            QX_PH_ADD(14, fit);
            (void)fit;
            for (int i = tid; i < n; i += 64) f2[i] = 2.0 * focc[i];
        }
        __syncthreads();
        ts = s_e[0];
        QX_PH(4);
        mfma_gemm<false, true>(n, n, n, Cc, lb, Cc, lb, P, lb, f2);
        __syncthreads();
        QX_PH(5);
        // row sums of P o M for M = S, D (3), Q (6), H0 as coalesced column
        // sweeps: wavefront w takes the matrices w, w + 4, w + 8 for all AOs,
        // so every lane keeps three independent sums (in nu order) and three
        // times the loads in flight of one sum per thread
        {
            static_assert(kNT == 256 && kRs <= 12, "three row-sum matrices per wavefront");
            const int c0 = wave, c1 = wave + 4, c2 = wave + 8;
            const bool has2 = c2 < kRs;
            const double* M0 = c0 == 0 ? S : Dm + (size_t)(c0 - 1) * ms;
            const double* M1 = Qm + (size_t)(c1 - 4) * ms;
            const double* M2 = !has2 ? M1 : c2 < 10 ? Qm + (size_t)(c2 - 4) * ms : H;
            for (int mu = lane; mu < n; mu += 64) {
                double a0 = 0.0, a1 = 0.0, a2 = 0.0;
                if (has2) {
#pragma unroll 8
                    for (int nu = 0; nu < n; ++nu) {
                        const double p = P[(size_t)nu * lb + mu];
                        const size_t o = (size_t)nu * ld + mu;
                        a0 += p * M0[o];
                        a1 += p * M1[o];
                        a2 += p * M2[o];
                    }
                } else {
#pragma unroll 8
                    for (int nu = 0; nu < n; ++nu) {
                        const double p = P[(size_t)nu * lb + mu];
                        const size_t o = (size_t)nu * ld + mu;
                        a0 += p * M0[o];
                        a1 += p * M1[o];
                    }
                }
                rs[(size_t)mu * kRs + c0] = a0;
                rs[(size_t)mu * kRs + c1] = a1;
                if (has2) rs[(size_t)mu * kRs + c2] = a2;
            }
        }
        __syncthreads();
        for (int s = tid; s < nsh; s += kNT) {
            const ShellDesc d = ss.sh[s];
            double pop = 0.0;
            for (int a = 0; a < d.nao; ++a) pop += rs[(size_t)(d.ao0 + a) * kRs];
            xout[s] = g_tab.e[d.el].refocc[d.k] - pop;
        }
        for (int t = tid; t < 9 * nat; t += kNT) {
            const int a = t / 9, c = t % 9;
            double v = 0.0;
            for (int mu = ss.aooff[a]; mu < ss.aooff[a + 1]; ++mu) v += rs[(size_t)mu * kRs + 1 + c];
            if (c < 3) xout[nsh + 3 * a + c] = -v;
            else xout[nsh + 3 * nat + 6 * a + (c - 3)] = -v;
        }
        double red2[2] = {0.0, 0.0};
        for (int mu = tid; mu < n; mu += kNT) red2[0] += rs[(size_t)mu * kRs + 10];
        __syncthreads();
        QX_PH(6);
        potentials(si, B, ss, xout, vs, vd, vq);
        QX_PH(7);
        for (int i = tid; i < nx; i += kNT) {
            const double dd = xout[i] - xin[i];
            red2[1] += dd * dd;
        }
        block_sums<2>(red2, ss.red);
        eband = red2[0];
        etot = eband + ss.eparts[0] + ss.eparts[1] + ss.eparts[2] + ss.eparts[3] + ss.eparts[4] + ts;
        const double err = sqrt(red2[1] / nx);
        errp = err;
        const double de = etot - eold;
        eold = etot;
        if (it > 1 && fabs(de) < opt.econv && err < opt.pconv) {
            conv = 1;
            break;
        }
        if (it == opt.maxiter) break;
        QX_PH(8);
        // modified Broyden update of xin with dq = xout - xin (stored into xout)
        if (it == 1 && hcount == 0) {
            for (int i = tid; i < nx; i += kNT) {
                const double dq = xout[i] - xin[i];
                hdql[i] = dq;
                hql[i] = xin[i];
                xin[i] = xin[i] + opt.damp * dq;
            }
            __syncthreads();
            QX_PH(9);
            continue;
        }
        // a new secant pair from this call's previous iterate (it >= 2) into
        // ring slot k1; on the first iteration of a continued call only the
        // kept pairs enter
        const bool add = it >= 2;
        const int k1 = hcount % ring;
        if (add) {
            double nr[2] = {0.0, 0.0};
            for (int i = tid; i < nx; i += kNT) {
                const double dq = xout[i] - xin[i];
                xout[i] = dq;
                const double d = dq - hdql[i];
                hdf[(size_t)k1 * nx + i] = d;
                nr[0] += d * d;
                nr[1] += dq * dq;
            }
            block_sums<2>(nr, ss.red);
            const double inv = nr[0] > 0.0 ? 1.0 / sqrt(nr[0]) : 0.0;
            for (int i = tid; i < nx; i += kNT) {
                const double d = hdf[(size_t)k1 * nx + i] * inv;
                hdf[(size_t)k1 * nx + i] = d;
                hu[(size_t)k1 * nx + i] = opt.damp * d + inv * (xin[i] - hql[i]);
            }
            if (tid == 0) {
                double w = sqrt(nr[1]);
                w = (w > 0.01 / 100000.0) ? 0.01 / w : 100000.0;
                if (w < 1.0) w = 1.0;
                hom[k1] = w;
            }
            ++hcount;
        } else {
            for (int i = tid; i < nx; i += kNT) xout[i] = xout[i] - xin[i];
        }
        const int it1 = hcount < ring ? hcount : ring;  // live pairs
        __syncthreads();
        // dot products df_i . df_k1 and df_i . dq: wavefront per i
        for (int i = wave; i < it1; i += kNT / 64) {
            double s = 0.0, c = 0.0;
            for (int j = lane; j < nx; j += 64) {
                const double di = hdf[(size_t)i * nx + j];
                if (add) s += di * hdf[(size_t)k1 * nx + j];
                c += di * xout[j];
            }
            s = wave_sum(s);
            c = wave_sum(c);
            if (lane == 0) {
                if (add) {
                    ha[(size_t)i * mem + k1] = s;
                    ha[(size_t)k1 * mem + i] = s;
                }
                hgam[i] = c;
            }
        }
        __syncthreads();
        // (w0^2 I + diag(om) a diag(om)) gamma = diag(om) c, Cholesky in LDS
        // (the Fock buffer is free here) when it fits, else in global memory
        const bool lds_solve = it1 * (it1 + 2) <= m * m;
        double* bt = lds_solve ? Ft : hbeta;
        const int lbt = lds_solve ? it1 : mem;
        double* gm = lds_solve ? Ft + (size_t)it1 * it1 : hgam;
        double* om = lds_solve ? gm + it1 : hom;
        if (lds_solve) {
            for (int i = tid; i < it1; i += kNT) {
                om[i] = hom[i];
                gm[i] = hgam[i];
            }
            __syncthreads();
        }
        for (int p = tid; p < it1 * it1; p += kNT) {
            const int i = p / it1, j = p % it1;
            bt[(size_t)i * lbt + j] = om[i] * om[j] * ha[(size_t)i * mem + j] + (i == j ? 1e-4 : 0.0);
        }
        __syncthreads();
        if (it1 <= kWaveSolve) {
            if (tid < 64) broyden_solve_wave(bt, lbt, gm, om, it1);
        } else if (tid == 0) {
            for (int i = 0; i < it1; ++i) gm[i] *= om[i];
            for (int j = 0; j < it1; ++j) {
                double d = bt[(size_t)j * lbt + j];
                for (int k = 0; k < j; ++k) d -= bt[(size_t)j * lbt + k] * bt[(size_t)j * lbt + k];
                d = sqrt(d);
                bt[(size_t)j * lbt + j] = d;
                for (int i = j + 1; i < it1; ++i) {
                    double v = bt[(size_t)i * lbt + j];
                    for (int k = 0; k < j; ++k) v -= bt[(size_t)i * lbt + k] * bt[(size_t)j * lbt + k];
                    bt[(size_t)i * lbt + j] = v / d;
                }
            }
            for (int i = 0; i < it1; ++i) {
                double v = gm[i];
                for (int k = 0; k < i; ++k) v -= bt[(size_t)i * lbt + k] * gm[k];
                gm[i] = v / bt[(size_t)i * lbt + i];
            }
            for (int i = it1 - 1; i >= 0; --i) {
                double v = gm[i];
                for (int k = i + 1; k < it1; ++k) v -= bt[(size_t)k * lbt + i] * gm[k];
                gm[i] = v / bt[(size_t)i * lbt + i];
            }
            for (int i = 0; i < it1; ++i) gm[i] *= om[i];
        }
        __syncthreads();
        for (int j = tid; j < nx; j += kNT) {
            const double dq = xout[j];
            double q = xin[j];
            hdql[j] = dq;
            hql[j] = q;
            q += opt.damp * dq;
            for (int i = 0; i < it1; ++i) q -= gm[i] * hu[(size_t)i * nx + j];
            xin[j] = q;
        }
        __syncthreads();
        QX_PH(9);
    }
    QX_PH_ADD(13, it > opt.maxiter ? opt.maxiter : it);
    if (opt.keep > 0 && tid == 0) hcnt[0] = (double)hcount;
    // results: moments, charges, potentials at the output moments
    for (int i = tid; i < nx; i += kNT) {
        xg[i] = xout[i];
        B.xo[si.x_off + i] = xout[i];
    }
    for (int a = tid; a < nat; a += kNT) B.qat[si.atom_off + a] = ss.qat[a];
    for (int s = tid; s < nsh; s += kNT) B.vs[si.shell_off + s] = vs[s];
    for (int i = tid; i < 3 * nat; i += kNT) B.vd[3 * (size_t)si.atom_off + i] = vd[i];
    for (int i = tid; i < 6 * nat; i += kNT) B.vq[6 * (size_t)si.atom_off + i] = vq[i];
    for (int i = tid; i < n; i += kNT) {
        B.eps[(size_t)sys * QX_MAX_NAO + i] = eps[i];
        B.focc[(size_t)sys * QX_MAX_NAO + i] = focc[i];
        f2[i] = 2.0 * focc[i] * eps[i];
    }
    double* Pg = B.P + si.mat_off;
    if (P != Pg)
        for (int p = tid; p < n * n; p += kNT) Pg[(size_t)(p / n) * ld + p % n] = P[(size_t)(p / n) * lb + p % n];
    __syncthreads();
    // W = C diag(2 f e) C^T ; U = L^T C for the next warm start
    mfma_gemm<false, true>(n, n, n, Cc, lb, Cc, lb, B.W + si.mat_off, ld, f2);
    mfma_gemm<true, false>(n, n, n, B.L + si.mat_off, ld, Cc, lb, B.U + si.mat_off, ld, nullptr);
    QX_PH(10);
    if (tid == 0) {
        double* ep = B.eparts + 10 * sys;
        ep[0] = eband;
        ep[1] = ss.eparts[0];
        ep[2] = ss.eparts[1];
        ep[3] = ss.eparts[2];
        ep[4] = ss.eparts[3];
        ep[5] = ss.eparts[4];
        ep[8] = ts;
        ep[9] = etot;
        B.energy[sys] = etot + ep[6] + ep[7];
        const int itn = it > opt.maxiter ? opt.maxiter : it;
        B.iters[sys] = itn;
        B.iters_total[sys] += (unsigned long long)itn;
        B.status[sys] = conv ? 0 : -1;
        B.warm[sys] = warm >= 5 ? 5 : warm + 1;
    }
}

// LDS of one system's SCC
struct SccLds {
    SysShared ss;
    double xin[kMaxNx], xout[kMaxNx], vs[QX_MAX_NSH], vd[3 * QX_MAX_NAT], vq[6 * QX_MAX_NAT];
    double eps[QX_MAX_NAO], focc[QX_MAX_NAO], f2[QX_MAX_NAO];
    double2 rot[QX_MAX_NAO / 2];
    int2 rpq[QX_MAX_NAO / 2];
    double s_e[4];
    int s_any[2];
};

// the whole SCC of system sys by the block (dyn: the launch's dynamic LDS)
//   EIG: with the MFMA eigen refinement (a separate instantiation: the code
//   alone, even when not taken, costs the Jacobi-only kernel a quarter of
//   its speed through the scheduling of the shared SCC body)
template <bool EIG>
QX_DEV void scc_system(const BatchPtrs& B, const SccOpts& opt, int sys, SccLds& L, double* dyn)
{
    const SysInfo si = B.info[sys];
    load_system(si, B, L.ss);
    const int m = (si.nao + 1) & ~1;
    if (m <= opt.lds_max_m) {
        // odd leading dimension: column-wise LDS accesses hit distinct banks
        const size_t mm = (size_t)m * (m + 1);
        // a fifth matrix after the row sums when the launch's LDS holds it
        double* E = EIG && 5 * mm + kRs * (size_t)m <= (size_t)opt.lds_doubles ? dyn + 4 * mm + kRs * m
                                                                            : nullptr;
        SccBuffers bf{dyn, dyn + mm, dyn + 2 * mm, dyn + 3 * mm, dyn + 4 * mm, m + 1, E};
        scc_body<true, EIG>(sys, si, B, opt, L.ss, bf, L.xin, L.xout, L.vs, L.vd, L.vq, L.eps, L.focc, L.f2,
                       L.rot, L.rpq, L.s_e, L.s_any);
    } else {
        // the energy-weighted density slot doubles as the row-sum scratch;
        // the MFMA eigen refinement's fifth matrix is the batch's E slot
        SccBuffers bf{B.T + si.mat_off, B.C + si.mat_off, B.U + si.mat_off, B.P + si.mat_off,
                      B.W + si.mat_off, si.ld, EIG ? B.E + si.mat_off : nullptr};
        scc_body<false, EIG>(sys, si, B, opt, L.ss, bf, L.xin, L.xout, L.vs, L.vd, L.vq, L.eps, L.focc,
                        L.f2, L.rot, L.rpq, L.s_e, L.s_any);
    }
}

template <bool EIG>
__global__ __launch_bounds__(kNT) void k_scc(BatchPtrs B, SccOpts opt)
{
    const int sys = opt.first_sys + blockIdx.x;
    if (B.active && !B.active[sys]) return;
    __shared__ SccLds L;
    extern __shared__ double dyn[];
    scc_system<EIG>(B, opt, sys, L, dyn);
}

}  // namespace qx
