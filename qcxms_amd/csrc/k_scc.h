// Self-consistent charge iterations of GFN2-xTB, one workgroup per system,
// all iterations inside one launch (data-dependent iteration counts never
// leave the device).
//
// Per iteration: potentials from the mixed moments -> Fock matrix built from
// row-local potentials G (F = H0 - (G + G^T)/2, coalesced row reads) ->
// A = C^T F C in the current orbital basis C (S-orthonormal) -> parallel
// cyclic Jacobi on A (round-robin pairing, rounds without a significant
// rotation skipped) -> C <- C V -> Fermi occupations (one wavefront) ->
// P = C diag(2f) C^T -> Mulliken shell charges and atomic dipole/quadrupole
// moments (wavefront per AO row) -> energy -> modified Broyden mixing
// (Johnson, PRB 38, 12807 (1988)).
// Systems with at most kFastM basis functions keep A, V, C, F and P in LDS;
// larger ones use per-system global scratch.  C starts as X U with X = L^-T
// and U the previous call's orthogonal basis (identity on a cold start).
#pragma once
#include "k_setup.h"

namespace qx {

constexpr int kMaxNx = 128 + 9 * QX_MAX_NAT;
constexpr int kMaxRef = QX_MAX_REF * QX_MAX_NAT;
constexpr int kFastM = 52;

// C[i][j] = sum_k opA[i][k] * d[k] * opB[k][j] over i < M, j < N, k < K
template <bool TA, bool TB>
QX_DEV void block_gemm(int M, int N, int K, const double* A, int lda, const double* Bm, int ldb,
                       double* Cm, int ldc, const double* d)
{
    const int tm = (M + 3) >> 2, tn = (N + 3) >> 2;
    for (int t = threadIdx.x; t < tm * tn; t += QX_BLOCK) {
        const int i0 = (t / tn) * 4, j0 = (t % tn) * 4;
        double acc[4][4];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
        for (int k = 0; k < K; ++k) {
            double av[4], bv[4];
            const double dk = d ? d[k] : 1.0;
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const int i = i0 + a < M ? i0 + a : M - 1;
                av[a] = (TA ? A[(size_t)k * lda + i] : A[(size_t)i * lda + k]) * dk;
            }
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int j = j0 + b < N ? j0 + b : N - 1;
                bv[b] = TB ? Bm[(size_t)j * ldb + k] : Bm[(size_t)k * ldb + j];
            }
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) acc[a][b] += av[a] * bv[b];
        }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b)
                if (i0 + a < M && j0 + b < N) Cm[(size_t)(i0 + a) * ldc + j0 + b] = acc[a][b];
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
    for (int i = 1; i < QX_BLOCK / 64; ++i) s = fmax(s, smem[i]);
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
        for (int i = 0; i < QX_BLOCK / 64; ++i) s += smem[i * K + k];
        v[k] = s;
    }
    __syncthreads();
}

// round-robin pairing: round r, slot k of m players
QX_DEV void rr_pair(int m, int r, int k, int& p, int& q)
{
    if (k == 0) {
        p = r;
        q = m - 1;
    } else {
        p = (r + k) % (m - 1);
        q = (r - k + (m - 1)) % (m - 1);
    }
}

// Parallel cyclic Jacobi on symmetric A (m x m, m even, leading dim lda),
// accumulating rotations into V (V must hold the identity on entry).
// Stops after a sweep in which no off-diagonal element exceeded tol.
QX_DEV int jacobi(double* A, double* V, int m, int lda, double tol, int maxsweep, double* rot,
                  int* rpq, double* red, int* s_any)
{
    const int half = m / 2;
    double mx = 0.0;
    for (int p = threadIdx.x; p < m * m; p += QX_BLOCK) {
        const int i = p / m, j = p % m;
        if (i != j) mx = fmax(mx, fabs(A[(size_t)i * lda + j]));
    }
    mx = block_max(mx, red);
    if (mx < tol) return 0;
    const double skip = 0.01 * tol;
    if (threadIdx.x == 0) {
        s_any[0] = 0;
        s_any[1] = 0;
    }
    __syncthreads();
    int sweep = 0;
    for (; sweep < maxsweep; ++sweep) {
        double smx = 0.0;
        for (int r = 0; r < m - 1; ++r) {
            const int f = r & 1;
            for (int k = threadIdx.x; k < half; k += QX_BLOCK) {
                int p, q;
                rr_pair(m, r, k, p, q);
                const double apq = A[(size_t)p * lda + q];
                smx = fmax(smx, fabs(apq));
                double c = 1.0, s = 0.0;
                if (fabs(apq) > skip) {
                    const double app = A[(size_t)p * lda + p], aqq = A[(size_t)q * lda + q];
                    const double tau = (aqq - app) / (2.0 * apq);
                    const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                    c = 1.0 / sqrt(1.0 + t * t);
                    s = t * c;
                    s_any[f] = 1;
                }
                rot[2 * k] = c;
                rot[2 * k + 1] = s;
                rpq[2 * k] = p;
                rpq[2 * k + 1] = q;
            }
            __syncthreads();
            if (s_any[f]) {
                const int nblk = half * (half + 1) / 2;
                for (int b = threadIdx.x; b < nblk; b += QX_BLOCK) {
                    int P = (int)((sqrt(8.0 * b + 1.0) - 1.0) * 0.5);
                    while (P * (P + 1) / 2 > b) --P;
                    while ((P + 1) * (P + 2) / 2 <= b) ++P;
                    const int R = b - P * (P + 1) / 2;  // R <= P
                    const double cp = rot[2 * P], sp = rot[2 * P + 1], cr = rot[2 * R], sr = rot[2 * R + 1];
                    if (sp == 0.0 && sr == 0.0) continue;
                    const int p = rpq[2 * P], q = rpq[2 * P + 1], r2 = rpq[2 * R], s2 = rpq[2 * R + 1];
                    const double b00 = A[(size_t)p * lda + r2], b01 = A[(size_t)p * lda + s2];
                    const double b10 = A[(size_t)q * lda + r2], b11 = A[(size_t)q * lda + s2];
                    const double m00 = b00 * cr - b01 * sr, m01 = b00 * sr + b01 * cr;
                    const double m10 = b10 * cr - b11 * sr, m11 = b10 * sr + b11 * cr;
                    double n00 = cp * m00 - sp * m10, n01 = cp * m01 - sp * m11;
                    double n10 = sp * m00 + cp * m10, n11 = sp * m01 + cp * m11;
                    if (P == R) {
                        n01 = 0.0;
                        n10 = 0.0;
                    }
                    A[(size_t)p * lda + r2] = n00;
                    A[(size_t)p * lda + s2] = n01;
                    A[(size_t)q * lda + r2] = n10;
                    A[(size_t)q * lda + s2] = n11;
                    if (P != R) {
                        A[(size_t)r2 * lda + p] = n00;
                        A[(size_t)s2 * lda + p] = n01;
                        A[(size_t)r2 * lda + q] = n10;
                        A[(size_t)s2 * lda + q] = n11;
                    }
                }
                for (int it = threadIdx.x; it < m * half; it += QX_BLOCK) {
                    const int i = it / half, k = it % half;
                    const double s = rot[2 * k + 1];
                    if (s == 0.0) continue;
                    const double c = rot[2 * k];
                    const int p = rpq[2 * k], q = rpq[2 * k + 1];
                    const double vp = V[(size_t)i * lda + p], vq = V[(size_t)i * lda + q];
                    V[(size_t)i * lda + p] = c * vp - s * vq;
                    V[(size_t)i * lda + q] = s * vp + c * vq;
                }
            }
            if (threadIdx.x == 0) s_any[f ^ 1] = 0;
            __syncthreads();
        }
        smx = block_max(smx, red);
        if (smx < tol) {
            ++sweep;
            break;
        }
    }
    return sweep;
}

// Fermi occupations per spin on wavefront 0 for nalpha electrons; returns -TS.
QX_DEV double fermi_wave(int n, const double* e, double kt, double nalpha, double* f)
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
    double mu = 0.0;
    for (int it = 0; it < 200; ++it) {
        mu = 0.5 * (lo + hi);
        double s = 0.0;
        for (int i = l; i < n; i += 64) s += 1.0 / (1.0 + exp((e[i] - mu) / kt));
        s = wave_sum(s);
        if (s > nalpha) hi = mu;
        else lo = mu;
        if (hi - lo < 1e-15 * (fabs(mu) + 1.0)) break;
    }
    for (int it = 0; it < 5; ++it) {
        double s = 0.0, ds = 0.0;
        for (int i = l; i < n; i += 64) {
            const double fi = 1.0 / (1.0 + exp((e[i] - mu) / kt));
            s += fi;
            ds += fi * (1.0 - fi) / kt;
        }
        s = wave_sum(s);
        ds = wave_sum(ds);
        if (ds <= 0.0) break;
        mu += (nalpha - s) / ds;
    }
    double ts = 0.0;
    for (int i = l; i < n; i += 64) {
        const double fi = 1.0 / (1.0 + exp((e[i] - mu) / kt));
        f[i] = fi;
        if (fi > 0.0 && fi < 1.0) ts += fi * log(fi) + (1.0 - fi) * log(1.0 - fi);
    }
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
    double red[8 * (QX_BLOCK / 64)];
    double eparts[6];  // es2, es3, aes, axc, d4, unused
};

// Potentials vs (per shell), vd (3 per atom), vq (6 per atom) and interaction
// energies (ES2, ES3, AES, AXC, D4 two-body) at moments x = [qsh, dp, qp].
QX_DEV void potentials(const SysInfo& si, const BatchPtrs& B, SysShared& ss, const double* x,
                       double* vs, double* vd, double* vq)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nat = si.nat, nsh = si.nsh;
    const double* qsh = x;
    const double* dp = x + nsh;
    const double* qp = x + nsh + 3 * nat;
    for (int a = tid; a < nat; a += QX_BLOCK) {
        double q = 0.0;
        for (int s = ss.shoff[a]; s < ss.shoff[a + 1]; ++s) q += qsh[s];
        ss.qat[a] = q;
    }
    __syncthreads();
    const int nd = si.ndim;
    for (int k = tid; k < nd; k += QX_BLOCK) {
        int a = 0;
        while (ss.refoff[a + 1] <= k) ++a;
        const ElemParams& e = g_tab.e[ss.sel[a]];
        const int r = k - ss.refoff[a];
        const double c = e.gam * gfn2::kD4Gc;
        ss.zr[k] = d4_zeta(gfn2::kD4Ga, c, e.refq[r] + e.zeff, ss.qat[a] + e.zeff);
        ss.dzr[k] = d4_dzeta(gfn2::kD4Ga, c, e.refq[r] + e.zeff, ss.qat[a] + e.zeff);
    }
    __syncthreads();
    // D4 response y = W z: wavefront per row, lanes over columns (coalesced)
    const double* WD = B.wd4 + si.d4_off;
    double ed4 = 0.0;
    for (int k = wave; k < nd; k += QX_BLOCK / 64) {
        double y = 0.0;
        for (int l = lane; l < nd; l += 64) y += WD[(size_t)k * nd + l] * ss.zr[l];
        y = wave_sum(y);
        if (lane == 0) {
            ss.yr[k] = y;
            ed4 += -0.5 * ss.zr[k] * y;
        }
    }
    __syncthreads();
    // AES and AXC: wavefront per atom i, lanes over partners j (nat <= 64)
    const double* ag = B.aesg + si.pair_off / 2;
    double eaes = 0.0, eaxc = 0.0;
    for (int i = wave; i < nat; i += QX_BLOCK / 64) {
        const double* mi = dp + 3 * i;
        const double* ti = qp + 6 * i;
        double acc[11];
#pragma unroll
        for (int k = 0; k < 11; ++k) acc[k] = 0.0;
        const int j = lane;
        if (j < nat && j != i) {
            const double* g = ag + (size_t)(i * nat + j) * 8;
            const double rv0 = g[0], rv1 = g[1], rv2 = g[2], h3 = g[3], h5 = g[4], h5b = g[5];
            const double* mj = dp + 3 * j;
            const double* tj = qp + 6 * j;
            const double mjr = mj[0] * rv0 + mj[1] * rv1 + mj[2] * rv2;
            const double tjrr = tj[0] * rv0 * rv0 + 2 * tj[1] * rv0 * rv1 + tj[2] * rv1 * rv1 +
                                2 * tj[3] * rv0 * rv2 + 2 * tj[4] * rv1 * rv2 + tj[5] * rv2 * rv2;
            const double qj = ss.qat[j];
            acc[0] = mjr * h3 + tjrr * h5;
            acc[1] = -qj * rv0 * h3 + mj[0] * h5b - 3.0 * rv0 * mjr * h5;
            acc[2] = -qj * rv1 * h3 + mj[1] * h5b - 3.0 * rv1 * mjr * h5;
            acc[3] = -qj * rv2 * h3 + mj[2] * h5b - 3.0 * rv2 * mjr * h5;
            acc[4] = qj * rv0 * rv0 * h5;
            acc[5] = qj * 2 * rv0 * rv1 * h5;
            acc[6] = qj * rv1 * rv1 * h5;
            acc[7] = qj * 2 * rv0 * rv2 * h5;
            acc[8] = qj * 2 * rv1 * rv2 * h5;
            acc[9] = qj * rv2 * rv2 * h5;
            if (j < i) {
                const double mir = mi[0] * rv0 + mi[1] * rv1 + mi[2] * rv2;
                const double mimj = mi[0] * mj[0] + mi[1] * mj[1] + mi[2] * mj[2];
                const double tirr = ti[0] * rv0 * rv0 + 2 * ti[1] * rv0 * rv1 + ti[2] * rv1 * rv1 +
                                    2 * ti[3] * rv0 * rv2 + 2 * ti[4] * rv1 * rv2 + ti[5] * rv2 * rv2;
                const double qi = ss.qat[i];
                acc[10] = qi * mjr * h3 - qj * mir * h3 + mimj * h5b - 3.0 * mir * mjr * h5 +
                          qi * tjrr * h5 + qj * tirr * h5;
            }
        }
#pragma unroll
        for (int k = 0; k < 11; ++k) acc[k] = wave_sum(acc[k]);
        if (lane == 0) {
            const ElemParams& ei = g_tab.e[ss.sel[i]];
            eaes += acc[10];
            eaxc += ei.dkernel * (mi[0] * mi[0] + mi[1] * mi[1] + mi[2] * mi[2]) +
                    ei.qkernel * (ti[0] * ti[0] + ti[2] * ti[2] + ti[5] * ti[5] +
                                  2.0 * (ti[1] * ti[1] + ti[3] * ti[3] + ti[4] * ti[4]));
            for (int c = 0; c < 3; ++c) vd[3 * i + c] = acc[1 + c] + 2.0 * ei.dkernel * mi[c];
            vq[6 * i + 0] = acc[4] + 2.0 * ei.qkernel * ti[0];
            vq[6 * i + 1] = acc[5] + 4.0 * ei.qkernel * ti[1];
            vq[6 * i + 2] = acc[6] + 2.0 * ei.qkernel * ti[2];
            vq[6 * i + 3] = acc[7] + 4.0 * ei.qkernel * ti[3];
            vq[6 * i + 4] = acc[8] + 4.0 * ei.qkernel * ti[4];
            vq[6 * i + 5] = acc[9] + 2.0 * ei.qkernel * ti[5];
            double vdisp = 0.0;
            for (int k = ss.refoff[i]; k < ss.refoff[i + 1]; ++k) vdisp += -ss.dzr[k] * ss.yr[k];
            ss.vat[i] = acc[0] + vdisp;
        }
    }
    __syncthreads();
    // ES2 + third order: wavefront per shell row of gamma
    const double* G = B.gam + si.gam_off;
    double es2 = 0.0, es3 = 0.0;
    for (int a = wave; a < nsh; a += QX_BLOCK / 64) {
        double v = 0.0;
        for (int b = lane; b < nsh; b += 64) v += G[(size_t)a * nsh + b] * qsh[b];
        v = wave_sum(v);
        if (lane == 0) {
            es2 += 0.5 * qsh[a] * v;
            const double g3 = g_tab.e[ss.sh[a].el].third[ss.sh[a].k];
            es3 += g3 * qsh[a] * qsh[a] * qsh[a] / 3.0;
            vs[a] = v + g3 * qsh[a] * qsh[a] + ss.vat[ss.sh[a].at];
        }
    }
    double e5[5] = {es2, es3, eaes, eaxc, ed4};
    block_sums<5>(e5, ss.red);
    if (tid == 0)
        for (int k = 0; k < 5; ++k) ss.eparts[k] = e5[k];
    __syncthreads();
}

QX_DEV void load_system(const SysInfo& si, const BatchPtrs& B, SysShared& ss)
{
    const int tid = threadIdx.x, nat = si.nat;
    for (int i = tid; i < 3 * nat; i += QX_BLOCK) ss.sx[i] = B.xyz[3 * (size_t)si.atom_off + i];
    for (int i = tid; i < nat; i += QX_BLOCK) {
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
    double *A, *V, *C, *F, *P;  // LDS or global
    int ld;                     // leading dimension of the five buffers
};

template <bool FAST>
QX_DEV void scc_body(const SysInfo& si, const BatchPtrs& B, const SccOpts& opt, SysShared& ss,
                     SccBuffers bf, double* xin, double* xout, double* vs, double* vd, double* vq,
                     double* eps, double* focc, double* f2, double* rot, int* rpq, double* s_e,
                     int* s_any)
{
    const int sys = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nat = si.nat, nsh = si.nsh, n = si.nao, ld = si.ld, nx = si.nx;
    const int m = (n + 1) & ~1;
    const int lb = bf.ld;
    const size_t ms = (size_t)ld * ld;
    const double* S = B.S + si.mat_off;
    const double* H = B.H0 + si.mat_off;
    const double* Dm = B.D + 3 * si.mat_off;
    const double* Qm = B.Q + 6 * si.mat_off;
    double* xg = B.x + si.x_off;
    double* xp = B.xprev + si.x_off;
    const int warm = B.warm[sys];
    double* A = bf.A;
    double* V = bf.V;
    double* Cc = bf.C;
    double* Ft = bf.F;
    double* P = bf.P;

    // initial moments (extrapolated in MD) and basis C = X U
    for (int i = tid; i < nx; i += QX_BLOCK) {
        double v = 0.0;
        if (warm >= 2 && opt.extrapolate) v = 2.0 * xg[i] - xp[i];
        else if (warm >= 1) v = xg[i];
        xin[i] = v;
        if (warm >= 1) xp[i] = xg[i];
    }
    if (warm) {
        block_gemm<false, false>(n, n, n, B.X + si.mat_off, ld, B.U + si.mat_off, ld, Cc, lb, nullptr);
    } else {
        const double* X = B.X + si.mat_off;
        for (int p = tid; p < n * n; p += QX_BLOCK) Cc[(size_t)(p / n) * lb + p % n] = X[(size_t)(p / n) * ld + p % n];
    }
    __syncthreads();

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

    const double nalpha = 0.5 * si.nel;
    double eold = 0.0, etot = 0.0, ts = 0.0, eband = 0.0;
    int it = 1, conv = 0;
    for (; it <= opt.maxiter; ++it) {
        potentials(si, B, ss, xin, vs, vd, vq);
        // G = row-local potential part of the Fock matrix
        for (int mu = wave; mu < n; mu += QX_BLOCK / 64) {
            const int am = ss.ao_at[mu];
            const double v0 = vs[ss.ao_sh[mu]];
            const double d0 = vd[3 * am], d1 = vd[3 * am + 1], d2 = vd[3 * am + 2];
            const double* q = vq + 6 * am;
            for (int nu = lane; nu < n; nu += 64) {
                const size_t o = (size_t)mu * ld + nu;
                const double g = S[o] * v0 + Dm[o] * d0 + Dm[ms + o] * d1 + Dm[2 * ms + o] * d2 +
                                 Qm[o] * q[0] + Qm[ms + o] * q[1] + Qm[2 * ms + o] * q[2] +
                                 Qm[3 * ms + o] * q[3] + Qm[4 * ms + o] * q[4] + Qm[5 * ms + o] * q[5];
                Ft[(size_t)mu * lb + nu] = g;
            }
        }
        __syncthreads();
        for (int p = tid; p < n * n; p += QX_BLOCK) {
            const int mu = p / n, nu = p % n;
            if (nu < mu) continue;
            const double gmn = Ft[(size_t)mu * lb + nu], gnm = Ft[(size_t)nu * lb + mu];
            const double f = H[(size_t)mu * ld + nu] - 0.5 * (gmn + gnm);
            Ft[(size_t)mu * lb + nu] = f;
            Ft[(size_t)nu * lb + mu] = f;
        }
        __syncthreads();
        // A = C^T F C
        block_gemm<false, false>(n, n, n, Ft, lb, Cc, lb, P, lb, nullptr);
        __syncthreads();
        block_gemm<true, false>(n, n, n, Cc, lb, P, lb, A, lb, nullptr);
        for (int p = tid; p < m * m; p += QX_BLOCK) {
            const int i = p / m, j = p % m;
            V[(size_t)i * lb + j] = (i == j) ? 1.0 : 0.0;
            if (i >= n || j >= n) A[(size_t)i * lb + j] = 0.0;
        }
        __syncthreads();
        jacobi(A, V, m, lb, opt.jacobi_tol, 50, rot, rpq, ss.red, s_any);
        for (int i = tid; i < n; i += QX_BLOCK) eps[i] = A[(size_t)i * lb + i];
        __syncthreads();
        if (tid < 64) {
            const double t = fermi_wave(n, eps, si.kt, nalpha, focc);
            if (tid == 0) s_e[0] = t;
        }
        // C <- C V (into the Fock buffer, then swap)
        block_gemm<false, false>(n, n, n, Cc, lb, V, lb, Ft, lb, nullptr);
        __syncthreads();
        ts = s_e[0];
        {
            double* t = Cc;
            Cc = Ft;
            Ft = t;
        }
        for (int i = tid; i < n; i += QX_BLOCK) f2[i] = 2.0 * focc[i];
        __syncthreads();
        block_gemm<false, true>(n, n, n, Cc, lb, Cc, lb, P, lb, f2);
        __syncthreads();
        // output moments: 10 row sums per AO row (wavefront per row), into A
        double* rs = A;
        for (int mu = wave; mu < n; mu += QX_BLOCK / 64) {
            double acc[10];
#pragma unroll
            for (int k = 0; k < 10; ++k) acc[k] = 0.0;
            for (int nu = lane; nu < n; nu += 64) {
                const size_t o = (size_t)mu * ld + nu;
                const double pm = P[(size_t)mu * lb + nu];
                acc[0] += pm * S[o];
                acc[1] += pm * Dm[o];
                acc[2] += pm * Dm[ms + o];
                acc[3] += pm * Dm[2 * ms + o];
#pragma unroll
                for (int k = 0; k < 6; ++k) acc[4 + k] += pm * Qm[k * ms + o];
            }
#pragma unroll
            for (int k = 0; k < 10; ++k) acc[k] = wave_sum(acc[k]);
            if (lane == 0)
                for (int k = 0; k < 10; ++k) rs[(size_t)mu * 10 + k] = acc[k];
        }
        __syncthreads();
        for (int s = tid; s < nsh; s += QX_BLOCK) {
            const ShellDesc d = ss.sh[s];
            double pop = 0.0;
            for (int a = 0; a < d.nao; ++a) pop += rs[(size_t)(d.ao0 + a) * 10];
            xout[s] = g_tab.e[d.el].refocc[d.k] - pop;
        }
        for (int t = tid; t < 9 * nat; t += QX_BLOCK) {
            const int a = t / 9, c = t % 9;
            double v = 0.0;
            for (int mu = ss.aooff[a]; mu < ss.aooff[a + 1]; ++mu) v += rs[(size_t)mu * 10 + 1 + c];
            if (c < 3) xout[nsh + 3 * a + c] = -v;
            else xout[nsh + 3 * nat + 6 * a + (c - 3)] = -v;
        }
        __syncthreads();
        potentials(si, B, ss, xout, vs, vd, vq);
        double red2[2] = {0.0, 0.0};
        for (int p = tid; p < n * n; p += QX_BLOCK) {
            const int i = p / n, j = p % n;
            red2[0] += P[(size_t)i * lb + j] * H[(size_t)i * ld + j];
        }
        for (int i = tid; i < nx; i += QX_BLOCK) {
            const double dd = xout[i] - xin[i];
            red2[1] += dd * dd;
        }
        block_sums<2>(red2, ss.red);
        eband = red2[0];
        etot = eband + ss.eparts[0] + ss.eparts[1] + ss.eparts[2] + ss.eparts[3] + ss.eparts[4] + ts;
        const double err = sqrt(red2[1] / nx);
        const double de = etot - eold;
        eold = etot;
        if (it > 1 && fabs(de) < opt.econv && err < opt.pconv) {
            conv = 1;
            break;
        }
        if (it == opt.maxiter) break;
        // modified Broyden update of xin with dq = xout - xin (stored into xout)
        if (it == 1) {
            for (int i = tid; i < nx; i += QX_BLOCK) {
                const double dq = xout[i] - xin[i];
                hdql[i] = dq;
                hql[i] = xin[i];
                xin[i] = xin[i] + opt.damp * dq;
            }
            __syncthreads();
            continue;
        }
        int it1 = it - 1;
        if (it1 > mem) it1 = mem;
        const int k1 = it1 - 1;
        double nr[2] = {0.0, 0.0};
        for (int i = tid; i < nx; i += QX_BLOCK) {
            const double dq = xout[i] - xin[i];
            xout[i] = dq;
            const double d = dq - hdql[i];
            hdf[(size_t)k1 * nx + i] = d;
            nr[0] += d * d;
            nr[1] += dq * dq;
        }
        block_sums<2>(nr, ss.red);
        const double inv = nr[0] > 0.0 ? 1.0 / sqrt(nr[0]) : 0.0;
        for (int i = tid; i < nx; i += QX_BLOCK) {
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
        __syncthreads();
        // dot products df_i . df_k1 and df_i . dq: wavefront per i
        for (int i = wave; i < it1; i += QX_BLOCK / 64) {
            double s = 0.0, c = 0.0;
            for (int j = lane; j < nx; j += 64) {
                const double di = hdf[(size_t)i * nx + j];
                s += di * hdf[(size_t)k1 * nx + j];
                c += di * xout[j];
            }
            s = wave_sum(s);
            c = wave_sum(c);
            if (lane == 0) {
                ha[(size_t)i * mem + k1] = s;
                ha[(size_t)k1 * mem + i] = s;
                hgam[i] = c;
            }
        }
        __syncthreads();
        if (tid == 0) {
            const double w0 = 0.01;
            for (int i = 0; i < it1; ++i)
                for (int j = 0; j < it1; ++j)
                    hbeta[(size_t)i * mem + j] = hom[i] * hom[j] * ha[(size_t)i * mem + j] + (i == j ? w0 * w0 : 0.0);
            for (int i = 0; i < it1; ++i) hgam[i] *= hom[i];
            for (int j = 0; j < it1; ++j) {
                double d = hbeta[(size_t)j * mem + j];
                for (int k = 0; k < j; ++k) d -= hbeta[(size_t)j * mem + k] * hbeta[(size_t)j * mem + k];
                d = sqrt(d);
                hbeta[(size_t)j * mem + j] = d;
                for (int i = j + 1; i < it1; ++i) {
                    double v = hbeta[(size_t)i * mem + j];
                    for (int k = 0; k < j; ++k) v -= hbeta[(size_t)i * mem + k] * hbeta[(size_t)j * mem + k];
                    hbeta[(size_t)i * mem + j] = v / d;
                }
            }
            for (int i = 0; i < it1; ++i) {
                double v = hgam[i];
                for (int k = 0; k < i; ++k) v -= hbeta[(size_t)i * mem + k] * hgam[k];
                hgam[i] = v / hbeta[(size_t)i * mem + i];
            }
            for (int i = it1 - 1; i >= 0; --i) {
                double v = hgam[i];
                for (int k = i + 1; k < it1; ++k) v -= hbeta[(size_t)k * mem + i] * hgam[k];
                hgam[i] = v / hbeta[(size_t)i * mem + i];
            }
            for (int i = 0; i < it1; ++i) hgam[i] *= hom[i];
        }
        __syncthreads();
        for (int j = tid; j < nx; j += QX_BLOCK) {
            const double dq = xout[j];
            double q = xin[j];
            hdql[j] = dq;
            hql[j] = q;
            q += opt.damp * dq;
            for (int i = 0; i < it1; ++i) q -= hgam[i] * hu[(size_t)i * nx + j];
            xin[j] = q;
        }
        __syncthreads();
    }
    // results: moments, charges, potentials at the output moments
    for (int i = tid; i < nx; i += QX_BLOCK) {
        xg[i] = xout[i];
        B.xo[si.x_off + i] = xout[i];
    }
    for (int a = tid; a < nat; a += QX_BLOCK) B.qat[si.atom_off + a] = ss.qat[a];
    for (int s = tid; s < nsh; s += QX_BLOCK) B.vs[si.shell_off + s] = vs[s];
    for (int i = tid; i < 3 * nat; i += QX_BLOCK) B.vd[3 * (size_t)si.atom_off + i] = vd[i];
    for (int i = tid; i < 6 * nat; i += QX_BLOCK) B.vq[6 * (size_t)si.atom_off + i] = vq[i];
    for (int i = tid; i < n; i += QX_BLOCK) {
        B.eps[(size_t)sys * QX_MAX_NAO + i] = eps[i];
        B.focc[(size_t)sys * QX_MAX_NAO + i] = focc[i];
        f2[i] = 2.0 * focc[i] * eps[i];
    }
    double* Pg = B.P + si.mat_off;
    if (P != Pg)
        for (int p = tid; p < n * n; p += QX_BLOCK) Pg[(size_t)(p / n) * ld + p % n] = P[(size_t)(p / n) * lb + p % n];
    __syncthreads();
    // W = C diag(2 f e) C^T ; U = L^T C for the next warm start
    block_gemm<false, true>(n, n, n, Cc, lb, Cc, lb, B.W + si.mat_off, ld, f2);
    block_gemm<true, false>(n, n, n, B.L + si.mat_off, ld, Cc, lb, B.U + si.mat_off, ld, nullptr);
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
        B.warm[sys] = warm >= 1 ? 2 : 1;
    }
}

__global__ __launch_bounds__(QX_BLOCK) void k_scc(BatchPtrs B, SccOpts opt)
{
    const int sys = blockIdx.x;
    if (B.active && !B.active[sys]) return;
    const SysInfo si = B.info[sys];
    __shared__ SysShared ss;
    __shared__ double xin[kMaxNx], xout[kMaxNx], vs[QX_MAX_NSH], vd[3 * QX_MAX_NAT],
        vq[6 * QX_MAX_NAT];
    __shared__ double eps[QX_MAX_NAO], focc[QX_MAX_NAO], f2[QX_MAX_NAO];
    __shared__ double rot[QX_MAX_NAO];
    __shared__ int rpq[QX_MAX_NAO];
    __shared__ double s_e[4];
    __shared__ int s_any[2];
    extern __shared__ double dyn[];
    load_system(si, B, ss);
    const int m = (si.nao + 1) & ~1;
    if (m <= kFastM) {
        const size_t mm = (size_t)m * m;
        SccBuffers bf{dyn, dyn + mm, dyn + 2 * mm, dyn + 3 * mm, dyn + 4 * mm, m};
        scc_body<true>(si, B, opt, ss, bf, xin, xout, vs, vd, vq, eps, focc, f2, rot, rpq, s_e, s_any);
    } else {
        SccBuffers bf{B.W + si.mat_off, B.U + si.mat_off, B.C + si.mat_off, B.T + si.mat_off,
                      B.P + si.mat_off, si.ld};
        scc_body<false>(si, B, opt, ss, bf, xin, xout, vs, vd, vq, eps, focc, f2, rot, rpq, s_e,
                        s_any);
    }
}

}  // namespace qx
