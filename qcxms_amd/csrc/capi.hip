// Host side of libqcxms_gpu.so: parameter tables, batch memory layout, kernel
// launches and the extern "C" entry points declared in include/qcxms_gpu.h.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/qcxms_gpu.h"
#include "common.h"
#include "d4_data.h"
#include "sto_ng_table.h"

__constant__ DevTables g_tab;

#include "k_grad.h"
#include "k_md.h"

namespace {

thread_local std::string g_err;
std::mutex g_mu;
bool g_ready = false;
int g_device = -1;
hipStream_t g_stream = nullptr;
DevTables g_host_tab;

int fail(const std::string& msg)
{
    g_err = msg;
    return QCXMS_INVALID;
}

#define HIPCHK(expr)                                                                   \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(std::string(#expr) + ": " + hipGetErrorString(e_));            \
    } while (0)

int elem_slot(int z)
{
    for (int e = 0; e < gfn2::kElem; ++e)
        if (gfn2::kZ[e] == z) return e;
    return -1;
}

// D4 reference polarisabilities of element slot e after the secondary-system
// hydride correction (dftd4.f90:524-595, GFN2 charge set)
void d4_refal(int e, double out[QX_MAX_REF][23], int refc[QX_MAX_REF])
{
    auto zeta = [](double a, double c, double qref, double qmod) {
        if (qmod < 0.0) return std::exp(a);
        return std::exp(a * (1.0 - std::exp(c * (1.0 - qref / qmod))));
    };
    int cncount[16] = {0};
    cncount[0] = 1;
    const int nref = d4_refn[e];
    for (int j = 0; j < nref; ++j) {
        const int is = d4_refsys[e][j];
        const double iz = d4_zeff[is - 1];
        const double zt = zeta(gfn2::kD4Ga, d4_gam[is - 1] * gfn2::kD4Gc, d4_secq[is - 1] + iz,
                               d4_refh[e][j] + iz);
        const int icn = (int)std::lrint(d4_refcn[e][j]);
        cncount[icn] += 1;
        for (int k = 0; k < 23; ++k) {
            const double sec = d4_sscale[is - 1] * d4_secaiw[is - 1][k] * zt;
            const double v = d4_ascale[e][j] * (d4_alphaiw[e][j][k] - d4_hcount[e][j] * sec);
            out[j][k] = v > 0.0 ? v : 0.0;
        }
    }
    for (int j = 0; j < nref; ++j) {
        const int icn = cncount[(int)std::lrint(d4_refcn[e][j])];
        refc[j] = icn * (icn + 1) / 2;
    }
}

// Casimir-Polder integral over the 23-point imaginary frequency grid as two
// interleaved trapezoidal rules (dftd4.f90:451-497)
double d4_trapzd(const double* pol)
{
    static const double freq[23] = {0.000001, 0.05, 0.1, 0.2, 0.3, 0.4, 0.5, 0.6, 0.7, 0.8, 0.9, 1.0,
                                    1.2, 1.4, 1.6, 1.8, 2.0, 2.5, 3.0, 4.0, 5.0, 7.5, 10.0};
    double t1 = 0.0, t2 = 0.0;
    for (int k = 0; k < 11; ++k) t1 += (freq[2 * k + 1] - freq[2 * k]) * (pol[2 * k + 1] + pol[2 * k]);
    for (int k = 0; k < 11; ++k) t2 += (freq[2 * k + 2] - freq[2 * k + 1]) * (pol[2 * k + 2] + pol[2 * k + 1]);
    return 0.5 * t1 + 0.5 * t2;
}

void build_tables(DevTables& T)
{
    std::memset(&T, 0, sizeof(T));
    static double refal[gfn2::kElem][QX_MAX_REF][23];
    for (int e = 0; e < gfn2::kElem; ++e) {
        ElemParams& p = T.e[e];
        p.z = gfn2::kZ[e];
        p.nsh = gfn2::kNsh[e];
        for (int k = 0; k < p.nsh; ++k) {
            const int ng = gfn2::kNprim[e][k], n = gfn2::kPqn[e][k], l = k;
            const double z = gfn2::kSlater[e][k];
            p.nprim[k] = ng;
            p.pqn[k] = n;
            double s = 0.0;
            for (int i = 0; i < ng; ++i) {
                const double a = sto_ng_alpha[n - 1][l][ng - 1][i] * z * z;
                double norm = std::pow(2.0 * a / M_PI, 0.75);
                if (l == 1) norm *= std::sqrt(4.0 * a);
                p.alpha[k][i] = a;
                p.coeff[k][i] = sto_ng_coeff[n - 1][l][ng - 1][i] * norm;
            }
            for (int i = 0; i < ng; ++i)
                for (int j = 0; j < ng; ++j) {
                    const double g = p.alpha[k][i] + p.alpha[k][j];
                    double v = std::pow(M_PI / g, 1.5);
                    if (l == 1) v /= 2.0 * g;
                    s += p.coeff[k][i] * p.coeff[k][j] * v;
                }
            for (int i = 0; i < ng; ++i) p.coeff[k][i] /= std::sqrt(s);
            p.slater[k] = z;
            p.hself0[k] = gfn2::kSelfEnergyEv[e][k] * gfn2::kEvtoAu;
            p.kcn[k] = gfn2::kKcnEv[e][k] * gfn2::kEvtoAu;
            p.shpoly[k] = gfn2::kShellPoly[e][k];
            p.refocc[k] = gfn2::kRefOcc[e][k];
            p.hard[k] = gfn2::kHubbard[e] * (1.0 + gfn2::kShellHubbard[e][k]);
            p.third[k] = gfn2::kHubbardDeriv[e] * gfn2::kThirdShell[k];
        }
        p.rep_alpha = gfn2::kRepAlpha[e];
        p.rep_zeff = gfn2::kRepZeff[e];
        p.en = gfn2::kEn[e];
        p.rad = gfn2::kAatoAu * gfn2::kAtomicRadAa[e];
        p.dkernel = gfn2::kDipKernel[e];
        p.qkernel = gfn2::kQuadKernel[e];
        p.mprad = gfn2::kMpRad[e];
        p.mpvcn = gfn2::kMpVcn[e];
        p.nval = gfn2::kValence[e];
        p.covrad = 4.0 / 3.0 * (gfn2::kAatoAu * gfn2::kCovRad2009Aa[e]);
        // D4
        p.nref = d4_refn[e];
        d4_refal(e, refal[e], p.refc);
        for (int j = 0; j < QX_MAX_REF; ++j) {
            p.refq[j] = d4_refq[e][j];
            p.refcovcn[j] = d4_refcovcn[e][j];
        }
        p.zeff = d4_zeff[p.z - 1];
        p.gam = d4_gam[p.z - 1];
        p.r4r2 = d4_r4r2[p.z - 1];
        p.rcov_d4 = d4_rcov[p.z - 1];
        p.en_d4 = d4_en[p.z - 1];
    }
    for (int a = 0; a < gfn2::kElem; ++a)
        for (int i = 0; i < T.e[a].nref; ++i)
            for (int b = 0; b < gfn2::kElem; ++b)
                for (int j = 0; j < T.e[b].nref; ++j) {
                    double pol[23];
                    for (int k = 0; k < 23; ++k) pol[k] = refal[a][i][k] * refal[b][j][k];
                    T.c6ref[a][i][b][j] = 3.0 / M_PI * d4_trapzd(pol);
                }
}

int ensure_init(int device)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ready) return QCXMS_OK;
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (ndev <= 0) return fail("no HIP device available");
    if (device < 0 || device >= ndev) return fail("device index out of range");
    HIPCHK(hipSetDevice(device));
    build_tables(g_host_tab);
    HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(g_tab), &g_host_tab, sizeof(DevTables)));
    HIPCHK(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
    HIPCHK(hipFuncSetAttribute((const void*)qx::k_scc, hipFuncAttributeMaxDynamicSharedMemorySize,
                               2 * QX_LDS_MAXM * QX_LDS_MAXM * (int)sizeof(double)));
    HIPCHK(hipFuncSetAttribute((const void*)qx::k_setup, hipFuncAttributeMaxDynamicSharedMemorySize,
                               QX_LDS_MAXM * QX_LDS_MAXM * (int)sizeof(double)));
    g_device = device;
    g_ready = true;
    return QCXMS_OK;
}

struct Timer {
    double ms = 0.0;
    int launches = 0;
};

}  // namespace

struct qcxms_batch {
    int nsys = 0;
    std::vector<SysInfo> info;
    std::vector<int> elem;
    int natoms = 0, nshells = 0, nrefs = 0, nxtot = 0;
    size_t mat_total = 0, gam_total = 0, d4_total = 0, hist_total = 0, pair_total = 0;
    int max_ld = 0, max_m = 0;
    SccOpts opts{1e-9, 1e-8, 0.4, 1e-11, 100, 100};
    int warm_start = 0;
    int profiling = 0;
    Timer t_setup, t_scc, t_grad, t_md_pre, t_md_post;
    std::vector<void*> allocs;
    BatchPtrs d{};
    SysInfo* d_info = nullptr;
    int* d_elem = nullptr;

    ~qcxms_batch()
    {
        for (void* p : allocs) (void)hipFree(p);
    }
};

namespace {

template <class T>
int dalloc(qcxms_batch* b, T** p, size_t n)
{
    if (n == 0) n = 1;
    HIPCHK(hipMalloc((void**)p, n * sizeof(T)));
    HIPCHK(hipMemsetAsync(*p, 0, n * sizeof(T), g_stream));
    b->allocs.push_back((void*)*p);
    return QCXMS_OK;
}

int layout(qcxms_batch* b, int nsys, const int32_t* offsets, const int32_t* num, const int32_t* charge)
{
    b->nsys = nsys;
    b->info.resize(nsys);
    long long mat = 0, gam = 0, d4 = 0, hist = 0, pair = 0;
    int at = 0, shl = 0, ref = 0, xo = 0;
    const int mem = b->opts.mem;
    for (int s = 0; s < nsys; ++s) {
        SysInfo& si = b->info[s];
        std::memset(&si, 0, sizeof(si));
        const int nat = offsets[s + 1] - offsets[s];
        if (nat < 1) return fail("system with no atoms");
        if (nat > QX_MAX_NAT) return fail("system exceeds the supported atom count (64)");
        int nsh = 0, nao = 0, nd = 0;
        double nel = 0.0;
        for (int i = 0; i < nat; ++i) {
            const int z = num[offsets[s] + i];
            const int e = elem_slot(z);
            if (e < 0) {
                g_err = "element Z=" + std::to_string(z) + " has no GFN2 parameters in this build";
                return QCXMS_UNKNOWN_METHOD;
            }
            b->elem.push_back(e);
            nsh += gfn2::kNsh[e];
            nao += gfn2::kNsh[e] == 2 ? 4 : 1;
            nd += g_host_tab.e[e].nref;
            nel += gfn2::kValence[e];
        }
        if (nsh > QX_MAX_NSH || nao > QX_MAX_NAO) return fail("system exceeds the supported basis size");
        si.nat = nat;
        si.nsh = nsh;
        si.nao = nao;
        si.ld = (nao + 15) & ~15;
        si.ndim = nd;
        si.nx = nsh + 9 * nat;
        si.atom_off = at;
        si.shell_off = shl;
        si.ref_off = ref;
        si.x_off = xo;
        si.mat_off = mat;
        si.gam_off = gam;
        si.d4_off = d4;
        si.hist_off = hist;
        si.pair_off = pair;
        si.charge = charge ? charge[s] : 0;
        si.nel = nel - si.charge;
        si.kt = 300.0 * gfn2::kKtoAu;
        at += nat;
        shl += nsh;
        ref += nd;
        xo += si.nx;
        mat += (long long)si.ld * si.ld;
        gam += (long long)nsh * nsh;
        d4 += (long long)nd * nd;
        hist += 2LL * mem * si.nx + 2LL * mem * mem + 2LL * mem + 2LL * si.nx + 8;
        pair += (long long)nat * nat * qx::kPairStride;
        b->max_ld = std::max(b->max_ld, si.ld);
        b->max_m = std::max(b->max_m, (nao + 1) & ~1);
    }
    b->natoms = at;
    b->nshells = shl;
    b->nrefs = ref;
    b->nxtot = xo;
    b->mat_total = mat;
    b->gam_total = gam;
    b->d4_total = d4;
    b->hist_total = hist;
    b->pair_total = pair;
    return QCXMS_OK;
}

int allocate(qcxms_batch* b)
{
    BatchPtrs& d = b->d;
    int st;
#define ALLOC(ptr, n)                                   \
    if ((st = dalloc(b, &(ptr), (size_t)(n))) != 0) return st;
    ALLOC(b->d_info, b->nsys);
    ALLOC(b->d_elem, b->natoms);
    HIPCHK(hipMemcpyAsync(b->d_info, b->info.data(), sizeof(SysInfo) * b->nsys, hipMemcpyHostToDevice, g_stream));
    HIPCHK(hipMemcpyAsync(b->d_elem, b->elem.data(), sizeof(int) * b->natoms, hipMemcpyHostToDevice, g_stream));
    d.nsys = b->nsys;
    d.info = b->d_info;
    d.elem = b->d_elem;
    const size_t na = b->natoms, ns = b->nshells, nr = b->nrefs, mt = b->mat_total;
    ALLOC(d.xyz, 3 * na);
    ALLOC(d.grad, 3 * na);
    ALLOC(d.qat, na);
    ALLOC(d.cn, na);
    ALLOC(d.cnd4, na);
    ALLOC(d.mrad, na);
    ALLOC(d.dmrad, na);
    ALLOC(d.hself, ns);
    ALLOC(d.gw, nr);
    ALLOC(d.dgw, nr);
    ALLOC(d.S, mt);
    ALLOC(d.H0, mt);
    ALLOC(d.D, 3 * mt);
    ALLOC(d.Q, 6 * mt);
    ALLOC(d.L, mt);
    ALLOC(d.X, mt);
    ALLOC(d.C, mt);
    ALLOC(d.T, mt);
    ALLOC(d.P, mt);
    ALLOC(d.W, mt);
    ALLOC(d.U, mt);
    ALLOC(d.gam, b->gam_total);
    ALLOC(d.wd4, b->d4_total);
    ALLOC(d.x, b->nxtot);
    ALLOC(d.xo, b->nxtot);
    ALLOC(d.vs, ns);
    ALLOC(d.vd, 3 * na);
    ALLOC(d.vq, 6 * na);
    ALLOC(d.eps, (size_t)b->nsys * QX_MAX_NAO);
    ALLOC(d.focc, (size_t)b->nsys * QX_MAX_NAO);
    ALLOC(d.hist, b->hist_total);
    ALLOC(d.pairbuf, b->pair_total);
    ALLOC(d.energy, b->nsys);
    ALLOC(d.eparts, 10 * (size_t)b->nsys);
    int* iters;
    int* status;
    int* warm;
    ALLOC(iters, b->nsys);
    ALLOC(status, b->nsys);
    ALLOC(warm, b->nsys);
    d.iters = iters;
    d.status = status;
    d.warm = warm;
    d.active = nullptr;
#undef ALLOC
    return QCXMS_OK;
}

struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
};

int launch_egrad(qcxms_batch* b, bool with_grad)
{
    // systems whose matrices fit in LDS use dynamic LDS of their own order;
    // size it for the largest such order present (at most QX_LDS_MAXM)
    const size_t ls = std::min(b->max_ld, QX_LDS_MAXM), lm = std::min(b->max_m, QX_LDS_MAXM);
    const size_t lds_setup = sizeof(double) * ls * ls;
    const size_t lds_scc = 2 * sizeof(double) * lm * lm;
    hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr, e3 = nullptr;
    if (b->profiling) {
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(hipEventCreate(&e2));
        HIPCHK(hipEventCreate(&e3));
        HIPCHK(hipEventRecord(e0, g_stream));
    }
    if (!b->warm_start) HIPCHK(hipMemsetAsync(b->d.warm, 0, sizeof(int) * b->nsys, g_stream));
    hipLaunchKernelGGL(qx::k_setup, dim3(b->nsys), dim3(QX_BLOCK), lds_setup, g_stream, b->d, 1);
    HIPCHK(hipGetLastError());
    if (b->profiling) HIPCHK(hipEventRecord(e1, g_stream));
    hipLaunchKernelGGL(qx::k_scc, dim3(b->nsys), dim3(QX_BLOCK), lds_scc, g_stream, b->d, b->opts);
    HIPCHK(hipGetLastError());
    if (b->profiling) HIPCHK(hipEventRecord(e2, g_stream));
    if (with_grad) {
        hipLaunchKernelGGL(qx::k_grad, dim3(b->nsys), dim3(QX_BLOCK), 0, g_stream, b->d);
        HIPCHK(hipGetLastError());
    }
    if (b->profiling) {
        HIPCHK(hipEventRecord(e3, g_stream));
        HIPCHK(hipEventSynchronize(e3));
        float t1, t2, t3;
        HIPCHK(hipEventElapsedTime(&t1, e0, e1));
        HIPCHK(hipEventElapsedTime(&t2, e1, e2));
        HIPCHK(hipEventElapsedTime(&t3, e2, e3));
        b->t_setup.ms += t1;
        b->t_setup.launches += 1;
        b->t_scc.ms += t2;
        b->t_scc.launches += 1;
        if (with_grad) {
            b->t_grad.ms += t3;
            b->t_grad.launches += 1;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipEventDestroy(e2);
        (void)hipEventDestroy(e3);
    }
    return QCXMS_OK;
}

int set_etemp(qcxms_batch* b, const double* etemp)
{
    bool changed = false;
    for (int s = 0; s < b->nsys; ++s) {
        const double kt = (etemp ? etemp[s] : 300.0) * gfn2::kKtoAu;
        if (!(kt > 0.0)) return fail("electronic temperature must be positive");
        if (b->info[s].kt != kt) {
            b->info[s].kt = kt;
            changed = true;
        }
    }
    if (changed)
        HIPCHK(hipMemcpyAsync(b->d_info, b->info.data(), sizeof(SysInfo) * b->nsys, hipMemcpyHostToDevice,
                              g_stream));
    return QCXMS_OK;
}

}  // namespace

extern "C" {

int qcxms_gpu_init(int device) { return ensure_init(device); }

const char* qcxms_gpu_last_error(void) { return g_err.c_str(); }

qcxms_batch* qcxms_batch_create(int32_t nsys, const int32_t* offsets, const int32_t* num, const int32_t* charge,
                                int32_t method)
{
    g_err.clear();
    if (ensure_init(0) != QCXMS_OK) return nullptr;
    if (method != QCXMS_GFN2_XTB) {
        g_err = "only the GFN2-xTB Hamiltonian (method 2) is implemented";
        return nullptr;
    }
    if (nsys < 1 || !offsets || !num) {
        g_err = "invalid batch arguments";
        return nullptr;
    }
    qcxms_batch* b = new qcxms_batch();
    if (layout(b, nsys, offsets, num, charge) != QCXMS_OK || allocate(b) != QCXMS_OK) {
        delete b;
        return nullptr;
    }
    if (hipStreamSynchronize(g_stream) != hipSuccess) {
        g_err = "stream synchronisation failed during batch creation";
        delete b;
        return nullptr;
    }
    return b;
}

void qcxms_batch_destroy(qcxms_batch* b)
{
    if (!b) return;
    (void)hipStreamSynchronize(g_stream);
    delete b;
}

int qcxms_batch_set_scc(qcxms_batch* b, double econv, double pconv, int32_t maxiter, double damping, double jacobi_tol)
{
    if (!b) return fail("null batch");
    if (maxiter < 2 || maxiter > b->opts.mem) return fail("maxiter must lie in [2, 100]");
    b->opts.econv = econv;
    b->opts.pconv = pconv;
    b->opts.maxiter = maxiter;
    b->opts.damp = damping;
    b->opts.jacobi_tol = jacobi_tol;
    return QCXMS_OK;
}

int qcxms_batch_set_warm_start(qcxms_batch* b, int32_t warm)
{
    if (!b) return fail("null batch");
    b->warm_start = warm ? 1 : 0;
    return QCXMS_OK;
}

int qcxms_batch_set_profiling(qcxms_batch* b, int32_t enable)
{
    if (!b) return fail("null batch");
    b->profiling = enable ? 1 : 0;
    return QCXMS_OK;
}

int qcxms_batch_kernel_time(qcxms_batch* b, const char* kernel, double* ms, int32_t* launches)
{
    if (!b || !kernel) return fail("null argument");
    Timer* t = nullptr;
    if (!std::strcmp(kernel, "setup")) t = &b->t_setup;
    else if (!std::strcmp(kernel, "scc")) t = &b->t_scc;
    else if (!std::strcmp(kernel, "grad")) t = &b->t_grad;
    else if (!std::strcmp(kernel, "md_pre")) t = &b->t_md_pre;
    else if (!std::strcmp(kernel, "md_post")) t = &b->t_md_post;
    else return fail("unknown kernel name");
    if (ms) *ms = t->ms;
    if (launches) *launches = t->launches;
    *t = Timer();
    return QCXMS_OK;
}

int qcxms_batch_egrad(qcxms_batch* b, const double* xyz, const double* etemp, double* energy, double* gradient,
                      double* qat, int32_t* stat)
{
    if (!b || !xyz) return fail("null argument");
    int st = set_etemp(b, etemp);
    if (st) return st;
    HIPCHK(hipMemcpyAsync(b->d.xyz, xyz, sizeof(double) * 3 * b->natoms, hipMemcpyHostToDevice, g_stream));
    st = launch_egrad(b, gradient != nullptr);
    if (st) return st;
    if (energy) HIPCHK(hipMemcpyAsync(energy, b->d.energy, sizeof(double) * b->nsys, hipMemcpyDeviceToHost, g_stream));
    if (gradient)
        HIPCHK(hipMemcpyAsync(gradient, b->d.grad, sizeof(double) * 3 * b->natoms, hipMemcpyDeviceToHost, g_stream));
    if (qat) HIPCHK(hipMemcpyAsync(qat, b->d.qat, sizeof(double) * b->natoms, hipMemcpyDeviceToHost, g_stream));
    std::vector<int> hs(b->nsys);
    HIPCHK(hipMemcpyAsync(hs.data(), b->d.status, sizeof(int) * b->nsys, hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    int worst = QCXMS_OK;
    for (int s = 0; s < b->nsys; ++s) {
        if (stat) stat[s] = hs[s];
        if (hs[s] != 0) worst = QCXMS_FATAL;
    }
    return worst;
}

int qcxms_batch_info(qcxms_batch* b, int32_t* iterations, double* eparts)
{
    if (!b) return fail("null batch");
    if (iterations)
        HIPCHK(hipMemcpyAsync(iterations, b->d.iters, sizeof(int) * b->nsys, hipMemcpyDeviceToHost, g_stream));
    if (eparts)
        HIPCHK(hipMemcpyAsync(eparts, b->d.eparts, sizeof(double) * 10 * b->nsys, hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    return QCXMS_OK;
}

int qcxms_batch_multipoles(qcxms_batch* b, double* dipoles, double* quadrupoles)
{
    if (!b) return fail("null batch");
    std::vector<double> x(b->nxtot);
    HIPCHK(hipMemcpyAsync(x.data(), b->d.xo, sizeof(double) * b->nxtot, hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    for (int s = 0; s < b->nsys; ++s) {
        const SysInfo& si = b->info[s];
        for (int a = 0; a < si.nat; ++a) {
            for (int c = 0; c < 3; ++c)
                if (dipoles) dipoles[3 * (si.atom_off + a) + c] = x[si.x_off + si.nsh + 3 * a + c];
            for (int k = 0; k < 6; ++k)
                if (quadrupoles)
                    quadrupoles[6 * (si.atom_off + a) + k] = x[si.x_off + si.nsh + 3 * si.nat + 6 * a + k];
        }
    }
    return QCXMS_OK;
}

int qcxms_egrad_batch(int32_t nsys, const int32_t* offsets, const int32_t* num, const double* xyz,
                      const int32_t* charge, const int32_t* multiplicity, int32_t method, const double* etemp,
                      double* qat, double* energy, double* gradient, int32_t* stat)
{
    (void)multiplicity;  // uhf = min(mult-1, 0) = 0 for every multiplicity (tblite.f90:111)
    if (method != QCXMS_GFN2_XTB) {
        g_err = "only the GFN2-xTB Hamiltonian (method 2) is implemented";
        if (stat)
            for (int s = 0; s < nsys; ++s) stat[s] = QCXMS_UNKNOWN_METHOD;
        return QCXMS_UNKNOWN_METHOD;
    }
    qcxms_batch* b = qcxms_batch_create(nsys, offsets, num, charge, method);
    if (!b) {
        const bool elem = g_err.find("no GFN2 parameters") != std::string::npos;
        if (stat)
            for (int s = 0; s < nsys; ++s) stat[s] = elem ? QCXMS_UNKNOWN_METHOD : QCXMS_INVALID;
        return elem ? QCXMS_UNKNOWN_METHOD : QCXMS_INVALID;
    }
    const int st = qcxms_batch_egrad(b, xyz, etemp, energy, gradient, qat, stat);
    qcxms_batch_destroy(b);
    return st;
}

int qcxms_get_xtb_egrad(int32_t nat, const int32_t* num, const double* xyz, int32_t charge, int32_t multiplicity,
                        int32_t method, double etemp, double* qat, double* energy, double* gradient, int32_t* stat)
{
    const int32_t offs[2] = {0, nat};
    int32_t s = 0;
    const int st = qcxms_egrad_batch(1, offs, num, xyz, &charge, &multiplicity, method, &etemp, qat, energy,
                                     gradient, &s);
    const int out = (st == QCXMS_UNKNOWN_METHOD) ? QCXMS_UNKNOWN_METHOD : (st == QCXMS_OK ? s : QCXMS_FATAL);
    if (stat) *stat = out;
    return out;
}

}  // extern "C"
