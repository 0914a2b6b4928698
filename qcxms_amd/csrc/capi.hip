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
#include "radii.h"
#include "sto_ng_table.h"

__constant__ DevTables g_tab;
__constant__ double g_rad[118];

// Device memory of the allocation-free entry point qcxms_egrad_batch_noalloc:
// the whole batch layout of one call is carved from this array.  Every GPU
// the library runs on holds its own copy until the process ends.  Host code
// never uses this name as an address: it takes the device address of the
// array from hipGetSymbolAddress once per device (g_scratch_dev).
constexpr size_t kScratchBytes = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// zero n doubles (a kernel rather than a memset node, so that callers may
// capture the launch)
__global__ __launch_bounds__(256) void k_fill_zero(double* p, size_t n)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0.0;
}

#include "k_grad.h"
#include "k_md.h"
#include "k_spec.h"
#include "k_cid.h"
#include "tuning_table.h"

namespace {

thread_local std::string g_err;
std::mutex g_mu;
bool g_ready = false;
int g_device = -1;
hipStream_t g_stream = nullptr;
size_t g_scc_dyn_max = 0;  // dynamic LDS the SCC kernels may take (set at init)
unsigned char* g_scratch_dev = nullptr;  // device address of scratch on g_device
std::mutex g_scratch_mu;                 // one allocation-free call at a time
size_t g_eig_dyn_max = 0;  // the same with the MFMA eigen refinement's fifth matrix
DevTables g_host_tab;

// Launch choices of the SCC for single-point batches by basis size: bin k
// takes the systems with nao <= max_nao (the first bin that fits); eig: MFMA
// eigen refinement, lds: SCC matrices in LDS where they fit.  A batch whose
// systems fall into one bin takes that bin's choice; a batch spanning several
// bins takes the mixed choice in one launch (eig >= 0) or one launch per run
// of systems of one bin (eig < 0).  The compiled-in table is the MI355X
// autotuner's (tuning_table.h); qcxms_set_tuning replaces it for batches
// created afterwards.  No bins: one launch with the batch's own options.
struct TuneBin {
    int max_nao, eig, lds;
};
std::vector<TuneBin> g_tune;
TuneBin g_tune_mixed{0, -1, 0};
bool g_tune_init = false;

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
    double rad[118];
    for (int z = 0; z < 118; ++z) rad[z] = gfn2::kAatoAu * kQcxmsRadAa[z];
    HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(g_rad), rad, sizeof(rad)));
    HIPCHK(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
    // the allocation-free entry point's array on this device
    void* sp = nullptr;
    HIPCHK(hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)));
    if (!sp) return fail("no device address for the scratch array");
    g_scratch_dev = (unsigned char*)sp;
    // dynamic LDS of the SCC kernels: whatever the workgroup's LDS leaves
    // beside the kernels' static LDS (at least the LDS-resident matrices of
    // kFastM basis functions)
    int lds_block = 0;
    HIPCHK(hipDeviceGetAttribute(&lds_block, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
    hipFuncAttributes fa;
    HIPCHK(hipFuncGetAttributes(&fa, (const void*)qx::k_scc<false>));
    size_t lds_static = fa.sharedSizeBytes;
    HIPCHK(hipFuncGetAttributes(&fa, (const void*)qx::k_scc<true>));
    lds_static = std::max(lds_static, fa.sharedSizeBytes);
    HIPCHK(hipFuncGetAttributes(&fa, (const void*)qx::k_md_fused<false>));
    lds_static = std::max(lds_static, fa.sharedSizeBytes);
    HIPCHK(hipFuncGetAttributes(&fa, (const void*)qx::k_md_fused<true>));
    lds_static = std::max(lds_static, fa.sharedSizeBytes);
    const size_t base_dyn = (4 * qx::kFastM * (qx::kFastM + 1) + qx::kRs * qx::kFastM) * sizeof(double);
    if ((size_t)lds_block < lds_static + base_dyn) return fail("the device's LDS per workgroup is too small");
    g_scc_dyn_max = (size_t)lds_block - lds_static;
    // the optional fifth matrix (MFMA eigen refinement) only where the
    // workgroup keeps 8 KiB of its LDS free: workgroups of up to 153 KB are
    // what the MI355X runs have exercised (m <= 50)
    constexpr size_t kLdsReserve = 8192;
    g_eig_dyn_max = g_scc_dyn_max > kLdsReserve ? g_scc_dyn_max - kLdsReserve : 0;
    HIPCHK(hipFuncSetAttribute((const void*)qx::k_scc<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)g_scc_dyn_max));
    HIPCHK(hipFuncSetAttribute((const void*)qx::k_scc<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)g_scc_dyn_max));
    HIPCHK(hipFuncSetAttribute((const void*)qx::k_setup, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (QX_LDS_MAXM * QX_LDS_MAXM + QX_LDS_MAXM * (QX_LDS_MAXM | 1)) *
                                   (int)sizeof(double)));
    HIPCHK(hipFuncSetAttribute((const void*)qx::k_grad, hipFuncAttributeMaxDynamicSharedMemorySize, 0));
    HIPCHK(hipFuncSetAttribute((const void*)qx::k_md_fused<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)g_scc_dyn_max));
    HIPCHK(hipFuncSetAttribute((const void*)qx::k_md_fused<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)g_scc_dyn_max));
    g_device = device;
    if (!g_tune_init) {
        for (const auto& t : qx::kTuneTable) g_tune.push_back({t[0], t[1], t[2]});
        g_tune_mixed = {0, qx::kTuneMixed[0], qx::kTuneMixed[1]};
        g_tune_init = true;
    }
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
    std::vector<int2> items;  // shell-pair work items of all systems
    int natoms = 0, nshells = 0, nrefs = 0, nxtot = 0;
    size_t mat_total = 0, gam_total = 0, d4_total = 0, hist_total = 0, pair_total = 0;
    int max_ld = 0, max_m = 0;
    SccOpts opts{1e-9, 1e-8, 0.4, 1e-11, 100, 100, 0, 0, 0, 0};
    int warm_start = 0;
    int profiling = 0;
    Timer t_setup, t_scc, t_grad, t_md_pre, t_md_post, t_fused, t_dblk;
    std::vector<void*> allocs;
    size_t alloc_bytes = 0;       // device memory the batch allocated
    double* dblk_mem = nullptr;   // stored derivative blocks (allocated on request)
    // SCC launch plan by basis-size bin (single-point batches; empty: one
    // launch): side streams forked from and joined to `stream`
    std::vector<TuneBin> tune;
    TuneBin tune_mixed{0, -1, 0};
    static constexpr int kSide = 4;
    hipStream_t side[kSide] = {};
    hipEvent_t fork = nullptr, join[kSide] = {};
    BatchPtrs d{};
    SysInfo* d_info = nullptr;
    int* d_elem = nullptr;
    // every batch (and the MD ensemble built on it) runs on a stream of its
    // own, so that the kernels of several ensembles overlap on the device
    hipStream_t stream = nullptr;
    bool owns_stream = true;
    // arena != null: device arrays are carved from [arena, arena + arena_cap)
    // instead of allocated (the allocation-free entry point)
    unsigned char* arena = nullptr;
    size_t arena_used = 0, arena_cap = 0;

    ~qcxms_batch()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        for (void* p : allocs) (void)hipFree(p);
        for (int k = 0; k < kSide; ++k) {
            if (side[k]) (void)hipStreamDestroy(side[k]);
            if (join[k]) (void)hipEventDestroy(join[k]);
        }
        if (fork) (void)hipEventDestroy(fork);
        if (stream && owns_stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

template <class T>
int dalloc(qcxms_batch* b, T** p, size_t n)
{
    if (n == 0) n = 1;
    if (b->arena) {
        // 256-byte aligned slices of the arena, zeroed together afterwards
        const size_t off = (b->arena_used + 255) & ~(size_t)255, bytes = n * sizeof(T);
        if (off + bytes > b->arena_cap)
            return fail("the batch needs more than the " + std::to_string(b->arena_cap >> 20) +
                        " MiB scratch of the allocation-free entry point; use qcxms_egrad_batch");
        *p = (T*)(b->arena + off);
        b->arena_used = off + bytes;
        return QCXMS_OK;
    }
    HIPCHK(hipMalloc((void**)p, n * sizeof(T)));
    HIPCHK(hipMemsetAsync(*p, 0, n * sizeof(T), b->stream));
    b->allocs.push_back((void*)*p);
    b->alloc_bytes += n * sizeof(T);
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
        // shell-pair work items: off-atom pairs (atom of ia > atom of ib)
        // grouped p-p, p-s, s-p, s-s so that wavefronts run one code path,
        // then the same-atom pairs (ia >= ib)
        {
            std::vector<int> sh_at, sh_l;
            for (int i = 0; i < nat; ++i) {
                const int e = b->elem[b->elem.size() - nat + i];
                for (int k = 0; k < gfn2::kNsh[e]; ++k) {
                    sh_at.push_back(i);
                    sh_l.push_back(k);
                }
            }
            si.item_off = (int)b->items.size();
            for (int cls = 3; cls >= 0; --cls)
                for (int ia = 0; ia < nsh; ++ia)
                    for (int ib = 0; ib < ia; ++ib)
                        if (sh_at[ia] > sh_at[ib] && sh_l[ia] * 2 + sh_l[ib] == cls)
                            b->items.push_back(make_int2(ia, ib));
            si.nitem_off = (int)b->items.size() - si.item_off;
            for (int ia = 0; ia < nsh; ++ia)
                for (int ib = 0; ib <= ia; ++ib)
                    if (sh_at[ia] == sh_at[ib]) b->items.push_back(make_int2(ia, ib));
            si.nitem = (int)b->items.size() - si.item_off;
            // the same pairs as rows of the first shell (a p shell has three),
            // row index in the high half of x; rows over a p second shell
            // (three AO pairs each) first
            si.rowitem_off = (int)b->items.size();
            for (int pass = 0; pass < 2; ++pass)
                for (int it = si.item_off; it < si.item_off + si.nitem; ++it) {
                    const int2 pr = b->items[it];
                    if ((sh_l[pr.y] != 0) != (pass == 0)) continue;
                    const int nrow = sh_l[pr.x] != 0 ? 3 : 1;
                    for (int r = 0; r < nrow; ++r) b->items.push_back(make_int2(pr.x | (r << 16), pr.y));
                }
            si.nrowitem = (int)b->items.size() - si.rowitem_off;
        }
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
    int2* d_items = nullptr;
    ALLOC(d_items, b->items.size());
    d.spitem = d_items;
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
    ALLOC(d.E, mt);
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
    ALLOC(d.spbuf, 5 * b->gam_total);
    ALLOC(d.aesg, b->pair_total / QX_PAIR_STRIDE * QX_AES_COMP);
    ALLOC(d.xprev, b->nxtot);
    ALLOC(d.xprev2, b->nxtot);
    ALLOC(d.xprev3, b->nxtot);
    ALLOC(d.xprev4, b->nxtot);
    ALLOC(d.energy, b->nsys);
    ALLOC(d.eparts, 10 * (size_t)b->nsys);
    int* iters;
    int* status;
    int* warm;
    ALLOC(iters, b->nsys);
    ALLOC(status, b->nsys);
    ALLOC(warm, b->nsys);
    unsigned long long* itot;
    ALLOC(itot, b->nsys);
    d.iters = iters;
    d.iters_total = itot;
    d.status = status;
    d.warm = warm;
    d.active = nullptr;
#undef ALLOC
    if (b->arena) {
        const size_t nd = (b->arena_used + 7) / 8;
        hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)std::min<size_t>((nd + 255) / 256, 4096)), dim3(256), 0,
                           b->stream, (double*)b->arena, nd);
        HIPCHK(hipGetLastError());
    }
    if (!b->items.empty())
        HIPCHK(hipMemcpyAsync(d_items, b->items.data(), sizeof(int2) * b->items.size(), hipMemcpyHostToDevice,
                              b->stream));
    HIPCHK(hipMemcpyAsync(b->d_info, b->info.data(), sizeof(SysInfo) * b->nsys, hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipMemcpyAsync(b->d_elem, b->elem.data(), sizeof(int) * b->natoms, hipMemcpyHostToDevice, b->stream));
    return QCXMS_OK;
}

struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
};

// enqueue one kernel, optionally bracketed by events whose elapsed time is
// added to timer t after the next synchronisation point
struct PendingTime {
    hipEvent_t a, b;
    Timer* t;
};
thread_local std::vector<PendingTime> g_pending;

template <class F>
int timed(qcxms_batch* b, Timer& t, F&& launch)
{
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (b->profiling) {
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(hipEventRecord(e0, b->stream));
    }
    launch();
    HIPCHK(hipGetLastError());
    if (b->profiling) {
        HIPCHK(hipEventRecord(e1, b->stream));
        g_pending.push_back({e0, e1, &t});
    }
    return QCXMS_OK;
}

int collect_times()
{
    for (auto& p : g_pending) {
        HIPCHK(hipEventSynchronize(p.b));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, p.a, p.b));
        p.t->ms += ms;
        p.t->launches += 1;
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    g_pending.clear();
    return QCXMS_OK;
}

size_t lds_setup_bytes(const qcxms_batch* b)
{
    // Cholesky factor and the columns of its inverse
    const size_t ls = std::min(b->max_ld, QX_LDS_MAXM);
    return sizeof(double) * (ls * ls + ls * (ls | 1));
}
// four m x m working matrices and the AO row sums for the systems that keep
// them in LDS
size_t lds_scc_bytes_range(const qcxms_batch* b, int s0, int s1, int lds_max_m, int eig)
{
    size_t lm = 0;
    for (int s = s0; s < s1; ++s) {
        const size_t m = (size_t)((b->info[s].nao + 1) & ~1);
        if (m <= (size_t)lds_max_m) lm = std::max(lm, m);
    }
    const size_t base = sizeof(double) * (4 * lm * (lm + 1) + qx::kRs * lm);
    // a fifth matrix for the MFMA eigen refinement when it is on and the LDS
    // holds it
    const size_t ext = base + sizeof(double) * lm * (lm + 1);
    return eig && ext <= g_eig_dyn_max ? ext : base;
}

size_t lds_scc_bytes(const qcxms_batch* b)
{
    return lds_scc_bytes_range(b, 0, b->nsys, qx::kFastM, b->opts.eig_mfma);
}

// the SCC options of one launch with its dynamic LDS size
SccOpts launch_opts(const qcxms_batch* b, size_t lds)
{
    SccOpts o = b->opts;
    o.lds_doubles = (int)(lds / sizeof(double));
    o.lds_max_m = qx::kFastM;
    o.first_sys = 0;
    return o;
}

// bin of a basis size in a tuning table (the last bin takes what is larger)
int tune_bin(const std::vector<TuneBin>& t, int nao)
{
    for (size_t k = 0; k < t.size(); ++k)
        if (nao <= t[k].max_nao) return (int)k;
    return (int)t.size() - 1;
}

// The SCC by launch plan: contiguous runs of systems in one bin, each run
// one launch with its bin's options and LDS sized for its own systems, the
// runs spread over side streams forked from the batch stream and joined back.
int enqueue_scc_plan(qcxms_batch* b)
{
    struct Run {
        int s0, s1, bin;
    };
    std::vector<Run> runs;
    for (int s = 0; s < b->nsys; ++s) {
        const int k = tune_bin(b->tune, b->info[s].nao);
        if (runs.empty() || runs.back().bin != k) runs.push_back({s, s + 1, k});
        else runs.back().s1 = s + 1;
    }
    auto launch = [&](const Run& r, hipStream_t st) {
        const TuneBin& t = r.bin >= 0 ? b->tune[r.bin] : b->tune_mixed;
        const int lmax = t.lds ? qx::kFastM : 0;
        const size_t lds = lds_scc_bytes_range(b, r.s0, r.s1, lmax, t.eig);
        SccOpts o = b->opts;
        o.eig_mfma = t.eig;
        o.lds_doubles = (int)(lds / sizeof(double));
        o.lds_max_m = lmax;
        o.first_sys = r.s0;
        hipLaunchKernelGGL(t.eig ? qx::k_scc<true> : qx::k_scc<false>, dim3(r.s1 - r.s0), dim3(QX_SCC_BLOCK),
                           lds, st, b->d, o);
        return hipGetLastError();
    };
    if (runs.size() == 1) {
        HIPCHK(launch(runs[0], b->stream));
        return QCXMS_OK;
    }
    if (b->tune_mixed.eig >= 0) {
        // systems of several bins in one launch with the mixed choice
        HIPCHK(launch(Run{0, b->nsys, -1}, b->stream));
        return QCXMS_OK;
    }
    if (!b->fork) {
        HIPCHK(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
        for (int k = 0; k < qcxms_batch::kSide; ++k) {
            HIPCHK(hipStreamCreateWithFlags(&b->side[k], hipStreamNonBlocking));
            HIPCHK(hipEventCreateWithFlags(&b->join[k], hipEventDisableTiming));
        }
    }
    HIPCHK(hipEventRecord(b->fork, b->stream));
    const int nside = std::min<int>((int)runs.size(), qcxms_batch::kSide);
    for (int k = 0; k < nside; ++k) HIPCHK(hipStreamWaitEvent(b->side[k], b->fork, 0));
    // largest systems first (the batch's order), dealt round-robin
    for (size_t r = 0; r < runs.size(); ++r) HIPCHK(launch(runs[r], b->side[r % nside]));
    for (int k = 0; k < nside; ++k) {
        HIPCHK(hipEventRecord(b->join[k], b->side[k]));
        HIPCHK(hipStreamWaitEvent(b->stream, b->join[k], 0));
    }
    return QCXMS_OK;
}

// setup -> SCC -> gradient for every system of the batch (masked by d.active)
int enqueue_egrad(qcxms_batch* b, bool with_grad)
{
    int st;
    const size_t l1 = lds_setup_bytes(b), l2 = lds_scc_bytes(b);
    if ((st = timed(b, b->t_setup, [&] {
             hipLaunchKernelGGL(qx::k_setup, dim3(b->nsys), dim3(QX_BLOCK), l1, b->stream, b->d, 1);
         })))
        return st;
    int plan_st = QCXMS_OK;
    if ((st = timed(b, b->t_scc, [&] {
             if (!b->tune.empty()) plan_st = enqueue_scc_plan(b);
             else
                 hipLaunchKernelGGL(b->opts.eig_mfma ? qx::k_scc<true> : qx::k_scc<false>, dim3(b->nsys),
                                    dim3(QX_SCC_BLOCK), l2, b->stream, b->d, launch_opts(b, l2));
         })))
        return st;
    if (plan_st) return plan_st;
    if (with_grad && b->d.dblk &&
        (st = timed(b, b->t_dblk, [&] {
             hipLaunchKernelGGL(qx::k_dblocks, dim3(b->nsys), dim3(QX_BLOCK), 0, b->stream, b->d);
         })))
        return st;
    if (with_grad &&
        (st = timed(b, b->t_grad, [&] {
             hipLaunchKernelGGL(qx::k_grad, dim3(b->nsys), dim3(QX_BLOCK), 0, b->stream, b->d);
         })))
        return st;
    return QCXMS_OK;
}

int launch_egrad(qcxms_batch* b, bool with_grad)
{
    // (an arena batch lives for one call and starts zeroed: no warm flags)
    if (!b->warm_start && !b->arena) HIPCHK(hipMemsetAsync(b->d.warm, 0, sizeof(int) * b->nsys, b->stream));
    int st = enqueue_egrad(b, with_grad);
    if (st) return st;
    return b->profiling ? collect_times() : QCXMS_OK;
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
                              b->stream));
    return QCXMS_OK;
}

}  // namespace


struct qcxms_md {
    qcxms_batch* b = nullptr;
    qx::MdParams p{};
    qx::MdPtrs m{};
    int ntraj = 0;
    bool use_graph = false, graph_ready = false;
    bool fused = false;  // persistent kernel: whole MD iterations per launch
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<void*> allocs;
    ~qcxms_md()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (void* p : allocs) (void)hipFree(p);
        delete b;
    }
};

namespace {

template <class T>
int mdalloc(qcxms_md* md, T** p, size_t n, const T* host = nullptr)
{
    if (n == 0) n = 1;
    HIPCHK(hipMalloc((void**)p, n * sizeof(T)));
    md->allocs.push_back((void*)*p);
    if (host) HIPCHK(hipMemcpyAsync(*p, host, n * sizeof(T), hipMemcpyHostToDevice, md->b->stream));
    else HIPCHK(hipMemsetAsync(*p, 0, n * sizeof(T), md->b->stream));
    return QCXMS_OK;
}

// one MD loop iteration for all running trajectories
int enqueue_md_step(qcxms_md* md)
{
    qcxms_batch* b = md->b;
    int st;
    if ((st = timed(b, b->t_md_pre, [&] {
             hipLaunchKernelGGL(qx::k_md_pre, dim3(md->ntraj), dim3(64), 0, b->stream, b->d, md->m, md->p);
         })))
        return st;
    if ((st = enqueue_egrad(b, true))) return st;
    return timed(b, b->t_md_post, [&] {
        hipLaunchKernelGGL(qx::k_md_post, dim3(md->ntraj), dim3(QX_BLOCK), 0, b->stream, b->d, md->m, md->p);
    });
}

double host_setetemp(int nfrag, double eimp, double ax, double ieetemp)
{
    double etemp = (double)5000.0f + (double)20000.0f * ax;
    if (eimp > 0 && nfrag <= 1) etemp = etemp + (eimp > 0.0 ? eimp : 0.0) * ieetemp;
    return etemp;
}

struct TmpBuf {
    std::vector<void*> v;
    ~TmpBuf()
    {
        for (void* p : v) (void)hipFree(p);
    }
    template <class T>
    int up(T** d, const T* h, size_t n)
    {
        HIPCHK(hipMalloc((void**)d, std::max<size_t>(n, 1) * sizeof(T)));
        v.push_back((void*)*d);
        if (h) HIPCHK(hipMemcpyAsync(*d, h, n * sizeof(T), hipMemcpyHostToDevice, g_stream));
        return QCXMS_OK;
    }
};

int check_offsets(int nsys, const int32_t* offsets, int maxnat)
{
    if (nsys < 1 || !offsets) return fail("invalid system count or offsets");
    for (int s = 0; s < nsys; ++s) {
        const int n = offsets[s + 1] - offsets[s];
        if (n < 1 || n > maxnat) return fail("system size outside 1..64 atoms");
    }
    return QCXMS_OK;
}

// Host description of an MD ensemble to build: systems in launch order with
// their atoms (offsets/num/charge), per-atom start data and MD records, and
// for every system the system of an older ensemble whose state it takes
// over (src >= 0; its host data here are then placeholders) or -1 (a new
// system that starts with an initial energy and gradient, md.f90:182-184).
struct MdLayout {
    std::vector<int32_t> offsets, num, charge, imass, src;
    std::vector<double> mass, xyz, velo, velof;
    std::vector<qx::MdTraj> traj;
};

// MD record of a new trajectory generation (md.f90:88-260)
qx::MdTraj new_md_record(const qcxms_md_params& p, double eimp, double tadd, int isec, int max_steps, double ttime)
{
    qx::MdTraj t;
    std::memset(&t, 0, sizeof(t));
    t.eimp = eimp;
    t.tadd = tadd;
    t.nfrag = 1;
    t.isec = isec;
    t.max_steps = max_steps;
    t.ttime = ttime;
    t.etemp = (p.etempin < 0.0) ? host_setetemp(1, eimp, p.ax, p.ieetemp) : p.etempin;
    return t;
}

__global__ void k_set_active(int* active, const int* src, int n)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n && src[s] >= 0) active[s] = 1;
}

// Build an ensemble from L (see MdLayout); old supplies the state of the
// systems with src >= 0 (gathered on the device by k_md_repack), params the
// MD and SCC options when old is null.
qcxms_md* md_build(const qcxms_md* old, const MdLayout& L, const qcxms_md_params* params)
{
    const int ntraj = (int)L.src.size();
    qcxms_batch* b = qcxms_batch_create(ntraj, L.offsets.data(), L.num.data(), L.charge.data(), QCXMS_GFN2_XTB);
    if (b) b->tune.clear();  // the MD kernels run one launch per step (or a persistent one)
    if (!b) return nullptr;
    qcxms_md* md = new qcxms_md();
    md->b = b;
    md->ntraj = ntraj;
    if (old) {
        md->use_graph = old->use_graph;
        md->fused = old->fused;
        md->p = old->p;
        b->opts = old->b->opts;
        b->profiling = old->b->profiling;
    } else {
        md->use_graph = params->use_graph != 0;
        md->fused = params->fused != 0;
        b->opts.eig_mfma = params->eigensolver == 2 ? 0 : 1;
        md->p.tstep = params->tstep;
        md->p.etempin = params->etempin;
        md->p.ax = params->ax;
        md->p.ieetemp = params->ieetemp;
        md->p.nfragexit = params->nfragexit;
        md->p.max_steps = params->max_steps;
        b->opts.keep = params->broyden_keep;
        // warm-started MD steps start close to the fixed point: a lighter
        // first mixing step (0.3 instead of 0.4) takes fewer SCC cycles
        b->opts.damp = 0.3;
        b->opts.extrapolate = params->extrapolate == 0 ? QX_DEFAULT_EXTRAPOLATE
                                                       : (params->extrapolate < 0 ? 0 : params->extrapolate);
        if (params->econv > 0.0) b->opts.econv = params->econv;
        if (params->pconv > 0.0) b->opts.pconv = params->pconv;
    }
    b->warm_start = 1;
    const int na = b->natoms;
    std::vector<int> act(ntraj), nm(L.num.begin(), L.num.end()), im(L.imass.begin(), L.imass.end());
    bool any_new = false;
    for (int s = 0; s < ntraj; ++s) {
        act[s] = L.src[s] < 0 ? 1 : 0;
        any_new = any_new || act[s];
        if (L.src[s] < 0) b->info[s].kt = L.traj[s].etemp * gfn2::kKtoAu;
    }
    std::vector<int> lst(na, 1);
    int st = QCXMS_OK;
    double* d_velof;
    double* d_mass;
    int* d_imass;
    int* d_num;
    int* d_src;
    if ((st = mdalloc(md, &md->m.traj, ntraj, L.traj.data())) ||
        (st = mdalloc(md, &md->m.velo, 3 * (size_t)na, L.velo.data())) ||
        (st = mdalloc(md, &d_velof, (size_t)na, L.velof.data())) ||
        (st = mdalloc(md, &d_mass, (size_t)na, L.mass.data())) ||
        (st = mdalloc(md, &d_imass, (size_t)na, im.data())) || (st = mdalloc(md, &d_num, (size_t)na, nm.data())) ||
        (st = mdalloc(md, &md->m.list, (size_t)na, lst.data())) || (st = mdalloc(md, &md->m.avchrg, (size_t)na)) ||
        (st = mdalloc(md, &md->m.avxyz, 3 * (size_t)na)) || (st = mdalloc(md, &md->m.avxyz2, 3 * (size_t)na)) ||
        (st = mdalloc(md, &md->m.store, 3 * (size_t)na)) ||
        (st = mdalloc(md, &md->m.active, (size_t)ntraj, act.data())) ||
        (st = mdalloc(md, &d_src, (size_t)ntraj, L.src.data()))) {
        delete md;
        return nullptr;
    }
    md->m.velof = d_velof;
    md->m.mass = d_mass;
    md->m.imass = d_imass;
    md->m.num = d_num;
    b->d.active = md->m.active;
    if (hipMemcpyAsync(b->d.xyz, L.xyz.data(), sizeof(double) * 3 * na, hipMemcpyHostToDevice, b->stream) !=
            hipSuccess ||
        hipMemcpyAsync(b->d_info, b->info.data(), sizeof(SysInfo) * ntraj, hipMemcpyHostToDevice, b->stream) !=
            hipSuccess) {
        g_err = "upload failed";
        delete md;
        return nullptr;
    }
    if (old) {
        // the old ensemble's work must be done before its state is read
        if (hipStreamSynchronize(old->b->stream) != hipSuccess) {
            g_err = "synchronisation with the previous ensemble failed";
            delete md;
            return nullptr;
        }
        hipLaunchKernelGGL(qx::k_md_repack, dim3(ntraj), dim3(64), 0, b->stream, old->b->d, old->m, b->d, md->m,
                           d_src, b->opts.keep > 0 ? 1 : 0, b->opts.mem);
        if (hipGetLastError() != hipSuccess) {
            g_err = "re-layout launch failed";
            delete md;
            return nullptr;
        }
    }
    // the initial energy and gradient of the new systems only (md.f90:182)
    if (any_new) {
        if (enqueue_egrad(b, true) != QCXMS_OK) {
            delete md;
            return nullptr;
        }
        hipLaunchKernelGGL(qx::k_md_init, dim3(ntraj), dim3(64), 0, b->stream, b->d, md->m, md->p);
    }
    hipLaunchKernelGGL(k_set_active, dim3((ntraj + 255) / 256), dim3(256), 0, b->stream, md->m.active, d_src, ntraj);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(b->stream) != hipSuccess) {
        g_err = "initial energy/gradient failed on the device";
        delete md;
        return nullptr;
    }
    if (b->profiling) collect_times();
    return md;
}

}  // namespace

extern "C" {

int qcxms_gpu_init(int device) { return ensure_init(device); }

#ifdef QX_PHASE_TIMING
// per-phase SCC cycle counters of the first systems (instrumented builds only)
int qcxms_debug_phase_read(unsigned long long* out, int nsys, int reset)
{
    if (!g_ready) return fail("library not initialised");
    const int n = std::min(nsys, qx::kPhaseSys) * qx::kPhaseSlots;
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpyFromSymbol(out, HIP_SYMBOL(qx::g_phase), sizeof(unsigned long long) * n));
    if (reset) {
        std::vector<unsigned long long> z(qx::kPhaseSys * qx::kPhaseSlots, 0ull);
        HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(qx::g_phase), z.data(), sizeof(unsigned long long) * z.size()));
    }
    return QCXMS_OK;
}
#endif

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
    {
        std::lock_guard<std::mutex> lk(g_mu);
        b->tune = g_tune;
        b->tune_mixed = g_tune_mixed;
    }
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        g_err = "stream creation failed";
        b->stream = nullptr;
        delete b;
        return nullptr;
    }
    if (layout(b, nsys, offsets, num, charge) != QCXMS_OK || allocate(b) != QCXMS_OK) {
        delete b;
        return nullptr;
    }
    if (hipStreamSynchronize(b->stream) != hipSuccess) {
        g_err = "stream synchronisation failed during batch creation";
        delete b;
        return nullptr;
    }
    return b;
}

void qcxms_batch_destroy(qcxms_batch* b)
{
    if (!b) return;
    (void)hipStreamSynchronize(b->stream);
    delete b;
}

int qcxms_batch_set_scc(qcxms_batch* b, double econv, double pconv, int32_t maxiter, double damping, double jacobi_tol)
{
    if (!b) return fail("null batch");
    if (maxiter < 2 || maxiter > b->opts.mem) return fail("maxiter must lie in [2, 100]");
    if (!(econv > 0.0) || !(pconv > 0.0) || !(jacobi_tol > 0.0))
        return fail("econv, pconv and jacobi_tol must be positive");
    if (!(damping > 0.0 && damping <= 1.0)) return fail("damping must lie in (0, 1]");
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
    else if (!std::strcmp(kernel, "fused")) t = &b->t_fused;
    else if (!std::strcmp(kernel, "dblocks")) t = &b->t_dblk;
    else return fail("unknown kernel name");
    if (ms) *ms = t->ms;
    if (launches) *launches = t->launches;
    *t = Timer();
    return QCXMS_OK;
}

int qcxms_set_tuning(int32_t nbins, const int32_t* max_nao, const int32_t* eig, const int32_t* lds,
                     int32_t mixed_eig, int32_t mixed_lds)
{
    if (mixed_eig < -1 || mixed_eig > 1 || (mixed_lds != 0 && mixed_lds != 1))
        return fail("tuning table: the mixed choice is eig -1 (binned launches), 0 or 1 and lds 0 or 1");
    if (nbins < 0 || nbins > 32) return fail("tuning table: 0 to 32 bins");
    if (nbins > 0 && (!max_nao || !eig || !lds)) return fail("null argument");
    std::vector<TuneBin> t;
    for (int k = 0; k < nbins; ++k) {
        if (max_nao[k] < 1 || (k > 0 && max_nao[k] <= max_nao[k - 1]))
            return fail("tuning table: bin limits must be positive and increasing");
        if ((eig[k] != 0 && eig[k] != 1) || (lds[k] != 0 && lds[k] != 1))
            return fail("tuning table: eig and lds are 0 or 1");
        t.push_back({max_nao[k], eig[k], lds[k]});
    }
    const int st = ensure_init(0);
    if (st) return st;
    std::lock_guard<std::mutex> lk(g_mu);
    g_tune = t;
    g_tune_mixed = {0, mixed_eig, mixed_lds};
    g_tune_init = true;
    return QCXMS_OK;
}

int qcxms_get_tuning(int32_t cap, int32_t* nbins, int32_t* max_nao, int32_t* eig, int32_t* lds,
                     int32_t* mixed_eig, int32_t* mixed_lds)
{
    if (!nbins) return fail("null argument");
    const int st = ensure_init(0);
    if (st) return st;
    std::lock_guard<std::mutex> lk(g_mu);
    *nbins = (int32_t)g_tune.size();
    if (mixed_eig) *mixed_eig = g_tune_mixed.eig;
    if (mixed_lds) *mixed_lds = g_tune_mixed.lds;
    if (cap < (int)g_tune.size()) return g_tune.empty() ? QCXMS_OK : fail("tuning buffers too small");
    for (size_t k = 0; k < g_tune.size(); ++k) {
        if (max_nao) max_nao[k] = g_tune[k].max_nao;
        if (eig) eig[k] = g_tune[k].eig;
        if (lds) lds[k] = g_tune[k].lds;
    }
    return QCXMS_OK;
}

int qcxms_batch_set_stored_derivatives(qcxms_batch* b, int32_t on)
{
    if (!b) return fail("null argument");
    if (b->arena) return fail("stored derivative blocks need a batch with its own device memory");
    if (on && !b->dblk_mem) {
        // 9 AO-pair slots of qx::kDblk values per shell-pair work item
        const size_t n = b->items.size() * 9 * (size_t)qx::kDblk;
        HIPCHK(hipMalloc((void**)&b->dblk_mem, n * sizeof(double)));
        b->allocs.push_back((void*)b->dblk_mem);
        b->alloc_bytes += n * sizeof(double);
    }
    b->d.dblk = on ? b->dblk_mem : nullptr;
    return QCXMS_OK;
}

int qcxms_batch_device_bytes(qcxms_batch* b, uint64_t* bytes)
{
    if (!b || !bytes) return fail("null argument");
    *bytes = b->alloc_bytes;
    return QCXMS_OK;
}

int qcxms_batch_egrad(qcxms_batch* b, const double* xyz, const double* etemp, double* energy, double* gradient,
                      double* qat, int32_t* stat)
{
    if (!b || !xyz) return fail("null argument");
    int st = set_etemp(b, etemp);
    if (st) return st;
    HIPCHK(hipMemcpyAsync(b->d.xyz, xyz, sizeof(double) * 3 * b->natoms, hipMemcpyHostToDevice, b->stream));
    st = launch_egrad(b, gradient != nullptr);
    if (st) return st;
    if (energy) HIPCHK(hipMemcpyAsync(energy, b->d.energy, sizeof(double) * b->nsys, hipMemcpyDeviceToHost, b->stream));
    if (gradient)
        HIPCHK(hipMemcpyAsync(gradient, b->d.grad, sizeof(double) * 3 * b->natoms, hipMemcpyDeviceToHost, b->stream));
    if (qat) HIPCHK(hipMemcpyAsync(qat, b->d.qat, sizeof(double) * b->natoms, hipMemcpyDeviceToHost, b->stream));
    std::vector<int> hs(b->nsys);
    HIPCHK(hipMemcpyAsync(hs.data(), b->d.status, sizeof(int) * b->nsys, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
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
        HIPCHK(hipMemcpyAsync(iterations, b->d.iters, sizeof(int) * b->nsys, hipMemcpyDeviceToHost, b->stream));
    if (eparts)
        HIPCHK(hipMemcpyAsync(eparts, b->d.eparts, sizeof(double) * 10 * b->nsys, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return QCXMS_OK;
}

int qcxms_batch_scc_iterations_total(qcxms_batch* b, uint64_t* total)
{
    if (!b || !total) return fail("null argument");
    HIPCHK(hipMemcpyAsync(total, b->d.iters_total, sizeof(uint64_t) * b->nsys, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return QCXMS_OK;
}

int qcxms_batch_multipoles(qcxms_batch* b, double* dipoles, double* quadrupoles)
{
    if (!b) return fail("null batch");
    std::vector<double> x(b->nxtot);
    HIPCHK(hipMemcpyAsync(x.data(), b->d.xo, sizeof(double) * b->nxtot, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
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

int qcxms_batch_orbitals(qcxms_batch* b, int32_t sys, int32_t cap, int32_t* nao, double* emo, double* focc,
                         double* coeff, double* overlap, int32_t* aoat)
{
    if (!b || !nao) return fail("null argument");
    if (sys < 0 || sys >= b->nsys) return fail("system index out of range");
    const SysInfo& si = b->info[sys];
    const int n = si.nao, ld = si.ld;
    *nao = n;
    if (cap < n) return fail("orbital buffers smaller than the basis");
    const size_t ms = (size_t)ld * ld;
    std::vector<double> S(ms), X(ms), U(ms), e(n), f(n);
    HIPCHK(hipMemcpyAsync(S.data(), b->d.S + si.mat_off, sizeof(double) * ms, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipMemcpyAsync(X.data(), b->d.X + si.mat_off, sizeof(double) * ms, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipMemcpyAsync(U.data(), b->d.U + si.mat_off, sizeof(double) * ms, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipMemcpyAsync(e.data(), b->d.eps + (size_t)sys * QX_MAX_NAO, sizeof(double) * n, hipMemcpyDeviceToHost,
                          b->stream));
    HIPCHK(hipMemcpyAsync(f.data(), b->d.focc + (size_t)sys * QX_MAX_NAO, sizeof(double) * n,
                          hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    // AO coefficients C = X U (X = L^-T, U = L^T C of the last SCC), orbitals
    // in ascending energy (the Jacobi eigensolver leaves them unordered)
    std::vector<int> ord(n);
    for (int i = 0; i < n; ++i) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int c) { return e[a] < e[c]; });
    for (int k = 0; k < n; ++k) {
        const int i = ord[k];
        if (emo) emo[k] = e[i];
        if (focc) focc[k] = 2.0 * f[i];
        if (coeff)
            for (int mu = 0; mu < n; ++mu) {
                double c = 0.0;
                for (int j = 0; j < n; ++j) c += X[(size_t)mu * ld + j] * U[(size_t)j * ld + i];
                coeff[(size_t)k * n + mu] = c;
            }
    }
    if (overlap)
        for (int mu = 0; mu < n; ++mu)
            for (int nu = 0; nu < n; ++nu) overlap[(size_t)nu * n + mu] = S[(size_t)mu * ld + nu];
    if (aoat) {
        int k = 0;
        for (int a = 0; a < si.nat; ++a) {
            const ElemParams& ep = g_host_tab.e[b->elem[si.atom_off + a]];
            for (int sh = 0; sh < ep.nsh; ++sh) {
                const int l = (ep.nsh == 2 && sh == 1) ? 1 : 0;
                for (int c = 0; c < 2 * l + 1; ++c) aoat[k++] = a + 1;
            }
        }
    }
    return QCXMS_OK;
}

int qcxms_batch_mo_spectrum(qcxms_batch* b, int32_t sys, int32_t cap, int32_t* nao, int32_t* ihomo, double* emo,
                            double* focc, double* qmo)
{
    if (!b || !nao || !ihomo || !emo || !focc || !qmo) return fail("null argument");
    if (sys < 0 || sys >= b->nsys) return fail("system index out of range");
    if (b->arena) return fail("the MO spectrum needs a batch with its own device memory");
    const SysInfo& si = b->info[sys];
    const int n = si.nao, nat = si.nat;
    *nao = n;
    if (cap < n) return fail("orbital buffers smaller than the basis");
    double* dv = nullptr;
    HIPCHK(hipMalloc((void**)&dv, sizeof(double) * (size_t)n * (2 + nat)));
    std::vector<double> host((size_t)n * (2 + nat));
    hipLaunchKernelGGL(qx::k_mo_pop, dim3(1), dim3(QX_BLOCK), 0, b->stream, b->d, (int)sys, dv, dv + n, dv + 2 * n);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(host.data(), dv, sizeof(double) * host.size(), hipMemcpyDeviceToHost,
                                            b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    (void)hipFree(dv);
    if (e != hipSuccess) return fail(std::string("k_mo_pop: ") + hipGetErrorString(e));
    std::copy(host.begin(), host.begin() + n, emo);
    std::copy(host.begin() + n, host.begin() + 2 * n, focc);
    std::copy(host.begin() + 2 * n, host.end(), qmo);
    // HOMO: the highest orbital holding electrons (1-based, at least 1)
    int h = 1;
    for (int k = 0; k < n; ++k)
        if (focc[k] > 1e-6) h = k + 1;
    *ihomo = h;
    return QCXMS_OK;
}

}  // extern "C"

namespace {

// a batch of one call whose device arrays live in the scratch array, run on
// the library stream (no device allocation, no stream creation)
qcxms_batch* arena_batch_create(int32_t nsys, const int32_t* offsets, const int32_t* num, const int32_t* charge)
{
    g_err.clear();
    if (ensure_init(0) != QCXMS_OK) return nullptr;
    qcxms_batch* b = new qcxms_batch();
    b->stream = g_stream;
    b->owns_stream = false;
    b->arena = g_scratch_dev;
    b->arena_cap = kScratchBytes;
    if (layout(b, nsys, offsets, num, charge) != QCXMS_OK || allocate(b) != QCXMS_OK) {
        delete b;
        return nullptr;
    }
    return b;
}

// The batch of the single-structure drop-in (qcxms_get_xtb_egrad), kept from
// call to call while the atoms and the charge stay the same: the MD of
// iniqm.f90:641-655 evaluates one structure per step, and rebuilding the
// device arrays every step would cost more than the evaluation.  Each call
// still starts from a fresh wavefunction (warm start off), as
// tblite.f90:133.  Held until the process ends.
struct SingleBatch {
    std::mutex mu;
    qcxms_batch* b = nullptr;
    std::vector<int32_t> num;
    int32_t charge = 0;
};
SingleBatch g_single;
constexpr double kEtempAufbau = 1.0e-3;  // K

// qcxms_egrad_batch and its allocation-free variant (arena != 0: every device
// array of the call is carved from the scratch array)
int egrad_batch_impl(bool arena, int32_t nsys, const int32_t* offsets, const int32_t* num, const double* xyz,
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
    if (!num || !xyz) {
        g_err = "null argument";
        return QCXMS_INVALID;
    }
    if (check_offsets(nsys, offsets, QX_MAX_NAT) != QCXMS_OK) {
        if (stat)
            for (int s = 0; s < nsys; ++s) stat[s] = QCXMS_INVALID;
        return QCXMS_INVALID;
    }
    // Launch order: largest basis first, so the longest workgroups start first
    // and the small systems fill the tail (longest-processing-time order).
    std::vector<int> order(nsys), nao(nsys);
    for (int s = 0; s < nsys; ++s) {
        order[s] = s;
        int n = 0;
        for (int i = offsets[s]; i < offsets[s + 1]; ++i) n += num[i] == 1 ? 1 : 4;
        nao[s] = n;
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return nao[a] > nao[c]; });
    std::vector<int32_t> po(nsys + 1, 0), pn, pc(nsys);
    std::vector<double> px, pe(nsys);
    for (int k = 0; k < nsys; ++k) {
        const int s = order[k];
        po[k + 1] = po[k] + (offsets[s + 1] - offsets[s]);
        pn.insert(pn.end(), num + offsets[s], num + offsets[s + 1]);
        px.insert(px.end(), xyz + 3 * (size_t)offsets[s], xyz + 3 * (size_t)offsets[s + 1]);
        pc[k] = charge ? charge[s] : 0;
        pe[k] = etemp ? etemp[s] : 300.0;
    }
    qcxms_batch* b = arena ? arena_batch_create(nsys, po.data(), pn.data(), pc.data())
                           : qcxms_batch_create(nsys, po.data(), pn.data(), pc.data(), method);
    if (!b) {
        const bool elem = g_err.find("no GFN2 parameters") != std::string::npos;
        if (stat)
            for (int s = 0; s < nsys; ++s) stat[s] = elem ? QCXMS_UNKNOWN_METHOD : QCXMS_INVALID;
        return elem ? QCXMS_UNKNOWN_METHOD : QCXMS_INVALID;
    }
    const int natoms = po[nsys];
    std::vector<double> e(nsys), g(3 * (size_t)natoms), q(natoms);
    std::vector<int32_t> ps(nsys);
    const int st = qcxms_batch_egrad(b, px.data(), pe.data(), e.data(), g.data(), q.data(), ps.data());
    qcxms_batch_destroy(b);
    for (int k = 0; k < nsys; ++k) {
        const int s = order[k];
        const int na = po[k + 1] - po[k];
        if (energy) energy[s] = e[k];
        if (stat) stat[s] = ps[k];
        if (qat) std::copy(q.begin() + po[k], q.begin() + po[k] + na, qat + offsets[s]);
        if (gradient)
            std::copy(g.begin() + 3 * (size_t)po[k], g.begin() + 3 * (size_t)(po[k] + na),
                      gradient + 3 * (size_t)offsets[s]);
    }
    return st;
}

}  // namespace

extern "C" {

int qcxms_egrad_batch(int32_t nsys, const int32_t* offsets, const int32_t* num, const double* xyz,
                      const int32_t* charge, const int32_t* multiplicity, int32_t method, const double* etemp,
                      double* qat, double* energy, double* gradient, int32_t* stat)
{
    return egrad_batch_impl(false, nsys, offsets, num, xyz, charge, multiplicity, method, etemp, qat, energy,
                            gradient, stat);
}

int qcxms_egrad_batch_noalloc(int32_t nsys, const int32_t* offsets, const int32_t* num, const double* xyz,
                              const int32_t* charge, const int32_t* multiplicity, int32_t method,
                              const double* etemp, double* qat, double* energy, double* gradient, int32_t* stat)
{
    // the scratch array is shared by every call in this process: one call at a time
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return egrad_batch_impl(true, nsys, offsets, num, xyz, charge, multiplicity, method, etemp, qat, energy,
                            gradient, stat);
}

int qcxms_get_xtb_egrad(int32_t nat, const int32_t* num, const double* xyz, int32_t charge, int32_t multiplicity,
                        int32_t method, double etemp, double* qat, double* energy, double* gradient, int32_t* stat)
{
    (void)multiplicity;  // uhf = min(mult-1, 0) = 0 for every multiplicity (tblite.f90:111)
    g_err.clear();
    auto done = [&](int s) {
        if (stat) *stat = s;
        return s;
    };
    if (method != QCXMS_GFN2_XTB) {
        g_err = "only the GFN2-xTB Hamiltonian (method 2) is implemented";
        return done(QCXMS_UNKNOWN_METHOD);
    }
    if (nat < 1 || !num || !xyz) {
        g_err = "null argument or no atoms";
        return done(QCXMS_FATAL);
    }
    // etemp 0 (getspec, mo_spec.f90:63) asks for integer occupations: the
    // Fermi filling at kEtempAufbau K gives them to within exp(-1e6)
    const double et = etemp > 0.0 ? etemp : kEtempAufbau;
    std::lock_guard<std::mutex> lk(g_single.mu);
    if (!g_single.b || g_single.charge != charge || g_single.num.size() != (size_t)nat ||
        !std::equal(g_single.num.begin(), g_single.num.end(), num)) {
        if (g_single.b) qcxms_batch_destroy(g_single.b);
        g_single.b = nullptr;
        g_single.num.assign(num, num + nat);
        g_single.charge = charge;
        const int32_t offs[2] = {0, nat};
        g_single.b = qcxms_batch_create(1, offs, num, &charge, method);
        if (!g_single.b) {
            const bool elem = g_err.find("no GFN2 parameters") != std::string::npos;
            return done(elem ? QCXMS_UNKNOWN_METHOD : QCXMS_FATAL);
        }
    }
    int32_t s = 0;
    const int st = qcxms_batch_egrad(g_single.b, xyz, &et, energy, gradient, qat, &s);
    return done(st == QCXMS_OK ? s : QCXMS_FATAL);
}

int qcxms_get_xtb_mo_spectrum(int32_t cap, int32_t* nao, int32_t* ihomo, double* emo, double* focc, double* qmo)
{
    std::lock_guard<std::mutex> lk(g_single.mu);
    if (!g_single.b) return fail("no structure evaluated by qcxms_get_xtb_egrad yet");
    return qcxms_batch_mo_spectrum(g_single.b, 0, cap, nao, ihomo, emo, focc, qmo);
}

qcxms_md* qcxms_md_create(int32_t ntraj, const int32_t* offsets, const int32_t* num, const int32_t* charge,
                          const double* mass, const int32_t* imass, const double* xyz, const double* velo,
                          const double* velof, const double* eimp, const double* tadd,
                          const qcxms_md_params* params)
{
    g_err.clear();
    if (!params || !mass || !xyz || !velo || !velof || !eimp || !tadd) {
        g_err = "null argument";
        return nullptr;
    }
    if (check_offsets(ntraj, offsets, QX_MAX_NAT) != QCXMS_OK) return nullptr;
    if (!(params->tstep > 0.0) || params->max_steps < 1) {
        g_err = "tstep and max_steps must be positive";
        return nullptr;
    }
    if (params->eigensolver < 0 || params->eigensolver > 2) {
        g_err = "eigensolver must be 0 (default), 1 (MFMA rotations, then Jacobi) or 2 (Jacobi)";
        return nullptr;
    }
    if (params->broyden_keep < 0 || params->broyden_keep > 100) {
        g_err = "broyden_keep must lie in [0, 100]";
        return nullptr;
    }
    if (params->extrapolate < -1 || params->extrapolate > 4) {
        g_err = "extrapolate must lie in -1..4";
        return nullptr;
    }
    MdLayout L;
    L.offsets.assign(offsets, offsets + ntraj + 1);
    const int na = offsets[ntraj];
    L.num.assign(num, num + na);
    L.charge.resize(ntraj);
    for (int s = 0; s < ntraj; ++s) L.charge[s] = charge ? charge[s] : 0;
    L.mass.assign(mass, mass + na);
    L.imass.assign(na, -1);
    if (imass) L.imass.assign(imass, imass + na);
    L.xyz.assign(xyz, xyz + 3 * (size_t)na);
    L.velo.assign(velo, velo + 3 * (size_t)na);
    L.velof.assign(velof, velof + na);
    L.src.assign(ntraj, -1);
    L.traj.resize(ntraj);
    for (int s = 0; s < ntraj; ++s) L.traj[s] = new_md_record(*params, eimp[s], tadd[s], 1, params->max_steps, 0.0);
    return md_build(nullptr, L, params);
}

void qcxms_md_destroy(qcxms_md* md)
{
    if (!md) return;
    (void)hipStreamSynchronize(md->b->stream);
    delete md;
}

qcxms_batch* qcxms_md_batch(qcxms_md* md) { return md ? md->b : nullptr; }

int qcxms_md_set_cid(qcxms_md* md, const int32_t* cid, const double* velo_cm, const int32_t* count_average)
{
    if (!md || !cid || !velo_cm) return fail("null argument");
    for (int s = 0; s < md->ntraj; ++s)
        if (cid[s] && !(velo_cm[s] >= 0.0 && velo_cm[s] < 1.0e12))
            return fail("velo_cm must be a finite speed >= 0 (m/s)");
    qcxms_batch* b = md->b;
    std::vector<qx::MdTraj> tr(md->ntraj);
    HIPCHK(hipMemcpyAsync(tr.data(), md->m.traj, sizeof(qx::MdTraj) * md->ntraj, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    for (int s = 0; s < md->ntraj; ++s) {
        if (!cid[s]) continue;
        if (tr[s].nstep != 0 || tr[s].cid) return fail("the CID mode is set once, before the first step");
    }
    for (int s = 0; s < md->ntraj; ++s) {
        if (!cid[s]) continue;
        tr[s].cid = 1;
        tr[s].new_velo = velo_cm[s];
        tr[s].count_average = count_average ? (count_average[s] != 0) : 0;
    }
    HIPCHK(hipMemcpyAsync(md->m.traj, tr.data(), sizeof(qx::MdTraj) * md->ntraj, hipMemcpyHostToDevice, b->stream));
    hipLaunchKernelGGL(qx::k_md_cid_start, dim3(md->ntraj), dim3(64), 0, b->stream, b->d, md->m);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(b->stream));
    return QCXMS_OK;
}

int qcxms_md_set_generation(qcxms_md* md, const int32_t* isec)
{
    if (!md || !isec) return fail("null argument");
    for (int s = 0; s < md->ntraj; ++s)
        if (isec[s] < 1) return fail("generation numbers start at 1");
    qcxms_batch* b = md->b;
    std::vector<qx::MdTraj> tr(md->ntraj);
    HIPCHK(hipMemcpyAsync(tr.data(), md->m.traj, sizeof(qx::MdTraj) * md->ntraj, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    for (int s = 0; s < md->ntraj; ++s) tr[s].isec = isec[s];
    HIPCHK(hipMemcpyAsync(md->m.traj, tr.data(), sizeof(qx::MdTraj) * md->ntraj, hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return QCXMS_OK;
}

}  // extern "C"

namespace {

// capture one MD step of md as a graph on its own stream (first use)
int md_graph(qcxms_md* md)
{
    qcxms_batch* b = md->b;
    if (md->graph_ready) return QCXMS_OK;
    HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    const int st = enqueue_md_step(md);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(b->stream, &g);
    if (st) return st;
    if (ec != hipSuccess) return fail(std::string("graph capture failed: ") + hipGetErrorString(ec));
    md->graph = g;
    HIPCHK(hipGraphInstantiate(&md->exec, md->graph, nullptr, nullptr, 0));
    md->graph_ready = true;
    return QCXMS_OK;
}

// enqueue one MD step of md on its stream (graph replay or eager launches)
int md_enqueue(qcxms_md* md)
{
    qcxms_batch* b = md->b;
    if (md->use_graph && !b->profiling) {
        HIPCHK(hipGraphLaunch(md->exec, b->stream));
        return QCXMS_OK;
    }
    return enqueue_md_step(md);
}

// up to nsteps MD iterations of every running trajectory in one persistent launch
int md_enqueue_fused(qcxms_md* md, int nsteps)
{
    qcxms_batch* b = md->b;
    const size_t lds = std::max(lds_setup_bytes(b), lds_scc_bytes(b));
    return timed(b, b->t_fused, [&] {
        hipLaunchKernelGGL(b->opts.eig_mfma ? qx::k_md_fused<true> : qx::k_md_fused<false>, dim3(md->ntraj),
                           dim3(QX_SCC_BLOCK), lds, b->stream, b->d, md->m, md->p, launch_opts(b, lds), nsteps);
    });
}

int md_running(qcxms_md* md, int32_t* running)
{
    std::vector<int> act(md->ntraj);
    HIPCHK(hipMemcpyAsync(act.data(), md->m.active, sizeof(int) * md->ntraj, hipMemcpyDeviceToHost, md->b->stream));
    HIPCHK(hipStreamSynchronize(md->b->stream));
    int n = 0;
    for (int v : act) n += v;
    *running = n;
    return QCXMS_OK;
}

}  // namespace

extern "C" {

int qcxms_md_run(qcxms_md* md, int32_t nsteps, int32_t* running)
{
    if (!md) return fail("null ensemble");
    return qcxms_md_run_group(&md, 1, nsteps, running);
}

int qcxms_md_run_group(qcxms_md** mds, int32_t n, int32_t nsteps, int32_t* running)
{
    if (!mds || n < 1) return fail("no ensembles");
    int st;
    for (int k = 0; k < n; ++k) {
        if (!mds[k]) return fail("null ensemble");
        if (!mds[k]->fused && mds[k]->use_graph && !mds[k]->b->profiling && (st = md_graph(mds[k])))
            return st;
    }
    // persistent ensembles: all steps in one launch each
    for (int k = 0; k < n; ++k)
        if (mds[k]->fused && nsteps > 0 && (st = md_enqueue_fused(mds[k], nsteps))) return st;
    // step s of every ensemble before step s + 1 of any: each ensemble's
    // stream keeps its own order, the device overlaps the streams
    for (int s = 0; s < nsteps; ++s)
        for (int k = 0; k < n; ++k)
            if (!mds[k]->fused && (st = md_enqueue(mds[k]))) return st;
    for (int k = 0; k < n; ++k) HIPCHK(hipStreamSynchronize(mds[k]->b->stream));
    bool prof = false;
    for (int k = 0; k < n; ++k) prof = prof || mds[k]->b->profiling;
    if (prof) collect_times();
    if (running)
        for (int k = 0; k < n; ++k)
            if ((st = md_running(mds[k], running + k))) return st;
    return QCXMS_OK;
}

int qcxms_md_read(qcxms_md* md, double* xyz, double* velo, double* grad, double* qat, int32_t* list,
                  qcxms_md_result* results)
{
    if (!md) return fail("null ensemble");
    qcxms_batch* b = md->b;
    const size_t na = b->natoms;
    if (xyz) HIPCHK(hipMemcpyAsync(xyz, b->d.xyz, sizeof(double) * 3 * na, hipMemcpyDeviceToHost, b->stream));
    if (velo) HIPCHK(hipMemcpyAsync(velo, md->m.velo, sizeof(double) * 3 * na, hipMemcpyDeviceToHost, b->stream));
    if (grad) HIPCHK(hipMemcpyAsync(grad, b->d.grad, sizeof(double) * 3 * na, hipMemcpyDeviceToHost, b->stream));
    if (qat) HIPCHK(hipMemcpyAsync(qat, b->d.qat, sizeof(double) * na, hipMemcpyDeviceToHost, b->stream));
    if (list) HIPCHK(hipMemcpyAsync(list, md->m.list, sizeof(int) * na, hipMemcpyDeviceToHost, b->stream));
    std::vector<qx::MdTraj> tr(md->ntraj);
    std::vector<int> it(md->ntraj), act(md->ntraj);
    HIPCHK(hipMemcpyAsync(tr.data(), md->m.traj, sizeof(qx::MdTraj) * md->ntraj, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipMemcpyAsync(it.data(), b->d.iters, sizeof(int) * md->ntraj, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipMemcpyAsync(act.data(), md->m.active, sizeof(int) * md->ntraj, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    if (results)
        for (int s = 0; s < md->ntraj; ++s) {
            const qx::MdTraj& t = tr[s];
            qcxms_md_result& r = results[s];
            std::memset(&r, 0, sizeof(r));
            r.nstep = t.nstep;
            r.nfrag = t.nfrag;
            r.fragstate = t.fragstate;
            r.mdok = t.mdok;
            r.exit_reason = t.exit_reason;
            r.running = act[s];
            r.scc_iterations = it[s];
            r.tav = t.nstep ? t.tav / t.nstep : 0.0;
            r.epav = t.nstep ? t.epav / t.nstep : 0.0;
            r.ekav = t.nstep ? t.ekav / t.nstep : 0.0;
            r.ttime = t.ttime;
            r.dtime = t.dtime;
            r.atlast = t.kdump ? t.atlast / t.kdump : 0.0;
            r.epot = t.epot;
            r.ekin = t.ekin;
            r.eerror = t.eerror;
            r.etemp = t.etemp;
            for (int f = 0; f < 10; ++f) {
                r.fragm[f] = t.fragm[f];
                r.fragT[f] = t.fragT[f];
            }
            r.check_fragmented = t.cid ? t.check_fragmented : 0;
            r.count_average = t.count_average;
            r.new_velo = t.new_velo;
        }
    return QCXMS_OK;
}

int qcxms_md_read_averages(qcxms_md* md, double* axyz, double* achrg)
{
    if (!md) return fail("null ensemble");
    qcxms_batch* b = md->b;
    const size_t na = b->natoms;
    std::vector<double> sx(3 * na), sq(na), so(3 * na);
    std::vector<qx::MdTraj> tr(md->ntraj);
    HIPCHK(hipMemcpyAsync(sx.data(), md->m.avxyz, sizeof(double) * 3 * na, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipMemcpyAsync(so.data(), md->m.store, sizeof(double) * 3 * na, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipMemcpyAsync(sq.data(), md->m.avchrg, sizeof(double) * na, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipMemcpyAsync(tr.data(), md->m.traj, sizeof(qx::MdTraj) * md->ntraj, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    // md.f90:691-698 (axyz = avxyz / kdump, or with CID after a new fragment
    // count the averaged structure store_avxyz); a trajectory that has not
    // completed a step yet has nothing averaged and gets zeros
    for (int s = 0; s < md->ntraj; ++s) {
        const int kd = tr[s].nstep > 0 ? tr[s].kdump : 0;
        const bool stored = tr[s].cid && tr[s].check_fragmented > 1;
        const SysInfo& si = b->info[s];
        for (int i = si.atom_off; i < si.atom_off + si.nat; ++i) {
            if (achrg) achrg[i] = kd ? sq[i] / kd : 0.0;
            if (axyz)
                for (int c = 0; c < 3; ++c) {
                    const size_t k = 3 * (size_t)i + c;
                    axyz[k] = stored ? so[k] : (kd ? sx[k] / kd : 0.0);
                }
        }
    }
    return QCXMS_OK;
}

int qcxms_fragments_batch(int32_t nsys, const int32_t* offsets, const int32_t* num, const double* xyz,
                          const double* mass, int32_t* list, int32_t* nfrag, double* fragm)
{
    g_err.clear();
    int st = ensure_init(0);
    if (st) return st;
    if ((st = check_offsets(nsys, offsets, QX_MAX_NAT))) return st;
    const size_t na = offsets[nsys];
    for (size_t i = 0; i < na; ++i)
        if (num[i] < 1 || num[i] > 118) return fail("atomic number outside 1..118");
    TmpBuf t;
    int *d_off, *d_num, *d_list, *d_nf;
    double *d_xyz, *d_mass, *d_fm;
    if ((st = t.up(&d_off, offsets, nsys + 1)) || (st = t.up(&d_num, num, na)) || (st = t.up(&d_xyz, xyz, 3 * na)) ||
        (st = t.up(&d_mass, mass, na)) || (st = t.up(&d_list, (const int*)nullptr, na)) ||
        (st = t.up(&d_nf, (const int*)nullptr, nsys)) || (st = t.up(&d_fm, (const double*)nullptr, 10 * (size_t)nsys)))
        return st;
    hipLaunchKernelGGL(qx::k_fragments_batch, dim3(nsys), dim3(64), 0, g_stream, d_off, d_num, d_xyz, d_mass, d_list,
                       d_nf, d_fm);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(list, d_list, sizeof(int) * na, hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipMemcpyAsync(nfrag, d_nf, sizeof(int) * nsys, hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipMemcpyAsync(fragm, d_fm, sizeof(double) * 10 * nsys, hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    return QCXMS_OK;
}

int qcxms_leapfrog_batch(int32_t nsys, const int32_t* offsets, const double* grad, const double* mass, double tstep,
                         double* xyz, double* velo, double* ekin)
{
    g_err.clear();
    int st = ensure_init(0);
    if (st) return st;
    if ((st = check_offsets(nsys, offsets, QX_MAX_NAT))) return st;
    const size_t na = offsets[nsys];
    TmpBuf t;
    int* d_off;
    double *d_g, *d_m, *d_x, *d_v, *d_e;
    if ((st = t.up(&d_off, offsets, nsys + 1)) || (st = t.up(&d_g, grad, 3 * na)) || (st = t.up(&d_m, mass, na)) ||
        (st = t.up(&d_x, xyz, 3 * na)) || (st = t.up(&d_v, velo, 3 * na)) ||
        (st = t.up(&d_e, (const double*)nullptr, nsys)))
        return st;
    hipLaunchKernelGGL(qx::k_leapfrog_batch, dim3(nsys), dim3(64), 0, g_stream, d_off, d_g, d_m, tstep, d_x, d_v, d_e);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(xyz, d_x, sizeof(double) * 3 * na, hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipMemcpyAsync(velo, d_v, sizeof(double) * 3 * na, hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipMemcpyAsync(ekin, d_e, sizeof(double) * nsys, hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    return QCXMS_OK;
}

int qcxms_impactscale_batch(int32_t nsys, const int32_t* offsets, double* velo, const double* mass,
                            const double* velof, const double* eimp, const double* ff, const double* e0,
                            int32_t* iterations)
{
    g_err.clear();
    int st = ensure_init(0);
    if (st) return st;
    if ((st = check_offsets(nsys, offsets, QX_MAX_NAT))) return st;
    const size_t na = offsets[nsys];
    TmpBuf t;
    int *d_off, *d_it;
    double *d_v, *d_m, *d_f, *d_ei, *d_ff, *d_e0;
    if ((st = t.up(&d_off, offsets, nsys + 1)) || (st = t.up(&d_v, velo, 3 * na)) || (st = t.up(&d_m, mass, na)) ||
        (st = t.up(&d_f, velof, na)) || (st = t.up(&d_ei, eimp, nsys)) || (st = t.up(&d_ff, ff, nsys)) ||
        (st = t.up(&d_e0, e0, nsys)) || (st = t.up(&d_it, (const int*)nullptr, nsys)))
        return st;
    hipLaunchKernelGGL(qx::k_impactscale_batch, dim3(nsys), dim3(64), 0, g_stream, d_off, d_v, d_m, d_f, d_ei, d_ff,
                       d_e0, d_it);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(velo, d_v, sizeof(double) * 3 * na, hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipMemcpyAsync(iterations, d_it, sizeof(int) * nsys, hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    return QCXMS_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Post-MD fragment analysis: analyse + manage_fragments (EI) for many
// finished trajectories; the single points of all fragments run as one batch.
#include "post_md.h"

extern "C" {

int qcxms_fragment_ips_batch(int32_t ntraj, const int32_t* offsets, const int32_t* num, const double* axyz,
                             const int32_t* list, const int32_t* nfrag, const int32_t* mchrg, double* fragip,
                             int32_t* ipok, int32_t* stat)
{
    using namespace qx_post;
    g_err.clear();
    if (ntraj < 0 || !offsets || !num || !axyz || !list || !nfrag || !mchrg || !fragip || !ipok || !stat)
        return fail("null argument");
    for (int t = 0; t < ntraj; ++t) {
        if (offsets[t + 1] < offsets[t]) return fail("offsets must be non-decreasing");
        if (nfrag[t] < 0 || nfrag[t] > kMaxFrag) return fail("nfrag outside 0..10");
        if (mchrg[t] == 0 || std::abs(mchrg[t]) > kMaxCharge) return fail("|mchrg| outside 1..3");
        for (int i = offsets[t]; i < offsets[t + 1]; ++i)
            if (list[i] < 1 || list[i] > nfrag[t]) return fail("fragment label outside 1..nfrag");
    }
    // avg_frag_struc (analyse.f90:452-502): fragment i = its atoms in input
    // order.  One job per (trajectory, fragment, charge 0..|mchrg|).
    struct Job {
        int t, frag, j, charge, sys;
        double e;
        int ok;
    };
    std::vector<Job> jobs;
    std::vector<int32_t> s_off{0}, s_num, s_chg, s_mult;
    std::vector<double> s_xyz, s_et;
    for (int t = 0; t < ntraj; ++t) {
        for (int k = 0; k < 10 * kMaxCharge; ++k) fragip[10 * kMaxCharge * (size_t)t + k] = 0.0;
        ipok[t] = 1;
        stat[t] = 0;
        if (nfrag[t] <= 1) continue;  // analyse.f90:125-130: no IPs for one fragment
        const int nch = std::abs(mchrg[t]), sgn = mchrg[t] < 0 ? -1 : 1;
        for (int f = 1; f <= nfrag[t]; ++f) {
            std::vector<int> z;
            std::vector<double> x;
            for (int i = offsets[t]; i < offsets[t + 1]; ++i)
                if (list[i] == f) {
                    z.push_back(num[i]);
                    x.insert(x.end(), axyz + 3 * (size_t)i, axyz + 3 * (size_t)i + 3);
                }
            for (int j = 0; j <= nch; ++j) {
                Job jb{t, f, j, sgn * j, -1, 0.0, 1};
                const int nel = valence_electrons(z.data(), (int)z.size(), jb.charge);
                if (nel == 0) {
                    if (!eself(z.data(), (int)z.size(), jb.e)) jb.ok = 0;
                } else if (nel < 0) {
                    jb.ok = 0;
                } else {
                    jb.sys = (int)s_chg.size();
                    s_num.insert(s_num.end(), z.begin(), z.end());
                    s_xyz.insert(s_xyz.end(), x.begin(), x.end());
                    s_off.push_back(s_off.back() + (int32_t)z.size());
                    s_chg.push_back(jb.charge);
                    s_mult.push_back(1);
                    s_et.push_back(kIpEtemp);
                }
                jobs.push_back(jb);
            }
        }
    }
    const int nsys = (int)s_chg.size();
    if (nsys > 0) {
        std::vector<double> e(nsys);
        std::vector<int32_t> st(nsys);
        const int rc = qcxms_egrad_batch(nsys, s_off.data(), s_num.data(), s_xyz.data(), s_chg.data(),
                                         s_mult.data(), QCXMS_GFN2_XTB, s_et.data(), nullptr, e.data(), nullptr,
                                         st.data());
        if (rc != QCXMS_OK && rc != QCXMS_FATAL) return rc;
        for (Job& jb : jobs)
            if (jb.sys >= 0) {
                jb.e = e[jb.sys];
                jb.ok = (st[jb.sys] == 0);
            }
    }
    // eqm (iniqm.f90:400-421) and analyse.f90:344-400 per trajectory
    size_t k = 0;
    while (k < jobs.size()) {
        const int t = jobs[k].t, nch = std::abs(mchrg[t]);
        double* ip = fragip + 10 * kMaxCharge * (size_t)t;
        bool good = true;
        for (; k < jobs.size() && jobs[k].t == t; k += nch + 1) {
            const Job& jn = jobs[k];
            if (!jn.ok) stat[t] = -1;
            if (std::fabs(jn.e) < 1.0e-10 || !jn.ok) good = false;
            for (int j = 1; j <= nch; ++j) {
                const Job& ji = jobs[k + j];
                if (!ji.ok) stat[t] = -1;
                if (std::fabs(ji.e) < 1.0e-10 || !ji.ok) good = false;
                if (ji.e != 0.0 && jn.e != 0.0) {
                    const double v = (ji.e - jn.e) * kAutoEV;
                    ip[(j - 1) * 10 + (jn.frag - 1)] = v;
                    if (mchrg[t] < 0) {
                        if (v > 40.0 || v < -35.0) good = false;
                    } else if (mchrg[t] == 1) {
                        if (v < 0.0 || v > 50.0) good = false;
                    } else if (v < 0.0 || v > 100.0) {
                        good = false;
                    }
                }
            }
        }
        if (!good) {
            // the second attempt repeats GFN2 and the third needs IPEA1
            // (not built): total failure, analyse.f90:406-414
            ipok[t] = 0;
            for (int j = 1; j <= mchrg[t]; ++j)
                for (int f = 0; f < nfrag[t]; ++f) ip[(j - 1) * 10 + f] = 0.0;
        }
    }
    return QCXMS_OK;
}

int qcxms_manage_fragments_batch(int32_t ntraj, const int32_t* offsets, const int32_t* num, const double* xyz,
                                 const double* axyz, const double* qat, const double* mass, const int32_t* imass,
                                 const double* atlast, const int32_t* itrj, int32_t isec, int32_t maxsec, double btf,
                                 int32_t* list, qcxms_fragment_report* report)
{
    using namespace qx_post;
    g_err.clear();
    if (ntraj < 0 || !offsets || !num || !xyz || !axyz || !qat || !mass || !atlast || !itrj || !list || !report)
        return fail("null argument");
    for (int t = 0; t < ntraj; ++t) {
        if (report[t].mchrg == 0 || std::abs(report[t].mchrg) > kMaxCharge) return fail("|mchrg| outside 1..3");
        for (int i = offsets[t]; i < offsets[t + 1]; ++i)
            if (imass && imass[i] > 99) return fail("isotope mass number above 99");
    }
    const size_t na = ntraj > 0 ? (size_t)offsets[ntraj] : 0;
    std::vector<int32_t> nf(ntraj);
    std::vector<double> fm(10 * (size_t)ntraj);
    int st;
    if (ntraj > 0 && (st = qcxms_fragments_batch(ntraj, offsets, num, xyz, mass, list, nf.data(), fm.data())))
        return st;
    // trajectories whose fragments get IPs: counted runs with 2..5 fragments
    std::vector<int> sel;
    std::vector<int32_t> o{0}, n, l, c, fnf;
    std::vector<double> x;
    for (int t = 0; t < ntraj; ++t) {
        qcxms_fragment_report& r = report[t];
        r.nfrag = nf[t];
        r.tcont = 0;
        r.ipok = 1;
        r.nres = 0;
        r.stat = 0;
        r.asave[0] = '\0';
        for (int f = 0; f < 10; ++f) {
            r.fragchrg3[f] = r.fragchrg[f] = 0.0;
            r.fragm[f] = fm[10 * (size_t)t + f];
        }
        for (int f = 0; f < 10 * kMaxCharge; ++f) r.fragip[f] = 0.0;
        if (nf[t] > 5) r.nfrag_ok = 0;  // write_fragments.f90:73-80
        if (!r.nfrag_ok || nf[t] <= 1) continue;
        sel.push_back(t);
        for (int i = offsets[t]; i < offsets[t + 1]; ++i) {
            n.push_back(num[i]);
            l.push_back(list[i]);
            x.insert(x.end(), axyz + 3 * (size_t)i, axyz + 3 * (size_t)i + 3);
        }
        o.push_back((int32_t)n.size());
        c.push_back(r.mchrg);
        fnf.push_back(nf[t]);
    }
    const int ns = (int)sel.size();
    std::vector<double> ips(10 * kMaxCharge * (size_t)ns);
    std::vector<int32_t> ok(ns), sst(ns);
    if (ns > 0 && (st = qcxms_fragment_ips_batch(ns, o.data(), n.data(), x.data(), l.data(), fnf.data(), c.data(),
                                                 ips.data(), ok.data(), sst.data())))
        return st;
    std::vector<int> slot(ntraj, -1);
    for (int k = 0; k < ns; ++k) slot[sel[k]] = k;
    for (int t = 0; t < ntraj; ++t) {
        qcxms_fragment_report& r = report[t];
        if (!r.nfrag_ok) continue;
        const int nfr = nf[t];
        // per-fragment atom counts, first element and element histogram
        // (fragmass, fragments.f90:27-35; isotopes at 100 + imass)
        int natf[kMaxFrag] = {0}, zfirst[kMaxFrag] = {0};
        std::vector<int> fragel(kMaxElem * kMaxFrag, 0);
        for (int i = offsets[t]; i < offsets[t + 1]; ++i) {
            const int f = list[i] - 1;
            if (natf[f]++ == 0) zfirst[f] = num[i];
            const int e = (imass && imass[i] > 0) ? 100 + imass[i] : num[i];
            fragel[f * kMaxElem + (e - 1)] += 1;
            r.fragchrg[f] = r.fragchrg[f] + qat[i];  // tm/orca/xtb branch, write_fragments.f90:320-343
        }
        if (slot[t] >= 0) {
            const double* ip = ips.data() + 10 * kMaxCharge * (size_t)slot[t];
            for (int f = 0; f < 10 * kMaxCharge; ++f) r.fragip[f] = ip[f];
            r.ipok = ok[slot[t]];
            r.stat = sst[slot[t]];
        }
        ChargeState cs;
        cs.nfrag = nfr;
        cs.mchrg = r.mchrg;
        cs.chrgcont = r.chrgcont;
        // fragip(nfrag, |mchrg|) column-major from the 10 x 3 slots
        double fip[kMaxFrag * kMaxCharge] = {0};
        for (int j = 0; j < std::abs(r.mchrg); ++j)
            for (int f = 0; f < nfr; ++f) fip[f + j * nfr] = r.fragip[j * 10 + f];
        assign_charges(cs, fip, natf, zfirst, atlast[t] * btf, isec, maxsec);
        r.mchrg = cs.mchrg;
        r.chrgcont = cs.chrgcont;
        r.tcont = cs.tcont;
        for (int f = 0; f < 10; ++f) r.fragchrg3[f] = cs.fragchrg3[f];
        // records (write_fragments.f90:368-450, EI)
        for (int j = 1; j <= nfr; ++j) {
            int zel[kMaxElem], cnt[kMaxElem], nl = 0;
            for (int e = 0; e < kMaxElem; ++e)
                if (fragel[(j - 1) * kMaxElem + e] != 0) {
                    zel[nl] = e + 1;
                    cnt[nl] = fragel[(j - 1) * kMaxElem + e];
                    ++nl;
                }
            const std::string rec = res_record(r.fragchrg3[j - 1], r.mchrg, itrj[t], r.icoll > 0 ? r.icoll : -1, isec, j, nl,
                                                   zel, cnt);
            char* dst = (r.tcont > 0 && r.tcont == j) ? r.asave : r.res[r.nres++];
            std::snprintf(dst, sizeof r.asave, "%s", rec.c_str());
        }
    }
    return QCXMS_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Fragmentation generations (src/main.F90:2208-2345, EI branch) for many
// trajectories, with the ensemble laid out again after every chunk of steps
// in which trajectories ended.

struct qcxms_ei {
    qcxms_md* md = nullptr;
    MdLayout L;  // host mirror of md's layout (num, charge, mass, imass)
    std::vector<int> slot;  // trajectory of each system of md
    qcxms_md_params params{};
    int maxsec = 7, repack = 100, repacks = 0;
    double btf = 1.0;
    struct Traj {
        int isec = 1, nmax = 0, mchrg = 1, nfrag_ok = 1, running = 1;
        double chrgcont = 1.0;
        int last_gen = -1;        // index in gens of the last finished generation
        std::vector<int> atoms;   // current atoms as indices into the first generation
    };
    std::vector<Traj> tr;
    std::vector<qcxms_ei_generation> gens;
    ~qcxms_ei() { delete md; }
};

namespace {

// one chunk's bookkeeping: generations that ended, manage_fragments for the
// finished runs, the next generations, and the new layout
int ei_after_chunk(qcxms_ei* ei)
{
    qcxms_md* md = ei->md;
    const int n = md->ntraj, na = md->b->natoms;
    std::vector<qcxms_md_result> res(n);
    std::vector<double> xyz(3 * (size_t)na), velo(3 * (size_t)na), axyz(3 * (size_t)na), achrg(na);
    int st;
    if ((st = qcxms_md_read(md, xyz.data(), velo.data(), nullptr, nullptr, nullptr, res.data()))) return st;
    std::vector<int> fin, keep;
    for (int s = 0; s < n; ++s) {
        if (!ei->tr[ei->slot[s]].running) continue;  // ended in an earlier chunk
        (res[s].running ? keep : fin).push_back(s);
    }
    if (fin.empty()) return QCXMS_OK;
    if ((st = qcxms_md_read_averages(md, axyz.data(), achrg.data()))) return st;
    const MdLayout& L = ei->L;
    // generation records
    std::vector<int> gidx(n, -1);
    for (int s : fin) {
        qcxms_ei::Traj& t = ei->tr[ei->slot[s]];
        qcxms_ei_generation g;
        std::memset(&g, 0, sizeof(g));
        const qcxms_md_result& r = res[s];
        g.traj = ei->slot[s];
        g.isec = t.isec;
        g.nat = L.offsets[s + 1] - L.offsets[s];
        g.nstep = r.nstep;
        g.exit_reason = r.exit_reason;
        g.mdok = r.mdok;
        g.fragstate = r.fragstate;
        g.nfrag = r.nfrag;
        g.max_steps = t.nmax;
        g.mchrg = t.mchrg;
        g.nfrag_ok = t.nfrag_ok;
        g.chrgcont = t.chrgcont;
        g.ttime = r.ttime;
        g.epot = r.epot;
        for (int i = 0; i < g.nat && i < 64; ++i) g.atoms[i] = t.atoms[i];
        if (!r.mdok) {
            // the run is not counted; the record of the fragment this run
            // continued goes to qcxms.res (main.F90:2340-2345)
            if (t.last_gen >= 0 && ei->gens[t.last_gen].asave[0]) ei->gens[t.last_gen].asave_written = 1;
            t.running = 0;
        }
        gidx[s] = (int)ei->gens.size();
        t.last_gen = gidx[s];
        ei->gens.push_back(g);
    }
    // manage_fragments for the counted runs, one call per generation number
    std::vector<int> isecs;
    for (int s : fin)
        if (res[s].mdok && std::find(isecs.begin(), isecs.end(), ei->tr[ei->slot[s]].isec) == isecs.end())
            isecs.push_back(ei->tr[ei->slot[s]].isec);
    std::vector<int> cont;  // finished systems whose trajectory goes on
    std::vector<qcxms_fragment_report> rep(n);
    for (int isec : isecs) {
        std::vector<int> grp;
        for (int s : fin)
            if (res[s].mdok && ei->tr[ei->slot[s]].isec == isec) grp.push_back(s);
        std::vector<int32_t> o{0}, nm, im, itrj, lst;
        std::vector<double> x, ax, q, ms, atl;
        std::vector<qcxms_fragment_report> rr(grp.size());
        for (size_t k = 0; k < grp.size(); ++k) {
            const int s = grp[k];
            const qcxms_ei::Traj& t = ei->tr[ei->slot[s]];
            for (int i = L.offsets[s]; i < L.offsets[s + 1]; ++i) {
                nm.push_back(L.num[i]);
                im.push_back(L.imass[i]);
                ms.push_back(L.mass[i]);
                q.push_back(achrg[i]);
                for (int c = 0; c < 3; ++c) {
                    x.push_back(xyz[3 * (size_t)i + c]);
                    ax.push_back(axyz[3 * (size_t)i + c]);
                }
            }
            o.push_back((int32_t)nm.size());
            itrj.push_back(ei->slot[s] + 1);
            atl.push_back(res[s].atlast);
            std::memset(&rr[k], 0, sizeof(rr[k]));
            rr[k].mchrg = t.mchrg;
            rr[k].nfrag_ok = t.nfrag_ok;
            rr[k].chrgcont = t.chrgcont;
        }
        lst.assign(nm.size(), 0);
        if ((st = qcxms_manage_fragments_batch((int32_t)grp.size(), o.data(), nm.data(), x.data(), ax.data(), q.data(),
                                               ms.data(), im.data(), atl.data(), itrj.data(), isec, ei->maxsec,
                                               ei->btf, lst.data(), rr.data())))
            return st;
        for (size_t k = 0; k < grp.size(); ++k) {
            const int s = grp[k];
            qcxms_ei::Traj& t = ei->tr[ei->slot[s]];
            const qcxms_fragment_report& r = rr[k];
            rep[s] = r;
            qcxms_ei_generation& g = ei->gens[gidx[s]];
            g.tcont = r.tcont;
            g.mchrg = t.mchrg;  // the MD's charge; the report carries the next one
            g.nfrag_ok = r.nfrag_ok;
            g.chrgcont = r.chrgcont;
            g.nres = r.nres;
            std::memcpy(g.asave, r.asave, sizeof(g.asave));
            for (int j = 0; j < r.nres && j < 10; ++j) std::memcpy(g.res[j], r.res[j], sizeof(g.res[j]));
            for (int i = 0; i < g.nat && i < 64; ++i) g.list[i] = lst[o[k] + i];
            t.mchrg = r.mchrg;
            t.chrgcont = r.chrgcont;
            t.nfrag_ok = r.nfrag_ok;
            if (!r.nfrag_ok || !(r.tcont > 0 && ei->maxsec > 0)) {
                t.running = 0;
                continue;
            }
            cont.push_back(s);
        }
    }
    // next generations (main.F90:2290-2345)
    struct NewSys {
        int traj;
        std::vector<int32_t> num, imass, atoms;
        std::vector<double> mass, xyz, velo;
        qx::MdTraj rec;
        int charge;
    };
    std::vector<NewSys> ns;
    const int nmax0 = ei->params.max_steps;
    for (int s : cont) {
        qcxms_ei::Traj& t = ei->tr[ei->slot[s]];
        qcxms_ei_generation& g = ei->gens[gidx[s]];
        int nmax = t.nmax - nmax0 / 5;
        nmax = std::max(nmax, nmax0 / 2);
        if (res[s].fragstate == 2) nmax = nmax0 / 3;
        NewSys y;
        y.traj = ei->slot[s];
        double dum = 0.0, cema[3] = {0.0, 0.0, 0.0};
        for (int i = L.offsets[s], a = 0; i < L.offsets[s + 1]; ++i, ++a) {
            if (g.list[a] != g.tcont) continue;
            y.num.push_back(L.num[i]);
            y.imass.push_back(L.imass[i]);
            y.mass.push_back(L.mass[i]);
            y.atoms.push_back(t.atoms[a]);
            for (int c = 0; c < 3; ++c) {
                y.xyz.push_back(xyz[3 * (size_t)i + c]);
                y.velo.push_back(velo[3 * (size_t)i + c]);
            }
            dum = dum + (double)L.num[i];
            for (int c = 0; c < 3; ++c) cema[c] = cema[c] + xyz[3 * (size_t)i + c] * (double)L.num[i];
        }
        for (int c = 0; c < 3; ++c) cema[c] = cema[c] / dum;
        const int nuc = (int)y.num.size();
        if (nuc <= 10) nmax = nmax0 / 10;
        if (nuc <= 3) {
            // run not continued because of small fragment(s)
            g.asave_written = g.asave[0] ? 1 : 0;
            t.running = 0;
            continue;
        }
        for (int i = 0; i < nuc; ++i)
            for (int c = 0; c < 3; ++c) y.xyz[3 * (size_t)i + c] = y.xyz[3 * (size_t)i + c] - cema[c];
        g.continued = 1;
        t.isec += 1;
        t.nmax = nmax;
        t.atoms = y.atoms;
        y.charge = t.mchrg;
        y.rec = new_md_record(ei->params, 0.0, 0.0, t.isec, nmax, res[s].ttime);
        ns.push_back(std::move(y));
    }
    // the new layout: running systems (state gathered on the device) and the
    // new generations, largest basis first
    struct Item {
        int nao, traj, kept_slot, new_idx;
    };
    std::vector<Item> items;
    auto nao_of = [](const int32_t* z, int cnt) {
        int k = 0;
        for (int i = 0; i < cnt; ++i) k += z[i] == 1 ? 1 : 4;
        return k;
    };
    for (int s : keep)
        items.push_back({nao_of(L.num.data() + L.offsets[s], L.offsets[s + 1] - L.offsets[s]), ei->slot[s], s, -1});
    for (size_t k = 0; k < ns.size(); ++k)
        items.push_back({nao_of(ns[k].num.data(), (int)ns[k].num.size()), ns[k].traj, -1, (int)k});
    if (items.empty()) {
        // every trajectory has ended: the last layout stays for reading
        return QCXMS_OK;
    }
    std::stable_sort(items.begin(), items.end(), [](const Item& a, const Item& c) {
        return a.nao != c.nao ? a.nao > c.nao : a.traj < c.traj;
    });
    MdLayout N;
    N.offsets.push_back(0);
    std::vector<int> slot;
    for (const Item& it : items) {
        slot.push_back(it.traj);
        if (it.kept_slot >= 0) {
            const int s = it.kept_slot;
            const int a0 = L.offsets[s], a1 = L.offsets[s + 1];
            N.num.insert(N.num.end(), L.num.begin() + a0, L.num.begin() + a1);
            N.imass.insert(N.imass.end(), L.imass.begin() + a0, L.imass.begin() + a1);
            N.mass.insert(N.mass.end(), L.mass.begin() + a0, L.mass.begin() + a1);
            N.xyz.insert(N.xyz.end(), 3 * (size_t)(a1 - a0), 0.0);
            N.velo.insert(N.velo.end(), 3 * (size_t)(a1 - a0), 0.0);
            N.velof.insert(N.velof.end(), (size_t)(a1 - a0), 0.0);
            N.charge.push_back(L.charge[s]);
            N.src.push_back(s);
            qx::MdTraj dummy;
            std::memset(&dummy, 0, sizeof(dummy));
            N.traj.push_back(dummy);
        } else {
            const NewSys& y = ns[it.new_idx];
            N.num.insert(N.num.end(), y.num.begin(), y.num.end());
            N.imass.insert(N.imass.end(), y.imass.begin(), y.imass.end());
            N.mass.insert(N.mass.end(), y.mass.begin(), y.mass.end());
            N.xyz.insert(N.xyz.end(), y.xyz.begin(), y.xyz.end());
            N.velo.insert(N.velo.end(), y.velo.begin(), y.velo.end());
            N.velof.insert(N.velof.end(), y.num.size(), 0.0);
            N.charge.push_back(y.charge);
            N.src.push_back(-1);
            N.traj.push_back(y.rec);
        }
        N.offsets.push_back((int32_t)N.num.size());
    }
    qcxms_md* nmd = md_build(md, N, nullptr);
    if (!nmd) return QCXMS_INVALID;
    delete md;
    ei->md = nmd;
    ei->L = std::move(N);
    ei->slot = std::move(slot);
    ei->repacks += 1;
    return QCXMS_OK;
}

}  // namespace

extern "C" {

qcxms_ei* qcxms_ei_create(int32_t ntraj, const int32_t* offsets, const int32_t* num, const int32_t* charge,
                          const double* mass, const int32_t* imass, const double* xyz, const double* velo,
                          const double* velof, const double* eimp, const double* tadd, const qcxms_md_params* params,
                          int32_t maxsec, double btf, int32_t repack_every)
{
    g_err.clear();
    if (repack_every < 1 || maxsec < 0 || !(btf > 0.0)) {
        g_err = "repack_every must be >= 1, maxsec >= 0 and btf > 0";
        return nullptr;
    }
    qcxms_md* md = qcxms_md_create(ntraj, offsets, num, charge, mass, imass, xyz, velo, velof, eimp, tadd, params);
    if (!md) return nullptr;
    qcxms_ei* ei = new qcxms_ei();
    ei->md = md;
    ei->params = *params;
    ei->maxsec = maxsec;
    ei->btf = btf;
    ei->repack = repack_every;
    const int na = offsets[ntraj];
    MdLayout& L = ei->L;
    L.offsets.assign(offsets, offsets + ntraj + 1);
    L.num.assign(num, num + na);
    L.imass.assign(na, -1);
    if (imass) L.imass.assign(imass, imass + na);
    L.mass.assign(mass, mass + na);
    L.charge.resize(ntraj);
    ei->tr.resize(ntraj);
    for (int t = 0; t < ntraj; ++t) {
        L.charge[t] = charge ? charge[t] : 0;
        ei->slot.push_back(t);
        qcxms_ei::Traj& T = ei->tr[t];
        T.nmax = params->max_steps;
        T.mchrg = L.charge[t];
        T.chrgcont = (double)T.mchrg;
        for (int i = 0; i < offsets[t + 1] - offsets[t]; ++i) T.atoms.push_back(i);
    }
    return ei;
}

void qcxms_ei_destroy(qcxms_ei* ei)
{
    if (!ei) return;
    if (ei->md) (void)hipStreamSynchronize(ei->md->b->stream);
    delete ei;
}

int qcxms_ei_run(qcxms_ei* ei, int32_t nsteps, int32_t* running)
{
    if (!ei || !ei->md) return fail("null ensemble");
    int left = nsteps, st;
    while (left > 0) {
        int anyrun = 0;
        for (const auto& t : ei->tr) anyrun |= t.running;
        if (!anyrun) break;
        const int chunk = std::min(left, ei->repack);
        if ((st = qcxms_md_run(ei->md, chunk, nullptr))) return st;
        left -= chunk;
        if ((st = ei_after_chunk(ei))) return st;
    }
    if (running) {
        int nr = 0;
        for (const auto& t : ei->tr) nr += t.running;
        *running = nr;
    }
    return QCXMS_OK;
}

int qcxms_ei_generations(qcxms_ei* ei, int32_t cap, qcxms_ei_generation* out, int32_t* count)
{
    if (!ei || !count) return fail("null argument");
    *count = (int32_t)ei->gens.size();
    for (int k = 0; k < cap && k < (int)ei->gens.size(); ++k) out[k] = ei->gens[k];
    return QCXMS_OK;
}

qcxms_md* qcxms_ei_md(qcxms_ei* ei) { return ei ? ei->md : nullptr; }

int qcxms_ei_steps(qcxms_ei* ei, uint64_t* total)
{
    if (!ei || !total || !ei->md) return fail("null argument");
    uint64_t n = 0;
    std::vector<char> seen(ei->tr.size(), 0);
    for (const qcxms_ei_generation& g : ei->gens) n += (uint64_t)g.nstep;
    // generations still in the current layout
    std::vector<qcxms_md_result> res(ei->md->ntraj);
    int st;
    if ((st = qcxms_md_read(ei->md, nullptr, nullptr, nullptr, nullptr, nullptr, res.data()))) return st;
    for (int s = 0; s < ei->md->ntraj; ++s)
        if (ei->tr[ei->slot[s]].running && res[s].running) n += (uint64_t)res[s].nstep;
    *total = n;
    return QCXMS_OK;
}

int qcxms_ei_repacks(qcxms_ei* ei, int32_t* count)
{
    if (!ei || !count) return fail("null argument");
    *count = ei->repacks;
    return QCXMS_OK;
}

}  // extern "C"

// ---- CID: the collision MD of src/cid.f90:724-1111 for many trajectories ----

struct qcxms_cid {
    qcxms_batch* b = nullptr;
    qx::CidParams p{};
    qx::CidPtrs c{};
    int ntraj = 0;
    std::vector<void*> allocs;
    ~qcxms_cid()
    {
        if (b && b->stream) (void)hipStreamSynchronize(b->stream);
        for (void* q : allocs) (void)hipFree(q);
        delete b;
    }
};

namespace {

template <class T>
int cidalloc(qcxms_cid* cd, T** p, size_t n, const T* host = nullptr)
{
    if (n == 0) n = 1;
    HIPCHK(hipMalloc((void**)p, n * sizeof(T)));
    cd->allocs.push_back((void*)*p);
    if (host) HIPCHK(hipMemcpyAsync(*p, host, n * sizeof(T), hipMemcpyHostToDevice, cd->b->stream));
    else HIPCHK(hipMemsetAsync(*p, 0, n * sizeof(T), cd->b->stream));
    return QCXMS_OK;
}

int cid_running(qcxms_cid* cd, int32_t* running)
{
    std::vector<int> act(cd->ntraj);
    HIPCHK(hipMemcpyAsync(act.data(), cd->c.active, sizeof(int) * cd->ntraj, hipMemcpyDeviceToHost, cd->b->stream));
    HIPCHK(hipStreamSynchronize(cd->b->stream));
    int n = 0;
    for (int v : act) n += v != 0;
    *running = n;
    return QCXMS_OK;
}

int cid_enqueue_step(qcxms_cid* cd)
{
    qcxms_batch* b = cd->b;
    hipLaunchKernelGGL(qx::k_cid_pre, dim3(cd->ntraj), dim3(64), 0, b->stream, b->d, cd->c, cd->p);
    HIPCHK(hipGetLastError());
    int st = enqueue_egrad(b, true);
    if (st) return st;
    hipLaunchKernelGGL(qx::k_cid_post, dim3(cd->ntraj), dim3(64), 0, b->stream, b->d, cd->c, cd->p);
    HIPCHK(hipGetLastError());
    return QCXMS_OK;
}

}  // namespace

extern "C" {

qcxms_cid* qcxms_cid_create(int32_t ntraj, const int32_t* offsets, const int32_t* nion, const int32_t* num,
                            const int32_t* charge, const double* mass, const int32_t* imass, const double* xyz,
                            const double* velo, const double* tinit, const double* beta, const double* velo_cm,
                            const qcxms_cid_params* params)
{
    g_err.clear();
    if (!nion || !num || !mass || !xyz || !velo || !tinit || !beta || !params) {
        g_err = "null argument";
        return nullptr;
    }
    if (check_offsets(ntraj, offsets, QX_MAX_NAT) != QCXMS_OK) return nullptr;
    for (int s = 0; s < ntraj; ++s) {
        const int nat = offsets[s + 1] - offsets[s];
        if (nion[s] < 1 || nion[s] >= nat) {
            g_err = "each system needs at least one ion atom followed by the gas atoms";
            return nullptr;
        }
    }
    if (!(params->tstep > 0.0) || !(params->etemp > 0.0) || params->ntot < 1) {
        g_err = "tstep, etemp and ntot must be positive";
        return nullptr;
    }
    qcxms_batch* b = qcxms_batch_create(ntraj, offsets, num, charge, QCXMS_GFN2_XTB);
    if (!b) return nullptr;
    qcxms_cid* cd = new qcxms_cid();
    cd->b = b;
    cd->ntraj = ntraj;
    b->tune.clear();  // one SCC launch per step, as the MD ensembles
    b->opts.eig_mfma = 1;
    b->opts.damp = 0.3;
    b->opts.extrapolate = QX_DEFAULT_EXTRAPOLATE;
    if (params->econv > 0.0) b->opts.econv = params->econv;
    if (params->pconv > 0.0) b->opts.pconv = params->pconv;
    b->warm_start = 1;
    cd->p.tstep = params->tstep;
    cd->p.dumpdist = 10;   // cid.f90:215
    cd->p.dumpavg = 50;    // cid.f90:216
    cd->p.cnt_steps = 50;  // cid.f90:231
    const int na = b->natoms;
    std::vector<qx::CidTraj> tr(ntraj);
    for (int s = 0; s < ntraj; ++s) {
        qx::CidTraj& t = tr[s];
        std::memset(&t, 0, sizeof(t));
        t.nuc = nion[s];
        t.nuc0 = offsets[s + 1] - offsets[s];
        t.ntot = params->ntot;
        // cid.f90:232-234
        t.add_steps = 0;
        if (t.nuc > 10) t.add_steps = (t.nuc / 10) * 500;
        if (t.nuc >= 40) t.add_steps = (t.nuc / 10) * 1000;
        t.tinit = tinit[s];
        t.beta = beta[s];
        t.velo_cm = velo_cm ? velo_cm[s] : 0.0;
        double sm = 0.0;
        for (int i = 0; i < t.nuc; ++i) sm = sm + mass[offsets[s] + i];
        t.summass = sm;
        t.total_steps = params->ntot;
        t.check_fragmented = 1;
        t.nfrag = 1;
        b->info[s].kt = params->etemp * gfn2::kKtoAu;
    }
    std::vector<int> act(ntraj, 1), lst(na, 1), im(na, -1);
    if (imass) im.assign(imass, imass + na);
    std::vector<int> nm(num, num + na);
    double* d_mass;
    int* d_imass;
    int* d_num;
    if (cidalloc(cd, &cd->c.traj, ntraj, tr.data()) || cidalloc(cd, &cd->c.velo, 3 * (size_t)na, velo) ||
        cidalloc(cd, &d_mass, (size_t)na, mass) || cidalloc(cd, &d_imass, (size_t)na, im.data()) ||
        cidalloc(cd, &d_num, (size_t)na, nm.data()) || cidalloc(cd, &cd->c.list, (size_t)na, lst.data()) ||
        cidalloc(cd, &cd->c.avxyz, 3 * (size_t)na) || cidalloc(cd, &cd->c.avxyz2, 3 * (size_t)na) ||
        cidalloc(cd, &cd->c.store, 3 * (size_t)na) || cidalloc(cd, &cd->c.active, (size_t)ntraj, act.data())) {
        delete cd;
        return nullptr;
    }
    cd->c.mass = d_mass;
    cd->c.imass = d_imass;
    cd->c.num = d_num;
    b->d.active = cd->c.active;
    if (hipMemcpyAsync(b->d.xyz, xyz, sizeof(double) * 3 * na, hipMemcpyHostToDevice, b->stream) != hipSuccess ||
        hipMemcpyAsync(b->d_info, b->info.data(), sizeof(SysInfo) * ntraj, hipMemcpyHostToDevice, b->stream) !=
            hipSuccess) {
        g_err = "upload failed";
        delete cd;
        return nullptr;
    }
    // the initial forces of ion + gas (cid.f90:703)
    if (enqueue_egrad(b, true) != QCXMS_OK) {
        delete cd;
        return nullptr;
    }
    hipLaunchKernelGGL(qx::k_cid_init, dim3(ntraj), dim3(64), 0, b->stream, b->d, cd->c);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(b->stream) != hipSuccess) {
        g_err = "initial energy/gradient failed on the device";
        delete cd;
        return nullptr;
    }
    if (b->profiling) collect_times();
    return cd;
}

void qcxms_cid_destroy(qcxms_cid* cd) { delete cd; }

int qcxms_cid_set_saved(qcxms_cid* cd, const int32_t* collided, const int32_t* count_average)
{
    if (!cd) return fail("null ensemble");
    qcxms_batch* b = cd->b;
    std::vector<qx::CidTraj> tr(cd->ntraj);
    HIPCHK(hipMemcpyAsync(tr.data(), cd->c.traj, sizeof(qx::CidTraj) * cd->ntraj, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    for (int s = 0; s < cd->ntraj; ++s)
        if (tr[s].nstep != 0) return fail("the saved flags are set before the first step");
    for (int s = 0; s < cd->ntraj; ++s) {
        if (collided) tr[s].collided = collided[s] != 0;
        if (count_average) tr[s].count_average = count_average[s] != 0;
    }
    HIPCHK(hipMemcpyAsync(cd->c.traj, tr.data(), sizeof(qx::CidTraj) * cd->ntraj, hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return QCXMS_OK;
}

int qcxms_cid_run(qcxms_cid* cd, int32_t nsteps, int32_t* running)
{
    if (!cd) return fail("null ensemble");
    int st, run = cd->ntraj;
    for (int s = 0; s < nsteps && run > 0; ++s) {
        if ((st = cid_enqueue_step(cd))) return st;
        // look for the end of the ensemble every 50 steps
        if ((s + 1) % 50 == 0 && (st = cid_running(cd, &run))) return st;
    }
    if ((st = cid_running(cd, &run))) return st;
    if (cd->b->profiling) collect_times();
    if (running) *running = run;
    return QCXMS_OK;
}

int qcxms_cid_read(qcxms_cid* cd, double* xyz, double* velo, double* grad, double* qat, int32_t* list,
                   double* axyz, qcxms_cid_result* results)
{
    if (!cd) return fail("null ensemble");
    qcxms_batch* b = cd->b;
    const size_t na = b->natoms;
    if (xyz) HIPCHK(hipMemcpyAsync(xyz, b->d.xyz, sizeof(double) * 3 * na, hipMemcpyDeviceToHost, b->stream));
    if (velo) HIPCHK(hipMemcpyAsync(velo, cd->c.velo, sizeof(double) * 3 * na, hipMemcpyDeviceToHost, b->stream));
    if (grad) HIPCHK(hipMemcpyAsync(grad, b->d.grad, sizeof(double) * 3 * na, hipMemcpyDeviceToHost, b->stream));
    if (qat) HIPCHK(hipMemcpyAsync(qat, b->d.qat, sizeof(double) * na, hipMemcpyDeviceToHost, b->stream));
    if (list) HIPCHK(hipMemcpyAsync(list, cd->c.list, sizeof(int) * na, hipMemcpyDeviceToHost, b->stream));
    std::vector<qx::CidTraj> tr(cd->ntraj);
    std::vector<double> av(3 * na), store(3 * na);
    HIPCHK(hipMemcpyAsync(tr.data(), cd->c.traj, sizeof(qx::CidTraj) * cd->ntraj, hipMemcpyDeviceToHost, b->stream));
    if (axyz) {
        HIPCHK(hipMemcpyAsync(av.data(), cd->c.avxyz, sizeof(double) * 3 * na, hipMemcpyDeviceToHost, b->stream));
        HIPCHK(hipMemcpyAsync(store.data(), cd->c.store, sizeof(double) * 3 * na, hipMemcpyDeviceToHost, b->stream));
    }
    HIPCHK(hipStreamSynchronize(b->stream));
    for (int s = 0; s < cd->ntraj; ++s) {
        const qx::CidTraj& t = tr[s];
        const size_t off = (size_t)b->info[s].atom_off;
        if (axyz) {
            // the structure cid returns (cid.f90:1097-1102)
            for (int k = 0; k < 3 * t.nuc0; ++k) axyz[3 * off + k] = 0.0;
            for (int k = 0; k < 3 * t.nuc; ++k)
                axyz[3 * off + k] = t.check_fragmented > 1 ? store[3 * off + k]
                                                           : av[3 * off + k] / (t.xyzavg_dump ? t.xyzavg_dump : 1);
        }
        if (results) {
            qcxms_cid_result& r = results[s];
            r.nstep = t.nstep;
            r.nfrag = t.nfrag;
            r.done = t.done;
            r.gradfail = t.gradfail;
            r.collided = t.collided;
            r.fragmented = t.fragmented;
            r.total_steps = t.total_steps;
            r.collision_step = t.collision_step;
            r.check_fragmented = t.check_fragmented;
            r.count_average = t.count_average;
            r.ttime = t.ttime;
            r.epot = t.epot;
            r.ekin = t.ekin;
            r.new_velo = t.new_velo;
            r.avgt = t.avgt;
            r.atlast = t.atlast;
            r.ecom = t.ecom;
            r.velo_diff = t.velo_diff;
            r.lowest_com = t.lowest;
            for (int f = 0; f < 10; ++f) {
                r.fragT[f] = t.fragT[f];
                r.eint[f] = t.eint[f];
            }
        }
    }
    return QCXMS_OK;
}

qcxms_batch* qcxms_cid_batch(qcxms_cid* cd) { return cd ? cd->b : nullptr; }

}  // extern "C"
