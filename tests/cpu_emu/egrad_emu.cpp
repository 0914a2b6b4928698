// Driver of the CPU-emulated library build: qcxms_egrad_batch on systems read
// from a binary file, results written to another.  With QX_EMU_STORED=1 in
// the environment the systems go through a persistent batch whose gradient
// contracts stored derivative blocks (qcxms_batch_set_stored_derivatives).
// QX_EMU_TUNING="max_nao,eig,lds;..." sets the SCC launch plan first
// (qcxms_set_tuning, one launch per run of systems of one bin).
//   input : int32 nsys, method; int32 offsets[nsys+1]; int32 num[nat];
//           double xyz[3 nat]; int32 charge[nsys], mult[nsys]; double etemp[nsys]
//   output: int32 rc; double energy[nsys]; double grad[3 nat]; double qat[nat];
//           int32 stat[nsys]
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/qcxms_gpu.h"

template <class T>
static bool rd(FILE* f, std::vector<T>& v, size_t n)
{
    v.resize(n);
    return std::fread(v.data(), sizeof(T), n, f) == n;
}

int main(int argc, char** argv)
{
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s input output\n", argv[0]);
        return 2;
    }
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    std::vector<int32_t> hdr, off, num, chg, mult;
    std::vector<double> xyz, et;
    if (!rd(f, hdr, 2)) return 2;
    const int nsys = hdr[0], method = hdr[1];
    if (!rd(f, off, nsys + 1)) return 2;
    const int nat = off[nsys];
    if (!rd(f, num, nat) || !rd(f, xyz, 3 * (size_t)nat) || !rd(f, chg, nsys) || !rd(f, mult, nsys) ||
        !rd(f, et, nsys))
        return 2;
    std::fclose(f);
    std::vector<double> e(nsys), g(3 * (size_t)nat), q(nat);
    std::vector<int32_t> st(nsys);
    int32_t rc;
    const char* stored = std::getenv("QX_EMU_STORED");
    const char* tuning = std::getenv("QX_EMU_TUNING");
    if (tuning && *tuning) {
        std::vector<int32_t> a, e, l;
        int x, y, z, n = 0;
        for (const char* p = tuning; std::sscanf(p, "%d,%d,%d%n", &x, &y, &z, &n) == 3;) {
            a.push_back(x);
            e.push_back(y);
            l.push_back(z);
            p += n;
            if (*p != ';') break;
            ++p;
        }
        rc = qcxms_set_tuning((int32_t)a.size(), a.data(), e.data(), l.data(), -1, 0);
        if (rc != 0) std::fprintf(stderr, "qcxms_set_tuning: %s\n", qcxms_gpu_last_error());
    }
    if (stored && !std::strcmp(stored, "1")) {
        qcxms_batch* b = qcxms_batch_create(nsys, off.data(), num.data(), chg.data(), method);
        rc = b ? qcxms_batch_set_stored_derivatives(b, 1) : -2;
        if (rc == 0) rc = qcxms_batch_set_scc(b, 1e-11, 1e-9, 100, 0.4, 1e-11);
        if (rc == 0) rc = qcxms_batch_egrad(b, xyz.data(), et.data(), e.data(), g.data(), q.data(), st.data());
        if (b) qcxms_batch_destroy(b);
    } else {
        rc = qcxms_egrad_batch(nsys, off.data(), num.data(), xyz.data(), chg.data(), mult.data(), method,
                               et.data(), q.data(), e.data(), g.data(), st.data());
    }
    if (rc != 0) std::fprintf(stderr, "qcxms_egrad_batch: %d %s\n", rc, qcxms_gpu_last_error());
    FILE* o = std::fopen(argv[2], "wb");
    if (!o) return 2;
    std::fwrite(&rc, sizeof(rc), 1, o);
    std::fwrite(e.data(), sizeof(double), e.size(), o);
    std::fwrite(g.data(), sizeof(double), g.size(), o);
    std::fwrite(q.data(), sizeof(double), q.size(), o);
    std::fwrite(st.data(), sizeof(int32_t), st.size(), o);
    std::fclose(o);
    return 0;
}
