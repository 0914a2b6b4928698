// Driver of the CPU-emulated library build: an MD ensemble (qcxms_md_create,
// qcxms_md_run, qcxms_md_read) from a binary input file, results written to
// another.
//   input : int32 ntraj, nsteps, fused, use_graph; int32 offsets[ntraj+1];
//           int32 num[nat], charge[ntraj], imass[nat]; double mass[nat],
//           xyz[3 nat], velo[3 nat], velof[nat], eimp[ntraj], tadd[ntraj]
//   output: int32 rc, running; double xyz[3 nat], velo[3 nat], qat[nat];
//           per trajectory: int32 nstep, exit_reason; double epot, ekin;
//           int32 nfrag, check_fragmented; double new_velo
#include <cstdio>
#include <cstdlib>
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
    if (argc != 3) return 2;
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    std::vector<int32_t> hdr, off, num, chg, imass;
    std::vector<double> mass, xyz, velo, velof, eimp, tadd;
    if (!rd(f, hdr, 4)) return 2;
    const int nt = hdr[0], nsteps = hdr[1];
    if (!rd(f, off, nt + 1)) return 2;
    const int nat = off[nt];
    if (!rd(f, num, nat) || !rd(f, chg, nt) || !rd(f, imass, nat) || !rd(f, mass, nat) ||
        !rd(f, xyz, 3 * (size_t)nat) || !rd(f, velo, 3 * (size_t)nat) || !rd(f, velof, nat) ||
        !rd(f, eimp, nt) || !rd(f, tadd, nt))
        return 2;
    std::fclose(f);
    qcxms_md_params p{};
    p.tstep = 0.5 * 41.3413733365614;
    p.etempin = -1.0;
    p.nfragexit = 3;
    p.max_steps = nsteps;
    p.use_graph = hdr[3];
    p.fused = hdr[2];
    // SCC thresholds of the GPU tests when given (QX_EMU_ECONV, QX_EMU_PCONV)
    if (const char* e = std::getenv("QX_EMU_ECONV")) p.econv = std::atof(e);
    if (const char* e = std::getenv("QX_EMU_PCONV")) p.pconv = std::atof(e);
    if (const char* e = std::getenv("QX_EMU_ETEMP")) p.etempin = std::atof(e);
    if (const char* e = std::getenv("QX_EMU_NFRAGEXIT")) p.nfragexit = std::atoi(e);
    // CID mode of every trajectory with these centre-of-mass speeds (m/s),
    // comma-separated, one per trajectory (QX_EMU_CID)
    std::vector<int32_t> cid;
    std::vector<double> vcm;
    if (const char* e = std::getenv("QX_EMU_CID")) {
        for (const char* c = e; *c;) {
            char* end;
            vcm.push_back(std::strtod(c, &end));
            c = *end ? end + 1 : end;
        }
        if ((int)vcm.size() != nt) return 2;
        cid.assign(nt, 1);
    }
    qcxms_md* md = qcxms_md_create(nt, off.data(), num.data(), chg.data(), mass.data(), imass.data(), xyz.data(),
                                   velo.data(), velof.data(), eimp.data(), tadd.data(), &p);
    int32_t rc = md ? 0 : -1, running = 0;
    std::vector<qcxms_md_result> res(nt);
    std::vector<double> q(nat);
    if (md && !cid.empty()) rc = qcxms_md_set_cid(md, cid.data(), vcm.data(), nullptr);
    if (md && !rc) {
        rc = qcxms_md_run(md, nsteps + 1, &running);
        if (!rc) rc = qcxms_md_read(md, xyz.data(), velo.data(), nullptr, q.data(), nullptr, res.data());
        qcxms_md_destroy(md);
    }
    if (rc) std::fprintf(stderr, "md: %d %s\n", rc, qcxms_gpu_last_error());
    FILE* o = std::fopen(argv[2], "wb");
    if (!o) return 2;
    std::fwrite(&rc, sizeof(rc), 1, o);
    std::fwrite(&running, sizeof(running), 1, o);
    std::fwrite(xyz.data(), sizeof(double), xyz.size(), o);
    std::fwrite(velo.data(), sizeof(double), velo.size(), o);
    std::fwrite(q.data(), sizeof(double), q.size(), o);
    for (int t = 0; t < nt; ++t) {
        std::fwrite(&res[t].nstep, sizeof(int32_t), 1, o);
        std::fwrite(&res[t].exit_reason, sizeof(int32_t), 1, o);
        std::fwrite(&res[t].epot, sizeof(double), 1, o);
        std::fwrite(&res[t].ekin, sizeof(double), 1, o);
        std::fwrite(&res[t].nfrag, sizeof(int32_t), 1, o);
        std::fwrite(&res[t].check_fragmented, sizeof(int32_t), 1, o);
        std::fwrite(&res[t].new_velo, sizeof(double), 1, o);
    }
    std::fclose(o);
    return 0;
}
