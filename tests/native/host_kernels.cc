// Test-only driver: the kernels' host-callable bodies (workspace layouts,
// run-time unranking, the FP64 Fisher-z in the reference's operation order,
// Perlman's pchisq) built for the host under ASan/UBSan by
// tests/test_host_sanitized.py, which compares their outputs with the oracle.
//
//   host_kernels layout                 every workspace carve stays inside its size, regions disjoint
//   host_kernels unrank N l count       1-based positions of ranks 0..count-1, one line each
//   host_kernels gauss_z IN OUT         IN: int32 p, int32 cases, double C[p*p], then per case
//                                       int32 vi, vj, l, S[8]; OUT: double z per case
//   host_kernels pchisq IN OUT          IN: int32 cases, then per case double x, int32 df; OUT: double
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../gpucsl_amd/csrc/job.h"

using namespace gpucsl;

namespace {

struct Region {
    const char *name;
    size_t off, bytes;
};

// every region inside [0, total), no two overlapping; each written end to end
bool check_regions(std::vector<Region> r, size_t total, const char *what, int p) {
    std::vector<unsigned char> buf(total);
    for (auto &x : r) {
        if (x.off + x.bytes > total) {
            std::fprintf(stderr, "%s p=%d: %s [%zu, %zu) beyond %zu\n", what, p, x.name, x.off, x.off + x.bytes, total);
            return false;
        }
        std::memset(buf.data() + x.off, 0x5a, x.bytes);
    }
    for (size_t a = 0; a < r.size(); ++a)
        for (size_t b = a + 1; b < r.size(); ++b) {
            const bool apart = r[a].off + r[a].bytes <= r[b].off || r[b].off + r[b].bytes <= r[a].off;
            if (!apart && r[a].bytes && r[b].bytes) {
                std::fprintf(stderr, "%s p=%d: %s and %s overlap\n", what, p, r[a].name, r[b].name);
                return false;
            }
        }
    return true;
}

size_t off_of(const void *base, const void *ptr) { return (size_t)((const char *)ptr - (const char *)base); }

int layout() {
    const int ps[] = {1, 2, 3, 5, 63, 64, 65, 127, 200, 256, 257, 500, 999, 1000, 1234, 2048};
    for (int p : ps) {
        const size_t total = job_ws_bytes(p);
        std::vector<unsigned char> mem(total);
        JobWs w = carve_job(mem.data(), p);
        const size_t I = sizeof(int32_t);
        std::vector<Region> r = {
            {"nbr", off_of(mem.data(), w.nbr), I * p * p},   {"deg", off_of(mem.data(), w.deg), I * p},
            {"cntA", off_of(mem.data(), w.cntA), I * p},     {"cntB", off_of(mem.data(), w.cntB), I * p},
            {"rowoffA", off_of(mem.data(), w.rowoffA), I * (p + 1)},
            {"rowoffB", off_of(mem.data(), w.rowoffB), I * (p + 1)},
            {"itemoffA", off_of(mem.data(), w.itemoffA), I * (p + 1)},
            {"itemoffB", off_of(mem.data(), w.itemoffB), I * (p + 1)},
            {"rthr", off_of(mem.data(), w.rthr), sizeof(float) * GPUCSL_MAX_LEVELS}};
        if (!check_regions(r, total, "job workspace", p)) return 1;
        // batched launch workspace for jobs of this p
        Job jobs[3];
        std::memset(jobs, 0, sizeof(jobs));
        for (auto &j : jobs) j.p = p;
        for (int nj = 1; nj <= 3; ++nj) {
            const size_t qcap = queue_capacity(jobs, nj), icap = item_capacity(jobs, nj);
            const size_t bt = batch_ws_bytes(nj, qcap, icap);
            std::vector<unsigned char> bm(bt);
            BatchWs b = carve_batch(bm.data(), nj, qcap, icap);
            std::vector<Region> br = {
                {"counters", off_of(bm.data(), b.counters), sizeof(unsigned int) * NUM_COUNTERS},
                {"bstats", off_of(bm.data(), b.bstats), sizeof(unsigned long long) * BSTAT_STRIPES * BSTAT_WORDS},
                {"itembaseA", off_of(bm.data(), b.itembaseA), I * (nj + 1)},
                {"itembaseB", off_of(bm.data(), b.itembaseB), I * (nj + 1)},
                {"queueA", off_of(bm.data(), b.queueA), sizeof(SlowItem) * qcap},
                {"queueB", off_of(bm.data(), b.queueB), sizeof(SlowItem) * qcap},
                {"itemsA", off_of(bm.data(), b.itemsA), sizeof(WorkItem) * icap},
                {"itemsB", off_of(bm.data(), b.itemsB), sizeof(WorkItem) * icap}};
            if (!check_regions(br, bt, "batch workspace", p)) return 1;
        }
        // correlation workspace
        const long long ns[] = {2, 63, 64, 65, 1000, 4097};
        for (long long n : ns)
            for (int ks : {1, 3, 16}) {
                const size_t ct = corr_job_ws_bytes(n, p, ks);
                std::vector<unsigned char> cm(ct);
                CorrWs c = carve_corr(cm.data(), n, p);
                const size_t split = sizeof(double) * (size_t)CORR_ROW_SPLITS_MAX * p;
                std::vector<Region> cr = {
                    {"partial", off_of(cm.data(), c.partial), split},
                    {"cmin", off_of(cm.data(), c.cmin), split},
                    {"cmax", off_of(cm.data(), c.cmax), split},
                    {"planes", off_of(cm.data(), c.planes), (size_t)I8_PLANES * i8_cols(p) * (size_t)i8_rows(n)},
                    {"gram", off_of(cm.data(), c.gram), sizeof(double) * (size_t)ks * p * p}};
                if (!check_regions(cr, ct, "correlation workspace", p)) return 1;
            }
    }
    std::printf("layout ok\n");
    return 0;
}

int unrank(int N, int l, long long count) {
    std::vector<int> pos(l > 0 ? l : 1);
    for (long long r = 0; r < count; ++r) {
        unrank_lex_rt(N, l, (uint64_t)r, pos.data());
        for (int k = 0; k < l; ++k) std::printf(k ? " %d" : "%d", pos[k]);
        std::printf("\n");
    }
    return 0;
}

template <typename T>
bool rd(FILE *f, T *v, size_t n = 1) { return std::fread(v, sizeof(T), n, f) == n; }

int gauss(const char *in, const char *out) {
    FILE *f = std::fopen(in, "rb");
    if (!f) return 2;
    int32_t p = 0, cases = 0;
    if (!rd(f, &p) || !rd(f, &cases) || p < 2) return 2;
    std::vector<double> C((size_t)p * p);
    if (!rd(f, C.data(), C.size())) return 2;
    std::vector<double> z(cases);
    for (int c = 0; c < cases; ++c) {
        int32_t h[3], S[8];
        if (!rd(f, h, 3) || !rd(f, S, 8)) return 2;
        const int l = h[2];
        switch (l) {
#define GPUCSL_Z(LV) \
    case LV: z[c] = ref64::gauss_z<LV>(C.data(), p, h[0], h[1], S); break;
            GPUCSL_Z(1) GPUCSL_Z(2) GPUCSL_Z(3) GPUCSL_Z(4) GPUCSL_Z(5) GPUCSL_Z(6) GPUCSL_Z(7) GPUCSL_Z(8)
#undef GPUCSL_Z
            default: return 2;
        }
    }
    std::fclose(f);
    FILE *o = std::fopen(out, "wb");
    std::fwrite(z.data(), sizeof(double), z.size(), o);
    std::fclose(o);
    return 0;
}

int chisq(const char *in, const char *out) {
    FILE *f = std::fopen(in, "rb");
    if (!f) return 2;
    int32_t cases = 0;
    if (!rd(f, &cases)) return 2;
    std::vector<double> pv(cases);
    for (int c = 0; c < cases; ++c) {
        double x;
        int32_t df;
        if (!rd(f, &x) || !rd(f, &df)) return 2;
        pv[c] = pchisq(x, df);
    }
    std::fclose(f);
    FILE *o = std::fopen(out, "wb");
    std::fwrite(pv.data(), sizeof(double), pv.size(), o);
    std::fclose(o);
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc >= 2 && !std::strcmp(argv[1], "layout")) return layout();
    if (argc == 5 && !std::strcmp(argv[1], "unrank")) return unrank(std::atoi(argv[2]), std::atoi(argv[3]), std::atoll(argv[4]));
    if (argc == 4 && !std::strcmp(argv[1], "gauss_z")) return gauss(argv[2], argv[3]);
    if (argc == 4 && !std::strcmp(argv[1], "pchisq")) return chisq(argv[2], argv[3]);
    std::fprintf(stderr, "usage: host_kernels layout | unrank N l count | gauss_z IN OUT | pchisq IN OUT\n");
    return 2;
}
