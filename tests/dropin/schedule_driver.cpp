// schedule_driver.cpp -- a roDV-style caller of the batch API, for the
// drop-in tests.  It declares pairhmm::schedule::schedule_pairhmm the way the
// reference's header does (src/core/pairhmm_schedule.h:14-25, restated here,
// not taken from this repository's include/), is compiled with g++ and linked
// against librodv_phmm.so: it links only if the library exports the symbol a
// reference caller needs.
//
//   schedule_driver <in> <out> [reps]
// in  (little-endian): u32 M, u32 N, u8 use_double, f64 idle_f, f64 idle_d,
//     u8 verbose; M x (u32 len, len bases); N x (u32 len, len bases, then
//     len bytes each of quality, insertion, deletion, gap continuation).
// out: u8 ok, u32 M, u32 N, M*N f64 result[hap][read]; then u32 reps and
//     reps f64 wall microseconds of each call (the first one creates the
//     engine's context).
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

namespace pairhmm {
namespace schedule {
bool schedule_pairhmm(const std::vector<std::vector<uint8_t>> &haplotypes,
                      const std::vector<std::vector<uint8_t>> &reads,
                      std::vector<std::vector<double>> &result,
                      const std::vector<std::vector<uint8_t>> &quality,
                      const std::vector<std::vector<uint8_t>> &insertion_qualities,
                      const std::vector<std::vector<uint8_t>> &deletion_qualities,
                      const std::vector<std::vector<uint8_t>> &gap_contiguous_qualities,
                      bool use_double = false, double max_idle_ratio_float = 0.1,
                      double max_idle_ratio_double = 0.1, bool verbose = false);
}  // namespace schedule
}  // namespace pairhmm

namespace {

template <typename T>
bool get(FILE *f, T *v) {
    return std::fread(v, sizeof(T), 1, f) == 1;
}

bool get_seq(FILE *f, uint32_t n, std::vector<uint8_t> &s) {
    s.resize(n);
    return n == 0 || std::fread(s.data(), 1, n, f) == n;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s <in> <out> [reps]\n", argv[0]);
        return 2;
    }
    const int reps = argc > 3 ? std::atoi(argv[3]) : 1;
    FILE *f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    uint32_t M = 0, N = 0;
    uint8_t use_double = 0, verbose = 0;
    double idle_f = 0.1, idle_d = 0.1;
    bool ok = get(f, &M) && get(f, &N) && get(f, &use_double) && get(f, &idle_f) && get(f, &idle_d) &&
              get(f, &verbose);
    std::vector<std::vector<uint8_t>> haps(M), reads(N), q(N), qi(N), qd(N), qc(N);
    for (uint32_t h = 0; ok && h < M; h++) {
        uint32_t n = 0;
        ok = get(f, &n) && get_seq(f, n, haps[h]);
    }
    for (uint32_t r = 0; ok && r < N; r++) {
        uint32_t n = 0;
        ok = get(f, &n) && get_seq(f, n, reads[r]) && get_seq(f, n, q[r]) && get_seq(f, n, qi[r]) &&
             get_seq(f, n, qd[r]) && get_seq(f, n, qc[r]);
    }
    std::fclose(f);
    if (!ok) {
        std::fprintf(stderr, "bad input file %s\n", argv[1]);
        return 2;
    }
    std::vector<std::vector<double>> result;
    std::vector<double> us;
    uint8_t res = 0;
    for (int k = 0; k < (reps > 0 ? reps : 1); k++) {
        const auto t0 = std::chrono::steady_clock::now();
        res = pairhmm::schedule::schedule_pairhmm(haps, reads, result, q, qi, qd, qc, use_double != 0, idle_f,
                                                  idle_d, verbose != 0);
        const auto t1 = std::chrono::steady_clock::now();
        us.push_back(std::chrono::duration<double, std::micro>(t1 - t0).count());
    }
    FILE *o = std::fopen(argv[2], "wb");
    if (!o) return 2;
    std::fwrite(&res, 1, 1, o);
    std::fwrite(&M, 4, 1, o);
    std::fwrite(&N, 4, 1, o);
    for (uint32_t h = 0; h < M; h++)
        for (uint32_t r = 0; r < N; r++) {
            const double v = (res && h < result.size() && r < result[h].size()) ? result[h][r] : 0.0;
            std::fwrite(&v, 8, 1, o);
        }
    const uint32_t nr = (uint32_t)us.size();
    std::fwrite(&nr, 4, 1, o);
    std::fwrite(us.data(), 8, us.size(), o);
    std::fclose(o);
    return res ? 0 : 1;
}
