// gpu_dtw <data> <time_out> <m> <r>
// Drop-in for the reference's GPU_dtw / CPU_dtw command line (README.md:13-34,
// GPU_version1/src/AAAA.cu:17-36): loads the series, derives w = MAX(round(m*r), 1), runs the
// reference pipeline once (ED matrix profile, seed from its 100 smallest entries, top-1 join),
// prints the reference's result lines and appends the seconds to <time_out>.
// Extension: "w=<int>" as the fourth argument passes the band radius directly.
//
// gpu_dtw <data> <time_out> [<m> <r>] --grid
// The reference CPU main's grid (CPU_version1/AAAA.cpp:30-38, which ignores m and r): band
// ratios 0.01, 0.02, 0.05, 0.1 (outer) times m = 32, 64, 128, 256, 512, 1024 (inner), one join
// per configuration over the whole series, each printing its lines and appending its seconds.
// A configuration the series is too short for is reported on stderr and skipped; the grid goes on.
//
// Exit status is non-zero on any error (the reference exits 0 on an unreadable file).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/dtwj.h"

int main(int argc, char **argv) {
    bool grid = false, verbose = false;
    int npos = 0;
    const char *pos[4] = {nullptr, nullptr, nullptr, nullptr};
    for (int k = 1; k < argc; ++k) {
        if (std::strcmp(argv[k], "--verbose") == 0)
            verbose = true;
        else if (std::strcmp(argv[k], "--grid") == 0)
            grid = true;
        else if (npos < 4)
            pos[npos++] = argv[k];
    }
    if (npos < (grid ? 2 : 4)) {
        std::fprintf(stderr, "usage: %s <data> <time_out> <m> <r | w=N> [--verbose]\n"
                             "       %s <data> <time_out> [<m> <r>] --grid [--verbose]\n", argv[0], argv[0]);
        return 2;
    }
    double *ts = nullptr;
    int64_t n = 0;
    int rc = dtwj_load_text(pos[0], &ts, &n);
    if (rc != DTWJ_OK) {
        std::fprintf(stderr, "%s\n", dtwj_last_error());
        return 3;
    }
    dtwj_options opt;
    dtwj_default_options(&opt);
    opt.verbose = verbose ? 1 : 0;
    auto run = [&](int m, int w) -> int {
        std::printf("subsequence length m=%d, band radius w=%d, n=%lld\n", m, w, (long long)n);
        std::fflush(stdout);
        dtwj_result r;
        const int e = dtwj_motif_discovery(ts, n, m, w, pos[1], &opt, &r);
        if (e != DTWJ_OK) std::fprintf(stderr, "error %d (m=%d, w=%d): %s\n", e, m, w, dtwj_last_error());
        std::fflush(stdout);
        return e;
    };
    int status = 0;
    if (grid) {
        static const int kM[] = {32, 64, 128, 256, 512, 1024};
        static const double kR[] = {0.01, 0.02, 0.05, 0.1};
        for (double r : kR)
            for (int m : kM)
                if (run(m, dtwj_band_from_ratio(m, r)) != DTWJ_OK) status = 5;
    } else {
        const int m = std::atoi(pos[2]);
        const int w = std::strncmp(pos[3], "w=", 2) == 0 ? std::atoi(pos[3] + 2)
                                                         : dtwj_band_from_ratio(m, std::atof(pos[3]));
        if (run(m, w) != DTWJ_OK) status = 4;
    }
    dtwj_free(ts);
    return status;
}
