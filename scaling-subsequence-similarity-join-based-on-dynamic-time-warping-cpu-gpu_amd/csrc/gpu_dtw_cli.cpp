// gpu_dtw <data> <time_out> <m> <r>
// Drop-in for the reference's GPU_dtw / CPU_dtw command line (README.md:13-34,
// GPU_version1/src/AAAA.cu:17-36): loads the series, derives w = MAX(round(m*r), 1), runs the
// top-1 join once, prints the reference's result lines and appends the seconds to <time_out>.
// Extension: "w=<int>" as the fourth argument passes the band radius directly.
// Exit status is non-zero on any error (the reference exits 0 on an unreadable file).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/dtwj.h"

int main(int argc, char **argv) {
    if (argc < 5) {
        std::fprintf(stderr, "usage: %s <data> <time_out> <m> <r | w=N> [--verbose]\n", argv[0]);
        return 2;
    }
    double *ts = nullptr;
    int64_t n = 0;
    int rc = dtwj_load_text(argv[1], &ts, &n);
    if (rc != DTWJ_OK) {
        std::fprintf(stderr, "%s\n", dtwj_last_error());
        return 3;
    }
    const int m = std::atoi(argv[3]);
    int w;
    if (std::strncmp(argv[4], "w=", 2) == 0)
        w = std::atoi(argv[4] + 2);
    else
        w = dtwj_band_from_ratio(m, std::atof(argv[4]));
    dtwj_options opt;
    dtwj_default_options(&opt);
    for (int k = 5; k < argc; ++k)
        if (std::strcmp(argv[k], "--verbose") == 0) opt.verbose = 1;
    std::printf("subsequence length m=%d, band radius w=%d, n=%lld\n", m, w, (long long)n);
    dtwj_result r;
    rc = dtwj_motif_discovery(ts, n, m, w, argv[2], &opt, &r);
    dtwj_free(ts);
    if (rc != DTWJ_OK) {
        std::fprintf(stderr, "error %d: %s\n", rc, dtwj_last_error());
        return 4;
    }
    return 0;
}
