// One-config driver for the reference CPU join, built by oracle/build_ref.sh.
// Test infrastructure only (the golden-vector generator runs it); never shipped.
//
// usage: ref_dtw <data> <time_out> <m> <w>
//   Calls the reference's DTWMotifDiscoveryGUI(TS, m, w, time_out)
//   (CPU_version1/DTWMotifDiscoveryGUI.cpp:12-27) for exactly one (m, w);
//   the reference main (CPU_version1/AAAA.cpp:30-38) ignores m and r and sweeps a grid.
//   w is passed explicitly so both w rules (round and floor of m*r) can be exercised.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

std::vector<double> loadfile(std::string filename);
void DTWMotifDiscoveryGUI(const std::vector<double> &TS, int subseqlen, int maxwarp,
                          const char *output_time_file_path);

int main(int argc, char **argv) {
    if (argc < 5) {
        std::fprintf(stderr, "usage: %s <data> <time_out> <m> <w>\n", argv[0]);
        return 2;
    }
    std::vector<double> ts = loadfile(argv[1]);
    int m = std::atoi(argv[3]);
    int w = std::atoi(argv[4]);
    DTWMotifDiscoveryGUI(ts, m, w, argv[2]);
    std::fflush(stdout);
    return 0;
}
