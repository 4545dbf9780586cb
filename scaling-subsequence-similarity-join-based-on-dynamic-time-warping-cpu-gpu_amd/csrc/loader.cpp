// Text loader: whitespace-separated decimal numbers -> doubles.  Mirrors the reference's
// loadfile (GPU_version1/src/loadfile.cu:14-33, CPU_version1/loadfile.cpp:14-33), which reads
// "inputFile >> value" until the first token that is not a number; unlike the reference it
// keeps FP64 (the GPU reference parses to FP32) and reports an unreadable file as DTWJ_EIO
// instead of exit(0).
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/dtwj.h"
#include "internal.hpp"

extern "C" int dtwj_load_text(const char *path, double **out, int64_t *n) {
    if (!path || !out || !n) {
        dtwj::set_error("null argument");
        return DTWJ_EINVAL;
    }
    *out = nullptr;
    *n = 0;
    FILE *f = std::fopen(path, "rb");
    if (!f) {
        dtwj::set_error(std::string("Error opening file: ") + path);
        return DTWJ_EIO;
    }
    std::string buf;
    char chunk[1 << 16];
    size_t got;
    while ((got = std::fread(chunk, 1, sizeof(chunk), f)) > 0) buf.append(chunk, got);
    std::fclose(f);
    std::vector<double> vals;
    vals.reserve(buf.size() / 8 + 1);
    const char *p = buf.c_str();
    for (;;) {
        while (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r' || *p == '\v' || *p == '\f') ++p;
        if (!*p) break;
        char *end = nullptr;
        errno = 0;
        double v = std::strtod(p, &end);
        if (end == p) break;  // first non-numeric token ends the series, as with operator>>
        vals.push_back(v);
        p = end;
    }
    double *mem = (double *)std::malloc(sizeof(double) * (vals.size() ? vals.size() : 1));
    if (!mem) {
        dtwj::set_error("out of host memory");
        return DTWJ_ENOMEM;
    }
    if (!vals.empty()) std::memcpy(mem, vals.data(), sizeof(double) * vals.size());
    *out = mem;
    *n = (int64_t)vals.size();
    return DTWJ_OK;
}
