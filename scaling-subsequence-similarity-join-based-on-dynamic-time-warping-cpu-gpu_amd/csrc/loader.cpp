// Text loader: whitespace-separated decimal numbers -> doubles.  Mirrors the reference's
// loadfile (GPU_version1/src/loadfile.cu:14-33, CPU_version1/loadfile.cpp:14-33), which reads
// "inputFile >> value" until the first extraction that fails; unlike the reference it keeps
// FP64 (the GPU reference parses to FP32) and reports an unreadable file as DTWJ_EIO instead of
// exit(0).
//
// What one extraction accepts is the C++ standard library's number grammar in the "C" locale
// (num_get): after skipping white space, the longest prefix of the form
//     [+-] digits [. digits] [(e|E) [+-] digits]
// where the mantissa needs at least one digit (before or after the point) and an exponent
// marker is only taken after a mantissa digit.  The collected characters are then converted
// as a whole; the extraction fails when they do not form a complete number ("1e", "+", "."),
// or when the value overflows to infinity.  Words such as nan or inf and hex floats are not
// numbers there ("0x1p3" reads 0 and then fails at 'x').  Consecutive numbers need no
// separator ("1.5.3" reads 1.5, then .3).
#include <cerrno>
#include <cmath>
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
    std::string tok;
    auto digit = [](char c) { return c >= '0' && c <= '9'; };
    for (;;) {
        while (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r' || *p == '\v' || *p == '\f') ++p;
        if (!*p) break;
        // collect the extraction's characters
        const char *q = p;
        bool mant = false, dot = false;
        if (*q == '+' || *q == '-') ++q;
        for (;; ++q) {
            if (digit(*q)) {
                mant = true;
            } else if (*q == '.' && !dot) {
                dot = true;
            } else {
                break;
            }
        }
        if (mant && (*q == 'e' || *q == 'E')) {
            ++q;
            if (*q == '+' || *q == '-') ++q;
            while (digit(*q)) ++q;
        }
        if (q == p) break;  // nothing collected: the extraction fails
        tok.assign(p, (size_t)(q - p));
        char *end = nullptr;
        errno = 0;
        const double v = std::strtod(tok.c_str(), &end);
        // incomplete collected text ("1e", "-", ".") or overflow: the extraction fails
        if (end != tok.c_str() + tok.size() || std::isinf(v)) break;
        vals.push_back(v);
        p = q;
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
