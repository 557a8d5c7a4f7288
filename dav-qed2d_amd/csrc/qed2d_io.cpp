// qed2d_io.cpp -- gauge-configuration files in the reference's text format
// (src/QED2dLoadSave.jl:1-19): DelimitedFiles.writedlm of hcat(real, imag) of
// reshape(U, 2V), i.e. one line per link in column-major order (i1 fastest,
// then i2, then the direction), "re<TAB>im".  Numbers are written in the
// shortest form that reads back to the same double.
#include <hip/hip_runtime.h>

#include <charconv>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/qed2d.h"

extern "C" int qed2d_write_links_text(const double *U_host, long nlinks, const char *path)
{
    if (!U_host || nlinks < 1 || !path) return QED2D_EINVAL;
    FILE *f = std::fopen(path, "w");
    if (!f) return QED2D_EINVAL;
    std::vector<char> buf(1 << 20);
    size_t pos = 0;
    char num[64];
    for (long i = 0; i < nlinks; ++i) {
        for (int k = 0; k < 2; ++k) {
            auto r = std::to_chars(num, num + sizeof num, U_host[2 * i + k]);
            size_t len = (size_t)(r.ptr - num);
            if (pos + len + 2 > buf.size()) {
                std::fwrite(buf.data(), 1, pos, f);
                pos = 0;
            }
            std::memcpy(buf.data() + pos, num, len);
            pos += len;
            buf[pos++] = (k == 0) ? '\t' : '\n';
        }
    }
    std::fwrite(buf.data(), 1, pos, f);
    int bad = std::ferror(f);
    bad |= std::fclose(f);
    return bad ? QED2D_EINVAL : QED2D_OK;
}

// Reads nlinks "re im" pairs (any blank separator); fails if the file holds
// fewer numbers or more than 2*nlinks numbers.
extern "C" int qed2d_read_links_text(double *U_host, long nlinks, const char *path)
{
    if (!U_host || nlinks < 1 || !path) return QED2D_EINVAL;
    FILE *f = std::fopen(path, "rb");
    if (!f) return QED2D_EINVAL;
    // Read in chunks until EOF: no seek, so pipes and FIFOs work too.
    std::vector<char> text;
    size_t got = 0;
    for (;;) {
        if (text.size() - got < (1u << 16)) text.resize(text.size() + (1u << 20));
        const size_t n = std::fread(text.data() + got, 1, text.size() - got, f);
        got += n;
        if (n == 0) break;
    }
    const int bad = std::ferror(f);
    std::fclose(f);
    if (bad) return QED2D_EINVAL;
    const char *p = text.data(), *end = text.data() + got;
    long k = 0;
    while (p < end) {
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
        if (p >= end) break;
        if (k >= 2 * nlinks) return QED2D_EINVAL;
        double v;
        auto r = std::from_chars(p, end, v);
        if (r.ec != std::errc()) return QED2D_EINVAL;
        U_host[k++] = v;
        p = r.ptr;
    }
    return k == 2 * nlinks ? QED2D_OK : QED2D_EINVAL;
}
