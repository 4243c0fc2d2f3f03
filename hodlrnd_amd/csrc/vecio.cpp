// vecio.cpp -- the reference's vector files: cereal's portable binary
// archive of a std::vector<double> (myHeaders.hpp:146-170; the data/
// fixtures x_10000.bin and rhs_10000.bin are such files).  Layout: one byte
// 0x01 (little-endian archive), u64 element count, the doubles.
//
// Host code only (no HIP): the sanitizer builds compile it with the host
// compiler.
#include <cstdio>

#include "../../include/hodlrnd.h"

extern "C" int hodlr_vec_save(const char *path, int64_t n, const double *v) {
    if (!path || n < 0 || (n > 0 && !v)) return HODLR_E_ARG;
    FILE *f = std::fopen(path, "wb");
    if (!f) return HODLR_E_ARG;
    const unsigned char flag = 1;
    const uint64_t cnt = (uint64_t)n;
    bool ok = std::fwrite(&flag, 1, 1, f) == 1 && std::fwrite(&cnt, 8, 1, f) == 1 &&
              (n == 0 || std::fwrite(v, sizeof(double), (size_t)n, f) == (size_t)n);
    ok = (std::fclose(f) == 0) && ok;
    return ok ? HODLR_OK : HODLR_E_ARG;
}

extern "C" int hodlr_vec_load(const char *path, int64_t cap, double *v, int64_t *n) {
    if (!path || !n || cap < 0) return HODLR_E_ARG;
    FILE *f = std::fopen(path, "rb");
    if (!f) return HODLR_E_ARG;
    unsigned char flag = 0;
    uint64_t cnt = 0;
    int rc = HODLR_OK;
    long size = -1;
    if (std::fseek(f, 0, SEEK_END) == 0) size = std::ftell(f);
    if (size < 9 || std::fseek(f, 0, SEEK_SET) != 0 || std::fread(&flag, 1, 1, f) != 1 || flag != 1 ||
        std::fread(&cnt, 8, 1, f) != 1) {
        rc = HODLR_E_ARG;
    } else if (cnt > (uint64_t)(size - 9) / sizeof(double)) {  // the count promises more than the file holds
        rc = HODLR_E_ARG;
    } else {
        *n = (int64_t)cnt;
        if (v && (int64_t)cnt > cap) rc = HODLR_E_ARG;
        else if (v && cnt && std::fread(v, sizeof(double), (size_t)cnt, f) != (size_t)cnt) rc = HODLR_E_ARG;
    }
    std::fclose(f);
    return rc;
}
