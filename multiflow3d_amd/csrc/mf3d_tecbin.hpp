// mf3d_tecbin.hpp -- host-side reader of the reference's checkpoint: the
// per-block tecbin####.bin files written by tecbin (post.for:496-548) and
// read back at restart (initial.for:708-748), and final_ctime.dat
// (written flosol.for:296-299, read initial.for:692-694).  Used to start
// the pressure solver from a real state of a MultiFlow3D run.
//
// tecbin####.bin is Fortran unformatted sequential: every record framed by
// its byte length as a little-endian 32-bit integer before and after.
//   record 1: ttc_i, ttc_j, ttc_k          (3 x int32)
//   record 2: pl                           (int32)
//   record 3: inind, jnind, knind          (3 x int32: -1 where the block
//             ends at a physical non-periodic east / north / top face)
//   then one record per cell, k outermost, then j, i fastest, of 19 FP64:
//   x(i), y(j), z(k), p, pm, ppm, vis, u, um, uum, v, vm, vvm, w, wm, wwm,
//   uvm, uwm, vwm.
// final_ctime.dat is formatted '(i8,3F15.6)' over ntime, ctime, forcn,
// qstpn, count, ntav1_count, ntav2_count: format reversion puts the first
// four on line 1 and the last three on line 2.
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace mf3d_io {

inline thread_local std::string last_error;

inline int fail(int code, const std::string &msg) {
    last_error = msg;
    return code;
}

struct File {
    FILE *f = nullptr;
    ~File() {
        if (f) std::fclose(f);
    }
};

// One framed record of exactly `len` payload bytes into dst.
inline int record(FILE *f, const char *path, uint32_t len, void *dst, const char *what) {
    uint32_t a = 0, b = 0;
    if (std::fread(&a, 4, 1, f) != 1) return fail(MF3D_EFORMAT, std::string(path) + ": file ends before " + what);
    if (a != len)
        return fail(MF3D_EFORMAT, std::string(path) + ": " + what + " record holds " + std::to_string(a) +
                                      " bytes, expected " + std::to_string(len) +
                                      " (not a little-endian tecbin file with 4-byte record markers?)");
    if (std::fread(dst, 1, len, f) != len || std::fread(&b, 4, 1, f) != 1 || b != len)
        return fail(MF3D_EFORMAT, std::string(path) + ": " + what + " record truncated or misframed");
    return 0;
}

inline int header(FILE *f, const char *path, mf3d_tecbin_hdr *h) {
    int rc;
    if ((rc = record(f, path, 12, h->tt, "extent")) || (rc = record(f, path, 4, &h->pl, "pl")) ||
        (rc = record(f, path, 12, h->ind, "boundary flag")))
        return rc;
    for (int d = 0; d < 3; ++d)
        if (h->tt[d] <= 0 || h->tt[d] > (1 << 20))
            return fail(MF3D_EFORMAT, std::string(path) + ": implausible extent " + std::to_string(h->tt[d]));
    if (h->pl < 0 || 2 * h->pl >= h->tt[0] || 2 * h->pl >= h->tt[1] || 2 * h->pl >= h->tt[2])
        return fail(MF3D_EFORMAT, std::string(path) + ": ghost layers " + std::to_string(h->pl) +
                                      " do not fit the extents");
    return 0;
}

}  // namespace mf3d_io

extern "C" {

const char *mf3d_io_last_error(void) { return mf3d_io::last_error.c_str(); }

int mf3d_tecbin_header(const char *path, mf3d_tecbin_hdr *hdr) {
    if (!path || !hdr) return MF3D_EINVAL;
    mf3d_io::File F;
    if (!(F.f = std::fopen(path, "rb"))) return mf3d_io::fail(MF3D_ENOENT, std::string(path) + ": cannot open");
    return mf3d_io::header(F.f, path, hdr);
}

int mf3d_tecbin_read(const char *path, mf3d_tecbin_hdr *hdr, double *const *fields, double *x, double *y,
                     double *z) {
    if (!path) return MF3D_EINVAL;
    mf3d_io::File F;
    if (!(F.f = std::fopen(path, "rb"))) return mf3d_io::fail(MF3D_ENOENT, std::string(path) + ": cannot open");
    mf3d_tecbin_hdr h{};
    int rc = mf3d_io::header(F.f, path, &h);
    if (rc) return rc;
    if (hdr) *hdr = h;
    const long long ni = h.tt[0], nj = h.tt[1], nk = h.tt[2];
    constexpr uint32_t kRec = 19 * 8;
    constexpr size_t kFrame = kRec + 8;
    const size_t chunk = 1 << 16;  // records per read
    std::vector<unsigned char> buf(chunk * kFrame);
    const long long total = ni * nj * nk;
    long long c = 0;
    while (c < total) {
        const size_t want = (size_t)std::min<long long>((long long)chunk, total - c);
        if (std::fread(buf.data(), kFrame, want, F.f) != want)
            return mf3d_io::fail(MF3D_EFORMAT, std::string(path) + ": file ends after " + std::to_string(c) +
                                                   " of " + std::to_string(total) + " cell records");
        for (size_t r = 0; r < want; ++r, ++c) {
            const unsigned char *rec = buf.data() + r * kFrame;
            uint32_t a, b;
            std::memcpy(&a, rec, 4);
            std::memcpy(&b, rec + 4 + kRec, 4);
            if (a != kRec || b != kRec)
                return mf3d_io::fail(MF3D_EFORMAT, std::string(path) + ": cell record " + std::to_string(c) +
                                                       " is not 19 FP64 values");
            double v[19];
            std::memcpy(v, rec + 4, kRec);
            const long long i = c % ni, j = (c / ni) % nj, k = c / (ni * nj);
            if (x && j == 0 && k == 0) x[i] = v[0];
            if (y && i == 0 && k == 0) y[j] = v[1];
            if (z && i == 0 && j == 0) z[k] = v[2];
            if (fields)
                for (int f = 0; f < 16; ++f)
                    if (fields[f]) fields[f][c] = v[3 + f];
        }
    }
    return 0;
}

int mf3d_final_ctime_read(const char *path, mf3d_final_ctime *out) {
    if (!path || !out) return MF3D_EINVAL;
    mf3d_io::File F;
    if (!(F.f = std::fopen(path, "r"))) return mf3d_io::fail(MF3D_ENOENT, std::string(path) + ": cannot open");
    char l1[256] = {}, l2[256] = {};
    if (!std::fgets(l1, sizeof l1, F.f) || !std::fgets(l2, sizeof l2, F.f))
        return mf3d_io::fail(MF3D_EFORMAT, std::string(path) + ": expected two lines");
    // fixed columns: i8, then F15.6 fields
    auto field = [](const char *line, int from, int width, std::string &s) {
        const int n = (int)std::strlen(line);
        s = from < n ? std::string(line + from, std::min(width, n - from)) : std::string();
    };
    std::string s[7];
    field(l1, 0, 8, s[0]);
    field(l1, 8, 15, s[1]);
    field(l1, 23, 15, s[2]);
    field(l1, 38, 15, s[3]);
    field(l2, 0, 8, s[4]);
    field(l2, 8, 15, s[5]);
    field(l2, 23, 15, s[6]);
    double v[7];
    for (int i = 0; i < 7; ++i) {
        char *end = nullptr;
        v[i] = std::strtod(s[i].c_str(), &end);
        bool blank = true;
        for (char ch : s[i]) blank = blank && (ch == ' ' || ch == '\n' || ch == '\r');
        if (blank || !end || end == s[i].c_str())
            return mf3d_io::fail(MF3D_EFORMAT, std::string(path) + ": field " + std::to_string(i + 1) +
                                                   " is not a number");
    }
    out->ntime = (int32_t)v[0];
    out->ctime = v[1];
    out->forcn = v[2];
    out->qstpn = v[3];
    out->count = (int32_t)v[4];
    out->ntav1_count = v[5];
    out->ntav2_count = v[6];
    return 0;
}

}  // extern "C"
