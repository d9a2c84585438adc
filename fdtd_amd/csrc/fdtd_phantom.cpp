// Voxel phantom input (fdtd.cu:875-939): the tissue table and the PGM
// slices <dir>/v1_%05d.pgm, read into uint8 tissue ids.
//
// Two read orders:
//   FDTD_PGM_REFERENCE  nx*ny codes in file order after the 6-token header,
//                       whatever the image width (the reference's loop,
//                       fdtd.cu:920-933, which shears the phantom when the
//                       grid width differs from the image width);
//   FDTD_PGM_IMAGE      the code of image pixel (column ix, row iy); cells
//                       outside the image are code 0 (air).
// Slices are parsed in parallel (one contiguous run of slices per thread)
// with a hand-written integer scanner; the first error in slice order is
// reported.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "fdtd_internal.h"

namespace {

bool read_file(const char *path, std::vector<char> &buf)
{
    FILE *f = fopen(path, "rb");
    if (!f) return false;
    fseek(f, 0, SEEK_END);
    const long len = ftell(f);
    fseek(f, 0, SEEK_SET);
    buf.resize((size_t)std::max(len, 0L) + 1);
    const size_t got = len > 0 ? fread(buf.data(), 1, (size_t)len, f) : 0;
    fclose(f);
    buf[got] = '\0';
    return true;
}

struct Scanner {
    const char *c;
    bool token(const char *&s, const char *&e)
    {
        while (*c && isspace((unsigned char)*c)) c++;
        s = c;
        while (*c && !isspace((unsigned char)*c)) c++;
        e = c;
        return s != e;
    }
    // Next decimal integer (optional sign); false at end or on a non-number.
    bool integer(long &v)
    {
        while (*c && isspace((unsigned char)*c)) c++;
        bool neg = false;
        if (*c == '-' || *c == '+') neg = *c++ == '-';
        if (!isdigit((unsigned char)*c)) return false;
        long x = 0;
        for (; isdigit((unsigned char)*c); c++)
            x = std::min(x * 10 + (*c - '0'), 1L << 40);   // saturates: every range check still fails
        v = neg ? -x : x;
        return true;
    }
    // PGM header with comments: magic, width, height, maxval.
    bool pgm_header(int &w, int &h)
    {
        const char *s, *e;
        if (!token(s, e) || e - s != 2 || s[0] != 'P' || s[1] != '2') return false;
        long v[3];
        for (int q = 0; q < 3; q++) {
            while (*c && isspace((unsigned char)*c)) c++;
            while (*c == '#') {
                while (*c && *c != '\n') c++;
                while (*c && isspace((unsigned char)*c)) c++;
            }
            if (!integer(v[q])) return false;
        }
        if (v[0] < 1 || v[1] < 1 || v[0] > (1L << 30) || v[1] > (1L << 30)) return false;
        w = (int)v[0];
        h = (int)v[1];
        return true;
    }
};

}  // namespace

int phantom_read_table(const char *path, float *specs, std::string &err)
{
    // fdtd.cu:887-904: up to 94 rows "code name sigma eps_s eps_inf tau";
    // a code >= 94 is stored and ends the table.  Codes are rows of a
    // 95-row table (the reference's specs[94] row is its out-of-bounds
    // write, kept here as parsed).
    std::fill(specs, specs + 95 * 4, 0.0f);
    std::vector<char> buf;
    if (!read_file(path, buf)) {
        err = std::string("Couldn't open file ") + path;
        return FDTD_EIO;
    }
    Scanner sc{buf.data()};
    int rows = 0;
    for (int r = 0; r < 94; r++) {
        const char *s[6], *e[6];
        bool ok = true;
        for (int q = 0; q < 6 && ok; q++) ok = sc.token(s[q], e[q]);
        if (!ok) break;
        std::string t0(s[0], e[0]);
        char *end = nullptr;
        const long code = strtol(t0.c_str(), &end, 10);
        if (*end != '\0') break;  // a comment line ends the table
        if (code < 0 || code > 94) {
            err = std::string("Tissue code out of range in ") + path;
            return FDTD_ERANGE;
        }
        for (int q = 0; q < 4; q++) specs[code * 4 + q] = strtof(std::string(s[2 + q], e[2 + q]).c_str(), nullptr);
        rows++;
        if (code >= 94) break;
    }
    return rows;
}

int phantom_read_slice(const char *path, int nx, int ny, int mode, unsigned char *ids, std::vector<char> &buf,
                       std::string &err)
{
    if (!read_file(path, buf)) {
        err = std::string("Couldn't open file ") + path;
        return FDTD_EIO;
    }
    Scanner sc{buf.data()};
    const size_t nxy = (size_t)nx * ny;
    long v = 0, top = 0;
    if (mode == FDTD_PGM_REFERENCE) {
        // fscanf("%s %s %s %d %d %s") then nx*ny "%d" in order
        const char *s, *e;
        for (int q = 0; q < 6; q++)
            if (!sc.token(s, e)) {
                err = std::string("Malformed phantom slice ") + path;
                return FDTD_EIO;
            }
        for (size_t t = 0; t < nxy; t++) {
            if (!sc.integer(v)) {
                err = std::string("Short phantom slice ") + path;
                return FDTD_EIO;
            }
            if (v < 0 || v > 94) {
                err = std::string("Tissue code out of range in ") + path;
                return FDTD_ERANGE;
            }
            ids[t] = (unsigned char)v;
            top = std::max(top, v);
        }
        return (int)top;
    }
    int w = 0, h = 0;
    if (!sc.pgm_header(w, h)) {
        err = std::string("Malformed phantom slice ") + path;
        return FDTD_EIO;
    }
    memset(ids, 0, nxy);
    for (int iy = 0; iy < h; iy++)
        for (int ix = 0; ix < w; ix++) {
            if (!sc.integer(v)) {
                err = std::string("Short phantom slice ") + path;
                return FDTD_EIO;
            }
            if (v < 0 || v > 94) {
                err = std::string("Tissue code out of range in ") + path;
                return FDTD_ERANGE;
            }
            if (ix < nx && iy < ny) {
                ids[(size_t)iy * nx + ix] = (unsigned char)v;
                top = std::max(top, v);
            }
        }
    return (int)top;
}

extern "C" int fdtd_load_tissue_table(const char *specs_path, float *table95)
{
    if (!specs_path || !table95) return fdtd_fail(FDTD_EINVAL, "null argument");
    std::string err;
    const int rc = phantom_read_table(specs_path, table95, err);
    return rc < 0 ? fdtd_fail(rc, "%s", err.c_str()) : rc;
}

extern "C" int fdtd_load_phantom_ids(const char *dir, int nx, int ny, int nz, int mode, int threads,
                                     unsigned char *ids)
{
    if (!dir || !ids || nx < 1 || ny < 1 || nz < 1) return fdtd_fail(FDTD_EINVAL, "bad phantom arguments");
    if (mode != FDTD_PGM_REFERENCE && mode != FDTD_PGM_IMAGE) return fdtd_fail(FDTD_EINVAL, "pgm mode %d", mode);
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    threads = std::min(threads, nz);
    std::vector<int> rc(nz, 0);
    std::vector<std::string> err(nz);
    auto work = [&](int z0, int z1) {
        std::vector<char> buf;
        char name[4096];
        for (int iz = z0; iz < z1; iz++) {
            snprintf(name, sizeof name, "%s/v1_%05d.pgm", dir, iz + 1);
            rc[iz] = phantom_read_slice(name, nx, ny, mode, ids + (size_t)iz * nx * ny, buf, err[iz]);
            if (rc[iz] < 0) return;
        }
    };
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; t++) {
        const int z0 = (int)((long long)nz * t / threads), z1 = (int)((long long)nz * (t + 1) / threads);
        pool.emplace_back(work, z0, z1);
    }
    for (auto &th : pool) th.join();
    int top = 0;
    for (int iz = 0; iz < nz; iz++) {
        if (rc[iz] < 0) return fdtd_fail(rc[iz], "%s", err[iz].c_str());
        top = std::max(top, rc[iz]);
    }
    return top;
}
