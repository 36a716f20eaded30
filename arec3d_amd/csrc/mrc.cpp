// MRC image files, the reference's reader and writer (arecImageIO.cpp:10-173,
// header layout arecImageIO.h:5-63): a 1024-byte header, then nx*ny*nz values
// x fastest.  Reading takes mode 2 (float32) and modes 1 and 6 (16-bit,
// converted as signed values, as the reference does); writing produces mode 2
// with the reference's header fields (cell 1x1x1 at 90 degrees, axes 1/2/3,
// min/max/mean of the data, "MAP", rms 1, no labels).  Host byte order, as the
// reference (no endian check).  Host code only: no device calls.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "arec_internal.hpp"

namespace arec {
namespace {

struct MrcHeader {
    int32_t nx, ny, nz, mode;
    int32_t nxstart, nystart, nzstart;
    int32_t mx, my, mz;
    float xlen, ylen, zlen;
    float alpha, beta, gamma;
    int32_t mapc, mapr, maps;
    float amin, amax, amean;
    int32_t ispg, nsymbt;
    int32_t user[25];
    float xorigin, yorigin, zorigin;
    char map[4];
    int32_t machinestamp;
    float rms;
    int32_t nlabels;
    char labels[10][80];
};
static_assert(sizeof(MrcHeader) == 1024, "MRC header is 1024 bytes");
static_assert(offsetof(MrcHeader, amin) == 76 && offsetof(MrcHeader, map) == 208 &&
                  offsetof(MrcHeader, nlabels) == 220,
              "MRC header field offsets");

struct File {
    FILE *f = nullptr;
    ~File() {
        if (f) fclose(f);
    }
};

int read_header(const char *path, MrcHeader *h, File *file) {
    file->f = fopen(path, "rb");
    if (!file->f) {
        set_error(std::string("MRC: failed to open ") + path);
        return -1;
    }
    if (fread(h, sizeof(MrcHeader), 1, file->f) != 1) {
        set_error(std::string("MRC: short header in ") + path);
        return -1;
    }
    return 0;
}

}  // namespace
}  // namespace arec

using namespace arec;

extern "C" int arec_mrc_read_header(const char *path, int32_t *nx, int32_t *ny, int32_t *nz, int32_t *mode) {
    if (!path) {
        set_error("arec_mrc_read_header: NULL path");
        return AREC_EINVAL;
    }
    MrcHeader h;
    File file;
    if (read_header(path, &h, &file)) return -1;
    if (nx) *nx = h.nx;
    if (ny) *ny = h.ny;
    if (nz) *nz = h.nz;
    if (mode) *mode = h.mode;
    return 0;
}

extern "C" int arec_mrc_read(const char *path, float *data, int64_t capacity) {
    if (!path || !data) {
        set_error("arec_mrc_read: NULL argument");
        return AREC_EINVAL;
    }
    MrcHeader h;
    File file;
    if (read_header(path, &h, &file)) return -1;
    const int64_t n = (int64_t)h.nx * h.ny * h.nz;
    if (h.nx < 0 || h.ny < 0 || h.nz < 0 || n > capacity) {
        set_error("arec_mrc_read: the image does not fit the buffer");
        return AREC_ESIZE;
    }
    if (h.mode == 1 || h.mode == 6) {
        std::vector<int16_t> s((size_t)n);
        if ((int64_t)fread(s.data(), sizeof(int16_t), (size_t)n, file.f) != n) {
            set_error(std::string("arec_mrc_read: failed to read ") + path);
            return -1;
        }
        for (int64_t i = 0; i < n; i++) data[i] = (float)s[i];
    } else if (h.mode == 2) {
        if ((int64_t)fread(data, sizeof(float), (size_t)n, file.f) != n) {
            set_error(std::string("arec_mrc_read: failed to read ") + path);
            return -1;
        }
    } else {
        set_error("arec_mrc_read: invalid mrc mode " + std::to_string(h.mode));
        return -1;
    }
    return 0;
}

extern "C" int arec_mrc_write(const char *path, int32_t nx, int32_t ny, int32_t nz, const float *data) {
    if (!path || (!data && (int64_t)nx * ny * nz > 0) || nx < 0 || ny < 0 || nz < 0) {
        set_error("arec_mrc_write: bad argument");
        return AREC_EINVAL;
    }
    const int64_t n = (int64_t)nx * ny * nz;
    MrcHeader h;
    memset(&h, 0, sizeof(h));
    h.nx = nx;
    h.ny = ny;
    h.nz = nz;
    h.mode = 2;
    h.mx = nx;
    h.my = ny;
    h.mz = nz;
    h.xlen = h.ylen = h.zlen = 1.0f;
    h.alpha = h.beta = h.gamma = 90.0f;
    h.mapc = 1;
    h.mapr = 2;
    h.maps = 3;
    // min, max and mean in double, mean divided by the count as a float
    // (arecImageIO.cpp:117-125; the count is taken in 64 bits, which differs
    // from the reference's int product only where that overflows)
    double amin = 1.0e20, amax = -1.0e20, mean = 0.0;
    for (int64_t i = 0; i < n; i++) {
        if (data[i] > amax) amax = data[i];
        if (data[i] < amin) amin = data[i];
        mean = mean + data[i];
    }
    mean = mean / (float)n;
    h.amin = (float)amin;
    h.amax = (float)amax;
    h.amean = (float)mean;
    for (int i = 0; i < 25; i++) h.user[i] = ' ';
    h.user[24] = 0;
    h.map[0] = 'M';
    h.map[1] = 'A';
    h.map[2] = 'P';
    h.machinestamp = 0x00004444;  // little-endian stamp (the reference leaves it unset)
    h.rms = 1.0f;
    FILE *f = fopen(path, "wb");
    if (!f) {
        set_error(std::string("arec_mrc_write: failed to open ") + path);
        return -1;
    }
    int rc = 0;
    if (fwrite(&h, sizeof(h), 1, f) != 1 || (int64_t)fwrite(data, sizeof(float), (size_t)n, f) != n) {
        set_error(std::string("arec_mrc_write: failed to write ") + path);
        rc = -1;
    }
    fclose(f);
    return rc;
}
