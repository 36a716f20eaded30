// Caller of the reference's MRC interface (arecImageIO.h:64-66) linked with
// libarec3d_dropin.so: writes a stack, reads its header and data back.  Host
// only.  argv[1] = a writable directory.  Exit status 0 = all checks passed.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "arec3d_dropin.h"

int main(int argc, char **argv) {
    if (argc < 2) return 2;
    const std::string path = std::string(argv[1]) + "/dropin.mrc";
    arecImage img;
    arecAllocateCBImage(&img, 7, 5, 3);
    for (int i = 0; i < 7 * 5 * 3; i++) img.data[i] = 0.25f * i - 9.0f;
    int fails = 0;
    if (arecWriteImage((char *)path.c_str(), img) != 0) fails++;
    int nx = 0, ny = 0, nz = 0;
    if (arecReadImageHeader((char *)path.c_str(), &nx, &ny, &nz) != 0 || nx != 7 || ny != 5 || nz != 3) fails++;
    arecImage back;
    memset(&back, 0, sizeof(back));
    if (arecReadImage((char *)path.c_str(), &back) != 0 || back.nx != 7 || back.ny != 5 || back.nz != 3 ||
        memcmp(back.data, img.data, sizeof(float) * 105) != 0)
        fails++;
    free(back.data);
    const std::string missing = std::string(argv[1]) + "/missing.mrc";
    if (arecReadImageHeader((char *)missing.c_str(), &nx, &ny, &nz) != -1) fails++;
    arecImageFree(&img);
    printf(fails ? "mrc driver: %d failures\n" : "mrc driver: all checks passed%.0d\n", fails);
    return fails ? 1 : 0;
}
