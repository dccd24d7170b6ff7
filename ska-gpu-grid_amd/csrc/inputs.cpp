// Host-side readers of the reference's input formats.
//
// UVW text (grid.cu:338-366): one visibility per line,
//   "%d,%d,%d: %f, %f, %f"   (three integers, then u, v, w),
// read as floats and scaled to the grid as the reference does,
//   x = (double)(float)(u * (float)img_dim) / 2048.0   (y from v likewise).
// The reference drops w; it is returned here (as read, in float precision).
// The reference's HDF5 layout (hdf52struct.cpp) needs the HDF5 library,
// which this build does not link.
#include <cstdio>
#include <cstring>

#include "internal.hpp"

using namespace migrid;

extern "C" int migrid_read_uvw_text(const char* path, int64_t img_dim, int64_t max_rows,
                                    double* uv_out, double* w_out, int64_t* nrows_out)
{
    clear_error();
    if (!path || !nrows_out || (max_rows > 0 && !uv_out))
        return fail(MIGRID_EINVAL, "NULL argument");
    if (img_dim < 1 || max_rows < 0) return fail(MIGRID_EINVAL, "img_dim/max_rows");
    FILE* f = std::fopen(path, "r");
    if (!f) return fail(MIGRID_EINVAL, "cannot open %s", path);
    int64_t n = 0, lineno = 0;
    char line[512];
    while (n < max_rows && std::fgets(line, sizeof line, f)) {
        lineno++;
        int a, b, c;
        float u, v, w;
        if (std::sscanf(line, "%d,%d,%d: %f, %f, %f", &a, &b, &c, &u, &v, &w) != 6) {
            bool blank = true;
            for (const char* p = line; *p; p++)
                if (*p != ' ' && *p != '\t' && *p != '\n' && *p != '\r') blank = false;
            if (blank) continue;
            std::fclose(f);
            *nrows_out = n;
            return fail(MIGRID_EINVAL, "%s: line %lld is not \"%%d,%%d,%%d: %%f, %%f, %%f\"", path,
                        (long long)lineno);
        }
        const float fi = (float)img_dim;
        uv_out[2 * n] = (double)(u * fi) / 2048.0;
        uv_out[2 * n + 1] = (double)(v * fi) / 2048.0;
        if (w_out) w_out[n] = (double)w;
        n++;
    }
    std::fclose(f);
    *nrows_out = n;
    return MIGRID_OK;
}
