// A caller written the way the reference's benchmark driver calls gridGPU
// (grid.cu:280-481): padded output buffer, glibc rand() inputs, init_gcf
// table, one gridGPU<double2,double2> call through include/migrid.hpp.
// Usage: dropin_main N S npts out.bin      (writes the padded buffer)
//        dropin_main --bad                 (an odd support: prints the error)
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "migrid.hpp"

struct double2 {
    double x, y;
};

static void make_gcf(double2* gcf, size_t size)  // the reference's init_gcf formula
{
    for (size_t sx = 0; sx < 8; sx++)
        for (size_t sy = 0; sy < 8; sy++)
            for (size_t x = 0; x < size; x++)
                for (size_t y = 0; y < size; y++) {
                    double tmp = sin(6.28 * x / size / 8) *
                                 exp(-(1.0 * x * x + 1.0 * y * y * sy) / size / size / 2);
                    double2& g = gcf[size * size * (sx + sy * 8) + x + y * size];
                    g.x = tmp * sin(1.0 * x * sx / (y + 1)) + 0.4;
                    g.y = tmp * cos(1.0 * x * sx / (y + 1)) - 0.2;
                }
}

int main(int argc, char** argv)
{
    if (argc > 1 && !strcmp(argv[1], "--bad")) {
        double2 dummy[4] = {};
        gridGPU(dummy, dummy, dummy, 1, 64, dummy, 7);
        return 0;
    }
    if (argc < 5) return 2;
    const size_t N = atoi(argv[1]), S = atoi(argv[2]), npts = atoi(argv[3]);
    const size_t P = N * N + 2 * N * S + 2 * S;
    double2* out = (double2*)malloc(sizeof(double2) * P);
    double2* in = (double2*)malloc(sizeof(double2) * npts);
    double2* vals = (double2*)malloc(sizeof(double2) * npts);
    double2* gcf = (double2*)malloc(sizeof(double2) * 64 * S * S);
    make_gcf(gcf, S);
    srand(1541617);
    for (size_t n = 0; n < npts; n++) {
        in[n].x = ((float)rand()) / RAND_MAX * N;
        in[n].y = ((float)rand()) / RAND_MAX * N;
        vals[n].x = ((float)rand()) / RAND_MAX;
        vals[n].y = ((float)rand()) / RAND_MAX;
    }
    gridGPU(out, in, vals, npts, N, gcf, S);
    FILE* f = fopen(argv[4], "wb");
    fwrite(out, sizeof(double2), P, f);
    fclose(f);
    free(out);
    free(in);
    free(vals);
    free(gcf);
    return 0;
}
