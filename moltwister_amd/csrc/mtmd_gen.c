/*
 * mtmd_gen.c - seeded workload generator helpers (libmtgen.so), used by
 * moltwister_amd/workloads.py for systems too large for Python loops.
 *
 * mtgen_rsa: random sequential addition of n points in a periodic box
 * [-L/2, L/2)^3 with minimum separation dmin, on a grid of cells of edge
 * >= dmin/sqrt(3) (at most one point per cell), candidates drawn from a
 * splitmix64 stream. Returns the number of points placed (n on success).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

static uint64_t sm64(uint64_t* s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static double u01(uint64_t* s) { return (double)(sm64(s) >> 11) * (1.0 / 9007199254740992.0); }

static int try_place(int* grid, int nc, double cs, int reach, double L, double d2, double x, double y, double z,
                     const double* out)
{
    int cx = (int)((x + 0.5 * L) / cs), cy = (int)((y + 0.5 * L) / cs), cz = (int)((z + 0.5 * L) / cs);
    if (cx >= nc) cx = nc - 1;
    if (cy >= nc) cy = nc - 1;
    if (cz >= nc) cz = nc - 1;
    if (grid[((size_t)cz * nc + cy) * nc + cx] >= 0) return -1;
    for (int dz = -reach; dz <= reach; dz++)
        for (int dy = -reach; dy <= reach; dy++)
            for (int dx = -reach; dx <= reach; dx++) {
                int ix = (cx + dx + nc) % nc, iy = (cy + dy + nc) % nc, iz = (cz + dz + nc) % nc;
                int j = grid[((size_t)iz * nc + iy) * nc + ix];
                if (j < 0) continue;
                double rx = out[3 * j] - x, ry = out[3 * j + 1] - y, rz = out[3 * j + 2] - z;
                rx -= L * rint(rx / L);
                ry -= L * rint(ry / L);
                rz -= L * rint(rz / L);
                if (rx * rx + ry * ry + rz * rz < d2) return -1;
            }
    return (int)(((size_t)cz * nc + cy) * nc + cx);
}

/* The same process, accelerated once plain draws mostly fail: candidates are
 * then drawn inside still-open grid cells (uniform within the cell), and a
 * cell is retired after `tries` consecutive failures inside it. */
int mtgen_rsa_fast(int n, double L, double dmin, uint64_t seed, int tries, double* out)
{
    int nc = (int)floor(L / (dmin / sqrt(3.0)));
    if (nc < 1) nc = 1;
    const double cs = L / nc;
    const int reach = (int)ceil(dmin / cs);
    const size_t ncell = (size_t)nc * nc * nc;
    int* grid = (int*)malloc(ncell * sizeof(int));
    if (!grid) return -1;
    for (size_t i = 0; i < ncell; i++) grid[i] = -1;
    const double d2 = dmin * dmin;
    uint64_t st = seed;
    int placed = 0;
    long long window = 0, acc = 0;
    /* phase 1: uniform draws until the acceptance over a window drops below 2% */
    while (placed < n) {
        double x = (u01(&st) - 0.5) * L, y = (u01(&st) - 0.5) * L, z = (u01(&st) - 0.5) * L;
        int c = try_place(grid, nc, cs, reach, L, d2, x, y, z, out);
        window++;
        if (c >= 0) {
            out[3 * placed] = x; out[3 * placed + 1] = y; out[3 * placed + 2] = z;
            grid[c] = placed++;
            acc++;
        }
        if (window == 20000) {
            if (acc < 400) break;
            window = acc = 0;
        }
    }
    /* phase 2: sample open cells */
    size_t nopen = 0;
    int* open = (int*)malloc(ncell * sizeof(int));
    if (!open) { free(grid); return -1; }
    for (size_t i = 0; i < ncell; i++) if (grid[i] < 0) open[nopen++] = (int)i;
    while (placed < n && nopen > 0) {
        size_t pick = (size_t)(sm64(&st) % nopen);
        int c = open[pick];
        int cx = c % nc, cy = (c / nc) % nc, cz = c / (nc * nc);
        int ok = 0;
        for (int t = 0; t < tries && !ok; t++) {
            double x = -0.5 * L + (cx + u01(&st)) * cs;
            double y = -0.5 * L + (cy + u01(&st)) * cs;
            double z = -0.5 * L + (cz + u01(&st)) * cs;
            int cc = try_place(grid, nc, cs, reach, L, d2, x, y, z, out);
            if (cc >= 0) {
                out[3 * placed] = x; out[3 * placed + 1] = y; out[3 * placed + 2] = z;
                grid[cc] = placed++;
                ok = 1;
            }
        }
        open[pick] = open[--nopen];   /* placed in, or retired */
    }
    free(open);
    free(grid);
    return placed;
}

int mtgen_rsa(int n, double L, double dmin, uint64_t seed, long long max_attempts, double* out)
{
    int nc = (int)floor(L / (dmin / sqrt(3.0)));
    if (nc < 1) nc = 1;
    const double cs = L / nc;
    const int reach = (int)ceil(dmin / cs);
    const size_t ncell = (size_t)nc * nc * nc;
    int* grid = (int*)malloc(ncell * sizeof(int));
    if (!grid) return -1;
    for (size_t i = 0; i < ncell; i++) grid[i] = -1;
    const double d2 = dmin * dmin;
    uint64_t st = seed;
    int placed = 0;
    long long att = 0;
    while (placed < n && att < max_attempts) {
        att++;
        double x = (u01(&st) - 0.5) * L, y = (u01(&st) - 0.5) * L, z = (u01(&st) - 0.5) * L;
        int cx = (int)((x + 0.5 * L) / cs), cy = (int)((y + 0.5 * L) / cs), cz = (int)((z + 0.5 * L) / cs);
        if (cx >= nc) cx = nc - 1;
        if (cy >= nc) cy = nc - 1;
        if (cz >= nc) cz = nc - 1;
        if (grid[((size_t)cz * nc + cy) * nc + cx] >= 0) continue;
        int ok = 1;
        for (int dz = -reach; dz <= reach && ok; dz++)
            for (int dy = -reach; dy <= reach && ok; dy++)
                for (int dx = -reach; dx <= reach && ok; dx++) {
                    int ix = (cx + dx + nc) % nc, iy = (cy + dy + nc) % nc, iz = (cz + dz + nc) % nc;
                    int j = grid[((size_t)iz * nc + iy) * nc + ix];
                    if (j < 0) continue;
                    double rx = out[3 * j] - x, ry = out[3 * j + 1] - y, rz = out[3 * j + 2] - z;
                    rx -= L * rint(rx / L);
                    ry -= L * rint(ry / L);
                    rz -= L * rint(rz / L);
                    if (rx * rx + ry * ry + rz * rz < d2) ok = 0;
                }
        if (!ok) continue;
        out[3 * placed] = x;
        out[3 * placed + 1] = y;
        out[3 * placed + 2] = z;
        grid[((size_t)cz * nc + cy) * nc + cx] = placed;
        placed++;
    }
    free(grid);
    return placed;
}
