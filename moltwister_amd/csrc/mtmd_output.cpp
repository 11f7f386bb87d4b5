// mtmd_output.cpp - trajectory and histogram files of mtmd_run (host only).
//
// Formats follow the reference byte for byte (Cmd/MolTwisterCmdMolDyn/MDLoop/
// MDLoop.cpp:229-258,325-366; Utilities/DCDFile.cpp:137-246,275-330,403-465,
// 670-729). DCD is the CHARMM binary layout with Fortran record markers,
// little-endian as written by an x86-64 host:
//   main header  84 | "CORD" nsets istart nsavc 0x6 delta(f32) 1 0x8 1 | 84
//                164 | 2 title[80] title[80] | 164      4 | natoms | 4
//   each frame   48 | boxX angX boxY angY angZ boxZ (f64) | 48
//                4n | x[n] (f32) | 4n   4n | y[n] | 4n   4n | z[n] | 4n
// The unit-cell doubles keep the reference's slot order (angZ before boxZ).
#include "mtmd_output.h"

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

struct Bytes {
    std::vector<unsigned char> b;
    template <class T> void put(T v)
    {
        unsigned char tmp[sizeof(T)];
        std::memcpy(tmp, &v, sizeof(T));
        b.insert(b.end(), tmp, tmp + sizeof(T));
    }
    void text(const char* s, size_t width)
    {
        // fixed-width field, NUL-padded (std::string::resize of the reference)
        const size_t len = std::min(std::strlen(s), width);
        b.insert(b.end(), s, s + len);
        b.insert(b.end(), width - len, (unsigned char)0);
    }
};

bool write_all(FILE* f, const Bytes& x)
{
    return std::fwrite(x.b.data(), 1, x.b.size(), f) == x.b.size();
}

} // namespace

int mt_xyz_append(const char* path, int n, int t, const double* xyz, const int32_t* type,
                  const char* const* type_names, int n_types)
{
    FILE* f = std::fopen(path, "a+");
    if (!f) return -1;
    std::fprintf(f, "%i\r\nFrame %u\r\n", n, (unsigned)t);
    char fallback[32];
    for (int k = 0; k < n; k++) {
        const int ty = type[k];
        const char* name = (type_names && ty >= 0 && ty < n_types && type_names[ty]) ? type_names[ty] : nullptr;
        if (!name) {
            std::snprintf(fallback, sizeof fallback, "%d", ty);
            name = fallback;
        }
        std::fprintf(f, "%-15s% -15g%-15g%-15g\r\n", name, xyz[3 * k], xyz[3 * k + 1], xyz[3 * k + 2]);
    }
    std::fclose(f);
    return 0;
}

int mt_dcd_create_if_missing(const char* path, int n_steps, int stride, double timestep_fs, int n_atoms)
{
    if (FILE* probe = std::fopen(path, "r")) {
        std::fclose(probe);
        return 0;
    }
    FILE* f = std::fopen(path, "w");
    if (!f) return -1;
    Bytes h;
    h.put<uint32_t>(84);
    h.text("CORD", 4);
    h.put<uint32_t>((uint32_t)(stride ? n_steps / stride : 0));   // frames announced
    h.put<uint32_t>(0);                                           // first step
    h.put<uint32_t>((uint32_t)stride);                            // steps between frames
    for (int i = 0; i < 6; i++) h.put<uint32_t>(0);
    h.put<float>((float)timestep_fs);
    const uint32_t tail[13] = {1, 0, 0, 0, 0, 0, 0, 0, 0, 1, 84, 164, 2};
    for (uint32_t v : tail) h.put<uint32_t>(v);
    h.text("Written by MolTwister", 80);
    h.text("---", 80);
    h.put<uint32_t>(164);
    h.put<uint32_t>(4);
    h.put<uint32_t>((uint32_t)n_atoms);
    h.put<uint32_t>(4);
    const bool ok = write_all(f, h);
    std::fclose(f);
    return ok ? 0 : -1;
}

int mt_dcd_append(const char* path, int n_atoms, const int box[3], const double* xyz)
{
    FILE* f = std::fopen(path, "a+");
    if (!f) return -1;
    Bytes r;
    r.b.reserve(56 + 3 * (8 + 4 * (size_t)n_atoms));
    r.put<uint32_t>(48);
    const double cell[6] = {(double)box[0], 0.0, (double)box[1], 0.0, 0.0, (double)box[2]};
    for (double v : cell) r.put<double>(v);
    r.put<uint32_t>(48);
    const uint32_t len = (uint32_t)(4 * (size_t)n_atoms);
    for (int axis = 0; axis < 3; axis++) {
        r.put<uint32_t>(len);
        for (int k = 0; k < n_atoms; k++) r.put<float>((float)xyz[3 * k + axis]);
        r.put<uint32_t>(len);
    }
    const bool ok = write_all(f, r);
    std::fclose(f);
    return ok ? 0 : -1;
}

int mt_store_pdistr(const char* path, const int* hist, int nbins, double maxP, int axis)
{
    // unit conversions Conv_P and Conv_Peta (Integrators/Constants.h:28-29)
    const double conv = (axis < 3) ? 1.0E-2 : 1.0E2;
    FILE* f = std::fopen(path, "w");
    if (!f) return -1;
    std::fprintf(f, "%-20s%-20s\r\n", axis < 3 ? "p[g*AA/(mol*fs)]" : "p[kJ*fs/mol]", "P");
    const double dN = (double)nbins;
    for (int i = 0; i < nbins; i++) {
        const double centre = ((maxP / dN) * (2.0 * (double)i + 1.0) - maxP) * conv;
        std::fprintf(f, "%-20g%-20i\r\n", centre, hist[i]);
    }
    std::fclose(f);
    return 0;
}

int mt_store_vdistr(const char* path, const int* hist, int nbins, double maxV)
{
    FILE* f = std::fopen(path, "w");
    if (!f) return -1;
    std::fprintf(f, "%-20s%-20s\r\n", "V[AA^3]", "P");
    for (int i = 0; i < nbins; i++) {
        const double centre = (maxV / (2.0 * (double)nbins)) * (double)(2 * i + 1);
        std::fprintf(f, "%-20g%-20i\r\n", centre, hist[i]);
    }
    std::fclose(f);
    return 0;
}
