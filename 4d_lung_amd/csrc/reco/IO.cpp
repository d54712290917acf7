// IO.cpp -- raw float32 readers and writers with the reference's layouts
// and error behaviour (print, exit(EXIT_FAILURE); IO.cpp:12-186).
#include "IO.h"

#include <cstdio>
#include <cstdlib>
#include <iostream>

namespace {

[[noreturn]] void die(const std::string& msg) {
    std::cout << msg;
    std::exit(EXIT_FAILURE);
}

std::FILE* open_or_die(const std::string& path, const char* mode) {
    std::FILE* f = std::fopen(path.c_str(), mode);
    if (!f) die("\nError: Could not open binary file " + path + "\n");
    return f;
}

}  // namespace

void IO::exportVectorBin(const std::vector<float>& v, std::string path) {
    std::FILE* f = open_or_die(path, "wb");
    std::fwrite(v.data(), sizeof(float), v.size(), f);
    std::fclose(f);
}

// IO.cpp:25-51: map{c}.bin under path, 2*N floats each.
void IO::importSensitivityMaps(std::vector<fftwf_complex*>& maps, const int N, const int numCh, std::string path) {
    std::cout << "Importing sensitivity maps\n";
    for (int c = 0; c < numCh; c++) {
        const std::string name = path + "map" + std::to_string(c) + ".bin";
        std::FILE* f = std::fopen(name.c_str(), "rb");
        if (!f) die("\nError: Could not open binary file " + path + "\n");
        const size_t got = std::fread(&maps[c][0][0], sizeof(float), 2 * (size_t)N, f);
        std::fclose(f);
        if (got != 2 * (size_t)N) {
            std::cout << "\nError: Number of imported data points is different from expected number!\nFilename: "
                      << name << "\n" << got << "\t" << 2 * N << "\n";
            std::exit(EXIT_FAILURE);
        }
    }
    std::cout << "Done.\n";
}

// IO.cpp:55-98: the file holds, per coil, the readout fastest, then the y
// line, then the z line; memory order is [c][readout][y][z].
bool IO::importDataBin(const Reco& r, fftwf_complex* v, const std::string& path) {
    std::FILE* f = open_or_die(path, "rb");
    const long sx = r.getSX(), sy = r.getSY(), sz = r.getSZ();
    const long per = sx * sy * sz;
    const long n = per * r.getNCh();
    std::vector<float> buf(2 * (size_t)n);
    const size_t got = std::fread(buf.data(), sizeof(float), buf.size(), f);
    std::fclose(f);
    if (got != buf.size()) {
        std::cout << "\nError: Number of imported data points is different from expected number!\nDoes expected "
                     "precision match file precision?\n";
        std::cout << "Path: " << path << "\n" << 2 * n << "  " << got << "\n";
        std::exit(EXIT_FAILURE);
    }
#pragma omp parallel for schedule(static)
    for (long k = 0; k < n; k++) {
        const long c = k / per, w = k % per;
        const long iz = w / (sx * sy), iy = (w / sx) % sy, ix = w % sx;
        const long m = c * per + iz + sz * iy + sy * sz * ix;
        v[m][0] = buf[2 * k];
        v[m][1] = buf[2 * k + 1];
    }
    return true;
}

void IO::exportDataBin(const Reco& r, const fftwf_complex* v, const std::string& path) {
    const long sx = r.getSX(), sy = r.getSY(), sz = r.getSZ();
    const long per = sx * sy * sz;
    const long n = per * r.getNCh();
    std::vector<float> buf(2 * (size_t)n);
    for (long k = 0; k < n; k++) {
        const long c = k / per, w = k % per;
        const long iz = w / (sx * sy), iy = (w / sx) % sy, ix = w % sx;
        const long m = c * per + iz + sz * iy + sy * sz * ix;
        buf[2 * k] = v[m][0];
        buf[2 * k + 1] = v[m][1];
    }
    std::FILE* f = open_or_die(path, "wb");
    std::fwrite(buf.data(), sizeof(float), buf.size(), f);
    std::fclose(f);
}

bool IO::importTrajectoryBin(mat2d<double>& traj, const Reco& r) {
    (void)traj;
    (void)r;
    return false;
}

std::vector<float> IO::importDensityCorr(const Reco& r) {
    return importDensityCorr(r.getDensityCorrectionPath(), r.getSX() * r.getSY() * r.getSZ());
}

// IO.cpp:129-156: N floats in memory order.
std::vector<float> IO::importDensityCorr(const std::string& path, const int& N) {
    std::cout << "Importing density correction matrix\n";
    std::vector<float> d(N);
    std::FILE* f = open_or_die(path, "rb");
    const size_t got = std::fread(d.data(), sizeof(float), (size_t)N, f);
    std::fclose(f);
    if (got != d.size()) {
        std::cout << "\nError: Number of imported data points is different from expected number!\nDoes expected "
                     "precision match file precision?\n" << d.size() << "  " << got << "\n";
        std::exit(EXIT_FAILURE);
    }
    std::cout << "Done importing density correction matrix\n";
    return d;
}

// IO.cpp:158-186, with the "_1.bin" fallback name when the first open fails.
void IO::exportFFTWComplexBin(fftwf_complex* w, const std::string& path, const int& N) {
    std::FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) {
        std::cout << "\nError: Could not open binary file " << path << "\n\tTrying to rename...";
        f = std::fopen((path + "_1.bin").c_str(), "wb");
        if (!f) die(" Failed!\n");
        std::cout << " Succeeded!\n";
    }
    std::fwrite(&w[0][0], sizeof(float), 2 * (size_t)N, f);
    std::fclose(f);
}
