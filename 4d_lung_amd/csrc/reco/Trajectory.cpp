// Trajectory.cpp -- kmax and the wave-CAIPI sample positions (the physics of
// the reference's Trajectory.cpp:9-184, restated).
//
// The wave gradients are sinusoids sampled on the 10 us gradient raster,
// held constant over each raster interval, and integrated with the trapezoid
// rule on a raster 10^4 times finer; a readout sample takes the k value at
// the nearest fine-raster point of its (delayed) acquisition time.  Positions
// are then normalised into [0, 1] with a border of 42 grid points (reduced by
// the wave amplitude along y and z) and permuted from acquisition order
// (readout fastest) into memory order (readout slowest).
#include <algorithm>
#include <cmath>
#include <iostream>

#include "Reco.h"
#include "Trajectory.h"
#include "globalVars.h"

// Trajectory::calculateKMax (Trajectory.cpp:9-16)
void Trajectory::calculateKMax() {
    const double ky = r->sY / (double)(r->fovY * 2.0 * gammaBar);
    const double kz = r->sZ / (double)(r->fovZ * 2.0 * gammaBar);
    kmax = {{0.0, 1.0}, {ky, -ky}, {kz, -kz}};
}

namespace {

// Trapezoid integral of a sample-and-hold waveform g (one value per raster
// interval) on a raster `os` times finer with step dt.
std::vector<double> integrate_held(const std::vector<double>& g, int os, double dt) {
    const size_t n = g.size() * (size_t)os;
    std::vector<double> k(n, 0.0);
    for (size_t i = 1; i < n; ++i) {
        const double a = g[(i - 1) / os];
        const double b = g[i / os];
        k[i] = k[i - 1] + dt * 0.5 * (a + b);
    }
    return k;
}

// Index of the fine-raster point nearest to the fraction t/T of the waveform,
// clamped to the last point (Trajectory.cpp:95-98).
size_t nearest_point(double t, double T, size_t n) {
    const long e = std::lround(t / T * (double)n);
    if (e < 0 || (size_t)e > n - 1) return n - 1;
    return (size_t)e;
}

}  // namespace

// WaveCaipi::calculateTraj (Trajectory.cpp:18-184)
void WaveCaipi::calculateTraj() {
    std::cout << "Calculating wave-CAIPI trajectory...\n";
    const int sx = r->sX, sy = r->sY, sz = r->sZ;
    const long ns = (long)sx * sy * sz;
    const double tAdc = dwellTime * sx;
    const double raster = 10 * std::pow(10, -6);
    const int os = 10000;
    const double dtFine = raster / (double)os;
    const int nGy = (int)(tAdc / raster);
    const double tZ = tAdc + 0.5 * tAdc / (double)nOsc;   // the z wave runs half a period longer
    const int nGz = (int)(tZ / raster);
    const double s2 = std::pow(std::sin(nOsc * PI / (double)(2 * nOsc)), 2);
    const double ampY = gradAmp * ampCorrY * tAdc * s2 / (double)(nOsc * PI);
    const double ampZ = gradAmp * ampCorrZ * tAdc * s2 / (double)(nOsc * PI);

    std::vector<double> gy(nGy), gz(nGz);
    for (int i = 0; i < nGy; i++) gy[i] = -gradAmp * ampCorrY * std::sin(2.0 * PI * nOsc * double(i) / double(nGy));
    for (int i = 0; i < nGz; i++) gz[i] = gradAmp * ampCorrZ * std::sin(2.0 * PI * (nOsc + 0.5) * double(i) / double(nGz));
    const std::vector<double> ky = integrate_held(gy, os, dtFine);
    const std::vector<double> kz = integrate_held(gz, os, dtFine);

    // helix along the readout
    std::vector<double> hx(sx), hy(sx), hz(sx);
    for (int i = 0; i < sx; i++) {
        const double ty = (i + 1) * dwellTime - shiftY;
        const double tz = tAdc / double(4.0 * nOsc) + (i + 1) * dwellTime - shiftZ;
        hx[i] = i / (double)(sx - 1.0);
        hy[i] = ty < 0 ? 0.0 : ky[nearest_point(ty, tAdc, ky.size())];
        hz[i] = kz[nearest_point(tz, tZ, kz.size())];
    }
    // Cartesian phase-encode positions (Trajectory.cpp:113-120)
    std::vector<double> by(sy), bz(sz);
    for (int i = 0; i < sy; i++) by[i] = kmax[1][0] - i / (double)(sy - 1.0) * (kmax[1][0] - kmax[1][1]);
    for (int i = 0; i < sz; i++) bz[i] = kmax[2][1] - i / (double)(sz - 1.0) * (kmax[2][1] - kmax[2][0]);

    // acquisition order: readout fastest, then y line, then z line
    mat2d<double> acq(ns, std::vector<double>(3));
    double minY = 0, maxY = 0, minZ = 0, maxZ = 0;   // extremes start at 0 (Trajectory.cpp:129-142)
    for (long i = 0; i < ns; i++) {
        const long ro = i % sx;
        acq[i][0] = hx[ro];
        acq[i][1] = by[(i / sx) % sy] + hy[ro];
        acq[i][2] = bz[i / ((long)sx * sy)] + hz[ro];
        maxY = std::max(maxY, acq[i][1]);
        minY = std::min(minY, acq[i][1]);
        maxZ = std::max(maxZ, acq[i][2]);
        minZ = std::min(minZ, acq[i][2]);
    }
    // border of 42 grid points, reduced by the wave amplitude (Trajectory.cpp:144-164)
    const double border = 42.0;
    const double gX = r->getGridSizeX(), gY = r->getGridSizeY(), gZ = r->getGridSizeZ();
    const double fX = (1.0 * border) / gX;
    const double fY = (1.0 * border - 0.405 * ampY / double(std::fabs(kmax[1][0]) + std::fabs(kmax[1][1])) * gY) / gY;
    const double fZ = (1.0 * border - 0.405 * ampZ / double(std::fabs(kmax[2][0]) + std::fabs(kmax[2][1])) * gZ) / gZ;
    for (long i = 0; i < ns; i++) {
        double* p = acq[i].data();
        p[1] -= minY;
        p[2] -= minZ;
        p[1] = p[1] / (double)(maxY - minY);
        p[2] = p[2] / (double)(maxZ - minZ);
        p[0] *= (1.0 - 2 * fX);
        p[1] *= (1.0 - 2 * fY);
        p[2] *= (1.0 - 2 * fZ);
        p[0] += fX;
        p[1] += fY;
        p[2] += fZ;
    }
    // memory order: readout slowest, z line fastest
    kSpaceTraj.assign(ns, std::vector<double>(3));
    for (long i = 0; i < ns; i++) {
        const long izl = i / ((long)sx * sy), iyl = (i / sx) % sy, ro = i % sx;
        kSpaceTraj[izl + sz * iyl + (long)sy * sz * ro] = acq[i];
    }
    std::cout << "Finished calculating trajectory\n";
}
