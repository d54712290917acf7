// IO.h -- the reference's raw float32 file formats (drop-in for IO.h:11-29).
// Formats (IO.cpp:12-186):
//   state files      float32 re,im x C*sX*sY*sZ, on disk coil slowest then
//                    z-line, y-line, readout fastest; read into memory order
//                    [c][readout][y][z]
//   map{c}.bin       float32 re,im x Ng, [x][y][z]
//   density file     float32 x sX*sY*sZ in memory order
//   result images    float32 re,im x Ng; float vectors as plain float32
#pragma once
#include <string>
#include <vector>

#include "Reco.h"
#include "cgs_complex.h"

class IO {
public:
    IO() {}
    ~IO() {}

    static void exportVectorBin(const std::vector<float>& v, std::string path);
    static void exportFFTWComplexBin(fftwf_complex* w, const std::string& path, const int& N);

    static void importSensitivityMaps(std::vector<fftwf_complex*>& maps, const int N, const int numCh,
                                      std::string path);
    static bool importDataBin(const Reco& r, fftwf_complex* v, const std::string& path);

    // Declared by the reference (IO.h:25) but never defined there: always
    // returns false (no trajectory file format exists).
    static bool importTrajectoryBin(mat2d<double>& traj, const Reco& r);

    static std::vector<float> importDensityCorr(const Reco& r);
    static std::vector<float> importDensityCorr(const std::string& path, const int& N);

    // Inverse of importDataBin: write [c][readout][y][z] data in the file order.
    static void exportDataBin(const Reco& r, const fftwf_complex* v, const std::string& path);
};
