// cgsense_main.cpp -- the reference's per-state driver (CGSense.cpp:27-291)
// written against the drop-in Reco / IO / Trajectory classes, reading its
// configuration from a directory instead of hard-coded paths.
//
//   cgsense_dropin DIR
//   DIR/params.txt : gridX gridY gridZ sX sY sZ nCh cgThreshold maxIter density nStates traj
//                    density: 0 none, 1 Voronoi (estimateDensityV3D over the whole
//                    trajectory), 2 import (DIR/dcf.bin); traj: 0 DIR/traj.bin
//                    (nS x 3 doubles, memory order), 1 wave-CAIPI (DIR/wave.txt:
//                    nOsc dwell gradAmp ampCorrY ampCorrZ shiftY shiftZ fovY fovZ)
//   DIR/state{n}.bin, DIR/sensitivityMaps/map{c}.bin
//   writes DIR/results/b_state{n}.bin, DIR/results/performedIterations.txt,
//          DIR/resultRssq.bin, DIR/traj_out.bin (the trajectory used)
//   cgsense_dropin DIR io : file round trip only (no device): reads state0.bin,
//          the maps and dcf.bin with IO and writes data_mem.bin (memory order),
//          maps_mem.bin, dcf_mem.bin and state_rt.bin (IO::exportDataBin)
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <future>
#include <iostream>
#include <string>
#include <utility>

#include "IO.h"
#include "Reco.h"
#include "Trajectory.h"
#include "functions.h"

class FileTrajectory : public Trajectory {
public:
    FileTrajectory(Reco* r, std::string p, long n) : Trajectory(r), path(std::move(p)), ns(n) {}
    void calculateTraj() override {
        std::vector<double> buf(3 * ns);
        std::FILE* f = std::fopen(path.c_str(), "rb");
        if (!f || std::fread(buf.data(), sizeof(double), buf.size(), f) != buf.size()) {
            std::cout << "\nError: could not read trajectory " << path << "\n";
            std::exit(EXIT_FAILURE);
        }
        std::fclose(f);
        kSpaceTraj.assign(ns, std::vector<double>(3));
        for (long i = 0; i < ns; i++)
            for (int a = 0; a < 3; a++) kSpaceTraj[i][a] = buf[3 * i + a];
    }

private:
    std::string path;
    long ns;
};

int main(int argc, char** argv) {
    if (argc < 2) {
        std::cout << "usage: cgsense_dropin DIR\n";
        return 2;
    }
    const std::string dir = std::string(argv[1]) + "/";
    std::ifstream pf(dir + "params.txt");
    int gx, gy, gz, sx, sy, sz, nch, maxIter, dens, nStates, trajKind;
    double thr;
    if (!(pf >> gx >> gy >> gz >> sx >> sy >> sz >> nch >> thr >> maxIter >> dens >> nStates >> trajKind)) {
        std::cout << "Error: bad params.txt\n";
        return 2;
    }
    Reco reco;
    reco.setGridSize(gx, gy, gz);
    reco.setSampling(sx, sy, sz);
    reco.setNCh(nch);
    std::vector<std::string> files;
    for (int n = 0; n < nStates; n++) files.push_back(dir + "state" + std::to_string(n) + ".bin");
    reco.setBatchFiles(files);
    reco.setOutputParentDir(dir);
    reco.setCGThreshold(thr);
    reco.setMaxIterations(maxIter);
    reco.setDensityCorrection(dens == 2   ? Reco::Density::IMPORT
                              : dens == 1 ? Reco::Density::VORONOI3D
                                          : Reco::Density::NONE);
    reco.setDensityCorrectionPath(dir + "dcf.bin");

    FileTrajectory* ftj = nullptr;
    WaveCaipi* wc = nullptr;
    if (trajKind == 1) {
        std::ifstream wf(dir + "wave.txt");
        int nOsc;
        double dwell, amp, acy, acz, shy, shz, fovY, fovZ;
        wf >> nOsc >> dwell >> amp >> acy >> acz >> shy >> shz >> fovY >> fovZ;
        reco.setFovY(fovY);
        reco.setFovZ(fovZ);
        wc = new WaveCaipi(&reco);
        wc->setNOsc(nOsc);
        wc->setDwellTime(dwell);
        wc->setGradAmp(amp);
        wc->setAmpCorrY(acy);
        wc->setAmpCorrZ(acz);
        wc->setShiftY(shy);
        wc->setShiftZ(shz);
        reco.setTrajectory(wc);
    } else {
        ftj = new FileTrajectory(&reco, dir + "traj.bin", (long)sx * sy * sz);
        reco.setTrajectory((Trajectory*)ftj);
    }
    {
        std::vector<double> t;
        for (auto& k : reco.trajectory->kSpaceTraj) t.insert(t.end(), k.begin(), k.end());
        std::FILE* f = std::fopen((dir + "traj_out.bin").c_str(), "wb");
        std::fwrite(t.data(), sizeof(double), t.size(), f);
        std::fclose(f);
    }
    if (nStates == 0) return 0;
    if (argc > 2 && std::string(argv[2]) == "io") {
        const long per = (long)sx * sy * sz;
        fftwf_complex* d = cgs_alloc_complex((size_t)per * nch);
        IO::importDataBin(reco, d, files[0]);
        IO::exportFFTWComplexBin(d, dir + "data_mem.bin", (int)(per * nch));
        IO::exportDataBin(reco, d, dir + "state_rt.bin");
        std::vector<fftwf_complex*> maps(nch);
        fftwf_complex* mb = cgs_alloc_complex((size_t)nch * gx * gy * gz);
        for (int c = 0; c < nch; c++) maps[c] = mb + (size_t)c * gx * gy * gz;
        IO::importSensitivityMaps(maps, gx * gy * gz, nch, dir + "sensitivityMaps/");
        IO::exportFFTWComplexBin(mb, dir + "maps_mem.bin", nch * gx * gy * gz);
        if (dens == 2) IO::exportVectorBin(IO::importDensityCorr(dir + "dcf.bin", (int)per), dir + "dcf_mem.bin");
        cgs_free_complex(d);
        cgs_free_complex(mb);
        delete wc;
        delete ftj;
        return 0;
    }

    reco.init();
    reco.check();
    const int nSamplesPerCh = sx * sy * sz;
    const int cGridSize = gx * gy * gz;
    reco.allocData(&reco.data);
    IO::importDataBin(reco, reco.data, reco.getBatchFileNum(0));
    reco.kernel->calculateKernelFT();
    reco.allocSMaps();
    reco.kernel->calculateWeights();
    if (reco.getDensityCorrection() == Reco::Density::IMPORT)
        reco.denCorr = IO::importDensityCorr(reco.getDensityCorrectionPath(), nSamplesPerCh);
    else if (reco.getDensityCorrection() == Reco::Density::VORONOI3D)
        reco.denCorr = estimateDensityV3D(reco.trajectory->kSpaceTraj, sx, sy, sz, false, 1.0);
    // Buffered data loading as in the reference's main (CGSense.cpp:204-216):
    // while a state is reconstructed, the next state's file is read and
    // permuted into reco.dataBuffer on a host thread; the two buffers swap at
    // the start of the next state.
    const bool multipleData = nStates > 1;
    if (multipleData) reco.allocData(&reco.dataBuffer);
    for (int n = 0; n < nStates; n++) {
        std::future<bool> nextImport;   // its destructor waits for the read
        if (multipleData) {
            if (n != 0) std::swap(reco.data, reco.dataBuffer);
            if (n < nStates - 1) {
                fftwf_complex* buf = reco.dataBuffer;
                const std::string path = reco.getBatchFileNum(n + 1);
                nextImport = std::async(std::launch::async,
                                        [&reco, buf, path]() { return IO::importDataBin(reco, buf, path); });
            }
        }
        if (n > 0) {
            reco.kernel->clearWeights();
            reco.kernel->calculateWeights();
        }
        reco.kernel->excludeMissingData();
        performDensityCorrection(reco, reco.data, reco.denCorr, nSamplesPerCh, nch);
        reco.allocImg();
        reco.nufftBWD();
        IO::exportVectorBin(rssq2DVecImg(reco.img, cGridSize), dir + "resultRssq.bin");
        if (n == 0) IO::importSensitivityMaps(reco.sMaps, cGridSize, nch, dir + "sensitivityMaps/");
        reco.cgSenseDynamic();
        reco.finalize();
        IO::exportFFTWComplexBin(reco.cgSenseImg, dir + "results/b_state" + std::to_string(n) + ".bin", cGridSize);
    }
    delete wc;
    delete ftj;
    std::cout << "Finished.\n";
    return 0;
}
