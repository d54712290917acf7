// Reco.h -- CG-SENSE reconstruction object, drop-in for the reference's Reco
// class (Reco.h:13-214) with the encoding operator, its adjoint and the CG
// loop running on the MI355X through the C-ABI of cgs.h.
//
// Public members, setters, getters and the call order of the reference's
// main (CGSense.cpp:42-282) are kept: calculateWeights -> excludeMissingData
// -> density correction of data -> nufftBWD (coil images in img) ->
// sensitivity maps in sMaps -> cgSenseDynamic (result in cgSenseImg, residuals
// in cgResiduals, pass count appended to numPerformedCGIterations).
// Errors print the reference's message and exit with EXIT_FAILURE, as the
// reference does.
#pragma once
#include <iostream>
#include <string>
#include <vector>

#include "Fourier.h"
#include "Trajectory.h"
#include "cgs_complex.h"

struct cgs_ctx;

template <class T>
using mat2d = std::vector<std::vector<T>>;

class Reco {
public:
    enum class Density { NONE, VORONOI3D, IMPORT, UNDEFINED };
    Reco();
    ~Reco();

    class ConvKernel {
    public:
        ConvKernel(Reco* r) : sizeX(r->gridSizeX), sizeY(r->gridSizeY), sizeZ(r->gridSizeZ), reco(r) {}
        ~ConvKernel();

        void calculateWeights();
        void calculateKernelFT();
        void clearWeights();
        void excludeMissingData();

        float getB() const { return b; }
        float getW() const { return w; }
        int getKernelOrder() const { return kernelOrder; }

        void setB(const float& b1);
        void setW(const float& w1);
        void setKernelOrder(const int& a);

        // The reference's per-sample neighbour lists.  The device path works
        // from separable tables instead; exportWeights() fills these two
        // members on request (same content as Reco.cpp:698-787).
        mat2d<float> weights;
        mat2d<int> indices;
        void exportWeights();
        fftwf_complex* kernelFT = nullptr;

    private:
        int sizeX;
        int sizeY;
        int sizeZ;
        Reco* reco;
        float b = 14.9086f;
        float w = 4.25f;
        int kernelOrder = 5;
        void pushKernel();
    };

    ConvKernel* kernel;
    Trajectory* trajectory;
    Fourier* ft;

    fftwf_complex* data;
    fftwf_complex* dataBuffer;

    std::vector<fftwf_complex*> sMaps;
    std::vector<float> denCorr;

    std::vector<int> numPerformedCGIterations;
    std::vector<double> cgResiduals;
    std::vector<fftwf_complex*> img;
    fftwf_complex* cgSenseImg;

    void check();
    void init();
    void finalize();
    bool importParameters(const std::string& path);
    bool createLogFile(const std::string& path);
    bool setDataInputPaths();

    void cgSenseDynamic();

    static inline int indexTransform(const int& i, const int& j, const int& k, const int& ni, const int& nj,
                                     const int& nk) {
        (void)ni;
        return k + nk * j + nk * nj * i;
    }
    static std::vector<int> indexBackTransform(const int& ind, const int& nx, const int& ny, const int& nz);
    static inline void indexBackTransformPt(int* ret, const int& ind, const int& nx, const int& ny, const int& nz) {
        (void)nx;
        ret[0] = ind / (ny * nz);
        ret[1] = (ind - ny * nz * ret[0]) / nz;
        ret[2] = ind - nz * ret[1] - ny * nz * ret[0];
    }

    void setTrajectory(WaveCaipi* t);
    // Extension point of Trajectory.h:11-14 without editing this class: any
    // Trajectory subclass (calculateKMax + calculateTraj are run as for WaveCaipi).
    void setTrajectory(Trajectory* t);
    void setSampling(const int& x, const int& y, const int& z);
    void setSX(const int& x) { sX = x; }
    void setSY(const int& y) { sY = y; }
    void setSZ(const int& z) { sZ = z; }
    void setNCh(const int& n) { nCh = n; }
    void setGridSize(const int& x, const int& y, const int& z);
    void setGridSizeX(const int& x) { gridSizeX = x; }
    void setGridSizeY(const int& y) { gridSizeY = y; }
    void setGridSizeZ(const int& z) { gridSizeZ = z; }
    void setOffcenter(const int& x, const int& y, const int& z);
    void setOffcenterX(const int& x) { offcenterX = x; }
    void setOffcenterY(const int& y) { offcenterY = y; }
    void setOffcenterZ(const int& z) { offcenterZ = z; }
    void setFovY(const double& d) { fovY = d; }
    void setFovZ(const double& d) { fovZ = d; }
    void setCGThreshold(const double& x) { cgThreshold = x; }
    void setBatch(const int& b) { batch = b; }
    void setBatchFiles(const std::vector<std::string>& s) { batchFiles = s; }
    void setDensityCorrection(const Density& d) { density = d; }
    void setDensityCorrectionPath(const std::string& s) { densityCorrPath = s; }
    void setOutputParentDir(const std::string& s) { outputParentDir = s; }
    void setNumThreads(const int& n) { numThreads = n; }

    int getSX() const { return sX; }
    int getSY() const { return sY; }
    int getSZ() const { return sZ; }
    int getNCh() const { return nCh; }
    int getGridSizeX() const { return gridSizeX; }
    int getGridSizeY() const { return gridSizeY; }
    int getGridSizeZ() const { return gridSizeZ; }
    int getOffcenterX() const { return offcenterX; }
    int getOffcenterY() const { return offcenterY; }
    int getOffcenterZ() const { return offcenterZ; }
    int getBatch() const { return batch; }
    int getBatchSize() const { return (int)batchFiles.size(); }
    std::string getBatchFileNum(const int& n) const { return batchFiles[n]; }
    std::vector<std::string> getBatchFiles() { return batchFiles; }
    double getFovY() const { return fovY; }
    double getFovZ() const { return fovZ; }
    double getCGThreshold() const { return cgThreshold; }
    int getTotalGridSize() const { return totalGridSize; }
    int getNumThreads() const { return numThreads; }
    int getNumThreadsFFTW() const { return numThreadsFFTW; }
    Density getDensityCorrection() const { return density; }
    std::string getDensityCorrectionPath() const { return densityCorrPath; }
    std::string getOutputParentDir() const { return outputParentDir; }

    void allocImg();
    void allocData(fftwf_complex** d);
    void allocSMaps();

    void freeImg();

    void nufftFWD();
    void nufftBWD();

    // Extensions (no counterpart in the reference):
    // the reference's nMaxIter (Reco.cpp:228), 15 by default: at most
    // maxIterations + 1 CG passes.
    void setMaxIterations(const int& n) { nMaxIter = n; }
    int getMaxIterations() const { return nMaxIter; }
    // reduceFOV box (Reco.cpp:626-671); init() sets it to the grid.
    void setImgRange(const int& x, const int& y, const int& z) { imgRange = {x, y, z}; }
    // device context behind this object (for the C-ABI of cgs.h)
    cgs_ctx* context() const { return ctx; }

private:
    int sX = 0;
    int sY = 0;
    int sZ = 0;
    int nCh = 0;
    int batch;

    double fovY = 0.0;
    double fovZ = 0.0;

    int gridSizeX = 0;
    int gridSizeY = 0;
    int gridSizeZ = 0;
    double cgThreshold = -1;
    int totalGridSize = 0;

    int offcenterX;
    int offcenterY;
    int offcenterZ;
    std::vector<int> imgRange;
    int numThreads;
    int numThreadsFFTW;

    Density density;

    std::string outputParentDir;
    std::string densityCorrPath;
    std::string dataPath;
    std::vector<std::string> batchFiles;
    std::string dataPathRoot;
    std::string dataPathFileName;
    std::string dataPathFileExtension;

    int nMaxIter = 15;
    cgs_ctx* ctx = nullptr;
    fftwf_complex* imgBlock = nullptr;    // img[ch] point into one [C][Ng] block
    fftwf_complex* mapsBlock = nullptr;   // sMaps[ch] point into one [C][Ng] block

    void ensureContext();
    void pushDeviceSettings();
    fftwf_complex* contiguousImg(bool& owned);
    fftwf_complex* contiguousMaps(bool& owned);
    [[noreturn]] void fatal(const std::string& msg) const;

    friend class Trajectory;
    friend class WaveCaipi;
};
