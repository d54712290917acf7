// Reco.cpp -- drop-in Reco on top of the C-ABI of cgs.h (host C++ only).
//
// The host-side state is the reference's (Reco.h:77-91); the hot path --
// ConvKernel tables, nufftFWD / nufftBWD and the CG loop -- runs on the
// device context created in init().  Messages of failures are the
// reference's, followed by exit(EXIT_FAILURE) as in Reco.cpp:194-195,
// 266-267, 333-334, 475-476.
#include "Reco.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>

#include "cgs.h"

fftwf_complex* cgs_alloc_complex(size_t n) {
    void* p = nullptr;
    if (posix_memalign(&p, 64, sizeof(fftwf_complex) * (n ? n : 1)) != 0) return nullptr;
    return (fftwf_complex*)p;
}

void cgs_free_complex(void* p) { std::free(p); }

Reco::Reco()
    : kernel(nullptr),
      trajectory(nullptr),
      ft(nullptr),
      data(nullptr),
      dataBuffer(nullptr),
      cgSenseImg(nullptr),
      batch(0),
      offcenterX(0),
      offcenterY(0),
      offcenterZ(0),
      numThreads(32),
      numThreadsFFTW(1),
      density(Density::NONE),
      dataPathFileExtension(".bin") {}

Reco::~Reco() {
    delete kernel;
    delete ft;
    if (data) cgs_free_complex(data);
    if (dataBuffer) cgs_free_complex(dataBuffer);
    if (mapsBlock) cgs_free_complex(mapsBlock);
    else
        for (auto p : sMaps)
            if (p) cgs_free_complex(p);
    if (imgBlock) cgs_free_complex(imgBlock);
    if (cgSenseImg) cgs_free_complex(cgSenseImg);
    if (ctx) cgs_destroy(ctx);
}

void Reco::fatal(const std::string& msg) const {
    std::cout << msg;
    if (ctx) std::cout << " (" << cgs_last_error(ctx) << ")";
    std::cout << "\n";
    std::exit(EXIT_FAILURE);
}

bool Reco::setDataInputPaths() {
    for (int i = 0; i < batch; i++) batchFiles.push_back(dataPathRoot + dataPathFileName + std::to_string(i) + dataPathFileExtension);
    return true;
}

bool Reco::createLogFile(const std::string& path) {
    (void)path;
    return true;
}

bool Reco::importParameters(const std::string& path) {
    (void)path;
    return true;
}

void Reco::check() {
    if (!batchFiles.empty()) {
        std::cout << "Checking all files...";
        for (const auto& p : batchFiles) {
            std::FILE* f = std::fopen(p.c_str(), "rb");
            if (!f) {
                std::cout << "\nError: Could not open binary data file " << p << "\n";
                std::exit(EXIT_FAILURE);
            }
            std::fclose(f);
        }
        std::cout << " Done!\n";
    }
}

void Reco::ensureContext() {
    if (ctx) return;
    const int grid[3] = {gridSizeX, gridSizeY, gridSizeZ};
    const int samp[3] = {sX, sY, sZ};
    if (cgs_create(grid, samp, nCh, &ctx) != CGS_OK) {
        ctx = nullptr;
        std::cout << "\nError: could not create the device context for grid " << gridSizeX << "x" << gridSizeY << "x"
                  << gridSizeZ << ", " << nCh << " channels\n";
        std::exit(EXIT_FAILURE);
    }
}

void Reco::pushDeviceSettings() {
    ensureContext();
    const int rng[3] = {imgRange[0], imgRange[1], imgRange[2]};
    const int off[3] = {offcenterX, offcenterY, offcenterZ};
    if (cgs_set_img_range(ctx, rng, off) != CGS_OK) fatal("\nError: invalid image range");
}

// Reco::init (Reco.cpp:374-400)
void Reco::init() {
    kernel = new ConvKernel(this);
    kernel->setB(14.9086f);
    kernel->setW(4.25f);
    kernel->setKernelOrder(5);
    ft = new Fourier(gridSizeX, gridSizeY, gridSizeZ);
    sMaps.assign(nCh, nullptr);
    totalGridSize = gridSizeX * gridSizeY * gridSizeZ;
    if (!createLogFile(outputParentDir + "log.txt"))
        std::cout << "\nWarning: Failed to create log-File. Could not open path " << outputParentDir + "log.txt" << "\n";
    imgRange = {gridSizeX, gridSizeY, gridSizeZ};
    batch = (int)batchFiles.size();
    ensureContext();
}

// Reco::finalize (Reco.cpp:402-412); paths use '/'.
void Reco::finalize() {
    std::ofstream os(outputParentDir + "results/performedIterations.txt");
    for (int n : numPerformedCGIterations) os << n << "\n";
}

void Reco::setSampling(const int& x, const int& y, const int& z) {
    sX = x;
    sY = y;
    sZ = z;
}

void Reco::setGridSize(const int& x, const int& y, const int& z) {
    gridSizeX = x;
    gridSizeY = y;
    gridSizeZ = z;
}

void Reco::setOffcenter(const int& x, const int& y, const int& z) {
    offcenterX = x;
    offcenterY = y;
    offcenterZ = z;
}

std::vector<int> Reco::indexBackTransform(const int& ind, const int& nx, const int& ny, const int& nz) {
    int r[3];
    indexBackTransformPt(r, ind, nx, ny, nz);
    return {r[0], r[1], r[2]};
}

// Reco::setTrajectory (Reco.cpp:324-336)
void Reco::setTrajectory(WaveCaipi* t) {
    t->calculateKMax();
    t->calculateTraj();
    trajectory = dynamic_cast<Trajectory*>(t);
    if (trajectory == nullptr) {
        std::cout << "\nError: Trajectory conversion failed in setTrajectory!\n";
        std::exit(EXIT_FAILURE);
    }
}

void Reco::setTrajectory(Trajectory* t) {
    t->calculateKMax();
    t->calculateTraj();
    trajectory = t;
}

// Per-coil buffers live in one [C][Ng] block so they cross the C-ABI in one piece.
void Reco::allocImg() {
    if (imgBlock) cgs_free_complex(imgBlock);
    imgBlock = cgs_alloc_complex((size_t)nCh * totalGridSize);
    img.assign(nCh, nullptr);
    for (int c = 0; c < nCh; c++) img[c] = imgBlock + (size_t)c * totalGridSize;
}

void Reco::allocData(fftwf_complex** d) { *d = cgs_alloc_complex((size_t)sX * sY * sZ * nCh); }

void Reco::allocSMaps() {
    if (mapsBlock) cgs_free_complex(mapsBlock);
    mapsBlock = cgs_alloc_complex((size_t)nCh * gridSizeX * gridSizeY * gridSizeZ);
    sMaps.assign(nCh, nullptr);
    for (int c = 0; c < nCh; c++) sMaps[c] = mapsBlock + (size_t)c * totalGridSize;
}

void Reco::freeImg() {
    if (imgBlock) cgs_free_complex(imgBlock);
    imgBlock = nullptr;
    img.clear();
    img.shrink_to_fit();
}

// The device calls need [C][Ng] contiguous; when the caller replaced the
// per-coil pointers, a packed copy is made.
fftwf_complex* Reco::contiguousImg(bool& owned) {
    owned = false;
    if ((int)img.size() != nCh) fatal("\nError: coil images not allocated!");
    bool packed = imgBlock != nullptr;
    for (int c = 0; c < nCh && packed; c++) packed = img[c] == imgBlock + (size_t)c * totalGridSize;
    if (packed) return imgBlock;
    fftwf_complex* b = cgs_alloc_complex((size_t)nCh * totalGridSize);
    for (int c = 0; c < nCh; c++) std::memcpy(b + (size_t)c * totalGridSize, img[c], sizeof(fftwf_complex) * totalGridSize);
    owned = true;
    return b;
}

fftwf_complex* Reco::contiguousMaps(bool& owned) {
    owned = false;
    if ((int)sMaps.size() != nCh) fatal("\nError: sensitivity maps not allocated!");
    bool packed = mapsBlock != nullptr;
    for (int c = 0; c < nCh && packed; c++) packed = sMaps[c] == mapsBlock + (size_t)c * totalGridSize;
    if (packed) return mapsBlock;
    fftwf_complex* b = cgs_alloc_complex((size_t)nCh * totalGridSize);
    for (int c = 0; c < nCh; c++) {
        if (!sMaps[c]) fatal("\nError: sensitivity map missing!");
        std::memcpy(b + (size_t)c * totalGridSize, sMaps[c], sizeof(fftwf_complex) * totalGridSize);
    }
    owned = true;
    return b;
}

// Reco::nufftBWD (Reco.cpp:468-519): data -> coil images img[ch].
void Reco::nufftBWD() {
    if (data == nullptr) {
        std::cout << "\nError in nufftBWD function: data is NULL pointer!\n";
        std::exit(EXIT_FAILURE);
    }
    if ((int)img.size() != nCh) allocImg();
    pushDeviceSettings();
    bool owned = false;
    fftwf_complex* out = contiguousImg(owned);
    if (cgs_nufft_bwd(ctx, &data[0][0], &out[0][0]) != CGS_OK) fatal("\nError in nufftBWD function");
    if (owned) {
        for (int c = 0; c < nCh; c++) std::memcpy(img[c], out + (size_t)c * totalGridSize, sizeof(fftwf_complex) * totalGridSize);
        cgs_free_complex(out);
    }
}

// Reco::nufftFWD (Reco.cpp:521-553): coil images img[ch] -> data.
void Reco::nufftFWD() {
    if (data == nullptr) allocData(&data);
    bool owned = false;
    fftwf_complex* in = contiguousImg(owned);
    if (cgs_nufft_fwd(ctx, &in[0][0], &data[0][0]) != CGS_OK) fatal("\nError in nufftFWD function");
    if (owned) cgs_free_complex(in);
}

// Reco::cgSenseDynamic (Reco.cpp:175-321), started from the coil images in img.
void Reco::cgSenseDynamic() {
    if (cgThreshold == -1) {
        std::cout << "\nError: no threshold set for CG SENSE!\n";
        std::exit(EXIT_FAILURE);
    }
    std::cout << "Begin cgSense\n";
    pushDeviceSettings();
    cgResiduals.clear();
    cgResiduals.shrink_to_fit();
    if (cgs_set_density(ctx, denCorr.empty() ? nullptr : denCorr.data()) != CGS_OK) fatal("\nError: density weights");
    bool mo = false, io = false;
    fftwf_complex* maps = contiguousMaps(mo);
    if (cgs_set_maps(ctx, &maps[0][0]) != CGS_OK) fatal("\nError: sensitivity maps");
    if (mo) cgs_free_complex(maps);
    fftwf_complex* im = contiguousImg(io);
    if (cgSenseImg) cgs_free_complex(cgSenseImg);
    cgSenseImg = cgs_alloc_complex(totalGridSize);
    std::vector<double> res(nMaxIter + 1, 0.0);
    int iters = 0;
    const int st = cgs_cg_sense_images(ctx, &im[0][0], cgThreshold, nMaxIter, &cgSenseImg[0][0], res.data(), &iters);
    if (io) cgs_free_complex(im);
    if (st == CGS_ERR_NONFINITE) {
        std::cout << "\nError in function cgSense: NaN or INF value in image detected!\n";
        std::exit(EXIT_FAILURE);
    }
    if (st != CGS_OK) fatal("\nError in function cgSense");
    for (int i = 0; i < iters; i++) {
        cgResiduals.push_back(res[i]);
        std::cout << "\n\tIteration:\t" << i << "\t" << res[i] << "\t" << cgThreshold << "\n";
    }
    numPerformedCGIterations.push_back(iters);
    std::cout << "\nEnd cgSense\n";
}

// ------------------------------------------------------------ ConvKernel
Reco::ConvKernel::~ConvKernel() {
    if (kernelFT) cgs_free_complex(kernelFT);
}

void Reco::ConvKernel::pushKernel() {
    if (!reco->ctx) return;
    if (cgs_set_kernel(reco->ctx, w, b, kernelOrder) != CGS_OK) reco->fatal("\nError: unsupported kernel parameters");
}

void Reco::ConvKernel::setB(const float& b1) {
    b = b1;
    pushKernel();
}

void Reco::ConvKernel::setW(const float& w1) {
    w = w1;
    pushKernel();
}

void Reco::ConvKernel::setKernelOrder(const int& a) {
    kernelOrder = a;
    pushKernel();
}

static std::vector<double> flatten_traj(const mat2d<double>& t) {
    std::vector<double> f(t.size() * 3);
    for (size_t i = 0; i < t.size(); i++) {
        f[3 * i] = t[i][0];
        f[3 * i + 1] = t[i][1];
        f[3 * i + 2] = t[i][2];
    }
    return f;
}

// ConvKernel::calculateWeights (Reco.cpp:698-787)
void Reco::ConvKernel::calculateWeights() {
    reco->ensureContext();
    pushKernel();
    if (!reco->trajectory) reco->fatal("\nError: no trajectory set!");
    const size_t ns = (size_t)reco->sX * reco->sY * reco->sZ;
    if (reco->trajectory->kSpaceTraj.size() != ns) reco->fatal("\nError: trajectory size does not match number of samples!");
    const std::vector<double> f = flatten_traj(reco->trajectory->kSpaceTraj);
    if (cgs_set_trajectory(reco->ctx, f.data()) != CGS_OK || cgs_calculate_weights(reco->ctx) != CGS_OK)
        reco->fatal("\nError in calculateWeights");
    weights.clear();
    indices.clear();
}

// ConvKernel::excludeMissingData (Reco.cpp:681-696), driven by reco->data.
void Reco::ConvKernel::excludeMissingData() {
    if (!reco->data) reco->fatal("\nError in excludeMissingData: data is NULL pointer!");
    if (cgs_exclude_missing_data(reco->ctx, &reco->data[0][0]) != CGS_OK) reco->fatal("\nError in excludeMissingData");
    weights.clear();
    indices.clear();
}

void Reco::ConvKernel::clearWeights() {
    weights.clear();
    weights.shrink_to_fit();
    indices.clear();
    indices.shrink_to_fit();
}

// ConvKernel::calculateKernelFT (Reco.cpp:560-601)
void Reco::ConvKernel::calculateKernelFT() {
    reco->ensureContext();
    pushKernel();
    const size_t ng = (size_t)sizeX * sizeY * sizeZ;
    std::vector<float> re(ng);
    if (cgs_calculate_kernel_ft(reco->ctx) != CGS_OK || cgs_get_kernel_ft(reco->ctx, re.data()) != CGS_OK)
        reco->fatal("\nError in calculateKernelFT");
    if (kernelFT) cgs_free_complex(kernelFT);
    kernelFT = cgs_alloc_complex(ng);
    for (size_t i = 0; i < ng; i++) {
        kernelFT[i][0] = re[i];
        kernelFT[i][1] = 0.0f;
    }
}

// The reference's neighbour lists from the device's separable tables: tap
// (x, y, z) of a sample carries float((vx*vy)*vz) and is kept above the
// threshold, in x-major candidate order.
void Reco::ConvKernel::exportWeights() {
    long long m = 0;
    int kw = 0, nt = 0;
    double thr = 0.0;
    if (cgs_bin_count(reco->ctx, 0, &m) != CGS_OK || cgs_kernel_info(reco->ctx, &kw, &nt, &thr) != CGS_OK)
        reco->fatal("\nError in exportWeights");
    std::vector<int32_t> perm(m), base(3 * m);
    std::vector<double> vals(3 * kw * m);
    if (cgs_bin_tables(reco->ctx, 0, nullptr, perm.data(), nullptr, base.data(), vals.data()) != CGS_OK)
        reco->fatal("\nError in exportWeights");
    const size_t ns = (size_t)reco->sX * reco->sY * reco->sZ;
    weights.assign(ns, {});
    indices.assign(ns, {});
    for (long long s = 0; s < m; s++) {
        const int i = perm[s];
        const double* v = &vals[3 * kw * s];
        for (int ox = 0; ox < kw; ox++)
            for (int oy = 0; oy < kw; oy++)
                for (int oz = 0; oz < kw; oz++) {
                    const float wv = (float)((v[ox] * v[kw + oy]) * v[2 * kw + oz]);
                    if ((double)wv > thr) {
                        const int x = base[3 * s] + ox, y = base[3 * s + 1] + oy, z = base[3 * s + 2] + oz;
                        indices[i].push_back(z + sizeZ * y + sizeY * sizeZ * x);
                        weights[i].push_back(wv);
                    }
                }
    }
}
