// Fourier.h -- centred 3D FFT of one grid on the MI355X (drop-in for the
// reference's Fourier class, Fourier.h:12-54).  transformFWD / transformBWD
// keep the reference's convention: diag((-1)^(x+y+z)) DFT diag((-1)^(x+y+z)),
// forward unnormalized, backward scaled by 1/sqrt(Nx*Ny*Nz) when rescale.
#pragma once
#include <string>
#include <vector>
#include "cgs_complex.h"

struct cgs_fft;

template <class T>
using mat3d = std::vector<std::vector<std::vector<T>>>;
template <class T>
using mat2d = std::vector<std::vector<T>>;

class Fourier {
public:
    Fourier(const int& x, const int& y, const int& z) : sizeX(x), sizeY(y), sizeZ(z) {}
    ~Fourier();
    // Creates the device plan (replaces the FFTW_MEASURE planning, Fourier.cpp:4-17).
    void initFT();

    // in place FFT (Fourier.cpp:19-42)
    void transformFWD(fftwf_complex* a) const;
    void transformBWD(fftwf_complex* a, bool rescale = true) const;

    // out of place FFT (declared by the reference, Fourier.h:33-34)
    void transformFWD(fftwf_complex* a, fftwf_complex* b) const;
    void transformBWD(fftwf_complex* a, fftwf_complex* b) const;

    void setSizeX(const int& x) { sizeX = x; }
    void setSizeY(const int& y) { sizeY = y; }
    void setSizeZ(const int& z) { sizeZ = z; }
    // Kept for interface compatibility; device plans need no wisdom file.
    void setWisdomFileName(const std::string& s) { wisdomFileName = s; }

private:
    int sizeX;
    int sizeY;
    int sizeZ;
    cgs_fft* plan = nullptr;
    std::string wisdomFileName;
    void ensurePlan() const;
};
