// Fourier.cpp -- drop-in Fourier (Fourier.cpp:4-77 of the reference) on the
// device FFT of cgs.h.
#include "Fourier.h"

#include <cstdlib>
#include <cstring>
#include <iostream>

#include "cgs.h"

Fourier::~Fourier() {
    if (plan) cgs_fft_destroy(plan);
}

void Fourier::initFT() {
    if (plan) {
        cgs_fft_destroy(plan);
        plan = nullptr;
    }
    const int dims[3] = {sizeX, sizeY, sizeZ};
    if (cgs_fft_create(dims, &plan) != CGS_OK) {
        std::cout << "\nError: FFT size " << sizeX << "x" << sizeY << "x" << sizeZ << " not supported\n";
        std::exit(EXIT_FAILURE);
    }
}

void Fourier::ensurePlan() const {
    if (!plan) const_cast<Fourier*>(this)->initFT();
}

void Fourier::transformFWD(fftwf_complex* a) const {
    ensurePlan();
    if (cgs_fft_forward(plan, &a[0][0]) != CGS_OK) {
        std::cout << "\nError: transformFWD failed\n";
        std::exit(EXIT_FAILURE);
    }
}

void Fourier::transformBWD(fftwf_complex* a, bool rescale) const {
    ensurePlan();
    if (cgs_fft_backward(plan, &a[0][0], rescale ? 1 : 0) != CGS_OK) {
        std::cout << "\nError: transformBWD failed\n";
        std::exit(EXIT_FAILURE);
    }
}

void Fourier::transformFWD(fftwf_complex* a, fftwf_complex* b) const {
    std::memcpy(b, a, sizeof(fftwf_complex) * (size_t)sizeX * sizeY * sizeZ);
    transformFWD(b);
}

void Fourier::transformBWD(fftwf_complex* a, fftwf_complex* b) const {
    std::memcpy(b, a, sizeof(fftwf_complex) * (size_t)sizeX * sizeY * sizeZ);
    transformBWD(b, true);
}
