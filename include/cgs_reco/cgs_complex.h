// cgs_complex.h -- the complex element type of the reference's Reco/Fourier/IO
// interface (fftwf_complex, an interleaved float[2]) and host allocation of
// arrays of it.  When FFTW's own header is already included its definition
// is used; the two are layout-identical.
#pragma once
#include <cstddef>

#ifndef FFTW3_H
typedef float fftwf_complex[2];
#endif

// Host arrays of n complex values, 64-byte aligned (Reco.cpp:436-456 uses
// fftwf_alloc_complex / fftwf_free for the same buffers).
fftwf_complex* cgs_alloc_complex(size_t n);
void cgs_free_complex(void* p);
