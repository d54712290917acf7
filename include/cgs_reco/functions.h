// functions.h -- the reference's helper functions of the reconstruction
// pipeline (CGSense/functions.h) that sit on the hot path, backed by the
// device library: Voronoi density estimation (estimateDensityV3D,
// functions.h:161-258), density correction of the data
// (performDensityCorrection, functions.h:319-335) and the root-sum-of-squares
// preview image (rssq2DVecImg, functions.h:290-316).  Same names, arguments
// and error behaviour (message on stdout, then exit(EXIT_FAILURE)).
#pragma once

#include <vector>

#include "Reco.h"
#include "cgs_complex.h"

// Arithmetic mean (functions.h:3-12).
double mean(const std::vector<float>& t);

// Voronoi cell volume of every trajectory point (0 for unbounded cells),
// optionally averaged over the readout index i % sX up to the cut-off, then
// divided by the maximum.  traj.size() must equal sX*sY*sZ.
std::vector<float> estimateDensityV3D(const mat2d<double>& traj, const int& sX, const int& sY, const int& sZ,
                                      const bool& average, const double& cutoff);

// data[ch*nS + i] *= d[i] for every channel; no-op for an empty d.
void performDensityCorrection(const Reco& r, fftwf_complex* data, const std::vector<float>& d,
                              const int& nSamplesPerCh, const int& nChannels, const int& nThreads = 1);

// sqrt(sum_ch |img[ch][i]|^2) over imgSize voxels.
std::vector<float> rssq2DVecImg(const std::vector<fftwf_complex*>& img, const int& imgSize);
