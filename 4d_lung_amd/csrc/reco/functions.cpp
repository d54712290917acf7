// functions.cpp -- drop-in helpers of CGSense/functions.h over the C-ABI.
#include "functions.h"

#include <cmath>
#include <cstdlib>
#include <iostream>

#include "cgs.h"

double mean(const std::vector<float>& t) {
    double s = 0.0;
    for (float v : t) s += v;
    return t.empty() ? 0.0 : s / (double)t.size();
}

// estimateDensityV3D (functions.h:161-258): cell volumes on the device
// (cgs_voronoi_density: half-space clipping instead of a CGAL triangulation,
// with the reference's rule: 0 only for the unbounded cells of points on the
// convex hull), averaging and normalisation as the reference.
std::vector<float> estimateDensityV3D(const mat2d<double>& traj, const int& sX, const int& sY, const int& sZ,
                                      const bool& average, const double& cutoff) {
    const long long nSamples = (long long)sX * sY * sZ;
    if ((long long)traj.size() != nSamples) {
        std::cout << "\nError in function estimateDensity: trajectory size does not match number of samples!\n";
        std::exit(EXIT_FAILURE);
    }
    std::cout << "Estimating k-space density... \n";
    std::vector<double> pts((size_t)nSamples * 3);
    for (long long i = 0; i < nSamples; i++)
        for (int a = 0; a < 3; a++) pts[3 * i + a] = traj[i][a];
    std::vector<float> d((size_t)nSamples);
    const int st = cgs_voronoi_density(pts.data(), nSamples, average ? 1 : 0, sX, cutoff, d.data());
    if (st != CGS_OK) {
        std::cout << "\nError in function estimateDensity: device Voronoi volumes failed (status " << st << ")\n";
        std::exit(EXIT_FAILURE);
    }
    std::cout << "Done estimating k-space density.\n";
    return d;
}

// performDensityCorrection (functions.h:319-335).
void performDensityCorrection(const Reco& r, fftwf_complex* data, const std::vector<float>& d,
                              const int& nSamplesPerCh, const int& nChannels, const int& nThreads) {
    (void)r;
    (void)nSamplesPerCh;
    (void)nThreads;
    if (d.empty()) return;
    const size_t dSize = d.size();
    for (int ch = 0; ch < nChannels; ch++)
        for (size_t i = 0; i < dSize; i++) {
            data[ch * dSize + i][0] *= d[i];
            data[ch * dSize + i][1] *= d[i];
        }
}

// rssq2DVecImg (functions.h:290-316).
std::vector<float> rssq2DVecImg(const std::vector<fftwf_complex*>& img, const int& imgSize) {
    std::vector<float> out((size_t)imgSize, 0.0f);
    for (size_t ch = 0; ch < img.size(); ch++)
        for (int i = 0; i < imgSize; i++) out[i] += img[ch][i][0] * img[ch][i][0] + img[ch][i][1] * img[ch][i][1];
    for (auto& v : out) v = std::sqrt(v);
    return out;
}
