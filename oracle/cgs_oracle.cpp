// cgs_oracle.cpp -- CPU restatement of the expRad/4d_lung CGSense hot path.
//
// TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
// never as the thing that is measured or shipped.  The product path
// (4d_lung_amd/) never links, loads or calls it.
//
// Every function follows the reference file:line it cites.  Two precisions:
//   prec 32 : the reference's own arithmetic (fftwf_complex FP32 vectors, FP64
//             CG scalars, float/double promotions exactly as written there);
//   prec 64 : the same operator evaluated in FP64 throughout (the accuracy
//             oracle).  The operator's *tables* (float KB tap weights, float
//             deapodization map) are shared by both modes, so prec 64 is the
//             exact-arithmetic version of the reference's operator.
//
// Third-party arithmetic the reference takes from libraries that are absent
// here is restated from published methods:
//   * Boost.Math 1.66 cyl_bessel_i(0, x)  -> power series, Abramowitz & Stegun
//     9.6.12 (I0(x) = sum_k (x^2/4)^k / (k!)^2), pinned against scipy.special.i0
//     and std::cyl_bessel_i in tests;
//   * FFTW3f plans (Fourier.cpp:12-13, 22, 29) -> own mixed-radix Stockham FFT,
//     pinned against numpy.fft and a brute-force DFT in tests.
// Build: oracle/Makefile (g++ -O2 -fopenmp -ffp-contract=off).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>
#include <omp.h>

namespace {

int g_verbose = 0;   // print one line per CG iteration to stderr

// ---------------------------------------------------------------------------
// Modified Bessel function I0 -- power series (Abramowitz & Stegun 9.6.12).
// Stands in for boost::math::cyl_bessel_i(0, x) used at Reco.h:238-240.
// The loop and its stopping rule are written so that the GPU setup kernel can
// evaluate the identical sequence of IEEE operations.
// ---------------------------------------------------------------------------
double i0_series(double x) {
    const double q = (0.25 * x) * x;
    double term = 1.0;
    double sum = 1.0;
    for (int k = 1; k < 500; ++k) {
        const double kk = (double)k * (double)k;
        term = (term * q) / kk;
        sum = sum + term;
        if (term < sum * 1.0e-17) break;
    }
    return sum;
}

// Per-axis factor of Reco::ConvKernel::convFunction (Reco.h:218-242).
//   dkf  = float(1.0 / float(N-1))         (Reco.h:221-223)
//   dxf  = float(distance)                  (Reco.h:228-230)
//   cut  : dxf > float(w/2) * dkf  -> 0     (Reco.h:225, 232)
//   val  = I0(double(b) * sqrt(1 - pow(2*dxf / float(w*dkf), 2)))
//          / double(float(100*w*dkf))       (Reco.h:238-240)
double kb_axis_value(float dxf, int N, float w, float b) {
    const float dkf = (float)(1.0 / (double)(float)(N - 1));
    const float wHalf = (float)((double)w / 2.0);
    if (dxf > wHalf * dkf) return 0.0;
    const float wdk = w * dkf;
    const float q = (2.0f * dxf) / wdk;
    const double q2 = (double)q * (double)q;   // pow(float, 2) promoted: exact
    const double arg = (double)b * std::sqrt(1.0 - q2);
    const float denom = (100.0f * w) * dkf;
    return i0_series(arg) / (double)denom;
}

// Reco::ConvKernel::convFunction (Reco.h:218-242) for a double distance triple.
float conv_function(const double* dist, const int* N, float w, float b) {
    const double vx = kb_axis_value((float)dist[0], N[0], w, b);
    const double vy = kb_axis_value((float)dist[1], N[1], w, b);
    const double vz = kb_axis_value((float)dist[2], N[2], w, b);
    if (vx == 0.0 || vy == 0.0 || vz == 0.0) return 0.0f;  // cut branch returns 0
    return (float)((vx * vy) * vz);
}

// ---------------------------------------------------------------------------
// Mixed-radix Stockham FFT (stands in for FFTW3f; Fourier.cpp:4-77).
// sign = -1 : exp(-2 pi i nk/N) (FFTW_FORWARD), +1 : exp(+...) (FFTW_BACKWARD).
// ---------------------------------------------------------------------------
template <typename T>
struct Cplx { T re, im; };

template <typename T>
struct Plan1D {
    int n = 0;
    std::vector<int> rad;
    std::vector<double> twr, twi;   // exp(-2 pi i k / n), k < n, computed in FP64
    std::vector<T> tr, ti;          // same, rounded to T
};

std::vector<int> factorize(int n) {
    std::vector<int> f;
    while (n % 4 == 0) { f.push_back(4); n /= 4; }
    while (n % 2 == 0) { f.push_back(2); n /= 2; }
    for (int p = 3; n > 1; p += 2) {
        while (n % p == 0) { f.push_back(p); n /= p; }
    }
    return f;
}

template <typename T>
Plan1D<T> make_plan(int n) {
    Plan1D<T> p;
    p.n = n;
    p.rad = factorize(n);
    p.twr.resize(n); p.twi.resize(n); p.tr.resize(n); p.ti.resize(n);
    const double two_pi = 6.283185307179586476925286766559;
    for (int k = 0; k < n; ++k) {
        const double a = -two_pi * (double)k / (double)n;
        p.twr[k] = std::cos(a); p.twi[k] = std::sin(a);
        p.tr[k] = (T)p.twr[k]; p.ti[k] = (T)p.twi[k];
    }
    return p;
}

// In-register DFT of length R (sign as in fft1); radix 2 and 4 written out,
// other radices by the definition with the plan's root table.
template <typename T>
inline void small_dft(const Plan1D<T>& P, int R, int sign, Cplx<T>* v, Cplx<T>* y) {
    if (R == 2) {
        y[0] = {v[0].re + v[1].re, v[0].im + v[1].im};
        y[1] = {v[0].re - v[1].re, v[0].im - v[1].im};
        return;
    }
    if (R == 4) {
        const Cplx<T> a0 = {v[0].re + v[2].re, v[0].im + v[2].im};
        const Cplx<T> a1 = {v[0].re - v[2].re, v[0].im - v[2].im};
        const Cplx<T> a2 = {v[1].re + v[3].re, v[1].im + v[3].im};
        const Cplx<T> d = {v[1].re - v[3].re, v[1].im - v[3].im};
        // d * (-i) for the forward sign, d * (+i) for the backward sign
        const Cplx<T> a3 = sign < 0 ? Cplx<T>{d.im, -d.re} : Cplx<T>{-d.im, d.re};
        y[0] = {a0.re + a2.re, a0.im + a2.im};
        y[2] = {a0.re - a2.re, a0.im - a2.im};
        y[1] = {a1.re + a3.re, a1.im + a3.im};
        y[3] = {a1.re - a3.re, a1.im - a3.im};
        return;
    }
    const int rstride = P.n / R;   // exp(-2 pi i u/R) = tw[u * n/R]
    for (int u = 0; u < R; ++u) {
        T sr = 0, si = 0;
        int t = 0;
        for (int i = 0; i < R; ++i) {
            const T wr = P.tr[t * rstride];
            const T wi = sign < 0 ? P.ti[t * rstride] : -P.ti[t * rstride];
            sr += v[i].re * wr - v[i].im * wi;
            si += v[i].re * wi + v[i].im * wr;
            t += u;
            if (t >= R) t -= R;
        }
        y[u] = {sr, si};
    }
}

template <typename T>
void fft1(const Plan1D<T>& P, Cplx<T>* x, Cplx<T>* work, int sign) {
    const int n = P.n;
    if (n == 1) return;
    Cplx<T>* in = x;
    Cplx<T>* out = work;
    int Ls = 1;
    Cplx<T> v[64], y[64];
    for (int R : P.rad) {
        const int m = n / R;
        const int tstride = n / (Ls * R);
        for (int j = 0; j < m; ++j) {
            const int k = j % Ls;
            for (int i = 0; i < R; ++i) {
                Cplx<T> a = in[j + i * m];
                if (i > 0 && k > 0) {
                    const int t = k * i * tstride;
                    const T wr = P.tr[t];
                    const T wi = sign < 0 ? P.ti[t] : -P.ti[t];
                    const T r = a.re * wr - a.im * wi;
                    const T im = a.re * wi + a.im * wr;
                    a.re = r; a.im = im;
                }
                v[i] = a;
            }
            small_dft(P, R, sign, v, y);
            const int base = (j / Ls) * Ls * R + k;
            for (int u = 0; u < R; ++u) out[base + u * Ls] = y[u];
        }
        std::swap(in, out);
        Ls *= R;
    }
    if (in != x) std::memcpy(x, in, sizeof(Cplx<T>) * n);
}

// Fourier::signAlternate3D (Fourier.cpp:45-72): negate a[i][j][k] where
// (i + j + k) is even.  Loop shape follows the reference (k += 2).
template <typename T>
void sign_alternate3d(Cplx<T>* a, const int* N) {
    for (int i = 0; i < N[0]; i++) {
        const int im = i % 2;
        for (int j = 0; j < N[1]; j++) {
            const int jm = j % 2;
            for (int k = 0; k < N[2]; k += 2) {
                long ind;
                if (jm == im) ind = (long)k + (long)N[2] * j + (long)N[2] * N[1] * i;
                else ind = (long)(k + 1) + (long)N[2] * j + (long)N[2] * N[1] * i;
                a[ind].re *= -1; a[ind].im *= -1;
            }
        }
    }
}

// Unnormalized 3D DFT along all three axes of an [x][y][z] (z fastest) array.
// Strided axes are transformed in blocks of ZB neighbouring lines so every
// gather reads contiguous runs.
template <typename T>
void fft3_raw(Cplx<T>* a, const int* N, int sign) {
    const int nx = N[0], ny = N[1], nz = N[2];
    Plan1D<T> px = make_plan<T>(nx), py = make_plan<T>(ny), pz = make_plan<T>(nz);
    const int maxn = std::max(nx, std::max(ny, nz));
    constexpr int ZB = 16;
    std::vector<Cplx<T>> blk((size_t)ZB * maxn), work(maxn);
    // z lines (contiguous)
    for (long xy = 0; xy < (long)nx * ny; ++xy) fft1(pz, a + xy * nz, work.data(), sign);
    // y lines
    for (int x = 0; x < nx; ++x)
        for (int z0 = 0; z0 < nz; z0 += ZB) {
            const int zb = std::min(ZB, nz - z0);
            Cplx<T>* base = a + (long)x * ny * nz + z0;
            for (int y = 0; y < ny; ++y)
                for (int j = 0; j < zb; ++j) blk[(size_t)j * ny + y] = base[(long)y * nz + j];
            for (int j = 0; j < zb; ++j) fft1(py, blk.data() + (size_t)j * ny, work.data(), sign);
            for (int y = 0; y < ny; ++y)
                for (int j = 0; j < zb; ++j) base[(long)y * nz + j] = blk[(size_t)j * ny + y];
        }
    // x lines
    const long st = (long)ny * nz;
    for (long yz0 = 0; yz0 < st; yz0 += ZB) {
        const int zb = (int)std::min<long>(ZB, st - yz0);
        Cplx<T>* base = a + yz0;
        for (int x = 0; x < nx; ++x)
            for (int j = 0; j < zb; ++j) blk[(size_t)j * nx + x] = base[x * st + j];
        for (int j = 0; j < zb; ++j) fft1(px, blk.data() + (size_t)j * nx, work.data(), sign);
        for (int x = 0; x < nx; ++x)
            for (int j = 0; j < zb; ++j) base[x * st + j] = blk[(size_t)j * nx + x];
    }
}

// Fourier::transformFWD (Fourier.cpp:19-24): sign-alternate, forward DFT, sign-alternate.
template <typename T>
void transform_fwd(Cplx<T>* a, const int* N) {
    sign_alternate3d(a, N);
    fft3_raw(a, N, -1);
    sign_alternate3d(a, N);
}

// Fourier::transformBWD (Fourier.cpp:26-42): sign-alternate, backward DFT,
// sign-alternate, optional scaling by 1/sqrt(Nx*Ny*Nz) (computed in double,
// applied as a[i] *= double).
template <typename T>
void transform_bwd(Cplx<T>* a, const int* N, bool rescale) {
    sign_alternate3d(a, N);
    fft3_raw(a, N, +1);
    sign_alternate3d(a, N);
    if (rescale) {
        const double f = 1.0 / std::sqrt((double)(N[0] * N[1] * N[2]));
        const long n = (long)N[0] * N[1] * N[2];
        for (long i = 0; i < n; i++) {
            a[i].re = (T)((double)a[i].re * f);
            a[i].im = (T)((double)a[i].im * f);
        }
    }
}

// ---------------------------------------------------------------------------
// The oracle's Reco state (mirrors the members of Reco.h:77-91 and
// ConvKernel.h:58-69 that the hot path reads).
// ---------------------------------------------------------------------------
struct Oracle {
    int N[3];
    long Ng;
    int S[3];
    long nS;
    int C;
    float w, b;
    int order;
    int nthreads;
    int imgRange[3];
    int offc[3];
    std::vector<double> traj;                      // nS x 3, memory order
    std::vector<std::vector<int>> idx;             // ConvKernel::indices (Reco.h:59)
    std::vector<std::vector<float>> wts;           // ConvKernel::weights (Reco.h:58)
    // Smooth-kernel mode (finite-difference checks of coordinate gradients
    // only): weights evaluated without the float roundings of the distance
    // and of q, kept in FP64 and used by the prec-64 operators.
    bool exact_w = false;
    std::vector<std::vector<double>> wts64;
    std::vector<float> kft;                        // ConvKernel::kernelFT real part (Reco.h:60)
    std::vector<float> dcf;                        // Reco::denCorr (Reco.h:86)
    std::vector<Cplx<float>> maps;                 // Reco::sMaps (Reco.h:84), C x Ng
    // Coil sum in the reference's own loop shape (float omp atomic adds of
    // each coil into q, in whatever order the threads reach them) instead of
    // the fixed coil order; used to time the reference's CPU loop.
    bool atomic_sum = false;
};

// Reco::ConvKernel::collectNeighborIndices (Reco.cpp:789-806)
void collect_neighbor_indices(const Oracle& o, const int* start, std::vector<int>& pts,
                              const std::vector<int>& kp) {
    const int ncand = (int)kp.size() / 3;
    for (int i = 0; i < ncand; i++) {
        const int a = start[0] + kp[3 * i], bb = start[1] + kp[3 * i + 1], c = start[2] + kp[3 * i + 2];
        if (a < 0 || a > o.N[0] - 1 || bb < 0 || bb > o.N[1] - 1 || c < 0 || c > o.N[2] - 1) continue;
        pts.push_back(c + o.N[2] * bb + o.N[1] * o.N[2] * a);
    }
}

// The per-axis factor of kb_axis_value as a smooth function of a double
// distance: the reference's constants (dkf, float w and b) but no float
// rounding of the distance or of q, so that finite differences in the sample
// coordinates see the kernel's derivative rather than rounding steps.  Not a
// statement of the reference; used only in smooth-kernel mode.
double kb_axis_exact(double d, int N, float w, float b) {
    const double dk = (double)(float)(1.0 / (double)(float)(N - 1));
    const double W = (double)w * dk;
    if (d > 0.5 * W) return 0.0;
    const double q = 2.0 * d / W;
    return i0_series((double)b * std::sqrt(1.0 - q * q)) / (100.0 * (double)w * dk);
}

// Threshold of Reco.cpp:716-717: convFunction({0,0,0}) * 1e-7 (float promoted).
double kernel_threshold(const Oracle& o) {
    const double zero[3] = {0.0, 0.0, 0.0};
    return (double)conv_function(zero, o.N, o.w, o.b) * 1e-7;
}

// Reco::ConvKernel::calculateWeights (Reco.cpp:698-787).  With a non-null
// mask, samples whose mask entry is 0 get empty lists, which is what
// excludeMissingData (Reco.cpp:681-696) leaves for them afterwards.
void calculate_weights(Oracle& o, const uint8_t* mask) {
    o.idx.assign(o.nS, {});
    o.wts.assign(o.nS, {});
    o.wts64.assign(o.exact_w ? o.nS : 0, {});
    const float dkx = (float)(1.0 / (double)(float)(o.N[0] - 1));
    const float dky = (float)(1.0 / (double)(float)(o.N[1] - 1));
    const float dkz = (float)(1.0 / (double)(float)(o.N[2] - 1));
    const double thr = kernel_threshold(o);
    std::vector<int> kp;
    for (int i1 = -o.order; i1 <= o.order; i1++)
        for (int i2 = -o.order; i2 <= o.order; i2++)
            for (int i3 = -o.order; i3 <= o.order; i3++) { kp.push_back(i1); kp.push_back(i2); kp.push_back(i3); }
    const double invYZ = 1.0 / double(o.N[1] * o.N[2]);
    const double invZ = 1.0 / double(o.N[2]);
    const int yz = o.N[1] * o.N[2];
    const int sizeZ = o.N[2];
#pragma omp parallel for num_threads(o.nthreads) schedule(static)
    for (long i = 0; i < o.nS; i++) {
        if (mask && !mask[i]) continue;
        std::vector<int> pts;
        pts.reserve(kp.size() / 3);
        const double* k = &o.traj[3 * i];
        int nearest[3] = {(int)((o.N[0] - 1) * k[0] + 0.5), (int)((o.N[1] - 1) * k[1] + 0.5),
                          (int)((o.N[2] - 1) * k[2] + 0.5)};
        collect_neighbor_indices(o, nearest, pts, kp);
        for (size_t j = 0; j < pts.size(); j++) {
            int ind[3];
            ind[0] = (int)(pts[j] * invYZ);
            ind[1] = (int)((pts[j] - yz * ind[0]) * invZ);
            ind[2] = pts[j] - sizeZ * ind[1] - yz * ind[0];
            const double ex = k[0] - (double)((float)ind[0] * dkx);
            const double ey = k[1] - (double)((float)ind[1] * dky);
            const double ez = k[2] - (double)((float)ind[2] * dkz);
            const double dist[3] = {std::sqrt(ex * ex), std::sqrt(ey * ey), std::sqrt(ez * ez)};
            if (o.exact_w) {
                const double kv = (kb_axis_exact(dist[0], o.N[0], o.w, o.b) * kb_axis_exact(dist[1], o.N[1], o.w, o.b)) *
                                  kb_axis_exact(dist[2], o.N[2], o.w, o.b);
                if (kv > thr) {
                    o.idx[i].push_back(pts[j]);
                    o.wts[i].push_back((float)kv);
                    o.wts64[i].push_back(kv);
                }
                continue;
            }
            const double kv = (double)conv_function(dist, o.N, o.w, o.b);
            if (kv > thr) {
                o.idx[i].push_back(pts[j]);
                o.wts[i].push_back((float)kv);
            }
        }
    }
}

// Reco::ConvKernel::calculateKernelFT (Reco.cpp:560-601).  The sampled kernel
// is separable, so its centred inverse DFT is the product of three 1D
// transforms; each is evaluated here in FP64 (the reference uses an FP32 FFTW
// plan).  kernelFT = re / (re*re + im*im) with im = 0 (Reco.cpp:594-597).
void calculate_kernel_ft(Oracle& o) {
    std::vector<double> axr[3], axi[3];
    for (int a = 0; a < 3; a++) {
        const int n = o.N[a];
        const double dk = 1.0 / (double)(n - 1);
        const int c = n / 2;
        axr[a].assign(n, 0.0);
        axi[a].assign(n, 0.0);
        // sampled per-axis kernel x_m at grid index m (Reco.cpp:571-585)
        std::vector<double> xs(n, 0.0);
        for (int m = 0; m < n; m++) {
            const double d = std::sqrt(dk * dk * (double)(m - c) * (double)(m - c));
            xs[m] = kb_axis_value((float)d, n, o.w, o.b);
        }
        // centred inverse DFT along this axis (transformBWD without scaling,
        // Reco.cpp:589): X_k = s_k sum_m s_m x_m e^{+2 pi i m k / n}, s = (-1)^index.
        const double two_pi = 6.283185307179586476925286766559;
        for (int kk = 0; kk < n; kk++) {
            double re = 0.0, im = 0.0;
            for (int m = 0; m < n; m++) {
                if (xs[m] == 0.0) continue;
                const double sgn = (((m + kk) & 1) ? -1.0 : 1.0);
                const double ang = two_pi * (double)(((long)m * kk) % n) / (double)n;
                re += sgn * xs[m] * std::cos(ang);
                im += sgn * xs[m] * std::sin(ang);
            }
            axr[a][kk] = re;
            axi[a][kk] = im;
        }
    }
    o.kft.assign(o.Ng, 0.0f);
    for (int i = 0; i < o.N[0]; i++)
        for (int j = 0; j < o.N[1]; j++) {
            const double xyr = axr[0][i] * axr[1][j] - axi[0][i] * axi[1][j];
            const double xyi = axr[0][i] * axi[1][j] + axi[0][i] * axr[1][j];
            for (int k = 0; k < o.N[2]; k++) {
                const double re = xyr * axr[2][k] - xyi * axi[2][k];   // Re of the 3D value
                o.kft[(long)k + (long)o.N[2] * j + (long)o.N[1] * o.N[2] * i] = (float)(re / (re * re));
            }
        }
}

// Reco::reduceFOV (Reco.cpp:626-671), inclusive upper bound kept.
template <typename T>
void reduce_fov(const Oracle& o, Cplx<T>* v) {
    if (o.N[0] == o.imgRange[0] && o.N[1] == o.imgRange[1] && o.N[2] == o.imgRange[2]) return;
    int first[3], last[3];
    for (int a = 0; a < 3; a++) {
        first[a] = (o.N[a] - o.imgRange[a]) / 2 + o.offc[a];
        last[a] = o.imgRange[a] + first[a];
    }
    for (int x = 0; x < o.N[0]; x++)
        for (int y = 0; y < o.N[1]; y++)
            for (int z = 0; z < o.N[2]; z++)
                if (x < first[0] || x > last[0] || y < first[1] || y > last[1] || z < first[2] || z > last[2]) {
                    const long i = (long)z + (long)o.N[2] * y + (long)o.N[1] * o.N[2] * x;
                    v[i].re = 0; v[i].im = 0;
                }
}

// ---------------------------------------------------------------------------
// NUFFT.  The FP32 arithmetic follows the reference statement by statement;
// T = double evaluates the same statements in FP64.
// ---------------------------------------------------------------------------

// Reco::nufftBWD (Reco.cpp:468-519) for one coil.
template <typename T>
void nufft_bwd_coil(const Oracle& o, const Cplx<T>* data, Cplx<T>* img) {
    for (long i = 0; i < o.Ng; i++) { img[i].re = 0; img[i].im = 0; }
    for (long i = 0; i < o.nS; i++) {
        const T d0 = data[i].re, d1 = data[i].im;
        const auto& ix = o.idx[i];
        const auto& wv = o.wts[i];
        const bool ex = o.exact_w && sizeof(T) == sizeof(double);
        for (size_t t = 0; t < ix.size(); t++) {
            const T kv = ex ? (T)o.wts64[i][t] : (T)wv[t];
            img[ix[t]].re += d0 * kv;
            img[ix[t]].im += d1 * kv;
        }
    }
    transform_bwd(img, o.N, true);
    const double rescale = 1e-3;
    for (long i = 0; i < o.Ng; i++) {
        const double gr = img[i].re, gi = img[i].im;
        const double kf = o.kft[i];
        img[i].re = (T)((gr * kf) * rescale);
        img[i].im = (T)((gi * kf) * rescale);
    }
    reduce_fov(o, img);
}

// Reco::nufftFWD (Reco.cpp:521-553) for one coil.  img is modified in place
// (deapodized and transformed), as in the reference.
template <typename T>
void nufft_fwd_coil(const Oracle& o, Cplx<T>* img, Cplx<T>* data) {
    for (long i = 0; i < o.Ng; i++) {
        const double gr = img[i].re, gi = img[i].im;
        const double kf = o.kft[i];
        img[i].re = (T)(gr * kf);
        img[i].im = (T)(gi * kf);
    }
    transform_fwd(img, o.N);
    for (long i = 0; i < o.nS; i++) {
        T a0 = 0, a1 = 0;
        const auto& ix = o.idx[i];
        const auto& wv = o.wts[i];
        const bool ex = o.exact_w && sizeof(T) == sizeof(double);
        for (size_t t = 0; t < ix.size(); t++) {
            const double kv = ex ? o.wts64[i][t] : (double)wv[t];
            a0 = (T)((double)a0 + (double)img[ix[t]].re * kv);
            a1 = (T)((double)a1 + (double)img[ix[t]].im * kv);
        }
        data[i].re = a0; data[i].im = a1;
    }
}

// Reco::conjScalarProduct (Reco.cpp:131-143): sum conj(x) y; the per-element
// term is formed in the vector precision and accumulated in double.
template <typename T>
void conj_dot(const Oracle& o, const Cplx<T>* x, const Cplx<T>* y, long n, double* out) {
    double r0 = 0, r1 = 0;
#pragma omp parallel for num_threads(o.nthreads) schedule(static) reduction(+ : r0, r1)
    for (long i = 0; i < n; i++) {
        const T t0 = x[i].re * y[i].re + x[i].im * y[i].im;
        const T t1 = x[i].re * y[i].im - x[i].im * y[i].re;
        r0 += t0;
        r1 += t1;
    }
    out[0] = r0; out[1] = r1;
}

// Reco::complexDivision (Reco.cpp:145-151)
void complex_division(const double* x, const double* y, double* r) {
    const double den = (double)(y[0] * y[0] + y[1] * y[1]);
    r[0] = (x[0] * y[0] + x[1] * y[1]) / den;
    r[1] = (x[1] * y[0] - x[0] * y[1]) / den;
}

}  // namespace

// ===========================================================================
// extern "C" surface used by tests/ and bench.py (ctypes)
// ===========================================================================
extern "C" {

double orc_i0(double x) { return i0_series(x); }

void orc_set_verbose(int v) { g_verbose = v; }

double orc_kb_axis(float dxf, int N, float w, float b) { return kb_axis_value(dxf, N, w, b); }

float orc_conv_function(const double* dist, const int* N, float w, float b) {
    return conv_function(dist, N, w, b);
}

// In-place centred 3D transforms on interleaved complex arrays.
// prec 64: double[2*n]; prec 32: float[2*n].  dir -1 = transformFWD,
// +1 = transformBWD(rescale).
int orc_fft3(void* a, const int* N, int prec, int dir, int rescale) {
    if (prec == 64) {
        if (dir < 0) transform_fwd((Cplx<double>*)a, N);
        else transform_bwd((Cplx<double>*)a, N, rescale != 0);
    } else if (prec == 32) {
        if (dir < 0) transform_fwd((Cplx<float>*)a, N);
        else transform_bwd((Cplx<float>*)a, N, rescale != 0);
    } else {
        return 1;
    }
    return 0;
}

// Plain (uncentred, unnormalized) 1D DFT of length n, FP64, for pinning tests.
int orc_fft1(double* a, int n, int sign) {
    Plan1D<double> p = make_plan<double>(n);
    std::vector<Cplx<double>> w(n);
    fft1(p, (Cplx<double>*)a, w.data(), sign);
    return 0;
}

void* orc_create(const int* grid, const int* samp, int ncoil, float w, float b, int order, int nthreads) {
    Oracle* o = new Oracle();
    for (int a = 0; a < 3; a++) { o->N[a] = grid[a]; o->S[a] = samp[a]; o->imgRange[a] = grid[a]; o->offc[a] = 0; }
    o->Ng = (long)grid[0] * grid[1] * grid[2];
    o->nS = (long)samp[0] * samp[1] * samp[2];
    o->C = ncoil;
    o->w = w; o->b = b; o->order = order;
    o->nthreads = nthreads > 0 ? nthreads : omp_get_max_threads();
    return o;
}

void orc_destroy(void* h) { delete (Oracle*)h; }

void orc_set_img_range(void* h, const int* range, const int* offc) {
    Oracle* o = (Oracle*)h;
    for (int a = 0; a < 3; a++) { o->imgRange[a] = range[a]; o->offc[a] = offc[a]; }
}

void orc_set_trajectory(void* h, const double* traj) {
    Oracle* o = (Oracle*)h;
    o->traj.assign(traj, traj + 3 * o->nS);
}

// calculateWeights (Reco.cpp:698-787).  data may be null; if given, samples
// whose coil-0 value is exactly 0+0i get empty lists (excludeMissingData,
// Reco.cpp:681-696).
void orc_calculate_weights(void* h, const float* data) {
    Oracle* o = (Oracle*)h;
    std::vector<uint8_t> mask;
    if (data) {
        mask.resize(o->nS);
        for (long i = 0; i < o->nS; i++) mask[i] = !(data[2 * i] == 0.0f && data[2 * i + 1] == 0.0f);
    }
    calculate_weights(*o, data ? mask.data() : nullptr);
}

// excludeMissingData (Reco.cpp:681-696) on already computed lists.
void orc_exclude_missing_data(void* h, const float* data) {
    Oracle* o = (Oracle*)h;
    for (long i = 0; i < o->nS; i++)
        if (data[2 * i] == 0.0f && data[2 * i + 1] == 0.0f) {
            o->idx[i].clear(); o->wts[i].clear();
            if (o->exact_w) o->wts64[i].clear();
        }
}

// Smooth-kernel mode for finite-difference checks (set before the weights).
void orc_set_exact_weights(void* h, int on) { ((Oracle*)h)->exact_w = on != 0; }
void orc_set_atomic_coil_sum(void* h, int on) { ((Oracle*)h)->atomic_sum = on != 0; }

long orc_num_taps(void* h) {
    Oracle* o = (Oracle*)h;
    long n = 0;
    for (long i = 0; i < o->nS; i++) n += (long)o->idx[i].size();
    return n;
}

// Export the per-sample lists as CSR: counts[nS], idx[total], w[total].
void orc_get_weights(void* h, int32_t* counts, int32_t* idx, float* w) {
    Oracle* o = (Oracle*)h;
    long p = 0;
    for (long i = 0; i < o->nS; i++) {
        counts[i] = (int32_t)o->idx[i].size();
        for (size_t t = 0; t < o->idx[i].size(); t++) { idx[p] = o->idx[i][t]; w[p] = o->wts[i][t]; p++; }
    }
}

double orc_kernel_threshold(void* h) { return kernel_threshold(*(Oracle*)h); }

void orc_calculate_kernel_ft(void* h) { calculate_kernel_ft(*(Oracle*)h); }

void orc_get_kernel_ft(void* h, float* out) {
    Oracle* o = (Oracle*)h;
    std::copy(o->kft.begin(), o->kft.end(), out);
}

void orc_set_density(void* h, const float* d) {
    Oracle* o = (Oracle*)h;
    if (d) o->dcf.assign(d, d + o->nS);
    else o->dcf.clear();
}

// Maps: C x Ng complex64, coil-major (IO::importSensitivityMaps, IO.cpp:25-51).
void orc_set_maps(void* h, const float* maps) {
    Oracle* o = (Oracle*)h;
    o->maps.resize((size_t)o->C * o->Ng);
    std::memcpy(o->maps.data(), maps, sizeof(float) * 2 * (size_t)o->C * o->Ng);
}

// performDensityCorrection (functions.h:319-335) on a C x nS complex64 array.
void orc_density_correct(void* h, float* data) {
    Oracle* o = (Oracle*)h;
    if (o->dcf.empty()) return;
    for (int c = 0; c < o->C; c++)
        for (long i = 0; i < o->nS; i++) {
            data[2 * (c * o->nS + i)] *= o->dcf[i];
            data[2 * (c * o->nS + i) + 1] *= o->dcf[i];
        }
}

// nufftFWD over all coils: img [C][Ng] -> data [C][nS]; img is not modified.
int orc_nufft_fwd(void* h, int prec, const void* img, void* data) {
    Oracle* o = (Oracle*)h;
    if (o->kft.empty()) return 2;
    const int nt = o->nthreads;
#pragma omp parallel for num_threads(nt) schedule(dynamic)
    for (int c = 0; c < o->C; c++) {
        if (prec == 64) {
            std::vector<Cplx<double>> tmp((const Cplx<double>*)img + (size_t)c * o->Ng,
                                          (const Cplx<double>*)img + (size_t)(c + 1) * o->Ng);
            nufft_fwd_coil(*o, tmp.data(), (Cplx<double>*)data + (size_t)c * o->nS);
        } else {
            std::vector<Cplx<float>> tmp((const Cplx<float>*)img + (size_t)c * o->Ng,
                                         (const Cplx<float>*)img + (size_t)(c + 1) * o->Ng);
            nufft_fwd_coil(*o, tmp.data(), (Cplx<float>*)data + (size_t)c * o->nS);
        }
    }
    return 0;
}

// nufftBWD over all coils: data [C][nS] -> img [C][Ng].
int orc_nufft_bwd(void* h, int prec, const void* data, void* img) {
    Oracle* o = (Oracle*)h;
    if (o->kft.empty()) return 2;
    const int nt = o->nthreads;
#pragma omp parallel for num_threads(nt) schedule(dynamic)
    for (int c = 0; c < o->C; c++) {
        if (prec == 64)
            nufft_bwd_coil(*o, (const Cplx<double>*)data + (size_t)c * o->nS, (Cplx<double>*)img + (size_t)c * o->Ng);
        else
            nufft_bwd_coil(*o, (const Cplx<float>*)data + (size_t)c * o->nS, (Cplx<float>*)img + (size_t)c * o->Ng);
    }
    return 0;
}

}  // extern "C"

namespace {

// One application q = A p of the CG system matrix (Reco.cpp:232-262):
// img_c = p (.) S_c; nufftFWD; data *= d; nufftBWD; q = sum_c conj(S_c) (.) img_c.
template <typename T>
int apply_normal(Oracle& o, const Cplx<T>* p, Cplx<T>* q, std::vector<Cplx<T>>& img, std::vector<Cplx<T>>& data) {
    const long Ng = o.Ng, nS = o.nS;
    const int nt = o.nthreads;
    // elementProduct (Reco.cpp:85-93, called 236-240)
#pragma omp parallel for num_threads(nt) schedule(static)
    for (int c = 0; c < o.C; c++) {
        const Cplx<float>* s = o.maps.data() + (size_t)c * Ng;
        Cplx<T>* ic = img.data() + (size_t)c * Ng;
        for (long i = 0; i < Ng; i++) {
            ic[i].re = p[i].re * (T)s[i].re - p[i].im * (T)s[i].im;
            ic[i].im = p[i].re * (T)s[i].im + p[i].im * (T)s[i].re;
        }
    }
    // nufftFWD (Reco.cpp:521-553)
#pragma omp parallel for num_threads(nt) schedule(dynamic)
    for (int c = 0; c < o.C; c++) nufft_fwd_coil(o, img.data() + (size_t)c * Ng, data.data() + (size_t)c * nS);
    // density correction (Reco.cpp:243-254)
    if (!o.dcf.empty()) {
#pragma omp parallel for num_threads(nt) schedule(static)
        for (int c = 0; c < o.C; c++)
            for (long i = 0; i < nS; i++) {
                data[(size_t)c * nS + i].re *= (T)o.dcf[i];
                data[(size_t)c * nS + i].im *= (T)o.dcf[i];
            }
    }
    // nufftBWD (Reco.cpp:468-519)
#pragma omp parallel for num_threads(nt) schedule(dynamic)
    for (int c = 0; c < o.C; c++) nufft_bwd_coil(o, data.data() + (size_t)c * nS, img.data() + (size_t)c * Ng);
    // mutliplyByConjSenseMap + addToSumOfImages (Reco.cpp:107-129, 257-262).
    // The reference sums coils with float atomics in arbitrary order; by
    // default the coil order is fixed here (0..C-1) and the grid is split
    // across threads.
    for (long i = 0; i < Ng; i++) { q[i].re = 0; q[i].im = 0; }
    if (o.atomic_sum) {
        // the reference's loop: one coil per thread, conj-map multiply in
        // place, then element-wise atomic adds into q (Reco.cpp:257-262)
#pragma omp parallel for num_threads(nt) schedule(static)
        for (int c = 0; c < o.C; c++) {
            const Cplx<float>* s = o.maps.data() + (size_t)c * Ng;
            Cplx<T>* ic = img.data() + (size_t)c * Ng;
            for (long i = 0; i < Ng; i++) {
                const double re = ic[i].re, im = ic[i].im;
                ic[i].re = (T)(re * s[i].re + im * s[i].im);
                ic[i].im = (T)(-re * s[i].im + im * s[i].re);
            }
            for (long i = 0; i < Ng; i++) {
#pragma omp atomic
                q[i].re += ic[i].re;
#pragma omp atomic
                q[i].im += ic[i].im;
            }
        }
        if (!std::isfinite((double)q[0].re) || !std::isfinite((double)q[0].im)) return 3;  // Reco.cpp:264-268
        return 0;
    }
#pragma omp parallel for num_threads(nt) schedule(static)
    for (long i = 0; i < Ng; i++) {
        for (int c = 0; c < o.C; c++) {
            const Cplx<float> s = o.maps[(size_t)c * Ng + i];
            const Cplx<T> v = img[(size_t)c * Ng + i];
            const double re = v.re, im = v.im;
            const T r0 = (T)(re * s.re + im * s.im);
            const T r1 = (T)(-re * s.im + im * s.re);
            q[i].re += r0;
            q[i].im += r1;
        }
    }
    if (!std::isfinite((double)q[0].re) || !std::isfinite((double)q[0].im)) return 3;  // Reco.cpp:264-268
    return 0;
}

// Reco::cgSenseDynamic (Reco.cpp:175-321) from the coil images img_c = E^H W y.
template <typename T>
int cg_sense(Oracle& o, std::vector<Cplx<T>>& img, double thresh, int max_iter, Cplx<T>* x,
             double* resid_out, int* iters_out, double* loop_seconds = nullptr, double* iter_seconds = nullptr) {
    const long Ng = o.Ng;
    const int nt = o.nthreads;
    std::vector<Cplx<T>> a(Ng), q(Ng), p(Ng), r(Ng);
    std::vector<Cplx<T>> data((size_t)o.C * o.nS);
    // initial image a = sum_c conj(S_c) img_c  (Reco.cpp:183-190)
    for (long i = 0; i < Ng; i++) {
        T a0 = 0, a1 = 0;
        for (int c = 0; c < o.C; c++) {
            const Cplx<float> s = o.maps[(size_t)c * Ng + i];
            const Cplx<T> v = img[(size_t)c * Ng + i];
            a0 += (T)s.re * v.re + (T)s.im * v.im;
            a1 += (T)s.re * v.im - (T)s.im * v.re;
        }
        a[i].re = a0; a[i].im = a1;
    }
    for (long i = 0; i < Ng; i++) { p[i] = a[i]; r[i] = a[i]; x[i].re = 0; x[i].im = 0; q[i].re = 0; q[i].im = 0; }
    int iteration = 0;
    double current = 1000;
    const double t0 = omp_get_wtime();
    double ti = t0;
    while (current > thresh) {
        const int st = apply_normal(o, p.data(), q.data(), img, data);
        if (st) { *iters_out = iteration; return st; }
        double rr[2], pq[2], al[2];
        conj_dot(o, r.data(), r.data(), Ng, rr);
        conj_dot(o, p.data(), q.data(), Ng, pq);
        complex_division(rr, pq, al);
        // addToArrayScaled (Reco.cpp:153-161): x += al p; r += (-al) q
#pragma omp parallel for num_threads(nt) schedule(static)
        for (long i = 0; i < Ng; i++) {
            x[i].re = (T)((double)x[i].re + (al[0] * (double)p[i].re - al[1] * (double)p[i].im));
            x[i].im = (T)((double)x[i].im + (al[0] * (double)p[i].im + al[1] * (double)p[i].re));
        }
        const double nal[2] = {-al[0], -al[1]};
#pragma omp parallel for num_threads(nt) schedule(static)
        for (long i = 0; i < Ng; i++) {
            r[i].re = (T)((double)r[i].re + (nal[0] * (double)q[i].re - nal[1] * (double)q[i].im));
            r[i].im = (T)((double)r[i].im + (nal[0] * (double)q[i].im + nal[1] * (double)q[i].re));
        }
        // residual (Reco.cpp:292-299): serial FP64 sum of vector-precision terms
        current = 0;
        for (long i = 0; i < Ng; i++) current += (T)(r[i].re * r[i].re + r[i].im * r[i].im);
        current = std::sqrt(current);
        if (resid_out) resid_out[iteration] = current;
        if (g_verbose) std::fprintf(stderr, "[oracle] CG iteration %d residual %.6e\n", iteration, current);
        double rn[2], be[2];
        conj_dot(o, r.data(), r.data(), Ng, rn);
        complex_division(rn, rr, be);
        // addArraysScaled (Reco.cpp:163-173): p = r + be p
#pragma omp parallel for num_threads(nt) schedule(static)
        for (long i = 0; i < Ng; i++) {
            const double pr = p[i].re, pi = p[i].im;
            p[i].re = (T)((double)r[i].re + be[0] * pr - be[1] * pi);
            p[i].im = (T)((double)r[i].im + be[0] * pi + be[1] * pr);
        }
        if (iter_seconds) {
            const double tn = omp_get_wtime();
            iter_seconds[iteration] = tn - ti;
            ti = tn;
        }
        iteration++;
        if (iteration > max_iter) break;
    }
    if (loop_seconds) *loop_seconds = omp_get_wtime() - t0;
    *iters_out = iteration;
    return 0;
}

}  // namespace

extern "C" {

// One CG system-matrix application q = A p (prec 64: complex128, 32: complex64).
int orc_apply_normal(void* h, int prec, const void* p, void* q) {
    Oracle* o = (Oracle*)h;
    if (o->kft.empty() || o->maps.empty()) return 2;
    if (prec == 64) {
        std::vector<Cplx<double>> img((size_t)o->C * o->Ng), data((size_t)o->C * o->nS);
        return apply_normal(*o, (const Cplx<double>*)p, (Cplx<double>*)q, img, data);
    }
    std::vector<Cplx<float>> img((size_t)o->C * o->Ng), data((size_t)o->C * o->nS);
    return apply_normal(*o, (const Cplx<float>*)p, (Cplx<float>*)q, img, data);
}

// The per-state sequence of CGSense.cpp:224-258 after the weights exist:
// performDensityCorrection on a copy of data, nufftBWD, cgSenseDynamic.
// data: C x nS complex64 (memory order).  x_out: Ng complex of prec.
// resid_out: at least max_iter+1 doubles.  max_iter is the reference's
// nMaxIter (Reco.cpp:228): at most max_iter+1 passes are made.
int orc_cg_sense_timed(void* h, int prec, const float* data, double thresh, int max_iter, void* x_out,
                       double* resid_out, int* iters_out, double* loop_seconds, double* iter_seconds) {
    Oracle* o = (Oracle*)h;
    if (o->kft.empty() || o->maps.empty()) return 2;
    const size_t nd = (size_t)o->C * o->nS;
    std::vector<float> dc(data, data + 2 * nd);
    orc_density_correct(h, dc.data());
    if (prec == 64) {
        std::vector<Cplx<double>> d(nd), img((size_t)o->C * o->Ng);
        for (size_t i = 0; i < nd; i++) { d[i].re = dc[2 * i]; d[i].im = dc[2 * i + 1]; }
        orc_nufft_bwd(h, 64, d.data(), img.data());
        return cg_sense(*o, img, thresh, max_iter, (Cplx<double>*)x_out, resid_out, iters_out, loop_seconds,
                        iter_seconds);
    }
    std::vector<Cplx<float>> img((size_t)o->C * o->Ng);
    orc_nufft_bwd(h, 32, dc.data(), img.data());
    return cg_sense(*o, img, thresh, max_iter, (Cplx<float>*)x_out, resid_out, iters_out, loop_seconds,
                    iter_seconds);
}

int orc_cg_sense(void* h, int prec, const float* data, double thresh, int max_iter, void* x_out,
                 double* resid_out, int* iters_out) {
    return orc_cg_sense_timed(h, prec, data, thresh, max_iter, x_out, resid_out, iters_out, nullptr, nullptr);
}

// std::cyl_bessel_i(0, x) of libstdc++ (C++17 special functions), a second
// independent I0 for the tests.
double orc_i0_std(double x) { return std::cyl_bessel_i(0.0, x); }

}  // extern "C"
