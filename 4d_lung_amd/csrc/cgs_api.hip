// cgs_api.hip -- host side of libcgs_hip.so: the C-ABI of include/cgs.h.
//
// One cgs_ctx holds the device state of one Reco (Reco.h:13-214): the
// trajectory, the KB tables, the deapodization map, the coil maps and one or
// more respiratory bins (datasets).  It owns one HIP stream; every public call
// returns a status and keeps an error message instead of exiting.
//
// Built with -ffp-contract=off: the host computes the KB constants and the
// per-axis deapodization transforms with the reference's operation order.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "cgs.h"
#include "cgs_internal.h"
#include "cgs_kernels.h"

using namespace cgs;

namespace {

struct BinData {
    long long m = 0;             // active samples
    int* perm = nullptr;         // [m] original index per sorted position
    int* tile_off = nullptr;     // [ntiles+1] relative offsets
    int4* base = nullptr;        // [m]
    double* vals = nullptr;      // [m][VS]
    float2* y = nullptr;         // [m][C] measured data (no density weights), or null
    float* dcf_full = nullptr;   // [nS] own density weights, or null (context's are used)
    bool built = false;
    bool has_y = false;          // values set (also for a bin without samples, m = 0)
};

struct ProfEntry {
    std::string name;
    hipEvent_t a, b;
};

}  // namespace

struct cgs_ctx {
    int N[3];
    long long ng;
    long long xp = 0, bp = 0;      // coil-grid x-plane and bin pitch (elements)
    int S[3];
    long long ns;
    int C;
    float w = 4.25f, b = 14.9086f;
    int order = 5;
    int imgRange[3];
    int offc[3];
    int tcfg = 0;
    Tiles tiles{};
    KbParams kb{};
    bool kb_ready = false;
    hipStream_t st = nullptr;      // stream every call runs on
    hipStream_t own_st = nullptr;  // the context's own stream (st unless cgs_set_stream chose another)
    std::string err;
    long long dev_bytes = 0;

    double* d_traj = nullptr;
    float* d_dcf_full = nullptr;   // context density weights (Reco::denCorr)
    float* d_kft = nullptr;
    bool kft_ready = false;
    float2* d_tw[3] = {nullptr, nullptr, nullptr};
    float2* d_maps = nullptr;
    bool maps_ready = false;

    // bins
    int nbins = 1;
    std::vector<BinData> bins;
    bool packed = false;
    std::vector<long long> bin_off;
    int4* p_base = nullptr;
    double* p_vals = nullptr;
    float* p_dcf = nullptr;
    int* p_sbin = nullptr;
    float2* p_y = nullptr;
    float2* p_yw = nullptr;
    int* p_tile_off = nullptr;
    bool p_alias = false;          // packed arrays alias bins[0]'s
    // column pruning of the bins' x passes (built with the packed arrays)
    int prune_mode = 1;            // 0 off, 1 when it skips enough, 2 always
    bool prune_on = false;
    long long pr_ngrp = 0, pr_gstride = 0;
    uint8_t* d_gact = nullptr;     // [nbins][pr_ngrp]
    int* d_glist = nullptr;        // [nbins][pr_gstride]
    int* d_gcnt = nullptr;         // [nbins]
    uint8_t* d_tneed = nullptr;    // [nbins][nty*ntz]
    // per-tile gridding candidate lists of the packed bins (built with them)
    int grid_lists = 1;            // 1: build and use them; 0: gridding scans every launch
    int zbpw = 64;                 // batched inverse z pass: bins per workgroup and cell line
    int* d_tlist = nullptr;
    long long tlist_n = 0;
    long long* d_tlist_off = nullptr;   // [nbins * ntiles + 1]
    // per-(y, z)-cell sample lists of the fused gridding (k_fft_xgrid)
    int xgrid = 1;                 // 1: gridding fused into the inverse x pass where supported
    bool xl_on = false;            // lists built for the packed bins
    int* d_xrec = nullptr;         // records of the sorted entries (xrec_stride(kw) ints each)
    long long xl_n = 0;
    long long* d_xoff = nullptr;   // [nbins][ny * nz + 1]
    std::vector<int> h_gcnt;
    std::vector<double> pr_gfrac, pr_tfrac;   // per bin: groups transformed, tile columns written

    // workspace
    float2* d_grid = nullptr;
    int grid_cap = 0;
    float2 *d_x = nullptr, *d_r = nullptr, *d_p = nullptr, *d_q = nullptr;
    int vec_bins = 0;
    double* d_part = nullptr;      // z-pass partials [nbins][nxy][2]
    double* d_part2 = nullptr;     // update partials [nbins][nblk_upd][2]
    int nblk_upd = 0;
    CgScalars* d_sc = nullptr;
    int* d_done = nullptr;
    double* d_hist = nullptr;
    int hist_stride = 0;
    int hist_bins = 0;
    double thresh = 0.0;
    int max_iter = 15;
    bool cg_ready = false;
    int group = 1;                 // bins per batched launch
    int group_cap = 0;             // caller's limit on bins per launch (0: memory only)

    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    double g_thresh = 0.0;         // loop-control values the graph was captured with
    int g_max_iter = -1;
    double* g_hist = nullptr;
    int g_group = 0;

    bool prof = false;
    std::vector<ProfEntry> pending;
    std::map<std::string, std::pair<long long, double>> acc;
};

struct cgs_fft {
    int N[3];
    long long ng;
    float2* d_a = nullptr;
    float2* d_tw[3] = {nullptr, nullptr, nullptr};
    hipStream_t st = nullptr;
};

// A free-standing set of m sample coordinates on a context's grid (the
// differentiable operators): tile-sorted tables like a bin's, plus the KB
// derivative table.  Valid for the tile shape and kernel it was built with.
struct cgs_pts {
    cgs_ctx* h = nullptr;
    long long m = 0;          // samples
    long long cap = 0;        // samples the coordinate buffer holds
    double* k64 = nullptr;    // [cap][3] coordinates widened to double
    int* perm = nullptr;      // [m]
    int* tile_off = nullptr;  // [ntiles + 1]
    int4* base = nullptr;     // [m]
    double* vals = nullptr;   // [m][VS]
    double* dvals = nullptr;  // [m][VS]
    int* sbin = nullptr;      // [m] zeros (every sample in bin 0)
    float2* ys = nullptr;     // [m][C] sorted sample values (scratch)
    int tcfg = -1;
    int kw = 0;
    float w = 0.0f, b = 0.0f;
    bool ready = false;
};

namespace {

int fail(cgs_ctx* h, int code, const std::string& msg) {
    if (h) h->err = msg;
    return code;
}

int fail_hip(cgs_ctx* h, const char* what, hipError_t e) {
    return fail(h, CGS_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// The captured batched iteration bakes in buffer addresses and the group
// split; it is dropped whenever one of them changes.
void drop_graph(cgs_ctx* h) {
    if (h->gexec) { (void)hipGraphExecDestroy(h->gexec); h->gexec = nullptr; }
    if (h->graph) { (void)hipGraphDestroy(h->graph); h->graph = nullptr; }
}

#define CK(call)                                                   \
    do {                                                           \
        hipError_t e_ = (call);                                    \
        if (e_ != hipSuccess) return fail_hip(h, #call, e_);       \
    } while (0)

template <typename T>
int dalloc(cgs_ctx* h, T** p, size_t count) {
    *p = nullptr;
    if (count == 0) return CGS_OK;
    void* q = nullptr;
    const hipError_t e = hipMalloc(&q, count * sizeof(T));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(h, CGS_ERR_MEMORY, "device allocation of " + std::to_string(count * sizeof(T)) + " bytes failed");
    }
    *p = (T*)q;
    if (h) h->dev_bytes += (long long)(count * sizeof(T));
    return CGS_OK;
}

template <typename T>
void dfree(cgs_ctx* h, T*& p, size_t count) {
    if (p) {
        (void)hipFree(p);
        if (h) h->dev_bytes -= (long long)(count * sizeof(T));
    }
    p = nullptr;
}

#define CKS(call)                 \
    do {                          \
        int s_ = (call);          \
        if (s_ != CGS_OK) return s_; \
    } while (0)

// ---------------------------------------------------------------- profiling
struct Prof {
    cgs_ctx* h;
    hipEvent_t a = nullptr;
    const char* name;
    Prof(cgs_ctx* hh, const char* n) : h(hh), name(n) {
        if (h->prof) {
            (void)hipEventCreate(&a);
            (void)hipEventRecord(a, h->st);
        }
    }
    ~Prof() {
        if (h->prof && a) {
            hipEvent_t b;
            (void)hipEventCreate(&b);
            (void)hipEventRecord(b, h->st);
            h->pending.push_back({name, a, b});
        }
    }
};

void prof_flush(cgs_ctx* h) {
    if (h->pending.empty()) return;
    (void)hipStreamSynchronize(h->st);
    for (auto& e : h->pending) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, e.a, e.b) == hipSuccess) {
            auto& slot = h->acc[e.name];
            slot.first += 1;
            slot.second += ms;
        }
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    h->pending.clear();
}

// ------------------------------------------------------------- KB constants
// Same statements as Reco::ConvKernel::convFunction (Reco.h:218-242).
double host_i0(double x) {
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

double host_kb_axis(float dxf, const AxisConst& c, double b) {
    if (dxf > c.cut) return 0.0;
    const float q = (2.0f * dxf) / c.wdk;
    const double q2 = (double)q * (double)q;
    const double arg = b * std::sqrt(1.0 - q2);
    return host_i0(arg) / (double)c.denom;
}

int setup_kb(cgs_ctx* h) {
    KbParams& kb = h->kb;
    const float wHalf = (float)((double)h->w / 2.0);
    for (int a = 0; a < 3; ++a) {
        AxisConst& c = kb.ax[a];
        c.dkf = (float)(1.0 / (double)(float)(h->N[a] - 1));
        c.cut = wHalf * c.dkf;
        c.wdk = h->w * c.dkf;
        c.denom = (100.0f * h->w) * c.dkf;
    }
    kb.b = (double)h->b;
    const double v0 = host_kb_axis(0.0f, kb.ax[0], kb.b);
    const double v1 = host_kb_axis(0.0f, kb.ax[1], kb.b);
    const double v2 = host_kb_axis(0.0f, kb.ax[2], kb.b);
    kb.thr = (double)(float)((v0 * v1) * v2) * 1e-7;   // Reco.cpp:716-717
    kb.thr_up = (float)kb.thr;
    if ((double)kb.thr_up <= kb.thr) kb.thr_up = std::nextafter(kb.thr_up, INFINITY);
    int hw = (int)std::floor((double)wHalf + 0.5 + 1e-4);
    if (hw > h->order) hw = h->order;
    kb.h = hw;
    kb.kw = 2 * hw + 1;
    kb.order = h->order;
    if (kb.kw != 3 && kb.kw != 5 && kb.kw != 7)
        return fail(h, CGS_ERR_ARG, "kernel width gives a " + std::to_string(kb.kw) +
                                        "-point window; supported windows are 3, 5 and 7 points");
    int tx, ty, tz;
    grid_tile_dims(h->tcfg, &tx, &ty, &tz);
    Tiles& t = h->tiles;
    t.tx = tx; t.ty = ty; t.tz = tz;
    t.ntx = (h->N[0] + tx - 1) / tx;
    t.nty = (h->N[1] + ty - 1) / ty;
    t.ntz = (h->N[2] + tz - 1) / tz;
    t.ntiles = t.ntx * t.nty * t.ntz;
    if (hw > std::min(tx, std::min(ty, tz)))
        return fail(h, CGS_ERR_ARG, "gridding tile smaller than the kernel half width");
    h->kb_ready = true;
    return CGS_OK;
}

// Pitch of the coil grids' x planes (elements).  The strided x pass reads a
// column segment from every plane: planes a multiple of 2 MiB apart put those
// reads on few HBM channels (512x512x320 grids of 32 coils, planes 40 MiB
// apart: x pass at 2.9 TB/s; the same planes 4 KiB further apart: 3.75 TB/s),
// so such planes are padded by 4 KiB.
#ifndef CGS_PITCH_MOD
#define CGS_PITCH_MOD (2LL << 20)
#endif
#ifndef CGS_PITCH_PAD
#define CGS_PITCH_PAD (4LL << 10)
#endif
long long coil_grid_pitch(int ny, int nz, int C) {
    long long p = (long long)ny * nz * C;
    if ((p * (long long)sizeof(float2)) % CGS_PITCH_MOD == 0) p += CGS_PITCH_PAD / (long long)sizeof(float2);
    return p;
}

Geom geom(const cgs_ctx* h) {
    Geom g;
    g.nx = h->N[0]; g.ny = h->N[1]; g.nz = h->N[2];
    g.C = h->C;
    g.ng = h->ng;
    g.xp = h->xp;
    g.bp = h->bp;
    return g;
}

int upload_twiddles(cgs_ctx* h, const int* N, float2** tw) {
    for (int a = 0; a < 3; ++a) {
        const int L = N[a];
        std::vector<float2> t(L);
        const double two_pi = 6.283185307179586476925286766559;
        for (int n = 0; n < L; ++n) {
            const double ang = -two_pi * (double)n / (double)L;
            t[n] = make_float2((float)std::cos(ang), (float)std::sin(ang));
        }
        CKS(dalloc(h, &tw[a], L));
        CK(hipMemcpy(tw[a], t.data(), sizeof(float2) * L, hipMemcpyHostToDevice));
    }
    return CGS_OK;
}

// ------------------------------------------------------------------ bins
void free_bin(cgs_ctx* h, BinData& b) {
    const int vs = val_stride(h->kb.kw);
    dfree(h, b.perm, b.m);
    dfree(h, b.tile_off, h->tiles.ntiles + 1);
    dfree(h, b.base, b.m);
    dfree(h, b.vals, b.m * vs);
    dfree(h, b.y, b.m * h->C);
    b.m = 0;
    b.built = false;
    b.has_y = false;
}

void free_packed(cgs_ctx* h) {
    // packed arrays hold at least one element (bins without samples)
    const long long mt = std::max(1LL, h->bin_off.empty() ? 0 : h->bin_off.back());
    if (!h->p_alias) {
        dfree(h, h->p_base, mt);
        dfree(h, h->p_vals, mt * val_stride(h->kb.kw));
        dfree(h, h->p_y, mt * h->C);
        dfree(h, h->p_tile_off, (long long)h->nbins * (h->tiles.ntiles + 1));
    }
    h->p_base = nullptr; h->p_vals = nullptr; h->p_y = nullptr; h->p_tile_off = nullptr;
    dfree(h, h->p_dcf, mt);
    dfree(h, h->p_sbin, mt);
    dfree(h, h->p_yw, mt * h->C);
    dfree(h, h->d_gact, h->nbins * h->pr_ngrp);
    dfree(h, h->d_glist, h->nbins * h->pr_gstride);
    dfree(h, h->d_gcnt, h->nbins);
    dfree(h, h->d_tneed, (long long)h->nbins * h->tiles.nty * h->tiles.ntz);
    dfree(h, h->d_tlist, std::max(1LL, h->tlist_n));
    dfree(h, h->d_tlist_off, (long long)h->nbins * h->tiles.ntiles + 1);
    h->tlist_n = 0;
    dfree(h, h->d_xrec, std::max(1LL, h->xl_n * xrec_stride(h->kb.kw)));
    dfree(h, h->d_xoff, (long long)h->nbins * ((long long)h->N[1] * h->N[2] + 1));
    h->xl_n = 0;
    h->xl_on = false;
    h->prune_on = false;
    h->pr_ngrp = h->pr_gstride = 0;
    h->h_gcnt.clear();
    h->pr_gfrac.clear();
    h->pr_tfrac.clear();
    h->p_alias = false;
    h->packed = false;
    drop_graph(h);
}

// Column pruning of the bins' x passes.  A bin's samples reach the grid
// cells (y, z) within the KB window of their nearest cell (clamped, as
// degridding reads them) -- every x.  The x pass transforms column groups of
// W*cpb consecutive (y, z, coil) columns; a group no window reaches is
// skipped by the forward x pass (degridding never reads it) and written as
// zeros by the inverse one (gridding leaves it zero), and gridding skips the
// tiles whose (y, z) columns all lie in such groups.  Worth it for the sparse
// line sets of respiratory bins; with mode 1 it is used only when it skips at
// least a tenth of the groups.
int build_prune(cgs_ctx* h) {
    const long long mt = h->bin_off.back();
    const int ny = h->N[1], nz = h->N[2], C = h->C;
    const long long nyz = (long long)ny * nz;
    int W = 0, cpb = 0;
    plain_geometry(h->N[0], h->xp, &W, &cpb);
    const long long gw = (long long)W * cpb;
    const long long ngrp = (nyz * C + gw - 1) / gw;
    const Tiles& t = h->tiles;
    const int kw = h->kb.kw;
    std::vector<int4> base(std::max(1LL, mt));
    if (mt) CK(hipMemcpy(base.data(), h->p_base, sizeof(int4) * mt, hipMemcpyDeviceToHost));
    std::vector<uint8_t> gact((size_t)h->nbins * ngrp, 0), tneed((size_t)h->nbins * t.nty * t.ntz, 0);
    std::vector<std::vector<int>> lists(h->nbins);
    std::vector<uint8_t> cell(nyz);
    long long active = 0;
    for (int b = 0; b < h->nbins; ++b) {
        std::fill(cell.begin(), cell.end(), 0);
        for (long long s = h->bin_off[b]; s < h->bin_off[b + 1]; ++s)
            for (int oy = 0; oy < kw; ++oy) {
                const long long y = std::min(std::max(base[s].y + oy, 0), ny - 1);
                for (int oz = 0; oz < kw; ++oz) cell[y * nz + std::min(std::max(base[s].z + oz, 0), nz - 1)] = 1;
            }
        uint8_t* ga = gact.data() + (size_t)b * ngrp;
        for (long long c = 0; c < nyz; ++c)
            if (cell[c])
                for (long long g = c * C / gw; g <= (c * C + C - 1) / gw; ++g) ga[g] = 1;
        uint8_t* tn = tneed.data() + (size_t)b * t.nty * t.ntz;
        for (long long g = 0; g < ngrp; ++g) {
            if (!ga[g]) continue;
            lists[b].push_back((int)g);
            const long long c1 = std::min((g * gw + gw - 1) / C, nyz - 1);
            for (long long c = g * gw / C; c <= c1; ++c) tn[(c / nz / t.ty) * t.ntz + (c % nz) / t.tz] = 1;
        }
        active += (long long)lists[b].size();
    }
    h->prune_on = h->prune_mode == 2 || (h->prune_mode == 1 && active * 10 <= 9 * ngrp * h->nbins);
    h->pr_gfrac.assign(h->nbins, 1.0);
    h->pr_tfrac.assign(h->nbins, 1.0);
    if (!h->prune_on) return CGS_OK;
    for (int b = 0; b < h->nbins; ++b) {
        const uint8_t* tn = tneed.data() + (size_t)b * t.nty * t.ntz;
        h->pr_gfrac[b] = (double)lists[b].size() / (double)ngrp;
        h->pr_tfrac[b] = (double)std::count(tn, tn + t.nty * t.ntz, (uint8_t)1) / (double)(t.nty * t.ntz);
    }
    long long stride = 1;
    for (auto& l : lists) stride = std::max<long long>(stride, (long long)l.size());
    std::vector<int> glist((size_t)h->nbins * stride, 0);
    h->h_gcnt.assign(h->nbins, 0);
    for (int b = 0; b < h->nbins; ++b) {
        std::copy(lists[b].begin(), lists[b].end(), glist.begin() + (size_t)b * stride);
        h->h_gcnt[b] = (int)lists[b].size();
    }
    h->pr_ngrp = ngrp;
    h->pr_gstride = stride;
    CKS(dalloc(h, &h->d_gact, h->nbins * ngrp));
    CKS(dalloc(h, &h->d_glist, h->nbins * stride));
    CKS(dalloc(h, &h->d_gcnt, h->nbins));
    CKS(dalloc(h, &h->d_tneed, (long long)h->nbins * t.nty * t.ntz));
    CK(hipMemcpy(h->d_gact, gact.data(), gact.size(), hipMemcpyHostToDevice));
    CK(hipMemcpy(h->d_glist, glist.data(), sizeof(int) * glist.size(), hipMemcpyHostToDevice));
    CK(hipMemcpy(h->d_gcnt, h->h_gcnt.data(), sizeof(int) * h->nbins, hipMemcpyHostToDevice));
    CK(hipMemcpy(h->d_tneed, tneed.data(), tneed.size(), hipMemcpyHostToDevice));
    return CGS_OK;
}

// Gridding candidate lists of every tile of the packed bins: the samples
// whose KB window meets the tile, in the order gridding's own scan of the 27
// neighbouring tiles finds them.  The trajectory is fixed across CG
// iterations, so the scan runs once here instead of in every launch.
GridArgs grid_args(cgs_ctx* h, int b0, int nb);
int build_grid_lists(cgs_ctx* h) {
    const long long nt = (long long)h->nbins * h->tiles.ntiles;
    GridArgs g = grid_args(h, 0, h->nbins);
    g.tlist = nullptr;
    g.tlist_off = nullptr;
    int* cnt = nullptr;
    CKS(dalloc(h, &cnt, nt));
    CK(launch_grid_lists(h->kb.kw, h->tcfg, g, cnt, h->st));
    std::vector<int> hc(nt);
    CK(hipMemcpyAsync(hc.data(), cnt, sizeof(int) * nt, hipMemcpyDeviceToHost, h->st));
    CK(hipStreamSynchronize(h->st));
    dfree(h, cnt, nt);
    std::vector<long long> off(nt + 1, 0);
    for (long long t = 0; t < nt; ++t) off[t + 1] = off[t] + hc[t];
    h->tlist_n = off[nt];
    CKS(dalloc(h, &h->d_tlist_off, nt + 1));
    CKS(dalloc(h, &h->d_tlist, std::max(1LL, h->tlist_n)));
    CK(hipMemcpy(h->d_tlist_off, off.data(), sizeof(long long) * (nt + 1), hipMemcpyHostToDevice));
    g.tlist = h->d_tlist;
    g.tlist_off = h->d_tlist_off;
    CK(launch_grid_lists(h->kb.kw, h->tcfg, g, nullptr, h->st));
    CK(hipStreamSynchronize(h->st));
    return CGS_OK;
}

// Per-(y, z)-cell sample lists of the fused gridding (cgs_xlists.hip), for
// every packed bin; only where the inverse x pass takes 16-coil chunks of a
// cell (xgrid_supported) and the bins are sparse enough to gain: the fused
// pass stages each sample once per (y, z) of its window and 16-coil group
// (16 values and a 32-byte record: about (8 + 2) * C bytes per (y, z)),
// where the separate pass writes and reads the coil grids (16 * C * Ng
// bytes);
// the fused pass is used when the densest bin stages at most a quarter of
// that (respiratory bins of a few percent of the lines stage a twentieth to
// a fifth; a single frame with all lines about three quarters, which
// measured slower fused).
int build_xlists(cgs_ctx* h) {
    h->xl_on = false;
    if (!h->xgrid || !xgrid_supported(h->N[0], h->C, h->xp) || !xlist_supported(h->N[0], h->N[1], h->N[2]))
        return CGS_OK;
    const int kw = h->kb.kw;
    const long long ncell = (long long)h->N[1] * h->N[2];
    long long total = 0, mmax = 0;
    for (int b = 0; b < h->nbins; ++b) {
        const long long n = (h->bin_off[b + 1] - h->bin_off[b]) * kw * kw;
        total += n;
        mmax = std::max(mmax, n);
    }
    const double staged = (double)mmax * 10.0 * h->C, round_trip = 16.0 * h->C * (double)h->ng;
    if (staged > 0.25 * round_trip || mmax >= (1LL << 31) - 256) return CGS_OK;   // (the sort takes int counts)
    const int rs = xrec_stride(kw);
    unsigned *ktmp = nullptr, *kout = nullptr;
    int *itmp = nullptr, *iout = nullptr;
    char* stmp = nullptr;
    const size_t sb = std::max<size_t>(1, xlist_sort_bytes(mmax));
    auto free_work = [&]() {
        dfree(h, ktmp, std::max(1LL, mmax));
        dfree(h, kout, std::max(1LL, mmax));
        dfree(h, itmp, std::max(1LL, mmax));
        dfree(h, iout, std::max(1LL, mmax));
        dfree(h, stmp, sb);
    };
    // the lists only speed up E^H: where their memory (about 32 bytes per
    // sample tap) is not available, the separate gridding pass is used
    h->xl_n = total;
    if (dalloc(h, &h->d_xrec, std::max(1LL, total * rs)) != CGS_OK ||
        dalloc(h, &h->d_xoff, h->nbins * (ncell + 1)) != CGS_OK || dalloc(h, &ktmp, std::max(1LL, mmax)) != CGS_OK ||
        dalloc(h, &kout, std::max(1LL, mmax)) != CGS_OK || dalloc(h, &itmp, std::max(1LL, mmax)) != CGS_OK ||
        dalloc(h, &iout, std::max(1LL, mmax)) != CGS_OK || dalloc(h, &stmp, sb) != CGS_OK) {
        free_work();
        dfree(h, h->d_xrec, std::max(1LL, total * rs));
        dfree(h, h->d_xoff, h->nbins * (ncell + 1));
        h->xl_n = 0;
        h->err.clear();
        return CGS_OK;
    }
    long long g0 = 0;
    for (int b = 0; b < h->nbins; ++b) {
        const long long m = h->bin_off[b + 1] - h->bin_off[b];
        hipError_t e = build_xlists_bin(kw, h->p_base, h->p_vals, h->bin_off[b], m, h->N[1], h->N[2], ktmp, itmp, stmp,
                                        sb, kout, iout, h->d_xrec + g0 * rs, h->d_xoff + b * (ncell + 1), g0, h->st);
        if (e == hipSuccess && b + 1 == h->nbins) e = hipStreamSynchronize(h->st);
        if (e != hipSuccess) {
            (void)hipStreamSynchronize(h->st);
            free_work();
            return fail_hip(h, "build_xlists_bin", e);
        }
        g0 += m * kw * kw;
    }
    if (h->nbins == 0) CK(hipStreamSynchronize(h->st));
    free_work();
    h->xl_on = true;
    return CGS_OK;
}

// Binning + stable tile sort + KB tables of the active samples among traj[ns]
// (coil-0 value of d_data0 != 0 when given, else d_mask != 0 when given, else
// all).  Allocates *perm, *tile_off, *base, *vals.
int sort_samples(cgs_ctx* h, const double* traj, long long ns, const float2* d_data0, const uint8_t* d_mask,
                 long long* m_out, int** perm, int** tile_off, int4** base, double** vals) {
    if (!h->kb_ready) CKS(setup_kb(h));
    const int nt = h->tiles.ntiles;
    int *key = nullptr, *counts = nullptr, *tmp = nullptr, *cursor = nullptr;
    CKS(dalloc(h, &key, std::max(1LL, ns)));
    CKS(dalloc(h, &counts, nt));
    CKS(dalloc(h, &cursor, nt));
    CKS(dalloc(h, tile_off, nt + 1));
    CK(hipMemsetAsync(counts, 0, sizeof(int) * nt, h->st));
    CK(hipMemsetAsync(cursor, 0, sizeof(int) * nt, h->st));
    if (ns > 0) CK(launch_bin_keys(traj, ns, d_data0, d_mask, geom(h), h->tiles, key, nullptr, counts, h->st));
    CK(launch_scan_offsets(counts, nt, *tile_off, h->st));
    int m = 0;
    CK(hipMemcpyAsync(&m, *tile_off + nt, sizeof(int), hipMemcpyDeviceToHost, h->st));
    CK(hipStreamSynchronize(h->st));
    *m_out = m;
    const int vs = val_stride(h->kb.kw);
    CKS(dalloc(h, &tmp, std::max(1, m)));
    CKS(dalloc(h, perm, m));
    CKS(dalloc(h, base, m));
    CKS(dalloc(h, vals, (size_t)m * vs));
    if (m > 0) {
        CK(launch_scatter(key, ns, *tile_off, cursor, tmp, h->st));
        CK(launch_segment_rank(tmp, *tile_off, nt, *perm, h->st));
        CK(launch_sample_tables(traj, *perm, m, geom(h), h->kb, *base, *vals, h->st));
    }
    CK(hipStreamSynchronize(h->st));
    dfree(h, key, std::max(1LL, ns));
    dfree(h, counts, nt);
    dfree(h, cursor, nt);
    dfree(h, tmp, std::max(1, m));
    return CGS_OK;
}

// Binning + stable tile sort + KB tables of one bin.
int build_bin(cgs_ctx* h, int bin, const float2* d_data0, const uint8_t* d_mask) {
    if (!h->d_traj) return fail(h, CGS_ERR_STATE, "no trajectory set");
    if (!h->kb_ready) CKS(setup_kb(h));
    free_packed(h);
    BinData& B = h->bins[bin];
    free_bin(h, B);
    CKS(sort_samples(h, h->d_traj, h->ns, d_data0, d_mask, &B.m, &B.perm, &B.tile_off, &B.base, &B.vals));
    B.built = true;
    return CGS_OK;
}

// Upload zero-filled [C][nS] data, gather the active samples into B.y.
int gather_bin_data(cgs_ctx* h, int bin, const float2* d_data) {
    BinData& B = h->bins[bin];
    dfree(h, B.y, B.m * h->C);
    CKS(dalloc(h, &B.y, (size_t)B.m * h->C));
    CK(launch_gather_data(d_data, h->ns, B.base, B.m, h->C, nullptr, nullptr, nullptr, B.y, h->st));
    B.has_y = true;
    return CGS_OK;
}

// Concatenate the bins' sample arrays (cached until a bin changes).
int pack_bins(cgs_ctx* h) {
    if (h->packed) return CGS_OK;
    for (int b = 0; b < h->nbins; ++b)
        if (!h->bins[b].built) return fail(h, CGS_ERR_STATE, "bin " + std::to_string(b) + " has no samples set");
    free_packed(h);
    const int nt = h->tiles.ntiles;
    const int vs = val_stride(h->kb.kw);
    h->bin_off.assign(h->nbins + 1, 0);
    for (int b = 0; b < h->nbins; ++b) h->bin_off[b + 1] = h->bin_off[b] + h->bins[b].m;
    const long long mt = h->bin_off.back();
    const long long mt1 = std::max(1LL, mt);   // a bin may have no samples
    bool all_y = true;
    for (int b = 0; b < h->nbins; ++b) all_y = all_y && h->bins[b].has_y;
    if (h->nbins == 1 && mt > 0) {
        h->p_alias = true;
        h->p_base = h->bins[0].base;
        h->p_vals = h->bins[0].vals;
        h->p_y = h->bins[0].y;
        h->p_tile_off = h->bins[0].tile_off;
    } else {
        CKS(dalloc(h, &h->p_base, mt1));
        CKS(dalloc(h, &h->p_vals, (size_t)mt1 * vs));
        if (all_y) CKS(dalloc(h, &h->p_y, (size_t)mt1 * h->C));
        CKS(dalloc(h, &h->p_tile_off, (size_t)h->nbins * (nt + 1)));
        std::vector<int> off(nt + 1);
        for (int b = 0; b < h->nbins; ++b) {
            const BinData& B = h->bins[b];
            const long long o = h->bin_off[b];
            if (B.m) {
                CK(hipMemcpyAsync(h->p_base + o, B.base, sizeof(int4) * B.m, hipMemcpyDeviceToDevice, h->st));
                CK(hipMemcpyAsync(h->p_vals + o * vs, B.vals, sizeof(double) * B.m * vs, hipMemcpyDeviceToDevice,
                                  h->st));
                if (all_y)
                    CK(hipMemcpyAsync(h->p_y + o * h->C, B.y, sizeof(float2) * B.m * h->C, hipMemcpyDeviceToDevice,
                                      h->st));
            }
            CK(hipMemcpy(off.data(), B.tile_off, sizeof(int) * (nt + 1), hipMemcpyDeviceToHost));
            for (auto& v : off) v += (int)o;
            CK(hipMemcpy(h->p_tile_off + (size_t)b * (nt + 1), off.data(), sizeof(int) * (nt + 1),
                         hipMemcpyHostToDevice));
        }
    }
    // bin index of every sample
    std::vector<int> sb(std::max(1LL, mt));
    for (int b = 0; b < h->nbins; ++b)
        for (long long i = h->bin_off[b]; i < h->bin_off[b + 1]; ++i) sb[i] = b;
    CKS(dalloc(h, &h->p_sbin, std::max(1LL, mt)));
    CK(hipMemcpy(h->p_sbin, sb.data(), sizeof(int) * std::max(1LL, mt), hipMemcpyHostToDevice));
    // density weights in sorted order (a bin's own weights, else the context's)
    bool any_dcf = h->d_dcf_full != nullptr;
    for (int b = 0; b < h->nbins; ++b) any_dcf = any_dcf || h->bins[b].dcf_full;
    if (any_dcf) {
        CKS(dalloc(h, &h->p_dcf, std::max(1LL, mt)));
        for (int b = 0; b < h->nbins; ++b) {
            const BinData& B = h->bins[b];
            const float* src = B.dcf_full ? B.dcf_full : h->d_dcf_full;
            if (!src) {
                std::vector<float> ones(B.m, 1.0f);
                if (B.m)
                    CK(hipMemcpy(h->p_dcf + h->bin_off[b], ones.data(), sizeof(float) * B.m, hipMemcpyHostToDevice));
                continue;
            }
            CK(launch_gather_data(nullptr, h->ns, h->p_base + h->bin_off[b], B.m, 1, src, nullptr,
                                  h->p_dcf + h->bin_off[b], nullptr, h->st));
        }
    }
    CKS(dalloc(h, &h->p_yw, std::max(1LL, mt) * h->C));
    CK(hipStreamSynchronize(h->st));
    CKS(build_prune(h));
    if (h->grid_lists) CKS(build_grid_lists(h));
    CKS(build_xlists(h));
    h->packed = true;
    h->cg_ready = false;
    return CGS_OK;
}

// --------------------------------------------------------------- workspace
int ensure_grid(cgs_ctx* h, int nb) {
    if (h->grid_cap >= nb) return CGS_OK;
    dfree(h, h->d_grid, (size_t)h->grid_cap * h->bp);
    h->grid_cap = 0;
    CKS(dalloc(h, &h->d_grid, (size_t)nb * h->bp));
    h->grid_cap = nb;
    drop_graph(h);
    return CGS_OK;
}

int ensure_vectors(cgs_ctx* h) {
    const int nb = h->nbins;
    if (h->vec_bins == nb) return CGS_OK;
    const size_t n = (size_t)h->vec_bins * h->ng;
    dfree(h, h->d_x, n); dfree(h, h->d_r, n); dfree(h, h->d_p, n); dfree(h, h->d_q, n);
    dfree(h, h->d_part, (size_t)h->vec_bins * h->N[0] * h->N[1] * 2);
    dfree(h, h->d_part2, (size_t)h->vec_bins * h->nblk_upd * 2);
    dfree(h, h->d_sc, h->vec_bins);
    dfree(h, h->d_done, h->vec_bins);
    h->vec_bins = 0;
    const size_t m = (size_t)nb * h->ng;
    CKS(dalloc(h, &h->d_x, m)); CKS(dalloc(h, &h->d_r, m)); CKS(dalloc(h, &h->d_p, m)); CKS(dalloc(h, &h->d_q, m));
    h->nblk_upd = (int)std::min<long long>(1024, (h->ng + 2047) / 2048);
    CKS(dalloc(h, &h->d_part, (size_t)nb * h->N[0] * h->N[1] * 2));
    CKS(dalloc(h, &h->d_part2, (size_t)nb * h->nblk_upd * 2));
    CKS(dalloc(h, &h->d_sc, nb));
    CKS(dalloc(h, &h->d_done, nb));
    CK(hipMemset(h->d_done, 0, sizeof(int) * nb));
    h->vec_bins = nb;
    drop_graph(h);
    return CGS_OK;
}

// -------------------------------------------------------------- operators
// Forward z pass (+ y, x passes): coil grids of bins b0..b0+nb-1 hold the
// centred FFT of (p * S_c or the grids themselves) * kft.
// bins: the grids are those of packed bins b0.. (x-pass column pruning applies).
int run_fwd_fft(cgs_ctx* h, int b0, int nb, const float2* p, bool deapod, const int* done, bool bins = true) {
    const int W = zpass_w(h->N[2], h->C);
    ZArgs z{};
    z.g = h->d_grid; z.tw = h->d_tw[2]; z.p = p; z.maps = p ? h->d_maps : nullptr;
    z.kft = deapod ? h->d_kft : nullptr; z.C = h->C; z.ng = h->ng; z.nxy = (long long)h->N[0] * h->N[1];
    z.xp = h->xp; z.bp = h->bp; z.ny = h->N[1]; z.done = done; z.nb = nb; z.bpw = h->zbpw;
    {
        Prof t(h, "fft_z_fwd");
        CK(launch_fft_zpre(h->N[2], W, z, (long long)nb * z.nxy, h->st));
    }
    PlainArgs y{};
    y.g = h->d_grid; y.os = h->xp; y.ls = (long long)h->N[2] * h->C;
    y.ncols = y.ls; y.tw = h->d_tw[1]; y.outer_per_bin = h->N[0]; y.done = done;
    {
        Prof t(h, "fft_y_fwd");
        CK(launch_fft_plain(h->N[1], -1, y, (long long)nb * h->N[0], h->st));
    }
    PlainArgs x{};
    x.g = h->d_grid; x.os = h->bp; x.ls = h->xp; x.ncols = (long long)h->N[1] * h->N[2] * h->C;
    x.tw = h->d_tw[0]; x.outer_per_bin = 1; x.done = done;
    if (bins && h->prune_on) {
        x.glist = h->d_glist + (size_t)b0 * h->pr_gstride;
        x.gcnt = h->d_gcnt + b0;
        x.gstride = h->pr_gstride;
    }
    {
        Prof t(h, "fft_x_fwd");
        CK(launch_fft_plain(h->N[0], -1, x, nb, h->st));
    }
    return CGS_OK;
}

// Inverse x, y passes and the fused inverse z pass.  With q != null the
// coils are combined with conj(S) into q (and <p, q> partials when p and part
// are given); otherwise the grids receive the coil images.
// With xg_y != null the coil grids are not read: the inverse x pass forms
// them from the samples (gridding fused in, k_fft_xgrid, xg_y the sample
// values, xg_dcf optional weights; the packed bins' cell lists must exist).
int run_inv_fft(cgs_ctx* h, int b0, int nb, float2* q, const float2* p, double* part, const int* done,
                bool bins = true, const float2* xg_y = nullptr, const float* xg_dcf = nullptr) {
    PlainArgs x{};
    x.g = h->d_grid; x.os = h->bp; x.ls = h->xp; x.ncols = (long long)h->N[1] * h->N[2] * h->C;
    x.tw = h->d_tw[0]; x.outer_per_bin = 1; x.done = done;
    if (bins && h->prune_on) x.gact = h->d_gact + (size_t)b0 * h->pr_ngrp;
    if (xg_y) {
        XGridArgs xa{};
        xa.x = x;
        xa.Y = xg_y;
        xa.gdcf = xg_dcf;
        xa.rec = h->d_xrec;
        xa.off = h->d_xoff;
        xa.C = h->C; xa.ny = h->N[1]; xa.nz = h->N[2];
        xa.thrf = h->kb.thr_up;
        xa.b0 = b0;
        Prof t(h, "fft_x_inv_grid");
        CK(launch_fft_xgrid(h->N[0], h->kb.kw, xa, nb, h->st));
    } else {
        Prof t(h, "fft_x_inv");
        CK(launch_fft_plain(h->N[0], +1, x, nb, h->st));
    }
    PlainArgs y{};
    y.g = h->d_grid; y.os = h->xp; y.ls = (long long)h->N[2] * h->C;
    y.ncols = y.ls; y.tw = h->d_tw[1]; y.outer_per_bin = h->N[0]; y.done = done;
    {
        Prof t(h, "fft_y_inv");
        CK(launch_fft_plain(h->N[1], +1, y, (long long)nb * h->N[0], h->st));
    }
    const int W = zpass_w(h->N[2], h->C);
    ZArgs z{};
    z.g = h->d_grid; z.tw = h->d_tw[2]; z.p = p; z.maps = q ? h->d_maps : nullptr; z.kft = h->d_kft;
    z.q = q; z.part = part;
    z.scale = (float)(1e-3 / std::sqrt((double)h->ng));
    z.C = h->C; z.ng = h->ng; z.nxy = (long long)h->N[0] * h->N[1]; z.ny = h->N[1]; z.done = done;
    z.xp = h->xp; z.bp = h->bp; z.nb = nb; z.bpw = h->zbpw;
    bool crop = false;
    for (int a = 0; a < 3; ++a) {
        if (h->imgRange[a] != h->N[a]) crop = true;
        z.lo[a] = (h->N[a] - h->imgRange[a]) / 2 + h->offc[a];
        z.hi[a] = h->imgRange[a] + z.lo[a];
    }
    z.crop = crop;
    {
        Prof t(h, "fft_z_inv");
        CK(launch_fft_zpost(h->N[2], W, z, (long long)nb * z.nxy, h->st));
    }
    return CGS_OK;
}

GridArgs grid_args(cgs_ctx* h, int b0, int nb) {
    GridArgs g{};
    g.G = h->d_grid;
    g.s.base = h->p_base;
    g.s.vals = h->p_vals;
    g.s.dcf = nullptr;
    g.s.sbin = h->p_sbin;
    g.s.first = h->bin_off[b0];
    g.s.m = h->bin_off[b0 + nb] - h->bin_off[b0];
    g.g = geom(h);
    g.t = h->tiles;
    g.tile_off = h->p_tile_off;
    g.gdcf = nullptr;
    g.thr = h->kb.thr;
    g.thrf = h->kb.thr_up;
    g.b0 = b0;
    g.nb = nb;
    g.done = nullptr;
    g.tneed = h->prune_on ? h->d_tneed : nullptr;
    g.tlist = h->d_tlist;
    g.tlist_off = h->d_tlist_off;
    return g;
}

// q = A p for bins b0..b0+nb-1 (vectors indexed by global bin).
int apply_normal_bins(cgs_ctx* h, int b0, int nb, bool use_done) {
    const int* done = use_done ? h->d_done + b0 : nullptr;
    const float2* p = h->d_p + (size_t)b0 * h->ng;
    CKS(run_fwd_fft(h, b0, nb, p, true, done));
    GridArgs g = grid_args(h, b0, nb);
    g.s.dcf = h->p_dcf;
    g.Y = h->p_yw;
    g.done = use_done ? h->d_done : nullptr;
    {
        Prof t(h, "degrid");
        CK(launch_degrid(h->kb.kw, g, h->st));
    }
    if (h->xl_on) {   // gridding fused into the inverse x pass
        CKS(run_inv_fft(h, b0, nb, h->d_q + (size_t)b0 * h->ng, p, h->d_part, done, true, h->p_yw));
        return CGS_OK;
    }
    {
        Prof t(h, "grid");
        CK(launch_grid(h->kb.kw, h->tcfg, g, h->st));
    }
    CKS(run_inv_fft(h, b0, nb, h->d_q + (size_t)b0 * h->ng, p, h->d_part, done));
    return CGS_OK;
}

int cg_step_bins(cgs_ctx* h, int b0, int nb) {
    const size_t o = (size_t)b0 * h->ng;
    const int nxy = h->N[0] * h->N[1];
    {
        Prof t(h, "cg_alpha");
        CK(launch_cg_alpha(h->d_sc + b0, h->d_part, nxy, nb, h->d_q + o, h->ng, h->d_done + b0, h->st));
    }
    {
        Prof t(h, "cg_update_xr");
        CK(launch_cg_update_xr(h->d_x + o, h->d_r + o, h->d_p + o, h->d_q + o, h->ng, nb, h->nblk_upd,
                               h->d_sc + b0, h->d_part2, h->st));
    }
    {
        Prof t(h, "cg_beta");
        CK(launch_cg_beta(h->d_sc + b0, h->d_part2, h->nblk_upd, nb, h->thresh, h->max_iter,
                          h->d_hist + (size_t)b0 * h->hist_stride, h->hist_stride, h->d_done + b0, h->st));
    }
    {
        Prof t(h, "cg_update_p");
        CK(launch_cg_update_p(h->d_p + o, h->d_r + o, h->ng, nb, h->d_sc + b0, h->st));
    }
    return CGS_OK;
}

int check_ready(cgs_ctx* h, bool need_maps) {
    if (!h->kft_ready) return fail(h, CGS_ERR_STATE, "kernelFT not computed (cgs_calculate_kernel_ft)");
    if (need_maps && !h->maps_ready) return fail(h, CGS_ERR_STATE, "sensitivity maps not set");
    return CGS_OK;
}

int choose_group(cgs_ctx* h) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) fr = 0;
    const double per = (double)h->ng * h->C * sizeof(float2);
    const double have = (double)fr + (double)h->grid_cap * per - 2.0e9;   // keep 2 GB headroom
    int g = (int)std::floor(have / per);
    g = std::max(1, std::min(g, h->nbins));
    if (h->group_cap > 0) g = std::min(g, h->group_cap);
    // equal groups
    const int ngroups = (h->nbins + g - 1) / g;
    h->group = (h->nbins + ngroups - 1) / ngroups;
    return CGS_OK;
}

int upload_complex(cgs_ctx* h, float2** d, const float* src, size_t n) {
    CKS(dalloc(h, d, n));
    CK(hipMemcpy(*d, src, sizeof(float2) * n, hipMemcpyHostToDevice));
    return CGS_OK;
}

}  // namespace

// ============================================================================
extern "C" {

const char* cgs_last_error(const cgs_ctx* h) { return h ? h->err.c_str() : "null context"; }

int cgs_create(const int grid[3], const int samp[3], int ncoil, cgs_ctx** out) {
    if (!out) return CGS_ERR_ARG;
    *out = nullptr;
    for (int a = 0; a < 3; ++a)
        if (grid[a] < 2 || samp[a] < 1) return CGS_ERR_ARG;
    if (ncoil < 1 || ncoil > CGS_MAX_COILS) return CGS_ERR_ARG;
    for (int a = 0; a < 3; ++a)
        if (!fft_len_supported(grid[a])) return CGS_ERR_ARG;
    cgs_ctx* h = new cgs_ctx();
    for (int a = 0; a < 3; ++a) {
        h->N[a] = grid[a];
        h->S[a] = samp[a];
        h->imgRange[a] = grid[a];
        h->offc[a] = 0;
    }
    h->ng = (long long)grid[0] * grid[1] * grid[2];
    h->ns = (long long)samp[0] * samp[1] * samp[2];
    h->C = ncoil;
    h->xp = coil_grid_pitch(grid[1], grid[2], ncoil);
    h->bp = h->xp * grid[0];
    h->bins.resize(1);
    if (hipStreamCreate(&h->own_st) != hipSuccess) {
        delete h;
        return CGS_ERR_HIP;
    }
    h->st = h->own_st;
    if (upload_twiddles(h, h->N, h->d_tw) != CGS_OK) {
        cgs_destroy(h);
        return CGS_ERR_HIP;
    }
    *out = h;
    return CGS_OK;
}

int cgs_destroy(cgs_ctx* h) {
    if (!h) return CGS_OK;
    if (h->st) (void)hipStreamSynchronize(h->st);
    prof_flush(h);
    free_packed(h);
    for (auto& b : h->bins) {
        free_bin(h, b);
        dfree(h, b.dcf_full, h->ns);
    }
    dfree(h, h->d_traj, h->ns * 3);
    dfree(h, h->d_dcf_full, h->ns);
    dfree(h, h->d_kft, h->ng);
    for (int a = 0; a < 3; ++a) dfree(h, h->d_tw[a], h->N[a]);
    dfree(h, h->d_maps, h->ng * h->C);
    dfree(h, h->d_grid, (size_t)h->grid_cap * h->bp);
    const size_t n = (size_t)h->vec_bins * h->ng;
    dfree(h, h->d_x, n); dfree(h, h->d_r, n); dfree(h, h->d_p, n); dfree(h, h->d_q, n);
    dfree(h, h->d_part, 1); dfree(h, h->d_part2, 1); dfree(h, h->d_sc, 1); dfree(h, h->d_done, 1);
    dfree(h, h->d_hist, 1);
    if (h->own_st) (void)hipStreamDestroy(h->own_st);
    delete h;
    return CGS_OK;
}

int cgs_set_kernel(cgs_ctx* h, float w, float b, int order) {
    if (!h) return CGS_ERR_ARG;
    if (!(w > 0.0f) || order < 1) return fail(h, CGS_ERR_ARG, "invalid kernel parameters");
    if (h->kb_ready && w == h->w && b == h->b && order == h->order) return CGS_OK;   // nothing changes
    h->w = w; h->b = b; h->order = order;
    h->kb_ready = false;
    h->kft_ready = false;
    for (auto& B : h->bins) B.built = false;
    free_packed(h);
    return setup_kb(h);
}

int cgs_set_img_range(cgs_ctx* h, const int range[3], const int offcenter[3]) {
    if (!h) return CGS_ERR_ARG;
    for (int a = 0; a < 3; ++a) {
        if (range[a] < 0 || range[a] > h->N[a]) return fail(h, CGS_ERR_ARG, "image range outside the grid");
        h->imgRange[a] = range[a];
        h->offc[a] = offcenter ? offcenter[a] : 0;
    }
    // the crop box is baked into the right-hand sides and into the captured
    // iteration: both are formed again by the next cgs_cg_prepare
    drop_graph(h);
    h->cg_ready = false;
    return CGS_OK;
}

int cgs_num_tile_configs(void) { return grid_tile_configs(); }

int cgs_set_tile_config(cgs_ctx* h, int tcfg) {
    if (!h) return CGS_ERR_ARG;
    if (tcfg < 0 || tcfg >= grid_tile_configs()) return fail(h, CGS_ERR_ARG, "tile config out of range");
    if (tcfg == h->tcfg && h->kb_ready) return CGS_OK;
    h->tcfg = tcfg;
    h->kb_ready = false;
    for (auto& B : h->bins) B.built = false;
    free_packed(h);
    return setup_kb(h);
}

int cgs_get_tile_config(const cgs_ctx* h, int* tcfg, int dims[3]) {
    if (!h) return CGS_ERR_ARG;
    if (tcfg) *tcfg = h->tcfg;
    if (dims) grid_tile_dims(h->tcfg, &dims[0], &dims[1], &dims[2]);
    return CGS_OK;
}

int cgs_set_trajectory(cgs_ctx* h, const double* ktraj) {
    if (!h || !ktraj) return CGS_ERR_ARG;
    if (!h->d_traj) CKS(dalloc(h, &h->d_traj, (size_t)h->ns * 3));
    CK(hipMemcpy(h->d_traj, ktraj, sizeof(double) * 3 * h->ns, hipMemcpyHostToDevice));
    for (auto& B : h->bins) B.built = false;
    free_packed(h);
    return CGS_OK;
}

int cgs_calculate_weights(cgs_ctx* h) {
    if (!h) return CGS_ERR_ARG;
    CKS(build_bin(h, 0, nullptr, nullptr));
    return CGS_OK;
}

int cgs_exclude_missing_data(cgs_ctx* h, const float* data) {
    if (!h || !data) return CGS_ERR_ARG;
    float2* d0 = nullptr;
    CKS(upload_complex(h, &d0, data, h->ns));   // coil 0 only
    const int st = build_bin(h, 0, d0, nullptr);
    dfree(h, d0, h->ns);
    return st;
}

int cgs_calculate_kernel_ft(cgs_ctx* h) {
    if (!h) return CGS_ERR_ARG;
    if (!h->kb_ready) CKS(setup_kb(h));
    // per-axis centred inverse DFT of the sampled kernel (Reco.cpp:571-589)
    std::vector<double> axr, axi;
    for (int a = 0; a < 3; ++a) {
        const int n = h->N[a];
        const double dk = 1.0 / (double)(n - 1);
        const int c = n / 2;
        std::vector<double> xs(n, 0.0);
        for (int m = 0; m < n; ++m) {
            const double d = std::sqrt(dk * dk * (double)(m - c) * (double)(m - c));
            xs[m] = host_kb_axis((float)d, h->kb.ax[a], h->kb.b);
        }
        const double two_pi = 6.283185307179586476925286766559;
        for (int k = 0; k < n; ++k) {
            double re = 0.0, im = 0.0;
            for (int m = 0; m < n; ++m) {
                if (xs[m] == 0.0) continue;
                const double sgn = (((m + k) & 1) ? -1.0 : 1.0);
                const double ang = two_pi * (double)(((long)m * k) % n) / (double)n;
                re += sgn * xs[m] * std::cos(ang);
                im += sgn * xs[m] * std::sin(ang);
            }
            axr.push_back(re);
            axi.push_back(im);
        }
    }
    double *dr = nullptr, *di = nullptr;
    const size_t na = axr.size();
    CKS(dalloc(h, &dr, na));
    CKS(dalloc(h, &di, na));
    CK(hipMemcpy(dr, axr.data(), sizeof(double) * na, hipMemcpyHostToDevice));
    CK(hipMemcpy(di, axi.data(), sizeof(double) * na, hipMemcpyHostToDevice));
    if (!h->d_kft) CKS(dalloc(h, &h->d_kft, h->ng));
    CK(launch_kernel_ft(dr, di, geom(h), h->d_kft, h->st));
    CK(hipStreamSynchronize(h->st));
    dfree(h, dr, na);
    dfree(h, di, na);
    h->kft_ready = true;
    return CGS_OK;
}

int cgs_get_kernel_ft(cgs_ctx* h, float* kft) {
    if (!h || !kft) return CGS_ERR_ARG;
    if (!h->kft_ready) return fail(h, CGS_ERR_STATE, "kernelFT not computed");
    CK(hipMemcpy(kft, h->d_kft, sizeof(float) * h->ng, hipMemcpyDeviceToHost));
    return CGS_OK;
}

int cgs_set_density(cgs_ctx* h, const float* dcf) {
    if (!h) return CGS_ERR_ARG;
    if (!dcf) {
        dfree(h, h->d_dcf_full, h->ns);
    } else {
        if (!h->d_dcf_full) CKS(dalloc(h, &h->d_dcf_full, h->ns));
        CK(hipMemcpy(h->d_dcf_full, dcf, sizeof(float) * h->ns, hipMemcpyHostToDevice));
    }
    free_packed(h);
    return CGS_OK;
}

int cgs_set_maps(cgs_ctx* h, const float* maps) {
    if (!h || !maps) return CGS_ERR_ARG;
    float2* tmp = nullptr;
    CKS(upload_complex(h, &tmp, maps, (size_t)h->C * h->ng));
    if (!h->d_maps) CKS(dalloc(h, &h->d_maps, (size_t)h->C * h->ng));
    CK(launch_transpose_coils(tmp, h->d_maps, h->ng, h->C, (long long)h->N[1] * h->N[2], (long long)h->N[1] * h->N[2] * h->C, 1, h->st));
    CK(hipStreamSynchronize(h->st));
    dfree(h, tmp, (size_t)h->C * h->ng);
    h->maps_ready = true;
    return CGS_OK;
}

int cgs_nufft_fwd(cgs_ctx* h, const float* img, float* data) {
    if (!h || !img || !data) return CGS_ERR_ARG;
    CKS(check_ready(h, false));
    CKS(pack_bins(h));
    CKS(ensure_grid(h, 1));
    float2* tmp = nullptr;
    CKS(upload_complex(h, &tmp, img, (size_t)h->C * h->ng));
    CK(launch_transpose_coils(tmp, h->d_grid, h->ng, h->C, (long long)h->N[1] * h->N[2], h->xp, 1, h->st));
    CKS(run_fwd_fft(h, 0, 1, nullptr, true, nullptr));
    GridArgs g = grid_args(h, 0, 1);
    g.Y = h->p_yw;
    CK(launch_degrid(h->kb.kw, g, h->st));
    // scatter back into the zero-filled [c][i] layout
    float2* dd = nullptr;
    CKS(dalloc(h, &dd, (size_t)h->C * h->ns));
    CK(hipMemsetAsync(dd, 0, sizeof(float2) * h->C * h->ns, h->st));
    CK(launch_scatter_data(h->p_yw, h->p_base, h->bin_off[1], h->C, h->ns, dd, h->st));
    CK(hipMemcpyAsync(data, dd, sizeof(float2) * h->C * h->ns, hipMemcpyDeviceToHost, h->st));
    CK(hipStreamSynchronize(h->st));
    dfree(h, dd, (size_t)h->C * h->ns);
    dfree(h, tmp, (size_t)h->C * h->ng);
    return CGS_OK;
}

int cgs_nufft_bwd(cgs_ctx* h, const float* data, float* img) {
    if (!h || !data || !img) return CGS_ERR_ARG;
    CKS(check_ready(h, false));
    CKS(pack_bins(h));
    CKS(ensure_grid(h, 1));
    float2* dd = nullptr;
    CKS(upload_complex(h, &dd, data, (size_t)h->C * h->ns));
    CK(launch_gather_data(dd, h->ns, h->p_base, h->bin_off[1], h->C, nullptr, nullptr, nullptr, h->p_yw, h->st));
    GridArgs g = grid_args(h, 0, 1);
    g.Y = h->p_yw;
    CK(launch_grid(h->kb.kw, h->tcfg, g, h->st));
    CKS(run_inv_fft(h, 0, 1, nullptr, nullptr, nullptr, nullptr));
    float2* out = nullptr;
    CKS(dalloc(h, &out, (size_t)h->C * h->ng));
    CK(launch_transpose_coils(h->d_grid, out, h->ng, h->C, (long long)h->N[1] * h->N[2], h->xp, 0, h->st));
    CK(hipMemcpyAsync(img, out, sizeof(float2) * h->C * h->ng, hipMemcpyDeviceToHost, h->st));
    CK(hipStreamSynchronize(h->st));
    dfree(h, out, (size_t)h->C * h->ng);
    dfree(h, dd, (size_t)h->C * h->ns);
    return CGS_OK;
}

int cgs_apply_normal(cgs_ctx* h, const float* p, float* q) {
    if (!h || !p || !q) return CGS_ERR_ARG;
    CKS(check_ready(h, true));
    CKS(pack_bins(h));
    CKS(ensure_grid(h, 1));
    CKS(ensure_vectors(h));
    CK(hipMemcpy(h->d_p, p, sizeof(float2) * h->ng, hipMemcpyHostToDevice));
    CKS(apply_normal_bins(h, 0, 1, false));
    CK(hipMemcpyAsync(q, h->d_q, sizeof(float2) * h->ng, hipMemcpyDeviceToHost, h->st));
    CK(hipStreamSynchronize(h->st));
    prof_flush(h);
    return CGS_OK;
}

int cgs_set_num_bins(cgs_ctx* h, int nbins) {
    if (!h || nbins < 1) return CGS_ERR_ARG;
    if (nbins == h->nbins) return CGS_OK;
    free_packed(h);
    for (int b = nbins; b < h->nbins; ++b) {
        free_bin(h, h->bins[b]);
        dfree(h, h->bins[b].dcf_full, h->ns);
    }
    h->bins.resize(nbins);
    h->nbins = nbins;
    h->cg_ready = false;
    return CGS_OK;
}

int cgs_set_pruning(cgs_ctx* h, int mode) {
    if (!h || mode < 0 || mode > 2) return CGS_ERR_ARG;
    if (mode == h->prune_mode) return CGS_OK;
    h->prune_mode = mode;
    free_packed(h);
    return CGS_OK;
}

int cgs_get_pruning(const cgs_ctx* h, int* active) {
    if (!h || !active) return CGS_ERR_ARG;
    *active = h->prune_on ? 1 : 0;
    return CGS_OK;
}

int cgs_get_prune_fractions(const cgs_ctx* h, int bin, double* groups, double* tiles) {
    if (!h || bin < 0 || bin >= h->nbins || !groups || !tiles) return CGS_ERR_ARG;
    if (!h->packed || (int)h->pr_gfrac.size() != h->nbins) return CGS_ERR_STATE;
    *groups = h->pr_gfrac[bin];
    *tiles = h->pr_tfrac[bin];
    return CGS_OK;
}

int cgs_set_grid_lists(cgs_ctx* h, int on) {
    if (!h || on < 0 || on > 1) return CGS_ERR_ARG;
    if (on == h->grid_lists) return CGS_OK;
    h->grid_lists = on;
    free_packed(h);
    return CGS_OK;
}

int cgs_set_fused_gridding(cgs_ctx* h, int on) {
    if (!h || on < 0 || on > 1) return CGS_ERR_ARG;
    if (on == h->xgrid) return CGS_OK;
    h->xgrid = on;
    free_packed(h);
    return CGS_OK;
}

int cgs_fused_gridding_active(cgs_ctx* h, int* on) {
    if (!h || !on) return CGS_ERR_ARG;
    *on = h->xl_on ? 1 : 0;
    return CGS_OK;
}

int cgs_set_zpass_bins(cgs_ctx* h, int bins) {
    if (!h || bins < 1 || bins > 64) return CGS_ERR_ARG;
    h->zbpw = bins;
    drop_graph(h);   // the captured iteration holds the z-pass launches
    return CGS_OK;
}

int cgs_set_bin_group(cgs_ctx* h, int max_bins) {
    if (!h || max_bins < 0) return CGS_ERR_ARG;
    h->group_cap = max_bins;
    h->cg_ready = false;
    return CGS_OK;
}

int cgs_get_bin_group(const cgs_ctx* h, int* bins_per_launch) {
    if (!h || !bins_per_launch) return CGS_ERR_ARG;
    *bins_per_launch = h->group;
    return CGS_OK;
}

int cgs_set_bin_data(cgs_ctx* h, int bin, const float* data, const float* dcf) {
    if (!h || !data || bin < 0 || bin >= h->nbins) return CGS_ERR_ARG;
    float2* dd = nullptr;
    CKS(upload_complex(h, &dd, data, (size_t)h->C * h->ns));
    int st = build_bin(h, bin, dd, nullptr);
    if (st == CGS_OK) st = gather_bin_data(h, bin, dd);
    BinData& B = h->bins[bin];
    if (st == CGS_OK) {
        dfree(h, B.dcf_full, h->ns);
        if (dcf) {
            st = dalloc(h, &B.dcf_full, h->ns);
            if (st == CGS_OK && hipMemcpy(B.dcf_full, dcf, sizeof(float) * h->ns, hipMemcpyHostToDevice) != hipSuccess)
                st = fail(h, CGS_ERR_HIP, "dcf upload failed");
        }
    }
    if (hipStreamSynchronize(h->st) != hipSuccess && st == CGS_OK) st = fail(h, CGS_ERR_HIP, "gather failed");
    dfree(h, dd, (size_t)h->C * h->ns);
    return st;
}

int cgs_set_bin_samples(cgs_ctx* h, int bin, long long m, const int32_t* idx, const float* y, const float* dcf) {
    if (!h || bin < 0 || bin >= h->nbins || m < 0 || (m > 0 && (!idx || !y))) return CGS_ERR_ARG;
    std::vector<uint8_t> mask(h->ns, 0);
    std::vector<long long> pos(h->ns, -1);
    for (long long j = 0; j < m; ++j) {
        if (idx[j] < 0 || idx[j] >= h->ns) return fail(h, CGS_ERR_ARG, "sample index out of range");
        if (mask[idx[j]]) return fail(h, CGS_ERR_ARG, "duplicate sample index");
        mask[idx[j]] = 1;
        pos[idx[j]] = j;
    }
    uint8_t* dm = nullptr;
    CKS(dalloc(h, &dm, h->ns));
    CK(hipMemcpy(dm, mask.data(), h->ns, hipMemcpyHostToDevice));
    int st = build_bin(h, bin, nullptr, dm);
    dfree(h, dm, h->ns);
    CKS(st);
    BinData& B = h->bins[bin];
    // sorted order on the host, then values and weights in that order
    std::vector<int> perm(B.m);
    if (B.m) CK(hipMemcpy(perm.data(), B.perm, sizeof(int) * B.m, hipMemcpyDeviceToHost));
    std::vector<float2> ys((size_t)B.m * h->C);
    for (long long s = 0; s < B.m; ++s) {
        const long long j = pos[perm[s]];
        for (int c = 0; c < h->C; ++c) ys[s * h->C + c] = make_float2(y[2 * (c * m + j)], y[2 * (c * m + j) + 1]);
    }
    dfree(h, B.y, B.m * h->C);
    CKS(dalloc(h, &B.y, (size_t)B.m * h->C));
    if (B.m) CK(hipMemcpy(B.y, ys.data(), sizeof(float2) * B.m * h->C, hipMemcpyHostToDevice));
    B.has_y = true;
    dfree(h, B.dcf_full, h->ns);
    if (dcf) {
        std::vector<float> full(h->ns, 0.0f);
        for (long long j = 0; j < m; ++j) full[idx[j]] = dcf[j];
        CKS(dalloc(h, &B.dcf_full, h->ns));
        CK(hipMemcpy(B.dcf_full, full.data(), sizeof(float) * h->ns, hipMemcpyHostToDevice));
    }
    return CGS_OK;
}

int cgs_cg_prepare(cgs_ctx* h, double thresh, int max_iter) {
    if (!h || max_iter < 0) return CGS_ERR_ARG;
    CKS(check_ready(h, true));
    CKS(pack_bins(h));
    for (int b = 0; b < h->nbins; ++b)
        if (!h->bins[b].has_y) return fail(h, CGS_ERR_STATE, "bin " + std::to_string(b) + " has no data");
    if (!h->p_y) return fail(h, CGS_ERR_STATE, "bin data not packed");
    CKS(ensure_vectors(h));
    CKS(choose_group(h));
    CKS(ensure_grid(h, h->group));
    h->thresh = thresh;
    h->max_iter = max_iter;
    if (!h->d_hist || h->hist_stride != max_iter + 1 || h->hist_bins != h->nbins) {
        dfree(h, h->d_hist, 1);
        h->hist_stride = max_iter + 1;
        h->hist_bins = h->nbins;
        CKS(dalloc(h, &h->d_hist, (size_t)h->nbins * h->hist_stride));
    }
    CK(hipMemsetAsync(h->d_hist, 0, sizeof(double) * h->nbins * h->hist_stride, h->st));
    CK(hipMemsetAsync(h->d_done, 0, sizeof(int) * h->nbins, h->st));
    for (int b0 = 0; b0 < h->nbins; b0 += h->group) {
        const int nb = std::min(h->group, h->nbins - b0);
        const size_t o = (size_t)b0 * h->ng;
        GridArgs g = grid_args(h, b0, nb);
        g.Y = h->p_y;
        g.gdcf = h->p_dcf;
        if (h->xl_on) {   // gridding fused into the inverse x pass, as in the iterations
            CKS(run_inv_fft(h, b0, nb, h->d_r + o, nullptr, nullptr, nullptr, true, h->p_y, h->p_dcf));
        } else {
            {
                Prof t(h, "grid");
                CK(launch_grid(h->kb.kw, h->tcfg, g, h->st));
            }
            CKS(run_inv_fft(h, b0, nb, h->d_r + o, nullptr, nullptr, nullptr));
        }
        CK(launch_copy_rhs(h->d_x + o, h->d_r + o, h->d_p + o, h->d_r + o, h->ng, nb, h->st));
        CK(launch_dot_partials(h->d_r + o, h->d_r + o, h->ng, nb, h->nblk_upd, h->d_part2, h->st));
        CK(launch_cg_init(h->d_sc + b0, h->d_part2, h->nblk_upd, nb, thresh, h->d_done + b0, h->st));
    }
    CK(hipStreamSynchronize(h->st));
    prof_flush(h);
    h->cg_ready = true;
    // The captured iteration bakes in the loop-control scalars and the history
    // buffer; keep it while those stay the same.
    if (h->gexec && (h->g_thresh != thresh || h->g_max_iter != max_iter || h->g_hist != h->d_hist ||
                     h->g_group != h->group))
        drop_graph(h);
    return CGS_OK;
}

static int iterate_all_groups(cgs_ctx* h) {
    for (int b0 = 0; b0 < h->nbins; b0 += h->group) {
        const int nb = std::min(h->group, h->nbins - b0);
        CKS(apply_normal_bins(h, b0, nb, true));
        CKS(cg_step_bins(h, b0, nb));
    }
    return CGS_OK;
}

int cgs_cg_iterate(cgs_ctx* h, int mode, int n_iter) {
    if (!h || n_iter < 0) return CGS_ERR_ARG;
    if (!h->cg_ready) return fail(h, CGS_ERR_STATE, "cgs_cg_prepare not called");
    if (mode == CGS_SOLVE_PLAIN) {
        // The reference's loop shape: one dataset at a time, the host reads the
        // residual after every iteration (Reco.cpp:230, 292-309).
        for (int b = 0; b < h->nbins; ++b) {
            for (int it = 0; it < n_iter; ++it) {
                CgScalars sc;
                CK(hipMemcpyAsync(&sc, h->d_sc + b, sizeof(sc), hipMemcpyDeviceToHost, h->st));
                CK(hipStreamSynchronize(h->st));
                if (sc.done) break;
                CKS(apply_normal_bins(h, b, 1, false));
                CKS(cg_step_bins(h, b, 1));
                prof_flush(h);
            }
        }
    } else if (mode == CGS_SOLVE_BATCHED) {
        if (h->prof) {
            for (int it = 0; it < n_iter; ++it) CKS(iterate_all_groups(h));
        } else {
            if (h->gexec && (h->g_thresh != h->thresh || h->g_max_iter != h->max_iter || h->g_hist != h->d_hist ||
                             h->g_group != h->group))
                drop_graph(h);
            if (!h->gexec) {
                CK(hipStreamBeginCapture(h->st, hipStreamCaptureModeThreadLocal));
                const int st = iterate_all_groups(h);
                hipGraph_t g = nullptr;
                const hipError_t e = hipStreamEndCapture(h->st, &g);
                if (st != CGS_OK) {
                    if (g) (void)hipGraphDestroy(g);
                    return st;
                }
                CK(e);
                h->graph = g;
                CK(hipGraphInstantiate(&h->gexec, g, nullptr, nullptr, 0));
                h->g_thresh = h->thresh;
                h->g_max_iter = h->max_iter;
                h->g_hist = h->d_hist;
                h->g_group = h->group;
            }
            for (int it = 0; it < n_iter; ++it) CK(hipGraphLaunch(h->gexec, h->st));
        }
    } else {
        return fail(h, CGS_ERR_ARG, "unknown solve mode");
    }
    CK(hipStreamSynchronize(h->st));
    prof_flush(h);
    return CGS_OK;
}

int cgs_cg_results(cgs_ctx* h, float* x_out, double* resid_out, int* iters_out) {
    if (!h) return CGS_ERR_ARG;
    if (!h->cg_ready) return fail(h, CGS_ERR_STATE, "no CG state");
    std::vector<CgScalars> sc(h->nbins);
    CK(hipMemcpy(sc.data(), h->d_sc, sizeof(CgScalars) * h->nbins, hipMemcpyDeviceToHost));
    if (x_out) CK(hipMemcpy(x_out, h->d_x, sizeof(float2) * h->ng * h->nbins, hipMemcpyDeviceToHost));
    if (resid_out)
        CK(hipMemcpy(resid_out, h->d_hist, sizeof(double) * h->nbins * h->hist_stride, hipMemcpyDeviceToHost));
    int nonfinite = 0;
    for (int b = 0; b < h->nbins; ++b) {
        if (iters_out) iters_out[b] = sc[b].iter + (sc[b].nonfinite ? 0 : 0);
        nonfinite |= sc[b].nonfinite;
    }
    if (nonfinite) return fail(h, CGS_ERR_NONFINITE, "NaN or INF value in image detected (q[0])");
    return CGS_OK;
}

int cgs_cg_history_length(const cgs_ctx* h, int* n) {
    if (!h || !n) return CGS_ERR_ARG;
    if (!h->cg_ready) return CGS_ERR_STATE;
    *n = h->hist_stride;
    return CGS_OK;
}

int cgs_cg_sense_bins(cgs_ctx* h, int mode, double thresh, int max_iter, float* x_out, double* resid_out,
                      int* iters_out) {
    CKS(cgs_cg_prepare(h, thresh, max_iter));
    CKS(cgs_cg_iterate(h, mode, max_iter + 1));
    return cgs_cg_results(h, x_out, resid_out, iters_out);
}

int cgs_cg_sense_images(cgs_ctx* h, const float* img, double thresh, int max_iter, float* x_out,
                        double* resid_out, int* iters_out) {
    if (!h || !img || max_iter < 0) return CGS_ERR_ARG;
    if (h->nbins != 1) return fail(h, CGS_ERR_STATE, "cgs_cg_sense_images works on a single dataset");
    CKS(check_ready(h, true));
    CKS(pack_bins(h));
    CKS(ensure_vectors(h));
    CKS(ensure_grid(h, 1));
    h->group = 1;
    h->thresh = thresh;
    h->max_iter = max_iter;
    // a captured batched iteration refers to the loop control and history
    // buffer of the last cgs_cg_prepare: this plain solve replaces both
    drop_graph(h);
    if (!h->d_hist || h->hist_stride != max_iter + 1 || h->hist_bins != 1) {
        dfree(h, h->d_hist, 1);
        h->hist_stride = max_iter + 1;
        h->hist_bins = 1;
        CKS(dalloc(h, &h->d_hist, (size_t)h->hist_stride));
    }
    CK(hipMemsetAsync(h->d_hist, 0, sizeof(double) * h->hist_stride, h->st));
    CK(hipMemsetAsync(h->d_done, 0, sizeof(int), h->st));
    float2* tmp = nullptr;
    CKS(upload_complex(h, &tmp, img, (size_t)h->C * h->ng));
    CK(launch_transpose_coils(tmp, h->d_grid, h->ng, h->C, (long long)h->N[1] * h->N[2], h->xp, 1, h->st));
    CK(launch_coil_combine(h->d_grid, h->d_maps, h->d_r, h->ng, h->C, (long long)h->N[1] * h->N[2], h->xp, h->st));
    CK(launch_copy_rhs(h->d_x, h->d_r, h->d_p, h->d_r, h->ng, 1, h->st));
    CK(launch_dot_partials(h->d_r, h->d_r, h->ng, 1, h->nblk_upd, h->d_part2, h->st));
    CK(launch_cg_init(h->d_sc, h->d_part2, h->nblk_upd, 1, thresh, h->d_done, h->st));
    CK(hipStreamSynchronize(h->st));
    dfree(h, tmp, (size_t)h->C * h->ng);
    h->cg_ready = true;
    CKS(cgs_cg_iterate(h, CGS_SOLVE_PLAIN, max_iter + 1));
    return cgs_cg_results(h, x_out, resid_out, iters_out);
}

int cgs_bin_count(cgs_ctx* h, int bin, long long* m) {
    if (!h || !m || bin < 0 || bin >= h->nbins) return CGS_ERR_ARG;
    *m = h->bins[bin].m;
    return CGS_OK;
}

int cgs_bin_tables(cgs_ctx* h, int bin, int32_t* nearest, int32_t* perm, int32_t* tile_off, int32_t* base,
                   double* vals) {
    if (!h || bin < 0 || bin >= h->nbins) return CGS_ERR_ARG;
    const BinData& B = h->bins[bin];
    if (!B.built) return fail(h, CGS_ERR_STATE, "bin not built");
    if (nearest) {
        int *key = nullptr, *cnt = nullptr, *nr = nullptr;
        CKS(dalloc(h, &key, h->ns));
        CKS(dalloc(h, &cnt, h->tiles.ntiles));
        CKS(dalloc(h, &nr, (size_t)h->ns * 3));
        CK(hipMemset(cnt, 0, sizeof(int) * h->tiles.ntiles));
        CK(launch_bin_keys(h->d_traj, h->ns, nullptr, nullptr, geom(h), h->tiles, key, nr, cnt, h->st));
        CK(hipMemcpyAsync(nearest, nr, sizeof(int) * 3 * h->ns, hipMemcpyDeviceToHost, h->st));
        CK(hipStreamSynchronize(h->st));
        dfree(h, key, h->ns); dfree(h, cnt, h->tiles.ntiles); dfree(h, nr, (size_t)h->ns * 3);
    }
    if (perm && B.m) CK(hipMemcpy(perm, B.perm, sizeof(int) * B.m, hipMemcpyDeviceToHost));
    if (tile_off) CK(hipMemcpy(tile_off, B.tile_off, sizeof(int) * (h->tiles.ntiles + 1), hipMemcpyDeviceToHost));
    if (base && B.m) {
        std::vector<int4> tb(B.m);
        CK(hipMemcpy(tb.data(), B.base, sizeof(int4) * B.m, hipMemcpyDeviceToHost));
        for (long long s = 0; s < B.m; ++s) { base[3 * s] = tb[s].x; base[3 * s + 1] = tb[s].y; base[3 * s + 2] = tb[s].z; }
    }
    if (vals && B.m) {
        const int kw = h->kb.kw, vs = val_stride(kw);
        std::vector<double> tv((size_t)B.m * vs);
        CK(hipMemcpy(tv.data(), B.vals, sizeof(double) * B.m * vs, hipMemcpyDeviceToHost));
        for (long long s = 0; s < B.m; ++s)
            for (int j = 0; j < 3 * kw; ++j) vals[s * 3 * kw + j] = tv[s * vs + j];
    }
    return CGS_OK;
}

int cgs_kernel_info(const cgs_ctx* h, int* kw, int* ntiles, double* thr) {
    if (!h || !h->kb_ready) return CGS_ERR_STATE;
    if (kw) *kw = h->kb.kw;
    if (ntiles) *ntiles = h->tiles.ntiles;
    if (thr) *thr = h->kb.thr;
    return CGS_OK;
}

int cgs_profile_enable(cgs_ctx* h, int on) {
    if (!h) return CGS_ERR_ARG;
    prof_flush(h);
    h->prof = on != 0;
    return CGS_OK;
}

int cgs_profile_reset(cgs_ctx* h) {
    if (!h) return CGS_ERR_ARG;
    prof_flush(h);
    h->acc.clear();
    return CGS_OK;
}

int cgs_profile_report(cgs_ctx* h, char* buf, int len) {
    if (!h) return -1;
    prof_flush(h);
    std::string s;
    char line[256];
    for (auto& kv : h->acc) {
        std::snprintf(line, sizeof(line), "%s %lld %.6f\n", kv.first.c_str(), kv.second.first, kv.second.second);
        s += line;
    }
    if (buf && len > 0) {
        std::strncpy(buf, s.c_str(), (size_t)len - 1);
        buf[len - 1] = 0;
    }
    return (int)s.size() + 1;
}

long long cgs_device_bytes(const cgs_ctx* h) { return h ? h->dev_bytes : 0; }

// ------------------------------------------------------------------ Voronoi
// One build of the listed cells (all when list is null; wide: from the wide
// box of the hull rule): small limits in LDS, the cells that outgrow them
// with medium limits in LDS, the rest with large limits in a device
// workspace.  Every tier gives a cell the same bits; the environment variable
// CGS_VORONOI_FIRST_TIER (1 or 2) starts every cell at a later tier, so that
// tests can check each tier on its own.  Returns the closed cells that
// outgrew the large limits in *over.
static int voronoi_pass(cgs_ctx* h, const double* dp, long long n, const int* list, bool wide, const VorHash& H,
                        double* dv, int* redo, int* redo2, int* dov, int* over, hipStream_t st) {
    const char* ft = std::getenv("CGS_VORONOI_FIRST_TIER");
    const int first = ft ? std::atoi(ft) : 0;
    CK(hipMemsetAsync(dov, 0, 3 * sizeof(int), st));
    // cells still to build: the list (n points, every point when list is null)
    long long todo = n;
    const int* tl = list;
    int cnt[3] = {0, 0, 0};
    if (first <= 0) {
        CK(launch_voronoi(dp, n, list, wide, H, dv, redo, dov, st));
        CK(hipMemcpyAsync(cnt, dov, sizeof(cnt), hipMemcpyDeviceToHost, st));
        CK(hipStreamSynchronize(st));
        todo = cnt[0];
        tl = redo;
        std::swap(redo, redo2);
    }
    if (todo > 0 && first <= 1) {
        CK(launch_voronoi_mid(dp, todo, tl, wide, H, dv, redo, dov + 2, st));
        CK(hipMemcpyAsync(cnt, dov, sizeof(cnt), hipMemcpyDeviceToHost, st));
        CK(hipStreamSynchronize(st));
        todo = cnt[2];
        tl = redo;
    }
    if (todo > 0) {
        // last tier: one wave per cell, each with a large-limit cell (about
        // 256 KB) in device memory, as many waves as there are cells, up to
        // what an eighth of the free memory (at most 4 GB) holds
        const size_t cell_b = voronoi_large_cell_bytes();
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) fr = 0;
        const long long budget = std::min<long long>((long long)(fr / 8), 4LL << 30);
        const long long fit = std::max<long long>(64, budget / (long long)cell_b);
        const int nwaves = (int)std::min<long long>(todo, fit);
        char* cells = nullptr;
        CKS(dalloc(h, &cells, cell_b * nwaves));
        CK(launch_voronoi_large(dp, wide, H, tl, todo, nwaves, cells, dv, dov + 1, st));
        CK(hipMemcpyAsync(cnt, dov, sizeof(cnt), hipMemcpyDeviceToHost, st));
        CK(hipStreamSynchronize(st));
        dfree(h, cells, cell_b * nwaves);
    }
    *over += cnt[1];
    return CGS_OK;
}

// Voronoi cell volumes of np points.  rule CGS_VORONOI_HULL: 0 only for
// unbounded cells (the reference); cells that keep a face of the first box
// are built again from the wide box.  CGS_VORONOI_CLIPPED: 0 for every cell
// that keeps a face of the first box.
static int voronoi_volumes_dev(cgs_ctx* h, const double* pts, long long np, int rule, double* vol_host,
                               hipStream_t st) {
    if (np < 2) {
        for (long long i = 0; i < np; i++) vol_host[i] = 0.0;
        return CGS_OK;
    }
    // uniform hash: about six points per cell, ascending point index per cell
    double lo[3], hi[3];
    for (int a = 0; a < 3; a++) { lo[a] = pts[a]; hi[a] = pts[a]; }
    for (long long i = 0; i < np; i++)
        for (int a = 0; a < 3; a++) {
            lo[a] = std::fmin(lo[a], pts[3 * i + a]);
            hi[a] = std::fmax(hi[a], pts[3 * i + a]);
        }
    double vol = 1.0;
    for (int a = 0; a < 3; a++) vol *= std::fmax(hi[a] - lo[a], 1e-12);
    VorHash H{};
    H.h = std::cbrt(vol * 6.0 / (double)np);
    for (int a = 0; a < 3; a++) { H.lo[a] = lo[a]; H.n[a] = (int)((hi[a] - lo[a]) / H.h) + 1; }
    for (int a = 0; a < 3; a++) { H.blo[a] = lo[a] - H.h; H.bhi[a] = hi[a] + H.h; }
    double ext = 0.0;
    for (int a = 0; a < 3; a++) ext = std::fmax(ext, hi[a] - lo[a]);
    const double wide = kVorWideBox * ext;
    for (int a = 0; a < 3; a++) { H.wlo[a] = lo[a] - wide; H.whi[a] = hi[a] + wide; }
    const long long nc = (long long)H.n[0] * H.n[1] * H.n[2];
    std::vector<int> start(nc + 1, 0), key(np), idx(np);
    for (long long i = 0; i < np; i++) {
        int c3[3];
        for (int a = 0; a < 3; a++) {
            int k = (int)((pts[3 * i + a] - H.lo[a]) / H.h);
            c3[a] = k < 0 ? 0 : (k >= H.n[a] ? H.n[a] - 1 : k);
        }
        key[i] = (c3[0] * H.n[1] + c3[1]) * H.n[2] + c3[2];
        start[key[i] + 1]++;
    }
    for (long long c = 0; c < nc; c++) start[c + 1] += start[c];
    std::vector<int> cur(start.begin(), start.end() - 1);
    for (long long i = 0; i < np; i++) idx[cur[key[i]]++] = (int)i;

    double *dp = nullptr, *dv = nullptr;
    int *ds = nullptr, *di = nullptr, *dov = nullptr, *redo = nullptr, *redo2 = nullptr, *dl = nullptr;
    CKS(dalloc(h, &dp, (size_t)np * 3));
    CKS(dalloc(h, &dv, (size_t)np));
    CKS(dalloc(h, &ds, (size_t)nc + 1));
    CKS(dalloc(h, &di, (size_t)np));
    CKS(dalloc(h, &dov, 3));          // [0] cells to rebuild, [1] final overflows, [2] rebuilt again
    CKS(dalloc(h, &redo, (size_t)np));
    CKS(dalloc(h, &redo2, (size_t)np));
    CK(hipMemcpy(dp, pts, sizeof(double) * 3 * np, hipMemcpyHostToDevice));
    CK(hipMemcpy(ds, start.data(), sizeof(int) * (nc + 1), hipMemcpyHostToDevice));
    CK(hipMemcpy(di, idx.data(), sizeof(int) * np, hipMemcpyHostToDevice));
    H.start = ds;
    H.idx = di;
    int over = 0;
    CKS(voronoi_pass(h, dp, np, nullptr, false, H, dv, redo, redo2, dov, &over, st));
    CK(hipMemcpyAsync(vol_host, dv, sizeof(double) * np, hipMemcpyDeviceToHost, st));
    CK(hipStreamSynchronize(st));
    if (rule == CGS_VORONOI_HULL) {
        // the cells still open in the first box, built again from the wide one
        std::vector<int> open;
        for (long long i = 0; i < np; i++)
            if (vol_host[i] == 0.0) open.push_back((int)i);
        if (!open.empty()) {
            CKS(dalloc(h, &dl, open.size()));
            CK(hipMemcpy(dl, open.data(), sizeof(int) * open.size(), hipMemcpyHostToDevice));
            CKS(voronoi_pass(h, dp, (long long)open.size(), dl, true, H, dv, redo, redo2, dov, &over, st));
            CK(hipMemcpyAsync(vol_host, dv, sizeof(double) * np, hipMemcpyDeviceToHost, st));
            CK(hipStreamSynchronize(st));
            dfree(h, dl, open.size());
        }
    }
    dfree(h, dp, (size_t)np * 3); dfree(h, dv, (size_t)np); dfree(h, ds, (size_t)nc + 1);
    dfree(h, di, (size_t)np); dfree(h, dov, 3); dfree(h, redo, (size_t)np); dfree(h, redo2, (size_t)np);
    if (over) return fail(h, CGS_ERR_ARG, std::to_string(over) + " Voronoi cells exceeded the face/vertex limits");
    return CGS_OK;
}

// functions.h:212-245: optional averaging over i mod sx, then / max.
static void normalise_density(const double* vol, long long np, int average, int sx, double cutoff, float* dcf) {
    std::vector<float> cv(np);
    for (long long i = 0; i < np; i++) cv[i] = (float)vol[i];
    if (average && sx > 0) {
        std::vector<float> avg(sx, 0.0f);
        for (long long i = 0; i < np; i++) avg[i % sx] += cv[i];
        double tot = 0.0;
        for (int i = 0; i < sx; i++) tot += avg[i];
        tot = 1.0 / tot;
        const int cut = (int)((sx - 1) * cutoff);
        for (long long i = 0; i < np; i++) cv[i] = (float)(tot * avg[std::min((int)(i % sx), cut)]);
    }
    double mx = 0.0;
    for (long long i = 0; i < np; i++)
        if (cv[i] > mx) mx = cv[i];
    mx = 1.0 / mx;
    for (long long i = 0; i < np; i++) dcf[i] = (float)(cv[i] * mx);
}

int cgs_voronoi_volumes_rule(const double* pts, long long np, int rule, double* vol) {
    if (!pts || !vol || np < 0 || (rule != CGS_VORONOI_HULL && rule != CGS_VORONOI_CLIPPED)) return CGS_ERR_ARG;
    return voronoi_volumes_dev(nullptr, pts, np, rule, vol, nullptr);
}

int cgs_voronoi_density_rule(const double* pts, long long np, int rule, int average, int sx, double cutoff,
                             float* dcf) {
    if (!pts || !dcf || np < 1 || (rule != CGS_VORONOI_HULL && rule != CGS_VORONOI_CLIPPED)) return CGS_ERR_ARG;
    std::vector<double> vol(np);
    const int st = voronoi_volumes_dev(nullptr, pts, np, rule, vol.data(), nullptr);
    if (st != CGS_OK) return st;
    normalise_density(vol.data(), np, average, sx, cutoff, dcf);
    return CGS_OK;
}

int cgs_voronoi_volumes(const double* pts, long long np, double* vol) {
    return cgs_voronoi_volumes_rule(pts, np, CGS_VORONOI_HULL, vol);
}

int cgs_voronoi_density(const double* pts, long long np, int average, int sx, double cutoff, float* dcf) {
    return cgs_voronoi_density_rule(pts, np, CGS_VORONOI_HULL, average, sx, cutoff, dcf);
}

int cgs_set_bin_voronoi(cgs_ctx* h, int bin) {
    if (!h || bin < 0 || bin >= h->nbins) return CGS_ERR_ARG;
    BinData& B = h->bins[bin];
    if (!B.built) return fail(h, CGS_ERR_STATE, "bin has no samples");
    std::vector<int> perm(B.m);
    std::vector<double> traj((size_t)h->ns * 3);
    if (B.m) CK(hipMemcpy(perm.data(), B.perm, sizeof(int) * B.m, hipMemcpyDeviceToHost));
    CK(hipMemcpy(traj.data(), h->d_traj, sizeof(double) * 3 * h->ns, hipMemcpyDeviceToHost));
    // the bin's active samples in original index order
    std::vector<int> act(perm);
    std::sort(act.begin(), act.end());
    std::vector<double> pts((size_t)B.m * 3);
    for (long long s = 0; s < B.m; s++)
        for (int a = 0; a < 3; a++) pts[3 * s + a] = traj[3 * (size_t)act[s] + a];
    std::vector<double> vol(B.m);
    CKS(voronoi_volumes_dev(h, pts.data(), B.m, CGS_VORONOI_CLIPPED, vol.data(), h->st));
    std::vector<float> w(B.m), full(h->ns, 0.0f);
    normalise_density(vol.data(), B.m, 0, 0, 1.0, w.data());
    for (long long s = 0; s < B.m; s++) full[act[s]] = w[s];
    if (!B.dcf_full) CKS(dalloc(h, &B.dcf_full, h->ns));
    CK(hipMemcpy(B.dcf_full, full.data(), sizeof(float) * h->ns, hipMemcpyHostToDevice));
    free_packed(h);
    return CGS_OK;
}

// ------------------------------------------------------ differentiable ops
int cgs_set_stream(cgs_ctx* h, void* stream) {
    if (!h) return CGS_ERR_ARG;
    CK(hipStreamSynchronize(h->st));
    h->st = stream ? (hipStream_t)stream : h->own_st;
    return CGS_OK;
}

namespace {

void pts_free_tables(cgs_pts* p) {
    cgs_ctx* h = p->h;
    const int vs = val_stride(p->kw > 0 ? p->kw : h->kb.kw);
    dfree(h, p->perm, p->m);
    dfree(h, p->tile_off, h->tiles.ntiles + 1);
    dfree(h, p->base, p->m);
    dfree(h, p->vals, p->m * vs);
    dfree(h, p->dvals, p->m * vs);
    dfree(h, p->sbin, std::max(1LL, p->m));
    dfree(h, p->ys, p->m * h->C);
    p->m = 0;
    p->ready = false;
}

int pts_check(cgs_pts* p, bool need_maps) {
    if (!p || !p->h) return CGS_ERR_ARG;
    cgs_ctx* h = p->h;
    if (!p->ready) return fail(h, CGS_ERR_STATE, "point set has no coordinates");
    if (p->tcfg != h->tcfg || p->kw != h->kb.kw || p->w != h->w || p->b != h->b || !h->kb_ready)
        return fail(h, CGS_ERR_STATE, "point set was built for another tile shape or kernel; set its coordinates again");
    if (need_maps && !h->maps_ready) return fail(h, CGS_ERR_STATE, "sensitivity maps not set");
    if (!h->kft_ready) return fail(h, CGS_ERR_STATE, "kernelFT not computed");
    CKS(ensure_grid(h, 1));
    return CGS_OK;
}

GridArgs pts_grid_args(cgs_pts* p) {
    cgs_ctx* h = p->h;
    GridArgs g{};
    g.G = h->d_grid;
    g.Y = p->ys;
    g.s.base = p->base;
    g.s.vals = p->vals;
    g.s.dcf = nullptr;
    g.s.sbin = p->sbin;
    g.s.first = 0;
    g.s.m = p->m;
    g.g = geom(h);
    g.t = h->tiles;
    g.tile_off = p->tile_off;
    g.gdcf = nullptr;
    g.thr = h->kb.thr;
    g.thrf = h->kb.thr_up;
    g.b0 = 0;
    g.nb = 1;
    g.done = nullptr;
    return g;
}

// Coordinate gradient of <B, degrid(grid)> on the context's grid buffer.
int pts_grad(cgs_pts* p, const float2* b_caller, float scale, float* gk) {
    cgs_ctx* h = p->h;
    CK(launch_permute_rows(b_caller, p->base, p->m, h->C, 0, p->ys, h->st));
    GradKArgs a{};
    a.A = h->d_grid;
    a.B = p->ys;
    a.base = p->base;
    a.vals = p->vals;
    a.dvals = p->dvals;
    a.m = p->m;
    a.g = geom(h);
    a.thr = h->kb.thr;
    a.thrf = h->kb.thr_up;
    a.scale = scale;
    a.gk = gk;
    {
        Prof t(h, "grad_k");
        CK(launch_grad_k(h->kb.kw, a, h->st));
    }
    return CGS_OK;
}

}  // namespace

int cgs_pts_create(cgs_ctx* h, cgs_pts** out) {
    if (!h || !out) return CGS_ERR_ARG;
    cgs_pts* p = new cgs_pts;
    p->h = h;
    *out = p;
    return CGS_OK;
}

int cgs_pts_destroy(cgs_pts* p) {
    if (!p) return CGS_OK;
    cgs_ctx* h = p->h;
    (void)hipStreamSynchronize(h->st);
    pts_free_tables(p);
    dfree(h, p->k64, (size_t)p->cap * 3);
    delete p;
    return CGS_OK;
}

int cgs_pts_set_coords(cgs_pts* p, long long m, const float* k_dev) {
    if (!p || !p->h || m < 0 || (m > 0 && !k_dev)) return CGS_ERR_ARG;
    cgs_ctx* h = p->h;
    if (m > (long long)INT32_MAX) return fail(h, CGS_ERR_ARG, "too many samples");
    if (!h->kb_ready) CKS(setup_kb(h));
    pts_free_tables(p);
    if (m > p->cap) {
        dfree(h, p->k64, (size_t)p->cap * 3);
        p->cap = 0;
        CKS(dalloc(h, &p->k64, (size_t)m * 3));
        p->cap = m;
    }
    CK(launch_widen_coords(k_dev, m * 3, p->k64, h->st));
    long long mm = 0;
    CKS(sort_samples(h, p->k64, m, nullptr, nullptr, &mm, &p->perm, &p->tile_off, &p->base, &p->vals));
    p->m = mm;   // every point is active: mm == m
    const int vs = val_stride(h->kb.kw);
    CKS(dalloc(h, &p->dvals, (size_t)mm * vs));
    CKS(dalloc(h, &p->sbin, std::max(1LL, mm)));
    CKS(dalloc(h, &p->ys, (size_t)mm * h->C));
    CK(hipMemsetAsync(p->sbin, 0, sizeof(int) * std::max(1LL, mm), h->st));
    CK(launch_sample_dtables(p->k64, p->perm, mm, geom(h), h->kb, p->dvals, h->st));
    p->tcfg = h->tcfg;
    p->kw = h->kb.kw;
    p->w = h->w;
    p->b = h->b;
    p->ready = true;
    return CGS_OK;
}

int cgs_pts_forward(cgs_pts* p, const float* x_dev, float* y_dev) {
    CKS(pts_check(p, true));
    cgs_ctx* h = p->h;
    CKS(run_fwd_fft(h, 0, 1, (const float2*)x_dev, true, nullptr, false));
    GridArgs g = pts_grid_args(p);
    {
        Prof t(h, "degrid");
        CK(launch_degrid(h->kb.kw, g, h->st));
    }
    CK(launch_permute_rows(p->ys, p->base, p->m, h->C, 1, (float2*)y_dev, h->st));
    return CGS_OK;
}

int cgs_pts_adjoint(cgs_pts* p, const float* y_dev, float* x_dev) {
    CKS(pts_check(p, true));
    cgs_ctx* h = p->h;
    CK(launch_permute_rows((const float2*)y_dev, p->base, p->m, h->C, 0, p->ys, h->st));
    GridArgs g = pts_grid_args(p);
    {
        Prof t(h, "grid");
        CK(launch_grid(h->kb.kw, h->tcfg, g, h->st));
    }
    CKS(run_inv_fft(h, 0, 1, (float2*)x_dev, nullptr, nullptr, nullptr, false));
    return CGS_OK;
}

int cgs_pts_grad_coords_fwd(cgs_pts* p, const float* x_dev, const float* gy_dev, float* gk_dev) {
    CKS(pts_check(p, true));
    cgs_ctx* h = p->h;
    CKS(run_fwd_fft(h, 0, 1, (const float2*)x_dev, true, nullptr, false));
    return pts_grad(p, (const float2*)gy_dev, 1.0f, gk_dev);
}

int cgs_pts_grad_coords_adj(cgs_pts* p, const float* y_dev, const float* gx_dev, float* gk_dev) {
    CKS(pts_check(p, true));
    cgs_ctx* h = p->h;
    CKS(run_fwd_fft(h, 0, 1, (const float2*)gx_dev, true, nullptr, false));
    const float s = (float)(1e-3 / std::sqrt((double)h->ng));
    return pts_grad(p, (const float2*)y_dev, s, gk_dev);
}

int cgs_pts_count(const cgs_pts* p, long long* m) {
    if (!p || !m) return CGS_ERR_ARG;
    *m = p->m;
    return CGS_OK;
}

// ------------------------------------------------------------------ Fourier
int cgs_fft_create(const int dims[3], cgs_fft** out) {
    if (!out) return CGS_ERR_ARG;
    *out = nullptr;
    for (int a = 0; a < 3; ++a)
        if (!fft_len_supported(dims[a])) return CGS_ERR_ARG;
    cgs_fft* f = new cgs_fft();
    for (int a = 0; a < 3; ++a) f->N[a] = dims[a];
    f->ng = (long long)dims[0] * dims[1] * dims[2];
    if (hipStreamCreate(&f->st) != hipSuccess) { delete f; return CGS_ERR_HIP; }
    if (upload_twiddles(nullptr, f->N, f->d_tw) != CGS_OK || hipMalloc(&f->d_a, sizeof(float2) * f->ng) != hipSuccess) {
        cgs_fft_destroy(f);
        return CGS_ERR_HIP;
    }
    *out = f;
    return CGS_OK;
}

int cgs_fft_destroy(cgs_fft* f) {
    if (!f) return CGS_OK;
    for (int a = 0; a < 3; ++a)
        if (f->d_tw[a]) (void)hipFree(f->d_tw[a]);
    if (f->d_a) (void)hipFree(f->d_a);
    if (f->st) (void)hipStreamDestroy(f->st);
    delete f;
    return CGS_OK;
}

static int fft_run(cgs_fft* f, float* a, int dir, int rescale) {
    if (!f || !a) return CGS_ERR_ARG;
    cgs_ctx* h = nullptr;
    CK(hipMemcpyAsync(f->d_a, a, sizeof(float2) * f->ng, hipMemcpyHostToDevice, f->st));
    CK(launch_fft_z1(f->N[2], dir, f->d_a, (long long)f->N[0] * f->N[1], f->d_tw[2], f->st));
    PlainArgs y{};
    y.g = f->d_a; y.os = (long long)f->N[1] * f->N[2]; y.ls = f->N[2]; y.ncols = f->N[2]; y.tw = f->d_tw[1];
    y.outer_per_bin = f->N[0];
    CK(launch_fft_plain(f->N[1], dir, y, f->N[0], f->st));
    PlainArgs x{};
    x.g = f->d_a; x.os = f->ng; x.ls = (long long)f->N[1] * f->N[2]; x.ncols = x.ls; x.tw = f->d_tw[0];
    x.outer_per_bin = 1;
    CK(launch_fft_plain(f->N[0], dir, x, 1, f->st));
    if (dir > 0 && rescale) CK(launch_scale_d(f->d_a, f->ng, 1.0 / std::sqrt((double)f->ng), f->st));
    CK(hipMemcpyAsync(a, f->d_a, sizeof(float2) * f->ng, hipMemcpyDeviceToHost, f->st));
    CK(hipStreamSynchronize(f->st));
    return CGS_OK;
}

int cgs_fft_forward(cgs_fft* f, float* a) { return fft_run(f, a, -1, 0); }

int cgs_fft_backward(cgs_fft* f, float* a, int rescale) { return fft_run(f, a, +1, rescale); }

}  // extern "C"
