// Host side of the rectv HIP library: FP64 geometry and tables, device memory
// plan, the ADMM driver and the extern "C" entry points declared in
// include/rectv_amd.h.
#include "kernels.hpp"
#include "../../include/rectv_amd.h"

#include <algorithm>
#include <climits>
#include <cstdint>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <numeric>
#include <string>
#include <vector>

namespace rtv {

thread_local std::string g_last_error;  // message of the last failing C-ABI call (all sources)

static const double PI_REF = 3.1415926535;  // kernels_radonusfft.cuh:1
static const double PI_EXACT = 3.14159265358979323846;

// a prime factor above 7 (planned as a direct stage)
static bool has_big_prime(int len) {
    for (int R : {2, 3, 5, 7})
        while (len % R == 0) len /= R;
    return len > 1;
}

// ---------------------------------------------------------------- kernel pickers
// FFT kernels are instantiated for power-of-two lengths (compile-time index
// math) and with a runtime plan for every other length: NP = 0 (at most FFT_WG
// threads per workgroup) or NP = -1 (longer transforms, up to FFT_THREADS), and
// NP = -2 for lengths with a prime factor above 7 (at most FFT_WG threads).
#define RTV_PICKER(NAME, K)                                                                        \
    static decltype(&K<0>) NAME(int len) {                                                         \
        switch (len) {                                                                             \
            case 64: return K<64>;                                                                 \
            case 128: return K<128>;                                                               \
            case 256: return K<256>;                                                               \
            case 512: return K<512>;                                                               \
            case 1024: return K<1024>;                                                             \
            case 2048: return K<2048>;                                                             \
            case 4096: return K<4096>;                                                             \
            case 192: return K<192>;                                                               \
            case 384: return K<384>;                                                               \
            case 768: return K<768>;                                                               \
            case 1536: return K<1536>;                                                             \
            case 3072: return K<3072>;                                                             \
            case 6144: return K<6144>;                                                             \
            case 320: return K<320>;                                                               \
            case 640: return K<640>;                                                               \
            case 1280: return K<1280>;                                                             \
            case 2560: return K<2560>;                                                             \
            case 5120: return K<5120>;                                                             \
            default: return has_big_prime(len) ? K<-2> : fft_tpt(len, FFT_E) > FFT_WG ? K<-1> : K<0>;   \
        }                                                                                          \
    }
RTV_PICKER(pick_fwd_rows, k_fwd_rows)
RTV_PICKER(pick_fwd_cols, k_fwd_cols)
RTV_PICKER(pick_adj_cols, k_adj_cols)
RTV_PICKER(pick_adj_rows, k_adj_rows)
#undef RTV_PICKER
// per-ray n-point transforms: powers of two at compile time, other lengths planned at run time
static decltype(&k_ray<0>) pick_ray(int len) {
    switch (len) {
        case 64: return k_ray<64>;
        case 128: return k_ray<128>;
        case 256: return k_ray<256>;
        case 512: return k_ray<512>;
        case 1024: return k_ray<1024>;
        case 2048: return k_ray<2048>;
        case 4096: return k_ray<4096>;
        default: return has_big_prime(len) ? k_ray<-2> : fft_tpt(len, FFT_E) > FFT_WG ? k_ray<-1> : k_ray<0>;
    }
}
// 2n-point passes with a compile-time kernel of length R0 * 2^k (R0 = 3, 5): the lengths above
static bool rp_kernel(int len) {
    switch (len) {
        case 192: case 384: case 768: case 1536: case 3072: case 6144:
        case 320: case 640: case 1280: case 2560: case 5120: return true;
        default: return false;
    }
}

// ---------------------------------------------------------------- FFT plans
static bool make_plan(int N, FftPlan &p) {
    p.n = N;
    p.nst = 0;
    int r = N;
    const int order[] = {8, 4, 2, 3, 5, 7};
    for (int R : order) {
        while (r % R == 0 && p.nst < FFT_MAX_STAGES) {
            p.radix[p.nst++] = R;
            r /= R;
        }
    }
    // remaining prime factors up to FFT_MAX_PRIME: direct stages
    for (int R = 11; R <= FFT_MAX_PRIME && r > 1; R += 2)
        while (r % R == 0 && p.nst < FFT_MAX_STAGES) {
            p.radix[p.nst++] = R;
            r /= R;
        }
    return r == 1;
}

// ---------------------------------------------------------------- geometry
struct HostGeom {
    std::vector<SampleGeom> geo;
    std::vector<int> fitems;  // forward items per origin tile for the matrix-core gather (-1 = padding)
    std::vector<unsigned short> fcells;  // per gather entry: its covered cells (count first)
    std::vector<int> asids;
    std::vector<AdjItem> aitems;  // asids with their window geometry
    std::vector<int2> arec;       // per adjoint record: {2 * sample | mirror, first coefficient of its row}
    std::vector<TileEntry> fmtiles, atiles;  // atiles[].pad: first chunk of the tile
    // adjoint chunks (SM_CH items of a tile): crange = (byte offset, byte length) of the
    // chunk's block lists in blist; brange per (chunk, 4x4 block) = (offset within the
    // chunk's lists, length): slots of the items whose window meets the block
    std::vector<int2> crange, brange;
    std::vector<unsigned char> blist;
    std::vector<int4> achunk;  // per adjoint chunk: {tile, first record, count, 0}
    std::vector<int2> runs;    // per row of G: circular run {lo, len} of the columns inside some window
    std::vector<float2> pre_f, post_a, cf;
    std::vector<float2> post_aw;    // post_a times the FBP filter (filtered back-projection)
    std::vector<float2> cf_static;  // pair coefficients of a basis selecting one static frame
    std::vector<float> post_f, pre_a, deap;
    float beta = 0.f;
};

// FBP filter of kernels_radonusfft.cuh:173-189 at detector frequency index s of the
// centred spectrum: rho = (s - n/2)/n, w = 0 at rho == 0, else |rho| n 4 sin(rho)/rho.
static double fbp_filter(int s, int n) {
    const double rho = (double)(s - n / 2) / (double)n;
    return rho == 0.0 ? 0.0 : std::fabs(rho) * n * 4 * std::sin(rho) / rho;
}

// USFFT constants, radonusfft.cu:10-12: mu = -log(eps)/(2 n^2) held in float.
static float usfft_mu(int n) {
    const float eps = 1e-3f;
    return (float)(-std::log((double)eps) / (2.0 * (double)n * (double)n));
}
static int usfft_K(int n, float mu) {
    const double m = mu, eps = (double)1e-3f;
    return (int)std::ceil(2.0 * n * 1 / PI_REF * std::sqrt(-m * std::log(eps) + (m * n) * (m * n) / 4));
}

#ifndef RTV_GM_BALANCE
#define RTV_GM_BALANCE 1  // gather block placement (build-time constant): 0 in order, 1 / 2 balanced
#endif
static void build_geometry(HostGeom &hg, int n, int nt, int M, const float *theta, const float *phi, float center) {
    const int N = 2 * n, nT = tiles_per_side(N), K = 4;
    const float mu = usfft_mu(n);
    if (usfft_K(n, mu) != K) throw Error(-1, "USFFT half-width differs from 4 for this n");
    const double mud = mu;
    hg.beta = (float)(PI_REF * PI_REF / (mud * 4.0 * n * n));
    const size_t ns = (size_t)nt * n;
    const size_t ni = 2 * ns;  // items: 2*id direct window, 2*id+1 point-mirrored window
    hg.geo.assign(ni, SampleGeom{0, 0, 0.f, 0.f});
    std::vector<int> ftile(ni, -1);
    // polar sample positions (takexy, kernels_radonusfft.cuh:66-79) in FP64
    for (int k = 0; k < nt; k++) {
        const double th = (double)theta[k];
        const double sn = std::sin(th), cs = std::cos(th);
        for (int s = 1; s < n; s++) {
            const double r = (double)(s - n / 2) / n;
            double x0 = r * sn, y0 = r * cs;
            if (x0 >= 0.5) x0 = 0.5 - 1e-5;
            if (y0 >= 0.5) y0 = 0.5 - 1e-5;
            const double X = 2.0 * n * x0, Y = 2.0 * n * y0;
            const double fxl = std::floor(X), fyl = std::floor(Y);
            // the oversampled spectra are stored transposed (G[kx][ky]): the window
            // "row" runs along kx (X) and the "column" along ky (Y)
            SampleGeom g;
            g.row0 = (int)(((long long)n + (long long)fxl - K) % N + N) % N;
            g.col0 = (int)(((long long)n + (long long)fyl - K) % N + N) % N;
            g.fy = (float)(X - fxl);
            g.fx = (float)(Y - fyl);
            // mirror: cells (2n - c) mod 2n of the window, taps reversed, fractions negated
            SampleGeom m;
            m.row0 = (2 * N - g.row0 - (KW - 1)) % N;
            m.col0 = (2 * N - g.col0 - (KW - 1)) % N;
            m.fy = -g.fy;
            m.fx = -g.fx;
            const size_t id = (size_t)k * n + s;
            hg.geo[2 * id] = g;
            hg.geo[2 * id + 1] = m;
            ftile[2 * id] = (g.row0 / TS) * nT + g.col0 / TS;
            ftile[2 * id + 1] = (m.row0 / TS) * nT + m.col0 / TS;
        }
    }
    // Work lists.  Forward: the items of the tile holding their window origin,
    // grouped by origin (ids ascending), packed into 16-item blocks and entries
    // of at most GM_ITEMS items; adjoint: every tile the 9x9 window intersects
    // (one entry per tile, empty tiles included).
    struct Key {
        unsigned long long k;  // (tile << 32) | origin
        int id;
        bool operator<(const Key &o) const { return k != o.k ? k < o.k : id < o.id; }
    };
    const int ntile = nT * nT;
    {
        std::vector<Key> keys;
        keys.reserve(ni);
        for (size_t id = 0; id < ni; id++) {
            if (ftile[id] < 0) continue;
            const SampleGeom &g = hg.geo[id];
            keys.push_back(Key{((unsigned long long)ftile[id] << 32) | (unsigned)(g.row0 * N + g.col0), (int)id});
        }
        std::sort(keys.begin(), keys.end());
        // flat lists for the matrix-core gather: the items of each origin tile packed
        // into 16-item blocks (one wave each) spanning at most 12 rows and 4 steps of
        // 4 cells, whose matrix-core work is rows x steps instructions.  Origins are
        // taken by item count (most first, then raster order); an origin's items
        // join the open block whose work grows least, if that growth per item stays
        // below 0.6 of a lone origin's (a 9-row, 3-step window over 16 items), else
        // they open a block.  Entries hold at most GM_ITEMS items.
        struct GBlock {
            int rlo, rhi, clo, chi, cnt;
            int ids[16];
        };
        auto gcost = [](int rlo, int rhi, int clo, int chi) {
            const int rows = rhi - rlo + KW, steps = (chi + KW - 1 - (clo & ~3)) / 4 + 1;
            return (rows > 12 || steps > GM_KS) ? -1 : rows * steps;
        };
        hg.fitems.clear();
        hg.fmtiles.clear();
        size_t a = 0;
        while (a < keys.size()) {
            const int t = (int)(keys[a].k >> 32);
            const int ty0 = (t / nT) * TS, tx0 = (t % nT) * TS;
            const int e0 = (int)hg.fitems.size();
            struct Origin {
                int ly0, lx0;
                size_t a, b;  // its items keys[a .. b), ids ascending
            };
            std::vector<Origin> org;
            while (a < keys.size() && (int)(keys[a].k >> 32) == t) {
                size_t b = a;
                while (b < keys.size() && keys[b].k == keys[a].k) b++;
                const SampleGeom &g = hg.geo[keys[a].id];
                org.push_back(Origin{(g.row0 - ty0 + N) % N, (g.col0 - tx0 + N) % N, a, b});
                a = b;
            }
            std::stable_sort(org.begin(), org.end(), [](const Origin &p, const Origin &q) { return p.b - p.a > q.b - q.a; });
            std::vector<GBlock> blocks;
            for (const Origin &o : org) {
                size_t c = o.a;
                while (c < o.b) {
                    const int left = (int)(o.b - c);
                    int best = -1, binc = 0;
                    for (int q = 0; q < (int)blocks.size(); q++) {
                        const GBlock &bk = blocks[q];
                        if (bk.cnt == 16) continue;
                        const int c1 = gcost(std::min(bk.rlo, o.ly0), std::max(bk.rhi, o.ly0), std::min(bk.clo, o.lx0),
                                             std::max(bk.chi, o.lx0));
                        if (c1 < 0) continue;
                        const int inc = c1 - gcost(bk.rlo, bk.rhi, bk.clo, bk.chi);
                        if (best < 0 || inc < binc) {
                            best = q;
                            binc = inc;
                        }
                    }
                    // join when inc / take < 0.6 * 27 / 16 (integer form)
                    const int take_b = best >= 0 ? std::min(left, 16 - blocks[best].cnt) : 0;
                    if (best < 0 || binc * 16 * 10 >= 27 * 6 * take_b) {
                        blocks.push_back(GBlock{o.ly0, o.ly0, o.lx0, o.lx0, 0, {}});
                        best = (int)blocks.size() - 1;
                    }
                    GBlock &bk = blocks[best];
                    bk.rlo = std::min(bk.rlo, o.ly0);
                    bk.rhi = std::max(bk.rhi, o.ly0);
                    bk.clo = std::min(bk.clo, o.lx0);
                    bk.chi = std::max(bk.chi, o.lx0);
                    const int take = std::min(left, 16 - bk.cnt);
                    for (int i = 0; i < take; i++) bk.ids[bk.cnt++] = keys[c + i].id;
                    c += take;
                }
            }
            for (const GBlock &bk : blocks)
                for (int i = 0; i < 16; i++) hg.fitems.push_back(i < bk.cnt ? bk.ids[i] : -1);
            const int e1 = (int)hg.fitems.size();
            // the kernel relies on every block spanning at most 12 rows and 4 steps of 4 cells;
            // a block's matrix-core work is rows x steps instructions per column block
            std::vector<int> cost;
            for (int b0 = e0; b0 < e1; b0 += 16) {
                int r0 = 1 << 30, r1 = -1, c0 = 1 << 30, c1 = -1;
                for (int i = b0; i < b0 + 16; i++) {
                    if (hg.fitems[i] < 0) continue;
                    const SampleGeom &g = hg.geo[hg.fitems[i]];
                    const int ly0 = (g.row0 - ty0 + N) % N, lx0 = (g.col0 - tx0 + N) % N;
                    r0 = std::min(r0, ly0);
                    r1 = std::max(r1, ly0);
                    c0 = std::min(c0, lx0);
                    c1 = std::max(c1, lx0);
                }
                cost.push_back(r1 < 0 ? 0 : (r1 - r0 + KW) * ((c1 + KW - 1 - (c0 & ~3)) / 4 + 1));
                if (r1 < 0) continue;
                if (r1 - r0 + KW > 12 || (c1 + KW - 1 - (c0 & ~3)) / 4 + 1 > GM_KS || r1 >= TS || c1 >= TS)
                    throw Error(-1, "internal: gather block exceeds its window bounds");
            }
#if RTV_GM_BALANCE
            // Each entry's blocks (one per wave) are placed so that the four SIMDs of the
            // CU get similar work between the kernel's per-slice barriers: longest
            // block first onto the least loaded SIMD (4 waves each), wave w running on
            // SIMD w mod 4.  Unused wave slots before the last used one hold empty
            // blocks.  Results do not depend on the placement.
            {
                std::vector<int> blk(hg.fitems.begin() + e0, hg.fitems.end());
                hg.fitems.resize(e0);
                const int nb = (int)cost.size();
                for (int q0 = 0; q0 < nb; q0 += GM_ITEMS / 16) {
                    const int q1 = std::min(nb, q0 + GM_ITEMS / 16);
                    std::vector<int> order;
                    for (int q = q0; q < q1; q++) order.push_back(q);
                    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
                    int load[4] = {0, 0, 0, 0}, used[4] = {0, 0, 0, 0};
                    int slot[GM_ITEMS / 16];
                    std::fill(slot, slot + GM_ITEMS / 16, -1);
                    for (int q : order) {
                        int s = -1;
                        for (int k = 0; k < 4; k++)
                            if (used[k] < GM_ITEMS / 64 && (s < 0 || load[k] < load[s])) s = k;
                        load[s] += cost[q];
#if RTV_GM_BALANCE == 2
                        slot[s * (GM_ITEMS / 64) + used[s]++] = q;
#else
                        slot[s + 4 * used[s]++] = q;
#endif
                    }
                    int last = GM_ITEMS / 16 - 1;
                    while (last > 0 && slot[last] < 0) last--;
                    const int s0 = (int)hg.fitems.size();
                    for (int w = 0; w <= last; w++)
                        for (int i = 0; i < 16; i++) hg.fitems.push_back(slot[w] < 0 ? -1 : blk[16 * slot[w] + i]);
                    hg.fmtiles.push_back(TileEntry{t, s0, (int)hg.fitems.size() - s0, 0});
                }
            }
#else
            for (int s0 = e0; s0 < e1; s0 += GM_ITEMS)
                hg.fmtiles.push_back(TileEntry{t, s0, std::min(GM_ITEMS, e1 - s0), 0});
#endif
        }
        // per entry, the cells of its TH x TH region that some item's 9 x 9 window
        // covers, in raster order (about half of the region at the benchmark
        // geometry): fcells[pad] = count, the cell indices r * TH + c follow.  The
        // gather stages only these; the others stay zero and meet zero taps only.
        hg.fcells.clear();
        bool cov[TH * TH];
        for (TileEntry &te : hg.fmtiles) {
            const int ty0 = (te.tile / nT) * TS, tx0 = (te.tile % nT) * TS;
            std::fill(cov, cov + TH * TH, false);
            for (int i = te.start; i < te.start + te.count; i++) {
                if (hg.fitems[i] < 0) continue;
                const SampleGeom &g = hg.geo[hg.fitems[i]];
                const int ly0 = (g.row0 - ty0 + N) % N, lx0 = (g.col0 - tx0 + N) % N;
                if (ly0 >= TS || lx0 >= TS) throw Error(-1, "internal: gather item outside its origin tile");
                for (int a = 0; a < KW; a++)
                    for (int b = 0; b < KW; b++) cov[(ly0 + a) * TH + lx0 + b] = true;
            }
            te.pad = (int)hg.fcells.size();
            hg.fcells.push_back(0);
            for (int c = 0; c < TH * TH; c++)
                if (cov[c]) hg.fcells.push_back((unsigned short)c);
            hg.fcells[te.pad] = (unsigned short)(hg.fcells.size() - te.pad - 1);
        }
    }
    // adjoint lists: every tile the 9x9 window intersects, each tile's items
    // sorted by (local window row, local window column, id)
    {
        std::vector<Key> keys;
        keys.reserve(3 * ni);
        for (size_t id = 0; id < ni; id++) {
            if (ftile[id] < 0) continue;
            const SampleGeom &g = hg.geo[id];
            // the tiles the 9 x 9 window meets: 2 per axis at most, 3 when a narrow
            // last tile (2n not a multiple of 16) lies inside the window
            int ty[3], tx[3], nty = 0, ntx = 0;
            for (int a = 0; a < KW; a++) {
                const int r = ((g.row0 + a) % N) / TS, c = ((g.col0 + a) % N) / TS;
                if (nty == 0 || ty[nty - 1] != r) ty[nty++] = r;
                if (ntx == 0 || tx[ntx - 1] != c) tx[ntx++] = c;
            }
            for (int a = 0; a < nty; a++)
                for (int b = 0; b < ntx; b++) {
                    int dy = (g.row0 - ty[a] * TS + N) % N, dx = (g.col0 - tx[b] * TS + N) % N;
                    const int ly0 = dy < TS ? dy : dy - N, lx0 = dx < TS ? dx : dx - N;
                    const unsigned loc = (unsigned)((ly0 + 64) << 8 | (lx0 + 64));
                    keys.push_back(Key{((unsigned long long)(ty[a] * nT + tx[b]) << 32) | loc, (int)id});
                }
        }
        std::sort(keys.begin(), keys.end());
        hg.asids.resize(keys.size());
        std::vector<int> cnt(ntile, 0), off(ntile + 1, 0);
        for (size_t a = 0; a < keys.size(); a++) {
            hg.asids[a] = keys[a].id;
            cnt[keys[a].k >> 32]++;
        }
        for (int t = 0; t < ntile; t++) off[t + 1] = off[t] + cnt[t];
        // Spread each tile's records over its chunks round robin (record k of the
        // sorted list goes to chunk k mod nchunks): a chunk then covers the whole tile
        // evenly, so the 16 waves (one 4 x 4 cell block each) get lists of similar
        // length between two barriers.  Sorted chunks put nearby window origins
        // together and leave most waves idle (busiest wave ~2.6x the mean at the full
        // workload, 1.3x interleaved).
        {
            std::vector<int> tmp;
            for (int t = 0; t < ntile; t++) {
                const int c = cnt[t], nchk = (c + SM_CH - 1) / SM_CH;
                if (nchk < 2) continue;
                tmp.assign(hg.asids.begin() + off[t], hg.asids.begin() + off[t] + c);
                int q = off[t];
                for (int ch = 0; ch < nchk; ch++)
                    for (int k = ch; k < c; k += nchk) hg.asids[q++] = tmp[k];
            }
        }
        hg.aitems.resize(hg.asids.size());
        hg.arec.resize(hg.asids.size());
        const int MPh = (M + 1) / 2;
        for (size_t a = 0; a < hg.asids.size(); a++) {
            const int pid = hg.asids[a], sid = pid >> 1, mir = pid & 1;
            const SampleGeom &g = hg.geo[pid];
            hg.aitems[a] = AdjItem{pid, g.row0 | (g.col0 << 16), g.fy, g.fx};
            hg.arec[a] = make_int2((2 * sid) | mir, ((sid / n) * 2 + mir) * MPh);
        }
        hg.atiles.resize(ntile);
        for (int t = 0; t < ntile; t++) hg.atiles[t] = TileEntry{t, off[t], cnt[t], 0};
        // per chunk of SM_CH items and 4 x 4 cell block: the slots of the items whose
        // 9 x 9 window meets the block, ascending, padded to a multiple of 4 with the
        // zero slot SM_CH
        hg.crange.clear();
        hg.brange.clear();
        hg.blist.clear();
        hg.achunk.clear();
        for (int t = 0; t < ntile; t++) {
            TileEntry &te = hg.atiles[t];
            te.pad = (int)hg.crange.size();
            const int ty0 = (t / nT) * TS, tx0 = (t % nT) * TS;
            for (int c0 = 0; c0 < te.count; c0 += SM_CH) {
                const int cn = std::min(SM_CH, te.count - c0);
                std::vector<int> ly(cn), lx(cn);
                for (int i = 0; i < cn; i++) {
                    const SampleGeom &g = hg.geo[hg.asids[te.start + c0 + i]];
                    const int dy = (g.row0 - ty0 + N) % N, dx = (g.col0 - tx0 + N) % N;
                    ly[i] = dy < TS ? dy : dy - N;
                    lx[i] = dx < TS ? dx : dx - N;
                }
                const int base = (int)hg.blist.size();
                for (int b = 0; b < 16; b++) {
                    const int r0 = 4 * (b / 4), q0 = 4 * (b % 4);
                    const int off = (int)hg.blist.size() - base;
                    // the block's slots by residue mod 4, emitted in groups of 4 whose
                    // k-th entry is k mod 4 while the residues last (distinct LDS bank
                    // groups for the 4 slots one matrix-core operand read covers)
                    std::vector<int> byres[4];
                    for (int i = 0; i < cn; i++)
                        if (ly[i] <= r0 + 3 && ly[i] + KW - 1 >= r0 && lx[i] <= q0 + 3 && lx[i] + KW - 1 >= q0)
                            byres[i & 3].push_back(i);
                    size_t head[4] = {0, 0, 0, 0};
                    auto left = [&](int k) { return byres[k].size() - head[k]; };
                    for (;;) {
                        int k0 = 0;
                        for (int k = 1; k < 4; k++)
                            if (left(k) > left(k0)) k0 = k;
                        if (left(k0) == 0) break;
                        for (int k = 0; k < 4; k++) {
                            int src = k;
                            if (left(src) == 0) {  // residue k exhausted: the fullest other one
                                src = 0;
                                for (int j = 1; j < 4; j++)
                                    if (left(j) > left(src)) src = j;
                            }
                            hg.blist.push_back((unsigned char)(left(src) ? byres[src][head[src]++] : SM_CH));
                        }
                    }
                    hg.brange.push_back(make_int2(off, (int)hg.blist.size() - base - off));
                }
                const int len = (int)hg.blist.size() - base;
                if (len > SM_LB) throw Error(-1, "internal: scatter block lists exceed their staging area");
                hg.crange.push_back(make_int2(base, len));
                hg.achunk.push_back(make_int4(t, te.start + c0, cn, 0));
            }
        }
        // the scatter stages SM_LB bytes from each chunk's first list
        hg.blist.insert(hg.blist.end(), SM_LB, (unsigned char)SM_CH);
    }
    // G cells inside some window (direct or mirrored), per row of G: the shortest
    // circular run of columns holding them.  The column passes skip the other
    // cells, and the scatter runs only over the tiles that meet a run.
    {
        std::vector<unsigned char> start((size_t)N * N, 0);  // window column starts per row
        for (size_t id = 0; id < ni; id++) {
            if (ftile[id] < 0) continue;
            const SampleGeom &g = hg.geo[id];
            for (int a = 0; a < KW; a++) start[(size_t)((g.row0 + a) % N) * N + g.col0] = 1;
        }
        hg.runs.assign(N, make_int2(0, 0));
        std::vector<unsigned char> cov(N);
        for (int r = 0; r < N; r++) {
            const unsigned char *s = &start[(size_t)r * N];
            int ncov = 0;
            for (int c = 0; c < N; c++) {
                unsigned char v = 0;
                for (int b = 0; b < KW; b++) v |= s[(c - b + N) % N];
                cov[c] = v;
                ncov += v;
            }
            if (ncov == 0) continue;
            if (ncov == N) {
                hg.runs[r] = make_int2(0, N);
                continue;
            }
            // the longest circular gap; scanning from a covered cell keeps gaps unbroken
            int c0 = 0;
            while (!cov[c0]) c0++;
            int best = 0, bstart = 0, run = 0, rstart = 0;
            for (int k = 1; k <= N; k++) {
                const int c = (c0 + k) % N;
                if (cov[c]) {
                    run = 0;
                    continue;
                }
                if (run == 0) rstart = c;
                if (++run > best) {
                    best = run;
                    bstart = rstart;
                }
            }
            hg.runs[r] = make_int2((bstart + best) % N, N - best);
        }
        auto in_run_h = [&](int r, int c) {
            int d = c - hg.runs[r].x;
            if (d < 0) d += N;
            return d < hg.runs[r].y;
        };
        std::vector<TileEntry> act;
        for (const TileEntry &te : hg.atiles) {
            const int ty0 = (te.tile / nT) * TS, tx0 = (te.tile % nT) * TS;
            bool meets = te.count > 0;
            for (int r = ty0; r < std::min(ty0 + TS, N) && !meets; r++)
                for (int c = tx0; c < std::min(tx0 + TS, N) && !meets; c++) meets = in_run_h(r, c);
            if (meets) act.push_back(te);
        }
        hg.atiles.swap(act);
    }
    // per-ray tables (takeshift kernels_radonusfft.cuh:191-198 with c = -/+(center - n/2),
    // fftshift signs, 1/n, PI/mu, 1/sqrt(n ntheta), ntheta/sqrt(M))
    const float cfl = (float)(center - n / 2.0);
    const double cfw = -(double)cfl, cad = (double)cfl;
    const double cst = (PI_REF / mud) / n / std::sqrt((double)n * nt) * nt / std::sqrt((double)M);
    hg.pre_f.resize(n);
    hg.post_f.resize(n);
    hg.pre_a.resize(n);
    hg.post_a.resize(n);
    hg.post_aw.resize(n);
    for (int s = 0; s < n; s++) {
        const double sg = 1 - 2 * ((s + 1) % 2);
        const double af = 2 * PI_REF * cfw * (s - n / 2.0) / n;
        const double aa = 2 * PI_REF * cad * (s - n / 2.0) / n;
        hg.pre_f[s] = s == 0 ? make_float2(0.f, 0.f) : make_float2((float)(sg * std::cos(af)), (float)(sg * std::sin(af)));
        hg.post_f[s] = (float)(sg * cst);
        hg.pre_a[s] = (float)sg;
        hg.post_a[s] = s == 0 ? make_float2(0.f, 0.f)
                              : make_float2((float)(sg * cst * std::cos(aa)), (float)(sg * cst * std::sin(aa)));
        // the filter acts on the centred spectrum right after the forward FFT, where
        // the `filter` argument of radonusfft::adjR (radonusfft.cu:106) would apply it;
        // the ray tables' basis scale ntheta/sqrt(M) is divided out (the static basis
        // below carries no such scale)
        const double wf = fbp_filter(s, n) * std::sqrt((double)M) / nt;
        hg.post_aw[s] = s == 0 ? make_float2(0.f, 0.f)
                               : make_float2((float)(sg * cst * wf * std::cos(aa)), (float)(sg * cst * wf * std::sin(aa)));
    }
    // deapodisation 1/(2n exp(-mu (x - n/2)^2)) per axis (divphi, kernels_radonusfft.cuh:11-22)
    hg.deap.resize(n);
    for (int x = 0; x < n; x++) {
        const double t = x - n / 2;
        hg.deap[x] = (float)(1.0 / (2.0 * n * std::exp(-mud * t * t)));
    }
    // pair coefficients.  D[k][i] = sum_j phi_j[k] conj(phi_j[t_i]), t_i = i*nt/M (int)
    // (mulphi with decphi / recphi, kernels_operators.cuh:99-149); pair p = (a, b) =
    // (2p, 2p+1): cf[k][0][p] = (D_a - i D_b)/2, cf[k][1][p] = conj((D_a + i D_b)/2).
    const int MP = (M + 1) / 2;
    hg.cf.assign((size_t)nt * 2 * MP, make_float2(0.f, 0.f));
    std::vector<double> pr((size_t)M * nt), pim((size_t)M * nt);
    for (int j = 0; j < M; j++)
        for (int k = 0; k < nt; k++) {
            pr[(size_t)j * nt + k] = phi[(size_t)j * 2 * nt + 2 * k];
            pim[(size_t)j * nt + k] = phi[(size_t)j * 2 * nt + 2 * k + 1];
        }
    // a basis selecting one static frame: D[k][0] = 1, D[k][i > 0] = 0, i.e. pair 0 with
    // alpha = beta = 1/2 (frame 0 of the adjoint frame pairs is Re R^* g)
    hg.cf_static.assign((size_t)nt * 2 * MP, make_float2(0.f, 0.f));
    for (int k = 0; k < nt; k++) {
        hg.cf_static[((size_t)k * 2 + 0) * MP] = make_float2(0.5f, 0.f);
        hg.cf_static[((size_t)k * 2 + 1) * MP] = make_float2(0.5f, 0.f);
    }
    for (int k = 0; k < nt; k++) {
        std::vector<double> dr(2 * MP, 0.0), di(2 * MP, 0.0);
        for (int i = 0; i < M; i++) {
            const int ti = (int)((long long)i * nt / M);
            for (int j = 0; j < M; j++) {
                const double ar = pr[(size_t)j * nt + k], ai = pim[(size_t)j * nt + k];
                const double br = pr[(size_t)j * nt + ti], bi = -pim[(size_t)j * nt + ti];
                dr[i] += ar * br - ai * bi;
                di[i] += ar * bi + ai * br;
            }
        }
        for (int p = 0; p < MP; p++) {
            const double ar = dr[2 * p], ai = di[2 * p], br = dr[2 * p + 1], bi = di[2 * p + 1];
            // alpha = (D_a - i D_b)/2, beta = (D_a + i D_b)/2
            hg.cf[((size_t)k * 2 + 0) * MP + p] = make_float2((float)(0.5 * (ar + bi)), (float)(0.5 * (ai - br)));
            hg.cf[((size_t)k * 2 + 1) * MP + p] = make_float2((float)(0.5 * (ar - bi)), (float)(-0.5 * (ai + br)));
        }
    }
}

// ---------------------------------------------------------------- device buffers
template <typename T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    void alloc(size_t count, size_t &total) {
        release(total);
        if (count == 0) return;
        hipError_t e = hipMalloc((void **)&p, count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            throw Error(-3, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed: " + hipGetErrorString(e));
        }
        n = count;
        total += count * sizeof(T);
    }
    void release(size_t &total) {
        if (p) {
            (void)hipFree(p);
            total -= n * sizeof(T);
        }
        p = nullptr;
        n = 0;
    }
};

template <typename T>
static void upload(DBuf<T> &b, const std::vector<T> &v, size_t &total, hipStream_t st) {
    b.alloc(std::max<size_t>(v.size(), 1), total);
    if (!v.empty()) RTV_HIP(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
}

struct Solver {
    int n, nt, M, nz, nzp, nch, ngpus, N;
    int MP = 1;  // frame pairs (spectra per slice)
    int dev = 0;
    hipStream_t st = nullptr;
    size_t bytes = 0;
    bool freed = false;
    int lds_max = 65536;
    // FFT
    FftPlan plan_n{}, plan_N{};
    bool odd_n = false, odd_N = false;  // lengths with factors 3, 5, 7
    bool p2N = false;                   // 2n has a compile-time power-of-two kernel
    bool rpN = false;                   // 2n has a compile-time kernel of length 3 * 2^k or 5 * 2^k
    DBuf<float2> d_tw;
    int T_rows = 1, T_cols = 1, T_fcols = 1, T_acols = 1, T_ray = 1, T_arows = 1, st_n = 0, st_N = 0;
    // state
    DBuf<float> d_f[3];
    int ia = 0;
    // d_psi holds chi = mu - rho0 psi (rho0 = d_rho[1], the penalty of the prox that
    // formed it), not psi: the dual part of the inner steps needs psi only through chi
    DBuf<float4> d_psi, d_mu;
    DBuf<float> d_dchi;  // div(mu - rho psi) for the current outer iteration
    DBuf<float> d_g;
    // Views of psi, mu, dchi and the data used by the ADMM kernels, which address
    // them by global chunk / slice index, and the chunk range [c_lo, c_lo + c_n)
    // they run over: the device arrays and all chunks when the state is resident;
    // out of core, the device window arrays offset back by the window's first
    // chunk (never dereferenced outside the window) and the window's chunks.
    float4 *v_psi = nullptr, *v_mu = nullptr;
    float *v_dchi = nullptr;
    const float *v_g = nullptr;
    int c_lo = 0, c_n = 0;
    // Out of core (volumes whose state does not fit in device memory, rectv.cu:12-16,
    // 143-197): f, fn, psi, mu and the data stay in pinned host memory and each outer
    // iteration streams windows of W chunks (plus one halo slice each side) through
    // the device; the norms of all windows are summed before the rho rule.
    bool ooc = false;
    int W = 0;
    float *h_f[2] = {nullptr, nullptr};  // h_f[hf]: current iterate, h_f[hf ^ 1]: the next
    int hf = 0;
    float4 *h_psi = nullptr, *h_mu = nullptr;
    float *h_g = nullptr;
    DBuf<double> d_acc;  // norms summed over the windows
    // a second window slot, so that the copies of the neighbouring windows run on
    // their own streams while a window computes (slot 0: d_f, d_psi, d_mu, d_dchi, d_g)
    DBuf<float> d_f2[3], d_dchi2, d_g2;
    DBuf<float4> d_psi2, d_mu2;
    hipStream_t s_in = nullptr, s_out = nullptr;
    hipEvent_t ev_in[2] = {nullptr, nullptr}, ev_comp[2] = {nullptr, nullptr}, ev_out[2] = {nullptr, nullptr};
    // penalty rho, its rule and the per-iteration norms live on the device
    // (k_rho_rule), so an outer iteration needs no host round trip
    DBuf<float> d_rho;
    DBuf<double> d_hist;  // [HIST_CAP][HIST_W] ring: r, s, rho used, ||fn - f||^2, ||fn||^2, data term, sum |grad fn|
    DBuf<int> d_iter;     // rows written since reset
    static constexpr int HIST_CAP = 1024;
    long long iters_done = 0;  // outer iterations enqueued since reset
    // one captured outer iteration per rotation state of the three f buffers
    bool use_graph = false;
    hipGraphExec_t gexec[3] = {nullptr, nullptr, nullptr};
    int gk_titer[3] = {0, 0, 0};
    float gk_lam0[3] = {0.f, 0.f, 0.f}, gk_lam1[3] = {0.f, 0.f, 0.f};
    // launch shapes chosen by autotune() (results are bitwise independent of them)
    int gather_zg = GM_ZG, scatter_zw = SM_ZW;
    // scratch (one batch of cb chunks = zb slices)
    int cb = 1, zb = 1;
    DBuf<float2> d_H, d_G, d_v, d_gq, d_q;
    DBuf<float> d_h1;
    DBuf<double> d_part, d_red;
    DBuf<double> d_rpart;  // per ray workgroup: ||R f - g||^2 partials of the first inner step
    // geometry
    bool inputs = false;
    DBuf<SampleGeom> d_geo;
    DBuf<int> d_fitems;
    DBuf<unsigned short> d_fcells;
    DBuf<AdjItem> d_aitems;
    DBuf<int2> d_arec;
    DBuf<TileEntry> d_fmtiles, d_atiles;
    DBuf<int2> d_runs;
    DBuf<int2> d_crange, d_brange;
    DBuf<unsigned char> d_blist;
    DBuf<int4> d_achunk;
    DBuf<float4> d_wtab;  // scatter weight rows per adjoint chunk (8 float4 per slot)
    DBuf<float2> d_cf, d_pre_f, d_post_a;
    DBuf<float2> d_post_aw, d_cf_static;  // filtered back-projection tables
    DBuf<int2> d_supp;                    // support mask: kept x range per row
    float support_r = -1.f;               // < 0: no support mask
    DBuf<float> d_post_f, d_pre_a, d_deap;
    int nfm_entries = 0, na_tiles = 0;
    float beta = 0.f;
    // profiling
    bool prof = false;
    struct Rec {
        const char *name;
        hipEvent_t a, b;
    };
    std::vector<Rec> pending;
    std::vector<hipEvent_t> evpool;
    std::map<std::string, std::pair<long long, double>> stats;

    Solver(int n_, int nt_, int m_, int nz_, int nzp_, int ngpus_) : n(n_), nt(nt_), M(m_), nz(nz_), nzp(nzp_) {
        RTV_REQUIRE(n > 0 && nt > 0 && M > 0 && nz > 0 && nzp > 0, "all sizes must be positive");
        RTV_REQUIRE(nz % nzp == 0, "nz must be a multiple of nzp (the reference silently drops the remainder)");
        RTV_REQUIRE(n >= 16, "n must be at least 16");
        nch = nz / nzp;
        MP = (M + 1) / 2;
        ngpus = std::max(1, std::min(ngpus_, nch));
        try {
            init_device();
        } catch (...) {
            release();
            throw;
        }
    }

    // Device bytes of a solver whose slice batch holds all chunks (what init_device
    // allocates when that fits), with an upper bound for the geometry tables.
    static size_t plan_bytes(int n, int nt, int M, int nz, int nzp) {
        const size_t N = 2 * (size_t)n, MP = (M + 1) / 2, nch = nz / nzp;
        const size_t V = (size_t)nz * M * n * n;
        const size_t V4 = nch * (nzp + 1) * (M + 1) * (size_t)(n + 1) * (n + 1);
        const size_t state = 3 * V * 4 + 2 * V4 * 16 + V * 4 + (size_t)nz * nt * n * 4;
        const size_t per_slice = MP * n * (N + 2) * 8 + MP * N * N * 8 + MP * n * n * 8 + (size_t)nt * n * 24 + (size_t)nt * n * 4;
        const size_t ni = 2 * (size_t)nt * n, ntile = (size_t)tiles_per_side(N) * tiles_per_side(N);
        // geo 16 B, gather slots <= 16 ints, adjoint lists <= 4 records of 16 + 8 B and
        // block lists <= 9 bytes per item (16 B spare); per adjoint chunk its range, 16
        // block ranges, up to 3 padding bytes per block, its record range and SM_CH weight
        // rows of 128 B; the gather entries' cell lists (at most ntile + ni / 16 entries of
        // at most TH * TH + 1 shorts); tables per ray / pixel and per row of the oversampled grid
        const size_t nchk = ntile + 4 * ni / SM_CH + 1;
        const size_t geom = ni * (16 + 16 + 64 + 32 + 64 + 9) + nchk * (8 + 16 * 8 + 16 * 3 + 16 + SM_CH * 128) +
                            3 * ntile * 16 + (ntile + ni / 16 + 1) * (TH * TH + 1) * 2 +
                            (size_t)nt * 2 * MP * 8 + 8 * (size_t)n * 8 + 2 * (N + n) * 8 + N * 8;
        // partial sums of the prox (per workgroup) and of the ray passes (per ray, at most)
        const size_t pb = (size_t)((n + 1 + PX_ROWS - 1) / PX_ROWS) * (M + 1) * (nzp + 1) * nch;
        const size_t parts = (PROX_W * (pb + (size_t)nz * nt) + 16 + (size_t)nz * nt + 1) * 8;
        return state + per_slice * nz + geom + parts + (1 << 20);
    }

    void init_device() {
        N = 2 * n;
        RTV_REQUIRE(make_plan(n, plan_n) && make_plan(N, plan_N),
                    "n must have no prime factor above 127 (FFT sizes n and 2n)");
        odd_n = (n & (n - 1)) != 0;
        odd_N = (N & (N - 1)) != 0;
        int ndev = 0;
        RTV_HIP(hipGetDeviceCount(&ndev));
        RTV_REQUIRE(ndev > 0, "no HIP device available");
        RTV_HIP(hipGetDevice(&dev));
        int lm = 0;
        if (hipDeviceGetAttribute(&lm, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && lm > 0)
            lds_max = lm;
        if (const char *e = std::getenv("RECTV_LDS_MAX")) lds_max = std::min(lds_max, std::atoi(e));
        lds_max -= 1024;  // room for the kernels' static shared arrays
        RTV_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        setup_fft();
        // persistent state: resident, or out of core in windows of W chunks when asked
        // (RECTV_STREAM_CHUNKS = chunks per window) or when the resident state and one
        // chunk of operator scratch would not fit in free device memory
        const size_t S = (size_t)M * n * n, D = (size_t)(nzp + 1) * (M + 1) * (size_t)(n + 1) * (n + 1);
        const size_t V = (size_t)nz * S, V4 = (size_t)nch * D;
        const size_t per_slice = (size_t)MP * n * (N + 2) * 8 + (size_t)MP * N * N * 8 + (size_t)MP * n * n * 8 +
                                 (size_t)nt * n * 24 + (size_t)nt * n * 4;
        // device bytes per chunk of state: f three times and div chi, psi, mu, data; out of
        // core a window is held twice (two slots)
        const size_t per_chunk = 4 * (size_t)nzp * S * 4 + 2 * D * 16 + (size_t)nzp * nt * n * 4;
        size_t freeb = 0, totb = 0;
        RTV_HIP(hipMemGetInfo(&freeb, &totb));
        W = nch;
        if (const char *e = std::getenv("RECTV_STREAM_CHUNKS")) {
            if (std::atoi(e) > 0) W = std::min(nch, std::atoi(e));
        } else if ((double)per_chunk * nch + (double)per_slice * nzp > 0.9 * (double)freeb) {
            W = (int)std::max(1.0, std::min((double)nch, std::floor(0.55 * (double)freeb /
                                                                     (2.0 * per_chunk + (double)per_slice * nzp))));
        }
        ooc = W < nch;
        const size_t ws = ooc ? (size_t)W * nzp + 2 : (size_t)nz;  // f slices on the device (window + 2 halos)
        for (auto &b : d_f) b.alloc(ws * S, bytes);
        d_psi.alloc(ooc ? W * D : V4, bytes);
        d_mu.alloc(ooc ? W * D : V4, bytes);
        d_dchi.alloc(ooc ? (size_t)W * nzp * S : V, bytes);
        d_g.alloc((ooc ? (size_t)W * nzp : (size_t)nz) * nt * n, bytes);
        if (ooc) {
            auto host = [&](void **p, size_t b) {
                if (hipHostMalloc(p, b, hipHostMallocDefault) != hipSuccess || !*p) {
                    *p = nullptr;
                    throw Error(-3, "pinned host allocation of " + std::to_string(b) +
                                        " bytes for the out-of-core state failed");
                }
            };
            for (auto &h : h_f) host((void **)&h, V * sizeof(float));
            host((void **)&h_psi, V4 * sizeof(float4));
            host((void **)&h_mu, V4 * sizeof(float4));
            host((void **)&h_g, (size_t)nz * nt * n * sizeof(float));
            d_acc.alloc(8, bytes);
            for (auto &b : d_f2) b.alloc(ws * S, bytes);
            d_psi2.alloc(W * D, bytes);
            d_mu2.alloc(W * D, bytes);
            d_dchi2.alloc((size_t)W * nzp * S, bytes);
            d_g2.alloc((size_t)W * nzp * nt * n, bytes);
            RTV_HIP(hipStreamCreateWithFlags(&s_in, hipStreamNonBlocking));
            RTV_HIP(hipStreamCreateWithFlags(&s_out, hipStreamNonBlocking));
            for (int q = 0; q < 2; q++)
                for (hipEvent_t *e : {&ev_in[q], &ev_comp[q], &ev_out[q]})
                    RTV_HIP(hipEventCreateWithFlags(e, hipEventDisableTiming));
        }
        v_psi = d_psi.p;
        v_mu = d_mu.p;
        v_dchi = d_dchi.p;
        v_g = d_g.p;
        c_lo = 0;
        c_n = nch;
        // scratch plan
        RTV_HIP(hipMemGetInfo(&freeb, &totb));
        double budget = 0.45 * (double)freeb;
        if (const char *e = std::getenv("RECTV_SCRATCH_MB")) budget = std::atof(e) * 1048576.0;
        cb = (int)std::max(1.0, std::min((double)W, std::floor(budget / ((double)per_slice * nzp))));
        zb = cb * nzp;
        d_H.alloc((size_t)zb * MP * n * ((N + 3) & ~3), bytes);  // whole 4-column groups (odd n: 2n = 2 mod 4)
        d_G.alloc((size_t)zb * MP * N * N, bytes);
        d_v.alloc((size_t)zb * MP * n * n, bytes);
        d_gq.alloc((size_t)zb * nt * n * 2, bytes);
        d_q.alloc((size_t)zb * nt * n, bytes);
        d_h1.alloc((size_t)zb * nt * n, bytes);
        const size_t pb = (size_t)((n + 1 + PX_ROWS - 1) / PX_ROWS) * (M + 1) * (nzp + 1) * nch;  // prox workgroups
        const size_t rb = (size_t)((zb * nt + T_ray - 1) / T_ray);
        d_part.alloc(PROX_W * std::max(pb, rb) + 16, bytes);
        d_rpart.alloc((size_t)((nch + cb - 1) / cb) * rb + 1, bytes);
        d_red.alloc(8, bytes);
        d_rho.alloc(2, bytes);  // rho, and the rho the stored chi was formed with
        d_hist.alloc((size_t)HIST_CAP * HIST_W, bytes);
        d_iter.alloc(1, bytes);
        RTV_HIP(hipStreamSynchronize(st));
        if (std::getenv("RECTV_VERBOSE"))
            fprintf(stderr,
                    "rectv: n=%d ntheta=%d m=%d nz=%d nzp=%d lds_max=%d T_rows=%d T_cols=%d/%d T_ray=%d T_arows=%d "
                    "chunks/batch=%d device bytes=%zu%s\n",
                    n, nt, M, nz, nzp, lds_max, T_rows, T_fcols, T_acols, T_ray, T_arows, cb, bytes,
                    ooc ? (" (out of core: " + std::to_string(W) + " chunks per window)").c_str() : "");
    }

    int fft_threads(int T, int len) const {
        int t = T * fft_tpt(len, FFT_E);
        t = ((t + 63) / 64) * 64;
        return std::max(64, t);
    }

    void setup_fft() {
        // twiddle tables exp(-2 pi i m / L) in FP64 -> float for L = n and 2n
        std::vector<float2> tw;
        auto add = [&](FftPlan &p) {
            p.tw_off = (int)tw.size();
            for (int m = 0; m < p.n; m++) {
                const double a = -2.0 * PI_EXACT * (double)m / p.n;
                tw.push_back(make_float2((float)std::cos(a), (float)std::sin(a)));
            }
        };
        add(plan_n);
        add(plan_N);
        rpN = rp_kernel(N);
        if (rpN) {  // the P-point table of the length-R0 * P kernels
            FftPlan sub{};
            sub.n = N / rp_radix(N);
            add(sub);
            plan_N.tw_sub = sub.tw_off;
        }
        upload(d_tw, tw, bytes, st);
        st_n = lp_stride(n);
        st_N = fft_row_pitch(N);
        // transforms per workgroup: bounded by LDS and by FFT_THREADS threads
        auto tmax = [&](int len, int stride) {
            return std::max(1, std::min(std::max(1, FFT_WG / fft_tpt(len, FFT_E)), lds_max / (8 * stride)));
        };
        RTV_REQUIRE(fft_tpt(N, FFT_E) <= FFT_THREADS, "2n-point transform needs too many threads");
        RTV_REQUIRE(!has_big_prime(N) || fft_tpt(N, FFT_E) <= FFT_WG,
                    "2n-point transforms with a prime factor above 7 are limited to 2n <= 8192");
        const int tN = tmax(N, st_N), tn = tmax(n, st_n);
        RTV_REQUIRE(tN >= 1 && 8 * st_N <= lds_max, "2n-point transform does not fit in LDS");
        T_cols = 1;
        while (T_cols * 2 <= std::min(tN, 16) && N % (T_cols * 2) == 0) T_cols *= 2;
        T_arows = T_rows = std::min(tN, 16);
        T_fcols = T_acols = T_cols;
        T_ray = std::min(tn, 16);
        // power-of-two lengths use compile-time row counts inside the kernels
        p2N = (N & (N - 1)) == 0 && N >= 64 && N <= 4096;
        if (rpN) {
            T_arows = T_rows = fft_rp_count(N, FFT_WG);
            T_fcols = fft_rp_count(N, RTV_FCOLS_THREADS);
            T_acols = fft_rp_count(N, RTV_ACOLS_THREADS);
            T_cols = std::max(T_fcols, T_acols);
            RTV_REQUIRE(8 * st_N * T_cols <= lds_max && 8 * st_N * T_rows <= lds_max,
                        "LDS too small for the 2n-point transforms");
        }
        if (p2N) {
            T_arows = T_rows = fft_rows_p2(N);
            T_fcols = fft_cols_p2(N, RTV_FCOLS_THREADS);
            T_acols = fft_cols_p2(N, RTV_ACOLS_THREADS);
            T_cols = std::max(T_fcols, T_acols);
            RTV_REQUIRE(8 * st_N * T_cols <= lds_max && 8 * st_N * T_rows <= lds_max,
                        "LDS too small for the 2n-point transforms");
        }
        if ((n & (n - 1)) == 0 && n >= 64 && n <= 4096) {
            T_ray = fft_rows_p2(n, RAY_WG);
            RTV_REQUIRE(8 * st_n * T_ray <= lds_max, "LDS too small for the n-point transforms");
        }
        const int big = lds_max;
        const void *fns[] = {(const void *)pick_fwd_rows(N), (const void *)pick_fwd_cols(N),
                             (const void *)pick_adj_cols(N), (const void *)pick_adj_rows(N), (const void *)pick_ray(n)};
        for (const void *fn : fns) RTV_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, big));
    }

    // ------------------------------------------------------------ profiling
    hipEvent_t ev() {
        if (evpool.empty()) {
            hipEvent_t e;
            RTV_HIP(hipEventCreate(&e));
            return e;
        }
        hipEvent_t e = evpool.back();
        evpool.pop_back();
        return e;
    }
    template <typename F>
    void launch(const char *name, F &&f) {
        if (!prof) {
            f();
        } else {
            hipEvent_t a = ev(), b = ev();
            RTV_HIP(hipEventRecord(a, st));
            f();
            RTV_HIP(hipEventRecord(b, st));
            pending.push_back(Rec{name, a, b});
        }
        RTV_HIP(hipGetLastError());
    }
    void collect() {
        if (pending.empty()) return;
        RTV_HIP(hipStreamSynchronize(st));
        for (auto &r : pending) {
            float ms = 0.f;
            RTV_HIP(hipEventElapsedTime(&ms, r.a, r.b));
            auto &s = stats[r.name];
            s.first += 1;
            s.second += ms;
            evpool.push_back(r.a);
            evpool.push_back(r.b);
        }
        pending.clear();
    }

    // ------------------------------------------------------------ inputs
    void set_inputs(const float *g, const float *theta, const float *phi, float center) {
        RTV_REQUIRE(g && theta && phi, "null input pointer");
        for (size_t i = 0; i < (size_t)M * 2 * nt; i++)
            RTV_REQUIRE(std::isfinite(phi[i]), "phi contains non-finite values");
        for (int k = 0; k < nt; k++) RTV_REQUIRE(std::isfinite(theta[k]), "theta contains non-finite values");
        RTV_REQUIRE(std::isfinite(center), "center must be finite");
        drop_graphs();  // the geometry buffers below are reallocated
        HostGeom hg;
        build_geometry(hg, n, nt, M, theta, phi, center);
        beta = hg.beta;
        upload(d_geo, hg.geo, bytes, st);
        upload(d_fitems, hg.fitems, bytes, st);
        upload(d_fcells, hg.fcells, bytes, st);
        upload(d_fmtiles, hg.fmtiles, bytes, st);
        nfm_entries = (int)hg.fmtiles.size();
        if (std::getenv("RECTV_VERBOSE")) {
            size_t real = 0;
            for (int v : hg.fitems) real += v >= 0;
            fprintf(stderr, "rectv: gather items %zu in %zu slots, %d entries\n", real, hg.fitems.size(), nfm_entries);
        }
        upload(d_aitems, hg.aitems, bytes, st);
        upload(d_arec, hg.arec, bytes, st);
        upload(d_atiles, hg.atiles, bytes, st);
        upload(d_crange, hg.crange, bytes, st);
        upload(d_brange, hg.brange, bytes, st);
        upload(d_blist, hg.blist, bytes, st);
        upload(d_achunk, hg.achunk, bytes, st);
        upload(d_cf, hg.cf, bytes, st);
        upload(d_pre_f, hg.pre_f, bytes, st);
        upload(d_post_f, hg.post_f, bytes, st);
        upload(d_pre_a, hg.pre_a, bytes, st);
        upload(d_post_a, hg.post_a, bytes, st);
        upload(d_post_aw, hg.post_aw, bytes, st);
        upload(d_cf_static, hg.cf_static, bytes, st);
        upload(d_deap, hg.deap, bytes, st);
        upload(d_runs, hg.runs, bytes, st);
        na_tiles = (int)hg.atiles.size();
        // G is zero outside the windows of this geometry (the column passes and the
        // scatter keep it so)
        RTV_HIP(hipMemsetAsync(d_G.p, 0, d_G.n * sizeof(float2), st));
        // scatter weight rows, computed once per geometry with the gather's window function
        const int nchunks = (int)hg.achunk.size();
        d_wtab.alloc((size_t)std::max(nchunks, 1) * SM_CH * 8, bytes);
        if (nchunks > 0)
            launch("scatter_weights", [&] {
                hipLaunchKernelGGL(k_scatter_weights, dim3(nchunks), dim3(1024), 0, st, d_aitems.p, d_achunk.p,
                                   d_wtab.p, n, beta);
            });
        if (ooc) std::memcpy(h_g, g, (size_t)nz * nt * n * sizeof(float));
        else RTV_HIP(hipMemcpyAsync(d_g.p, g, (size_t)nz * nt * n * sizeof(float), hipMemcpyHostToDevice, st));
        RTV_HIP(hipStreamSynchronize(st));
        inputs = true;
    }

    void reset() {
        const size_t V = (size_t)nz * M * n * n;
        if (ooc) {
            std::memset(h_f[0], 0, V * sizeof(float));
            std::memset(h_psi, 0, (size_t)nch * dual_chunk() * sizeof(float4));
            std::memset(h_mu, 0, (size_t)nch * dual_chunk() * sizeof(float4));
            hf = 0;
        }
        for (auto &b : d_f) RTV_HIP(hipMemsetAsync(b.p, 0, b.n * sizeof(float), st));
        RTV_HIP(hipMemsetAsync(d_psi.p, 0, d_psi.n * sizeof(float4), st));
        RTV_HIP(hipMemsetAsync(d_mu.p, 0, d_mu.n * sizeof(float4), st));
        RTV_HIP(hipMemsetAsync(d_dchi.p, 0, d_dchi.n * sizeof(float), st));  // psi = mu = 0
        static const float rho0[2] = {0.5f, 0.5f};  // rectv.cu:125
        static const int zero = 0;
        RTV_HIP(hipMemcpyAsync(d_rho.p, rho0, 2 * sizeof(float), hipMemcpyHostToDevice, st));
        RTV_HIP(hipMemcpyAsync(d_iter.p, &zero, sizeof(int), hipMemcpyHostToDevice, st));
        ia = 0;
        iters_done = 0;
        RTV_HIP(hipStreamSynchronize(st));
    }

    void drop_graphs() {
        for (auto &g : gexec) {
            if (g) (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    }

    // ------------------------------------------------------------ operator pipeline
    // forward: f slices [z0, z0+zc) -> polar samples in d_gq (and h1 / residual / q in d_q per mode)
    void forward_to_rays(const float *f, int zc, const float *gdat, float *h1out, int mode, double *partial) {
        const size_t S = (size_t)M * n * n;
        (void)S;
        {
            // power-of-two lengths: (slice, pair, T_rows image rows); else (slice, row, T_rows pairs)
            const int blocks = (p2N || rpN) ? zc * MP * (n / T_rows) : zc * n * ((MP + T_rows - 1) / T_rows);
            const int thr = fft_threads(T_rows, N);
            const size_t lds = (size_t)T_rows * st_N * sizeof(float2);
            launch("fwd_rows", [&] {
                hipLaunchKernelGGL(pick_fwd_rows(N), dim3(blocks), dim3(thr), lds, st, f, d_H.p, d_deap.p, plan_N,
                                   d_tw.p, n, M, MP, T_rows, st_N);
            });
        }
        {
            const int blocks = zc * MP * (N / T_fcols);
            const int thr = fft_threads(T_fcols, N);
            const size_t lds = (size_t)T_fcols * st_N * sizeof(float2);
            launch("fwd_cols", [&] {
                hipLaunchKernelGGL(pick_fwd_cols(N), dim3(blocks), dim3(thr), lds, st, d_H.p, d_G.p, plan_N, d_tw.p,
                                   d_runs.p, n, MP, T_fcols, st_N);
            });
        }
        gather(zc, gather_zg);
        ray(zc, gdat, nullptr, h1out, mode, partial);
    }

    // Pairs per gather launch: at most 16 (two 16-column blocks), and few enough that
    // the kernel's 32-bit byte offsets within one slice's pair planes stay in range.
    int gather_group() const {
        const unsigned long long plane = 8ull * (2ull * n) * (2ull * n);
        return (int)std::max(1ull, std::min(16ull, 0xFFFFFFFFull / plane));
    }
    void gather(int zc, int zg) {
        const int blocks = nfm_entries * ((zc + zg - 1) / zg);
        if (blocks <= 0) return;
        const int pg = gather_group();
        for (int p0 = 0; p0 < MP; p0 += pg) {
            const int np = std::min(pg, MP - p0), acc = p0 > 0 ? 1 : 0;
            // few pairs: 2 or 4 slices share the 16 matrix-core columns
            const int pk = MP <= 2 ? 4 : MP <= 4 ? 2 : 1;
            launch("gather", [&] {
#define RTV_GATHER(NB_, PK_)                                                                                          \
    hipLaunchKernelGGL((k_gather_mfma<NB_, PK_>), dim3(blocks), dim3(GM_THREADS), 0, st, d_G.p, d_gq.p, d_geo.p,      \
                       d_fitems.p, d_fmtiles.p, d_fcells.p, d_cf.p, n, nt, MP, p0, acc, beta, nfm_entries, zc, zg)
                if (pk == 4) RTV_GATHER(1, 4);
                else if (pk == 2) RTV_GATHER(1, 2);
                else if (np <= 8) RTV_GATHER(1, 1);
                else RTV_GATHER(2, 1);
#undef RTV_GATHER
            });
        }
    }

    // one launch per group of at most 16 pairs (one or two 16-column blocks)
    template <int ZW>
    void scatter_t(int zc, const float2 *cf) {
        const int blocks = na_tiles * ((zc + ZW - 1) / ZW);
        for (int p0 = 0; p0 < MP; p0 += 16) {
            const int np = std::min(16, MP - p0);
            // few pairs: 2 or 4 of the workgroup's slices share the 16 matrix-core columns
            const int pk = (MP <= 2 && ZW >= 4) ? 4 : MP <= 4 ? 2 : 1;
            launch("scatter", [&] {
#define RTV_SCATTER(NB_, PK_)                                                                                   \
    hipLaunchKernelGGL((k_scatter<NB_, ZW, PK_>), dim3(blocks), dim3(1024), 0, st, d_G.p, d_q.p, d_arec.p,      \
                       d_atiles.p, d_crange.p, d_brange.p, d_blist.p, cf, d_wtab.p, n, nt, MP, p0, na_tiles, zc)
                if (pk == 4) RTV_SCATTER(1, (ZW >= 4 ? 4 : 1));
                else if (pk == 2) RTV_SCATTER(1, 2);
                else if (np <= 8) RTV_SCATTER(1, 1);
                else RTV_SCATTER(2, 1);
#undef RTV_SCATTER
            });
        }
    }
    void scatter(int zc, int zw, const float2 *cf) {
        if (zw == 2) scatter_t<2>(zc, cf);
        else if (zw == 8) scatter_t<8>(zc, cf);
        else scatter_t<4>(zc, cf);
    }

    int ray_blocks(int zc) const { return (zc * nt + T_ray - 1) / T_ray; }
    void ray(int zc, const float *gdat, const float *h1in, float *h1out, int mode, double *partial,
             const float2 *post_a = nullptr) {
        const int nrays = zc * nt;
        const int blocks = ray_blocks(zc);
        const int thr = fft_threads(T_ray, n);
        const size_t lds = (size_t)T_ray * st_n * sizeof(float2);
        const float2 *pa = post_a ? post_a : d_post_a.p;
        launch("ray", [&] {
            hipLaunchKernelGGL(pick_ray(n), dim3(blocks), dim3(thr), lds, st, d_gq.p, d_q.p, gdat, h1in, h1out, d_pre_f.p,
                               d_post_f.p, d_pre_a.p, pa, plan_n, d_tw.p, n, nrays, T_ray, st_n, mode, partial);
        });
    }

    // adjoint: d_q (q samples of zc slices) -> d_v images (cf: pair coefficients,
    // default the solver's basis)
    void rays_to_images(int zc, const float2 *cf = nullptr) {
        scatter(zc, scatter_zw, cf ? cf : d_cf.p);
        {
            const int blocks = zc * MP * (N / T_acols);
            const int thr = fft_threads(T_acols, N);
            const size_t lds = (size_t)T_acols * st_N * sizeof(float2);
            launch("adj_cols", [&] {
                hipLaunchKernelGGL(pick_adj_cols(N), dim3(blocks), dim3(thr), lds, st, d_G.p, d_H.p, plan_N, d_tw.p,
                                   d_runs.p, n, MP, T_acols, st_N);
            });
        }
        {
            const int nyt = (n + T_arows - 1) / T_arows;
            const int blocks = zc * MP * nyt;
            const int thr = fft_threads(T_arows, N);
            const size_t lds = (size_t)T_arows * st_N * sizeof(float2);
            launch("adj_rows", [&] {
                hipLaunchKernelGGL(pick_adj_rows(N), dim3(blocks), dim3(thr), lds, st, d_H.p, d_v.p, d_deap.p, plan_N,
                                   d_tw.p, n, MP, T_arows, st_N);
            });
        }
    }

    // workgroups of the TV kernels over an m x m (x, y) plane per chunk, for nc chunks
    static int tv_blocks(int m, int nc) { return ((m + 63) / 64) * ((m + TV_ROWS - 1) / TV_ROWS) * nc; }
    // slice segments per chunk for the column-marching kernels (k_update, k_divchi):
    // where the (x, y) tiles of the chunks give fewer workgroups than the device holds
    // at once (two per CU), as many segments as still fit in that one round, of at
    // least 2 slices each (each segment reads the slice before it once more)
    static int tv_zseg(int m, int nc, int nzp) {
        const int b = tv_blocks(m, nc);
        return std::max(1, std::min(nzp / 2, 2 * 256 / b));
    }

    TvArgs tv_args(float lam1, float lam0) const {
        TvArgs a;
        a.n = n;
        a.M = M;
        a.nzp = nzp;
        a.nchunk = nch;
        a.chunk0 = 0;
        const double c2 = std::cbrt(2.0);
        a.gs = (float)(1.0 / (c2 * lam1));
        a.gt = (float)(1.0 / c2);
        a.ks = (float)(1.0 / (2.0 * c2 * lam1));
        a.kt = (float)(1.0 / (2.0 * c2));
        a.rho = d_rho.p;
        a.lam0 = lam0;
        a.supp = support_r >= 0.f ? d_supp.p : nullptr;
        return a;
    }

    template <int MM>
    void launch_update(const float *cur, const float *A, float *nxt, const TvArgs &ta, int ncb, int zv0) {
        TvArgs tb = ta;
        tb.zseg = tv_zseg(n, ncb, nzp);
        const dim3 grid(tv_blocks(n, ncb) * tb.zseg);
        for (int i0 = 0; i0 < M; i0 += MM)  // groups of MM frames (one group unless M > 32)
            launch("update", [&] {
                hipLaunchKernelGGL(k_update<MM>, grid, dim3(TV_THREADS), 0, st, cur, A, nxt, v_dchi, d_v.p, tb, zv0,
                                   i0);
            });
    }
    void update(const float *cur, const float *A, float *nxt, const TvArgs &ta, int ncb, int zv0) {
        if (M <= 4) launch_update<4>(cur, A, nxt, ta, ncb, zv0);
        else if (M <= 8) launch_update<8>(cur, A, nxt, ta, ncb, zv0);
        else if (M <= 16) launch_update<16>(cur, A, nxt, ta, ncb, zv0);
        else launch_update<32>(cur, A, nxt, ta, ncb, zv0);
    }
    // prox over chunks [c_lo, c_lo + c_n), its norm partials reduced into d_red[0..4]
    void prox_partial(const float *cur, const float *A, const TvArgs &ta) {
        const dim3 grid(((n + 1 + PX_ROWS - 1) / PX_ROWS) * (M + 1) * (nzp + 1) * c_n);
        TvArgs tb = ta;
        tb.chunk0 = c_lo;
        launch("prox", [&] {
            hipLaunchKernelGGL(k_prox<true>, grid, dim3(64 * PX_WAVES), 0, st, cur, A, v_psi, v_mu, tb, d_part.p);
        });
        const int cnt = grid.x;
        launch("reduce", [&] { hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, st, d_part.p, d_red.p, cnt, PROX_W); });
    }
    void rho_rule(bool has_data) {
        launch("rho", [&] {
            hipLaunchKernelGGL(k_rho_rule, dim3(1), dim3(64), 0, st, d_red.p, d_rho.p, d_hist.p, d_iter.p, HIST_CAP,
                               has_data ? 1 : 0, d_rho.p + 1);
        });
    }
    void prox(const float *cur, const float *A, const TvArgs &ta, bool has_data) {
        prox_partial(cur, A, ta);
        rho_rule(has_data);
    }
    template <int MM>
    void launch_divchi(const TvArgs &ta) {
        TvArgs tb = ta;
        tb.chunk0 = c_lo;
        tb.zseg = tv_zseg(n, c_n, nzp);
        const dim3 grid(tv_blocks(n, c_n) * tb.zseg);
        for (int i0 = 0; i0 < M; i0 += MM)
            launch("divchi", [&] {
                hipLaunchKernelGGL((k_divchi<MM, true>), grid, dim3(TV_THREADS), 0, st, v_psi, v_mu, v_dchi, tb, i0);
            });
    }
    void divchi(const TvArgs &ta) {
        if (M <= 4) launch_divchi<4>(ta);
        else if (M <= 8) launch_divchi<8>(ta);
        else if (M <= 16) launch_divchi<16>(ta);
        else launch_divchi<32>(ta);
    }

    // ------------------------------------------------------------ ADMM driver
    // Rotation of the three f buffers over one outer iteration: inner step k reads
    // `cur` and writes `nxt`; the previous iterate A = d_f[ia] stays untouched
    // (halo slabs and the prox's f_old).
    static int next_ia(int ia0, int titer) {
        const int w1 = (ia0 + 1) % 3, w2 = (ia0 + 2) % 3;
        int cur = ia0;
        for (int k = 0; k < titer; k++) cur = (k == 0) ? w1 : (cur == w1 ? w2 : w1);
        return cur;
    }

    // One outer ADMM iteration (rectv.cu:140-246) enqueued on the stream: titer
    // inner steps over all chunks, prox/dual update with the norms, the rho rule
    // (k_rho_rule) and the dual part of the next iteration's inner steps.
    void one_iteration(int titer, float lam0, float lam1) {
        const size_t S = (size_t)M * n * n;
        const TvArgs ta = tv_args(lam1, lam0);
        const int w1 = (ia + 1) % 3, w2 = (ia + 2) % 3;
        float *A = d_f[ia].p;
        int cur = ia;
        for (int k = 0; k < titer; k++) {
            const int nxt = (k == 0) ? w1 : (cur == w1 ? w2 : w1);
            inner_step(d_f[cur].p, A, d_f[nxt].p, ta, k == 0);
            cur = nxt;
        }
        prox(d_f[cur].p, A, ta, titer > 0);  // + reduce + rho rule
        divchi(ta);                          // reads the updated rho
        ia = cur;
    }

    // Out of core: one outer iteration window by window.  A window of w chunks
    // holds its f slices in device slots 1 .. w nzp of the three f buffers, the
    // halo slices of the neighbouring chunks in slots 0 and w nzp + 1 of the first
    // (the previous iterate A), and its psi, mu, dchi and data.  Per window:
    // dchi = div(mu - rho psi) with the current rho, the titer inner steps, the
    // prox (norm partials added into d_acc), then fn, psi and mu go back to the
    // host.  After the last window the rho rule sees the norms of all chunks
    // (rectv.cu:216-237).  Chunks read only f of other chunks, so windows are
    // independent within an iteration, as the reference's chunk loop is.
    template <typename T>
    static T *shift_back(T *p, size_t elems) {  // p - elems, as an address (never dereferenced below p)
        return reinterpret_cast<T *>(reinterpret_cast<uintptr_t>(p) - elems * sizeof(T));
    }
    void one_iteration_ooc(int titer, float lam0, float lam1) {
        const size_t S = (size_t)M * n * n, D = dual_chunk(), G1 = (size_t)nt * n;
        const TvArgs ta = tv_args(lam1, lam0);
        const float *hin = h_f[hf];
        float *hout = h_f[hf ^ 1];
        struct Slot {
            float *f[3];
            float4 *psi, *mu;
            float *dchi, *g;
        };
        const Slot slot[2] = {{{d_f[0].p, d_f[1].p, d_f[2].p}, d_psi.p, d_mu.p, d_dchi.p, d_g.p},
                              {{d_f2[0].p, d_f2[1].p, d_f2[2].p}, d_psi2.p, d_mu2.p, d_dchi2.p, d_g2.p}};
        RTV_HIP(hipMemsetAsync(d_acc.p, 0, 8 * sizeof(double), st));
        // window k uses slot k & 1: its uploads (s_in) wait for the downloads of window
        // k - 2 from that slot, its kernels (st) for its uploads, its downloads (s_out)
        // for its kernels; the three streams overlap neighbouring windows
        int k = 0;
        for (int g0 = 0; g0 < nch; g0 += W, k++) {
            const int q = k & 1;
            const Slot &sl = slot[q];
            const int w = std::min(W, nch - g0);
            const size_t z0 = (size_t)g0 * nzp, zw = (size_t)w * nzp;
            const size_t lo = g0 > 0 ? z0 - 1 : z0, hi = g0 + w < nch ? z0 + zw + 1 : z0 + zw;
            if (k >= 2) RTV_HIP(hipStreamWaitEvent(s_in, ev_out[q], 0));
            RTV_HIP(hipMemcpyAsync(sl.f[0] + (lo + 1 - z0) * S, hin + lo * S, (hi - lo) * S * sizeof(float),
                                   hipMemcpyHostToDevice, s_in));
            RTV_HIP(hipMemcpyAsync(sl.psi, h_psi + g0 * D, w * D * sizeof(float4), hipMemcpyHostToDevice, s_in));
            RTV_HIP(hipMemcpyAsync(sl.mu, h_mu + g0 * D, w * D * sizeof(float4), hipMemcpyHostToDevice, s_in));
            RTV_HIP(hipMemcpyAsync(sl.g, h_g + z0 * G1, zw * G1 * sizeof(float), hipMemcpyHostToDevice, s_in));
            RTV_HIP(hipEventRecord(ev_in[q], s_in));
            RTV_HIP(hipStreamWaitEvent(st, ev_in[q], 0));
            // views addressed by global chunk / slice index
            float *fv[3];
            for (int b = 0; b < 3; b++) fv[b] = shift_back(sl.f[b] + S, z0 * S);
            v_psi = shift_back(sl.psi, g0 * D);
            v_mu = shift_back(sl.mu, g0 * D);
            v_dchi = shift_back(sl.dchi, z0 * S);
            v_g = shift_back(sl.g, z0 * G1);
            c_lo = g0;
            c_n = w;
            divchi(ta);
            int cur = 0;
            for (int it = 0; it < titer; it++) {
                const int nxt = (it == 0) ? 1 : (cur == 1 ? 2 : 1);
                inner_step(fv[cur], fv[0], fv[nxt], ta, it == 0);
                if (it == 0)
                    launch("accumulate", [&] {
                        hipLaunchKernelGGL(k_accumulate, dim3(1), dim3(64), 0, st, d_acc.p + 5, d_red.p + 5, 1);
                    });
                cur = nxt;
            }
            prox_partial(fv[cur], fv[0], ta);
            launch("accumulate", [&] {
                hipLaunchKernelGGL(k_accumulate, dim3(1), dim3(64), 0, st, d_acc.p, d_red.p, PROX_W);
            });
            RTV_HIP(hipEventRecord(ev_comp[q], st));
            RTV_HIP(hipStreamWaitEvent(s_out, ev_comp[q], 0));
            RTV_HIP(hipMemcpyAsync(hout + z0 * S, sl.f[cur] + S, zw * S * sizeof(float), hipMemcpyDeviceToHost, s_out));
            RTV_HIP(hipMemcpyAsync(h_psi + g0 * D, sl.psi, w * D * sizeof(float4), hipMemcpyDeviceToHost, s_out));
            RTV_HIP(hipMemcpyAsync(h_mu + g0 * D, sl.mu, w * D * sizeof(float4), hipMemcpyDeviceToHost, s_out));
            RTV_HIP(hipEventRecord(ev_out[q], s_out));
        }
        v_psi = d_psi.p;
        v_mu = d_mu.p;
        v_dchi = d_dchi.p;
        v_g = d_g.p;
        c_lo = 0;
        c_n = nch;
        RTV_HIP(hipMemcpyAsync(d_red.p, d_acc.p, 8 * sizeof(double), hipMemcpyDeviceToDevice, st));
        rho_rule(titer > 0);
        hf ^= 1;
        // the host arrays are read and written by the copies above: finish them before
        // the caller can read the result or the next iteration reads hout as its input
        RTV_HIP(hipStreamSynchronize(s_in));
        RTV_HIP(hipStreamSynchronize(st));
        RTV_HIP(hipStreamSynchronize(s_out));
    }

    // One inner ADMM step (admm.cu:85-97) over all chunks: nxt = cur - 1/2 div(rho
    // grad cur + chi) - 1/2 R^*(R cur - g), halo slabs from A.  With data_term the
    // ray passes also leave the partial sums of ||R cur - g||^2, reduced into d_red[5].
    void inner_step(const float *cur, const float *A, float *nxt, const TvArgs &ta, bool data_term) {
        const size_t S = (size_t)M * n * n;
        size_t roff = 0;
        for (int c0 = c_lo; c0 < c_lo + c_n; c0 += cb) {
            const int ncb = std::min(cb, c_lo + c_n - c0);
            const int z0 = c0 * nzp, zc = ncb * nzp;
            forward_to_rays(cur + z0 * S, zc, v_g + (size_t)z0 * nt * n, nullptr, RAY_FWD | RAY_SUBG | RAY_ADJ,
                            data_term ? d_rpart.p + roff : nullptr);
            roff += ray_blocks(zc);
            rays_to_images(zc);
            TvArgs tb = ta;
            tb.chunk0 = c0;
            update(cur, A, nxt, tb, ncb, z0);
        }
        if (data_term)
            launch("reduce", [&] {
                hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, st, d_rpart.p, d_red.p + 5, (int)roff, 1);
            });
    }

    void graph_iteration(int titer, float lam0, float lam1) {
        hipGraphExec_t &ge = gexec[ia];
        if (ge && (gk_titer[ia] != titer || gk_lam0[ia] != lam0 || gk_lam1[ia] != lam1)) {
            (void)hipGraphExecDestroy(ge);
            ge = nullptr;
        }
        if (!ge) {
            const int ia0 = ia;
            hipGraph_t g = nullptr;
            RTV_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            try {
                one_iteration(titer, lam0, lam1);
            } catch (...) {
                hipGraph_t junk = nullptr;
                (void)hipStreamEndCapture(st, &junk);
                if (junk) (void)hipGraphDestroy(junk);
                ia = ia0;
                throw;
            }
            RTV_HIP(hipStreamEndCapture(st, &g));
            ia = ia0;
            const hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                ge = nullptr;
                RTV_HIP(e);
            }
            gk_titer[ia] = titer;
            gk_lam0[ia] = lam0;
            gk_lam1[ia] = lam1;
        }
        RTV_HIP(hipGraphLaunch(ge, st));
        ia = next_ia(ia, titer);
    }

    // Enqueue niter outer iterations; no host synchronisation.
    void enqueue(int niter, int titer, float lam0, float lam1) {
        RTV_REQUIRE(inputs, "inputs not set (call set_inputs / run first)");
        RTV_REQUIRE(niter >= 0 && titer >= 0, "niter and titer must be >= 0");
        RTV_REQUIRE(std::isfinite(lam0) && std::isfinite(lam1) && lam1 != 0.f, "lambda1 must be finite and nonzero");
        for (int it = 0; it < niter; it++) {
            if (ooc) one_iteration_ooc(titer, lam0, lam1);
            else if (use_graph && !prof) graph_iteration(titer, lam0, lam1);
            else one_iteration(titer, lam0, lam1);
            iters_done++;
        }
    }

    // Rows [first, first + count) of the per-iteration norms (waits for the stream).
    void history(long long first, int count, double *out) {
        RTV_REQUIRE(out || count == 0, "null history buffer");
        RTV_REQUIRE(first >= 0 && count >= 0 && first + count <= iters_done,
                    "history rows requested beyond the iterations run since reset");
        RTV_REQUIRE(iters_done - first <= HIST_CAP, "history rows no longer held (ring of 1024 iterations)");
        if (count == 0) return;
        std::vector<double> ring((size_t)HIST_CAP * HIST_W);
        RTV_HIP(hipMemcpyAsync(ring.data(), d_hist.p, ring.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        RTV_HIP(hipStreamSynchronize(st));
        for (int q = 0; q < count; q++) {
            const size_t row = (size_t)((first + q) % HIST_CAP);
            for (int c = 0; c < HIST_W; c++) out[HIST_W * (size_t)q + c] = ring[HIST_W * row + c];
        }
    }

    void iterate(int niter, int titer, float lam0, float lam1, double *hist) {
        const bool quiet = std::getenv("RECTV_QUIET") && std::atoi(std::getenv("RECTV_QUIET")) != 0;
        RTV_REQUIRE(niter >= 0, "niter must be >= 0");
        int done = 0;
        std::vector<double> h;
        while (done < niter) {
            // the reference prints one line per outer iteration (rectv.cu:238)
            const int blk = quiet ? std::min(HIST_CAP, niter - done) : 1;
            const long long first = iters_done;
            enqueue(blk, titer, lam0, lam1);
            h.resize((size_t)HIST_W * blk);
            history(first, blk, h.data());
            collect();
            for (int q = 0; q < blk; q++) {
                const int it = done + q;
                if (hist)
                    for (int c = 0; c < 4; c++) hist[4 * (size_t)it + c] = h[HIST_W * (size_t)q + c];
                if (!quiet)
                    fprintf(stderr, "iter (%d/%d), rdiff ||f(k+1)-f(k)||:%f \n", it, niter, h[HIST_W * (size_t)q + 3]);
            }
            done += blk;
        }
    }

    void get_result(float *fres) {
        RTV_REQUIRE(fres, "null output pointer");
        if (ooc) {
            RTV_HIP(hipStreamSynchronize(st));
            std::memcpy(fres, h_f[hf], (size_t)nz * M * n * n * sizeof(float));
            return;
        }
        RTV_HIP(hipMemcpyAsync(fres, d_f[ia].p, (size_t)nz * M * n * n * sizeof(float), hipMemcpyDeviceToHost, st));
        RTV_HIP(hipStreamSynchronize(st));
    }
    size_t dual_chunk() const { return (size_t)(nzp + 1) * (M + 1) * (size_t)(n + 1) * (n + 1); }
    void require_resident(const char *what) const {
        if (ooc)
            throw Error(-1, std::string(what) + " needs the state in device memory (this solver holds it in host "
                                                "memory: out of core)");
    }

    // ------------------------------------------------------------ state (checkpoint / restore)
    // f: the previous iterate A [nz][M][n][n]; psi, mu [nch][nzp+1][M+1][n+1][n+1][4]
    // (NULL: zero); rho > 0.  div(mu - rho psi) is formed for lambda1, the outer
    // iteration counter and history restart at zero.  Replaces the zero start of
    // rectv::run (rectv.cu:119-125) for callers that resume a saved state.
    // The solver holds chi = mu - rho psi in place of psi (k_prox<true>): the import
    // forms it, the export recovers psi = (mu - chi) / rho (to FP32 rounding).
    void set_state(const float *f, const float *psi, const float *mu, float rho, float lam1) {
        require_resident("set_state");
        RTV_REQUIRE(std::isfinite(rho) && rho > 0.f, "rho must be positive and finite");
        RTV_REQUIRE(std::isfinite(lam1) && lam1 != 0.f, "lambda1 must be finite and nonzero");
        const size_t V = (size_t)nz * M * n * n;
        if (f) RTV_HIP(hipMemcpyAsync(d_f[ia].p, f, V * sizeof(float), hipMemcpyHostToDevice, st));
        else RTV_HIP(hipMemsetAsync(d_f[ia].p, 0, V * sizeof(float), st));
        if (psi) RTV_HIP(hipMemcpyAsync(d_psi.p, psi, d_psi.n * sizeof(float4), hipMemcpyHostToDevice, st));
        else RTV_HIP(hipMemsetAsync(d_psi.p, 0, d_psi.n * sizeof(float4), st));
        if (mu) RTV_HIP(hipMemcpyAsync(d_mu.p, mu, d_mu.n * sizeof(float4), hipMemcpyHostToDevice, st));
        else RTV_HIP(hipMemsetAsync(d_mu.p, 0, d_mu.n * sizeof(float4), st));
        hipLaunchKernelGGL(k_form_chi, dim3((unsigned)((d_psi.n + 255) / 256)), dim3(256), 0, st, d_psi.p, d_mu.p, rho,
                           d_psi.n);
        RTV_HIP(hipGetLastError());
        static const int zero = 0;
        const float rr[2] = {rho, rho};
        RTV_HIP(hipMemcpyAsync(d_rho.p, rr, 2 * sizeof(float), hipMemcpyHostToDevice, st));
        RTV_HIP(hipMemcpyAsync(d_iter.p, &zero, sizeof(int), hipMemcpyHostToDevice, st));
        iters_done = 0;
        divchi(tv_args(lam1, 0.f));
        RTV_HIP(hipStreamSynchronize(st));
        collect();
    }

    void get_state(float *f, float *psi, float *mu, float *dchi, float *rho) {
        require_resident("get_state");
        const size_t V = (size_t)nz * M * n * n;
        if (f) RTV_HIP(hipMemcpyAsync(f, d_f[ia].p, V * sizeof(float), hipMemcpyDeviceToHost, st));
        if (mu) RTV_HIP(hipMemcpyAsync(mu, d_mu.p, d_mu.n * sizeof(float4), hipMemcpyDeviceToHost, st));
        if (dchi) RTV_HIP(hipMemcpyAsync(dchi, d_dchi.p, V * sizeof(float), hipMemcpyDeviceToHost, st));
        float rr[2] = {0.f, 0.f};
        RTV_HIP(hipMemcpyAsync(rr, d_rho.p, 2 * sizeof(float), hipMemcpyDeviceToHost, st));
        RTV_HIP(hipStreamSynchronize(st));
        if (rho) *rho = rr[0];
        if (psi) {
            // psi = (mu - chi) / rho_chi, one chunk of dual entries at a time
            const size_t D = dual_chunk();
            std::vector<float4> hc(D), hm(D);
            for (int c = 0; c < nch; c++) {
                RTV_HIP(hipMemcpyAsync(hc.data(), d_psi.p + c * D, D * sizeof(float4), hipMemcpyDeviceToHost, st));
                RTV_HIP(hipMemcpyAsync(hm.data(), d_mu.p + c * D, D * sizeof(float4), hipMemcpyDeviceToHost, st));
                RTV_HIP(hipStreamSynchronize(st));
                float4 *o = reinterpret_cast<float4 *>(psi) + c * D;
                const float r0 = rr[1];
                for (size_t i = 0; i < D; i++)
                    o[i] = make_float4((hm[i].x - hc[i].x) / r0, (hm[i].y - hc[i].y) / r0, (hm[i].z - hc[i].z) / r0,
                                       (hm[i].w - hc[i].w) / r0);
            }
        }
    }

    // One inner step of the hot path on a given iterate fcur (host), against the
    // resident previous iterate, psi, mu, rho and data: fout = fcur - 1/2 div(rho grad
    // fcur + mu - rho psi) - 1/2 R^*(R fcur - g) (admm.cu:85-97).
    void apply_inner_step(const float *fcur, float *fout, float lam1) {
        require_resident("apply_inner_step");
        RTV_REQUIRE(inputs, "inputs not set");
        RTV_REQUIRE(fcur && fout, "null array");
        const size_t V = (size_t)nz * M * n * n;
        const int w1 = (ia + 1) % 3, w2 = (ia + 2) % 3;
        RTV_HIP(hipMemcpyAsync(d_f[w1].p, fcur, V * sizeof(float), hipMemcpyHostToDevice, st));
        inner_step(d_f[w1].p, d_f[ia].p, d_f[w2].p, tv_args(lam1, 0.f), false);
        RTV_HIP(hipMemcpyAsync(fout, d_f[w2].p, V * sizeof(float), hipMemcpyDeviceToHost, st));
        RTV_HIP(hipStreamSynchronize(st));
        collect();
    }

    // The final step of an outer iteration on a given iterate fcur (host): prox and
    // dual update of the resident psi, mu (admm.cu:101-117), the norms, the rho rule
    // (rectv.cu:224-232) and div(mu - rho psi) for the next iteration.  The resident
    // previous iterate is kept.  row receives the history row (HIST_W doubles).
    void apply_prox(const float *fcur, float lam0, float lam1, double *row) {
        require_resident("apply_prox");
        RTV_REQUIRE(fcur, "null array");
        RTV_REQUIRE(std::isfinite(lam0) && std::isfinite(lam1) && lam1 != 0.f, "lambda1 must be finite and nonzero");
        const size_t V = (size_t)nz * M * n * n;
        const int w1 = (ia + 1) % 3;
        RTV_HIP(hipMemcpyAsync(d_f[w1].p, fcur, V * sizeof(float), hipMemcpyHostToDevice, st));
        const TvArgs ta = tv_args(lam1, lam0);
        prox(d_f[w1].p, d_f[ia].p, ta, false);
        divchi(ta);
        iters_done++;
        if (row) history(iters_done - 1, 1, row);
        RTV_HIP(hipStreamSynchronize(st));
        collect();
    }

    // ------------------------------------------------------------ filtered back-projection, support
    // Support mask circ(r) of kernels_radonusfft.cuh:146-159 (r < 0: none): kept pixels
    // satisfy 4 xf^2 + 4 yf^2 < 1 - r with xf = (x - n/2)/float(n), evaluated in float
    // without contractions, one contiguous column range per row.  Applied to the
    // filtered back-projection and after every inner step (a support constraint).
    void set_support(float r) {
        RTV_REQUIRE(std::isfinite(r), "support radius parameter must be finite");
        drop_graphs();
        support_r = r;
        if (r < 0.f) return;
        std::vector<int2> rows(n, make_int2(0, 0));
        const volatile float fn_ = (float)n, lim = 1.0f - r;
        for (int y = 0; y < n; y++) {
            int lo = -1, hi = -1;
            for (int x = 0; x < n; x++) {
                const volatile float xf = (float)(x - n / 2) / fn_, yf = (float)(y - n / 2) / fn_;
                const volatile float ax = 4.0f * xf, ay = 4.0f * yf;
                const volatile float qx = ax * xf, qy = ay * yf;
                const volatile float q = qx + qy;
                if (q < lim) {
                    if (lo < 0) lo = x;
                    RTV_REQUIRE(hi < 0 || hi == x, "internal: support row is not one interval");
                    hi = x + 1;
                }
            }
            if (lo >= 0) rows[y] = make_int2(lo, hi);
        }
        upload(d_supp, rows, bytes, st);
        RTV_HIP(hipStreamSynchronize(st));
    }

    // f (device, [nz][M][n][n]) = every frame pi/(4 sqrt(M)) Re R^*_w g: the filtered
    // back-projection of each slice as a static object (radonusfft::adjR with its
    // filter argument set, the applyfilter weights of kernels_radonusfft.cuh:173-189),
    // masked by the support when one is set.
    void fbp_dev(const float *gdev, float *fdev) {
        const size_t S = (size_t)M * n * n;
        const float kappa = (float)(PI_EXACT / (4.0 * std::sqrt((double)M)));
        const int2 *supp = support_r >= 0.f ? d_supp.p : nullptr;
        for (int z0 = 0; z0 < nz; z0 += zb) {
            const int zc = std::min(zb, nz - z0);
            ray(zc, nullptr, gdev + (size_t)z0 * nt * n, nullptr, RAY_ADJ, nullptr, d_post_aw.p);
            rays_to_images(zc, d_cf_static.p);
            dim3 grid((unsigned)(((size_t)n * n + 255) / 256), zc);
            launch("fbp", [&] {
                hipLaunchKernelGGL(k_fbp_frames, grid, dim3(256), 0, st, fdev + z0 * S, d_v.p, n, M, kappa, supp);
            });
        }
    }

    void apply_fbp(const float *g, float *f) {
        RTV_REQUIRE(inputs, "inputs not set");
        RTV_REQUIRE(g && f, "null array");
        const size_t S = (size_t)M * n * n;
        DBuf<float> tf, th;
        size_t tmp = 0;
        tf.alloc((size_t)nz * S, tmp);
        th.alloc((size_t)nz * nt * n, tmp);
        RTV_HIP(hipMemcpyAsync(th.p, g, (size_t)nz * nt * n * sizeof(float), hipMemcpyHostToDevice, st));
        fbp_dev(th.p, tf.p);
        RTV_HIP(hipMemcpyAsync(f, tf.p, (size_t)nz * S * sizeof(float), hipMemcpyDeviceToHost, st));
        RTV_HIP(hipStreamSynchronize(st));
        collect();
        tf.release(tmp);
        th.release(tmp);
    }

    // Starting state from the filtered back-projection of the resident data: f = FBP,
    // psi = mu = 0, rho = 0.5 (the reference starts from f = 0, rectv.cu:119-125).
    void fbp_start() {
        RTV_REQUIRE(inputs, "inputs not set");
        require_resident("the filtered back-projection start");
        reset();
        fbp_dev(d_g.p, d_f[ia].p);
        RTV_HIP(hipStreamSynchronize(st));
        collect();
    }

    // ------------------------------------------------------------ operator entry points
    void apply_radon(const float *f, float *g) {
        RTV_REQUIRE(inputs, "inputs not set");
        const size_t S = (size_t)M * n * n;
        DBuf<float> tf, th;
        size_t tmp = 0;
        tf.alloc((size_t)nz * S, tmp);
        th.alloc((size_t)nz * nt * n, tmp);
        RTV_HIP(hipMemcpyAsync(tf.p, f, (size_t)nz * S * sizeof(float), hipMemcpyHostToDevice, st));
        for (int z0 = 0; z0 < nz; z0 += zb) {
            const int zc = std::min(zb, nz - z0);
            forward_to_rays(tf.p + z0 * S, zc, nullptr, th.p + (size_t)z0 * nt * n, RAY_FWD, nullptr);
        }
        RTV_HIP(hipMemcpyAsync(g, th.p, (size_t)nz * nt * n * sizeof(float), hipMemcpyDeviceToHost, st));
        RTV_HIP(hipStreamSynchronize(st));
        collect();
        tf.release(tmp);
        th.release(tmp);
    }

    void apply_radon_adj_dev(const float *gdev, float *fdev, float tau, int accumulate) {
        const size_t S = (size_t)M * n * n;
        for (int z0 = 0; z0 < nz; z0 += zb) {
            const int zc = std::min(zb, nz - z0);
            ray(zc, nullptr, gdev + (size_t)z0 * nt * n, nullptr, RAY_ADJ, nullptr);
            rays_to_images(zc);
            dim3 grid((unsigned)(((size_t)n * n + 255) / 256), zc);
            launch("unpack", [&] {
                hipLaunchKernelGGL(k_unpack, grid, dim3(256), 0, st, fdev + z0 * S, d_v.p, n, M, tau, accumulate);
            });
        }
    }

    void apply_radon_adj(const float *g, float *f) {
        RTV_REQUIRE(inputs, "inputs not set");
        const size_t S = (size_t)M * n * n;
        DBuf<float> tf, th;
        size_t tmp = 0;
        tf.alloc((size_t)nz * S, tmp);
        th.alloc((size_t)nz * nt * n, tmp);
        RTV_HIP(hipMemcpyAsync(th.p, g, (size_t)nz * nt * n * sizeof(float), hipMemcpyHostToDevice, st));
        apply_radon_adj_dev(th.p, tf.p, 1.0f, 0);
        RTV_HIP(hipMemcpyAsync(f, tf.p, (size_t)nz * S * sizeof(float), hipMemcpyDeviceToHost, st));
        RTV_HIP(hipStreamSynchronize(st));
        collect();
        tf.release(tmp);
        th.release(tmp);
    }

    void apply_gradient(const float *fext, int iz, float lam1, float *h2) {
        RTV_REQUIRE(iz >= 0 && iz < nch, "chunk index out of range");
        const size_t S = (size_t)M * n * n;
        const size_t D = (size_t)(nzp + 1) * (M + 1) * (n + 1) * (n + 1);
        DBuf<float> tf;
        DBuf<float4> th;
        size_t tmp = 0;
        tf.alloc((size_t)(nzp + 2) * S, tmp);
        th.alloc(D, tmp);
        RTV_HIP(hipMemcpyAsync(tf.p, fext, (size_t)(nzp + 2) * S * sizeof(float), hipMemcpyHostToDevice, st));
        const TvArgs ta = tv_args(lam1, 0.f);
        launch("grad", [&] {
            hipLaunchKernelGGL(k_grad, dim3((unsigned)((D + 255) / 256)), dim3(256), 0, st, tf.p, th.p, n, M, nzp,
                               iz != 0 ? 1 : 0, iz != nch - 1 ? 1 : 0, ta.gs, ta.gt);
        });
        RTV_HIP(hipMemcpyAsync(h2, th.p, D * sizeof(float4), hipMemcpyDeviceToHost, st));
        RTV_HIP(hipStreamSynchronize(st));
        collect();
        tf.release(tmp);
        th.release(tmp);
    }

    void apply_divergence(const float *h2, float lam1, float tau, float *f) {
        const size_t S = (size_t)M * n * n;
        const size_t D = (size_t)(nzp + 1) * (M + 1) * (n + 1) * (n + 1);
        DBuf<float> tf;
        DBuf<float4> th;
        size_t tmp = 0;
        tf.alloc((size_t)nzp * S, tmp);
        th.alloc(D, tmp);
        RTV_HIP(hipMemcpyAsync(tf.p, f, (size_t)nzp * S * sizeof(float), hipMemcpyHostToDevice, st));
        RTV_HIP(hipMemcpyAsync(th.p, h2, D * sizeof(float4), hipMemcpyHostToDevice, st));
        const TvArgs ta = tv_args(lam1, 0.f);
        const size_t tot = (size_t)nzp * S;
        launch("div", [&] {
            hipLaunchKernelGGL(k_div, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, tf.p, th.p, n, M, nzp,
                               ta.ks, ta.kt, tau);
        });
        RTV_HIP(hipMemcpyAsync(f, tf.p, (size_t)nzp * S * sizeof(float), hipMemcpyDeviceToHost, st));
        RTV_HIP(hipStreamSynchronize(st));
        collect();
        tf.release(tmp);
        th.release(tmp);
    }

    // ------------------------------------------------------------ autotuner
    // Chooses the slices per workgroup of the gather (zg) and of the scatter (zw)
    // for this problem shape by timing each candidate on the solver's own
    // geometry (HIP events, best of 3 after a warm-up launch).  Every candidate
    // does the same arithmetic in the same order, so results do not depend on
    // the choice.  Choices are cached per shape for the life of the process.
    // Tuned shapes are shared per (n, M, projection-count class): the choice moves
    // little with the slice count and the exact projection count, and a batch of
    // uneven jobs then tunes once per class rather than once per job.
    struct TuneKey {
        int n, ntc, M;
        bool operator<(const TuneKey &o) const { return std::tie(n, ntc, M) < std::tie(o.n, o.ntc, o.M); }
    };
    static int nt_class(int nt) {
        int c = 0;
        while ((1 << (c + 1)) <= nt) c++;
        return c;  // floor(log2 nt)
    }
    static std::map<TuneKey, std::pair<int, int>> &tune_cache() {
        static std::map<TuneKey, std::pair<int, int>> c;
        return c;
    }
    static std::mutex &tune_mutex() {
        static std::mutex m;
        return m;
    }
    template <typename F>
    float time_best(F &&f) {
        hipEvent_t a = ev(), b = ev();
        f();  // warm-up
        float best = 1e30f;
        for (int r = 0; r < 2; r++) {
            RTV_HIP(hipEventRecord(a, st));
            f();
            RTV_HIP(hipEventRecord(b, st));
            RTV_HIP(hipEventSynchronize(b));
            float ms = 0.f;
            RTV_HIP(hipEventElapsedTime(&ms, a, b));
            best = std::min(best, ms);
        }
        evpool.push_back(a);
        evpool.push_back(b);
        return best;
    }
    void autotune(int *zg_out, int *zw_out) {
        RTV_REQUIRE(inputs, "inputs not set (call set_inputs first)");
        const TuneKey key{n, nt_class(nt), M};
        std::pair<int, int> choice{-1, -1};
        {
            std::lock_guard<std::mutex> lk(tune_mutex());
            auto it = tune_cache().find(key);
            if (it != tune_cache().end()) choice = it->second;
        }
        if (choice.first < 0) {
            const bool was = prof;
            prof = false;
            const int zc = std::min(zb, nz);
            // real spectra and samples of the current iterate in G and gq
            forward_to_rays(d_f[ia].p, zc, d_g.p, nullptr, RAY_FWD | RAY_SUBG | RAY_ADJ, nullptr);
            float bestg = 1e30f, bests = 1e30f;
            int zg = GM_ZG, zw = SM_ZW;
            for (int c : {4, 8, 16, 32, 64}) {
                if (c > 4 && c / 2 >= zc) break;
                const float t = time_best([&] { gather(zc, c); });
                if (t < bestg) {
                    bestg = t;
                    zg = c;
                }
            }
            ray(zc, d_g.p, nullptr, nullptr, RAY_FWD | RAY_SUBG | RAY_ADJ, nullptr);  // q for the scatter timings
            for (int c : {2, 4, 8}) {
                const float t = time_best([&] { scatter(zc, c, d_cf.p); });
                if (t < bests) {
                    bests = t;
                    zw = c;
                }
            }
            prof = was;
            choice = {zg, zw};
            std::lock_guard<std::mutex> lk(tune_mutex());
            tune_cache()[key] = choice;
        }
        gather_zg = choice.first;
        scatter_zw = choice.second;
        drop_graphs();
        if (zg_out) *zg_out = gather_zg;
        if (zw_out) *zw_out = scatter_zw;
    }

    // rectv::check_adjoints, rectv.cu:255-289
    void check_adjoints(double *res, const float *g, const float *theta, const float *phi, float lam1, float center) {
        RTV_REQUIRE(nz == nzp, "check_adjoints requires nz == nzp (one chunk), as in the reference");
        set_inputs(g, theta, phi, center);
        const size_t S = (size_t)M * n * n;
        const size_t V = (size_t)nz * S, Gs = (size_t)nz * nt * n;
        std::vector<float> f(V), h1(Gs), fn(V);
        apply_radon_adj(g, f.data());
        apply_radon(f.data(), h1.data());
        for (int q = 0; q < 6; q++) res[q] = 0.0;
        for (size_t k = 0; k < V; k++) res[0] += (double)(f[k] * f[k]);
        for (size_t k = 0; k < Gs; k++) res[1] += (double)(g[k] * h1[k]);
        for (size_t k = 0; k < Gs; k++) res[2] += (double)(h1[k] * h1[k]);
        std::vector<float> fext((size_t)(nzp + 2) * S, 0.f);
        std::memcpy(fext.data() + S, f.data(), V * sizeof(float));
        const size_t D = (size_t)(nzp + 1) * (M + 1) * (n + 1) * (n + 1);
        std::vector<float> h2(4 * D);
        apply_gradient(fext.data(), 0, lam1, h2.data());
        std::fill(fn.begin(), fn.end(), 0.f);
        apply_divergence(h2.data(), lam1, 1.0f, fn.data());
        const size_t q4 = D;  // first (n+1)^2 (m+1)(nz+1) floats of the float4 array
        for (size_t k = 0; k < V; k++) res[3] += (double)(f[k] * fn[k]);
        for (size_t k = 0; k < q4; k++) res[4] += (double)(h2[k] * h2[k]);
        for (size_t k = 0; k < V; k++) res[5] += (double)(fn[k] * fn[k]);
    }

    void release() {
        if (freed) return;
        (void)hipStreamSynchronize(st);
        for (auto &r : pending) {
            (void)hipEventDestroy(r.a);
            (void)hipEventDestroy(r.b);
        }
        pending.clear();
        for (auto e : evpool) (void)hipEventDestroy(e);
        evpool.clear();
        drop_graphs();
        size_t &t = bytes;
        d_tw.release(t);
        d_rho.release(t);
        d_hist.release(t);
        d_iter.release(t);
        for (auto &b : d_f) b.release(t);
        for (auto &h : h_f) {
            if (h) (void)hipHostFree(h);
            h = nullptr;
        }
        for (float4 **h : {&h_psi, &h_mu}) {
            if (*h) (void)hipHostFree(*h);
            *h = nullptr;
        }
        if (h_g) (void)hipHostFree(h_g);
        h_g = nullptr;
        d_acc.release(t);
        for (auto &b : d_f2) b.release(t);
        d_psi2.release(t);
        d_mu2.release(t);
        d_dchi2.release(t);
        d_g2.release(t);
        for (hipStream_t *q : {&s_in, &s_out}) {
            if (*q) {
                (void)hipStreamSynchronize(*q);
                (void)hipStreamDestroy(*q);
            }
            *q = nullptr;
        }
        for (int q = 0; q < 2; q++)
            for (hipEvent_t *e : {&ev_in[q], &ev_comp[q], &ev_out[q]}) {
                if (*e) (void)hipEventDestroy(*e);
                *e = nullptr;
            }
        d_psi.release(t);
        d_mu.release(t);
        d_dchi.release(t);
        d_g.release(t);
        d_H.release(t);
        d_G.release(t);
        d_v.release(t);
        d_gq.release(t);
        d_q.release(t);
        d_h1.release(t);
        d_part.release(t);
        d_rpart.release(t);
        d_red.release(t);
        d_geo.release(t);
        d_aitems.release(t);
        d_arec.release(t);
        d_fitems.release(t);
        d_fcells.release(t);
        d_fmtiles.release(t);
        d_atiles.release(t);
        d_runs.release(t);
        d_crange.release(t);
        d_brange.release(t);
        d_blist.release(t);
        d_achunk.release(t);
        d_wtab.release(t);
        d_cf.release(t);
        d_pre_f.release(t);
        d_post_a.release(t);
        d_post_f.release(t);
        d_pre_a.release(t);
        d_deap.release(t);
        d_post_aw.release(t);
        d_cf_static.release(t);
        d_supp.release(t);
        if (st) (void)hipStreamDestroy(st);
        st = nullptr;
        freed = true;
    }
    ~Solver() { release(); }
};

}  // namespace rtv

// ------------------------------------------------------- allocation-free closing phase
// Scratch of rtv_prox_noalloc: the per-workgroup partial sums of k_prox, the
// reduced sums and the history-row counter of k_rho_rule.  Each device holds its
// own copy for the life of the process; host code uses the address returned by
// hipGetSymbolAddress for the current device, never the array's name.
constexpr size_t NOALLOC_DOUBLES = (size_t)1 << 20;  // 8 MiB
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];
static constexpr int NOALLOC_MAX_DEV = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[NOALLOC_MAX_DEV];

// Device address of `scratch` on the current device (looked up once per device).
static double *scratch_on_current_device() {
    int dev = 0;
    RTV_HIP(hipGetDevice(&dev));
    RTV_REQUIRE(dev >= 0 && dev < NOALLOC_MAX_DEV, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        RTV_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        RTV_REQUIRE(p != nullptr, "no device address for the scratch array");
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    return g_scratch_ptr[dev];
}

// rtv_apply_prox on caller-owned device arrays, with the scratch above.
static void prox_noalloc(const float *fcur, const float *fprev, float *psi, float *mu, float *dchi, float *rho,
                         double *row, int n, int m, int nz, int nzp, float lam0, float lam1, hipStream_t stream) {
    using namespace rtv;
    RTV_REQUIRE(fcur && fprev && psi && mu && dchi && rho && row, "null device pointer");
    RTV_REQUIRE(n > 0 && m > 0 && nz > 0 && nzp > 0 && nz % nzp == 0, "invalid sizes (nz must be a multiple of nzp)");
    RTV_REQUIRE(((uintptr_t)psi % 16) == 0 && ((uintptr_t)mu % 16) == 0, "psi and mu must be 16-byte aligned");
    RTV_REQUIRE(std::isfinite(lam0) && std::isfinite(lam1) && lam1 != 0.f, "lambda1 must be finite and nonzero");
    const int nch = nz / nzp;
    const size_t blocks = (size_t)((n + 1 + PX_ROWS - 1) / PX_ROWS) * (m + 1) * (nzp + 1) * nch;
    const size_t red_off = ((PROX_W * blocks + 31) / 32) * 32;  // 256-byte aligned
    const size_t need = red_off + 8 + 1;                         // reduced sums, row counter
    if (need > NOALLOC_DOUBLES || blocks > (size_t)INT_MAX)
        throw Error(RTV_ERR_ARG, "rtv_prox_noalloc: sizes need " + std::to_string(need * 8) +
                                     " bytes of scratch, more than its " + std::to_string(NOALLOC_DOUBLES * 8) +
                                     " (use rtv_apply_prox)");
    double *sc = scratch_on_current_device();
    double *part = sc, *red = sc + red_off;
    int *counter = reinterpret_cast<int *>(red + 8);
    TvArgs a;
    a.n = n;
    a.M = m;
    a.nzp = nzp;
    a.nchunk = nch;
    a.chunk0 = 0;
    const double c2 = std::cbrt(2.0);
    a.gs = (float)(1.0 / (c2 * lam1));
    a.gt = (float)(1.0 / c2);
    a.ks = (float)(1.0 / (2.0 * c2 * lam1));
    a.kt = (float)(1.0 / (2.0 * c2));
    a.rho = rho;
    a.lam0 = lam0;
    float4 *p4 = reinterpret_cast<float4 *>(psi), *m4 = reinterpret_cast<float4 *>(mu);
    hipLaunchKernelGGL(k_prox<false>, dim3((unsigned)blocks), dim3(64 * PX_WAVES), 0, stream, fcur, fprev, p4, m4, a,
                       part);
    RTV_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, stream, part, red, (int)blocks, PROX_W);
    RTV_HIP(hipGetLastError());
    // one history row (capacity 1: the counter's value does not matter)
    hipLaunchKernelGGL(k_rho_rule, dim3(1), dim3(64), 0, stream, red, rho, row, counter, 1, 0, nullptr);
    RTV_HIP(hipGetLastError());
    const dim3 grid(Solver::tv_blocks(n, nch)), thr(TV_THREADS);
    if (m <= 4) hipLaunchKernelGGL((k_divchi<4, false>), grid, thr, 0, stream, p4, m4, dchi, a, 0);
    else if (m <= 8) hipLaunchKernelGGL((k_divchi<8, false>), grid, thr, 0, stream, p4, m4, dchi, a, 0);
    else if (m <= 16) hipLaunchKernelGGL((k_divchi<16, false>), grid, thr, 0, stream, p4, m4, dchi, a, 0);
    else
        for (int i0 = 0; i0 < m; i0 += 32)  // groups of 32 frames
            hipLaunchKernelGGL((k_divchi<32, false>), grid, thr, 0, stream, p4, m4, dchi, a, i0);
    RTV_HIP(hipGetLastError());
}

// ====================================================================== C ABI
struct rtv_solver {
    rtv::Solver *s;
};

template <typename F>
static int guarded(F &&f) {
    try {
        f();
        rtv::g_last_error.clear();
        return RTV_OK;
    } catch (const rtv::Error &e) {
        rtv::g_last_error = e.what();
        return e.code;
    } catch (const std::exception &e) {
        rtv::g_last_error = e.what();
        return RTV_ERR_HIP;
    }
}

static rtv::Solver *live(rtv_solver *h) {
    if (!h || !h->s || h->s->freed) throw rtv::Error(RTV_ERR_STATE, "solver handle is null or freed");
    return h->s;
}

extern "C" {

const char *rtv_last_error(void) { return rtv::g_last_error.c_str(); }

int rtv_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int rtv_create(int n, int ntheta, int m, int nz, int nzp, int ngpus, rtv_solver **out) {
    return guarded([&] {
        RTV_REQUIRE(out, "null output handle");
        *out = nullptr;
        rtv::Solver *s = new rtv::Solver(n, ntheta, m, nz, nzp, ngpus);
        *out = new rtv_solver{s};
    });
}

int rtv_free(rtv_solver *h) {
    return guarded([&] {
        if (!h) return;
        if (h->s) {
            h->s->release();
            delete h->s;
        }
        delete h;
    });
}

int rtv_set_inputs(rtv_solver *h, const float *g, const float *theta, const float *phi, float center) {
    return guarded([&] { live(h)->set_inputs(g, theta, phi, center); });
}
int rtv_reset(rtv_solver *h) {
    return guarded([&] { live(h)->reset(); });
}
int rtv_iterate(rtv_solver *h, int niter, int titer, float lambda0, float lambda1, double *hist) {
    return guarded([&] { live(h)->iterate(niter, titer, lambda0, lambda1, hist); });
}
int rtv_get_result(rtv_solver *h, float *fres) {
    return guarded([&] { live(h)->get_result(fres); });
}
int rtv_synchronize(rtv_solver *h) {
    return guarded([&] { RTV_HIP(hipStreamSynchronize(live(h)->st)); });
}

int rtv_run(rtv_solver *h, float *fres, const float *g, const float *theta, const float *phi, float center,
            float lambda0, float lambda1, int niter, int titer) {
    return guarded([&] {
        rtv::Solver *s = live(h);
        s->set_inputs(g, theta, phi, center);
        s->reset();
        s->iterate(niter, titer, lambda0, lambda1, nullptr);
        s->get_result(fres);
    });
}

int rtv_check_adjoints(rtv_solver *h, double *res, const float *g, const float *theta, const float *phi,
                       float lambda1, float center) {
    return guarded([&] {
        RTV_REQUIRE(res, "null result pointer");
        live(h)->check_adjoints(res, g, theta, phi, lambda1, center);
    });
}

int rtv_set_state(rtv_solver *h, const float *f, const float *psi, const float *mu, float rho, float lambda1) {
    return guarded([&] { live(h)->set_state(f, psi, mu, rho, lambda1); });
}
int rtv_get_state(rtv_solver *h, float *f, float *psi, float *mu, float *dchi, float *rho) {
    return guarded([&] { live(h)->get_state(f, psi, mu, dchi, rho); });
}
int rtv_apply_inner_step(rtv_solver *h, const float *fcur, float *fout, float lambda1) {
    return guarded([&] { live(h)->apply_inner_step(fcur, fout, lambda1); });
}
int rtv_apply_prox(rtv_solver *h, const float *fcur, float lambda0, float lambda1, double *row) {
    return guarded([&] { live(h)->apply_prox(fcur, lambda0, lambda1, row); });
}

int rtv_prox_noalloc(const float *fcur, const float *fprev, float *psi, float *mu, float *dchi, float *rho,
                     double *row, int n, int m, int nz, int nzp, float lambda0, float lambda1, void *stream) {
    return guarded([&] {
        prox_noalloc(fcur, fprev, psi, mu, dchi, rho, row, n, m, nz, nzp, lambda0, lambda1,
                     static_cast<hipStream_t>(stream));
    });
}

int rtv_set_support(rtv_solver *h, float r) {
    return guarded([&] { live(h)->set_support(r); });
}
int rtv_apply_fbp(rtv_solver *h, const float *g, float *f) {
    return guarded([&] { live(h)->apply_fbp(g, f); });
}
int rtv_fbp_start(rtv_solver *h) {
    return guarded([&] { live(h)->fbp_start(); });
}

int rtv_apply_radon(rtv_solver *h, const float *f, float *g) {
    return guarded([&] { live(h)->apply_radon(f, g); });
}
int rtv_apply_radon_adj(rtv_solver *h, const float *g, float *f) {
    return guarded([&] { live(h)->apply_radon_adj(g, f); });
}
int rtv_apply_gradient(rtv_solver *h, const float *fext, int iz, float lambda1, float *h2) {
    return guarded([&] { live(h)->apply_gradient(fext, iz, lambda1, h2); });
}
int rtv_apply_divergence(rtv_solver *h, const float *h2, float lambda1, float tau, float *f) {
    return guarded([&] { live(h)->apply_divergence(h2, lambda1, tau, f); });
}

int rtv_set_profiling(rtv_solver *h, int on) {
    return guarded([&] {
        rtv::Solver *s = live(h);
        s->collect();
        s->prof = on != 0;
        if (on) s->stats.clear();
    });
}

int rtv_get_profile(rtv_solver *h, char *buf, int len) {
    return guarded([&] {
        rtv::Solver *s = live(h);
        s->collect();
        std::string out;
        char line[256];
        for (auto &kv : s->stats) {
            snprintf(line, sizeof line, "%s %lld %.6f\n", kv.first.c_str(), kv.second.first, kv.second.second);
            out += line;
        }
        RTV_REQUIRE(buf && len > 0, "null profile buffer");
        std::strncpy(buf, out.c_str(), (size_t)len - 1);
        buf[len - 1] = 0;
    });
}

long long rtv_device_bytes(rtv_solver *h) {
    if (!h || !h->s) return -1;
    return (long long)h->s->bytes;
}

int rtv_iterate_async(rtv_solver *h, int niter, int titer, float lambda0, float lambda1) {
    return guarded([&] { live(h)->enqueue(niter, titer, lambda0, lambda1); });
}

int rtv_query(rtv_solver *h, int *idle) {
    return guarded([&] {
        RTV_REQUIRE(idle, "null output pointer");
        const hipError_t e = hipStreamQuery(live(h)->st);
        if (e == hipSuccess) *idle = 1;
        else if (e == hipErrorNotReady) *idle = 0;
        else RTV_HIP(e);
    });
}

int rtv_iterations(rtv_solver *h, long long *count) {
    return guarded([&] {
        RTV_REQUIRE(count, "null output pointer");
        *count = live(h)->iters_done;
    });
}

int rtv_history(rtv_solver *h, long long first, int count, double *rows) {
    return guarded([&] { live(h)->history(first, count, rows); });
}

int rtv_set_graph(rtv_solver *h, int on) {
    return guarded([&] {
        rtv::Solver *s = live(h);
        s->use_graph = on != 0;
        if (!on) s->drop_graphs();
    });
}

int rtv_autotune(rtv_solver *h, int *gather_slices, int *scatter_slices) {
    return guarded([&] { live(h)->autotune(gather_slices, scatter_slices); });
}

int rtv_batch_iterate(rtv_solver **hs, int count, const int *niter, const int *titer, const float *lambda0,
                      const float *lambda1) {
    return guarded([&] {
        RTV_REQUIRE(count >= 0 && (count == 0 || (hs && niter && titer && lambda0 && lambda1)), "null batch arrays");
        std::vector<rtv::Solver *> ss(count);
        int most = 0;
        for (int j = 0; j < count; j++) {
            ss[j] = live(hs[j]);
            for (int i = 0; i < j; i++) RTV_REQUIRE(ss[i] != ss[j], "a solver appears twice in the batch");
            RTV_REQUIRE(niter[j] >= 0, "niter must be >= 0");
            most = std::max(most, niter[j]);
        }
        // round robin: iteration k of every job is enqueued (each on its own
        // stream) before iteration k + 1 of any, so the jobs' kernels overlap
        for (int k = 0; k < most; k++)
            for (int j = 0; j < count; j++)
                if (k < niter[j]) ss[j]->enqueue(1, titer[j], lambda0[j], lambda1[j]);
        for (int j = 0; j < count; j++) RTV_HIP(hipStreamSynchronize(ss[j]->st));
    });
}

int rtv_footprint(int n, int ntheta, int m, int nz, int nzp, long long *bytes) {
    return guarded([&] {
        RTV_REQUIRE(bytes, "null output pointer");
        RTV_REQUIRE(n > 0 && ntheta > 0 && m > 0 && nz > 0 && nzp > 0 && nz % nzp == 0, "invalid sizes");
        *bytes = (long long)rtv::Solver::plan_bytes(n, ntheta, m, nz, nzp);
    });
}

}  // extern "C"
