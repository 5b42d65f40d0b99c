// HIP kernels of the rectv reconstruction path (gfx950).
//
// The M time frames of a slice are real images, so they are transformed in
// pairs: Z_p = FFT2(f_{2p} + i f_{2p+1}) holds both frame spectra (F_a(c) =
// (Z(c) + conj Z(-c))/2, F_b(c) = (Z(c) - conj Z(-c))/(2i)).  The basis
// contraction sum_j phi_j[k] conj(phi_j[t_i]) = D[k][i] becomes, per pair, one
// coefficient on Z at the sample's own 9x9 window ("direct" item) and one on Z
// at the point-mirrored window ("mirror" item, conjugated); the adjoint is the
// transpose and yields Hermitian-symmetric grids whose inverse transforms are
// the frame pairs themselves.  MP = ceil(M/2) spectra per slice.
//
// Data layouts in HBM (z = slice within the current slice batch):
//   f        [nz][M][n][n]                 f32   time frames of each slice
//   H        [zb][MP][2n/g][n][g]          c64   row-transformed spectra (forward, g = HG),
//                                                 column-transformed (adjoint, g = HGA
//                                                 or 2n): groups of g adjacent columns kx
//   G        [zb][MP][2n][2n]              c64   oversampled spectra, centred, stored
//                                                 transposed: G[kx][ky] (column FFT rows)
//   gq       [zb][ntheta][n][2]            c64   per sample: direct / mirror gather
//                                                 sums (forward)
//   q        [zb][ntheta][n]               c64   per sample: the scatter's input (adjoint)
//   v        [zb][MP][n][n]                c64   adjoint frame pairs (re = f_{2p}, im = f_{2p+1})
//   psi, mu  [nchunk][nzp+1][M+1][n+1][n+1] f32x4 dual variables (x, y, t, z)
// The oversampled grid index kb in [0, 2n) carries the centred frequency
// kb - n (the mirror of kb is (2n - kb) mod 2n); the checkerboard (-1)^(a+b)
// before and after the 2D FFT (the reference's fftshift on the (2n+2K)^2 grid,
// kernels_radonusfft.cuh:26-38) is applied separably in the row and column
// passes, and the periodic border of that grid (wrap, kernels_radonusfft.cuh:40-64)
// becomes index arithmetic mod 2n.
#pragma once
#include "common.hpp"
#include <type_traits>
#include "fft_lds.hpp"

namespace rtv {

constexpr int FFT_E = 16;       // complex elements per thread in the LDS FFT
constexpr int FFT_THREADS = 1024;
constexpr int TS = 16;          // spectral tile edge (cells) for gather/scatter
constexpr int KW = 9;           // Gaussian window taps per axis, 2K+1 with K = 4
constexpr int TH = TS + KW - 1; // tile + halo rows loaded by the gather
// Tiles per side of the 2n x 2n grid.  When 2n is not a multiple of TS the last
// tile row / column is narrower: its cells past 2n wrap to the first tile, so
// the gather reads them modulo 2n and the scatter writes only the cells < 2n.
__host__ __device__ constexpr int tiles_per_side(int N) { return (N + TS - 1) / TS; }

__device__ __forceinline__ float sgn(int i) { return (i & 1) ? -1.0f : 1.0f; }

// Element loop of the runtime-planned FFT passes: for idx = tid + q*nthr < total,
// ld(idx) (global loads) for CNT elements at a time, all issued before st(idx, value)
// uses any of them.  The loads run unconditionally on clamped indices (total > 0), so
// that the compiler's wait counts stay exact and the loads are in flight together;
// a loop that waits for every load before the next costs a full memory latency per
// element.
template <int CNT, typename V, typename LD, typename ST>
__device__ __forceinline__ void staged_elems(int tid, int nthr, int total, LD ld, ST st) {
    for (int base = 0; base < total; base += CNT * nthr) {
        V tmp[CNT];
#pragma unroll
        for (int q = 0; q < CNT; q++) tmp[q] = ld(min(base + tid + q * nthr, total - 1));
#pragma unroll
        for (int q = 0; q < CNT; q++) {
            const int idx = base + tid + q * nthr;
            if (idx < total) st(idx, tmp[q]);
        }
    }
}

// H holds each (slice, pair) spectrum as groups of g adjacent columns kx, each
// group [n rows y][g] (ceil(2n/g) groups: for odd n the last group is half used);
// g = 2n is the plain row-major layout.  The forward
// direction uses g = HG: the column pass then reads contiguous blocks, and the row
// pass writes g-element row segments, which the TT rows of one workgroup complete
// to whole cache lines.  The adjoint uses g = HGA (0: row-major): its column pass
// writes the groups, and its row pass reads each group's block of rows whole
// through LDS.  Build-time tuning constants.
#ifndef RTV_HG
#define RTV_HG 4
#endif
#ifndef RTV_HGA
#define RTV_HGA 4
#endif
constexpr int HG = RTV_HG, HGA = RTV_HGA;
__host__ __device__ constexpr int h_group_adj(int N) { return HGA > 0 ? HGA : N; }
__device__ __forceinline__ size_t h_index(size_t zj, int n, int y, int kx, int g) {
    return ((zj * (size_t)((2 * n + g - 1) / g) + (size_t)(kx / g)) * (size_t)n + y) * g + (kx % g);
}
// element kx = jl + m * TPT of row y (row passes of a power-of-two length)
template <int TPT, int G>
__device__ __forceinline__ size_t h_row_off(size_t zj, int nn, int y, int jl, int m) {
    if constexpr (TPT % G == 0) return h_index(zj, nn, y, jl, G) + (size_t)m * (TPT / G) * nn * G;
    else if constexpr (G >= 16 * TPT) return h_index(zj, nn, y, 0, G) + jl + m * TPT;  // one group per row
    else return h_index(zj, nn, y, jl + m * TPT, G);
}

typedef float v4f __attribute__((ext_vector_type(4)));  // MFMA accumulator fragment

// Transforms per FFT workgroup for a compile-time power-of-two length L:
// FFT_WG / (L / FFT_E) rows, at most 16 (the host uses the same rule).  FFT_WG
// = 512 threads keeps LDS near 70 KB so that two workgroups share a CU and one
// transforms while the other loads or stores.
constexpr int FFT_WG = 512;
// Per-ray transforms: 256-thread workgroups (4 rays at n = 1024) measured 7%
// faster than 512 threads (128 threads: the same as 256); the 2n-point row
// passes are fastest at 512.
#ifndef RTV_RAY_WG
#define RTV_RAY_WG 256
#endif
constexpr int RAY_WG = RTV_RAY_WG;  // threads of the per-ray transform workgroups
// Launch bound of the FFT pass kernels by their length template NP: compile-time
// powers of two keep FFT_THREADS; the runtime-planned path (NP = 0) runs at most
// FFT_WG threads, which lets it hold a stage's butterflies in registers without
// spilling (up to 256 VGPRs); NP = -1 is the same path for transforms that need
// more than FFT_WG threads (length > 16 FFT_WG).
__host__ __device__ constexpr int fft_bound(int NP) {
    return NP == 0 || NP == -2 || (NP > 0 && rp_radix(NP) > 0) ? FFT_WG : FFT_THREADS;
}
// NP = -2: runtime plan with a prime factor above 7 (direct stages), at most FFT_WG threads;
// compile-time lengths R0 * 2^k run at most 384 threads (fft_rp_count) under the FFT_WG bound,
// which leaves them the registers of the radix-R0 combine
__host__ __device__ constexpr int fft_rows_p2(int L, int wg = FFT_WG) {
    return (wg / (L / FFT_E)) < 1 ? 1 : ((wg / (L / FFT_E)) < 16 ? (wg / (L / FFT_E)) : 16);
}
// Column passes: threads / (L / FFT_E) adjacent columns per workgroup (at most
// 16).  Build-time tuning constants: the forward pass reads HG-column groups of H
// (256 threads at L = 2048: 2 columns, LDS ~35 KB, four workgroups per CU); the
// adjoint pass writes HGA-column groups of H (256 threads as well).
#ifndef RTV_FCOLS_THREADS
#define RTV_FCOLS_THREADS 256
#endif
#ifndef RTV_ACOLS_THREADS
#define RTV_ACOLS_THREADS 256
#endif
__host__ __device__ constexpr int fft_cols_p2(int L, int threads) {
    return (threads / (L / FFT_E)) < 1 ? 1 : ((threads / (L / FFT_E)) < 16 ? (threads / (L / FFT_E)) : 16);
}
// Lengths L = R0 * 2^k (rp_radix(L) != 0): transforms per workgroup, a power of
// two (it divides L and L/2), at most 16, and enough for whole waves (t * L/16 a
// multiple of 64 threads); beyond that as many as fit in `threads` threads.
__host__ __device__ constexpr int fft_rp_count(int L, int threads) {
    int t = 1;
    while (t < 16 && (t * (L / FFT_E)) % 64 != 0) t *= 2;
    while (2 * t <= 16 && 2 * t * (L / FFT_E) <= threads) t *= 2;
    return t;
}

// Cells of G inside some gather window, per row kx of G: the shortest circular
// run of columns [lo, lo + len) (mod 2n) holding them, len = 0 for none.  The
// forward column pass stores only these cells and the adjoint column pass reads
// only these; every other cell of G stays zero (G is cleared when the geometry
// is set, and the scatter writes exact zeros outside the windows).
__device__ __forceinline__ bool in_run(int2 run, int c, int N) {
    int d = c - run.x;
    if (d < 0) d += N;
    return d < run.y;
}

struct SampleGeom {
    int row0;   // first window row in the 2n grid (already mod 2n)
    int col0;   // first window column
    float fy;   // Y - floor(Y), Y = 2n * y0
    float fx;   // X - floor(X)
};

// An adjoint list entry: the item (2*sample + mirror) with its window geometry,
// stored in list order so that staging a chunk reads consecutive records.
struct AdjItem {
    int pid;    // item id
    int rc;     // row0 | col0 << 16
    float fy;
    float fx;
};

struct TileEntry {
    int tile;
    int start;
    int count;
    int pad;
};

// ---------------------------------------------------------------------------
// F1: frame pairs f_{2p} + i f_{2p+1} + deapodisation + zero-padded row FFT
// (makecomplex/decphi, kernels_operators.cuh:86-116, is replaced by the pair
// packing; divphi FWD + the first half of the 2D cuFFT, kernels_radonusfft.cuh:4-24,
// radonusfft.cu:85-89).  One workgroup per (slice z, pair p, TT image rows) for
// power-of-two lengths, per (slice z, image row y, T pairs) otherwise.
template <int NP>
__global__ void __launch_bounds__(fft_bound(NP))
k_fwd_rows(const float *__restrict__ f, float2 *__restrict__ H, const float *__restrict__ deap, FftPlan plan,
           const float2 *__restrict__ tw, int n, int M, int MP, int T, int stride) {
    extern __shared__ float2 lds[];
    const int tid = threadIdx.x;
    const int L = xcd_remap(blockIdx.x, gridDim.x);
    if constexpr (NP > 0 && rp_radix(NP) > 0) {
        // length R0 * P: thread (t, g, jl) of image row yt*TT + t supplies the pixels
        // x = R0 (jl + m TPS) + g at positions x + n/2 (m = 4..11 of fft_rp_reg) and
        // receives kx = jl + (16 g + m) TPS
        constexpr int N = NP, nn = NP / 2, TPT = NP / 16, R0 = rp_radix(NP), P = NP / R0, TPS = P / 16;
        constexpr int TT = fft_rp_count(NP, FFT_WG);
        constexpr int nyt = nn / TT;
        const int yt = L % nyt;
        const int zp = L / nyt;  // z*MP + p
        const int z = zp / MP, p = zp - z * MP;
        const int t = tid / TPT, r = tid % TPT, g = r / TPS, jl = r % TPS;
        const int y = yt * TT + t;
        const size_t plane = (size_t)nn * nn;
        const float sgi = sgn(jl + g), sgo = sgn(jl);  // (-1)^position of inputs / outputs
        float2 v[16];
#pragma unroll
        for (int m = 0; m < 16; m++) v[m] = make_float2(0.f, 0.f);
        {
            const int i0 = 2 * p;
            const float *fa = f + ((size_t)z * M + i0) * plane + (size_t)y * nn;
            const float sy = deap[y] * sgi;
            const float *fb = fa + (i0 + 1 < M ? plane : 0);
            float va[8], vb[8], dx[8];
#pragma unroll
            for (int m = 0; m < 8; m++) {
                const int x = R0 * (jl + m * TPS) + g;
                va[m] = fa[x];
                vb[m] = fb[x];
                dx[m] = deap[x];
            }
#pragma unroll
            for (int m = 0; m < 8; m++)
                v[m + 4] = make_float2(va[m] * dx[m] * sy, i0 + 1 < M ? vb[m] * dx[m] * sy : 0.f);
        }
        fft_rp_reg<R0, P, false>(v, lds + t * stride, true, g, jl, tw + plan.tw_sub, tw + plan.tw_off);
#pragma unroll
        for (int m = 0; m < 16; m++) H[h_row_off<TPS, HG>(zp, nn, y, jl, 16 * g + m)] = cscale(v[m], sgo);
    } else if constexpr (NP > 0) {
        // thread (t, jl): image row yt*TT + t; element positions jl + m*TPT of
        // the 2n-point row; the n image pixels sit at positions 4*TPT .. 12*TPT - 1
        constexpr int N = NP, nn = NP / 2, TPT = NP / 16, TT = fft_rows_p2(NP);
        constexpr int nyt = nn / TT;
        const int yt = L % nyt;
        const int zp = L / nyt;  // z*MP + p
        const int z = zp / MP, p = zp - z * MP;
        const int t = tid / TPT, jl = tid % TPT;
        const int y = yt * TT + t;
        const size_t plane = (size_t)nn * nn;
        const float sg = sgn(jl);  // (-1)^position: TPT is even
        float2 v[16];
#pragma unroll
        for (int m = 0; m < 16; m++) v[m] = make_float2(0.f, 0.f);
        {
            const int i0 = 2 * p;
            const float *fa = f + ((size_t)z * M + i0) * plane + (size_t)y * nn + jl;
            const float sy = deap[y] * sg;
            float va[8], vb[8], dx[8];
#pragma unroll
            for (int m = 0; m < 8; m++) {
                va[m] = fa[m * TPT];
                vb[m] = i0 + 1 < M ? fa[plane + m * TPT] : 0.f;
                dx[m] = deap[jl + m * TPT];
            }
#pragma unroll
            for (int m = 0; m < 8; m++) v[m + 4] = make_float2(va[m] * dx[m] * sy, vb[m] * dx[m] * sy);
        }
        fft_p2_reg<NP, false>(v, lds + t * stride, true, jl, tw + plan.tw_off);
#pragma unroll
        for (int m = 0; m < 16; m++) H[h_row_off<TPT, HG>(zp, nn, y, jl, m)] = cscale(v[m], sg);
    } else {
        const int nthr = blockDim.x;
        const int ngr = (MP + T - 1) / T;
        const int pg = L % ngr;
        const int y = (L / ngr) % n;
        const int z = L / ngr / n;
        const int p0 = pg * T;
        const int Tc = min(T, MP - p0);
        const int N = 2 * n, h = N / 4;
        const size_t plane = (size_t)n * n;
        const float *fz = f + (size_t)z * M * plane + (size_t)y * n;
        const float dy = deap[y];
        staged_elems<8, float2>(
            tid, nthr, Tc * n,
            [&](int idx) {
                const int pt = idx / n, x = idx - pt * n;
                const int i0 = 2 * (p0 + pt);
                const float va = fz[(size_t)i0 * plane + x];
                const float vb = fz[(size_t)min(i0 + 1, M - 1) * plane + x];
                const float sc = deap[x] * dy * sgn(x + h);
                return make_float2(va * sc, i0 + 1 < M ? vb * sc : 0.f);
            },
            [&](int idx, float2 val) {
                const int pt = idx / n, x = idx - pt * n;
                lds[pt * stride + lp(x + h)] = val;
                const int q = x < h ? x : x + n;  // zero padding
                lds[pt * stride + lp(q)] = make_float2(0.f, 0.f);
            });
        __syncthreads();
        fft_lds<FFT_E, false, true, NP == -2>(lds, stride, Tc, plan, tw, tid, nthr);
        for (int idx = tid; idx < Tc * N; idx += nthr) {
            const int pt = idx / N, kx = idx - pt * N;
            H[h_index((size_t)z * MP + p0 + pt, n, y, kx, HG)] = cscale(lds[pt * stride + lp(kx)], sgn(kx));
        }
    }
}

// F2: zero-padded column FFT -> oversampled spectrum G (second half of the 2D
// cuFFT of radonusfft.cu:89 with the post-checkerboard).  One workgroup per
// (z, j, tile of T adjacent columns).
template <int NP>
__global__ void __launch_bounds__(fft_bound(NP))
k_fwd_cols(const float2 *__restrict__ H, float2 *__restrict__ G, FftPlan plan, const float2 *__restrict__ tw,
           const int2 *__restrict__ runs, int n, int M, int T, int stride) {
    extern __shared__ float2 lds[];
    const int tid = threadIdx.x;
    const int L = xcd_remap(blockIdx.x, gridDim.x);
    if constexpr (NP > 0 && rp_radix(NP) > 0) {
        // length R0 * P: thread (t, g, jl), t fastest, supplies the rows
        // y = R0 (jl + m TPS) + g and receives ky = jl + (16 g + m) TPS
        constexpr int N = NP, nn = NP / 2, TPT = NP / 16, R0 = rp_radix(NP), P = NP / R0, TPS = P / 16;
        constexpr int TT = fft_rp_count(NP, RTV_FCOLS_THREADS);
        const int ntile = N / TT;
        const int kt = L % ntile;
        const int zj = L / ntile;  // z*MP + p
        const int t = tid % TT, r = tid / TT, g = r / TPS, jl = r % TPS;
        const int kx = kt * TT + t;
        const float sgi = sgn(jl + g), sgo = sgn(jl);
        float2 v[16];
#pragma unroll
        for (int m = 0; m < 16; m++) v[m] = make_float2(0.f, 0.f);
#pragma unroll
        for (int m = 0; m < 8; m++) v[m + 4] = cscale(H[h_index(zj, nn, R0 * (jl + m * TPS) + g, kx, HG)], sgi);
        fft_rp_reg<R0, P, false>(v, lds + t * stride, true, g, jl, tw + plan.tw_sub, tw + plan.tw_off);
        const int2 run = runs[kx];
        float2 *Gs = G + (size_t)zj * N * N + (size_t)kx * N + jl;
#pragma unroll
        for (int m = 0; m < 16; m++) {
            const int ky = jl + (16 * g + m) * TPS;
            if (in_run(run, ky, N)) st_nt(Gs + (16 * g + m) * TPS, cscale(v[m], sgo));
        }
    } else if constexpr (NP > 0) {
        // thread (t, jl), t fastest: adjacent lanes read adjacent columns; the n
        // input rows sit at positions 4*TPT .. 12*TPT - 1
        constexpr int N = NP, nn = NP / 2, TPT = NP / 16, TT = fft_cols_p2(NP, RTV_FCOLS_THREADS);
        const int ntile = N / TT;
        const int kt = L % ntile;
        const int zj = L / ntile;  // z*MP + p
        const int t = tid % TT, jl = tid / TT;
        const float sg = sgn(jl);
        const float2 *Hs = H + h_index(zj, nn, jl, kt * TT + t, HG);  // rows jl + m*TPT: stride HG
        float2 v[16];
#pragma unroll
        for (int m = 0; m < 16; m++) v[m] = make_float2(0.f, 0.f);
#pragma unroll
        for (int m = 0; m < 8; m++) v[m + 4] = cscale(Hs[m * TPT * HG], sg);
        fft_p2_reg<NP, false>(v, lds + t * stride, true, jl, tw + plan.tw_off);
        // transposed store: the transform of column kx is row kx of G (window cells only)
        const int2 run = runs[kt * TT + t];
        float2 *Gs = G + (size_t)zj * N * N + (size_t)(kt * TT + t) * N + jl;
#pragma unroll
        for (int m = 0; m < 16; m++)
            if (in_run(run, jl + m * TPT, N)) st_nt(Gs + m * TPT, cscale(v[m], sg));
    } else {
        const int nthr = blockDim.x;
        const int N = 2 * n, h = N / 4;
        const int ntile = N / T;
        const int kt = L % ntile;
        const int zj = L / ntile;  // z*MP + p
        const int kx0 = kt * T;
        staged_elems<8, float2>(
            tid, nthr, T * n,
            [&](int idx) {
                const int t = idx % T, y = idx / T;
                return cscale(H[h_index(zj, n, y, kx0 + t, HG)], sgn(y + h));
            },
            [&](int idx, float2 val) {
                const int t = idx % T, y = idx / T;
                lds[t * stride + lp(y + h)] = val;
                const int q = y < h ? y : y + n;
                lds[t * stride + lp(q)] = make_float2(0.f, 0.f);
            });
        __syncthreads();
        fft_lds<FFT_E, false, true, NP == -2>(lds, stride, T, plan, tw, tid, nthr);
        float2 *Gs = G + (size_t)zj * N * N + (size_t)kx0 * N;
        staged_elems<16, int2>(
            tid, nthr, T * N, [&](int idx) { return runs[kx0 + idx / N]; },
            [&](int idx, int2 run) {
                const int t = idx / N, ky = idx - t * N;
                if (in_run(run, ky, N)) Gs[(size_t)t * N + ky] = cscale(lds[t * stride + lp(ky)], sgn(ky));
            });
    }
}

// One separable Gaussian window factor exp(-beta d^2) (kernels_radonusfft.cuh:114-118;
// the PI/mu factor is folded into the ray tables).  exp through the hardware
// exp2: relative error a few 1e-7 for the taps that matter (beta d^2 < 10); the
// forward and adjoint use the same function, so they stay exact transposes.
__device__ __forceinline__ float gauss(float beta, float d) { return __expf(-beta * d * d); }

// Gaussian window weights exp(-beta (a - K - frac)^2), a = 0..8.
__device__ __forceinline__ void window_weights(float frac, float beta, float *w) {
#pragma unroll
    for (int a = 0; a < KW; a++) w[a] = gauss(beta, (float)(a - 4) - frac);
}

// F3: tile-binned Gaussian gather from the MP pair spectra on the matrix cores
// (replaces gather FWD, kernels_radonusfft.cuh:81-144, run once per basis
// function, plus mulphi/addreal's sum over j, operators.cu:22-31).
// Items: sample (k, s) has a direct item 2*(k*n+s) (its own window) and a mirror
// item 2*(k*n+s)+1 (the point-mirrored window, fractions negated).  Item sums
// are contracted with cf[k][kind][p] (alpha_p[k] direct, conj(beta_p[k]) mirror)
// and written to gq[z][item]; the per-ray kernel forms direct + conj(mirror).
// Rays repeated every 2 pi, and rays theta+pi read in reverse detector order,
// share window origins.  One workgroup per (tile entry of <= GM_ITEMS items
// whose window origin lies in the tile, group of zg slices); items are sorted
// by local origin (row, column) and packed by the host into 16-item blocks (one
// wave each) spanning at most 12 rows and 4 steps of 4 cells.  The 9 window
// taps per axis of every item are staged once; per slice the tile plus its
// 8-cell halo of all MP pair spectra is staged as tile[row][cell][col], col = p
// (real part) and HALF + p (imaginary part), the next slice's tile being loaded
// into registers meanwhile.  Per window row r of its block and per 4 cells, a
// wave runs
//   D[col][item] += sum_c tile[r][c][col] * wy[item][r] * wx[item][c]
// as v_mfma_f32_16x16x4f32 (16 columns x 16 items x 4 cells); each lane then
// holds 4 columns of one item, contracts them with cf (alpha_p[k] direct,
// conj(beta_p[k]) mirror) and two cross-lane steps finish the sum over pairs.
constexpr int GM_ITEMS = 256;  // items per workgroup: 16 waves x 16
constexpr int GM_THREADS = 1024;
constexpr int GM_KS = 4;       // max steps of 4 cells per block (the host packs blocks so)
constexpr int GM_ZG = 8;       // default slices per workgroup (runtime argument zg)
constexpr int GM_WYP = 17;     // row-tap entries per item (3 zeros, 9 taps, 5 zeros); odd, so
                               // that the 16 items of a block read 16 different LDS banks
// More than 16 pairs (M > 32) run as one launch per group of at most 8*NB pairs
// p0 .. p0 + 8*NB - 1 of the MPt pairs: the item sums are linear in the pairs, so
// launches after the first add their part to gq (accum != 0).
// With few pairs (PK = 2: at most 4, PK = 4: at most 2; NB = 1) PK slices share
// the 16 columns: slice s of the group holds the real parts of its pairs in
// columns s*CW + p and the imaginary parts in s*CW + CW/2 + p (CW = 16 / PK), so
// one pass of matrix-core instructions serves PK slices.
template <int NB, int PK = 1>  // 16-column blocks: pairs of this launch <= 8*NB; PK slices per pass
__global__ void __launch_bounds__(GM_THREADS)
k_gather_mfma(const float2 *__restrict__ G, float2 *__restrict__ gq, const SampleGeom *__restrict__ geo,
              const int *__restrict__ items, const TileEntry *__restrict__ tiles,
              const unsigned short *__restrict__ cells, const float2 *__restrict__ cf,
              int n, int nt, int MPt, int p0, int accum, float beta, int nentries, int nzb, int zg) {
    static_assert(NB == 1 || NB == 2, "one or two column blocks");
    static_assert(PK == 1 || (NB == 1 && (PK == 2 || PK == 4)), "slice packing needs one column block");
    constexpr int NC = 16 * NB, HALF = 8 * NB, NCP = NC + 1;
    constexpr int CW = NC / PK, HP = CW / 2;  // columns per packed slice; its pairs
    constexpr int CELLS = TH * TH;
    constexpr int TPS = GM_THREADS / HALF;          // staging threads per pair slot
    constexpr int NLD = (CELLS + TPS - 1) / TPS;    // tile loads per thread and pass
    const int MP = min(HP, MPt - p0);  // pairs of this launch
    // 4 spare cells: the last 4-cell step of row TH-1 may read up to 3 cells past
    // the region (their weights are zero; the spare cells are zeroed)
    __shared__ float tile[(CELLS + 4) * NCP];
    // the 9 window taps along rows, stored at [3 + a] with zeros at 0-2 and 12-15, so
    // that the row loop reads them at fixed offsets without a bounds check; along cells
    __shared__ float wty[GM_ITEMS][GM_WYP];
    __shared__ float wtx[GM_ITEMS][KW];
    __shared__ int sid[GM_ITEMS];
    __shared__ int sly[GM_ITEMS], slx[GM_ITEMS];
    const int tid = threadIdx.x;
    const int N = 2 * n, nT = tiles_per_side(N);
    const int L = xcd_remap(blockIdx.x, gridDim.x);
    const int e = L % nentries;
    const int z0 = (L / nentries) * zg;
    const int zc = min(zg, nzb - z0);
    const TileEntry te = tiles[e];
    const int ty0 = (te.tile / nT) * TS, tx0 = (te.tile % nT) * TS;
    const int cnt = te.count;
    // Tile staging: thread tid serves pair slot sp = tid / TPS (pair sp % HP of packed
    // slice sp / HP) and the cells k0 + TPS*q of the entry's cell list (the cells of
    // the region some item's window covers, raster order; the other cells stay zero
    // and meet zero taps only).  Byte offsets within one slice's spectra (all pairs of
    // a slice span MP*4n^2*8 < 2^32 bytes).  A thread's loads stop at the list's end
    // (nq, equal across the workgroup); entries below that bound but past the list,
    // or of absent pairs, read the slice's first element and are not stored.
    const int sp = tid / TPS, k0 = tid % TPS;
    const int spp = sp % HP, tsl = sp / HP;
    const int tcol = tsl * CW + spp;  // tile column of the real part
    const unsigned short *cl = cells + te.pad;
    const int ncell = cl[0];
    unsigned tofs[NLD];
    unsigned rcp[(NLD + 1) / 2];      // the cells, two 16-bit indices per word
    unsigned keep = 0;                // bit q: entry q is stored
#pragma unroll
    for (int q = 0; q < (NLD + 1) / 2; q++) rcp[q] = 0u;
#pragma unroll
    for (int q = 0; q < NLD; q++) {
        const int k = k0 + TPS * q;
        const int rc = cl[1 + max(min(k, ncell - 1), 0)];
        const int r = rc / TH, c = rc - r * TH;
        int gy = ty0 + r;
        if (gy >= N) gy -= N;
        int gx = tx0 + c;
        if (gx >= N) gx -= N;
        const bool ok = k < ncell && spp < MP;
        tofs[q] = ok ? 8u * (unsigned)((spp * N + gy) * N + gx) : 0u;
        keep |= ok ? (1u << q) : 0u;
        rcp[q / 2] |= (unsigned)rc << (16 * (q & 1));
    }
    // loads per thread that meet the list (equal across the workgroup)
    const int nq = __builtin_amdgcn_readfirstlane((ncell + TPS - 1) / TPS);
    float2 stv[NLD];
    // the tile of the pass starting at slice z, of which nv slices exist (the others
    // read the last one: finite values in columns whose results are not written)
    auto load_tile = [&](int z, int nv) {
        const char *Gz = reinterpret_cast<const char *>(G + ((size_t)z * MPt + p0) * N * N) +
                         (PK == 1 ? 0 : (size_t)min(tsl, nv - 1) * ((size_t)MPt * N * N * sizeof(float2)));
#pragma unroll
        for (int q = 0; q < NLD; q++)
            if (q < nq) stv[q] = *reinterpret_cast<const float2 *>(Gz + tofs[q]);
    };
    auto store_tile = [&] {
#pragma unroll
        for (int q = 0; q < NLD; q++) {
            if (keep & (1u << q)) {
                const int rc = (rcp[q / 2] >> (16 * (q & 1))) & 0xffff;
                tile[rc * NCP + tcol] = stv[q].x;
                tile[rc * NCP + HP + tcol] = stv[q].y;
            }
        }
    };
    // cells outside the list, columns of absent pairs (MP < HP) and the spare cells stay zero
    for (int i = tid; i < (CELLS + 4) * NCP; i += GM_THREADS) tile[i] = 0.f;
    load_tile(z0, min(PK, zc));
    // stage the items: 4 threads per item (rows taps 0-4 / 5-8, cell taps 0-4 / 5-8)
    {
        const int it = tid >> 2, part = tid & 3;
        const int id = it < cnt ? items[te.start + it] : -1;
        SampleGeom g = {0, 0, 0.f, 0.f};
        int ly0 = 64, lx0 = 64;  // empty slot: window outside the region
        if (id >= 0) {
            g = geo[id];
            ly0 = g.row0 - ty0;
            if (ly0 < 0) ly0 += N;
            lx0 = g.col0 - tx0;
            if (lx0 < 0) lx0 += N;
        }
        if (part == 0) {
            sid[it] = id;
            sly[it] = ly0;
            slx[it] = lx0;
        }
        const float fr = (part < 2) ? g.fy : g.fx;
        float *wt = (part < 2) ? &wty[it][3] : &wtx[it][0];
        const int a0 = (part & 1) ? 5 : 0, a1 = (part & 1) ? KW : 5;
        for (int a = a0; a < a1; a++) wt[a] = gauss(beta, (float)(a - 4) - fr);
        if (part == 0)
            for (int a = 0; a < 3; a++) wty[it][a] = 0.f;
        if (part == 1)
            for (int a = 3 + KW; a < GM_WYP; a++) wty[it][a] = 0.f;
    }
    lds_barrier();
    const int w = tid >> 6, l = tid & 63;
    const int li = l & 15, lk = l >> 4;
    const int b0 = w * 16;
    // block window: rows [rlo, rhi], 4-cell steps from clo (padding slots do not count)
    int rlo = 64, rhi = -1, clo = 64, chi = -1;
#pragma unroll 4
    for (int i = 0; i < 16; i++) {
        if (b0 + i < cnt && sid[b0 + i] >= 0) {
            rlo = min(rlo, sly[b0 + i]);
            rhi = max(rhi, sly[b0 + i] + KW - 1);
            clo = min(clo, slx[b0 + i]);
            chi = max(chi, slx[b0 + i] + KW - 1);
        }
    }
    // the bounds are equal across the wave: keep them in scalar registers
    rlo = __builtin_amdgcn_readfirstlane(rlo);
    rhi = __builtin_amdgcn_readfirstlane(rhi);  // -1: no item in this wave's block
    clo = __builtin_amdgcn_readfirstlane(clo) & ~3;
    const int nks = __builtin_amdgcn_readfirstlane((chi - clo) / 4 + 1);  // 3 or 4
    // this lane's item (column li of D) and its coefficients for columns 4*lk + v
    const int oid = b0 + li < cnt ? sid[b0 + li] : -1;
    float2 cfv[NB][4];
    {
        const float2 *crow = cf + (oid >= 0 ? ((size_t)((oid >> 1) / n) * 2 + (oid & 1)) * MPt + p0 : 0);
#pragma unroll
        for (int b = 0; b < NB; b++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int pp = ((16 * b + 4 * lk + v) % CW) % HP;
                cfv[b][v] = (oid >= 0 && pp < MP) ? crow[pp] : make_float2(0.f, 0.f);
            }
    }
    float wxv[GM_KS];
#pragma unroll
    for (int ks = 0; ks < GM_KS; ks++) {
        const int ax = clo + 4 * ks + lk - slx[b0 + li];
        wxv[ks] = (ks < nks && (unsigned)ax < (unsigned)KW && rhi >= 0) ? wtx[b0 + li][ax] : 0.f;
    }
    const int myly = sly[b0 + li];
    const int bp32 = 4 * (l ^ 32);  // ds_bpermute address of the other half's lane
    // row taps of this lane's item from the block's first row rlo (empty slots: any
    // in-range row; their cell taps are zero) and the tile row of rlo
    const float *wrow = &wty[b0 + li][3 + max(rlo - myly, -3)];
    const float *tb = tile + (rlo * TH + clo + lk) * NCP + li;
    float2 *out = gq + (size_t)z0 * nt * n * 2;
    for (int zz = 0; zz < zc; zz += PK) {
        if (zz > 0) lds_barrier();  // previous pass's tile fully read
        store_tile();
        if (zz + PK < zc) load_tile(z0 + zz + PK, min(PK, zc - zz - PK));
        lds_barrier();
        if (rhi >= 0) {
            v4f acc0[NB], acc1[NB];
#pragma unroll
            for (int b = 0; b < NB; b++) acc0[b] = acc1[b] = (v4f){0.f, 0.f, 0.f, 0.f};
            // a 9-cell window spans exactly 3 steps of 4 cells; blocks span 3 or 4
            // one fully unrolled row loop per block height (9 to 12 rows), so that the
            // LDS operands sit at immediate offsets from two base addresses
            auto rows = [&](auto nr) {
#pragma unroll
                for (int rr = 0; rr < decltype(nr)::value; rr++) {
                    const float wy = wrow[rr];
                    const float *trow = tb + rr * TH * NCP;
                    float bv[GM_KS][NB];
#pragma unroll
                    for (int ks = 0; ks < 3; ks++)
#pragma unroll
                        for (int b = 0; b < NB; b++) bv[ks][b] = trow[4 * ks * NCP + 16 * b];
                    if (nks > 3) {
#pragma unroll
                        for (int b = 0; b < NB; b++) bv[3][b] = trow[12 * NCP + 16 * b];
                    }
#pragma unroll
                    for (int ks = 0; ks < 3; ks++) {
                        const float av = wy * wxv[ks];
#pragma unroll
                        for (int b = 0; b < NB; b++) {
                            if (ks & 1) acc1[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(bv[ks][b], av, acc1[b], 0, 0, 0);
                            else acc0[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(bv[ks][b], av, acc0[b], 0, 0, 0);
                        }
                    }
                    if (nks > 3) {
                        const float av = wy * wxv[3];
#pragma unroll
                        for (int b = 0; b < NB; b++)
                            acc1[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(bv[3][b], av, acc1[b], 0, 0, 0);
                    }
                }
            };
            switch (rhi - rlo) {
                case 8: rows(std::integral_constant<int, 9>{}); break;
                case 9: rows(std::integral_constant<int, 10>{}); break;
                case 10: rows(std::integral_constant<int, 11>{}); break;
                default: rows(std::integral_constant<int, 12>{}); break;
            }
            // lane (li, lk): D[16*b + 4*lk + v][li]; column c is pair (c % CW) % HP of
            // packed slice c / CW, imaginary part for c % CW >= HP
            float sx = 0.f, sy = 0.f;
#pragma unroll
            for (int b = 0; b < NB; b++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const float d = acc0[b][v] + acc1[b][v];
                    if ((16 * b + 4 * lk + v) % CW < HP) {
                        sx += cfv[b][v].x * d;
                        sy += cfv[b][v].y * d;
                    } else {  // cf * (i d)
                        sx -= cfv[b][v].y * d;
                        sy += cfv[b][v].x * d;
                    }
                }
            // the lanes of one packed slice hold its CW columns: PK = 1: lk = 0..3
            // (xor 16, xor 32), PK = 2: lk = 2s, 2s + 1 (xor 16), PK = 4: lk = s
            if (PK <= 2) {
                sx = add_xor16(sx);
                sy = add_xor16(sy);
            }
            if (PK == 1) {
                sx = add_xor32(sx, bp32);
                sy = add_xor32(sy, bp32);
            }
            const int zs = zz + (PK == 1 ? 0 : lk * PK / 4);  // this lane's slice
            if ((lk % (4 / PK)) == 0 && oid >= 0 && zs < zc) {
                float2 *o = out + (size_t)zs * nt * n * 2 + oid;
                if (accum) {
                    const float2 pv = *o;
                    sx += pv.x;
                    sy += pv.y;
                }
                *o = make_float2(sx, sy);
            }
        }
    }
}

// F4: per-ray 1D transforms (one workgroup per T rays).
//   FWD: h1[s'] = Re{post_f[s'] * IFFT(pre_f[s] * (P[s] + conj Q[s]))} (- g), with P, Q
//        the direct and mirror gather sums of sample s   (shift, fftshift,
//        CUFFT_INVERSE, fftshift, mulr, mulphi scale: radonusfft.cu:96-102,
//        kernels_operators.cuh:135-149), optional residual norm partials.
//   ADJ: q[s] = post_a[s] * FFT(pre_a[s'] * h1[s'])  (radonusfft.cu:115-121 and
//        the 1/n, s == 0 rule of the adjoint gather, kernels_radonusfft.cuh:102-108),
//        written to q[s] (8 bytes per sample: whole cache lines).
enum { RAY_FWD = 1, RAY_SUBG = 2, RAY_ADJ = 4 };
template <int NP>
__global__ void __launch_bounds__(fft_bound(NP))
k_ray(const float2 *__restrict__ gq, float2 *__restrict__ q, const float *__restrict__ g, const float *__restrict__ h1in,
      float *__restrict__ h1out, const float2 *__restrict__ pre_f, const float *__restrict__ post_f,
      const float *__restrict__ pre_a, const float2 *__restrict__ post_a, FftPlan plan,
      const float2 *__restrict__ tw, int n, int nrays, int T, int stride, int mode,
      double *__restrict__ partial) {
    extern __shared__ float2 lds[];
    __shared__ double red[FFT_THREADS / 64];
    const int tid = threadIdx.x, nthr = blockDim.x;
    double acc = 0.0;
    if constexpr (NP > 0) {
        // thread (t, jl) of ray r0 + t holds detector positions s = jl + m*TPT,
        // which are both the outputs of one transform and the inputs of the next
        constexpr int nn = NP, TPT = NP / 16, TT = fft_rows_p2(NP, RAY_WG);
        const int t = tid / TPT, jl = tid % TPT;
        const int ray = blockIdx.x * TT + t;
        const bool active = ray < nrays;
        float2 *row = lds + t * stride;
        const float2 *twn = tw + plan.tw_off;
        float2 v[16];
#pragma unroll
        for (int m = 0; m < 16; m++) v[m] = make_float2(0.f, 0.f);
        const size_t o0 = (size_t)ray * nn + jl;
        if (mode & RAY_FWD) {
            if (active) {
                const float4 *pq4 = reinterpret_cast<const float4 *>(gq) + o0;
#pragma unroll
                for (int m = 0; m < 16; m++) {
                    const float4 pq = pq4[m * TPT];
                    // slot pair of s == 0 is never written by the gather
                    if (m > 0 || jl > 0) v[m] = cmul(make_float2(pq.x + pq.z, pq.y - pq.w), pre_f[jl + m * TPT]);
                }
            }
            fft_p2_reg<NP, true>(v, row, active, jl, twn);
            if (active) {
#pragma unroll
                for (int m = 0; m < 16; m++) {
                    const int sd = jl + m * TPT;
                    float val = v[m].x * post_f[sd];
                    if (mode & RAY_SUBG) val -= g[o0 + m * TPT];
                    acc += (double)val * val;
                    if (h1out) h1out[o0 + m * TPT] = val;
                    v[m] = make_float2(val * pre_a[sd], 0.f);
                }
            }
        } else if (active) {
#pragma unroll
            for (int m = 0; m < 16; m++) v[m] = make_float2(h1in[o0 + m * TPT] * pre_a[jl + m * TPT], 0.f);
        }
        if (mode & RAY_ADJ) {
            __syncthreads();  // the row is reused
            fft_p2_reg<NP, false>(v, row, active, jl, twn);
            if (active) {
#pragma unroll
                for (int m = 0; m < 16; m++)
                    q[o0 + m * TPT] = (m > 0 || jl > 0) ? cmul(v[m], post_a[jl + m * TPT]) : make_float2(0.f, 0.f);
            }
        }
    } else {
    const int r0 = blockIdx.x * T;
    const int Tc = min(T, nrays - r0);
    if (mode & RAY_FWD) {
        // the slot pair of s == 0 is never written by the gather: loaded, not used
        staged_elems<16, float2>(
            tid, nthr, Tc * n,
            [&](int idx) {
                const int t = idx / n, s = idx - t * n;
                const float4 pq = reinterpret_cast<const float4 *>(gq)[(size_t)(r0 + t) * n + s];
                return cmul(make_float2(pq.x + pq.z, pq.y - pq.w), pre_f[s]);
            },
            [&](int idx, float2 val) {
                const int t = idx / n, s = idx - t * n;
                lds[t * stride + lp(s)] = s != 0 ? val : make_float2(0.f, 0.f);
            });
        __syncthreads();
        fft_lds<FFT_E, true, true, NP == -2>(lds, stride, Tc, plan, tw, tid, nthr);
        staged_elems<16, float>(
            tid, nthr, Tc * n,
            [&](int idx) { return (mode & RAY_SUBG) ? g[(size_t)r0 * n + idx] : 0.f; },
            [&](int idx, float gv) {
                const int t = idx / n, s = idx - t * n;
                const size_t o = (size_t)(r0 + t) * n + s;
                const float val = lds[t * stride + lp(s)].x * post_f[s] - gv;
                acc += (double)val * val;
                if (h1out) h1out[o] = val;
                lds[t * stride + lp(s)] = make_float2(val * pre_a[s], 0.f);
            });
    } else {
        staged_elems<16, float>(
            tid, nthr, Tc * n, [&](int idx) { return h1in[(size_t)r0 * n + idx]; },
            [&](int idx, float hv) {
                const int t = idx / n, s = idx - t * n;
                lds[t * stride + lp(s)] = make_float2(hv * pre_a[s], 0.f);
            });
    }
    if (mode & RAY_ADJ) {
        __syncthreads();
        fft_lds<FFT_E, false, true, NP == -2>(lds, stride, Tc, plan, tw, tid, nthr);
        staged_elems<16, float2>(
            tid, nthr, Tc * n, [&](int idx) { return post_a[idx % n]; },
            [&](int idx, float2 pa) {
                const int t = idx / n, s = idx - t * n;
                q[(size_t)r0 * n + idx] = s != 0 ? cmul(lds[t * stride + lp(s)], pa) : make_float2(0.f, 0.f);
            });
    }
    }
    if (partial) {
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
        if ((tid & 63) == 0) red[tid >> 6] = acc;
        __syncthreads();
        if (tid == 0) {
            double s = 0.0;
            for (int w = 0; w < (nthr + 63) / 64; w++) s += red[w];
            partial[blockIdx.x] = s;
        }
    }
}


// A1: tile-owned Gaussian scatter on the FP32 matrix cores (replaces gather ADJ
// with global atomicAdd, kernels_radonusfft.cuh:126-131, and wrap ADJ, :59-62,
// run once per basis function after mulphi(-1), operators.cu:47-56).
// One workgroup (16 waves) owns one TS x TS tile of the 2n grid for ZW
// slices; wave w owns the 4 x 4 cell block (rows 4*(w/4).., cells 4*(w%4)..).
// The items (direct and mirror windows of the samples, see F3) whose 9x9
// window intersects the tile are staged in chunks of SM_CH in LDS once for the
// ZW slices: 16-wide weight rows wyr[s][r] = wy[r-ly0], wxr[s][i] = wx[i-lx0]
// (zero outside the window) and, per slice, P[s][col] = conj(cf[k][kind][p]) * q'
// (q' = q direct, conj(q) mirror) with the real part in column p and the
// imaginary part in column 8*NB + p.  The result is the Hermitian-symmetrised
// pair grid whose inverse transform is f_{2p} + i f_{2p+1}.  The host lists,
// per chunk and block, the slots of the items whose window intersects the
// block (ascending, padded to a multiple of 4 with the zero slot SM_CH); the
// wave walks its list and accumulates
//   D[cell][col] += sum_k wyr[k][row(cell)] * wxr[k][x(cell)] * P[k][col]
// with v_mfma_f32_16x16x4f32 (16 cells x 16 columns x 4 items per instruction).
// A 4 x 4 block keeps 56% of the products nonzero against 37.5% for a 16-cell
// tile row (9-tap window).  Lane (li, lk) ends with cells 0..3 of block row lk
// for column li and writes them straight to HBM.  No atomics; fixed summation
// order.
constexpr int SM_CH = 128;  // items staged per chunk (8 staging threads each)
constexpr int SM_ZW = 4;    // default slices per workgroup (template argument ZW)
constexpr int SM_LB = 16 * (SM_CH + 4);  // list bytes per chunk (16 blocks, padded)

// Weight rows of the scatter's adjoint chunks, once per geometry: for chunk ci
// (achunk[ci] = {tile, first record, count}) and slot s < count, 16-wide rows
// wy[r] = gauss(r - ly0 - 4 - fy), wx[c] = gauss(c - lx0 - 4 - fx) over the tile's
// rows / cells, zero outside the window, stored as 8 float4 per slot (wy then wx).
// The gather computes its taps with the same function, so the two operators stay
// exact transposes of each other.
__global__ void __launch_bounds__(1024) k_scatter_weights(const AdjItem *__restrict__ aitems,
                                                          const int4 *__restrict__ achunk, float4 *__restrict__ wtab,
                                                          int n, float beta) {
    const int ci = blockIdx.x;
    const int4 ch = achunk[ci];
    const int sm = threadIdx.x >> 3, part = threadIdx.x & 7;
    if (sm >= ch.z) return;
    const int N = 2 * n, nT = tiles_per_side(N);
    const int ty0 = (ch.x / nT) * TS, tx0 = (ch.x % nT) * TS;
    const AdjItem g = aitems[ch.y + sm];
    int dy = (g.rc & 0xffff) - ty0;
    if (dy < 0) dy += N;
    int dx = (g.rc >> 16) - tx0;
    if (dx < 0) dx += N;
    const int l0 = part < 4 ? (dy < TS ? dy : dy - N) : (dx < TS ? dx : dx - N);
    const float fr = part < 4 ? g.fy : g.fx;
    float w[4];
#pragma unroll
    for (int h = 0; h < 4; h++) {
        const int a = 4 * (part & 3) + h - l0;
        w[h] = ((unsigned)a < (unsigned)KW) ? gauss(beta, (float)(a - 4) - fr) : 0.f;
    }
    wtab[((size_t)ci * SM_CH + sm) * 8 + part] = make_float4(w[0], w[1], w[2], w[3]);
}

// With few pairs (PK = 2: at most 4, PK = 4: at most 2; NB = 1) PK slices share the
// 16 columns of P and of the accumulators: slice zz of the workgroup goes to
// plane zz / PK, columns (zz % PK)*CW + p (real) and + CW/2 (imaginary), CW = 16 / PK.
template <int NB, int ZW, int PK = 1>  // 16-column blocks: MP <= 8*NB (NB = 1 or 2); ZW slices per workgroup
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8, 8)))  // two workgroups per CU
k_scatter(float2 *__restrict__ G, const float2 *__restrict__ qs, const int2 *__restrict__ arec,
          const TileEntry *__restrict__ tiles, const int2 *__restrict__ crange,
          const int2 *__restrict__ brange, const unsigned char *__restrict__ blist, const float2 *__restrict__ cf,
          const float4 *__restrict__ wtab, int n, int nt, int MPt, int p0, int ntiles, int nzb) {
    static_assert(NB == 1 || NB == 2, "one or two column blocks");
    static_assert(PK == 1 || (NB == 1 && (PK == 2 || PK == 4) && ZW % PK == 0), "slice packing: one column block");
    constexpr int NG = ZW / PK;        // planes of P and accumulators
    constexpr int CW = 16 * NB / PK, HP = CW / 2;  // columns per packed slice; its pairs
    // this launch writes the G planes of pairs p0 .. p0 + MP - 1 of the MPt pairs
    // (more than 16 pairs run as one launch per group)
    const int MP = min(HP, MPt - p0);
    constexpr int NC = 16 * NB;
    constexpr int HALF = 8 * NB;       // column of the imaginary part of pair 0 (PK = 1)
    // P rows: NC columns, unpadded.  The host orders each block's list in groups of
    // 4 slots that are distinct mod 4 wherever it can, so one matrix-core operand
    // read (4 slots x 16 columns) meets 4 different 16-bank groups; for NB = 2 the
    // column halves of slots with bit 1 set are swapped (col ^ 16), which the reads
    // undo per lane.  The weight rows (16 floats per slot) fall likewise.
    constexpr int NCP = NC;
    constexpr int NPP = HALF / 8;      // pairs per staging thread
    constexpr int NS = SM_CH + 1;      // slots incl. the zero slot
    __shared__ float wyr[NS][TS];
    __shared__ float wxr[NS][TS];
    __shared__ float sP[NG][NS][NCP];
    __shared__ unsigned char sl[SM_LB];
    const int tid = threadIdx.x;
    const int N = 2 * n, nT = tiles_per_side(N);
    const int L = xcd_remap(blockIdx.x, gridDim.x);
    const int e = L % ntiles;
    const int z0 = (L / ntiles) * ZW;
    const int zc = min(ZW, nzb - z0);
    const TileEntry te = tiles[e];
    const int ty0 = (te.tile / nT) * TS, tx0 = (te.tile % nT) * TS;
    const int w = tid >> 6, l = tid & 63;
    const int li = l & 15, lk = l >> 4;
    const int br0 = 4 * (w >> 2), bc0 = 4 * (w & 3);   // this wave's block
    const int ar = br0 + (li >> 2), ac = bc0 + (li & 3);  // cell of A row li
    const int sm = tid >> 3, part = tid & 7;  // staging: item slot, eighth
    // the zero slot (never overwritten by the chunk staging)
    if (tid < TS) {
        wyr[SM_CH][tid] = 0.f;
        wxr[SM_CH][tid] = 0.f;
    }
    if (tid < NG * NCP) sP[tid / NCP][SM_CH][tid % NCP] = 0.f;
    v4f acc[NG][NB];
#pragma unroll
    for (int g = 0; g < NG; g++)
#pragma unroll
        for (int b = 0; b < NB; b++) acc[g][b] = (v4f){0.f, 0.f, 0.f, 0.f};
    // chunk staging data are prefetched into registers one chunk ahead, so the
    // dependent global loads (item -> q, coefficients) overlap the MFMA loop; the
    // weight rows come precomputed (k_scatter_weights), one float4 per thread
    // Every load below is unconditional (indices clamped into range; values past a
    // chunk's end are staged but never listed), so that the compiler's wait counts
    // stay exact and the loads of the next chunk remain in flight through the
    // matrix-core loop.  The record of a slot is loaded two chunks ahead, so the
    // dependent loads (q, coefficients) can be issued at once one chunk ahead.
    bool has = false;  // this thread's slot holds a record of the chunk
    bool mir = false;  // ... and that record is a mirrored sample
    float4 pw = make_float4(0.f, 0.f, 0.f, 0.f);
    const size_t sq = (size_t)nt * n;  // float2 per slice of q
    float2 pq[ZW], pph[NPP];
    constexpr int NLW = SM_LB / 4;     // list words per chunk staging (blist is padded by SM_LB bytes)
    static_assert(NLW <= 1024, "one list word per thread");
    unsigned plw = 0u;
    int2 nrec = make_int2(0, 0);       // record of slot sm in the chunk after the prefetched one
    int2 ncr = make_int2(0, 0);        // list range of that chunk
    auto rec_of = [&](int c0) {
        const int cnt = min(SM_CH, te.count - c0);
        return arec[te.start + c0 + min(sm, cnt - 1)];
    };
    auto prefetch = [&](int c0, int2 rec, int2 cr) {
        const int ci = te.pad + c0 / SM_CH;
        has = sm < min(SM_CH, te.count - c0);
        mir = rec.x & 1;
        // slot sm's weight rows (8 float4 per slot: thread tid holds part tid & 7)
        pw = wtab[(size_t)ci * SM_CH * 8 + tid];
        // record: item 2 * sample | mirror flag, coefficient row
        const unsigned qb = 4u * (unsigned)(rec.x & ~1);  // byte offset of the sample's q in a slice
#pragma unroll
        for (int zz = 0; zz < ZW; zz++)
            pq[zz] = *reinterpret_cast<const float2 *>(
                reinterpret_cast<const char *>(qs + (size_t)min(z0 + zz, nzb - 1) * sq) + qb);
#pragma unroll
        for (int q = 0; q < NPP; q++) pph[q] = cf[rec.y + p0 + min(part + 8 * q, MP - 1)];
        // the chunk's block lists: SM_LB bytes from its first (4-byte aligned) list
        plw = reinterpret_cast<const unsigned *>(blist + __builtin_amdgcn_readfirstlane(cr.x))[min(tid, NLW - 1)];
    };
    if (te.count > 0) {
        const int2 r0 = rec_of(0);
        prefetch(0, r0, crange[te.pad]);
        if (SM_CH < te.count) {
            nrec = rec_of(SM_CH);
            ncr = crange[te.pad + 1];
        }
    }
    for (int c0 = 0; c0 < te.count; c0 += SM_CH) {
        // this wave's list within the chunk: (byte offset, length, a multiple of 4)
        const int2 lst = brange[((size_t)te.pad + c0 / SM_CH) * 16 + __builtin_amdgcn_readfirstlane(w)];
        lds_barrier();
        if (tid < NLW) reinterpret_cast<unsigned *>(sl)[tid] = plw;
        if (has) {
            // weight rows: parts 0-3 hold wy[4 part ..], parts 4-7 wx[4 (part - 4) ..]
            float *wd = part < 4 ? &wyr[sm][4 * part] : &wxr[sm][4 * (part - 4)];
            *reinterpret_cast<float4 *>(wd) = pw;
#pragma unroll
            for (int zz = 0; zz < ZW; zz++) {
                const float2 qv = make_float2(pq[zz].x, mir ? -pq[zz].y : pq[zz].y);
#pragma unroll
                for (int q = 0; q < NPP; q++) {
                    const int j = part + 8 * q;
                    // q' * conj(cf); zero for pairs j >= MP
                    const float2 p = j < MP ? cmulc(qv, pph[q]) : make_float2(0.f, 0.f);
                    if (PK == 1) {
                        const int sw = NB == 2 ? 16 * ((sm >> 1) & 1) : 0;
                        sP[zz][sm][j ^ sw] = p.x;
                        sP[zz][sm][(HALF + j) ^ sw] = p.y;
                    } else if (j < HP) {
                        sP[zz / PK][sm][(zz % PK) * CW + j] = p.x;
                        sP[zz / PK][sm][(zz % PK) * CW + HP + j] = p.y;
                    }
                }
            }
        }
        // slots of the chunk past its item count are never listed
        if (c0 + SM_CH < te.count) {
            prefetch(c0 + SM_CH, nrec, ncr);
            if (c0 + 2 * SM_CH < te.count) {
                nrec = rec_of(c0 + 2 * SM_CH);
                ncr = crange[te.pad + c0 / SM_CH + 2];
            }
        }
        lds_barrier();
        const unsigned char *ls = sl + lst.x + lk;
        const int len = __builtin_amdgcn_readfirstlane(lst.y);
#pragma unroll 2
        for (int k0 = 0; k0 < len; k0 += 4) {
            const int slot = ls[k0];
            const float av = wyr[slot][ar] * wxr[slot][ac];
            const int sh = NB == 2 ? (slot >> 1) & 1 : 0;  // swapped halves
            const float *pp = &sP[0][slot][li + 16 * sh];
#pragma unroll
            for (int g = 0; g < NG; g++) {
#pragma unroll
                for (int b = 0; b < NB; b++) {
                    const float bv = pp[g * NS * NCP + b * (16 - 32 * sh)];
                    acc[g][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc[g][b], 0, 0, 0);
                }
            }
        }
    }
    // acc[g][b][v] = D_g[cell 4*lk + v][col 16*b + li] = block row lk, cell v: the
    // real part of pair p of packed slice s in column s*CW + p, its imaginary part in
    // column s*CW + HP + p (PK = 1: s = 0, HP = HALF)
    float2 *Gt = G + (size_t)(ty0 + br0 + lk) * N + tx0 + bc0;
    const int ls = PK == 1 ? 0 : li / CW, lp_ = PK == 1 ? li : li % CW;  // this lane's packed slice, pair
#pragma unroll
    for (int g = 0; g < NG; g++) {
        float re[4], im[4];
#pragma unroll
        for (int v = 0; v < 4; v++) {
            if constexpr (NB == 1) {
                re[v] = acc[g][0][v];
                im[v] = __shfl(acc[g][0][v], (l & 48) | ((li + HP) & 15), 64);
            } else {
                re[v] = acc[g][0][v];
                im[v] = acc[g][1][v];
            }
        }
        const int zz = g * PK + ls;
        // cells past 2n (a narrow last tile) are not this tile's; 2n and tx0 + bc0 are even
        if (zz < zc && lp_ < MP && ty0 + br0 + lk < N) {
            float4 *o = reinterpret_cast<float4 *>(Gt + ((size_t)(z0 + zz) * MPt + p0 + lp_) * N * N);
            if (tx0 + bc0 < N) o[0] = make_float4(re[0], im[0], re[1], im[1]);
            if (tx0 + bc0 + 2 < N) o[1] = make_float4(re[2], im[2], re[3], im[3]);
        }
    }
}

// A2: column inverse FFT of G keeping the n centre rows (cuFFT INVERSE of
// radonusfft.cu:128 with the checkerboards, first half).
template <int NP>
__global__ void __launch_bounds__(fft_bound(NP))
k_adj_cols(const float2 *__restrict__ G, float2 *__restrict__ H, FftPlan plan, const float2 *__restrict__ tw,
           const int2 *__restrict__ runs, int n, int M, int T, int stride) {
    extern __shared__ float2 lds[];
    const int tid = threadIdx.x;
    const int L = xcd_remap(blockIdx.x, gridDim.x);
    if constexpr (NP > 0 && rp_radix(NP) > 0) {
        // length R0 * P: thread (t, g, jl), t fastest, supplies ky = R0 (jl + m TPS) + g
        // and receives the positions jl + (16 g + m) TPS, of which the n centre ones
        // (16 g + m in [4 R0, 12 R0)) are kept as rows y of H
        constexpr int N = NP, nn = NP / 2, TPT = NP / 16, R0 = rp_radix(NP), P = NP / R0, TPS = P / 16;
        constexpr int TT = fft_rp_count(NP, RTV_ACOLS_THREADS);
        const int ntile = N / TT;
        const int kt = L % ntile;
        const int zj = L / ntile;
        const int t = tid % TT, r = tid / TT, g = r / TPS, jl = r % TPS;
        const int kx = kt * TT + t;
        const float sgi = sgn(jl + g), sgo = sgn(jl);
        const int2 run = runs[kx];
        const float2 *Gs = G + (size_t)zj * N * N + (size_t)kx * N;
        const float2 *G0 = G + (size_t)zj * N * N;
        float2 v[16];
#pragma unroll
        for (int m = 0; m < 16; m++) {
            const int ky = R0 * (jl + m * TPS) + g;
            const bool in = in_run(run, ky, N);
            v[m] = cscale(*(in ? Gs + ky : G0), in ? sgi : 0.f);
        }
        fft_rp_reg<R0, P, true>(v, lds + t * stride, true, g, jl, tw + plan.tw_sub, tw + plan.tw_off);
        constexpr int GA = h_group_adj(NP);
#pragma unroll
        for (int m = 0; m < 16; m++) {
            const int q = 16 * g + m - 4 * R0;  // centre row index / TPS
            if (q >= 0 && q < 8 * R0) H[h_index(zj, nn, jl + q * TPS, kx, GA)] = cscale(v[m], sgo);
        }
    } else if constexpr (NP > 0) {
        // thread (t, jl), t fastest; output positions 4*TPT .. 12*TPT - 1 are kept
        constexpr int N = NP, nn = NP / 2, TPT = NP / 16, TT = fft_cols_p2(NP, RTV_ACOLS_THREADS);
        const int ntile = N / TT;
        const int kt = L % ntile;
        const int zj = L / ntile;
        const int t = tid % TT, jl = tid / TT;
        const float sg = sgn(jl);
        const float2 *Gs = G + (size_t)zj * N * N + (size_t)(kt * TT + t) * N + jl;  // row kx of G
        // cells outside the windows are zero: their lanes all read the spectrum's
        // first cell instead (one cache line per wave; a finite value, scaled by
        // zero, so that the loads stay unpredicated and are issued together)
        const int2 run = runs[kt * TT + t];
        const float2 *G0 = G + (size_t)zj * N * N;
        float2 v[16];
#pragma unroll
        for (int m = 0; m < 16; m++) {
            const bool in = in_run(run, jl + m * TPT, N);
            v[m] = cscale(*(in ? Gs + m * TPT : G0), in ? sg : 0.f);
        }
        fft_p2_reg<NP, true>(v, lds + t * stride, true, jl, tw + plan.tw_off);
        constexpr int GA = h_group_adj(NP);
        float2 *Hs = H + h_index(zj, nn, jl, kt * TT + t, GA);  // rows jl + (m-4)*TPT: stride GA
#pragma unroll
        for (int m = 4; m < 12; m++) Hs[(size_t)(m - 4) * TPT * GA] = cscale(v[m], sg);
    } else {
        const int nthr = blockDim.x;
        const int N = 2 * n, h = N / 4;
        const int ntile = N / T;
        const int kt = L % ntile;
        const int zj = L / ntile;
        const int kx0 = kt * T;
        const float2 *Gs = G + (size_t)zj * N * N + (size_t)kx0 * N;
        staged_elems<16, float2>(
            tid, nthr, T * N,
            [&](int idx) {
                const int t = idx / N, ky = idx - t * N;
                const float2 g = Gs[(size_t)t * N + ky];  // finite: cells outside the runs hold zero
                return in_run(runs[kx0 + t], ky, N) ? cscale(g, sgn(ky)) : make_float2(0.f, 0.f);
            },
            [&](int idx, float2 val) {
                const int t = idx / N, ky = idx - t * N;
                lds[t * stride + lp(ky)] = val;
            });
        __syncthreads();
        fft_lds<FFT_E, true, true, NP == -2>(lds, stride, T, plan, tw, tid, nthr);
        const int ga = h_group_adj(N);
        for (int idx = tid; idx < T * n; idx += nthr) {
            const int t = idx % T, y = idx / T;
            H[h_index(zj, n, y, kx0 + t, ga)] = cscale(lds[t * stride + lp(y + h)], sgn(y + h));
        }
    }
}

// A3: row inverse FFT keeping the n centre columns, deapodisation (divphi ADJ,
// kernels_radonusfft.cuh:19-23).  One workgroup per (z, j, T rows).
template <int NP>
__global__ void __launch_bounds__(fft_bound(NP))
k_adj_rows(const float2 *__restrict__ H, float2 *__restrict__ v, const float *__restrict__ deap, FftPlan plan,
           const float2 *__restrict__ tw, int n, int M, int T, int stride) {
    extern __shared__ float2 lds[];
    const int tid = threadIdx.x;
    const int L = xcd_remap(blockIdx.x, gridDim.x);
    if constexpr (NP > 0 && rp_radix(NP) > 0) {
        // length R0 * P: the workgroup's TT rows of H are read whole into the LDS row
        // buffers; thread (t, g, jl) takes kx = R0 (jl + m TPS) + g and receives the
        // positions jl + (16 g + m) TPS, of which the n centre ones are the pixels
        constexpr int N = NP, nn = NP / 2, TPT = NP / 16, R0 = rp_radix(NP), P = NP / R0, TPS = P / 16;
        constexpr int TT = fft_rp_count(NP, FFT_WG);
        constexpr int nyt = nn / TT;
        const int yt = L % nyt;
        const int zj = L / nyt;
        const int t = tid / TPT, r = tid % TPT, g = r / TPS, jl = r % TPS;
        const int y = yt * TT + t;
        const float sgi = sgn(jl + g), sgo = sgn(jl);
        constexpr int GA = h_group_adj(NP);
        const int y0 = yt * TT;
        staged_elems<16, float2>(
            tid, TT * TPT, TT * N,
            [&](int idx) {
                const int rr = idx / N, kx = idx - rr * N;
                return H[h_index(zj, nn, y0 + rr, kx, GA)];
            },
            [&](int idx, float2 val) {
                const int rr = idx / N, kx = idx - rr * N;
                lds[rr * stride + lp(kx)] = val;
            });
        __syncthreads();
        float2 u[16];
#pragma unroll
        for (int m = 0; m < 16; m++) u[m] = cscale(lds[t * stride + lp(R0 * (jl + m * TPS) + g)], sgi);
        __syncthreads();
        fft_rp_reg<R0, P, true>(u, lds + t * stride, true, g, jl, tw + plan.tw_sub, tw + plan.tw_off);
        const float sy = sgo * deap[y];
        float2 *vr = v + ((size_t)zj * nn + y) * nn;
#pragma unroll
        for (int m = 0; m < 16; m++) {
            const int q = 16 * g + m - 4 * R0;
            if (q >= 0 && q < 8 * R0) {
                const int x = jl + q * TPS;
                st_nt(vr + x, cscale(u[m], sy * deap[x]));
            }
        }
    } else if constexpr (NP > 0) {
        constexpr int N = NP, nn = NP / 2, TPT = NP / 16, TT = fft_rows_p2(NP);
        constexpr int nyt = nn / TT;
        const int yt = L % nyt;
        const int zj = L / nyt;
        const int t = tid / TPT, jl = tid % TPT;
        const int y = yt * TT + t;
        const float sg = sgn(jl);
        constexpr int GA = h_group_adj(NP);
        float2 u[16];
        if constexpr (GA < N) {
            // grouped layout: each group of GA columns holds the workgroup's TT rows as one
            // contiguous block, so the rows are read block by block (16 bytes per lane,
            // consecutive lanes on consecutive bytes) into the LDS row buffers, from
            // which each thread takes its 16 elements
            static_assert(GA % 2 == 0, "16-byte reads of whole column pairs");
            constexpr int NTH = TT * TPT, GROUP = TT * GA;
            const int y0 = yt * TT;
            const float2 *Hz = H + (size_t)zj * N * nn;
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const int f = 2 * (tid + k * NTH);  // float2 index in the workgroup's rows, group-major
                const int g = f / GROUP, rem = f - g * GROUP;
                const int r = rem / GA, c = rem - r * GA;
                const float4 q = *reinterpret_cast<const float4 *>(Hz + ((size_t)g * nn + y0 + r) * GA + c);
                const int kx = g * GA + c;
                lds[r * stride + lp(kx)] = make_float2(q.x, q.y);
                lds[r * stride + lp(kx + 1)] = make_float2(q.z, q.w);
            }
            __syncthreads();
#pragma unroll
            for (int m = 0; m < 16; m++) u[m] = cscale(lds[t * stride + lp(jl + m * TPT)], sg);
            __syncthreads();
        } else {
#pragma unroll
            for (int m = 0; m < 16; m++) u[m] = cscale(H[h_row_off<TPT, GA>(zj, nn, y, jl, m)], sg);
        }
        fft_p2_reg<NP, true>(u, lds + t * stride, true, jl, tw + plan.tw_off);
        const float sy = sg * deap[y];
        float2 *vr = v + ((size_t)zj * nn + y) * nn + jl;
#pragma unroll
        for (int m = 4; m < 12; m++) st_nt(vr + (m - 4) * TPT, cscale(u[m], sy * deap[jl + (m - 4) * TPT]));
    } else {
        const int nthr = blockDim.x;
        const int N = 2 * n, h = N / 4;
        const int nyt = (n + T - 1) / T;
        const int yt = L % nyt;
        const int zj = L / nyt;
        const int y0 = yt * T;
        const int Tc = min(T, n - y0);
        const int ga = h_group_adj(N);
        staged_elems<16, float2>(
            tid, nthr, Tc * N,
            [&](int idx) {
                const int t = idx / N, kx = idx - t * N;
                return cscale(H[h_index(zj, n, y0 + t, kx, ga)], sgn(kx));
            },
            [&](int idx, float2 val) {
                const int t = idx / N, kx = idx - t * N;
                lds[t * stride + lp(kx)] = val;
            });
        __syncthreads();
        fft_lds<FFT_E, true, true, NP == -2>(lds, stride, Tc, plan, tw, tid, nthr);
        float2 *vs = v + ((size_t)zj * n + y0) * n;
        staged_elems<8, float>(
            tid, nthr, Tc * n,
            [&](int idx) {
                const int t = idx / n, x = idx - t * n;
                return sgn(x + h) * deap[x] * deap[y0 + t];
            },
            [&](int idx, float sc) {
                const int t = idx / n, x = idx - t * n;
                vs[(size_t)t * n + x] = cscale(lds[t * stride + lp(x + h)], sc);
            });
    }
}

// Pixel (x, y) lies in the support mask (circ, kernels_radonusfft.cuh:146-159): the
// host evaluates 4 xf^2 + 4 yf^2 < 1 - r in float per pixel and stores each row's
// kept columns as one range, so the kernels compare integers only.
__device__ __forceinline__ bool in_support(const int2 *supp, int x, int y) {
    if (!supp) return true;
    const int2 r = supp[y];
    return x >= r.x && x < r.y;
}

// Filtered back-projection as the starting iterate: every frame i of slice z gets
// kappa * Re v[z][0] (the adjoint of the filtered rays through a basis that selects
// one static frame), masked by the support when one is set.
__global__ void k_fbp_frames(float *__restrict__ f, const float2 *__restrict__ v, int n, int M, float kappa,
                             const int2 *__restrict__ supp) {
    const size_t plane = (size_t)n * n;
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int z = blockIdx.y;
    if (p >= plane) return;
    const int MP = (M + 1) / 2;
    const int y = (int)(p / n), x = (int)(p - (size_t)y * n);
    const float val = in_support(supp, x, y) ? kappa * v[(size_t)z * MP * plane + p].x : 0.f;
    float *fz = f + (size_t)z * M * plane + p;
    for (int i = 0; i < M; i++) fz[(size_t)i * plane] = val;
}

// f[z][i] (+)= tau * frame i of the adjoint pairs v (re of pair i/2 for even i,
// im for odd i): recphi + addreal, kernels_operators.cuh:118-133 and :75-84.
// Used by the standalone R^* entry point.
__global__ void k_unpack(float *__restrict__ f, const float2 *__restrict__ v, int n, int M, float tau, int accumulate) {
    const size_t plane = (size_t)n * n;
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int z = blockIdx.y;
    if (p >= plane) return;
    const int MP = (M + 1) / 2;
    const float2 *vz = v + (size_t)z * MP * plane + p;
    float *fz = f + (size_t)z * M * plane + p;
    for (int i = 0; i < M; i++) {
        const float2 a = vz[(size_t)(i >> 1) * plane];
        const float val = tau * ((i & 1) ? a.y : a.x);
        fz[(size_t)i * plane] = accumulate ? fz[(size_t)i * plane] + val : val;
    }
}

// ---------------------------------------------------------------------------
// TV / ADMM kernels.  Scales (kernels_operators.cuh:40-43, 69-72):
//   gradient: x, y, z: 1/(cbrt2*lambda1); t: 1/cbrt2
//   div:      x, y, z: 1/(2 cbrt2 lambda1); t: 1/(2 cbrt2)
struct TvArgs {
    int n, M, nzp, nchunk;  // nchunk = total chunks (border flags)
    int chunk0;             // first chunk handled by this launch
    float gs, gt, ks, kt;
    const float *rho;  // device-resident penalty (updated by k_rho_rule between outer iterations)
    float lam0;
    const int2 *supp = nullptr;  // support mask (kept x range [lo, hi) per row y), null: none
    int zseg = 1;                // slice segments per chunk of the column-marching kernels
};


// TV kernels: workgroups of 64 x TV_ROWS threads over (x, y) tiles of one chunk.
// The 1D grid (x tiles fastest, then y tiles, then chunks) is remapped so that
// consecutive tiles share an XCD and its L2, which then serves the rows and
// columns that neighbouring tiles read across their edges.
#ifndef RTV_TV_ROWS
#define RTV_TV_ROWS 16  // rows per TV workgroup (build-time tuning constant)
#endif
constexpr int TV_ROWS = RTV_TV_ROWS;
constexpr int TV_THREADS = 64 * TV_ROWS;
struct TvTile {
    int x, y, c, b;  // column, row, chunk, logical workgroup index
    int z0, z1;      // the chunk's slices [z0, z1) this column marches over
};
// Column-marching kernels split each chunk's slices into a.zseg segments (more
// workgroups when n^2 x chunks alone would not fill the device); the grid holds
// (x tiles, y tiles, segments, chunks) with x tiles fastest.
__device__ __forceinline__ TvTile tv_tile(const TvArgs &a) {
    const int gx = (a.n + 63) / 64, gy = (a.n + TV_ROWS - 1) / TV_ROWS;
    const int L = xcd_remap(blockIdx.x, gridDim.x);
    const int bx = L % gx, by = (L / gx) % gy, bz = L / (gx * gy);
    const int sg = bz % a.zseg;
    return TvTile{bx * 64 + (int)(threadIdx.x & 63), by * TV_ROWS + (int)(threadIdx.x >> 6), a.chunk0 + bz / a.zseg, L,
                  sg * a.nzp / a.zseg, (sg + 1) * a.nzp / a.zseg};
}

__device__ __forceinline__ size_t dual_idx(const TvArgs &a, int c, int zl, int i, int y, int x) {
    return ((((size_t)c * (a.nzp + 1) + zl) * (a.M + 1) + i) * (a.n + 1) + y) * (a.n + 1) + x;
}

// U0: the dual part of the inner step's divergence.  In the inner steps
// (admm.cu:85-97) h2 = rho*(grad fm - psi) + mu = rho*grad fm + chi with
// chi = mu - rho*psi fixed for the whole outer iteration, and div is linear, so
//   dchi[z][i][y][x] = div(chi) = ks (chi_x[x'-1] - chi_x[x']) + ks (chi_y[y'-1] - chi_y[y'])
//                    + kt (chi_t[i'-1] - chi_t[i']) + ks (chi_z[z'-1] - chi_z[z'])
// at the dual position (z', i', y', x') = (zl+1, i+1, y+1, x+1) of the voxel
// (div, kernels_operators.cuh:46-73) is computed once per outer iteration, after
// the rho rule.  Thread per (chunk, y, x) column; the t and z neighbours are the
// previous frame's / slice's own values and are carried.
// Frames i0 .. i0 + MM - 1 (one launch per group of MM frames when M > MM; i0 even).
// With CHI the first array holds chi0 = mu - rho0 psi as k_prox<true> stored it,
// rho0 = a.rho[1] the penalty of that prox and rho = a.rho[0] the one after the
// rule: chi = mu - rho psi = (1 - c) mu + c chi0 with c = rho / rho0 (1, 2 or 1/2),
// so mu is read only when the rule changed rho.
template <int MM, bool CHI>
__global__ void __launch_bounds__(TV_THREADS)
k_divchi(const float4 *__restrict__ psi, const float4 *__restrict__ mu, float *__restrict__ dchi, TvArgs a, int i0) {
    const TvTile tt = tv_tile(a);
    const int x = tt.x, y = tt.y, c = tt.c;
    const int n = a.n, M = a.M, nzp = a.nzp;
    if (x >= n || y >= n) return;
    const size_t plane = (size_t)n * n, S = (size_t)M * plane;
    const size_t col = (size_t)y * n + x;
    const float r = *a.rho;
    // the group's frames seen from frame i0: dual entries shift by i0 planes of
    // (n+1)^2 (strides keep the full frame count), outputs by i0 frame planes
    psi += (size_t)i0 * (n + 1) * (n + 1);
    mu += (size_t)i0 * (n + 1) * (n + 1);
    dchi += (size_t)i0 * plane;
    const int Mg = M - i0;
    const float cr = CHI ? r / a.rho[1] : 1.f;
    const bool mix = CHI && cr != 1.f;  // uniform
    const float cm = 1.f - cr;
    auto chi = [&](size_t d) {
        const float4 p = psi[d];
        if (CHI) {
            if (!mix) return p;
            const float4 m = mu[d];
            return make_float4(cm * m.x + cr * p.x, cm * m.y + cr * p.y, cm * m.z + cr * p.z, cm * m.w + cr * p.w);
        }
        const float4 m = mu[d];
        return make_float4(m.x - r * p.x, m.y - r * p.y, m.z - r * p.z, m.w - r * p.w);
    };
    auto chi1 = [&](size_t d, int k) {  // component k of chi at entry d
        const float *pp = reinterpret_cast<const float *>(psi + d) + k;
        if (CHI) {
            if (!mix) return *pp;
            return cm * reinterpret_cast<const float *>(mu + d)[k] + cr * *pp;
        }
        return reinterpret_cast<const float *>(mu + d)[k] - r * *pp;
    };
    float czc[MM];  // chi_z of the previous slice plane (the ghost plane z' = 0 for the chunk's first slice)
#pragma unroll
    for (int i = 0; i < MM; i++)
        if (i < Mg) czc[i] = chi(dual_idx(a, c, tt.z0, i + 1, y + 1, x + 1)).w;
    for (int zl = tt.z0; zl < tt.z1; zl++) {
        float ctc = chi(dual_idx(a, c, zl + 1, 0, y + 1, x + 1)).z;  // frame i0 - 1 (the ghost frame for i0 = 0)
        float *o = dchi + ((size_t)c * nzp + zl) * S + col;
#pragma unroll
        for (int i = 0; i < MM; i++) {
            if (i < Mg) {
                const size_t d = dual_idx(a, c, zl + 1, i + 1, y + 1, x + 1);
                const float4 c0 = chi(d);
                const float cxm = chi1(d - 1, 0);
                const float cym = chi1(d - (n + 1), 1);
                o[(size_t)i * plane] = a.ks * (cxm - c0.x) + a.ks * (cym - c0.y) + a.kt * (ctc - c0.z) +
                                       a.ks * (czc[i] - c0.w);
                ctc = c0.z;
                czc[i] = c0.w;
            }
        }
    }
}

// U: one inner ADMM step of the TV part fused with the R^* update (admm.cu:85-97):
//   out = fm - 0.5*(rho*div(grad fm) + dchi) - 0.5*(R^* h1)_i
// with (R^* h1)_i = Re sum_j v_j phi_j[t_i] read directly from the frame pairs v.
// Thread per (chunk, y, x) column marching over the chunk's slices and frames.
// Halo slabs (chunk neighbours) come from A, the previous outer iterate
// (rectv.cu:170-171; clamped borders kernels_operators.cuh:4-20).  The
// gradient values needed by the backward differences in z and t are the
// previous slice's / frame's own values, so they are carried.  dchi and v are
// read, and out written, as non-temporal streams: the caches keep fm, whose
// x / y neighbours other lanes and rows read again.
// Frames i0 .. i0 + MM - 1 (one launch per group of MM frames when M > MM; i0 even).
template <int MM>
__global__ void __launch_bounds__(TV_THREADS)
k_update(const float *__restrict__ fm, const float *__restrict__ A, float *__restrict__ out,
         const float *__restrict__ dchi, const float2 *__restrict__ v, TvArgs a, int zv0, int i0) {
    const TvTile tt = tv_tile(a);
    const int x = tt.x, y = tt.y, c = tt.c;
    const int n = a.n, M = a.M, nzp = a.nzp;
    if (x >= n || y >= n) return;
    const bool flgl = c != 0, flgr = c != a.nchunk - 1;
    const unsigned pb = (unsigned)n * n * 4;  // bytes per frame plane (< 2^32 per slice)
    const size_t S = (size_t)M * n * n;
    const size_t zb = (size_t)c * nzp;
    // per-lane byte offsets within a frame plane; slice bases go into buffer
    // descriptors and frame offsets into the uniform offset
    const unsigned oc = 4u * (y * n + x);
    const unsigned oxp = 4u * (y * n + (x < n - 1 ? x + 1 : x)), oxm = 4u * (y * n + (x > 0 ? x - 1 : x));
    const unsigned oyp = 4u * ((y < n - 1 ? y + 1 : y) * n + x), oym = 4u * ((y > 0 ? y - 1 : y) * n + x);
    const float r = *a.rho;
    const float rgs = r * a.gs, rgt = r * a.gt;
    const float keep = in_support(a.supp, x, y) ? 1.f : 0.f;  // projection onto the support
    float fz[MM], fzp[MM], hzc[MM];
    // own column at the segment's first slice z0 and the z-gradient values on its
    // plane (against slice z0 - 1; for z0 = 0 the left halo, or zero at the border)
    {
        const auto rf = buf_rsrc(fm + (zb + tt.z0) * S);
        const auto ra = buf_rsrc(tt.z0 > 0 ? fm + (zb + tt.z0 - 1) * S : flgl ? A + (zb - 1) * S : fm + zb * S);
#pragma unroll
        for (int i = 0; i < MM; i++) {
            if (i0 + i < M) {
                fz[i] = buf_ld(rf, oc, (i0 + i) * pb);
                hzc[i] = rgs * (fz[i] - buf_ld(ra, oc, (i0 + i) * pb));
            }
        }
    }
    for (int zl = tt.z0; zl < tt.z1; zl++) {
        const size_t zg = zb + zl;
        {
            const auto rn = buf_rsrc(zl + 1 < nzp ? fm + (zg + 1) * S : (flgr ? A + (zg + 1) * S : fm + zg * S));
#pragma unroll
            for (int i = 0; i < MM; i++)
                if (i0 + i < M) fzp[i] = buf_ld(rn, oc, (i0 + i) * pb);
        }
        const auto rf = buf_rsrc(fm + zg * S);
        const auto rd = buf_rsrc(dchi + zg * S);
        const auto ro = buf_rsrc(out + zg * S);
        const auto rv = buf_rsrc(v ? (const void *)(v + (size_t)(zg - zv0) * ((M + 1) / 2) * n * n) : (const void *)dchi);
        // t-gradient value on frame i0 - 1: E(t'=1) - E(t'=0) = 0 for i0 = 0; the
        // frame after the group (t-gradient of its last frame)
        float htc = i0 > 0 ? rgt * (fz[0] - buf_ld(rf, oc, (i0 - 1) * pb)) : 0.f;
        const float fnx = i0 + MM < M ? buf_ld(rf, oc, (i0 + MM) * pb) : 0.f;
        float2 vp = make_float2(0.f, 0.f);
#pragma unroll
        for (int i = 0; i < MM; i++) {
            if (i0 + i < M) {
                // R^* part: frame i0 + i of the adjoint pairs (re: even, im: odd; i0 is even)
                if ((i & 1) == 0 && v) vp = buf_ld2_nt(rv, 2 * oc, ((i0 + i) >> 1) * 2 * pb);
                const float rpi = (i & 1) ? vp.y : vp.x;
                const float f0 = fz[i];
                const unsigned so = (i0 + i) * pb;
                const float fxp = buf_ld(rf, oxp, so), fxm = buf_ld(rf, oxm, so);
                const float fyp = buf_ld(rf, oyp, so), fym = buf_ld(rf, oym, so);
                // register indices stay compile-time constants (i is unrolled)
                const float ftp = (i0 + i + 1 < M) ? ((i + 1 < MM) ? fz[(i + 1 < MM) ? i + 1 : i] : fnx) : f0;
                const float hx0 = rgs * (fxp - f0), hxm = rgs * (f0 - fxm);
                const float hy0 = rgs * (fyp - f0), hym = rgs * (f0 - fym);
                const float ht0 = rgt * (ftp - f0);
                const float hz0 = rgs * (fzp[i] - f0);
                const float dv = a.ks * (hxm - hx0) + a.ks * (hym - hy0) + a.kt * (htc - ht0) + a.ks * (hzc[i] - hz0) +
                                 buf_ld_nt(rd, oc, so);
                buf_st_nt((f0 - 0.5f * dv - 0.5f * rpi) * keep, ro, oc, so);
                htc = ht0;
                hzc[i] = hz0;
            }
        }
#pragma unroll
        for (int i = 0; i < MM; i++)
            if (i0 + i < M) fz[i] = fzp[i];
    }
}

// P1: after the inner steps: h2 = grad fm; psi = shrink(h2 + mu/rho, lam0/rho);
// mu += rho (h2 - psi); partial sums r = |h2 - psi|^2, s = |grad(f_old - fm)|^2
// and ||fm - f_old||^2 (admm.cu:101-117, rectv.cu:224-237), plus ||fm||^2 (the
// scale of the relative change used by callers that stop early) and the sum of
// the group norms |h2| (the penalty term of the objective), over all dual
// entries, ghost entries included.  One thread per dual entry, no state carried
// between entries: workgroup = (chunk, z', t') plane and a band of PX_ROWS rows
// y'; wave w takes rows y0 + w, y0 + w + PX_WAVES, ..., lanes take x' = lane + 64 k.
// Each entry reads the five extension values it needs (own, +x, +y, +t, +z) of
// fm and of the previous iterate A straight from memory (the neighbours are
// mostly cache hits: blocks run plane after plane over one band of rows, so the +t
// and +z planes are read again by the next blocks).  Halo slabs come from A
// (rectv.cu:170-171).
// psi itself is needed only through chi = mu' - rho psi (the dual part of the next
// inner steps, k_divchi): with CHI the kernel stores chi in the psi array, so that
// k_divchi reads one dual array instead of two; psi = (mu' - chi) / rho is
// recovered where a caller asks for it (state export).
constexpr int PROX_W = 5;  // partial sums per workgroup: r, s, ||fm - f_old||^2, ||fm||^2, sum |h2|
constexpr int PX_ROWS = 16;  // rows y' per workgroup
#ifndef RTV_PX_WAVES
#define RTV_PX_WAVES 16
#endif
constexpr int PX_WAVES = RTV_PX_WAVES;  // waves per workgroup
template <bool CHI>
__global__ void __launch_bounds__(64 * PX_WAVES)
k_prox(const float *__restrict__ fm, const float *__restrict__ A, float4 *__restrict__ psi,
            float4 *__restrict__ mu, TvArgs a, double *__restrict__ partial) {
    __shared__ double red[PROX_W][PX_WAVES];
    const int n = a.n, M = a.M, nzp = a.nzp;
    const int nb = (n + 1 + PX_ROWS - 1) / PX_ROWS;  // row bands per plane
    const int L = xcd_remap(blockIdx.x, gridDim.x);
    // planes fastest, row bands slowest: the +t and +z neighbour planes of a block's
    // rows are read by the blocks that start right after it, while they are in cache
    const int npl = (M + 1) * (nzp + 1) * (int)(gridDim.x / (nb * (M + 1) * (nzp + 1)));
    const int band = L / npl;
    const int pl = L % npl;                 // plane: (chunk, z', t')
    const int iq = pl % (M + 1);
    const int zq = (pl / (M + 1)) % (nzp + 1);
    const int c = a.chunk0 + pl / ((M + 1) * (nzp + 1));
    const int w = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
    const float r = *a.rho;
    const float rinv = 1.0f / r;
    const float th = a.lam0 / r;
    const bool flgl = c != 0, flgr = c != a.nchunk - 1;
    const size_t plane = (size_t)n * n, S = (size_t)M * plane;
    const size_t zb = (size_t)c * nzp;
    // slab of extension index z' (E index) and whether it is a halo
    auto slab = [&](int zlq, bool &halo) -> size_t {
        int q = zlq - 1;
        const int lo = flgl ? -1 : 0, hi = flgr ? nzp : nzp - 1;
        q = q < lo ? lo : (q > hi ? hi : q);
        halo = (q < 0 || q >= nzp);
        return zb + q;
    };
    bool hc, hn;
    const size_t sc0 = slab(zq, hc) * S, sn = slab(zq + 1, hn) * S;
    const int ic = iq > 0 ? iq - 1 : 0, it = iq < M ? iq : M - 1;
    const size_t fc = (size_t)ic * plane, ft = (size_t)it * plane;
    // the three planes of extension values an entry reads (own, +t, +z) as buffer
    // descriptors: uniform bases, 32-bit lane offsets
    const auto rFc = buf_rsrc(fm + sc0 + fc), rAc = buf_rsrc(A + sc0 + fc);
    const auto rFt = buf_rsrc(fm + sc0 + ft), rAt = buf_rsrc(A + sc0 + ft);
    const auto rFn = buf_rsrc(fm + sn + fc), rAn = buf_rsrc(A + sn + fc);
    float4 *const mup = mu + dual_idx(a, c, zq, iq, 0, 0);
    float4 *const psp = psi + dual_idx(a, c, zq, iq, 0, 0);
    const float th2 = th * th;
    // per-entry terms are formed in FP32 and summed in FP64
    double sr = 0.0, ss = 0.0, sc = 0.0, sf = 0.0, st = 0.0;
    const int y1 = min(n + 1, (band + 1) * PX_ROWS);
    for (int yq = band * PX_ROWS + w; yq < y1; yq += PX_WAVES) {
        const int yc = yq > 0 ? yq - 1 : 0, yn = yq < n ? yq : n - 1;
        for (int xq = lane; xq <= n; xq += 64) {
            const int xc = xq > 0 ? xq - 1 : 0, xn = xq < n ? xq : n - 1;
            const unsigned o0 = 4u * (unsigned)(yc * n + xc), ox = 4u * (unsigned)(yc * n + xn),
                           oy = 4u * (unsigned)(yn * n + xc);
            // extension values (F) and the change against the previous iterate (D)
            auto FD = [&](__amdgpu_buffer_rsrc_t rf, __amdgpu_buffer_rsrc_t ra, unsigned o, bool h, float &F,
                          float &D) {
                const float av = buf_ld(ra, o, 0), fv = buf_ld(rf, o, 0);
                F = h ? av : fv;
                D = h ? 0.f : av - fv;
            };
            float e0, d0, xF, xD, yF, yD, tF, tD, zF, zD;
            FD(rFc, rAc, o0, hc, e0, d0);
            FD(rFc, rAc, ox, hc, xF, xD);
            FD(rFc, rAc, oy, hc, yF, yD);
            FD(rFt, rAt, o0, hc, tF, tD);
            FD(rFn, rAn, o0, hn, zF, zD);
            const float hx = a.gs * (xF - e0), hy = a.gs * (yF - e0);
            const float ht = a.gt * ((iq < M ? tF : e0) - e0), hz = a.gs * (zF - e0);
            const float dx = a.gs * (xD - d0), dy = a.gs * (yD - d0);
            const float dt = a.gt * ((iq < M ? tD : d0) - d0), dz = a.gs * (zD - d0);
            ss += (double)(dx * dx + dy * dy + dt * dt + dz * dz);
            st += (double)__builtin_amdgcn_sqrtf(hx * hx + hy * hy + ht * ht + hz * hz);
            if (zq >= 1 && iq >= 1 && yq >= 1 && xq >= 1) {
                sc += (double)(d0 * d0);
                sf += (double)(e0 * e0);
            }
            const unsigned dl = (unsigned)(yq * (n + 1) + xq);
            // mu, psi: streamed (non-temporal), so the caches keep the extension values of
            // fm and A that neighbouring entries read again
            const v4f mv = __builtin_nontemporal_load((const v4f *)(mup + dl));
            float4 m4 = make_float4(mv[0], mv[1], mv[2], mv[3]);
            float4 p4 = make_float4(hx + m4.x * rinv, hy + m4.y * rinv, ht + m4.z * rinv, hz + m4.w * rinv);
            // shrink: |p| <= th -> 0, else p (1 - th / |p|)
            const float q2 = p4.x * p4.x + p4.y * p4.y + p4.z * p4.z + p4.w * p4.w;
            if (q2 <= th2) {
                p4 = make_float4(0.f, 0.f, 0.f, 0.f);
            } else {
                const float sh = th * __builtin_amdgcn_rsqf(q2);
                p4.x -= sh * p4.x;
                p4.y -= sh * p4.y;
                p4.z -= sh * p4.z;
                p4.w -= sh * p4.w;
            }
            m4.x += r * hx - r * p4.x;
            m4.y += r * hy - r * p4.y;
            m4.z += r * ht - r * p4.z;
            m4.w += r * hz - r * p4.w;
            if (CHI)  // chi = mu' - rho psi in place of psi: k_divchi then reads one dual array
                __builtin_nontemporal_store((v4f){m4.x - r * p4.x, m4.y - r * p4.y, m4.z - r * p4.z, m4.w - r * p4.w},
                                            (v4f *)(psp + dl));
            else
                __builtin_nontemporal_store((v4f){p4.x, p4.y, p4.z, p4.w}, (v4f *)(psp + dl));
            __builtin_nontemporal_store((v4f){m4.x, m4.y, m4.z, m4.w}, (v4f *)(mup + dl));
            const float ex = hx - p4.x, ey = hy - p4.y, et = ht - p4.z, ez = hz - p4.w;
            sr += (double)(ex * ex + ey * ey + et * et + ez * ez);
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        sr += __shfl_down(sr, o, 64);
        ss += __shfl_down(ss, o, 64);
        sc += __shfl_down(sc, o, 64);
        sf += __shfl_down(sf, o, 64);
        st += __shfl_down(st, o, 64);
    }
    if (lane == 0) {
        red[0][w] = sr;
        red[1][w] = ss;
        red[2][w] = sc;
        red[3][w] = sf;
        red[4][w] = st;
    }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int q = 0; q < PROX_W; q++) {
            double acc = 0.0;
            for (int k = 0; k < PX_WAVES; k++) acc += red[q][k];
            partial[PROX_W * (size_t)L + q] = acc;
        }
}

// The penalty rule between outer iterations (rectv.cu:220-232), on the device so
// that an outer iteration needs no host round trip (and can be replayed as one
// graph): with red = {sum |h2 - psi|^2, sum |grad(f_old - fm)|^2, ||fn - f||^2,
// ||fn||^2, sum |h2|, ||R f - g||^2}, r = red[0], s = rho^2 red[1]; rho *= 2 if
// r > 10 s, rho /= 2 if s > 10 r.  Row *iter of hist (capacity cap rows of
// HIST_W, wrapping) receives {r, s, rho used, ||fn - f||^2, ||fn||^2,
// 1/2 ||R f - g||^2 of the entering iterate (NaN when no inner step ran),
// sum |grad fn|}; *iter is then incremented.
constexpr int HIST_W = 7;
// dst[i] += src[i], i < cnt (norms of the out-of-core windows)
__global__ void k_accumulate(double *__restrict__ dst, const double *__restrict__ src, int cnt) {
    const int i = threadIdx.x;
    if (i < cnt) dst[i] += src[i];
}
// rho_chi (may be null): receives the penalty used by the prox that formed the
// stored chi (k_divchi<.., true> rescales chi when the rule changed rho).
__global__ void k_rho_rule(const double *__restrict__ red, float *__restrict__ rho, double *__restrict__ hist,
                           int *__restrict__ iter, int cap, int has_data, float *__restrict__ rho_chi) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const float rh = *rho;
    const double r = red[0], s = (double)rh * rh * red[1];
    const int it = *iter;
    double *h = hist + (size_t)(it % cap) * HIST_W;
    h[0] = r;
    h[1] = s;
    h[2] = rh;
    h[3] = red[2];
    h[4] = red[3];
    h[5] = has_data ? 0.5 * red[5] : __builtin_nan("");
    h[6] = red[4];
    float nr = rh;
    if (r > 10 * s) nr = rh * 2;
    else if (s > 10 * r) nr = rh * 0.5f;
    *rho = nr;
    if (rho_chi) *rho_chi = rh;
    *iter = it + 1;
}

// Deterministic final reduction of `count` groups of `width` doubles (fixed order).
__global__ void __launch_bounds__(1024) k_reduce(const double *__restrict__ partial, double *__restrict__ out,
                                                 int count, int width) {
    __shared__ double red[1024];
    for (int q = 0; q < width; q++) {
        double s = 0.0;
        for (int b = threadIdx.x; b < count; b += blockDim.x) s += partial[(size_t)b * width + q];
        red[threadIdx.x] = s;
        __syncthreads();
        for (int o = blockDim.x / 2; o > 0; o >>= 1) {
            if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[q] = red[0];
        __syncthreads();
    }
}

// Standalone gradient of one chunk (extendf + gradf, kernels_operators.cuh:4-44):
// fext [nzp+2][M][n][n] with halo slabs, h2 AoS float4 [nzp+1][M+1][n+1][n+1].
__global__ void k_grad(const float *__restrict__ fext, float4 *__restrict__ h2, int n, int M, int nzp, int flgl,
                       int flgr, float gs, float gt) {
    const size_t total = (size_t)(nzp + 1) * (M + 1) * (n + 1) * (n + 1);
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= total) return;
    const int xq = p % (n + 1);
    const int yq = (p / (n + 1)) % (n + 1);
    const int iq = (p / ((size_t)(n + 1) * (n + 1))) % (M + 1);
    const int zq = p / ((size_t)(n + 1) * (n + 1) * (M + 1));
    const size_t plane = (size_t)n * n, S = (size_t)M * plane;
    auto cl = [](int q, int lo, int hi) { return q < lo ? lo : (q > hi ? hi : q); };
    auto E = [&](int zz, int ii, int yy, int xx) {
        const int zs = cl(zz - 1, -flgl, nzp - 1 + flgr) + 1;
        return fext[zs * S + cl(ii - 1, 0, M - 1) * plane + (size_t)cl(yy - 1, 0, n - 1) * n + cl(xx - 1, 0, n - 1)];
    };
    const float e0 = E(zq, iq, yq, xq);
    h2[p] = make_float4(gs * (E(zq, iq, yq, xq + 1) - e0), gs * (E(zq, iq, yq + 1, xq) - e0),
                        gt * (E(zq, iq + 1, yq, xq) - e0), gs * (E(zq + 1, iq, yq, xq) - e0));
}

// State import: chi = mu - rho psi in place of psi (the dual array of k_prox<true>).
__global__ void k_form_chi(float4 *__restrict__ psi, const float4 *__restrict__ mu, float rho, size_t cnt) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cnt) return;
    const float4 p = psi[i], m = mu[i];
    psi[i] = make_float4(m.x - rho * p.x, m.y - rho * p.y, m.z - rho * p.z, m.w - rho * p.w);
}

// Standalone divergence (div, kernels_operators.cuh:46-73): f += tau * div(h2).
__global__ void k_div(float *__restrict__ f, const float4 *__restrict__ h2, int n, int M, int nzp, float ks,
                      float kt, float tau) {
    const size_t total = (size_t)nzp * M * n * n;
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= total) return;
    const int x = p % n;
    const int y = (p / n) % n;
    const int i = (p / ((size_t)n * n)) % M;
    const int z = p / ((size_t)n * n * M);
    auto H = [&](int zz, int ii, int yy, int xx) {
        return h2[(((size_t)zz * (M + 1) + ii) * (n + 1) + yy) * (n + 1) + xx];
    };
    const float4 c = H(z + 1, i + 1, y + 1, x + 1);
    float val = f[p];
    val += tau * (H(z + 1, i + 1, y + 1, x).x - c.x) * ks;
    val += tau * (H(z + 1, i + 1, y, x + 1).y - c.y) * ks;
    val += tau * (H(z + 1, i, y + 1, x + 1).z - c.z) * kt;
    val += tau * (H(z, i + 1, y + 1, x + 1).w - c.w) * ks;
    f[p] = val;
}

}  // namespace rtv
