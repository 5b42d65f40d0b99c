// LDS-staged mixed-radix Stockham FFT executed cooperatively by one workgroup.
//
// T transforms of length N live in LDS at buf[t*stride + i].  Each stage reads
// every butterfly's R inputs into registers, synchronises, applies the radix-R
// DFT and writes the autosorted outputs back into the same buffer, so one
// buffer of T*stride complex values is enough.  Radices 2,3,4,5,7,8 (and any
// larger prime through a direct O(R) per output stage); twiddles
// come from a table tw[m] = exp(-2 pi i m / N) computed in FP64 on the host.
// Forward = exp(-2 pi i jk/N), inverse = exp(+2 pi i jk/N), both unnormalised.
// Requirement: blockDim.x >= T * ceil(N / E) (E = max elements held per thread).
#pragma once
#include "common.hpp"

namespace rtv {

constexpr int FFT_MAX_STAGES = 16;
constexpr int FFT_MAX_PRIME = 127;  // largest prime factor of a transform length
struct FftPlan {
    int n;
    int nst;
    int radix[FFT_MAX_STAGES];
    int tw_off;  // offset of this size's twiddle table in the shared table buffer
    int tw_sub;  // lengths R0 * P (R0 = 3, 5; P a power of two): offset of the P-point table
};

// LDS position of element i of a transform row: one pad slot per 16 elements
// breaks the power-of-two strides of the Stockham stores (bank conflicts).
__host__ __device__ __forceinline__ int lp(int i) { return i + (i >> 4); }
// Row pitch (in float2) of a padded transform of length N: = 4 (mod 16), so that
// 4 transforms read by one 16-lane group (column passes) fall on disjoint banks.
__host__ __device__ __forceinline__ int lp_stride(int N) {
    const int s = N + (N >> 4);
    return s + ((4 - s) & 15);
}

template <bool INV>
__device__ inline float2 mul_mi(float2 a) {  // a * (-i) forward, a * (+i) inverse
    return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

template <bool INV>
__device__ inline void dft2(float2 *v) {
    float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <bool INV>
__device__ inline void dft4(float2 *v) {
    float2 s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
    float2 s13 = cadd(v[1], v[3]), d13 = mul_mi<INV>(csub(v[1], v[3]));
    v[0] = cadd(s02, s13);
    v[2] = csub(s02, s13);
    v[1] = cadd(d02, d13);
    v[3] = csub(d02, d13);
}

template <bool INV>
__device__ inline void dft8(float2 *v) {
    // decimation in frequency: a_n = x_n + x_{n+4}, b_n = (x_n - x_{n+4}) W8^n
    const float r2 = 0.70710678118654752440f;
    float2 a[4], b[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        a[i] = cadd(v[i], v[i + 4]);
        b[i] = csub(v[i], v[i + 4]);
    }
    // W8^1 = (1 -/+ i)/sqrt2, W8^2 = -/+ i, W8^3 = (-1 -/+ i)/sqrt2
    {
        float2 t = b[1];
        b[1] = INV ? make_float2(r2 * (t.x - t.y), r2 * (t.x + t.y)) : make_float2(r2 * (t.x + t.y), r2 * (t.y - t.x));
        b[2] = mul_mi<INV>(b[2]);
        t = b[3];
        b[3] = INV ? make_float2(-r2 * (t.x + t.y), r2 * (t.x - t.y)) : make_float2(r2 * (t.y - t.x), -r2 * (t.x + t.y));
    }
    dft4<INV>(a);
    dft4<INV>(b);
#pragma unroll
    for (int m = 0; m < 4; m++) {
        v[2 * m] = a[m];
        v[2 * m + 1] = b[m];
    }
}

// 16-point DFT as 4 x 4: inner DFT-4 over x[n2 + 4 n1], twiddle W16^(n2 k1),
// outer DFT-4 over n2; output X[k1 + 4 k2].
template <bool INV>
__device__ inline void dft16(float2 *v) {
    float2 y[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; n2++) {
        float2 t[4] = {v[n2], v[n2 + 4], v[n2 + 8], v[n2 + 12]};
        dft4<INV>(t);
#pragma unroll
        for (int k1 = 0; k1 < 4; k1++) y[n2][k1] = t[k1];
    }
    // W16^m, m = n2*k1 in {0,1,2,3,4,6,9}
    const float c1 = 0.92387953251128675613f, s1 = 0.38268343236508977173f, r2 = 0.70710678118654752440f;
#pragma unroll
    for (int n2 = 1; n2 < 4; n2++) {
#pragma unroll
        for (int k1 = 1; k1 < 4; k1++) {
            const int m = n2 * k1;
            float cr, ci;  // exp(-2 pi i m / 16)
            if (m == 1) { cr = c1; ci = -s1; }
            else if (m == 2) { cr = r2; ci = -r2; }
            else if (m == 3) { cr = s1; ci = -c1; }
            else if (m == 4) { cr = 0.f; ci = -1.f; }
            else if (m == 6) { cr = -r2; ci = -r2; }
            else { cr = -c1; ci = s1; }  // m == 9: exp(-2 pi i 9/16)
            if (INV) ci = -ci;
            y[n2][k1] = cmul(y[n2][k1], make_float2(cr, ci));
        }
    }
#pragma unroll
    for (int k1 = 0; k1 < 4; k1++) {
        float2 t[4] = {y[0][k1], y[1][k1], y[2][k1], y[3][k1]};
        dft4<INV>(t);
#pragma unroll
        for (int k2 = 0; k2 < 4; k2++) v[k1 + 4 * k2] = t[k2];
    }
}

// Direct odd-radix DFT: y_k = sum_r x_r exp(-/+ 2 pi i r k / R).
// cos/sin(2 pi m / R) for m < R as literals (folded once the loops are unrolled).
template <int R>
__device__ __forceinline__ float odd_cos(int m) {
    if constexpr (R == 3) {
        return m == 0 ? 1.0f : -0.5f;
    } else if constexpr (R == 5) {
        return m == 0 ? 1.0f : (m == 1 || m == 4) ? 0.30901699437494742410f : -0.80901699437494742410f;
    } else {
        return m == 0 ? 1.0f
               : (m == 1 || m == 6) ? 0.62348980185873353053f
               : (m == 2 || m == 5) ? -0.22252093395631440429f
                                    : -0.90096886790241912624f;
    }
}
template <int R>
__device__ __forceinline__ float odd_sin(int m) {
    if constexpr (R == 3) {
        return m == 0 ? 0.0f : (m == 1 ? 0.86602540378443864676f : -0.86602540378443864676f);
    } else if constexpr (R == 5) {
        return m == 0 ? 0.0f
               : m == 1 ? 0.95105651629515357212f
               : m == 2 ? 0.58778525229247312917f
               : m == 3 ? -0.58778525229247312917f
                        : -0.95105651629515357212f;
    } else {
        return m == 0 ? 0.0f
               : m == 1 ? 0.78183148246802980871f
               : m == 2 ? 0.97492791218182360702f
               : m == 3 ? 0.43388373911755812048f
               : m == 4 ? -0.43388373911755812048f
               : m == 5 ? -0.97492791218182360702f
                        : -0.78183148246802980871f;
    }
}

template <int R, bool INV>
__device__ inline void dft_odd(float2 *v) {
    float2 y[R];
#pragma unroll
    for (int k = 0; k < R; k++) {
        float2 acc = v[0];
#pragma unroll
        for (int r = 1; r < R; r++) {
            const int m = (r * k) % R;
            // exp(-i 2 pi m/R) = c - i s (forward); c + i s (inverse)
            const float c = odd_cos<R>(m);
            const float s = INV ? odd_sin<R>(m) : -odd_sin<R>(m);
            acc.x += v[r].x * c - v[r].y * s;
            acc.y += v[r].x * s + v[r].y * c;
        }
        y[k] = acc;
    }
#pragma unroll
    for (int k = 0; k < R; k++) v[k] = y[k];
}

template <int R, bool INV>
__device__ inline void dft(float2 *v) {
    if constexpr (R == 2) dft2<INV>(v);
    else if constexpr (R == 4) dft4<INV>(v);
    else if constexpr (R == 8) dft8<INV>(v);
    else if constexpr (R == 16) dft16<INV>(v);
    else dft_odd<R, INV>(v);
}

// One radix-R Stockham stage.  Threads are split into groups of tpt threads per
// transform (t = tid / tpt); a thread handles butterflies j = jl + b*tpt.
template <int R, int E, bool INV>
__device__ void fft_stage(float2 *buf, int stride, int T, int N, int Ns, int lgNs, const float2 *__restrict__ tw,
                          int t, int jl, int tpt) {
    constexpr int MB = (E + R - 1) / R;
    float2 v[MB][R];
    const int nb = N / R;
    const int twstep = N / (Ns * R);
    const bool active = t < T;
    float2 *row = buf + t * stride;
#pragma unroll
    for (int b = 0; b < MB; b++) {
        const int j = jl + b * tpt;
        if (active && j < nb) {
#pragma unroll
            for (int r = 0; r < R; r++) v[b][r] = row[lp(j + r * nb)];
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < MB; b++) {
        const int j = jl + b * tpt;
        if (active && j < nb) {
            const int jq = lgNs >= 0 ? (j >> lgNs) : j / Ns;
            const int k = j - jq * Ns;
            if (Ns > 1) {
#pragma unroll
                for (int r = 1; r < R; r++) {
                    float2 w = tw[k * r * twstep];
                    if (INV) w.y = -w.y;
                    v[b][r] = cmul(v[b][r], w);
                }
            }
            dft<R, INV>(v[b]);
            const int d0 = jq * Ns * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) row[lp(d0 + r * Ns)] = v[b][r];
        }
    }
    __syncthreads();
}

// A stage of any radix R (the prime factors above 7, which have no butterfly of
// their own): every thread forms its E outputs of the stage directly,
//   y[jq Ns R + k + q Ns] = sum_r x[j + r N/R] W_N^(r (k N/(Ns R) + q N/R)),  j = jq Ns + k,
// O(R) LDS reads and table twiddles per output, then all outputs are written back.
template <int E, bool INV>
__device__ void fft_stage_gen(float2 *buf, int stride, int T, int N, int Ns, int R, const float2 *__restrict__ tw,
                              int t, int jl, int tpt) {
    const int nb = N / R;
    const int twstep = N / (Ns * R);
    const bool active = t < T;
    float2 *row = buf + t * stride;
    float2 y[E];
#pragma unroll
    for (int e = 0; e < E; e++) {
        const int d = jl + e * tpt;
        y[e] = make_float2(0.f, 0.f);
        if (active && d < N) {
            const int k = d % Ns, q = (d / Ns) % R, jq = d / (Ns * R);
            const int j = jq * Ns + k;
            const int step = (k * twstep + q * nb) % N;
            int m = 0;
            float2 acc = make_float2(0.f, 0.f);
            for (int r = 0; r < R; r++) {
                float2 w = tw[m];
                if (INV) w.y = -w.y;
                const float2 x = row[lp(j + r * nb)];
                acc.x += x.x * w.x - x.y * w.y;
                acc.y += x.x * w.y + x.y * w.x;
                m += step;
                if (m >= N) m -= N;
            }
            y[e] = acc;
        }
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; e++) {
        const int d = jl + e * tpt;
        if (active && d < N) row[lp(d)] = y[e];
    }
    __syncthreads();
}

// Threads per transform for length N (each thread holds at most E elements).
__host__ __device__ inline int fft_tpt(int N, int E) { return (N + E - 1) / E; }

// Transform T rows of length p.n held in LDS (element i at buf[t*stride + lp(i)]).  Needs
// blockDim.x >= T * fft_tpt(p.n, E).  Caller must __syncthreads() before.
// ODD selects the instantiation that also carries radices 3, 5, 7 (needed only
// for lengths with those factors; the power-of-two variant keeps register use
// low enough to avoid spills); GEN adds the direct stage for primes above 7
// (its own kernel variant, so that the common lengths do not pay for it).
template <int E, bool INV, bool ODD, bool GEN = false>
__device__ void fft_lds(float2 *buf, int stride, int T, const FftPlan &p, const float2 *__restrict__ twall,
                        int tid, int nthr) {
    (void)nthr;
    const float2 *tw = twall + p.tw_off;
    const int tpt = fft_tpt(p.n, E);
    const int t = tid / tpt, jl = tid - t * tpt;
    int Ns = 1;
    for (int st = 0; st < p.nst; st++) {
        const int R = p.radix[st];
        const int lg = (Ns & (Ns - 1)) == 0 ? __builtin_ctz(Ns) : -1;
        switch (R) {
            case 8: fft_stage<8, E, INV>(buf, stride, T, p.n, Ns, lg, tw, t, jl, tpt); break;
            case 4: fft_stage<4, E, INV>(buf, stride, T, p.n, Ns, lg, tw, t, jl, tpt); break;
            case 2: fft_stage<2, E, INV>(buf, stride, T, p.n, Ns, lg, tw, t, jl, tpt); break;
            default:
                if constexpr (ODD) {
                    if (R == 3) fft_stage<3, E, INV>(buf, stride, T, p.n, Ns, lg, tw, t, jl, tpt);
                    else if (R == 5) fft_stage<5, E, INV>(buf, stride, T, p.n, Ns, lg, tw, t, jl, tpt);
                    else if (R == 7) fft_stage<7, E, INV>(buf, stride, T, p.n, Ns, lg, tw, t, jl, tpt);
                    else if constexpr (GEN) fft_stage_gen<E, INV>(buf, stride, T, p.n, Ns, R, tw, t, jl, tpt);
                }
                break;
        }
        Ns *= R;
    }
}

// ---------------------------------------------------------------------------
// Power-of-two lengths with the length known at compile time: radix-16 stages
// (a final radix-8, -4 or -2), every index computed from constants and shifts.
// Threads per transform TPT = N / E; thread (t, jl) handles butterflies
// j = jl + b * TPT.  Needs blockDim.x >= T * TPT.
template <int R, int N, int Ns, int E, bool INV>
__device__ __forceinline__ void p2_stage(float2 *row, bool active, int jl, const float2 *__restrict__ tw) {
    constexpr int TPT = N / E;
    constexpr int MB = E / R;
    constexpr int nb = N / R;
    constexpr int twstep = N / (Ns * R);
    float2 v[MB][R];
    if (active) {
#pragma unroll
        for (int b = 0; b < MB; b++) {
            const int j = jl + b * TPT;
#pragma unroll
            for (int r = 0; r < R; r++) v[b][r] = row[lp(j + r * nb)];
        }
    }
    __syncthreads();
    if (active) {
#pragma unroll
        for (int b = 0; b < MB; b++) {
            const int j = jl + b * TPT;
            const int k = j & (Ns - 1);
            const int jq = j / Ns;
            if constexpr (Ns > 1) {
#pragma unroll
                for (int r = 1; r < R; r++) {
                    float2 w = tw[k * r * twstep];
                    if (INV) w.y = -w.y;
                    v[b][r] = cmul(v[b][r], w);
                }
            }
            dft<R, INV>(v[b]);
            const int d0 = jq * Ns * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) row[lp(d0 + r * Ns)] = v[b][r];
        }
    }
    __syncthreads();
}

template <int N, int Ns, int E, bool INV>
__device__ __forceinline__ void p2_stages(float2 *row, bool active, int jl, const float2 *__restrict__ tw) {
    if constexpr (Ns < N) {
        constexpr int rem = N / Ns;
        constexpr int R = (rem >= 16 && E >= 16) ? 16 : (rem >= 8 ? 8 : rem);
        p2_stage<R, N, Ns, E, INV>(row, active, jl, tw);
        p2_stages<N, Ns * R, E, INV>(row, active, jl, tw);
    }
}

// Transform T rows of compile-time power-of-two length N in LDS.
template <int N, int E, bool INV>
__device__ __forceinline__ void fft_p2(float2 *buf, int stride, int T, const float2 *__restrict__ tw, int tid) {
    static_assert((N & (N - 1)) == 0 && N >= E, "power-of-two length >= E");
    constexpr int TPT = N / E;
    const int t = tid / TPT, jl = tid & (TPT - 1);
    p2_stages<N, 1, E, INV>(buf + t * stride, t < T, jl, tw);
}

// Register-in / register-out power-of-two transform (compile-time N >= 64).
// The TPT = N/16 threads of a transform each hold 16 elements: thread jl supplies
// x[jl + m*TPT] in v[m] and receives X[jl + m*TPT] in v[m], m = 0..15.  The
// first radix-16 stage (no twiddles) runs on the registers, the radix-16 stages
// between exchange through the transform's LDS row, and the last stage (radix
// N/Ns_last, 16/RL butterflies per thread) leaves its outputs in registers:
// butterfly j = jl + b*TPT writes X[j + r*Ns_last] = X[jl + (b + MB*r)*TPT].
// Every thread of the workgroup must call it (it synchronises); `active`
// masks threads without a transform.  The caller synchronises before the call
// when the row may still be read by a previous transform.
template <int N>
__host__ __device__ constexpr int p2_last_ns() {
    int ns = 16;
    while (N / ns > 16) ns *= 16;
    return ns;
}
template <int N, int Ns, bool INV>
__device__ __forceinline__ void p2_mid(float2 *row, bool active, int jl, const float2 *__restrict__ tw) {
    if constexpr (N / Ns > 16) {
        p2_stage<16, N, Ns, 16, INV>(row, active, jl, tw);
        p2_mid<N, Ns * 16, INV>(row, active, jl, tw);
    }
}
template <int N, bool INV>
__device__ __forceinline__ void fft_p2_reg(float2 (&v)[16], float2 *row, bool active, int jl,
                                           const float2 *__restrict__ tw) {
    static_assert((N & (N - 1)) == 0 && N >= 64, "power-of-two length >= 64");
    constexpr int TPT = N / 16;
    constexpr int NS = p2_last_ns<N>();
    constexpr int RL = N / NS;
    constexpr int MB = 16 / RL;
    if (active) {
        dft16<INV>(v);
        float2 *o = row + lp(16 * jl);  // outputs 16*jl + r: constant offsets r + r/16
#pragma unroll
        for (int r = 0; r < 16; r++) o[r] = v[r];
    }
    __syncthreads();
    p2_mid<N, 16, INV>(row, active, jl, tw);
    if (active) {
        float2 u[MB][RL];
#pragma unroll
        for (int b = 0; b < MB; b++) {
            const int j = jl + b * TPT;
#pragma unroll
            for (int r = 0; r < RL; r++) u[b][r] = row[lp(j + r * NS)];
        }
#pragma unroll
        for (int b = 0; b < MB; b++) {
            const int j = jl + b * TPT;
#pragma unroll
            for (int r = 1; r < RL; r++) {
                float2 w = tw[j * r];
                if (INV) w.y = -w.y;
                u[b][r] = cmul(u[b][r], w);
            }
            dft<RL, INV>(u[b]);
        }
#pragma unroll
        for (int b = 0; b < MB; b++)
#pragma unroll
            for (int r = 0; r < RL; r++) v[b + MB * r] = u[b][r];
    }
}

// ---------------------------------------------------------------------------
// Compile-time lengths N = R0 * P (R0 = 3 or 5, P a power of two >= 64): R0
// sub-groups of TPS = P/16 threads (N/16 threads per transform, as for a power of
// two).  Thread (g, jl) supplies x[R0*(jl + m*TPS) + g] in v[m] and receives
// X[jl + (16 g + m)*TPS] in v[m], m = 0..15:
//   Y_g = P-point DFT of x[R0 j + g] (fft_p2_reg on the sub-row g),
//   X[k + P q] = sum_g W_N^(g k) w_R0^(g q) Y_g[k]  (k = jl + m*TPS, q = g),
// the twiddled Y_g exchanged through the transform's LDS row (R0 sub-rows of
// pitch lp_stride(P); the exchange uses positions lp(g*P + k) of the same row).
__host__ __device__ constexpr int rp_radix(int N) {
    return (N % 3 == 0 && ((N / 3) & (N / 3 - 1)) == 0 && N / 3 >= 64) ? 3
           : (N % 5 == 0 && ((N / 5) & (N / 5 - 1)) == 0 && N / 5 >= 64) ? 5 : 0;
}
// LDS row pitch (float2) of one transform of length N on either path
__host__ __device__ inline int fft_row_pitch(int N) {
    const int r0 = rp_radix(N);
    const int a = lp_stride(N);
    const int b = r0 ? r0 * lp_stride(N / r0) : 0;
    return a > b ? a : b;
}
template <int R0, int P, bool INV>
__device__ __forceinline__ void fft_rp_reg(float2 (&v)[16], float2 *row, bool active, int g, int jl,
                                           const float2 *__restrict__ twP, const float2 *__restrict__ twN) {
    constexpr int TPS = P / 16;
    const int sub = lp_stride(P);
    fft_p2_reg<P, INV>(v, row + g * sub, active, jl, twP);
    __syncthreads();  // the sub-rows become the exchange row
    if (active) {
#pragma unroll
        for (int m = 0; m < 16; m++) {
            const int k = jl + m * TPS;
            float2 w = twN[g * k];
            if (INV) w.y = -w.y;
            row[lp(g * P + k)] = cmul(v[m], w);
        }
    }
    __syncthreads();
    if (active) {
        // w_R0^(g' g): the row g of the R0-point DFT matrix
        float wc[R0], ws[R0];
#pragma unroll
        for (int gg = 0; gg < R0; gg++) {
            const int e = (gg * g) % R0;
            wc[gg] = odd_cos<R0>(e);
            ws[gg] = INV ? odd_sin<R0>(e) : -odd_sin<R0>(e);
        }
#pragma unroll
        for (int m = 0; m < 16; m++) {
            const int k = jl + m * TPS;
            float2 acc = make_float2(0.f, 0.f);
#pragma unroll
            for (int gg = 0; gg < R0; gg++) {
                const float2 y = row[lp(gg * P + k)];
                acc.x += y.x * wc[gg] - y.y * ws[gg];
                acc.y += y.x * ws[gg] + y.y * wc[gg];
            }
            v[m] = acc;
        }
    }
}

// NP > 0: compile-time power-of-two length NP; NP == 0: runtime plan (radices
// 2, 3, 4, 5, 7, 8).
template <int NP, int E, bool INV>
__device__ __forceinline__ void fft_any(float2 *buf, int stride, int T, const FftPlan &p, const float2 *__restrict__ twall,
                                        int tid, int nthr) {
    if constexpr (NP > 0) fft_p2<NP, E, INV>(buf, stride, T, twall + p.tw_off, tid);
    else fft_lds<E, INV, true>(buf, stride, T, p, twall, tid, nthr);
}

}  // namespace rtv
