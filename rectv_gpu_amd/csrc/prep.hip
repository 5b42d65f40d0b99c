// Pre-processing of projection data before the solver, on the GPU: dark/flat
// normalisation with -log (examples/tao/tao.py:34-43), the stripe removal of
// examples/tao/tao.py:58-71 (after V. Titarenko), and single-material phase
// retrieval (Paganin et al., J. Microsc. 206 (2002) 33-40) in the form tomopy's
// retrieve_phase applies it (tomocupy's --retrieve-phase-method paganin,
// examples/tao/rec_line.txt:1).  Host arrays in and out; extern "C" entry points
// declared in include/rectv_amd.h.
#include "common.hpp"
#include "fft_lds.hpp"
#include "../../include/rectv_amd.h"

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

namespace rtv {

extern thread_local std::string g_last_error;

namespace {

// A device array owned by one call.
template <typename T>
struct DevArr {
    T *p = nullptr;
    explicit DevArr(size_t count) {
        const hipError_t e = hipMalloc((void **)&p, std::max<size_t>(count, 1) * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            throw Error(e == hipErrorOutOfMemory ? -3 : -2,
                        "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed: " + hipGetErrorString(e));
        }
    }
    ~DevArr() {
        if (p) (void)hipFree(p);
    }
    DevArr(const DevArr &) = delete;
    DevArr &operator=(const DevArr &) = delete;
};

template <typename F>
int guarded_prep(F &&f) {
    try {
        f();
        g_last_error.clear();
        return RTV_OK;
    } catch (const Error &e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return RTV_ERR_HIP;
    }
}

// ---------------------------------------------------------------- kernels
// out[p] = mean over the frames of a[f][p] (FP64 sum)
__global__ void k_frame_mean(const float *__restrict__ a, int nframes, size_t plane, float *__restrict__ out) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= plane) return;
    double s = 0.0;
    for (int f = 0; f < nframes; f++) s += a[(size_t)f * plane + p];
    out[p] = (float)(s / nframes);
}

// out[z][k][x] = -log((data[k][z][x] - dm[z][x]) / (1e-3 + fm[z][x] - dm[z][x])) (or
// the ratio itself when take_log == 0): normalisation and the swap of the
// projection and row axes of tao.py:34-43.  Grid (x blocks, k, z).
__global__ void k_darkflat(const float *__restrict__ data, const float *__restrict__ dm, const float *__restrict__ fm,
                           float *__restrict__ out, int nt, int nz, int n, int take_log) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y, z = blockIdx.z;
    if (x >= n) return;
    const size_t zx = (size_t)z * n + x;
    const float d = dm[zx];
    const float r = (data[((size_t)k * nz + z) * n + x] - d) / (1e-3f + fm[zx] - d);
    out[((size_t)z * nt + k) * n + x] = take_log ? -logf(r) : r;
}

// v[z][x] = mean over projections of g[z][k][x] (FP64)
__global__ void k_sino_mean(const float *__restrict__ g, int nt, int n, double *__restrict__ v) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int z = blockIdx.y;
    if (x >= n) return;
    const float *gz = g + (size_t)z * nt * n + x;
    double s = 0.0;
    for (int k = 0; k < nt; k++) s += gz[(size_t)k * n];
    v[(size_t)z * n + x] = s / nt;
}

// corr[z][x] = mask[x] * sum_j (v[z][j] - v[z][0]) gamma[(x - j) mod n]: the circular
// convolution that tao.py:58-71 forms with rfft / irfft.  gamma staged in LDS.
constexpr int STRIPE_MAX_N = 4096;
__global__ void __launch_bounds__(256) k_stripe_conv(const double *__restrict__ v, const double *__restrict__ gamma,
                                                     int n, int lo, int hi, float *__restrict__ corr) {
    __shared__ double gs[STRIPE_MAX_N];
    __shared__ double vs[STRIPE_MAX_N];
    const int z = blockIdx.y;
    const double *vz = v + (size_t)z * n;
    const double v0 = vz[0];
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        gs[i] = gamma[i];
        vs[i] = vz[i] - v0;
    }
    __syncthreads();
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    double c = 0.0;
    for (int j = 0; j < n; j++) {
        int d = x - j;
        if (d < 0) d += n;
        c += vs[j] * gs[d];
    }
    corr[(size_t)z * n + x] = (x >= lo && x < hi) ? (float)c : 0.f;
}

// g[z][k][x] += corr[z][x]
__global__ void k_stripe_add(float *__restrict__ g, const float *__restrict__ corr, int nt, int n) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y, z = blockIdx.z;
    if (x >= n) return;
    g[((size_t)z * nt + k) * n + x] += corr[(size_t)z * n + x];
}

// Paganin, pass 1: padded row r of projection b (edge-extended: rows and columns
// clamped into the image), forward FFT of length pz_len, into buf[b][r][:].
__global__ void __launch_bounds__(256) k_pag_rows_fwd(const float *__restrict__ proj, int ny, int nz, int py, int pz,
                                                      int PY, FftPlan plan, const float2 *__restrict__ tw,
                                                      float2 *__restrict__ buf) {
    extern __shared__ float2 lds[];
    const int r = blockIdx.x, b = blockIdx.y;
    const int PZ = plan.n;
    const int yr = min(max(r - py, 0), ny - 1);
    const float *src = proj + ((size_t)b * ny + yr) * nz;
    for (int i = threadIdx.x; i < PZ; i += blockDim.x) {
        const int xc = min(max(i - pz, 0), nz - 1);
        lds[lp(i)] = make_float2(src[xc], 0.f);
    }
    __syncthreads();
    fft_lds<16, false, true>(lds, 0, 1, plan, tw, threadIdx.x, blockDim.x);
    float2 *dst = buf + ((size_t)b * PY + r) * PZ;
    for (int i = threadIdx.x; i < PZ; i += blockDim.x) dst[i] = lds[lp(i)];
}

// Pass 2: T adjacent columns of projection b: forward FFT of length PY, times the
// normalised filter H = norm / (c1 (ry[i] + rz[j]) + alpha), inverse FFT.
__global__ void __launch_bounds__(256) k_pag_cols(float2 *__restrict__ buf, int PZ, int T, int stride, FftPlan plan,
                                                  const float2 *__restrict__ tw, const float *__restrict__ ry,
                                                  const float *__restrict__ rz, float c1, float alpha, float norm) {
    extern __shared__ float2 lds[];
    const int PY = plan.n;
    const int c0 = blockIdx.x * T, b = blockIdx.y;
    float2 *base = buf + (size_t)b * PY * PZ + c0;
    for (int idx = threadIdx.x; idx < T * PY; idx += blockDim.x) {
        const int t = idx % T, r = idx / T;
        lds[t * stride + lp(r)] = base[(size_t)r * PZ + t];
    }
    __syncthreads();
    fft_lds<16, false, true>(lds, stride, T, plan, tw, threadIdx.x, blockDim.x);
    for (int idx = threadIdx.x; idx < T * PY; idx += blockDim.x) {
        const int t = idx % T, r = idx / T;
        const float h = norm / (c1 * (ry[r] + rz[c0 + t]) + alpha);
        lds[t * stride + lp(r)] = cscale(lds[t * stride + lp(r)], h);
    }
    __syncthreads();
    fft_lds<16, true, true>(lds, stride, T, plan, tw, threadIdx.x, blockDim.x);
    for (int idx = threadIdx.x; idx < T * PY; idx += blockDim.x) {
        const int t = idx % T, r = idx / T;
        base[(size_t)r * PZ + t] = lds[t * stride + lp(r)];
    }
}

// Pass 3: inverse FFT of padded row py + y, real part of the unpadded columns,
// scaled by 1/(PY PZ), into proj[b][y][:].
__global__ void __launch_bounds__(256) k_pag_rows_inv(const float2 *__restrict__ buf, int ny, int nz, int py, int pz,
                                                      int PY, FftPlan plan, const float2 *__restrict__ tw,
                                                      float *__restrict__ proj) {
    extern __shared__ float2 lds[];
    const int y = blockIdx.x, b = blockIdx.y;
    const int PZ = plan.n;
    const float2 *src = buf + ((size_t)b * PY + py + y) * PZ;
    for (int i = threadIdx.x; i < PZ; i += blockDim.x) lds[lp(i)] = src[i];
    __syncthreads();
    fft_lds<16, true, true>(lds, 0, 1, plan, tw, threadIdx.x, blockDim.x);
    const float sc = 1.0f / ((float)PY * (float)PZ);
    float *dst = proj + ((size_t)b * ny + y) * nz;
    for (int i = threadIdx.x; i < nz; i += blockDim.x) dst[i] = lds[lp(i + pz)].x * sc;
}

// ---------------------------------------------------------------- host helpers
const double PI_TOMOPY = 3.14159265359;  // tomopy/prep/phase.py
const double PLANCK_KEV_S = 6.58211928e-19, SPEED_OF_LIGHT_CM = 299792458e+2;

bool plan_for(int N, FftPlan &p) {
    p.n = N;
    p.nst = 0;
    p.tw_off = 0;
    int r = N;
    const int order[] = {8, 4, 2, 3, 5, 7};
    for (int R : order)
        while (r % R == 0 && p.nst < FFT_MAX_STAGES) {
            p.radix[p.nst++] = R;
            r /= R;
        }
    return r == 1;
}

std::vector<float2> twiddles(int N) {
    std::vector<float2> tw(N);
    for (int m = 0; m < N; m++) {
        const double a = -2.0 * 3.14159265358979323846 * (double)m / N;
        tw[m] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    return tw;
}

void sync_check(hipStream_t st) {
    RTV_HIP(hipGetLastError());
    RTV_HIP(hipStreamSynchronize(st));
}

double wavelength_cm(double energy) { return 2 * PI_TOMOPY * PLANCK_KEV_S * SPEED_OF_LIGHT_CM / energy; }

int pad_width(int dim, double pixel, double dist, double energy) {
    const double pad_pix = std::ceil(PI_TOMOPY * wavelength_cm(energy) * dist / (pixel * pixel));
    return (int)((std::pow(2.0, std::ceil(std::log2(dim + pad_pix))) - dim) * 0.5);
}

// squared reciprocal coordinates of tomopy's _reciprocal_coord, fftshift-ed:
// entry i holds rc[(i - N/2) mod N]^2, rc[m] = (2m - (N-1)) * 2 pi / ((N-1) pixel)
std::vector<double> recip_sq_shifted(int N, double pixel) {
    std::vector<double> out(N);
    const double step = 2 * PI_TOMOPY / ((N - 1) * pixel);
    for (int i = 0; i < N; i++) {
        const int m = ((i - N / 2) % N + N) % N;
        const double rc = (2.0 * m - (N - 1)) * step;
        out[i] = rc * rc;
    }
    return out;
}

}  // namespace
}  // namespace rtv

using namespace rtv;

extern "C" {

int rtv_prep_darkflat(const float *data, const float *dark, const float *flat, int ntheta, int nz, int n, int ndark,
                      int nflat, int take_log, float *out) {
    return guarded_prep([&] {
        RTV_REQUIRE(data && dark && flat && out, "null array");
        RTV_REQUIRE(ntheta > 0 && nz > 0 && n > 0 && ndark > 0 && nflat > 0, "sizes must be positive");
        RTV_REQUIRE(ntheta <= 65535 && nz <= 65535, "ntheta and nz must be at most 65535");
        const size_t plane = (size_t)nz * n, tot = (size_t)ntheta * plane;
        DevArr<float> dd(tot), dk((size_t)ndark * plane), df((size_t)nflat * plane), dm(plane), fm(plane), o(tot);
        RTV_HIP(hipMemcpy(dd.p, data, tot * 4, hipMemcpyHostToDevice));
        RTV_HIP(hipMemcpy(dk.p, dark, (size_t)ndark * plane * 4, hipMemcpyHostToDevice));
        RTV_HIP(hipMemcpy(df.p, flat, (size_t)nflat * plane * 4, hipMemcpyHostToDevice));
        const unsigned gp = (unsigned)((plane + 255) / 256);
        hipLaunchKernelGGL(k_frame_mean, dim3(gp), dim3(256), 0, 0, dk.p, ndark, plane, dm.p);
        hipLaunchKernelGGL(k_frame_mean, dim3(gp), dim3(256), 0, 0, df.p, nflat, plane, fm.p);
        hipLaunchKernelGGL(k_darkflat, dim3((n + 255) / 256, ntheta, nz), dim3(256), 0, 0, dd.p, dm.p, fm.p, o.p, ntheta,
                           nz, n, take_log);
        sync_check(0);
        RTV_HIP(hipMemcpy(out, o.p, tot * 4, hipMemcpyDeviceToHost));
    });
}

int rtv_prep_stripe_ti(float *g, int nz, int ntheta, int n, float beta, int mask_size) {
    return guarded_prep([&] {
        RTV_REQUIRE(g, "null array");
        RTV_REQUIRE(nz > 0 && ntheta > 0 && n > 0 && n % 2 == 0, "sizes must be positive, n even");
        RTV_REQUIRE(n <= STRIPE_MAX_N && nz <= 65535 && ntheta <= 65535, "n at most 4096, nz and ntheta at most 65535");
        RTV_REQUIRE(beta > 0.f && beta < 1.f && mask_size >= 0, "beta in (0, 1), mask_size >= 0");
        // gamma_k = beta ((1 - beta)/(1 + beta))^|k| over the signed frequencies k, gamma_0 - 1
        std::vector<double> gamma(n);
        const double q = (1.0 - beta) / (1.0 + beta);
        for (int i = 0; i < n; i++) {
            const int k = i < (n + 1) / 2 ? i : i - n;
            gamma[i] = (double)beta * std::pow(q, std::abs(k));
        }
        gamma[0] -= 1.0;
        // columns [n//2 - ms//2, n//2 + ms//2) under Python's slice rules, ms = mask_size * n
        const long long ms = (long long)mask_size * n;
        auto norm = [n](long long i) {
            if (i < 0) i += n;
            return (int)std::min<long long>(std::max<long long>(i, 0), n);
        };
        const int lo = norm(n / 2 - ms / 2), hi = norm(n / 2 + ms / 2);
        const size_t tot = (size_t)nz * ntheta * n;
        DevArr<float> dg(tot), corr((size_t)nz * n);
        DevArr<double> v((size_t)nz * n), dgam(n);
        RTV_HIP(hipMemcpy(dg.p, g, tot * 4, hipMemcpyHostToDevice));
        RTV_HIP(hipMemcpy(dgam.p, gamma.data(), n * 8, hipMemcpyHostToDevice));
        const unsigned gx = (unsigned)((n + 255) / 256);
        hipLaunchKernelGGL(k_sino_mean, dim3(gx, nz), dim3(256), 0, 0, dg.p, ntheta, n, v.p);
        hipLaunchKernelGGL(k_stripe_conv, dim3(gx, nz), dim3(256), 0, 0, v.p, dgam.p, n, lo, hi, corr.p);
        hipLaunchKernelGGL(k_stripe_add, dim3(gx, ntheta, nz), dim3(256), 0, 0, dg.p, corr.p, ntheta, n);
        sync_check(0);
        RTV_HIP(hipMemcpy(g, dg.p, tot * 4, hipMemcpyDeviceToHost));
    });
}

int rtv_prep_paganin(float *proj, int ntheta, int ny, int nz, float pixel_size, float dist, float energy, float alpha) {
    return guarded_prep([&] {
        RTV_REQUIRE(proj, "null array");
        RTV_REQUIRE(ntheta > 0 && ny > 0 && nz > 0, "sizes must be positive");
        RTV_REQUIRE(pixel_size > 0.f && dist > 0.f && energy > 0.f && alpha > 0.f,
                    "pixel_size, dist, energy and alpha must be positive");
        const int py = pad_width(ny, pixel_size, dist, energy), pz = pad_width(nz, pixel_size, dist, energy);
        const int PY = ny + 2 * py, PZ = nz + 2 * pz;
        FftPlan ply{}, plz{};
        RTV_REQUIRE(plan_for(PY, ply) && plan_for(PZ, plz),
                    "padded projection sizes must factor into 2, 3, 5, 7 (" + std::to_string(PY) + " x " +
                        std::to_string(PZ) + ")");
        RTV_REQUIRE(fft_tpt(PY, 16) <= 256 && fft_tpt(PZ, 16) <= 256, "padded projection sizes above 4096");
        RTV_REQUIRE(PY <= 65535 && ntheta <= 65535, "too many rows or projections");
        // filter H = norm / (c1 w2 + alpha), c1 = lambda dist / (4 pi), normalised to max 1
        const std::vector<double> ry = recip_sq_shifted(PY, pixel_size), rz = recip_sq_shifted(PZ, pixel_size);
        const double c1 = wavelength_cm(energy) * dist / (4 * PI_TOMOPY);
        const double w2min = *std::min_element(ry.begin(), ry.end()) + *std::min_element(rz.begin(), rz.end());
        const float norm = (float)(c1 * w2min + alpha);
        std::vector<float> fy(ry.begin(), ry.end()), fz(rz.begin(), rz.end());
        std::vector<float2> tw = twiddles(PY), twz = twiddles(PZ);
        DevArr<float> dry(PY), drz(PZ);
        DevArr<float2> dtwy(PY), dtwz(PZ);
        RTV_HIP(hipMemcpy(dry.p, fy.data(), PY * 4, hipMemcpyHostToDevice));
        RTV_HIP(hipMemcpy(drz.p, fz.data(), PZ * 4, hipMemcpyHostToDevice));
        RTV_HIP(hipMemcpy(dtwy.p, tw.data(), PY * 8, hipMemcpyHostToDevice));
        RTV_HIP(hipMemcpy(dtwz.p, twz.data(), PZ * 8, hipMemcpyHostToDevice));
        // projections in batches whose padded spectra fit 2 GiB
        const size_t per = (size_t)PY * PZ * 8;
        const int B = (int)std::max<size_t>(1, std::min<size_t>((size_t)ntheta, ((size_t)2 << 30) / per));
        DevArr<float> dp((size_t)B * ny * nz);
        DevArr<float2> buf((size_t)B * PY * PZ);
        const int sy = lp_stride(PY), szl = lp_stride(PZ);
        int T = 16;
        while (T > 1 && ((size_t)T * sy * 8 > 60 * 1024 || PZ % T || T * fft_tpt(PY, 16) > 256)) T /= 2;
        const int thr_z = std::max(64, ((fft_tpt(PZ, 16) + 63) / 64) * 64);
        const int thr_c = std::max(64, ((T * fft_tpt(PY, 16) + 63) / 64) * 64);
        RTV_REQUIRE(thr_c <= 256, "column transform too long for one workgroup");
        for (int b0 = 0; b0 < ntheta; b0 += B) {
            const int nb = std::min(B, ntheta - b0);
            const size_t cnt = (size_t)nb * ny * nz;
            RTV_HIP(hipMemcpy(dp.p, proj + (size_t)b0 * ny * nz, cnt * 4, hipMemcpyHostToDevice));
            hipLaunchKernelGGL(k_pag_rows_fwd, dim3(PY, nb), dim3(thr_z), (size_t)szl * 8, 0, dp.p, ny, nz, py, pz, PY,
                               plz, dtwz.p, buf.p);
            hipLaunchKernelGGL(k_pag_cols, dim3(PZ / T, nb), dim3(thr_c), (size_t)T * sy * 8, 0, buf.p, PZ, T, sy, ply,
                               dtwy.p, dry.p, drz.p, (float)c1, alpha, norm);
            hipLaunchKernelGGL(k_pag_rows_inv, dim3(ny, nb), dim3(thr_z), (size_t)szl * 8, 0, buf.p, ny, nz, py, pz, PY,
                               plz, dtwz.p, dp.p);
            sync_check(0);
            RTV_HIP(hipMemcpy(proj + (size_t)b0 * ny * nz, dp.p, cnt * 4, hipMemcpyDeviceToHost));
        }
    });
}

}  // extern "C"
