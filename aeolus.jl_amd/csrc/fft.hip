// Direct Poisson solvers on gfx950: hand-written FFTs in LDS, FP64 Thomas.
//
// FFTPoissonSolver (src/PressureSolvers.jl:103-174): r2c along x, complex
// DFT along y, then along z fused with the division by the exact 7-point
// eigenvalues and the inverse z DFT (one LDS round trip), inverse y, c2r x.
//
// ColumnTridiagonalSolver (src/PressureSolvers.jl:544-805), periodic x-y:
// r2c along x and a complex DFT along y per level, the Thomas algorithm per
// horizontal wavenumber in FP64 (one work item per column, coalesced across
// kx), the λ = 0 column integrated as a flux with the gauge mean
// (_mean_mode!), and the inverse transforms with the 1/(Nx Ny) of the
// unnormalised pair. The reference's half-complex R2HC basis and this
// complex basis diagonalise the same operator with the same eigenvalue per
// wavenumber, so the solution is the same discrete one.
//
// Spectrum layout: hat[kx + LD*(ky + Ny*kz)], kx = 0 .. Nx/2, LD = Nx/2+1
// rounded up to a multiple of 16 (aligned column groups for the y/z passes).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "kernels.h"

namespace aeo {

// One transform length: L = 2^a 3^b 5^c, its radices in stage order (DIT:
// first stage first) and the table tw[m] = exp(−2πi m/L), m < L.
constexpr int FFT_MAXS = 16;
struct FFTLen {
    int L, ns;
    int r[FFT_MAXS];
    const double2 *tw;
    const int *perm;  // perm[x] = fft_pos(x), precomputed per length
};

struct SpectralPlan {
    int Nx, Ny, Nz, LD;
    double2 *hat = nullptr;
    double *cp = nullptr;  // Thomas scratch c'_k per column
    std::vector<std::pair<int, double2 *>> tables;  // twiddle tables by length
    std::vector<std::pair<int, int *>> perms;       // digit-reversal tables by length
    // FFT solver: squared modified wavenumbers and 1/(Nx Ny Nz)
    double *kx2 = nullptr, *ky2 = nullptr, *kz2 = nullptr;
    double nrm = 0;
    // column solver
    double *lx = nullptr, *ly = nullptr;  // eigenvalue magnitudes per wavenumber
    double *lower = nullptr, *upper = nullptr, *hweight = nullptr, *measure = nullptr, *icoup = nullptr,
           *gauge = nullptr;
    double measure_sum = 0, gauge_sum = 0, norm = 0;
    int mode = -1;  // 0 fft, 1 column
    int neu_x = 0, neu_y = 0;  // Neumann lateral axes: REDFT10 / REDFT01 (DCT-II / DCT-III)
    int nkx = 0;               // spectral columns along x: Nx/2+1 (periodic) or Nx (Neumann)
};

// the radices of L (4s first, then 2, 3, 5), or ns = 0 when L has another prime factor
static FFTLen fft_factor(int L) {
    FFTLen f{};
    f.L = L;
    int n = L;
    while (n % 4 == 0 && f.ns < FFT_MAXS) { f.r[f.ns++] = 4; n /= 4; }
    while (n % 2 == 0 && f.ns < FFT_MAXS) { f.r[f.ns++] = 2; n /= 2; }
    while (n % 3 == 0 && f.ns < FFT_MAXS) { f.r[f.ns++] = 3; n /= 3; }
    while (n % 5 == 0 && f.ns < FFT_MAXS) { f.r[f.ns++] = 5; n /= 5; }
    if (n != 1 || L < 1) f.ns = 0;
    if (L == 1) { f.ns = 0; }
    return f;
}
static bool fft_ok(int L) { return L == 1 || fft_factor(L).ns > 0; }

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// position of input element x in the digit-reversed order of the DIT stages
// (the output position of element x after the DIF stages): with M_0 = 1,
// M_{i+1} = M_i r_i, the digits of x taken least significant first in the
// radices r_{s−1}, …, r_0 weigh M_{s−1}, …, M_0 (bit reversal when every r = 2)
__device__ __forceinline__ int fft_pos(int x, const FFTLen &f) {
    int p = 0, M = f.L;
    for (int i = f.ns - 1; i >= 0; i--) {
        const int r = f.r[i];
        M /= r;
        const int q = x / r;
        p += (x - q * r) * M;
        x = q;
    }
    return p;
}

// the radix-R DFT of a[0..R) in place; sign −1 forward, +1 inverse
template <int R>
__device__ __forceinline__ void small_dft(double2 *a, double sg);
template <>
__device__ __forceinline__ void small_dft<2>(double2 *a, double) {
    const double2 a0 = a[0], a1 = a[1];
    a[0] = make_double2(a0.x + a1.x, a0.y + a1.y);
    a[1] = make_double2(a0.x - a1.x, a0.y - a1.y);
}
template <>
__device__ __forceinline__ void small_dft<4>(double2 *a, double sg) {
    const double2 s02 = make_double2(a[0].x + a[2].x, a[0].y + a[2].y);
    const double2 d02 = make_double2(a[0].x - a[2].x, a[0].y - a[2].y);
    const double2 s13 = make_double2(a[1].x + a[3].x, a[1].y + a[3].y);
    const double2 d13 = make_double2(a[1].x - a[3].x, a[1].y - a[3].y);
    const double2 j13 = make_double2(-sg * d13.y, sg * d13.x);  // sign·i·d13
    a[0] = make_double2(s02.x + s13.x, s02.y + s13.y);
    a[2] = make_double2(s02.x - s13.x, s02.y - s13.y);
    a[1] = make_double2(d02.x + j13.x, d02.y + j13.y);
    a[3] = make_double2(d02.x - j13.x, d02.y - j13.y);
}
template <>
__device__ __forceinline__ void small_dft<3>(double2 *a, double sg) {
    // y1,2 = a0 − (a1 + a2)/2 ± sign·i·(√3/2)(a1 − a2)
    const double s3 = 0.86602540378443864676;
    const double2 t = make_double2(a[1].x + a[2].x, a[1].y + a[2].y);
    const double2 d = make_double2(a[1].x - a[2].x, a[1].y - a[2].y);
    const double2 m = make_double2(a[0].x - 0.5 * t.x, a[0].y - 0.5 * t.y);
    const double2 j = make_double2(-sg * s3 * d.y, sg * s3 * d.x);
    a[0] = make_double2(a[0].x + t.x, a[0].y + t.y);
    a[1] = make_double2(m.x + j.x, m.y + j.y);
    a[2] = make_double2(m.x - j.x, m.y - j.y);
}
template <>
__device__ __forceinline__ void small_dft<5>(double2 *a, double sg) {
    const double c1 = 0.30901699437494742410, c2 = -0.80901699437494742410;
    const double s1 = 0.95105651629515357212, s2 = 0.58778525229247312917;
    const double2 b1 = make_double2(a[1].x + a[4].x, a[1].y + a[4].y);
    const double2 b2 = make_double2(a[2].x + a[3].x, a[2].y + a[3].y);
    const double2 d1 = make_double2(a[1].x - a[4].x, a[1].y - a[4].y);
    const double2 d2 = make_double2(a[2].x - a[3].x, a[2].y - a[3].y);
    const double2 a0 = a[0];
    const double2 m1 = make_double2(a0.x + c1 * b1.x + c2 * b2.x, a0.y + c1 * b1.y + c2 * b2.y);
    const double2 m2 = make_double2(a0.x + c2 * b1.x + c1 * b2.x, a0.y + c2 * b1.y + c1 * b2.y);
    // e1 = s1 d1 + s2 d2, e2 = s2 d1 − s1 d2; y = m ± sign·i·e
    const double2 e1 = make_double2(s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y);
    const double2 e2 = make_double2(s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y);
    const double2 j1 = make_double2(-sg * e1.y, sg * e1.x), j2 = make_double2(-sg * e2.y, sg * e2.x);
    a[0] = make_double2(a0.x + b1.x + b2.x, a0.y + b1.y + b2.y);
    a[1] = make_double2(m1.x + j1.x, m1.y + j1.y);
    a[4] = make_double2(m1.x - j1.x, m1.y - j1.y);
    a[2] = make_double2(m2.x + j2.x, m2.y + j2.y);
    a[3] = make_double2(m2.x - j2.x, m2.y - j2.y);
}

// One radix-R stage of `nb` interleaved sequences s[m * ld + b] over blocks
// of m = M_i; DIT: twiddle W_len^{qj} before the small DFT, DIF: after it.
// R is a template argument so the butterfly stays in registers.
template <int R, bool DIT>
__device__ __forceinline__ void fft_stage_r(double2 *s, const double2 *__restrict__ tw, int L, int m, int nb, int ld,
                                            int sign) {
    // nb divides blockDim.x: thread t always works on sequence b = t % nb
    const int len = m * R, step = L / len;
    const int nbf = L / R, bstep = blockDim.x / nb;
    const double sg = (double)sign;
    const int b = threadIdx.x % nb;
    for (int bf = threadIdx.x / nb; bf < nbf; bf += bstep) {
        const int g = bf / m, q = bf - g * m;
        const int base = g * len + q;
        double2 a[R];
#pragma unroll
        for (int j = 0; j < R; j++) a[j] = s[(base + j * m) * ld + b];
        if (DIT) {
#pragma unroll
            for (int j = 1; j < R; j++) {
                double2 w = tw[q * j * step];
                w.y *= -sg;  // conj for the inverse
                a[j] = cmul(a[j], w);
            }
        }
        small_dft<R>(a, sg);
        if (!DIT) {
#pragma unroll
            for (int j = 1; j < R; j++) {
                double2 w = tw[q * j * step];
                w.y *= -sg;
                a[j] = cmul(a[j], w);
            }
        }
#pragma unroll
        for (int j = 0; j < R; j++) s[(base + j * m) * ld + b] = a[j];
    }
    __syncthreads();
}
template <bool DIT>
__device__ __forceinline__ void fft_stage(double2 *s, const FFTLen &f, int i, int m, int nb, int ld, int sign) {
    switch (f.r[i]) {
    case 4: fft_stage_r<4, DIT>(s, f.tw, f.L, m, nb, ld, sign); break;
    case 2: fft_stage_r<2, DIT>(s, f.tw, f.L, m, nb, ld, sign); break;
    case 3: fft_stage_r<3, DIT>(s, f.tw, f.L, m, nb, ld, sign); break;
    default: fft_stage_r<5, DIT>(s, f.tw, f.L, m, nb, ld, sign); break;
    }
}

// In-LDS mixed-radix FFT, DIT: input in fft_pos order, output natural
__device__ void lds_fft(double2 *s, const FFTLen &f, int nb, int ld, int sign) {
    int m = 1;
    for (int i = 0; i < f.ns; i++) {
        fft_stage<true>(s, f, i, m, nb, ld, sign);
        m *= f.r[i];
    }
}
// DIF: input natural, output in fft_pos order
__device__ void lds_fft_dif(double2 *s, const FFTLen &f, int nb, int ld, int sign) {
    int m = f.L;
    for (int i = f.ns - 1; i >= 0; i--) {
        m /= f.r[i];
        fft_stage<false>(s, f, i, m, nb, ld, sign);
    }
}

// ---- r2c along x: rows of N = 2L reals -> N/2+1 complex ---------------------
// z[m] = x[2m] + i x[2m+1]; Z = DFT_L(z); X[k] = E[k] + W^k O[k] with
// E = (Z[k] + conj Z[L-k]) / 2, O = (Z[k] - conj Z[L-k]) / 2i, W = e^{-2 pi i/N}.
// the rows of the solver's source buffer
struct SrcRows {
    const double *src;
    __device__ __forceinline__ double2 operator()(int row, int m, int N) const {
        return reinterpret_cast<const double2 *>(src + (size_t)row * N)[m];
    }
};
// the rows of ∇·u* computed where they are read: cells 2m and 2m+1 of row
// (j, k), the expression of div_into_kernel (_div_at, src/FiniteVolume.jl:312-315)
template <typename T> struct DivRows {
    const T *u, *v, *w;
    Metrics<double> md;
    __device__ __forceinline__ double at(size_t c, int K) const {
        const size_t sx = md.sx, sxy = (size_t)md.sx * md.sy;
        const double du = (double)u[c + 1] - (double)u[c];
        const double dv = (double)v[c + sx] - (double)v[c];
        const double dw = (double)w[c + sxy] - (double)w[c];
        return (du * md.idx + dv * md.idy) + dw * md.m_idz(K);
    }
    __device__ __forceinline__ double2 operator()(int row, int m, int) const {
        const int j = row % md.Ny, k = row / md.Ny;
        const int K = k + md.H;
        const size_t c = (size_t)(2 * m + md.H) + (size_t)md.sx * ((size_t)(j + md.H) + (size_t)md.sy * K);
        return make_double2(at(c, K), at(c + 1, K));
    }
};

template <class Rows>
__global__ void r2c_rows_kernel(double2 *__restrict__ hat, Rows rows, FFTLen f, int rpb, int nrows, int LD,
                                const double2 *__restrict__ twN) {
    extern __shared__ double2 sm[];
    const int row0 = blockIdx.x * rpb;
    const int L = f.L, N = 2 * L;
    for (int t = threadIdx.x; t < rpb * L; t += blockDim.x) {
        int r = t / L, m = t - r * L;
        int row = row0 + r;
        double2 z = make_double2(0, 0);
        if (row < nrows) z = rows(row, m, N);
        sm[f.perm[m] * rpb + r] = z;
    }
    __syncthreads();
    lds_fft(sm, f, rpb, rpb, -1);
    for (int t = threadIdx.x; t < rpb * (L + 1); t += blockDim.x) {
        int r = t / (L + 1), k = t - r * (L + 1);
        int row = row0 + r;
        if (row >= nrows) continue;
        double2 a = sm[(k == L ? 0 : k) * rpb + r];
        double2 b = sm[(k == 0 ? 0 : L - k) * rpb + r];
        b.y = -b.y;  // conj Z[L-k]
        double2 e = make_double2(0.5 * (a.x + b.x), 0.5 * (a.y + b.y));
        double2 d = make_double2(a.x - b.x, a.y - b.y);
        double2 o = make_double2(0.5 * d.y, -0.5 * d.x);
        double2 w = twN[k];
        double2 wo = cmul(w, o);
        hat[(size_t)row * LD + k] = make_double2(e.x + wo.x, e.y + wo.y);
    }
}

// ---- c2r along x: N/2+1 Hermitian-half complex -> N reals, times scale ------
__global__ void c2r_rows_kernel(double *__restrict__ dst, const double2 *__restrict__ hat, FFTLen f, int rpb,
                                int nrows, int LD, const double2 *__restrict__ twN, double scale) {
    extern __shared__ double2 sm[];
    const int row0 = blockIdx.x * rpb;
    const int L = f.L, N = 2 * L;
    for (int t = threadIdx.x; t < rpb * L; t += blockDim.x) {
        int r = t / L, k = t - r * L;
        int row = row0 + r;
        double2 z = make_double2(0, 0);
        if (row < nrows) {
            double2 a = hat[(size_t)row * LD + k];
            double2 b = hat[(size_t)row * LD + (L - k)];
            if (k == 0) { a.y = 0; b.y = 0; }  // DC and Nyquist are real
            b.y = -b.y;
            double2 s = make_double2(a.x + b.x, a.y + b.y);
            double2 d = make_double2(a.x - b.x, a.y - b.y);
            double2 w = twN[k];
            w.y = -w.y;  // W^{-k}
            double2 wd = cmul(w, d);
            z = make_double2(s.x - wd.y, s.y + wd.x);  // s + i·wd
        }
        sm[f.perm[k] * rpb + r] = z;
    }
    __syncthreads();
    lds_fft(sm, f, rpb, rpb, +1);
    for (int t = threadIdx.x; t < rpb * L; t += blockDim.x) {
        int r = t / L, m = t - r * L;
        int row = row0 + r;
        if (row >= nrows) continue;
        double2 z = sm[m * rpb + r];
        reinterpret_cast<double2 *>(dst + (size_t)row * N)[m] = make_double2(z.x * scale, z.y * scale);
    }
}

// ---- complex DFT along a strided axis -------------------------------------
// Sequence (g, o): element m at base + m*es + b, b < TB, base = o*os + g*TB.
// mode 0: one transform of `sign`; mode 1: forward, divide by the 7-point
// eigenvalue (FFTPoissonSolver, the zero mode annihilated), inverse.
struct EigArgs {
    const double *kx2, *ky2, *kz2;
    double nrm;
    int nkx;
};

template <int TB>
__global__ void c2c_cols_kernel(double2 *__restrict__ d, FFTLen f, long es, long os, int ncols, int sign, int mode,
                                EigArgs ea) {
    extern __shared__ double2 sm[];
    const int L = f.L;
    const int g = blockIdx.x, o = blockIdx.y;
    const long base = (long)o * os + (long)g * TB;
    for (int t = threadIdx.x; t < L * TB; t += blockDim.x) {
        int m = t / TB, b = t - m * TB;
        double2 z = make_double2(0, 0);
        if (g * TB + b < ncols) z = d[base + (long)m * es + b];
        sm[f.perm[m] * TB + b] = z;
    }
    __syncthreads();
    lds_fft(sm, f, TB, TB, sign);
    if (mode == 1) {
        // here the axis is z (m = kz), o = ky, column g*TB + b = kx
        for (int t = threadIdx.x; t < L * TB; t += blockDim.x) {
            int m = t / TB, b = t - m * TB;
            int kx = g * TB + b;
            double2 z = sm[m * TB + b];
            double l = kx < ea.nkx ? -((ea.kx2[kx] + ea.ky2[o]) + ea.kz2[m]) : 1.0;
            double inv = l == 0 ? 0.0 : ea.nrm / l;
            z.x *= inv;
            z.y *= inv;
            sm[m * TB + b] = z;
        }
        __syncthreads();
        // the inverse by decimation in frequency: natural order in, fft_pos order out
        lds_fft_dif(sm, f, TB, TB, +1);
        for (int t = threadIdx.x; t < L * TB; t += blockDim.x) {
            int m = t / TB, b = t - m * TB;
            if (g * TB + b < ncols) d[base + (long)m * es + b] = sm[f.perm[m] * TB + b];
        }
        return;
    }
    for (int t = threadIdx.x; t < L * TB; t += blockDim.x) {
        int m = t / TB, b = t - m * TB;
        if (g * TB + b < ncols) d[base + (long)m * es + b] = sm[m * TB + b];
    }
}

// ---- the Thomas sweeps (_thomas_slab!, src/PressureSolvers.jl:754-772) ------
// One work item per wavenumber column (kx, ky); the factorisation of the
// reference's constructor (:648-660) is evaluated on the fly, the same
// operations in the same order, with c'_k kept in `cp` for the back
// substitution. The (0, 0) column is the mean mode (_mean_mode!, :780-805).
// wave64 scans in a fixed order (Hillis-Steele over shuffles): the same
// association for every run
__device__ __forceinline__ double wave_incl_scan(double x) {
    const int l = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const double y = __shfl_up(x, o, 64);
        if (l >= o) x += y;
    }
    return x;
}

// The mean-mode column of the column solve with one wave (blockDim = 64, all
// lanes present), in chunks of 64 levels, lane l on level kb + l:
//   mb   = Σ_k h_k m_k / Σ m
//   F_k  = Σ_{j<k} (h_j − mb) m_j,  φ_k = Σ_{1≤j≤k} F_j icoup_j   (φ_0 = 0)
//   φ_k ← φ_k − Σ φ g / Σ g
// The prefix sums are wave scans with the chunk carries added in level order.
// φ goes through cp (this lane's own levels) between the passes.
__device__ void mean_mode_wave(double2 *__restrict__ hat, double *__restrict__ cp, size_t plane, int Nz,
                               const double *__restrict__ measure, const double *__restrict__ icoup,
                               const double *__restrict__ gauge, double measure_sum, double gauge_sum) {
    const int l = threadIdx.x & 63;
    double mb = 0.0;
    for (int kb = 0; kb < Nz; kb += 64) {
        const int k = kb + l;
        const double t = k < Nz ? hat[plane * k].x * measure[k] : 0.0;
        mb += __shfl(wave_incl_scan(t), 63, 64);
    }
    mb /= measure_sum;
    double Fc = 0.0, Pc = 0.0, mphi = 0.0;
    for (int kb = 0; kb < Nz; kb += 64) {
        const int k = kb + l;
        const bool in = k < Nz;
        const double a = in ? (hat[plane * k].x - mb) * measure[k] : 0.0;
        const double ai = wave_incl_scan(a);
        double ex = __shfl_up(ai, 1, 64);
        if (l == 0) ex = 0.0;
        const double F = Fc + ex;  // Σ_{j<k} (h_j − mb) m_j
        const double b = (in && k >= 1) ? F * icoup[k] : 0.0;
        const double ph = Pc + wave_incl_scan(b);
        if (in) cp[plane * k] = ph;
        mphi += __shfl(wave_incl_scan(in ? ph * gauge[k] : 0.0), 63, 64);
        Fc += __shfl(ai, 63, 64);
        Pc = __shfl(ph, 63, 64);
    }
    const double sh = mphi / gauge_sum;
    for (int kb = 0; kb < Nz; kb += 64) {
        const int k = kb + l;
        if (k < Nz) hat[plane * k] = make_double2(cp[plane * k] - sh, 0.0);
    }
}

__global__ void __launch_bounds__(64) thomas_kernel(double2 *__restrict__ hat, double *__restrict__ cp, int nkx, int Ny, int Nz, int LD,
                              const double *__restrict__ lx, const double *__restrict__ ly,
                              const double *__restrict__ lower, const double *__restrict__ upper,
                              const double *__restrict__ hweight, const double *__restrict__ measure,
                              const double *__restrict__ icoup, const double *__restrict__ gauge, double measure_sum,
                              double gauge_sum) {
    const int kx = blockIdx.x * blockDim.x + threadIdx.x;
    const int ky = blockIdx.y;
    const size_t plane = (size_t)LD * Ny;
    const size_t c0 = (size_t)kx + (size_t)LD * ky;
    if (blockIdx.x == 0 && ky == 0) {
        // the mean mode (kx = ky = 0), by the whole wave of this block before
        // its lanes 1..63 solve their own columns: remove the m-weighted mean,
        // integrate the flux up, gauge (_mean_mode!, src/PressureSolvers.jl:780-805)
        mean_mode_wave(hat, cp, plane, Nz, measure, icoup, gauge, measure_sum, gauge_sum);
        if (kx == 0) return;
    }
    if (kx >= nkx) return;
    const double lam = lx[kx] + ly[ky];
    // The sweeps take TB levels at a time and are software-pipelined: the
    // loads of the next batch are issued before the current one is worked
    // (the recurrence itself is sequential; with one column per thread there
    // are few waves, so memory latency is hidden by the loads in flight).
    constexpr int TB = 16;
    double cprev = 0.0;
    double2 dprev = make_double2(0, 0);
    double2 cur[TB], nxt[TB];
#pragma unroll
    for (int q = 0; q < TB; q++)
        if (q < Nz) cur[q] = hat[c0 + plane * q];
    for (int kb = 0; kb < Nz; kb += TB) {
        if (kb + TB < Nz) {
#pragma unroll
            for (int q = 0; q < TB; q++)
                if (kb + TB + q < Nz) nxt[q] = hat[c0 + plane * (kb + TB + q)];
        }
#pragma unroll
        for (int q = 0; q < TB; q++) {
            const int k = kb + q;
            if (k < Nz) {
                const double b = -(lower[k] + upper[k]) - lam * hweight[k];
                const double den = b - lower[k] * cprev;
                const double inv_den = 1 / den;
                cprev = upper[k] / den;
                cp[c0 + plane * k] = cprev;
                if (k == 0) {
                    cur[q].x = cur[q].x * inv_den;
                    cur[q].y = cur[q].y * inv_den;
                } else {
                    const double a = lower[k];
                    cur[q].x = (cur[q].x - a * dprev.x) * inv_den;
                    cur[q].y = (cur[q].y - a * dprev.y) * inv_den;
                }
                dprev = cur[q];
            }
        }
#pragma unroll
        for (int q = 0; q < TB; q++) {
            if (kb + q < Nz) hat[c0 + plane * (kb + q)] = cur[q];
            cur[q] = nxt[q];
        }
    }
    // back substitution, levels Nz−2 … 0 in batches of TB, pipelined likewise
    double2 xn = dprev;
    double cc[TB], cn[TB];
    {
        const int kt = Nz - 2;
#pragma unroll
        for (int q = 0; q < TB; q++)
            if (kt - q >= 0) {
                cur[q] = hat[c0 + plane * (kt - q)];
                cc[q] = cp[c0 + plane * (kt - q)];
            }
    }
    for (int kt = Nz - 2; kt >= 0; kt -= TB) {
        if (kt - TB >= 0) {
#pragma unroll
            for (int q = 0; q < TB; q++)
                if (kt - TB - q >= 0) {
                    nxt[q] = hat[c0 + plane * (kt - TB - q)];
                    cn[q] = cp[c0 + plane * (kt - TB - q)];
                }
        }
#pragma unroll
        for (int q = 0; q < TB; q++)
            if (kt - q >= 0) {
                cur[q].x -= cc[q] * xn.x;
                cur[q].y -= cc[q] * xn.y;
                xn = cur[q];
            }
#pragma unroll
        for (int q = 0; q < TB; q++) {
            if (kt - q >= 0) hat[c0 + plane * (kt - q)] = cur[q];
            cur[q] = nxt[q];
            cc[q] = cn[q];
        }
    }
}

// ---- DCT-II (REDFT10) and DCT-III (REDFT01) through one complex FFT --------
// REDFT10: X_k = 2 Σ_n x_n cos(πk(2n+1)/(2N)); with v_n = x_{2n}, v_{N-1-n} =
// x_{2n+1} and V = DFT_N(v), X_k = 2 Re(e^{-iπk/(2N)} V_k) (Makhoul, IEEE
// TASSP 28 (1980) 27). REDFT01: x_n = X_0 + 2 Σ_{k≥1} X_k cos(πk(2n+1)/(2N));
// with W_k = (X_k − i X_{N−k}) e^{iπk/(2N)} (X_N = 0), v = N·IDFT_N(W) and
// x_{2n} = v_n, x_{2n+1} = v_{N-1-n}. REDFT01 ∘ REDFT10 = 2N.
__device__ __forceinline__ int dct_pos(int m, int N) { return m < N / 2 ? 2 * m : 2 * (N - 1 - m) + 1; }

// rows along x of N reals: forward src → hat (X_k, 0), k < N; inverse Re(hat) → dst · scale
__global__ void dct_rows_kernel(double2 *__restrict__ hat, const double *__restrict__ src, double *__restrict__ dst,
                                FFTLen f, int rpb, int nrows, int LD, int dir, double scale) {
    extern __shared__ double2 sm[];
    const int N = f.L;
    const int row0 = blockIdx.x * rpb;
    for (int t = threadIdx.x; t < rpb * N; t += blockDim.x) {
        const int r = t / N, m = t - r * N, row = row0 + r;
        double2 z = make_double2(0, 0);
        if (row < nrows) {
            if (dir > 0) {
                z.x = src[(size_t)row * N + dct_pos(m, N)];
            } else {
                const double xk = hat[(size_t)row * LD + m].x;
                const double xn = m ? hat[(size_t)row * LD + (N - m)].x : 0.0;
                double sn, cs;
                sincospi((double)m / (2.0 * N), &sn, &cs);
                z = make_double2(xk * cs + xn * sn, xk * sn - xn * cs);
            }
        }
        sm[f.perm[m] * rpb + r] = z;
    }
    __syncthreads();
    lds_fft(sm, f, rpb, rpb, dir > 0 ? -1 : +1);
    for (int t = threadIdx.x; t < rpb * N; t += blockDim.x) {
        const int r = t / N, m = t - r * N, row = row0 + r;
        if (row >= nrows) continue;
        const double2 v = sm[m * rpb + r];
        if (dir > 0) {
            double sn, cs;
            sincospi((double)m / (2.0 * N), &sn, &cs);
            hat[(size_t)row * LD + m] = make_double2(2.0 * (cs * v.x + sn * v.y), 0.0);
        } else {
            dst[(size_t)row * N + dct_pos(m, N)] = v.x * scale;
        }
    }
}

// complex columns along a strided axis (element m at base + m*es + b), in
// place: the transform of the real and the imaginary parts
template <int TB>
__global__ void dct_cols_kernel(double2 *__restrict__ d, FFTLen f, long es, long os, int ncols, int dir) {
    extern __shared__ double2 sm[];
    const int L = f.L;
    const int g = blockIdx.x, o = blockIdx.y;
    const long base = (long)o * os + (long)g * TB;
    for (int t = threadIdx.x; t < L * TB; t += blockDim.x) {
        const int m = t / TB, b = t - m * TB;
        double2 z = make_double2(0, 0);
        if (g * TB + b < ncols) {
            if (dir > 0) {
                z = d[base + (long)dct_pos(m, L) * es + b];
            } else {
                const double2 zk = d[base + (long)m * es + b];
                const double2 zn = m ? d[base + (long)(L - m) * es + b] : make_double2(0, 0);
                const double2 a = make_double2(zk.x + zn.y, zk.y - zn.x);  // Z_k − i Z_{L−k}
                double sn, cs;
                sincospi((double)m / (2.0 * L), &sn, &cs);
                z = make_double2(a.x * cs - a.y * sn, a.x * sn + a.y * cs);
            }
        }
        sm[f.perm[m] * TB + b] = z;
    }
    __syncthreads();
    lds_fft(sm, f, TB, TB, dir > 0 ? -1 : +1);
    for (int t = threadIdx.x; t < L * TB; t += blockDim.x) {
        const int m = t / TB, b = t - m * TB;
        if (g * TB + b >= ncols) continue;
        if (dir > 0) {
            // split V into the transforms of the real and imaginary parts
            const double2 v = sm[m * TB + b];
            const double2 w = sm[(m ? L - m : 0) * TB + b];
            const double2 va = make_double2(0.5 * (v.x + w.x), 0.5 * (v.y - w.y));  // (V + conj W)/2
            const double2 vb = make_double2(0.5 * (v.y + w.y), -0.5 * (v.x - w.x)); // (V − conj W)/(2i)
            double sn, cs;
            sincospi((double)m / (2.0 * L), &sn, &cs);
            d[base + (long)m * es + b] = make_double2(2.0 * (cs * va.x + sn * va.y), 2.0 * (cs * vb.x + sn * vb.y));
        } else {
            d[base + (long)dct_pos(m, L) * es + b] = sm[m * TB + b];
        }
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
#define HCHK(x)                                                                                   \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) {                                                                   \
            snprintf(err, errlen, "%s: %s", #x, hipGetErrorString(e_));                           \
            return nullptr;                                                                       \
        }                                                                                         \
    } while (0)

static hipError_t upload_table(double2 **dst, int n) {
    std::vector<double2> h(n);
    for (int m = 0; m < n; m++) h[m] = make_double2(cos(2.0 * M_PI * m / n), -sin(2.0 * M_PI * m / n));
    hipError_t e = hipMalloc(dst, sizeof(double2) * n);
    if (e != hipSuccess) return e;
    return upload_dev(*dst, h.data(), sizeof(double2) * n);
}

// the table of length L (created at plan set-up, never during a solve)
static const double2 *table(const SpectralPlan *p, int L) {
    for (auto &t : p->tables)
        if (t.first == L) return t.second;
    return nullptr;
}
static const int *perm_table(const SpectralPlan *p, int L) {
    for (auto &t : p->perms)
        if (t.first == L) return t.second;
    return nullptr;
}
// the host form of fft_pos
static int fft_pos_host(int x, const FFTLen &f) {
    int pos = 0, M = f.L;
    for (int i = f.ns - 1; i >= 0; i--) {
        M /= f.r[i];
        pos += (x % f.r[i]) * M;
        x /= f.r[i];
    }
    return pos;
}
static hipError_t ensure_table(SpectralPlan *p, int L) {
    if (L < 1 || table(p, L)) return hipSuccess;
    double2 *d = nullptr;
    hipError_t e = upload_table(&d, L);
    if (e != hipSuccess) return e;
    p->tables.emplace_back(L, d);
    const FFTLen f = fft_factor(L);
    std::vector<int> h(L);
    for (int x = 0; x < L; x++) h[x] = f.ns ? fft_pos_host(x, f) : x;
    int *dp = nullptr;
    if ((e = hipMalloc(&dp, sizeof(int) * L)) != hipSuccess) return e;
    p->perms.emplace_back(L, dp);
    return upload_dev(dp, h.data(), sizeof(int) * L);
}
static FFTLen fftlen(const SpectralPlan *p, int L) {
    FFTLen f = fft_factor(L);
    f.tw = table(p, L);
    f.perm = perm_table(p, L);
    return f;
}

SpectralPlan *spectral_create(int Nx, int Ny, int Nz, hipStream_t st, char *err, size_t errlen) {
    if (Nx % 2 || !fft_ok(Nx / 2) || !fft_ok(Ny) || Nx < 4 || Ny < 2 || Nx > 2048 || Ny > 512) {
        snprintf(err, errlen,
                 "spectral solver: Nx (%d) must be even with Nx/2 a product of 2, 3 and 5, and Ny (%d) a product "
                 "of 2, 3 and 5 (Nx in 4..2048, Ny in 2..512)",
                 Nx, Ny);
        return nullptr;
    }
    SpectralPlan *p = new SpectralPlan();
    p->Nx = Nx; p->Ny = Ny; p->Nz = Nz;
    p->LD = ((Nx / 2 + 1) + 15) / 16 * 16;
    size_t nh = (size_t)p->LD * Ny * Nz;
    HCHK(hipMalloc(&p->hat, sizeof(double2) * nh));
    HCHK(hipMemsetAsync(p->hat, 0, sizeof(double2) * nh, st));
    HCHK(hipMalloc(&p->cp, sizeof(double) * nh));
    const int need[4] = {Nx, Nx / 2, Ny, fft_ok(Nz) ? Nz : 0};
    for (int q = 0; q < 4; q++) HCHK(ensure_table(p, need[q]));
    HCHK(hipStreamSynchronize(st));
    return p;
}

void spectral_destroy(SpectralPlan *p) {
    if (!p) return;
    hipFree(p->hat); hipFree(p->cp);
    for (auto &t : p->tables) hipFree(t.second);
    for (auto &t : p->perms) hipFree(t.second);
    hipFree(p->kx2); hipFree(p->ky2); hipFree(p->kz2);
    hipFree(p->lx); hipFree(p->ly);
    hipFree(p->lower); hipFree(p->upper); hipFree(p->hweight); hipFree(p->measure); hipFree(p->icoup); hipFree(p->gauge);
    delete p;
}

static int upload(double **dst, const std::vector<double> &h) {
    if (hipMalloc(dst, sizeof(double) * h.size()) != hipSuccess) return -1;
    return upload_dev(*dst, h.data(), sizeof(double) * h.size()) == hipSuccess ? 0 : -1;
}

// FFTPoissonSolver constructor, src/PressureSolvers.jl:121-138
int spectral_fft_setup(SpectralPlan *p, double dx, double dy, double dz) {
    if (!fft_ok(p->Nz) || p->Nz < 2 || p->Nz > 512) return -2;
    int Nx = p->Nx, Ny = p->Ny, Nz = p->Nz;
    std::vector<double> kx(Nx / 2 + 1), ky(Ny), kz(Nz);
    for (int m = 0; m <= Nx / 2; m++) { double s = sin(M_PI * m / Nx); kx[m] = 4 / (dx * dx) * (s * s); }
    for (int m = 0; m < Ny; m++) { double s = sin(M_PI * m / Ny); ky[m] = 4 / (dy * dy) * (s * s); }
    for (int m = 0; m < Nz; m++) { double s = sin(M_PI * m / Nz); kz[m] = 4 / (dz * dz) * (s * s); }
    p->nrm = 1 / (double)((long long)Nx * Ny * Nz);
    p->nkx = Nx / 2 + 1;
    if (upload(&p->kx2, kx) || upload(&p->ky2, ky) || upload(&p->kz2, kz)) return -1;
    p->mode = 0;
    return 0;
}

// ColumnTridiagonalSolver constructor (periodic x-y), src/PressureSolvers.jl:626-666
int spectral_column_setup(SpectralPlan *p, double dx, double dy, const double *lower, const double *upper,
                          const double *hweight, const double *measure, const double *inv_coupling,
                          const double *gauge, int neu_x, int neu_y) {
    int Nx = p->Nx, Ny = p->Ny, Nz = p->Nz;
    p->neu_x = neu_x;
    p->neu_y = neu_y;
    p->nkx = neu_x ? Nx : Nx / 2 + 1;
    if (neu_x && p->LD < Nx) {
        // the REDFT10 spectrum along x has Nx real coefficients per row
        hipFree(p->hat);
        hipFree(p->cp);
        p->LD = (Nx + 15) / 16 * 16;
        size_t nh = (size_t)p->LD * Ny * Nz;
        if (hipMalloc(&p->hat, sizeof(double2) * nh) != hipSuccess) return -1;
        if (hipMalloc(&p->cp, sizeof(double) * nh) != hipSuccess) return -1;
        if (zero_dev(p->hat, sizeof(double2) * nh) != hipSuccess) return -1;
    }
    std::vector<double> lx(p->nkx), ly(Ny);
    // _r2r_eigen (src/PressureSolvers.jl:618-624): R2HC index q <-> frequency
    // min(q, N-q), 4/Δ² sin²(πf/N); REDFT10 index q: 4/Δ² sin²(πq/(2N))
    for (int q = 0; q < p->nkx; q++) {
        double s = neu_x ? sin(M_PI * (double)q / (2.0 * Nx)) : sin(M_PI * (double)q / Nx);
        lx[q] = 4 / (dx * dx) * (s * s);
    }
    for (int q = 0; q < Ny; q++) {
        double s;
        if (neu_y) s = sin(M_PI * (double)q / (2.0 * Ny));
        else {
            int f = q < Ny - q ? q : Ny - q;
            s = sin(M_PI * (double)f / Ny);
        }
        ly[q] = 4 / (dy * dy) * (s * s);
    }
    std::vector<double> lo(lower, lower + Nz), up(upper, upper + Nz), hw(hweight, hweight + Nz),
        ms(measure, measure + Nz), ic(inv_coupling, inv_coupling + Nz + 1), ga(gauge, gauge + Nz);
    double msum = 0, gsum = 0;
    for (int k = 0; k < Nz; k++) msum += ms[k];
    for (int k = 0; k < Nz; k++) gsum += ga[k];
    p->measure_sum = msum;
    p->gauge_sum = gsum;
    // norm = 1/(s_x N_x s_y N_y), s = 2 on a REDFT10 axis (:660-664)
    p->norm = 1.0 / (double)((long long)(neu_x ? 2 : 1) * Nx * (neu_y ? 2 : 1) * Ny);
    if (upload(&p->lx, lx) || upload(&p->ly, ly) || upload(&p->lower, lo) || upload(&p->upper, up) ||
        upload(&p->hweight, hw) || upload(&p->measure, ms) || upload(&p->icoup, ic) || upload(&p->gauge, ga))
        return -1;
    p->mode = 1;
    return 0;
}

// rows kernels: tpr threads per row (the butterflies of a radix-4 stage), rpb
// rows per block, bt = rpb·tpr threads (a multiple of rpb: fft_stage_r
// keeps each thread on one row)
static void row_shape(int L, int *tpr, int *rpb, int *bt) {
    *tpr = L / 4 > 0 ? L / 4 : 1;
    *rpb = *tpr >= 256 ? 1 : 256 / *tpr;
    *bt = *rpb * *tpr;
}

template <class Rows> static void xfwd_rows(SpectralPlan *p, hipStream_t st, const Rows &rows) {
    int L = p->Nx / 2;
    int tpr, rpb, bt;
    row_shape(L, &tpr, &rpb, &bt);
    int nrows = p->Ny * p->Nz;
    size_t lds = sizeof(double2) * (size_t)rpb * L;
    hipLaunchKernelGGL(r2c_rows_kernel<Rows>, dim3((nrows + rpb - 1) / rpb), dim3(bt), lds, st, p->hat, rows,
                       fftlen(p, L), rpb, nrows, p->LD, table(p, 2 * L));
}
static void xfwd(SpectralPlan *p, hipStream_t st, const double *src) { xfwd_rows(p, st, SrcRows{src}); }
static void xinv(SpectralPlan *p, hipStream_t st, double *dst, double scale) {
    int L = p->Nx / 2;
    int tpr, rpb, bt;
    row_shape(L, &tpr, &rpb, &bt);
    int nrows = p->Ny * p->Nz;
    size_t lds = sizeof(double2) * (size_t)rpb * L;
    hipLaunchKernelGGL(c2r_rows_kernel, dim3((nrows + rpb - 1) / rpb), dim3(bt), lds, st, dst, p->hat, fftlen(p, L),
                       rpb, nrows, p->LD, table(p, 2 * L), scale);
}
// a complex DFT along y (axis = 1) or z (axis = 2) of every kx column group
static void cols(SpectralPlan *p, hipStream_t st, int axis, int sign, int mode) {
    int L = axis == 1 ? p->Ny : p->Nz;
    const FFTLen f = fftlen(p, L);
    long es = axis == 1 ? (long)p->LD : (long)p->LD * p->Ny;
    long os = axis == 1 ? (long)p->LD * p->Ny : (long)p->LD;
    int nouter = axis == 1 ? p->Nz : p->Ny;
    int ncols = p->mode == 1 ? p->nkx : p->Nx / 2 + 1;
    EigArgs ea{p->kx2, p->ky2, p->kz2, p->nrm, ncols};
    // 8 columns per block (128-byte row segments; 32 KB of LDS at L = 256, so
    // five blocks per CU) over the valid columns only
    constexpr int TB = 8;
    size_t lds = sizeof(double2) * (size_t)L * TB;
    hipLaunchKernelGGL(c2c_cols_kernel<TB>, dim3((ncols + TB - 1) / TB, nouter), dim3(256), lds, st, p->hat, f, es,
                       os, ncols, sign, mode, ea);
}

int spectral_fft_solve(SpectralPlan *p, hipStream_t st, double *phi, const double *src) {
    if (p->mode != 0) return -1;
    xfwd(p, st, src);
    cols(p, st, 1, -1, 0);
    cols(p, st, 2, -1, 1);  // forward z, 1/λ, inverse z
    cols(p, st, 1, +1, 0);
    xinv(p, st, phi, 1.0);
    return 0;
}

// REDFT10 (dir +1) or REDFT01 (dir −1) along x of every row
static void xdct(SpectralPlan *p, hipStream_t st, const double *src, double *dst, int dir, double scale) {
    int N = p->Nx;
    int tpr, rpb, bt;
    row_shape(N, &tpr, &rpb, &bt);
    int nrows = p->Ny * p->Nz;
    size_t lds = sizeof(double2) * (size_t)rpb * N;
    hipLaunchKernelGGL(dct_rows_kernel, dim3((nrows + rpb - 1) / rpb), dim3(bt), lds, st, p->hat, src, dst,
                       fftlen(p, N), rpb, nrows, p->LD, dir, scale);
}
// the y transform of the column solver: complex DFT (periodic) or the DCT pair (Neumann)
static void ytrans(SpectralPlan *p, hipStream_t st, int dir) {
    if (!p->neu_y) {
        cols(p, st, 1, dir > 0 ? -1 : +1, 0);
        return;
    }
    int L = p->Ny;
    const FFTLen f = fftlen(p, L);
    long es = (long)p->LD, os = (long)p->LD * p->Ny;
    constexpr int TB = 8;
    size_t lds = sizeof(double2) * (size_t)L * TB;
    hipLaunchKernelGGL(dct_cols_kernel<TB>, dim3((p->nkx + TB - 1) / TB, p->Nz), dim3(256), lds, st, p->hat, f, es,
                       os, p->nkx, dir);
}

static void column_tail(SpectralPlan *p, hipStream_t st, double *phi);

int spectral_column_solve(SpectralPlan *p, hipStream_t st, double *phi, const double *src) {
    if (p->mode != 1) return -1;
    if (p->neu_x) xdct(p, st, src, nullptr, +1, 1.0);
    else xfwd(p, st, src);
    column_tail(p, st, phi);
    return 0;
}

// the column solve of ∇·u* without the source buffer: the divergence is formed
// in the forward row transform (periodic x only; -1 otherwise)
template <typename T>
int spectral_column_solve_div(SpectralPlan *p, hipStream_t st, double *phi, const T *u, const T *v, const T *w,
                              const Metrics<double> &md) {
    if (p->mode != 1 || p->neu_x) return -1;
    xfwd_rows(p, st, DivRows<T>{u, v, w, md});
    column_tail(p, st, phi);
    return 0;
}
template int spectral_column_solve_div<double>(SpectralPlan *, hipStream_t, double *, const double *, const double *,
                                               const double *, const Metrics<double> &);
template int spectral_column_solve_div<float>(SpectralPlan *, hipStream_t, double *, const float *, const float *,
                                              const float *, const Metrics<double> &);

static void column_tail(SpectralPlan *p, hipStream_t st, double *phi) {
    ytrans(p, st, +1);
    int nkx = p->nkx;
    hipLaunchKernelGGL(thomas_kernel, dim3((nkx + 63) / 64, p->Ny), dim3(64), 0, st, p->hat, p->cp, nkx, p->Ny, p->Nz,
                       p->LD, p->lx, p->ly, p->lower, p->upper, p->hweight, p->measure, p->icoup, p->gauge,
                       p->measure_sum, p->gauge_sum);
    ytrans(p, st, -1);
    if (p->neu_x) xdct(p, st, nullptr, phi, -1, p->norm);
    else xinv(p, st, phi, p->norm);
}

}  // namespace aeo
