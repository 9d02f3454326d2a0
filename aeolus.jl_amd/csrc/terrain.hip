// Terrain-following (Gal-Chen) operators, src/StretchedGrids.jl:608-1343.
//
//   z(x, y, σ) = h + σ (H − h)/H,  J = (H − h)/H,
//   σ_x|z = −(H − σ) h_x/(H − h),  σ_y|z likewise.
//
// The metrics are evaluated on the fly from the terrain plane h and its
// slopes (the centred _slopes of the storage plane) with the reference's
// operations (terrain_following_grid, :777-835): x-face terrain and slope
// (h_i + h_{i−1})/2 and (h_i − h_{i−1})/Δx, the column levels σ_c, σ_f. Only
// three planes and four column vectors are read, never the eight 3-D metric
// arrays. Kernels:
//   tf_div_kernel   _b_tf_divergence! with _tf_Ω (:1180-1205)
//   tf_lap_kernel   the Laplace–Beltrami operator of a cell-centred field,
//                   _b_tf_lap_fx!/_fy!/_fs! then _b_tf_lap_div! (:1285-1327),
//                   each face flux evaluated by both of its cells with the
//                   same arithmetic
//   tf_apply        J·∇² on interior-sized vectors with the boundary maps
//                   (periodic x, y; mirror σ) and zero σ-flux through the
//                   walls: the operator of the terrain Poisson solve (not
//                   symmetric: the metric cross terms are evaluated at
//                   different points, hence flexible CG)
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace aeo {

__device__ __forceinline__ double tf_J(const TFView &t, int i, int j) { return (t.Hl - t.h[i + t.sx * j]) / t.Hl; }
__device__ __forceinline__ double tf_hxf(const TFView &t, int i, int j) {
    const int im = i - 1 < 0 ? 0 : i - 1;
    return (t.h[i + t.sx * j] + t.h[im + t.sx * j]) / 2;
}
__device__ __forceinline__ double tf_hyf(const TFView &t, int i, int j) {
    const int jm = j - 1 < 0 ? 0 : j - 1;
    return (t.h[i + t.sx * j] + t.h[i + t.sx * jm]) / 2;
}
__device__ __forceinline__ double tf_Jx(const TFView &t, int i, int j) { return (t.Hl - tf_hxf(t, i, j)) / t.Hl; }
__device__ __forceinline__ double tf_Jy(const TFView &t, int i, int j) { return (t.Hl - tf_hyf(t, i, j)) / t.Hl; }
// dσdx_xf, dσdy_yf (x-/y-face, centre level k)
__device__ __forceinline__ double tf_sxf(const TFView &t, int i, int j, int k) {
    const int im = i - 1 < 0 ? 0 : i - 1;
    const double hxf = tf_hxf(t, i, j), sl = (t.h[i + t.sx * j] - t.h[im + t.sx * j]) / t.dx;
    return -(t.Hl - t.sigc[k]) * sl / (t.Hl - hxf);
}
__device__ __forceinline__ double tf_syf(const TFView &t, int i, int j, int k) {
    const int jm = j - 1 < 0 ? 0 : j - 1;
    const double hyf = tf_hyf(t, i, j), sl = (t.h[i + t.sx * j] - t.h[i + t.sx * jm]) / t.dy;
    return -(t.Hl - t.sigc[k]) * sl / (t.Hl - hyf);
}
// dσdx_zf, dσdy_zf (σ-face k)
__device__ __forceinline__ double tf_sxz(const TFView &t, int i, int j, int k) {
    return -(t.Hl - t.sigf[k]) * t.hx[i + t.sx * j] / (t.Hl - t.h[i + t.sx * j]);
}
__device__ __forceinline__ double tf_syz(const TFView &t, int i, int j, int k) {
    return -(t.Hl - t.sigf[k]) * t.hy[i + t.sx * j] / (t.Hl - t.h[i + t.sx * j]);
}

// field access: storage layout (F = nullptr map) or interior-sized with the boundary maps
struct FieldRef {
    const double *d;
    int interior;
};
__device__ __forceinline__ int tfmap(int a, int n, int bc) {
    if (a < 0) return bc == BC_PERIODIC ? a + n : -a - 1;
    if (a >= n) return bc == BC_PERIODIC ? a - n : 2 * n - 1 - a;
    return a;
}
// (i, j, k) storage indices
__device__ __forceinline__ double fget(const FieldRef &f, const TFView &t, int i, int j, int k) {
    if (!f.interior) return f.d[(size_t)i + (size_t)t.sx * (j + (size_t)t.sy * k)];
    const int a = tfmap(i - t.H, t.Nx, t.bcx), b = tfmap(j - t.H, t.Ny, t.bcy), c = tfmap(k - t.H, t.Nz, BC_NOFLUX);
    return f.d[(size_t)a + (size_t)t.Nx * (b + (size_t)t.Ny * c)];
}

__device__ __forceinline__ double tf_omega(const TFView &t, const double *u, const double *v, const double *w, int i,
                                           int j, int k) {
    const size_t sx = t.sx, sxy = (size_t)t.sx * t.sy;
    const size_t c = (size_t)i + sx * (j + (size_t)t.sy * k);
    const double wlo = t.dzc[k] / (t.dzc[k - 1] + t.dzc[k]);
    const double whi = t.dzc[k - 1] / (t.dzc[k - 1] + t.dzc[k]);
    const double ucm = (u[c - sxy] + u[c - sxy + 1]) / 2, uck = (u[c] + u[c + 1]) / 2;
    const double vcm = (v[c - sxy] + v[c - sxy + sx]) / 2, vck = (v[c] + v[c + sx]) / 2;
    const double ub = wlo * ucm + whi * uck;
    const double vb = wlo * vcm + whi * vck;
    return w[c] + tf_J(t, i, j) * (tf_sxz(t, i, j, k) * ub + tf_syz(t, i, j, k) * vb);
}

__global__ void tf_div_kernel(double *__restrict__ out, const double *__restrict__ u, const double *__restrict__ v,
                              const double *__restrict__ w, TFView t) {
    const int i0 = blockIdx.x * blockDim.x + threadIdx.x, j0 = blockIdx.y, k0 = blockIdx.z;
    if (i0 >= t.Nx) return;
    const int i = i0 + t.H, j = j0 + t.H, k = k0 + t.H;
    const size_t sx = t.sx;
    const size_t c = (size_t)i + sx * (j + (size_t)t.sy * k);
    const double Wb = tf_omega(t, u, v, w, i, j, k), Wt = tf_omega(t, u, v, w, i, j, k + 1);
    const double hxf = (tf_Jx(t, i + 1, j) * u[c + 1] - tf_Jx(t, i, j) * u[c]) * (1 / t.dx);
    const double hyf = (tf_Jy(t, i, j + 1) * v[c + sx] - tf_Jy(t, i, j) * v[c]) * (1 / t.dy);
    const double vzf = (Wt - Wb) / t.dzc[k];
    out[c] = ((hxf + hyf) + vzf) / tf_J(t, i, j);
}

__device__ __forceinline__ void dzw(double hm, double hp, double &wm, double &wc, double &wp) {
    const double s = hm + hp;
    wm = -hp / (hm * s);
    wc = (hp - hm) / (hm * hp);
    wp = hm / (hp * s);
}
__device__ __forceinline__ double dsig_c(const FieldRef &f, const TFView &t, int i, int j, int k) {
    double wm, wc, wp;
    dzw(t.dzf[k], t.dzf[k + 1], wm, wc, wp);
    return (wm * fget(f, t, i, j, k - 1) + wc * fget(f, t, i, j, k)) + wp * fget(f, t, i, j, k + 1);
}
// the x-face flux at face i (west face of cell i), level k
__device__ __forceinline__ double tf_Fx(const FieldRef &f, const TFView &t, int i, int j, int k) {
    const double dfds = (dsig_c(f, t, i - 1, j, k) + dsig_c(f, t, i, j, k)) / 2;
    return tf_Jx(t, i, j) * ((fget(f, t, i, j, k) - fget(f, t, i - 1, j, k)) / t.dx + tf_sxf(t, i, j, k) * dfds);
}
__device__ __forceinline__ double tf_Fy(const FieldRef &f, const TFView &t, int i, int j, int k) {
    const double dfds = (dsig_c(f, t, i, j - 1, k) + dsig_c(f, t, i, j, k)) / 2;
    return tf_Jy(t, i, j) * ((fget(f, t, i, j, k) - fget(f, t, i, j - 1, k)) / t.dy + tf_syf(t, i, j, k) * dfds);
}
__device__ __forceinline__ double tf_Fs(const FieldRef &f, const TFView &t, int i, int j, int k) {
    const double wlo = t.dzc[k] / (t.dzc[k - 1] + t.dzc[k]);
    const double whi = t.dzc[k - 1] / (t.dzc[k - 1] + t.dzc[k]);
    const double dfdx = wlo * (fget(f, t, i + 1, j, k - 1) - fget(f, t, i - 1, j, k - 1)) / (2 * t.dx) +
                        whi * (fget(f, t, i + 1, j, k) - fget(f, t, i - 1, j, k)) / (2 * t.dx);
    const double dfdy = wlo * (fget(f, t, i, j + 1, k - 1) - fget(f, t, i, j - 1, k - 1)) / (2 * t.dy) +
                        whi * (fget(f, t, i, j + 1, k) - fget(f, t, i, j - 1, k)) / (2 * t.dy);
    const double dfds = (fget(f, t, i, j, k) - fget(f, t, i, j, k - 1)) / t.dzf[k];
    const double sxx = tf_sxz(t, i, j, k), syy = tf_syz(t, i, j, k), Jf = tf_J(t, i, j);
    const double Gss = (sxx * sxx + syy * syy) + 1 / (Jf * Jf);
    return Jf * ((sxx * dfdx + syy * dfdy) + Gss * dfds);
}
// wall0: the σ-flux through the bottom and top walls is zero (the no-flux
// Poisson operator: Σ Δσ_c·J∇²f then telescopes to zero, so the range is
// orthogonal to the constants and a zero-mean source is compatible)
__device__ __forceinline__ double tf_lap_at(const FieldRef &f, const TFView &t, int i, int j, int k, bool wall0) {
    const double ft = (wall0 && k + 1 == t.H + t.Nz) ? 0.0 : tf_Fs(f, t, i, j, k + 1);
    const double fb = (wall0 && k == t.H) ? 0.0 : tf_Fs(f, t, i, j, k);
    return (((tf_Fx(f, t, i + 1, j, k) - tf_Fx(f, t, i, j, k)) / t.dx +
             (tf_Fy(f, t, i, j + 1, k) - tf_Fy(f, t, i, j, k)) / t.dy) +
            (ft - fb) / t.dzc[k]) /
           tf_J(t, i, j);
}

// out (storage) = ∇²f (storage, halo filled)
__global__ void tf_lap_kernel(double *__restrict__ out, const double *__restrict__ f, TFView t) {
    const int i0 = blockIdx.x * blockDim.x + threadIdx.x, j0 = blockIdx.y, k0 = blockIdx.z;
    if (i0 >= t.Nx) return;
    const int i = i0 + t.H, j = j0 + t.H, k = k0 + t.H;
    FieldRef F{f, 0};
    out[(size_t)i + (size_t)t.sx * (j + (size_t)t.sy * k)] = tf_lap_at(F, t, i, j, k, false);
}
// out (interior-sized) = J·∇²x (x interior-sized, boundary maps, no-flux walls)
__global__ void tf_apply_kernel(double *__restrict__ out, const double *__restrict__ x, TFView t) {
    const int i0 = blockIdx.x * blockDim.x + threadIdx.x, j0 = blockIdx.y, k0 = blockIdx.z;
    if (i0 >= t.Nx) return;
    const int i = i0 + t.H, j = j0 + t.H, k = k0 + t.H;
    FieldRef F{x, 1};
    out[(size_t)i0 + (size_t)t.Nx * (j0 + (size_t)t.Ny * k0)] = tf_J(t, i, j) * tf_lap_at(F, t, i, j, k, true);
}

void launch_tf(hipStream_t st, int op, const TFView &t, double *out, const double *a, const double *b,
               const double *c) {
    dim3 grid((t.Nx + 127) / 128, t.Ny, t.Nz), blk(128);
    if (op == 0) hipLaunchKernelGGL(tf_div_kernel, grid, blk, 0, st, out, a, b, c, t);
    else if (op == 1) hipLaunchKernelGGL(tf_lap_kernel, grid, blk, 0, st, out, a, t);
    else hipLaunchKernelGGL(tf_apply_kernel, grid, blk, 0, st, out, a, t);
}

}  // namespace aeo
