// The sound-proof (mass-weighted) core on a regular grid: the fused
// tendency + predictor of _mass_weighted_step! (src/Anelastic.jl:1791-1876)
// and weighted_divergence! (:1082-1110). Weights are the halo-padded
// profiles of a ReferenceState (storage-level index): Wc at cell centres,
// Wf at z faces (entry k: the bottom face of level k), θ̄ at faces for the
// anelastic buoyancy. Operations in the reference's order (centred weighted
// flux forms, _wmom_u/_v/_w :1192-1271, _wscal :1297-1311).
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace aeo {

template <typename T>
__global__ void __launch_bounds__(256) mw_tendency_kernel(T *__restrict__ us, T *__restrict__ vs, T *__restrict__ ws,
                                                          T *__restrict__ tn, const T *__restrict__ u,
                                                          const T *__restrict__ v, const T *__restrict__ w,
                                                          const T *__restrict__ th, const double *__restrict__ Wc,
                                                          const double *__restrict__ Wf,
                                                          const double *__restrict__ thf, Metrics<T> m, MWPhys ph) {
    const int i0 = blockIdx.x * blockDim.x + threadIdx.x, j0 = blockIdx.y, k0 = blockIdx.z;
    if (i0 >= m.Nx) return;
    const int H = m.H;
    const int k = k0 + H;
    const uint32_t sx = m.sx, sxy = (uint32_t)m.sx * (uint32_t)m.sy;
    const uint32_t c = lin(i0 + H, j0 + H, k, m.sx, m.sy);
    const T idx = m.idx, idy = m.idy, idz = m.idz;
    const T h = T(0.5);
#define Uo(o) u[c + (o)]
#define Vo(o) v[c + (o)]
#define Wo(o) w[c + (o)]
#define Fo(o) th[c + (o)]
    const T iwc = T(1) / T(Wc[k]), wfb = T(Wf[k]), wft = T(Wf[k + 1]);
    // _wmom_u
    T tu, tv, tw, tt;
    {
        const T uce = h * (Uo(0) + Uo(1)), ucw = h * (Uo(-1) + Uo(0));
        const T d_uu_dx = (uce * uce - ucw * ucw) * idx;
        const T vun = h * (Vo(sx - 1) + Vo(sx)), uvn = h * (Uo(0) + Uo(sx));
        const T vus = h * (Vo(-1) + Vo(0)), uvs = h * (Uo(-sx) + Uo(0));
        const T d_vu_dy = (vun * uvn - vus * uvs) * idy;
        const T wut = h * (Wo(sxy - 1) + Wo(sxy)), uwt = h * (Uo(0) + Uo(sxy));
        const T wub = h * (Wo(-1) + Wo(0)), uwb = h * (Uo(-sxy) + Uo(0));
        const T d_wu_dz = (wft * wut * uwt - wfb * wub * uwb) * idz * iwc;
        tu = T(0) - ((d_uu_dx + d_vu_dy) + d_wu_dz);
    }
    // _wmom_v
    {
        const T uve = h * (Uo(1 - sx) + Uo(1)), vue = h * (Vo(0) + Vo(1));
        const T uvw = h * (Uo(-sx) + Uo(0)), vuw = h * (Vo(-1) + Vo(0));
        const T d_uv_dx = (uve * vue - uvw * vuw) * idx;
        const T vcn = h * (Vo(0) + Vo(sx)), vcs = h * (Vo(-sx) + Vo(0));
        const T d_vv_dy = (vcn * vcn - vcs * vcs) * idy;
        const T wvt = h * (Wo(sxy - sx) + Wo(sxy)), vwt = h * (Vo(0) + Vo(sxy));
        const T wvb = h * (Wo(-sx) + Wo(0)), vwb = h * (Vo(-sxy) + Vo(0));
        const T d_wv_dz = (wft * wvt * vwt - wfb * wvb * vwb) * idz * iwc;
        tv = T(0) - ((d_uv_dx + d_vv_dy) + d_wv_dz);
    }
    // _wmom_w
    {
        const T iwf = T(1) / T(Wf[k]), wcb = T(Wc[k - 1]), wct = T(Wc[k]);
        const T uwe = h * (Uo(1 - sxy) + Uo(1)), wue = h * (Wo(0) + Wo(1));
        const T uww = h * (Uo(-sxy) + Uo(0)), wuw = h * (Wo(-1) + Wo(0));
        const T d_uw_dx = (uwe * wue - uww * wuw) * idx;
        const T vwn = h * (Vo(sx - sxy) + Vo(sx)), wvn = h * (Wo(0) + Wo(sx));
        const T vws = h * (Vo(-sxy) + Vo(0)), wvs = h * (Wo(-sx) + Wo(0));
        const T d_vw_dy = (vwn * wvn - vws * wvs) * idy;
        const T wtop = h * (Wo(0) + Wo(sxy)), wbot = h * (Wo(-sxy) + Wo(0));
        const T d_ww_dz = (wct * wtop * wtop - wcb * wbot * wbot) * idz * iwf;
        tw = T(0) - ((d_uw_dx + d_vw_dy) + d_ww_dz);
    }
    // _wscal
    {
        const T ufe = Uo(1) * h * (Fo(0) + Fo(1)), ufw = Uo(0) * h * (Fo(-1) + Fo(0));
        const T vfn = Vo(sx) * h * (Fo(0) + Fo(sx)), vfs = Vo(0) * h * (Fo(-sx) + Fo(0));
        const T wftf = wft * Wo(sxy) * h * (Fo(0) + Fo(sxy)), wfbf = wfb * Wo(0) * h * (Fo(-sxy) + Fo(0));
        tt = T(0) - (((ufe - ufw) * idx + (vfn - vfs) * idy) + (wftf - wfbf) * idz * iwc);
    }
    if (ph.nu != 0) {
        // axpy!(t, ν, ∇²f), the uniform 7-point Laplacian (_lap_at, src/FiniteVolume.jl:271-283)
        const T nu = T(ph.nu);
        auto lap = [&](const T *f) {
            const T cc = f[c];
            return (((f[c + 1] - T(2) * cc) + f[c - 1]) * m.idx2 + ((f[c + sx] - T(2) * cc) + f[c - sx]) * m.idy2) +
                   ((f[c + sxy] - T(2) * cc) + f[c - sxy]) * m.idz2;
        };
        tu = tu + nu * lap(u);
        tv = tv + nu * lap(v);
        tw = tw + nu * lap(w);
        tt = tt + nu * lap(th);
    }
    if (ph.buoyancy) {
        // _anelastic_buoyancy (src/Anelastic.jl:1352-1360): g (θ_face − θ̄_f)/θ̄_f
        const T tb = T(thf[k]);
        const T coef = T(ph.g_accel) / tb;
        const T tface = h * (Fo(-sxy) + Fo(0));
        tw += coef * (tface - tb);
    }
    if (ph.fcor != 0) {
        // f-plane Coriolis, _b_coriolis! (src/FiniteVolume.jl:900-911)
        const T f = T(ph.fcor);
        tu += f * (T(0.25) * (((Vo(-1) + Vo(0)) + Vo(sx - 1)) + Vo(sx)));
        tv -= f * (T(0.25) * (((Uo(-sx) + Uo(1 - sx)) + Uo(0)) + Uo(1)));
    }
    if (k == H) tt += T(ph.fb);
    if (k == H + m.Nz - 1) tt -= T(ph.ft);
    if (ph.sponge) {
        // Rayleigh sponge −λ(z)(f − f_target), _b_damping! (src/FiniteVolume.jl:639-647)
        const T *sp = (const T *)ph.sponge;
        const int sz = m.Nz + 2 * H;
        tu -= sp[k] * (Uo(0) - sp[2 * sz + k]);
        tv -= sp[k] * (Vo(0) - sp[3 * sz + k]);
        tw -= sp[sz + k] * (Wo(0) - sp[4 * sz + k]);
    }
#undef Uo
#undef Vo
#undef Wo
#undef Fo
    const T dt = T(ph.dt);
    us[c] = u[c] + dt * tu;
    vs[c] = v[c] + dt * tv;
    ws[c] = w[c] + dt * tw;
    tn[c] = th[c] + dt * tt;
}

// weighted_divergence!(…; normalize) into the FP64 interior-sized buffer or a storage field
template <typename T>
__global__ void wdiv_kernel(double *__restrict__ buf, T *__restrict__ out, const T *__restrict__ u,
                            const T *__restrict__ v, const T *__restrict__ w, const double *__restrict__ Wc,
                            const double *__restrict__ Wf, Metrics<double> m, int normalize) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k0 = blockIdx.z;
    if (i >= m.Nx) return;
    const int k = k0 + m.H;
    const size_t sx = m.sx, sxy = (size_t)m.sx * m.sy;
    const size_t c = lin(i + m.H, j + m.H, k, m.sx, m.sy);
    const double wc = Wc[k], wfb = Wf[k], wft = Wf[k + 1];
    const double scale = normalize ? 1 / wc : 1.0;
    const double d = wc * (((double)u[c + 1] - (double)u[c]) * m.idx + ((double)v[c + sx] - (double)v[c]) * m.idy) +
                     (wft * (double)w[c + sxy] - wfb * (double)w[c]) * m.idz;
    if (buf) buf[(size_t)i + (size_t)m.Nx * (j + (size_t)m.Ny * k0)] = scale * d;
    if (out) out[c] = (T)(scale * d);
}

template <typename T>
void launch_mw_tendency(hipStream_t st, const Metrics<T> &m, T *us, T *vs, T *ws, T *tn, const T *u, const T *v,
                        const T *w, const T *th, const double *Wc, const double *Wf, const double *thf,
                        const MWPhys &ph) {
    dim3 grid((m.Nx + 255) / 256, m.Ny, m.Nz);
    hipLaunchKernelGGL(mw_tendency_kernel<T>, grid, dim3(256), 0, st, us, vs, ws, tn, u, v, w, th, Wc, Wf, thf, m, ph);
}
template <typename T>
void launch_wdiv(hipStream_t st, const Metrics<double> &m, double *buf, T *out, const T *u, const T *v, const T *w,
                 const double *Wc, const double *Wf, int normalize) {
    dim3 grid((m.Nx + 255) / 256, m.Ny, m.Nz);
    hipLaunchKernelGGL(wdiv_kernel<T>, grid, dim3(256), 0, st, buf, out, u, v, w, Wc, Wf, m, normalize);
}
template void launch_mw_tendency<double>(hipStream_t, const Metrics<double> &, double *, double *, double *, double *,
                                         const double *, const double *, const double *, const double *,
                                         const double *, const double *, const double *, const MWPhys &);
template void launch_mw_tendency<float>(hipStream_t, const Metrics<float> &, float *, float *, float *, float *,
                                        const float *, const float *, const float *, const float *, const double *,
                                        const double *, const double *, const MWPhys &);
template void launch_wdiv<double>(hipStream_t, const Metrics<double> &, double *, double *, const double *,
                                  const double *, const double *, const double *, const double *, int);
template void launch_wdiv<float>(hipStream_t, const Metrics<double> &, double *, float *, const float *,
                                 const float *, const float *, const double *, const double *, int);

}  // namespace aeo
