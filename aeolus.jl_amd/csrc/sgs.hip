// The static Smagorinsky closure of the Boussinesq step (StaticSmagorinsky,
// src/operators/sgs_closures.jl:403-425, with sgs_stress = :laplacian):
//
//   νt = (C_s Δ_k)² |S|                       _b_smagorinsky! (src/FiniteVolume.jl:930-944)
//   |S| = √(2 S_ij S_ij) from the cell-centre velocity gradient
//                                             _velocity_gradient_at / _strain_at
//                                             (src/Operators.jl:448-496)
//   t_u += ∂_j(νt ∂_j u), likewise v, w       _b_eddy_u!/_v!/_w! (src/FiniteVolume.jl:966-1042)
//   t_θ += (1/Pr_t) ∂_j(νt ∂_j θ)             _b_eddy_scalar! (src/FiniteVolume.jl:1280-1300)
//
// with νt averaged to the flux location (2-point means on cell faces,
// 4-point means on cell edges). The per-level filter widths and the
// off-diagonal vertical metric come precomputed: sgs.coef[k] = (C_s Δ_k)²,
// sgs.idz_off[k] (storage level k).
//
// The eddy terms are added to the predictor outputs after the fused
// tendency kernel: u* += Δt·t_u (and θ' += b·Δt·t_θ with the stage weight b
// of the θ combination), which is the reference's u + Δt(t + t_sgs) up to
// the order of two additions.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace aeo {

template <typename T>
__global__ void smag_nut_kernel(T *__restrict__ nut, const T *__restrict__ u, const T *__restrict__ v,
                                const T *__restrict__ w, Metrics<T> m, const T *__restrict__ coef,
                                const T *__restrict__ idz_off) {
    const int i0 = blockIdx.x * blockDim.x + threadIdx.x, j0 = blockIdx.y, k0 = blockIdx.z;
    if (i0 >= m.Nx) return;
    const int H = m.H, k = k0 + H;
    const size_t sx = m.sx, sxy = (size_t)m.sx * m.sy;
    const size_t c = lin(i0 + H, j0 + H, k, m.sx, m.sy);
    const T idx = m.idx, idy = m.idy, idz = m.m_idz(k), ioff = idz_off[k];
    const T q = T(0.25);
    // _velocity_gradient_at
    const T du_dx = (u[c + 1] - u[c]) * idx;
    const T dv_dy = (v[c + sx] - v[c]) * idy;
    const T dw_dz = (w[c + sxy] - w[c]) * idz;
    const T du_dy = ((u[c + sx] + u[c + sx + 1]) - (u[c - sx] + u[c - sx + 1])) * (q * idy);
    const T dv_dx = ((v[c + 1] + v[c + 1 + sx]) - (v[c - 1] + v[c - 1 + sx])) * (q * idx);
    const T du_dz = ((u[c + sxy] + u[c + sxy + 1]) - (u[c - sxy] + u[c - sxy + 1])) * (q * ioff);
    const T dw_dx = ((w[c + 1] + w[c + 1 + sxy]) - (w[c - 1] + w[c - 1 + sxy])) * (q * idx);
    const T dv_dz = ((v[c + sxy] + v[c + sxy + sx]) - (v[c - sxy] + v[c - sxy + sx])) * (q * ioff);
    const T dw_dy = ((w[c + sx] + w[c + sx + sxy]) - (w[c - sx] + w[c - sx + sxy])) * (q * idy);
    // _strain_at
    const T h = T(0.5);
    const T S12 = h * (du_dy + dv_dx), S13 = h * (du_dz + dw_dx), S23 = h * (dv_dz + dw_dy);
    const T S2 = T(2) * ((du_dx * du_dx + dv_dy * dv_dy) + dw_dz * dw_dz) +
                 T(4) * ((S12 * S12 + S13 * S13) + S23 * S23);
    nut[c] = coef[k] * sqrt(S2);
}

// the eddy-viscosity and eddy-diffusion tendencies, added to the predictor
// outputs: us += dt·t_u, vs += dt·t_v, ws += dt·t_w, so += bt·dt·s·(…)_θ
template <typename T>
__global__ void smag_apply_kernel(T *__restrict__ us, T *__restrict__ vs, T *__restrict__ ws, T *__restrict__ so,
                                  const T *__restrict__ u, const T *__restrict__ v, const T *__restrict__ w,
                                  const T *__restrict__ th, const T *__restrict__ n, Metrics<T> m, T dt, T bt,
                                  T sc) {
    const int i0 = blockIdx.x * blockDim.x + threadIdx.x, j0 = blockIdx.y, k0 = blockIdx.z;
    if (i0 >= m.Nx) return;
    const int H = m.H, k = k0 + H;
    const size_t sx = m.sx, sxy = (size_t)m.sx * m.sy;
    const size_t c = lin(i0 + H, j0 + H, k, m.sx, m.sy);
    const T idx = m.idx, idy = m.idy;
    const T idzk = m.m_idz(k), idzf_k = m.m_idzf(k), idzf_kp = m.m_idzf(k + 1);
    const T q = T(0.25);
#define N(o) n[c + (o)]
    {  // _b_eddy_u!
        const T fxE = N(0) * (u[c + 1] - u[c]) * idx;
        const T fxW = N(-1) * (u[c] - u[c - 1]) * idx;
        const T eN = q * (((N(-1) + N(0)) + N(sx - 1)) + N(sx));
        const T eS = q * (((N(-1 - sx) + N(-sx)) + N(-1)) + N(0));
        const T fyN = eN * (u[c + sx] - u[c]) * idy;
        const T fyS = eS * (u[c] - u[c - sx]) * idy;
        const T eT = q * (((N(-1) + N(0)) + N(sxy - 1)) + N(sxy));
        const T eB = q * (((N(-1 - sxy) + N(-sxy)) + N(-1)) + N(0));
        const T fzT = eT * (u[c + sxy] - u[c]) * idzf_kp;
        const T fzB = eB * (u[c] - u[c - sxy]) * idzf_k;
        const T t = ((fxE - fxW) * idx + (fyN - fyS) * idy) + (fzT - fzB) * idzk;
        us[c] = us[c] + dt * t;
    }
    {  // _b_eddy_v!
        const T fyN = N(0) * (v[c + sx] - v[c]) * idy;
        const T fyS = N(-sx) * (v[c] - v[c - sx]) * idy;
        const T eE = q * (((N(-sx) + N(1 - sx)) + N(0)) + N(1));
        const T eW = q * (((N(-1 - sx) + N(-sx)) + N(-1)) + N(0));
        const T fxE = eE * (v[c + 1] - v[c]) * idx;
        const T fxW = eW * (v[c] - v[c - 1]) * idx;
        const T eT = q * (((N(-sx) + N(0)) + N(sxy - sx)) + N(sxy));
        const T eB = q * (((N(-sx - sxy) + N(-sxy)) + N(-sx)) + N(0));
        const T fzT = eT * (v[c + sxy] - v[c]) * idzf_kp;
        const T fzB = eB * (v[c] - v[c - sxy]) * idzf_k;
        const T t = ((fxE - fxW) * idx + (fyN - fyS) * idy) + (fzT - fzB) * idzk;
        vs[c] = vs[c] + dt * t;
    }
    {  // _b_eddy_w!
        const T fzT = N(0) * (w[c + sxy] - w[c]) * idzk;
        const T fzB = N(-sxy) * (w[c] - w[c - sxy]) * m.m_idz(k - 1);
        const T eE = q * (((N(-sxy) + N(1 - sxy)) + N(0)) + N(1));
        const T eW = q * (((N(-1 - sxy) + N(-sxy)) + N(-1)) + N(0));
        const T fxE = eE * (w[c + 1] - w[c]) * idx;
        const T fxW = eW * (w[c] - w[c - 1]) * idx;
        const T eN = q * (((N(-sxy) + N(sx - sxy)) + N(0)) + N(sx));
        const T eS = q * (((N(-sx - sxy) + N(-sxy)) + N(-sx)) + N(0));
        const T fyN = eN * (w[c + sx] - w[c]) * idy;
        const T fyS = eS * (w[c] - w[c - sx]) * idy;
        const T t = ((fxE - fxW) * idx + (fyN - fyS) * idy) + (fzT - fzB) * idzf_k;
        ws[c] = ws[c] + dt * t;
    }
    {  // _b_eddy_scalar! with s = 1/Pr_t
        const T h = T(0.5);
        const T nE = h * (N(0) + N(1)), nW = h * (N(-1) + N(0));
        const T nN = h * (N(0) + N(sx)), nS = h * (N(-sx) + N(0));
        const T nT = h * (N(0) + N(sxy)), nB = h * (N(-sxy) + N(0));
        const T fxE = nE * (th[c + 1] - th[c]) * idx, fxW = nW * (th[c] - th[c - 1]) * idx;
        const T fyN = nN * (th[c + sx] - th[c]) * idy, fyS = nS * (th[c] - th[c - sx]) * idy;
        const T fzT = nT * (th[c + sxy] - th[c]) * idzf_kp, fzB = nB * (th[c] - th[c - sxy]) * idzf_k;
        const T t = sc * (((fxE - fxW) * idx + (fyN - fyS) * idy) + (fzT - fzB) * idzk);
        so[c] = so[c] + bt * (dt * t);
    }
#undef N
}

// the sound-proof cores' form (flux_weights = constraint_weights):
// the vertical parts (1/W) ∂z(W νt ∂z u_i), W at the z faces for u, v
// (wf[k+1]·F_T − wf[k]·F_B)·(1/Δz)/wc[k], at the centres for w
// (wc[k]·F_T − wc[k−1]·F_B)·(1/Δz_f)/wf[k] (_vertical_divergence[_w],
// src/FiniteVolume.jl:1556-1563); the scalar with κt = νt/Pr_t pointwise
// (_b_eddy_scalar_masked! without a mask, :1473-1492). Predictor outputs
// updated as us += dt·t_u, …, tn += dt·t_θ.
template <typename T>
__global__ void smag_apply_w_kernel(T *__restrict__ us, T *__restrict__ vs, T *__restrict__ ws, T *__restrict__ tn,
                                    const T *__restrict__ u, const T *__restrict__ v, const T *__restrict__ w,
                                    const T *__restrict__ th, const T *__restrict__ n, const T *__restrict__ kap,
                                    const double *__restrict__ Wc, const double *__restrict__ Wf, Metrics<T> m,
                                    T dt) {
    const int i0 = blockIdx.x * blockDim.x + threadIdx.x, j0 = blockIdx.y, k0 = blockIdx.z;
    if (i0 >= m.Nx) return;
    const int H = m.H, k = k0 + H;
    const size_t sx = m.sx, sxy = (size_t)m.sx * m.sy;
    const size_t c = lin(i0 + H, j0 + H, k, m.sx, m.sy);
    const T idx = m.idx, idy = m.idy;
    const T idzk = m.m_idz(k), idzf_k = m.m_idzf(k), idzf_kp = m.m_idzf(k + 1);
    const T wck = T(Wc[k]), wckm = T(Wc[k - 1]), wfk = T(Wf[k]), wfkp = T(Wf[k + 1]);
    const T q = T(0.25);
#define N(o) n[c + (o)]
    {
        const T fxE = N(0) * (u[c + 1] - u[c]) * idx;
        const T fxW = N(-1) * (u[c] - u[c - 1]) * idx;
        const T eN = q * (((N(-1) + N(0)) + N(sx - 1)) + N(sx));
        const T eS = q * (((N(-1 - sx) + N(-sx)) + N(-1)) + N(0));
        const T fyN = eN * (u[c + sx] - u[c]) * idy;
        const T fyS = eS * (u[c] - u[c - sx]) * idy;
        const T eT = q * (((N(-1) + N(0)) + N(sxy - 1)) + N(sxy));
        const T eB = q * (((N(-1 - sxy) + N(-sxy)) + N(-1)) + N(0));
        const T fzT = eT * (u[c + sxy] - u[c]) * idzf_kp;
        const T fzB = eB * (u[c] - u[c - sxy]) * idzf_k;
        const T t = ((fxE - fxW) * idx + (fyN - fyS) * idy) + (wfkp * fzT - wfk * fzB) * idzk / wck;
        us[c] = us[c] + dt * t;
    }
    {
        const T fyN = N(0) * (v[c + sx] - v[c]) * idy;
        const T fyS = N(-sx) * (v[c] - v[c - sx]) * idy;
        const T eE = q * (((N(-sx) + N(1 - sx)) + N(0)) + N(1));
        const T eW = q * (((N(-1 - sx) + N(-sx)) + N(-1)) + N(0));
        const T fxE = eE * (v[c + 1] - v[c]) * idx;
        const T fxW = eW * (v[c] - v[c - 1]) * idx;
        const T eT = q * (((N(-sx) + N(0)) + N(sxy - sx)) + N(sxy));
        const T eB = q * (((N(-sx - sxy) + N(-sxy)) + N(-sx)) + N(0));
        const T fzT = eT * (v[c + sxy] - v[c]) * idzf_kp;
        const T fzB = eB * (v[c] - v[c - sxy]) * idzf_k;
        const T t = ((fxE - fxW) * idx + (fyN - fyS) * idy) + (wfkp * fzT - wfk * fzB) * idzk / wck;
        vs[c] = vs[c] + dt * t;
    }
    {
        const T fzT = N(0) * (w[c + sxy] - w[c]) * idzk;
        const T fzB = N(-sxy) * (w[c] - w[c - sxy]) * m.m_idz(k - 1);
        const T eE = q * (((N(-sxy) + N(1 - sxy)) + N(0)) + N(1));
        const T eW = q * (((N(-1 - sxy) + N(-sxy)) + N(-1)) + N(0));
        const T fxE = eE * (w[c + 1] - w[c]) * idx;
        const T fxW = eW * (w[c] - w[c - 1]) * idx;
        const T eN = q * (((N(-sxy) + N(sx - sxy)) + N(0)) + N(sx));
        const T eS = q * (((N(-sx - sxy) + N(-sxy)) + N(-sx)) + N(0));
        const T fyN = eN * (w[c + sx] - w[c]) * idy;
        const T fyS = eS * (w[c] - w[c - sx]) * idy;
        const T t = ((fxE - fxW) * idx + (fyN - fyS) * idy) + (wck * fzT - wckm * fzB) * idzf_k / wfk;
        ws[c] = ws[c] + dt * t;
    }
#undef N
    {
        const T h = T(0.5);
        const T *K = kap;
        const T fxE = h * (K[c] + K[c + 1]) * (th[c + 1] - th[c]) * idx;
        const T fxW = h * (K[c - 1] + K[c]) * (th[c] - th[c - 1]) * idx;
        const T fyN = h * (K[c] + K[c + sx]) * (th[c + sx] - th[c]) * idy;
        const T fyS = h * (K[c - sx] + K[c]) * (th[c] - th[c - sx]) * idy;
        const T fzT = h * (K[c] + K[c + sxy]) * (th[c + sxy] - th[c]) * idzk;
        const T fzB = h * (K[c - sxy] + K[c]) * (th[c] - th[c - sxy]) * idzk;
        const T t = ((fxE - fxW) * idx + (fyN - fyS) * idy) + (wfkp * fzT - wfk * fzB) * idzk / wck;
        tn[c] = tn[c] + dt * t;
    }
}

// κt = νt/Pr_t over the whole storage (_divide_storage!, src/operators/sgs_closures.jl:369-378)
template <typename T>
__global__ void divide_kernel(T *__restrict__ out, const T *__restrict__ a, T d, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = a[i] / d;
}

template <typename T>
void launch_smag_apply_weighted(hipStream_t st, const Metrics<T> &m, T *us, T *vs, T *ws, T *tn, const T *u,
                                const T *v, const T *w, const T *th, const T *nut, T *kap, const double *Wc,
                                const double *Wf, double Pr_t, double dt) {
    const size_t n = (size_t)m.sx * m.sy * m.sz;
    hipLaunchKernelGGL(divide_kernel<T>, dim3(1024), dim3(256), 0, st, kap, nut, T(Pr_t), n);
    dim3 grid((m.Nx + 127) / 128, m.Ny, m.Nz);
    hipLaunchKernelGGL(smag_apply_w_kernel<T>, grid, dim3(128), 0, st, us, vs, ws, tn, u, v, w, th, nut,
                       (const T *)kap, Wc, Wf, m, T(dt));
}
template void launch_smag_apply_weighted<double>(hipStream_t, const Metrics<double> &, double *, double *, double *,
                                                 double *, const double *, const double *, const double *,
                                                 const double *, const double *, double *, const double *,
                                                 const double *, double, double);
template void launch_smag_apply_weighted<float>(hipStream_t, const Metrics<float> &, float *, float *, float *,
                                                float *, const float *, const float *, const float *, const float *,
                                                const float *, float *, const double *, const double *, double,
                                                double);

template <typename T>
void launch_smag_nut(hipStream_t st, const Metrics<T> &m, T *nut, const T *u, const T *v, const T *w, const T *coef,
                     const T *idz_off) {
    dim3 grid((m.Nx + 127) / 128, m.Ny, m.Nz);
    hipLaunchKernelGGL(smag_nut_kernel<T>, grid, dim3(128), 0, st, nut, u, v, w, m, coef, idz_off);
}
template <typename T>
void launch_smag_apply(hipStream_t st, const Metrics<T> &m, T *us, T *vs, T *ws, T *so, const T *u, const T *v,
                       const T *w, const T *th, const T *nut, double dt, double bt, double sc) {
    dim3 grid((m.Nx + 127) / 128, m.Ny, m.Nz);
    hipLaunchKernelGGL(smag_apply_kernel<T>, grid, dim3(128), 0, st, us, vs, ws, so, u, v, w, th, nut, m, T(dt),
                       T(bt), T(sc));
}
template void launch_smag_nut<double>(hipStream_t, const Metrics<double> &, double *, const double *, const double *,
                                      const double *, const double *, const double *);
template void launch_smag_nut<float>(hipStream_t, const Metrics<float> &, float *, const float *, const float *,
                                     const float *, const float *, const float *);
template void launch_smag_apply<double>(hipStream_t, const Metrics<double> &, double *, double *, double *, double *,
                                        const double *, const double *, const double *, const double *,
                                        const double *, double, double, double);
template void launch_smag_apply<float>(hipStream_t, const Metrics<float> &, float *, float *, float *, float *,
                                       const float *, const float *, const float *, const float *, const float *,
                                       double, double, double);

}  // namespace aeo
