// Positivity-preserving scalar advection (advect_positive!,
// src/AdvectionSchemes.jl:758-1160): the high-order face fluxes are split
// into a first-order upwind part L and an antidiffusive remainder A = F − L,
// and A is scaled on every face by the ratio of the cell it flows out of,
//
//   P⁻_c = Δt Σ_faces max(outflowing A, 0)/Δx_d
//   Q_c  = max(ϕ_c − Δt ∇·L, 0)
//   R_c  = P⁻_c > 0 ? min(1, Q_c/P⁻_c) : 1,
//
// so a forward-Euler update with the same Δt cannot take ϕ below zero.
//
// Three passes, the grid-wide dependencies of the algorithm:
//   1. pos_flux_kernel: F and L on the interior faces plus the closing face
//      of each axis (_b_rflux_*, _b_uflux_*; :858-909), one thread per face
//      triple, written to the workspace;
//   2. pos_ratio_kernel: R of every interior cell (_b_limit_ratio!, :1026-1053);
//   then R's halo (fill_halos! with the caller's BCs, or the one-layer
//   periodic wrap of _wrap_one_layer!, :944-958);
//   3. pos_div_kernel: each cell limits its six faces from (F, L, R) and
//      differences them into tend (_b_limit_*!, :1056-1081, fused with
//      _b_flux_div!, :912-928). Both cells of a face evaluate the same
//      expression, so the flux telescopes exactly and Σϕ is conserved.
//
// Workspace (storage-sized, precision T): Fx Fy Fz Lx Ly Lz R.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace aeo {

// Julia's max(x, 0) and min(1, x): NaN propagates
template <typename T> __device__ __forceinline__ T jl_max0(T x) { return x != x ? x : (x > T(0) ? x : T(0)); }
template <typename T> __device__ __forceinline__ T jl_min1(T x) { return x != x ? x : (x < T(1) ? x : T(1)); }

template <int S, int P, typename T>
__global__ void __launch_bounds__(256) pos_flux_kernel(T *__restrict__ Fx, T *__restrict__ Fy, T *__restrict__ Fz,
                                                       T *__restrict__ Lx, T *__restrict__ Ly, T *__restrict__ Lz,
                                                       const T *__restrict__ u, const T *__restrict__ v,
                                                       const T *__restrict__ w, const T *__restrict__ phi,
                                                       Metrics<T> m) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
    if (i > m.Nx || j > m.Ny) return;
    constexpr int R = SchemeInfo<S>::R;
    const uint32_t sx = m.sx, sxy = (uint32_t)m.sx * m.sy;
    const uint32_t c = lin(i + m.H, j + m.H, k + m.H, m.sx, m.sy);  // face c: west/south/bottom face of cell c
    constexpr uint32_t E = sizeof(T);  // byte offsets for face_value
    if (j < m.Ny && k < m.Nz) {
        const T vel = u[c];
        Fx[c] = vel * face_value<S, P, T>(phi, (c - R) * E, E, vel, m.teno_ct);
        Lx[c] = vel * (vel >= T(0) ? phi[c - 1] : phi[c]);
    }
    if (i < m.Nx && k < m.Nz) {
        const T vel = v[c];
        Fy[c] = vel * face_value<S, P, T>(phi, (c - R * sx) * E, sx * E, vel, m.teno_ct);
        Ly[c] = vel * (vel >= T(0) ? phi[c - sx] : phi[c]);
    }
    if (i < m.Nx && j < m.Ny) {
        const T vel = w[c];
        Fz[c] = vel * face_value<S, P, T>(phi, (c - R * sxy) * E, sxy * E, vel, m.teno_ct);
        Lz[c] = vel * (vel >= T(0) ? phi[c - sxy] : phi[c]);
    }
}

template <typename T>
__global__ void __launch_bounds__(256) pos_ratio_kernel(T *__restrict__ Rd, const T *__restrict__ Fx,
                                                        const T *__restrict__ Fy, const T *__restrict__ Fz,
                                                        const T *__restrict__ Lx, const T *__restrict__ Ly,
                                                        const T *__restrict__ Lz, const T *__restrict__ phi, T dt,
                                                        Metrics<T> m) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
    if (i >= m.Nx || j >= m.Ny) return;
    const int K = k + m.H;
    const uint32_t sx = m.sx, sxy = (uint32_t)m.sx * m.sy;
    const uint32_t c = lin(i + m.H, j + m.H, K, m.sx, m.sy);
    const T idx = m.idx, idy = m.idy, idz = m.m_idz(K);
    const T lxe = Lx[c + 1], lxw = Lx[c], lyn = Ly[c + sx], lys = Ly[c], lzt = Lz[c + sxy], lzb = Lz[c];
    const T axe = Fx[c + 1] - lxe, axw = Fx[c] - lxw;
    const T ayn = Fy[c + sx] - lyn, ays = Fy[c] - lys;
    const T azt = Fz[c + sxy] - lzt, azb = Fz[c] - lzb;
    const T Pm = dt * (((jl_max0(axe) + jl_max0(-axw)) * idx + (jl_max0(ayn) + jl_max0(-ays)) * idy) +
                       (jl_max0(azt) + jl_max0(-azb)) * idz);
    const T div_lo = ((lxe - lxw) * idx + (lyn - lys) * idy) + (lzt - lzb) * idz;
    const T Q = jl_max0(phi[c] - dt * div_lo);
    Rd[c] = Pm > T(0) ? jl_min1(Q / Pm) : T(1);
}

// the limited flux of a face: lo + c·a, a = F − lo, c the ratio of the cell
// below the face when a ≥ 0, of the cell above otherwise
template <typename T> __device__ __forceinline__ T pos_limited(T F, T lo, T r_below, T r_above) {
    const T a = F - lo;
    return lo + (a >= T(0) ? r_below : r_above) * a;
}

template <typename T>
__global__ void __launch_bounds__(256) pos_div_kernel(T *__restrict__ tend, const T *__restrict__ Fx,
                                                      const T *__restrict__ Fy, const T *__restrict__ Fz,
                                                      const T *__restrict__ Lx, const T *__restrict__ Ly,
                                                      const T *__restrict__ Lz, const T *__restrict__ Rd,
                                                      Metrics<T> m, int accumulate) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
    if (i >= m.Nx || j >= m.Ny) return;
    const int K = k + m.H;
    const uint32_t sx = m.sx, sxy = (uint32_t)m.sx * m.sy;
    const uint32_t c = lin(i + m.H, j + m.H, K, m.sx, m.sy);
    const T rc = Rd[c];
    const T fxw = pos_limited(Fx[c], Lx[c], Rd[c - 1], rc);
    const T fxe = pos_limited(Fx[c + 1], Lx[c + 1], rc, Rd[c + 1]);
    const T fys = pos_limited(Fy[c], Ly[c], Rd[c - sx], rc);
    const T fyn = pos_limited(Fy[c + sx], Ly[c + sx], rc, Rd[c + sx]);
    const T fzb = pos_limited(Fz[c], Lz[c], Rd[c - sxy], rc);
    const T fzt = pos_limited(Fz[c + sxy], Lz[c + sxy], rc, Rd[c + sxy]);
    const T d = ((fxe - fxw) * m.idx + (fyn - fys) * m.idy) + (fzt - fzb) * m.m_idz(K);
    tend[c] = accumulate ? tend[c] - d : -d;
}

template <int S, int P, typename T>
static void launch_flux(hipStream_t st, T *const ws[7], const T *u, const T *v, const T *w, const T *phi,
                        const Metrics<T> &m) {
    const dim3 blk(64, 4), grd((m.Nx + 1 + 63) / 64, (m.Ny + 1 + 3) / 4, m.Nz + 1);
    hipLaunchKernelGGL((pos_flux_kernel<S, P, T>), grd, blk, 0, st, ws[0], ws[1], ws[2], ws[3], ws[4], ws[5], u, v, w,
                       phi, m);
}

template <typename T>
int launch_advect_positive(hipStream_t st, const Metrics<T> &m, int scheme, int p, T *tend, const T *u, const T *v,
                           const T *w, const T *phi, T dt, T *const ws[7], const int *bcs, int accumulate) {
    switch (scheme) {
    case SCH_CENTERED: launch_flux<SCH_CENTERED, 0, T>(st, ws, u, v, w, phi, m); break;
    case SCH_WENO_JS5: launch_flux<SCH_WENO_JS5, 0, T>(st, ws, u, v, w, phi, m); break;
    case SCH_WENO_Z5:
        if (p == 1) launch_flux<SCH_WENO_Z5, 1, T>(st, ws, u, v, w, phi, m);
        else launch_flux<SCH_WENO_Z5, 2, T>(st, ws, u, v, w, phi, m);
        break;
    case SCH_WENO_Z7:
        if (p == 1) launch_flux<SCH_WENO_Z7, 1, T>(st, ws, u, v, w, phi, m);
        else launch_flux<SCH_WENO_Z7, 2, T>(st, ws, u, v, w, phi, m);
        break;
    case SCH_TENO5: launch_flux<SCH_TENO5, 0, T>(st, ws, u, v, w, phi, m); break;
    case SCH_TENO8: launch_flux<SCH_TENO8, 0, T>(st, ws, u, v, w, phi, m); break;
    default: return -1;
    }
    const dim3 blk(64, 4), grd((m.Nx + 63) / 64, (m.Ny + 3) / 4, m.Nz);
    hipLaunchKernelGGL(pos_ratio_kernel<T>, grd, blk, 0, st, ws[6], ws[0], ws[1], ws[2], ws[3], ws[4], ws[5], phi, dt,
                       m);
    // R's halo: the caller's BCs, or the periodic wrap (whose first layer is
    // the one-layer wrap of the reference; only that layer is read)
    if (bcs) launch_halo<T>(st, ws[6], m.Nx, m.Ny, m.Nz, m.H, LOC_CC, bcs[0], bcs[1], bcs[2]);
    else launch_halo<T>(st, ws[6], m.Nx, m.Ny, m.Nz, m.H, LOC_CC, BC_PERIODIC, BC_PERIODIC, BC_PERIODIC);
    hipLaunchKernelGGL(pos_div_kernel<T>, grd, blk, 0, st, tend, ws[0], ws[1], ws[2], ws[3], ws[4], ws[5], ws[6], m,
                       accumulate);
    return 0;
}
template int launch_advect_positive<double>(hipStream_t, const Metrics<double> &, int, int, double *, const double *,
                                            const double *, const double *, const double *, double,
                                            double *const[7], const int *, int);
template int launch_advect_positive<float>(hipStream_t, const Metrics<float> &, int, int, float *, const float *,
                                           const float *, const float *, const float *, float, float *const[7],
                                           const int *, int);

}  // namespace aeo
