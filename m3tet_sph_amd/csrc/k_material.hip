// The steps either side of the Stokes solve in the time loop, on the device:
// the material law (mfiles_M3TET/rock_prop.m:31-118) that turns the nodal
// temperature into the nodal density and viscosity the assembly reads, and
// the advective time-step limit (mfiles_M3TET/calculate_dt.m:40-56) from the
// solve's velocity.  Both are one pass over nodal arrays (HBM-bound, a few
// microseconds at c1).
#include "m3_internal.h"

#include <algorithm>
#include <cmath>

namespace {

struct RockArgs {
  int64_t n;
  const double *T, *X;
  double r_max;
  m3_rock_params p;
  double *visc, *dens;
};

__global__ void k_rock(RockArgs a) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const m3_rock_params &p = a.p;
  const double T = a.T[i];
  double rho = p.Dens0 * (1.0 - p.alpha * (T - p.T_ref));
  if (p.has_maxDens) rho = fmin(rho, p.maxDens);
  if (p.has_minDens) rho = fmax(rho, p.minDens);
  const double R = 8.314472, T0K = p.T_ref + 273.0;
  double f = 1.0;
  if (p.Ea > 0) {
    const double eta = p.Ea / (R * T0K);
    f = exp(eta * (T0K / (T + 273.0) - 1.0));
    f = fmax(fmin(f, 1e8), 1e-8);
  }
  double fP = f;
  if (p.Va > 0) {
    const double x = a.X[3 * i], y = a.X[3 * i + 1], z = a.X[3 * i + 2];
    const double depth = a.r_max - sqrt(x * x + y * y + z * z);
    const double P = fabs(p.g * p.Dens0 * p.L0 * depth);
    fP *= exp(P * p.Va / (R * T0K));
  }
  double fL = f;
  if (p.a_UM2LM > 0) fL *= p.a_UM2LM;
  double v = fmin(fP, fL) * p.Visc0;
  if (p.has_maxVisc) v = fmin(v, p.maxVisc);
  if (p.has_minVisc) v = fmax(v, p.minVisc);
  a.visc[i] = v;
  a.dens[i] = rho;
}

// per-block maxima of |U_x|, |U_y|, |U_z| (Cartesian, from the rotated U)
__global__ void k_umax(int64_t n, const double *frames, const double *U, double *part) {
  double m[3] = {0, 0, 0};
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double *R = frames + 9 * i;
    const double a = U[3 * i], b = U[3 * i + 1], c = U[3 * i + 2];
    for (int k = 0; k < 3; ++k) m[k] = fmax(m[k], fabs(R[3 * k] * a + R[3 * k + 1] * b + R[3 * k + 2] * c));
  }
  __shared__ double sh[3][256];
  for (int k = 0; k < 3; ++k) sh[k][threadIdx.x] = m[k];
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o)
      for (int k = 0; k < 3; ++k) sh[k][threadIdx.x] = fmax(sh[k][threadIdx.x], sh[k][threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x < 3) part[3 * blockIdx.x + threadIdx.x] = sh[threadIdx.x][0];
}

}  // namespace

namespace m3 {

void rock_prop(m3_ctx *c, const double *T, const m3_rock_params *rp, double *Visc, double *Dens) {
  M3_REQUIRE(!c->L.empty(), M3_E_STATE, "m3_rock_prop before m3_set_mesh");
  M3_REQUIRE(T && rp, M3_E_ARG, "null argument");
  M3_REQUIRE(!(rp->Ea > 0 && rp->Ea < 1000), M3_E_ARG,
             " Activation energy PHYSICS.Ea seems to be in kJ/mol. Convert to J/mol.");
  const int64_t n = c->L[0].nnod;
  double rmax = 0;
  for (int64_t i = 0; i < n; ++i) {
    const double *x = &c->h_coord[3 * i];
    rmax = std::max(rmax, std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]));
  }
  hipStream_t s = c->stream;
  DBuf<double> Td;
  Td.alloc(n);
  HIPCHECK(hipMemcpyAsync(Td.p, T, sizeof(double) * n, hipMemcpyHostToDevice, s));
  RockArgs a{n, Td.p, c->X.p, rmax, *rp, c->visc.p, c->dens.p};
  hipLaunchKernelGGL(k_rock, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
  HIPCHECK(hipGetLastError());
  if (Visc) HIPCHECK(hipMemcpyAsync(Visc, c->visc.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  if (Dens) HIPCHECK(hipMemcpyAsync(Dens, c->dens.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  HIPCHECK(hipStreamSynchronize(s));
  c->have_fields = true;
}

double calculate_dt(m3_ctx *c, const double *U, double dt_adv_limit, double dtmax) {
  M3_REQUIRE(!c->L.empty(), M3_E_STATE, "m3_calculate_dt before m3_set_mesh");
  M3_REQUIRE(U, M3_E_ARG, "null argument");
  const int64_t n = c->L[0].nnod;
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int64_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], c->h_coord[3 * i + k]);
      hi[k] = std::max(hi[k], c->h_coord[3 * i + k]);
    }
  hipStream_t s = c->stream;
  constexpr int NB = 256;
  DBuf<double> Ud, part;
  Ud.alloc(3 * n);
  part.alloc(3 * NB);
  HIPCHECK(hipMemcpyAsync(Ud.p, U, sizeof(double) * 3 * n, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_umax, dim3(NB), dim3(256), 0, s, n, c->frames.p, Ud.p, part.p);
  HIPCHECK(hipGetLastError());
  std::vector<double> h(3 * NB);
  HIPCHECK(hipMemcpyAsync(h.data(), part.p, sizeof(double) * 3 * NB, hipMemcpyDeviceToHost, s));
  HIPCHECK(hipStreamSynchronize(s));
  double mx[3] = {0, 0, 0};
  for (int b = 0; b < NB; ++b)
    for (int k = 0; k < 3; ++k) mx[k] = std::max(mx[k], h[3 * b + k]);
  if (mx[0] == 0 && mx[1] == 0 && mx[2] == 0) return dtmax;
  double dta = 1e300;
  for (int k = 0; k < 3; ++k)
    if (mx[k] > 0) dta = std::min(dta, (hi[k] - lo[k]) / mx[k]);
  return std::min(dtmax, dt_adv_limit * dta);
}

}  // namespace m3
