// gp_run.hpp -- host-side launch sequences for one (P, C, real type); included by
// the instantiation files gp_inst_p*.hip.
#pragma once

#include <vector>

#include "gp_ctx.hpp"
#include "gp_fused.hpp"

namespace gpk {

// Tile of the fused kernel: the largest of (8,64), (4,64), (8,32), (4,32),
// (2,32) interior cells whose LDS staging fits in 60 KiB.
template <int P, int C, class R>
struct FusedTile {
  template <int TX, int TY>
  static constexpr bool fits = sizeof(FusedSmem<P, C, R, TX, TY>) <= 60 * 1024;
  static constexpr int TX = fits<8, 64> ? 8 : fits<4, 64> ? 4 : fits<8, 32> ? 8 : fits<4, 32> ? 4 : 2;
  static constexpr int TY = fits<8, 64> || fits<4, 64> ? 64 : 32;
};

template <int P, int C, class R>
int run_steps(gp_ctx *c, long nsteps) {
  const DevParams<R> *prm = (const DevParams<R> *)c->dprm;
  if (c->path == 0) {
    using T = FusedTile<P, C, R>;
    dim3 g((c->hp.mesh_y + T::TY - 1) / T::TY, (c->hp.mesh_x + T::TX - 1) / T::TX);
    for (long t = 0; t < nsteps; t++) {
      int i = c->cur, o = c->cur ^ 1;
      hipLaunchKernelGGL((k_fused<P, C, R, T::TX, T::TY>), g, dim3(256), 0, c->stream, prm, (const R *)c->phi[i],
                         (const R *)c->mu[i], (const R *)c->mulo[i], (R *)c->phi[o], (R *)c->mu[o], (R *)c->mulo[o]);
      c->cur = o;
    }
    GP_HIP_CHECK(hipGetLastError());
    return GP_OK;
  }
  dim3 blk(BX, BY);
  dim3 g_phase((c->ry - 4 + BX - 1) / BX, (c->rx - 4 + BY - 1) / BY);
  dim3 g_conc((c->hp.mesh_y + BX - 1) / BX, (c->hp.mesh_x + BY - 1) / BY);
  for (long t = 0; t < nsteps; t++) {
    const R *phi = (const R *)c->phi[c->cur], *mu = (const R *)c->mu[c->cur];
    R *phi_o = (R *)c->phi[c->cur ^ 1], *mu_o = (R *)c->mu[c->cur ^ 1];
    hipLaunchKernelGGL((k_phase<P, C, R>), g_phase, blk, 0, c->stream, prm, phi, mu, (R *)c->dphi);
    hipLaunchKernelGGL((k_conc<P, C, R>), g_conc, blk, 0, c->stream, prm, phi, mu, (const R *)c->dphi, phi_o, mu_o);
    c->cur ^= 1;
  }
  GP_HIP_CHECK(hipGetLastError());
  return GP_OK;
}

template <int P, int C, class R>
int run_stats(gp_ctx *c, gp_stats *st) {
  constexpr int NV = 3 * P + 3 * C + 1;
  const DevParams<R> *prm = (const DevParams<R> *)c->dprm;
  hipLaunchKernelGGL((k_stats<P, C, R>), dim3(c->nstat_blocks), dim3(256), 0, c->stream, prm,
                     (const R *)c->phi[c->cur], (const R *)c->mu[c->cur], c->dpart);
  GP_HIP_CHECK(hipGetLastError());
  std::vector<double> h((size_t)c->nstat_blocks * NV);
  GP_HIP_CHECK(hipMemcpyAsync(h.data(), c->dpart, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  GP_HIP_CHECK(hipStreamSynchronize(c->stream));
  for (int a = 0; a < P; a++) { st->phi_max[a] = -INFINITY; st->phi_min[a] = INFINITY; st->phi_sum[a] = 0; }
  for (int k = 0; k < C; k++) { st->mu_max[k] = -INFINITY; st->mu_min[k] = INFINITY; st->solute[k] = 0; }
  st->nonfinite = 0;
  for (int b = 0; b < c->nstat_blocks; b++) {
    const double *v = &h[(size_t)b * NV];
    for (int a = 0; a < P; a++) {
      st->phi_max[a] = fmax(st->phi_max[a], v[a]);
      st->phi_min[a] = fmin(st->phi_min[a], v[P + a]);
      st->phi_sum[a] += v[2 * P + a];
    }
    for (int k = 0; k < C; k++) {
      st->mu_max[k] = fmax(st->mu_max[k], v[3 * P + k]);
      st->mu_min[k] = fmin(st->mu_min[k], v[3 * P + C + k]);
      st->solute[k] += v[3 * P + 2 * C + k];
    }
    st->nonfinite += v[NV - 1];
  }
  return GP_OK;
}

}  // namespace gpk

#define GP_INSTANTIATE(P, C, R)                             \
  template int gpk::run_steps<P, C, R>(gp_ctx *, long);    \
  template int gpk::run_stats<P, C, R>(gp_ctx *, gp_stats *);
