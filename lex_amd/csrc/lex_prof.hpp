// Optional per-kernel timing of the device step driver with HIP events recorded on the
// launch stream around each kernel.  Off by default; when on, every launch of the step
// driver is bracketed by an event pair taken from a fixed pool, and lex_profile_read()
// synchronises and sums the elapsed times per kernel kind.  Host-side state is
// process-global and not thread-safe (one profiling client at a time).
#pragma once
#include <hip/hip_runtime.h>
#include <vector>

namespace lex {

enum KernelId {
  KID_SMAG_DEFORM = 0,
  KID_SMAG_NM,
  KID_SMAG_K,
  KID_SMAG_UVW,
  KID_SMAG_SCALAR,
  KID_DIV_RHO,
  KID_RHS_UV,
  KID_RHS_W,
  KID_THETA_RHO,
  KID_ACOUSTIC,
  KID_SCALAR_UPDATE,
  KID_COPY,
  KID_STAGE_SCALARS,
  KID_STAGE_MOMENTUM,
  KID_SMAG_FUSED,
  KID_MOM_V,
  KID_MOM_W,
  KID_ACOUSTIC_VI,
  KID_ACOUSTIC2,
  KID_HEATING,
  KID_DAMPING,
  KID_SFC,
  KID_BICG_INIT,
  KID_BICG_PQ,
  KID_BICG_ST,
  KID_BICG_XR,
  KID_SMAG_ONE,
  KID_COUNT
};

inline const char* kernel_name(int id) {
  static const char* names[KID_COUNT] = {
      "k_smag_deform", "k_smag_nm",     "k_smag_k",    "k_smag_uvw",         "k_smag_scalar",
      "k_div_rho",     "k_rhs_uv",      "k_rhs_w",     "k_theta_rho",        "k_acoustic",
      "k_scalar_update", "k_copy",      "k_stage_scalars", "k_mom_u", "k_smag_tau", "k_mom_v", "k_mom_w",
      "k_acoustic_vi", "k_acoustic2", "k_newton_heating", "k_rayleigh",
      "k_atm_ocn_flux", "k_bicg_init", "k_bicg_pq", "k_bicg_st", "k_bicg_xr", "k_smag_fused"};
  return (id >= 0 && id < KID_COUNT) ? names[id] : "";
}

struct Prof {
  bool on = false;
  std::vector<hipEvent_t> ev;
  std::vector<int> ids;
  int used = 0;
  hipStream_t last_stream = nullptr;
};

inline Prof& prof() {
  static Prof p;
  return p;
}

inline int prof_start(int id, hipStream_t s) {
  Prof& p = prof();
  if (!p.on || p.used + 2 > (int)p.ev.size()) return -1;
  int slot = p.used;
  p.ids[slot / 2] = id;
  (void)hipEventRecord(p.ev[slot], s);
  return slot;
}

inline void prof_stop(int slot, hipStream_t s) {
  if (slot < 0) return;
  Prof& p = prof();
  (void)hipEventRecord(p.ev[slot + 1], s);
  p.used = slot + 2;
  p.last_stream = s;
}

#define LEX_LAUNCH(id, kern, grid, block, shm, stream, ...)                 \
  do {                                                                     \
    int _lex_slot = ::lex::prof_start((id), (stream));                     \
    hipLaunchKernelGGL(kern, grid, block, shm, stream, __VA_ARGS__);       \
    ::lex::prof_stop(_lex_slot, (stream));                                 \
  } while (0)

}  // namespace lex
