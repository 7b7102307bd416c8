// cg_timer.hpp -- per-kernel timing of a run (host side).
#pragma once

#include <hip/hip_runtime.h>

// Per-kernel timing of a run: HIP events recorded on the launching stream
// around every kernel of the step (only when timing is switched on; the
// graph path never records them).
enum KernelId {
  KID_SCAN = 0, KID_FLUX_X1, KID_FLUX_X2, KID_FLUX_X3, KID_DIVERGENCE,
  KID_IMPLICIT_FWD, KID_IMPLICIT_BWD, KID_UPDATE, KID_BC_X1, KID_BC_X2,
  KID_BC_X3, KID_FLUX32, KID_KESSLER, KID_SATADJ, KID_NEW_DT, KID_IMPLICIT_FUSED, KID_MIX_FUSED, KID_LINES, KID_COUNT
};
static const char* kKernelNames[KID_COUNT] = {
    "k_hydrostatic_scan", "k_flux_x1",        "k_flux_h<x2>",
    "k_flux_h<x3>",       "k_divergence",     "k_implicit_fwd",
    "k_implicit_bwd",     "k_update",         "k_bc_x1",
    "k_bc_x2",            "k_bc_x3",          "k_flux32",
    "k_kessler",          "k_satadj_run_mlp", "k_new_dt",
    "k_implicit_fused",   "k_mix_fused",      "k_line_*"};

struct KTimer {
  bool on = false;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  double ms[KID_COUNT] = {0};
  long long cnt[KID_COUNT] = {0};
  void begin() {
    if (on) hipEventRecord(e0, s);
  }
  void end(int id) {
    if (!on) return;
    hipEventRecord(e1, s);
    hipEventSynchronize(e1);
    float t = 0.f;
    hipEventElapsedTime(&t, e0, e1);
    ms[id] += t;
    cnt[id] += 1;
  }
};
#define TIMED(tm, id, launch) \
  do {                        \
    if (tm) (tm)->begin();    \
    launch;                   \
    if (tm) (tm)->end(id);    \
  } while (0)

