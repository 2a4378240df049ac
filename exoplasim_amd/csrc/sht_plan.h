// sht_plan.h -- the plan object behind exo_sht_plan_t (library-internal).
#pragma once
#include <mutex>

#include "sht_internal.h"

struct exo_sht_plan_s {
  exo::Tables tb;
  void* scratch = nullptr;  // partial direct sums of multi-block truncations
  size_t scratch_bytes = 0;
  void* stage = nullptr;    // device staging of the host drop-in calls
  size_t stage_bytes = 0;
  void* pk = nullptr;       // pre-split spectral operand of the split-BF16 inverse
  size_t pk_bytes = 0;
  int nlev = 0;
  // device address of the file-scope `scratch` array on the device the
  // plan was created on (exo_sht_roundtrip_noalloc)
  char* na_scratch = nullptr;
  // tile variants per precision mode, chosen by exo_sht_autotune
  int inv_var[4] = {0, 0, 0, 0}, uv_var[4] = {0, 0, 0, 0}, fwd_var[4] = {0, 0, 0, 0};
  std::mutex mu;
};

// grow a device buffer to at least `need` bytes (contents not kept)
int plan_grow(void** p, size_t* have, size_t need);
