// sht_launch.h -- per-mode launchers of the transform kernels, instantiated
// one mode per translation unit (sht_transforms_m<mode>.hip).
#pragma once
#include "sht_internal.h"

namespace exo {

template <int MODE>
int inv_mode(const Tables& tb, bool fused, const InvArgs& a, hipStream_t s);
template <int MODE>
int inv_rt_mode(const Tables& tb, const InvArgs& aq, const InvArgs& aj, const InvArgs& au,
                hipStream_t s);
template <int MODE>
int fwd_mode(const Tables& tb, bool fused, const FwdArgs& a, void* scr, size_t sb, hipStream_t s);
template <int MODE>
size_t scr_mode(const Tables& tb, int nmem, int nfield);

}  // namespace exo
