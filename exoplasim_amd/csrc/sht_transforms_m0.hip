// sht_transforms_m0.hip -- the transform kernels of MODE_F64.
#include "sht_transforms_impl.h"

namespace exo {
template int inv_mode<MODE_F64>(const Tables&, bool, const InvArgs&, hipStream_t);
template int inv_rt_mode<MODE_F64>(const Tables&, const InvArgs&, const InvArgs&, const InvArgs&,
                                  hipStream_t);
template int fwd_mode<MODE_F64>(const Tables&, bool, const FwdArgs&, void*, size_t, hipStream_t);
template size_t scr_mode<MODE_F64>(const Tables&, int, int);
}  // namespace exo
