// x-row kernels instantiated for 64-thread workgroups (templates in kw_xrow.inc).
#include "kw_xrow.inc"

namespace kw {
KW_ROW_INSTANTIATE(64)
}  // namespace kw
