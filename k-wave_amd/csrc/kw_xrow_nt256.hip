// x-row kernels instantiated for 256-thread workgroups (templates in kw_xrow.inc).
#include "kw_xrow.inc"

namespace kw {
KW_ROW_INSTANTIATE(256)
}  // namespace kw
