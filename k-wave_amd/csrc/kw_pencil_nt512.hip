// Pencil kernels instantiated for 512-thread workgroups (templates in kw_pencil.inc).
#include "kw_pencil.inc"

namespace kw {
KW_PENCIL_INSTANTIATE(512)
}  // namespace kw
