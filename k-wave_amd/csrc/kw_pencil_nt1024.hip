// Pencil kernels instantiated for 1024-thread workgroups (templates in kw_pencil.inc).
#include "kw_pencil.inc"

namespace kw {
KW_PENCIL_INSTANTIATE(1024)
}  // namespace kw
