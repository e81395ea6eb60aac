// The instrumented variants of the unstructured-mesh render kernel (kCount:
// they also count point location's node visits, element tests and bytes),
// instantiated here so that they compile beside umesh.hip's plain variants.
#include "umesh_render.h"

namespace bnamd {
#define BNAMD_UMESH_COUNT_VARIANT(L, M)                                                                 \
  template void launch_render_umesh_t<L, M, true>(const UMeshDD &, const UMeshOpts &, const CameraDD &, \
                                                  const RenderDD &, int, float4 *, unsigned long long *, \
                                                  size_t, hipStream_t);
BNAMD_UMESH_COUNT_VARIANT(8, false)
BNAMD_UMESH_COUNT_VARIANT(4, false)
BNAMD_UMESH_COUNT_VARIANT(2, false)
BNAMD_UMESH_COUNT_VARIANT(1, false)
BNAMD_UMESH_COUNT_VARIANT(4, true)
BNAMD_UMESH_COUNT_VARIANT(1, true)
#undef BNAMD_UMESH_COUNT_VARIANT
}  // namespace bnamd
