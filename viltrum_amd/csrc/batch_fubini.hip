/* The batch pipeline for the Fubini form (FP32): (region dims, inner family, inner
   dims) = (2, blobs, 3), (2, lobes5, 5), (2, poly, 3), (1, poly, 2). */
#include "run_batch.h"

using FubiniBlobs3 = vmx::Fubini<float, vmx::Blobs<float>, 3, 2>;
using FubiniLobes5 = vmx::Fubini<float, vmx::Lobes5<float>, 5, 2>;
using FubiniPoly3 = vmx::Fubini<float, vmx::Poly<float>, 3, 2>;
using FubiniPoly2 = vmx::Fubini<float, vmx::Poly<float>, 2, 1>;
VMX_INSTANTIATE_BATCH(float, 2, FubiniBlobs3, vmx::FastMath)
VMX_INSTANTIATE_BATCH(float, 2, FubiniLobes5, vmx::FastMath)
VMX_INSTANTIATE_BATCH(float, 2, FubiniPoly3, vmx::FastMath)
VMX_INSTANTIATE_BATCH(float, 1, FubiniPoly2, vmx::FastMath)
