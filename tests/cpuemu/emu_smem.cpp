// TEST INFRASTRUCTURE: backing store of the emulated LDS (see hip/hip_runtime.h).
#include "hip/hip_runtime.h"
namespace hfx {
alignas(64) char hfx_smem[HFX_EMU_SMEM_BYTES];
}
