// Square batched LU for n <= 128 with the matrix held in registers as 16x16
// MFMA tiles (placeholder until the tiled kernel lands).
#include "magma_internal.h"
#include "getrf_kernels.h"

namespace mgi {

bool getrf_tile_supported(int n) { (void)n; return false; }

magma_int_t getrf_tile_launch(int, double**, int, int**, int*, int, hipStream_t)
{
    return MAGMA_ERR_NOT_SUPPORTED;
}

}  // namespace mgi
