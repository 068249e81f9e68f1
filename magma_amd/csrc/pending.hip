// Entry points declared in include/magma_mi355x.h whose kernels are not in
// this build yet: each reports MAGMA_ERR_NOT_IMPLEMENTED after the argument
// checks, so callers get an explicit error rather than a wrong answer.
#include "magma_internal.h"

extern "C" {

magma_int_t magma_dhgesv_iteref_gpu(magma_trans_t, magma_int_t, magma_int_t, magmaDouble_ptr, magma_int_t,
                                    magma_int_t*, magmaInt_ptr, magmaDouble_ptr, magma_int_t, magmaDouble_ptr,
                                    magma_int_t, magmaDouble_ptr, magmaFloat_ptr, magma_int_t*, magma_int_t*)
{
    return MAGMA_ERR_NOT_IMPLEMENTED;
}

magma_int_t magma_dbgesv_iteref_gpu(magma_trans_t, magma_int_t, magma_int_t, magmaDouble_ptr, magma_int_t,
                                    magma_int_t*, magmaInt_ptr, magmaDouble_ptr, magma_int_t, magmaDouble_ptr,
                                    magma_int_t, magmaDouble_ptr, magmaFloat_ptr, magma_int_t*, magma_int_t*)
{
    return MAGMA_ERR_NOT_IMPLEMENTED;
}

magma_int_t magma_dxgesv_gmres_gpu(magma_trans_t, magma_int_t, magma_int_t, magmaDouble_ptr, magma_int_t,
                                   magma_int_t*, magmaInt_ptr, magmaDouble_ptr, magma_int_t, magmaDouble_ptr,
                                   magma_int_t, magmaDouble_ptr, magmaFloat_ptr, magma_refinement_t,
                                   magma_refinement_t, magma_int_t*, magma_int_t*, real_Double_t*)
{
    return MAGMA_ERR_NOT_IMPLEMENTED;
}

magma_int_t magma_dxgesv_iteref_batched(magma_int_t, magma_int_t, double**, magma_int_t, double**, magma_int_t,
                                        double**, magma_int_t, magma_refinement_t, magma_refinement_t,
                                        magma_int_t, magma_int_t*, magma_int_t*, magma_int_t, magma_queue_t)
{
    return MAGMA_ERR_NOT_IMPLEMENTED;
}

}  // extern "C"
