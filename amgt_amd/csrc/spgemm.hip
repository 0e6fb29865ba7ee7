// mBSR SpGEMM (placeholder until the symbolic/numeric kernels land).
#include "internal.hpp"

namespace amgt {

int spgemm(const Mbsr &A, const Mbsr &B, hipStream_t s, Mbsr *C) {
    (void)A;
    (void)B;
    (void)s;
    (void)C;
    set_error("SpGEMM not built yet");
    return AMGT_ERR_UNSUPPORTED;
}

}  // namespace amgt
