// es_sieve_k2.hip -- sieve and batch kernel instances with 7..9 word masks (see
// es_sieve_dev.h).
#include "es_sieve_dev.h"

namespace es_dev {

KernelFn pick_kernel_p2(int F, uint32_t W, uint32_t G, uint32_t C) {
  switch (F) {
    case 7: return kernel_for<7>(W, G, C);
    case 8: return kernel_for<8>(W, G, C);
    case 9: return kernel_for<9>(W, G, C);
    default: return nullptr;
  }
}

GridBatchFn pick_grid_batch_p2(int F, uint32_t W, uint32_t G, uint32_t C) {
  switch (F) {
    case 7: return grid_batch_for<7>(W, G, C);
    case 8: return grid_batch_for<8>(W, G, C);
    case 9: return grid_batch_for<9>(W, G, C);
    default: return nullptr;
  }
}

RowsFn pick_rows_p2(int F, uint32_t C, uint32_t W, bool wave8) {
  switch (F) {
    case 7: return rows_kernel_for<7>(C, W, wave8);
    case 8: return rows_kernel_for<8>(C, W, wave8);
    case 9: return rows_kernel_for<9>(C, W, wave8);
    default: return nullptr;
  }
}

RowsBatchFn pick_rows_batch_p2(int F, uint32_t C, uint32_t W, bool wave8) {
  switch (F) {
    case 7: return rows_batch_for<7>(C, W, wave8);
    case 8: return rows_batch_for<8>(C, W, wave8);
    case 9: return rows_batch_for<9>(C, W, wave8);
    default: return nullptr;
  }
}

}  // namespace es_dev
