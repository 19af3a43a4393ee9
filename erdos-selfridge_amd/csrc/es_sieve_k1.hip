// es_sieve_k1.hip -- sieve and batch kernel instances with 4..6 word masks (see
// es_sieve_dev.h).
#include "es_sieve_dev.h"

namespace es_dev {

KernelFn pick_kernel_p1(int F, uint32_t W, uint32_t G, uint32_t C) {
  switch (F) {
    case 4: return kernel_for<4>(W, G, C);
    case 5: return kernel_for<5>(W, G, C);
    case 6: return kernel_for<6>(W, G, C);
    default: return nullptr;
  }
}

GridBatchFn pick_grid_batch_p1(int F, uint32_t W, uint32_t G, uint32_t C) {
  switch (F) {
    case 4: return grid_batch_for<4>(W, G, C);
    case 5: return grid_batch_for<5>(W, G, C);
    case 6: return grid_batch_for<6>(W, G, C);
    default: return nullptr;
  }
}

RowsFn pick_rows_p1(int F, uint32_t C, uint32_t W, bool wave8) {
  switch (F) {
    case 4: return rows_kernel_for<4>(C, W, wave8);
    case 5: return rows_kernel_for<5>(C, W, wave8);
    case 6: return rows_kernel_for<6>(C, W, wave8);
    default: return nullptr;
  }
}

RowsBatchFn pick_rows_batch_p1(int F, uint32_t C, uint32_t W, bool wave8) {
  switch (F) {
    case 4: return rows_batch_for<4>(C, W, wave8);
    case 5: return rows_batch_for<5>(C, W, wave8);
    case 6: return rows_batch_for<6>(C, W, wave8);
    default: return nullptr;
  }
}

}  // namespace es_dev
